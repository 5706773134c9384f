// Test driver (host code only) for the sanitizer builds of the host set-up: runs the checks of reuse_check.cpp
// and air_flip_check.cpp on cases read from a file, as one executable (so the sanitizer runtime is linked
// statically into it rather than loaded into a Python process). Compiled by tests/test_host_sanitizers.py.
//
// Case file: little-endian records, each an int32 header of 8 words then its arrays:
//   {1, nV, nF, nFixed, nVt1, nFa1, nVt2, nFa2} F[3nF] fixed[nFixed] Fa1[3nFa1] Fa2[3nFa2]   reuse_check
//   {2, nV, nF, nFixed, nVt, nFa, nThreads, 0}  F[3nF] fixed[nFixed] Fa[3nFa]                concurrent_check
//   {3, nP, nF, 0, 0, 0, 0, 0}                  P[2nP] (float64) F[3nF]                      air_flips
// Exit status 0 when every case passes; otherwise the failing case's number and message go to stderr.
#include "reuse_check.cpp"
#include "air_flip_check.cpp"

#include <cstdio>

namespace {

bool read_ints(std::FILE* f, std::vector<int>& v, long n)
{
    v.resize(n);
    return n == 0 || std::fread(v.data(), sizeof(int), n, f) == (size_t)n;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s cases.bin\n", argv[0]);
        return 2;
    }
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    int h[8];
    int caseNo = 0;
    char what[256];
    while (std::fread(h, sizeof(int), 8, f) == 8) {
        ++caseNo;
        what[0] = 0;
        int rc = 0;
        if (h[0] == 1) {
            std::vector<int> F, fx, Fa1, Fa2;
            if (!read_ints(f, F, 3L * h[2]) || !read_ints(f, fx, h[3]) || !read_ints(f, Fa1, 3L * h[5]) ||
                !read_ints(f, Fa2, 3L * h[7]))
                return 3;
            int stats[4];
            rc = reuse_check(h[1], h[2], F.data(), h[3], fx.data(), h[4], h[5], Fa1.data(), h[6], h[7], Fa2.data(),
                             stats, what, sizeof(what));
        } else if (h[0] == 2) {
            std::vector<int> F, fx, Fa;
            if (!read_ints(f, F, 3L * h[2]) || !read_ints(f, fx, h[3]) || !read_ints(f, Fa, 3L * h[5])) return 3;
            rc = concurrent_check(h[1], h[2], F.data(), h[3], fx.data(), h[4], h[5], Fa.data(), h[6], what,
                                  sizeof(what));
        } else if (h[0] == 3) {
            std::vector<double> P(2L * h[1]);
            std::vector<int> F;
            if (std::fread(P.data(), sizeof(double), P.size(), f) != P.size() || !read_ints(f, F, 3L * h[2]))
                return 3;
            double minAngle = 0.0;
            const long n = air_flips(h[1], P.data(), h[2], F.data(), 10000000L, &minAngle);
            if (n < 0 || !(minAngle > 0.0)) {
                std::snprintf(what, sizeof(what), "air_flips returned %ld, min angle %g", n, minAngle);
                rc = 1;
            }
        } else {
            std::fprintf(stderr, "case %d: unknown record %d\n", caseNo, h[0]);
            return 3;
        }
        if (rc) {
            std::fprintf(stderr, "case %d (record %d) failed: %s\n", caseNo, h[0], what);
            return 1;
        }
    }
    std::fclose(f);
    std::printf("%d cases passed\n", caseNo);
    return 0;
}
