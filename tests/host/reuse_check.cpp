// Test driver (host code only): builds the Newton-system pattern, its slot
// maps and the multigrid structure for a mesh with one air mesh, then for a
// second air mesh twice, from scratch and reusing the first build's rows, and
// reports the first array that differs. Compiled by tests/test_host_reuse.py.
#include "../../optcuts_amd/csrc/oc_system_pattern.hpp"

#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Build {
    std::vector<int> vcPtr, vcSrc;
    oc::HostPattern P;
    oc::SystemMaps M;
    std::vector<oc::AmgHostLevel> L;
    std::vector<unsigned char> touched;
};

// Combined system of nV mesh vertices with triangles F (nF) and air triangles
// Fa (nFa, combined ids) over nVt vertices; fixed mesh vertices fixedV.
// storage (optional): levels of another build whose arrays the new levels
// take over (rewritten), as the library reuses the hierarchy two set-ups old.
void build(int nV, int nF, const int* F, int nVt, int nFa, const int* Fa, int nFixed, const int* fixedV,
           const Build* prev, Build& B, std::string& err, std::vector<oc::AmgHostLevel>* storage = nullptr)
{
    std::vector<int> Ft(F, F + 3 * (size_t)nF);
    Ft.insert(Ft.end(), Fa, Fa + 3 * (size_t)nFa);
    std::vector<unsigned char> fixed(nVt, 0);
    for (int k = 0; k < nFixed; ++k) fixed[fixedV[k]] = 1;
    B.touched.assign(nV, 0);
    for (long k = 0; k < 3L * nFa; ++k)
        if (Fa[k] < nV) B.touched[Fa[k]] = 1;
    std::vector<unsigned char> keep;
    if (prev) {
        keep.resize(nV);
        for (int v = 0; v < nV; ++v) keep[v] = !B.touched[v] && !prev->touched[v];
    }
    oc::build_vertex_lists(nVt, (int)(Ft.size() / 3), Ft.data(), B.vcPtr, B.vcSrc);
    oc::build_system_maps(nVt, Ft, fixed.data(), B.vcPtr, B.vcSrc, B.P, B.M, prev ? &prev->P : nullptr,
                          prev ? &prev->M : nullptr, prev ? keep.data() : nullptr, (int)keep.size());
    oc::AmgPrevious ap;
    if (prev)
        for (const auto& l : prev->L) ap.levels.push_back(&l);
    if (storage) B.L = std::move(*storage);
    // as the library does: the vertices of air triangles are aggregated last
    std::vector<unsigned char> late;
    if (nFa) {
        late.assign(nVt, 0);
        for (long k = 0; k < 3L * nFa; ++k) late[Fa[k]] = 1;
    }
    if (oc::amg_build_structure(nVt, B.P.rowPtr, B.P.colIdx, fixed, 64, 10, B.L, err, prev ? &ap : nullptr,
                                late.empty() ? nullptr : &late) != 0)
        B.L.clear();
}

template <typename T>
bool same(const std::vector<T>& a, const std::vector<T>& b)
{
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

// 0 when identical, else a code naming the first array that differs.
int compare(const Build& a, const Build& b, char* what, int whatLen)
{
    std::string w;
    if (!same(a.vcPtr, b.vcPtr) || !same(a.vcSrc, b.vcSrc)) w = "vertex lists";
    else if (!same(a.P.rowPtr, b.P.rowPtr) || !same(a.P.colIdx, b.P.colIdx)) w = "pattern";
    else if (!same(a.P.diagPos, b.P.diagPos)) w = "diagPos";
    else if (!same(a.M.ubRow, b.M.ubRow) || !same(a.M.ubCol, b.M.ubCol)) w = "upper blocks";
    else if (!same(a.M.ubPosU, b.M.ubPosU) || !same(a.M.ubPosL, b.M.ubPosL)) w = "upper block positions";
    else if (!same(a.M.ubCPtr, b.M.ubCPtr) || !same(a.M.ubCSrc, b.M.ubCSrc)) w = "slot map";
    else if (a.L.size() != b.L.size()) w = "level count";
    else
        for (size_t l = 0; l < a.L.size() && w.empty(); ++l) {
            const oc::AmgHostLevel &x = a.L[l], &y = b.L[l];
            const std::string lv = "level " + std::to_string(l) + " ";
            if (x.n != y.n || x.nc != y.nc || !same(x.rowPtr, y.rowPtr) || !same(x.colIdx, y.colIdx) ||
                !same(x.diagPos, y.diagPos))
                w = lv + "matrix pattern";
            else if (!same(x.agg, y.agg) || !same(x.aggPtr, y.aggPtr) || !same(x.aggMem, y.aggMem)) w = lv + "aggregates";
            else if (!same(x.pPtr, y.pPtr) || !same(x.pCol, y.pCol) || !same(x.pRow, y.pRow) ||
                     !same(x.pIsTent, y.pIsTent))
                w = lv + "prolongator pattern";
            else if (!same(x.psPtr, y.psPtr) || !same(x.psA, y.psA)) w = lv + "smoothing map";
            else if (!same(x.rPtr, y.rPtr) || !same(x.rP, y.rP)) w = lv + "restriction lists";
            else if (!same(x.tPtr, y.tPtr) || !same(x.tCol, y.tCol) || !same(x.tcPtr, y.tcPtr) ||
                     !same(x.tcA, y.tcA) || !same(x.tcP, y.tcP))
                w = lv + "A P map";
            else if (!same(x.acPtr, y.acPtr) || !same(x.acP, y.acP) || !same(x.acT, y.acT)) w = lv + "P^T A P map";
        }
    if (w.empty()) return 0;
    std::strncpy(what, w.c_str(), whatLen - 1);
    what[whatLen - 1] = 0;
    return 1;
}

}  // namespace

extern "C" {

// Air mesh 1 then air mesh 2 (nFa2 = 0: no air mesh), the second from scratch and incrementally (reusing the
// first build's rows, written into storage that held an earlier build); stats[0..3] = levels,
// kept mesh rows, level-0 aggregates, level-0 nodes of the second build.
int reuse_check(int nV, int nF, const int* F, int nFixed, const int* fixedV, int nVt1, int nFa1, const int* Fa1,
                int nVt2, int nFa2, const int* Fa2, int* stats, char* what, int whatLen)
{
    std::string err;
    Build b1, full, donor;
    build(nV, nF, F, nVt1, nFa1, Fa1, nFixed, fixedV, nullptr, b1, err);
    build(nV, nF, F, nVt2, nFa2, Fa2, nFixed, fixedV, nullptr, full, err);
    build(nV, nF, F, nVt1, nFa1, Fa1, nFixed, fixedV, nullptr, donor, err);
    // as the library does, the incremental build writes into storage that holds an earlier build's arrays
    Build inc = donor;
    build(nV, nF, F, nVt2, nFa2, Fa2, nFixed, fixedV, &b1, inc, err, &donor.L);
    int kept = 0;
    for (int v = 0; v < nV; ++v) kept += !b1.touched[v] && !full.touched[v];
    stats[0] = (int)full.L.size();
    stats[1] = kept;
    stats[2] = full.L.empty() ? 0 : full.L[0].nc;
    stats[3] = full.L.empty() ? 0 : full.L[0].n;
    if (full.L.empty()) {
        std::strncpy(what, err.c_str(), whatLen - 1);
        what[whatLen - 1] = 0;
        return 2;
    }
    return compare(full, inc, what, whatLen);
}

// nThreads host threads build the same system (with air mesh 1) at once, so
// that all but one find the worker pool busy and start their own threads;
// every result must equal a build on its own.
int concurrent_check(int nV, int nF, const int* F, int nFixed, const int* fixedV, int nVt, int nFa, const int* Fa,
                     int nThreads, char* what, int whatLen)
{
    std::string err;
    Build ref;
    build(nV, nF, F, nVt, nFa, Fa, nFixed, fixedV, nullptr, ref, err);
    std::vector<Build> out(nThreads);
    std::vector<std::string> errs(nThreads);
    std::vector<std::thread> th;
    for (int i = 0; i < nThreads; ++i)
        th.emplace_back([&, i] { build(nV, nF, F, nVt, nFa, Fa, nFixed, fixedV, nullptr, out[i], errs[i]); });
    for (auto& t : th) t.join();
    for (int i = 0; i < nThreads; ++i)
        if (int r = compare(ref, out[i], what, whatLen)) return r;
    return 0;
}

}
