// Test driver (host code only) for the air mesh's Lawson re-mesh (optcuts_amd/csrc/oc_air_mesh.hpp): runs the
// flips on a caller's triangulation in place and reports the smallest angle. Compiled by
// tests/test_air_mesh.py.
#include "../../optcuts_amd/csrc/oc_air_mesh.hpp"

#include <vector>

extern "C" long air_flips(int nP, const double* P, int nF, int* F, long maxFlips, double* minAngle)
{
    std::vector<double> p(P, P + 2 * (size_t)nP);
    std::vector<int> f(F, F + 3 * (size_t)nF);
    const long n = oc::lawson_flips(p, f, maxFlips);
    for (size_t k = 0; k < f.size(); ++k) F[k] = f[k];
    if (minAngle) *minAngle = oc::air_min_angle(p, f);
    return n;
}
