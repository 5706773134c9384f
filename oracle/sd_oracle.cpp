// CPU oracle for the OptCuts geometry-step hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is a plain FP64 CPU restatement of the
// reference algorithm. It is the checker that the HIP path is compared with:
// only tests, the smoke check and the bench's cpu_baseline leg load it. The
// product library never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the OptCuts source tree). Pinning of each function (what makes a mistake in
// it fail a test) is listed in tests/test_oracle_*.py.
//
// Layout conventions of this C ABI (chosen for numpy interop, not the
// reference's Eigen col-major storage):
//   Vrest  nV x 3 row-major doubles      F     nF x 3 row-major int32
//   UV     nV x 2 row-major doubles (== the interleaved DOF vector [u0,v0,u1,v1,...])
//   DOF vectors (gradient, search direction) length 2*nV, interleaved.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <vector>

extern "C" {

// ---------------------------------------------------------------------------
// a1: per-triangle rest features.  TriMesh::computeFeatures, TriMesh.cpp:343-404.
// areaThres > 0 enables the air-mesh degeneracy clamp of TriMesh.cpp:373-383.
// avgEdgeLen follows igl::avg_edge_length (every face half-edge counted,
// ext/libigl/include/igl/avg_edge_length.cpp:17-30).
// Returns -1 if a triangle has zero rest area (TriMesh.cpp:368-372,399-402).
int orc_rest_features(int nV, const double* Vrest, int nF, const int* F, double areaThres,
                      double* triArea, double* triAreaSq, double* e0SqLen, double* e1SqLen,
                      double* e0dote1, double* e0SqLen_div_dbAreaSq, double* e1SqLen_div_dbAreaSq,
                      double* e0dote1_div_dbAreaSq, double* surfaceArea, double* avgEdgeLen)
{
    (void)nV;
    double sa = 0.0;
    bool invalid = false;
    for (int t = 0; t < nF; ++t) {
        const double* P1 = Vrest + 3 * F[3 * t + 0];
        const double* P2 = Vrest + 3 * F[3 * t + 1];
        const double* P3 = Vrest + 3 * F[3 * t + 2];
        const double a[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
        const double b[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
        const double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        triArea[t] = 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (triArea[t] == 0.0) invalid = true;
        if (triArea[t] < areaThres) {
            triArea[t] = areaThres;
            sa += areaThres;
            triAreaSq[t] = areaThres * areaThres;
            e0SqLen[t] = e1SqLen[t] = 4.0 / std::sqrt(3.0) * areaThres;
            e0dote1[t] = e0SqLen[t] / 2.0;
            e0SqLen_div_dbAreaSq[t] = e1SqLen_div_dbAreaSq[t] = 2.0 / std::sqrt(3.0) / areaThres;
            e0dote1_div_dbAreaSq[t] = e0SqLen_div_dbAreaSq[t] / 2.0;
        } else {
            sa += triArea[t];
            triAreaSq[t] = triArea[t] * triArea[t];
            e0SqLen[t] = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
            e1SqLen[t] = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
            e0dote1[t] = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
            e0SqLen_div_dbAreaSq[t] = e0SqLen[t] / 2. / triAreaSq[t];
            e1SqLen_div_dbAreaSq[t] = e1SqLen[t] / 2. / triAreaSq[t];
            e0dote1_div_dbAreaSq[t] = e0dote1[t] / 2. / triAreaSq[t];
        }
    }
    *surfaceArea = sa;
    double el = 0.0;
    long cnt = 0;
    for (int t = 0; t < nF; ++t)
        for (int j = 0; j < 3; ++j) {
            const double* p = Vrest + 3 * F[3 * t + j];
            const double* q = Vrest + 3 * F[3 * t + (j + 1) % 3];
            const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
            el += std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            ++cnt;
        }
    *avgEdgeLen = cnt ? el / (double)cnt : 0.0;
    return invalid ? -1 : 0;
}

// ---------------------------------------------------------------------------
// a2: per-element symmetric Dirichlet energy.
// SymDirichletEnergy::getEnergyValPerElem, SymDirichletEnergy.cpp:24-46
// (same arithmetic order as :42-44). uniform != 0 gives w = 1.
void orc_energy_per_elem(int nF, const int* F, const double* UV, const double* triArea,
                         const double* triAreaSq, const double* e0SqLen, const double* e1SqLen,
                         const double* e0dote1, double surfaceArea, int uniform, double* E)
{
    for (int t = 0; t < nF; ++t) {
        const double* U1 = UV + 2 * F[3 * t + 0];
        const double* U2 = UV + 2 * F[3 * t + 1];
        const double* U3 = UV + 2 * F[3 * t + 2];
        const double U2m1[2] = {U2[0] - U1[0], U2[1] - U1[1]};
        const double U3m1[2] = {U3[0] - U1[0], U3[1] - U1[1]};
        const double area_U = 0.5 * (U2m1[0] * U3m1[1] - U2m1[1] * U3m1[0]);
        const double w = uniform ? 1.0 : (triArea[t] / surfaceArea);
        const double sq3 = U3m1[0] * U3m1[0] + U3m1[1] * U3m1[1];
        const double sq2 = U2m1[0] * U2m1[0] + U2m1[1] * U2m1[1];
        const double dt = U3m1[0] * U2m1[0] + U3m1[1] * U2m1[1];
        E[t] = w * (1.0 + triAreaSq[t] / area_U / area_U) *
               ((sq3 * e0SqLen[t] + sq2 * e1SqLen[t]) / 4 / triAreaSq[t] - dt * e0dote1[t] / 2 / triAreaSq[t]);
    }
}

// Energy::computeEnergyVal, Energy.cpp:35-40 (sum of the per-element values).
// Eigen's .sum() uses a packet-wise order; this sums sequentially, which
// differs only at rounding level.
double orc_energy(int nF, const int* F, const double* UV, const double* triArea, const double* triAreaSq,
                  const double* e0SqLen, const double* e1SqLen, const double* e0dote1, double surfaceArea,
                  int uniform)
{
    std::vector<double> E(nF);
    orc_energy_per_elem(nF, F, UV, triArea, triAreaSq, e0SqLen, e1SqLen, e0dote1, surfaceArea, uniform, E.data());
    double s = 0.0;
    for (double e : E) s += e;
    return s;
}

// ---------------------------------------------------------------------------
// Per-triangle gradient rows w (dLeft_k rightTerm + dRight_k leftTerm) of
// SymDirichletEnergy::computeGradient (SymDirichletEnergy.cpp:258-304), the
// same rows computeLocalGradient stores (:215-256). r: 3 x 2.
static void local_gradient(int t, const int* F, const double* UV, const double* triArea, const double* triAreaSq,
                           const double* e0SqLen, const double* e1SqLen, const double* e0dote1, double surfaceArea,
                           int uniform, double r[6])
{
    const int i1 = F[3 * t + 0], i2 = F[3 * t + 1], i3 = F[3 * t + 2];
    const double* U1 = UV + 2 * i1;
    const double* U2 = UV + 2 * i2;
    const double* U3 = UV + 2 * i3;
    const double U2m1[2] = {U2[0] - U1[0], U2[1] - U1[1]};
    const double U3m1[2] = {U3[0] - U1[0], U3[1] - U1[1]};
    const double area_U = 0.5 * (U2m1[0] * U3m1[1] - U2m1[1] * U3m1[0]);
    const double leftTerm = 1.0 + triAreaSq[t] / area_U / area_U;
    const double sq3 = U3m1[0] * U3m1[0] + U3m1[1] * U3m1[1];
    const double sq2 = U2m1[0] * U2m1[0] + U2m1[1] * U2m1[1];
    const double dt = U3m1[0] * U2m1[0] + U3m1[1] * U2m1[1];
    const double rightTerm = (sq3 * e0SqLen[t] + sq2 * e1SqLen[t]) / 4 / triAreaSq[t] -
                             dt * e0dote1[t] / 2 / triAreaSq[t];
    const double areaRatio = triAreaSq[t] / area_U / area_U / area_U;
    const double w = uniform ? 1.0 : (triArea[t] / surfaceArea);
    const double e01 = e0dote1[t], e0 = e0SqLen[t], e1 = e1SqLen[t], A2 = triAreaSq[t];

    // vertex 1: edge opposite is U3-U2
    {
        const double eo[2] = {U3[0] - U2[0], U3[1] - U2[1]};
        const double dL[2] = {areaRatio * eo[1], areaRatio * -eo[0]};
        const double dR[2] = {((e01 - e0) * U3m1[0] + (e01 - e1) * U2m1[0]) / 2.0 / A2,
                              ((e01 - e0) * U3m1[1] + (e01 - e1) * U2m1[1]) / 2.0 / A2};
        r[0] = w * (dL[0] * rightTerm + dR[0] * leftTerm);
        r[1] = w * (dL[1] * rightTerm + dR[1] * leftTerm);
    }
    // vertex 2: edge opposite is U1-U3
    {
        const double eo[2] = {U1[0] - U3[0], U1[1] - U3[1]};
        const double dL[2] = {areaRatio * eo[1], areaRatio * -eo[0]};
        const double dR[2] = {(e1 * U2m1[0] - e01 * U3m1[0]) / 2.0 / A2, (e1 * U2m1[1] - e01 * U3m1[1]) / 2.0 / A2};
        r[2] = w * (dL[0] * rightTerm + dR[0] * leftTerm);
        r[3] = w * (dL[1] * rightTerm + dR[1] * leftTerm);
    }
    // vertex 3: edge opposite is U2-U1
    {
        const double eo[2] = {U2[0] - U1[0], U2[1] - U1[1]};
        const double dL[2] = {areaRatio * eo[1], areaRatio * -eo[0]};
        const double dR[2] = {(e0 * U3m1[0] - e01 * U2m1[0]) / 2.0 / A2, (e0 * U3m1[1] - e01 * U2m1[1]) / 2.0 / A2};
        r[4] = w * (dL[0] * rightTerm + dR[0] * leftTerm);
        r[5] = w * (dL[1] * rightTerm + dR[1] * leftTerm);
    }
}

// a3: analytic gradient, scatter-add in triangle order, fixed rows zeroed.
// SymDirichletEnergy::computeGradient, SymDirichletEnergy.cpp:258-304.
void orc_gradient(int nV, int nF, const int* F, const double* UV, const double* triArea,
                  const double* triAreaSq, const double* e0SqLen, const double* e1SqLen,
                  const double* e0dote1, double surfaceArea, int uniform, int nFixed, const int* fixedV,
                  double* g)
{
    std::memset(g, 0, sizeof(double) * 2 * (size_t)nV);
    for (int t = 0; t < nF; ++t) {
        double r[6];
        local_gradient(t, F, UV, triArea, triAreaSq, e0SqLen, e1SqLen, e0dote1, surfaceArea, uniform, r);
        for (int k = 0; k < 3; ++k) {
            g[2 * F[3 * t + k] + 0] += r[2 * k];
            g[2 * F[3 * t + k] + 1] += r[2 * k + 1];
        }
    }
    for (int k = 0; k < nFixed; ++k) {
        g[2 * fixedV[k] + 0] = 0.0;
        g[2 * fixedV[k] + 1] = 0.0;
    }
}

// SymDirichletEnergy::computeLocalGradient, SymDirichletEnergy.cpp:215-256
// (weights triArea / surfaceArea): out nF x 3 x 2.
void orc_local_gradients(int nF, const int* F, const double* UV, const double* triArea, const double* triAreaSq,
                         const double* e0SqLen, const double* e1SqLen, const double* e0dote1, double surfaceArea,
                         double* out)
{
    for (int t = 0; t < nF; ++t)
        local_gradient(t, F, UV, triArea, triAreaSq, e0SqLen, e1SqLen, e0dote1, surfaceArea, 0, out + 6 * (size_t)t);
}

// SymDirichletEnergy::computeDivGradPerVert, SymDirichletEnergy.cpp:108-154
// (the STANDARD_DEVIATION_FILTERING branch it compiles): per vertex the mean of
// its triangles' local gradient rows (accumulated in triangle order), then
// sqrt(sum |row - mean|^2 / (count - 1)); 0 for a vertex in one triangle.
void orc_div_grad_per_vert(int nV, int nF, const int* F, const double* UV, const double* triArea,
                           const double* triAreaSq, const double* e0SqLen, const double* e1SqLen,
                           const double* e0dote1, double surfaceArea, double* out)
{
    std::vector<double> lg(6 * (size_t)nF), mean(2 * (size_t)nV, 0.0), sd(nV, 0.0);
    std::vector<int> cnt(nV, 0);
    orc_local_gradients(nF, F, UV, triArea, triAreaSq, e0SqLen, e1SqLen, e0dote1, surfaceArea, lg.data());
    for (int t = 0; t < nF; ++t)
        for (int i = 0; i < 3; ++i) {
            const int v = F[3 * t + i];
            mean[2 * v] += lg[6 * t + 2 * i];
            mean[2 * v + 1] += lg[6 * t + 2 * i + 1];
            cnt[v]++;
        }
    for (int v = 0; v < nV; ++v) {
        mean[2 * v] /= cnt[v];
        mean[2 * v + 1] /= cnt[v];
    }
    for (int t = 0; t < nF; ++t)
        for (int i = 0; i < 3; ++i) {
            const int v = F[3 * t + i];
            const double dx = lg[6 * t + 2 * i] - mean[2 * v], dy = lg[6 * t + 2 * i + 1] - mean[2 * v + 1];
            sd[v] += dx * dx + dy * dy;
        }
    for (int v = 0; v < nV; ++v) out[v] = cnt[v] == 1 ? 0.0 : std::sqrt(sd[v] / (cnt[v] - 1.0));
}

// The candidate filter of TriMesh::querySplit (TriMesh.cpp:556-633) for
// propagate = false. vNeighbor and edge2Tri as TriMesh::computeFeatures builds
// them (TriMesh.cpp:437-446); isBoundaryVert as TriMesh.cpp:2069-2083.
// Boundary splits: boundary vertices with more than two neighbours; interior
// splits: non-boundary vertices with more than two neighbours and no boundary
// neighbour. Each goes into a std::map keyed by -divGrad / vertWeight (a later
// vertex with an equal key replaces the earlier), and the first
// max(2, (int)(size^e)) in key order are the candidates (e = 0.8 for boundary
// splits, filterExpIn for interior ones). cand: the candidates in that order;
// returns their number; *nSorted = map size.
int orc_split_candidates(int nV, int nF, const int* F, const double* divGrad, const double* vertWeight,
                         int splitInterior, double filterExpIn, int* cand, int* nSorted)
{
    std::vector<std::set<int>> nbr(nV);
    std::set<std::pair<int, int>> edge2Tri;
    for (int t = 0; t < nF; ++t)
        for (int i = 0; i < 3; ++i) {
            const int a = F[3 * t + i], b = F[3 * t + (i + 1) % 3];
            edge2Tri.insert({a, b});
            nbr[a].insert(b);
            nbr[b].insert(a);
        }
    std::vector<char> bnd(nV, 0);
    for (int v = 0; v < nV; ++v)
        for (int w : nbr[v])
            if (!edge2Tri.count({v, w}) || !edge2Tri.count({w, v})) {
                bnd[v] = 1;
                break;
            }
    std::map<double, int> sorted;
    for (int v = 0; v < nV; ++v) {
        if (nbr[v].size() <= 2) continue;
        if (splitInterior) {
            if (bnd[v]) continue;
            bool toBound = false;
            for (int w : nbr[v])
                if (bnd[w]) {
                    toBound = true;
                    break;
                }
            if (toBound) continue;
        } else if (!bnd[v]) {
            continue;
        }
        sorted[-divGrad[v] / vertWeight[v]] = v;
    }
    *nSorted = (int)sorted.size();
    if (sorted.empty()) return 0;
    int amt = splitInterior ? static_cast<int>(std::pow(sorted.size(), filterExpIn))
                            : static_cast<int>(std::pow(sorted.size(), 0.8) * 1.0);
    if (amt < 2) amt = 2;
    int n = 0;
    for (const auto& kv : sorted) {
        cand[n++] = kv.second;
        if (n >= amt) break;
    }
    return n;
}

// The per-pair screen of TriMesh::queryMerge (TriMesh.cpp:801-905) for propagate = false, before its
// computeLocalLDec: a cohesive row [a', b', a'', b''] is a "zipper bottom" when a' == a'' (fork 0) or
// b' == b'' (fork 1); the two vertices to merge are row[1 - fork] and row[3 - fork]. Their incident triangles
// come from vNeighbor (ascending) and edge2Tri (TriMesh.cpp:836-848); the merged position starts at their
// midpoint; if that inverts an incident triangle (checkInversion, TriMesh.cpp:1710-1756: doubled area < 0), it is
// projected by the relaxation method for the half-planes of the opposite edges (at most 70 sweeps,
// eps = 1e-6 avgEdgeLen, built with both merge vertices at the midpoint, TriMesh.cpp:861-899) and the pair is
// dropped if the projected position still inverts one (:901-908). status: 0 candidate (merged = its position),
// 1 not a zipper bottom, 2 no feasible position. Returns the number of candidates.
int orc_merge_screen(int nV, int nF, const int* F, const double* UV, double avgEdgeLen, int nCoh, const int* cohE,
                     int* status, double* merged)
{
    std::vector<std::set<int>> nbr(nV);
    std::map<std::pair<int, int>, int> edge2Tri;
    for (int t = 0; t < nF; ++t)
        for (int i = 0; i < 3; ++i) {
            const int a = F[3 * t + i], b = F[3 * t + (i + 1) % 3];
            edge2Tri[{a, b}] = t;
            nbr[a].insert(b);
            nbr[b].insert(a);
        }
    std::vector<double> V(UV, UV + 2 * (size_t)nV);
    const auto inverted = [&](const std::vector<int>& tris) {
        for (int t : tris) {
            const int* f = F + 3 * (size_t)t;
            const double e0x = V[2 * f[1]] - V[2 * f[0]], e0y = V[2 * f[1] + 1] - V[2 * f[0] + 1];
            const double e1x = V[2 * f[2]] - V[2 * f[0]], e1y = V[2 * f[2] + 1] - V[2 * f[0] + 1];
            const double dbArea = e0x * e1y - e0y * e1x;
            if (dbArea < 0.0) return true;
        }
        return false;
    };
    int nCand = 0;
    for (int r = 0; r < nCoh; ++r) {
        const int* row = cohE + 4 * (size_t)r;
        int fork;
        if (row[0] == row[2]) fork = 0;
        else if (row[1] == row[3]) fork = 1;
        else {
            status[r] = 1;
            continue;
        }
        const int va = row[1 - fork], vb = row[3 - fork];
        std::vector<int> tris;
        int nFirst = 0;
        for (int pass = 0; pass < 2; ++pass) {
            const int v = pass ? vb : va;
            for (int w : nbr[v]) {
                auto it = edge2Tri.find({v, w});
                if (it != edge2Tri.end()) tris.push_back(it->second);
            }
            if (!pass) nFirst = (int)tris.size();
        }
        double mx = (V[2 * va] + V[2 * vb]) / 2.0, my = (V[2 * va + 1] + V[2 * vb + 1]) / 2.0;
        const double b0x = V[2 * va], b0y = V[2 * va + 1], b1x = V[2 * vb], b1y = V[2 * vb + 1];
        V[2 * va] = V[2 * vb] = mx;
        V[2 * va + 1] = V[2 * vb + 1] = my;
        bool ok = true;
        if (inverted(tris)) {
            // half-plane rows (unit normal, offset) of the edge opposite the merge vertex, one per triangle
            std::vector<double> nx(tris.size()), ny(tris.size()), c(tris.size());
            for (size_t k = 0; k < tris.size(); ++k) {
                const int t = tris[k], vm = (int)k < nFirst ? va : vb;
                for (int i = 0; i < 3; ++i)
                    if (F[3 * t + i] == vm) {
                        const int p1 = F[3 * t + (i + 1) % 3], p2 = F[3 * t + (i + 2) % 3];
                        const double x1 = V[2 * p1], y1 = V[2 * p1 + 1], x2 = V[2 * p2], y2 = V[2 * p2 + 1];
                        const double cx = y2 - y1, cy = x1 - x2;
                        const double len = std::sqrt(cx * cx + cy * cy);
                        nx[k] = cx / len;
                        ny[k] = cy / len;
                        c[k] = (x1 * y2 - y1 * x2) / len;
                        break;
                    }
            }
            const double eps = 1.0e-6 * avgEdgeLen;
            for (int iter = 0; iter < 70; ++iter) {
                double maxRes = -__DBL_MAX__;
                for (size_t k = 0; k < tris.size(); ++k) {
                    const double res = (nx[k] * mx + ny[k] * my) - c[k];
                    if (res > -eps) {
                        const double s = res + eps;
                        mx -= s * nx[k];
                        my -= s * ny[k];
                    }
                    if (res > maxRes) maxRes = res;
                }
                if (maxRes < 0.0) break;
            }
            V[2 * va] = V[2 * vb] = mx;
            V[2 * va + 1] = V[2 * vb + 1] = my;
            ok = !inverted(tris);
        }
        V[2 * va] = b0x;
        V[2 * va + 1] = b0y;
        V[2 * vb] = b1x;
        V[2 * vb + 1] = b1y;
        status[r] = ok ? 0 : 2;
        merged[2 * r] = mx;
        merged[2 * r + 1] = my;
        nCand += ok;
    }
    return nCand;
}

// ---------------------------------------------------------------------------
// a5: symmetric eigen-decomposition of a small dense matrix by the cyclic
// Jacobi method (Golub & Van Loan, Matrix Computations 4th ed., Alg. 8.5.3),
// eigenvalues returned in ascending order with matching eigenvector columns.
// The reference uses Eigen's tridiagonal-QR SelfAdjointEigenSolver; the
// projection built from it below is unique, so any accurate FP64 solver
// gives the same PSD matrix to rounding.
static void jacobi_eigh(int n, const double* Ain, double* lam, double* Vout)
{
    double A[36], V[36];
    std::memcpy(A, Ain, sizeof(double) * n * n);
    for (int i = 0; i < n * n; ++i) V[i] = 0.0;
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    double fro = 0.0;
    for (int i = 0; i < n * n; ++i) fro += A[i] * A[i];
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
        if (off <= 1e-36 * fro || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (apq == 0.0) continue;
                const double tau = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                // A <- J^T A J with J the (p,q) rotation
                for (int k = 0; k < n; ++k) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    int idx[6];
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx, idx + n, [&](int a, int b) { return A[a * n + a] < A[b * n + b]; });
    for (int j = 0; j < n; ++j) {
        lam[j] = A[idx[j] * n + idx[j]];
        for (int k = 0; k < n; ++k) Vout[k * n + j] = V[k * n + idx[j]];
    }
}

// IglUtils::makePD, IglUtils.hpp:71-90: if the smallest eigenvalue is >= 0
// the matrix is returned unchanged; otherwise every negative eigenvalue is
// set to 0 and V D V^T is rebuilt. Row-major n x n, n <= 6.
void orc_make_pd(int n, double* H)
{
    double lam[6], V[36];
    jacobi_eigh(n, H, lam, V);
    if (lam[0] >= 0.0) return;
    for (int i = 0; i < n; ++i) {
        if (lam[i] < 0.0) lam[i] = 0.0;
        else break;
    }
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += V[r * n + k] * lam[k] * V[c * n + k];
            H[r * n + c] = s;
        }
}

void orc_eigh(int n, const double* A, double* lam, double* V) { jacobi_eigh(n, A, lam, V); }

// ---------------------------------------------------------------------------
// a4: 6x6 element Hessian (row-major, DOF order u1x,u1y,u2x,u2y,u3x,u3y).
// SymDirichletEnergy::computeHessian element part, SymDirichletEnergy.cpp:438-523,
// followed by makePD (:525) when project != 0.
void orc_element_hessians(int nF, const int* F, const double* UV, const double* triArea,
                          const double* triAreaSq, const double* e0d, const double* e1d,
                          const double* e01d, double surfaceArea, int uniform, int project, double* Hout)
{
    for (int t = 0; t < nF; ++t) {
        const double* U1 = UV + 2 * F[3 * t + 0];
        const double* U2 = UV + 2 * F[3 * t + 1];
        const double* U3 = UV + 2 * F[3 * t + 2];
        const double U2m1[2] = {U2[0] - U1[0], U2[1] - U1[1]};
        const double U3m1[2] = {U3[0] - U1[0], U3[1] - U1[1]};
        const double area_U = 0.5 * (U2m1[0] * U3m1[1] - U2m1[1] * U3m1[0]);
        const double areaRatio = triAreaSq[t] / area_U / area_U / area_U;
        const double dAR = 3.0 / 2.0 * areaRatio / area_U;
        const double w = uniform ? 1.0 : (triArea[t] / surfaceArea);
        const double E0 = e0d[t], E1 = e1d[t], E01 = e01d[t];
        const double leftTerm = 1.0 + triAreaSq[t] / area_U / area_U;
        const double sq3 = U3m1[0] * U3m1[0] + U3m1[1] * U3m1[1];
        const double sq2 = U2m1[0] * U2m1[0] + U2m1[1] * U2m1[1];
        const double dt = U3m1[0] * U2m1[0] + U3m1[1] * U2m1[1];
        const double rightTerm = (sq3 * E0 + sq2 * E1) / 2. - dt * E01;

        const double eo1[2] = {U3[0] - U2[0], U3[1] - U2[1]};
        const double eo2[2] = {U1[0] - U3[0], U1[1] - U3[1]};
        const double eo3[2] = {U2[0] - U1[0], U2[1] - U1[1]};
        const double o[3][2] = {{eo1[1], -eo1[0]}, {eo2[1], -eo2[0]}, {eo3[1], -eo3[0]}};
        // dOrtho/dU = [[0,-1],[1,0]]
        const double R[2][2] = {{0.0, -1.0}, {1.0, 0.0}};

        double dL[3][2], dR[3][2];
        for (int k = 0; k < 3; ++k) { dL[k][0] = areaRatio * o[k][0]; dL[k][1] = areaRatio * o[k][1]; }
        dR[0][0] = (E01 - E0) * U3m1[0] + (E01 - E1) * U2m1[0];
        dR[0][1] = (E01 - E0) * U3m1[1] + (E01 - E1) * U2m1[1];
        dR[1][0] = E1 * U2m1[0] - E01 * U3m1[0];
        dR[1][1] = E1 * U2m1[1] - E01 * U3m1[1];
        dR[2][0] = E0 * U3m1[0] - E01 * U2m1[0];
        dR[2][1] = E0 * U3m1[1] - E01 * U2m1[1];

        // d2Right_ij scalars (:365,372,379,386,392,399) and the sign of the
        // areaRatio*dOrtho term in d2Left_ij (:370-371 +, :377-378 -, :390-391 +).
        const double d2R[3][3] = {{E0 + E1 - 2.0 * E01, E01 - E1, E01 - E0},
                                  {E01 - E1, E1, -E01},
                                  {E01 - E0, -E01, E0}};
        const double rsign[3][3] = {{0.0, 1.0, -1.0}, {0.0, 0.0, 1.0}, {0.0, 0.0, 0.0}};

        double* H = Hout + 36 * (size_t)t;
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) {
                double blk[2][2];
                for (int r = 0; r < 2; ++r)
                    for (int c = 0; c < 2; ++c) {
                        const double d2L = dAR * o[i][r] * o[j][c] + areaRatio * rsign[i][j] * R[r][c];
                        double v;
                        if (i == j) {
                            // :367-368 etc: d2L*right + dL_i dR_i^T + d2R*left*I + (dL_i dR_i^T)^T
                            v = d2L * rightTerm + dL[i][r] * dR[i][c] + (r == c ? d2R[i][j] * leftTerm : 0.0) +
                                dL[i][c] * dR[i][r];
                        } else {
                            // :373-374: d2L*right + dL_i dR_j^T + d2R*left*I + dR_i dL_j^T
                            v = d2L * rightTerm + dL[i][r] * dR[j][c] + (r == c ? d2R[i][j] * leftTerm : 0.0) +
                                dR[i][r] * dL[j][c];
                        }
                        blk[r][c] = w * v;
                    }
                for (int r = 0; r < 2; ++r)
                    for (int c = 0; c < 2; ++c) {
                        H[(2 * i + r) * 6 + (2 * j + c)] = blk[r][c];
                        H[(2 * j + c) * 6 + (2 * i + r)] = blk[r][c];
                    }
            }
        if (project) orc_make_pd(6, H);
    }
}

// ---------------------------------------------------------------------------
// a6: triplet emission in the reference order.
// SymDirichletEnergy::computeHessian(V,I,J), SymDirichletEnergy.cpp:527-548 with
// IglUtils::addBlockToMatrix (IglUtils.cpp:396-451) and
// IglUtils::addDiagonalToMatrix (IglUtils.cpp:361-394); scale multiplies every
// value (Optimizer.cpp:825 'V *= energyParams[eI]').
// Pass I/J/V = NULL to get the count. fixedV must be sorted ascending
// (std::set order).
int orc_hessian_triplets(int nF, const int* F, const double* Helem, int nFixed, const int* fixedV,
                         double scale, int* I, int* J, double* Vv)
{
    std::set<int> fixedSet(fixedV, fixedV + nFixed);
    int cnt = 0;
    for (int t = 0; t < nF; ++t) {
        int vInd[3];
        int nfree = 0;
        for (int k = 0; k < 3; ++k) {
            vInd[k] = fixedSet.count(F[3 * t + k]) ? -1 : F[3 * t + k];
            if (vInd[k] >= 0) ++nfree;
        }
        if (!nfree) continue;
        const double* H = Helem + 36 * (size_t)t;
        for (int a = 0; a < 3; ++a) {
            if (vInd[a] < 0) continue;
            for (int b = 0; b < 3; ++b) {
                if (vInd[b] < 0) continue;
                for (int r = 0; r < 2; ++r)
                    for (int c = 0; c < 2; ++c) {
                        if (Vv) {
                            Vv[cnt] = H[(2 * a + r) * 6 + (2 * b + c)] * scale;
                            I[cnt] = 2 * vInd[a] + r;
                            J[cnt] = 2 * vInd[b] + c;
                        }
                        ++cnt;
                    }
            }
        }
    }
    for (int v : fixedSet)
        for (int d = 0; d < 2; ++d) {
            if (Vv) {
                Vv[cnt] = 1.0 * scale;
                I[cnt] = J[cnt] = 2 * v + d;
            }
            ++cnt;
        }
    return cnt;
}

// ---------------------------------------------------------------------------
// a8: upper-triangle CSR over 2x2 vertex blocks, 0-based.
// LinSysSolver::set_pattern, LinSysSolver.hpp:37-135, with the -1 shift of
// EigenLibSolver::set_pattern (EigenLibSolver.cpp:34-35).
// nbrPtr/nbr: vNeighbor as CSR (each vertex's neighbor set, any order; it is
// sorted here as std::set would be). Pass ia=NULL to get nnz only.
int orc_set_pattern(int nV, const int* nbrPtr, const int* nbr, int nFixed, const int* fixedV, int* ia,
                    int* ja)
{
    std::vector<char> isFixed(nV, 0);
    for (int k = 0; k < nFixed; ++k) isFixed[fixedV[k]] = 1;
    std::vector<int> jav;
    std::vector<int> iav(2 * (size_t)nV + 1);
    iav[0] = 0;
    for (int v = 0; v < nV; ++v) {
        if (!isFixed[v]) {
            std::vector<int> nb(nbr + nbrPtr[v], nbr + nbrPtr[v + 1]);
            std::sort(nb.begin(), nb.end());
            nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
            std::vector<int> rowx = {2 * v, 2 * v + 1};
            for (int c : nb)
                if (!isFixed[c] && c > v) {
                    rowx.push_back(2 * c);
                    rowx.push_back(2 * c + 1);
                }
            for (int c : rowx) jav.push_back(c);
            iav[2 * v + 1] = iav[2 * v] + (int)rowx.size();
            for (size_t k = 1; k < rowx.size(); ++k) jav.push_back(rowx[k]);
            iav[2 * v + 2] = iav[2 * v + 1] + (int)rowx.size() - 1;
        } else {
            jav.push_back(2 * v);
            jav.push_back(2 * v + 1);
            iav[2 * v + 1] = iav[2 * v] + 1;
            iav[2 * v + 2] = iav[2 * v + 1] + 1;
        }
    }
    if (ia) {
        std::memcpy(ia, iav.data(), sizeof(int) * iav.size());
        std::memcpy(ja, jav.data(), sizeof(int) * jav.size());
    }
    return (int)jav.size();
}

// a9: LinSysSolver::update_a, LinSysSolver.hpp:138-159: a = 0, then for each
// triplet with i <= j, a[slot(i,j)] += v in triplet order. Returns -1 if a
// triplet (i<=j) has no slot (the reference asserts, compiled out in Release).
int orc_update_a(int n, const int* ia, const int* ja, int nT, const int* I, const int* J, const double* Vv,
                 double* a)
{
    const int nnz = ia[n];
    std::memset(a, 0, sizeof(double) * nnz);
    for (int k = 0; k < nT; ++k) {
        const int i = I[k], j = J[k];
        if (i > j) continue;
        const int* b = ja + ia[i];
        const int* e = ja + ia[i + 1];
        const int* p = std::lower_bound(b, e, j);
        if (p == e || *p != j) return -1;
        a[p - ja] += Vv[k];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// a10: sparse LDL^T solve of the SPD system given by its upper-triangle CSR
// (== lower CSC, as EigenLibSolver reads it, EigenLibSolver.cpp:32-37,71-108).
// Fill-reducing order: recursive graph bisection by BFS level sets (George,
// "Nested dissection of a regular finite element mesh", 1973). Factorization:
// up-looking simplicial LDL^T via the elimination tree (Davis, "Algorithm 849:
// A concise sparse Cholesky factorization package", ACM TOMS 2005).
// The reference orders with AMD; the direction is the same to rounding.

namespace {

struct Graph {
    int n;
    std::vector<int> ptr, adj;
};

void nd_order(const Graph& G, std::vector<int>& nodes, std::vector<int>& part, int myPart, int& nextPart,
              std::vector<int>& out, std::vector<int>& level)
{
    // nodes: vertices of this subgraph, all with part[v] == myPart.
    if (nodes.size() <= 32) {
        for (int v : nodes) out.push_back(v);
        return;
    }
    auto bfs = [&](int root, std::vector<int>& order) -> int {
        order.clear();
        order.push_back(root);
        level[root] = 0;
        int maxLev = 0;
        for (size_t h = 0; h < order.size(); ++h) {
            const int v = order[h];
            for (int p = G.ptr[v]; p < G.ptr[v + 1]; ++p) {
                const int u = G.adj[p];
                if (part[u] == myPart && level[u] < 0) {
                    level[u] = level[v] + 1;
                    maxLev = std::max(maxLev, level[u]);
                    order.push_back(u);
                }
            }
        }
        return maxLev;
    };
    std::vector<int> order;
    int root = nodes[0];
    int lastDepth = -1;
    for (int it = 0; it < 4; ++it) {  // pseudo-peripheral root search
        for (int v : nodes) level[v] = -1;
        const int d = bfs(root, order);
        if (d <= lastDepth) break;
        lastDepth = d;
        root = order.back();
    }
    for (int v : nodes) level[v] = -1;
    const int depth = bfs(root, order);
    if (order.size() < nodes.size()) {
        // disconnected: order the reached component and the rest separately
        const int pa = nextPart++, pb = nextPart++;
        std::vector<int> A, B;
        for (int v : nodes) {
            if (level[v] >= 0) { part[v] = pa; A.push_back(v); }
            else { part[v] = pb; B.push_back(v); }
        }
        nd_order(G, A, part, pa, nextPart, out, level);
        nd_order(G, B, part, pb, nextPart, out, level);
        return;
    }
    if (depth < 2) {
        for (int v : nodes) out.push_back(v);
        return;
    }
    // choose the level whose cumulative count first reaches half
    std::vector<int> cnt(depth + 1, 0);
    for (int v : nodes) cnt[level[v]]++;
    int acc = 0, sepLev = 1;
    for (int l = 0; l <= depth; ++l) {
        acc += cnt[l];
        if (2 * acc >= (int)nodes.size()) { sepLev = std::max(1, std::min(l, depth - 1)); break; }
    }
    const int pa = nextPart++, pb = nextPart++, ps = nextPart++;
    std::vector<int> A, B, S;
    for (int v : nodes) {
        if (level[v] < sepLev) { part[v] = pa; A.push_back(v); }
        else if (level[v] > sepLev) { part[v] = pb; B.push_back(v); }
        else { part[v] = ps; S.push_back(v); }
    }
    nd_order(G, A, part, pa, nextPart, out, level);
    nd_order(G, B, part, pb, nextPart, out, level);
    for (int v : S) out.push_back(v);
}

}  // namespace

struct OrcLDLT {
    int n = 0;
    std::vector<int> perm, iperm;  // perm[new] = old
    std::vector<int> parent, Lp, Li, Lnz;
    std::vector<double> Lx, D;
};

// Analyze + factorize. Returns a handle (NULL on failure: zero pivot).
void* orc_ldlt_factor(int n, const int* ia, const int* ja, const double* a)
{
    OrcLDLT* F = new OrcLDLT();
    F->n = n;
    // full symmetric adjacency (no diagonal)
    Graph G;
    G.n = n;
    std::vector<int> deg(n, 0);
    for (int i = 0; i < n; ++i)
        for (int p = ia[i]; p < ia[i + 1]; ++p)
            if (ja[p] != i) { deg[i]++; deg[ja[p]]++; }
    G.ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) G.ptr[i + 1] = G.ptr[i] + deg[i];
    G.adj.resize(G.ptr[n]);
    std::vector<int> fill(G.ptr.begin(), G.ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int p = ia[i]; p < ia[i + 1]; ++p)
            if (ja[p] != i) { G.adj[fill[i]++] = ja[p]; G.adj[fill[ja[p]]++] = i; }
    std::vector<int> nodes(n), part(n, 0), level(n, -1);
    for (int i = 0; i < n; ++i) nodes[i] = i;
    int nextPart = 1;
    F->perm.clear();
    F->perm.reserve(n);
    nd_order(G, nodes, part, 0, nextPart, F->perm, level);
    F->iperm.assign(n, 0);
    for (int k = 0; k < n; ++k) F->iperm[F->perm[k]] = k;

    // permuted upper part by columns: for new column k, entries (i<=k) of C = P A P^T
    std::vector<int> Cp(n + 1, 0), Ci;
    std::vector<double> Cx;
    {
        std::vector<int> cnt(n, 0);
        for (int i = 0; i < n; ++i)
            for (int p = ia[i]; p < ia[i + 1]; ++p) {
                const int r = F->iperm[i], c = F->iperm[ja[p]];
                cnt[std::max(r, c)]++;
            }
        for (int k = 0; k < n; ++k) Cp[k + 1] = Cp[k] + cnt[k];
        Ci.resize(Cp[n]);
        Cx.resize(Cp[n]);
        std::vector<int> pos(Cp.begin(), Cp.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int p = ia[i]; p < ia[i + 1]; ++p) {
                const int r = F->iperm[i], c = F->iperm[ja[p]];
                const int col = std::max(r, c), row = std::min(r, c);
                Ci[pos[col]] = row;
                Cx[pos[col]++] = a[p];
            }
    }
    // symbolic
    F->parent.assign(n, -1);
    F->Lnz.assign(n, 0);
    std::vector<int> flag(n, -1);
    for (int k = 0; k < n; ++k) {
        flag[k] = k;
        for (int p = Cp[k]; p < Cp[k + 1]; ++p) {
            int i = Ci[p];
            if (i >= k) continue;
            for (; flag[i] != k; i = F->parent[i]) {
                if (F->parent[i] == -1) F->parent[i] = k;
                F->Lnz[i]++;
                flag[i] = k;
            }
        }
    }
    F->Lp.assign(n + 1, 0);
    for (int k = 0; k < n; ++k) F->Lp[k + 1] = F->Lp[k] + F->Lnz[k];
    F->Li.resize(F->Lp[n]);
    F->Lx.resize(F->Lp[n]);
    F->D.assign(n, 0.0);
    // numeric (up-looking)
    std::vector<double> Y(n, 0.0);
    std::vector<int> pattern(n);
    std::fill(flag.begin(), flag.end(), -1);
    std::fill(F->Lnz.begin(), F->Lnz.end(), 0);
    for (int k = 0; k < n; ++k) {
        Y[k] = 0.0;
        int top = n;
        flag[k] = k;
        for (int p = Cp[k]; p < Cp[k + 1]; ++p) {
            int i = Ci[p];
            Y[i] += Cx[p];
            int len = 0;
            for (; flag[i] != k; i = F->parent[i]) {
                pattern[len++] = i;
                flag[i] = k;
            }
            while (len > 0) pattern[--top] = pattern[--len];
        }
        F->D[k] = Y[k];
        Y[k] = 0.0;
        for (; top < n; ++top) {
            const int i = pattern[top];
            const double yi = Y[i];
            Y[i] = 0.0;
            const int p2 = F->Lp[i] + F->Lnz[i];
            for (int p = F->Lp[i]; p < p2; ++p) Y[F->Li[p]] -= F->Lx[p] * yi;
            const double lki = yi / F->D[i];
            F->D[k] -= lki * yi;
            F->Li[p2] = k;
            F->Lx[p2] = lki;
            F->Lnz[i]++;
        }
        if (F->D[k] == 0.0 || !std::isfinite(F->D[k])) {
            delete F;
            return nullptr;
        }
    }
    return F;
}

void orc_ldlt_solve(void* h, const double* b, double* x)
{
    const OrcLDLT* F = static_cast<const OrcLDLT*>(h);
    const int n = F->n;
    std::vector<double> y(n);
    for (int k = 0; k < n; ++k) y[k] = b[F->perm[k]];
    for (int j = 0; j < n; ++j)
        for (int p = F->Lp[j]; p < F->Lp[j + 1]; ++p) y[F->Li[p]] -= F->Lx[p] * y[j];
    for (int j = 0; j < n; ++j) y[j] /= F->D[j];
    for (int j = n - 1; j >= 0; --j)
        for (int p = F->Lp[j]; p < F->Lp[j + 1]; ++p) y[j] -= F->Lx[p] * y[F->Li[p]];
    for (int k = 0; k < n; ++k) x[F->perm[k]] = y[k];
}

long orc_ldlt_nnz(void* h) { return static_cast<OrcLDLT*>(h)->Lp.back(); }

void orc_ldlt_free(void* h) { delete static_cast<OrcLDLT*>(h); }

// ---------------------------------------------------------------------------
// a11: largest injectivity-safe step. SymDirichletEnergy::initStepSize,
// SymDirichletEnergy.cpp:551-610: per triangle, the first positive root of
// det(U + alpha V) = a alpha^2 + b alpha + c, min over triangles (in/out).
double orc_init_step_size(int nF, const int* F, const double* UV, const double* dir, double stepSize)
{
    for (int t = 0; t < nF; ++t) {
        const int i1 = F[3 * t + 0], i2 = F[3 * t + 1], i3 = F[3 * t + 2];
        const double U2m1[2] = {UV[2 * i2] - UV[2 * i1], UV[2 * i2 + 1] - UV[2 * i1 + 1]};
        const double U3m1[2] = {UV[2 * i3] - UV[2 * i1], UV[2 * i3 + 1] - UV[2 * i1 + 1]};
        const double V2m1[2] = {dir[2 * i2] - dir[2 * i1], dir[2 * i2 + 1] - dir[2 * i1 + 1]};
        const double V3m1[2] = {dir[2 * i3] - dir[2 * i1], dir[2 * i3 + 1] - dir[2 * i1 + 1]};
        const double a = V2m1[0] * V3m1[1] - V2m1[1] * V3m1[0];
        const double b = U2m1[0] * V3m1[1] - U2m1[1] * V3m1[0] + V2m1[0] * U3m1[1] - V2m1[1] * U3m1[0];
        const double c = U2m1[0] * U3m1[1] - U2m1[1] * U3m1[0];
        const double delta = b * b - 4.0 * a * c;
        double bound = stepSize;
        if (a > 0.0) {
            if ((b < 0.0) && (delta >= 0.0)) bound = 2.0 * c / (-b + std::sqrt(delta));
        } else if (a < 0.0) {
            if (b < 0.0) bound = 2.0 * c / (-b + std::sqrt(delta));
            else bound = (-b - std::sqrt(delta)) / 2.0 / a;
        } else {
            if (b < 0.0) bound = -c / b;
        }
        if (bound < stepSize) stepSize = bound;
    }
    return stepSize;
}

// TriMesh::checkInversion, TriMesh.cpp:1710-1756: 1 if every doubled signed
// UV area is >= 0.
int orc_check_inversion(int nF, const int* F, const double* UV)
{
    for (int t = 0; t < nF; ++t) {
        const int i0 = F[3 * t + 0], i1 = F[3 * t + 1], i2 = F[3 * t + 2];
        const double e0[2] = {UV[2 * i1] - UV[2 * i0], UV[2 * i1 + 1] - UV[2 * i0 + 1]};
        const double e1[2] = {UV[2 * i2] - UV[2 * i0], UV[2 * i2 + 1] - UV[2 * i0 + 1]};
        const double dbArea = e0[0] * e1[1] - e0[1] * e1[0];
        if (dbArea < 0.0) return 0;
    }
    return 1;
}

// a14: TriMesh::computeSeamSparsity, TriMesh.cpp:1542-1558. cohE is nC x 4
// row-major; boundary cohesive edges (any index < 0) are skipped.
double orc_seam_sparsity(int nC, const int* cohE, const double* Vrest, const double* UV, double avgEdgeLen,
                         double initSeamLen, int triSoup)
{
    const double thres = 1.0e-2;
    double s = 0.0;
    for (int c = 0; c < nC; ++c) {
        const int* e = cohE + 4 * c;
        if (std::min(std::min(e[0], e[1]), std::min(e[2], e[3])) < 0) continue;
        auto d2 = [&](int p, int q) {
            const double dx = UV[2 * p] - UV[2 * q], dy = UV[2 * p + 1] - UV[2 * q + 1];
            return std::sqrt(dx * dx + dy * dy);
        };
        if (!triSoup || d2(e[0], e[2]) / avgEdgeLen > thres || d2(e[1], e[3]) / avgEdgeLen > thres) {
            const double* p = Vrest + 3 * e[0];
            const double* q = Vrest + 3 * e[1];
            const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            s += std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    return s + initSeamLen;
}

// ---------------------------------------------------------------------------
// a12/a13: the Newton loop without scaffold, as Optimizer::precompute
// (Optimizer.cpp:154-201), Optimizer::solve (:203-261), solve_oneStep
// (:505-573), lineSearch (:575-652), stepForward (:654-673) and
// updateTargetGRes (:675-678) run it for a single SymDirichlet term with
// weight a0 and fixed vertex 0 (TriMesh.cpp:325-328).
//
// trace: (maxIter + 1) rows of 6 doubles:
//   row 0         : [E_total, E_SD, 0, 0, 0, 0]          (after precompute)
//   row k (k>=1)  : [E_total, E_SD, |g|^2, stepSize, halvings, stopped]
// Returns the number of Newton iterations that produced a row (>= 0), or a
// negative status: -1 factorization failed, -2 invalid rest mesh,
// -3 inverted initial UV.
// *exitReason: 1 gradient tolerance reached, 2 line search stop, 0 maxIter.
int orc_newton_run(int nV, const double* Vrest, int nF, const int* F, double* UV, double a0, int maxIter,
                   double relGL2Tol, int allowEDecRelTol, double* trace, int* exitReason)
{
    std::vector<double> tA(nF), tA2(nF), e0(nF), e1(nF), e01(nF), e0d(nF), e1d(nF), e01d(nF);
    double sa, ael;
    if (orc_rest_features(nV, Vrest, nF, F, 0.0, tA.data(), tA2.data(), e0.data(), e1.data(), e01.data(),
                          e0d.data(), e1d.data(), e01d.data(), &sa, &ael) != 0)
        return -2;
    if (!orc_check_inversion(nF, F, UV)) return -3;
    const int fixedV[1] = {0};
    // vNeighbor from F (TriMesh.cpp:436-447)
    std::vector<std::set<int>> nbs(nV);
    for (int t = 0; t < nF; ++t)
        for (int k = 0; k < 3; ++k) {
            const int a = F[3 * t + k], b = F[3 * t + (k + 1) % 3];
            nbs[a].insert(b);
            nbs[b].insert(a);
        }
    std::vector<int> nptr(nV + 1, 0), nbr;
    for (int v = 0; v < nV; ++v) {
        for (int u : nbs[v]) nbr.push_back(u);
        nptr[v + 1] = (int)nbr.size();
    }
    const int n = 2 * nV;
    const int nnz = orc_set_pattern(nV, nptr.data(), nbr.data(), 1, fixedV, nullptr, nullptr);
    std::vector<int> ia(n + 1), ja(nnz);
    orc_set_pattern(nV, nptr.data(), nbr.data(), 1, fixedV, ia.data(), ja.data());
    std::vector<double> aval(nnz);

    const double targetGRes = a0 * (double)(nV - 1) / (double)nV * relGL2Tol;
    auto energySD = [&](const double* X) {
        return orc_energy(nF, F, X, tA.data(), tA2.data(), e0.data(), e1.data(), e01.data(), sa, 0);
    };
    double ESD = energySD(UV);
    double lastE = a0 * ESD;
    trace[0] = lastE; trace[1] = ESD;
    for (int k = 2; k < 6; ++k) trace[k] = 0.0;

    std::vector<double> g(n), H(36 * (size_t)nF), UV0(n), d(n), rhs(n);
    const int nT = orc_hessian_triplets(nF, F, H.data(), 1, fixedV, a0, nullptr, nullptr, nullptr);
    std::vector<int> I(nT), J(nT);
    std::vector<double> Vt(nT);
    *exitReason = 0;
    int it = 0;
    for (; it < maxIter; ++it) {
        orc_gradient(nV, nF, F, UV, tA.data(), tA2.data(), e0.data(), e1.data(), e01.data(), sa, 0, 1, fixedV,
                     g.data());
        double sqn = 0.0;
        for (int k = 0; k < n; ++k) { g[k] *= a0; sqn += g[k] * g[k]; }
        if (sqn < targetGRes) { *exitReason = 1; break; }
        orc_element_hessians(nF, F, UV, tA.data(), tA2.data(), e0d.data(), e1d.data(), e01d.data(), sa, 0, 1,
                             H.data());
        orc_hessian_triplets(nF, F, H.data(), 1, fixedV, a0, I.data(), J.data(), Vt.data());
        orc_update_a(n, ia.data(), ja.data(), nT, I.data(), J.data(), Vt.data(), aval.data());
        void* fac = orc_ldlt_factor(n, ia.data(), ja.data(), aval.data());
        if (!fac) return -1;
        for (int k = 0; k < n; ++k) rhs[k] = -g[k];
        orc_ldlt_solve(fac, rhs.data(), d.data());
        orc_ldlt_free(fac);

        // line search
        bool stopped = false;
        double step = orc_init_step_size(nF, F, UV, d.data(), 1.0);
        step *= 0.99;
        std::memcpy(UV0.data(), UV, sizeof(double) * n);
        auto stepForward = [&](double s) {
            for (int k = 0; k < n; ++k) UV[k] = UV0[k] + s * d[k];
        };
        stepForward(step);
        ESD = energySD(UV);
        double testE = a0 * ESD;
        int halvings = 0;
        while (testE > lastE) {
            step /= 2.0;
            ++halvings;
            if (step == 0.0) { stopped = true; break; }
            stepForward(step);
            ESD = energySD(UV);
            testE = a0 * ESD;
        }
        while (!orc_check_inversion(nF, F, UV)) {
            step /= 2.0;
            ++halvings;
            if (step == 0.0) { stopped = true; break; }
            stepForward(step);
            ESD = energySD(UV);
            testE = a0 * ESD;
        }
        const double lastEDec = lastE - testE;
        if (allowEDecRelTol && (lastEDec / lastE < 1.0e-6 * step) && (step > 1.0e-3)) stopped = true;
        lastE = testE;
        double* row = trace + 6 * (size_t)(it + 1);
        row[0] = lastE; row[1] = ESD; row[2] = sqn; row[3] = step; row[4] = halvings; row[5] = stopped;
        if (stopped) { ++it; *exitReason = 2; break; }
    }
    return it;
}

}  // extern "C"
