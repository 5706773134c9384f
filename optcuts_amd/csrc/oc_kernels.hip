// Geometry-step kernels for gfx950 (CDNA4).
//
// Layout in HBM:
//   UV, gradient, search direction: interleaved per vertex (double2), the
//     reference's DOF order [u0, v0, u1, v1, ...].
//   Triangles: SoA vertex ids f0/f1/f2 and SoA rest features (coalesced).
//   Hessian: full symmetric 2x2-block CSR (one block per vertex pair,
//     row-major [xx, xy, yx, yy]).
//   Element buffers: elemG[t*3 + k] (double2), elemH[t*6 + blk] (4 values).
// Every sum is deterministic: element contributions are gathered per vertex
// or per block in triangle order, and grid reductions finish in the last
// block in block-index order.
#include "oc_common.hpp"
#include "oc_kernels.hpp"
#include "oc_sd_element.hpp"

#include <cfloat>
#include <type_traits>

namespace oc {

// ----------------------------------------------------------------------------
// Grid-wide sums. Each block publishes its compensated partials with plain
// stores; a one-block finisher kernel launched next on the same stream
// reduces them in block-index order (deterministic). The kernel boundary
// provides the cross-XCD visibility, so no kernel issues a device-scope
// fence (which on a multi-XCD part writes back and invalidates the L2).
template <int NV>
__device__ __forceinline__ void finish_csum(const double* partials, int nblk, double* out)
{
    CSum acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = CSum{0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            acc[k] = csum_add(acc[k], CSum{partials[(size_t)b * 2 * NV + 2 * k], partials[(size_t)b * 2 * NV + 2 * k + 1]});
    block_csum<NV>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = acc[k].s + acc[k].c;
}

template <int NV>
__global__ void __launch_bounds__(kBlock) k_finish_csum(const double* __restrict__ partials, int nblk, double* out)
{
    finish_csum<NV>(partials, nblk, out);
}

template <int NV>
__device__ __forceinline__ void publish_csum(CSum (&acc)[NV], double* partials)
{
    block_csum<NV>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            partials[(size_t)blockIdx.x * 2 * NV + 2 * k] = acc[k].s;
            partials[(size_t)blockIdx.x * 2 * NV + 2 * k + 1] = acc[k].c;
        }
}

// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_rest_features(int nF, const int* __restrict__ f0,
                                                          const int* __restrict__ f1, const int* __restrict__ f2,
                                                          const double* __restrict__ Vrest, double areaThres,
                                                          double* triArea, double* A2, double* E0, double* E1,
                                                          double* E01, double* partials)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    CSum acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (t < nF) {
        const int i[3] = {f0[t], f1[t], f2[t]};
        double P[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int d = 0; d < 3; ++d) P[k][d] = Vrest[3 * (size_t)i[k] + d];
        const double a[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
        const double b[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
        const double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        double area = 0.5 * sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        acc[2].s = (area == 0.0) ? 1.0 : 0.0;
        if (area < areaThres) {
            area = areaThres;
            A2[t] = areaThres * areaThres;
            const double e = 2.0 / sqrt(3.0) / areaThres;
            E0[t] = e;
            E1[t] = e;
            E01[t] = e / 2.0;
        } else {
            const double a2 = area * area;
            A2[t] = a2;
            E0[t] = (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) / 2. / a2;
            E1[t] = (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) / 2. / a2;
            E01[t] = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / 2. / a2;
        }
        triArea[t] = area;
        acc[0].s = area;
        double el = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int q = (k + 1) % 3;
            const double dx = P[k][0] - P[q][0], dy = P[k][1] - P[q][1], dz = P[k][2] - P[q][2];
            el += sqrt(dx * dx + dy * dy + dz * dz);
        }
        acc[1].s = el;
    }
    publish_csum<3>(acc, partials);
}

void launch_rest_features(hipStream_t s, int nF, const int* f0, const int* f1, const int* f2, const double* Vrest,
                          double areaThres, double* triArea, double* A2, double* E0, double* E1, double* E01,
                          double* partials, double* sums)
{
    const int g = grid_for(nF);
    hipLaunchKernelGGL(k_rest_features, dim3(g), dim3(kBlock), 0, s, nF, f0, f1, f2, Vrest, areaThres, triArea, A2,
                       E0, E1, E01, partials);
    hipLaunchKernelGGL(k_finish_csum<3>, dim3(1), dim3(kBlock), 0, s, partials, g, sums);
}

// ----------------------------------------------------------------------------
__device__ __forceinline__ RestT<double> load_rest(const TriSet& T, int t, bool isMesh)
{
    RestT<double> r;
    r.w = (isMesh && !T.uniformMesh) ? T.triArea[t] / T.surfaceArea[T.triChart ? T.triChart[t] : 0] : 1.0;
    r.A2 = T.A2[t];
    r.E0 = T.E0[t];
    r.E1 = T.E1[t];
    r.E01 = T.E01[t];
    return r;
}

// Element precision EP = double follows the reference arithmetic on absolute
// UV coordinates. EP = float (the mixed form) evaluates the element in FP32 on
// vertex positions relative to the first vertex (differences taken in FP64,
// so large UV offsets cost no accuracy) and accumulates energies in
// compensated FP64.
template <typename EP>
__device__ __forceinline__ void element_coords(double2 a, double2 b, double2 c, EP (&U)[3][2])
{
    if constexpr (std::is_same<EP, double>::value) {
        U[0][0] = a.x;
        U[0][1] = a.y;
        U[1][0] = b.x;
        U[1][1] = b.y;
        U[2][0] = c.x;
        U[2][1] = c.y;
    } else {
        U[0][0] = EP(0);
        U[0][1] = EP(0);
        U[1][0] = EP(b.x - a.x);
        U[1][1] = EP(b.y - a.y);
        U[2][0] = EP(c.x - a.x);
        U[2][1] = EP(c.y - a.y);
    }
}

template <typename EP>
__device__ __forceinline__ RestT<EP> rest_as(const RestT<double>& r)
{
    return RestT<EP>{EP(r.w), EP(r.A2), EP(r.E0), EP(r.E1), EP(r.E01)};
}

template <typename EP>
using Vec4 = typename std::conditional<std::is_same<EP, float>::value, float4, double4>::type;

template <typename EP, bool WANT_G, bool WANT_H>
__global__ void __launch_bounds__(kBlock) k_element(TriSet T, const double2* __restrict__ UV, double* Eelem,
                                                    double2* elemG, Vec4<EP>* elemH, int project, double* partials)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    CSum acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (t < T.nF) {
        const bool isMesh = t < T.nFmesh;
        const double2 a = UV[T.f0[t]], b = UV[T.f1[t]], c = UV[T.f2[t]];
        const RestT<EP> r = rest_as<EP>(load_rest(T, t, isMesh));
        EP U[3][2];
        element_coords<EP>(a, b, c, U);
        EP g[6], h[6][4];
        const double E = (double)sd_element<EP, WANT_H>(U, r, g, h);
        if (Eelem) Eelem[t] = E;
        acc[0].s = isMesh ? E : 0.0;
        acc[1].s = isMesh ? 0.0 : E;
        const double sc = isMesh ? T.sMesh : T.sAir;
        if (WANT_G) {
#pragma unroll
            for (int k = 0; k < 3; ++k)
                elemG[3 * (size_t)t + k] = make_double2(sc * (double)g[2 * k], sc * (double)g[2 * k + 1]);
        }
        if (WANT_H) {
            if (project) sd_project_psd<EP>(h);
            const EP se = (EP)sc;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                Vec4<EP> o;
                o.x = se * h[k][0];
                o.y = se * h[k][1];
                o.z = se * h[k][2];
                o.w = se * h[k][3];
                elemH[6 * (size_t)t + k] = o;
            }
        }
    }
    if (partials) publish_csum<2>(acc, partials);
}

template <typename EP>
void launch_element_t(hipStream_t s, const TriSet& T, const double* UV, double* Eelem, double* elemG, void* elemH,
                      int project, double* partials, double* esum)
{
    const dim3 grid(grid_for(T.nF)), blk(kBlock);
    const double2* uv = reinterpret_cast<const double2*>(UV);
    double* part = esum ? partials : nullptr;
    double2* eg = reinterpret_cast<double2*>(elemG);
    Vec4<EP>* eh = reinterpret_cast<Vec4<EP>*>(elemH);
    if (elemH)
        hipLaunchKernelGGL((k_element<EP, true, true>), grid, blk, 0, s, T, uv, Eelem, eg, eh, project, part);
    else if (elemG)
        hipLaunchKernelGGL((k_element<EP, true, false>), grid, blk, 0, s, T, uv, Eelem, eg, nullptr, project, part);
    else
        hipLaunchKernelGGL((k_element<EP, false, false>), grid, blk, 0, s, T, uv, Eelem, nullptr, nullptr, project,
                           part);
    if (esum) hipLaunchKernelGGL(k_finish_csum<2>, dim3(1), dim3(kBlock), 0, s, partials, (int)grid.x, esum);
}

void launch_element(hipStream_t s, const TriSet& T, const double* UV, double* Eelem, double* elemG, void* elemH,
                    int project, double* partials, double* esum, int f32)
{
    if (f32)
        launch_element_t<float>(s, T, UV, Eelem, elemG, elemH, project, partials, esum);
    else
        launch_element_t<double>(s, T, UV, Eelem, elemG, elemH, project, partials, esum);
}

// ----------------------------------------------------------------------------
// Reference-order triplets for one triangle block (mesh only; unscaled).
__global__ void __launch_bounds__(kBlock) k_triplets(TriSet T, const double2* __restrict__ UV,
                                                     const long long* __restrict__ tripOff,
                                                     const unsigned char* __restrict__ vFree, int* I, int* J,
                                                     double* V)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.nF) return;
    const unsigned char fm = vFree[t];
    if (!fm) return;
    const bool isMesh = t < T.nFmesh;
    const int vi[3] = {T.f0[t], T.f1[t], T.f2[t]};
    const double2 a = UV[vi[0]], b = UV[vi[1]], c = UV[vi[2]];
    const RestT<double> r = load_rest(T, t, isMesh);
    const double U[3][2] = {{a.x, a.y}, {b.x, b.y}, {c.x, c.y}};
    double g[6], h[6][4];
    sd_element<double, true>(U, r, g, h);
    sd_project_psd<double>(h);
    long long o = tripOff[t];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        if (!((fm >> p) & 1)) continue;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (!((fm >> q) & 1)) continue;
            const int lo = p < q ? p : q, hi = p < q ? q : p;
            const int id = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
#pragma unroll
            for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    V[o] = p <= q ? h[id][2 * rr + cc] : h[id][2 * cc + rr];
                    I[o] = 2 * vi[p] + rr;
                    J[o] = 2 * vi[q] + cc;
                    ++o;
                }
        }
    }
}

void launch_triplets(hipStream_t s, const TriSet& T, const double* UV, const long long* tripOff,
                     const unsigned char* vFree, int* I, int* J, double* V)
{
    hipLaunchKernelGGL(k_triplets, dim3(grid_for(T.nF)), dim3(kBlock), 0, s, T,
                       reinterpret_cast<const double2*>(UV), tripOff, vFree, I, J, V);
}

// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_gather_grad(int nV, const int* __restrict__ vcPtr,
                                                        const int* __restrict__ vcSrc,
                                                        const unsigned char* __restrict__ fixedMask,
                                                        const double2* __restrict__ elemG, double2* g,
                                                        double* partials)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    double sq[1] = {0.0};
    if (v < nV) {
        double2 s = make_double2(0.0, 0.0);
        if (!fixedMask[v]) {
            for (int k = vcPtr[v]; k < vcPtr[v + 1]; ++k) {
                const double2 e = elemG[vcSrc[k]];
                s.x += e.x;
                s.y += e.y;
            }
        }
        g[v] = s;
        sq[0] = s.x * s.x + s.y * s.y;
    }
    if (partials) {
        CSum c[1] = {CSum{sq[0], 0.0}};
        publish_csum<1>(c, partials);
    }
}

void launch_gather_grad(hipStream_t s, int nV, const int* vcPtr, const int* vcSrc, const unsigned char* fixedMask,
                        const double* elemG, double* g, double* partials, double* gsq)
{
    const int grid = grid_for(nV);
    hipLaunchKernelGGL(k_gather_grad, dim3(grid), dim3(kBlock), 0, s, nV, vcPtr, vcSrc, fixedMask,
                       reinterpret_cast<const double2*>(elemG), reinterpret_cast<double2*>(g),
                       gsq ? partials : nullptr);
    if (gsq) hipLaunchKernelGGL(k_finish_csum<1>, dim3(1), dim3(kBlock), 0, s, partials, grid, gsq);
}

// ----------------------------------------------------------------------------
// Seam length (TriMesh::computeSeamSparsity, TriMesh.cpp:1542-1558): each
// interior cohesive edge (no negative id) adds its rest length, in tri-soup
// mode only when one of its vertex pairs (0, 2), (1, 3) lies more than
// 1e-2 avgEdgeLen apart in UV. Per-block compensated sums, then one finisher.
__global__ void __launch_bounds__(kBlock) k_seam_len(int nC, const int4* __restrict__ cohE,
                                                     const double* __restrict__ Vrest, const double2* __restrict__ UV,
                                                     double avgEdgeLen, int triSoup, double* partials)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    CSum acc[1] = {{0.0, 0.0}};
    if (c < nC) {
        const int4 e = cohE[c];
        if (min(min(e.x, e.y), min(e.z, e.w)) >= 0) {
            bool take = !triSoup;
            if (!take) {
                const double2 a = UV[e.x], b = UV[e.z], p = UV[e.y], q = UV[e.w];
                const double d0 = sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y));
                const double d1 = sqrt((p.x - q.x) * (p.x - q.x) + (p.y - q.y) * (p.y - q.y));
                take = d0 / avgEdgeLen > 1.0e-2 || d1 / avgEdgeLen > 1.0e-2;
            }
            if (take) {
                const double* u = Vrest + 3 * (size_t)e.x;
                const double* w = Vrest + 3 * (size_t)e.y;
                const double dx = u[0] - w[0], dy = u[1] - w[1], dz = u[2] - w[2];
                acc[0].s = sqrt(dx * dx + dy * dy + dz * dz);
            }
        }
    }
    publish_csum<1>(acc, partials);
}

void launch_seam_len(hipStream_t s, int nC, const int* cohE, const double* Vrest, const double* UV, double avgEdgeLen,
                     int triSoup, double* partials, double* out)
{
    const int grid = grid_for(nC);
    hipLaunchKernelGGL(k_seam_len, dim3(grid), dim3(kBlock), 0, s, nC, reinterpret_cast<const int4*>(cohE), Vrest,
                       reinterpret_cast<const double2*>(UV), avgEdgeLen, triSoup, partials);
    hipLaunchKernelGGL(k_finish_csum<1>, dim3(1), dim3(kBlock), 0, s, partials, grid, out);
}

// ----------------------------------------------------------------------------
// Per-pair screen of the merge query (TriMesh::queryMerge, TriMesh.cpp:836-908), one lane per zipper-bottom
// cohesive pair: the two vertices to merge (va, vb) move to their midpoint; when that inverts one of their
// incident triangles (doubled signed area < 0, TriMesh.cpp:1710-1756), the midpoint is projected onto the
// half-planes of the edges opposite the merge vertex (relaxation method: at most 70 sweeps in list order, a
// violated or nearly active row moves the point to eps inside; half-planes built with both merge vertices at
// the midpoint) and the pair is infeasible when the result still inverts a triangle. Triangle lists
// tl[off[r]..off[r+1]) hold va's triangles, then vb's, each with the corner of its own merge vertex in cl.
// Operations in the reference's order, uncontracted, so positions and decisions are bit-identical.
// status: 0 candidate, 2 infeasible; merged: the position.
__global__ void __launch_bounds__(kBlock) k_merge_screen(int nRows, const int* __restrict__ rA,
                                                         const int* __restrict__ rB, const int* __restrict__ off,
                                                         const int* __restrict__ tl,
                                                         const int* __restrict__ cl, const int* __restrict__ F0,
                                                         const int* __restrict__ F1, const int* __restrict__ F2,
                                                         const double2* __restrict__ UV, double eps, int* status,
                                                         double2* merged)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nRows) return;
    const int va = rA[r], vb = rB[r], k0 = off[r], k1 = off[r + 1];
    const double2 pa = UV[va], pb = UV[vb];
    double2 m = make_double2((pa.x + pb.x) / 2.0, (pa.y + pb.y) / 2.0);
    const double2 mid = m;
    const auto pos = [&](int v, double2 at) { return (v == va || v == vb) ? at : UV[v]; };
    const auto corner = [&](int t, int k) { return k == 0 ? F0[t] : (k == 1 ? F1[t] : F2[t]); };
    const auto inverts = [&](double2 at) {
        for (int k = k0; k < k1; ++k) {
            const int t = tl[k];
            const double2 p0 = pos(F0[t], at), p1 = pos(F1[t], at), p2 = pos(F2[t], at);
            const double e0x = p1.x - p0.x, e0y = p1.y - p0.y, e1x = p2.x - p0.x, e1y = p2.y - p0.y;
            const double dbArea = e0x * e1y - e0y * e1x;
            if (dbArea < 0.0) return true;
        }
        return false;
    };
    int st = 0;
    if (inverts(m)) {
        for (int it = 0; it < 70; ++it) {
            double maxRes = -DBL_MAX;
            for (int k = k0; k < k1; ++k) {
                const int t = tl[k], i = cl[k];
                const double2 v1 = pos(corner(t, (i + 1) % 3), mid), v2 = pos(corner(t, (i + 2) % 3), mid);
                const double cx = v2.y - v1.y, cy = v1.x - v2.x;
                const double len = sqrt(cx * cx + cy * cy);
                const double nx = cx / len, ny = cy / len, c = (v1.x * v2.y - v1.y * v2.x) / len;
                const double res = (nx * m.x + ny * m.y) - c;
                if (res > -eps) {
                    const double s = res + eps;
                    m.x -= s * nx;
                    m.y -= s * ny;
                }
                if (res > maxRes) maxRes = res;
            }
            if (maxRes < 0.0) break;
        }
        st = inverts(m) ? 2 : 0;
    }
    status[r] = st;
    merged[r] = m;
}

void launch_merge_screen(hipStream_t s, int nRows, const int* rA, const int* rB, const int* off, const int* tl,
                         const int* cl, const int* F0, const int* F1, const int* F2, const double* UV, double eps,
                         int* status, double* merged)
{
    if (nRows <= 0) return;
    hipLaunchKernelGGL(k_merge_screen, dim3(grid_for(nRows)), dim3(kBlock), 0, s, nRows, rA, rB, off, tl, cl, F0,
                       F1, F2, reinterpret_cast<const double2*>(UV), eps, status,
                       reinterpret_cast<double2*>(merged));
}

// ----------------------------------------------------------------------------
// Candidate key of the split query: per vertex, the sample standard deviation
// of its triangles' gradient rows (SymDirichletEnergy::computeDivGradPerVert,
// SymDirichletEnergy.cpp:108-154): the mean accumulated in triangle order and
// divided by the count, then sqrt(sum |row - mean|^2 / (count - 1)); 0 for a
// vertex in one triangle. Operations in the reference's order, uncontracted.
__global__ void __launch_bounds__(kBlock) k_div_grad(int nV, const int* __restrict__ vcPtr,
                                                     const int* __restrict__ vcSrc,
                                                     const double2* __restrict__ elemG, double* out)
{
#pragma clang fp contract(off)
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    const int k0 = vcPtr[v], k1 = vcPtr[v + 1], n = k1 - k0;
    double mx = 0.0, my = 0.0;
    for (int k = k0; k < k1; ++k) {
        const double2 e = elemG[vcSrc[k]];
        mx += e.x;
        my += e.y;
    }
    mx /= n;
    my /= n;
    double sd = 0.0;
    for (int k = k0; k < k1; ++k) {
        const double2 e = elemG[vcSrc[k]];
        const double dx = e.x - mx, dy = e.y - my;
        sd += dx * dx + dy * dy;
    }
    out[v] = n == 1 ? 0.0 : sqrt(sd / (n - 1.0));
}

void launch_div_grad(hipStream_t s, int nV, const int* vcPtr, const int* vcSrc, const double* elemG, double* out)
{
    hipLaunchKernelGGL(k_div_grad, dim3(grid_for(nV)), dim3(kBlock), 0, s, nV, vcPtr, vcSrc,
                       reinterpret_cast<const double2*>(elemG), out);
}

// ----------------------------------------------------------------------------
template <typename OV>
__device__ __forceinline__ OV cvt4(double x, double y, double z, double w)
{
    OV o;
    o.x = (decltype(o.x))x;
    o.y = (decltype(o.y))y;
    o.z = (decltype(o.z))z;
    o.w = (decltype(o.w))w;
    return o;
}

__device__ __forceinline__ double4 load_block(const double4* H, int src) { return H[src]; }
__device__ __forceinline__ double4 load_block(const float4* H, int src)
{
    const float4 f = H[src];
    return make_double4(f.x, f.y, f.z, f.w);
}

__device__ __forceinline__ void jacobi_block(const double4& d, bool blockJacobi, double* M)
{
    if (blockJacobi) {
        const double det = d.x * d.w - d.y * d.z;
        M[0] = d.w / det;
        M[1] = -d.y / det;
        M[2] = -d.z / det;
        M[3] = d.x / det;
    } else {
        M[0] = 1.0 / d.x;
        M[1] = 0.0;
        M[2] = 0.0;
        M[3] = 1.0 / d.w;
    }
}

template <typename HV, typename OV>
__global__ void __launch_bounds__(kBlock) k_assemble(int nUB, const int* __restrict__ ubRow,
                                                     const int* __restrict__ ubCol, const int* __restrict__ ubPosU,
                                                     const int* __restrict__ ubPosL, const int* __restrict__ ubCPtr,
                                                     const int* __restrict__ ubCSrc,
                                                     const unsigned char* __restrict__ fixedMask,
                                                     const double* __restrict__ fixedDiag, const HV* __restrict__ elemH,
                                                     OV* vals, double* Minv, int blockJacobi)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nUB) return;
    const int u = ubRow[b], w = ubCol[b];
    double4 s = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int k = ubCPtr[b]; k < ubCPtr[b + 1]; ++k) {
        const unsigned int code = (unsigned int)ubCSrc[k];
        double4 e = load_block(elemH, (int)(code & 0x7fffffffu));
        if (code >> 31) {
            const double tmp = e.y;
            e.y = e.z;
            e.z = tmp;
        }
        s.x += e.x;
        s.y += e.y;
        s.z += e.z;
        s.w += e.w;
    }
    if (u == w) {
        if (fixedMask[u]) {
            const double fd = fixedDiag[u];
            s = make_double4(fd, 0.0, 0.0, fd);
        } else {
            s.z = s.y;  // the lower in-block entry mirrors the accumulated upper one
        }
        vals[ubPosU[b]] = cvt4<OV>(s.x, s.y, s.z, s.w);
        jacobi_block(s, blockJacobi != 0, Minv + 4 * (size_t)u);
    } else {
        vals[ubPosU[b]] = cvt4<OV>(s.x, s.y, s.z, s.w);
        vals[ubPosL[b]] = cvt4<OV>(s.x, s.z, s.y, s.w);
    }
}

template <typename HT>
void launch_assemble(hipStream_t s, int nUB, const int* ubRow, const int* ubCol, const int* ubPosU,
                     const int* ubPosL, const int* ubCPtr, const int* ubCSrc, const unsigned char* fixedMask,
                     const double* fixedDiag, const HT* elemH, void* vals, int vals32, double* Minv, int blockJacobi)
{
    using HV = typename std::conditional<std::is_same<HT, float>::value, float4, double4>::type;
    const HV* eh = reinterpret_cast<const HV*>(elemH);
    if (vals32)
        hipLaunchKernelGGL((k_assemble<HV, float4>), dim3(grid_for(nUB)), dim3(kBlock), 0, s, nUB, ubRow, ubCol,
                           ubPosU, ubPosL, ubCPtr, ubCSrc, fixedMask, fixedDiag, eh, reinterpret_cast<float4*>(vals),
                           Minv, blockJacobi);
    else
        hipLaunchKernelGGL((k_assemble<HV, double4>), dim3(grid_for(nUB)), dim3(kBlock), 0, s, nUB, ubRow, ubCol,
                           ubPosU, ubPosL, ubCPtr, ubCSrc, fixedMask, fixedDiag, eh, reinterpret_cast<double4*>(vals),
                           Minv, blockJacobi);
}
template void launch_assemble<double>(hipStream_t, int, const int*, const int*, const int*, const int*, const int*,
                                      const int*, const unsigned char*, const double*, const double*, void*, int,
                                      double*, int);
template void launch_assemble<float>(hipStream_t, int, const int*, const int*, const int*, const int*, const int*,
                                     const int*, const unsigned char*, const double*, const float*, void*, int,
                                     double*, int);

// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_step_bound(TriSet T, const double2* __restrict__ UV,
                                                       const double2* __restrict__ dir, unsigned long long* ordMin,
                                                       unsigned long long* ordMinAir, double cap)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double bound = cap;
    if (t < T.nF) {
        const int vi[3] = {T.f0[t], T.f1[t], T.f2[t]};
        double U[3][2], Vd[3][2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double2 u = UV[vi[k]], d = dir[vi[k]];
            U[k][0] = u.x;
            U[k][1] = u.y;
            Vd[k][0] = d.x;
            Vd[k][1] = d.y;
        }
        bound = sd_step_bound(U, Vd, cap);
    }
    // NaN bounds never lower the step (the reference's 'bound < stepSize' is false)
    bound = bound != bound ? cap : bound;
    if (ordMinAir) {
        // the air triangles' and the mesh triangles' own minima as well: how far the air mesh held the step back
        const double b = wave_min((t < T.nF && t >= T.nFmesh) ? bound : cap);
        if ((threadIdx.x % kWave) == 0 && b < cap) atomicMin(ordMinAir, dbl_to_ordered(b));
        const double m = wave_min(t < T.nFmesh ? bound : cap);
        if ((threadIdx.x % kWave) == 0 && m < cap) atomicMin(ordMinAir + 1, dbl_to_ordered(m));
    }
    bound = wave_min(bound);
    if ((threadIdx.x % kWave) == 0 && bound < cap) atomicMin(ordMin, dbl_to_ordered(bound));
}

void launch_step_bound(hipStream_t s, const TriSet& T, const double* UV, const double* dir,
                       unsigned long long* ordMin, double cap, unsigned long long* ordMinAir)
{
    hipLaunchKernelGGL(k_step_bound, dim3(grid_for(T.nF)), dim3(kBlock), 0, s, T,
                       reinterpret_cast<const double2*>(UV), reinterpret_cast<const double2*>(dir), ordMin,
                       ordMinAir, cap);
}

// ----------------------------------------------------------------------------
template <typename EP>
__global__ void __launch_bounds__(kBlock) k_linesearch(TriSet T, const double2* __restrict__ UV,
                                                       const double2* __restrict__ dir, double alphaK0,
                                                       const unsigned long long* __restrict__ boundOrd,
                                                       double* partials)
{
    constexpr int K = kLSCand;
    if (boundOrd) alphaK0 = ordered_to_dbl(*boundOrd) * 0.99;  // Optimizer.cpp:578: 0.99 x injectivity bound
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    CSum acc[3 * K];
#pragma unroll
    for (int k = 0; k < 3 * K; ++k) acc[k] = CSum{0.0, 0.0};
    if (t < T.nF) {
        const bool isMesh = t < T.nFmesh;
        const int vi[3] = {T.f0[t], T.f1[t], T.f2[t]};
        double2 u0[3], d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u0[k] = UV[vi[k]];
            d[k] = dir[vi[k]];
        }
        const RestT<EP> r = rest_as<EP>(load_rest(T, t, isMesh));
        double alpha = alphaK0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double p[3][2];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                p[q][0] = add_rn(u0[q].x, mul_rn(alpha, d[q].x));
                p[q][1] = add_rn(u0[q].y, mul_rn(alpha, d[q].y));
            }
            double E, dbA;
            if constexpr (std::is_same<EP, double>::value) {
                E = sd_energy<double>(p[0][0], p[0][1], p[1][0], p[1][1], p[2][0], p[2][1], r, dbA);
            } else {
                // FP32 energy on relative positions; the inversion test stays FP64
                const double ax = p[1][0] - p[0][0], ay = p[1][1] - p[0][1];
                const double bx = p[2][0] - p[0][0], by = p[2][1] - p[0][1];
                EP dbAf;
                E = (double)sd_energy<EP>(EP(0), EP(0), EP(ax), EP(ay), EP(bx), EP(by), r, dbAf);
                dbA = ax * by - ay * bx;
            }
            acc[k].s = isMesh ? E : 0.0;
            acc[K + k].s = isMesh ? 0.0 : E;
            acc[2 * K + k].s = (dbA < 0.0) ? 1.0 : 0.0;  // inversion count (exact small integers)
            alpha = alpha / 2.0;
        }
    }
    publish_csum<3 * K>(acc, partials);
}

__global__ void __launch_bounds__(kBlock) k_finish_linesearch(const double* __restrict__ partials, int nblk,
                                                              double sM, double sA, double* out)
{
    constexpr int K = kLSCand;
    __shared__ double res[3 * K];
    finish_csum<3 * K>(partials, nblk, res);
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            out[k] = sM * res[k] + sA * res[K + k];
            out[K + k] = res[k];
            out[2 * K + k] = res[K + k];
            out[3 * K + k] = res[2 * K + k];
        }
    }
}

void launch_linesearch(hipStream_t s, const TriSet& T, const double* UV, const double* dir, double alpha0, int k0,
                       double* partials, double* out, int f32, const unsigned long long* boundOrd)
{
    double a = alpha0;
    for (int k = 0; k < k0; ++k) a = a / 2.0;
    const int grid = grid_for(T.nF);
    if (f32)
        hipLaunchKernelGGL(k_linesearch<float>, dim3(grid), dim3(kBlock), 0, s, T,
                           reinterpret_cast<const double2*>(UV), reinterpret_cast<const double2*>(dir), a, boundOrd,
                           partials);
    else
        hipLaunchKernelGGL(k_linesearch<double>, dim3(grid), dim3(kBlock), 0, s, T,
                           reinterpret_cast<const double2*>(UV), reinterpret_cast<const double2*>(dir), a, boundOrd,
                           partials);
    hipLaunchKernelGGL(k_finish_linesearch, dim3(1), dim3(kBlock), 0, s, partials, grid, T.sMesh, T.sAir, out);
}

// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_axpy_commit(long n, double alpha, const double* __restrict__ d,
                                                        double* UV)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) UV[i] = add_rn(UV[i], mul_rn(alpha, d[i]));
}

void launch_axpy_commit(hipStream_t s, long n, double alpha, const double* d, double* UV)
{
    hipLaunchKernelGGL(k_axpy_commit, dim3(grid_for(n)), dim3(kBlock), 0, s, n, alpha, d, UV);
}

// The reference's line-search rule (Optimizer.cpp:575-629) applied to the K candidates of one k_linesearch
// launch from the bound: halve while E(alpha) > lastE, then while a triangle is inverted; alpha reaching 0 stops.
// lastE is the host's value, or (esum non-null: the scaffold changed since it was computed, :588-592)
// sMesh esum[0] + sAir esum[1] of the energy pass at the current UV. One thread.
__global__ void k_ls_decide(const double* __restrict__ cand, const unsigned long long* __restrict__ boundOrd,
                            const unsigned long long* __restrict__ airOrd, const double* __restrict__ esum,
                            double lastEHost, double sMesh, double sAir, LsDecision* out)
{
    constexpr int K = kLSCand;
    LsDecision r{};
    if (esum) {
        r.lastESD = esum[0];
        r.lastEScaf = esum[1] * sAir;
        r.lastE = sMesh * esum[0] + r.lastEScaf;
    } else {
        r.lastE = lastEHost;
    }
    r.bound = ordered_to_dbl(*boundOrd);
    r.airBound = airOrd ? ordered_to_dbl(airOrd[0]) : r.bound;
    r.meshBound = airOrd ? ordered_to_dbl(airOrd[1]) : r.bound;
    double alpha = r.bound * 0.99, alphaCommit = alpha;
    int k = 0, halv = 0, status = 0;
    double testE = cand[0];
    bool inv = cand[3 * K] > 0.0;
    for (int phase = 0; phase < 2 && status == 0; ++phase) {
        while (status == 0 && (phase == 0 ? (testE > r.lastE) : inv)) {
            const double na = alpha / 2.0;
            ++halv;
            if (na == 0.0) {
                status = 2;
                alpha = 0.0;
                break;
            }
            alpha = na;
            alphaCommit = alpha;
            if (++k >= K) {
                status = 1;
                break;
            }
            testE = cand[k];
            inv = cand[3 * K + k] > 0.0;
        }
    }
    r.alpha = alpha;
    r.alphaCommit = alphaCommit;
    const int kk = k < K ? k : K - 1;
    r.testE = testE;
    r.testESD = cand[K + kk];
    r.testES = cand[2 * K + kk] * sAir;
    r.halvings = halv;
    r.status = status;
    *out = r;
}

__global__ void __launch_bounds__(kBlock) k_axpy_commit_dev(long n, const LsDecision* __restrict__ dec,
                                                            const double* __restrict__ d, double* UV)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && dec->status != 1) UV[i] = add_rn(UV[i], mul_rn(dec->alphaCommit, d[i]));
}

void launch_ls_decide_commit(hipStream_t s, const double* cand, const unsigned long long* boundOrd,
                             const unsigned long long* airOrd, const double* esum, double lastEHost, double sMesh,
                             double sAir, LsDecision* dec, long n, const double* d, double* UV)
{
    hipLaunchKernelGGL(k_ls_decide, dim3(1), dim3(1), 0, s, cand, boundOrd, airOrd, esum, lastEHost, sMesh, sAir,
                       dec);
    hipLaunchKernelGGL(k_axpy_commit_dev, dim3(grid_for(n)), dim3(kBlock), 0, s, n, dec, d, UV);
}

// ----------------------------------------------------------------------------
// PCG (preconditioned conjugate gradients, Hestenes & Stiefel 1952) with a
// 2x2-block Jacobi preconditioner, two kernels per iteration i:
//   spmv(i)  : rz_i, rr_i = sums of update(i-1)'s per-block partials (every
//              block reduces them itself, in the same order, so all blocks
//              agree); convergence test rr_i <= tol^2 bb; beta = rz_i/rz_{i-1};
//              p_new = z + beta p_old (own rows); q = A p_new, the neighbours'
//              p_new formed on the fly from z and p_old, so no grid-wide wait
//              sits between the update of p and A p; partials of p.q.
//   update(i): pq = sum of spmv(i)'s partials; alpha = rz_i/pq; x += alpha p;
//              r -= alpha q; z = M r; partials of r.z and r.r.
// With an external preconditioner (Minv == nullptr) init/update leave z
// alone and the caller applies M between update(i) and spmv(i+1), handing
// its own r.z partial records to spmv.
// Scalars shared across kernels (rz_i, bb, done) are written by block 0 only,
// into slots that no other block of the same kernel reads. No atomics and no
// device-scope fences: kernel boundaries order everything.
// Rows are split into one contiguous chunk per workgroup, so the neighbour
// gathers of a chunk stay in the L2 of the XCD that runs it.
__device__ __forceinline__ double2 apply_M(const double* Minv, int v, double2 r)
{
    const double4 m = reinterpret_cast<const double4*>(Minv)[v];
    return make_double2(m.x * r.x + m.y * r.y, m.z * r.x + m.w * r.y);
}

// Sum of NV-value partial records (record b at part[b*NV + k]) over nblk
// records, identical in every block.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nblk, double (&out)[NV])
{
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += part[(size_t)b * NV + k];
    block_sum<NV>(acc);
    __shared__ double sh[NV];
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k] = acc[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = sh[k];
}

// Sum of one value per record (record b at part[b*stride]), identical in every block.
__device__ __forceinline__ double reduce_strided(const double* __restrict__ part, int nrec, int stride)
{
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < nrec; b += blockDim.x) acc[0] += part[(size_t)b * stride];
    block_sum<1>(acc);
    __shared__ double sh;
    if (threadIdx.x == 0) sh = acc[0];
    __syncthreads();
    return sh;
}

__device__ __forceinline__ void row_chunk(int nV, int& r0, int& r1)
{
    const int chunk = (nV + (int)gridDim.x - 1) / (int)gridDim.x;
    r0 = (int)blockIdx.x * chunk;
    r1 = min(nV, r0 + chunk);
}

template <bool APPLY_M>
__global__ void __launch_bounds__(kBlock) k_pcg_init(int nV, const double2* __restrict__ b,
                                                     const double* __restrict__ Minv, double2* x, double2* r,
                                                     double2* z, double* partU, PcgCtrl* ctrl)
{
    int r0, r1;
    row_chunk(nV, r0, r1);
    double acc[2] = {0.0, 0.0};
    for (int v = r0 + threadIdx.x; v < r1; v += blockDim.x) {
        const double2 bv = b[v];
        x[v] = make_double2(0.0, 0.0);
        r[v] = bv;
        if (APPLY_M) {
            const double2 zv = apply_M(Minv, v, bv);
            z[v] = zv;
            acc[0] += bv.x * zv.x + bv.y * zv.y;
        }
        acc[1] += bv.x * bv.x + bv.y * bv.y;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partU[2 * blockIdx.x] = acc[0];
        partU[2 * blockIdx.x + 1] = acc[1];
        if (blockIdx.x == 0) {
            ctrl->done = 0;
            ctrl->iter = 0;
            ctrl->itNext = 0;
            ctrl->itCur = 0;
        }
    }
}

int pcg_grid_update(int nV) { return std::max(1, std::min((nV + kBlock - 1) / kBlock, kPcgMaxGrid)); }
// Workgroups of the PCG product: one pass of kPcgRows rows per kPcgLanes-lane
// group, at most kPcgMaxGrid, a multiple of 8 (XCD-contiguous chunk order).
int pcg_grid_spmv(int nV)
{
    constexpr int rows = (kBlock / kPcgLanes) * kPcgRows;
    const int g = std::max(1, std::min((nV + rows - 1) / rows, kPcgMaxGrid));
    return (g + 7) & ~7;
}

void launch_pcg_init(hipStream_t s, int nV, const double* b, const double* Minv, double* x, double* r, double* z,
                     double* partU, PcgCtrl* ctrl)
{
    if (Minv)
        hipLaunchKernelGGL(k_pcg_init<true>, dim3(pcg_grid_update(nV)), dim3(kBlock), 0, s, nV,
                           reinterpret_cast<const double2*>(b), Minv, reinterpret_cast<double2*>(x),
                           reinterpret_cast<double2*>(r), reinterpret_cast<double2*>(z), partU, ctrl);
    else
        hipLaunchKernelGGL(k_pcg_init<false>, dim3(pcg_grid_update(nV)), dim3(kBlock), 0, s, nV,
                           reinterpret_cast<const double2*>(b), Minv, reinterpret_cast<double2*>(x),
                           reinterpret_cast<double2*>(r), reinterpret_cast<double2*>(z), partU, ctrl);
}

// q_v = sum_k A_vk (z_w + beta p_w) over the row's blocks, kSpmvLanes lanes per row.
template <typename V4>
__device__ __forceinline__ double2 row_apply(const int* __restrict__ rowPtr, const int* __restrict__ colIdx,
                                             const V4* __restrict__ vals, const double2* __restrict__ z,
                                             const double2* __restrict__ pOld, double beta, int v, int lane)
{
    double yx = 0.0, yy = 0.0;
    const int k1 = rowPtr[v + 1];
    for (int k = rowPtr[v] + lane; k < k1; k += kSpmvLanes) {
        const int w = colIdx[k];
        const V4 av = vals[k];
        const double4 a = make_double4((double)av.x, (double)av.y, (double)av.z, (double)av.w);
        const double2 zw = z[w];
        double px = zw.x, py = zw.y;
        if (pOld) {
            const double2 pw = pOld[w];
            px += beta * pw.x;
            py += beta * pw.y;
        }
        yx += a.x * px + a.y * py;
        yy += a.z * px + a.w * py;
    }
#pragma unroll
    for (int off = kSpmvLanes / 2; off > 0; off >>= 1) {
        yx += __shfl_xor(yx, off, kSpmvLanes);
        yy += __shfl_xor(yy, off, kSpmvLanes);
    }
    return make_double2(yx, yy);
}

__device__ __forceinline__ double4 to_double4(const double4& v) { return v; }
__device__ __forceinline__ double4 to_double4(const float4& v)
{
    return make_double4((double)v.x, (double)v.y, (double)v.z, (double)v.w);
}

// PCG direction and product: p_new = z + beta p_old, q = A p_new, partial
// records of p_new.q. beta needs the previous update's reductions, so the
// row loads (blocks, gathered z and p_old) are issued first and the records
// reduced while they are in flight; p_w = z_w + beta p_old_w is formed per
// gathered entry. kPcgRows rows per kPcgLanes-lane group; workgroup b takes
// row chunk (b % 8) * (G / 8) + b / 8 so each XCD owns a contiguous eighth of
// the rows (G = pcg_grid_spmv, a multiple of 8).
template <typename V4>
__global__ void __launch_bounds__(kBlock) k_pcg_spmv(int nV, const int* __restrict__ rowPtr,
                                                     const int* __restrict__ colIdx,
                                                     const V4* __restrict__ vals,
                                                     const double2* __restrict__ z,
                                                     const double2* __restrict__ pOld, double2* pNew, double2* q,
                                                     const double* __restrict__ partRz, int nRz, int sRz,
                                                     const double* __restrict__ partRr, int nRr, int sRr,
                                                     double* partS, PcgCtrl* ctrl)
{
    if (ctrl->done) return;
    const int it = ctrl->itNext;
    constexpr int LANES = kPcgLanes, UNROLL = kPcgRows, RPP = kBlock / LANES;
    const int G = gridDim.x;
    const int c = (nV + G - 1) / G;
    const int r0 = ((int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3)) * c, r1 = min(nV, r0 + c);
    const int sub = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    const int passes = max(1, (c + RPP * UNROLL - 1) / (RPP * UNROLL));
    const bool first = it == 0;
    double beta = 0.0;
    double acc[1] = {0.0};
    for (int pass = 0; pass < passes; ++pass) {
        int u[UNROLL], kb[UNROLL], ke[UNROLL];
        double4 a[UNROLL];
        double2 gz[UNROLL], gp[UNROLL], oz[UNROLL], op[UNROLL];
#pragma unroll
        for (int r = 0; r < UNROLL; ++r) {
            u[r] = r0 + sub + (pass * UNROLL + r) * RPP;
            const bool valid = u[r] < r1;
            kb[r] = valid ? rowPtr[u[r]] : 0;
            ke[r] = valid ? rowPtr[u[r] + 1] : 0;
        }
#pragma unroll
        for (int r = 0; r < UNROLL; ++r) {
            const int k = kb[r] + lane;
            const bool has = k < ke[r];
            const int kk = has ? k : 0;
            const double4 av = to_double4(vals[kk]);
            a[r] = has ? av : make_double4(0.0, 0.0, 0.0, 0.0);
            const int w = has ? colIdx[kk] : 0;
            gz[r] = z[w];
            gp[r] = first ? make_double2(0.0, 0.0) : pOld[w];
            oz[r] = op[r] = make_double2(0.0, 0.0);
            if (lane == 0 && u[r] < r1) {
                oz[r] = z[u[r]];
                if (!first) op[r] = pOld[u[r]];
            }
        }
        if (pass == 0) {
            const double rz = reduce_strided(partRz, nRz, sRz);
            const double rr = reduce_strided(partRr, nRr, sRr);
            const double bb = first ? rr : ctrl->bb;
            const bool conv = (bb == 0.0) || (rr <= ctrl->tol2 * bb);
            const bool capped = it >= ctrl->maxit;
            const bool indef = !conv && !(rz > 0.0);  // preconditioner not positive definite on r
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                if (first) ctrl->bb = bb;
                ctrl->rzSlot[it & 1] = rz;
                ctrl->rr = rr;
                ctrl->iter = it;
                ctrl->itCur = it;
                if (conv) ctrl->done = 1;
                else if (indef) ctrl->done = 3;
                else if (capped) ctrl->done = 2;
            }
            if (conv || capped || indef) return;
            beta = first ? 0.0 : rz / ctrl->rzSlot[(it + 1) & 1];
        }
#pragma unroll
        for (int r = 0; r < UNROLL; ++r) {
            double px = gz[r].x + beta * gp[r].x, py = gz[r].y + beta * gp[r].y;
            double yx = a[r].x * px + a[r].y * py;
            double yy = a[r].z * px + a[r].w * py;
            for (int k = kb[r] + lane + LANES; k < ke[r]; k += LANES) {
                const int w = colIdx[k];
                const double4 ak = to_double4(vals[k]);
                const double2 zw = z[w];
                px = zw.x;
                py = zw.y;
                if (!first) {
                    const double2 pw = pOld[w];
                    px += beta * pw.x;
                    py += beta * pw.y;
                }
                yx += ak.x * px + ak.y * py;
                yy += ak.z * px + ak.w * py;
            }
#pragma unroll
            for (int off = LANES / 2; off > 0; off >>= 1) {
                yx += __shfl_xor(yx, off, LANES);
                yy += __shfl_xor(yy, off, LANES);
            }
            if (lane == 0 && u[r] < r1) {
                const double2 pn = make_double2(oz[r].x + beta * op[r].x, oz[r].y + beta * op[r].y);
                pNew[u[r]] = pn;
                q[u[r]] = make_double2(yx, yy);
                acc[0] += pn.x * yx + pn.y * yy;
            }
        }
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) partS[blockIdx.x] = acc[0];
}

void launch_pcg_spmv(hipStream_t s, int nV, const int* rowPtr, const int* colIdx, const void* vals,
                     int vals32, const double* z, const double* pOld, double* pNew, double* q, const double* partRz,
                     int nRz, int sRz, const double* partRr, int nRr, int sRr, double* partS, PcgCtrl* ctrl)
{
    if (vals32)
        hipLaunchKernelGGL(k_pcg_spmv<float4>, dim3(pcg_grid_spmv(nV)), dim3(kBlock), 0, s, nV, rowPtr, colIdx,
                           reinterpret_cast<const float4*>(vals), reinterpret_cast<const double2*>(z),
                           reinterpret_cast<const double2*>(pOld), reinterpret_cast<double2*>(pNew),
                           reinterpret_cast<double2*>(q), partRz, nRz, sRz, partRr, nRr, sRr, partS, ctrl);
    else
        hipLaunchKernelGGL(k_pcg_spmv<double4>, dim3(pcg_grid_spmv(nV)), dim3(kBlock), 0, s, nV, rowPtr, colIdx,
                           reinterpret_cast<const double4*>(vals), reinterpret_cast<const double2*>(z),
                           reinterpret_cast<const double2*>(pOld), reinterpret_cast<double2*>(pNew),
                           reinterpret_cast<double2*>(q), partRz, nRz, sRz, partRr, nRr, sRr, partS, ctrl);
}

template <bool APPLY_M>
__global__ void __launch_bounds__(kBlock) k_pcg_update(int nV, const double2* __restrict__ p,
                                                       const double2* __restrict__ q,
                                                       const double* __restrict__ Minv, double2* x, double2* r,
                                                       double2* z, const double* __restrict__ partS, int nPartS,
                                                       double* partUout, PcgCtrl* ctrl)
{
    if (ctrl->done) return;
    double pq[1];
    reduce_partials<1>(partS, nPartS, pq);
    if (!(pq[0] > 0.0)) {  // breakdown: A not positive definite along p
        if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->done = 3;
        return;
    }
    const int it = ctrl->itCur;
    const double alpha = ctrl->rzSlot[it & 1] / pq[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->itNext = it + 1;
    int r0, r1;
    row_chunk(nV, r0, r1);
    double acc[2] = {0.0, 0.0};
    for (int v = r0 + threadIdx.x; v < r1; v += blockDim.x) {
        const double2 pv = p[v], qv = q[v];
        double2 xv = x[v], rv = r[v];
        xv.x += alpha * pv.x;
        xv.y += alpha * pv.y;
        rv.x -= alpha * qv.x;
        rv.y -= alpha * qv.y;
        x[v] = xv;
        r[v] = rv;
        if (APPLY_M) {
            const double2 zv = apply_M(Minv, v, rv);
            z[v] = zv;
            acc[0] += rv.x * zv.x + rv.y * zv.y;
        }
        acc[1] += rv.x * rv.x + rv.y * rv.y;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partUout[2 * blockIdx.x] = acc[0];
        partUout[2 * blockIdx.x + 1] = acc[1];
    }
}

void launch_pcg_update(hipStream_t s, int nV, const double* pNew, const double* q, const double* Minv,
                       double* x, double* r, double* z, const double* partS, double* partUout, PcgCtrl* ctrl)
{
    if (Minv)
        hipLaunchKernelGGL(k_pcg_update<true>, dim3(pcg_grid_update(nV)), dim3(kBlock), 0, s, nV,
                           reinterpret_cast<const double2*>(pNew), reinterpret_cast<const double2*>(q), Minv,
                           reinterpret_cast<double2*>(x), reinterpret_cast<double2*>(r),
                           reinterpret_cast<double2*>(z), partS, pcg_grid_spmv(nV), partUout, ctrl);
    else
        hipLaunchKernelGGL(k_pcg_update<false>, dim3(pcg_grid_update(nV)), dim3(kBlock), 0, s, nV,
                           reinterpret_cast<const double2*>(pNew), reinterpret_cast<const double2*>(q), Minv,
                           reinterpret_cast<double2*>(x), reinterpret_cast<double2*>(r),
                           reinterpret_cast<double2*>(z), partS, pcg_grid_spmv(nV), partUout, ctrl);
}

// ----------------------------------------------------------------------------
template <typename V4>
__global__ void __launch_bounds__(kBlock) k_spmv(int nV, const int* __restrict__ rowPtr,
                                                 const int* __restrict__ colIdx, const V4* __restrict__ vals,
                                                 const double2* __restrict__ x, double2* y)
{
    int r0, r1;
    row_chunk(nV, r0, r1);
    const int sub = threadIdx.x / kSpmvLanes, lane = threadIdx.x % kSpmvLanes;
    constexpr int rowsPerPass = kBlock / kSpmvLanes;
    for (int v = r0 + sub; v < r1; v += rowsPerPass) {
        const double2 s = row_apply(rowPtr, colIdx, vals, x, nullptr, 0.0, v, lane);
        if (lane == 0) y[v] = s;
    }
}

void launch_spmv(hipStream_t s, int nV, const int* rowPtr, const int* colIdx, const void* vals, int vals32,
                 const double* x, double* y)
{
    if (vals32)
        hipLaunchKernelGGL(k_spmv<float4>, dim3(pcg_grid_spmv(nV)), dim3(kBlock), 0, s, nV, rowPtr, colIdx,
                           reinterpret_cast<const float4*>(vals), reinterpret_cast<const double2*>(x),
                           reinterpret_cast<double2*>(y));
    else
        hipLaunchKernelGGL(k_spmv<double4>, dim3(pcg_grid_spmv(nV)), dim3(kBlock), 0, s, nV, rowPtr, colIdx,
                           reinterpret_cast<const double4*>(vals), reinterpret_cast<const double2*>(x),
                           reinterpret_cast<double2*>(y));
}

__global__ void __launch_bounds__(kBlock) k_expand_upper(int nnz, const double* __restrict__ a,
                                                         const int* __restrict__ dst, const int* __restrict__ mir,
                                                         double* vals)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const double v = a[k];
    vals[dst[k]] = v;
    if (mir[k] >= 0) vals[mir[k]] = v;
}

void launch_expand_upper(hipStream_t s, int nnz, const double* a, const int* dst, const int* mir, double* vals)
{
    hipLaunchKernelGGL(k_expand_upper, dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, a, dst, mir, vals);
}

__global__ void __launch_bounds__(kBlock) k_jacobi_from_bsr(int nV, const int* __restrict__ diagPos,
                                                            const double4* __restrict__ vals, double* Minv,
                                                            int blockJacobi)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    jacobi_block(vals[diagPos[v]], blockJacobi != 0, Minv + 4 * (size_t)v);
}

void launch_jacobi_from_bsr(hipStream_t s, int nV, const int* diagPos, const double* vals, double* Minv,
                            int blockJacobi)
{
    hipLaunchKernelGGL(k_jacobi_from_bsr, dim3(grid_for(nV)), dim3(kBlock), 0, s, nV, diagPos,
                       reinterpret_cast<const double4*>(vals), Minv, blockJacobi);
}

}  // namespace oc

namespace oc {

__global__ void __launch_bounds__(kBlock) k_count_inverted(TriSet T, const double2* __restrict__ UV,
                                                           unsigned int* count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int inv = 0;
    if (t < T.nF) {
        const double2 a = UV[T.f0[t]], b = UV[T.f1[t]], c = UV[T.f2[t]];
        const double e0x = b.x - a.x, e0y = b.y - a.y, e1x = c.x - a.x, e1y = c.y - a.y;
        inv = (__dsub_rn(mul_rn(e0x, e1y), mul_rn(e0y, e1x)) < 0.0) ? 1 : 0;
    }
    const unsigned long long m = __ballot(inv);
    if ((threadIdx.x % kWave) == 0 && m) atomicAdd(count, (unsigned int)__popcll(m));
}

void launch_count_inverted(hipStream_t s, const TriSet& T, const double* UV, unsigned int* count)
{
    hipLaunchKernelGGL(k_count_inverted, dim3(grid_for(T.nF)), dim3(kBlock), 0, s, T,
                       reinterpret_cast<const double2*>(UV), count);
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

void launch_set_u64(hipStream_t s, unsigned long long* p, unsigned long long v)
{
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, p, v);
}

__global__ void __launch_bounds__(kBlock) k_scale_copy(long n, double sign, const double* __restrict__ g,
                                                       double* b)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = sign * g[i];
}

void launch_scale_copy(hipStream_t s, long n, double sign, const double* g, double* b)
{
    hipLaunchKernelGGL(k_scale_copy, dim3(grid_for(n)), dim3(kBlock), 0, s, n, sign, g, b);
}

__global__ void __launch_bounds__(kBlock) k_uv_gather(int n, const int* __restrict__ ids,
                                                      const double2* __restrict__ UV, double* out3, double2* out2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double2 p = UV[ids ? ids[k] : k];
    if (out3) {
        out3[3 * (size_t)k] = p.x;
        out3[3 * (size_t)k + 1] = p.y;
        out3[3 * (size_t)k + 2] = 0.0;
    }
    if (out2) out2[k] = p;
}

// The Newton matrix mapped onto an earlier pattern (the one a multigrid structure was built for), per row u of
// that pattern: each block takes the new matrix's block of the same (u, w) (map >= 0) or zero; the blocks of
// row u that the earlier pattern lacks (drop list) are left out and their Frobenius norms added to both diagonal
// entries of u's diagonal block. Dropping a symmetric pair of off-diagonal blocks B, B^T and adding |B| I to both
// diagonal blocks adds [[|B| I, -B], [-B^T, |B| I]], which is positive semidefinite: the result stays symmetric
// positive definite when the new matrix is.
__global__ void __launch_bounds__(kBlock) k_remap_blocks(int n, const int* __restrict__ oldRowPtr,
                                                         const int* __restrict__ map, const int* __restrict__ dropPtr,
                                                         const int* __restrict__ drop, const int* __restrict__ oldDiag,
                                                         const double4* __restrict__ newVals, double4* oldVals)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    for (int k = oldRowPtr[u]; k < oldRowPtr[u + 1]; ++k) {
        const int m = map[k];
        oldVals[k] = m >= 0 ? newVals[m] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    double comp = 0.0;
    for (int q = dropPtr[u]; q < dropPtr[u + 1]; ++q) {
        const double4 b = newVals[drop[q]];
        comp += sqrt(b.x * b.x + b.y * b.y + b.z * b.z + b.w * b.w);
    }
    if (comp > 0.0) {
        double4 d = oldVals[oldDiag[u]];
        d.x += comp;
        d.w += comp;
        oldVals[oldDiag[u]] = d;
    }
}

void launch_remap_blocks(hipStream_t s, int n, const int* oldRowPtr, const int* map, const int* dropPtr,
                         const int* drop, const int* oldDiag, const double* newVals, double* oldVals)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_remap_blocks, dim3(grid_for(n)), dim3(kBlock), 0, s, n, oldRowPtr, map, dropPtr, drop,
                       oldDiag, reinterpret_cast<const double4*>(newVals), reinterpret_cast<double4*>(oldVals));
}

// ---- iterative refinement ----------------------------------------------------------------------------------
// Row u of the FP32 rounding of a 2x2-block matrix, with factor * (row abs sum) added to the diagonal entries of
// u's diagonal block before rounding (a Gershgorin guard: with factor 4 * 2^-24 it exceeds what the rounding can
// take from any eigenvalue, so a positive semidefinite FP64 matrix rounds to a positive definite FP32 one).
__global__ void __launch_bounds__(kBlock) k_round_blocks(int n, const int* __restrict__ rowPtr,
                                                         const int* __restrict__ diagPos,
                                                         const double4* __restrict__ in, float4* out, double factor)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    double rs0 = 0.0, rs1 = 0.0;
    for (int k = rowPtr[u]; k < rowPtr[u + 1]; ++k) {
        const double4 v = in[k];
        rs0 += fabs(v.x) + fabs(v.y);
        rs1 += fabs(v.z) + fabs(v.w);
    }
    const int d = diagPos[u];
    for (int k = rowPtr[u]; k < rowPtr[u + 1]; ++k) {
        double4 v = in[k];
        if (k == d) {
            v.x += factor * rs0;
            v.w += factor * rs1;
        }
        out[k] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
    }
}

void launch_round_blocks(hipStream_t s, int nV, const int* rowPtr, const int* diagPos, const double* in, float* out,
                         double factor)
{
    if (nV <= 0) return;
    hipLaunchKernelGGL(k_round_blocks, dim3(grid_for(nV)), dim3(kBlock), 0, s, nV, rowPtr, diagPos,
                       reinterpret_cast<const double4*>(in), reinterpret_cast<float4*>(out), factor);
}

// r = b - y (y null: r = b) over n doubles and per-workgroup sums of r.r in part (grid_for(n) records).
__global__ void __launch_bounds__(kBlock) k_residual_norm(long n, const double* __restrict__ b,
                                                          const double* __restrict__ y, double* r, double* part)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (k < n) {
        const double v = y ? b[k] - y[k] : b[k];
        if (r) r[k] = v;
        acc[0] = v * v;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

int launch_residual_norm(hipStream_t s, long n, const double* b, const double* y, double* r, double* part)
{
    const int g = grid_for(n);
    hipLaunchKernelGGL(k_residual_norm, dim3(g), dim3(kBlock), 0, s, n, b, y, r, part);
    return g;
}

void launch_uv_gather(hipStream_t s, int n, const int* ids, const double* UV, double* out3, double* out2)
{
    hipLaunchKernelGGL(k_uv_gather, dim3(grid_for(n)), dim3(kBlock), 0, s, n, ids,
                       reinterpret_cast<const double2*>(UV), out3, reinterpret_cast<double2*>(out2));
}

}  // namespace oc

// ============================================================================ batched charts
namespace oc {

// out[s] = compensated sum of vals over [segStart[s], segStart[s+1]) (one block per segment)
__global__ void __launch_bounds__(kBlock) k_chart_seg_sum(const int* __restrict__ segStart,
                                                          const double* __restrict__ vals, double* out)
{
    const int sg = blockIdx.x;
    CSum acc[1] = {{0.0, 0.0}};
    for (int i = segStart[sg] + threadIdx.x; i < segStart[sg + 1]; i += blockDim.x)
        acc[0] = csum_add(acc[0], CSum{vals[i], 0.0});
    block_csum<1>(acc);
    if (threadIdx.x == 0) out[sg] = acc[0].s + acc[0].c;
}

void launch_chart_seg_sum(hipStream_t s, int nSeg, const int* segStart, const double* vals, double* out)
{
    hipLaunchKernelGGL(k_chart_seg_sum, dim3(nSeg), dim3(kBlock), 0, s, segStart, vals, out);
}

// out[c] = |g|^2 over chart c's vertices (one block per chart)
__global__ void __launch_bounds__(kBlock) k_chart_gsq(const int* __restrict__ vStart, const double2* __restrict__ g,
                                                      double* out)
{
    const int c = blockIdx.x;
    double acc[1] = {0.0};
    for (int v = vStart[c] + threadIdx.x; v < vStart[c + 1]; v += blockDim.x) {
        const double2 x = g[v];
        acc[0] += x.x * x.x + x.y * x.y;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) out[c] = acc[0];
}

void launch_chart_gsq(hipStream_t s, int nCharts, const int* vStart, const double* g, double* out)
{
    hipLaunchKernelGGL(k_chart_gsq, dim3(nCharts), dim3(kBlock), 0, s, vStart, reinterpret_cast<const double2*>(g),
                       out);
}

// Line-search energies of the mesh triangles of chart-aligned blocks at K
// halvings of each chart's own step; block records of (E_k, inverted_k).
__global__ void __launch_bounds__(kBlock) k_chart_linesearch(TriSet T, const double2* __restrict__ UV,
                                                             const double2* __restrict__ dir,
                                                             const int2* __restrict__ blkRange,
                                                             const int* __restrict__ blkChart,
                                                             const double* __restrict__ alphaChart, double* partials)
{
    constexpr int K = kLSCand;
    const int2 rg = blkRange[blockIdx.x];
    const int t = rg.x + threadIdx.x;
    CSum acc[2 * K];
#pragma unroll
    for (int k = 0; k < 2 * K; ++k) acc[k] = CSum{0.0, 0.0};
    if (t < rg.y) {
        const int vi[3] = {T.f0[t], T.f1[t], T.f2[t]};
        double2 u0[3], d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u0[k] = UV[vi[k]];
            d[k] = dir[vi[k]];
        }
        const RestT<double> r = load_rest(T, t, true);
        double alpha = alphaChart[blkChart[blockIdx.x]];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double p[3][2];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                p[q][0] = add_rn(u0[q].x, mul_rn(alpha, d[q].x));
                p[q][1] = add_rn(u0[q].y, mul_rn(alpha, d[q].y));
            }
            double dbA;
            acc[k].s = sd_energy<double>(p[0][0], p[0][1], p[1][0], p[1][1], p[2][0], p[2][1], r, dbA);
            acc[K + k].s = dbA < 0.0 ? 1.0 : 0.0;
            alpha = alpha / 2.0;
        }
    }
    publish_csum<2 * K>(acc, partials);
}

__global__ void __launch_bounds__(kBlock) k_chart_linesearch_finish(const double* __restrict__ partials,
                                                                    const int* __restrict__ chartBlk, double* out)
{
    constexpr int K = kLSCand;
    const int c = blockIdx.x;
    const int b0 = chartBlk[c], nb = chartBlk[c + 1] - chartBlk[c];
    __shared__ double res[2 * K];
    finish_csum<2 * K>(partials + (size_t)b0 * 4 * K, nb, res);
    __syncthreads();
    if (threadIdx.x < 2 * K) out[(size_t)c * 2 * K + threadIdx.x] = res[threadIdx.x];
}

void launch_chart_linesearch(hipStream_t s, const TriSet& T, const double* UV, const double* dir, int nBlk,
                             const int2* blkRange, const int* blkChart, const double* alphaChart, int nCharts,
                             const int* chartBlk, double* partials, double* out)
{
    hipLaunchKernelGGL(k_chart_linesearch, dim3(nBlk), dim3(kBlock), 0, s, T, reinterpret_cast<const double2*>(UV),
                       reinterpret_cast<const double2*>(dir), blkRange, blkChart, alphaChart, partials);
    hipLaunchKernelGGL(k_chart_linesearch_finish, dim3(nCharts), dim3(kBlock), 0, s, partials, chartBlk, out);
}

// per-chart minimum of the injectivity bound (one atomicMin per block)
__global__ void __launch_bounds__(kBlock) k_chart_step_bound(TriSet T, const double2* __restrict__ UV,
                                                             const double2* __restrict__ dir,
                                                             const int2* __restrict__ blkRange,
                                                             const int* __restrict__ blkChart,
                                                             unsigned long long* ordMin, double cap)
{
    const int2 rg = blkRange[blockIdx.x];
    const int t = rg.x + threadIdx.x;
    double bound = cap;
    if (t < rg.y) {
        const int vi[3] = {T.f0[t], T.f1[t], T.f2[t]};
        double U[3][2], Vd[3][2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double2 u = UV[vi[k]], d = dir[vi[k]];
            U[k][0] = u.x;
            U[k][1] = u.y;
            Vd[k][0] = d.x;
            Vd[k][1] = d.y;
        }
        bound = sd_step_bound(U, Vd, cap);
    }
    bound = wave_min(bound != bound ? cap : bound);
    __shared__ double wmin[kBlock / kWave];
    if (threadIdx.x % kWave == 0) wmin[threadIdx.x / kWave] = bound;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        for (int w = 1; w < kBlock / kWave; ++w) m = fmin(m, wmin[w]);
        if (m < cap) atomicMin(ordMin + blkChart[blockIdx.x], dbl_to_ordered(m));
    }
}

void launch_chart_step_bound(hipStream_t s, const TriSet& T, const double* UV, const double* dir, int nBlk,
                             const int2* blkRange, const int* blkChart, unsigned long long* ordMin, double cap)
{
    hipLaunchKernelGGL(k_chart_step_bound, dim3(nBlk), dim3(kBlock), 0, s, T, reinterpret_cast<const double2*>(UV),
                       reinterpret_cast<const double2*>(dir), blkRange, blkChart, ordMin, cap);
}

__global__ void __launch_bounds__(kBlock) k_chart_rhs(int nV, const int* __restrict__ vChart,
                                                      const unsigned char* __restrict__ active,
                                                      const double2* __restrict__ g, double2* b)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    const double2 x = g[v];
    b[v] = active[vChart[v]] ? make_double2(-x.x, -x.y) : make_double2(0.0, 0.0);
}

void launch_chart_rhs(hipStream_t s, int nV, const int* vChart, const unsigned char* active, const double* g,
                      double* b)
{
    hipLaunchKernelGGL(k_chart_rhs, dim3(grid_for(nV)), dim3(kBlock), 0, s, nV, vChart, active,
                       reinterpret_cast<const double2*>(g), reinterpret_cast<double2*>(b));
}

__global__ void __launch_bounds__(kBlock) k_chart_commit(int nV, const int* __restrict__ vChart,
                                                         const double* __restrict__ alphaChart,
                                                         const double2* __restrict__ d, double2* UV)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    const double a = alphaChart[vChart[v]];
    const double2 dv = d[v];
    double2 u = UV[v];
    u.x = add_rn(u.x, mul_rn(a, dv.x));
    u.y = add_rn(u.y, mul_rn(a, dv.y));
    UV[v] = u;
}

void launch_chart_commit(hipStream_t s, int nV, const int* vChart, const double* alphaChart, const double* d,
                         double* UV)
{
    hipLaunchKernelGGL(k_chart_commit, dim3(grid_for(nV)), dim3(kBlock), 0, s, nV, vChart, alphaChart,
                       reinterpret_cast<const double2*>(d), reinterpret_cast<double2*>(UV));
}

}  // namespace oc
