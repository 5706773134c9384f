// Per-triangle symmetric Dirichlet terms, evaluated in registers.
//
// Follows SymDirichletEnergy.cpp: energy (:24-46), gradient (:258-304) and the
// 6x6 element Hessian (:438-523) with its PSD projection IglUtils::makePD
// (IglUtils.hpp:71-90). All quantities use the per-triangle rest features
// w, A^2 (triAreaSq) and e0SqLen/(2A^2), e1SqLen/(2A^2), e0dote1/(2A^2)
// (TriMesh.cpp:386-393), the set the reference's Hessian uses.
//
// PSD projection. The element energy depends only on U2-U1 and U3-U1, so both
// rigid translations lie in the kernel of H. With B the orthonormal 6x4 basis
// of the translation-free subspace (per coordinate: (1,-1,0)/sqrt2 and
// (1,1,-2)/sqrt6), H = B S B^T with S = B^T H B (4x4). The eigenpairs of H
// with nonzero eigenvalue are (lambda, B q) for the eigenpairs (lambda, q)
// of S, so clamping H's negative eigenvalues (makePD) equals
//     H_psd = H - sum_{lambda<0} lambda (B q)(B q)^T,
// computed here with a cyclic Jacobi eigensolver on S (Golub & Van Loan,
// Alg. 8.5.3). A 4x4 Cholesky test skips the eigensolver when S is positive
// definite, where makePD leaves H unchanged.
#pragma once

#include "oc_common.hpp"

namespace oc {

template <typename T>
struct SDConst;
template <>
struct SDConst<double> {
    static constexpr double inv_sqrt2 = 0.70710678118654752440;
    static constexpr double inv_sqrt6 = 0.40824829046386301637;
    static constexpr double jac_tol = 1e-30;   // off^2 <= tol * |S|_F^2 (~ (eps)^2)
    static constexpr int max_sweeps = 8;
};
template <>
struct SDConst<float> {
    static constexpr float inv_sqrt2 = 0.70710678118654752440f;
    static constexpr float inv_sqrt6 = 0.40824829046386301637f;
    static constexpr float jac_tol = 1e-13f;
    static constexpr int max_sweeps = 6;
};

// Rest features of one triangle.
template <typename T>
struct RestT {
    T w, A2, E0, E1, E01;
};

// Energy of one triangle at UV positions (U1,U2,U3) (SymDirichletEnergy.cpp:39-44
// written with the /(2A^2) features: e0SqLen/4/A^2 == E0/2).
template <typename T>
__device__ __forceinline__ T sd_energy(T u1x, T u1y, T u2x, T u2y, T u3x, T u3y, const RestT<T>& r, T& dbArea)
{
    const T ax = u2x - u1x, ay = u2y - u1y;  // U2m1
    const T bx = u3x - u1x, by = u3y - u1y;  // U3m1
    dbArea = ax * by - ay * bx;
    const T area_U = T(0.5) * dbArea;
    const T left = T(1) + r.A2 / area_U / area_U;
    const T right = ((bx * bx + by * by) * r.E0 + (ax * ax + ay * ay) * r.E1) / T(2) - (bx * ax + by * ay) * r.E01;
    return r.w * left * right;
}

// Energy, gradient (6) and the upper 2x2 blocks of the 6x6 Hessian (before
// projection), all multiplied by w. Block ids: 0:(1,1) 1:(1,2) 2:(1,3)
// 3:(2,2) 4:(2,3) 5:(3,3); each block row-major [xx, xy, yx, yy].
template <typename T, bool WANT_H>
__device__ __forceinline__ T sd_element(const T (&U)[3][2], const RestT<T>& r, T (&g)[6], T (&h)[6][4])
{
    const T U2m1[2] = {U[1][0] - U[0][0], U[1][1] - U[0][1]};
    const T U3m1[2] = {U[2][0] - U[0][0], U[2][1] - U[0][1]};
    const T area_U = T(0.5) * (U2m1[0] * U3m1[1] - U2m1[1] * U3m1[0]);
    const T areaRatio = r.A2 / area_U / area_U / area_U;
    const T dAR = T(3) / T(2) * areaRatio / area_U;
    const T left = T(1) + r.A2 / area_U / area_U;
    const T right = ((U3m1[0] * U3m1[0] + U3m1[1] * U3m1[1]) * r.E0 +
                     (U2m1[0] * U2m1[0] + U2m1[1] * U2m1[1]) * r.E1) / T(2) -
                    (U3m1[0] * U2m1[0] + U3m1[1] * U2m1[1]) * r.E01;
    // edge opposite vertex k, rotated by -90 degrees: (e.y, -e.x)
    const T o[3][2] = {{U[2][1] - U[1][1], -(U[2][0] - U[1][0])},
                       {U[0][1] - U[2][1], -(U[0][0] - U[2][0])},
                       {U[1][1] - U[0][1], -(U[1][0] - U[0][0])}};
    T dL[3][2], dR[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dL[k][0] = areaRatio * o[k][0];
        dL[k][1] = areaRatio * o[k][1];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        dR[0][c] = (r.E01 - r.E0) * U3m1[c] + (r.E01 - r.E1) * U2m1[c];
        dR[1][c] = r.E1 * U2m1[c] - r.E01 * U3m1[c];
        dR[2][c] = r.E0 * U3m1[c] - r.E01 * U2m1[c];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        g[2 * k + 0] = r.w * (dL[k][0] * right + dR[k][0] * left);
        g[2 * k + 1] = r.w * (dL[k][1] * right + dR[k][1] * left);
    }
    if (WANT_H) {
        const T E0 = r.E0, E1 = r.E1, E01 = r.E01;
        // d2Right_ij (SymDirichletEnergy.cpp:486,493,500,507,513,520)
        const T d2R[6] = {E0 + E1 - T(2) * E01, E01 - E1, E01 - E0, E1, -E01, E0};
        // sign of the areaRatio * dOrtho/dU term in d2Left_ij (:491-492 +, :498-499 -, :511-512 +)
        const T rs[6] = {T(0), T(1), T(-1), T(0), T(1), T(0)};
        const int bi[6] = {0, 0, 0, 1, 1, 2}, bj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            const int i = bi[b], j = bj[b];
#pragma unroll
            for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    // dOrtho/dU = [[0,-1],[1,0]]
                    const T Rrc = (rr == 0 && cc == 1) ? T(-1) : ((rr == 1 && cc == 0) ? T(1) : T(0));
                    const T d2L = dAR * o[i][rr] * o[j][cc] + areaRatio * (rs[b] * Rrc);
                    T v = d2L * right;
                    if (i == j) v = v + dL[i][rr] * dR[i][cc];
                    else v = v + dL[i][rr] * dR[j][cc];
                    if (rr == cc) v = v + d2R[b] * left;
                    if (i == j) v = v + dL[i][cc] * dR[i][rr];
                    else v = v + dR[i][rr] * dL[j][cc];
                    h[b][2 * rr + cc] = r.w * v;
                }
        }
    }
    return r.w * left * right;
}

// ----------------------------------------------------------------------------
// 4x4 symmetric Jacobi rotation on pair (p,q) of a (full storage) with
// eigenvector accumulation in v.
template <typename T, int p, int q>
__device__ __forceinline__ void jacobi_rot(T (&a)[4][4], T (&v)[4][4])
{
    const T apq = a[p][q];
    if (apq == T(0)) return;
    const T theta = (a[q][q] - a[p][p]) / (T(2) * apq);
    T t;
    const T at = theta < T(0) ? -theta : theta;
    if (at > T(1e18)) t = T(0.5) / theta;
    else t = (theta >= T(0) ? T(1) : T(-1)) / (at + sqrt(theta * theta + T(1)));
    const T c = T(1) / sqrt(t * t + T(1));
    const T s = t * c;
    a[p][p] = a[p][p] - t * apq;
    a[q][q] = a[q][q] + t * apq;
    a[p][q] = a[q][p] = T(0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r == p || r == q) continue;
        const T arp = a[r][p], arq = a[r][q];
        a[r][p] = a[p][r] = c * arp - s * arq;
        a[r][q] = a[q][r] = s * arp + c * arq;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const T vrp = v[r][p], vrq = v[r][q];
        v[r][p] = c * vrp - s * vrq;
        v[r][q] = s * vrp + c * vrq;
    }
}

// Project the element Hessian (upper blocks h, already scaled by w) onto the
// PSD cone in place. Returns true when the matrix was changed.
template <typename T>
__device__ __forceinline__ bool sd_project_psd(T (&h)[6][4])
{
    using K = SDConst<T>;
    // basis coefficients b_k[i]
    const T b[2][3] = {{K::inv_sqrt2, -K::inv_sqrt2, T(0)}, {K::inv_sqrt6, K::inv_sqrt6, T(-2) * K::inv_sqrt6}};
    // full block access Hb(i,j) (2x2) from the upper blocks
    auto blk = [&](int i, int j, int rr, int cc) -> T {
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const int id = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
        return i <= j ? h[id][2 * rr + cc] : h[id][2 * cc + rr];
    };
    // M_k(j) = sum_i b_k[i] Hb(i,j)
    T M[2][3][4];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rr = e >> 1, cc = e & 1;
                M[k][j][e] = b[k][0] * blk(0, j, rr, cc) + b[k][1] * blk(1, j, rr, cc) + b[k][2] * blk(2, j, rr, cc);
            }
    // S(k,l) = sum_j M_k(j) b_l[j]; index (k,d) -> 2k+d
    T a[4][4];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int l = 0; l < 2; ++l)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rr = e >> 1, cc = e & 1;
                a[2 * k + rr][2 * l + cc] = M[k][0][e] * b[l][0] + M[k][1][e] * b[l][1] + M[k][2][e] * b[l][2];
            }
    // symmetrize (rounding-level asymmetry from the two-sided product)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = r + 1; c < 4; ++c) {
            const T m = T(0.5) * (a[r][c] + a[c][r]);
            a[r][c] = a[c][r] = m;
        }
    // Cholesky test: positive pivots <=> positive definite
    {
        T L[4][4];
        bool pd = true;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T d = a[j][j];
#pragma unroll
            for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
            pd = pd && (d > T(0));
            const T ljj = sqrt(d > T(0) ? d : T(1));
            L[j][j] = ljj;
#pragma unroll
            for (int i = j + 1; i < 4; ++i) {
                T s = a[i][j];
#pragma unroll
                for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
                L[i][j] = s / ljj;
            }
        }
        if (pd) return false;
    }
    T v[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] = (r == c) ? T(1) : T(0);
    T fro = T(0);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) fro += a[r][c] * a[r][c];
    for (int sweep = 0; sweep < K::max_sweeps; ++sweep) {
        const T off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[0][3] * a[0][3] + a[1][2] * a[1][2] +
                      a[1][3] * a[1][3] + a[2][3] * a[2][3];
        if (off <= K::jac_tol * fro) break;
        jacobi_rot<T, 0, 1>(a, v);
        jacobi_rot<T, 0, 2>(a, v);
        jacobi_rot<T, 0, 3>(a, v);
        jacobi_rot<T, 1, 2>(a, v);
        jacobi_rot<T, 1, 3>(a, v);
        jacobi_rot<T, 2, 3>(a, v);
    }
    bool changed = false;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const T lam = a[m][m];
        if (!(lam < T(0))) continue;
        changed = true;
        // u = B v_m: u[(i,d)] = sum_k b_k[i] v[(k,d)][m]
        T u[3][2];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int d = 0; d < 2; ++d) u[i][d] = b[0][i] * v[d][m] + b[1][i] * v[2 + d][m];
        const int bi[6] = {0, 0, 0, 1, 1, 2}, bj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int id = 0; id < 6; ++id)
#pragma unroll
            for (int e = 0; e < 4; ++e) h[id][e] -= lam * u[bi[id]][e >> 1] * u[bj[id]][e & 1];
    }
    // keep the diagonal blocks exactly symmetric
#pragma unroll
    for (int id = 0; id < 6; id += 3) h[id][2] = h[id][1];
    h[5][2] = h[5][1];
    return changed;
}

// Largest injectivity-safe step for one triangle (SymDirichletEnergy.cpp:567-604),
// FP64 always. Returns cap when no root limits the step.
__device__ __forceinline__ double sd_step_bound(const double (&U)[3][2], const double (&Vd)[3][2], double cap)
{
    const double U2m1[2] = {U[1][0] - U[0][0], U[1][1] - U[0][1]};
    const double U3m1[2] = {U[2][0] - U[0][0], U[2][1] - U[0][1]};
    const double V2m1[2] = {Vd[1][0] - Vd[0][0], Vd[1][1] - Vd[0][1]};
    const double V3m1[2] = {Vd[2][0] - Vd[0][0], Vd[2][1] - Vd[0][1]};
    const double a = __dsub_rn(mul_rn(V2m1[0], V3m1[1]), mul_rn(V2m1[1], V3m1[0]));
    // evaluated left to right as ((p1 - p2) + p3) - p4, each product rounded once
    const double b = __dsub_rn(add_rn(__dsub_rn(mul_rn(U2m1[0], V3m1[1]), mul_rn(U2m1[1], V3m1[0])),
                                      mul_rn(V2m1[0], U3m1[1])),
                               mul_rn(V2m1[1], U3m1[0]));
    const double c = __dsub_rn(mul_rn(U2m1[0], U3m1[1]), mul_rn(U2m1[1], U3m1[0]));
    const double delta = __dsub_rn(mul_rn(b, b), mul_rn(mul_rn(4.0, a), c));
    double bound = cap;
    if (a > 0.0) {
        if ((b < 0.0) && (delta >= 0.0)) bound = 2.0 * c / (-b + sqrt(delta));
    } else if (a < 0.0) {
        if (b < 0.0) bound = 2.0 * c / (-b + sqrt(delta));
        else bound = (-b - sqrt(delta)) / 2.0 / a;
    } else {
        if (b < 0.0) bound = -c / b;
    }
    return bound;
}

}  // namespace oc
