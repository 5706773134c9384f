// Smoothed-aggregation AMG on the GPU (see oc_amg.hpp for the method).
#include "oc_amg.hpp"
#include "oc_amg_structure.hpp"
#include "oc_common.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <type_traits>

namespace oc {

namespace {

constexpr int NS = 3;              // near-null-space vectors: two translations and a rotation
constexpr int kAmgMaxGrid = 1024;  // workgroups of a row kernel (bounds partial records)
#ifndef OC_COARSE_MAX_DOFS
#define OC_COARSE_MAX_DOFS 64
#endif
// Coarsening stops at the first level with at most this many unknowns; its dense inverse is formed in LDS (up to
// 128 x 128 doubles = 128 KiB of gfx950's 160 KiB would fit). Measured at 200k faces: 128 (a 38-node coarsest
// level instead of 4) saves 13 us per V-cycle but its Gauss-Jordan inverse, one workgroup with a barrier pair per
// pivot, takes 0.58 ms per set-up, so the disc step slows by 9%; 64 stays.
constexpr int kCoarseMaxDofs = OC_COARSE_MAX_DOFS;
constexpr int kLanczosSteps = 15;       // Lanczos steps of the lambda_max estimate (fine level; array size)
constexpr int kLanczosStepsCoarse = 10;  // on coarse levels
constexpr int kMaxLevels = 10;

__device__ __forceinline__ void chunk_rows(int n, int& r0, int& r1)
{
    const int c = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    r0 = (int)blockIdx.x * c;
    r1 = min(n, r0 + c);
}

int rows_grid(int n, int rowsPerPass)
{
    return std::max(1, std::min((n + rowsPerPass - 1) / rowsPerPass, kAmgMaxGrid));
}

// y = sum over the row of A_uw x_w, LANES lanes per row (result in every lane of the group)
template <typename T, int BS, int LANES, typename X>
__device__ __forceinline__ void row_mv(const int* __restrict__ rowPtr, const int* __restrict__ colIdx,
                                       const T* __restrict__ vals, const X* __restrict__ x, int u, int lane,
                                       T (&y)[BS])
{
#pragma unroll
    for (int i = 0; i < BS; ++i) y[i] = T(0);
    const int k1 = rowPtr[u + 1];
    for (int k = rowPtr[u] + lane; k < k1; k += LANES) {
        const int w = colIdx[k];
        const T* a = vals + (size_t)k * BS * BS;
        T xw[BS];
#pragma unroll
        for (int j = 0; j < BS; ++j) xw[j] = (T)x[(size_t)w * BS + j];
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int j = 0; j < BS; ++j) y[i] += a[i * BS + j] * xw[j];
    }
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < BS; ++i) y[i] += __shfl_xor(y[i], off, LANES);
}

// Chebyshev parameters on [lmax/30, lmax], lmax = 1.1 * lambda estimate
struct ChebP {
    double theta, delta, sigma;
};
__device__ __forceinline__ ChebP cheb_params(double lam)
{
    const double lmax = 1.1 * lam, lmin = lmax / 30.0;
    ChebP p;
    p.theta = 0.5 * (lmax + lmin);
    p.delta = 0.5 * (lmax - lmin);
    p.sigma = p.theta / p.delta;
    return p;
}

// ---------------------------------------------------------------------------- numeric set-up kernels
template <typename TI, typename T>
__global__ void __launch_bounds__(kBlock) k_cvt(long n, const TI* __restrict__ in, T* out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)in[i];
}

// Inverse of a small SPD block in double: closed form for 2x2, Gauss-Jordan
// without pivoting otherwise.
template <int BS>
__device__ __forceinline__ void block_inverse(const double (&m)[BS][BS], double (&inv)[BS][BS])
{
    if (BS == 2) {
        const double det = m[0][0] * m[1][1] - m[0][1] * m[1][0];
        inv[0][0] = m[1][1] / det;
        inv[0][1] = -m[0][1] / det;
        inv[1][0] = -m[1][0] / det;
        inv[1][1] = m[0][0] / det;
    } else {
        double w[BS][2 * BS];
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int j = 0; j < 2 * BS; ++j) w[i][j] = j < BS ? m[i][j] : (j - BS == i ? 1.0 : 0.0);
#pragma unroll
        for (int p = 0; p < BS; ++p) {
            const double piv = 1.0 / w[p][p];
#pragma unroll
            for (int j = 0; j < 2 * BS; ++j) w[p][j] *= piv;
#pragma unroll
            for (int i = 0; i < BS; ++i) {
                if (i == p) continue;
                const double f = w[i][p];
#pragma unroll
                for (int j = 0; j < 2 * BS; ++j) w[i][j] -= f * w[p][j];
            }
        }
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int j = 0; j < BS; ++j) inv[i][j] = w[i][BS + j];
    }
}

// Inverse of each diagonal block; an exactly zero diagonal entry (a coarse
// DOF whose prolongator column vanished) is set to 1 in A as well.
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_blockinv(int n, const int* __restrict__ diagPos, T* vals, T* Dinv)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    T* a = vals + (size_t)diagPos[u] * BS * BS;
    double m[BS][BS];
#pragma unroll
    for (int i = 0; i < BS; ++i) {
#pragma unroll
        for (int j = 0; j < BS; ++j) m[i][j] = (double)a[i * BS + j];
        if (m[i][i] == 0.0) {
            m[i][i] = 1.0;
            a[i * BS + i] = T(1);
        }
    }
    double inv[BS][BS];
    block_inverse<BS>(m, inv);
    T* o = Dinv + (size_t)u * BS * BS;
#pragma unroll
    for (int i = 0; i < BS; ++i)
#pragma unroll
        for (int j = 0; j < BS; ++j) o[i * BS + j] = (T)inv[i][j];
}

// FP32-stored levels: adds to each diagonal entry `factor` times its scalar row's absolute sum. Rounding a
// symmetric positive definite matrix to FP32 perturbs each entry by at most 2^-24 of itself, a perturbation E
// with sum_j |E_ij| <= 2^-24 sum_j |A_ij|; with factor >= 2^-24 (plus the rounding of the shifted diagonal
// itself) E + shift is diagonally dominant, so the stored matrix stays positive definite whatever the
// condition number (a V-cycle on an indefinite level breaks PCG). The shift only changes the preconditioner:
// PCG multiplies by the FP64 matrix.
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_shift_diag(int n, const int* __restrict__ rowPtr,
                                                       const int* __restrict__ diagPos, T* vals, double factor)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    double rs[BS] = {};
    for (int k = rowPtr[u]; k < rowPtr[u + 1]; ++k) {
        const T* b = vals + (size_t)k * BS * BS;
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int j = 0; j < BS; ++j) rs[i] += fabs((double)b[i * BS + j]);
    }
    T* d = vals + (size_t)diagPos[u] * BS * BS;
#pragma unroll
    for (int i = 0; i < BS; ++i) d[i * BS + i] = (T)((double)d[i * BS + i] + factor * rs[i]);
}

__device__ __forceinline__ double hash_unit(unsigned int i)
{
    i ^= i >> 16;
    i *= 0x7feb352du;
    i ^= i >> 15;
    i *= 0x846ca68bu;
    i ^= i >> 16;
    return (double)(i & 0xffffff) / (double)0x800000 - 1.0;
}

// lambda_max(D^-1 A) by Lanczos in the D inner product (D^-1 A is self-adjoint
// there; D is the block diagonal with zero diagonal entries set to 1, as in
// k_blockinv). Ritz values approach the extreme eigenvalues far faster than
// power iteration does when the top of the spectrum is clustered (large
// meshes), and they never exceed lambda_max.
//
// One kernel per step j with every grid reduction folded into the next step.
// With M = D^-1 A, W_{j-1} = M v_{j-1} (W_{-1} = D^-1 h for a fixed
// pseudo-random h), alpha_{j-1} = v_{j-1}^T A v_{j-1} and gamma_{j-1} =
// |W_{j-1}|_D^2 (partial records of the previous step):
//   beta_j^2 = gamma_{j-1} - alpha_{j-1}^2 - beta_{j-1}^2,
//   v_j = (W_{j-1} - alpha_{j-1} v_{j-1} - beta_{j-1} v_{j-2}) / beta_j,
//   W_j = M v_j = (M W_{j-1} - alpha_{j-1} W_{j-1} - beta_{j-1} W_{j-2}) / beta_j,
// so the product gathers the single vector W_{j-1}; everything else is the
// row's own entries. A v_j = D W_j gives this step's records v_j^T D W_j and
// W_j^T D W_j. Every workgroup reduces the previous step's records itself
// (same order, same result) after issuing its row loads, so the reduction
// overlaps their latency; workgroup 0 keeps the coefficients. coef:
// alpha[kLanczosSteps], beta[kLanczosSteps], then the number of valid steps
// (fewer after an invariant subspace is found).
struct LanczosBufs {
    double* v[3];  // v_j in v[j % 3]
    double* W[3];  // W_j in W[(j + 1) % 3] (W_{-1} in W[0])
};

// Lane-group shape of the row kernels (Lanczos step, Chebyshev steps,
// residual): lanes per block row and rows per lane group per pass. 2x2 rows
// hold about 7 blocks: 4 lanes and two rows in flight; 3x3 coarse rows hold
// 10-30 blocks: 16 lanes, one row.
template <int BS>
struct RowShape;
template <>
struct RowShape<2> {
    static constexpr int lanes = 4, unroll = 2;
};
template <>
struct RowShape<3> {
    static constexpr int lanes = 16, unroll = 1;
};

// Workgroups of a row kernel: enough for one pass over the rows, at most
// kAmgMaxGrid, a multiple of 8 (the XCD-contiguous chunk order below).
int row_grid(int n, int bs)
{
    const int rows = bs == 2 ? (kBlock / RowShape<2>::lanes) * RowShape<2>::unroll
                             : (kBlock / RowShape<3>::lanes) * RowShape<3>::unroll;
    const int g = std::max(1, std::min((n + rows - 1) / rows, kAmgMaxGrid));
    return (g + 7) & ~7;
}

// Lanczos steps per level: the fine level's top spectrum is clustered on large
// meshes and needs all kLanczosSteps; coarse levels settle sooner. At 200k
// faces 8 coarse steps lose nothing and 6 under-estimate often enough to force
// a safe-mode retry; 10 keeps a margin.
#ifndef OC_LANCZOS_FINE
#define OC_LANCZOS_FINE kLanczosSteps
#endif
#ifndef OC_LANCZOS_COARSE
#define OC_LANCZOS_COARSE kLanczosStepsCoarse
#endif
int lanczos_steps(int level) { return level == 0 ? OC_LANCZOS_FINE : OC_LANCZOS_COARSE; }

// Workgroup b is dispatched to XCD b % 8; chunk c of the rows goes to
// workgroup (c % (G/8)) * 8 + c / (G/8), so each XCD owns one contiguous
// eighth of the rows and its L2 serves most neighbour gathers.
__device__ __forceinline__ int xcd_chunk(int b, int G) { return (b & 7) * (G >> 3) + (b >> 3); }

// Block, vector-entry and store helpers: one 16-byte load per pair of doubles
// or per 2x2 float block where the layout allows.
template <int BS, typename T, typename R>
__device__ __forceinline__ void ld_block(const T* __restrict__ p, R (&a)[BS * BS])
{
    if constexpr (BS == 2 && std::is_same<T, double>::value) {
        const double2 x = reinterpret_cast<const double2*>(p)[0], y = reinterpret_cast<const double2*>(p)[1];
        a[0] = (R)x.x;
        a[1] = (R)x.y;
        a[2] = (R)y.x;
        a[3] = (R)y.y;
    } else if constexpr (BS == 2 && std::is_same<T, float>::value) {
        const float4 x = reinterpret_cast<const float4*>(p)[0];
        a[0] = (R)x.x;
        a[1] = (R)x.y;
        a[2] = (R)x.z;
        a[3] = (R)x.w;
    } else {
#pragma unroll
        for (int i = 0; i < BS * BS; ++i) a[i] = (R)p[i];
    }
}

template <int BS, typename T, typename R>
__device__ __forceinline__ void ld_vec(const T* __restrict__ p, size_t u, R (&x)[BS])
{
    if constexpr (BS == 2 && std::is_same<T, double>::value) {
        const double2 t = reinterpret_cast<const double2*>(p)[u];
        x[0] = (R)t.x;
        x[1] = (R)t.y;
    } else if constexpr (BS == 2 && std::is_same<T, float>::value) {
        const float2 t = reinterpret_cast<const float2*>(p)[u];
        x[0] = (R)t.x;
        x[1] = (R)t.y;
    } else {
#pragma unroll
        for (int i = 0; i < BS; ++i) x[i] = (R)p[u * BS + i];
    }
}

template <int BS, typename T>
__device__ __forceinline__ void st_vec(T* p, size_t u, const T (&x)[BS])
{
    if constexpr (BS == 2 && std::is_same<T, double>::value) {
        reinterpret_cast<double2*>(p)[u] = make_double2(x[0], x[1]);
    } else if constexpr (BS == 2 && std::is_same<T, float>::value) {
        reinterpret_cast<float2*>(p)[u] = make_float2(x[0], x[1]);
    } else {
#pragma unroll
        for (int i = 0; i < BS; ++i) p[u * BS + i] = x[i];
    }
}

// One pass of a row kernel: UNROLL block rows per LANES-lane group. load()
// issues the row pointers, the first block of every lane (a row of at most
// LANES blocks needs no other) and its gathered x entries, so the caller can
// issue its own-row loads before any of them is consumed; mv() finishes
// y = (A x)_u (remaining blocks, lane-group reduction; y in every lane).
template <typename T, int BS, int UNROLL>
struct RowPass {
    static constexpr int LANES = RowShape<BS>::lanes;
    int u[UNROLL], kb[UNROLL], ke[UNROLL];
    T a[UNROLL][BS * BS], g[UNROLL][BS];

    template <typename X>
    __device__ __forceinline__ void load(int first, int stride, int r1, const int* __restrict__ rowPtr,
                                         const int* __restrict__ colIdx, const T* __restrict__ vals,
                                         const X* __restrict__ x, int lane)
    {
#pragma unroll
        for (int r = 0; r < UNROLL; ++r) {
            u[r] = first + r * stride;
            const bool valid = u[r] < r1;
            kb[r] = valid ? rowPtr[u[r]] : 0;
            ke[r] = valid ? rowPtr[u[r] + 1] : 0;
        }
#pragma unroll
        for (int r = 0; r < UNROLL; ++r) {
            const int k = kb[r] + lane;
            const bool has = k < ke[r];
            const int kk = has ? k : 0;
            ld_block<BS>(vals + (size_t)kk * BS * BS, a[r]);
            ld_vec<BS>(x, has ? colIdx[kk] : 0, g[r]);
            if (!has)
#pragma unroll
                for (int i = 0; i < BS * BS; ++i) a[r][i] = T(0);
        }
    }

    template <typename X, typename Y>
    __device__ __forceinline__ void mv(int r, const int* __restrict__ colIdx, const T* __restrict__ vals,
                                       const X* __restrict__ x, int lane, Y (&y)[BS]) const
    {
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            y[i] = Y(0);
#pragma unroll
            for (int q = 0; q < BS; ++q) y[i] += (Y)a[r][i * BS + q] * (Y)g[r][q];
        }
        for (int k = kb[r] + lane + LANES; k < ke[r]; k += LANES) {
            T ak[BS * BS];
            Y gk[BS];
            ld_block<BS>(vals + (size_t)k * BS * BS, ak);
            ld_vec<BS>(x, colIdx[k], gk);
#pragma unroll
            for (int i = 0; i < BS; ++i)
#pragma unroll
                for (int q = 0; q < BS; ++q) y[i] += (Y)ak[i * BS + q] * gk[q];
        }
#pragma unroll
        for (int off = LANES / 2; off > 0; off >>= 1)
#pragma unroll
            for (int i = 0; i < BS; ++i) y[i] += __shfl_xor(y[i], off, LANES);
    }
};

// Rows of this workgroup's chunk and the pass count (block-uniform).
struct RowChunk {
    int r0, r1, passes;
};
// Row chunk of workgroup b of G (the V-cycle kernels' bodies also run as one
// workgroup covering every row: b = 0, G = 1).
template <int BS>
__device__ __forceinline__ RowChunk row_chunk(int n, int b, int G)
{
    constexpr int RPP = kBlock / RowShape<BS>::lanes, UNROLL = RowShape<BS>::unroll;
    const int c = (n + G - 1) / G;
    RowChunk rc;
    rc.r0 = xcd_chunk(b, G) * c;
    rc.r1 = min(n, rc.r0 + c);
    rc.passes = max(1, (c + RPP * UNROLL - 1) / (RPP * UNROLL));
    return rc;
}

__device__ __forceinline__ void lanczos_coeffs(int j, const double* __restrict__ partIn, int nPartIn,
                                               const double* __restrict__ coef, double& aprev, double& bprev,
                                               double& bj)
{
    double acc[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nPartIn; b += blockDim.x) {
        acc[0] += partIn[2 * b];
        acc[1] += partIn[2 * b + 1];
    }
    block_sum<2>(acc);
    __shared__ double sh[2];
    if (threadIdx.x == 0) {
        sh[0] = acc[0];
        sh[1] = acc[1];
    }
    __syncthreads();
    aprev = j == 0 ? 0.0 : sh[0];
    bprev = j <= 1 ? 0.0 : coef[kLanczosSteps + j - 1];
    const double gamma = sh[1];
    const double b2 = gamma - aprev * aprev - bprev * bprev;
    bj = (gamma > 0.0 && b2 > 1e-10 * gamma) ? sqrt(b2) : 0.0;
}

// Launch with gridDim.x a multiple of 8 (row_grid).
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_lanczos(int j, int nsteps, int n, const int* __restrict__ rowPtr,
                                                    const int* __restrict__ colIdx, const T* __restrict__ vals,
                                                    const int* __restrict__ diagPos, LanczosBufs buf,
                                                    const double* __restrict__ partIn, int nPartIn, double* partOut,
                                                    double* coef)
{
    constexpr int LANES = RowShape<BS>::lanes, UNROLL = RowShape<BS>::unroll, RPP = kBlock / LANES;
    if (j > 0 && coef[2 * kLanczosSteps] < (double)nsteps) return;  // stopped earlier
    const double* __restrict__ Wp = buf.W[j % 3];
    const double* __restrict__ Wpp = buf.W[(j + 2) % 3];
    const double* __restrict__ v1 = buf.v[(j + 2) % 3];
    const double* __restrict__ v2 = buf.v[(j + 1) % 3];
    double* vj = buf.v[j % 3];
    double* Wj = buf.W[(j + 1) % 3];
    const int G = gridDim.x;
    const int c = (n + G - 1) / G;
    const int r0 = xcd_chunk(blockIdx.x, G) * c, r1 = min(n, r0 + c);
    const int sub = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    const int passes = max(1, (c + RPP * UNROLL - 1) / (RPP * UNROLL));
    double aprev = 0.0, bprev = 0.0, inv = 1.0;
    double acc[2] = {0.0, 0.0};
    for (int p = 0; p < passes; ++p) {
        int u[UNROLL], kb[UNROLL], ke[UNROLL];
        double a[UNROLL][BS * BS], g[UNROLL][BS];
        double oW[UNROLL][BS], oWp[UNROLL][BS], o1[UNROLL][BS], o2[UNROLL][BS], D[UNROLL][BS * BS];
#pragma unroll
        for (int r = 0; r < UNROLL; ++r) {
            u[r] = r0 + sub + (p * UNROLL + r) * RPP;
            const bool valid = u[r] < r1;
            kb[r] = valid ? rowPtr[u[r]] : 0;
            ke[r] = valid ? rowPtr[u[r] + 1] : 0;
        }
        // first slice of every row (a row of at most LANES blocks needs no other)
#pragma unroll
        for (int r = 0; r < UNROLL; ++r) {
            const int k = kb[r] + lane;
            const bool has = k < ke[r];
            const int kk = has ? k : 0;
            ld_block<BS>(vals + (size_t)kk * BS * BS, a[r]);
            ld_vec<BS>(Wp, has ? colIdx[kk] : 0, g[r]);
            if (!has)
#pragma unroll
                for (int i = 0; i < BS * BS; ++i) a[r][i] = 0.0;
#pragma unroll
            for (int i = 0; i < BS; ++i) oW[r][i] = oWp[r][i] = o1[r][i] = o2[r][i] = 0.0;
#pragma unroll
            for (int i = 0; i < BS * BS; ++i) D[r][i] = 0.0;
            if (lane == 0 && u[r] < r1) {
                ld_vec<BS>(Wp, u[r], oW[r]);
                if (j >= 1) {
                    ld_vec<BS>(Wpp, u[r], oWp[r]);
                    ld_vec<BS>(v1, u[r], o1[r]);
                }
                if (j >= 2) ld_vec<BS>(v2, u[r], o2[r]);
                ld_block<BS>(vals + (size_t)diagPos[u[r]] * BS * BS, D[r]);
            }
        }
        if (p == 0) {
            double bj;
            lanczos_coeffs(j, partIn, nPartIn, coef, aprev, bprev, bj);
            if (bj == 0.0) {  // the Krylov space is invariant: T so far is exact
                if (blockIdx.x == 0 && threadIdx.x == 0) {
                    if (j > 0) coef[j - 1] = aprev;
                    coef[2 * kLanczosSteps] = (double)j;
                }
                return;
            }
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                if (j > 0) coef[j - 1] = aprev;
                coef[kLanczosSteps + j] = bj;
            }
            inv = 1.0 / bj;
        }
#pragma unroll
        for (int r = 0; r < UNROLL; ++r) {
            double y[BS];
#pragma unroll
            for (int i = 0; i < BS; ++i) {
                y[i] = 0.0;
#pragma unroll
                for (int q = 0; q < BS; ++q) y[i] += a[r][i * BS + q] * g[r][q];
            }
            for (int k = kb[r] + lane + LANES; k < ke[r]; k += LANES) {
                double ak[BS * BS], gk[BS];
                ld_block<BS>(vals + (size_t)k * BS * BS, ak);
                ld_vec<BS>(Wp, colIdx[k], gk);
#pragma unroll
                for (int i = 0; i < BS; ++i)
#pragma unroll
                    for (int q = 0; q < BS; ++q) y[i] += ak[i * BS + q] * gk[q];
            }
#pragma unroll
            for (int off = LANES / 2; off > 0; off >>= 1)
#pragma unroll
                for (int i = 0; i < BS; ++i) y[i] += __shfl_xor(y[i], off, LANES);
            if (lane == 0 && u[r] < r1) {  // y = (A W_{j-1})_u
                double m[BS][BS], mi[BS][BS];
#pragma unroll
                for (int i = 0; i < BS; ++i)
#pragma unroll
                    for (int q = 0; q < BS; ++q) m[i][q] = D[r][i * BS + q];
                block_inverse<BS>(m, mi);
                double w[BS], x[BS];
#pragma unroll
                for (int i = 0; i < BS; ++i) {
                    double t = 0.0;
#pragma unroll
                    for (int q = 0; q < BS; ++q) t += mi[i][q] * y[q];
                    t -= aprev * oW[r][i];
                    if (j >= 1) t -= bprev * oWp[r][i];
                    w[i] = t * inv;
                    double xv = oW[r][i];
                    if (j >= 1) xv -= aprev * o1[r][i];
                    if (j >= 2) xv -= bprev * o2[r][i];
                    x[i] = xv * inv;
                }
#pragma unroll
                for (int i = 0; i < BS; ++i) {
                    double z = 0.0;  // (A v_j)_i = (D W_j)_i
#pragma unroll
                    for (int q = 0; q < BS; ++q) z += m[i][q] * w[q];
                    acc[0] += x[i] * z;
                    acc[1] += w[i] * z;
                }
                st_vec<BS>(vj, u[r], x);
                st_vec<BS>(Wj, u[r], w);
            }
        }
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partOut[2 * blockIdx.x] = acc[0];
        partOut[2 * blockIdx.x + 1] = acc[1];
    }
}

// Start vector and Gershgorin bound. g_u = max over the vertex's rows of
// sum_j |(D^-1 A)_ij| bounds lambda_max(D^-1 A) by max_u g_u. The largest
// eigenvalues of D^-1 A on distorted meshes belong to eigenvectors localised
// on a few vertices, which a uniform random start barely touches (Lanczos
// then needs many more steps); weighting the start by g_u^4 puts its mass
// where those vectors live. h = hash * g^4, W_{-1} = D^-1 h, partial records
// of h^T D^-1 h = |W_{-1}|_D^2 and of max g (gPart).
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_lanczos_start(int n, const int* __restrict__ rowPtr,
                                                          const T* __restrict__ vals, const T* __restrict__ Dinv,
                                                          double* W, double* partOut, double* gPart, double* coef,
                                                          int nsteps)
{
    int r0, r1;
    chunk_rows(n, r0, r1);
    double acc[2] = {0.0, 0.0};
    double gmax = 0.0;
    for (int u = r0 + (int)threadIdx.x; u < r1; u += kBlock) {
        const T* D = Dinv + (size_t)u * BS * BS;
        double rs[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) rs[i] = 0.0;
        for (int k = rowPtr[u]; k < rowPtr[u + 1]; ++k) {
            const T* a = vals + (size_t)k * BS * BS;
#pragma unroll
            for (int i = 0; i < BS; ++i)
#pragma unroll
                for (int j = 0; j < BS; ++j) {
                    double s = 0.0;
#pragma unroll
                    for (int c = 0; c < BS; ++c) s += (double)D[i * BS + c] * (double)a[c * BS + j];
                    rs[i] += fabs(s);
                }
        }
        double gu = 0.0;
#pragma unroll
        for (int i = 0; i < BS; ++i) gu = fmax(gu, rs[i]);
        gmax = fmax(gmax, gu);
        const double g2 = gu * gu, w4 = g2 * g2;
        double h[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) h[i] = hash_unit((unsigned)(u * BS + i)) * w4;
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < BS; ++c) s += (double)D[i * BS + c] * h[c];
            W[(size_t)u * BS + i] = s;
            acc[1] += h[i] * s;
        }
    }
    block_sum<2>(acc);
    // block maximum of g (values are >= 0)
    gmax = wave_max(gmax);
    __shared__ double gsh[kBlock / kWave];
    if (threadIdx.x % kWave == 0) gsh[threadIdx.x / kWave] = gmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        partOut[2 * blockIdx.x] = acc[0];
        partOut[2 * blockIdx.x + 1] = acc[1];
        double m = gsh[0];
        for (int w = 1; w < kBlock / kWave; ++w) m = fmax(m, gsh[w]);
        gPart[blockIdx.x] = m;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) coef[2 * kLanczosSteps] = (double)nsteps;
}

// lam[1] = the Gershgorin bound; lam[0] = the largest eigenvalue of the
// Lanczos tridiagonal (Sturm-count bisection), or, in the safe mode (no
// Lanczos steps were run), the Gershgorin bound / 1.1 so that the smoother's
// interval ends exactly at a guaranteed upper bound.
__global__ void __launch_bounds__(kBlock) k_lanczos_finish(const double* __restrict__ partIn, int nPartIn,
                                                           const double* __restrict__ gPart, int safe, int nsteps,
                                                           double* coef, double* lam)
{
    __shared__ double red[kBlock];
    {
        double g = 0.0;
        for (int b = threadIdx.x; b < nPartIn; b += blockDim.x) g = fmax(g, gPart[b]);
        red[threadIdx.x] = g;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            lam[1] = red[0];
            if (safe) lam[0] = red[0] / 1.1;
        }
        if (safe) return;
    }
    const int m = (int)coef[2 * kLanczosSteps];
    if (m == nsteps) {
        double acc[2] = {0.0, 0.0};
        for (int b = threadIdx.x; b < nPartIn; b += blockDim.x) {
            acc[0] += partIn[2 * b];
            acc[1] += partIn[2 * b + 1];
        }
        block_sum<2>(acc);
        if (threadIdx.x == 0) coef[nsteps - 1] = acc[0];
        __syncthreads();
    }
    if (m == 0) {  // start vector in the null space of A: no information
        if (threadIdx.x == 0) lam[0] = 1.0;
        return;
    }
    // T: al[0..m), be[i] couples rows i-1 and i (i >= 1)
    __shared__ double al[kLanczosSteps], be[kLanczosSteps];
    if ((int)threadIdx.x < m) {
        al[threadIdx.x] = coef[threadIdx.x];
        be[threadIdx.x] = coef[kLanczosSteps + threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x >= kWave) return;  // the rest runs in wave 0
    const int l = threadIdx.x;
    // Gershgorin interval of T
    double lo = 1e300, hi = -1e300;
    if (l < m) {
        const double r = (l > 0 ? fabs(be[l]) : 0.0) + (l + 1 < m ? fabs(be[l + 1]) : 0.0);
        lo = al[l] - r;
        hi = al[l] + r;
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    // Multisection on the Sturm count (number of eigenvalues of T below x):
    // 64 points per round, one per lane, shrink [lo, hi] around lambda_max by
    // 65x; counts grow with x, so the first point below which all m eigenvalues
    // lie brackets lambda_max.
    for (int round = 0; round < 6; ++round) {
        const double h = (hi - lo) / (kWave + 1);
        const double x = lo + (l + 1) * h;
        int cnt = 0;
        double d = 1.0;
        for (int i = 0; i < m; ++i) {
            d = al[i] - x - (i > 0 ? be[i] * be[i] / d : 0.0);
            if (d == 0.0) d = -1e-300;
            if (d < 0.0) ++cnt;
        }
        const unsigned long long hits = __ballot(cnt == m);
        const int k = hits ? __ffsll((long long)hits) - 1 : kWave;
        const double base = lo;
        lo = base + k * h;
        if (k < kWave) hi = base + (k + 1) * h;
    }
    if (l == 0) lam[0] = hi;
}

// Lanes per output of the set-up kernels: the contribution lists of one
// output are split over a lane group and reduced by shuffles, so their
// scattered block loads are in flight together.
constexpr int kSetupLanes = 8;

template <int LANES, int NV>
__device__ __forceinline__ void group_sum(double (&v)[NV])
{
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] += __shfl_xor(v[i], off, LANES);
}

int setup_grid(long nOut, int lanes) { return (int)std::max<long>(1, (nOut * lanes + kBlock - 1) / kBlock); }

// Tentative prolongator: per aggregate, B (rows of its members) = Q R by
// Cholesky of the 3x3 Gram matrix (columns whose remaining norm vanishes are
// dropped: zero Q column, zero R row). Q rows go to Pt (per member node),
// R to Bout (the coarse node's near-null-space rows). kSetupLanes lanes per
// aggregate share its members.
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_tentative(int nc, const int* __restrict__ aggPtr,
                                                      const int* __restrict__ aggMem,
                                                      const double2* __restrict__ UV,
                                                      const double* __restrict__ Bin, T* Pt, double* Bout)
{
    constexpr int LANES = kSetupLanes;
    const int c = (blockIdx.x * blockDim.x + threadIdx.x) / LANES, lane = threadIdx.x % LANES;
    const bool active = c < nc;  // whole lane groups are active or not
    if (!active) return;
    const int m0 = aggPtr[c], m1 = aggPtr[c + 1];
    double cu = 0.0, cv = 0.0;
    if (BS == 2) {
        double cs[2] = {0.0, 0.0};
        for (int m = m0 + lane; m < m1; m += LANES) {
            const double2 p = UV[aggMem[m]];
            cs[0] += p.x;
            cs[1] += p.y;
        }
        group_sum<LANES, 2>(cs);
        cu = cs[0] / (double)(m1 - m0);
        cv = cs[1] / (double)(m1 - m0);
    }
    auto brows = [&](int u, double (&B)[BS][NS]) {
        if (BS == 2) {
            const double2 p = UV[u];
            B[0][0] = 1.0; B[0][1] = 0.0; B[0][2] = -(p.y - cv);
            B[1][0] = 0.0; B[1][1] = 1.0; B[1][2] = (p.x - cu);
        } else {
#pragma unroll
            for (int i = 0; i < BS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j) B[i][j] = Bin[(size_t)u * BS * NS + i * NS + j];
        }
    };
    double Gv[NS * NS] = {};
    for (int m = m0 + lane; m < m1; m += LANES) {
        double B[BS][NS];
        brows(aggMem[m], B);
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int a = 0; a < NS; ++a)
#pragma unroll
                for (int b = 0; b < NS; ++b) Gv[a * NS + b] += B[i][a] * B[i][b];
    }
    group_sum<LANES, NS * NS>(Gv);
    double R[NS][NS] = {};
    bool keep[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double d = Gv[j * NS + j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= R[k][j] * R[k][j];
        keep[j] = Gv[j * NS + j] > 0.0 && d > 1e-12 * Gv[j * NS + j];
        if (keep[j]) {
            R[j][j] = sqrt(d);
#pragma unroll
            for (int k = j + 1; k < NS; ++k) {
                double s = Gv[j * NS + k];
#pragma unroll
                for (int i = 0; i < j; ++i) s -= R[i][j] * R[i][k];
                R[j][k] = s / R[j][j];
            }
        }
    }
    for (int m = m0 + lane; m < m1; m += LANES) {
        const int u = aggMem[m];
        double B[BS][NS];
        brows(u, B);
        T* o = Pt + (size_t)u * BS * NS;
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            double q[NS];
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                double s = B[i][j];
#pragma unroll
                for (int k = 0; k < j; ++k) s -= q[k] * R[k][j];
                q[j] = keep[j] ? s / R[j][j] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < NS; ++j) o[i * NS + j] = (T)q[j];
        }
    }
    // Coarse near-null space in one global basis: the rows above rotate about
    // the aggregate centroid (well conditioned), while the next level needs the
    // rotation about one common point (the UV origin), so R becomes R C with
    // C = [[1, 0, -c_v], [0, 1, c_u], [0, 0, 1]] (rotation about the origin =
    // rotation about c + translation (-c_v, c_u)).
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            double* o = Bout + (size_t)c * NS * NS + i * NS;
            o[0] = R[i][0];
            o[1] = R[i][1];
            o[2] = BS == 2 ? (-cv * R[i][0] + cu * R[i][1] + R[i][2]) : R[i][2];
        }
}

// P(u,c) = [c == agg(u)] Pt_u - omega Dinv_u sum_k A_k Pt_{col(k)}, omega = 4/(3 lambda);
// kSetupLanes lanes per P block
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_smooth_p(int nP, const int* __restrict__ pRow,
                                                     const unsigned char* __restrict__ pIsTent,
                                                     const int* __restrict__ psPtr, const int* __restrict__ psA,
                                                     const int* __restrict__ colIdx, const T* __restrict__ vals,
                                                     const T* __restrict__ Dinv, const T* __restrict__ Pt,
                                                     const double* __restrict__ lam, T* P)
{
    constexpr int LANES = kSetupLanes;
    const int j = (blockIdx.x * blockDim.x + threadIdx.x) / LANES, lane = threadIdx.x % LANES;
    if (j >= nP) return;
    const int u = pRow[j];
    double S[BS * NS] = {};
    for (int q = psPtr[j] + lane; q < psPtr[j + 1]; q += LANES) {
        const int k = psA[q];
        const T* a = vals + (size_t)k * BS * BS;
        const T* pt = Pt + (size_t)colIdx[k] * BS * NS;
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int l = 0; l < BS; ++l)
#pragma unroll
                for (int c = 0; c < NS; ++c) S[i * NS + c] += (double)a[i * BS + l] * (double)pt[l * NS + c];
    }
    group_sum<LANES, BS * NS>(S);
    if (lane != 0) return;
    const double omega = 4.0 / (3.0 * *lam);
    const T* D = Dinv + (size_t)u * BS * BS;
    const T* ptu = Pt + (size_t)u * BS * NS;
    const bool tent = pIsTent[j] != 0;
    T* o = P + (size_t)j * BS * NS;
#pragma unroll
    for (int i = 0; i < BS; ++i)
#pragma unroll
        for (int c = 0; c < NS; ++c) {
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < BS; ++l) s += (double)D[i * BS + l] * S[l * NS + c];
            o[i * NS + c] = (T)((tent ? (double)ptu[i * NS + c] : 0.0) - omega * s);
        }
}

// T block = sum over (A block, P block) pairs of A_k P_j (BS x NS); kSetupLanes lanes per block
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_ap(int nT, const int* __restrict__ cPtr, const int* __restrict__ cA,
                                               const int* __restrict__ cP, const T* __restrict__ vals,
                                               const T* __restrict__ P, T* Tv)
{
    constexpr int LANES = kSetupLanes;
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) / LANES, lane = threadIdx.x % LANES;
    if (t >= nT) return;
    double S[BS * NS] = {};
    for (int q = cPtr[t] + lane; q < cPtr[t + 1]; q += LANES) {
        const T* a = vals + (size_t)cA[q] * BS * BS;
        const T* p = P + (size_t)cP[q] * BS * NS;
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int l = 0; l < BS; ++l)
#pragma unroll
                for (int c = 0; c < NS; ++c) S[i * NS + c] += (double)a[i * BS + l] * (double)p[l * NS + c];
    }
    group_sum<LANES, BS * NS>(S);
    if (lane != 0) return;
    T* o = Tv + (size_t)t * BS * NS;
#pragma unroll
    for (int i = 0; i < BS * NS; ++i) o[i] = (T)S[i];
}

// coarse block = sum over (P block, T block) pairs of P_j^T T_t (NS x NS); kSetupLanes lanes per block
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_ptap(int nAc, const int* __restrict__ cPtr, const int* __restrict__ cP,
                                                 const int* __restrict__ cT, const T* __restrict__ P,
                                                 const T* __restrict__ Tv, T* Ac)
{
    constexpr int LANES = kSetupLanes;
    const int a = (blockIdx.x * blockDim.x + threadIdx.x) / LANES, lane = threadIdx.x % LANES;
    if (a >= nAc) return;
    double S[NS * NS] = {};
    for (int q = cPtr[a] + lane; q < cPtr[a + 1]; q += LANES) {
        const T* p = P + (size_t)cP[q] * BS * NS;
        const T* t = Tv + (size_t)cT[q] * BS * NS;
#pragma unroll
        for (int l = 0; l < BS; ++l)
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j) S[i * NS + j] += (double)p[l * NS + i] * (double)t[l * NS + j];
    }
    group_sum<LANES, NS * NS>(S);
    if (lane != 0) return;
    T* o = Ac + (size_t)a * NS * NS;
#pragma unroll
    for (int i = 0; i < NS * NS; ++i) o[i] = (T)S[i];
}

// Dense coarsest matrix (double) and its inverse by in-place Gauss-Jordan in
// LDS, one workgroup (SPD: pivots on the diagonal):
//   a_pp' = 1/a_pp, a_pj' = a_pj/a_pp, a_ip' = -a_ip/a_pp, a_ij' = a_ij - a_ip a_pj/a_pp.
template <typename T>
__global__ void __launch_bounds__(256) k_dense_inverse(int nn, int bs, const int* __restrict__ rowPtr,
                                                       const int* __restrict__ colIdx, const T* __restrict__ vals,
                                                       double* Inv)
{
    __shared__ double W[kCoarseMaxDofs * kCoarseMaxDofs];
    const int N = nn * bs;
    for (int i = threadIdx.x; i < N * N; i += blockDim.x) W[i] = 0.0;
    __syncthreads();
    for (int u = 0; u < nn; ++u)
        for (int k = rowPtr[u] + threadIdx.x; k < rowPtr[u + 1]; k += blockDim.x) {
            const int w = colIdx[k];
            for (int i = 0; i < bs; ++i)
                for (int j = 0; j < bs; ++j) W[(u * bs + i) * N + w * bs + j] = (double)vals[(size_t)k * bs * bs + i * bs + j];
        }
    __syncthreads();
    // Gauss-Jordan without pivoting (the matrix is symmetric positive definite), two barriers per pivot. For
    // pivot p, colp[p & 1] holds column p with its own entry zeroed (so that the elimination leaves row p alone)
    // and pv[p & 1] its reciprocal pivot. Thread pair (j, half) eliminates column j over one half of the rows,
    // 8 rows per batch with all loads before the stores so that the LDS accesses overlap; then row p is scaled,
    // column p set, and column p + 1 and its pivot taken for the next step (final once p is eliminated).
    __shared__ double colp[2][kCoarseMaxDofs];
    __shared__ double pv[2];
    const int j = threadIdx.x >> 1, half = threadIdx.x & 1;
    const int hN = (N + 1) / 2, i0 = half * hN, i1 = min(N, i0 + hN);
    for (int i = threadIdx.x; i < N; i += blockDim.x) colp[0][i] = i == 0 ? 0.0 : W[i * N];
    if (threadIdx.x == 0) pv[0] = W[0] != 0.0 ? 1.0 / W[0] : 0.0;
    __syncthreads();
    for (int p = 0; p < N; ++p) {
        const double* cp = colp[p & 1];
        const double piv = pv[p & 1];
        if (j < N && j != p) {
            const double wpj = W[p * N + j] * piv;
            for (int ib = i0; ib < i1; ib += 8) {
                double w[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) w[q] = ib + q < i1 ? W[(ib + q) * N + j] : 0.0;
#pragma unroll
                for (int q = 0; q < 8; ++q) w[q] -= (ib + q < i1 ? cp[ib + q] : 0.0) * wpj;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (ib + q < i1 && ib + q != p) W[(ib + q) * N + j] = w[q];
            }
        }
        __syncthreads();
        const int pn = p + 1;
        // row p scaled (its entry in column p + 1 by the column pass below, which reads it)
        if (j < N && j != p && j != pn && half == 0) W[p * N + j] *= piv;
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            W[i * N + p] = i == p ? piv : -cp[i] * piv;
            if (pn < N) {
                double v = W[i * N + pn];
                if (i == p) {
                    v *= piv;
                    W[i * N + pn] = v;
                }
                colp[pn & 1][i] = i == pn ? 0.0 : v;
                if (i == pn) pv[pn & 1] = v != 0.0 ? 1.0 / v : 0.0;
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < N * N; i += blockDim.x) Inv[i] = W[i];
}

// ---------------------------------------------------------------------------- V-cycle kernels
// V-cycle kernels store vectors and matrices in T and compute in FP64: they are
// bound by memory, and FP64 arithmetic keeps the FP32-stored V-cycle as close
// to a fixed symmetric operator as its storage allows.
//
// pre-smoothing, first Chebyshev step from x = 0: r = D^-1 b, d = r/theta, x = d
template <typename T, int BS, typename BIN>
__global__ void __launch_bounds__(kBlock) k_cheb_pre1(const int* __restrict__ stop, int n, const BIN* __restrict__ bin, T* b,
                                                      const T* __restrict__ Dinv, const double* __restrict__ lam,
                                                      T* x, T* r, T* d)
{
    if (stop && *stop) return;
    const ChebP cp = cheb_params(*lam);
    int r0, r1;
    chunk_rows(n, r0, r1);
    for (int u = r0 + threadIdx.x; u < r1; u += blockDim.x) {
        double bu[BS], Dm[BS * BS];
        ld_vec<BS>(bin, u, bu);
        ld_block<BS>(Dinv + (size_t)u * BS * BS, Dm);
        if ((const void*)bin != (const void*)b) {
            T bt[BS];
#pragma unroll
            for (int i = 0; i < BS; ++i) bt[i] = (T)bu[i];
            st_vec<BS>(b, u, bt);
        }
        T rt[BS], dt[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < BS; ++j) t += Dm[i * BS + j] * bu[j];
            rt[i] = (T)t;
            dt[i] = (T)(t / cp.theta);
        }
        st_vec<BS>(r, u, rt);
        st_vec<BS>(d, u, dt);
        st_vec<BS>(x, u, dt);
    }
}

// Chebyshev step k (k = 0 after pre1/post1): w = A d; r -= D^-1 w;
// d' = rho1 rho0 d + 2 rho1/delta r; x += d'. d is read for the neighbours,
// so the new d goes to dout. FINAL: z = x (FP64) and partial records of
// rFinal.z per workgroup. Launch with row_grid workgroups.
template <typename T, int BS, bool FINAL>
__device__ __forceinline__ void cheb_step_rows(int blk, int G, int n, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colIdx, const T* __restrict__ vals,
                                               const T* __restrict__ Dinv, const double* __restrict__ lam,
                                               const T* __restrict__ din, T* dout, T* r, T* x,
                                               const double* __restrict__ rFinal, double* zFinal, double* rzPart)
{
    constexpr int LANES = RowShape<BS>::lanes, UNROLL = RowShape<BS>::unroll, RPP = kBlock / LANES;
    const ChebP cp = cheb_params(*lam);
    const double rho0 = 1.0 / cp.sigma;
    const double rho1 = 1.0 / (2.0 * cp.sigma - rho0);
    const double c1 = rho1 * rho0, c2 = 2.0 * rho1 / cp.delta;
    const RowChunk rc = row_chunk<BS>(n, blk, G);
    const int sub = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    double acc[1] = {0.0};
    for (int p = 0; p < rc.passes; ++p) {
        RowPass<T, BS, UNROLL> rp;
        rp.load(rc.r0 + sub + p * UNROLL * RPP, RPP, rc.r1, rowPtr, colIdx, vals, din, lane);
        double Dm[UNROLL][BS * BS], ro[UNROLL][BS], dd[UNROLL][BS], xo[UNROLL][BS], rf[UNROLL][BS];
#pragma unroll
        for (int q = 0; q < UNROLL; ++q)
            if (lane == 0 && rp.u[q] < rc.r1) {
                ld_block<BS>(Dinv + (size_t)rp.u[q] * BS * BS, Dm[q]);
                ld_vec<BS>(r, rp.u[q], ro[q]);
                ld_vec<BS>(din, rp.u[q], dd[q]);
                ld_vec<BS>(x, rp.u[q], xo[q]);
                if (FINAL) ld_vec<BS>(rFinal, rp.u[q], rf[q]);
            }
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
            double w[BS];
            rp.mv(q, colIdx, vals, din, lane, w);
            if (lane == 0 && rp.u[q] < rc.r1) {
                double xn[BS];
                T rt[BS], dt[BS], xt[BS];
#pragma unroll
                for (int i = 0; i < BS; ++i) {
                    double dw = 0.0;
#pragma unroll
                    for (int j = 0; j < BS; ++j) dw += Dm[q][i * BS + j] * w[j];
                    const double rn = ro[q][i] - dw;
                    const double dn = c1 * dd[q][i] + c2 * rn;
                    xn[i] = xo[q][i] + dn;
                    rt[i] = (T)rn;
                    dt[i] = (T)dn;
                    xt[i] = (T)xn[i];
                }
                st_vec<BS>(r, rp.u[q], rt);
                st_vec<BS>(dout, rp.u[q], dt);
                st_vec<BS>(x, rp.u[q], xt);
                if (FINAL) {
#pragma unroll
                    for (int i = 0; i < BS; ++i) acc[0] += rf[q][i] * xn[i];
                    st_vec<BS>(zFinal, rp.u[q], xn);
                }
            }
        }
    }
    if (FINAL) {
        block_sum<1>(acc);
        if (threadIdx.x == 0) rzPart[blk] = acc[0];
    }
}

template <typename T, int BS, bool FINAL>
__global__ void __launch_bounds__(kBlock) k_cheb_step(const int* __restrict__ stop, int n, const int* __restrict__ rowPtr,
                                                      const int* __restrict__ colIdx, const T* __restrict__ vals,
                                                      const T* __restrict__ Dinv, const double* __restrict__ lam,
                                                      const T* __restrict__ din, T* dout, T* r, T* x,
                                                      const double* __restrict__ rFinal, double* zFinal,
                                                      double* rzPart)
{
    if (stop && *stop) return;
    cheb_step_rows<T, BS, FINAL>(blockIdx.x, gridDim.x, n, rowPtr, colIdx, vals, Dinv, lam, din, dout, r, x, rFinal,
                                 zFinal, rzPart);
}

// residual res = b - A x (row_grid workgroups)
template <typename T, int BS>
__device__ __forceinline__ void residual_rows(int blk, int G, int n, const int* __restrict__ rowPtr,
                                              const int* __restrict__ colIdx, const T* __restrict__ vals,
                                              const T* __restrict__ b, const T* __restrict__ x, T* res)
{
    constexpr int LANES = RowShape<BS>::lanes, UNROLL = RowShape<BS>::unroll, RPP = kBlock / LANES;
    const RowChunk rc = row_chunk<BS>(n, blk, G);
    const int sub = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    for (int p = 0; p < rc.passes; ++p) {
        RowPass<T, BS, UNROLL> rp;
        rp.load(rc.r0 + sub + p * UNROLL * RPP, RPP, rc.r1, rowPtr, colIdx, vals, x, lane);
        double bo[UNROLL][BS];
#pragma unroll
        for (int q = 0; q < UNROLL; ++q)
            if (lane == 0 && rp.u[q] < rc.r1) ld_vec<BS>(b, rp.u[q], bo[q]);
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
            double y[BS];
            rp.mv(q, colIdx, vals, x, lane, y);
            if (lane == 0 && rp.u[q] < rc.r1) {
                T o[BS];
#pragma unroll
                for (int i = 0; i < BS; ++i) o[i] = (T)(bo[q][i] - y[i]);
                st_vec<BS>(res, rp.u[q], o);
            }
        }
    }
}

template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_residual(const int* __restrict__ stop, int n, const int* __restrict__ rowPtr,
                                                     const int* __restrict__ colIdx, const T* __restrict__ vals,
                                                     const T* __restrict__ b, const T* __restrict__ x, T* res)
{
    if (stop && *stop) return;
    residual_rows<T, BS>(blockIdx.x, gridDim.x, n, rowPtr, colIdx, vals, b, x, res);
}

// post-smoothing first step from the prolongated x: r = D^-1 (b - A x); d = r/theta;
// xout = x + d (row_grid workgroups)
template <typename T, int BS>
__device__ __forceinline__ void cheb_post1_rows(int blk, int G, int n, const int* __restrict__ rowPtr,
                                                const int* __restrict__ colIdx, const T* __restrict__ vals,
                                                const T* __restrict__ Dinv, const double* __restrict__ lam,
                                                const T* __restrict__ b, const T* __restrict__ xin, T* xout, T* r,
                                                T* d)
{
    constexpr int LANES = RowShape<BS>::lanes, UNROLL = RowShape<BS>::unroll, RPP = kBlock / LANES;
    const ChebP cp = cheb_params(*lam);
    const RowChunk rc = row_chunk<BS>(n, blk, G);
    const int sub = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    for (int p = 0; p < rc.passes; ++p) {
        RowPass<T, BS, UNROLL> rp;
        rp.load(rc.r0 + sub + p * UNROLL * RPP, RPP, rc.r1, rowPtr, colIdx, vals, xin, lane);
        double Dm[UNROLL][BS * BS], bo[UNROLL][BS], xo[UNROLL][BS];
#pragma unroll
        for (int q = 0; q < UNROLL; ++q)
            if (lane == 0 && rp.u[q] < rc.r1) {
                ld_block<BS>(Dinv + (size_t)rp.u[q] * BS * BS, Dm[q]);
                ld_vec<BS>(b, rp.u[q], bo[q]);
                ld_vec<BS>(xin, rp.u[q], xo[q]);
            }
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
            double y[BS];
            rp.mv(q, colIdx, vals, xin, lane, y);
            if (lane == 0 && rp.u[q] < rc.r1) {
                double res[BS];
                T ru[BS], dv[BS], xv[BS];
#pragma unroll
                for (int i = 0; i < BS; ++i) res[i] = bo[q][i] - y[i];
#pragma unroll
                for (int i = 0; i < BS; ++i) {
                    double t = 0.0;
#pragma unroll
                    for (int j = 0; j < BS; ++j) t += Dm[q][i * BS + j] * res[j];
                    const double dvi = t / cp.theta;
                    ru[i] = (T)t;
                    dv[i] = (T)dvi;
                    xv[i] = (T)(xo[q][i] + dvi);
                }
                st_vec<BS>(r, rp.u[q], ru);
                st_vec<BS>(d, rp.u[q], dv);
                st_vec<BS>(xout, rp.u[q], xv);
            }
        }
    }
}

template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_cheb_post1(const int* __restrict__ stop, int n, const int* __restrict__ rowPtr,
                                                       const int* __restrict__ colIdx, const T* __restrict__ vals,
                                                       const T* __restrict__ Dinv, const double* __restrict__ lam,
                                                       const T* __restrict__ b, const T* __restrict__ xin,
                                                       T* xout, T* r, T* d)
{
    if (stop && *stop) return;
    cheb_post1_rows<T, BS>(blockIdx.x, gridDim.x, n, rowPtr, colIdx, vals, Dinv, lam, b, xin, xout, r, d);
}

// bc_c = sum over P blocks (u,c) of P^T res_u, LN lanes per coarse node (the launcher picks LN from the mean
// length of the restriction lists: on the small coarse levels a node gathers ~100 P blocks, and with few nodes
// the kernel is one latency chain per lane, so the lanes per node grow with the list).
// PRE1: the next level's first pre-smoothing step (k_cheb_pre1 from x = 0:
// r = D^-1 b, d = r/theta, x = d) is row-local, so it runs here on each
// restricted row instead of in a launch of its own.
template <typename T, int BS, bool PRE1, int LN>
__device__ __forceinline__ void restrict_rows(int blk, int G, int nc, const int* __restrict__ rPtr,
                                              const int* __restrict__ rP, const int* __restrict__ pRow,
                                              const T* __restrict__ P, const T* __restrict__ res, T* bc,
                                              const T* __restrict__ DinvN, const double* __restrict__ lamN, T* xN,
                                              T* rN, T* dN)
{
    const int cc = (nc + G - 1) / G;
    const int c0 = blk * cc, c1 = min(nc, c0 + cc);
    constexpr int RPP = kBlock / LN;
    const int sub = threadIdx.x / LN, lane = threadIdx.x % LN;
    for (int c = c0 + sub; c < c1; c += RPP) {
        double s[NS] = {};
        for (int q = rPtr[c] + lane; q < rPtr[c + 1]; q += LN) {
            const int j = rP[q];
            const T* p = P + (size_t)j * BS * NS;
            const T* ru = res + (size_t)pRow[j] * BS;
#pragma unroll
            for (int l = 0; l < BS; ++l)
#pragma unroll
                for (int i = 0; i < NS; ++i) s[i] += (double)p[l * NS + i] * (double)ru[l];
        }
#pragma unroll
        for (int off = LN / 2; off > 0; off >>= 1)
#pragma unroll
            for (int i = 0; i < NS; ++i) s[i] += __shfl_xor(s[i], off, LN);
        if (lane == 0) {
            T bt[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) bt[i] = (T)s[i];
            st_vec<NS>(bc, c, bt);
            if (PRE1) {
                const double theta = cheb_params(*lamN).theta;
                double Dm[NS * NS];
                ld_block<NS>(DinvN + (size_t)c * NS * NS, Dm);
                T rt[NS], dt[NS];
#pragma unroll
                for (int i = 0; i < NS; ++i) {
                    double t = 0.0;
#pragma unroll
                    for (int j = 0; j < NS; ++j) t += Dm[i * NS + j] * (double)bt[j];
                    rt[i] = (T)t;
                    dt[i] = (T)(t / theta);
                }
                st_vec<NS>(rN, c, rt);
                st_vec<NS>(dN, c, dt);
                st_vec<NS>(xN, c, dt);
            }
        }
    }
}

template <typename T, int BS, bool PRE1, int LN>
__global__ void __launch_bounds__(kBlock) k_restrict(const int* __restrict__ stop, int nc, const int* __restrict__ rPtr,
                                                     const int* __restrict__ rP, const int* __restrict__ pRow,
                                                     const T* __restrict__ P, const T* __restrict__ res, T* bc,
                                                     const T* __restrict__ DinvN, const double* __restrict__ lamN,
                                                     T* xN, T* rN, T* dN)
{
    if (stop && *stop) return;
    restrict_rows<T, BS, PRE1, LN>(blockIdx.x, gridDim.x, nc, rPtr, rP, pRow, P, res, bc, DinvN, lamN, xN, rN, dN);
}

// x_u += sum over P blocks (u,c) of P xc_c, kProlongLanes lanes per fine row (one P block each per pass, the
// partial sums combined by shuffles): the few blocks of a row are loaded together instead of one after another
constexpr int kProlongLanes = 4;
template <typename T, int BS>
__device__ __forceinline__ void prolong_rows(int blk, int n, const int* __restrict__ pPtr,
                                             const int* __restrict__ pCol, const T* __restrict__ P,
                                             const T* __restrict__ xc, T* x)
{
    constexpr int LN = kProlongLanes;
    const int t = blk * kBlock + threadIdx.x;
    const int u = t / LN, lane = t % LN;
    double s[BS] = {};
    if (u < n) {
        for (int j = pPtr[u] + lane; j < pPtr[u + 1]; j += LN) {
            const T* p = P + (size_t)j * BS * NS;
            const T* xcv = xc + (size_t)pCol[j] * NS;
#pragma unroll
            for (int i = 0; i < BS; ++i)
#pragma unroll
                for (int c = 0; c < NS; ++c) s[i] += (double)p[i * NS + c] * (double)xcv[c];
        }
    }
#pragma unroll
    for (int off = LN / 2; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < BS; ++i) s[i] += __shfl_xor(s[i], off, LN);
    if (u < n && lane == 0) {
#pragma unroll
        for (int i = 0; i < BS; ++i) x[(size_t)u * BS + i] = (T)((double)x[(size_t)u * BS + i] + s[i]);
    }
}

template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_prolong(const int* __restrict__ stop, int n, const int* __restrict__ pPtr, const int* __restrict__ pCol,
                                                    const T* __restrict__ P, const T* __restrict__ xc, T* x)
{
    if (stop && *stop) return;
    prolong_rows<T, BS>(blockIdx.x, n, pPtr, pCol, P, xc, x);
}

template <typename T>
__device__ __forceinline__ void dense_apply_rows(int N, const double* __restrict__ Inv, const T* __restrict__ b, T* x)
{
    __shared__ double bs[kCoarseMaxDofs];
    for (int j = threadIdx.x; j < N; j += blockDim.x) bs[j] = (double)b[j];
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < N; ++j) s += Inv[(size_t)j * N + i] * bs[j];  // symmetric: column i = row i
        x[i] = (T)s;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_dense_apply(const int* __restrict__ stop, int N,
                                                     const double* __restrict__ Inv, const T* __restrict__ b, T* x)
{
    if (stop && *stop) return;
    dense_apply_rows<T>(N, Inv, b, x);
}

// z = x (FP64) and partial records of r.z over the same row chunks as the level-0 step kernel
template <typename T>
__global__ void __launch_bounds__(kBlock) k_finish_z(const int* __restrict__ stop, int n2, const T* __restrict__ x, const double* __restrict__ r,
                                                     double* z, double* rzPart)
{
    if (stop && *stop) return;
    int r0, r1;
    chunk_rows(n2, r0, r1);
    double acc[1] = {0.0};
    for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        const double zi = (double)x[i];
        z[i] = zi;
        acc[0] += r[i] * zi;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) rzPart[blockIdx.x] = acc[0];
}

template <typename P, typename T>
void dalloc(P& pool, T*& p, size_t n)
{
    p = static_cast<T*>(pool.get(sizeof(T) * std::max<size_t>(n, 1)));
}

template <typename P, typename V>
void up(P& pool, hipStream_t s, int*& d, const std::vector<V>& h)
{
    dalloc(pool, d, h.size());
    if (d && !h.empty()) (void)hipMemcpyAsync(d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice, s);
}

}  // namespace

// ============================================================================ host structure
template <typename T>
struct Amg<T>::Level : AmgHostLevel {
    // device
    int *dRowPtr = nullptr, *dColIdx = nullptr, *dDiagPos = nullptr;
    int *dAggPtr = nullptr, *dAggMem = nullptr, *dPPtr = nullptr, *dPCol = nullptr, *dPRow = nullptr,
        *dPsPtr = nullptr, *dPsA = nullptr, *dRPtr = nullptr, *dRP = nullptr, *dTcPtr = nullptr, *dTcA = nullptr,
        *dTcP = nullptr, *dAcPtr = nullptr, *dAcP = nullptr, *dAcT = nullptr;
    unsigned char* dPIsTent = nullptr;
    T *vals = nullptr, *Dinv = nullptr, *Pt = nullptr, *P = nullptr, *Tv = nullptr;
    T *b = nullptr, *x = nullptr, *x2 = nullptr, *r = nullptr, *d = nullptr, *d2 = nullptr;
    double *Bnext = nullptr, *lz = nullptr, *lam = nullptr;  // lz: 6 Lanczos vectors, then alpha/beta/count
    // hand every buffer back to the owner's pool
    void give_back(Pool& pool)
    {
        for (void* p : {(void*)dRowPtr, (void*)dColIdx, (void*)dDiagPos, (void*)dAggPtr, (void*)dAggMem, (void*)dPPtr,
                        (void*)dPCol, (void*)dPRow, (void*)dPsPtr, (void*)dPsA, (void*)dRPtr, (void*)dRP,
                        (void*)dTcPtr, (void*)dTcA, (void*)dTcP, (void*)dAcPtr, (void*)dAcP, (void*)dAcT,
                        (void*)dPIsTent, (void*)vals, (void*)Dinv, (void*)Pt, (void*)P, (void*)Tv, (void*)b,
                        (void*)x, (void*)x2, (void*)r, (void*)d, (void*)d2, (void*)Bnext, (void*)lz, (void*)lam})
            if (p) pool.put(p);
    }
};

template <typename T>
void* Amg<T>::Pool::get(size_t bytes)
{
    // smallest free buffer that fits and wastes at most half of itself
    auto it = free.lower_bound(bytes);
    if (it != free.end() && it->first <= 2 * bytes) {
        void* p = it->second;
        free.erase(it);
        return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    size[p] = bytes;
    return p;
}

template <typename T>
void Amg<T>::Pool::put(void* p)
{
    free.emplace(size.at(p), p);
}

template <typename T>
void Amg<T>::Pool::trim()
{
    for (auto& kv : free) {
        (void)hipFree(kv.second);
        size.erase(kv.second);
    }
    free.clear();
}

template <typename T>
void Amg<T>::Pool::clear()
{
    for (auto& kv : size) (void)hipFree(kv.first);
    free.clear();
    size.clear();
}

template <typename T>
Amg<T>::~Amg()
{
    release();
    pool_.clear();
}

template <typename T>
void Amg<T>::release()
{
    for (Level* l : L) {
        l->give_back(pool_);
        delete l;
    }
    L.clear();
    if (dCoarseInv) pool_.put(dCoarseInv);
    if (dPartials) pool_.put(dPartials);
    dCoarseInv = nullptr;
    dPartials = nullptr;
    ready_ = false;
}

template <typename T>
int Amg<T>::rz_records() const
{
    return row_grid(L.empty() ? 1 : L[0]->n, 2);
}

template <typename T>
long Amg<T>::level_rows(int l) const
{
    return (long)L[l]->n * L[l]->bs;
}

template <typename T>
void Amg<T>::lanczos0_prepare(hipStream_t s)
{
    Level& A = *L[0];
    const int g = row_grid(A.n, 2);
    double* part[2] = {dPartials, dPartials + nPartials};
    const size_t nv = (size_t)A.n * 2;
    double* coef = A.lz + 6 * nv;
    LanczosBufs lb{{A.lz, A.lz + nv, A.lz + 2 * nv}, {A.lz + 3 * nv, A.lz + 4 * nv, A.lz + 5 * nv}};
    hipLaunchKernelGGL((k_lanczos_start<T, 2>), dim3(g), dim3(kBlock), 0, s, A.n, A.dRowPtr, A.vals, A.Dinv, lb.W[0],
                       part[1], dPartials + 2 * (size_t)nPartials, coef, kLanczosSteps);
    for (int j = 0; j < 2; ++j)
        hipLaunchKernelGGL((k_lanczos<T, 2>), dim3(g), dim3(kBlock), 0, s, j, kLanczosSteps, A.n, A.dRowPtr,
                           A.dColIdx, A.vals, A.dDiagPos, lb, part[(j + 1) & 1], g, part[j & 1], coef);
}

template <typename T>
void Amg<T>::lanczos0_step(hipStream_t s)
{
    // step 2 reads part[1] (written by step 1), W[2], W[1], v[1], v[0] and
    // writes part[0], v[2], W[0]: repeatable
    Level& A = *L[0];
    const int g = row_grid(A.n, 2);
    const size_t nv = (size_t)A.n * 2;
    LanczosBufs lb{{A.lz, A.lz + nv, A.lz + 2 * nv}, {A.lz + 3 * nv, A.lz + 4 * nv, A.lz + 5 * nv}};
    hipLaunchKernelGGL((k_lanczos<T, 2>), dim3(g), dim3(kBlock), 0, s, 2, kLanczosSteps, A.n, A.dRowPtr, A.dColIdx,
                       A.vals, A.dDiagPos, lb, dPartials + nPartials, g, dPartials, A.lz + 6 * nv);
}

template <typename T>
void Amg<T>::cheb_step_level(hipStream_t s, int l)
{
    Level& A = *L[l];
    if (A.bs == 2)
        hipLaunchKernelGGL((k_cheb_step<T, 2, false>), dim3(row_grid(A.n, 2)), dim3(kBlock), 0, s, nullptr, A.n,
                           A.dRowPtr, A.dColIdx, A.vals, A.Dinv, A.lam, A.d, A.d2, A.r, A.x, nullptr, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_cheb_step<T, 3, false>), dim3(row_grid(A.n, 3)), dim3(kBlock), 0, s, nullptr, A.n,
                           A.dRowPtr, A.dColIdx, A.vals, A.Dinv, A.lam, A.d, A.d2, A.r, A.x, nullptr, nullptr, nullptr);
}

template <typename T>
int Amg<T>::level_block_size(int l) const
{
    return L[l]->bs;
}

template <typename T>
long Amg<T>::level_blocks(int l) const
{
    return (long)L[l]->colIdx.size();
}

template <typename T>
int Amg<T>::setup(hipStream_t s, int nV, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                  const std::vector<unsigned char>& fixedMask, std::string& err,
                  const std::vector<unsigned char>* late)
{
    ++generation_;
    tailMax_ = tailPending_;
    // The previous hierarchy's rows are reused where they come out the same
    // (an air-mesh change leaves most of the mesh rows as they were).
    std::vector<AmgHostLevel> H = std::move(spare_);
    spare_.clear();
    AmgPrevious prev;
    for (Level* l : L) prev.levels.push_back(l);
    const int st =
        amg_build_structure(nV, rowPtr, colIdx, fixedMask, kCoarseMaxDofs, kMaxLevels, H, err, &prev, late);
    for (Level* l : L) spare_.push_back(std::move(static_cast<AmgHostLevel&>(*l)));
    release();
    if (st != 0) return -1;
    for (auto& h : H) {
        Level* lv = new Level();
        static_cast<AmgHostLevel&>(*lv) = std::move(h);
        L.push_back(lv);
    }
    // ---- device storage
    for (size_t l = 0; l < L.size(); ++l) {
        Level& A = *L[l];
        const int bs = A.bs;
        up(pool_, s, A.dRowPtr, A.rowPtr);
        up(pool_, s, A.dColIdx, A.colIdx);
        up(pool_, s, A.dDiagPos, A.diagPos);
        dalloc(pool_, A.vals, A.colIdx.size() * bs * bs);
        dalloc(pool_, A.Dinv, (size_t)A.n * bs * bs);
        dalloc(pool_, A.b, (size_t)A.n * bs);
        dalloc(pool_, A.x, (size_t)A.n * bs);
        dalloc(pool_, A.x2, (size_t)A.n * bs);
        dalloc(pool_, A.r, (size_t)A.n * bs);
        dalloc(pool_, A.d, (size_t)A.n * bs);
        dalloc(pool_, A.d2, (size_t)A.n * bs);
        dalloc(pool_, A.lam, 2);
        dalloc(pool_, A.lz, (size_t)6 * A.n * bs + 2 * kLanczosSteps + 1);
        if (l + 1 < L.size()) {
            up(pool_, s, A.dAggPtr, A.aggPtr);
            up(pool_, s, A.dAggMem, A.aggMem);
            up(pool_, s, A.dPPtr, A.pPtr);
            up(pool_, s, A.dPCol, A.pCol);
            up(pool_, s, A.dPRow, A.pRow);
            up(pool_, s, A.dPsPtr, A.psPtr);
            up(pool_, s, A.dPsA, A.psA);
            up(pool_, s, A.dRPtr, A.rPtr);
            up(pool_, s, A.dRP, A.rP);
            up(pool_, s, A.dTcPtr, A.tcPtr);
            up(pool_, s, A.dTcA, A.tcA);
            up(pool_, s, A.dTcP, A.tcP);
            up(pool_, s, A.dAcPtr, A.acPtr);
            up(pool_, s, A.dAcP, A.acP);
            up(pool_, s, A.dAcT, A.acT);
            dalloc(pool_, A.dPIsTent, A.pIsTent.size());
            if (A.dPIsTent && !A.pIsTent.empty())
                (void)hipMemcpyAsync(A.dPIsTent, A.pIsTent.data(), A.pIsTent.size(), hipMemcpyHostToDevice, s);
            dalloc(pool_, A.Pt, (size_t)A.n * bs * NS);
            (void)hipMemsetAsync(A.Pt, 0, sizeof(T) * (size_t)A.n * bs * NS, s);
            dalloc(pool_, A.P, A.pCol.size() * bs * NS);
            dalloc(pool_, A.Tv, A.tCol.size() * bs * NS);
            dalloc(pool_, A.Bnext, (size_t)A.nc * NS * NS);
        }
    }
    const Level& C = *L.back();
    nCoarse = C.n * C.bs;
    coarseDiag = true;
    for (int u = 0; u < C.n; ++u) coarseDiag = coarseDiag && C.rowPtr[u + 1] - C.rowPtr[u] == 1;
    if (nCoarse > kCoarseMaxDofs && !coarseDiag) {
        err = "AMG: coarsening stalled above the dense coarsest size";
        release();
        return -1;
    }
    if (!coarseDiag) dalloc(pool_, dCoarseInv, (size_t)nCoarse * nCoarse);
    nPartials = 2 * kAmgMaxGrid;  // two doubles per workgroup record
    dalloc(pool_, dPartials, (size_t)nPartials * 2 + kAmgMaxGrid);  // two record buffers + Gershgorin records
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
        err = "AMG: device allocation or upload failed";
        release();
        return -1;
    }
    pool_.trim();  // buffers of the previous hierarchy this one did not reuse
    ready_ = true;
    return 0;
}

template <typename T>
void Amg<T>::numeric(hipStream_t s, const void* A0vals, int vals32, const double* UV, int safe)
{
    for (size_t l = 0; l < L.size(); ++l) {
        Level& A = *L[l];
        const int bs = A.bs;
        const long nvals = (long)A.colIdx.size() * bs * bs;
        if (l == 0) {
            if (vals32)
                hipLaunchKernelGGL((k_cvt<float, T>), dim3(grid_for(nvals)), dim3(kBlock), 0, s, nvals,
                                   static_cast<const float*>(A0vals), A.vals);
            else
                hipLaunchKernelGGL((k_cvt<double, T>), dim3(grid_for(nvals)), dim3(kBlock), 0, s, nvals,
                                   static_cast<const double*>(A0vals), A.vals);
        }
        if (std::is_same<T, float>::value) {  // keep the FP32-stored level positive definite
            constexpr double kShift = 4.0 / 16777216.0;
            if (bs == 2)
                hipLaunchKernelGGL((k_shift_diag<T, 2>), dim3(grid_for(A.n)), dim3(kBlock), 0, s, A.n, A.dRowPtr,
                                   A.dDiagPos, A.vals, kShift);
            else
                hipLaunchKernelGGL((k_shift_diag<T, 3>), dim3(grid_for(A.n)), dim3(kBlock), 0, s, A.n, A.dRowPtr,
                                   A.dDiagPos, A.vals, kShift);
        }
        if (bs == 2)
            hipLaunchKernelGGL((k_blockinv<T, 2>), dim3(grid_for(A.n)), dim3(kBlock), 0, s, A.n, A.dDiagPos, A.vals,
                               A.Dinv);
        else
            hipLaunchKernelGGL((k_blockinv<T, 3>), dim3(grid_for(A.n)), dim3(kBlock), 0, s, A.n, A.dDiagPos, A.vals,
                               A.Dinv);
        if (l + 1 == L.size()) break;
        // lambda_max(D^-1 A) by Lanczos
        const int g = row_grid(A.n, bs);
        double* part[2] = {dPartials, dPartials + nPartials};
        const size_t nv = (size_t)A.n * bs;
        double* coef = A.lz + 6 * nv;
        LanczosBufs lb{{A.lz, A.lz + nv, A.lz + 2 * nv}, {A.lz + 3 * nv, A.lz + 4 * nv, A.lz + 5 * nv}};
        double* gPart = dPartials + 2 * (size_t)nPartials;
        const int nsteps = lanczos_steps(l);
        const int lsafe = safe || nsteps == 0;
        if (bs == 2)
            hipLaunchKernelGGL((k_lanczos_start<T, 2>), dim3(g), dim3(kBlock), 0, s, A.n, A.dRowPtr, A.vals, A.Dinv,
                               lb.W[0], part[1], gPart, coef, nsteps);
        else
            hipLaunchKernelGGL((k_lanczos_start<T, 3>), dim3(g), dim3(kBlock), 0, s, A.n, A.dRowPtr, A.vals, A.Dinv,
                               lb.W[0], part[1], gPart, coef, nsteps);
        for (int j = 0; j < (lsafe ? 0 : nsteps); ++j) {
            if (bs == 2)
                hipLaunchKernelGGL((k_lanczos<T, 2>), dim3(g), dim3(kBlock), 0, s, j, nsteps, A.n, A.dRowPtr,
                                   A.dColIdx, A.vals, A.dDiagPos, lb, part[(j + 1) & 1], g, part[j & 1], coef);
            else
                hipLaunchKernelGGL((k_lanczos<T, 3>), dim3(g), dim3(kBlock), 0, s, j, nsteps, A.n, A.dRowPtr,
                                   A.dColIdx, A.vals, A.dDiagPos, lb, part[(j + 1) & 1], g, part[j & 1], coef);
        }
        hipLaunchKernelGGL(k_lanczos_finish, dim3(1), dim3(kBlock), 0, s, part[(nsteps + 1) & 1], g, gPart, lsafe,
                           nsteps, coef, A.lam);
        // tentative prolongator, smoothing, Galerkin product
        Level& N = *L[l + 1];
        const int nP = (int)A.pCol.size(), nT = (int)A.tCol.size(), nAc = (int)N.colIdx.size();
        const double* Bin = l == 0 ? nullptr : L[l - 1]->Bnext;
        if (bs == 2) {
            hipLaunchKernelGGL((k_tentative<T, 2>), dim3(setup_grid(A.nc, kSetupLanes)), dim3(kBlock), 0, s, A.nc, A.dAggPtr,
                               A.dAggMem, reinterpret_cast<const double2*>(UV), Bin, A.Pt, A.Bnext);
            hipLaunchKernelGGL((k_smooth_p<T, 2>), dim3(setup_grid(nP, kSetupLanes)), dim3(kBlock), 0, s, nP, A.dPRow, A.dPIsTent,
                               A.dPsPtr, A.dPsA, A.dColIdx, A.vals, A.Dinv, A.Pt, A.lam, A.P);
            hipLaunchKernelGGL((k_ap<T, 2>), dim3(setup_grid(nT, kSetupLanes)), dim3(kBlock), 0, s, nT, A.dTcPtr, A.dTcA, A.dTcP,
                               A.vals, A.P, A.Tv);
            hipLaunchKernelGGL((k_ptap<T, 2>), dim3(setup_grid(nAc, kSetupLanes)), dim3(kBlock), 0, s, nAc, A.dAcPtr, A.dAcP,
                               A.dAcT, A.P, A.Tv, N.vals);
        } else {
            hipLaunchKernelGGL((k_tentative<T, 3>), dim3(setup_grid(A.nc, kSetupLanes)), dim3(kBlock), 0, s, A.nc, A.dAggPtr,
                               A.dAggMem, reinterpret_cast<const double2*>(UV), Bin, A.Pt, A.Bnext);
            hipLaunchKernelGGL((k_smooth_p<T, 3>), dim3(setup_grid(nP, kSetupLanes)), dim3(kBlock), 0, s, nP, A.dPRow, A.dPIsTent,
                               A.dPsPtr, A.dPsA, A.dColIdx, A.vals, A.Dinv, A.Pt, A.lam, A.P);
            hipLaunchKernelGGL((k_ap<T, 3>), dim3(setup_grid(nT, kSetupLanes)), dim3(kBlock), 0, s, nT, A.dTcPtr, A.dTcA, A.dTcP,
                               A.vals, A.P, A.Tv);
            hipLaunchKernelGGL((k_ptap<T, 3>), dim3(setup_grid(nAc, kSetupLanes)), dim3(kBlock), 0, s, nAc, A.dAcPtr, A.dAcP,
                               A.dAcT, A.P, A.Tv, N.vals);
        }
    }
    const Level& C = *L.back();
    if (!coarseDiag)
        hipLaunchKernelGGL(k_dense_inverse<T>, dim3(1), dim3(256), 0, s, C.n, C.bs, C.dRowPtr, C.dColIdx, C.vals,
                           dCoarseInv);
}

// x = D^-1 b on a block-diagonal coarsest level (exact)
template <typename T, int BS>
__device__ __forceinline__ void coarse_diag_rows(int blk, int n, const T* __restrict__ Dinv, const T* __restrict__ b,
                                                 T* x)
{
    const int u = blk * kBlock + threadIdx.x;
    if (u >= n) return;
    const T* D = Dinv + (size_t)u * BS * BS;
#pragma unroll
    for (int i = 0; i < BS; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < BS; ++j) acc += (double)D[i * BS + j] * (double)b[(size_t)u * BS + j];
        x[(size_t)u * BS + i] = (T)acc;
    }
}

template <typename T, int BS>
__global__ void __launch_bounds__(kBlock) k_coarse_diag(const int* __restrict__ stop, int n, const T* __restrict__ Dinv,
                                                        const T* __restrict__ b, T* x)
{
    if (stop && *stop) return;
    coarse_diag_rows<T, BS>(blockIdx.x, n, Dinv, b, x);
}

// The last smoothed level's part of the V-cycle, with 3x3 blocks, in one
// workgroup: the level's remaining pre-smoothing step, its residual, the
// restriction, the coarsest solve (dense inverse, or block diagonal with
// 2x2 or 3x3 blocks: cbs), the prolongation and both post-smoothing steps.
// The phases are the bodies of the kernels the V-cycle launches on larger
// levels, run over every row by this one workgroup (each row's arithmetic,
// lane groups and shuffle order included, is the same as in those launches,
// so the result is bit-identical), with barriers where the kernel boundaries
// were. On a level of a few dozen nodes each of those launches is one chain of
// dependent loads; here the chains follow each other inside one launch.
#ifndef OC_TAIL_MAX_NODES
#define OC_TAIL_MAX_NODES 128
#endif
constexpr int kTailMaxNodes = OC_TAIL_MAX_NODES;  // largest last smoothed level run as one workgroup (0: never)
struct TailArgs {
    // smoothed level (3x3 blocks)
    int n;
    const int *rowPtr, *colIdx, *rPtr, *rP, *pRow, *pPtr, *pCol;
    const void *vals, *Dinv, *P;
    const double* lam;
    void *b, *x, *x2, *r, *d, *d2;
    // coarsest level
    int nc, cbs, dense;
    const void* DinvC;
    const double* Inv;
    void *bC, *xC;
};

template <typename T, int LN>
__global__ void __launch_bounds__(kBlock) k_vcycle_tail(const int* __restrict__ stop, TailArgs a)
{
    if (stop && *stop) return;
    const T* vals = static_cast<const T*>(a.vals);
    const T* Dinv = static_cast<const T*>(a.Dinv);
    const T* P = static_cast<const T*>(a.P);
    T *b = static_cast<T*>(a.b), *x = static_cast<T*>(a.x), *x2 = static_cast<T*>(a.x2);
    T *r = static_cast<T*>(a.r), *d = static_cast<T*>(a.d), *d2 = static_cast<T*>(a.d2);
    T *bC = static_cast<T*>(a.bC), *xC = static_cast<T*>(a.xC);
    cheb_step_rows<T, 3, false>(0, 1, a.n, a.rowPtr, a.colIdx, vals, Dinv, a.lam, d, d2, r, x, nullptr, nullptr,
                                nullptr);
    __syncthreads();
    residual_rows<T, 3>(0, 1, a.n, a.rowPtr, a.colIdx, vals, b, x, r);
    __syncthreads();
    restrict_rows<T, 3, false, LN>(0, 1, a.nc, a.rPtr, a.rP, a.pRow, P, r, bC, nullptr, nullptr, nullptr, nullptr,
                                   nullptr);
    __syncthreads();
    if (a.dense) {
        dense_apply_rows<T>(a.nc * a.cbs, a.Inv, bC, xC);
    } else {
        const T* DinvC = static_cast<const T*>(a.DinvC);
        for (int blk = 0; blk * kBlock < a.nc; ++blk) {
            if (a.cbs == 2) coarse_diag_rows<T, 2>(blk, a.nc, DinvC, bC, xC);
            else coarse_diag_rows<T, 3>(blk, a.nc, DinvC, bC, xC);
        }
    }
    __syncthreads();
    for (int blk = 0; blk * kBlock < a.n * kProlongLanes; ++blk) prolong_rows<T, 3>(blk, a.n, a.pPtr, a.pCol, P, xC, x);
    __syncthreads();
    cheb_post1_rows<T, 3>(0, 1, a.n, a.rowPtr, a.colIdx, vals, Dinv, a.lam, b, x, x2, r, d);
    __syncthreads();
    cheb_step_rows<T, 3, false>(0, 1, a.n, a.rowPtr, a.colIdx, vals, Dinv, a.lam, d, d2, r, x2, nullptr, nullptr,
                                nullptr);
}

template <typename T>
void Amg<T>::coarse_solve(hipStream_t s, const int* stop)
{
    Level& C = *L.back();
    if (!coarseDiag)
        hipLaunchKernelGGL(k_dense_apply<T>, dim3(1), dim3(256), 0, s, stop, nCoarse, dCoarseInv, C.b, C.x);
    else if (C.bs == 2)
        hipLaunchKernelGGL((k_coarse_diag<T, 2>), dim3(grid_for(C.n)), dim3(kBlock), 0, s, stop, C.n, C.Dinv, C.b, C.x);
    else
        hipLaunchKernelGGL((k_coarse_diag<T, 3>), dim3(grid_for(C.n)), dim3(kBlock), 0, s, stop, C.n, C.Dinv, C.b, C.x);
}

// Restriction of level l's residual to level l + 1's right-hand side, with
// level l + 1's first pre-smoothing step fused in when that level is smoothed
// (not the coarsest).
template <typename T>
void Amg<T>::restrict_to(hipStream_t s, int l, const int* stop)
{
    Level& A = *L[l];
    Level& N = *L[l + 1];
    const bool fuse = l + 2 < (int)L.size();
    // lanes per coarse node: the power of two nearest the mean restriction-list length, 16 to 64
    const long meanLen = A.nc > 0 ? (long)A.rP.size() / A.nc : 0;
    const int ln = meanLen >= 48 ? 64 : (meanLen >= 24 ? 32 : 16);
    const dim3 g(rows_grid(A.nc, kBlock / ln));
    const auto launch = [&](auto kern) {
        if (fuse)
            hipLaunchKernelGGL(kern, g, dim3(kBlock), 0, s, stop, A.nc, A.dRPtr, A.dRP, A.dPRow, A.P, A.r, N.b, N.Dinv,
                               N.lam, N.x, N.r, N.d);
        else
            hipLaunchKernelGGL(kern, g, dim3(kBlock), 0, s, stop, A.nc, A.dRPtr, A.dRP, A.dPRow, A.P, A.r, N.b,
                               nullptr, nullptr, nullptr, nullptr, nullptr);
    };
#define OC_RESTRICT(BS_, F_)                                                      \
    (ln == 64 ? launch(k_restrict<T, BS_, F_, 64>)                                \
              : (ln == 32 ? launch(k_restrict<T, BS_, F_, 32>) : launch(k_restrict<T, BS_, F_, 16>)))
    if (A.bs == 2) {
        if (fuse) OC_RESTRICT(2, true);
        else OC_RESTRICT(2, false);
    } else {
        if (fuse) OC_RESTRICT(3, true);
        else OC_RESTRICT(3, false);
    }
#undef OC_RESTRICT
}

template <typename T>
void Amg<T>::vcycle_tail(hipStream_t s, int l, const int* stop)
{
    Level& A = *L[l];
    Level& C = *L[l + 1];
    TailArgs a;
    a.n = A.n;
    a.rowPtr = A.dRowPtr;
    a.colIdx = A.dColIdx;
    a.rPtr = A.dRPtr;
    a.rP = A.dRP;
    a.pRow = A.dPRow;
    a.pPtr = A.dPPtr;
    a.pCol = A.dPCol;
    a.vals = A.vals;
    a.Dinv = A.Dinv;
    a.P = A.P;
    a.lam = A.lam;
    a.b = A.b;
    a.x = A.x;
    a.x2 = A.x2;
    a.r = A.r;
    a.d = A.d;
    a.d2 = A.d2;
    a.nc = C.n;
    a.cbs = C.bs;
    a.dense = coarseDiag ? 0 : 1;
    a.DinvC = C.Dinv;
    a.Inv = dCoarseInv;
    a.bC = C.b;
    a.xC = C.x;
    // lanes per coarse node as restrict_to picks them
    const long meanLen = A.nc > 0 ? (long)A.rP.size() / A.nc : 0;
    const int ln = meanLen >= 48 ? 64 : (meanLen >= 24 ? 32 : 16);
    if (ln == 64) hipLaunchKernelGGL((k_vcycle_tail<T, 64>), dim3(1), dim3(kBlock), 0, s, stop, a);
    else if (ln == 32) hipLaunchKernelGGL((k_vcycle_tail<T, 32>), dim3(1), dim3(kBlock), 0, s, stop, a);
    else hipLaunchKernelGGL((k_vcycle_tail<T, 16>), dim3(1), dim3(kBlock), 0, s, stop, a);
}

template <typename T>
void Amg<T>::vcycle(hipStream_t s, int l, const int* stop)
{
    Level& A = *L[l];
    if (l + 1 == (int)L.size()) {
        coarse_solve(s, stop);
        return;
    }
    Level& N = *L[l + 1];
    const int gUp = rows_grid(A.n, kBlock);
    if (A.bs == 2) {
        const int g = row_grid(A.n, 2);
        (void)gUp;  // level 0 starts from k_cheb_pre1 in apply(); lower levels from the fused restriction
        hipLaunchKernelGGL((k_cheb_step<T, 2, false>), dim3(g), dim3(kBlock), 0, s, stop, A.n, A.dRowPtr, A.dColIdx,
                           A.vals, A.Dinv, A.lam, A.d, A.d2, A.r, A.x, nullptr, nullptr, nullptr);
        hipLaunchKernelGGL((k_residual<T, 2>), dim3(g), dim3(kBlock), 0, s, stop, A.n, A.dRowPtr, A.dColIdx, A.vals,
                           A.b, A.x, A.r);
        restrict_to(s, l, stop);
        vcycle(s, l + 1, stop);
        hipLaunchKernelGGL((k_prolong<T, 2>), dim3(grid_for((long)A.n * kProlongLanes)), dim3(kBlock), 0, s, stop, A.n, A.dPPtr, A.dPCol, A.P,
                           N.x, A.x);
        hipLaunchKernelGGL((k_cheb_post1<T, 2>), dim3(g), dim3(kBlock), 0, s, stop, A.n, A.dRowPtr, A.dColIdx, A.vals,
                           A.Dinv, A.lam, A.b, A.x, A.x2, A.r, A.d);
    } else {
        const int g = row_grid(A.n, 3);
        if (l == 0)  // levels below the finest start from the restriction's fused first step
            hipLaunchKernelGGL((k_cheb_pre1<T, 3, T>), dim3(gUp), dim3(kBlock), 0, s, stop, A.n, A.b, A.b, A.Dinv, A.lam,
                               A.x, A.r, A.d);
        if (l + 2 == (int)L.size() && A.n <= (tailMax_ < 0 ? kTailMaxNodes : tailMax_)) {
            vcycle_tail(s, l, stop);
            std::swap(A.x, A.x2);
            return;
        }
        hipLaunchKernelGGL((k_cheb_step<T, 3, false>), dim3(g), dim3(kBlock), 0, s, stop, A.n, A.dRowPtr, A.dColIdx,
                           A.vals, A.Dinv, A.lam, A.d, A.d2, A.r, A.x, nullptr, nullptr, nullptr);
        hipLaunchKernelGGL((k_residual<T, 3>), dim3(g), dim3(kBlock), 0, s, stop, A.n, A.dRowPtr, A.dColIdx, A.vals,
                           A.b, A.x, A.r);
        restrict_to(s, l, stop);
        vcycle(s, l + 1, stop);
        hipLaunchKernelGGL((k_prolong<T, 3>), dim3(grid_for((long)A.n * kProlongLanes)), dim3(kBlock), 0, s, stop, A.n, A.dPPtr, A.dPCol, A.P,
                           N.x, A.x);
        hipLaunchKernelGGL((k_cheb_post1<T, 3>), dim3(g), dim3(kBlock), 0, s, stop, A.n, A.dRowPtr, A.dColIdx, A.vals,
                           A.Dinv, A.lam, A.b, A.x, A.x2, A.r, A.d);
        hipLaunchKernelGGL((k_cheb_step<T, 3, false>), dim3(g), dim3(kBlock), 0, s, stop, A.n, A.dRowPtr, A.dColIdx,
                           A.vals, A.Dinv, A.lam, A.d, A.d2, A.r, A.x2, nullptr, nullptr, nullptr);
        std::swap(A.x, A.x2);
    }
}

template <typename T>
void Amg<T>::apply(hipStream_t s, const double* r, double* z, double* rzPart, const int* stop)
{
    Level& A = *L[0];
    const int g = row_grid(A.n, 2);
    if (L.size() == 1) {
        // single level: the dense inverse is the whole preconditioner
        hipLaunchKernelGGL((k_cvt<double, T>), dim3(grid_for((long)A.n * 2)), dim3(kBlock), 0, s, (long)A.n * 2, r, A.b);
        coarse_solve(s, stop);
        hipLaunchKernelGGL(k_finish_z<T>, dim3(g), dim3(kBlock), 0, s, stop, A.n * 2, A.x, r, z, rzPart);
        return;
    }
    // level 0 pre-smoothing reads the FP64 residual and keeps a T copy in A.b
    hipLaunchKernelGGL((k_cheb_pre1<T, 2, double>), dim3(rows_grid(A.n, kBlock)), dim3(kBlock), 0, s, stop, A.n, r, A.b,
                       A.Dinv, A.lam, A.x, A.r, A.d);
    vcycle(s, 0, stop);
    // final post-smoothing step writes z (FP64) and the r.z partial records
    hipLaunchKernelGGL((k_cheb_step<T, 2, true>), dim3(g), dim3(kBlock), 0, s, stop, A.n, A.dRowPtr, A.dColIdx, A.vals,
                       A.Dinv, A.lam, A.d, A.d2, A.r, A.x2, r, z, rzPart);
}

template class Amg<float>;
template class Amg<double>;

}  // namespace oc
