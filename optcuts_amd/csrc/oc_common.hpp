// Shared device helpers for the gfx950 geometry-step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oc {

constexpr int kWave = 64;     // CDNA wavefront width
constexpr int kBlock = 256;   // threads per workgroup for 1-D kernels (4 waves)

// ---------------------------------------------------------------------------
// Exact-rounding helpers: a*b and a+b each rounded once, never contracted
// into an FMA, so expressions that the reference evaluates as two IEEE
// operations round the same way here.
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }

// ---------------------------------------------------------------------------
// Compensated (TwoSum) accumulation: value + running error, combined
// pairwise in a fixed tree order so every reduction is deterministic.
struct CSum {
    double s, c;
};

__device__ __forceinline__ CSum csum_add(CSum a, CSum b)
{
    const double s = add_rn(a.s, b.s);
    const double bp = __dsub_rn(s, a.s);
    const double err = add_rn(__dsub_rn(a.s, __dsub_rn(s, bp)), __dsub_rn(b.s, bp));
    return CSum{s, a.c + b.c + err};
}

__device__ __forceinline__ CSum wave_csum(CSum v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        CSum o{__shfl_xor(v.s, off), __shfl_xor(v.c, off)};
        v = csum_add(v, o);
    }
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ float wave_sumf(float v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    return v;
}

// Block reduction of NV independent compensated sums; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_csum(CSum (&v)[NV])
{
    __shared__ CSum sh[kBlock / kWave][NV];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_csum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            CSum a = sh[0][k];
            for (int w = 1; w < (int)(blockDim.x / kWave); ++w) a = csum_add(a, sh[w][k]);
            v[k] = a;
        }
    }
    __syncthreads();
}

// Block reduction of NV plain sums; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double sh[kBlock / kWave][NV];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double a = sh[0][k];
            for (int w = 1; w < (int)(blockDim.x / kWave); ++w) a += sh[w][k];
            v[k] = a;
        }
    }
    __syncthreads();
}

// Order-preserving map of a double to uint64 (for atomicMin over doubles of
// either sign); NaN inputs are filtered by the caller.
__device__ __forceinline__ unsigned long long dbl_to_ordered(double x)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double ordered_to_dbl(unsigned long long u)
{
    u = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
    return __longlong_as_double((long long)u);
}

inline double ordered_to_dbl_host(unsigned long long u)
{
    u = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
    double d;
    __builtin_memcpy(&d, &u, sizeof(d));
    return d;
}

inline unsigned long long dbl_to_ordered_host(double x)
{
    unsigned long long u;
    __builtin_memcpy(&u, &x, sizeof(u));
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

inline int grid_for(long n, int block = kBlock)
{
    long g = (n + block - 1) / block;
    return (int)(g < 1 ? 1 : g);
}

}  // namespace oc
