// Learned preconditioner on the GPU (see oc_gnn.hpp for the model): feature
// extraction, forward pass, factor application inside PCG, the probe loss,
// backpropagation and Adam, plus the oc_gnn_* C ABI.
//
// Matrix-core products use v_mfma_f32_32x32x16_bf16 on 32-row tiles, one
// tile per wavefront: lane l feeds row (l % 32) of the A operand and column
// (l % 32) of the B operand with 8 consecutive K values starting at
// 16 s + 8 (l / 32) for K step s; the accumulator holds D[row][l % 32] for
// row = 8 (v / 4) + 4 (l / 32) + (v % 4), v = 0..15. A and B use the same
// (lane, element) -> K assignment, so the product does not depend on the
// order in which the hardware walks K inside one instruction.
#include "../../include/optcuts_gpu.h"
#include "oc_common.hpp"
#include "oc_gnn.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace oc {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kGnnMaxGrid = 512;     // workgroups of a partial-producing kernel
constexpr int kWStrideM = kGnnH * kGnnKM + kGnnH;  // message weight-gradient record (W_m, b_m)
constexpr int kWStrideH = kGnnH * kGnnKH + kGnnH;  // update weight-gradient record (W_h, b_h)
constexpr int kWStrideHead = 4 * kGnnKM + 4 + 3 * kGnnH + 3;  // W_o, b_o, W_d, b_d
constexpr int kWStrideEnc = kGnnH * kGnnNX + kGnnH;             // W_enc, b_enc

// float -> bfloat16, round to nearest, ties to even (finite inputs)
__device__ __forceinline__ __bf16 to_bf16(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (unsigned short)(u >> 16));
}

__device__ __forceinline__ bf16x8 pack8(const float* __restrict__ p)
{
    const float4 a = reinterpret_cast<const float4*>(p)[0];
    const float4 b = reinterpret_cast<const float4*>(p)[1];
    bf16x8 r;
    r[0] = to_bf16(a.x);
    r[1] = to_bf16(a.y);
    r[2] = to_bf16(a.z);
    r[3] = to_bf16(a.w);
    r[4] = to_bf16(b.x);
    r[5] = to_bf16(b.y);
    r[6] = to_bf16(b.z);
    r[7] = to_bf16(b.w);
    return r;
}

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = to_bf16(0.0f);
    return r;
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x16 zero16()
{
    f32x16 z;
#pragma unroll
    for (int v = 0; v < 16; ++v) z[v] = 0.0f;
    return z;
}

__device__ __forceinline__ int acc_row(int v, int half) { return 8 * (v / 4) + 4 * half + (v % 4); }

int blocks_for(long items, int perBlock)
{
    const long b = (items + perBlock - 1) / perBlock;
    return (int)std::max(1L, std::min<long>(b, kGnnMaxGrid));
}
int tile_blocks(long rows) { return blocks_for((rows + 31) / 32, 4); }

__device__ __forceinline__ void chunk(long n, long& r0, long& r1)
{
    const long c = (n + gridDim.x - 1) / gridDim.x;
    r0 = (long)blockIdx.x * c;
    r1 = min(n, r0 + c);
}

// 2x2 Cholesky factor inverse of a symmetric block (lower: [[p, 0], [q, r]]).
__device__ __forceinline__ void chol_inv(double a, double b, double d, double& p, double& q, double& r)
{
    double c00 = a > 0.0 ? sqrt(a) : 1.0;
    const double c10 = b / c00;
    const double s = d - c10 * c10;
    const double c11 = s > 0.0 ? sqrt(s) : 1.0;
    p = 1.0 / c00;
    r = 1.0 / c11;
    q = -c10 / (c00 * c11);
}

template <typename TV>
__device__ __forceinline__ void load_block(const TV* vals, long slot, double& a, double& b, double& c, double& d)
{
    a = (double)vals[4 * slot + 0];
    b = (double)vals[4 * slot + 1];
    c = (double)vals[4 * slot + 2];
    d = (double)vals[4 * slot + 3];
}

// ---------------------------------------------------------------------------- features
template <typename TV>
__global__ void __launch_bounds__(kBlock) k_gnn_node_prep(int nV, const int* __restrict__ diagPos,
                                                          const TV* __restrict__ vals,
                                                          const unsigned char* __restrict__ fixed, float4* Cinv,
                                                          float* logTr, double* part)
{
    long r0, r1;
    chunk(nV, r0, r1);
    double acc[2] = {0.0, 0.0};
    for (long u = r0 + threadIdx.x; u < r1; u += blockDim.x) {
        double a, b, c, d;
        load_block(vals, diagPos[u], a, b, c, d);
        double p, q, r;
        chol_inv(a, 0.5 * (b + c), d, p, q, r);
        Cinv[u] = make_float4((float)p, 0.0f, (float)q, (float)r);
        const double lt = log(a + d);
        logTr[u] = (float)lt;
        if (!fixed[u]) {
            acc[0] += lt;
            acc[1] += 1.0;
        }
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = acc[0];
        part[2 * blockIdx.x + 1] = acc[1];
    }
}

// node inputs x and the encoder h0 = relu(W_enc x + b_enc); one thread per (vertex, unit)
__global__ void __launch_bounds__(kBlock) k_gnn_node_feat(int nV, const int* __restrict__ rowE,
                                                          const unsigned char* __restrict__ fixed,
                                                          const float* __restrict__ logTr,
                                                          const double* __restrict__ part, int nPart,
                                                          const float* __restrict__ theta, float4* X, float* h0)
{
    __shared__ double mean;
    {
        double acc[2] = {0.0, 0.0};
        for (int b = threadIdx.x; b < nPart; b += blockDim.x) {
            acc[0] += part[2 * b];
            acc[1] += part[2 * b + 1];
        }
        block_sum<2>(acc);
        if (threadIdx.x == 0) mean = acc[1] > 0.0 ? acc[0] / acc[1] : 0.0;
        __syncthreads();
    }
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nV * kGnnH) return;
    const int u = (int)(id / kGnnH), c = (int)(id % kGnnH);
    const bool fx = fixed[u] != 0;
    const float x[4] = {1.0f, (float)(rowE[u + 1] - rowE[u]) / 8.0f, fx ? 0.0f : (float)((double)logTr[u] - mean),
                        fx ? 1.0f : 0.0f};
    if (c == 0) X[u] = make_float4(x[0], x[1], x[2], x[3]);
    const float* W = theta + kGnnOffWenc;
    float z = theta[kGnnOffBenc + c];
#pragma unroll
    for (int i = 0; i < kGnnNX; ++i) z += W[c * kGnnNX + i] * x[i];
    h0[id] = fmaxf(z, 0.0f);
}

template <typename TV>
__global__ void __launch_bounds__(kBlock) k_gnn_edge_prep(int E, const int* __restrict__ eu, const int* __restrict__ ew,
                                                          const int* __restrict__ eslot,
                                                          const int* __restrict__ diagPos,
                                                          const TV* __restrict__ vals, float4* Ahat, float* ef)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int u = eu[e], w = ew[e];
    double a, b, c, d, pu, qu, ru, pw, qw, rw;
    load_block(vals, diagPos[u], a, b, c, d);
    chol_inv(a, 0.5 * (b + c), d, pu, qu, ru);
    load_block(vals, diagPos[w], a, b, c, d);
    chol_inv(a, 0.5 * (b + c), d, pw, qw, rw);
    load_block(vals, eslot[e], a, b, c, d);
    // L = Ci_u * A (Ci_u = [[pu, 0], [qu, ru]])
    const double l00 = pu * a, l01 = pu * b;
    const double l10 = qu * a + ru * c, l11 = qu * b + ru * d;
    // A^ = L * Ci_w^T (Ci_w^T = [[pw, qw], [0, rw]])
    const double h00 = l00 * pw, h01 = l00 * qw + l01 * rw;
    const double h10 = l10 * pw, h11 = l10 * qw + l11 * rw;
    Ahat[e] = make_float4((float)h00, (float)h01, (float)h10, (float)h11);
    float4* o = reinterpret_cast<float4*>(ef + (size_t)e * kGnnNEs);
    o[0] = make_float4((float)h00, (float)h01, (float)h10, (float)h11);
    o[1] = make_float4((float)sqrt(h00 * h00 + h01 * h01 + h10 * h10 + h11 * h11), w < u ? 1.0f : 0.0f,
                       w > u ? 1.0f : 0.0f, 0.0f);
}

// ---------------------------------------------------------------------------- forward layers (MFMA)
// m_e = relu(bf(h_u | h_w | e_e) bf(W_m)^T + b_m) for 32-edge tiles
__global__ void __launch_bounds__(kBlock) k_gnn_message(int E, const int* __restrict__ eu, const int* __restrict__ ew,
                                                        const float* __restrict__ ef, const float* __restrict__ h,
                                                        const float* __restrict__ W, const float* __restrict__ bias,
                                                        float* m)
{
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int n = lane % 32, half = lane / 32;
    bf16x8 wb[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) wb[s] = pack8(W + n * kGnnKM + 16 * s + 8 * half);
    const float bn = bias[n];
    for (long tile = (long)blockIdx.x * 4 + wave; tile * 32 < E; tile += (long)gridDim.x * 4) {
        const long r = tile * 32 + n;
        bf16x8 a[5];
        if (r < E) {
            const float* hu = h + (size_t)eu[r] * kGnnH;
            const float* hw = h + (size_t)ew[r] * kGnnH;
            a[0] = pack8(hu + 8 * half);
            a[1] = pack8(hu + 16 + 8 * half);
            a[2] = pack8(hw + 8 * half);
            a[3] = pack8(hw + 16 + 8 * half);
            a[4] = half ? zero8() : pack8(ef + (size_t)r * kGnnNEs);  // inputs 72..79 are zero
        } else {
#pragma unroll
            for (int s = 0; s < 5; ++s) a[s] = zero8();
        }
        f32x16 acc = zero16();
#pragma unroll
        for (int s = 0; s < 5; ++s) acc = mfma(a[s], wb[s], acc);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const long e = tile * 32 + acc_row(v, half);
            if (e < E) m[e * kGnnH + n] = fmaxf(acc[v] + bn, 0.0f);
        }
    }
}

// a_u = mean of m over the row; R = relu(bf(h_u | a_u) bf(W_h)^T + b_h); h_out = h + R
__global__ void __launch_bounds__(kBlock) k_gnn_update(int nV, const int* __restrict__ rowE,
                                                       const float* __restrict__ m, const float* __restrict__ h,
                                                       const float* __restrict__ W, const float* __restrict__ bias,
                                                       float* aOut, float* R, float* hOut)
{
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int n = lane % 32, half = lane / 32;
    bf16x8 wb[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) wb[s] = pack8(W + n * kGnnKH + 16 * s + 8 * half);
    const float bn = bias[n];
    for (long tile = (long)blockIdx.x * 4 + wave; tile * 32 < nV; tile += (long)gridDim.x * 4) {
        const long u = tile * 32 + n;
        bf16x8 a[4];
        if (u < nV) {
            const float* hu = h + (size_t)u * kGnnH;
            a[0] = pack8(hu + 8 * half);
            a[1] = pack8(hu + 16 + 8 * half);
            float s2[8], s3[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) s2[j] = s3[j] = 0.0f;
            const int e0 = rowE[u], e1 = rowE[u + 1];
            for (int e = e0; e < e1; ++e) {
                const float* me = m + (size_t)e * kGnnH;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    s2[j] += me[8 * half + j];
                    s3[j] += me[16 + 8 * half + j];
                }
            }
            const float inv = 1.0f / (float)max(e1 - e0, 1);
            float* ao = aOut + (size_t)u * kGnnH;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                s2[j] *= inv;
                s3[j] *= inv;
                ao[8 * half + j] = s2[j];
                ao[16 + 8 * half + j] = s3[j];
                a[2][j] = to_bf16(s2[j]);
                a[3][j] = to_bf16(s3[j]);
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) a[s] = zero8();
        }
        f32x16 acc = zero16();
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma(a[s], wb[s], acc);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const long node = tile * 32 + acc_row(v, half);
            if (node < nV) {
                const float r = fmaxf(acc[v] + bn, 0.0f);
                R[node * kGnnH + n] = r;
                hOut[node * kGnnH + n] = h[node * kGnnH + n] + r;
            }
        }
    }
}

// ---------------------------------------------------------------------------- heads (VALU, FP32)
__global__ void __launch_bounds__(kBlock) k_gnn_heads_edge(int E, const int* __restrict__ eu, const int* __restrict__ ew,
                                                           const float* __restrict__ ef, const float* __restrict__ h,
                                                           const float* __restrict__ theta, float4* Go)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int u = eu[e], w = ew[e];
    if (w > u) {
        Go[e] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float* Wo = theta + kGnnOffWo;
    float o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = theta[kGnnOffBo + i];
    const float* hu = h + (size_t)u * kGnnH;
    const float* hw = h + (size_t)w * kGnnH;
    const float* er = ef + (size_t)e * kGnnNEs;
    for (int k = 0; k < kGnnH; ++k) {
        const float xu = hu[k], xw = hw[k];
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] += Wo[i * kGnnKM + k] * xu + Wo[i * kGnnKM + kGnnH + k] * xw;
    }
    for (int k = 0; k < kGnnNEs; ++k) {
        const float x = er[k];
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] += Wo[i * kGnnKM + 2 * kGnnH + k] * x;
    }
    Go[e] = make_float4(o[0], o[1], o[2], o[3]);
}

__global__ void __launch_bounds__(kBlock) k_gnn_heads_node(int nV, const float* __restrict__ h,
                                                           const float* __restrict__ theta,
                                                           const unsigned char* __restrict__ fixed, float4* Gd)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nV) return;
    if (fixed[u]) {
        Gd[u] = make_float4(1.0f, 0.0f, 0.0f, 1.0f);
        return;
    }
    const float* Wd = theta + kGnnOffWd;
    float d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) d[i] = theta[kGnnOffBd + i];
    const float* hu = h + (size_t)u * kGnnH;
    for (int k = 0; k < kGnnH; ++k) {
        const float x = hu[k];
#pragma unroll
        for (int i = 0; i < 3; ++i) d[i] += Wd[i * kGnnH + k] * x;
    }
    Gd[u] = make_float4(expf(d[0]), 0.0f, d[1], expf(d[2]));
}

// ---------------------------------------------------------------------------- factor products (FP64 arithmetic)
__device__ __forceinline__ double2 bmv(float4 B, double2 v)
{
    return make_double2((double)B.x * v.x + (double)B.y * v.y, (double)B.z * v.x + (double)B.w * v.y);
}
__device__ __forceinline__ double2 bmtv(float4 B, double2 v)
{
    return make_double2((double)B.x * v.x + (double)B.z * v.y, (double)B.y * v.x + (double)B.w * v.y);
}

__device__ __forceinline__ unsigned int probe_hash(unsigned int i, unsigned int seed, unsigned int j)
{
    unsigned int x = i * 0x9E3779B1u + seed * 0x85EBCA77u + j * 0xC2B2AE3Du;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ double2 probe(int u, unsigned int seed, unsigned int j)
{
    return make_double2((probe_hash(2u * u, seed, j) & 1u) ? 1.0 : -1.0,
                        (probe_hash(2u * u + 1u, seed, j) & 1u) ? 1.0 : -1.0);
}

enum { kApplyPrecond = 0, kApplyLoss = 1 };

// out = G^ s, s_x = C^-1 in_x (precond) or in_x (loss: then out = 2 norm (G^ t - w), loss records)
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_gnn_apply_G(int nV, const int* __restrict__ rowE,
                                                        const int* __restrict__ ew, const float4* __restrict__ Cinv,
                                                        const float4* __restrict__ Go, const float4* __restrict__ Gd,
                                                        const double2* __restrict__ in, double2* out, double norm,
                                                        unsigned int seed, unsigned int pj, double* part,
                                                        const int* stop)
{
    if (stop && *stop) return;
    long r0, r1;
    chunk(nV, r0, r1);
    double acc[1] = {0.0};
    for (long u = r0 + threadIdx.x; u < r1; u += blockDim.x) {
        double2 su = in[u];
        if (MODE == kApplyPrecond) su = bmv(Cinv[u], su);
        double2 o = bmv(Gd[u], su);
        for (int e = rowE[u]; e < rowE[u + 1]; ++e) {
            const int w = ew[e];
            if (w >= u) continue;
            double2 sw = in[w];
            if (MODE == kApplyPrecond) sw = bmv(Cinv[w], sw);
            const double2 g = bmv(Go[e], sw);
            o.x += g.x;
            o.y += g.y;
        }
        if (MODE == kApplyLoss) {
            const double2 wv = probe((int)u, seed, pj);
            o.x -= wv.x;
            o.y -= wv.y;
            acc[0] += o.x * o.x + o.y * o.y;
            o.x *= 2.0 * norm;
            o.y *= 2.0 * norm;
        }
        out[u] = o;
    }
    if (MODE == kApplyLoss) {
        block_sum<1>(acc);
        if (threadIdx.x == 0) part[blockIdx.x] = acc[0] * norm;
    }
}

enum { kGTPrecond = 0, kGTProbe = 1, kGTPlain = 2 };

// y = G^T in (in = probe when MODE == kGTProbe); precond: z = C^-T y and r.z records
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_gnn_apply_GT(int nV, const int* __restrict__ rowE,
                                                         const int* __restrict__ ew, const int* __restrict__ rev,
                                                         const float4* __restrict__ Cinv,
                                                         const float4* __restrict__ Go, const float4* __restrict__ Gd,
                                                         const double2* __restrict__ in, double2* out,
                                                         const double2* __restrict__ r, double* rzPart,
                                                         unsigned int seed, unsigned int pj, const int* stop)
{
    if (stop && *stop) return;
    long r0, r1;
    chunk(nV, r0, r1);
    double acc[1] = {0.0};
    for (long u = r0 + threadIdx.x; u < r1; u += blockDim.x) {
        const double2 iu = MODE == kGTProbe ? probe((int)u, seed, pj) : in[u];
        double2 y = bmtv(Gd[u], iu);
        for (int e = rowE[u]; e < rowE[u + 1]; ++e) {
            const int w = ew[e];
            if (w <= u) continue;
            const double2 iw = MODE == kGTProbe ? probe(w, seed, pj) : in[w];
            const double2 g = bmtv(Go[rev[e]], iw);
            y.x += g.x;
            y.y += g.y;
        }
        if (MODE == kGTPrecond) {
            y = bmtv(Cinv[u], y);
            const double2 rv = r[u];
            acc[0] += rv.x * y.x + rv.y * y.y;
        }
        out[u] = y;
    }
    if (MODE == kGTPrecond) {
        block_sum<1>(acc);
        if (threadIdx.x == 0) rzPart[blockIdx.x] = acc[0];
    }
}

// out = A^ in (identity diagonal blocks)
__global__ void __launch_bounds__(kBlock) k_gnn_ahat(int nV, const int* __restrict__ rowE, const int* __restrict__ ew,
                                                     const float4* __restrict__ Ahat, const double2* __restrict__ in,
                                                     double2* out)
{
    const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nV) return;
    double2 o = in[u];
    for (int e = rowE[u]; e < rowE[u + 1]; ++e) {
        const double2 g = bmv(Ahat[e], in[ew[e]]);
        o.x += g.x;
        o.y += g.y;
    }
    out[u] = o;
}

// dG^_uv += gs_u t_v^T + w_u q_v^T on lower edges and diagonal blocks
__global__ void __launch_bounds__(kBlock) k_gnn_dG(int E, int nV, const int* __restrict__ eu,
                                                   const int* __restrict__ ew, const double2* __restrict__ gs,
                                                   const double2* __restrict__ t, const double2* __restrict__ q,
                                                   unsigned int seed, unsigned int pj, float4* dGo, float4* dGd)
{
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int u, v;
    float4* dst;
    if (id < E) {
        u = eu[id];
        v = ew[id];
        if (v > u) return;
        dst = dGo + id;
    } else if (id < (long)E + nV) {
        u = v = (int)(id - E);
        dst = dGd + u;
    } else {
        return;
    }
    const double2 a = gs[u], tv = t[v], wv = probe(u, seed, pj), qv = q[v];
    float4 d = *dst;
    d.x += (float)(a.x * tv.x + wv.x * qv.x);
    d.y += (float)(a.x * tv.y + wv.x * qv.y);
    d.z += (float)(a.y * tv.x + wv.y * qv.x);
    d.w += (float)(a.y * tv.y + wv.y * qv.y);
    *dst = d;
}

// ---------------------------------------------------------------------------- backward
// dh_u (final layer) from both heads; one thread per (vertex, unit)
__global__ void __launch_bounds__(kBlock) k_gnn_heads_bwd(int nV, const int* __restrict__ rowE,
                                                          const int* __restrict__ ew, const int* __restrict__ rev,
                                                          const unsigned char* __restrict__ fixed,
                                                          const float4* __restrict__ Gd, const float4* __restrict__ dGd,
                                                          const float4* __restrict__ dGo,
                                                          const float* __restrict__ theta, float* dh)
{
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nV * kGnnH) return;
    const int u = (int)(id / kGnnH), j = (int)(id % kGnnH);
    const float* Wd = theta + kGnnOffWd;
    const float* Wo = theta + kGnnOffWo;
    float acc = 0.0f;
    if (!fixed[u]) {
        const float4 g = Gd[u], dg = dGd[u];
        acc = dg.x * g.x * Wd[j] + dg.z * Wd[kGnnH + j] + dg.w * g.w * Wd[2 * kGnnH + j];
    }
    for (int e = rowE[u]; e < rowE[u + 1]; ++e) {
        const int w = ew[e];
        if (w < u) {
            const float4 d = dGo[e];
            acc += d.x * Wo[j] + d.y * Wo[kGnnKM + j] + d.z * Wo[2 * kGnnKM + j] + d.w * Wo[3 * kGnnKM + j];
        } else {
            const float4 d = dGo[rev[e]];
            acc += d.x * Wo[kGnnH + j] + d.y * Wo[kGnnKM + kGnnH + j] + d.z * Wo[2 * kGnnKM + kGnnH + j] +
                   d.w * Wo[3 * kGnnKM + kGnnH + j];
        }
    }
    dh[id] = acc;
}

// per-workgroup records of the head weight gradients [W_o (4x80), b_o, W_d (3x32), b_d]
__global__ void __launch_bounds__(512) k_gnn_heads_wgrad(int E, int nV, const int* __restrict__ eu,
                                                         const int* __restrict__ ew, const float* __restrict__ ef,
                                                         const float* __restrict__ h,
                                                         const unsigned char* __restrict__ fixed,
                                                         const float4* __restrict__ Gd,
                                                         const float4* __restrict__ dGd,
                                                         const float4* __restrict__ dGo, float* part)
{
    const int j = threadIdx.x;
    if (j >= kWStrideHead) return;
    float acc = 0.0f;
    if (j < 4 * kGnnKM + 4) {
        long e0, e1;
        chunk(E, e0, e1);
        const int i = j < 4 * kGnnKM ? j / kGnnKM : j - 4 * kGnnKM;
        const int k = j < 4 * kGnnKM ? j % kGnnKM : -1;
        for (long e = e0; e < e1; ++e) {
            const int u = eu[e], w = ew[e];
            if (w > u) continue;
            const float4 d = dGo[e];
            const float di = i == 0 ? d.x : (i == 1 ? d.y : (i == 2 ? d.z : d.w));
            float x = 1.0f;
            if (k >= 0)
                x = k < kGnnH ? h[(size_t)u * kGnnH + k]
                              : (k < 2 * kGnnH ? h[(size_t)w * kGnnH + k - kGnnH]
                                               : (k < 2 * kGnnH + kGnnNEs ? ef[(size_t)e * kGnnNEs + k - 2 * kGnnH]
                                                                          : 0.0f));
            acc += di * x;
        }
    } else {
        long u0, u1;
        chunk(nV, u0, u1);
        const int jj = j - (4 * kGnnKM + 4);
        const int i = jj < 3 * kGnnH ? jj / kGnnH : jj - 3 * kGnnH;
        const int k = jj < 3 * kGnnH ? jj % kGnnH : -1;
        for (long u = u0; u < u1; ++u) {
            if (fixed[u]) continue;
            const float4 g = Gd[u], dg = dGd[u];
            const float dd = i == 0 ? dg.x * g.x : (i == 1 ? dg.z : dg.w * g.w);
            acc += dd * (k >= 0 ? h[(size_t)u * kGnnH + k] : 1.0f);
        }
    }
    part[(size_t)blockIdx.x * kWStrideHead + j] = acc;
}

// Update-layer backward on 32-vertex tiles: dZ = dh * [R > 0];
// dh += bf(dZ) bf(W_h)[:, :32]; da = bf(dZ) bf(W_h)[:, 32:];
// records of dW_h = bf(dZ)^T bf(h | a) and db_h.
__global__ void __launch_bounds__(kBlock) k_gnn_update_bwd(int nV, const float* __restrict__ h,
                                                           const float* __restrict__ a, const float* __restrict__ R,
                                                           const float* __restrict__ W, float* dh, float* da,
                                                           float* part)
{
    __shared__ float sdz[4][32][33];
    __shared__ float sred[4][kWStrideH];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int n = lane % 32, half = lane / 32;
    // B fragments of the data product: B[k = c][col] = W_h[c][32 nt + col]
    bf16x8 wb[2][2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int j = 0; j < 8; ++j) wb[nt][s][j] = to_bf16(W[(16 * s + 8 * half + j) * kGnnKH + 32 * nt + n]);
    f32x16 accW[2] = {zero16(), zero16()};
    float accB = 0.0f;
    const long nTiles = ((long)nV + 31) / 32;
    for (long base = (long)blockIdx.x * 4; base < nTiles; base += (long)gridDim.x * 4) {
        const long tile = base + wave;
        {
            const long u = tile * 32 + n;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int cc = 16 * half + c;
                float z = 0.0f;
                if (u < nV && R[u * kGnnH + cc] > 0.0f) z = dh[u * kGnnH + cc];
                sdz[wave][n][cc] = z;
            }
        }
        __syncthreads();
        // data product
        bf16x8 ad[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int j = 0; j < 8; ++j) ad[s][j] = to_bf16(sdz[wave][n][16 * s + 8 * half + j]);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            f32x16 acc = zero16();
#pragma unroll
            for (int s = 0; s < 2; ++s) acc = mfma(ad[s], wb[nt][s], acc);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long node = tile * 32 + acc_row(v, half);
                if (node < nV) {
                    if (nt == 0)
                        dh[node * kGnnH + n] += acc[v];
                    else
                        da[node * kGnnH + n] = acc[v];
                }
            }
        }
        // weight product: A[c][i] = dZ[i][c], B[i][col] = (h | a)[i][32 nt + col]
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            const int i0 = 16 * ch + 8 * half;
            bf16x8 aw, bh, ba;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                aw[j] = to_bf16(sdz[wave][i0 + j][n]);
                const long node = tile * 32 + i0 + j;
                bh[j] = to_bf16(node < nV ? h[node * kGnnH + n] : 0.0f);
                ba[j] = to_bf16(node < nV ? a[node * kGnnH + n] : 0.0f);
            }
            accW[0] = mfma(aw, bh, accW[0]);
            accW[1] = mfma(aw, ba, accW[1]);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) accB += sdz[wave][16 * half + i][n];
        __syncthreads();
    }
    // workgroup record: sum of the four waves
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int v = 0; v < 16; ++v) sred[wave][acc_row(v, half) * kGnnKH + 32 * nt + n] = accW[nt][v];
    // db: lanes n and n + 32 hold halves of column n
    const float other = __shfl_xor(accB, 32);
    if (half == 0) sred[wave][kGnnH * kGnnKH + n] = accB + other;
    __syncthreads();
    for (int k = threadIdx.x; k < kWStrideH; k += blockDim.x)
        part[(size_t)blockIdx.x * kWStrideH + k] = sred[0][k] + sred[1][k] + sred[2][k] + sred[3][k];
}

// Message-layer backward on 32-edge tiles: dZ = (da_u / deg_u) * [m > 0];
// dXe[:, :64] = bf(dZ) bf(W_m)[:, :64]; records of dW_m = bf(dZ)^T bf(h_u | h_w | e) and db_m.
__global__ void __launch_bounds__(kBlock) k_gnn_message_bwd(int E, const int* __restrict__ eu,
                                                            const int* __restrict__ ew, const int* __restrict__ rowE,
                                                            const float* __restrict__ ef, const float* __restrict__ h,
                                                            const float* __restrict__ m, const float* __restrict__ da,
                                                            const float* __restrict__ W, float* dXe, float* part)
{
    __shared__ float sdz[4][32][33];
    __shared__ float sred[4][kWStrideM];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int n = lane % 32, half = lane / 32;
    bf16x8 wb[2][2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int j = 0; j < 8; ++j) wb[nt][s][j] = to_bf16(W[(16 * s + 8 * half + j) * kGnnKM + 32 * nt + n]);
    f32x16 accW[3] = {zero16(), zero16(), zero16()};
    float accB = 0.0f;
    const long nTiles = ((long)E + 31) / 32;
    for (long base = (long)blockIdx.x * 4; base < nTiles; base += (long)gridDim.x * 4) {
        const long tile = base + wave;
        {
            const long e = tile * 32 + n;
            float sc = 0.0f;
            int u = 0;
            if (e < E) {
                u = eu[e];
                sc = 1.0f / (float)max(rowE[u + 1] - rowE[u], 1);
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int cc = 16 * half + c;
                float z = 0.0f;
                if (e < E && m[e * kGnnH + cc] > 0.0f) z = da[(size_t)u * kGnnH + cc] * sc;
                sdz[wave][n][cc] = z;
            }
        }
        __syncthreads();
        bf16x8 ad[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int j = 0; j < 8; ++j) ad[s][j] = to_bf16(sdz[wave][n][16 * s + 8 * half + j]);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            f32x16 acc = zero16();
#pragma unroll
            for (int s = 0; s < 2; ++s) acc = mfma(ad[s], wb[nt][s], acc);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long e = tile * 32 + acc_row(v, half);
                if (e < E) dXe[e * (2 * kGnnH) + 32 * nt + n] = acc[v];
            }
        }
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            const int i0 = 16 * ch + 8 * half;
            bf16x8 aw, b0, b1, b2;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                aw[j] = to_bf16(sdz[wave][i0 + j][n]);
                const long e = tile * 32 + i0 + j;
                float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f;
                if (e < E) {
                    x0 = h[(size_t)eu[e] * kGnnH + n];
                    x1 = h[(size_t)ew[e] * kGnnH + n];
                    x2 = n < kGnnNEs ? ef[e * kGnnNEs + n] : 0.0f;
                }
                b0[j] = to_bf16(x0);
                b1[j] = to_bf16(x1);
                b2[j] = to_bf16(x2);
            }
            accW[0] = mfma(aw, b0, accW[0]);
            accW[1] = mfma(aw, b1, accW[1]);
            accW[2] = mfma(aw, b2, accW[2]);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) accB += sdz[wave][16 * half + i][n];
        __syncthreads();
    }
#pragma unroll
    for (int nt = 0; nt < 3; ++nt)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int col = 32 * nt + n;
            if (col < kGnnKM) sred[wave][acc_row(v, half) * kGnnKM + col] = accW[nt][v];
        }
    const float other = __shfl_xor(accB, 32);
    if (half == 0) sred[wave][kGnnH * kGnnKM + n] = accB + other;
    __syncthreads();
    for (int k = threadIdx.x; k < kWStrideM; k += blockDim.x)
        part[(size_t)blockIdx.x * kWStrideM + k] = sred[0][k] + sred[1][k] + sred[2][k] + sred[3][k];
}

// dh_u += sum over the row of dXe[e][:32] + dXe[rev e][32:]
__global__ void __launch_bounds__(kBlock) k_gnn_scatter_dh(int nV, const int* __restrict__ rowE,
                                                           const int* __restrict__ rev, const float* __restrict__ dXe,
                                                           float* dh)
{
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nV * kGnnH) return;
    const int u = (int)(id / kGnnH), j = (int)(id % kGnnH);
    float acc = 0.0f;
    for (int e = rowE[u]; e < rowE[u + 1]; ++e)
        acc += dXe[(size_t)e * (2 * kGnnH) + j] + dXe[(size_t)rev[e] * (2 * kGnnH) + kGnnH + j];
    dh[id] += acc;
}

// records of the encoder gradients [W_enc (32x4), b_enc]
__global__ void __launch_bounds__(kBlock) k_gnn_enc_wgrad(int nV, const float4* __restrict__ X,
                                                          const float* __restrict__ h0, const float* __restrict__ dh,
                                                          float* part)
{
    const int j = threadIdx.x;
    if (j >= kWStrideEnc) return;
    long u0, u1;
    chunk(nV, u0, u1);
    const int c = j < kGnnH * kGnnNX ? j / kGnnNX : j - kGnnH * kGnnNX;
    const int i = j < kGnnH * kGnnNX ? j % kGnnNX : -1;
    float acc = 0.0f;
    for (long u = u0; u < u1; ++u) {
        if (!(h0[u * kGnnH + c] > 0.0f)) continue;
        const float4 x = X[u];
        const float xi = i == 0 ? x.x : (i == 1 ? x.y : (i == 2 ? x.z : (i == 3 ? x.w : 1.0f)));
        acc += dh[u * kGnnH + c] * xi;
    }
    part[(size_t)blockIdx.x * kWStrideEnc + j] = acc;
}

// grad[j] += sum over records b of part[b * stride + j] (record order)
__global__ void __launch_bounds__(kBlock) k_gnn_reduce(int nRec, int stride, int nVals, const float* __restrict__ part,
                                                       float* grad)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nVals) return;
    float s = 0.0f;
    for (int b = 0; b < nRec; ++b) s += part[(size_t)b * stride + j];
    grad[j] += s;
}

__global__ void __launch_bounds__(kBlock) k_gnn_sum_to(const double* __restrict__ part, int n, double* out)
{
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < n; b += blockDim.x) acc[0] += part[b];
    block_sum<1>(acc);
    if (threadIdx.x == 0) *out += acc[0];
}

__global__ void __launch_bounds__(kBlock) k_gnn_fill(long n, float* p)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void __launch_bounds__(kBlock) k_gnn_adam(int n, float* theta, const float* __restrict__ g, float* m,
                                                     float* v, float lr, float b1, float b2, float eps, float c1,
                                                     float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    theta[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

template <typename T>
int dalloc(T*& p, size_t n)
{
    return hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * std::max<size_t>(n, 1)) == hipSuccess ? 0 : -1;
}
template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

// ============================================================================ host side
Gnn::~Gnn()
{
    release_pattern();
    dfree(dTheta);
    dfree(dGrad);
    dfree(dAdamM);
    dfree(dAdamV);
}

int Gnn::init(std::string& err)
{
    if (dTheta) return 0;
    if (dalloc(dTheta, kGnnParams) || dalloc(dGrad, kGnnParams) || dalloc(dAdamM, kGnnParams) ||
        dalloc(dAdamV, kGnnParams)) {
        err = "GNN: parameter allocation failed";
        return -1;
    }
    for (float* p : {dTheta, dGrad, dAdamM, dAdamV})
        if (hipMemset(p, 0, sizeof(float) * kGnnParams) != hipSuccess) {
            err = "GNN: parameter initialisation failed";
            return -1;
        }
    return 0;
}

void Gnn::release_pattern()
{
    for (int** p : {&dRowE, &dEu, &dEw, &dRev, &dESlot, &dDiagPos}) dfree(*p);
    dfree(dFixed);
    for (float** p : {&dCinv, &dLogTr, &dX, &dAhat, &dEf, &dGo, &dGd, &dDGo, &dDGd, &dDh, &dDa, &dDXe, &dPart}) dfree(*p);
    for (int l = 0; l <= kGnnLayers; ++l) dfree(dH[l]);
    for (int l = 0; l < kGnnLayers; ++l) {
        dfree(dA[l]);
        dfree(dR[l]);
        dfree(dM[l]);
    }
    for (double** p : {&dY, &dT, &dGs, &dQ, &dTmp, &dPartD}) dfree(*p);
    nV = E = 0;
}

int Gnn::set_pattern(hipStream_t s, int nVin, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                     const std::vector<unsigned char>& fixedMask, std::string& err)
{
    ++generation_;
    release_pattern();
    const int n = nVin;
    if (n <= 0 || (int)rowPtr.size() != n + 1 || (int)fixedMask.size() != n) {
        err = "GNN: bad pattern";
        return -1;
    }
    std::vector<int> rowE(n + 1, 0), eu, ew, eslot, diagPos(n, -1);
    for (int u = 0; u < n; ++u) {
        for (int k = rowPtr[u]; k < rowPtr[u + 1]; ++k) {
            const int w = colIdx[k];
            if (w == u) {
                diagPos[u] = k;
                continue;
            }
            eu.push_back(u);
            ew.push_back(w);
            eslot.push_back(k);
        }
        rowE[u + 1] = (int)eu.size();
    }
    for (int u = 0; u < n; ++u)
        if (diagPos[u] < 0) {
            err = "GNN: pattern row without its diagonal block";
            return -1;
        }
    const int nE = (int)eu.size();
    // reverse edges: rows are sorted, so (w, u) is found by binary search in row w
    std::vector<int> rev(nE, -1);
    for (int e = 0; e < nE; ++e) {
        const int u = eu[e], w = ew[e];
        int lo = rowE[w], hi = rowE[w + 1];
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (ew[mid] < u) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= rowE[w + 1] || ew[lo] != u) {
            err = "GNN: pattern is not symmetric";
            return -1;
        }
        rev[e] = lo;
    }
    nV = n;
    E = nE;
    auto up = [&](int*& d, const std::vector<int>& h) -> int {
        if (dalloc(d, h.size())) return -1;
        if (!h.empty() && hipMemcpyAsync(d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s) != hipSuccess)
            return -1;
        return 0;
    };
    int bad = up(dRowE, rowE) | up(dEu, eu) | up(dEw, ew) | up(dRev, rev) | up(dESlot, eslot) | up(dDiagPos, diagPos);
    bad |= dalloc(dFixed, n);
    if (!bad) bad |= hipMemcpyAsync(dFixed, fixedMask.data(), n, hipMemcpyHostToDevice, s) != hipSuccess;
    const size_t nh = (size_t)n * kGnnH, mh = (size_t)std::max(nE, 1) * kGnnH;
    bad |= dalloc(dCinv, 4 * (size_t)n) | dalloc(dLogTr, n) | dalloc(dX, 4 * (size_t)n) |
           dalloc(dAhat, 4 * (size_t)std::max(nE, 1)) | dalloc(dEf, (size_t)std::max(nE, 1) * kGnnNEs);
    for (int l = 0; l <= kGnnLayers; ++l) bad |= dalloc(dH[l], nh);
    for (int l = 0; l < kGnnLayers; ++l) bad |= dalloc(dA[l], nh) | dalloc(dR[l], nh) | dalloc(dM[l], mh);
    bad |= dalloc(dGo, 4 * (size_t)std::max(nE, 1)) | dalloc(dGd, 4 * (size_t)n);
    bad |= dalloc(dY, 2 * (size_t)n) | dalloc(dT, 2 * (size_t)n) | dalloc(dGs, 2 * (size_t)n) |
           dalloc(dQ, 2 * (size_t)n) | dalloc(dTmp, 2 * (size_t)n);
    bad |= dalloc(dDGo, 4 * (size_t)std::max(nE, 1)) | dalloc(dDGd, 4 * (size_t)n) | dalloc(dDh, nh) |
           dalloc(dDa, nh) | dalloc(dDXe, 2 * mh);
    nPart = (size_t)kGnnMaxGrid * (size_t)std::max(std::max(kWStrideM, kWStrideH), std::max(kWStrideHead, kWStrideEnc));
    bad |= dalloc(dPart, nPart) | dalloc(dPartD, 2 * (size_t)kGnnMaxGrid);
    if (bad || hipStreamSynchronize(s) != hipSuccess) {
        release_pattern();
        err = "GNN: device allocation or upload failed";
        return -1;
    }
    return 0;
}

void Gnn::forward(hipStream_t s, const void* vals, int vals32)
{
    const int gN = blocks_for(nV, kBlock);
    if (vals32)
        hipLaunchKernelGGL(k_gnn_node_prep<float>, dim3(gN), dim3(kBlock), 0, s, nV, dDiagPos,
                           static_cast<const float*>(vals), dFixed, reinterpret_cast<float4*>(dCinv), dLogTr, dPartD);
    else
        hipLaunchKernelGGL(k_gnn_node_prep<double>, dim3(gN), dim3(kBlock), 0, s, nV, dDiagPos,
                           static_cast<const double*>(vals), dFixed, reinterpret_cast<float4*>(dCinv), dLogTr, dPartD);
    const long nh = (long)nV * kGnnH;
    hipLaunchKernelGGL(k_gnn_node_feat, dim3((nh + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nV, dRowE, dFixed,
                       dLogTr, dPartD, gN, dTheta, reinterpret_cast<float4*>(dX), dH[0]);
    if (E > 0) {
        if (vals32)
            hipLaunchKernelGGL(k_gnn_edge_prep<float>, dim3((E + kBlock - 1) / kBlock), dim3(kBlock), 0, s, E, dEu,
                               dEw, dESlot, dDiagPos, static_cast<const float*>(vals),
                               reinterpret_cast<float4*>(dAhat), dEf);
        else
            hipLaunchKernelGGL(k_gnn_edge_prep<double>, dim3((E + kBlock - 1) / kBlock), dim3(kBlock), 0, s, E, dEu,
                               dEw, dESlot, dDiagPos, static_cast<const double*>(vals),
                               reinterpret_cast<float4*>(dAhat), dEf);
    }
    for (int l = 0; l < kGnnLayers; ++l) {
        if (E > 0)
            hipLaunchKernelGGL(k_gnn_message, dim3(tile_blocks(E)), dim3(kBlock), 0, s, E, dEu, dEw, dEf, dH[l],
                               dTheta + gnn_off_Wm(l), dTheta + gnn_off_bm(l), dM[l]);
        hipLaunchKernelGGL(k_gnn_update, dim3(tile_blocks(nV)), dim3(kBlock), 0, s, nV, dRowE, dM[l], dH[l],
                           dTheta + gnn_off_Wh(l), dTheta + gnn_off_bh(l), dA[l], dR[l], dH[l + 1]);
    }
    if (E > 0)
        hipLaunchKernelGGL(k_gnn_heads_edge, dim3((E + kBlock - 1) / kBlock), dim3(kBlock), 0, s, E, dEu, dEw, dEf,
                           dH[kGnnLayers], dTheta, reinterpret_cast<float4*>(dGo));
    hipLaunchKernelGGL(k_gnn_heads_node, dim3((nV + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nV, dH[kGnnLayers],
                       dTheta, dFixed, reinterpret_cast<float4*>(dGd));
}

int Gnn::rz_records() const { return blocks_for(nV, kBlock); }

void Gnn::apply(hipStream_t s, const double* r, double* z, double* rzPart, const int* stop)
{
    const int g = blocks_for(nV, kBlock);
    const float4* Ci = reinterpret_cast<const float4*>(dCinv);
    const float4* Go = reinterpret_cast<const float4*>(dGo);
    const float4* Gd = reinterpret_cast<const float4*>(dGd);
    hipLaunchKernelGGL(k_gnn_apply_G<kApplyPrecond>, dim3(g), dim3(kBlock), 0, s, nV, dRowE, dEw, Ci, Go, Gd,
                       reinterpret_cast<const double2*>(r), reinterpret_cast<double2*>(dTmp), 0.0, 0u, 0u, nullptr,
                       stop);
    hipLaunchKernelGGL(k_gnn_apply_GT<kGTPrecond>, dim3(g), dim3(kBlock), 0, s, nV, dRowE, dEw, dRev, Ci, Go, Gd,
                       reinterpret_cast<const double2*>(dTmp), reinterpret_cast<double2*>(z),
                       reinterpret_cast<const double2*>(r), rzPart, 0u, 0u, stop);
}

void Gnn::zero_grad(hipStream_t s)
{
    hipLaunchKernelGGL(k_gnn_fill, dim3((kGnnParams + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (long)kGnnParams,
                       dGrad);
}

void Gnn::loss_grad(hipStream_t s, int nProbes, unsigned seed, double* dLoss)
{
    const int g = blocks_for(nV, kBlock);
    const float4* Ci = reinterpret_cast<const float4*>(dCinv);
    const float4* Go = reinterpret_cast<const float4*>(dGo);
    const float4* Gd = reinterpret_cast<const float4*>(dGd);
    const float4* Ah = reinterpret_cast<const float4*>(dAhat);
    double2* y = reinterpret_cast<double2*>(dY);
    double2* t = reinterpret_cast<double2*>(dT);
    double2* gs = reinterpret_cast<double2*>(dGs);
    double2* q = reinterpret_cast<double2*>(dQ);
    double2* tmp = reinterpret_cast<double2*>(dTmp);
    const int gv = (nV + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_gnn_fill, dim3((4L * std::max(E, 1) + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       4L * std::max(E, 1), dDGo);
    hipLaunchKernelGGL(k_gnn_fill, dim3((4L * nV + kBlock - 1) / kBlock), dim3(kBlock), 0, s, 4L * nV, dDGd);
    const double norm = 1.0 / ((double)nProbes * 2.0 * nV);
    for (int j = 0; j < nProbes; ++j) {
        hipLaunchKernelGGL(k_gnn_apply_GT<kGTProbe>, dim3(g), dim3(kBlock), 0, s, nV, dRowE, dEw, dRev, Ci, Go, Gd,
                           nullptr, y, nullptr, nullptr, seed, (unsigned)j, nullptr);
        hipLaunchKernelGGL(k_gnn_ahat, dim3(gv), dim3(kBlock), 0, s, nV, dRowE, dEw, Ah, y, t);
        hipLaunchKernelGGL(k_gnn_apply_G<kApplyLoss>, dim3(g), dim3(kBlock), 0, s, nV, dRowE, dEw, Ci, Go, Gd, t, gs,
                           norm, seed, (unsigned)j, dPartD, nullptr);
        hipLaunchKernelGGL(k_gnn_sum_to, dim3(1), dim3(kBlock), 0, s, dPartD, g, dLoss);
        hipLaunchKernelGGL(k_gnn_apply_GT<kGTPlain>, dim3(g), dim3(kBlock), 0, s, nV, dRowE, dEw, dRev, Ci, Go, Gd, gs,
                           tmp, nullptr, nullptr, 0u, 0u, nullptr);
        hipLaunchKernelGGL(k_gnn_ahat, dim3(gv), dim3(kBlock), 0, s, nV, dRowE, dEw, Ah, tmp, q);
        hipLaunchKernelGGL(k_gnn_dG, dim3(((long)E + nV + kBlock - 1) / kBlock), dim3(kBlock), 0, s, E, nV, dEu, dEw,
                           gs, t, q, seed, (unsigned)j, reinterpret_cast<float4*>(dDGo),
                           reinterpret_cast<float4*>(dDGd));
    }
    // heads
    const long nh = (long)nV * kGnnH;
    hipLaunchKernelGGL(k_gnn_heads_bwd, dim3((nh + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nV, dRowE, dEw, dRev,
                       dFixed, Gd, reinterpret_cast<const float4*>(dDGd), reinterpret_cast<const float4*>(dDGo),
                       dTheta, dDh);
    const int gH = blocks_for(std::max(E, nV), 2048);
    hipLaunchKernelGGL(k_gnn_heads_wgrad, dim3(gH), dim3(512), 0, s, E, nV, dEu, dEw, dEf, dH[kGnnLayers], dFixed,
                       Gd, reinterpret_cast<const float4*>(dDGd), reinterpret_cast<const float4*>(dDGo), dPart);
    hipLaunchKernelGGL(k_gnn_reduce, dim3((kWStrideHead + kBlock - 1) / kBlock), dim3(kBlock), 0, s, gH, kWStrideHead,
                       kWStrideHead, dPart, dGrad + kGnnOffWo);
    // layers, last to first
    for (int l = kGnnLayers - 1; l >= 0; --l) {
        const int gU = tile_blocks(nV);
        hipLaunchKernelGGL(k_gnn_update_bwd, dim3(gU), dim3(kBlock), 0, s, nV, dH[l], dA[l], dR[l],
                           dTheta + gnn_off_Wh(l), dDh, dDa, dPart);
        hipLaunchKernelGGL(k_gnn_reduce, dim3((kWStrideH + kBlock - 1) / kBlock), dim3(kBlock), 0, s, gU, kWStrideH,
                           kWStrideH, dPart, dGrad + gnn_off_Wh(l));
        if (E > 0) {
            const int gM = tile_blocks(E);
            hipLaunchKernelGGL(k_gnn_message_bwd, dim3(gM), dim3(kBlock), 0, s, E, dEu, dEw, dRowE, dEf, dH[l], dM[l],
                               dDa, dTheta + gnn_off_Wm(l), dDXe, dPart);
            hipLaunchKernelGGL(k_gnn_reduce, dim3((kWStrideM + kBlock - 1) / kBlock), dim3(kBlock), 0, s, gM,
                               kWStrideM, kWStrideM, dPart, dGrad + gnn_off_Wm(l));
            hipLaunchKernelGGL(k_gnn_scatter_dh, dim3((nh + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nV, dRowE, dRev,
                               dDXe, dDh);
        }
    }
    const int gE = blocks_for(nV, 1024);
    hipLaunchKernelGGL(k_gnn_enc_wgrad, dim3(gE), dim3(kBlock), 0, s, nV, reinterpret_cast<const float4*>(dX), dH[0],
                       dDh, dPart);
    hipLaunchKernelGGL(k_gnn_reduce, dim3(1), dim3(kBlock), 0, s, gE, kWStrideEnc, kWStrideEnc, dPart,
                       dGrad + kGnnOffWenc);
}

void Gnn::adam(hipStream_t s, float lr, float b1, float b2, float eps, int t)
{
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, t)));
    hipLaunchKernelGGL(k_gnn_adam, dim3((kGnnParams + kBlock - 1) / kBlock), dim3(kBlock), 0, s, kGnnParams, dTheta,
                       dGrad, dAdamM, dAdamV, lr, b1, b2, eps, c1, c2);
}

}  // namespace oc

// ============================================================================ C ABI
#define GNN_HIP(g, call)                                                    \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            (g)->err = std::string(#call) + ": " + hipGetErrorString(e_);   \
            return OC_ERR_HIP;                                              \
        }                                                                   \
    } while (0)

#define GNN_ARG(g, cond, msg)  \
    do {                       \
        if (!(cond)) {         \
            (g)->err = (msg);  \
            return OC_ERR_ARG; \
        }                      \
    } while (0)

extern "C" {

int oc_gnn_num_params(void) { return oc::kGnnParams; }

int oc_gnn_create(int device, oc_gnn** out)
{
    if (!out) return OC_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return OC_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return OC_ERR_NO_DEVICE;
    oc_gnn* g = new oc_gnn();
    g->device = device;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return OC_ERR_HIP;
    }
    std::string e;
    if (g->net.init(e) != 0 || hipMalloc(reinterpret_cast<void**>(&g->dLoss), sizeof(double)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&g->dStop), sizeof(int)) != hipSuccess ||
        hipMemset(g->dStop, 0, sizeof(int)) != hipSuccess) {
        (void)hipStreamDestroy(g->stream);
        delete g;
        return OC_ERR_HIP;
    }
    *out = g;
    return OC_OK;
}

int oc_gnn_destroy(oc_gnn* g)
{
    if (!g) return OC_ERR_ARG;
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->stream);
    for (void* p : {(void*)g->dVals, (void*)g->dLoss, (void*)g->dR, (void*)g->dZ, (void*)g->dRz, (void*)g->dStop})
        if (p) (void)hipFree(p);
    (void)hipStreamDestroy(g->stream);
    delete g;
    return OC_OK;
}

const char* oc_gnn_last_error(const oc_gnn* g) { return g ? g->err.c_str() : "null network"; }

int oc_gnn_set_params(oc_gnn* g, const float* theta, int n)
{
    if (!g) return OC_ERR_ARG;
    GNN_ARG(g, theta && n == oc::kGnnParams, "oc_gnn_set_params: need oc_gnn_num_params() values");
    (void)hipSetDevice(g->device);
    GNN_HIP(g, hipMemcpyAsync(g->net.params(), theta, sizeof(float) * n, hipMemcpyHostToDevice, g->stream));
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    return OC_OK;
}

int oc_gnn_get_params(oc_gnn* g, float* theta, int n)
{
    if (!g) return OC_ERR_ARG;
    GNN_ARG(g, theta && n == oc::kGnnParams, "oc_gnn_get_params: need oc_gnn_num_params() values");
    (void)hipSetDevice(g->device);
    GNN_HIP(g, hipMemcpyAsync(theta, g->net.params(), sizeof(float) * n, hipMemcpyDeviceToHost, g->stream));
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    return OC_OK;
}

int oc_gnn_get_grads(oc_gnn* g, float* grad, int n)
{
    if (!g) return OC_ERR_ARG;
    GNN_ARG(g, grad && n == oc::kGnnParams, "oc_gnn_get_grads: need oc_gnn_num_params() values");
    (void)hipSetDevice(g->device);
    GNN_HIP(g, hipMemcpyAsync(grad, g->net.grads(), sizeof(float) * n, hipMemcpyDeviceToHost, g->stream));
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    return OC_OK;
}

int oc_gnn_set_system(oc_gnn* g, int nV, const int* rowPtr, const int* colIdx, const double* vals,
                      const unsigned char* fixedMask)
{
    if (!g) return OC_ERR_ARG;
    GNN_ARG(g, nV > 0 && rowPtr && colIdx && vals && fixedMask && rowPtr[0] == 0 && rowPtr[nV] > 0,
            "oc_gnn_set_system: bad arguments");
    (void)hipSetDevice(g->device);
    const int nnzb = rowPtr[nV];
    for (int u = 0; u < nV; ++u)
        GNN_ARG(g, rowPtr[u + 1] > rowPtr[u], "oc_gnn_set_system: empty row");
    for (int k = 0; k < nnzb; ++k) GNN_ARG(g, colIdx[k] >= 0 && colIdx[k] < nV, "oc_gnn_set_system: column out of range");
    for (int u = 0; u < nV; ++u)
        for (int k = rowPtr[u] + 1; k < rowPtr[u + 1]; ++k)
            GNN_ARG(g, colIdx[k] > colIdx[k - 1], "oc_gnn_set_system: rows must be sorted and unique");
    std::vector<int> rp(rowPtr, rowPtr + nV + 1), ci(colIdx, colIdx + nnzb);
    std::vector<unsigned char> fm(fixedMask, fixedMask + nV);
    if (!g->net.has_pattern() || rp != g->rowPtr || ci != g->colIdx || fm != g->fixed) {
        std::string e;
        if (g->net.set_pattern(g->stream, nV, rp, ci, fm, e) != 0) {
            g->err = e;
            return OC_ERR_PATTERN;
        }
        g->rowPtr = rp;
        g->colIdx = ci;
        g->fixed = fm;
        for (void** p : {(void**)&g->dR, (void**)&g->dZ, (void**)&g->dRz})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
        GNN_HIP(g, hipMalloc(reinterpret_cast<void**>(&g->dR), sizeof(double) * 2 * (size_t)nV));
        GNN_HIP(g, hipMalloc(reinterpret_cast<void**>(&g->dZ), sizeof(double) * 2 * (size_t)nV));
        GNN_HIP(g, hipMalloc(reinterpret_cast<void**>(&g->dRz), sizeof(double) * (size_t)g->net.rz_records()));
    }
    const size_t bytes = sizeof(double) * 4 * (size_t)nnzb;
    if (g->valsBytes < bytes) {
        if (g->dVals) (void)hipFree(g->dVals);
        g->dVals = nullptr;
        g->valsBytes = 0;
        GNN_HIP(g, hipMalloc(&g->dVals, bytes));
        g->valsBytes = bytes;
    }
    GNN_HIP(g, hipMemcpyAsync(g->dVals, vals, bytes, hipMemcpyHostToDevice, g->stream));
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    return OC_OK;
}

int oc_gnn_forward(oc_gnn* g, float* Gblocks)
{
    if (!g) return OC_ERR_ARG;
    if (!g->net.has_pattern()) {
        g->err = "oc_gnn_forward: call oc_gnn_set_system first";
        return OC_ERR_STATE;
    }
    (void)hipSetDevice(g->device);
    g->net.forward(g->stream, g->dVals, 0);
    GNN_HIP(g, hipGetLastError());
    if (Gblocks) {
        const int nV = g->net.num_vertices(), E = g->net.num_edges();
        std::vector<float> go(4 * (size_t)std::max(E, 1)), gd(4 * (size_t)nV);
        GNN_HIP(g, hipMemcpyAsync(go.data(), g->net.factor_lower(), sizeof(float) * 4 * (size_t)E,
                                  hipMemcpyDeviceToHost, g->stream));
        GNN_HIP(g, hipMemcpyAsync(gd.data(), g->net.factor_diag(), sizeof(float) * 4 * (size_t)nV,
                                  hipMemcpyDeviceToHost, g->stream));
        GNN_HIP(g, hipStreamSynchronize(g->stream));
        int e = 0;
        for (int u = 0; u < nV; ++u)
            for (int k = g->rowPtr[u]; k < g->rowPtr[u + 1]; ++k) {
                const int w = g->colIdx[k];
                const float* src = w == u ? &gd[4 * (size_t)u] : &go[4 * (size_t)e];
                if (w != u) ++e;
                const bool keep = w <= u;
                for (int i = 0; i < 4; ++i) Gblocks[4 * (size_t)k + i] = keep ? src[i] : 0.0f;
            }
    }
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    return OC_OK;
}

int oc_gnn_zero_grad(oc_gnn* g)
{
    if (!g) return OC_ERR_ARG;
    (void)hipSetDevice(g->device);
    g->net.zero_grad(g->stream);
    GNN_HIP(g, hipGetLastError());
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    return OC_OK;
}

int oc_gnn_loss_grad(oc_gnn* g, int nProbes, unsigned int seed, double* loss)
{
    if (!g) return OC_ERR_ARG;
    GNN_ARG(g, nProbes > 0 && nProbes <= 1024, "oc_gnn_loss_grad: 1..1024 probes");
    if (!g->net.has_pattern()) {
        g->err = "oc_gnn_loss_grad: call oc_gnn_set_system first";
        return OC_ERR_STATE;
    }
    (void)hipSetDevice(g->device);
    GNN_HIP(g, hipMemsetAsync(g->dLoss, 0, sizeof(double), g->stream));
    g->net.forward(g->stream, g->dVals, 0);
    g->net.loss_grad(g->stream, nProbes, seed, g->dLoss);
    GNN_HIP(g, hipGetLastError());
    double l = 0.0;
    GNN_HIP(g, hipMemcpyAsync(&l, g->dLoss, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    if (loss) *loss = l;
    return OC_OK;
}

int oc_gnn_adam_step(oc_gnn* g, float lr, float beta1, float beta2, float eps, int t)
{
    if (!g) return OC_ERR_ARG;
    GNN_ARG(g, t >= 1 && lr > 0.0f && beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f && eps > 0.0f,
            "oc_gnn_adam_step: bad hyper-parameters");
    (void)hipSetDevice(g->device);
    g->net.adam(g->stream, lr, beta1, beta2, eps, t);
    GNN_HIP(g, hipGetLastError());
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    return OC_OK;
}

int oc_gnn_precond_apply(oc_gnn* g, const double* r, double* z)
{
    if (!g) return OC_ERR_ARG;
    GNN_ARG(g, r && z, "oc_gnn_precond_apply: null vector");
    if (!g->net.has_pattern()) {
        g->err = "oc_gnn_precond_apply: call oc_gnn_set_system and oc_gnn_forward first";
        return OC_ERR_STATE;
    }
    (void)hipSetDevice(g->device);
    const size_t n = 2 * (size_t)g->net.num_vertices();
    GNN_HIP(g, hipMemcpyAsync(g->dR, r, sizeof(double) * n, hipMemcpyHostToDevice, g->stream));
    g->net.apply(g->stream, g->dR, g->dZ, g->dRz, g->dStop);
    GNN_HIP(g, hipGetLastError());
    GNN_HIP(g, hipMemcpyAsync(z, g->dZ, sizeof(double) * n, hipMemcpyDeviceToHost, g->stream));
    GNN_HIP(g, hipStreamSynchronize(g->stream));
    return OC_OK;
}

}  // extern "C"
