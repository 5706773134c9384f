// Block IC(0) preconditioner (see oc_ic0.hpp): fixed-point factor sweeps and Jacobi-swept triangular solves.
#include "oc_common.hpp"
#include "oc_ic0.hpp"

#include <algorithm>
#include <cmath>

namespace oc {
namespace {

constexpr int kIcMaxGrid = 1024;  // workgroups of the row kernels (bounds the r.z records)

int row_grid(long n) { return (int)std::max(1L, std::min<long>((n + kBlock - 1) / kBlock, kIcMaxGrid)); }

// 2x2 blocks are double4 (x, y; z, w), row-major.
__device__ __forceinline__ double4 sub_abT(double4 s, double4 a, double4 b)
{
    // s - a b^T
    s.x -= a.x * b.x + a.y * b.y;
    s.y -= a.x * b.z + a.y * b.w;
    s.z -= a.z * b.x + a.w * b.y;
    s.w -= a.z * b.z + a.w * b.w;
    return s;
}

// Lower Cholesky factor (x, 0; z, w) of the symmetric part of s; false when it is not positive definite.
__device__ __forceinline__ bool chol2(double4 s, double4& l)
{
    const double off = 0.5 * (s.y + s.z);
    if (!(s.x > 0.0)) return false;
    const double l00 = sqrt(s.x);
    const double l10 = off / l00;
    const double t = s.w - l10 * l10;
    if (!(t > 0.0)) return false;
    l = make_double4(l00, 0.0, l10, sqrt(t));
    return true;
}

// chol2 with the diagonal-part fallback (positive diagonal entries of a, else 1).
__device__ __forceinline__ double4 chol2_safe(double4 s, double4 a, unsigned int* fail)
{
    double4 l;
    if (chol2(s, l)) return l;
    if (fail) atomicAdd(fail, 1u);
    return make_double4(a.x > 0.0 ? sqrt(a.x) : 1.0, 0.0, 0.0, a.w > 0.0 ? sqrt(a.w) : 1.0);
}

// s L^-T for lower L = (p, 0; q, r)
__device__ __forceinline__ double4 mul_inv_lt(double4 s, double4 L)
{
    const double p = L.x, q = L.z, r = L.w;
    const double c0 = 1.0 / p, c1 = -q / (p * r), c2 = 1.0 / r;  // L^-T = (c0, c1; 0, c2)
    return make_double4(s.x * c0, s.x * c1 + s.y * c2, s.z * c0, s.z * c1 + s.w * c2);
}

// L^-1 v (forward substitution) and L^-T v (backward) for lower L = (p, 0; q, r)
__device__ __forceinline__ double2 solve_l(double4 L, double2 v)
{
    const double y0 = v.x / L.x;
    return make_double2(y0, (v.y - L.z * y0) / L.w);
}
__device__ __forceinline__ double2 solve_lt(double4 L, double2 v)
{
    const double z1 = v.y / L.w;
    return make_double2((v.x - L.z * z1) / L.x, z1);
}

// One factor sweep over the lower blocks (first: the starting factor, no products).
__global__ void __launch_bounds__(kBlock) k_ic_factor(long nLow, const int* __restrict__ low,
                                                      const int* __restrict__ rowOf, const int* __restrict__ rowPtr,
                                                      const int* __restrict__ colIdx, const int* __restrict__ diag,
                                                      const double4* __restrict__ A, const double4* __restrict__ Lold,
                                                      double4* __restrict__ Lnew, int first, unsigned int* fail)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nLow) return;
    const int e = low[i];
    const int u = rowOf[e], w = colIdx[e];
    double4 s = A[e];
    if (!first) {
        // sum over k < w of L_uk L_wk^T: merge rows u and w (sorted columns)
        int p = rowPtr[u], q = rowPtr[w];
        const int qe = diag[w];
        while (q < qe) {
            const int kp = colIdx[p];
            if (kp >= w) break;
            const int kq = colIdx[q];
            if (kp == kq) {
                s = sub_abT(s, Lold[p], Lold[q]);
                ++p;
                ++q;
            } else if (kp < kq) {
                ++p;
            } else {
                ++q;
            }
        }
    }
    if (u == w) {
        Lnew[e] = chol2_safe(s, A[e], fail);
    } else {
        const int dw = diag[w];
        const double4 Lw = first ? chol2_safe(A[dw], A[dw], nullptr) : Lold[dw];
        Lnew[e] = mul_inv_lt(s, Lw);
    }
}

// Forward sweep: yout_u = L_uu^-1 (r_u - sum_{w<u} L_uw yin_w) (yin null: the diagonal solve alone).
__global__ void __launch_bounds__(kBlock) k_ic_fwd(int n, const int* __restrict__ rowPtr,
                                                   const int* __restrict__ colIdx, const int* __restrict__ diag,
                                                   const double4* __restrict__ L, const double2* __restrict__ r,
                                                   const double2* __restrict__ yin, double2* __restrict__ yout,
                                                   const int* stop)
{
    if (stop && *stop) return;
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        double2 acc = r[u];
        const int du = diag[u];
        if (yin)
            for (int e = rowPtr[u]; e < du; ++e) {
                const double4 b = L[e];
                const double2 y = yin[colIdx[e]];
                acc.x -= b.x * y.x + b.y * y.y;
                acc.y -= b.z * y.x + b.w * y.y;
            }
        yout[u] = solve_l(L[du], acc);
    }
}

// Backward sweep: zout_u = L_uu^-T (y_u - sum_{w>u} L_wu^T zin_w) (zin null: the diagonal solve alone); with
// rzPart, per-workgroup sums of r.zout.
__global__ void __launch_bounds__(kBlock) k_ic_bwd(int n, const int* __restrict__ rowPtr,
                                                   const int* __restrict__ colIdx, const int* __restrict__ diag,
                                                   const int* __restrict__ rev, const double4* __restrict__ L,
                                                   const double2* __restrict__ y, const double2* __restrict__ zin,
                                                   double2* __restrict__ zout, const double2* __restrict__ r,
                                                   double* rzPart, const int* stop)
{
    if (stop && *stop) return;
    double acc[1] = {0.0};
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        double2 v = y[u];
        const int du = diag[u];
        if (zin)
            for (int e = du + 1; e < rowPtr[u + 1]; ++e) {
                const double4 b = L[rev[e]];  // L_wu; its transpose multiplies z_w
                const double2 z = zin[colIdx[e]];
                v.x -= b.x * z.x + b.z * z.y;
                v.y -= b.y * z.x + b.w * z.y;
            }
        const double2 z = solve_lt(L[du], v);
        zout[u] = z;
        if (rzPart) {
            const double2 rv = r[u];
            acc[0] += rv.x * z.x + rv.y * z.y;
        }
    }
    if (rzPart) {
        block_sum<1>(acc);
        if (threadIdx.x == 0) rzPart[blockIdx.x] = acc[0];
    }
}

template <typename T>
hipError_t upload_vec(hipStream_t s, T*& d, const std::vector<T>& h)
{
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), sizeof(T) * std::max<size_t>(1, h.size()));
    if (e != hipSuccess) return e;
    if (h.empty()) return hipSuccess;
    return hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s);
}

}  // namespace

Ic0::~Ic0() { release(); }

void Ic0::release()
{
    for (int* p : {dRowPtr_, dColIdx_, dDiag_, dRev_, dLow_, dRowOf_})
        if (p) (void)hipFree(p);
    for (double* p : {dL_[0], dL_[1], dT_[0], dT_[1], dT_[2], dT_[3]})
        if (p) (void)hipFree(p);
    if (dFail_) (void)hipFree(dFail_);
    dRowPtr_ = dColIdx_ = dDiag_ = dRev_ = dLow_ = dRowOf_ = nullptr;
    dL_[0] = dL_[1] = nullptr;
    dT_[0] = dT_[1] = dT_[2] = dT_[3] = nullptr;
    dFail_ = nullptr;
    nV_ = 0;
    nnzb_ = nLow_ = 0;
}

int Ic0::setup(hipStream_t s, int nV, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
               std::string& err)
{
    (void)hipStreamSynchronize(s);
    release();
    ++generation_;
    if (nV <= 0 || (int)rowPtr.size() != nV + 1 || (long)colIdx.size() != (long)rowPtr[nV]) {
        err = "IC(0): bad pattern";
        return -1;
    }
    const long nnzb = rowPtr[nV];
    std::vector<int> diag(nV, -1), rev(nnzb, -1), rowOf(nnzb), low;
    low.reserve((nnzb + nV) / 2);
    for (int u = 0; u < nV; ++u)
        for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
            const int w = colIdx[e];
            if (w < 0 || w >= nV || (e > rowPtr[u] && colIdx[e - 1] >= w)) {
                err = "IC(0): pattern rows must be sorted and in range";
                return -1;
            }
            rowOf[e] = u;
            if (w == u) diag[u] = e;
            if (w <= u) low.push_back(e);
            const int* b = colIdx.data() + rowPtr[w];
            const int* f = colIdx.data() + rowPtr[w + 1];
            const int* p = std::lower_bound(b, f, u);
            if (p == f || *p != u) {
                err = "IC(0): pattern is not symmetric";
                return -1;
            }
            rev[e] = (int)(p - colIdx.data());
        }
    for (int u = 0; u < nV; ++u)
        if (diag[u] < 0) {
            err = "IC(0): a row has no diagonal block";
            return -1;
        }
    hipError_t e = hipSuccess;
    if ((e = upload_vec(s, dRowPtr_, rowPtr)) != hipSuccess || (e = upload_vec(s, dColIdx_, colIdx)) != hipSuccess ||
        (e = upload_vec(s, dDiag_, diag)) != hipSuccess || (e = upload_vec(s, dRev_, rev)) != hipSuccess ||
        (e = upload_vec(s, dLow_, low)) != hipSuccess || (e = upload_vec(s, dRowOf_, rowOf)) != hipSuccess) {
        err = std::string("IC(0): ") + hipGetErrorString(e);
        release();
        return -1;
    }
    for (double*& p : dL_)
        if ((e = hipMalloc(reinterpret_cast<void**>(&p), sizeof(double) * 4 * (size_t)nnzb)) != hipSuccess) break;
    if (e == hipSuccess)
        for (double*& p : dT_)
            if ((e = hipMalloc(reinterpret_cast<void**>(&p), sizeof(double) * 2 * (size_t)nV)) != hipSuccess) break;
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dFail_), sizeof(unsigned int));
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host vectors end with this call
    if (e != hipSuccess) {
        err = std::string("IC(0): ") + hipGetErrorString(e);
        release();
        return -1;
    }
    nV_ = nV;
    nnzb_ = nnzb;
    nLow_ = (long)low.size();
    return 0;
}

void Ic0::set_sweeps(int factor, int apply)
{
    factorSweeps_ = std::max(0, factor);
    applySweeps_ = std::max(0, apply);
    ++generation_;  // apply() may now read the other factor buffer
}

void Ic0::numeric(hipStream_t s, const double* vals)
{
    if (!ready()) return;
    (void)hipMemsetAsync(dFail_, 0, sizeof(unsigned int), s);
    const double4* A = reinterpret_cast<const double4*>(vals);
    const int g = (int)((nLow_ + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_ic_factor, dim3(g), dim3(kBlock), 0, s, nLow_, dLow_, dRowOf_, dRowPtr_, dColIdx_, dDiag_,
                       A, nullptr, reinterpret_cast<double4*>(dL_[0]), 1, dFail_);
    for (int k = 1; k <= factorSweeps_; ++k) {
        // only the last sweep's diagonal breakdowns count
        if (k == factorSweeps_) (void)hipMemsetAsync(dFail_, 0, sizeof(unsigned int), s);
        hipLaunchKernelGGL(k_ic_factor, dim3(g), dim3(kBlock), 0, s, nLow_, dLow_, dRowOf_, dRowPtr_, dColIdx_,
                           dDiag_, A, reinterpret_cast<const double4*>(dL_[(k - 1) & 1]),
                           reinterpret_cast<double4*>(dL_[k & 1]), 0, dFail_);
    }
}

int Ic0::rz_records() const { return row_grid(nV_); }

void Ic0::apply(hipStream_t s, const double* r, double* z, double* rzPart, const int* stop)
{
    const int n = nV_, g = row_grid(n), m = applySweeps_;
    const double4* L = reinterpret_cast<const double4*>(dL_[factorSweeps_ & 1]);
    const double2* r2 = reinterpret_cast<const double2*>(r);
    double2* T[4];
    for (int k = 0; k < 4; ++k) T[k] = reinterpret_cast<double2*>(dT_[k]);
    // forward: T0 = D^-1 r, then sweeps alternating T1, T0, ...
    hipLaunchKernelGGL(k_ic_fwd, dim3(g), dim3(kBlock), 0, s, n, dRowPtr_, dColIdx_, dDiag_, L, r2,
                       (const double2*)nullptr, T[0], stop);
    for (int k = 1; k <= m; ++k)
        hipLaunchKernelGGL(k_ic_fwd, dim3(g), dim3(kBlock), 0, s, n, dRowPtr_, dColIdx_, dDiag_, L, r2,
                           (const double2*)T[(k - 1) & 1], T[k & 1], stop);
    const double2* y = T[m & 1];
    // backward: diagonal solve then sweeps alternating T2, T3; the last one writes z and the r.z records
    double2* z2 = reinterpret_cast<double2*>(z);
    for (int k = 0; k <= m; ++k) {
        const double2* zin = k == 0 ? nullptr : T[2 + ((k - 1) & 1)];
        double2* zout = k == m ? z2 : T[2 + (k & 1)];
        hipLaunchKernelGGL(k_ic_bwd, dim3(g), dim3(kBlock), 0, s, n, dRowPtr_, dColIdx_, dDiag_, dRev_, L, y, zin,
                           zout, r2, k == m ? rzPart : nullptr, stop);
    }
}

int Ic0::breakdowns(hipStream_t s)
{
    if (!ready()) return 0;
    unsigned int h = 0;
    if (hipMemcpyAsync(&h, dFail_, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    return (int)h;
}

}  // namespace oc
