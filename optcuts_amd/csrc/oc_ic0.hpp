// Block incomplete Cholesky IC(0) on the 2x2-block Newton pattern, factored and applied on the device.
//
//   A ~= L L^T, L block lower triangular on the lower half of A's pattern (no fill), diagonal blocks lower
//   triangular: the IC(0) factor, whose product L L^T equals A on every block of the pattern.
//
// Factor: the fine-grained fixed-point iteration of Chow and Patel (SIAM J. Sci. Comput. 37(2), C169-C193,
// 2015). Every lower block is recomputed from the previous sweep's blocks, all blocks in parallel:
//   L_uw = (A_uw - sum_{k<w} L_uk L_wk^T) L_ww^-T   (w < u),
//   L_uu = chol(A_uu - sum_{k<u} L_uk L_uk^T),
// starting from L_uw = A_uw chol(A_ww)^-T, L_uu = chol(A_uu). The sequential IC(0) factor is its fixed point and
// is reached exactly after as many sweeps as the longest dependency chain; a few sweeps give a usable factor.
// A diagonal block whose Schur complement is not positive definite (IC(0) can break down on matrices that are
// not M-matrices) takes the Cholesky factor of its diagonal part instead; such blocks are counted.
//
// Apply: z = M^-1 r with the two triangular solves replaced by a fixed number of block-Jacobi sweeps each
// (iterative sparse triangular solves: Anzt, Chow and Dongarra, ISC 2015),
//   y_0 = D^-1 r,   y_k = D^-1 (r - N y_{k-1});   z_0 = D^-T y,   z_k = D^-T (y - N^T z_{k-1}),
// with D the diagonal blocks of L and N its strict lower part. Then y = B r and z = B^T y for one fixed
// polynomial B in D^-1 N, so M^-1 = B^T B is symmetric positive definite for every sweep count (PCG stays
// valid) and equals (L L^T)^-1 once the sweeps reach the dependency depth.
#pragma once

#include <hip/hip_runtime.h>

#include "oc_precond.hpp"

#include <string>
#include <vector>

namespace oc {

constexpr int kIcFactorSweeps = 3;  // fixed-point sweeps after the starting factor
constexpr int kIcApplySweeps = 2;   // Jacobi sweeps per triangular solve after the diagonal one

class Ic0 : public Preconditioner {
public:
    Ic0() = default;
    ~Ic0() override;
    Ic0(const Ic0&) = delete;
    Ic0& operator=(const Ic0&) = delete;

    // Pattern of a full symmetric 2x2-block matrix (rows sorted, each with its diagonal). Returns 0 or -1.
    int setup(hipStream_t s, int nV, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
              std::string& err);
    bool ready() const { return nV_ > 0; }
    // Factor of the FP64 block values (double4 per block of the pattern) on the stream.
    void numeric(hipStream_t s, const double* vals);
    void apply(hipStream_t s, const double* r, double* z, double* rzPart, const int* stop) override;
    int rz_records() const override;
    // Diagonal blocks that took the diagonal-part fallback in the last numeric() (synchronises the stream).
    int breakdowns(hipStream_t s);
    // Sweep counts (factor >= 0, apply >= 0); take effect at the next numeric() / apply().
    void set_sweeps(int factor, int apply);
    int factor_sweeps() const { return factorSweeps_; }
    int apply_sweeps() const { return applySweeps_; }

private:
    void release();
    int nV_ = 0;
    long nnzb_ = 0, nLow_ = 0;
    int factorSweeps_ = kIcFactorSweeps, applySweeps_ = kIcApplySweeps;
    int* dRowPtr_ = nullptr;
    int* dColIdx_ = nullptr;
    int* dDiag_ = nullptr;  // position of (u, u)
    int* dRev_ = nullptr;   // position of (w, u) for the block at (u, w)
    int* dLow_ = nullptr;   // positions of the lower blocks (w <= u)
    int* dRowOf_ = nullptr;  // row of each block position
    double* dL_[2] = {nullptr, nullptr};  // factor blocks at their pattern positions (lower half used)
    double* dT_[4] = {nullptr, nullptr, nullptr, nullptr};  // apply sweep vectors
    unsigned int* dFail_ = nullptr;
};

}  // namespace oc
