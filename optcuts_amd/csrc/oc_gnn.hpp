// Learned preconditioner for the Newton systems: a 3-layer edge-message
// graph network over the 2x2-block sparsity graph predicts a sparse
// lower-triangular factor G^ on the Hessian pattern, and
//     M^-1 = C^-T G^T G C^-1      (D_u = A_uu = C_u C_u^T)
// preconditions PCG. With zero output heads G^ = I and M is block Jacobi.
//
// Network (hidden width 32; bf() = round to bfloat16 for the MFMA products):
//   node input x_u = [1, deg_u/8, log tr D_u - mean, fixed_u]
//   edge input e_uw = [A^_uw (4), |A^_uw|_F, w<u, w>u, 0 x 9], A^ = C^-1 A C^-T
//   h = relu(W_enc x + b_enc)
//   3 x { m_uw = relu(bf(h_u|h_w|e_uw) bf(W_m)^T + b_m);  a_u = mean_w m_uw;
//         h_u += relu(bf(h_u|a_u) bf(W_h)^T + b_h) }
//   G^_uw = (h_u|h_w|e_uw) W_o^T + b_o (w < u);  G^_uu = [[e^d0, 0], [d1, e^d2]],
//   d = W_d h_u + b_d (identity on fixed vertices).
// Training loss (self-supervised, per system): the Hutchinson estimate
//   L = sum_j |G^ A^ G^T w_j - w_j|^2 / (k 2 nV)  over Rademacher probes w_j
// of |G^ A^ G^T - I|_F^2 / (2 nV). The message and update products (forward,
// data-gradient and weight-gradient) run on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16: bf16 operands, FP32 accumulation); heads,
// features and the sparse products run on the vector ALUs.
#pragma once

#include <hip/hip_runtime.h>

#include "oc_precond.hpp"

#include <string>
#include <vector>

namespace oc {

constexpr int kGnnH = 32;
constexpr int kGnnNX = 4;
constexpr int kGnnNE = 16;   // edge-feature inputs of the message and head layers (columns of W_m, W_o)
constexpr int kGnnNEs = 8;   // edge features stored per edge: the 7 defined + 1 zero; inputs kGnnNEs.. are zero
constexpr int kGnnKM = 2 * kGnnH + kGnnNE;  // 80
constexpr int kGnnKH = 2 * kGnnH;           // 64
constexpr int kGnnLayers = 3;
constexpr int kGnnLayerParams = kGnnH * kGnnKM + kGnnH + kGnnH * kGnnKH + kGnnH;
// flat parameter layout
constexpr int kGnnOffWenc = 0;
constexpr int kGnnOffBenc = kGnnH * kGnnNX;
constexpr int gnn_off_Wm(int l) { return kGnnOffBenc + kGnnH + l * kGnnLayerParams; }
constexpr int gnn_off_bm(int l) { return gnn_off_Wm(l) + kGnnH * kGnnKM; }
constexpr int gnn_off_Wh(int l) { return gnn_off_bm(l) + kGnnH; }
constexpr int gnn_off_bh(int l) { return gnn_off_Wh(l) + kGnnH * kGnnKH; }
constexpr int kGnnOffWo = gnn_off_Wm(kGnnLayers);
constexpr int kGnnOffBo = kGnnOffWo + 4 * kGnnKM;
constexpr int kGnnOffWd = kGnnOffBo + 4;
constexpr int kGnnOffBd = kGnnOffWd + 3 * kGnnH;
constexpr int kGnnParams = kGnnOffBd + 3;

class Gnn : public Preconditioner {
public:
    Gnn() = default;
    ~Gnn() override;
    Gnn(const Gnn&) = delete;
    Gnn& operator=(const Gnn&) = delete;

    // Parameters and optimizer state (allocated once).
    int init(std::string& err);
    float* params() const { return dTheta; }
    float* grads() const { return dGrad; }

    // Edge maps of a full symmetric 2x2-block pattern (rows sorted, with
    // diagonal); allocates the activations. Returns 0 or -1.
    int set_pattern(hipStream_t s, int nV, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                    const std::vector<unsigned char>& fixedMask, std::string& err);
    bool has_pattern() const { return nV > 0; }
    int num_vertices() const { return nV; }
    int num_edges() const { return E; }

    // Features and forward pass for device block values (double, or float
    // when vals32) on the pattern: fills the factor G^.
    void forward(hipStream_t s, const void* vals, int vals32);
    // z = M^-1 r for FP64 vectors of 2 nV; per-workgroup records of r.z in
    // rzPart (rz_records() of them). Returns at once while *stop != 0.
    void apply(hipStream_t s, const double* r, double* z, double* rzPart, const int* stop) override;
    int rz_records() const override;

    // Training (after forward): adds dL/dtheta to grads(), *dLoss (device) += L.
    void zero_grad(hipStream_t s);
    void loss_grad(hipStream_t s, int nProbes, unsigned seed, double* dLoss);
    void adam(hipStream_t s, float lr, float b1, float b2, float eps, int t);

    // Factor values (device): lower edge blocks per edge slot (zero on
    // upper slots), diagonal blocks per vertex.
    const float* factor_lower() const { return dGo; }
    const float* factor_diag() const { return dGd; }
    // edge e -> position in the block CSR
    const int* edge_slots() const { return dESlot; }

private:
    int nV = 0, E = 0;
    // parameters
    float *dTheta = nullptr, *dGrad = nullptr, *dAdamM = nullptr, *dAdamV = nullptr;
    // pattern maps
    int *dRowE = nullptr, *dEu = nullptr, *dEw = nullptr, *dRev = nullptr, *dESlot = nullptr, *dDiagPos = nullptr;
    unsigned char* dFixed = nullptr;
    // features
    float *dCinv = nullptr, *dLogTr = nullptr, *dX = nullptr, *dAhat = nullptr, *dEf = nullptr;
    // activations
    float* dH[kGnnLayers + 1] = {};
    float* dA[kGnnLayers] = {};
    float* dR[kGnnLayers] = {};
    float* dM[kGnnLayers] = {};
    float *dGo = nullptr, *dGd = nullptr;
    // training
    double *dY = nullptr, *dT = nullptr, *dGs = nullptr, *dQ = nullptr, *dTmp = nullptr;
    float *dDGo = nullptr, *dDGd = nullptr, *dDh = nullptr, *dDa = nullptr, *dDXe = nullptr;
    float* dPart = nullptr;
    double* dPartD = nullptr;
    size_t nPart = 0;
    void release_pattern();
};

}  // namespace oc

// C-ABI handle (include/optcuts_gpu.h): a network on one device with its own stream.
struct oc_gnn {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    oc::Gnn net;
    // host-uploaded system (oc_gnn_set_system)
    void* dVals = nullptr;
    size_t valsBytes = 0;
    double* dLoss = nullptr;
    double *dR = nullptr, *dZ = nullptr, *dRz = nullptr;
    int* dStop = nullptr;
    std::vector<int> rowPtr, colIdx;
    std::vector<unsigned char> fixed;
};
