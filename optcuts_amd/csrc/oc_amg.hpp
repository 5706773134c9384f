// Smoothed-aggregation algebraic multigrid preconditioner for the Newton
// systems of the geometry step (the chosen solve form).
//
// Method (Vanek, Mandel & Brezina, "Algebraic multigrid by smoothed
// aggregation for second and fourth order elliptic problems", 1996):
//   - aggregates: greedy aggregation of the vertex graph, once per sparsity
//     pattern (pattern only, so it is reused across Newton iterations);
//   - near-null space: the rigid motions of the UV plane per vertex,
//     [[1, 0, -(v - v_a)], [0, 1, (u - u_a)]] about the aggregate centroid,
//     orthonormalised per aggregate (tentative prolongator) with the R
//     factors as the coarse near-null space;
//   - prolongator P = (I - 4/(3 lambda) D^-1 A) P_tent with lambda the
//     largest eigenvalue of D^-1 A (Lanczos in the D inner product, D = block
//     diagonal);
//   - Galerkin coarse operators P^T A P;
//   - V-cycle with Chebyshev (degree 2) smoothing on D^-1 A over
//     [lambda/30, 1.1 lambda] and an explicit inverse on the coarsest level.
// Every numeric step runs on the GPU from index maps built on the host once
// per pattern, so a Newton iteration re-derives the whole hierarchy without a
// host round trip. Values are stored in T (float for the chosen form); the
// coarsest inverse and the QR factorisations are computed in double.
#pragma once

#include <hip/hip_runtime.h>

#include "oc_amg_structure.hpp"
#include "oc_precond.hpp"

#include <string>
#include <map>
#include <vector>

namespace oc {

// Precision-independent interface of the hierarchy (the Newton driver holds
// one of the instantiations below).
class AmgBase : public Preconditioner {
public:
    ~AmgBase() override = default;
    // late (may be null): nodes aggregated after all others (see aggregate_nodes), e.g. the rows an air-mesh
    // change can touch, so that the rest of the hierarchy keeps its numbering across such changes
    virtual int setup(hipStream_t s, int nV, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                      const std::vector<unsigned char>& fixedMask, std::string& err,
                      const std::vector<unsigned char>* late = nullptr) = 0;
    // safe != 0: smoother interval and prolongator damping from the
    // Gershgorin bound of D^-1 A (guaranteed, slower convergence) instead of
    // the Lanczos estimate.
    virtual void numeric(hipStream_t s, const void* A0vals, int vals32, const double* UV, int safe = 0) = 0;
    virtual int levels() const = 0;
    virtual long level_rows(int l) const = 0;
    virtual long level_blocks(int l) const = 0;
    // Prepares the level-0 Lanczos recurrence (start and two steps) of the
    // current hierarchy; after it, lanczos0_step() launches one further
    // level-0 step that reads and writes the same buffers every time (for
    // timing). Call after numeric().
    virtual void lanczos0_prepare(hipStream_t s) = 0;
    virtual void lanczos0_step(hipStream_t s) = 0;
    // One V-cycle Chebyshev step on level l (a level with a transfer), on that level's own vectors, as the
    // V-cycle launches it; repeatable (for timing). Call after numeric().
    virtual void cheb_step_level(hipStream_t s, int l) = 0;
    virtual int level_block_size(int l) const = 0;
    virtual bool ready() const = 0;
    // Largest last smoothed level (nodes) whose part of the V-cycle runs as one workgroup launch (0: never);
    // takes effect at the next setup().
    virtual void set_tail_max(int nodes) = 0;
};

template <typename T>
class Amg : public AmgBase {
public:
    Amg() = default;
    ~Amg() override;
    Amg(const Amg&) = delete;
    Amg& operator=(const Amg&) = delete;

    // Structure of the hierarchy for the 2x2-block pattern (rowPtr/colIdx
    // over nV vertices, sorted rows with diagonals). Fixed vertices take no
    // part in aggregation. Allocates all device storage. Returns 0 or -1
    // (message in err).
    int setup(hipStream_t s, int nV, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
              const std::vector<unsigned char>& fixedMask, std::string& err,
              const std::vector<unsigned char>* late = nullptr) override;

    // Numeric hierarchy from the fine-level block values (same layout as the
    // pattern; double, or float when vals32) and the current UV (for the
    // rigid near-null space).
    void numeric(hipStream_t s, const void* A0vals, int vals32, const double* UV, int safe = 0) override;

    // z = M r (one V-cycle) for FP64 vectors of 2*nV; writes per-workgroup
    // partial records of r.z into rzPart (rz_records() of them). Every
    // kernel returns at once while *stop is nonzero (a converged PCG).
    void apply(hipStream_t s, const double* r, double* z, double* rzPart, const int* stop) override;

    int rz_records() const override;
    int levels() const override { return (int)L.size(); }
    long level_rows(int l) const override;
    long level_blocks(int l) const override;
    void lanczos0_prepare(hipStream_t s) override;
    void lanczos0_step(hipStream_t s) override;
    void cheb_step_level(hipStream_t s, int l) override;
    int level_block_size(int l) const override;
    bool ready() const override { return ready_; }
    void set_tail_max(int nodes) override { tailPending_ = nodes; }

private:
    struct Level;
    std::vector<Level*> L;
    int nCoarse = 0;           // DOFs of the coarsest level
    bool coarseDiag = false;   // coarsest level block diagonal: solved by its diagonal-block inverses
    double* dCoarseInv = nullptr;  // nCoarse x nCoarse, double
    double* dPartials = nullptr;
    int nPartials = 0;
    // Device buffers of released hierarchies, reused by the next set-up (a new
    // pattern of similar size, e.g. after the caller replaces the air mesh,
    // then needs no hipFree/hipMalloc).
    struct Pool {
        std::multimap<size_t, void*> free;
        std::map<void*, size_t> size;
        void* get(size_t bytes);
        void put(void* p);
        void trim();   // free the buffers nobody took back
        void clear();  // free everything
    } pool_;
    // Host storage of the hierarchy before the previous one: the next
    // structure build writes into it instead of fresh allocations.
    std::vector<AmgHostLevel> spare_;
    bool ready_ = false;
    int tailMax_ = -1, tailPending_ = -1;  // set_tail_max (-1: the built-in limit), active since the last setup()
    void release();
    void vcycle(hipStream_t s, int l, const int* stop);
    void vcycle_tail(hipStream_t s, int l, const int* stop);
    void restrict_to(hipStream_t s, int l, const int* stop);
    void coarse_solve(hipStream_t s, const int* stop);
};

extern template class Amg<float>;
extern template class Amg<double>;

}  // namespace oc
