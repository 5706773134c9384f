// Interface of an external PCG preconditioner (multigrid, learned factor).
#pragma once

#include <hip/hip_runtime.h>

namespace oc {

struct Preconditioner {
    virtual ~Preconditioner() = default;
    // z = M^-1 r for FP64 vectors on the stream; writes rz_records()
    // per-workgroup partial sums of r.z into rzPart. Every kernel returns at
    // once while *stop is nonzero (a converged PCG); stop may be null.
    virtual void apply(hipStream_t s, const double* r, double* z, double* rzPart, const int* stop) = 0;
    virtual int rz_records() const = 0;
    // Changes whenever the device buffers apply() launches on may have
    // changed (a new set-up); launches captured into a graph stay valid while
    // it holds.
    unsigned long generation() const { return generation_; }

protected:
    unsigned long generation_ = 0;
};

}  // namespace oc
