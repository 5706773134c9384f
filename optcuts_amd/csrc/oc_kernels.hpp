// Launch wrappers of the geometry-step kernels (host-callable).
#pragma once

#include <hip/hip_runtime.h>

namespace oc {

// Triangle set of one system, SoA in HBM. Triangles [0, nFmesh) are mesh
// triangles (weight triArea/surfaceArea, scale sMesh), [nFmesh, nF) are air
// triangles (uniform weight 1, scale sAir).
struct TriSet {
    int nF = 0, nFmesh = 0;
    const int* f0 = nullptr;
    const int* f1 = nullptr;
    const int* f2 = nullptr;
    const double* triArea = nullptr;  // mesh: rest area (w = triArea / *surfaceArea)
    const double* A2 = nullptr;
    const double* E0 = nullptr;
    const double* E1 = nullptr;
    const double* E01 = nullptr;
    const double* surfaceArea = nullptr;  // device scalar (mesh block), or per-chart array with triChart
    const int* triChart = nullptr;        // batched charts: chart of each mesh triangle (w = area / its chart's area)
    int uniformMesh = 0;                  // mesh triangles use w = 1
    double sMesh = 1.0, sAir = 1.0;
};

// Per-triangle rest features (TriMesh.cpp:343-404). Vrest is nV x 3
// row-major. Triangles with rest area < areaThres get the equilateral
// clamp (:373-383). sums[0] += sum of areas, sums[1] += sum of half-edge
// lengths, sums[2] += count of zero-area triangles (compensated; reduced by
// a one-block finisher kernel).
void launch_rest_features(hipStream_t s, int nF, const int* f0, const int* f1, const int* f2, const double* Vrest,
                          double areaThres, double* triArea, double* A2, double* E0, double* E1, double* E01,
                          double* partials, double* sums);

// Element pass. Writes (when non-null): per-triangle energy Eelem[t] (w
// applied, block scale not applied), gradient contributions elemG[t*6..]
// and projected Hessian upper blocks elemH[t*24..] (both multiplied by w
// and the block scale). When esum is non-null, esum[0] = sum of mesh
// energies, esum[1] = sum of air energies (compensated). f32: the element
// terms are evaluated in FP32 (elemH then holds floats); sums stay FP64.
void launch_element(hipStream_t s, const TriSet& T, const double* UV, double* Eelem, double* elemG, void* elemH,
                    int project, double* partials, double* esum, int f32 = 0);

// Triplets in the reference order (SymDirichletEnergy.cpp:527-548 with
// IglUtils::addBlockToMatrix), unscaled: triangle t writes at tripOff[t];
// freeMask bit k of vFree[t] tells whether local vertex k is free.
void launch_triplets(hipStream_t s, const TriSet& T, const double* UV, const long long* tripOff,
                     const unsigned char* vFree, int* I, int* J, double* V);

// g[v] = sum of contributions (in triangle order) unless vertex v is fixed;
// gsq = ||g||^2 (last block).
void launch_gather_grad(hipStream_t s, int nV, const int* vcPtr, const int* vcSrc, const unsigned char* fixedMask,
                        const double* elemG, double* g, double* partials, double* gsq);

// out = total rest length of the cohesive edges that count as seam (cohE:
// nC x 4 ids, rows with a negative id skipped; tri-soup mode: only edges
// whose sides lie apart in UV), via per-block partials.
void launch_seam_len(hipStream_t s, int nC, const int* cohE, const double* Vrest, const double* UV, double avgEdgeLen,
                     int triSoup, double* partials, double* out);

// The merge query's per-pair screen (TriMesh::queryMerge, TriMesh.cpp:836-908) over nRows zipper-bottom pairs
// (merge vertices rA/rB, incident triangles tl[off[r]..off[r+1]) with the merge vertex's corner in cl): status 0
// candidate / 2 infeasible, merged position (2 doubles per pair).
void launch_merge_screen(hipStream_t s, int nRows, const int* rA, const int* rB, const int* off, const int* tl,
                         const int* cl, const int* F0, const int* F1, const int* F2, const double* UV, double eps,
                         int* status, double* merged);

// out[v] = sample standard deviation of the contributions of vertex v's
// triangles (0 for a vertex in one triangle): the split-query key.
void launch_div_grad(hipStream_t s, int nV, const int* vcPtr, const int* vcSrc, const double* elemG, double* out);

// Upper-block assembly into the full 2x2-block CSR; diagonal blocks also
// produce the Jacobi preconditioner (blockJacobi: inverse 2x2 block,
// else inverse diagonal entries). Sums are FP64; vals is float when vals32.
template <typename HT>
void launch_assemble(hipStream_t s, int nUB, const int* ubRow, const int* ubCol, const int* ubPosU,
                     const int* ubPosL, const int* ubCPtr, const int* ubCSrc, const unsigned char* fixedMask,
                     const double* fixedDiag, const HT* elemH, void* vals, int vals32, double* Minv,
                     int blockJacobi);

// Minimum over triangles of the injectivity step bound, folded into
// *ordMin (order-preserving uint64 of a double; min is order-independent);
// with ordMinAir, the minimum over the air triangles alone as well.
void launch_step_bound(hipStream_t s, const TriSet& T, const double* UV, const double* dir,
                       unsigned long long* ordMin, double cap, unsigned long long* ordMinAir = nullptr);

// Line-search energies at K candidate steps alpha_k = alpha0 / 2^(k0+k),
// k = 0..K-1: out[k] = sMesh*E_mesh + sAir*E_air, out[K+k] = E_mesh,
// out[2K+k] = E_air, out[3K+k] = number of triangles with negative
// doubled area (always FP64). f32: FP32 element energies, FP64 sums.
// boundOrd non-null: alpha0 = 0.99 x the step bound it holds (order-preserving uint64, as launch_step_bound
// leaves it), read on the device; k0 must then be 0.
constexpr int kLSCand = 8;
void launch_linesearch(hipStream_t s, const TriSet& T, const double* UV, const double* dir, double alpha0, int k0,
                       double* partials, double* out, int f32 = 0, const unsigned long long* boundOrd = nullptr);

// Outcome of the line-search rule applied on the device to one batch of candidates (k_ls_decide).
struct LsDecision {
    double alphaCommit;  // step the UV moved by (status 0 and 2)
    double alpha;        // accepted step (0 when stopped)
    double testE, testESD, testES;  // energies at the accepted candidate
    double lastE, lastESD, lastEScaf;  // the energy the candidates were compared with
    double bound;        // the injectivity bound (before the 0.99 factor)
    double airBound;     // its minimum over the air triangles alone (cap when none)
    double meshBound;    // and over the mesh triangles alone
    int halvings;
    int status;  // 0 accepted, 1 more than kLSCand halvings needed (nothing committed), 2 alpha underflowed to 0
    int pad[2];
};
// Applies Optimizer::lineSearch's rule (:597-629) to the candidates `cand` of a launch_linesearch from boundOrd
// and, unless more candidates are needed, commits UV += alphaCommit d (n doubles). lastE: lastEHost, or with
// esum non-null sMesh esum[0] + sAir esum[1]. airOrd (may be null): the air triangles' and the mesh triangles'
// bounds (two words), reported.
void launch_ls_decide_commit(hipStream_t s, const double* cand, const unsigned long long* boundOrd,
                             const unsigned long long* airOrd, const double* esum, double lastEHost, double sMesh,
                             double sAir, LsDecision* dec, long n, const double* d, double* UV);

// UV[i] = UV[i] + alpha * d[i] for i < n (two roundings, as the reference's
// stepForward, Optimizer.cpp:666-669).
void launch_axpy_commit(hipStream_t s, long n, double alpha, const double* d, double* UV);

// ---- PCG on the 2x2-block CSR ------------------------------------------------
constexpr int kSpmvLanes = 8;     // lanes per block-row in the SpMV
constexpr int kPcgLanes = 4;      // lanes per block-row in the PCG product
constexpr int kPcgRows = 2;       // block rows per lane group and pass in the PCG product
constexpr int kPcgMaxGrid = 1024;  // workgroups of the PCG kernels (bounds the partial records)

// The iteration number lives here, not in launch arguments, so that every
// iteration of one parity launches identical arguments (graph replay):
// k_pcg_init sets itNext = 0; k_pcg_spmv reads itNext and records it in itCur
// (and iter); k_pcg_update reads itCur and sets itNext = itCur + 1. Each field
// is written by one kernel and read only by the next.
struct PcgCtrl {
    double bb, tol2, rr, pad0;
    double rzSlot[2];
    int iter, maxit, done, itNext;
    int itCur, pad1, pad2, pad3;
};

// Workgroups of the update/init kernels and of the SpMV kernel for nV rows.
int pcg_grid_update(int nV);
int pcg_grid_spmv(int nV);

// partU: per-block (r.z, r.r) records written by init / update(i) (r.z only
// when Minv is given: the block-Jacobi M is applied inline); partS: per-block
// p.q records of spmv(i) read by update(i). spmv(i) sums r.z over nRz
// records of stride sRz and r.r over nRr records of stride sRr.
void launch_pcg_init(hipStream_t s, int nV, const double* b, const double* Minv, double* x, double* r, double* z,
                     double* partU, PcgCtrl* ctrl);
// vals: double4 blocks, or float4 blocks when vals32 (FP64 arithmetic either way).
void launch_pcg_spmv(hipStream_t s, int nV, const int* rowPtr, const int* colIdx, const void* vals,
                     int vals32, const double* z, const double* pOld, double* pNew, double* q, const double* partRz,
                     int nRz, int sRz, const double* partRr, int nRr, int sRr, double* partS, PcgCtrl* ctrl);
void launch_pcg_update(hipStream_t s, int nV, const double* pNew, const double* q, const double* Minv,
                       double* x, double* r, double* z, const double* partS, double* partUout, PcgCtrl* ctrl);

// y = A x (2x2-block CSR; float blocks when vals32).
void launch_spmv(hipStream_t s, int nV, const int* rowPtr, const int* colIdx, const void* vals, int vals32,
                 const double* x, double* y);

// Expand the reference's upper-triangle CSR values (per DOF) into the full
// block CSR: vals[dst[k]] = a[k] and, for off-diagonal-block or strictly
// lower-in-block entries, vals[mir[k]] = a[k].
void launch_expand_upper(hipStream_t s, int nnz, const double* a, const int* dst, const int* mir, double* vals);
void launch_jacobi_from_bsr(hipStream_t s, int nV, const int* diagPos, const double* vals, double* Minv,
                            int blockJacobi);

// ---- batched charts ----------------------------------------------------------
// Charts are contiguous vertex and triangle ranges of one combined mesh. Chart
// kernels use a chart-aligned block map: block b covers items
// [blkRange[b].x, blkRange[b].y) of chart blkChart[b] (at most kBlock items),
// the blocks of chart c are [chartBlk[c], chartBlk[c+1]).
void launch_chart_seg_sum(hipStream_t s, int nSeg, const int* segStart, const double* vals, double* out);
void launch_chart_gsq(hipStream_t s, int nCharts, const int* vStart, const double* g, double* out);
// out[c*2K + k] = E_mesh of chart c at alpha_c / 2^k, out[c*2K + K + k] = inverted triangles
void launch_chart_linesearch(hipStream_t s, const TriSet& T, const double* UV, const double* dir, int nBlk,
                             const int2* blkRange, const int* blkChart, const double* alphaChart, int nCharts,
                             const int* chartBlk, double* partials, double* out);
void launch_chart_step_bound(hipStream_t s, const TriSet& T, const double* UV, const double* dir, int nBlk,
                             const int2* blkRange, const int* blkChart, unsigned long long* ordMin, double cap);
// b[v] = active[chart(v)] ? -g[v] : 0 (per DOF pair)
void launch_chart_rhs(hipStream_t s, int nV, const int* vChart, const unsigned char* active, const double* g,
                      double* b);
// UV[v] += alpha[chart(v)] * d[v] (two roundings, as the single-chart commit)
void launch_chart_commit(hipStream_t s, int nV, const int* vChart, const double* alphaChart, const double* d,
                         double* UV);

}  // namespace oc

namespace oc {
// *count += number of triangles with negative doubled signed UV area.
void launch_count_inverted(hipStream_t s, const TriSet& T, const double* UV, unsigned int* count);
// *p = v (one thread; keeps device scalars out of hipMemset so streams stay capturable).
void launch_set_u64(hipStream_t s, unsigned long long* p, unsigned long long v);
// The 2x2-block matrix newVals (current pattern) mapped onto an earlier pattern of the same n rows (oldRowPtr;
// map: per earlier block the current block of the same (u, w) or -1; dropPtr/drop: per row the current blocks the
// earlier pattern lacks, whose Frobenius norms go onto both diagonal entries of the row's diagonal block, oldDiag).
void launch_remap_blocks(hipStream_t s, int n, const int* oldRowPtr, const int* map, const int* dropPtr,
                         const int* drop, const int* oldDiag, const double* newVals, double* oldVals);
// Iterative refinement: out = the FP32 rounding of the 2x2-block matrix in (nV block rows) with factor times each
// row's absolute sum added to its diagonal entries first; r = b - y (y null: r = b) over n doubles with
// per-workgroup sums of r.r in part (returns the record count, grid_for(n); r may be null).
void launch_round_blocks(hipStream_t s, int nV, const int* rowPtr, const int* diagPos, const double* in, float* out,
                         double factor);
int launch_residual_norm(hipStream_t s, long n, const double* b, const double* y, double* r, double* part);
// b = sign * g for n doubles.
void launch_scale_copy(hipStream_t s, long n, double sign, const double* g, double* b);
// For k < n, p = UV row ids[k] (row k when ids is null): out3[k] = (p.x, p.y, 0) and out2[k] = p (either may be
// null). Turns UV rows into the 3-vector rest positions of a planar mesh, or copies out a subset of rows.
void launch_uv_gather(hipStream_t s, int n, const int* ids, const double* UV, double* out3, double* out2);
}  // namespace oc
