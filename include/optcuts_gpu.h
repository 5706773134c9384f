/*
 * optcuts_gpu.h — C ABI of the MI355X (gfx950) geometry-step library.
 *
 * One context (oc_ctx) owns a HIP stream and the device state of one
 * parameterization problem: a mesh ("block 0", the TriMesh handed to
 * SymDirichletEnergy), an optional bijectivity air mesh ("block 1", the
 * Scaffold's airMesh), the assembled Hessian and the solver workspace.
 * All pointers in this header are HOST pointers to plain arrays; the library
 * copies to and from HBM itself. The set-up calls (oc_set_mesh, oc_set_air_mesh,
 * oc_newton_init, oc_solver_set_pattern, oc_solver_factorize) size the device
 * buffers; later calls reuse them and allocate only when a size grows (an air
 * re-mesh with more blocks, a multigrid rebuild with larger levels, the first
 * refinement pass or IC(0) fallback). oc_energy_noalloc never allocates.
 * Every function returns an OC_* status; nothing throws.
 *
 * Array conventions (row-major, C order):
 *   Vrest  nV x 3 doubles       F      nF x 3 int32 (counter-clockwise in UV)
 *   UV     nV x 2 doubles       DOF vectors (gradient, directions) 2*nV
 *   doubles interleaved [u0, v0, u1, v1, ...] (the reference's order,
 *   SymDirichletEnergy.cpp:287,301).
 * Reference locations are relative to the OptCuts source tree.
 */
#ifndef OPTCUTS_GPU_H
#define OPTCUTS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OC_OK = 0,
    OC_ERR_ARG = -1,            /* bad argument or size */
    OC_ERR_HIP = -2,            /* HIP runtime error (message in oc_last_error) */
    OC_ERR_STATE = -3,          /* call out of order (e.g. no mesh set) */
    OC_ERR_PATTERN = -4,        /* triplet (i<=j) outside the sparsity pattern */
    OC_ERR_INVERTED = -5,       /* a UV triangle has negative signed area */
    OC_ERR_DEGENERATE = -6,     /* a rest triangle has zero area */
    OC_ERR_NOT_CONVERGED = -7,  /* linear solve stopped before its tolerance */
    OC_ERR_BREAKDOWN = -8,      /* PCG breakdown: matrix not positive definite */
    OC_ERR_NO_DEVICE = -9       /* no usable gfx950 device */
};

/* Newton-system preconditioners: point Jacobi (the plain form), 2x2 block Jacobi, smoothed-aggregation
 * algebraic multigrid under FP64 PCG (needs the UV map for its rigid near-null space), the learned sparse
 * factor of an oc_gnn network (oc_newton_set_learned), and block incomplete Cholesky IC(0) on the 2x2-block
 * pattern (factored by fixed-point sweeps, applied by Jacobi-swept triangular solves; oc_ic0_set_sweeps). LEARNED
 * is available to the device-resident Newton driver only; the others to the drop-in oc_solver_* as well. None has
 * a reference counterpart: the reference factorizes every system with SimplicialLDLT (EigenLibSolver.cpp:71-108).
 * IC(0)-PCG is also the learned form's fallback before Jacobi-PCG (pcg_fallback 7). */
enum { OC_PRECOND_JACOBI = 0, OC_PRECOND_BLOCK_JACOBI = 1, OC_PRECOND_AMG = 2, OC_PRECOND_LEARNED = 3,
       OC_PRECOND_IC0 = 4 };

/* Arithmetic of the device-resident Newton driver.
 * FP64: every element term, the Hessian values and the PCG in FP64 (the plain form; matches the reference's
 *   FP64 run).
 * MIXED: element energy, gradient and projected Hessian in FP32 (on vertex positions relative to each
 *   triangle's first vertex), energy and gradient sums in compensated FP64, the multigrid hierarchy in FP32;
 *   the assembled Hessian (summed in FP64 from the FP32 element blocks), the injectivity step-bound roots, the
 *   inversion test, the PCG matrix products, vectors, residual recurrence and scalars stay FP64. The FP32
 *   element Hessians carry rounding noise of ~1e-7 of each block; Newton matrices of distorted maps reach
 *   condition numbers near 1e15, so that noise perturbs their smooth spectrum and can make them indefinite:
 *   long Krylov runs may then break down (the driver falls back as pcg_fallback reports). */
enum { OC_PRECISION_FP64 = 0, OC_PRECISION_MIXED = 1 };

typedef struct oc_ctx oc_ctx;
typedef struct oc_gnn oc_gnn;

/* ---- context ------------------------------------------------------------ */
int oc_ctx_create(int device, oc_ctx** out);
int oc_ctx_destroy(oc_ctx* ctx);
/* Last error message of this context (static storage, valid until the next call). */
const char* oc_last_error(const oc_ctx* ctx);
/* Name of the device the context runs on, e.g. "gfx950". */
int oc_device_arch(const oc_ctx* ctx, char* buf, int buflen);

/* ---- mesh (TriMesh data of SymDirichletEnergy) -------------------------- */
/* Replaces the TriMesh fields the energy reads: V_rest, F, fixedVert and the
 * rest features computed by TriMesh::computeFeatures (TriMesh.cpp:323-404;
 * areaThres is TriMesh::areaThres_AM, 0 for a surface mesh).
 * fixedV: vertex ids, any order (the reference fixes vertex 0,
 * TriMesh.cpp:325-328). Returns OC_ERR_DEGENERATE on a zero-area triangle
 * (TriMesh.cpp:368-372,399-402). */
int oc_set_mesh(oc_ctx* ctx, int nV, const double* Vrest, int nF, const int* F, int nFixed, const int* fixedV,
                double areaThres);
/* TriMesh::V (the UV coordinates), nV x 2. */
int oc_set_uv(oc_ctx* ctx, const double* UV);
int oc_get_uv(oc_ctx* ctx, double* UV);
/* surfaceArea and avgEdgeLen as TriMesh::computeFeatures computes them (TriMesh.cpp:385,403). */
int oc_mesh_stats(oc_ctx* ctx, double* surfaceArea, double* avgEdgeLen);

/* ---- Energy interface (Energy.hpp:16-46, SymDirichletEnergy.hpp:15-42) ---- */
/* Energy::computeEnergyVal (Energy.cpp:35-40): sum of per-element SD. */
int oc_energy(oc_ctx* ctx, int uniformWeight, double* E);
/* oc_energy without any device allocation, for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). No reference counterpart. The same kernels and result as
 * oc_energy, with the reduction's workgroup records in a fixed device array of the library (2 MiB per device, held
 * until the process ends) instead of the context's scratch: meshes of at most 16,777,216 triangles (OC_ERR_ARG
 * beyond). Every call on a device shares that array, so calls that overlap on one device (contexts on other
 * streams or threads) race. */
int oc_energy_noalloc(oc_ctx* ctx, int uniformWeight, double* E);
/* SymDirichletEnergy::getEnergyValPerElem (SymDirichletEnergy.cpp:24-46), nF values. */
int oc_energy_per_elem(oc_ctx* ctx, int uniformWeight, double* Eelem);
/* SymDirichletEnergy::computeGradient (SymDirichletEnergy.cpp:258-304), 2*nV values, fixed rows 0. */
int oc_gradient(oc_ctx* ctx, int uniformWeight, double* g);
/* SymDirichletEnergy::computeHessian(data, V, I, J) (SymDirichletEnergy.cpp:429-549): the PSD-projected
 * element Hessians as triplets in the reference's order (IglUtils::addBlockToMatrix, IglUtils.cpp:396-451,
 * then the identity on fixed DOFs, IglUtils.cpp:361-394). With I == NULL only *count is set. capacity is the
 * length of I/J/V. */
int oc_hessian_triplets(oc_ctx* ctx, int uniformWeight, long long capacity, int* I, int* J, double* V,
                        long long* count);
/* The per-triangle terms of SymDirichletEnergy (SymDirichletEnergy.cpp:24-46, 258-304, 438-523 with
 * IglUtils::makePD) as the device Newton driver's element pass computes them in the given precision
 * (OC_PRECISION_FP64, or OC_PRECISION_MIXED: FP32 element arithmetic on positions relative to each triangle's first
 * vertex), weights triArea / surfaceArea: E nF energies, G nF x 6 gradient rows [u0 v0 u1 v1 u2 v2], H nF x 36
 * projected Hessians (row-major 6x6). Any pointer may be null. No reference counterpart: it exposes the element
 * arithmetic of both precisions for error measurement against an FP64 evaluation. */
int oc_element_terms(oc_ctx* ctx, int precision, double* E, double* G, double* H);
/* SymDirichletEnergy::initStepSize (SymDirichletEnergy.cpp:551-610): *stepSize = min(*stepSize, first
 * positive root of det(U + a dir) over the triangles); dir has 2*nV values. */
int oc_init_step_size(oc_ctx* ctx, const double* dir, double* stepSize);
/* TriMesh::checkInversion (TriMesh.cpp:1736-1756): *ok = 1 if no triangle has negative signed UV area. */
int oc_check_inversion(oc_ctx* ctx, int* ok);

/* ---- seam length ----------------------------------------------------------- */
/* TriMesh::computeSeamSparsity (TriMesh.cpp:1542-1558) at the context's UV, the E_se column of the energy trace
 * (Optimizer.cpp:706-728): the sum of the rest lengths (edgeLen, TriMesh.cpp:340) of the cohesive edges cohE
 * (nCoh x 4 vertex ids; a row with a negative id is a boundary edge and is skipped, TriMesh.cpp:334-339), in
 * tri-soup mode (triSoup != 0) only of those with a vertex pair (cohE[0], cohE[2]) or (cohE[1], cohE[3]) more
 * than 1e-2 avgEdgeLen apart in UV, plus initSeamLen. */
int oc_seam_sparsity(oc_ctx* ctx, int nCoh, const int* cohE, double initSeamLen, int triSoup, double* sparsity);

/* ---- split-query pre-screen (topology step) ------------------------------ */
/* SymDirichletEnergy::computeDivGradPerVert (SymDirichletEnergy.cpp:108-154, the STANDARD_DEVIATION_FILTERING
 * branch it compiles) at the context's UV: per mesh vertex the sample standard deviation of its triangles'
 * local gradient rows (computeLocalGradient, :215-256, weights triArea / surfaceArea), 0 for a vertex in one
 * triangle. out: nV doubles. */
int oc_div_grad_per_vert(oc_ctx* ctx, double* out);
/* The candidate filter of TriMesh::querySplit (TriMesh.cpp:556-633) for propagate = false, which picks the
 * vertices whose splits the topology step then evaluates (computeLocalLDec): vertices with more than two
 * neighbours that are on the boundary (splitInterior = 0) or neither on nor next to it (splitInterior = 1),
 * keyed by -divGrad / vertWeight (vertWeight null: all 1, TriMesh.cpp:203) in a std::map (a later vertex with
 * an equal key replaces the earlier), the first max(2, (int)(size^e)) keys in order, e = 0.8 for boundary
 * splits and filterExpIn (the reference's filterExp_in, main.cpp:70) for interior ones. cand receives up to
 * capacity of the *nCand candidates in that order; *nSorted = the map's size (inSplitTotalAmt for interior
 * splits). The keys come from the GPU; the selection runs on the host. */
int oc_split_candidates(oc_ctx* ctx, int splitInterior, double filterExpIn, const double* vertWeight, int capacity,
                        int* cand, int* nCand, int* nSorted);
/* The per-pair screen of TriMesh::queryMerge (TriMesh.cpp:801-908) for propagate = false, before each pair's
 * computeLocalLDec, at the context's UV over the cohesive edges cohE (nCoh x 4, [a' b' a'' b''], mesh ids):
 * a row is a zipper bottom when a' == a'' or b' == b''; its two other vertices move to their midpoint, which is
 * projected by the reference's relaxation method (70 sweeps, eps = 1e-6 avgEdgeLen) when it inverts one of their
 * triangles. status[r]: 0 candidate (mergedPos[2r..2r+1] = the position computeLocalLDec starts from), 1 not a
 * zipper bottom, 2 no feasible position (skipped by the reference). *nCand = rows with status 0. Positions and
 * decisions are bit-identical to the reference's arithmetic. */
int oc_merge_candidates(oc_ctx* ctx, int nCoh, const int* cohE, int* status, double* mergedPos, int* nCand);

/* ---- bijectivity scaffold (Scaffold.cpp:27-313) ------------------------- */
/* The air mesh built by the Scaffold constructor: nVair vertices with UV airUV (nVair x 2; its rest shape is
 * this UV with z = 0, Scaffold.cpp:174-177), nFair triangles Fair in air-mesh indices, the first nBnd air
 * vertices are the mesh vertices bnd[0..nBnd) (Scaffold::bnd, localVI2Global for l < nBnd) and the others are
 * appended after the mesh vertices (Scaffold.cpp:179-184). areaThres is airMesh.areaThres_AM. Air vertices that
 * are fixed mesh vertices are fixed (Scaffold.cpp:186-192). wScaf is the scaffold weight w_scaf
 * (Optimizer.cpp:87); each air energy is scaled by wScaf / nFair (Optimizer.cpp:776,795,839). Pass nFair = 0 to
 * drop the scaffold. The next oc_newton_init builds the new pattern; on the same mesh it copies the pattern,
 * slot-map and multigrid rows of mesh vertices outside the old and new air triangles from the previous build
 * (identical result, a fraction of the time). */
int oc_set_air_mesh(oc_ctx* ctx, int nVair, const double* airUV, int nFair, const int* Fair, int nBnd,
                    const int* bnd, double areaThres, double wScaf);
/* The air mesh between Newton steps, in place of the reference's new Triangle air mesh after every step
 * (Optimizer.cpp:251-257, Scaffold.cpp:27-176). The air vertices have moved with the steps (the step bound keeps
 * the air triangles positively oriented). Every mode resets the air mesh's rest shape to its current UV (z = 0,
 * areaThres_AM clamp as set), as a freshly built air mesh has, and evaluates *E = a0 E_SD + scaffold term.
 *   OC_SCAFFOLD_REFRESH: the triangulation is kept; the Newton system's pattern is unchanged.
 *   OC_SCAFFOLD_REMESH_WHEN_NEEDED: when the air mesh held the last step back (the air triangles' injectivity bound
 *     was below half of what the mesh triangles alone allowed, capped at a full step), when maxInterval > 0 calls
 *     have passed since the last re-mesh, or when minAngleDeg > 0
 *     and the smallest angle of the air triangles is below it, the air mesh is first re-meshed by Lawson edge flips
 *     (the Delaunay triangulation of the current air vertices constrained by the map's boundary and the outer
 *     box) and the system re-patterned (only air rows change).
 *   OC_SCAFFOLD_REMESH: flips at every call.
 * *flips (may be null) = edges flipped (0: pattern unchanged). Needs oc_newton_init with an air mesh. Writes no
 * trace line and keeps the state oc_newton_rebuild keeps. */
enum { OC_SCAFFOLD_REFRESH = 0, OC_SCAFFOLD_REMESH_WHEN_NEEDED = 1, OC_SCAFFOLD_REMESH = 2 };
int oc_scaffold_update(oc_ctx* ctx, int mode, double minAngleDeg, int maxInterval, double* E, int* flips);
/* The current air mesh: airUV nVair x 2 (air-mesh numbering: the boundary vertices first), Fair nFair x 3 (air
 * ids, after any flips). Either pointer may be null. */
int oc_air_mesh_get(oc_ctx* ctx, double* airUV, int* Fair);
/* Host wall times (ms) of the last Newton-system set-up (oc_newton_init, oc_newton_rebuild or a re-meshing
 * oc_scaffold_update): ms[0] vertex lists, [1] pattern and slot maps, [2] their upload, [3] multigrid structure
 * (or, after a re-mesh that kept it, the map of the Newton pattern onto it), [4] learned network's graph, [5] the
 * whole set-up; and of the last oc_scaffold_update: [6] air-vertex download and edge flips (0 for a refresh), [7]
 * the whole update; ms[8] = air re-meshes the current multigrid structure has been kept across (0: built for the
 * current Newton pattern; at most 4, then it is rebuilt). ms holds 9 doubles. No reference counterpart
 * (diagnostics). */
int oc_setup_times(oc_ctx* ctx, double* ms);

/* ---- LinSysSolver interface (LinSysSolver.hpp:22-256) -------------------- */
/* LinSysSolver::set_pattern(vNeighbor, fixedVert) (LinSysSolver.hpp:37-135) + EigenLibSolver::set_pattern
 * (EigenLibSolver.cpp:21-39): nbrPtr/nbr is vNeighbor as CSR over nV vertices. Builds the upper-triangle CSR
 * over 2*nV DOFs (0-based, the reference's ia/ja) and the device block-CSR it maps to. *nnz = ja length. */
int oc_solver_set_pattern(oc_ctx* ctx, int nV, const int* nbrPtr, const int* nbr, int nFixed, const int* fixedV,
                          int* nnz);
int oc_solver_get_pattern(oc_ctx* ctx, int* ia, int* ja);
/* LinSysSolver::update_a (LinSysSolver.hpp:138-159): a = 0; a[slot(i,j)] += v for triplets with i <= j,
 * in triplet order. OC_ERR_PATTERN if such a triplet has no slot. */
int oc_solver_update_a(oc_ctx* ctx, long long nT, const int* I, const int* J, const double* V);
int oc_solver_get_a(oc_ctx* ctx, double* a);
/* Set the upper-triangle CSR values directly (nnz = ja length), as EigenLibSolver::update_a copies LinSysSolver::a
 * into its matrix (EigenLibSolver.cpp:62-67); for callers that accumulate a themselves. */
int oc_solver_set_a(oc_ctx* ctx, const double* a);
/* analyze_pattern / factorize (EigenLibSolver.cpp:71-94): no symbolic phase; factorize builds the
 * preconditioner of the selected kind from the current values. */
int oc_solver_analyze_pattern(oc_ctx* ctx);
int oc_solver_factorize(oc_ctx* ctx);
/* EigenLibSolver::solve (EigenLibSolver.cpp:96-108): x = A^-1 rhs by PCG to relTol. OC_ERR_NOT_CONVERGED when
 * maxIter is reached first (x holds the last iterate). */
int oc_solver_solve(oc_ctx* ctx, const double* rhs, double* x);
/* LinSysSolver::multiply (LinSysSolver.hpp:172-187): Ax = A x with the full symmetric matrix. */
int oc_solver_multiply(oc_ctx* ctx, const double* x, double* Ax);
/* The solve's parameters (EigenLibSolver has none: its factorization is exact): OC_PRECOND_JACOBI,
 * OC_PRECOND_BLOCK_JACOBI, OC_PRECOND_AMG (smoothed-aggregation multigrid built by factorize; a breakdown
 * rebuilds it with the Gershgorin interval, then falls back to Jacobi-PCG) or OC_PRECOND_IC0 (block IC(0) built
 * by factorize; a breakdown falls back to Jacobi-PCG), relative residual relTol, at most maxIter PCG
 * iterations. */
int oc_solver_set_params(oc_ctx* ctx, int precond, double relTol, int maxIter);
/* z = M^-1 r (2*nV doubles each) with the preconditioner of the last factorize (OC_PRECOND_AMG or
 * OC_PRECOND_IC0). No reference counterpart (diagnostics and tests). */
int oc_solver_apply_precond(oc_ctx* ctx, const double* r, double* z);
/* Sweep counts of OC_PRECOND_IC0 (the Newton driver's and the drop-in's): factorSweeps fixed-point sweeps after
 * the starting factor (default 3; the exact IC(0) factor once they reach the dependency depth) and applySweeps
 * Jacobi sweeps per triangular solve after the diagonal one (default 2; exact triangular solves at the depth).
 * Takes effect at the next factorization (oc_solver_factorize, or the next Newton step). No reference
 * counterpart. */
int oc_ic0_set_sweeps(oc_ctx* ctx, int factorSweeps, int applySweeps);
/* Largest last smoothed multigrid level (nodes) whose part of the V-cycle (its smoothing, residual, transfers and
 * the coarsest solve) runs as one single-workgroup launch instead of seven launches; the result is bit-identical
 * either way. 0 turns it off, -1 restores the built-in limit (128). Takes effect when a multigrid structure is
 * next built (oc_newton_init, a re-mesh that rebuilds it, oc_solver_set_pattern + oc_solver_factorize). No
 * reference counterpart (diagnostics and tests). */
int oc_amg_set_tail(oc_ctx* ctx, int maxNodes);
/* Coordinates (2*nV, [u0 v0 u1 v1 ...], e.g. the current UV of TriMesh::V) for OC_PRECOND_AMG's near-null
 * space (translations and the rotation of the plane per aggregate); without them the multigrid uses
 * translations only. Takes effect at the next factorize. No reference counterpart: LinSysSolver passes no
 * geometry. */
int oc_solver_set_coords(oc_ctx* ctx, const double* UV);
int oc_solver_last_stats(oc_ctx* ctx, int* iters, double* relResidual);
/* What the last oc_solver_solve fell back to: 0 none, 1 multigrid rebuilt with the Gershgorin interval,
 * 2 Jacobi-PCG. No reference counterpart. */
int oc_solver_last_fallback(oc_ctx* ctx, int* fallback);

/* ---- device-resident Newton driver (Optimizer.cpp:154-261,505-704) ------- */
typedef struct {
    double E_total;   /* lastEnergyVal after the step: a0*E_SD + scaffold term */
    double E_SD;      /* mesh SD energy (energyVal_ET[0]) */
    double E_scaf;    /* scaffold term (energyVal_scaffold) */
    double grad_sq;   /* ||g||^2 of the combined gradient at the start of the iteration */
    double step;      /* accepted step size */
    int halvings;     /* step halvings taken in the line search */
    int stopped;      /* 1 if the line search or the relative-decrease test stopped the solve */
    int converged;    /* 1 if ||g||^2 < targetGRes (no step taken) */
    int pcg_iters;    /* PCG iterations of this Newton step */
    double pcg_rel_res;
    int pcg_fallback; /* 0: the selected preconditioner solved the step. Otherwise its first attempt broke down
                         (r.z or p.Ap not positive) or hit pcgMaxIter, and 4 (OC_PRECOND_LEARNED with
                         OC_PRECISION_MIXED): the matrix rebuilt from FP64 element terms with a newly predicted
                         factor solved it (the Newton system's element terms then stay FP64 until the next
                         oc_newton_init; line-search energies stay FP32); 1: multigrid
                         rebuilt with its guaranteed smoother interval solved it; 2: Jacobi-PCG solved it; 3:
                         Jacobi-PCG (the plain form's own solve, or the fallback) met p.Ap <= 0 (a singular Newton
                         matrix, as for one triangle free to rotate about its fixed vertex, or FP32 element
                         Hessians that leave it slightly indefinite) and the step uses its last iterate, or -g
                         when it had none (truncated Newton); 5: Jacobi-PCG was used directly because
                         multigrid failed (code 2) within the last 3 steps (it is tried again after them); 6:
                         multigrid-PCG reached pcgMaxIter without a breakdown and the step uses its last iterate
                         (a descent direction: it minimises the quadratic model over the Krylov space so far);
                         7: OC_PRECOND_LEARNED failed and block IC(0)-PCG at the same tolerance solved it */
    int pcg_iters_first; /* iterations of the first attempt when pcg_fallback != 0 (pcg_iters counts all) */
    int precond_setups;  /* multigrid hierarchies built on this step (1, or 2 after a safe-interval retry) */
    int ir_passes;       /* passes of the iterative refinement (oc_newton_set_refinement), 0 when it is off */
} oc_iter_stats;

/* Optimizer::precompute (Optimizer.cpp:154-201) for energyParams = {a0}: sets targetGRes =
 * a0 (nV-nFixed)/nV relGL2Tol (Optimizer.cpp:675-678), builds the combined pattern (mesh + air mesh) and
 * evaluates the initial energy (*E0 = a0 E_SD + scaffold term). allowEDecRelTol: Optimizer::setAllowEDecRelTol.
 * pcgRelTol/pcgMaxIter/precond select the Newton-system solve. */
int oc_newton_init(oc_ctx* ctx, double a0, double relGL2Tol, int allowEDecRelTol, int precond, double pcgRelTol,
                   int pcgMaxIter, double* E0);
/* The scaffold rebuild inside Optimizer::solve (Optimizer.cpp:251-257, re-patterned at :522-530): after
 * oc_set_air_mesh replaced the air mesh of an initialised problem, rebuilds the Newton system with the parameters
 * of the last oc_newton_init and evaluates *E = a0 E_SD + scaffold term. Unlike oc_newton_init it writes no trace
 * line and keeps the solver state that carries across iterations (the previous solve's iteration count, the
 * mixed learned form's switch to FP64 element terms). */
int oc_newton_rebuild(oc_ctx* ctx, double* E);
/* PCG iterations of the Newton solve (and of the other solves through this context) as replays of captured HIP
 * graphs (enable != 0, the default: an even iteration, an odd one and an even+odd pair are captured once per
 * set of buffers) or launched kernel by kernel (0). The results are identical either way. No reference
 * counterpart. */
int oc_pcg_graphs(oc_ctx* ctx, int enable);
/* Selects OC_PRECISION_FP64 (default) or OC_PRECISION_MIXED for the next oc_newton_init. No reference
 * counterpart (the reference is FP64 throughout). */
int oc_newton_set_precision(oc_ctx* ctx, int precision);
/* Mixed-precision iterative refinement of the Newton solve (maxPasses > 0; 0, the default, turns it off): the PCG
 * iterations multiply by the FP32 rounding of the FP64 Newton matrix (half the matrix traffic of each product)
 * while the residual and the solution stay FP64: x = 0, r = b; each pass runs PCG on A32 d = r to
 * min(0.5, pcgRelTol |b| / |r|), then x += d and r = b - A x with the FP64 matrix, until |r| <= pcgRelTol |b|, maxPasses
 * passes, pcgMaxIter iterations in all, or a pass that does not halve |r| (the rounding can perturb the matrix's
 * near-null directions, the rigid rotation about a single fixed vertex, beyond what refinement corrects; the solve
 * then ends as one at its iteration cap). Breakdowns of a pass keep the iterate it reached; the fallbacks of
 * pcg_fallback apply to the refined solve as to a plain one. pcg_rel_res then reports the FP64 residual and
 * ir_passes the passes. Works with every precision and preconditioner; takes effect at the next step. No
 * reference counterpart. */
int oc_newton_set_refinement(oc_ctx* ctx, int maxPasses);
/* Step-dependent PCG tolerance (looseTol > 0; 0, the default, turns it off): after a Newton step shorter than
 * stepBelow (the injectivity bound cut it short, so the direction's accuracy goes mostly unused), and for the
 * first step after oc_newton_init, the next solve runs to max(pcgRelTol, looseTol); otherwise to pcgRelTol.
 * Takes effect at the next step. No reference counterpart (the reference solves exactly). */
int oc_newton_set_forcing(oc_ctx* ctx, double looseTol, double stepBelow);
/* Binds the network used by OC_PRECOND_LEARNED from the next oc_newton_init on (null unbinds). The context
 * does not own it; while bound, the network's pattern and activations belong to the context's Newton system,
 * and it must not be used from another context or through oc_gnn_set_system. No reference counterpart. */
int oc_newton_set_learned(oc_ctx* ctx, oc_gnn* net);
/* One Optimizer::solve(1) iteration (Optimizer.cpp:206-231 with solve_oneStep :505-573 and lineSearch
 * :575-652): gradient and convergence test, Hessian assembly, PCG direction, injectivity bound, line search,
 * step. *done = 1 when the solve returned 1 (gradient tolerance reached or line search stopped). */
int oc_newton_step(oc_ctx* ctx, oc_iter_stats* stats, int* done);

/* ---- initial map ------------------------------------------------------------------------------------
 * Tutte's embedding as the reference initialises UV (main.cpp:1313-1364 and 1380-1395, 1486-1532). The
 * pinned loops are
 *   OC_TUTTE_COMPONENTS (mesh without input UV): per edge-connected triangle component (igl::facet_components,
 *     numbered by lowest triangle), the first longest loop of igl::boundary_loop over its triangles;
 *   OC_TUTTE_ALL_LOOPS (input UV re-initialised): every loop of igl::boundary_loop over the whole mesh.
 * Loop k is placed on the unit circle by arc length of the rest positions (igl::map_vertices_to_circle) and
 * shifted by (k % d, k / d) * 2.1, d the smallest integer with d*d >= number of loops; the rest solves the
 * harmonic system of the uniform Laplacian (IglUtils::computeUniformLaplacian, IglUtils.cpp:39-60;
 * igl::harmonic with k = 1) — here by PCG on the GPU (OC_PRECOND_JACOBI or OC_PRECOND_AMG, relative
 * tolerance relTol) instead of a sparse Cholesky; pinned vertices take their circle positions exactly. The
 * map becomes the context's UV (as oc_set_uv) and is copied to UVout (nV x 2) when non-null; *nLoops = loops
 * pinned, *nBnd = vertices pinned, *iters = PCG iterations. Errors: a component without boundary
 * (OC_ERR_STATE), an edge of three or more triangles, a vertex on two pinned loops or in no triangle
 * (OC_ERR_ARG). Call oc_newton_init afterwards. */
enum { OC_TUTTE_COMPONENTS = 0, OC_TUTTE_ALL_LOOPS = 1 };
int oc_tutte(oc_ctx* ctx, int layout, int precond, double relTol, int maxIter, double* UVout, int* nLoops,
             int* nBnd, int* iters);

/* ---- files: OBJ meshes and per-iteration traces ----------------------------------------------------
 * Host-side I/O in the formats OptCuts reads and writes. No GPU is needed for the OBJ calls.
 * oc_obj_read_sizes / oc_obj_read replace igl::readOBJ(path, V, UV, N, F, FUV, FN) as main.cpp:1153 calls it:
 * V nV x 3, UV nUV x 2 (texture coordinates), F nF x 3 and FUV nF x 3 (0-based; *hasFUV = 0 when faces carry
 * no texture indices). Triangles with 3-coordinate vertices only. Errors: OC_ERR_ARG with
 * oc_io_last_error(). */
const char* oc_io_last_error(void);
int oc_obj_read_sizes(const char* path, int* nV, int* nUV, int* nF, int* hasFUV);
int oc_obj_read(const char* path, double* V, double* UV, int* F, int* FUV);
/* TriMesh::save(path, V, F, UV, FUV) (TriMesh.cpp:1758-1798): "v"/"vt" lines in %e, "f v/t v/t v/t";
 * FUV null writes F for both indices. */
int oc_obj_write(const char* path, int nV, const double* V, int nF, const int* F, int nUV, const double* UV,
                 const int* FUV);
/* TriMesh::saveAsMesh(path, F0, scaleUV) (TriMesh.cpp:1939-1969, written by main.cpp:341-342): the cut mesh
 * (nV rest positions Vrest, UV, triangles F) is written on the input mesh's vertices F0 (same triangle
 * order) with the cut mesh's UV as texture coordinates, normalised to the unit square when scaleUV. */
int oc_obj_write_mesh(const char* path, int nV, const double* Vrest, const double* UV, int nF, const int* F,
                      const int* F0, int scaleUV);
/* Per-iteration traces of the device Newton loop in the files Optimizer writes (Optimizer.cpp:61-62,
 * 706-739): folder/energyValPerIter.txt gets "E_total E_SD seam a0" at oc_newton_init and after every step
 * that did not stop; folder/gradientPerIter.txt gets "|g|^2 |g_SD|^2" at the start of every step. The seam
 * column is the caller's value (TriMesh::computeSeamSparsity / virtualRadius, Optimizer.cpp:708-710; 0 until
 * set). Single-mesh contexts only (not with oc_set_charts). oc_trace_close (or oc_ctx_destroy) closes the
 * files. */
int oc_trace_open(oc_ctx* ctx, const char* folder);
int oc_trace_set_seam(oc_ctx* ctx, double seamTerm);
/* The seam column as Optimizer::writeEnergyValToFile computes it (Optimizer.cpp:706-728): every energy line
 * recomputes E_se = TriMesh::computeSeamSparsity(E_se, !fractureMode) / virtualRadius at the current UV over the
 * cohesive edges cohE (nCoh x 4, as oc_seam_sparsity) plus initSeamLen, virtualRadius = sqrt(surfaceArea / pi)
 * (TriMesh.cpp:404); in fracture mode the first column is lastEnergyVal + (1 - a0) E_se. Replaces the constant of
 * oc_trace_set_seam (and vice versa). */
int oc_trace_set_seams(oc_ctx* ctx, int nCoh, const int* cohE, double initSeamLen, int fractureMode);
int oc_trace_close(oc_ctx* ctx);
/* info.txt as saveInfoForPresent writes it (main.cpp:346-387) at the context's UV: the caller's counts (input
 * vertices and faces, iterations, topology iterations), lambda_init and timings[15] = {timing_total, secPast, topo,
 * desc, scaf, enUp, mtrComp, mtrAssem, symFac, numFac, backSolve, lineSearch, bSplit, iSplit, cMerge}; the
 * library's distortion (E_SD), seam length / virtualRadius (half the boundary length when a0 = 1, else the seams and
 * mode of oc_trace_set_seams), TriMesh::outputStandardStretch (TriMesh.cpp:1585-1658) and the initial seams
 * (nInitSeams x 4, printed as Eigen prints a matrix). 1 - energyParams[0] is 1 - a0 of the last oc_newton_init. */
int oc_info_write(oc_ctx* ctx, const char* path, int vertAmtInput, int nFInitial, int iterNum, int topoIter,
                  double lambdaInit, const double* timings, int nInitSeams, const int* initSeams);

/* ---- batched charts ----------------------------------------------------------------------------------
 * Many independent charts in one context: after oc_set_mesh with the union of the charts (chart k owns the
 * vertices [vertexStart[k], vertexStart[k+1]) and the triangles [faceStart[k], faceStart[k+1]), triangles
 * referencing only their own chart's vertices; fixed vertices as usual), oc_set_charts makes every later
 * oc_newton_init / oc_newton_step treat each chart as its own problem: its own surface-area weights, gradient
 * tolerance, injectivity bound, line search, step and stop tests — the results of separate
 * Optimizer::solve(1) calls per chart (Optimizer.cpp:206-231, 575-652) — while each kernel launch and the
 * (block-diagonal) Newton solve cover all charts at once. The step reports sums over charts; *done = 1 once
 * no chart is active. oc_set_mesh clears the charts; an air mesh is not supported with charts. No reference
 * counterpart (the reference runs charts one at a time). */
int oc_set_charts(oc_ctx* ctx, int nCharts, const int* vertexStart, const int* faceStart);
/* Per chart after the last oc_newton_init / oc_newton_step: E_SD, |g|^2, step size, state (0 active, 1 gradient
 * tolerance reached, 2 line search stopped). Any pointer may be null; arrays of nCharts. */
int oc_charts_state(oc_ctx* ctx, double* E_SD, double* gradSq, double* step, int* state);

/* ---- timing of the dominant kernels (HIP events on the context stream) ---- */
/* Average device time in ms (HIP events around `reps` back-to-back launches) of: 0 element pass (E, g,
 * projected H), 1 block assembly, 2 one PCG iteration (both kernels), 3 line-search candidates, 4 step bound,
 * 5 PCG SpMV kernel, 6 PCG update kernel, 7 one application of the multigrid or learned preconditioner, 8 its
 * set-up (multigrid hierarchy or network forward pass), 9 one fine-level Lanczos step of the multigrid set-up
 * (the lambda_max estimate), 10 + l one Chebyshev smoothing step of the V-cycle on multigrid level l (a level
 * with a transfer, l < levels - 1). Needs oc_newton_init (7-9 and 10+ with that preconditioner and one
 * oc_newton_step); the current search direction is kept. */
int oc_time_kernel(oc_ctx* ctx, int which, int reps, double* ms);
/* The multigrid hierarchy of the last Newton system: *nLevels, and per level (up to capacity) its unknowns (nodes
 * times block size), stored blocks and block size (2 on the finest level, 3 below). *nLevels = 0 without multigrid. No reference
 * counterpart (diagnostics for the per-level roofline). */
int oc_amg_levels(oc_ctx* ctx, int capacity, int* nLevels, long long* rows, long long* blocks, int* blockSize);
/* Sizes used for algorithmic-byte accounting: triangles, vertices (combined), stored blocks, upper blocks. */
int oc_system_sizes(oc_ctx* ctx, long long* nF, long long* nV, long long* nnzBlocks, long long* nUpperBlocks);
/* Copies the Newton system of the last oc_newton_step to the host (any pointer may be null): the full
 * symmetric 2x2-block CSR (rowPtr nV+1, colIdx nnzBlocks; nV and nnzBlocks from oc_system_sizes), its block
 * values as doubles (4 per block, row-major [xx, xy, yx, yy]; widened from FP32 under OC_PRECISION_MIXED), the
 * right-hand side -g (2 nV) and the fixed-vertex mask (nV bytes). No reference counterpart: it exports the
 * method's own Hessians (training data of the learned preconditioner, solver diagnostics). */
int oc_newton_get_system(oc_ctx* ctx, int* rowPtr, int* colIdx, double* vals, double* rhs, unsigned char* fixedMask);

/* ---- learned preconditioner (no reference counterpart) --------------------------------------------------
 * A 3-layer edge-message graph network (hidden width 32) over the vertex graph of a Newton system predicts a
 * sparse lower-triangular block factor G on the Hessian pattern; PCG then uses M^-1 = C^-T G^T G C^-1 with
 * C_u C_u^T the diagonal blocks (G = I is block Jacobi). Message and update products run on BF16 matrix cores
 * with FP32 accumulation. Training minimises the probe estimate of |G A^ G^T - I|_F^2 / (2 nV),
 * A^ = C^-1 A C^-T, with the library's own backward pass and Adam. Parameters are FP32, flat, in the order
 * W_enc (32x4), b_enc, then per layer W_m (32x80), b_m, W_h (32x64), b_h, then W_o (4x80), b_o, W_d (3x32),
 * b_d; oc_gnn_num_params() of them. */
int oc_gnn_num_params(void);
int oc_gnn_create(int device, oc_gnn** out); /* parameters and optimizer state start at zero */
int oc_gnn_destroy(oc_gnn* net);
const char* oc_gnn_last_error(const oc_gnn* net);
int oc_gnn_set_params(oc_gnn* net, const float* theta, int n);
int oc_gnn_get_params(oc_gnn* net, float* theta, int n);
int oc_gnn_get_grads(oc_gnn* net, float* grad, int n);
/* A system to train on or to precondition: full symmetric 2x2-block CSR (rows sorted with their diagonal
 * block; nnzBlocks = rowPtr[nV]), block values as doubles (4 per block, row-major), fixed-vertex mask. The
 * edge maps are rebuilt only when the pattern changes. */
int oc_gnn_set_system(oc_gnn* net, int nV, const int* rowPtr, const int* colIdx, const double* vals,
                      const unsigned char* fixedMask);
/* Forward pass; when Gblocks is non-null it receives the factor per CSR block (4 floats: lower and diagonal
 * blocks hold G, upper blocks zero). */
int oc_gnn_forward(oc_gnn* net, float* Gblocks);
int oc_gnn_zero_grad(oc_gnn* net);
/* Forward + loss over nProbes Rademacher probes drawn from seed; adds the gradient to the accumulated one. */
int oc_gnn_loss_grad(oc_gnn* net, int nProbes, unsigned int seed, double* loss);
/* theta -= lr * m_hat / (sqrt(v_hat) + eps) with the accumulated gradient (Adam; t counts from 1). */
int oc_gnn_adam_step(oc_gnn* net, float lr, float beta1, float beta2, float eps, int t);
/* z = M^-1 r for the current system after oc_gnn_forward (2 nV doubles each). */
int oc_gnn_precond_apply(oc_gnn* net, const double* r, double* z);

#ifdef __cplusplus
}
#endif

#endif /* OPTCUTS_GPU_H */
