// Harness around the reference's own, unmodified hot-path code that compiles
// without stand-ins in this image (TEST INFRASTRUCTURE ONLY; built into
// oracle/_ref/ by oracle/Makefile from the sources where they lie under the
// reference tree, never copied):
//   - src/LinSysSolver/EigenLibSolver.cpp + LinSysSolver.hpp: the CSR pattern
//     builder (set_pattern), triplet accumulation (update_a) and the Eigen
//     SimplicialLDLT analyze/factorize/solve (EigenLibSolver.cpp:21-108);
//   - src/Utils/IglUtils.hpp: IglUtils::makePD (IglUtils.hpp:71-90),
//     header-only, on Eigen's SelfAdjointEigenSolver;
//   - the vendored libigl's header-only igl::facet_components,
//     igl::boundary_loop, igl::map_vertices_to_circle and igl::harmonic, called
//     as the reference's Tutte initialisation calls them (main.cpp:1313-1364,
//     1380-1395, 1486-1532);
//   - igl::readOBJ as main.cpp:1153 calls it, and std::ostream formatting of
//     doubles with the flags TriMesh::save (std::scientific) and the
//     Optimizer trace writers (defaults) use. Its Laplacian
//     comes from IglUtils::computeUniformLaplacian (IglUtils.cpp:39-60), a
//     translation unit that needs TBB; ref_tutte assembles the same triplets
//     (1 per directed triangle edge, -2 per corner) itself.
// The remaining hot-path translation units (SymDirichletEnergy.cpp,
// IglUtils.cpp, TriMesh.cpp, Scaffold.cpp) include <tbb/tbb.h>, which this
// image lacks, so they are not built.

#include "EigenLibSolver.hpp"
#include "IglUtils.hpp"

#include <igl/boundary_loop.h>
#include <igl/facet_components.h>
#include <igl/harmonic.h>
#include <igl/map_vertices_to_circle.h>
#include <igl/readOBJ.h>

#include <cstring>
#include <iomanip>
#include <set>
#include <sstream>
#include <vector>

using Solver = OptCuts::EigenLibSolver<Eigen::VectorXi, Eigen::VectorXd>;

extern "C" {

void* ref_solver_create(void)
{
    Solver* s = new Solver();
    s->set_type(1, -2);
    return s;
}

void ref_solver_destroy(void* h) { delete static_cast<Solver*>(h); }

// Builds vNeighbor/fixedVert exactly as the reference containers and calls
// EigenLibSolver::set_pattern. Returns nnz (upper triangle).
int ref_set_pattern(void* h, int nV, const int* nbrPtr, const int* nbr, int nFixed, const int* fixedV)
{
    std::vector<std::set<int>> vNeighbor(nV);
    for (int v = 0; v < nV; ++v)
        for (int p = nbrPtr[v]; p < nbrPtr[v + 1]; ++p) vNeighbor[v].insert(nbr[p]);
    std::set<int> fixedVert(fixedV, fixedV + nFixed);
    Solver* s = static_cast<Solver*>(h);
    s->set_pattern(vNeighbor, fixedVert);
    return s->getNumNonzeros();
}

// 0-based ia (n+1) and ja (nnz), as EigenLibSolver leaves them (-1 shift).
void ref_get_pattern(void* h, int* ia, int* ja)
{
    Solver* s = static_cast<Solver*>(h);
    const Eigen::VectorXi& IA = s->get_ia();
    const Eigen::VectorXi& JA = s->get_ja();
    for (int k = 0; k < IA.size(); ++k) ia[k] = IA[k];
    for (int k = 0; k < JA.size(); ++k) ja[k] = JA[k];
}

void ref_update_a(void* h, int nT, const int* I, const int* J, const double* V)
{
    Eigen::VectorXi II(nT), JJ(nT);
    Eigen::VectorXd VV(nT);
    for (int k = 0; k < nT; ++k) { II[k] = I[k]; JJ[k] = J[k]; VV[k] = V[k]; }
    static_cast<Solver*>(h)->update_a(II, JJ, VV);
}

void ref_get_a(void* h, double* a)
{
    const Eigen::VectorXd& A = static_cast<const Solver*>(h)->get_a();
    std::memcpy(a, A.data(), sizeof(double) * A.size());
}

int ref_analyze_factorize(void* h)
{
    Solver* s = static_cast<Solver*>(h);
    s->analyze_pattern();
    return s->factorize() ? 0 : -1;
}

void ref_solve(void* h, int n, const double* rhs, double* x)
{
    Eigen::VectorXd b(n), r;
    for (int k = 0; k < n; ++k) b[k] = rhs[k];
    static_cast<Solver*>(h)->solve(b, r);
    for (int k = 0; k < n; ++k) x[k] = r[k];
}

// IglUtils::makePD on a row-major 6x6 (symmetric, so layout is immaterial).
void ref_make_pd6(double* H)
{
    Eigen::Matrix<double, 6, 6> M;
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) M(r, c) = H[6 * r + c];
    OptCuts::IglUtils::makePD(M);
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) H[6 * r + c] = M(r, c);
}

// Eigen's SelfAdjointEigenSolver eigenvalues (ascending) of a row-major 6x6.
void ref_eigvals6(const double* H, double* lam)
{
    Eigen::Matrix<double, 6, 6> M;
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) M(r, c) = H[6 * r + c];
    Eigen::SelfAdjointEigenSolver<Eigen::Matrix<double, 6, 6>> es(M);
    for (int k = 0; k < 6; ++k) lam[k] = es.eigenvalues()[k];
}

// Tutte initial UV, laid out as the reference does it:
//   layout 0 (main.cpp:1380-1395,1486-1532, no input UV): facet components by
//     igl::facet_components, per component the first longest loop of
//     igl::boundary_loop over that component's triangles (in mesh order);
//   layout 1 (main.cpp:1313-1364, input UV re-initialised): every loop of
//     igl::boundary_loop over the whole mesh.
// Loop k goes on the unit circle (igl::map_vertices_to_circle, arc length of the
// rest positions) shifted by (k % d, k / d) * 2.1 with d the smallest integer
// with d*d >= number of loops; the interior is igl::harmonic (k = 1) of the
// uniform Laplacian over the whole mesh. UV: nV x 2 row-major. bndOut receives
// the stacked loop vertices, loopLen the loop lengths. Returns the number of
// loops (-1: a component without boundary).
int ref_tutte(int nV, const double* Vrest, int nF, const int* F, int layout, double* UV, int* bndOut,
              int* loopLen)
{
    Eigen::MatrixXd V(nV, 3);
    for (int i = 0; i < nV; ++i)
        for (int j = 0; j < 3; ++j) V(i, j) = Vrest[3 * i + j];
    Eigen::MatrixXi Fm(nF, 3);
    for (int t = 0; t < nF; ++t)
        for (int j = 0; j < 3; ++j) Fm(t, j) = F[3 * t + j];
    std::vector<std::vector<int>> loops;
    if (layout == 0) {
        Eigen::VectorXi C;
        igl::facet_components(Fm, C);
        const int nc = C.maxCoeff() + 1;
        std::vector<std::vector<int>> rows(nc);
        for (int t = 0; t < nF; ++t) rows[C[t]].push_back(t);
        for (int k = 0; k < nc; ++k) {
            Eigen::MatrixXi Fc((int)rows[k].size(), 3);
            for (int r = 0; r < (int)rows[k].size(); ++r) Fc.row(r) = Fm.row(rows[k][r]);
            std::vector<std::vector<int>> all;
            igl::boundary_loop(Fc, all);
            if (all.empty()) return -1;
            int best = 0;
            for (int b = 1; b < (int)all.size(); ++b)
                if (all[best].size() < all[b].size()) best = b;
            loops.push_back(all[best]);
        }
    } else {
        igl::boundary_loop(Fm, loops);
        if (loops.empty()) return -1;
    }
    int d = 0;
    do {
        ++d;
    } while (d * d < (int)loops.size());
    int nb = 0;
    for (auto& l : loops) nb += (int)l.size();
    Eigen::VectorXi bnd(nb);
    Eigen::MatrixXd bnd_uv(nb, 2);
    int off = 0;
    for (int k = 0; k < (int)loops.size(); ++k) {
        const int n = (int)loops[k].size();
        for (int i = 0; i < n; ++i) bnd[off + i] = loops[k][i];
        Eigen::MatrixXd circ;
        igl::map_vertices_to_circle(V, bnd.segment(off, n), circ);
        const double xo = k % d * 2.1, yo = k / d * 2.1;
        for (int i = 0; i < n; ++i) {
            bnd_uv(off + i, 0) = circ(i, 0) + xo;
            bnd_uv(off + i, 1) = circ(i, 1) + yo;
        }
        if (loopLen) loopLen[k] = n;
        off += n;
    }
    std::vector<Eigen::Triplet<double>> trip;
    trip.reserve(9 * (size_t)nF);
    for (int t = 0; t < nF; ++t) {
        const int a = Fm(t, 0), b = Fm(t, 1), c = Fm(t, 2);
        trip.emplace_back(a, b, 1.0);
        trip.emplace_back(b, a, 1.0);
        trip.emplace_back(b, c, 1.0);
        trip.emplace_back(c, b, 1.0);
        trip.emplace_back(c, a, 1.0);
        trip.emplace_back(a, c, 1.0);
        trip.emplace_back(a, a, -2.0);
        trip.emplace_back(b, b, -2.0);
        trip.emplace_back(c, c, -2.0);
    }
    Eigen::SparseMatrix<double> A(nV, nV), M;
    A.setFromTriplets(trip.begin(), trip.end());
    Eigen::MatrixXd W;
    igl::harmonic(A, M, bnd, bnd_uv, 1, W);
    for (int i = 0; i < nV; ++i) {
        UV[2 * i] = W(i, 0);
        UV[2 * i + 1] = W(i, 1);
    }
    if (bndOut)
        for (int i = 0; i < nb; ++i) bndOut[i] = bnd[i];
    return (int)loops.size();
}

// igl::readOBJ(path, V, TC, N, F, FTC, FN) into caller arrays: the sizes are
// always returned (sizes[0..6] = V rows, V cols, TC rows, TC cols, F rows,
// F cols, FTC rows); arrays are filled when non-null. Returns 1 on success.
int ref_read_obj(const char* path, int* sizes, double* V, double* TC, int* F, int* FTC)
{
    Eigen::MatrixXd Vm, TCm, Nm;
    Eigen::MatrixXi Fm, FTCm, FNm;
    if (!igl::readOBJ(path, Vm, TCm, Nm, Fm, FTCm, FNm)) return 0;
    sizes[0] = (int)Vm.rows();
    sizes[1] = (int)Vm.cols();
    sizes[2] = (int)TCm.rows();
    sizes[3] = (int)TCm.cols();
    sizes[4] = (int)Fm.rows();
    sizes[5] = (int)Fm.cols();
    sizes[6] = (int)FTCm.rows();
    if (V)
        for (int i = 0; i < Vm.rows(); ++i)
            for (int j = 0; j < Vm.cols(); ++j) V[i * Vm.cols() + j] = Vm(i, j);
    if (TC)
        for (int i = 0; i < TCm.rows(); ++i)
            for (int j = 0; j < TCm.cols(); ++j) TC[i * TCm.cols() + j] = TCm(i, j);
    if (F)
        for (int i = 0; i < Fm.rows(); ++i)
            for (int j = 0; j < Fm.cols(); ++j) F[i * Fm.cols() + j] = Fm(i, j);
    if (FTC)
        for (int i = 0; i < FTCm.rows(); ++i)
            for (int j = 0; j < FTCm.cols(); ++j) FTC[i * FTCm.cols() + j] = FTCm(i, j);
    return 1;
}

// A double as an std::ostream writes it with default flags (scientific = 0)
// or after std::scientific (scientific = 1); returns the string length.
int ref_stream_double(double x, int scientific, char* out, int cap)
{
    std::ostringstream os;
    if (scientific) os << std::scientific;
    os << x;
    const std::string t = os.str();
    std::snprintf(out, cap, "%s", t.c_str());
    return (int)t.size();
}

}  // extern "C"
