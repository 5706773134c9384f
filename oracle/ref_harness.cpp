// Harness around the reference's own, unmodified hot-path code that compiles
// without stand-ins in this image (TEST INFRASTRUCTURE ONLY; built into
// oracle/_ref/ by oracle/Makefile from the sources where they lie under the
// reference tree, never copied):
//   - src/LinSysSolver/EigenLibSolver.cpp + LinSysSolver.hpp: the CSR pattern
//     builder (set_pattern), triplet accumulation (update_a) and the Eigen
//     SimplicialLDLT analyze/factorize/solve (EigenLibSolver.cpp:21-108);
//   - src/Utils/IglUtils.hpp: IglUtils::makePD (IglUtils.hpp:71-90),
//     header-only, on Eigen's SelfAdjointEigenSolver.
// The remaining hot-path translation units (SymDirichletEnergy.cpp,
// IglUtils.cpp, TriMesh.cpp, Scaffold.cpp) include <tbb/tbb.h>, which this
// image lacks, so they are not built.

#include "EigenLibSolver.hpp"
#include "IglUtils.hpp"

#include <cstring>
#include <set>
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

}  // extern "C"
