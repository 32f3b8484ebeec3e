// poisson2d.cpp -- drop-in PoissonSolver2D: the CSR Krylov solvers of the reference
// (src/poisson2d.cpp:230-575) on the GPU through asd2_csr_pcg / asd2_csr_bicgstab.
//
// Host side keeps the reference's gather/scatter between the padded pressure array and the
// compact unknown vector (x fastest, the pinned cell (0,0) removed when every P side is
// Neumann-like, :236-250), its console report and its fatal NaN/inf check (:257-276).
#include "poisson2d.h"

#include <algorithm>
#include <map>

#include "host_util.h"

using namespace asd2host;

namespace {

bool neumann_like(BC2D k) { return k == BC2D::NEUMANN || k == BC2D::AXISYM || k == BC2D::WALL; }

bool pinned(const BoundaryCondition2D& B) {
    return neumann_like(B.m_BC_PW) && neumann_like(B.m_BC_PE) && neumann_like(B.m_BC_PS) && neumann_like(B.m_BC_PN);
}

// CSR arrays on the device for one solve
struct DeviceCSR {
    int n = 0;
    int* rp = nullptr;
    int* cols = nullptr;
    double* vals = nullptr;
    double* mdiag = nullptr;
    double* b = nullptr;
    double* x = nullptr;
    ~DeviceCSR() {
        for (void* p : {static_cast<void*>(rp), static_cast<void*>(cols), static_cast<void*>(vals),
                        static_cast<void*>(mdiag), static_cast<void*>(b), static_cast<void*>(x)})
            if (p) asd2_free(p);
    }
};

template <typename T>
T* upload(const T* h, size_t n) {
    void* d = nullptr;
    check(asd2_malloc(&d, std::max<size_t>(n, 1) * sizeof(T)), "asd2_malloc");
    if (n) check(asd2_memcpy_h2d(d, h, n * sizeof(T)), "upload");
    return static_cast<T*>(d);
}

// The Jacobi preconditioner arrives as a diagonal CSR matrix (mac2d.cpp:1641-1643); its
// values become one entry per row.
std::vector<double> diagonal_of(int n, const std::vector<double>& MVals, const std::vector<MKL_INT>& MCols,
                                const std::vector<MKL_INT>& MRowIdx) {
    if (static_cast<int>(MRowIdx.size()) != n + 1) fatal("PoissonSolver2D: preconditioner row count mismatch");
    std::vector<double> d(static_cast<size_t>(n), 0.0);
    for (int r = 0; r < n; r++) {
        for (MKL_INT e = MRowIdx[r]; e < MRowIdx[r + 1]; e++) {
            if (MCols[e] != r) fatal("PoissonSolver2D: the preconditioner must be diagonal (Jacobi)");
            d[r] = MVals[e];
        }
    }
    return d;
}

}  // namespace

PoissonSolver2D::PoissonSolver2D(int nx, int ny, int num_bc_grid) : kNx(nx), kNy(ny), kNumBCGrid(num_bc_grid) {}

PoissonSolver2D::~PoissonSolver2D() {
    if (m_ctx) asd2_destroy(m_ctx);
}

asd2_ctx* PoissonSolver2D::Context() {
    if (!m_ctx) m_ctx = make_ctx(kNx, kNy, kNumBCGrid, 0.0, 0.0, 1.0, 1.0);
    return m_ctx;
}

int PoissonSolver2D::CG_2FUniformP_2D(std::vector<double>& ps, const std::vector<double>& rhs,
                                      std::vector<double>& AVals, std::vector<MKL_INT>& ACols,
                                      std::vector<MKL_INT>& ARowIdx, std::vector<double>& MVals,
                                      std::vector<MKL_INT>& MCols, std::vector<MKL_INT>& MRowIdx,
                                      const std::shared_ptr<BoundaryCondition2D>& PBC, int64_t size,
                                      const int maxIter) {
    std::vector<double> b(static_cast<size_t>(size), 0.0), x(static_cast<size_t>(size), 0.0);
    const bool pin = pinned(*PBC);
    for (int j = 0; j < kNy; j++)
        for (int i = 0; i < kNx; i++) {
            int64_t k = i + static_cast<int64_t>(j) * kNx;
            if (pin) {
                if (i == 0 && j == 0) continue;
                k--;
            }
            b[k] = rhs[idx(i + kNumBCGrid, j + kNumBCGrid)];
        }
    CG_2FUniform_2D(x.data(), b.data(), AVals, ACols, ARowIdx, MVals, MCols, MRowIdx, PBC, size, maxIter);
    for (int j = 0; j < kNy; j++)
        for (int i = 0; i < kNx; i++) {
            int64_t k = i + static_cast<int64_t>(j) * kNx;
            if (pin) {
                if (i == 0 && j == 0) continue;
                k--;
            }
            const double val = x[k];
            ps[idx(i + kNumBCGrid, j + kNumBCGrid)] = val;
            if (std::isnan(val) || std::isinf(val)) {
                std::cout << "poisson equation nan/inf error : " << i + kNumBCGrid << " " << j + kNumBCGrid << " "
                          << val << std::endl;
                std::exit(1);
            }
        }
    return 0;
}

int PoissonSolver2D::CG_2FUniform_2D(double* x, double* b, std::vector<double>& AVals, std::vector<MKL_INT>& ACols,
                                     std::vector<MKL_INT>& ARowIdx, std::vector<double>& MVals,
                                     std::vector<MKL_INT>& MCols, std::vector<MKL_INT>& MRowIdx,
                                     const std::shared_ptr<BoundaryCondition2D>& PBC, int64_t size,
                                     const int maxIter) {
    (void)PBC;
    const int n = static_cast<int>(size);
    if (static_cast<int64_t>(ARowIdx.size()) - 1 != size) std::cout << "the # of rows is invalid!" << std::endl;
    const std::vector<double> md = diagonal_of(n, MVals, MCols, MRowIdx);
    DeviceCSR D;
    D.n = n;
    D.rp = upload(ARowIdx.data(), ARowIdx.size());
    D.cols = upload(ACols.data(), ACols.size());
    D.vals = upload(AVals.data(), AVals.size());
    D.mdiag = upload(md.data(), md.size());
    D.b = upload(b, static_cast<size_t>(n));
    D.x = upload(x, static_cast<size_t>(n));
    asd2_solve_opts o{};
    o.tol = m_tol;
    o.max_iter = maxIter;
    o.mode = m_mode;
    o.precision = ASD2_PREC_FP64;
    o.precond = ASD2_PRECOND_JACOBI;
    asd2_solve_stats st{};
    const int rc = asd2_csr_pcg(Context(), n, D.rp, D.cols, D.vals, D.mdiag, D.b, D.x, &o, &st);
    if (rc == ASD2_ERR_NONFINITE) {
        std::cout << "poisson equation nan/inf error : " << asd2_last_error() << std::endl;
        std::exit(1);
    }
    if (rc != ASD2_OK && rc != ASD2_ERR_NOT_CONVERGED) fatal("CG_2FUniform_2D");
    check(asd2_memcpy_d2h(x, D.x, static_cast<size_t>(n) * 8), "download");
    m_lastIters = st.iterations;
    m_lastRes = st.rel_residual;
    // the reference's report line (src/poisson2d.cpp:398); running out of iterations is not fatal
    std::cout << "CG : " << st.iterations << " " << maxIter << " Err : " << st.rel_residual << std::endl;
    return 0;
}

int PoissonSolver2D::BiCGStab_2FUniform_2D(std::vector<double>& ps, const std::vector<double>& rhs,
                                           std::vector<double>& AVals, std::vector<MKL_INT>& ACols,
                                           std::vector<MKL_INT>& ARowIdx, std::vector<double>& MVals,
                                           std::vector<MKL_INT>& MCols, std::vector<MKL_INT>& MRowIdx,
                                           const std::shared_ptr<BoundaryCondition2D>& PBC, int64_t size,
                                           const int maxiter) {
    // Unknown order as in CG_2FUniformP_2D: the reference's BiCGStab gathers all nx*ny cells
    // even when the matrix has nx*ny - 1 rows (:427-439), which overruns its arrays.
    const int n = static_cast<int>(size);
    if (static_cast<int64_t>(ARowIdx.size()) - 1 != size) std::cout << "the # of rows is invalid!" << std::endl;
    const bool pin = pinned(*PBC) && size == static_cast<int64_t>(kNx) * kNy - 1;
    std::vector<double> b(static_cast<size_t>(n), 0.0), x(static_cast<size_t>(n), 0.0);
    for (int j = 0; j < kNy; j++)
        for (int i = 0; i < kNx; i++) {
            int64_t k = i + static_cast<int64_t>(j) * kNx;
            if (pin) {
                if (i == 0 && j == 0) continue;
                k--;
            }
            if (k < n) b[k] = rhs[idx(i + kNumBCGrid, j + kNumBCGrid)];
        }
    const std::vector<double> md = diagonal_of(n, MVals, MCols, MRowIdx);
    DeviceCSR D;
    D.n = n;
    D.rp = upload(ARowIdx.data(), ARowIdx.size());
    D.cols = upload(ACols.data(), ACols.size());
    D.vals = upload(AVals.data(), AVals.size());
    D.mdiag = upload(md.data(), md.size());
    D.b = upload(b.data(), b.size());
    D.x = upload(x.data(), x.size());
    asd2_solve_opts o{};
    o.tol = m_tol;
    o.max_iter = maxiter;
    o.mode = ASD2_CG_STANDARD;
    asd2_solve_stats st{};
    const int rc = asd2_csr_bicgstab(Context(), n, D.rp, D.cols, D.vals, D.mdiag, D.b, D.x, &o, &st);
    if (rc != ASD2_OK && rc != ASD2_ERR_NOT_CONVERGED && rc != ASD2_ERR_NONFINITE) fatal("BiCGStab_2FUniform_2D");
    check(asd2_memcpy_d2h(x.data(), D.x, static_cast<size_t>(n) * 8), "download");
    m_lastIters = st.iterations;
    m_lastRes = st.rel_residual;
    std::cout << "BiCG : " << st.iterations << " " << maxiter << " Err : " << st.rel_residual << std::endl;
    for (int j = 0; j < kNy; j++)
        for (int i = 0; i < kNx; i++) {
            int64_t k = i + static_cast<int64_t>(j) * kNx;
            if (pin) {
                if (i == 0 && j == 0) continue;
                k--;
            }
            if (k >= n) continue;
            const double val = x[k];
            ps[idx(i + kNumBCGrid, j + kNumBCGrid)] = val;
            if (std::isnan(val) || std::isinf(val)) {
                std::cout << "poisson equation nan/inf error : " << i + kNumBCGrid << " " << j + kNumBCGrid << " "
                          << val << std::endl;
                std::exit(1);
            }
        }
    return 0;
}

// Gauss-Seidel sweeps with relaxation 1.2 to a relative change of 1e-6 (src/poisson2d.cpp:7-88):
// the unknowns are the nx*ny interior cells, i + j nx, starting from ps; the caller's CSR has a row
// per unknown.
int PoissonSolver2D::GS_2FUniform_2D(std::vector<double>& ps, const std::vector<double>& rhs,
                                     std::vector<double>& AVals, std::vector<MKL_INT>& ACols,
                                     std::vector<MKL_INT>& ARowIdx, const double, const double, const double,
                                     const double, const std::shared_ptr<BoundaryCondition2D>&, const int maxIter) {
    const int n = kNx * kNy;
    if (static_cast<int>(ARowIdx.size()) - 1 < n) {
        std::cerr << "GS_2FUniform_2D: the matrix has " << ARowIdx.size() - 1 << " rows, " << n << " needed"
                  << std::endl;
        return -1;
    }
    std::vector<double> b(static_cast<size_t>(n)), x(static_cast<size_t>(n));
    for (int i = 0; i < kNx; i++)
        for (int j = 0; j < kNy; j++) {
            b[i + static_cast<size_t>(j) * kNx] = rhs[idx(i + kNumBCGrid, j + kNumBCGrid)];
            x[i + static_cast<size_t>(j) * kNx] = ps[idx(i + kNumBCGrid, j + kNumBCGrid)];
        }
    DeviceCSR D;
    D.n = n;
    D.rp = upload(ARowIdx.data(), static_cast<size_t>(n) + 1);
    D.cols = upload(ACols.data(), ACols.size());
    D.vals = upload(AVals.data(), AVals.size());
    D.b = upload(b.data(), b.size());
    D.x = upload(x.data(), x.size());
    asd2_solve_stats st{};
    if (asd2_csr_gs(Context(), n, D.rp, D.cols, D.vals, D.b, D.x, maxIter, &st) != ASD2_OK) fatal("GS_2FUniform_2D");
    check(asd2_memcpy_d2h(x.data(), D.x, static_cast<size_t>(n) * 8), "download");
    m_lastIters = st.iterations;
    m_lastRes = st.rel_residual;
    for (int i = 0; i < kNx; i++)
        for (int j = 0; j < kNy; j++) ps[idx(i + kNumBCGrid, j + kNumBCGrid)] = x[i + static_cast<size_t>(j) * kNx];
    for (int i = kNumBCGrid; i < kNx + kNumBCGrid; i++)
        for (int j = kNumBCGrid; j < kNy + kNumBCGrid; j++)
            if (std::isnan(ps[idx(i, j)]) || std::isinf(ps[idx(i, j)]))
                std::cout << "Pseudo-p nan/inf error : " << i << " " << j << " " << ps[idx(i, j)] << std::endl;
    return 0;
}

// src/poisson2d.cpp:90-228: the constant-coefficient Poisson solve the reference hands to the MKL
// Poisson Library, here asd2_fast_poisson (an eigenbasis transform along y and tridiagonal
// solves along x on the GPU).  The reference maps only Dirichlet, Neumann and periodic pressure
// sides to MKL's side types; periodic and the other kinds stop with the reference's "MKL Error"
// convention (perror + exit(1)).
int PoissonSolver2D::MKL_2FUniform_2D(std::vector<double>& ps, const std::vector<double>& rhs, double lenX,
                                      double lenY, double dx, double dy, std::shared_ptr<BoundaryCondition2D> PBC) {
    const size_t n = static_cast<size_t>(kNx + 2 * kNumBCGrid) * static_cast<size_t>(kNy + 2 * kNumBCGrid);
    if (ps.size() != n || rhs.size() != n) fatal("PoissonSolver2D::MKL_2FUniform_2D: array size mismatch");
    asd2_ctx* ctx = Context();
    PBC->Configure(ctx);
    DevVec d_ps(ps), d_rhs(rhs);
    if (asd2_fast_poisson(ctx, d_rhs.p(), d_ps.p(), lenX, lenY, dx, dy, 0.0) != ASD2_OK) {
        std::cerr << "MKL Error:d_Helmholtz_2D: " << asd2_last_error() << std::endl;
        std::exit(1);
    }
    d_ps.get(ps);
    return 0;
}

// Declared by the reference header without a definition: true when A equals its transpose.
bool PoissonSolver2D::IsSymmetric(std::vector<double>& AVals, std::vector<MKL_INT>& ACols,
                                  std::vector<MKL_INT>& ARowIdx) {
    const int n = static_cast<int>(ARowIdx.size()) - 1;
    std::map<std::pair<int, int>, double> entries;
    for (int r = 0; r < n; r++)
        for (MKL_INT e = ARowIdx[r]; e < ARowIdx[r + 1]; e++) entries[{r, ACols[e]}] += AVals[e];
    for (const auto& kv : entries) {
        auto it = entries.find({kv.first.second, kv.first.first});
        const double other = (it == entries.end()) ? 0.0 : it->second;
        if (other != kv.second) return false;
    }
    return true;
}

// src/poisson2d.cpp:577-585
std::vector<double> PoissonSolver2D::InvertMatrixDiagonal(const std::vector<double>& M) {
    std::vector<double> inv(M.size(), 0.0);
    for (size_t i = 0; i < M.size(); i++) inv[i] = 1.0 / M[i];
    return inv;
}
