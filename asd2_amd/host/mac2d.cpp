// mac2d.cpp -- drop-in MACSolver2D on the gfx950 kernels (asd2_mac_*, asd2_gfm_*,
// asd2_poisson_*).  Reference: src/mac2d.cpp.
//
// The public fields stay host std::vectors because the reference's drivers read and write
// them directly (tests/LargeTests/test_mac2d_large.cpp:451-457,488,534-535).  By default each
// method uploads what it reads into the solver's named device buffers, runs the kernels, and
// downloads what it writes; in resident mode (asd2_resident.h) the vectors it writes keep their
// newest contents on the device between calls.  SolvePoisson never builds the CSR matrix: the
// ghost-fluid coefficients are assembled matrix-free and solved on the GPU; PoissonSolver2D
// keeps its CSR entry points for callers that build their own matrices.
#include "mac2d.h"

#include "host_util.h"

using namespace asd2host;

struct MACSolver2D::Device : SolverDevice {};

// src/mac2d.cpp:3-26 (dimensionless groups given; g = Fr L / U^2 as written there)
MACSolver2D::MACSolver2D(double Re, double We, double Fr, GAXISENUM2D GAxis, double L, double U, double sigma,
                         double densityRatio, double viscosityRatio, double rhoH, double muH, int nx, int ny,
                         double baseX, double baseY, double lenX, double lenY, TIMEORDERENUM timeOrder, double cfl,
                         double maxtime, int maxIter, int niterskip, int num_bc_grid, bool writeVTK)
    : kNx(nx), kNy(ny), kNumBCGrid(num_bc_grid),
      kArrSize(static_cast<int64_t>(nx + 2 * num_bc_grid) * static_cast<int64_t>(ny + 2 * num_bc_grid)),
      kBaseX(baseX), kBaseY(baseY), kLenX(lenX), kLenY(lenY), kDx(lenX / static_cast<double>(nx)),
      kDy(lenY / static_cast<double>(ny)), kRe(Re), kWe(We), kFr(Fr), kLScale(L), kUScale(U), kSigma(sigma),
      kG(Fr * L / (U * U)), kMuScale((We * sigma / (L * U * U)) * L * U / Re), kRhoScale(We * sigma / (L * U * U)),
      kRhoH(rhoH), kRhoL(rhoH * densityRatio), kRhoRatio(densityRatio), kMuH(muH), kMuL(muH * viscosityRatio),
      kMuRatio(viscosityRatio), kMaxIter(maxIter), kNIterSkip(niterskip), kTimeOrder(timeOrder), kCFL(cfl),
      kMaxTime(maxtime), kWriteVTK(writeVTK), m_PoissonSolverType(POISSONTYPE::CG), m_PLTType(PLTTYPE::NONE),
      kGAxis(GAxis) {
    m_iter = 0;
    m_curTime = 0.0;
    m_dt = std::min(0.005, kCFL * std::min(lenX / static_cast<double>(nx), lenY / static_cast<double>(ny)) / U);
}

// src/mac2d.cpp:28-49 (dimensional properties given)
MACSolver2D::MACSolver2D(double rhoH, double rhoL, double muH, double muL, double gConstant, GAXISENUM2D GAxis,
                         double L, double U, double sigma, int nx, int ny, double baseX, double baseY, double lenX,
                         double lenY, TIMEORDERENUM timeOrder, double cfl, double maxtime, int maxIter, int niterskip,
                         int num_bc_grid, bool writeVTK)
    : kNx(nx), kNy(ny), kNumBCGrid(num_bc_grid),
      kArrSize(static_cast<int64_t>(nx + 2 * num_bc_grid) * static_cast<int64_t>(ny + 2 * num_bc_grid)),
      kBaseX(baseX), kBaseY(baseY), kLenX(lenX), kLenY(lenY), kDx(lenX / static_cast<double>(nx)),
      kDy(lenY / static_cast<double>(ny)), kRe(rhoH * L * U / muH), kWe(rhoH * L * U * U / sigma),
      kFr(U * U / (gConstant * L)), kLScale(L), kUScale(U), kSigma(sigma), kG(gConstant), kMuScale(muH),
      kRhoScale(rhoH), kRhoH(rhoH), kRhoL(rhoL), kRhoRatio(rhoL / rhoH), kMuH(muH), kMuL(muL),
      kMuRatio(muL / muH), kMaxIter(maxIter), kNIterSkip(niterskip), kTimeOrder(timeOrder), kCFL(cfl),
      kMaxTime(maxtime), kWriteVTK(writeVTK), m_PoissonSolverType(POISSONTYPE::CG), m_PLTType(PLTTYPE::NONE),
      kGAxis(GAxis) {
    m_iter = 0;
    m_curTime = 0.0;
    m_dt = std::min(0.005, kCFL * std::min(lenX / static_cast<double>(nx), lenY / static_cast<double>(ny)) / U);
}

// the members' device mirrors (resident mode) go with the members
void MACSolver2D::ForgetMembers() {
    for (auto* a : {&m_u, &m_v, &m_rho, &m_kappa, &m_mu, &m_ps, &m_p, &m_J11, &m_J12, &m_J21, &m_J22, &m_nx, &m_ny,
                    &m_t1x, &m_t1y, &m_t1z, &m_t2x, &m_t2y, &m_t2z})
        mirror::forget(*a);
}

MACSolver2D::~MACSolver2D() { ForgetMembers(); }

MACSolver2D::Device& MACSolver2D::Dev() {
    if (!m_dev) {
        m_dev.reset(new Device);
        m_dev->ctx = make_ctx(kNx, kNy, kNumBCGrid, kBaseX, kBaseY, kDx, kDy);
        m_dev->n = static_cast<size_t>(kArrSize);
        m_dev->who = "MACSolver2D";
    }
    // the reference creates m_BC in SetBC_*_2D (src/mac2d.cpp:1883-1911); without one all
    // sides are CUSTOM
    if (!m_BC) m_BC = std::make_shared<BoundaryCondition2D>(kNx, kNy, kNumBCGrid);
    m_BC->Configure(m_dev->ctx);
    return *m_dev;
}

asd2_ctx* MACSolver2D::Context() { return Dev().ctx; }

void MACSolver2D::FluidParams(void* out) const {
    asd2_fluid& f = *static_cast<asd2_fluid*>(out);
    f.rho_h = kRhoH;
    f.rho_l = kRhoL;
    f.mu_h = kMuH;
    f.mu_l = kMuL;
    f.sigma = kSigma;
    f.dt = m_dt;
    f.g = kG;
    f.g_axis = (kGAxis == GAXISENUM2D::X) ? 0 : (kGAxis == GAXISENUM2D::Y ? 1 : 2);
    f.gravity_fix = m_gravityFix ? 1 : 0;
    f.re = kRe;
    f.we = kWe;
    f.fr = kFr;
    f.cfl = kCFL;
}

void MACSolver2D::SetPoissonOptions(double tol, int mode, int precision, int precond) {
    m_tol = tol;
    m_mode = mode;
    m_precision = precision;
    m_precond = precond;
}

// src/mac2d.cpp:59-85
int MACSolver2D::AllocateVariables() {
    ForgetMembers();
    for (auto* a : {&m_u, &m_v, &m_rho, &m_kappa, &m_mu, &m_ps, &m_p, &m_J11, &m_J12, &m_J21, &m_J22, &m_nx, &m_ny,
                    &m_t1x, &m_t1y, &m_t1z, &m_t2x, &m_t2y, &m_t2z})
        *a = std::vector<double>(static_cast<size_t>(kArrSize), 0.0);
    return 0;
}

// ------------------------------------------------------------------ interface fields
std::vector<double> MACSolver2D::UpdateHeavisideFunc(const std::vector<double>& ls) {
    Device& D = Dev();
    check(asd2_mac_heaviside(D.ctx, D.in("ls", ls), D.scratch("H"), 0), "UpdateHeavisideFunc");
    return D.result("H");
}

std::vector<double> MACSolver2D::UpdateSmoothHeavisideFunc(const std::vector<double>& ls) {
    Device& D = Dev();
    check(asd2_mac_heaviside(D.ctx, D.in("ls", ls), D.scratch("Hs"), 1), "UpdateSmoothHeavisideFunc");
    return D.result("Hs");
}

void MACSolver2D::UpdateHeavisideFunc(const std::vector<double>& ls, std::vector<double>& H) {
    Device& D = Dev();
    double* d_ls = D.in("ls", ls);
    check(asd2_mac_heaviside(D.ctx, d_ls, D.out("H", H), 0), "UpdateHeavisideFunc");
    D.wrote("H", H);
}

void MACSolver2D::UpdateSmoothHeavisideFunc(const std::vector<double>& ls, std::vector<double>& H) {
    Device& D = Dev();
    double* d_ls = D.in("ls", ls);
    check(asd2_mac_heaviside(D.ctx, d_ls, D.out("Hs", H), 1), "UpdateSmoothHeavisideFunc");
    D.wrote("Hs", H);
}

// src/mac2d.cpp:221-252: m_kappa's interior is rewritten, its ghosts are kept
int MACSolver2D::UpdateKappa(const std::vector<double>& ls) {
    Device& D = Dev();
    if (m_kappa.size() != static_cast<size_t>(kArrSize)) m_kappa.assign(static_cast<size_t>(kArrSize), 0.0);
    double* d_ls = D.in("ls", ls);
    check(asd2_mac_kappa(D.ctx, d_ls, D.inout("kappa", m_kappa)), "UpdateKappa");
    D.wrote("kappa", m_kappa);
    return 0;
}

// src/mac2d.cpp:124-219.  Normals of the "simple level set" cells (|1 - |grad phi|| <=
// 1e-4 with the ENO2 gradient) from central differences; the reference leaves the other
// branch (curve fitting) unimplemented, so those cells keep their values.  Host loop on
// the |grad phi| computed by the GPU: the routine is not on the solver path.
int MACSolver2D::UpdateNTK(const std::shared_ptr<LevelSetSolver2D>& LSolver, const std::vector<double>& ls,
                           std::tuple<std::vector<double>&, std::vector<double>&> normalVec,
                           std::tuple<std::vector<double>&, std::vector<double>&, std::vector<double>&>,
                           std::tuple<std::vector<double>&, std::vector<double>&, std::vector<double>&>,
                           std::vector<double>&) {
    const double eta = 0.0001, eps = 1.0e-100;
    const std::vector<double> absd = LSolver->ENO_DerivAbsLS_2D(ls, ls);
    SolverDevice::host_read(ls);
    std::vector<double>& nX = std::get<0>(normalVec);
    std::vector<double>& nY = std::get<1>(normalVec);
    SolverDevice::host_read(nX);  // cells that are not rewritten keep their values
    SolverDevice::host_read(nY);
    for (int i = kNumBCGrid; i < kNx + kNumBCGrid; i++)
        for (int j = kNumBCGrid; j < kNy + kNumBCGrid; j++) {
            if (std::fabs(1.0 - absd[idx(i, j)]) > eta) continue;
            const double dX = (ls[idx(i + 1, j)] - ls[idx(i - 1, j)]) / (2.0 * kDx);
            const double dY = (ls[idx(i, j + 1)] - ls[idx(i, j - 1)]) / (2.0 * kDy);
            nX[idx(i, j)] = 1.0 / std::sqrt(dX * dX + dY * dY + eps) * dX;
            nY[idx(i, j)] = 1.0 / std::sqrt(dX * dX + dY * dY + eps) * dY;
        }
    asd2::ModifiedHost(nX);
    asd2::ModifiedHost(nY);
    return 0;
}

// ------------------------------------------------------------------ momentum predictor
std::vector<double> MACSolver2D::MomentumTerm(int grid_kind, int term, const std::vector<double>* ls,
                                              const std::vector<double>& u, const std::vector<double>& v,
                                              const std::vector<double>* H, const char* where) {
    Device& D = Dev();
    asd2_fluid f;
    FluidParams(&f);
    // the convection term reads neither ls nor H
    double* d_ls = ls ? D.in("ls", *ls) : D.scratch("ls");
    double* d_u = D.in("u", u);
    double* d_v = D.in("v", v);
    double* d_H = H ? D.in("Hs", *H) : D.scratch("Hs");
    check(asd2_mac_momentum_term(D.ctx, &f, grid_kind, term, d_ls, d_u, d_v, d_H, D.scratch("out")), where);
    return D.result("out");
}

std::vector<double> MACSolver2D::UpdateFU(const std::shared_ptr<LevelSetSolver2D>&, const std::vector<double>& ls,
                                          const std::vector<double>& u, const std::vector<double>& v,
                                          const std::vector<double>& H) {
    return MomentumTerm(ASD2_GRID_U, ASD2_MOM_RHS, &ls, u, v, &H, "UpdateFU");
}

std::vector<double> MACSolver2D::UpdateFV(const std::shared_ptr<LevelSetSolver2D>&, const std::vector<double>& ls,
                                          const std::vector<double>& u, const std::vector<double>& v,
                                          const std::vector<double>& H) {
    return MomentumTerm(ASD2_GRID_V, ASD2_MOM_RHS, &ls, u, v, &H, "UpdateFV");
}

std::vector<double> MACSolver2D::AddConvectionFU(const std::vector<double>& u, const std::vector<double>& v) {
    return MomentumTerm(ASD2_GRID_U, ASD2_MOM_CONVECTION, nullptr, u, v, nullptr, "AddConvectionFU");
}

std::vector<double> MACSolver2D::AddConvectionFV(const std::vector<double>& u, const std::vector<double>& v) {
    return MomentumTerm(ASD2_GRID_V, ASD2_MOM_CONVECTION, nullptr, u, v, nullptr, "AddConvectionFV");
}

std::vector<double> MACSolver2D::AddViscosityFU(const std::vector<double>& u, const std::vector<double>& v,
                                                const std::vector<double>& ls, const std::vector<double>& H) {
    return MomentumTerm(ASD2_GRID_U, ASD2_MOM_VISCOSITY, &ls, u, v, &H, "AddViscosityFU");
}

std::vector<double> MACSolver2D::AddViscosityFV(const std::vector<double>& u, const std::vector<double>& v,
                                                const std::vector<double>& ls, const std::vector<double>& H) {
    return MomentumTerm(ASD2_GRID_V, ASD2_MOM_VISCOSITY, &ls, u, v, &H, "AddViscosityFV");
}

int MACSolver2D::UnitHJWENO5(const std::vector<double>& F, std::vector<double>& FP, std::vector<double>& FM,
                             const double d, const int n) {
    const size_t len = static_cast<size_t>(n + 2 * kNumBCGrid);
    if (F.size() < len) fatal("UnitHJWENO5: line shorter than n + 2g");
    DevVec dF(len), dP(len), dM(len);
    dF.put(F.data(), len);
    check(asd2_ls_unit_weno5(Dev().ctx, dF.p(), dP.p(), dM.p(), d, n, kNumBCGrid), "UnitHJWENO5");
    const std::vector<double> hp = dP.get(), hm = dM.get();
    if (FP.size() < len) FP.resize(len);
    if (FM.size() < len) FM.resize(len);
    for (int i = kNumBCGrid; i < n + kNumBCGrid; i++) {
        FP[i] = hp[i];
        FM[i] = hm[i];
    }
    return 0;
}

// src/mac2d.cpp:1043-1079: a constant on the interior faces of the gravity axis (no kernel
// needed; the fused momentum kernels add it themselves)
std::vector<double> MACSolver2D::AddGravityFU() {
    std::vector<double> gU(static_cast<size_t>(kArrSize), 0.0);
    if ((kFr != 0 || kG != 0.0) && !std::isnan(kFr) && !std::isinf(kFr) && kGAxis == GAXISENUM2D::X)
        for (int i = kNumBCGrid + 1; i < kNx + kNumBCGrid; i++)
            for (int j = kNumBCGrid; j < kNy + kNumBCGrid; j++) gU[idx(i, j)] = -kG;
    return gU;
}

std::vector<double> MACSolver2D::AddGravityFV() {
    std::vector<double> gV(static_cast<size_t>(kArrSize), 0.0);
    if ((kFr != 0 || kG != 0.0) && !std::isnan(kFr) && !std::isinf(kFr) && kGAxis == GAXISENUM2D::Y)
        for (int i = kNumBCGrid; i < kNx + kNumBCGrid; i++)
            for (int j = kNumBCGrid + 1; j < kNy + kNumBCGrid; j++) gV[idx(i, j)] = -kG;
    return gV;
}

int MACSolver2D::GetIntermediateVel(const std::shared_ptr<LevelSetSolver2D>&, const std::vector<double>& ls,
                                    const std::vector<double>& u, const std::vector<double>& v,
                                    std::vector<double>& uhat, std::vector<double>& vhat,
                                    const std::vector<double>& H) {
    Device& D = Dev();
    asd2_fluid f;
    FluidParams(&f);
    const int order = (kTimeOrder == TIMEORDERENUM::EULER) ? 0 : (kTimeOrder == TIMEORDERENUM::RK2 ? 1 : 2);
    double* d_ls = D.in("ls", ls);
    double* d_u = D.in("u", u);
    double* d_v = D.in("v", v);
    double* d_H = D.in("Hs", H);
    double* du = D.inout("uhat", uhat);
    double* dv = D.inout("vhat", vhat);
    check(asd2_mac_intermediate_vel(D.ctx, &f, order, d_ls, d_u, d_v, du, dv, d_H), "GetIntermediateVel");
    D.wrote("uhat", uhat);
    D.wrote("vhat", vhat);
    return 0;
}

// ------------------------------------------------------------------ pressure
int MACSolver2D::SetPoissonSolver(POISSONTYPE type) {
    m_PoissonSolverType = type;
    if (!m_Poisson) m_Poisson = std::make_shared<PoissonSolver2D>(kNx, kNy, kNumBCGrid);
    return 0;
}

// m_p = ps / dt on the interior (src/mac2d.cpp:1722-1724), on the device
void MACSolver2D::PressureFromPs(double* d_ps) {
    Device& D = Dev();
    if (m_p.size() != static_cast<size_t>(kArrSize)) m_p.assign(static_cast<size_t>(kArrSize), 0.0);
    check(asd2_mac_pressure(D.ctx, d_ps, D.inout("p", m_p), m_dt), "SolvePoisson (p = ps / dt)");
    D.wrote("p", m_p);
}

// src/mac2d.cpp:1249-1727.  ps is zeroed, kappa refreshed from ls when the Weber number is
// finite, the ghost-fluid system assembled on the GPU, solved, ps gets the P BCs and
// m_p = ps / dt on the interior.
int MACSolver2D::SolvePoisson(std::vector<double>& ps, const std::vector<double>& div, const std::vector<double>& ls,
                              const std::vector<double>&, const std::vector<double>&, const std::vector<double>&,
                              const std::vector<double>& H, const int maxIter) {
    if (!m_Poisson)
        perror("Solver method for Poisson equations are not set. Please add SetPoissonSolver Method to running code");
    if (m_PoissonSolverType == POISSONTYPE::GS || m_PoissonSolverType == POISSONTYPE::ICPCG) {
        std::cerr << "MACSolver2D::SolvePoisson: POISSONTYPE::GS and POISSONTYPE::ICPCG have no GPU path"
                  << std::endl;
        std::exit(1);
    }
    Device& D = Dev();
    if (m_kappa.size() != static_cast<size_t>(kArrSize)) m_kappa.assign(static_cast<size_t>(kArrSize), 0.0);
    double* d_ls = D.in("ls", ls);
    double* d_kappa = D.inout("kappa", m_kappa);
    if (kWe != 0.0 && !std::isnan(kWe) && !std::isinf(kWe)) {
        check(asd2_mac_kappa(D.ctx, d_ls, d_kappa), "UpdateKappa");
        check(asd2_bc_apply(D.ctx, ASD2_GRID_P, d_kappa), "ApplyBC_P_2D(kappa)");
        D.wrote("kappa", m_kappa);
    }
    const size_t n = static_cast<size_t>(kArrSize);
    if (m_PoissonSolverType == POISSONTYPE::MKL) {
        // src/mac2d.cpp:1658-1689: the constant-coefficient solve with rhs = -div on the interior
        // (ghosts 0) and ps zeroed; the assembled ghost-fluid system is not used on this path
        SolverDevice::host_read(div);
        std::vector<double> rhs(n, 0.0);
        for (int j = kNumBCGrid; j < kNy + kNumBCGrid; j++)
            for (int i = kNumBCGrid; i < kNx + kNumBCGrid; i++) rhs[idx(i, j)] = -div[idx(i, j)];
        double* d_ps = D.out("ps", ps);
        check(asd2_fill(D.ctx, d_ps, n, 0.0), "SolvePoisson (ps = 0)");
        if (asd2_fast_poisson(D.ctx, D.in("rhs", rhs), d_ps, kLenX, kLenY, kDx, kDy, 0.0) != ASD2_OK) {
            std::cerr << "MKL Error:d_Helmholtz_2D: " << asd2_last_error() << std::endl;
            std::exit(1);
        }
        check(asd2_bc_apply(D.ctx, ASD2_GRID_P, d_ps), "ApplyBC_P_2D(ps)");
        D.wrote("ps", ps);
        SolverDevice::host_read(ps);  // the ghost update leaves the interior the reference checks unchanged
        for (int j = kNumBCGrid; j < kNy + kNumBCGrid; j++)
            for (int i = kNumBCGrid; i < kNx + kNumBCGrid; i++)
                if (std::isnan(ps[idx(i, j)]) || std::isinf(ps[idx(i, j)]))
                    std::cout << "Pseudo-p nan/inf error : " << i << " " << j << " " << ps[idx(i, j)] << std::endl;
        PressureFromPs(d_ps);
        return 0;
    }
    asd2_fluid f;
    FluidParams(&f);
    check(asd2_gfm_assemble(D.ctx, &f, d_ls, d_kappa, D.in("H", H), D.in("div", div)), "SolvePoisson assembly");
    double* d_ps = D.out("ps", ps);
    check(asd2_fill(D.ctx, d_ps, n, 0.0), "SolvePoisson (ps = 0)");
    asd2_solve_opts o{};
    o.tol = m_tol;
    o.max_iter = maxIter;
    o.mode = m_mode;
    o.precision = m_precision;
    o.precond = m_precond;
    asd2_solve_stats st{};
    const int rc = (m_PoissonSolverType == POISSONTYPE::BICGSTAB) ? asd2_poisson_bicgstab(D.ctx, &o, d_ps, &st)
                                                                   : asd2_poisson_solve(D.ctx, &o, d_ps, &st);
    if (rc == ASD2_ERR_NONFINITE) {
        std::cout << "poisson equation nan/inf error : " << asd2_last_error() << std::endl;
        std::exit(1);
    }
    if (rc != ASD2_OK && rc != ASD2_ERR_NOT_CONVERGED) fatal("SolvePoisson");
    m_lastIters = st.iterations;
    std::cout << (m_PoissonSolverType == POISSONTYPE::BICGSTAB ? "BiCG : " : "CG : ") << st.iterations << " "
              << maxIter << " Err : " << st.rel_residual << std::endl;
    check(asd2_bc_apply(D.ctx, ASD2_GRID_P, d_ps), "ApplyBC_P_2D(ps)");
    D.wrote("ps", ps);
    PressureFromPs(d_ps);
    return 0;
}

std::vector<double> MACSolver2D::GetDivergence(const std::vector<double>& u, const std::vector<double>& v) {
    Device& D = Dev();
    double* d_u = D.in("u", u);
    check(asd2_mac_divergence(D.ctx, d_u, D.in("v", v), D.scratch("div")), "GetDivergence");
    return D.result("div");
}

void MACSolver2D::GetDivergence(const std::vector<double>& u, const std::vector<double>& v, std::vector<double>& div) {
    Device& D = Dev();
    double* d_u = D.in("u", u);
    double* d_v = D.in("v", v);
    check(asd2_mac_divergence(D.ctx, d_u, d_v, D.out("div", div)), "GetDivergence");
    D.wrote("div", div);
}

int MACSolver2D::UpdateVel(std::vector<double>& u, std::vector<double>& v, const std::vector<double>& us,
                           const std::vector<double>& vs, const std::vector<double>& ps, const std::vector<double>& ls,
                           const std::vector<double>&, const std::vector<double>& H) {
    Device& D = Dev();
    asd2_fluid f;
    FluidParams(&f);
    if (m_kappa.size() != static_cast<size_t>(kArrSize)) m_kappa.assign(static_cast<size_t>(kArrSize), 0.0);
    double* d_us = D.in("us", us);
    double* d_vs = D.in("vs", vs);
    double* d_ps = D.in("ps", ps);
    double* d_ls = D.in("ls", ls);
    double* d_k = D.in("kappa", m_kappa);
    double* d_H = D.in("H", H);
    double* du = D.inout("u", u);
    double* dv = D.inout("v", v);
    check(asd2_mac_update_vel(D.ctx, &f, du, dv, d_us, d_vs, d_ps, d_ls, d_k, d_H), "UpdateVel");
    D.wrote("u", u);
    D.wrote("v", v);
    return 0;
}

double MACSolver2D::UpdateDt(const std::vector<double>& u, const std::vector<double>& v) {
    Device& D = Dev();
    asd2_fluid f;
    FluidParams(&f);
    double dt = 0.0;
    double* d_u = D.in("u", u);
    check(asd2_mac_update_dt(D.ctx, &f, d_u, D.in("v", v), &dt), "UpdateDt");
    if (std::isnan(dt) || std::isinf(dt)) std::cout << "dt nan/inf error" << std::endl;
    return dt;
}

// ------------------------------------------------------------------ boundary conditions
// The ghost fills run on the solver's own context (configured with m_BC's rules), so a resident
// vector stays on this context's stream.  An LS-free pressure side extrapolates with the BC
// object's grid spacing, which the reference's MACSolver2D never sets (src/mac2d.cpp:1884-1904
// build m_BC without dx, dy): those fills go through m_BC, whose rounding follows the reference.
int MACSolver2D::ApplyBC(int grid_kind, std::vector<double>& arr, const char* where) {
    Device& D = Dev();
    if (grid_kind == ASD2_GRID_P) {
        for (BC2D k : {m_BC->m_BC_PW, m_BC->m_BC_PE, m_BC->m_BC_PS, m_BC->m_BC_PN})
            if (k == BC2D::LSFREE) return m_BC->ApplyBC_P_2D(arr);
    }
    check(asd2_bc_apply(D.ctx, grid_kind, D.inout("bc", arr)), where);
    D.wrote("bc", arr);
    return 0;
}

int MACSolver2D::SetBC_U_2D(std::string W, std::string E, std::string S, std::string N) {
    if (!m_BC) m_BC = std::make_shared<BoundaryCondition2D>(kNx, kNy, kNumBCGrid);
    return m_BC->SetBC_U_2D(W, E, S, N);
}
int MACSolver2D::SetBC_V_2D(std::string W, std::string E, std::string S, std::string N) {
    if (!m_BC) m_BC = std::make_shared<BoundaryCondition2D>(kNx, kNy, kNumBCGrid);
    return m_BC->SetBC_V_2D(W, E, S, N);
}
int MACSolver2D::SetBC_P_2D(std::string W, std::string E, std::string S, std::string N) {
    if (!m_BC) m_BC = std::make_shared<BoundaryCondition2D>(kNx, kNy, kNumBCGrid);
    return m_BC->SetBC_P_2D(W, E, S, N);
}
int MACSolver2D::ApplyBC_U_2D(std::vector<double>& arr) { return ApplyBC(ASD2_GRID_U, arr, "ApplyBC_U_2D"); }
int MACSolver2D::ApplyBC_V_2D(std::vector<double>& arr) { return ApplyBC(ASD2_GRID_V, arr, "ApplyBC_V_2D"); }
int MACSolver2D::ApplyBC_P_2D(std::vector<double>& arr) { return ApplyBC(ASD2_GRID_P, arr, "ApplyBC_P_2D"); }
void MACSolver2D::SetBCConstantUW(double c) { Dev(); m_BC->SetBCConstantUW(c); }
void MACSolver2D::SetBCConstantUE(double c) { Dev(); m_BC->SetBCConstantUE(c); }
void MACSolver2D::SetBCConstantUS(double c) { Dev(); m_BC->SetBCConstantUS(c); }
void MACSolver2D::SetBCConstantUN(double c) { Dev(); m_BC->SetBCConstantUN(c); }
void MACSolver2D::SetBCConstantVW(double c) { Dev(); m_BC->SetBCConstantVW(c); }
void MACSolver2D::SetBCConstantVE(double c) { Dev(); m_BC->SetBCConstantVE(c); }
void MACSolver2D::SetBCConstantVS(double c) { Dev(); m_BC->SetBCConstantVS(c); }
void MACSolver2D::SetBCConstantVN(double c) { Dev(); m_BC->SetBCConstantVN(c); }
void MACSolver2D::SetBCConstantPW(double c) { Dev(); m_BC->SetBCConstantPW(c); }
void MACSolver2D::SetBCConstantPE(double c) { Dev(); m_BC->SetBCConstantPE(c); }
void MACSolver2D::SetBCConstantPS(double c) { Dev(); m_BC->SetBCConstantPS(c); }
void MACSolver2D::SetBCConstantPN(double c) { Dev(); m_BC->SetBCConstantPN(c); }

// ------------------------------------------------------------------ utilities and output
// Declared by the reference header without a definition: the Thomas algorithm for a
// tridiagonal system (a sub-, b main, c super-diagonal, d right-hand side -> solution).
void MACSolver2D::TDMA(double* a, double* b, double* c, double* d, int n) {
    for (int k = 1; k < n; k++) {
        const double m = a[k] / b[k - 1];
        b[k] -= m * c[k - 1];
        d[k] -= m * d[k - 1];
    }
    d[n - 1] /= b[n - 1];
    for (int k = n - 2; k >= 0; k--) d[k] = (d[k] - c[k] * d[k + 1]) / b[k];
}

int MACSolver2D::sign(const double& val) { return (val > 0) - (val < 0); }

int MACSolver2D::SetPLTType(PLTTYPE type) {
    m_PLTType = type;
    return 0;
}

// src/mac2d.cpp:1990-2050: Tecplot ASCII zones of cell-centred U, V, LS, PS and of LS, DIV,
// PS.  The binary .szplt writer needs TECIO, which is not linked here.
int MACSolver2D::OutRes(const int iter, const double curTime, const std::string fname_vel_base,
                        const std::string fname_div_base, const std::vector<double>& u, const std::vector<double>& v,
                        const std::vector<double>& ps, const std::vector<double>& ls) {
    if (m_PLTType == PLTTYPE::BINARY || m_PLTType == PLTTYPE::BOTH) {
        static bool told = false;
        if (!told) std::cerr << "MACSolver2D::OutRes: Tecplot binary output (TECIO) is not available" << std::endl;
        told = true;
    }
    if (m_PLTType != PLTTYPE::ASCII && m_PLTType != PLTTYPE::BOTH) return 0;
    const std::string fname_vel(fname_vel_base + "_ASCII.plt");
    const std::string fname_div(fname_div_base + "_ASCII.plt");
    std::ofstream outF;
    if (m_iter == 0) {
        outF.open(fname_vel.c_str(), std::ios::out);
        outF << "TITLE = VEL" << std::endl;
        outF << "VARIABLES = \"X\", \"Y\", \"U\", \"V\", \"LS\", \"PS\" " << std::endl;
        outF.close();
        outF.open(fname_div.c_str(), std::ios::out);
        outF << "TITLE = DIV" << std::endl;
        outF << "VARIABLES = \"X\", \"Y\", \"LS\", \"DIV\", \"PS\" " << std::endl;
        outF.close();
    }
    const std::vector<double> div = GetDivergence(u, v);
    for (const std::vector<double>* a : {&u, &v, &ps, &ls}) SolverDevice::host_read(*a);
    auto zone = [&](std::ofstream& o) {
        o << "ZONE T=\"" << iter << "\", I=" << kNx << ", J=" << kNy << ", SOLUTIONTIME=" << curTime
          << ", STRANDID=" << iter + 1 << std::endl;
    };
    outF.open(fname_vel.c_str(), std::ios::app);
    zone(outF);
    for (int j = kNumBCGrid; j < kNy + kNumBCGrid; j++)
        for (int i = kNumBCGrid; i < kNx + kNumBCGrid; i++)
            outF << kBaseX + (i + 0.5 - kNumBCGrid) * kDx << "," << kBaseY + (j + 0.5 - kNumBCGrid) * kDy << ","
                 << (u[idx(i, j)] + u[idx(i + 1, j)]) * 0.5 << "," << (v[idx(i, j)] + v[idx(i, j + 1)]) * 0.5 << ","
                 << ls[idx(i, j)] << "," << ps[idx(i, j)] << std::endl;
    outF.close();
    outF.open(fname_div.c_str(), std::ios::app);
    zone(outF);
    for (int j = kNumBCGrid; j < kNy + kNumBCGrid; j++)
        for (int i = kNumBCGrid; i < kNx + kNumBCGrid; i++)
            outF << kBaseX + (i + 0.5 - kNumBCGrid) * kDx << "," << kBaseY + (j + 0.5 - kNumBCGrid) * kDy << ","
                 << ls[idx(i, j)] << "," << div[idx(i, j)] << "," << ps[idx(i, j)] << std::endl;
    outF.close();
    return 0;
}

int MACSolver2D::OutResClose() { return 0; }
