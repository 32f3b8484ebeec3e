// levelset2d.cpp -- drop-in LevelSetSolver2D on the gfx950 level-set kernels.
// Reference: src/levelset2d.cpp.  By default every std::vector method uploads its inputs into the
// solver's device buffers, runs the asd2_ls_* entry point and downloads the result; in resident
// mode (asd2_resident.h) the level set stays on the device between calls.  The pseudo-time loops
// of the reinitialisations become an iteration count computed with the reference's own
// floating-point loop (asd2_ls_reinit_iterations).
#include "levelset2d.h"

#include <limits>

#include "host_util.h"

using namespace asd2host;

struct LevelSetSolver2D::Device : SolverDevice {};

// src/levelset2d.cpp:3-13: eps = min(dx, dy), thickness 1.5 eps, dtau = eps / 2,
// default end time 1.5 * thickness
LevelSetSolver2D::LevelSetSolver2D(int nx, int ny, int num_bc_grid, double baseX, double baseY, double dx, double dy)
    : LevelSetSolver2D(nx, ny, num_bc_grid, baseX, baseY, dx, dy, 1.5 * (1.5 * std::min(dx, dy))) {}

LevelSetSolver2D::LevelSetSolver2D(int nx, int ny, int num_bc_grid, double baseX, double baseY, double dx, double dy,
                                   double maxTime)
    : kEps(std::min(dx, dy)), kThickness(1.5 * kEps), kMaxATime(maxTime), kNx(nx), kNy(ny),
      kNumBCGrid(num_bc_grid), kENOSpatialOrder(2),
      kArrSize(static_cast<int64_t>(nx + 2 * num_bc_grid) * static_cast<int64_t>(ny + 2 * num_bc_grid)),
      kBaseX(baseX), kBaseY(baseY), kDx(dx), kDy(dy), m_atime(0.0), kAdt(kEps * 0.5) {}

LevelSetSolver2D::~LevelSetSolver2D() = default;

LevelSetSolver2D::Device& LevelSetSolver2D::Dev() {
    if (!m_dev) {
        m_dev.reset(new Device);
        m_dev->ctx = make_ctx(kNx, kNy, kNumBCGrid, kBaseX, kBaseY, kDx, kDy);
        m_dev->n = static_cast<size_t>(kArrSize);
        m_dev->who = "LevelSetSolver2D";
    }
    EnsureBC();
    m_BC->Configure(m_dev->ctx);
    return *m_dev;
}

// The reference creates its BoundaryCondition2D in SetBC_*_2D (src/levelset2d.cpp:925-953);
// a solver used before any SetBC call gets one whose sides are all CUSTOM (no ghost writes).
void LevelSetSolver2D::EnsureBC() {
    if (!m_BC) m_BC = std::make_shared<BoundaryCondition2D>(kNx, kNy, kDx, kDy, kNumBCGrid);
}

asd2_ctx* LevelSetSolver2D::Context() { return Dev().ctx; }

int LevelSetSolver2D::ReinitIterations() const { return asd2_ls_reinit_iterations(kEps, kMaxATime); }

// ------------------------------------------------------------------ transport
int LevelSetSolver2D::Solve_LevelSet_2D(std::vector<double>& ls, const std::vector<double>& u,
                                        const std::vector<double>& v, double dt) {
    Device& D = Dev();
    double* d_u = D.in("u", u);
    double* d_v = D.in("v", v);
    check(asd2_ls_advect(D.ctx, D.inout("ls", ls), d_u, d_v, dt), "Solve_LevelSet_2D");
    D.wrote("ls", ls);
    return 0;
}

int LevelSetSolver2D::Solve_LevelSet_2D_Device(double* d_ls, const double* d_u, const double* d_v, double dt) {
    check(asd2_ls_advect(Dev().ctx, d_ls, d_u, d_v, dt), "Solve_LevelSet_2D");
    return 0;
}

std::vector<double> LevelSetSolver2D::HJWENO5_LS_2D(std::vector<double>& ls, const std::vector<double>& u,
                                                    const std::vector<double>& v) {
    Device& D = Dev();
    double* d_u = D.in("u", u);
    double* d_v = D.in("v", v);
    check(asd2_ls_weno_rhs(D.ctx, D.inout("ls", ls), d_u, d_v, D.scratch("out")), "HJWENO5_LS_2D");
    D.wrote("ls", ls);  // the BC fill of ls is visible to the caller, as in the reference (:99)
    return D.result("out");
}

int LevelSetSolver2D::UnitHJWENO5(const std::vector<double>& F, std::vector<double>& FP, std::vector<double>& FM,
                                  const double d, const int n) {
    const size_t len = static_cast<size_t>(n + 2 * kNumBCGrid);
    if (F.size() < len) fatal("UnitHJWENO5: line shorter than n + 2g");
    DevVec dF(len), dP(len), dM(len);
    dF.put(F.data(), len);
    if (FP.size() >= len) dP.put(FP.data(), len);
    if (FM.size() >= len) dM.put(FM.data(), len);
    check(asd2_ls_unit_weno5(Dev().ctx, dF.p(), dP.p(), dM.p(), d, n, kNumBCGrid), "UnitHJWENO5");
    std::vector<double> hp = dP.get(), hm = dM.get();
    if (FP.size() < len) FP.resize(len);
    if (FM.size() < len) FM.resize(len);
    for (int i = kNumBCGrid; i < n + kNumBCGrid; i++) {
        FP[i] = hp[i];
        FM[i] = hm[i];
    }
    return 0;
}

// ------------------------------------------------------------------ reinitialisation
// method 0 = Reinit_Original_2D, 1 = Reinit_Sussman_2D, 2 = Reinit_MinRK2_2D, 3 = first-time Sussman
int LevelSetSolver2D::Reinit(int method, std::vector<double>& ls, int iterations, const char* where) {
    Device& D = Dev();
    double* d = D.inout("ls", ls);
    int rc = ASD2_OK;
    switch (method) {
        case 0: rc = asd2_ls_reinit_original(D.ctx, d, iterations); break;
        case 1: rc = asd2_ls_reinit_sussman(D.ctx, d, iterations); break;
        case 2: rc = asd2_ls_reinit_minrk2(D.ctx, d, iterations); break;
        default: rc = asd2_ls_reinit_first_time(D.ctx, d, iterations); break;
    }
    check(rc, where);
    D.wrote("ls", ls);
    m_atime = iterations * kAdt;
    return 0;
}

int LevelSetSolver2D::Reinit_Original_2D(std::vector<double>& ls) {
    return Reinit(0, ls, ReinitIterations(), "Reinit_Original_2D");
}

int LevelSetSolver2D::Reinit_Sussman_2D(std::vector<double>& ls) {
    return Reinit(1, ls, ReinitIterations(), "Reinit_Sussman_2D");
}

int LevelSetSolver2D::Reinit_Sussman_2D_Device(double* d_ls) {
    check(asd2_ls_reinit_sussman(Dev().ctx, d_ls, ReinitIterations()), "Reinit_Sussman_2D");
    m_atime = ReinitIterations() * kAdt;
    return 0;
}

int LevelSetSolver2D::Reinit_MinRK2_2D(std::vector<double>& ls) {
    return Reinit(2, ls, ReinitIterations(), "Reinit_MinRK2_2D");
}

int LevelSetSolver2D::Reinit_MinRK2_2D_Device(double* d_ls) {
    check(asd2_ls_reinit_minrk2(Dev().ctx, d_ls, ReinitIterations()), "Reinit_MinRK2_2D");
    m_atime = ReinitIterations() * kAdt;
    return 0;
}

// src/levelset2d.cpp:377-431: the loop runs until the pseudo time reaches max(dx nx, dy ny)
int LevelSetSolver2D::FirstTimeOnlyReinit_Sussman_2D(std::vector<double>& ls) {
    return Reinit(3, ls, asd2_ls_reinit_iterations(kEps, std::max(kDx * kNx, kDy * kNy)),
                  "FirstTimeOnlyReinit_Sussman_2D");
}

// Declared by the reference header (include/levelset2d.h:30) without a definition; here it
// is the first-time Sussman reinitialisation.
int LevelSetSolver2D::Reinit_Sussman_FirstTime_2D(std::vector<double>& ls) { return FirstTimeOnlyReinit_Sussman_2D(ls); }

std::vector<double> LevelSetSolver2D::GetSussmanReinitConstraint(const std::vector<double>& ls,
                                                                 const std::vector<double>& lsInit,
                                                                 const std::vector<double>& heavisideDeriv) {
    Device& D = Dev();
    double* d_ls = D.in("ls", ls);
    double* d_ls0 = D.in("ls0", lsInit);
    double* d_hd = D.in("hd", heavisideDeriv);
    check(asd2_ls_sussman_constraint(D.ctx, d_ls, d_ls0, d_hd, D.scratch("out")), "GetSussmanReinitConstraint");
    return D.result("out");
}

// ------------------------------------------------------------------ |grad phi| and pointwise helpers
// op 0 = HJWENO5 |grad phi|, 1 = ENO, 2 = subcell ENO, 3 = Heaviside derivative, 4 = smoothed sign
std::vector<double> LevelSetSolver2D::Pointwise(int op, const std::vector<double>& ls,
                                                const std::vector<double>* ls0, const char* where) {
    Device& D = Dev();
    double* d_ls = D.in("ls", ls);
    double* d_ls0 = ls0 ? D.in("ls0", *ls0) : nullptr;
    double* out = D.scratch("out");
    int rc = ASD2_OK;
    switch (op) {
        case 0: rc = asd2_ls_weno_absgrad(D.ctx, d_ls, d_ls0, out); break;
        case 1: rc = asd2_ls_eno_absgrad(D.ctx, d_ls, d_ls0, out); break;
        case 2: rc = asd2_ls_subcell_absgrad(D.ctx, d_ls, d_ls0, out); break;
        case 3: rc = asd2_ls_heaviside_deriv(D.ctx, d_ls, out); break;
        default: rc = asd2_ls_smoothed_sign(D.ctx, d_ls, out); break;
    }
    check(rc, where);
    return D.result("out");
}

std::vector<double> LevelSetSolver2D::HJWENO5_DerivAbsLS_2D(const std::vector<double>& ls,
                                                            const std::vector<double>& lsInit) {
    return Pointwise(0, ls, &lsInit, "HJWENO5_DerivAbsLS_2D");
}

std::vector<double> LevelSetSolver2D::ENO_DerivAbsLS_2D(const std::vector<double>& ls,
                                                        const std::vector<double>& lsInit) {
    return Pointwise(1, ls, &lsInit, "ENO_DerivAbsLS_2D");
}

std::vector<double> LevelSetSolver2D::SubcellENO_Min_DerivAbsLS_2D(const std::vector<double>& ls,
                                                                   const std::vector<double>& lsInit) {
    return Pointwise(2, ls, &lsInit, "SubcellENO_Min_DerivAbsLS_2D");
}

std::vector<double> LevelSetSolver2D::UpdateHeavisideFuncDeriv(const std::vector<double>& ls) {
    return Pointwise(3, ls, nullptr, "UpdateHeavisideFuncDeriv");
}

std::vector<double> LevelSetSolver2D::GetSmoothedSignFunc(const std::vector<double>& ls) {
    return Pointwise(4, ls, nullptr, "GetSmoothedSignFunc");
}

// src/levelset2d.cpp:895-912 (integer result, host only: it is not used by any solver path)
std::vector<int> LevelSetSolver2D::GetSignedLSNormalized(const std::vector<double>& v) {
    std::vector<int> r(v.size());
    for (int i = 0; i < kNx + 2 * kNumBCGrid; i++)
        for (int j = 0; j < kNy + 2 * kNumBCGrid; j++) {
            const double x = v[idx(i, j)];
            if (std::fabs(x) <= kEps)
                r[idx(i, j)] = static_cast<int>(2.0 * ((0.5 + std::cos(M_PI * x / kEps)) / kEps - 0.5));
            else if (x > 0)
                r[idx(i, j)] = 1;
            else if (x < 0)
                r[idx(i, j)] = -1;
        }
    return r;
}

int LevelSetSolver2D::sign(const double& val) { return (val > 0) - (val < 0); }

double LevelSetSolver2D::MinAbs(const double& val1, const double& val2) {
    return (std::fabs(val1) < std::fabs(val2)) ? val1 : val2;
}

double LevelSetSolver2D::MinMod(const double& val1, const double& val2) {
    if (val1 * val2 <= 0) return 0.0;
    if (val1 * val2 > 0) return (std::fabs(val1) <= std::fabs(val2)) ? val1 : val2;
    return std::numeric_limits<double>::quiet_NaN();  // a NaN operand
}

// ------------------------------------------------------------------ boundary conditions
int LevelSetSolver2D::SetBC_U_2D(std::string W, std::string E, std::string S, std::string N) {
    EnsureBC();
    return m_BC->SetBC_U_2D(W, E, S, N);
}
int LevelSetSolver2D::SetBC_V_2D(std::string W, std::string E, std::string S, std::string N) {
    EnsureBC();
    return m_BC->SetBC_V_2D(W, E, S, N);
}
int LevelSetSolver2D::SetBC_P_2D(std::string W, std::string E, std::string S, std::string N) {
    EnsureBC();
    return m_BC->SetBC_P_2D(W, E, S, N);
}
// the ghost fills run on the solver's own context, configured with m_BC's rules and the same
// grid spacing (src/levelset2d.cpp:925-953 build m_BC with dx, dy)
int LevelSetSolver2D::ApplyBC(int grid_kind, std::vector<double>& arr, const char* where) {
    Device& D = Dev();
    check(asd2_bc_apply(D.ctx, grid_kind, D.inout("bc", arr)), where);
    D.wrote("bc", arr);
    return 0;
}
int LevelSetSolver2D::ApplyBC_U_2D(std::vector<double>& arr) { return ApplyBC(ASD2_GRID_U, arr, "ApplyBC_U_2D"); }
int LevelSetSolver2D::ApplyBC_V_2D(std::vector<double>& arr) { return ApplyBC(ASD2_GRID_V, arr, "ApplyBC_V_2D"); }
int LevelSetSolver2D::ApplyBC_P_2D(std::vector<double>& arr) { return ApplyBC(ASD2_GRID_P, arr, "ApplyBC_P_2D"); }
void LevelSetSolver2D::SetBCConstantUW(double c) { EnsureBC(); m_BC->SetBCConstantUW(c); }
void LevelSetSolver2D::SetBCConstantUE(double c) { EnsureBC(); m_BC->SetBCConstantUE(c); }
void LevelSetSolver2D::SetBCConstantUS(double c) { EnsureBC(); m_BC->SetBCConstantUS(c); }
void LevelSetSolver2D::SetBCConstantUN(double c) { EnsureBC(); m_BC->SetBCConstantUN(c); }
void LevelSetSolver2D::SetBCConstantVW(double c) { EnsureBC(); m_BC->SetBCConstantVW(c); }
void LevelSetSolver2D::SetBCConstantVE(double c) { EnsureBC(); m_BC->SetBCConstantVE(c); }
void LevelSetSolver2D::SetBCConstantVS(double c) { EnsureBC(); m_BC->SetBCConstantVS(c); }
void LevelSetSolver2D::SetBCConstantVN(double c) { EnsureBC(); m_BC->SetBCConstantVN(c); }
void LevelSetSolver2D::SetBCConstantPW(double c) { EnsureBC(); m_BC->SetBCConstantPW(c); }
void LevelSetSolver2D::SetBCConstantPE(double c) { EnsureBC(); m_BC->SetBCConstantPE(c); }
void LevelSetSolver2D::SetBCConstantPS(double c) { EnsureBC(); m_BC->SetBCConstantPS(c); }
void LevelSetSolver2D::SetBCConstantPN(double c) { EnsureBC(); m_BC->SetBCConstantPN(c); }
