// levelset2d.cpp -- drop-in LevelSetSolver2D on the gfx950 level-set kernels.
// Reference: src/levelset2d.cpp.  Every std::vector method uploads its inputs into the
// solver's device buffers, runs the asd2_ls_* entry point and downloads the result; the
// pseudo-time loops of the reinitialisations become an iteration count computed with the
// reference's own floating-point loop (asd2_ls_reinit_iterations).
#include "levelset2d.h"

#include <limits>

#include "host_util.h"

using namespace asd2host;

struct LevelSetSolver2D::Device {
    asd2_ctx* ctx = nullptr;
    std::unique_ptr<DevVec> buf[5];
    size_t n = 0;
    ~Device() {
        for (auto& b : buf) b.reset();
        if (ctx) asd2_destroy(ctx);
    }
    double* get(int k) {
        if (!buf[k]) buf[k].reset(new DevVec(n));
        return buf[k]->p();
    }
    double* put(int k, const std::vector<double>& h) {
        if (h.size() != n) {
            std::cerr << "LevelSetSolver2D: array of " << h.size() << " values, expected " << n << std::endl;
            std::exit(1);
        }
        double* p = get(k);
        buf[k]->put(h);
        return p;
    }
    void take(int k, std::vector<double>& h) { buf[k]->get(h); }
};

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
    double* d_ls = D.put(0, ls);
    check(asd2_ls_advect(D.ctx, d_ls, D.put(1, u), D.put(2, v), dt), "Solve_LevelSet_2D");
    D.take(0, ls);
    return 0;
}

int LevelSetSolver2D::Solve_LevelSet_2D_Device(double* d_ls, const double* d_u, const double* d_v, double dt) {
    check(asd2_ls_advect(Dev().ctx, d_ls, d_u, d_v, dt), "Solve_LevelSet_2D");
    return 0;
}

std::vector<double> LevelSetSolver2D::HJWENO5_LS_2D(std::vector<double>& ls, const std::vector<double>& u,
                                                    const std::vector<double>& v) {
    Device& D = Dev();
    double* d_ls = D.put(0, ls);
    check(asd2_ls_weno_rhs(D.ctx, d_ls, D.put(1, u), D.put(2, v), D.get(3)), "HJWENO5_LS_2D");
    D.take(0, ls);  // the BC fill of ls is visible to the caller, as in the reference (:99)
    std::vector<double> L;
    D.take(3, L);
    return L;
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
int LevelSetSolver2D::Reinit_Original_2D(std::vector<double>& ls) {
    Device& D = Dev();
    double* d = D.put(0, ls);
    check(asd2_ls_reinit_original(D.ctx, d, ReinitIterations()), "Reinit_Original_2D");
    D.take(0, ls);
    m_atime = ReinitIterations() * kAdt;
    return 0;
}

int LevelSetSolver2D::Reinit_Sussman_2D(std::vector<double>& ls) {
    Device& D = Dev();
    check(Reinit_Sussman_2D_Device(D.put(0, ls)), "Reinit_Sussman_2D");
    D.take(0, ls);
    return 0;
}

int LevelSetSolver2D::Reinit_Sussman_2D_Device(double* d_ls) {
    check(asd2_ls_reinit_sussman(Dev().ctx, d_ls, ReinitIterations()), "Reinit_Sussman_2D");
    m_atime = ReinitIterations() * kAdt;
    return 0;
}

int LevelSetSolver2D::Reinit_MinRK2_2D(std::vector<double>& ls) {
    Device& D = Dev();
    check(Reinit_MinRK2_2D_Device(D.put(0, ls)), "Reinit_MinRK2_2D");
    D.take(0, ls);
    return 0;
}

int LevelSetSolver2D::Reinit_MinRK2_2D_Device(double* d_ls) {
    check(asd2_ls_reinit_minrk2(Dev().ctx, d_ls, ReinitIterations()), "Reinit_MinRK2_2D");
    m_atime = ReinitIterations() * kAdt;
    return 0;
}

// src/levelset2d.cpp:377-431: the loop runs until the pseudo time reaches max(dx nx, dy ny)
int LevelSetSolver2D::FirstTimeOnlyReinit_Sussman_2D(std::vector<double>& ls) {
    Device& D = Dev();
    const int iters = asd2_ls_reinit_iterations(kEps, std::max(kDx * kNx, kDy * kNy));
    double* d = D.put(0, ls);
    check(asd2_ls_reinit_first_time(D.ctx, d, iters), "FirstTimeOnlyReinit_Sussman_2D");
    D.take(0, ls);
    m_atime = iters * kAdt;
    return 0;
}

// Declared by the reference header (include/levelset2d.h:30) without a definition; here it
// is the first-time Sussman reinitialisation.
int LevelSetSolver2D::Reinit_Sussman_FirstTime_2D(std::vector<double>& ls) { return FirstTimeOnlyReinit_Sussman_2D(ls); }

std::vector<double> LevelSetSolver2D::GetSussmanReinitConstraint(const std::vector<double>& ls,
                                                                 const std::vector<double>& lsInit,
                                                                 const std::vector<double>& heavisideDeriv) {
    Device& D = Dev();
    check(asd2_ls_sussman_constraint(D.ctx, D.put(0, ls), D.put(1, lsInit), D.put(2, heavisideDeriv), D.get(3)),
          "GetSussmanReinitConstraint");
    std::vector<double> out;
    D.take(3, out);
    return out;
}

// ------------------------------------------------------------------ |grad phi| and pointwise helpers
std::vector<double> LevelSetSolver2D::HJWENO5_DerivAbsLS_2D(const std::vector<double>& ls,
                                                            const std::vector<double>& lsInit) {
    Device& D = Dev();
    check(asd2_ls_weno_absgrad(D.ctx, D.put(0, ls), D.put(1, lsInit), D.get(3)), "HJWENO5_DerivAbsLS_2D");
    std::vector<double> out;
    D.take(3, out);
    return out;
}

std::vector<double> LevelSetSolver2D::ENO_DerivAbsLS_2D(const std::vector<double>& ls,
                                                        const std::vector<double>& lsInit) {
    Device& D = Dev();
    check(asd2_ls_eno_absgrad(D.ctx, D.put(0, ls), D.put(1, lsInit), D.get(3)), "ENO_DerivAbsLS_2D");
    std::vector<double> out;
    D.take(3, out);
    return out;
}

std::vector<double> LevelSetSolver2D::SubcellENO_Min_DerivAbsLS_2D(const std::vector<double>& ls,
                                                                   const std::vector<double>& lsInit) {
    Device& D = Dev();
    check(asd2_ls_subcell_absgrad(D.ctx, D.put(0, ls), D.put(1, lsInit), D.get(3)), "SubcellENO_Min_DerivAbsLS_2D");
    std::vector<double> out;
    D.take(3, out);
    return out;
}

std::vector<double> LevelSetSolver2D::UpdateHeavisideFuncDeriv(const std::vector<double>& ls) {
    Device& D = Dev();
    check(asd2_ls_heaviside_deriv(D.ctx, D.put(0, ls), D.get(3)), "UpdateHeavisideFuncDeriv");
    std::vector<double> out;
    D.take(3, out);
    return out;
}

std::vector<double> LevelSetSolver2D::GetSmoothedSignFunc(const std::vector<double>& ls) {
    Device& D = Dev();
    check(asd2_ls_smoothed_sign(D.ctx, D.put(0, ls), D.get(3)), "GetSmoothedSignFunc");
    std::vector<double> out;
    D.take(3, out);
    return out;
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
int LevelSetSolver2D::ApplyBC_U_2D(std::vector<double>& arr) { EnsureBC(); return m_BC->ApplyBC_U_2D(arr); }
int LevelSetSolver2D::ApplyBC_V_2D(std::vector<double>& arr) { EnsureBC(); return m_BC->ApplyBC_V_2D(arr); }
int LevelSetSolver2D::ApplyBC_P_2D(std::vector<double>& arr) { EnsureBC(); return m_BC->ApplyBC_P_2D(arr); }
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
