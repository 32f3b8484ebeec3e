// bc2d.cpp -- drop-in BoundaryCondition2D on the gfx950 ghost-fill kernels.
// Reference: src/bc2d.cpp.  Side-specific methods (BC_NeumannPW, ...) run the same
// kernels with the rule on that side only.
#include "bc2d.h"

#include "host_util.h"

using namespace asd2host;

BoundaryCondition2D::BoundaryCondition2D(int nx, int ny, int num_bc_grid)
    : BoundaryCondition2D(nx, ny, -1.0, -1.0, num_bc_grid) {}

BoundaryCondition2D::BoundaryCondition2D(int nx, int ny, double dx, double dy, int num_bc_grid)
    : kNx(nx), kNy(ny), kNumBCGrid(num_bc_grid), kDx(dx), kDy(dy), m_AmbientPressure(0.0),
      m_BC_UW(BC2D::CUSTOM), m_BC_UE(BC2D::CUSTOM), m_BC_US(BC2D::CUSTOM), m_BC_UN(BC2D::CUSTOM),
      m_BC_VW(BC2D::CUSTOM), m_BC_VE(BC2D::CUSTOM), m_BC_VS(BC2D::CUSTOM), m_BC_VN(BC2D::CUSTOM),
      m_BC_PW(BC2D::CUSTOM), m_BC_PE(BC2D::CUSTOM), m_BC_PS(BC2D::CUSTOM), m_BC_PN(BC2D::CUSTOM),
      m_BC_DirichletConstantUW(0), m_BC_DirichletConstantUE(0), m_BC_DirichletConstantUS(0),
      m_BC_DirichletConstantUN(0), m_BC_DirichletConstantVW(0), m_BC_DirichletConstantVE(0),
      m_BC_DirichletConstantVS(0), m_BC_DirichletConstantVN(0), m_BC_DirichletConstantPW(0),
      m_BC_DirichletConstantPE(0), m_BC_DirichletConstantPS(0), m_BC_DirichletConstantPN(0) {}

BoundaryCondition2D::~BoundaryCondition2D() {
    if (m_ctx) {
        mirror::release(m_ctx);
        asd2_destroy(m_ctx);
    }
}

namespace {
// the ghost fill of arr on ctx: through arr's mirror in resident mode, else upload / download
void fill_ghosts(asd2_ctx* ctx, int grid_kind, std::vector<double>& arr, const char* where) {
    if (mirror::resident()) {
        check(asd2_bc_apply(ctx, grid_kind, mirror::acquire(ctx, arr, true)), where);
        mirror::written(ctx, arr);
        return;
    }
    DevVec d(arr);
    check(asd2_bc_apply(ctx, grid_kind, d.p()), where);
    d.get(arr);
}
}  // namespace

int BoundaryCondition2D::SetBC_U_2D(std::string W, std::string E, std::string S, std::string N) {
    for (auto* s : {&W, &E, &S, &N})
        if (lower(*s) == "pressure") {
            std::cout << "Boundary Condition Error : constant pressure BC is not allowed at velocity" << std::endl;
            std::exit(1);
        }
    m_BC_UW = parse_bc(W, false);
    m_BC_UE = parse_bc(E, false);
    m_BC_US = parse_bc(S, false);
    m_BC_UN = parse_bc(N, false);
    // wall sets a zero constant on every side, axisym on the normal-velocity sides
    if (m_BC_UW == BC2D::WALL || m_BC_UW == BC2D::AXISYM) m_BC_DirichletConstantUW = 0.0;
    if (m_BC_UE == BC2D::WALL || m_BC_UE == BC2D::AXISYM) m_BC_DirichletConstantUE = 0.0;
    if (m_BC_US == BC2D::WALL) m_BC_DirichletConstantUS = 0.0;
    if (m_BC_UN == BC2D::WALL) m_BC_DirichletConstantUN = 0.0;
    return 0;
}

int BoundaryCondition2D::SetBC_V_2D(std::string W, std::string E, std::string S, std::string N) {
    for (auto* s : {&W, &E, &S, &N})
        if (lower(*s) == "pressure") {
            std::cout << "Boundary Condition Error : constant pressure BC is not allowed at velocity" << std::endl;
            std::exit(1);
        }
    m_BC_VW = parse_bc(W, false);
    m_BC_VE = parse_bc(E, false);
    m_BC_VS = parse_bc(S, false);
    m_BC_VN = parse_bc(N, false);
    if (m_BC_VW == BC2D::WALL) m_BC_DirichletConstantVW = 0.0;
    if (m_BC_VE == BC2D::WALL) m_BC_DirichletConstantVE = 0.0;
    if (m_BC_VS == BC2D::WALL || m_BC_VS == BC2D::AXISYM) m_BC_DirichletConstantVS = 0.0;
    if (m_BC_VN == BC2D::WALL || m_BC_VN == BC2D::AXISYM) m_BC_DirichletConstantVN = 0.0;
    return 0;
}

int BoundaryCondition2D::SetBC_P_2D(std::string W, std::string E, std::string S, std::string N) {
    m_BC_PW = parse_bc(W, true);
    m_BC_PE = parse_bc(E, true);
    m_BC_PS = parse_bc(S, true);
    m_BC_PN = parse_bc(N, true);
    return 0;
}

void BoundaryCondition2D::Configure(asd2_ctx* ctx) const {
    const int u[4] = {code(m_BC_UW), code(m_BC_UE), code(m_BC_US), code(m_BC_UN)};
    const int v[4] = {code(m_BC_VW), code(m_BC_VE), code(m_BC_VS), code(m_BC_VN)};
    const int p[4] = {code(m_BC_PW), code(m_BC_PE), code(m_BC_PS), code(m_BC_PN)};
    const double cu[4] = {m_BC_DirichletConstantUW, m_BC_DirichletConstantUE, m_BC_DirichletConstantUS,
                          m_BC_DirichletConstantUN};
    const double cv[4] = {m_BC_DirichletConstantVW, m_BC_DirichletConstantVE, m_BC_DirichletConstantVS,
                          m_BC_DirichletConstantVN};
    const double cp[4] = {m_BC_DirichletConstantPW, m_BC_DirichletConstantPE, m_BC_DirichletConstantPS,
                          m_BC_DirichletConstantPN};
    check(asd2_set_bc(ctx, ASD2_GRID_U, u, cu), "SetBC U");
    check(asd2_set_bc(ctx, ASD2_GRID_V, v, cv), "SetBC V");
    check(asd2_set_bc(ctx, ASD2_GRID_P, p, cp), "SetBC P");
}

void BoundaryCondition2D::ApplyAll(int grid_kind, std::vector<double>& arr) {
    if (!m_ctx) m_ctx = make_ctx(kNx, kNy, kNumBCGrid, 0.0, 0.0, kDx, kDy);
    Configure(m_ctx);
    fill_ghosts(m_ctx, grid_kind, arr, "ApplyBC");
}

void BoundaryCondition2D::ApplySide(int grid_kind, int side, int kind, std::vector<double>& arr) {
    if (!m_ctx) m_ctx = make_ctx(kNx, kNy, kNumBCGrid, 0.0, 0.0, kDx, kDy);
    Configure(m_ctx);
    int kinds[4] = {ASD2_BC_CUSTOM, ASD2_BC_CUSTOM, ASD2_BC_CUSTOM, ASD2_BC_CUSTOM};
    kinds[side] = kind;
    const double* all[3][4] = {
        {&m_BC_DirichletConstantPW, &m_BC_DirichletConstantPE, &m_BC_DirichletConstantPS, &m_BC_DirichletConstantPN},
        {&m_BC_DirichletConstantUW, &m_BC_DirichletConstantUE, &m_BC_DirichletConstantUS, &m_BC_DirichletConstantUN},
        {&m_BC_DirichletConstantVW, &m_BC_DirichletConstantVE, &m_BC_DirichletConstantVS, &m_BC_DirichletConstantVN}};
    double consts[4];
    for (int s = 0; s < 4; s++) consts[s] = *all[grid_kind][s];
    check(asd2_set_bc(m_ctx, grid_kind, kinds, consts), "SetBC side");
    fill_ghosts(m_ctx, grid_kind, arr, "BC side");
}

int BoundaryCondition2D::ApplyBC_U_2D(std::vector<double>& arr) { ApplyAll(ASD2_GRID_U, arr); return 0; }
int BoundaryCondition2D::ApplyBC_V_2D(std::vector<double>& arr) { ApplyAll(ASD2_GRID_V, arr); return 0; }
int BoundaryCondition2D::ApplyBC_P_2D(std::vector<double>& arr) { ApplyAll(ASD2_GRID_P, arr); return 0; }

// dispatching single sides (reference src/bc2d.cpp:355-573)
void BoundaryCondition2D::BC_UW(std::vector<double>& a) { ApplySide(ASD2_GRID_U, 0, code(m_BC_UW), a); }
void BoundaryCondition2D::BC_UE(std::vector<double>& a) { ApplySide(ASD2_GRID_U, 1, code(m_BC_UE), a); }
void BoundaryCondition2D::BC_US(std::vector<double>& a) { ApplySide(ASD2_GRID_U, 2, code(m_BC_US), a); }
void BoundaryCondition2D::BC_UN(std::vector<double>& a) { ApplySide(ASD2_GRID_U, 3, code(m_BC_UN), a); }
void BoundaryCondition2D::BC_VW(std::vector<double>& a) { ApplySide(ASD2_GRID_V, 0, code(m_BC_VW), a); }
void BoundaryCondition2D::BC_VE(std::vector<double>& a) { ApplySide(ASD2_GRID_V, 1, code(m_BC_VE), a); }
void BoundaryCondition2D::BC_VS(std::vector<double>& a) { ApplySide(ASD2_GRID_V, 2, code(m_BC_VS), a); }
void BoundaryCondition2D::BC_VN(std::vector<double>& a) { ApplySide(ASD2_GRID_V, 3, code(m_BC_VN), a); }
void BoundaryCondition2D::BC_PW(std::vector<double>& a) { ApplySide(ASD2_GRID_P, 0, code(m_BC_PW), a); }
void BoundaryCondition2D::BC_PE(std::vector<double>& a) { ApplySide(ASD2_GRID_P, 1, code(m_BC_PE), a); }
void BoundaryCondition2D::BC_PS(std::vector<double>& a) { ApplySide(ASD2_GRID_P, 2, code(m_BC_PS), a); }
void BoundaryCondition2D::BC_PN(std::vector<double>& a) { ApplySide(ASD2_GRID_P, 3, code(m_BC_PN), a); }

// fixed rules (reference src/bc2d.cpp:575-934)
#define ASD2_SIDE(name, grid, side, kind) \
    void BoundaryCondition2D::name(std::vector<double>& a) { ApplySide(grid, side, kind, a); }
ASD2_SIDE(BC_PeriodicUW, ASD2_GRID_U, 0, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicUE, ASD2_GRID_U, 1, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicUS, ASD2_GRID_U, 2, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicUN, ASD2_GRID_U, 3, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicVW, ASD2_GRID_V, 0, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicVE, ASD2_GRID_V, 1, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicVS, ASD2_GRID_V, 2, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicVN, ASD2_GRID_V, 3, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicPW, ASD2_GRID_P, 0, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicPE, ASD2_GRID_P, 1, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicPS, ASD2_GRID_P, 2, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_PeriodicPN, ASD2_GRID_P, 3, ASD2_BC_PERIODIC)
ASD2_SIDE(BC_NeumannUW, ASD2_GRID_U, 0, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannUE, ASD2_GRID_U, 1, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannUS, ASD2_GRID_U, 2, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannUN, ASD2_GRID_U, 3, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannVW, ASD2_GRID_V, 0, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannVE, ASD2_GRID_V, 1, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannVS, ASD2_GRID_V, 2, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannVN, ASD2_GRID_V, 3, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannPW, ASD2_GRID_P, 0, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannPE, ASD2_GRID_P, 1, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannPS, ASD2_GRID_P, 2, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_NeumannPN, ASD2_GRID_P, 3, ASD2_BC_NEUMANN)
ASD2_SIDE(BC_DirichletUW, ASD2_GRID_U, 0, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletUE, ASD2_GRID_U, 1, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletUS, ASD2_GRID_U, 2, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletUN, ASD2_GRID_U, 3, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletVW, ASD2_GRID_V, 0, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletVE, ASD2_GRID_V, 1, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletVS, ASD2_GRID_V, 2, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletVN, ASD2_GRID_V, 3, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletPW, ASD2_GRID_P, 0, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletPE, ASD2_GRID_P, 1, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletPS, ASD2_GRID_P, 2, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_DirichletPN, ASD2_GRID_P, 3, ASD2_BC_DIRICHLET)
ASD2_SIDE(BC_LSFreeBoundaryPW, ASD2_GRID_P, 0, ASD2_BC_LSFREE)
ASD2_SIDE(BC_LSFreeBoundaryPE, ASD2_GRID_P, 1, ASD2_BC_LSFREE)
ASD2_SIDE(BC_LSFreeBoundaryPS, ASD2_GRID_P, 2, ASD2_BC_LSFREE)
ASD2_SIDE(BC_LSFreeBoundaryPN, ASD2_GRID_P, 3, ASD2_BC_LSFREE)
#undef ASD2_SIDE

void BoundaryCondition2D::SetBCConstantUW(double c) { m_BC_DirichletConstantUW = c; }
void BoundaryCondition2D::SetBCConstantUE(double c) { m_BC_DirichletConstantUE = c; }
void BoundaryCondition2D::SetBCConstantUS(double c) { m_BC_DirichletConstantUS = c; }
void BoundaryCondition2D::SetBCConstantUN(double c) { m_BC_DirichletConstantUN = c; }
void BoundaryCondition2D::SetBCConstantVW(double c) { m_BC_DirichletConstantVW = c; }
void BoundaryCondition2D::SetBCConstantVE(double c) { m_BC_DirichletConstantVE = c; }
void BoundaryCondition2D::SetBCConstantVS(double c) { m_BC_DirichletConstantVS = c; }
void BoundaryCondition2D::SetBCConstantVN(double c) { m_BC_DirichletConstantVN = c; }
void BoundaryCondition2D::SetBCConstantPW(double c) { m_BC_DirichletConstantPW = c; }
void BoundaryCondition2D::SetBCConstantPE(double c) { m_BC_DirichletConstantPE = c; }
void BoundaryCondition2D::SetBCConstantPS(double c) { m_BC_DirichletConstantPS = c; }
void BoundaryCondition2D::SetBCConstantPN(double c) { m_BC_DirichletConstantPN = c; }

// reference src/bc2d.cpp:851-866
void BoundaryCondition2D::SetAmbientPressure(double p) {
    m_AmbientPressure = p;
    auto amb = [](BC2D k) { return k == BC2D::INLET || k == BC2D::OUTLET || k == BC2D::PRESSURE; };
    if (amb(m_BC_PW)) SetBCConstantPW(p);
    if (amb(m_BC_PE)) SetBCConstantPE(p);
    if (amb(m_BC_PS)) SetBCConstantPS(p);
    if (amb(m_BC_PN)) SetBCConstantPN(p);
}
