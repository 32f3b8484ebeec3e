// ref_wrapper.cpp -- C entry points over the reference's own LevelSetSolver2D and
// BoundaryCondition2D, compiled together with src/levelset2d.cpp and src/bc2d.cpp
// where they lie under the reference checkout (see oracle/Makefile).
//
// TEST INFRASTRUCTURE ONLY: the output library oracle/_ref/libasd2ref.so validates the
// C restatement in asd2_oracle.c and serves as the "reference" CPU baseline.  Nothing
// from the reference is copied here; this file only calls its public methods.
#include "levelset2d.h"
#include "bc2d.h"
#include "asd2_oracle.h"

#include <memory>
#include <string>
#include <vector>

namespace {

const char* bc_name(int k) {
    static const char* names[] = {"periodic", "neumann", "dirichlet", "axisym", "wall",
                                  "inlet", "outlet", "pressure", "lsfree", "custom"};
    return (k >= 0 && k <= 9) ? names[k] : "custom";
}

std::unique_ptr<LevelSetSolver2D> make_ls(const or_grid* gr, double max_time) {
    std::unique_ptr<LevelSetSolver2D> s;
    if (max_time > 0.0)
        s.reset(new LevelSetSolver2D(gr->nx, gr->ny, gr->g, 0.0, 0.0, gr->dx, gr->dy, max_time));
    else
        s.reset(new LevelSetSolver2D(gr->nx, gr->ny, gr->g, 0.0, 0.0, gr->dx, gr->dy));
    s->SetBC_P_2D(bc_name(gr->bc_p[0]), bc_name(gr->bc_p[1]), bc_name(gr->bc_p[2]), bc_name(gr->bc_p[3]));
    s->SetBCConstantPW(gr->c_p[0]);
    s->SetBCConstantPE(gr->c_p[1]);
    s->SetBCConstantPS(gr->c_p[2]);
    s->SetBCConstantPN(gr->c_p[3]);
    return s;
}

std::vector<double> vec(const double* p, int64_t n) { return std::vector<double>(p, p + n); }

void out(const std::vector<double>& v, double* p) { std::copy(v.begin(), v.end(), p); }

int64_t grid_size(const or_grid* gr) {
    return static_cast<int64_t>(gr->nx + 2 * gr->g) * static_cast<int64_t>(gr->ny + 2 * gr->g);
}

}  // namespace

extern "C" {

int ref_bc_p(const or_grid* gr, double* a) {
    BoundaryCondition2D bc(gr->nx, gr->ny, gr->dx, gr->dy, gr->g);
    bc.SetBC_P_2D(bc_name(gr->bc_p[0]), bc_name(gr->bc_p[1]), bc_name(gr->bc_p[2]), bc_name(gr->bc_p[3]));
    bc.SetBCConstantPW(gr->c_p[0]); bc.SetBCConstantPE(gr->c_p[1]);
    bc.SetBCConstantPS(gr->c_p[2]); bc.SetBCConstantPN(gr->c_p[3]);
    std::vector<double> v = vec(a, grid_size(gr));
    bc.ApplyBC_P_2D(v);
    out(v, a);
    return 0;
}

int ref_bc_u(const or_grid* gr, double* a) {
    BoundaryCondition2D bc(gr->nx, gr->ny, gr->dx, gr->dy, gr->g);
    bc.SetBC_U_2D(bc_name(gr->bc_u[0]), bc_name(gr->bc_u[1]), bc_name(gr->bc_u[2]), bc_name(gr->bc_u[3]));
    bc.SetBCConstantUW(gr->c_u[0]); bc.SetBCConstantUE(gr->c_u[1]);
    bc.SetBCConstantUS(gr->c_u[2]); bc.SetBCConstantUN(gr->c_u[3]);
    std::vector<double> v = vec(a, grid_size(gr));
    bc.ApplyBC_U_2D(v);
    out(v, a);
    return 0;
}

int ref_bc_v(const or_grid* gr, double* a) {
    BoundaryCondition2D bc(gr->nx, gr->ny, gr->dx, gr->dy, gr->g);
    bc.SetBC_V_2D(bc_name(gr->bc_v[0]), bc_name(gr->bc_v[1]), bc_name(gr->bc_v[2]), bc_name(gr->bc_v[3]));
    bc.SetBCConstantVW(gr->c_v[0]); bc.SetBCConstantVE(gr->c_v[1]);
    bc.SetBCConstantVS(gr->c_v[2]); bc.SetBCConstantVN(gr->c_v[3]);
    std::vector<double> v = vec(a, grid_size(gr));
    bc.ApplyBC_V_2D(v);
    out(v, a);
    return 0;
}

int ref_unit_hjweno5(const double* F, double* FP, double* FM, double d, int n, int g) {
    LevelSetSolver2D s(n, n, g, 0.0, 0.0, d, d);
    std::vector<double> f(F, F + n + 2 * g), fp(n + 2 * g, 0.0), fm(n + 2 * g, 0.0);
    s.UnitHJWENO5(f, fp, fm, d, n);
    out(fp, FP);
    out(fm, FM);
    return 0;
}

int ref_hjweno5_ls(const or_grid* gr, double* ls, const double* uc, const double* vc, double* L) {
    auto s = make_ls(gr, 0.0);
    const int64_t n = grid_size(gr);
    std::vector<double> l = vec(ls, n), u = vec(uc, n), v = vec(vc, n);
    std::vector<double> r = s->HJWENO5_LS_2D(l, u, v);
    out(l, ls);
    out(r, L);
    return 0;
}

int ref_solve_levelset(const or_grid* gr, double* ls, const double* u, const double* v, double dt) {
    auto s = make_ls(gr, 0.0);
    const int64_t n = grid_size(gr);
    std::vector<double> l = vec(ls, n), uu = vec(u, n), vv = vec(v, n);
    s->Solve_LevelSet_2D(l, uu, vv, dt);
    out(l, ls);
    return 0;
}

int ref_eno_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* o) {
    auto s = make_ls(gr, 0.0);
    const int64_t n = grid_size(gr);
    out(s->ENO_DerivAbsLS_2D(vec(ls, n), vec(ls0, n)), o);
    return 0;
}

int ref_subcell_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* o) {
    auto s = make_ls(gr, 0.0);
    const int64_t n = grid_size(gr);
    out(s->SubcellENO_Min_DerivAbsLS_2D(vec(ls, n), vec(ls0, n)), o);
    return 0;
}

int ref_smoothed_sign(const or_grid* gr, const double* ls, double* o) {
    auto s = make_ls(gr, 0.0);
    out(s->GetSmoothedSignFunc(vec(ls, grid_size(gr))), o);
    return 0;
}

int ref_heaviside_deriv(const or_grid* gr, const double* ls, double* o) {
    auto s = make_ls(gr, 0.0);
    out(s->UpdateHeavisideFuncDeriv(vec(ls, grid_size(gr))), o);
    return 0;
}

int ref_sussman_constraint(const or_grid* gr, const double* ls, const double* ls0, const double* hd, double* o) {
    auto s = make_ls(gr, 0.0);
    const int64_t n = grid_size(gr);
    out(s->GetSussmanReinitConstraint(vec(ls, n), vec(ls0, n), vec(hd, n)), o);
    return 0;
}

int ref_reinit_minrk2(const or_grid* gr, double* ls, double max_time) {
    auto s = make_ls(gr, max_time);
    std::vector<double> l = vec(ls, grid_size(gr));
    s->Reinit_MinRK2_2D(l);
    out(l, ls);
    return 0;
}

int ref_reinit_original(const or_grid* gr, double* ls, double max_time) {
    auto s = make_ls(gr, max_time);
    std::vector<double> l = vec(ls, grid_size(gr));
    s->Reinit_Original_2D(l);
    out(l, ls);
    return 0;
}

int ref_reinit_first_time(const or_grid* gr, double* ls) {
    auto s = make_ls(gr, 0.0);
    std::vector<double> l = vec(ls, grid_size(gr));
    s->FirstTimeOnlyReinit_Sussman_2D(l);
    out(l, ls);
    return 0;
}

int ref_weno_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* o) {
    auto s = make_ls(gr, 0.0);
    const int64_t n = grid_size(gr);
    out(s->HJWENO5_DerivAbsLS_2D(vec(ls, n), vec(ls0, n)), o);
    return 0;
}

int ref_reinit_sussman(const or_grid* gr, double* ls, double max_time) {
    auto s = make_ls(gr, max_time);
    std::vector<double> l = vec(ls, grid_size(gr));
    s->Reinit_Sussman_2D(l);
    out(l, ls);
    return 0;
}

}  // extern "C"
