// dropin_driver.cpp -- exercises the drop-in C++ classes (include/levelset2d.h, mac2d.h,
// poisson2d.h, bc2d.h) exactly as the reference's own drivers do, built against the
// headers and libasd2_hip.so only.  The Python tests feed it inputs and compare what it
// writes with the oracle.
//
//   dropin_driver step <dir> <n> <steps> <mode> [resident]
//                                           stationary bubble (tests/LargeTests/
//                                           test_mac2d_large.cpp:376-540) on an n x n grid;
//                                           mode 0 keeps the reference CG (recurrence and
//                                           1e-6 tolerance), mode 1 solves to 1e-10 with
//                                           the textbook recurrence; resident 1 runs the same
//                                           loop with the arrays kept on the device
//                                           (asd2_resident.h) and the in-place overloads
//   dropin_driver bench <n> <steps> <resident>
//                                           the droplet step of the benchmark (unit box, drop
//                                           r = 0.25 heavy inside, density 1000 : 1, viscosity
//                                           1e-2 : 1e-4, sigma = 1, 3 MinRK2 iterations,
//                                           mixed-precision multilevel PCG to 1e-8) through the
//                                           classes in the reference loop's order; one warmup
//                                           step, then a line "BENCH {json}" with ms/step and
//                                           the host <-> device bytes per step
//   dropin_driver ls <dir> <nx> <ny>        LevelSetSolver2D methods on <dir>/phi.bin, u.bin, v.bin
//   dropin_driver csr <dir> <nx> <ny> <all_neumann 0|1>
//                                           PoissonSolver2D CG + BiCGStab on a CSR system
//   dropin_driver outres <dir> <nx> <ny>     MACSolver2D::OutRes (Tecplot ASCII) of <dir>/u.bin, v.bin,
//                                           ps.bin, ls.bin at iterations 0 and 1 into <dir>/vel_ASCII.plt
//                                           and <dir>/div_ASCII.plt (grid of 1 x 1.5, base (0.25, -0.5))
//   dropin_driver fast <dir> <nx> <ny> <sides>
//                                           PoissonSolver2D::MKL_2FUniform_2D on <dir>/rhs.bin,
//                                           ps.bin, consts.bin (W E S N) with pressure sides
//                                           <sides> (4 letters N/D), lenX = 1, lenY = ny / nx;
//                                           then MACSolver2D::SolvePoisson with POISSONTYPE::MKL
//                                           on <dir>/div.bin (all sides Neumann)
//
// Arrays are raw little-endian float64 (int32 for CSR indices) in the padded layout.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "asd2_hip.h"
#include "bc2d.h"
#include "levelset2d.h"
#include "mac2d.h"
#include "poisson2d.h"

namespace {

template <typename T>
std::vector<T> load(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) {
        std::cerr << "cannot read " << path << std::endl;
        std::exit(2);
    }
    const std::streamsize bytes = f.tellg();
    f.seekg(0);
    std::vector<T> v(static_cast<size_t>(bytes) / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), bytes);
    return v;
}

template <typename T>
void save(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

int run_step(const std::string& dir, int n, int steps, int mode, bool resident) {
    asd2::SetDeviceResident(resident);
    const double rhoL = 1.226, muL = 1.78e-5, rhoH = 1000, muH = 1.137e-3, sigma = 0.0728, gConstant = 0.0;
    const double L = 1.0, U = 1.0, baseX = 0.0, baseY = 0.0, lenX = 0.04, lenY = 0.04, cfl = 0.3, radius = 0.01;
    const int nx = n, ny = n, num_bc_grid = 3, poissonMaxIter = 20000;
    const double dx = lenX / nx, dy = lenY / ny;
    auto MSolver = std::make_unique<MACSolver2D>(rhoH, rhoL, muH, muL, gConstant, GAXISENUM2D::Y, L, U, sigma, nx,
                                                 ny, baseX, baseY, lenX, lenY, TIMEORDERENUM::RK3, cfl, 5.0, 20, 1,
                                                 num_bc_grid, false);
    MSolver->SetBC_U_2D("dirichlet", "dirichlet", "dirichlet", "dirichlet");
    MSolver->SetBC_V_2D("dirichlet", "dirichlet", "dirichlet", "dirichlet");
    MSolver->SetBC_P_2D("neumann", "neumann", "neumann", "neumann");
    MSolver->SetBCConstantUW(0.0);
    MSolver->SetBCConstantUE(0.0);
    MSolver->SetBCConstantUS(0.0);
    MSolver->SetBCConstantUN(0.0);
    MSolver->SetBCConstantVW(0.0);
    MSolver->SetBCConstantVE(0.0);
    MSolver->SetBCConstantVS(0.0);
    MSolver->SetBCConstantVN(0.0);
    MSolver->SetPLTType(PLTTYPE::NONE);
    MSolver->SetPoissonSolver(POISSONTYPE::CG);
    if (mode == 1) MSolver->SetPoissonOptions(1e-10, ASD2_CG_STANDARD, ASD2_PREC_FP64, ASD2_PRECOND_JACOBI);
    auto LSolver = std::make_shared<LevelSetSolver2D>(nx, ny, num_bc_grid, baseX, baseY, dx, dy);
    const size_t N = static_cast<size_t>((nx + 2 * num_bc_grid) * (ny + 2 * num_bc_grid));
    std::vector<double> lsB(N, 0.0), ls(N, 0.0), H(N, 0.0), HSmooth(N, 0.0);
    MSolver->AllocateVariables();
    MSolver->ApplyBC_U_2D(MSolver->m_u);
    MSolver->ApplyBC_V_2D(MSolver->m_v);
    MSolver->ApplyBC_P_2D(MSolver->m_ps);
    MSolver->ApplyBC_P_2D(MSolver->m_p);
    LSolver->SetBC_P_2D("neumann", "neumann", "neumann", "neumann");
    LSolver->SetBCConstantPW(0.0);
    LSolver->SetBCConstantPE(0.0);
    LSolver->SetBCConstantPS(0.0);
    LSolver->SetBCConstantPN(0.0);
    for (int j = 0; j < ny + 2 * num_bc_grid; j++)
        for (int i = 0; i < nx + 2 * num_bc_grid; i++) {
            const double x = baseX + (i + 0.5 - num_bc_grid) * dx, y = baseY + (j + 0.5 - num_bc_grid) * dy;
            ls[LSolver->idx(i, j)] = std::sqrt((x - 0.02) * (x - 0.02) + (y - 0.02) * (y - 0.02)) - radius;
        }
    LSolver->ApplyBC_P_2D(ls);
    LSolver->Reinit_MinRK2_2D(ls);
    LSolver->ApplyBC_P_2D(ls);
    asd2::SyncHost(ls);  // resident: the host copy is refreshed only on request
    save(dir + "/ls0.bin", ls);
    MSolver->m_dt = cfl * std::min(dx, dy) / U;
    std::vector<double> uhat(N), vhat(N), div(N);
    std::vector<double> dts;
    std::vector<int> iters;
    for (int s = 0; s < steps; s++) {
        MSolver->m_iter++;
        lsB = ls;
        LSolver->Solve_LevelSet_2D(ls, MSolver->m_u, MSolver->m_v, MSolver->m_dt);
        LSolver->ApplyBC_P_2D(ls);
        LSolver->Reinit_MinRK2_2D(ls);
        LSolver->ApplyBC_P_2D(ls);
        MSolver->UpdateKappa(ls);
        if (resident) {
            MSolver->UpdateHeavisideFunc(ls, H);
            MSolver->UpdateSmoothHeavisideFunc(ls, HSmooth);
        } else {
            H = MSolver->UpdateHeavisideFunc(ls);
            HSmooth = MSolver->UpdateSmoothHeavisideFunc(ls);
        }
        MSolver->GetIntermediateVel(LSolver, ls, MSolver->m_u, MSolver->m_v, uhat, vhat, HSmooth);
        MSolver->ApplyBC_U_2D(uhat);
        MSolver->ApplyBC_V_2D(vhat);
        if (resident)
            MSolver->GetDivergence(uhat, vhat, div);
        else
            div = MSolver->GetDivergence(uhat, vhat);
        MSolver->ApplyBC_P_2D(MSolver->m_ps);
        LSolver->ApplyBC_P_2D(ls);
        MSolver->SolvePoisson(MSolver->m_ps, div, ls, lsB, MSolver->m_u, MSolver->m_v, H, poissonMaxIter);
        MSolver->ApplyBC_P_2D(MSolver->m_ps);
        MSolver->UpdateVel(MSolver->m_u, MSolver->m_v, uhat, vhat, MSolver->m_ps, ls, lsB, H);
        MSolver->ApplyBC_U_2D(MSolver->m_u);
        MSolver->ApplyBC_V_2D(MSolver->m_v);
        dts.push_back(MSolver->m_dt);
        iters.push_back(MSolver->LastPoissonIterations());
        MSolver->m_dt = MSolver->UpdateDt(MSolver->m_u, MSolver->m_v);
        MSolver->m_curTime += MSolver->m_dt;
        std::fill(uhat.begin(), uhat.end(), 0.0);
        std::fill(vhat.begin(), vhat.end(), 0.0);
        asd2::ModifiedHost(uhat);  // resident: the next call uploads the zeros
        asd2::ModifiedHost(vhat);
    }
    asd2::SyncHostAll();
    dts.push_back(MSolver->m_dt);
    save(dir + "/ls.bin", ls);
    save(dir + "/u.bin", MSolver->m_u);
    save(dir + "/v.bin", MSolver->m_v);
    save(dir + "/ps.bin", MSolver->m_ps);
    save(dir + "/p.bin", MSolver->m_p);
    save(dir + "/kappa.bin", MSolver->m_kappa);
    save(dir + "/dt.bin", dts);
    save(dir + "/iters.bin", iters);
    return 0;
}

int run_bench(int n, int steps, bool resident) {
    asd2::SetDeviceResident(resident);
    const double rhoH = 1000.0, rhoL = 1.0, muH = 1.0e-2, muL = 1.0e-4, sigma = 1.0, cfl = 0.3, radius = 0.25;
    const int nx = n, ny = n, g = 3, poissonMaxIter = 100000;
    const double lenX = 1.0, lenY = 1.0, dx = lenX / nx, dy = lenY / ny;
    auto MSolver = std::make_unique<MACSolver2D>(rhoH, rhoL, muH, muL, 0.0, GAXISENUM2D::Y, 1.0, 1.0, sigma, nx, ny,
                                                 0.0, 0.0, lenX, lenY, TIMEORDERENUM::RK3, cfl, 1.0e9, 1 << 30, 1, g,
                                                 false);
    MSolver->SetBC_U_2D("dirichlet", "dirichlet", "dirichlet", "dirichlet");
    MSolver->SetBC_V_2D("dirichlet", "dirichlet", "dirichlet", "dirichlet");
    MSolver->SetBC_P_2D("neumann", "neumann", "neumann", "neumann");
    MSolver->SetPLTType(PLTTYPE::NONE);
    MSolver->SetPoissonSolver(POISSONTYPE::CG);
    MSolver->SetPoissonOptions(1e-8, ASD2_CG_STANDARD, ASD2_PREC_MIXED, ASD2_PRECOND_MULTILEVEL);
    // 3 MinRK2 iterations: a pseudo end time of 2.5 pseudo steps (eps / 2 each)
    const double eps = std::min(dx, dy);
    auto LSolver = std::make_shared<LevelSetSolver2D>(nx, ny, g, 0.0, 0.0, dx, dy, 1.25 * eps);
    if (LSolver->ReinitIterations() != 3) {
        std::cerr << "reinit iterations " << LSolver->ReinitIterations() << ", expected 3" << std::endl;
        return 1;
    }
    LSolver->SetBC_P_2D("neumann", "neumann", "neumann", "neumann");
    const size_t N = static_cast<size_t>((nx + 2 * g) * (ny + 2 * g));
    std::vector<double> lsB(N, 0.0), ls(N, 0.0), H(N, 0.0), HSmooth(N, 0.0), uhat(N, 0.0), vhat(N, 0.0), div(N, 0.0);
    MSolver->AllocateVariables();
    for (int i = 0; i < nx + 2 * g; i++)
        for (int j = 0; j < ny + 2 * g; j++) {
            const double x = (i + 0.5 - g) * dx, y = (j + 0.5 - g) * dy;
            ls[LSolver->idx(i, j)] = radius - std::sqrt((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5));
        }
    LSolver->ApplyBC_P_2D(ls);
    LSolver->Reinit_MinRK2_2D(ls);
    LSolver->ApplyBC_P_2D(ls);
    MSolver->m_dt = cfl * std::min(dx, dy);
    std::vector<int> iters;
    auto step = [&]() {
        MSolver->m_iter++;
        if (!resident) lsB = ls;  // read by neither class (SolvePoisson and UpdateVel ignore lsB)
        LSolver->Solve_LevelSet_2D(ls, MSolver->m_u, MSolver->m_v, MSolver->m_dt);
        LSolver->ApplyBC_P_2D(ls);
        LSolver->Reinit_MinRK2_2D(ls);
        LSolver->ApplyBC_P_2D(ls);
        MSolver->UpdateKappa(ls);
        if (resident) {
            MSolver->UpdateHeavisideFunc(ls, H);
            MSolver->UpdateSmoothHeavisideFunc(ls, HSmooth);
        } else {
            H = MSolver->UpdateHeavisideFunc(ls);
            HSmooth = MSolver->UpdateSmoothHeavisideFunc(ls);
        }
        MSolver->GetIntermediateVel(LSolver, ls, MSolver->m_u, MSolver->m_v, uhat, vhat, HSmooth);
        MSolver->ApplyBC_U_2D(uhat);
        MSolver->ApplyBC_V_2D(vhat);
        if (resident)
            MSolver->GetDivergence(uhat, vhat, div);
        else
            div = MSolver->GetDivergence(uhat, vhat);
        MSolver->ApplyBC_P_2D(MSolver->m_ps);
        LSolver->ApplyBC_P_2D(ls);
        MSolver->SolvePoisson(MSolver->m_ps, div, ls, lsB, MSolver->m_u, MSolver->m_v, H, poissonMaxIter);
        MSolver->ApplyBC_P_2D(MSolver->m_ps);
        MSolver->UpdateVel(MSolver->m_u, MSolver->m_v, uhat, vhat, MSolver->m_ps, ls, lsB, H);
        MSolver->ApplyBC_U_2D(MSolver->m_u);
        MSolver->ApplyBC_V_2D(MSolver->m_v);
        MSolver->m_dt = MSolver->UpdateDt(MSolver->m_u, MSolver->m_v);
        MSolver->m_curTime += MSolver->m_dt;
        iters.push_back(MSolver->LastPoissonIterations());
    };
    step();  // warmup
    const asd2::TransferCounts c0 = asd2::Transfers();
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; s++) step();
    const auto t1 = std::chrono::steady_clock::now();
    const asd2::TransferCounts c1 = asd2::Transfers();
    const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
    double it = 0.0;
    for (int s = 1; s <= steps; s++) it += iters[s];
    asd2::SyncHostAll();
    double mass = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) mass += ls[LSolver->idx(i, j)] >= 0.0 ? dx * dy : 0.0;
    std::printf("BENCH {\"n\": %d, \"steps\": %d, \"resident\": %d, \"ms_per_step\": %.4f, \"pcg_iters_per_step\": %.2f, "
                "\"upload_MB_per_step\": %.3f, \"download_MB_per_step\": %.3f, \"uploads_per_step\": %.2f, "
                "\"downloads_per_step\": %.2f, \"mirrors\": %lld, \"mass\": %.12f, \"dt\": %.12e}\n",
                n, steps, resident ? 1 : 0, ms, it / steps, (c1.upload_bytes - c0.upload_bytes) / 1e6 / steps,
                (c1.download_bytes - c0.download_bytes) / 1e6 / steps,
                static_cast<double>(c1.uploads - c0.uploads) / steps,
                static_cast<double>(c1.downloads - c0.downloads) / steps, c1.mirrors, mass, MSolver->m_dt);
    std::fflush(stdout);
    return 0;
}

int run_ls(const std::string& dir, int nx, int ny) {
    const double dx = 1.0 / nx, dy = dx;
    LevelSetSolver2D S(nx, ny, 3, 0.0, 0.0, dx, dy);
    S.SetBC_P_2D("neumann", "dirichlet", "wall", "outlet");
    S.SetBCConstantPE(0.1);
    S.SetBCConstantPN(-0.05);
    const std::vector<double> phi = load<double>(dir + "/phi.bin"), u = load<double>(dir + "/u.bin"),
                              v = load<double>(dir + "/v.bin");
    std::vector<double> a = phi;
    S.Solve_LevelSet_2D(a, u, v, 0.4 * dx);
    save(dir + "/advect.bin", a);
    std::vector<double> b = phi;
    S.Reinit_MinRK2_2D(b);
    save(dir + "/minrk2.bin", b);
    b = phi;
    S.Reinit_Sussman_2D(b);
    save(dir + "/sussman.bin", b);
    b = phi;
    S.Reinit_Original_2D(b);
    save(dir + "/original.bin", b);
    b = phi;
    S.FirstTimeOnlyReinit_Sussman_2D(b);
    save(dir + "/first_time.bin", b);
    save(dir + "/subcell.bin", S.SubcellENO_Min_DerivAbsLS_2D(a, phi));
    save(dir + "/eno.bin", S.ENO_DerivAbsLS_2D(a, phi));
    save(dir + "/weno_absgrad.bin", S.HJWENO5_DerivAbsLS_2D(a, phi));
    save(dir + "/sign.bin", S.GetSmoothedSignFunc(phi));
    save(dir + "/hderiv.bin", S.UpdateHeavisideFuncDeriv(a));
    std::vector<double> p = phi;
    save(dir + "/weno_rhs.bin", S.HJWENO5_LS_2D(p, u, v));
    save(dir + "/weno_rhs_phi.bin", p);
    std::vector<double> bc = u;
    S.ApplyBC_P_2D(bc);
    save(dir + "/bc_p.bin", bc);
    std::vector<double> ints = {static_cast<double>(S.ReinitIterations()),
                                static_cast<double>(S.sign(-2.0)), S.MinMod(-1.0, -3.0), S.MinMod(1.0, -3.0),
                                S.MinAbs(-1.0, 0.5)};
    save(dir + "/scalars.bin", ints);
    return 0;
}

int run_csr(const std::string& dir, int nx, int ny, bool all_neumann) {
    auto vals = load<double>(dir + "/AVals.bin");
    auto cols = load<MKL_INT>(dir + "/ACols.bin");
    auto rows = load<MKL_INT>(dir + "/ARowIdx.bin");
    auto mv = load<double>(dir + "/MVals.bin");
    auto mc = load<MKL_INT>(dir + "/MCols.bin");
    auto mr = load<MKL_INT>(dir + "/MRowIdx.bin");
    const auto rhs = load<double>(dir + "/rhs.bin");
    auto bc = std::make_shared<BoundaryCondition2D>(nx, ny, 3);
    if (all_neumann)
        bc->SetBC_P_2D("neumann", "neumann", "neumann", "neumann");
    else
        bc->SetBC_P_2D("neumann", "neumann", "dirichlet", "neumann");
    PoissonSolver2D P(nx, ny, 3);
    const int64_t size = static_cast<int64_t>(rows.size()) - 1;
    std::vector<double> ps(rhs.size(), 0.0);
    P.CG_2FUniformP_2D(ps, rhs, vals, cols, rows, mv, mc, mr, bc, size, 100000);
    save(dir + "/ps_cg.bin", ps);
    std::vector<int> it = {P.LastIterations()};
    P.SetRecurrence(1);
    P.SetTolerance(1e-12);
    std::fill(ps.begin(), ps.end(), 0.0);
    P.CG_2FUniformP_2D(ps, rhs, vals, cols, rows, mv, mc, mr, bc, size, 100000);
    save(dir + "/ps_cg_tight.bin", ps);
    it.push_back(P.LastIterations());
    std::fill(ps.begin(), ps.end(), 0.0);
    P.BiCGStab_2FUniform_2D(ps, rhs, vals, cols, rows, mv, mc, mr, bc, size, 100000);
    save(dir + "/ps_bicg.bin", ps);
    it.push_back(P.LastIterations());
    save(dir + "/iters.bin", it);
    std::vector<double> inv = P.InvertMatrixDiagonal(mv);
    save(dir + "/minv.bin", inv);
    std::vector<int> sym = {P.IsSymmetric(vals, cols, rows) ? 1 : 0};
    save(dir + "/sym.bin", sym);
    return 0;
}

int run_outres(const std::string& dir, int nx, int ny) {
    auto M = std::make_unique<MACSolver2D>(1000.0, 1.0, 1.0, 0.01, 0.0, GAXISENUM2D::Y, 1.0, 1.0, 0.0, nx, ny, 0.25,
                                           -0.5, 1.0, 1.5, TIMEORDERENUM::RK3, 0.3, 5.0, 20, 1, 3, false);
    M->SetBC_U_2D("dirichlet", "dirichlet", "dirichlet", "dirichlet");
    M->SetBC_V_2D("dirichlet", "dirichlet", "dirichlet", "dirichlet");
    M->SetBC_P_2D("neumann", "neumann", "neumann", "neumann");
    M->SetPLTType(PLTTYPE::ASCII);
    M->AllocateVariables();
    const auto u = load<double>(dir + "/u.bin"), v = load<double>(dir + "/v.bin");
    const auto ps = load<double>(dir + "/ps.bin"), ls = load<double>(dir + "/ls.bin");
    M->m_iter = 0;
    M->OutRes(0, 0.0, dir + "/vel", dir + "/div", u, v, ps, ls);
    M->m_iter = 1;
    M->OutRes(7, 0.125, dir + "/vel", dir + "/div", u, v, ps, ls);
    return 0;
}

int run_fast(const std::string& dir, int nx, int ny, const std::string& sides) {
    const double lenX = 1.0, lenY = static_cast<double>(ny) / nx, dx = lenX / nx, dy = lenY / ny;
    const auto rhs = load<double>(dir + "/rhs.bin");
    auto ps = load<double>(dir + "/ps.bin");
    const auto c = load<double>(dir + "/consts.bin");
    auto word = [](char k) { return std::string(k == 'D' ? "dirichlet" : "neumann"); };
    auto bc = std::make_shared<BoundaryCondition2D>(nx, ny, 3);
    bc->SetBC_P_2D(word(sides[0]), word(sides[1]), word(sides[2]), word(sides[3]));
    bc->SetBCConstantPW(c[0]);
    bc->SetBCConstantPE(c[1]);
    bc->SetBCConstantPS(c[2]);
    bc->SetBCConstantPN(c[3]);
    PoissonSolver2D P(nx, ny, 3);
    P.MKL_2FUniform_2D(ps, rhs, lenX, lenY, dx, dy, bc);
    save(dir + "/ps_fast.bin", ps);

    // the MAC path: rhs = -div, ps zeroed, P BCs, m_p = ps / dt
    auto M = std::make_unique<MACSolver2D>(1000.0, 1.0, 1.0, 0.01, 0.0, GAXISENUM2D::Y, 1.0, 1.0, 0.0, nx, ny, 0.0,
                                           0.0, lenX, lenY, TIMEORDERENUM::RK3, 0.3, 5.0, 20, 1, 3, false);
    M->SetBC_U_2D("dirichlet", "dirichlet", "dirichlet", "dirichlet");
    M->SetBC_V_2D("dirichlet", "dirichlet", "dirichlet", "dirichlet");
    M->SetBC_P_2D("neumann", "neumann", "neumann", "neumann");
    M->SetPLTType(PLTTYPE::NONE);
    M->SetPoissonSolver(POISSONTYPE::MKL);
    M->AllocateVariables();
    M->m_dt = 0.01;
    const auto div = load<double>(dir + "/div.bin");
    const std::vector<double> zero(div.size(), 0.0);
    M->SolvePoisson(M->m_ps, div, zero, zero, zero, zero, zero, 100);
    save(dir + "/mac_ps.bin", M->m_ps);
    save(dir + "/mac_p.bin", M->m_p);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: dropin_driver step|ls|csr <dir> ..." << std::endl;
        return 2;
    }
    const std::string cmd = argv[1], dir = argv[2];
    if (cmd == "step" && (argc == 6 || argc == 7))
        return run_step(dir, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), argc == 7 && std::atoi(argv[6]));
    if (cmd == "bench" && argc == 5) return run_bench(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]) != 0);
    if (cmd == "ls" && argc == 5) return run_ls(dir, std::atoi(argv[3]), std::atoi(argv[4]));
    if (cmd == "csr" && argc == 6) return run_csr(dir, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]) != 0);
    if (cmd == "outres" && argc == 5) return run_outres(dir, std::atoi(argv[3]), std::atoi(argv[4]));
    if (cmd == "fast" && argc == 6 && std::string(argv[5]).size() == 4)
        return run_fast(dir, std::atoi(argv[3]), std::atoi(argv[4]), argv[5]);
    std::cerr << "bad arguments" << std::endl;
    return 2;
}
