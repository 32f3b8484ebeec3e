// mac2d.h -- drop-in MACSolver2D (reference include/mac2d.h:31-168).
//
// Same constructors, public data members and methods.  The fields live on the host in
// the public std::vectors (drivers read and write them directly); every method uploads
// its inputs, runs the gfx950 kernels of asd2_hip.h and downloads its outputs.
// SolvePoisson assembles the ghost-fluid system matrix-free on the GPU (no CSR) and
// solves it with the configured pressure path.  Tecplot binary output is not available
// (no TECIO); OutRes writes Tecplot ASCII.
#ifndef ASD2_MAC2D_H
#define ASD2_MAC2D_H

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "bc2d.h"
#include "common.h"
#include "levelset2d.h"
#include "poisson2d.h"

struct asd2_ctx;

class MACSolver2D {
public:
    const int kNx, kNy, kNumBCGrid;
    const int64_t kArrSize;
    const double kBaseX, kBaseY, kLenX, kLenY, kDx, kDy;

    const double kRe, kWe, kFr;
    const double kLScale, kUScale, kSigma, kG, kMuScale, kRhoScale;
    const double kRhoH, kRhoL, kRhoRatio;
    const double kMuH, kMuL, kMuRatio;

    const int kMaxIter, kNIterSkip;
    const TIMEORDERENUM kTimeOrder;
    const double kCFL, kMaxTime;
    const bool kWriteVTK;
    POISSONTYPE m_PoissonSolverType;
    PLTTYPE m_PLTType;
    GAXISENUM2D kGAxis;

    const double kEps_div = 1.0e-6;

    std::shared_ptr<BoundaryCondition2D> m_BC;
    std::shared_ptr<PoissonSolver2D> m_Poisson;
    int m_iter;
    double m_dt, m_curTime;
    std::string m_outFname;

    std::vector<double> m_u, m_v;
    std::vector<double> m_rho, m_kappa, m_mu;
    std::vector<double> m_ps, m_p;
    std::vector<double> m_J11, m_J12, m_J21, m_J22;
    std::vector<double> m_nx, m_ny;
    std::vector<double> m_t1x, m_t1y, m_t1z;
    std::vector<double> m_t2x, m_t2y, m_t2z;

    MACSolver2D(double Re, double We, double Fr, GAXISENUM2D GAxis, double L, double U, double sigma,
                double densityRatio, double viscosityRatio, double rhoI, double mu1, int nx, int ny, double baseX,
                double baseY, double lenX, double lenY, TIMEORDERENUM RKOrder, double cfl, double maxtime,
                int maxiter, int niterskip, int num_bc_grid, bool writeVTK);
    MACSolver2D(double rhoI, double rhoO, double muI, double muO, double gConstant, GAXISENUM2D GAxis, double L,
                double U, double sigma, int nx, int ny, double baseX, double baseY, double lenX, double lenY,
                TIMEORDERENUM RKOrder, double cfl, double maxtime, int maxiter, int niterskip, int num_bc_grid,
                bool writeVTK);
    ~MACSolver2D();
    MACSolver2D(const MACSolver2D&) = delete;
    MACSolver2D& operator=(const MACSolver2D&) = delete;

    int AllocateVariables();

    std::vector<double> UpdateSmoothHeavisideFunc(const std::vector<double>& ls);
    std::vector<double> UpdateHeavisideFunc(const std::vector<double>& ls);
    int UpdateNTK(const std::shared_ptr<LevelSetSolver2D>& LSolver, const std::vector<double>& ls,
                  std::tuple<std::vector<double>&, std::vector<double>&> normalVec,
                  std::tuple<std::vector<double>&, std::vector<double>&, std::vector<double>&> t1Vec,
                  std::tuple<std::vector<double>&, std::vector<double>&, std::vector<double>&> t2Vec,
                  std::vector<double>& kappa);
    int UpdateKappa(const std::vector<double>& ls);

    std::vector<double> UpdateFU(const std::shared_ptr<LevelSetSolver2D>& LSolver, const std::vector<double>& ls,
                                 const std::vector<double>& u, const std::vector<double>& v,
                                 const std::vector<double>& H);
    std::vector<double> UpdateFV(const std::shared_ptr<LevelSetSolver2D>& LSolver, const std::vector<double>& ls,
                                 const std::vector<double>& u, const std::vector<double>& v,
                                 const std::vector<double>& H);

    std::vector<double> AddConvectionFU(const std::vector<double>& u, const std::vector<double>& v);
    std::vector<double> AddConvectionFV(const std::vector<double>& u, const std::vector<double>& v);
    int UnitHJWENO5(const std::vector<double>& F, std::vector<double>& FP, std::vector<double>& FM, const double d,
                    const int n);

    std::vector<double> AddViscosityFU(const std::vector<double>& u, const std::vector<double>& v,
                                       const std::vector<double>& ls, const std::vector<double>& H);
    std::vector<double> AddViscosityFV(const std::vector<double>& u, const std::vector<double>& v,
                                       const std::vector<double>& ls, const std::vector<double>& H);

    std::vector<double> AddGravityFU();
    std::vector<double> AddGravityFV();

    int GetIntermediateVel(const std::shared_ptr<LevelSetSolver2D>& LSolver, const std::vector<double>& ls,
                           const std::vector<double>& u, const std::vector<double>& v, std::vector<double>& uhat,
                           std::vector<double>& vhat, const std::vector<double>& H);

    int SetPoissonSolver(POISSONTYPE type);
    int SolvePoisson(std::vector<double>& ps, const std::vector<double>& div, const std::vector<double>& ls,
                     const std::vector<double>& lsB, const std::vector<double>& u, const std::vector<double>& v,
                     const std::vector<double>& H, const int maxiter);

    std::vector<double> GetDivergence(const std::vector<double>& u, const std::vector<double>& v);
    int UpdateVel(std::vector<double>& u, std::vector<double>& v, const std::vector<double>& us,
                  const std::vector<double>& vs, const std::vector<double>& ps, const std::vector<double>& ls,
                  const std::vector<double>& lsB, const std::vector<double>& H);
    double UpdateDt(const std::vector<double>& u, const std::vector<double>& v);

    int SetBC_U_2D(std::string BC_W, std::string BC_E, std::string BC_S, std::string BC_N);
    int SetBC_V_2D(std::string BC_W, std::string BC_E, std::string BC_S, std::string BC_N);
    int SetBC_P_2D(std::string BC_W, std::string BC_E, std::string BC_S, std::string BC_N);
    int ApplyBC_U_2D(std::vector<double>& arr);
    int ApplyBC_V_2D(std::vector<double>& arr);
    int ApplyBC_P_2D(std::vector<double>& arr);
    void SetBCConstantUW(double BC_ConstantW);
    void SetBCConstantUE(double BC_ConstantE);
    void SetBCConstantUS(double BC_ConstantS);
    void SetBCConstantUN(double BC_ConstantN);
    void SetBCConstantVW(double BC_ConstantW);
    void SetBCConstantVE(double BC_ConstantE);
    void SetBCConstantVS(double BC_ConstantS);
    void SetBCConstantVN(double BC_ConstantN);
    void SetBCConstantPW(double BC_ConstantW);
    void SetBCConstantPE(double BC_ConstantE);
    void SetBCConstantPS(double BC_ConstantS);
    void SetBCConstantPN(double BC_ConstantN);

    void TDMA(double* a, double* b, double* c, double* d, int n);
    int sign(const double& val);
    int SetPLTType(PLTTYPE type);
    int OutRes(const int iter, const double curTime, const std::string fname_vel_base,
               const std::string fname_div_base, const std::vector<double>& u, const std::vector<double>& v,
               const std::vector<double>& ps, const std::vector<double>& ls);
    int OutResClose();

    int idx(int i, int j) { return j + (kNy + 2 * kNumBCGrid) * i; }

    // -- additions (no reference counterpart) --------------------------------------
    // pressure path of SolvePoisson: tolerance (the reference's CG stops at 1e-6), the CG
    // recurrence (ASD2_CG_*), the precision (ASD2_PREC_*) and the preconditioner
    // (ASD2_PRECOND_*); gravity on the V equation (the reference adds the U-grid term,
    // src/mac2d.cpp:294; 1 applies the V-grid term)
    void SetPoissonOptions(double tol, int mode, int precision, int precond);
    void SetGravityFix(bool fix) { m_gravityFix = fix; }
    int LastPoissonIterations() const { return m_lastIters; }
    asd2_ctx* Context();
    // in-place forms of the by-value UpdateHeavisideFunc / UpdateSmoothHeavisideFunc /
    // GetDivergence: the result goes into H / div (sized to the grid if it is not), so in
    // resident mode (asd2_resident.h) it stays on the device and no host vector is allocated
    void UpdateHeavisideFunc(const std::vector<double>& ls, std::vector<double>& H);
    void UpdateSmoothHeavisideFunc(const std::vector<double>& ls, std::vector<double>& H);
    void GetDivergence(const std::vector<double>& u, const std::vector<double>& v, std::vector<double>& div);

private:
    struct Device;  // device context and named workspace buffers
    Device& Dev();
    void FluidParams(void* out) const;
    void ForgetMembers();
    std::vector<double> MomentumTerm(int grid_kind, int term, const std::vector<double>* ls,
                                     const std::vector<double>& u, const std::vector<double>& v,
                                     const std::vector<double>* H, const char* where);
    void PressureFromPs(double* d_ps);
    int ApplyBC(int grid_kind, std::vector<double>& arr, const char* where);
    std::unique_ptr<Device> m_dev;
    double m_tol = 1.0e-6;
    int m_mode = 0, m_precision = 0, m_precond = 0;
    bool m_gravityFix = false;
    int m_lastIters = 0;
    std::ofstream m_outVel;
};

#endif
