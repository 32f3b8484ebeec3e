// levelset2d.h -- drop-in LevelSetSolver2D (reference include/levelset2d.h:17-87).
//
// Same constructors, public constants and methods.  Each call uploads its arrays, runs the
// gfx950 kernels of asd2_hip.h (asd2_ls_*), and downloads the results; the solver keeps
// one device context and its workspace for its lifetime.  Device-pointer variants
// (suffix _Device) skip the transfers for callers that keep their fields on the GPU.
#ifndef ASD2_LEVELSET2D_H
#define ASD2_LEVELSET2D_H

#include <algorithm>
#include <cmath>
#include <functional>
#include <iostream>
#include <memory>
#include <tuple>
#include <vector>

#include "asd2_resident.h"
#include "bc2d.h"
#include "common.h"

struct asd2_ctx;

class LevelSetSolver2D {
public:
    LevelSetSolver2D(int nx, int ny, int num_bc_grid, double baseX, double baseY, double dx, double dy);
    LevelSetSolver2D(int nx, int ny, int num_bc_grid, double baseX, double baseY, double dx, double dy,
                     double maxTime);
    ~LevelSetSolver2D();
    LevelSetSolver2D(const LevelSetSolver2D&) = delete;
    LevelSetSolver2D& operator=(const LevelSetSolver2D&) = delete;

    std::vector<double> UpdateHeavisideFuncDeriv(const std::vector<double>& ls);

    int Solve_LevelSet_2D(std::vector<double>& ls, const std::vector<double>& u, const std::vector<double>& v,
                          double dt);

    int Reinit_Original_2D(std::vector<double>& ls);
    int Reinit_Sussman_2D(std::vector<double>& ls);
    int Reinit_Sussman_FirstTime_2D(std::vector<double>& ls);
    int FirstTimeOnlyReinit_Sussman_2D(std::vector<double>& ls);
    int Reinit_MinRK2_2D(std::vector<double>& ls);

    std::vector<double> GetSussmanReinitConstraint(const std::vector<double>& ls, const std::vector<double>& lsInit,
                                                   const std::vector<double>& heavisideDeriv);

    int sign(const double& val);
    double MinAbs(const double& val1, const double& val2);
    double MinMod(const double& val1, const double& val2);
    std::vector<int> GetSignedLSNormalized(const std::vector<double>& ls);
    std::vector<double> GetSmoothedSignFunc(const std::vector<double>& ls);

    std::vector<double> HJWENO5_LS_2D(std::vector<double>& ls, const std::vector<double>& u,
                                      const std::vector<double>& v);
    int UnitHJWENO5(const std::vector<double>& F, std::vector<double>& FP, std::vector<double>& FM, const double d,
                    const int n);

    std::vector<double> HJWENO5_DerivAbsLS_2D(const std::vector<double>& ls, const std::vector<double>& lsInit);
    std::vector<double> ENO_DerivAbsLS_2D(const std::vector<double>& ls, const std::vector<double>& lsInit);
    std::vector<double> SubcellENO_Min_DerivAbsLS_2D(const std::vector<double>& ls, const std::vector<double>& lsInit);

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

    int idx(int i, int j) { return j + (kNy + 2 * kNumBCGrid) * i; }

    const double kEps, kThickness, kMaxATime;
    const int kNx, kNy, kNumBCGrid, kENOSpatialOrder;
    const int64_t kArrSize;
    const double kBaseX, kBaseY;
    const double kDx, kDy;

    double m_atime;
    const double kAdt;

    std::shared_ptr<BoundaryCondition2D> m_BC;

    // -- additions (no reference counterpart) --------------------------------------
    // device-resident variants: d_* are device pointers in the padded layout
    int Solve_LevelSet_2D_Device(double* d_ls, const double* d_u, const double* d_v, double dt);
    int Reinit_MinRK2_2D_Device(double* d_ls);
    int Reinit_Sussman_2D_Device(double* d_ls);
    // number of pseudo-time iterations of the reinitialisation loops (kMaxATime / kAdt)
    int ReinitIterations() const;
    asd2_ctx* Context();

private:
    struct Device;  // device context, workspace buffers
    void EnsureBC();
    Device& Dev();
    std::vector<double> Pointwise(int op, const std::vector<double>& ls, const std::vector<double>* ls0,
                                  const char* where);
    int Reinit(int method, std::vector<double>& ls, int iterations, const char* where);
    int ApplyBC(int grid_kind, std::vector<double>& arr, const char* where);
    std::unique_ptr<Device> m_dev;
};

#endif
