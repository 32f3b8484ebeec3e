// bc2d.h -- drop-in BoundaryCondition2D (reference include/bc2d.h:20-127).
//
// Same public members and methods; every Apply/BC_* call runs the gfx950 ghost-fill
// kernels of asd2_hip.h on a device copy of the array (upload, fill, download).  Periodic
// P-grid sides are not supported on the GPU path and terminate with a message, as the
// reference does on fatal conditions.
#ifndef ASD2_BC2D_H
#define ASD2_BC2D_H

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "common.h"

struct asd2_ctx;

class BoundaryCondition2D {
public:
    BoundaryCondition2D(int nx, int ny, int num_bc_grid);
    BoundaryCondition2D(int nx, int ny, double dx, double dy, int num_bc_grid);
    ~BoundaryCondition2D();
    BoundaryCondition2D(const BoundaryCondition2D&) = delete;
    BoundaryCondition2D& operator=(const BoundaryCondition2D&) = delete;

    const int kNx, kNy;
    const int kNumBCGrid;
    const double kDx, kDy;
    double m_AmbientPressure;

    BC2D m_BC_UW, m_BC_UE, m_BC_US, m_BC_UN;
    BC2D m_BC_VW, m_BC_VE, m_BC_VS, m_BC_VN;
    BC2D m_BC_PW, m_BC_PE, m_BC_PS, m_BC_PN;
    double m_BC_DirichletConstantUW, m_BC_DirichletConstantUE, m_BC_DirichletConstantUS, m_BC_DirichletConstantUN;
    double m_BC_DirichletConstantVW, m_BC_DirichletConstantVE, m_BC_DirichletConstantVS, m_BC_DirichletConstantVN;
    double m_BC_DirichletConstantPW, m_BC_DirichletConstantPE, m_BC_DirichletConstantPS, m_BC_DirichletConstantPN;

    inline int idx(int i, int j) { return j + (kNy + 2 * kNumBCGrid) * i; }

    int SetBC_U_2D(std::string BC_W, std::string BC_E, std::string BC_S, std::string BC_N);
    int SetBC_V_2D(std::string BC_W, std::string BC_E, std::string BC_S, std::string BC_N);
    int SetBC_P_2D(std::string BC_W, std::string BC_E, std::string BC_S, std::string BC_N);
    int ApplyBC_U_2D(std::vector<double>& arr);
    int ApplyBC_V_2D(std::vector<double>& arr);
    int ApplyBC_P_2D(std::vector<double>& arr);

    void BC_UW(std::vector<double>& arr);
    void BC_UE(std::vector<double>& arr);
    void BC_US(std::vector<double>& arr);
    void BC_UN(std::vector<double>& arr);
    void BC_VW(std::vector<double>& arr);
    void BC_VE(std::vector<double>& arr);
    void BC_VS(std::vector<double>& arr);
    void BC_VN(std::vector<double>& arr);
    void BC_PW(std::vector<double>& arr);
    void BC_PE(std::vector<double>& arr);
    void BC_PS(std::vector<double>& arr);
    void BC_PN(std::vector<double>& arr);

    void BC_PeriodicUW(std::vector<double>& arr);
    void BC_PeriodicUE(std::vector<double>& arr);
    void BC_PeriodicUS(std::vector<double>& arr);
    void BC_PeriodicUN(std::vector<double>& arr);
    void BC_PeriodicVW(std::vector<double>& arr);
    void BC_PeriodicVE(std::vector<double>& arr);
    void BC_PeriodicVS(std::vector<double>& arr);
    void BC_PeriodicVN(std::vector<double>& arr);
    void BC_PeriodicPW(std::vector<double>& arr);
    void BC_PeriodicPE(std::vector<double>& arr);
    void BC_PeriodicPS(std::vector<double>& arr);
    void BC_PeriodicPN(std::vector<double>& arr);

    void BC_NeumannUW(std::vector<double>& arr);
    void BC_NeumannUE(std::vector<double>& arr);
    void BC_NeumannUS(std::vector<double>& arr);
    void BC_NeumannUN(std::vector<double>& arr);
    void BC_NeumannVW(std::vector<double>& arr);
    void BC_NeumannVE(std::vector<double>& arr);
    void BC_NeumannVS(std::vector<double>& arr);
    void BC_NeumannVN(std::vector<double>& arr);
    void BC_NeumannPW(std::vector<double>& arr);
    void BC_NeumannPE(std::vector<double>& arr);
    void BC_NeumannPS(std::vector<double>& arr);
    void BC_NeumannPN(std::vector<double>& arr);

    void BC_DirichletUW(std::vector<double>& arr);
    void BC_DirichletUE(std::vector<double>& arr);
    void BC_DirichletUS(std::vector<double>& arr);
    void BC_DirichletUN(std::vector<double>& arr);
    void BC_DirichletVW(std::vector<double>& arr);
    void BC_DirichletVE(std::vector<double>& arr);
    void BC_DirichletVS(std::vector<double>& arr);
    void BC_DirichletVN(std::vector<double>& arr);
    void BC_DirichletPW(std::vector<double>& arr);
    void BC_DirichletPE(std::vector<double>& arr);
    void BC_DirichletPS(std::vector<double>& arr);
    void BC_DirichletPN(std::vector<double>& arr);

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

    void BC_LSFreeBoundaryPW(std::vector<double>& arr);
    void BC_LSFreeBoundaryPE(std::vector<double>& arr);
    void BC_LSFreeBoundaryPS(std::vector<double>& arr);
    void BC_LSFreeBoundaryPN(std::vector<double>& arr);

    void SetAmbientPressure(double ambientPressure);

    // -- additions (no reference counterpart) --------------------------------------
    // pushes the current kinds/constants of all three grids into a device context
    void Configure(asd2_ctx* ctx) const;

private:
    // applies one side rule (kind on `side`, nothing elsewhere) of one grid on the GPU
    void ApplySide(int grid_kind, int side, int kind, std::vector<double>& arr);
    void ApplyAll(int grid_kind, std::vector<double>& arr);
    asd2_ctx* m_ctx = nullptr;
};

#endif
