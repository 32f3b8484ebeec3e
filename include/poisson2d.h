// poisson2d.h -- drop-in PoissonSolver2D (reference include/poisson2d.h:31-61).
//
// Same CSR-based entry points; the Krylov iterations run on the GPU (asd2_csr_pcg and the
// CSR BiCGStab of asd2_hip.h).  MKL is not used: MKL_INT is a 32-bit int as in MKL's LP64
// interface.  GS_2FUniform_2D runs the reference's Gauss-Seidel sweeps on the GPU
// (asd2_csr_gs); the FFT Helmholtz solver (MKL_2FUniform_2D) is outside the GPU path and
// reports an error.
#ifndef ASD2_POISSON2D_H
#define ASD2_POISSON2D_H

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "bc2d.h"
#include "common.h"

#ifndef MKL_INT
typedef int MKL_INT;
#define MKL_INT MKL_INT
#endif

struct asd2_ctx;

class PoissonSolver2D {
    int kNx, kNy, kNumBCGrid;

public:
    PoissonSolver2D(int nx, int ny, int num_bc_grid);
    ~PoissonSolver2D();
    PoissonSolver2D(const PoissonSolver2D&) = delete;
    PoissonSolver2D& operator=(const PoissonSolver2D&) = delete;

    int GS_2FUniform_2D(std::vector<double>& ps, const std::vector<double>& rhs, std::vector<double>& AVals,
                        std::vector<MKL_INT>& ACols, std::vector<MKL_INT>& ARowIdx, const double lenX,
                        const double lenY, const double dx, const double dy,
                        const std::shared_ptr<BoundaryCondition2D>& PBC, const int maxiter);
    int MKL_2FUniform_2D(std::vector<double>& phi, const std::vector<double>& rhs, double lenX, double lenY,
                         double dx, double dy, std::shared_ptr<BoundaryCondition2D> PBC);

    int CG_2FUniformP_2D(std::vector<double>& ps, const std::vector<double>& rhs, std::vector<double>& AVals,
                         std::vector<MKL_INT>& ACols, std::vector<MKL_INT>& ARowIdx, std::vector<double>& MVals,
                         std::vector<MKL_INT>& MCols, std::vector<MKL_INT>& MRowIdx,
                         const std::shared_ptr<BoundaryCondition2D>& PBC, int64_t size, const int maxIter);
    int CG_2FUniform_2D(double* x, double* b, std::vector<double>& AVals, std::vector<MKL_INT>& ACols,
                        std::vector<MKL_INT>& ARowIdx, std::vector<double>& MVals, std::vector<MKL_INT>& MCols,
                        std::vector<MKL_INT>& MRowIdx, const std::shared_ptr<BoundaryCondition2D>& PBC,
                        int64_t size, const int maxIter);

    int BiCGStab_2FUniform_2D(std::vector<double>& ps, const std::vector<double>& rhs, std::vector<double>& AVals,
                              std::vector<MKL_INT>& ACols, std::vector<MKL_INT>& ARowIdx, std::vector<double>& M,
                              std::vector<MKL_INT>& MCols, std::vector<MKL_INT>& MRowIdx,
                              const std::shared_ptr<BoundaryCondition2D>& PBC, int64_t size, const int maxiter);

    bool IsSymmetric(std::vector<double>& AVals, std::vector<MKL_INT>& ACols, std::vector<MKL_INT>& ARowIdx);

    std::vector<double> InvertMatrixDiagonal(const std::vector<double>& M);
    int idx(int i, int j) { return j + (kNy + 2 * kNumBCGrid) * i; }

    // -- additions (no reference counterpart) --------------------------------------
    // relative residual tolerance of the Krylov loops (the reference hard-codes 1e-6,
    // src/poisson2d.cpp:311) and the recurrence (ASD2_CG_REFERENCE / ASD2_CG_STANDARD)
    void SetTolerance(double tol) { m_tol = tol; }
    void SetRecurrence(int mode) { m_mode = mode; }
    int LastIterations() const { return m_lastIters; }
    double LastResidual() const { return m_lastRes; }

private:
    asd2_ctx* Context();
    asd2_ctx* m_ctx = nullptr;
    double m_tol = 1.0e-6;
    int m_mode = 0;
    int m_lastIters = 0;
    double m_lastRes = 0.0;
};

#endif
