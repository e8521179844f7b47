// One Chorin projection time step on the GPU.
//
// Replaces the time loop of solveNsChorinCuda
// (cpp/execs/NavierStokes/src/chorinCuda.cu:154-379):
//   1. C(u) assembled into the fixed velocity pattern, A = nu K + C
//      (fcSpmv + cublasSgeam, :241-255)
//   2. rhs = A [ux uy]; solve M a = rhs (2 channels, cold start); u -= tau a;
//      re-impose Dirichlet velocity (:263-278)
//   3. d = -(1/tau) VPD u, gathered on the internal pressure nodes; solve
//      Kp_int p = d (warm start); scatter p (:297-300, PressureSolver :77-152)
//   4. g = PVD p; solve M a = g (cold); u -= tau a; re-impose Dirichlet (:341-356)
// The convection values come from an element-wise kernel (36 local values per
// P2 element, then a deterministic gather per CSR entry) instead of the
// reference's (45.5 V x 8 V) FastConvection SpMV; both compute the same
// linear map u -> C(u) under the same 6-point quadrature.
#pragma once

#include <memory>
#include <string>

#include "common.hpp"
#include "context.hpp"
#include "ops.hpp"
#include "solver.hpp"

namespace nsfem
{
    struct ChorinSolverConfig
    {
        std::string method = "gs";
        int maxIterations = 200;
        double target = 1e-6;
        int checkInterval = 1;
        int lanes = 0;        // 0: from the mean row length
        int rowsPerBlock = 0; // 0: solver default
        int sweepLanes = 0;   // colour sweeps (0: as lanes / rowsPerBlock)
        int sweepRowsPerBlock = 0;
    };

    struct ChorinConfig
    {
        ChorinSolverConfig velocity;
        ChorinSolverConfig pressure;
        int dtype = F64;        // state and matrix precision (F32 = the reference's precision)
        double innerTol = 1e-4; // pcg_mixed inner tolerance
        bool graphs = true;     // solver iterations replayed as captured graphs
    };

    struct ChorinStepStats
    {
        double mseTentX, mseTentY, msePressure, mseFinalX, mseFinalY;
        double relTent, relPressure, relFinal;
        int itersTent, itersPressure, itersFinal;
        int outerTent, outerPressure, outerFinal;
        double msConvection, msTentative, msPressure, msFinal, msTotal;
    };

    class ChorinGpu
    {
    public:
        ChorinGpu(const ChorinContext & ctx, const ChorinConfig & cfg);
        ~ChorinGpu();

        void setTimeStep(double tau) { tau_ = tau; }
        double timeStep() const { return tau_; }
        void step(ChorinStepStats * stats);

        // Convection assembly alone: A = nu K + C(u) (and C if cOut != nullptr).
        void assembleConvection(const void * u, void * aOut, void * cOut, hipStream_t s);

        int dtype() const { return cfg_.dtype; }
        int numVelocityNodes() const { return N_; }
        int numPressureNodes() const { return V_; }
        void * velocity() { return u_.get(); }
        void * pressure() { return p_.get(); }
        void resetState();
        void imposeDirichlet(hipStream_t s);

        // Output path: a device snapshot of (u, p) after the last step,
        // copied to library-owned pinned host memory on a stream of its own
        // while later steps run; downloadWait waits for it and copies it to
        // the caller's buffers (either may be null).  One at a time.
        void downloadAsync();
        void downloadWait(void * velocity, void * pressure);
        bool downloadPending() const { return dlPending_; }

        // Average launch duration (ms) and algorithmic bytes per launch of a
        // step kernel: "convection" (element + gather kernels), "spmm" (A u),
        // "velocity_<k>" / "pressure_<k>" with k a solver kernel name.
        double timeKernel(const std::string & which, int reps, double * bytesPerLaunch, double * launchesPerIteration);

        McgsSolver & velocitySolver() { return *vSolver_; }
        McgsSolver & pressureSolver() { return *pSolver_; }

    private:
        struct DevCsr
        {
            int rows = 0, cols = 0, nnz = 0;
            DeviceArray<int> rowStart, column;
            DeviceArray<char> values;
            CsrView view(int dt) const { return CsrView{rows, cols, nnz, rowStart.get(), column.get(), values.get(), dt}; }
        };
        void upload(DevCsr & d, const CsrD & m, const std::vector<double> * scaledValues = nullptr);

        ChorinConfig cfg_;
        int N_ = 0, V_ = 0, E_ = 0;
        double tau_ = 1e-3;
        double nu_ = 0;
        size_t tsize_ = 8;

        DevCsr nuK_, A_, vpd_, pvd_;
        DeviceArray<int> elemNodes_, contribStart_, contribIndex_;
        DeviceArray<double> elemGeom_;
        DeviceArray<char> elemC_;

        DeviceArray<char> u_, p_;
        DeviceArray<int> dirIds_, internalP_;
        DeviceArray<char> dirValues_;
        int nDir_ = 0, nIntP_ = 0;

        std::unique_ptr<McgsSolver> vSolver_, pSolver_;
        hipEvent_t ev_[5];

        DeviceArray<char> snapU_, snapP_;
        void * hostU_ = nullptr; // pinned
        void * hostP_ = nullptr;
        hipStream_t copyStream_ = nullptr;
        hipEvent_t snapDone_ = nullptr, copyDone_ = nullptr;
        bool dlPending_ = false;
    };
} // namespace nsfem
