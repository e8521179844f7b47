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

    // What a time step beyond the explicit scheme's stability limit does.
    // The step is forward Euler in diffusion and convection
    // (chorinCuda.cu:263-278): it stays bounded only while tau nu
    // lambda_max(M^-1 K) <= 2.  The limit used here is half that,
    // tau_max = 1 / lambda_max(M^-1 nu K), leaving room for convection.
    enum class TimeStepPolicy : int
    {
        Reject = 0,   // an unstable tau is an error (std::invalid_argument)
        Clamp = 1,    // an unstable tau is replaced by tau_max
        Unchecked = 2 // any tau runs (the reference's behaviour)
    };

    struct ChorinConfig
    {
        ChorinSolverConfig velocity;
        ChorinSolverConfig pressure;
        int dtype = F64;        // state and matrix precision (F32 = the reference's precision)
        double innerTol = 1e-4; // pcg_mixed inner tolerance
        bool graphs = true;     // solver iterations replayed as captured graphs
        TimeStepPolicy timeStepPolicy = TimeStepPolicy::Reject;
        int lambdaIterations = 40; // power iterations of the lambda_max estimate
        // The whole step as one captured graph (pcg solvers without
        // mixedRestart); see ChorinGpu::step.
        bool stepGraph = false;
        // Ensemble: this many independent instances (1, 2, 4, 8 or 16) on the
        // context's one mesh, instance b's inflow (the Dirichlet velocity)
        // scaled by inflowScale[b] (empty: 1).  They share every matrix: the
        // velocity solves are one solve of 2B right-hand sides, the pressure
        // solve one of B; the convection is each instance's own.
        int instances = 1;
        std::vector<double> inflowScale;
        // Pressure solve's initial guess: false = the last pressure (as the
        // reference's in-place solve), true = 2 p_n - p_n-1 (linear
        // extrapolation in time; the first step after a reset uses p_n).
        bool pressureExtrapolation = false;
        // Velocity solves' initial guess (tentative and final each from its
        // own history): 0 = zero (the reference's accel.memsetZero()),
        // 1 = the same solve's solution one step before, 2 = 2 a_n - a_n-1.
        int velocityStart = 0;
    };

    struct ChorinStepStats
    {
        double mseTentX, mseTentY, msePressure, mseFinalX, mseFinalY;
        double relTent, relPressure, relFinal;
        int itersTent, itersPressure, itersFinal;
        int outerTent, outerPressure, outerFinal;
        double msConvection, msTentative, msPressure, msFinal, msTotal;
        double tauRequested, tauUsed;
        int capturedStep; // 1: ran as one captured graph (phase times are 0)
    };

    class ChorinGpu
    {
    public:
        ChorinGpu(const ChorinContext & ctx, const ChorinConfig & cfg);
        ~ChorinGpu();

        // Sets the time step under the configured policy: the requested value,
        // or tau_max (Clamp), or an std::invalid_argument (Reject, the
        // previous step stays in force) when it exceeds the stability limit.
        void setTimeStep(double tau);
        double timeStep() const { return tau_; }
        double requestedTimeStep() const { return tauRequested_; }
        // lambda_max(M^-1 nu K) (the largest over the instances of a batch),
        // estimated on first use by power iteration with the step's velocity
        // mass solver: y = M^-1 nu K x, Rayleigh quotient y.(nu K y) / y.(M y)
        // with M y = nu K x, x = y normalised.  Overwrites the velocity
        // solver's io vectors, not the state.
        double diffusionLambda();
        double stableTimeStep() { return 1.0 / diffusionLambda(); }
        // One projection step.  With the stepGraph option, after a first
        // regular step that measures the iteration counts, the whole step -
        // a device-side snapshot of the velocity and the pressure warm start,
        // convection, the tentative solve, the pressure solve, the final
        // solve, the updates and Dirichlet impositions - is one captured
        // graph launched once per step: every solve runs a fixed budget of
        // iterations (the last counts plus a margin) whose kernels do
        // nothing once that system has converged, so convergence is decided
        // on the device and the host reads the statistics only after the
        // step.  A solve that ran out of budget (still iterating below
        // maxIters) rolls the state back to the snapshot and the step runs
        // again on the regular path; the graph is then rebuilt with budgets
        // from the new counts (also when the budgets grow far above need, or
        // tau changes).
        void step(ChorinStepStats * stats);
        void setStepGraph(bool on);
        // Drops the captured step, which is captured again at the next step
        // (after a solver option changed what its launches are).
        void invalidateStepGraph() { destroyStepGraph(); }
        bool stepGraph() const { return cfg_.stepGraph; }
        int stepGraphBuilds() const { return graphBuilds_; }
        int stepGraphRollbacks() const { return graphRollbacks_; }
        double stepGraphBuildSeconds() const { return graphBuildSeconds_; }
        int stepGraphBudget(int k) const { return budget_[k]; }
        // Sets the iteration budgets of the next captured step (they adapt
        // again from its counts).
        void setStepGraphBudgets(int tent, int pres, int fin);

        // Convection assembly alone: A = nu K + C(u) (and C if cOut != nullptr).
        void assembleConvection(const void * u, void * aOut, void * cOut, hipStream_t s);

        int dtype() const { return cfg_.dtype; }
        int instances() const { return B_; }
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
        int B_ = 1; // ensemble instances: state [u_0 .. u_B-1] (2N each), [p_0 .. p_B-1] (V each)
        DeviceArray<int> pScatterIds_; // internal pressure ids of every instance (instance b: + b V)
        double tau_ = 1e-3;
        double tauRequested_ = 1e-3;
        bool tauChecked_ = false;
        double lambdaNu_ = -1;
        double nu_ = 0;
        size_t tsize_ = 8;

        void stepRegular(ChorinStepStats * stats);
        // the step's launches around the solves (every instance)
        void launchConvectionRhs(hipStream_t s);
        void launchPressureRhs(hipStream_t s);
        void launchScatterGradient(hipStream_t s);
        void startPressure(hipStream_t s);
        // velocity solve `kind` (0 tentative, 1 final): the initial guess
        // before it (returns whether it starts from zero), the history after
        bool startVelocity(int kind, hipStream_t s);
        void keepVelocity(int kind, hipStream_t s);
        DeviceArray<char> vHist_[2][2]; // [kind][0 last solution, 1 the one before]
        DeviceArray<char> pPrev_; // the pressure solution of the step before (extrapolation)
        bool stepCaptured(ChorinStepStats * stats);
        void buildStepGraph();
        void destroyStepGraph();
        void setBudgets(int tent, int pres, int fin);

        DevCsr nuK_, A_, vpd_, pvd_;
        hipGraphExec_t stepExec_ = nullptr;
        double graphTau_ = -1;
        int budget_[3] = {0, 0, 0}; // tentative, pressure, final iterations
        int graphBuilds_ = 0, graphRollbacks_ = 0;
        double graphBuildSeconds_ = 0;
        DeviceArray<char> stepSnapU_, stepSnapP_;
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
