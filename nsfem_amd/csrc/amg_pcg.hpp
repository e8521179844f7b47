// CG preconditioned by a smoothed-aggregation V-cycle, for the pressure
// Poisson solve.  The smoother of a level is one multicolour Gauss-Seidel
// sweep (forward on the way down, backward on the way up) or the Chebyshev
// iteration on D^-1 A of a chosen degree (the same polynomial on both
// ways); either keeps the preconditioner symmetric.
//
// No counterpart in the reference (its pressure solve is multicolour GS,
// cpp/libs/CudaUtils/src/gaussSeidel.cu:121-427, or cuDSS); the stopping
// rule is the pcg solver's, |r| <= target |b| in FP64, and the Krylov
// vectors, products and reductions are FP64.  Precision of the V-cycle:
// FP64, or FP32 (mixed: the FP32 V-cycle on an FP32 copy of every level).
// 1, 2, 4, 8 or 16 right-hand sides (channels, interleaved per row: entry
// (i, c) at i * channels + c) share every matrix read; each channel has its
// own CG scalars and stops on its own, the solve when all have stopped.
//
// Launches per iteration: the direction (p, q = A p and p.q partials), the
// update (x, r, |r|^2 partials, the FP32 residual copy), the V-cycle (per
// Gauss-Seidel level 2 x colours + 3, per Chebyshev level of degree k
// 2 k + 2 (+1 on the finest, whose first iterate is not written by a
// restriction), one dense product on the coarsest), the r.z partials and a
// one-block finalize (rho, convergence).  Every kernel
// returns at once after the solve has stopped, so whole iterations are
// replayed from one captured graph and polled one iteration behind.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "host/amg.hpp"
#include "kernels.hpp"

namespace nsfem
{
    class AmgPcg
    {
    public:
        // Finest level: the caller's colour-major system (stripped rows with
        // rowSplit, diagonal, colour boundaries); a64 its FP64 device copy
        // (the CG products), level0 the same matrix in the V-cycle precision
        // (FP64 runs pass a64 again, mixed an FP32 copy).
        AmgPcg(const CsrD & stripped, const std::vector<int> & rowSplit, const std::vector<double> & diag,
               const std::vector<int> & colorStart, bool mixed, int numCh, McgsMatrix<double> a64, McgsMatrix<float> level0F32,
               hipStream_t s);
        ~AmgPcg();

        struct Result
        {
            int iterations = 0;            // of the channel that iterated longest
            std::vector<double> bb, rr;    // per channel |b|^2 and the final |r|^2
            bool nonFinite = false;
            bool converged = false;        // every channel
        };
        // Solves A x = b (device, colour order, FP64, channels interleaved);
        // x holds the initial guess unless zeroGuess.  Blocking.
        Result solve(const double * b, double * x, bool zeroGuess, double target, int maxIters, bool useGraphs);
        // The whole solve on the stream with no host synchronisation (for
        // graph capture): the prologue and exactly `budget` iterations, whose
        // kernels do nothing once every channel has stopped.  The device
        // scalars then hold the statistics (deviceScalars: per channel blocks
        // of rho, rho before, |b|^2, |r|^2; deviceInts: stop, iterations,
        // zero right-hand sides).
        void enqueueSolve(hipStream_t s, const double * b, double * x, bool zeroGuess, double target, int budget);
        const double * deviceScalars() const { return scal_.get(); }
        const int * deviceInts() const { return ints_.get(); }
        int numCh() const { return nch_; }

        // Average duration (ms) of one launch of `which`, HIP events around
        // reps back-to-back replays on the solver's stream: "direction",
        // "update", "rho", "vcycle" (the V-cycle's launches in sequence) or
        // "iteration" (every launch of one CG iteration, replayed from its
        // captured graph); the algorithmic bytes of one launch and the
        // launches per iteration.  x is the caller's solution vector
        // (overwritten, as are the internal work vectors).
        double timeKernel(const std::string & which, int reps, double * bytes, double * launches, double * x);

        // Smoothers: levels from `chebyshevFrom` on (but the coarsest, which
        // is solved densely) take `degree` steps of the Chebyshev iteration
        // on D^-1 A over [fraction, 1] x lambda_max (the level's bound:
        // 1.1 x its power-iteration estimate, at most the Gershgorin bound;
        // degree 2 to 8);
        // the levels before it one multicolour Gauss-Seidel sweep.  A first
        // level beyond the last smoothing level keeps Gauss-Seidel everywhere.
        void setSmoother(int chebyshevFrom, int degree, double fraction);
        int chebyshevFrom() const { return chebFrom_; }
        int chebyshevDegree() const { return chebDegree_; }
        double chebyshevFraction() const { return chebFraction_; }

        int numLevels() const { return static_cast<int>(levels_.size()); }
        int levelRows(int l) const { return levels_[l].n; }
        int levelColors(int l) const { return static_cast<int>(levels_[l].colorStart.size()) - 1; }
        double operatorComplexity() const { return complexity_; }
        double setupSeconds() const { return setupSeconds_; }

    private:
        struct Level
        {
            int n = 0;
            std::vector<int> colorStart;
            int lanes = 4, pLanes = 4, rLanes = 8;
            double nnz = 0, nnzLower = 0, nnzP = 0, nnzR = 0; // entry counts (algorithmic bytes)
            // this level's matrix (levels >= 1; level 0 uses the caller's)
            DeviceArray<int> rowStart, rowSplit, column;
            DeviceArray<double> values64, diag64, invDiag64;
            DeviceArray<float> values32, diag32, invDiag32;
            // prolongation from / restriction to the next level
            int pRows = 0, rRows = 0;
            DeviceArray<int> pRowStart, pColumn, rRowStart, rColumn;
            DeviceArray<double> pValues64, rValues64;
            DeviceArray<float> pValues32, rValues32;
            // vectors in the V-cycle precision (bytes)
            DeviceArray<char> b, x, r;
            // Chebyshev smoothing (chebDegree 0: Gauss-Seidel): step k forms
            // d = chebC1[k] d + chebC2[k] D^-1 (b - A x), x += d; xa is the
            // alternate iterate (steps read one vector and write the other),
            // d the direction
            double lambdaMax = 0;
            int chebDegree = 0;
            std::vector<double> chebC1, chebC2;
            DeviceArray<char> xa, d;
        };

        template <typename T>
        McgsMatrix<T> levelMatrix(int l) const;
        template <typename T, int NCH>
        void enqueueVcycle(hipStream_t s, const T * b0, T * x0);
        template <typename P, int NCH>
        void enqueueInit(hipStream_t s, const double * b, double * x, bool zeroGuess, double tol2);
        template <typename P, int NCH>
        void enqueueIteration(hipStream_t s, double tol2);
        // calls f.template operator()<P, NCH>() for this solver's precision and channels
        template <typename F>
        void typed(F && f);
        void destroyGraphs();

        bool mixed_ = false;
        int n_ = 0;
        int nch_ = 1;
        McgsMatrix<double> a64_{};
        McgsMatrix<float> a32_{};
        std::vector<Level> levels_;
        int coarseN_ = 0;
        DeviceArray<double> coarse64_;
        DeviceArray<float> coarse32_;
        double complexity_ = 0, setupSeconds_ = 0;
        int chebFrom_ = 1, chebDegree_ = 2;
        double chebFraction_ = 0.2;

        // Krylov vectors (FP64) and the V-cycle operands (P)
        DeviceArray<double> r_, p_, q_;
        DeviceArray<char> rP_, z_;
        int parts_ = 0;
        DeviceArray<double> scal_, partPq_, partRr_, partRho_, partBb_;
        DeviceArray<int> ints_;   // stop, iterations, channels with a zero right-hand side (bit mask), unused
        DeviceArray<int> active_; // per channel: 1 while it iterates
        int * hostPoll_ = nullptr; // pinned, 2 slots x 4 ints
        hipEvent_t pollEvents_[2] = {nullptr, nullptr};
        hipStream_t stream_ = nullptr;

        double * xDevice_ = nullptr; // the solution vector the captured iteration updates
        hipGraphExec_t iterExec_ = nullptr;
        double graphTol2_ = -1;
        int predicted_ = 0;
    };
} // namespace nsfem
