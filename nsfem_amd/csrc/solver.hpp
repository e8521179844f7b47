// Multicolour Gauss-Seidel / PCG solver engine on the GPU.
//
// Replaces cu::GaussSeidel (cpp/libs/CudaUtils/src/gaussSeidel.cu:121-427)
// behind cu::AbstractSolver (include/cu/solvers/abstractSolver.h:8-32) and
// adds the preconditioned CG path:
//   "gs"  - the reference's solver: plain multicolour GS sweeps, residual
//           RMS |b - Ax| / sqrt(n) checked every mseCheckInterval sweeps,
//           stop when every channel is below targetMse.
//   "pcg" - CG preconditioned by one symmetric multicolour GS sweep, stop
//           when |r| <= targetMse * |b| for every channel (relative).
// Precision: F64, F32, or Mixed (pcg only): FP64 outer refinement on the
// FP64 matrix around FP32 inner PCG solves; reductions are FP64 throughout.
//
// One solver can hold several independent systems (block-diagonal batch):
// each system is coloured on its own, the colour classes are concatenated
// colour-major, and convergence is tracked per system and channel.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "csr.hpp"
#include "kernels.hpp"
#include "ops.hpp"

namespace nsfem
{
    enum class SolverMethod
    {
        GaussSeidel,
        Pcg
    };

    enum class SolverPrecision
    {
        F32,
        F64,
        Mixed
    };

    struct SolverOptions
    {
        SolverMethod method = SolverMethod::GaussSeidel;
        SolverPrecision precision = SolverPrecision::F64;
        int numCh = 1;
        int maxIters = 200;
        double target = 1e-6;
        int mseCheckInterval = 1;
        int lanes = 0;           // 0: choose from the mean row length
        int rowsPerBlock = 0;    // 0: 4 row slabs per block
        int checkChunk = 16;     // iterations between host convergence polls
        double innerTol = 1e-4;  // mixed: relative tolerance of each inner solve
        int maxOuter = 40;       // mixed: refinement steps
        // Pure-Neumann (singular, zero row sum) systems with zero-mean right
        // hand sides: the solution is returned with zero mean, and the FP32
        // inner right-hand sides of the mixed path are projected to zero mean.
        bool projectMean = false;
    };

    SolverOptions parseSolverName(const std::string & name);

    template <typename T>
    struct DeviceMcgs
    {
        DeviceArray<int> rowStart, rowSplit, column;
        DeviceArray<T> values, diag, invDiag;
        McgsMatrix<T> view(int n) const
        {
            return McgsMatrix<T>{n, rowStart.get(), rowSplit.get(), column.get(), values.get(), diag.get(), invDiag.get()};
        }
    };

    struct DeviceState
    {
        DeviceArray<double> scal;    // 8 arrays of S*NCH doubles
        DeviceArray<int> ints;       // active, iters (S*NCH each), flags[4]
        DeviceArray<double> partials;
        DeviceArray<double> totals;
        SolveState view(int k, double tol2);
        void init(int numBlocks, int numSystems, int numCh);
    };

    class McgsSolver
    {
    public:
        McgsSolver(const std::vector<const CsrD *> & systems, const SolverOptions & opt);
        ~McgsSolver();

        // io vectors: natural order, systems concatenated, channels planar
        // (channel c of system s row i at [c * n + offset(s) + i]).
        void * ioRhs() { return ioRhs_.get(); }
        void * ioSol() { return ioSol_.get(); }
        int ioDtype() const { return opt_.precision == SolverPrecision::F32 ? F32 : F64; }
        int n() const { return n_; }
        int numCh() const { return opt_.numCh; }
        int numSystems() const { return numSystems_; }
        int numColors() const { return numColors_; }
        int systemOffset(int s) const { return sysOffset_[s]; }
        const std::vector<int> & coloring(int s) const { return colorings_[s]; }
        const std::vector<int> & partitionStart(int s) const { return partitions_[s]; }

        // Runs the configured method on the current io vectors (blocking).
        void solve();
        // Enqueues the solve without host synchronisation where possible.
        void solveAsync(hipStream_t stream);

        // Statistics of the last solve (per system s, channel c).
        double lastMse(int s, int c) const { return lastMse_[s * opt_.numCh + c]; }
        double lastRelRes(int s, int c) const { return lastRel_[s * opt_.numCh + c]; }
        int lastIterations(int s) const { return lastIters_[s]; }
        int lastOuterIterations() const { return lastOuter_; }
        bool lastNonFinite() const { return lastNonFinite_; }

        void setMaxIters(int n) { opt_.maxIters = n; }
        void setTarget(double t) { opt_.target = t; }
        void setMseCheckInterval(int k);
        void setInnerTol(double t) { opt_.innerTol = t; }
        void setProjectMean(bool on) { opt_.projectMean = on; }
        const SolverOptions & options() const { return opt_; }

        hipStream_t stream() const { return stream_; }
        void setUseGraphs(bool on) { useGraphs_ = on; }

        // Average duration (ms) of one launch of a solver kernel over reps
        // launches, timed with HIP events on the solver's stream, and the
        // algorithmic bytes one launch moves.  which: "direction" (A z fused
        // with the p/q updates), "sweep" (one full-row GS pass over all
        // colours, per colour launch averaged), "update".  Leaves the io
        // vectors untouched; internal work vectors are overwritten.
        double timeKernel(const std::string & which, int reps, double * bytesPerLaunch);

    private:
        template <typename T>
        void runGs(hipStream_t s);
        template <typename T>
        int runPcg(hipStream_t s, const DeviceMcgs<T> & A, SolveState st, const T * b, T * x, T * r, T * z, T * p, T * q,
                   bool zeroGuess, const int * gate, int maxIters, bool sync, bool usePrediction);
        void runMixed(hipStream_t s);
        template <typename T>
        void sgsSweeps(hipStream_t s, const DeviceMcgs<T> & A, SolveState st, const T * r, T * z);
        template <typename T>
        void launchRho(hipStream_t s, const McgsMatrix<T> & Av, SolveState st, const T * r, T * z, double * totals);
        template <typename T>
        void pcgIteration(hipStream_t s, const DeviceMcgs<T> & A, SolveState st, T * x, T * r, T * z, T * p, T * q);
        void readStats(hipStream_t s, DeviceState & st, bool gs, int launched);
        template <typename T>
        void projectMean(hipStream_t s, T * v);
        bool pollDone(hipStream_t s, DeviceState & st);

        SolverOptions opt_;
        int n_ = 0;
        int numSystems_ = 0;
        int numColors_ = 0;
        int lanes_ = 8;
        int color0End_ = 0;
        std::vector<int> sysOffset_, sysRows_;
        std::vector<std::vector<int>> colorings_, partitions_;
        std::vector<int> colorBlockStart_; // numColors + 1
        int numBlocks_ = 0;

        DeviceArray<int> perm_;
        DeviceArray<int2> blocks_;
        DeviceArray<int> blockSystem_, sysBlockStart_, sysBlocks_, systemRows_;
        BlockMap bm_{};

        DeviceMcgs<double> A64_;
        DeviceMcgs<float> A32_;

        // io (natural order) and internal (colour order) vectors
        DeviceArray<char> ioRhs_, ioSol_;
        DeviceArray<double> b64_, x64_, r64_, z64_, p64_, q64_;
        DeviceArray<float> b32_, x32_, r32_, z32_, p32_, q32_;

        DeviceState inner_, outer_;
        DeviceArray<double> means_;
        int * hostFlags_ = nullptr; // pinned
        double * hostScal_ = nullptr;
        int * hostInts_ = nullptr;

        std::vector<double> lastMse_, lastRel_;
        std::vector<int> lastIters_;
        int lastOuter_ = 0;
        bool lastNonFinite_ = false;
        hipStream_t stream_;

        // One captured graph of checkChunk PCG iterations per precision path.
        struct ChunkGraph
        {
            hipGraphExec_t exec = nullptr;
            int k = 0;
            double tol2 = -1;
        };
        ChunkGraph graph64_, graph32_;
        bool useGraphs_ = true;
        int predicted64_ = 0, predicted32_ = 0;
    };
} // namespace nsfem
