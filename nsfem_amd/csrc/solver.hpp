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
//   "amg_pcg" - the same stopping rule, CG preconditioned by a smoothed-
//           aggregation V-cycle smoothed by multicolour GS (one system, one
//           channel; "amg_pcg_mixed": the V-cycle in FP32).
// Precision: F64, F32, or Mixed (pcg only): FP64 outer refinement on the
// FP64 matrix around FP32 inner PCG solves; reductions are FP64 throughout.
//
// One solver can hold several independent systems (block-diagonal batch):
// each system is coloured on its own, the colour classes are concatenated
// colour-major, and convergence is tracked per system and channel.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "csr.hpp"
#include "amg_pcg.hpp"
#include "bcg_kernels.hpp"
#include "kernels.hpp"
#include "ops.hpp"

namespace nsfem
{
    enum class SolverMethod
    {
        GaussSeidel,
        Pcg,
        BlockPcg, // one Krylov space per tile of 4-32 right-hand sides
        AmgPcg    // CG preconditioned by a smoothed-aggregation V-cycle with multicolour GS smoothing
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
        int sweepLanes = 0;      // colour sweeps' own shape (0: as above)
        int sweepRowsPerBlock = 0;
        bool localityOrder = true; // rows inside each colour by locality rank
        int checkChunk = 16;     // iterations between host convergence polls
        // Mixed precision.  Default: FP64 Krylov vectors, matrix products and
        // reductions, the symmetric GS preconditioner in FP32 (its sweeps on
        // an FP32 copy of the matrix).  restart = true: iterative refinement,
        // FP32 PCG corrections from zero to innerTol on the FP64 residual,
        // at most maxOuter of them.
        bool restart = false;
        double innerTol = 1e-4;
        int maxOuter = 40;
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
        DeviceArray<int> ints;       // active, iters (S*NCH each), flags[4], activeNext
        DeviceArray<double> partials, partials2, partials3;
        DeviceArray<double> totals;
        SolveState view(int k, double tol2);
        void init(int numBlocks, int numSystems, int numCh);
    };

    // Device operands of one PCG solve: Krylov vectors and matrix products
    // in T (matrix A), the symmetric GS preconditioner in P (matrix M, the
    // residual copy rP and the preconditioned vector z).  P == T: M is A and
    // rP aliases r.
    template <typename T, typename P>
    struct PcgOperands
    {
        McgsMatrix<T> A;
        McgsMatrix<P> M;
        const T * b;
        T * x;
        T * r;
        P * rP;
        P * z;
        T * p;
        T * q;
        // the factorized approximate inverse (values null unless it is the
        // preconditioner) and its intermediate vector t = G r
        McgsMatrix<P> G;
        P * t;
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
        SolverMethod method() const { return opt_.method; }
        int numSystems() const { return numSystems_; }
        int numColors() const { return numColors_; }
        int systemOffset(int s) const { return sysOffset_[s]; }
        const std::vector<int> & coloring(int s) const { return colorings_[s]; }
        const std::vector<int> & partitionStart(int s) const { return partitions_[s]; }

        // Runs the configured method on the current io vectors (blocking).
        // zeroInitial: the initial guess is zero whatever the io solution
        // vector holds (pcg: the first residual is b, no matrix product).
        void solve(bool zeroInitial = false);
        // Enqueues the solve without host synchronisation where possible.
        void solveAsync(hipStream_t stream);

        // Captured solves (pcg methods without mixedRestart): enqueueSolve
        // puts a whole solve on the stream - permutations, the PCG prologue
        // and exactly `budget` iterations (capped at maxIters) whose kernels
        // do nothing once the solve has converged - with no host
        // synchronisation, so it can be captured into a graph.
        // snapshotStats copies the statistics of the solve just enqueued to
        // device slot `slot` (< kStatSlots); collectStats reads a slot after
        // the stream has finished, sets the last-solve statistics and returns
        // false if some system was still iterating when the budget ran out
        // (below maxIters), i.e. the solve needs more iterations.
        static constexpr int kStatSlots = 4;
        void enqueueSolve(hipStream_t s, bool zeroInitial, int budget);
        void snapshotStats(hipStream_t s, int slot);
        bool collectStats(int slot);

        // Statistics of the last solve (per system s, channel c).
        double lastMse(int s, int c) const { return lastMse_[s * opt_.numCh + c]; }
        double lastRelRes(int s, int c) const { return lastRel_[s * opt_.numCh + c]; }
        int lastIterations(int s) const { return lastIters_[s]; }
        int lastOuterIterations() const { return lastOuter_; }
        bool lastNonFinite() const { return lastNonFinite_; }
        // amg_pcg: the hierarchy (nullptr for the other methods)
        const AmgPcg * amg() const { return amg_.get(); }
        AmgPcg * amg() { return amg_.get(); }
        // block PCG: whether the last solve dropped a direction (a rank-
        // deficient block)
        bool lastDroppedDirection() const { return lastDropped_; }

        void setMaxIters(int n) { opt_.maxIters = n; }
        void setTarget(double t) { opt_.target = t; }
        void setMseCheckInterval(int k);
        void setInnerTol(double t) { opt_.innerTol = t; }
        void setMixedRestart(bool on) { opt_.restart = on; }
        void setProjectMean(bool on) { opt_.projectMean = on; }
        // block PCG: relative pivot threshold of the rho = Z^T R
        // factorisation (0: 1e-13 with an FP64 preconditioner, 1e-10 with FP32)
        void setBlockPivot(double t) { blockPivot_ = t; }
        const SolverOptions & options() const { return opt_; }

        hipStream_t stream() const { return stream_; }
        void setUseGraphs(bool on) { useGraphs_ = on; }
        // Fused reductions: the direction kernel finishes the previous rho and
        // convergence test and the update kernel the step length, each block
        // from its system's partial sums, instead of two one-block finalize
        // kernels per iteration.  mode 1 or -1 (default): on; 0: off.
        void setFusedReductions(int mode);
        bool fused() const { return fusedMode_ != 0; }
        // Rows per workgroup of the direction, update and rho kernels of the
        // fused mode on one system (0: the main shape's); the sums of their
        // partials are redone by every block of the next kernel, so fewer,
        // larger blocks pay there.  reshape resets them.
        void setKindRows(int direction, int update, int rho);
        // Rows per workgroup of the approximate inverse's two launches
        // (preconditioner = 1; fused reductions on one system): 0 (default)
        // shares the direction kernel's map, else a positive multiple of 256.
        void setPreconditionerRows(int rows);
        int preconditionerRows() const { return rowsPrecond_; }
        int kindRows(int kind) const; // 0 direction, 1 update, 2 rho
        // Lanes per row of the direction kernel alone (0: the main shape's;
        // 1 and 2 channels).  reshape resets it.
        void setDirectionLanes(int lanes);
        int directionLanes() const { return lanesDir_ > 0 && opt_.numCh <= 2 ? lanesDir_ : lanes_; }
        // Region-major direction map (fused reductions, one system): the
        // rows of every colour are cut into eight stretches of the locality
        // order, and the workgroups one XCD runs take stretch k of every
        // colour, so that XCD gathers from one eighth of the vector.
        // 1 on (default), 0 off; reshape keeps it.
        void setRegionedDirection(int on);
        // Update kernel fused with the first forward sweep (fused reductions,
        // one system, 1-2 channels, two colours or more): 1 on (default), 0
        // off.  One launch fewer per iteration.  The update workgroups that
        // hold colour-1 rows sweep them in slabs of kBlock / sweep lanes
        // rows, so it applies only while that is at most two slabs (more
        // leave the launch waiting on those workgroups).
        void setFusedUpdateSweep(int on);
        bool mergedUpdate() const
        {
            return mergeMode_ && fused() && flat() && opt_.numCh <= 2 && numColors_ >= 2 && precond_ == 0 &&
                   bmUpd_.rowsPerBlock * sweepLanes_ <= 2 * kBlock;
        }
        // PCG preconditioner: 0 one symmetric multicolour GS sweep (the
        // reference's); 1 the factorized sparse approximate inverse
        // G^T G ~ A^-1 on the lower-colour pattern plus the diagonal (built on
        // the host when first chosen; pcg methods without mixedRestart, any
        // channel count; a block-diagonal batch gets a block-diagonal factor,
        // since no row has a neighbour in another system): two launches per
        // application instead of 2 colours - 2 dependent sweeps.
        void setPreconditioner(int kind);
        int preconditioner() const { return precond_; }
        int regionedDirection() const { return regionDir_; }
        // Fewest rows per workgroup of the default and autotuned shapes.
        int minRowsPerBlock() const;

        // Average duration (ms) of one launch of a solver kernel, timed with
        // HIP events on the solver's stream around reps back-to-back
        // replays of the launches of one PCG iteration, the algorithmic bytes of one launch and the
        // launches per iteration.  which: "direction", "sweep", "rho",
        // "update", "finalize".  Leaves the io vectors untouched; internal
        // work vectors are overwritten.
        double timeKernel(const std::string & which, int reps, double * bytesPerLaunch, double * launchesPerIteration);
        // Sum over one PCG iteration's launches of their timed durations (ms).
        double iterationTime(int reps);

        // Launch shape: lanes per matrix row (1, 2, 4, 8, 16; 0 picks from the mean
        // row length; wide channel tiles always run one lane per channel) and
        // rows per workgroup (0: four row slabs, for wide tiles at least
        // minRowsPerBlock()).  Rebuilds the block map only; the colouring and
        // matrices are kept.
        void reshape(int lanes, int rowsPerBlock);
        int lanes() const { return lanes_; }
        int rowsPerBlock() const { return rowsPerBlock_; }
        // The colour sweeps' own shape (they write no partial sums, so any
        // rows per workgroup go); reshape sets it to the main shape.
        void reshapeSweeps(int lanes, int rowsPerBlock);
        int sweepLanes() const { return sweepLanes_; }
        int sweepRowsPerBlock() const { return sweepRows_; }
        // Times one graph-replayed PCG iteration for lanes {1, 2, 4, 8, 16} x
        // {1, 2, 4, 8} row slabs per workgroup (wide tiles: their own lanes x
        // {1, 2, 4, 8} times minRowsPerBlock()) and keeps the fastest shape,
        // then the sweeps' own shape over lanes x {1, 2, 4} slabs, then (fused
        // reductions) {2, 4, 8, 16} times the rows for the direction, update
        // and rho kernels in turn; returns (lanes, rowsPerBlock).
        // Overwrites the internal work vectors (not the io vectors).
        std::pair<int, int> autotune(int reps);

    private:
        // One captured graph of a chunk of PCG iterations per precision path.
        struct ChunkGraph
        {
            hipGraphExec_t exec = nullptr;
            int k = 0;
            double tol2 = -1;
        };

        template <typename T>
        void runGs(hipStream_t s);
        template <typename T, typename P>
        PcgOperands<T, P> operands();
        template <typename T, typename P>
        ChunkGraph & chunkGraph();
        template <typename T, typename P>
        int & predicted();
        template <typename T, typename P>
        int runPcg(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, bool zeroGuess, const int * gate, int maxIters, bool sync,
                   bool usePrediction);
        template <typename T, typename P>
        int iterate(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, int maxIters, bool sync, bool usePrediction);
        template <typename T, typename P>
        void pcgIteration(hipStream_t s, const PcgOperands<T, P> & v, SolveState st);
        template <typename T, typename P>
        void enqueuePcg(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, bool zeroGuess, int budget);
        template <typename P>
        void sgsSweeps(hipStream_t s, const McgsMatrix<P> & M, SolveState st, const P * r, P * z, int firstForward = 1);
        template <typename T, typename P>
        void launchRho(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, double * totals, bool withUpdate);
        template <typename T, typename P>
        void launchUpdateSweep(hipStream_t s, const PcgOperands<T, P> & v, SolveState st);
        // z = preconditioner(rP) with the rho = r.z partial sums (and the
        // rho finalize without fused reductions)
        template <typename T, typename P>
        void launchPrecond(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, double * totals, bool withUpdate);
        // colour-0 rows whose z the init and update kernels set to r / d
        // (the start of the forward sweep; none for the approximate inverse)
        int zColour0End() const { return precond_ == 0 ? color0End_ : 0; }
        // blocks whose rho partials the next direction kernel sums (the
        // approximate inverse runs on the direction kernel's map)
        int rhoParts() const { return precond_ == 0 ? nbRho_ : nbPre_; }
        template <typename P>
        McgsMatrix<P> fsaiView() const;
        void runMixedRestart(hipStream_t s);
        template <typename P>
        void runBlockPcg(hipStream_t s, const McgsMatrix<P> & M, P * rP, P * z, bool zeroGuess);
        template <typename P>
        void bcgIteration(hipStream_t s, const BcgState & b, SolveState st, const McgsMatrix<P> & M, P * rP, P * z);
        BcgState bcgState(SolveState st);
        void readStats(hipStream_t s, DeviceState & st, bool gs, int launched);
        template <typename T>
        void projectMean(hipStream_t s, T * v);
        bool pollDone(hipStream_t s, DeviceState & st);
        // The io vectors (original order, channel-major) to or from the
        // solver's colour-major rows (inverse), same precision dt on both
        // sides, 1-2 channels: into colour order over the region-major
        // direction map when there is one, out of it as a gather in original
        // order; otherwise one element per thread.
        void permuteRows(int dt, const void * src, void * dst, bool inverse, hipStream_t s);

        SolverOptions opt_;
        int n_ = 0;
        int numSystems_ = 0;
        int numColors_ = 0;
        int lanes_ = 8;
        int color0End_ = 0;
        std::vector<int> sysOffset_, sysRows_;
        std::vector<std::vector<int>> colorings_, partitions_;
        struct RowChunk
        {
            int color, system, begin, end; // rows [begin, end) of the global order
        };
        std::vector<RowChunk> chunks_;     // colour-major
        double meanRow_ = 0;               // off-diagonal entries per row
        int rowsPerBlock_ = 0;
        std::vector<int> colorRowStart_; // numColors + 1 (global colour-major order)
        // one system: colour c's rows split into 8 stretches of the locality
        // order, [colorRegionStart_[9c + k], colorRegionStart_[9c + k + 1])
        std::vector<int> colorRegionStart_;
        int regionDir_ = 1;
        int mergeMode_ = 1;
        DeviceArray<int2> dirTable_;
        int color0Blocks_ = 0;           // main-map blocks holding colour-0 rows (leading)
        // one system: flat workgroup maps (ranges computed in the kernels)
        bool flat() const { return numSystems_ == 1; }
        struct ColorStats
        {
            double rows = 0, lowerNnz = 0, fullNnz = 0, lowerCols = 0, fullCols = 0;
        };
        std::vector<ColorStats> colorStats_;
        int numBlocks_ = 0;
        int fusedMode_ = -1;
        // per-kernel-kind maps: all bm_ unless fused (flat, own row counts)
        int rowsDir_ = 0, rowsUpd_ = 0, rowsRho_ = 0, lanesDir_ = 0;
        BlockMap bmDir_{}, bmUpd_{}, bmRho_{};
        int nbDir_ = 0, nbUpd_ = 0, nbRho_ = 0, color0BlocksRho_ = 0;
        // the approximate inverse's map: the direction map, or (rowsPrecond_
        // > 0, fused reductions on one system) a region-major map of its own
        int rowsPrecond_ = 0;
        BlockMap bmPre_{};
        int nbPre_ = 0;
        DeviceArray<int2> preTable_;
        void setupKindMaps();
        // the colour sweeps' workgroup map
        int sweepLanes_ = 8, sweepRows_ = 0;
        std::vector<int> colorSweepBlockStart_;
        DeviceArray<int4> sweepBlocks_;
        BlockMap sweepBm_{};
        std::pair<int, int> resolveShape(int lanes, int rowsPerBlock, bool sweeps) const;
        void buildBlocks(int rpb, std::vector<int4> & blocks, std::vector<int> & blockSystem, std::vector<int> & colorStart) const;
        void destroyGraphs();

        DeviceArray<int> perm_;
        DeviceArray<int> iperm_; // the inverse: original row -> colour-major row
        DeviceArray<int4> blocks_;
        DeviceArray<int> sysBlockStart_, sysBlocks_, systemRows_;
        BlockMap bm_{};

        DeviceMcgs<double> A64_;
        DeviceMcgs<float> A32_;

        // io (natural order) and internal (colour order) vectors
        DeviceArray<char> ioRhs_, ioSol_;
        DeviceArray<double> b64_, x64_, r64_, z64_, p64_, q64_;
        DeviceArray<float> b32_, x32_, r32_, z32_, p32_, q32_;

        DeviceState inner_, outer_;
        DeviceArray<char> statSlots_; // kStatSlots x (8K doubles, 3K + 4 ints)
        // block PCG: flat row partition of its own (few workgroups, so the
        // NCH x NCH Gram partials stay small), Gram and norm partials, the
        // small matrices, the iteration counter
        int bcgRows_ = 0, bcgBlocks_ = 0;
        DeviceArray<double> bcgGram_, bcgNorm_, bcgSmall_;
        DeviceArray<int> bcgInts_;
        bool lastDropped_ = false;
        double blockPivot_ = 0;
        DeviceArray<double> means_;
        int * hostFlags_ = nullptr; // pinned
        int * pollRing_ = nullptr;  // pinned, two slots of the solve flags
        hipEvent_t pollEvents_[2] = {nullptr, nullptr};
        double * hostScal_ = nullptr;
        int * hostInts_ = nullptr;

        std::vector<double> lastMse_, lastRel_;
        std::vector<int> lastIters_;
        int lastOuter_ = 0;
        bool lastNonFinite_ = false;
        hipStream_t stream_;

        int precond_ = 0;
        DeviceArray<double> fsai64_, fsaiDiag64_;
        DeviceArray<float> fsai32_, fsaiDiag32_;
        DeviceArray<char> fsaiT_;

        std::unique_ptr<AmgPcg> amg_;
        ChunkGraph graph64_, graph32_, graphMixed_;
        bool useGraphs_ = true;
        int predicted64_ = 0, predicted32_ = 0, predictedMixed_ = 0;
        // iterations the last pcg solve took (its slowest channel): the
        // next solve polls the stop flag only from a chunk before that count
        int lastSolveIters_ = 0;
    };
} // namespace nsfem
