#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>

namespace nsfem
{
    namespace
    {
        template <typename F>
        void dispatchLC(int lanes, int nch, F && f)
        {
            auto byC = [&]<int L>()
            {
                if (nch == 1)
                {
                    f.template operator()<L, 1>();
                }
                else if (nch == 2)
                {
                    f.template operator()<L, 2>();
                }
                else
                {
                    throw std::invalid_argument("Only 1 and 2 channel solvers are supported");
                }
            };
            switch (lanes)
            {
            case 4: byC.template operator()<4>(); break;
            case 8: byC.template operator()<8>(); break;
            case 16: byC.template operator()<16>(); break;
            default: throw std::invalid_argument("solver lanes must be 4, 8 or 16");
            }
        }

        template <typename T>
        void uploadMatrix(DeviceMcgs<T> & d, const std::vector<int> & rowStart, const std::vector<int> & rowSplit,
                          const std::vector<int> & column, const std::vector<double> & values, const std::vector<double> & diag,
                          const std::vector<double> & invDiag)
        {
            d.rowStart.upload(rowStart);
            d.rowSplit.upload(rowSplit);
            d.column.upload(column.empty() ? std::vector<int>{0} : column);
            auto cast = [](const std::vector<double> & v)
            {
                std::vector<T> o(v.size());
                for (size_t i = 0; i < v.size(); i++)
                {
                    o[i] = static_cast<T>(v[i]);
                }
                if (o.empty())
                {
                    o.push_back(T(0));
                }
                return o;
            };
            d.values.upload(cast(values));
            d.diag.upload(cast(diag));
            std::vector<T> inv(invDiag.size());
            for (size_t i = 0; i < inv.size(); i++)
            {
                // The FP32 inverse is taken of the FP32 diagonal, as a float
                // solver would compute it (gaussSeidel.cpp:43).
                inv[i] = T(1) / static_cast<T>(diag[i]);
            }
            d.invDiag.upload(inv);
        }
    } // namespace

    SolverOptions parseSolverName(const std::string & name)
    {
        SolverOptions o;
        if (name == "gs")
        {
            o.method = SolverMethod::GaussSeidel;
        }
        else if (name == "pcg")
        {
            o.method = SolverMethod::Pcg;
        }
        else if (name == "pcg_mixed")
        {
            o.method = SolverMethod::Pcg;
            o.precision = SolverPrecision::Mixed;
        }
        else if (name == "dss")
        {
            throw std::invalid_argument("Solver 'dss' (cuDSS direct solve) is not available in this build; use 'gs', 'pcg' or 'pcg_mixed'");
        }
        else
        {
            throw std::invalid_argument("Invalid solver name: " + name);
        }
        return o;
    }

    void DeviceState::init(int numBlocks, int numSystems, int numCh)
    {
        const int K = numSystems * numCh;
        scal.resize(static_cast<size_t>(7) * K);
        ints.resize(static_cast<size_t>(2) * K + 4);
        partials.resize(static_cast<size_t>(std::max(1, numBlocks)) * 2 * numCh);
        totals.resize(static_cast<size_t>(numSystems) * 2 * numCh);
        scal.zero();
        ints.zero();
    }

    SolveState DeviceState::view(int K, double tol2)
    {
        SolveState s;
        double * d = scal.get();
        s.rho = d;
        s.alpha = d + K;
        s.beta = d + 2 * K;
        s.rr = d + 3 * K;
        s.bb = d + 4 * K;
        s.ref = d + 5 * K;
        s.mse = d + 6 * K;
        s.active = ints.get();
        s.iters = ints.get() + K;
        s.flags = ints.get() + 2 * K;
        s.partials = partials.get();
        s.tol2 = tol2;
        return s;
    }

    McgsSolver::McgsSolver(const std::vector<const CsrD *> & systems, const SolverOptions & opt) : opt_(opt)
    {
        if (systems.empty())
        {
            throw std::invalid_argument("Solver needs at least one matrix");
        }
        if (opt_.numCh < 1 || opt_.numCh > 2)
        {
            throw std::invalid_argument("Only 1 and 2 ch solvers are supported");
        }
        if (opt_.mseCheckInterval < 1)
        {
            throw std::invalid_argument("MSE check interval should be at least 1");
        }
        if (opt_.method == SolverMethod::GaussSeidel && opt_.precision == SolverPrecision::Mixed)
        {
            throw std::invalid_argument("Mixed precision needs the pcg method");
        }
        stream_ = libraryStream();
        numSystems_ = static_cast<int>(systems.size());

        std::vector<McgsLayout> layouts;
        layouts.reserve(numSystems_);
        double nnzOff = 0;
        for (int s = 0; s < numSystems_; s++)
        {
            layouts.push_back(buildMcgsLayout(*systems[s]));
            sysOffset_.push_back(n_);
            sysRows_.push_back(layouts.back().n);
            n_ += layouts.back().n;
            nnzOff += layouts.back().stripped.nnz();
            numColors_ = std::max(numColors_, layouts.back().numColors());
            colorings_.push_back(layouts.back().coloring);
            partitions_.push_back(layouts.back().partitionStart);
        }
        lanes_ = opt_.lanes > 0 ? opt_.lanes : std::min(16, lanesForMeanRow(nnzOff / n_));
        if (lanes_ != 4 && lanes_ != 8 && lanes_ != 16)
        {
            throw std::invalid_argument("solver lanes must be 4, 8 or 16");
        }
        const int slab = kBlock / lanes_;
        int rowsPerBlock = opt_.rowsPerBlock > 0 ? opt_.rowsPerBlock : 4 * slab;
        rowsPerBlock = std::max(slab, rowsPerBlock / slab * slab);

        // Colour-major global order: colour c, then system s, then the
        // system's own colour order.
        std::vector<int> perm(n_);
        std::vector<std::vector<int>> gpos(numSystems_);
        for (int s = 0; s < numSystems_; s++)
        {
            gpos[s].resize(layouts[s].n);
        }
        std::vector<int2> blocks;
        std::vector<int> blockSystem;
        colorBlockStart_.assign(numColors_ + 1, 0);
        int g = 0;
        for (int c = 0; c < numColors_; c++)
        {
            colorBlockStart_[c] = static_cast<int>(blocks.size());
            for (int s = 0; s < numSystems_; s++)
            {
                const auto & L = layouts[s];
                if (c >= L.numColors())
                {
                    continue;
                }
                const int l0 = L.partitionStart[c], l1 = L.partitionStart[c + 1];
                const int g0 = g;
                for (int l = l0; l < l1; l++)
                {
                    gpos[s][l] = g;
                    perm[g] = sysOffset_[s] + L.coloring[l];
                    g++;
                }
                for (int b0 = g0; b0 < g; b0 += rowsPerBlock)
                {
                    blocks.push_back(int2{b0, std::min(g, b0 + rowsPerBlock)});
                    blockSystem.push_back(s);
                }
            }
            if (c == 0)
            {
                color0End_ = g;
            }
        }
        colorBlockStart_[numColors_] = static_cast<int>(blocks.size());
        numBlocks_ = static_cast<int>(blocks.size());

        // Global stripped matrix in colour-major order.
        std::vector<int> rowStart(n_ + 1, 0), rowSplit(n_), column;
        std::vector<double> values, diag(n_), invDiag(n_);
        column.reserve(static_cast<size_t>(nnzOff));
        values.reserve(static_cast<size_t>(nnzOff));
        {
            std::vector<std::pair<int, int>> owner(n_); // global row -> (system, local permuted row)
            for (int s = 0; s < numSystems_; s++)
            {
                for (int l = 0; l < layouts[s].n; l++)
                {
                    owner[gpos[s][l]] = {s, l};
                }
            }
            for (int r = 0; r < n_; r++)
            {
                const auto [s, l] = owner[r];
                const auto & L = layouts[s];
                rowStart[r] = static_cast<int>(column.size());
                for (int j = L.stripped.rowStart[l]; j < L.stripped.rowStart[l + 1]; j++)
                {
                    column.push_back(gpos[s][L.stripped.column[j]]);
                    values.push_back(L.stripped.values[j]);
                }
                rowSplit[r] = rowStart[r] + (L.rowSplit[l] - L.stripped.rowStart[l]);
                diag[r] = L.diag[l];
                invDiag[r] = L.invDiag[l];
            }
            rowStart[n_] = static_cast<int>(column.size());
        }

        if (opt_.precision != SolverPrecision::F32)
        {
            uploadMatrix(A64_, rowStart, rowSplit, column, values, diag, invDiag);
        }
        if (opt_.precision != SolverPrecision::F64)
        {
            uploadMatrix(A32_, rowStart, rowSplit, column, values, diag, invDiag);
        }

        // Block map
        std::vector<int> sysBlockStart(numSystems_ + 1, 0), sysBlocks(numBlocks_);
        for (int b = 0; b < numBlocks_; b++)
        {
            sysBlockStart[blockSystem[b] + 1]++;
        }
        for (int s = 0; s < numSystems_; s++)
        {
            sysBlockStart[s + 1] += sysBlockStart[s];
        }
        {
            std::vector<int> fillPos(sysBlockStart.begin(), sysBlockStart.end() - 1);
            for (int b = 0; b < numBlocks_; b++)
            {
                sysBlocks[fillPos[blockSystem[b]]++] = b;
            }
        }
        perm_.upload(perm);
        blocks_.upload(blocks);
        blockSystem_.upload(blockSystem);
        sysBlockStart_.upload(sysBlockStart);
        sysBlocks_.upload(sysBlocks);
        systemRows_.upload(sysRows_);
        bm_ = BlockMap{blocks_.get(), blockSystem_.get(), sysBlockStart_.get(), sysBlocks_.get(), systemRows_.get(), numSystems_};

        const size_t nv = static_cast<size_t>(n_) * opt_.numCh;
        const size_t ioBytes = nv * (ioDtype() == F64 ? 8 : 4);
        ioRhs_.resize(ioBytes);
        ioSol_.resize(ioBytes);
        ioRhs_.zero();
        ioSol_.zero();
        if (opt_.precision == SolverPrecision::F64)
        {
            for (auto * a : {&b64_, &x64_, &r64_, &z64_, &p64_, &q64_})
            {
                a->resize(nv);
                a->zero();
            }
        }
        else if (opt_.precision == SolverPrecision::F32)
        {
            for (auto * a : {&b32_, &x32_, &r32_, &z32_, &p32_, &q32_})
            {
                a->resize(nv);
                a->zero();
            }
        }
        else
        {
            for (auto * a : {&b64_, &x64_, &r64_, &z64_})
            {
                a->resize(nv);
                a->zero();
            }
            for (auto * a : {&b32_, &x32_, &r32_, &z32_, &p32_, &q32_})
            {
                a->resize(nv);
                a->zero();
            }
            outer_.init(numBlocks_, numSystems_, opt_.numCh);
        }
        inner_.init(numBlocks_, numSystems_, opt_.numCh);
        means_.resize(static_cast<size_t>(numSystems_) * opt_.numCh);

        const int K = numSystems_ * opt_.numCh;
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&hostFlags_), 4 * sizeof(int), hipHostMallocDefault));
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&hostScal_), 7 * K * sizeof(double), hipHostMallocDefault));
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&hostInts_), (2 * K + 4) * sizeof(int), hipHostMallocDefault));
        lastMse_.assign(K, -1.0);
        lastRel_.assign(K, -1.0);
        lastIters_.assign(numSystems_, -1);
        syncLibraryStream();
    }

    McgsSolver::~McgsSolver()
    {
        (void)hipStreamSynchronize(stream_);
        for (ChunkGraph * g : {&graph64_, &graph32_})
        {
            if (g->exec)
            {
                (void)hipGraphExecDestroy(g->exec);
            }
        }
        if (hostFlags_)
        {
            (void)hipHostFree(hostFlags_);
        }
        if (hostScal_)
        {
            (void)hipHostFree(hostScal_);
        }
        if (hostInts_)
        {
            (void)hipHostFree(hostInts_);
        }
    }

    void McgsSolver::setMseCheckInterval(int k)
    {
        if (k < 1)
        {
            throw std::invalid_argument("MSE check interval should be at least 1");
        }
        opt_.mseCheckInterval = k;
    }

    bool McgsSolver::pollDone(hipStream_t s, DeviceState & st)
    {
        const int K = numSystems_ * opt_.numCh;
        NSFEM_HIP(hipMemcpyAsync(hostFlags_, st.ints.get() + 2 * K, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipStreamSynchronize(s));
        return hostFlags_[0] != 0;
    }

    void McgsSolver::readStats(hipStream_t s, DeviceState & st, bool gs, int launched)
    {
        const int K = numSystems_ * opt_.numCh;
        NSFEM_HIP(hipMemcpyAsync(hostScal_, st.scal.get(), 7 * K * sizeof(double), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipMemcpyAsync(hostInts_, st.ints.get(), (2 * K + 4) * sizeof(int), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipStreamSynchronize(s));
        const double * rr = hostScal_ + 3 * K;
        const double * ref = hostScal_ + 5 * K;
        const double * mse = hostScal_ + 6 * K;
        const int * active = hostInts_;
        const int * iters = hostInts_ + K;
        lastNonFinite_ = hostInts_[2 * K + 2] != 0;
        for (int sys = 0; sys < numSystems_; sys++)
        {
            int it = 0;
            for (int c = 0; c < opt_.numCh; c++)
            {
                const int k = sys * opt_.numCh + c;
                if (gs)
                {
                    lastMse_[k] = mse[k];
                    lastRel_[k] = -1.0;
                    it = active[k] ? launched : iters[k];
                }
                else
                {
                    lastMse_[k] = std::sqrt(rr[k] / sysRows_[sys]);
                    lastRel_[k] = ref[k] > 0 ? std::sqrt(rr[k] / ref[k]) : 0.0;
                    it = std::max(it, iters[k]);
                }
            }
            lastIters_[sys] = it;
        }
    }

    template <typename T>
    void McgsSolver::runGs(hipStream_t s)
    {
        const int nch = opt_.numCh;
        const int K = numSystems_ * nch;
        const bool f64 = std::is_same_v<T, double>;
        const DeviceMcgs<T> & A = [&]() -> const DeviceMcgs<T> &
        {
            if constexpr (std::is_same_v<T, double>)
            {
                return A64_;
            }
            else
            {
                return A32_;
            }
        }();
        T * b = f64 ? reinterpret_cast<T *>(b64_.get()) : reinterpret_cast<T *>(b32_.get());
        T * x = f64 ? reinterpret_cast<T *>(x64_.get()) : reinterpret_cast<T *>(x32_.get());
        const int dt = f64 ? F64 : F32;

        SolveState st = inner_.view(K, opt_.target);
        fillInt(K, 1, st.active, s);
        fillInt(K + 4, 0, st.iters, s); // iters + flags
        fill(F64, K, -1.0, st.mse, s);
        permute(dt, dt, n_, nch, perm_.get(), ioRhs_.get(), b, false, s);
        permute(dt, dt, n_, nch, perm_.get(), ioSol_.get(), x, false, s);

        const McgsMatrix<T> Av = A.view(n_);
        int iter = 0;
        for (; iter < opt_.maxIters; iter++)
        {
            for (int c = 0; c < numColors_; c++)
            {
                const int cnt = colorBlockStart_[c + 1] - colorBlockStart_[c];
                if (cnt == 0)
                {
                    continue;
                }
                dispatchLC(lanes_, nch, [&]<int L, int C>()
                           { mcgsColorKernel<T, L, C><<<cnt, kBlock, 0, s>>>(Av, bm_, colorBlockStart_[c], st, b, x, 0); });
            }
            if (iter % opt_.mseCheckInterval == 0)
            {
                dispatchLC(lanes_, nch, [&]<int L, int C>()
                           {
                               gsResidualKernel<T, L, C><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, b, x);
                               gsResidualFinalizeKernel<C><<<1, kBlock, 0, s>>>(bm_, st, iter, inner_.totals.get());
                           });
            }
            NSFEM_HIP(hipGetLastError());
            if ((iter + 1) % opt_.checkChunk == 0 && pollDone(s, inner_))
            {
                iter++;
                break;
            }
        }
        readStats(s, inner_, true, opt_.maxIters);
        permute(dt, dt, n_, nch, perm_.get(), x, ioSol_.get(), true, s);
        NSFEM_HIP(hipStreamSynchronize(s));
    }

    template <typename T>
    void McgsSolver::sgsSweeps(hipStream_t s, const DeviceMcgs<T> & A, SolveState st, const T * r, T * z)
    {
        const McgsMatrix<T> Av = A.view(n_);
        const int nch = opt_.numCh;
        // forward colours 1 .. C-1 on the lower part (colour 0 was set to r/d)
        for (int c = 1; c < numColors_; c++)
        {
            const int cnt = colorBlockStart_[c + 1] - colorBlockStart_[c];
            if (cnt == 0)
            {
                continue;
            }
            dispatchLC(lanes_, nch, [&]<int L, int C>()
                       { mcgsColorKernel<T, L, C><<<cnt, kBlock, 0, s>>>(Av, bm_, colorBlockStart_[c], st, r, z, 1); });
        }
        // backward colours C-2 .. 1 on full rows (colour C-1 is unchanged)
        for (int c = numColors_ - 2; c >= 1; c--)
        {
            const int cnt = colorBlockStart_[c + 1] - colorBlockStart_[c];
            if (cnt == 0)
            {
                continue;
            }
            dispatchLC(lanes_, nch, [&]<int L, int C>()
                       { mcgsColorKernel<T, L, C><<<cnt, kBlock, 0, s>>>(Av, bm_, colorBlockStart_[c], st, r, z, 0); });
        }
    }

    template <typename T>
    void McgsSolver::launchRho(hipStream_t s, const McgsMatrix<T> & Av, SolveState st, const T * r, T * z, double * totals)
    {
        dispatchLC(lanes_, opt_.numCh, [&]<int L, int C>()
                   {
                       pcgRhoKernel<T, L, C><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, r, z, color0End_);
                       pcgRhoFinalizeKernel<C><<<1, kBlock, 0, s>>>(bm_, st, totals);
                   });
    }

    template <typename T>
    void McgsSolver::pcgIteration(hipStream_t s, const DeviceMcgs<T> & A, SolveState st, T * x, T * r, T * z, T * p, T * q)
    {
        const McgsMatrix<T> Av = A.view(n_);
        const int nch = opt_.numCh;
        double * totals = inner_.totals.get();
        dispatchLC(lanes_, nch, [&]<int L, int C>()
                   {
                       pcgDirectionKernel<T, L, C><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, z, p, q);
                       pcgDirectionFinalizeKernel<C><<<1, kBlock, 0, s>>>(bm_, st, totals);
                   });
        if (nch == 1)
        {
            pcgUpdateKernel<T, 1><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, p, q, x, r, z, color0End_);
            pcgUpdateFinalizeKernel<1><<<1, kBlock, 0, s>>>(bm_, st, totals);
        }
        else
        {
            pcgUpdateKernel<T, 2><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, p, q, x, r, z, color0End_);
            pcgUpdateFinalizeKernel<2><<<1, kBlock, 0, s>>>(bm_, st, totals);
        }
        sgsSweeps<T>(s, A, st, r, z);
        launchRho<T>(s, Av, st, r, z, totals);
    }

    template <typename T>
    int McgsSolver::runPcg(hipStream_t s, const DeviceMcgs<T> & A, SolveState st, const T * b, T * x, T * r, T * z, T * p, T * q,
                           bool zeroGuess, const int * gate, int maxIters, bool sync, bool usePrediction)
    {
        const int nch = opt_.numCh;
        DeviceState & ds = inner_;
        const McgsMatrix<T> Av = A.view(n_);
        double * totals = ds.totals.get();
        fillInt(4, 0, st.flags, s);
        dispatchLC(lanes_, nch, [&]<int L, int C>()
                   {
                       pcgInitKernel<T, T, L, C><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, b, x, r, z, color0End_, zeroGuess ? 1 : 0);
                       pcgInitFinalizeKernel<C><<<1, kBlock, 0, s>>>(bm_, st, gate, totals);
                   });
        sgsSweeps<T>(s, A, st, r, z);
        launchRho<T>(s, Av, st, r, z, totals);
        NSFEM_HIP(hipGetLastError());

        // Iterations run in chunks of checkChunk; a full chunk is one launch
        // of a captured graph.  Kernels after convergence exit at once.
        ChunkGraph & g = std::is_same_v<T, double> ? graph64_ : graph32_;
        const int chunk = std::max(1, opt_.checkChunk);
        if (useGraphs_ && (g.exec == nullptr || g.k != chunk || g.tol2 != st.tol2))
        {
            if (g.exec)
            {
                NSFEM_HIP(hipGraphExecDestroy(g.exec));
                g.exec = nullptr;
            }
            hipGraph_t graph = nullptr;
            NSFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < chunk; i++)
            {
                pcgIteration<T>(s, A, st, x, r, z, p, q);
            }
            NSFEM_HIP(hipStreamEndCapture(s, &graph));
            NSFEM_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
            NSFEM_HIP(hipGraphDestroy(graph));
            g.k = chunk;
            g.tol2 = st.tol2;
        }
        // The first host poll comes one chunk before the iteration count of
        // the previous solve on this path (warm-started solves repeat it).
        int & predicted = std::is_same_v<T, double> ? predicted64_ : predicted32_;
        const int firstPoll = usePrediction ? std::max(chunk, predicted - chunk) : chunk;
        int launched = 0;
        while (launched < maxIters)
        {
            const int k = std::min(chunk, maxIters - launched);
            if (useGraphs_ && k == chunk)
            {
                NSFEM_HIP(hipGraphLaunch(g.exec, s));
            }
            else
            {
                for (int i = 0; i < k; i++)
                {
                    pcgIteration<T>(s, A, st, x, r, z, p, q);
                }
            }
            NSFEM_HIP(hipGetLastError());
            launched += k;
            if (sync && launched >= firstPoll && pollDone(s, ds))
            {
                break;
            }
        }
        if (usePrediction)
        {
            predicted = launched;
        }
        return launched;
    }

    void McgsSolver::runMixed(hipStream_t s)
    {
        const int nch = opt_.numCh;
        const int K = numSystems_ * nch;
        const long nv = static_cast<long>(n_) * nch;
        permute(F64, F64, n_, nch, perm_.get(), ioRhs_.get(), b64_.get(), false, s);
        permute(F64, F64, n_, nch, perm_.get(), ioSol_.get(), x64_.get(), false, s);
        SolveState outer = outer_.view(K, opt_.target * opt_.target);
        SolveState inner = inner_.view(K, opt_.innerTol * opt_.innerTol);
        const McgsMatrix<double> A64 = A64_.view(n_);
        std::vector<int> totalIters(numSystems_, 0);
        int outerIt = 0;
        for (;; outerIt++)
        {
            fillInt(4, 0, outer.flags, s);
            dispatchLC(lanes_, nch, [&]<int L, int C>()
                       {
                           pcgInitKernel<double, double, L, C><<<numBlocks_, kBlock, 0, s>>>(A64, bm_, outer, b64_.get(), x64_.get(), r64_.get(),
                                                                                             z64_.get(), color0End_, 0);
                           pcgInitFinalizeKernel<C><<<1, kBlock, 0, s>>>(bm_, outer, nullptr, outer_.totals.get());
                       });
            NSFEM_HIP(hipGetLastError());
            if (pollDone(s, outer_) || outerIt >= opt_.maxOuter)
            {
                break;
            }
            convert(F64, F32, nv, r64_.get(), b32_.get(), s);
            if (opt_.projectMean)
            {
                projectMean<float>(s, b32_.get());
            }
            runPcg<float>(s, A32_, inner, b32_.get(), x32_.get(), r32_.get(), z32_.get(), p32_.get(), q32_.get(), true,
                          outer.active, opt_.maxIters, true, false);
            // accumulate inner iteration counts
            NSFEM_HIP(hipMemcpyAsync(hostInts_, inner_.ints.get(), (2 * K + 4) * sizeof(int), hipMemcpyDeviceToHost, s));
            addConvertKernel<float><<<elementwiseGrid(nv), 256, 0, s>>>(nv, x32_.get(), x64_.get());
            NSFEM_HIP(hipGetLastError());
            NSFEM_HIP(hipStreamSynchronize(s));
            for (int sys = 0; sys < numSystems_; sys++)
            {
                int it = 0;
                for (int c = 0; c < nch; c++)
                {
                    it = std::max(it, hostInts_[K + sys * nch + c]);
                }
                totalIters[sys] += it;
            }
        }
        readStats(s, outer_, false, 0);
        for (int sys = 0; sys < numSystems_; sys++)
        {
            lastIters_[sys] = totalIters[sys];
        }
        lastOuter_ = outerIt;
        if (opt_.projectMean)
        {
            projectMean<double>(s, x64_.get());
        }
        permute(F64, F64, n_, nch, perm_.get(), x64_.get(), ioSol_.get(), true, s);
        NSFEM_HIP(hipStreamSynchronize(s));
    }

    void McgsSolver::solve()
    {
        hipStream_t s = stream_;
        const int nch = opt_.numCh;
        const int K = numSystems_ * nch;
        lastOuter_ = 0;
        if (opt_.method == SolverMethod::GaussSeidel)
        {
            if (opt_.precision == SolverPrecision::F64)
            {
                runGs<double>(s);
            }
            else
            {
                runGs<float>(s);
            }
            return;
        }
        if (opt_.precision == SolverPrecision::Mixed)
        {
            runMixed(s);
            return;
        }
        const double tol2 = opt_.target * opt_.target;
        if (opt_.precision == SolverPrecision::F64)
        {
            permute(F64, F64, n_, nch, perm_.get(), ioRhs_.get(), b64_.get(), false, s);
            permute(F64, F64, n_, nch, perm_.get(), ioSol_.get(), x64_.get(), false, s);
            SolveState st = inner_.view(K, tol2);
            runPcg<double>(s, A64_, st, b64_.get(), x64_.get(), r64_.get(), z64_.get(), p64_.get(), q64_.get(), false, nullptr,
                           opt_.maxIters, true, true);
            readStats(s, inner_, false, 0);
            if (opt_.projectMean)
            {
                projectMean<double>(s, x64_.get());
            }
            permute(F64, F64, n_, nch, perm_.get(), x64_.get(), ioSol_.get(), true, s);
        }
        else
        {
            permute(F32, F32, n_, nch, perm_.get(), ioRhs_.get(), b32_.get(), false, s);
            permute(F32, F32, n_, nch, perm_.get(), ioSol_.get(), x32_.get(), false, s);
            SolveState st = inner_.view(K, tol2);
            runPcg<float>(s, A32_, st, b32_.get(), x32_.get(), r32_.get(), z32_.get(), p32_.get(), q32_.get(), false, nullptr,
                          opt_.maxIters, true, true);
            readStats(s, inner_, false, 0);
            if (opt_.projectMean)
            {
                projectMean<float>(s, x32_.get());
            }
            permute(F32, F32, n_, nch, perm_.get(), x32_.get(), ioSol_.get(), true, s);
        }
        NSFEM_HIP(hipStreamSynchronize(s));
    }

    template <typename T>
    void McgsSolver::projectMean(hipStream_t s, T * v)
    {
        // reuse the inner state's reduction scratch (no solve is in flight)
        if (opt_.numCh == 1)
        {
            meanKernel<T, 1><<<numBlocks_, kBlock, 0, s>>>(n_, bm_, v, inner_.partials.get());
            meanFinalizeKernel<1><<<1, kBlock, 0, s>>>(bm_, inner_.partials.get(), inner_.totals.get(), means_.get());
            subtractMeanKernel<T, 1><<<numBlocks_, kBlock, 0, s>>>(n_, bm_, v, means_.get());
        }
        else
        {
            meanKernel<T, 2><<<numBlocks_, kBlock, 0, s>>>(n_, bm_, v, inner_.partials.get());
            meanFinalizeKernel<2><<<1, kBlock, 0, s>>>(bm_, inner_.partials.get(), inner_.totals.get(), means_.get());
            subtractMeanKernel<T, 2><<<numBlocks_, kBlock, 0, s>>>(n_, bm_, v, means_.get());
        }
        NSFEM_HIP(hipGetLastError());
    }

    void McgsSolver::solveAsync(hipStream_t) { solve(); }

    double McgsSolver::timeKernel(const std::string & which, int reps, double * bytesPerLaunch)
    {
        if (reps < 1)
        {
            throw std::invalid_argument("reps must be positive");
        }
        if (opt_.method != SolverMethod::Pcg)
        {
            throw std::invalid_argument("kernel timing is available for pcg solvers");
        }
        hipStream_t s = stream_;
        const int nch = opt_.numCh;
        const int K = numSystems_ * nch;
        const bool f64 = opt_.precision == SolverPrecision::F64;
        const double ts = f64 ? 8.0 : 4.0;
        SolveState st = inner_.view(K, 0.0);
        double * totals = inner_.totals.get();
        long nnzOff = 0;
        {
            int last = 0;
            NSFEM_HIP(hipMemcpy(&last, (f64 ? A64_.rowStart.get() : A32_.rowStart.get()) + n_, sizeof(int), hipMemcpyDeviceToHost));
            nnzOff = last;
        }
        const double n = n_;
        const double matrixBytes = nnzOff * (ts + 4.0) + (n + 1) * 4.0 + n * ts;
        double bytes = 0;
        std::function<void()> launch;
        auto run = [&](auto tag)
        {
            using T = decltype(tag);
            const DeviceMcgs<T> & A = [&]() -> const DeviceMcgs<T> &
            {
                if constexpr (std::is_same_v<T, double>) return A64_;
                else return A32_;
            }();
            T * x = f64 ? reinterpret_cast<T *>(x64_.get()) : reinterpret_cast<T *>(x32_.get());
            T * r = f64 ? reinterpret_cast<T *>(r64_.get()) : reinterpret_cast<T *>(r32_.get());
            T * z = f64 ? reinterpret_cast<T *>(z64_.get()) : reinterpret_cast<T *>(z32_.get());
            T * p = f64 ? reinterpret_cast<T *>(p64_.get()) : reinterpret_cast<T *>(p32_.get());
            T * q = f64 ? reinterpret_cast<T *>(q64_.get()) : reinterpret_cast<T *>(q32_.get());
            const McgsMatrix<T> Av = A.view(n_);
            if (which == "direction")
            {
                // z read once, p and q read and written
                bytes = matrixBytes + 5.0 * nch * n * ts;
                launch = [=, this]()
                {
                    dispatchLC(lanes_, nch, [&]<int L, int C>()
                               { pcgDirectionKernel<T, L, C><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, z, p, q); });
                };
            }
            else if (which == "sweep")
            {
                // one full-row GS pass over every colour (numColors launches)
                bytes = matrixBytes + 3.0 * nch * n * ts;
                launch = [=, this]()
                {
                    for (int c = 0; c < numColors_; c++)
                    {
                        const int cnt = colorBlockStart_[c + 1] - colorBlockStart_[c];
                        if (cnt == 0) continue;
                        dispatchLC(lanes_, nch, [&]<int L, int C>()
                                   { mcgsColorKernel<T, L, C><<<cnt, kBlock, 0, s>>>(Av, bm_, colorBlockStart_[c], st, r, z, 0); });
                    }
                };
            }
            else if (which == "update")
            {
                bytes = (6.0 * nch * n + n) * ts;
                launch = [=, this]()
                {
                    if (nch == 1)
                        pcgUpdateKernel<T, 1><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, p, q, x, r, z, color0End_);
                    else
                        pcgUpdateKernel<T, 2><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, p, q, x, r, z, color0End_);
                };
            }
            else
            {
                throw std::invalid_argument("unknown kernel '" + which + "' (direction, sweep, update)");
            }
        };
        if (f64)
        {
            run(double{});
        }
        else
        {
            run(float{});
        }
        // A state in which every kernel does its full work: all channels
        // active, not the first iteration, zero step lengths.
        fillInt(K, 1, st.active, s);
        fillInt(4, 0, st.flags, s);
        fill(F64, 3L * K, 0.0, st.rho, s); // rho, alpha, beta
        fill(F64, 1L * K, 1.0, st.ref, s);
        launch();
        NSFEM_HIP(hipGetLastError());
        hipEvent_t e0, e1;
        NSFEM_HIP(hipEventCreate(&e0));
        NSFEM_HIP(hipEventCreate(&e1));
        NSFEM_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < reps; i++)
        {
            launch();
        }
        NSFEM_HIP(hipEventRecord(e1, s));
        NSFEM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        NSFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        NSFEM_HIP(hipEventDestroy(e0));
        NSFEM_HIP(hipEventDestroy(e1));
        if (bytesPerLaunch)
        {
            *bytesPerLaunch = bytes;
        }
        return static_cast<double>(ms) / reps;
    }

    template void McgsSolver::runGs<double>(hipStream_t);
    template void McgsSolver::runGs<float>(hipStream_t);
} // namespace nsfem
