#include "solver.hpp"

#include "host/fsai.hpp"

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
            // wide channel tiles: one lane per channel, the lane count is the
            // channel count
            switch (nch)
            {
            case 4: f.template operator()<4, 4>(); return;
            case 8: f.template operator()<8, 8>(); return;
            case 16: f.template operator()<16, 16>(); return;
            case 32: f.template operator()<32, 32>(); return;
            default: break;
            }
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
                    throw std::invalid_argument("solver channels must be 1, 2, 4, 8, 16 or 32");
                }
            };
            switch (lanes)
            {
            case 1: byC.template operator()<1>(); break;
            case 2: byC.template operator()<2>(); break;
            case 4: byC.template operator()<4>(); break;
            case 8: byC.template operator()<8>(); break;
            case 16: byC.template operator()<16>(); break;
            default: throw std::invalid_argument("solver lanes must be 1, 2, 4, 8 or 16");
            }
        }

        // f<C>() for the channel count
        template <typename F>
        void dispatchC(int nch, F && f)
        {
            switch (nch)
            {
            case 1: f.template operator()<1>(); break;
            case 2: f.template operator()<2>(); break;
            case 4: f.template operator()<4>(); break;
            case 8: f.template operator()<8>(); break;
            case 16: f.template operator()<16>(); break;
            case 32: f.template operator()<32>(); break;
            default: throw std::invalid_argument("solver channels must be 1, 2, 4, 8, 16 or 32");
            }
        }

        bool validChannelCount(int nch) { return nch == 1 || nch == 2 || nch == 4 || nch == 8 || nch == 16 || nch == 32; }

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
        else if (name == "bpcg")
        {
            o.method = SolverMethod::BlockPcg;
        }
        else if (name == "bpcg_mixed")
        {
            o.method = SolverMethod::BlockPcg;
            o.precision = SolverPrecision::Mixed;
        }
        else if (name == "amg_pcg")
        {
            o.method = SolverMethod::AmgPcg;
        }
        else if (name == "amg_pcg_mixed")
        {
            o.method = SolverMethod::AmgPcg;
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
        scal.resize(static_cast<size_t>(8) * K);
        ints.resize(static_cast<size_t>(3) * K + 4);
        partials.resize(static_cast<size_t>(std::max(1, numBlocks)) * 2 * numCh);
        partials2.resize(static_cast<size_t>(std::max(1, numBlocks)) * numCh);
        partials3.resize(static_cast<size_t>(std::max(1, numBlocks)) * numCh);
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
        s.activeNext = ints.get() + 2 * K + 4;
        s.rhoNext = d + 7 * K;
        s.partials = partials.get();
        s.partials2 = partials2.get();
        s.partials3 = partials3.get();
        s.tol2 = tol2;
        return s;
    }

    McgsSolver::McgsSolver(const std::vector<const CsrD *> & systems, const SolverOptions & opt) : opt_(opt)
    {
        if (systems.empty())
        {
            throw std::invalid_argument("Solver needs at least one matrix");
        }
        if (!validChannelCount(opt_.numCh))
        {
            throw std::invalid_argument("solver channels must be 1, 2, 4, 8, 16 or 32");
        }
        if (opt_.mseCheckInterval < 1)
        {
            throw std::invalid_argument("MSE check interval should be at least 1");
        }
        if (opt_.method == SolverMethod::GaussSeidel && opt_.precision == SolverPrecision::Mixed)
        {
            throw std::invalid_argument("Mixed precision needs the pcg method");
        }
        if (opt_.method == SolverMethod::BlockPcg)
        {
            if (opt_.numCh < 4)
            {
                throw std::invalid_argument("block PCG needs a tile of 4, 8, 16 or 32 right-hand sides");
            }
            if (systems.size() != 1)
            {
                throw std::invalid_argument("block PCG solves one system");
            }
            if (opt_.precision == SolverPrecision::F32)
            {
                throw std::invalid_argument("block PCG keeps FP64 Krylov vectors: use an FP64 matrix (bpcg or bpcg_mixed)");
            }
        }
        if (opt_.method == SolverMethod::AmgPcg)
        {
            if (systems.size() != 1 || opt_.numCh > 16)
            {
                throw std::invalid_argument("amg_pcg solves one system with 1, 2, 4, 8 or 16 right-hand sides");
            }
            if (opt_.precision == SolverPrecision::F32)
            {
                throw std::invalid_argument("amg_pcg keeps FP64 Krylov vectors: use amg_pcg or amg_pcg_mixed");
            }
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
        meanRow_ = n_ > 0 ? nnzOff / n_ : 0.0;
        if (opt_.lanes != 0 && opt_.lanes != 1 && opt_.lanes != 2 && opt_.lanes != 4 && opt_.lanes != 8 && opt_.lanes != 16)
        {
            throw std::invalid_argument("solver lanes must be 1, 2, 4, 8 or 16");
        }

        // Colour-major global order: colour c, then system s, then inside the
        // (colour, system) chunk the rows by locality rank (rows of one colour
        // are independent, so their order changes no result beyond rounding;
        // neighbouring rows then gather neighbouring entries of the other
        // colours).  Each chunk is one range of rows; workgroups never
        // straddle chunks.  The exported colouring keeps the reference's order.
        std::vector<int> perm(n_);
        std::vector<std::vector<int>> gpos(numSystems_);
        std::vector<std::vector<int>> key(numSystems_);
        for (int s = 0; s < numSystems_; s++)
        {
            gpos[s].resize(layouts[s].n);
            if (opt_.localityOrder)
            {
                const std::vector<int> rank = localityRank(systems[s]->rowStart, systems[s]->column, layouts[s].n);
                key[s].resize(layouts[s].n);
                for (int l = 0; l < layouts[s].n; l++)
                {
                    key[s][l] = rank[layouts[s].coloring[l]];
                }
            }
        }
        std::vector<int> order;
        chunks_.clear();
        int g = 0;
        for (int c = 0; c < numColors_; c++)
        {
            for (int s = 0; s < numSystems_; s++)
            {
                const auto & L = layouts[s];
                if (c >= L.numColors())
                {
                    continue;
                }
                const int l0 = L.partitionStart[c], l1 = L.partitionStart[c + 1];
                const int g0 = g;
                order.resize(l1 - l0);
                for (int l = l0; l < l1; l++)
                {
                    order[l - l0] = l;
                }
                if (opt_.localityOrder)
                {
                    const std::vector<int> & k = key[s];
                    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return k[a] < k[b]; });
                }
                if (numSystems_ == 1)
                {
                    // eight stretches of the colour: by locality rank
                    // (rank < j n / 8), else by count
                    const long m = static_cast<long>(order.size());
                    for (int j = 0; j < 8; j++)
                    {
                        long at = m * j / 8;
                        if (opt_.localityOrder)
                        {
                            const std::vector<int> & k = key[s];
                            const long bound = static_cast<long>(L.n) * j / 8;
                            at = std::partition_point(order.begin(), order.end(), [&](int a) { return k[a] < bound; }) -
                                 order.begin();
                        }
                        colorRegionStart_.push_back(g0 + static_cast<int>(at));
                    }
                    colorRegionStart_.push_back(g0 + static_cast<int>(m));
                }
                for (int l : order)
                {
                    gpos[s][l] = g;
                    perm[g] = sysOffset_[s] + L.coloring[l];
                    g++;
                }
                if (g > g0)
                {
                    chunks_.push_back(RowChunk{c, s, g0, g});
                }
            }
            if (c == 0)
            {
                color0End_ = g;
            }
            colorRowStart_.push_back(g); // end of colour c
        }
        colorRowStart_.insert(colorRowStart_.begin(), 0);

        // Global stripped matrix in colour-major order.
        std::vector<int> rowStart(n_ + 1, 0), rowSplit(n_), column;
        std::vector<double> values, diag(n_), invDiag(n_);
        column.reserve(static_cast<size_t>(nnzOff));
        values.reserve(static_cast<size_t>(nnzOff));
        {
            std::vector<std::pair<int, int>> owner(n_); // global row -> (system, local permuted row)
            std::vector<std::pair<int, double>> e;     // row-part sort buffer
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
                if (opt_.localityOrder)
                {
                    // ascending columns inside the lower-colour and the
                    // upper part of the row (the split stays where it is)
                    auto sortPart = [&](int j0, int j1)
                    {
                        e.clear();
                        for (int j = j0; j < j1; j++)
                        {
                            e.emplace_back(column[j], values[j]);
                        }
                        std::sort(e.begin(), e.end(), [](const auto & a, const auto & b) { return a.first < b.first; });
                        for (int j = j0; j < j1; j++)
                        {
                            column[j] = e[j - j0].first;
                            values[j] = e[j - j0].second;
                        }
                    };
                    sortPart(rowStart[r], rowSplit[r]);
                    sortPart(rowSplit[r], static_cast<int>(column.size()));
                }
            }
            rowStart[n_] = static_cast<int>(column.size());
        }

        // Per-colour sizes for algorithmic byte counts: rows, off-diagonal
        // entries of the lower-colour part and of the full row, and the
        // number of distinct columns each part gathers (x read once).
        colorStats_.assign(numColors_, ColorStats{});
        {
            std::vector<int> colorOfRow(n_, 0), seenLower(n_, -1), seenFull(n_, -1);
            for (const RowChunk & ch : chunks_)
            {
                for (int r = ch.begin; r < ch.end; r++)
                {
                    colorOfRow[r] = ch.color;
                }
            }
            for (int r = 0; r < n_; r++)
            {
                const int c = colorOfRow[r];
                ColorStats & cs = colorStats_[c];
                cs.rows += 1;
                cs.lowerNnz += rowSplit[r] - rowStart[r];
                cs.fullNnz += rowStart[r + 1] - rowStart[r];
                for (int j = rowStart[r]; j < rowStart[r + 1]; j++)
                {
                    const int col = column[j];
                    if (j < rowSplit[r] && seenLower[col] != c)
                    {
                        seenLower[col] = c;
                        cs.lowerCols += 1;
                    }
                    if (seenFull[col] != c)
                    {
                        seenFull[col] = c;
                        cs.fullCols += 1;
                    }
                }
            }
        }

        if (opt_.precision != SolverPrecision::F32)
        {
            uploadMatrix(A64_, rowStart, rowSplit, column, values, diag, invDiag);
        }
        if (opt_.precision != SolverPrecision::F64)
        {
            uploadMatrix(A32_, rowStart, rowSplit, column, values, diag, invDiag);
        }

        perm_.upload(perm);
        {
            std::vector<int> iperm(n_);
            for (int i = 0; i < n_; i++)
            {
                iperm[perm[i]] = i;
            }
            iperm_.upload(iperm);
        }
        systemRows_.upload(sysRows_);
        if (opt_.method == SolverMethod::AmgPcg)
        {
            CsrD s0;
            s0.rows = s0.cols = n_;
            s0.rowStart = rowStart;
            s0.column = column;
            s0.values = values;
            const bool mixed = opt_.precision == SolverPrecision::Mixed;
            amg_ = std::make_unique<AmgPcg>(s0, rowSplit, diag, colorRowStart_, mixed, opt_.numCh, A64_.view(n_),
                                            mixed ? A32_.view(n_) : McgsMatrix<float>{}, stream_);
        }

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
            for (auto * a : {&b64_, &x64_, &r64_, &z64_, &p64_, &q64_})
            {
                a->resize(nv);
                a->zero();
            }
            for (auto * a : {&b32_, &x32_, &r32_, &z32_, &p32_, &q32_})
            {
                a->resize(nv);
                a->zero();
            }
        }
        means_.resize(static_cast<size_t>(numSystems_) * opt_.numCh);
        reshape(opt_.lanes, opt_.rowsPerBlock);
        if (opt_.sweepLanes != 0 || opt_.sweepRowsPerBlock != 0)
        {
            reshapeSweeps(opt_.sweepLanes != 0 ? opt_.sweepLanes : lanes_,
                          opt_.sweepRowsPerBlock != 0 ? opt_.sweepRowsPerBlock : (opt_.sweepLanes != 0 ? 0 : rowsPerBlock_));
        }

        const int K = numSystems_ * opt_.numCh;
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&hostFlags_), 4 * sizeof(int), hipHostMallocDefault));
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&pollRing_), 8 * sizeof(int), hipHostMallocDefault));
        for (hipEvent_t & e : pollEvents_)
        {
            NSFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&hostScal_), 7 * K * sizeof(double), hipHostMallocDefault));
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&hostInts_), (2 * K + 4) * sizeof(int), hipHostMallocDefault));
        lastMse_.assign(K, -1.0);
        lastRel_.assign(K, -1.0);
        lastIters_.assign(numSystems_, -1);
        // statistics slots of captured solves (allocated here: nothing may be
        // allocated while a stream is being captured)
        statSlots_.resize(kStatSlots * (8 * static_cast<size_t>(K) * sizeof(double) + (3 * static_cast<size_t>(K) + 4) * sizeof(int)));
        if (opt_.method == SolverMethod::BlockPcg)
        {
            // at most 2048 workgroups (eight per CU), each a whole number of
            // row slabs; their Gram partials are summed by a reduction kernel
            const int nch = opt_.numCh;
            const int slab = kBlock / nch;
            int rows = ceilDiv(n_, 2048);
            rows = std::max(slab, ceilDiv(rows, slab) * slab);
            bcgRows_ = rows;
            bcgBlocks_ = std::max(1, ceilDiv(n_, rows));
            bcgGram_.resize(static_cast<size_t>(bcgBlocks_) * nch * nch);
            bcgNorm_.resize(static_cast<size_t>(bcgBlocks_) * 2 * nch);
            bcgSmall_.resize((2 + kBcgSlices) * static_cast<size_t>(nch) * nch + (2 + 2 * kBcgSlices) * nch + 2);
            bcgInts_.resize(5);
        }
        syncLibraryStream();
    }

    McgsSolver::~McgsSolver()
    {
        (void)hipStreamSynchronize(stream_);
        for (ChunkGraph * g : {&graph64_, &graph32_, &graphMixed_})
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
        if (pollRing_)
        {
            (void)hipHostFree(pollRing_);
        }
        for (hipEvent_t e : pollEvents_)
        {
            if (e)
            {
                (void)hipEventDestroy(e);
            }
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

    std::pair<int, int> McgsSolver::resolveShape(int lanes, int rowsPerBlock, bool sweeps) const
    {
        if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 && !(opt_.numCh > 2 && lanes == opt_.numCh))
        {
            throw std::invalid_argument("solver lanes must be 1, 2, 4, 8 or 16");
        }
        if (rowsPerBlock < 0)
        {
            throw std::invalid_argument("rowsPerBlock must be >= 0");
        }
        // wide channel tiles run one lane per channel: the lane count is fixed
        const int L = opt_.numCh > 2 ? opt_.numCh : (lanes > 0 ? lanes : std::min(16, lanesForMeanRow(meanRow_)));
        const int slab = kBlock / L;
        // the sweeps write no partial sums: any row count goes
        int rpb = rowsPerBlock > 0 ? rowsPerBlock : (sweeps ? 4 * slab : std::max(4 * slab, minRowsPerBlock()));
        rpb = std::max(slab, rpb / slab * slab);
        return {L, rpb};
    }

    // Workgroups of at most rpb rows, colour-major, never straddling a
    // (colour, system) chunk; colorStart[c] = first workgroup of colour c.
    void McgsSolver::buildBlocks(int rpb, std::vector<int4> & blocks, std::vector<int> & blockSystem, std::vector<int> & colorStart) const
    {
        blocks.clear();
        blockSystem.clear();
        colorStart.assign(numColors_ + 1, 0);
        size_t next = 0;
        for (int c = 0; c < numColors_; c++)
        {
            colorStart[c] = static_cast<int>(blocks.size());
            for (; next < chunks_.size() && chunks_[next].color == c; next++)
            {
                const RowChunk & ch = chunks_[next];
                for (int b0 = ch.begin; b0 < ch.end; b0 += rpb)
                {
                    blocks.push_back(int4{b0, std::min(ch.end, b0 + rpb), ch.system, -1});
                    blockSystem.push_back(ch.system);
                }
            }
        }
        colorStart[numColors_] = static_cast<int>(blocks.size());
    }

    void McgsSolver::reshapeSweeps(int lanes, int rowsPerBlock)
    {
        const auto [L, rpb] = resolveShape(lanes, rowsPerBlock, true);
        NSFEM_HIP(hipStreamSynchronize(stream_));
        sweepLanes_ = L;
        sweepRows_ = rpb;
        std::vector<int4> blocks;
        std::vector<int> blockSystem;
        buildBlocks(rpb, blocks, blockSystem, colorSweepBlockStart_);
        sweepBlocks_.upload(blocks.empty() ? std::vector<int4>{int4{0, 0, 0, 0}} : blocks);
        sweepBm_ = bm_;
        sweepBm_.blocks = sweepBlocks_.get();
        sweepBm_.numBlocks = static_cast<int>(blocks.size());
        sweepBm_.rowsPerBlock = rpb;
        destroyGraphs();
        syncLibraryStream();
    }

    void McgsSolver::destroyGraphs()
    {
        for (ChunkGraph * g : {&graph64_, &graph32_, &graphMixed_})
        {
            if (g->exec)
            {
                NSFEM_HIP(hipGraphExecDestroy(g->exec));
                g->exec = nullptr;
            }
        }
    }

    void McgsSolver::reshape(int lanes, int rowsPerBlock)
    {
        const auto [L, rpb] = resolveShape(lanes, rowsPerBlock, false);
        NSFEM_HIP(hipStreamSynchronize(stream_));
        lanes_ = L;
        rowsPerBlock_ = rpb;

        std::vector<int4> blocks;
        std::vector<int> blockSystem;
        if (flat())
        {
            // one system: the row-wise kernels cut [0, n) evenly, whatever
            // the colours (a block may straddle two)
            for (int b0 = 0; b0 < n_; b0 += rpb)
            {
                blocks.push_back(int4{b0, std::min(n_, b0 + rpb), 0, -1});
                blockSystem.push_back(0);
            }
            color0Blocks_ = (color0End_ + rpb - 1) / rpb;
        }
        else
        {
            std::vector<int> colorStart;
            buildBlocks(rpb, blocks, blockSystem, colorStart);
            color0Blocks_ = colorStart.size() > 1 ? colorStart[1] : 0;
        }
        numBlocks_ = static_cast<int>(blocks.size());

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
                sysBlocks[fillPos[blockSystem[b]]] = b;
                blocks[b].w = fillPos[blockSystem[b]]++; // partial slot
            }
        }
        blocks_.upload(blocks.empty() ? std::vector<int4>{int4{0, 0, 0, 0}} : blocks);
        sysBlockStart_.upload(sysBlockStart);
        sysBlocks_.upload(sysBlocks.empty() ? std::vector<int>{0} : sysBlocks);
        bm_ = BlockMap{blocks_.get(), sysBlockStart_.get(), sysBlocks_.get(), systemRows_.get(), numSystems_, numBlocks_,
                       flat() ? n_ : 0, rpb};

        // reduction scratch sized by the block count; captured graphs hold
        // the old launch shapes and buffers
        rowsDir_ = rowsUpd_ = rowsRho_ = 0;
        lanesDir_ = 0;
        setupKindMaps();
        // the sweeps take the same shape until reshapeSweeps
        reshapeSweeps(lanes_, rowsPerBlock_);
    }

    int McgsSolver::minRowsPerBlock() const
    {
        // Wide channel tiles do NCH times the work per row: the default and
        // autotuned shapes keep at most about 16k workgroups, so that the
        // one-workgroup finalizes read a bounded number of partial sums.
        if (opt_.numCh <= 2)
        {
            return 1;
        }
        const int slab = kBlock / opt_.numCh;
        const int need = (n_ + 16383) / 16384;
        return (need + slab - 1) / slab * slab;
    }

    void McgsSolver::setFusedReductions(int mode)
    {
        if (mode < -1 || mode > 1)
        {
            throw std::invalid_argument("fusedReductions must be -1 (auto), 0 or 1");
        }
        NSFEM_HIP(hipStreamSynchronize(stream_));
        fusedMode_ = mode;
        setupKindMaps();
        destroyGraphs();
    }

    void McgsSolver::setKindRows(int direction, int update, int rho)
    {
        if (direction < 0 || update < 0 || rho < 0)
        {
            throw std::invalid_argument("rows per workgroup must be >= 0");
        }
        NSFEM_HIP(hipStreamSynchronize(stream_));
        // the direction and rho kernels work in row slabs of their lanes
        const int slab = kBlock / lanes_;
        auto slabs = [&](int r) { return r == 0 ? 0 : std::max(slab, r / slab * slab); };
        rowsDir_ = slabs(direction);
        rowsUpd_ = update;
        rowsRho_ = slabs(rho);
        setupKindMaps();
        destroyGraphs();
    }

    void McgsSolver::setPreconditionerRows(int rows)
    {
        if (rows < 0 || rows % kBlock != 0)
        {
            throw std::invalid_argument("rowsPreconditioner must be 0 or a positive multiple of 256");
        }
        NSFEM_HIP(hipStreamSynchronize(stream_));
        rowsPrecond_ = rows;
        setupKindMaps();
        destroyGraphs();
    }

    void McgsSolver::setDirectionLanes(int lanes)
    {
        if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16)
        {
            throw std::invalid_argument("direction lanes must be 0, 1, 2, 4, 8 or 16");
        }
        NSFEM_HIP(hipStreamSynchronize(stream_));
        lanesDir_ = lanes;
        destroyGraphs();
    }

    void McgsSolver::setRegionedDirection(int on)
    {
        if (on != 0 && on != 1)
        {
            throw std::invalid_argument("regionDirection must be 0 or 1");
        }
        NSFEM_HIP(hipStreamSynchronize(stream_));
        regionDir_ = on;
        setupKindMaps();
        destroyGraphs();
    }

    int McgsSolver::kindRows(int kind) const
    {
        if (!(fused() && flat()))
        {
            return rowsPerBlock_;
        }
        switch (kind)
        {
        case 0: return bmDir_.rowsPerBlock;
        case 1: return bmUpd_.rowsPerBlock;
        case 2: return bmRho_.rowsPerBlock;
        default: throw std::invalid_argument("kind must be 0 (direction), 1 (update) or 2 (rho)");
        }
    }

    void McgsSolver::setupKindMaps()
    {
        int maxBlocks = numBlocks_;
        if (fused() && flat())
        {
            // Unset rows per workgroup: every workgroup of the next kernel
            // re-sums all partials of this one, so the default keeps the
            // direction and update kernels at about 1,024 workgroups and the
            // rho kernel at about 4,096 (in whole multiples of the main rows)
            auto defaultRows = [&](int maxBlocksWanted)
            {
                int rows = rowsPerBlock_;
                while (static_cast<long>(rows) * maxBlocksWanted < n_)
                {
                    rows *= 2;
                }
                return rows;
            };
            auto flatMap = [&](int rows, int & nb, int maxBlocksWanted)
            {
                BlockMap m = bm_;
                m.rowsPerBlock = rows > 0 ? rows : defaultRows(maxBlocksWanted);
                nb = (n_ + m.rowsPerBlock - 1) / m.rowsPerBlock;
                m.numBlocks = nb;
                maxBlocks = std::max(maxBlocks, nb);
                return m;
            };
            // stretch k of every colour, then stretch k + 1: the XCD that
            // runs a contiguous range of blocks (logicalBlock) gathers from
            // one stretch of the locality order
            auto regionMap = [&](BlockMap & m, int & nb, DeviceArray<int2> & table)
            {
                const int rpb = m.rowsPerBlock;
                std::vector<int2> t;
                for (int k = 0; k < 8; k++)
                {
                    for (int c = 0; c < numColors_; c++)
                    {
                        const int r0 = colorRegionStart_[9 * c + k], r1 = colorRegionStart_[9 * c + k + 1];
                        for (int b0 = r0; b0 < r1; b0 += rpb)
                        {
                            t.push_back(int2{b0, std::min(r1, b0 + rpb)});
                        }
                    }
                }
                table.upload(t);
                nb = static_cast<int>(t.size());
                m.numBlocks = nb;
                m.rowTable = table.get();
                maxBlocks = std::max(maxBlocks, nb);
            };
            bmDir_ = flatMap(rowsDir_, nbDir_, 1024);
            bmUpd_ = flatMap(rowsUpd_, nbUpd_, 1024);
            bmRho_ = flatMap(rowsRho_, nbRho_, 4096);
            const bool ownPrecondMap = rowsPrecond_ > 0 && rowsPrecond_ != bmDir_.rowsPerBlock;
            if (ownPrecondMap)
            {
                bmPre_ = flatMap(rowsPrecond_, nbPre_, 1024);
            }
            if (regionDir_ && n_ > 0)
            {
                regionMap(bmDir_, nbDir_, dirTable_);
                if (ownPrecondMap)
                {
                    regionMap(bmPre_, nbPre_, preTable_);
                }
            }
            if (!ownPrecondMap)
            {
                bmPre_ = bmDir_;
                nbPre_ = nbDir_;
            }
            color0BlocksRho_ = (color0End_ + bmRho_.rowsPerBlock - 1) / bmRho_.rowsPerBlock;
        }
        else
        {
            bmDir_ = bmUpd_ = bmRho_ = bmPre_ = bm_;
            nbDir_ = nbUpd_ = nbRho_ = nbPre_ = numBlocks_;
            color0BlocksRho_ = color0Blocks_;
        }
        // reduction scratch sized by the largest block count
        inner_.init(maxBlocks, numSystems_, opt_.numCh);
        if (opt_.precision == SolverPrecision::Mixed)
        {
            outer_.init(maxBlocks, numSystems_, opt_.numCh);
        }
    }

    double McgsSolver::iterationTime(int reps)
    {
        double total = 0;
        for (const char * k : {"direction", "sweep", "rho", "update", "finalize"})
        {
            double bytes = 0, launches = 0;
            const double ms = timeKernel(k, reps, &bytes, &launches);
            total += ms * launches;
        }
        return total;
    }

    std::pair<int, int> McgsSolver::autotune(int reps)
    {
        if (opt_.method != SolverMethod::Pcg)
        {
            throw std::invalid_argument("autotuning is available for pcg solvers");
        }
        // The objective is one PCG iteration replayed from a captured graph,
        // as the solver runs it: kernel durations and the gaps between the
        // dependent launches.
        // (Eager replays when graphs are off, e.g. under a kernel tracer.)
        const char * kind = useGraphs_ ? "iteration_graph" : "iteration";
        auto iterationMs = [&]() { return timeKernel(kind, reps, nullptr, nullptr); };
        const std::vector<int> laneChoices = opt_.numCh > 2 ? std::vector<int>{opt_.numCh} : std::vector<int>{1, 2, 4, 8, 16};
        // 1. one shape for every kernel
        double best = 0;
        int bestLanes = lanes_, bestRows = rowsPerBlock_;
        for (int lanes : laneChoices)
        {
            for (int slabs : {1, 2, 4, 8})
            {
                const int slab = kBlock / lanes;
                // wide tiles: from the fewest rows per workgroup that keep the
                // finalize kernels' partial sums bounded
                const int base = opt_.numCh > 2 ? std::max(slab, minRowsPerBlock()) : slab;
                reshape(lanes, slabs * base);
                const double t = iterationMs();
                if (best == 0 || t < best)
                {
                    best = t;
                    bestLanes = lanes;
                    bestRows = rowsPerBlock_;
                }
            }
        }
        reshape(bestLanes, bestRows);
        // 2. the colour sweeps (no partial sums) on a shape of their own
        int sweepLanes = bestLanes, sweepRows = bestRows;
        for (int lanes : laneChoices)
        {
            for (int slabs : {1, 2, 4})
            {
                reshapeSweeps(lanes, slabs * (kBlock / lanes));
                const double t = iterationMs();
                if (t < best)
                {
                    best = t;
                    sweepLanes = sweepLanes_;
                    sweepRows = sweepRows_;
                }
            }
        }
        reshapeSweeps(sweepLanes, sweepRows);
        // 3. fused reductions on one system: the direction, update and rho
        // kernels on row counts of their own (every block of the next kernel
        // re-sums their partials, so fewer, larger blocks can pay)
        if (fused() && flat())
        {
            int rows[3] = {0, 0, 0};
            for (int kind = 0; kind < 3; kind++)
            {
                int keep = rows[kind];
                for (int mult : {2, 4, 8, 16})
                {
                    int trial[3] = {rows[0], rows[1], rows[2]};
                    trial[kind] = mult * rowsPerBlock_;
                    setKindRows(trial[0], trial[1], trial[2]);
                    const double t = iterationMs();
                    if (t < best)
                    {
                        best = t;
                        keep = trial[kind];
                    }
                }
                rows[kind] = keep;
                setKindRows(rows[0], rows[1], rows[2]);
            }
        }
        // 4. the direction kernel's own lanes per row (its rows are whole
        // slabs of any lane count up to 16 once they are multiples of 256)
        if (opt_.numCh <= 2)
        {
            int keep = 0;
            for (int lanes : {1, 2, 4, 8, 16})
            {
                if (lanes == lanes_)
                {
                    continue;
                }
                setDirectionLanes(lanes);
                const double t = iterationMs();
                if (t < best)
                {
                    best = t;
                    keep = lanes;
                }
            }
            setDirectionLanes(keep);
        }
        // 5. fused reductions against the finalize kernels (on a large system
        // the fused mode needs large direction workgroups to keep its partial
        // sums few; the finalize kernels let every kernel keep the main shape)
        if (fused() && fusedMode_ == -1)
        {
            const int keepDir = lanesDir_;
            const int keepRows[3] = {rowsDir_, rowsUpd_, rowsRho_};
            fusedMode_ = 0;
            setupKindMaps();
            destroyGraphs();
            double unfused = 0;
            int unfusedLanes = 0;
            for (int lanes : {0, 1, 2, 4, 8, 16})
            {
                setDirectionLanes(lanes);
                const double t = iterationMs();
                if (unfused == 0 || t < unfused)
                {
                    unfused = t;
                    unfusedLanes = lanes;
                }
            }
            if (unfused < best)
            {
                setDirectionLanes(unfusedLanes);
            }
            else
            {
                fusedMode_ = -1;
                lanesDir_ = keepDir;
                setKindRows(keepRows[0], keepRows[1], keepRows[2]);
            }
        }
        return {bestLanes, bestRows};
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
        permuteRows(dt, ioRhs_.get(), b, false, s);
        permuteRows(dt, ioSol_.get(), x, false, s);

        const McgsMatrix<T> Av = A.view(n_);
        int iter = 0;
        for (; iter < opt_.maxIters; iter++)
        {
            for (int c = 0; c < numColors_; c++)
            {
                const int cnt = colorSweepBlockStart_[c + 1] - colorSweepBlockStart_[c];
                if (cnt == 0)
                {
                    continue;
                }
                dispatchLC(sweepLanes_, nch, [&]<int L, int C>()
                           {
                               mcgsColorKernel<T, L, C><<<cnt, kBlock, 0, s>>>(Av, sweepBm_, colorSweepBlockStart_[c], colorRowStart_[c],
                                                                              colorRowStart_[c + 1], st, b, x, 0);
                           });
            }
            if (iter % opt_.mseCheckInterval == 0)
            {
                dispatchLC(lanes_, nch, [&]<int L, int C>()
                           {
                               gsResidualKernel<T, L, C><<<numBlocks_, kBlock, 0, s>>>(Av, bm_, st, b, x);
                               gsResidualFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bm_, st, iter, inner_.totals.get());
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
        permuteRows(dt, x, ioSol_.get(), true, s);
        NSFEM_HIP(hipStreamSynchronize(s));
    }

    template <>
    PcgOperands<double, double> McgsSolver::operands<double, double>()
    {
        return {A64_.view(n_), A64_.view(n_), b64_.get(), x64_.get(), r64_.get(), r64_.get(), z64_.get(), p64_.get(), q64_.get(),
                fsaiView<double>(), reinterpret_cast<double *>(fsaiT_.get())};
    }

    template <>
    PcgOperands<float, float> McgsSolver::operands<float, float>()
    {
        return {A32_.view(n_), A32_.view(n_), b32_.get(), x32_.get(), r32_.get(), r32_.get(), z32_.get(), p32_.get(), q32_.get(),
                fsaiView<float>(), reinterpret_cast<float *>(fsaiT_.get())};
    }

    template <>
    PcgOperands<double, float> McgsSolver::operands<double, float>()
    {
        return {A64_.view(n_), A32_.view(n_), b64_.get(), x64_.get(), r64_.get(), r32_.get(), z32_.get(), p64_.get(), q64_.get(),
                fsaiView<float>(), reinterpret_cast<float *>(fsaiT_.get())};
    }

    template <typename T, typename P>
    McgsSolver::ChunkGraph & McgsSolver::chunkGraph()
    {
        if constexpr (std::is_same_v<T, P>)
        {
            return std::is_same_v<T, double> ? graph64_ : graph32_;
        }
        else
        {
            return graphMixed_;
        }
    }

    template <typename T, typename P>
    int & McgsSolver::predicted()
    {
        if constexpr (std::is_same_v<T, P>)
        {
            return std::is_same_v<T, double> ? predicted64_ : predicted32_;
        }
        else
        {
            return predictedMixed_;
        }
    }

    template <typename P>
    void McgsSolver::sgsSweeps(hipStream_t s, const McgsMatrix<P> & M, SolveState st, const P * r, P * z, int firstForward)
    {
        const int nch = opt_.numCh;
        // forward colours firstForward .. C-1 on the lower part (colour 0 was
        // set to r/d; colour 1 may have been swept with the update)
        for (int c = firstForward; c < numColors_; c++)
        {
            const int cnt = colorSweepBlockStart_[c + 1] - colorSweepBlockStart_[c];
            if (cnt == 0)
            {
                continue;
            }
            dispatchLC(sweepLanes_, nch, [&]<int L, int C>()
                       {
                           mcgsColorKernel<P, L, C><<<cnt, kBlock, 0, s>>>(M, sweepBm_, colorSweepBlockStart_[c], colorRowStart_[c],
                                                                          colorRowStart_[c + 1], st, r, z, 1);
                       });
        }
        // backward colours C-2 .. 1 on full rows (colour C-1 is unchanged)
        for (int c = numColors_ - 2; c >= 1; c--)
        {
            const int cnt = colorSweepBlockStart_[c + 1] - colorSweepBlockStart_[c];
            if (cnt == 0)
            {
                continue;
            }
            dispatchLC(sweepLanes_, nch, [&]<int L, int C>()
                       {
                           mcgsColorKernel<P, L, C><<<cnt, kBlock, 0, s>>>(M, sweepBm_, colorSweepBlockStart_[c], colorRowStart_[c],
                                                                          colorRowStart_[c + 1], st, r, z, 0);
                       });
        }
    }

    template <typename T, typename P>
    void McgsSolver::launchRho(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, double * totals, bool withUpdate)
    {
        // fused reductions: the next direction kernel finishes rho
        const bool fin = !fused();
        dispatchLC(lanes_, opt_.numCh, [&]<int L, int C>()
                   {
                       pcgRhoKernel<P, T, L, C><<<nbRho_, kBlock, 0, s>>>(v.M, bmRho_, st, v.rP, v.r, v.z, color0End_,
                                                                          color0BlocksRho_, v.q, withUpdate && mergedUpdate() ? 1 : 0);
                       if (fin)
                       {
                           pcgRhoFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bm_, st, totals, withUpdate ? 1 : 0);
                       }
                   });
    }

    template <typename T, typename P>
    void McgsSolver::launchUpdateSweep(hipStream_t s, const PcgOperands<T, P> & v, SolveState st)
    {
        // colour 1 = rows [colorRowStart_[1], colorRowStart_[2]); the update
        // map's workgroups that hold them are interleaved among the others
        const int rpb = bmUpd_.rowsPerBlock;
        const int c1Begin = colorRowStart_[1], c1End = colorRowStart_[2];
        const int spBegin = c1Begin / rpb;
        const int spCount = c1End > c1Begin ? (c1End + rpb - 1) / rpb - spBegin : 0;
        dispatchLC(sweepLanes_, opt_.numCh, [&]<int L, int C>()
                   {
                       if constexpr (!kWide<C>)
                       {
                           pcgUpdateSweepKernel<T, P, L, C><<<nbUpd_, kBlock, 0, s>>>(v.M, bmUpd_, st, v.p, v.q, v.x, v.r, v.rP, v.z,
                                                                                     color0End_, c1Begin, c1End, spBegin, spCount,
                                                                                     nbDir_);
                       }
                   });
    }

    void McgsSolver::setFusedUpdateSweep(int on)
    {
        if (on != 0 && on != 1)
        {
            throw std::invalid_argument("fuseUpdateSweep must be 0 or 1");
        }
        NSFEM_HIP(hipStreamSynchronize(stream_));
        mergeMode_ = on;
        destroyGraphs();
    }

    template <typename T, typename P>
    void McgsSolver::pcgIteration(hipStream_t s, const PcgOperands<T, P> & v, SolveState st)
    {
        const int nch = opt_.numCh;
        double * totals = inner_.totals.get();
        const int fz = fused() ? 1 : 0;
        dispatchLC(directionLanes(), nch, [&]<int L, int C>()
                   {
                       pcgDirectionKernel<T, P, L, C><<<nbDir_, kBlock, 0, s>>>(v.A, bmDir_, st, v.z, v.p, v.q, fz, rhoParts(), nbUpd_);
                       if (!fz)
                       {
                           pcgDirectionFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bm_, st, totals);
                       }
                   });
        if (mergedUpdate())
        {
            launchUpdateSweep<T, P>(s, v, st);
        }
        else
        {
            dispatchC(nch, [&]<int C>()
                      {
                          pcgUpdateKernel<T, P, C><<<nbUpd_, kBlock, 0, s>>>(n_, bmUpd_, st, v.p, v.q, v.x, v.r, v.rP, v.z, v.M.invDiag,
                                                                            zColour0End(), fz, nbDir_);
                      });
        }
        // the update's convergence test is finished together with rho
        launchPrecond<T, P>(s, v, st, totals, true);
    }

    template <typename T, typename P>
    void McgsSolver::launchPrecond(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, double * totals, bool withUpdate)
    {
        if (precond_ == 0)
        {
            sgsSweeps<P>(s, v.M, st, v.rP, v.z, withUpdate && mergedUpdate() ? 2 : 1);
            launchRho<T, P>(s, v, st, totals, withUpdate);
            return;
        }
        // on a region-major map (the workgroups of one XCD take the same
        // stretch of every colour and gather from about an eighth of the
        // vector): the direction kernel's, or one with rows of its own
        // (rowsPreconditioner); the next direction kernel reads nbPre_ rho
        // partials
        const bool fin = !fused();
        dispatchLC(directionLanes(), opt_.numCh, [&]<int L, int C>()
                   {
                       fsaiLowerKernel<P, L, C><<<nbPre_, kBlock, 0, s>>>(v.G, bmPre_, st, v.rP, v.t);
                       fsaiUpperKernel<P, T, L, C><<<nbPre_, kBlock, 0, s>>>(v.G, bmPre_, st, v.t, v.z, v.r);
                       if (fin)
                       {
                           pcgRhoFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bm_, st, totals, withUpdate ? 1 : 0);
                       }
                   });
    }

    template <typename P>
    McgsMatrix<P> McgsSolver::fsaiView() const
    {
        // the structure of the preconditioner's matrix, the factor's values
        if constexpr (std::is_same_v<P, double>)
        {
            McgsMatrix<double> G = A64_.view(n_);
            G.values = fsai64_.get();
            G.diag = fsaiDiag64_.get();
            G.invDiag = nullptr;
            return G;
        }
        else
        {
            McgsMatrix<float> G = A32_.view(n_);
            G.values = fsai32_.get();
            G.diag = fsaiDiag32_.get();
            G.invDiag = nullptr;
            return G;
        }
    }

    void McgsSolver::setPreconditioner(int kind)
    {
        if (kind != 0 && kind != 1)
        {
            throw std::invalid_argument("preconditioner must be 0 (symmetric multicolour GS) or 1 (factorized approximate inverse)");
        }
        if (kind == 1)
        {
            const bool pcg = opt_.method == SolverMethod::Pcg;
            if (!pcg || opt_.restart)
            {
                throw std::invalid_argument("the factorized approximate inverse preconditions pcg / pcg_mixed (without "
                                            "mixedRestart)");
            }
            const bool p64 = opt_.precision == SolverPrecision::F64;
            if ((p64 && !fsai64_.get()) || (!p64 && !fsai32_.get()))
            {
                // the colour-major matrix from the device (its FP64 copy where there is one)
                std::vector<int> rowStart, rowSplit, column;
                std::vector<double> values, diag;
                if (A64_.values.get())
                {
                    rowStart = A64_.rowStart.download();
                    rowSplit = A64_.rowSplit.download();
                    column = A64_.column.download();
                    values = A64_.values.download();
                    diag = A64_.diag.download();
                }
                else
                {
                    rowStart = A32_.rowStart.download();
                    rowSplit = A32_.rowSplit.download();
                    column = A32_.column.download();
                    const std::vector<float> v32 = A32_.values.download(), d32 = A32_.diag.download();
                    values.assign(v32.begin(), v32.end());
                    diag.assign(d32.begin(), d32.end());
                }
                std::vector<double> g, gd;
                buildFsai(n_, rowStart, rowSplit, column, values, diag, g, gd);
                if (column.empty())
                {
                    g.assign(1, 0.0); // a diagonal matrix: keep a valid pointer
                }
                if (p64)
                {
                    fsai64_.upload(g);
                    fsaiDiag64_.upload(gd);
                }
                else
                {
                    fsai32_.upload(std::vector<float>(g.begin(), g.end()));
                    fsaiDiag32_.upload(std::vector<float>(gd.begin(), gd.end()));
                }
            }
            const size_t bytes = static_cast<size_t>(n_) * opt_.numCh * (p64 ? 8 : 4);
            if (fsaiT_.size() < bytes)
            {
                fsaiT_.resize(bytes);
            }
        }
        NSFEM_HIP(hipStreamSynchronize(stream_));
        precond_ = kind;
        destroyGraphs();
    }

    template <typename T, typename P>
    int McgsSolver::runPcg(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, bool zeroGuess, const int * gate, int maxIters,
                           bool sync, bool usePrediction)
    {
        const int nch = opt_.numCh;
        double * totals = inner_.totals.get();
        fillInt(4, 0, st.flags, s);
        // r = b - A x is the direction kernel's product: its map and lanes
        dispatchLC(directionLanes(), nch, [&]<int L, int C>()
                   {
                       pcgInitKernel<T, P, L, C><<<nbDir_, kBlock, 0, s>>>(v.A, bmDir_, st, v.b, v.x, v.r, v.rP, v.z, v.M.invDiag,
                                                                          zColour0End(), zeroGuess ? 1 : 0);
                       pcgInitFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bmDir_, st, gate, totals);
                   });
        launchPrecond<T, P>(s, v, st, totals, false);
        NSFEM_HIP(hipGetLastError());
        return iterate<T, P>(s, v, st, maxIters, sync, usePrediction);
    }

    template <typename T, typename P>
    int McgsSolver::iterate(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, int maxIters, bool sync, bool usePrediction)
    {
        DeviceState & ds = inner_;
        // Iterations run in chunks; a full chunk is one launch of a captured
        // graph.  Kernels after convergence exit at once.
        ChunkGraph & g = chunkGraph<T, P>();
        int & predictedIters = predicted<T, P>();
        // Chunks of up to checkChunk iterations, shorter for solves expected
        // to take few iterations (at least 8 chunks per solve, at least 2
        // iterations per chunk), so the host polls keep pace and little runs
        // past convergence.
        int chunk = std::max(1, opt_.checkChunk);
        if (usePrediction)
        {
            while (chunk > 2 && 8 * chunk > std::max(predictedIters, 1))
            {
                chunk /= 2;
            }
        }
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
                pcgIteration<T, P>(s, v, st);
            }
            NSFEM_HIP(hipStreamEndCapture(s, &graph));
            NSFEM_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
            NSFEM_HIP(hipGraphDestroy(graph));
            g.k = chunk;
            g.tol2 = st.tol2;
        }
        // Lagged convergence polls: after each chunk an asynchronous copy of
        // the flags into pinned host memory and an event are queued; before
        // the next chunk is queued the host waits for the previous chunk's
        // copy, which has finished by the time the current chunk runs, so the
        // GPU never drains and at most one chunk runs past convergence (its
        // kernels exit at once).  Each copy idles the GPU for some 18 µs
        // (the copy and the barriers around it), so the chunks up to one
        // before the last solve's count less two go without (at most 64
        // iterations: after a solve that took many more, no-op iterations
        // would cost more than the polls they save).
        const int quiet = usePrediction ? std::min(lastSolveIters_ - 2 - chunk, 64) : 0;
        int launched = 0;
        int slot = 0;
        bool pending[2] = {false, false};
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
                    pcgIteration<T, P>(s, v, st);
                }
            }
            NSFEM_HIP(hipGetLastError());
            launched += k;
            if (!sync || launched <= quiet)
            {
                continue;
            }
            NSFEM_HIP(hipMemcpyAsync(pollRing_ + 4 * slot, st.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
            NSFEM_HIP(hipEventRecord(pollEvents_[slot], s));
            pending[slot] = true;
            const int prev = slot ^ 1;
            if (pending[prev])
            {
                NSFEM_HIP(hipEventSynchronize(pollEvents_[prev]));
                pending[prev] = false;
                if (pollRing_[4 * prev] != 0)
                {
                    break; // converged before the chunk just queued
                }
            }
            slot = prev;
        }
        if (fused())
        {
            // the last update's convergence test (no direction kernel follows)
            dispatchC(opt_.numCh, [&]<int C>() { pcgUpdateCheckKernel<C><<<1, kFinBlock, 0, s>>>(bmUpd_, st, ds.totals.get()); });
        }
        if (usePrediction)
        {
            predictedIters = launched; // only sizes the chunks of the next solve
        }
        return launched;
    }

    // Iterative refinement: FP64 residual, FP32 PCG correction from zero to
    // innerTol, x64 += correction; at most maxOuter steps.
    void McgsSolver::runMixedRestart(hipStream_t s)
    {
        const int nch = opt_.numCh;
        const int K = numSystems_ * nch;
        const long nv = static_cast<long>(n_) * nch;
        permuteRows(F64, ioRhs_.get(), b64_.get(), false, s);
        permuteRows(F64, ioSol_.get(), x64_.get(), false, s);
        SolveState outer = outer_.view(K, opt_.target * opt_.target);
        SolveState inner = inner_.view(K, opt_.innerTol * opt_.innerTol);
        const PcgOperands<double, double> v64 = operands<double, double>();
        const PcgOperands<float, float> v32 = operands<float, float>();
        std::vector<int> totalIters(numSystems_, 0);
        int outerIt = 0;
        for (;; outerIt++)
        {
            fillInt(4, 0, outer.flags, s);
            dispatchLC(lanes_, nch, [&]<int L, int C>()
                       {
                           pcgInitKernel<double, double, L, C><<<numBlocks_, kBlock, 0, s>>>(v64.A, bm_, outer, v64.b, v64.x, v64.r, v64.r,
                                                                                             v64.z, v64.M.invDiag, color0End_, 0);
                           pcgInitFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bm_, outer, nullptr, outer_.totals.get());
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
            runPcg<float, float>(s, v32, inner, true, outer.active, opt_.maxIters, true, false);
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
        permuteRows(F64, x64_.get(), ioSol_.get(), true, s);
        NSFEM_HIP(hipStreamSynchronize(s));
    }

    BcgState McgsSolver::bcgState(SolveState st)
    {
        const int nch = opt_.numCh;
        BcgState b;
        b.gramParts = bcgGram_.get();
        b.normParts = bcgNorm_.get();
        b.rho = bcgSmall_.get();
        b.step = b.rho + nch * nch;
        b.bb = b.step + nch * nch;
        b.rr = b.bb + nch;
        b.best = b.rr + nch;
        b.gramTotal = b.best + 2;
        b.normTotal = b.gramTotal + kBcgSlices * nch * nch;
        b.flags = st.flags;
        b.iters = bcgInts_.get();
        b.tol2 = opt_.target * opt_.target;
        b.rhoThreshold = blockPivot_ > 0 ? blockPivot_ : (opt_.precision == SolverPrecision::Mixed ? 1e-10 : 1e-13);
        b.n = n_;
        b.blocks = bcgBlocks_;
        b.rowsPerBlock = bcgRows_;
        return b;
    }

    namespace
    {
        // f<C>() for the wide channel counts of block PCG
        template <typename F>
        void dispatchWide(int nch, F && f)
        {
            switch (nch)
            {
            case 4: f.template operator()<4>(); break;
            case 8: f.template operator()<8>(); break;
            case 16: f.template operator()<16>(); break;
            case 32: f.template operator()<32>(); break;
            default: throw std::invalid_argument("block PCG tiles have 4, 8, 16 or 32 columns");
            }
        }
    } // namespace

    template <typename P>
    void McgsSolver::bcgIteration(hipStream_t s, const BcgState & b, SolveState st, const McgsMatrix<P> & M, P * rP, P * z)
    {
        const McgsMatrix<double> A = A64_.view(n_);
        const int g = bcgBlocks_;
        dispatchWide(opt_.numCh, [&]<int C>()
                     {
                         constexpr int red = (C * C + 63) / 64 + 1;
                         bcgProductKernel<double, C><<<g, kBlock, 0, s>>>(A, b, p64_.get(), q64_.get());
                         bcgGramReduceKernel<C><<<(red - 1) * kBcgSlices, kBlock, 0, s>>>(b, 0);
                         bcgAlphaKernel<C><<<1, kBcgSmallBlock, 0, s>>>(b);
                         bcgUpdateKernel<double, P, C><<<g, kBlock, 0, s>>>(b, M.invDiag, p64_.get(), q64_.get(), x64_.get(), r64_.get(),
                                                                             rP, z, color0End_);
                     });
        sgsSweeps<P>(s, M, st, rP, z);
        dispatchWide(opt_.numCh, [&]<int C>()
                     {
                         constexpr int red = (C * C + 63) / 64 + 1;
                         bcgRhoKernel<P, double, C><<<g, kBlock, 0, s>>>(M, b, rP, r64_.get(), z, color0End_);
                         bcgGramReduceKernel<C><<<red * kBcgSlices, kBlock, 0, s>>>(b, 1);
                         bcgBetaKernel<C><<<1, kBcgSmallBlock, 0, s>>>(b, 0);
                         bcgDirectionKernel<double, P, C><<<g, kBlock, 0, s>>>(b, z, p64_.get(), 0);
                     });
    }

    template <typename P>
    void McgsSolver::runBlockPcg(hipStream_t s, const McgsMatrix<P> & M, P * rP, P * z, bool zeroGuess)
    {
        const int nch = opt_.numCh;
        const int K = nch;
        SolveState st = inner_.view(K, opt_.target * opt_.target);
        const BcgState b = bcgState(st);
        const McgsMatrix<double> A = A64_.view(n_);
        const int g = bcgBlocks_;
        fillInt(K, 1, st.active, s); // the sweeps run while the solve does
        fillInt(5, 0, b.iters, s);
        // cycles: a (re)start on the true residual B - A X, then iterations in
        // chunks with lagged polls of the stop flag (as the single-RHS
        // solver); the kernels after a stop exit at once.  A stop with a
        // restart or verification request starts the next cycle.
        constexpr int kMaxCycles = 200;
        const int chunk = std::max(1, opt_.checkChunk);
        int launched = 0;
        int cycles = 0;
        int hostFlags[4] = {0, 0, 0, 0};
        for (;; cycles++)
        {
            if (cycles > 0 && opt_.projectMean)
            {
                // verify (and restart from) the zero-mean solution that is
                // returned: A 1 = 0 only up to rounding, so a mean removed
                // after the check would change the residual
                projectMean<double>(s, x64_.get());
            }
            fillInt(4, 0, st.flags, s);
            fillInt(4, 0, b.iters + 1, s); // the cycle's rank bookkeeping
            const int first = cycles == 0;
            dispatchWide(nch, [&]<int C>()
                         {
                             bcgInitKernel<double, P, C><<<g, kBlock, 0, s>>>(A, b, b64_.get(), x64_.get(), r64_.get(), rP, z, M.invDiag,
                                                                               color0End_, first && zeroGuess ? 1 : 0);
                         });
            sgsSweeps<P>(s, M, st, rP, z);
            dispatchWide(nch, [&]<int C>()
                         {
                             constexpr int red = (C * C + 63) / 64 + 1;
                             bcgRhoKernel<P, double, C><<<g, kBlock, 0, s>>>(M, b, rP, r64_.get(), z, color0End_);
                             bcgGramReduceKernel<C><<<red * kBcgSlices, kBlock, 0, s>>>(b, 1);
                             bcgBetaKernel<C><<<1, kBcgSmallBlock, 0, s>>>(b, first ? 2 : 1);
                             bcgDirectionKernel<double, P, C><<<g, kBlock, 0, s>>>(b, z, p64_.get(), 1);
                         });
            NSFEM_HIP(hipGetLastError());
            int slot = 0;
            bool pending[2] = {false, false};
            bool stopped = false;
            while (launched < opt_.maxIters)
            {
                const int k = std::min(chunk, opt_.maxIters - launched);
                for (int i = 0; i < k; i++)
                {
                    bcgIteration<P>(s, b, st, M, rP, z);
                }
                NSFEM_HIP(hipGetLastError());
                launched += k;
                NSFEM_HIP(hipMemcpyAsync(pollRing_ + 4 * slot, st.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
                NSFEM_HIP(hipEventRecord(pollEvents_[slot], s));
                pending[slot] = true;
                const int prev = slot ^ 1;
                if (pending[prev])
                {
                    NSFEM_HIP(hipEventSynchronize(pollEvents_[prev]));
                    pending[prev] = false;
                    if (pollRing_[4 * prev] != 0)
                    {
                        stopped = true;
                        break;
                    }
                }
                slot = prev;
            }
            int done = 0;
            NSFEM_HIP(hipMemcpyAsync(hostFlags, st.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
            NSFEM_HIP(hipMemcpyAsync(&done, b.iters, sizeof(int), hipMemcpyDeviceToHost, s));
            NSFEM_HIP(hipStreamSynchronize(s));
            (void)stopped;
            launched = done; // the budget counts updates, not the idle launches after a stop
            const bool again = hostFlags[0] != 0 && hostFlags[2] == 0 && (hostFlags[3] == 1 || hostFlags[3] == 2);
            if (!again || launched >= opt_.maxIters || cycles + 1 >= kMaxCycles)
            {
                break;
            }
        }
        lastOuter_ = cycles; // restarts and verifications after the first cycle
        // statistics: |b_c|^2, |r_c|^2, updates done, dropped directions, flags
        std::vector<double> h(2 * nch);
        int ints[2] = {0, 0};
        int flags[4] = {0, 0, 0, 0};
        NSFEM_HIP(hipMemcpyAsync(h.data(), b.bb, 2 * nch * sizeof(double), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipMemcpyAsync(ints, b.iters, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipMemcpyAsync(flags, st.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipStreamSynchronize(s));
        bool finite = true;
        for (int c = 0; c < nch; c++)
        {
            const double bb = h[c], rr = h[nch + c];
            finite = finite && std::isfinite(rr);
            lastMse_[c] = std::sqrt(rr / n_);
            lastRel_[c] = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        }
        lastIters_[0] = ints[0];
        lastDropped_ = ints[1] != 0;
        lastNonFinite_ = !finite || flags[2] != 0;
    }

    void McgsSolver::solve(bool zeroInitial)
    {
        hipStream_t s = stream_;
        const int nch = opt_.numCh;
        const int K = numSystems_ * nch;
        lastOuter_ = 0;
        if (opt_.method == SolverMethod::AmgPcg)
        {
            permuteRows(F64, ioRhs_.get(), b64_.get(), false, s);
            if (!zeroInitial)
            {
                permuteRows(F64, ioSol_.get(), x64_.get(), false, s);
            }
            const AmgPcg::Result r = amg_->solve(b64_.get(), x64_.get(), zeroInitial, opt_.target, opt_.maxIters, useGraphs_);
            for (int c = 0; c < nch; c++)
            {
                lastMse_[c] = std::sqrt(r.rr[c] / n_);
                lastRel_[c] = r.bb[c] > 0 ? std::sqrt(r.rr[c] / r.bb[c]) : 0.0;
            }
            lastIters_[0] = r.iterations;
            lastNonFinite_ = r.nonFinite;
            if (opt_.projectMean)
            {
                projectMean<double>(s, x64_.get());
            }
            permuteRows(F64, x64_.get(), ioSol_.get(), true, s);
            NSFEM_HIP(hipStreamSynchronize(s));
            return;
        }
        if (opt_.method == SolverMethod::BlockPcg)
        {
            permuteRows(F64, ioRhs_.get(), b64_.get(), false, s);
            if (!zeroInitial)
            {
                permuteRows(F64, ioSol_.get(), x64_.get(), false, s);
            }
            if (opt_.precision == SolverPrecision::F64)
            {
                runBlockPcg<double>(s, A64_.view(n_), r64_.get(), z64_.get(), zeroInitial);
            }
            else
            {
                runBlockPcg<float>(s, A32_.view(n_), r32_.get(), z32_.get(), zeroInitial);
            }
            if (opt_.projectMean)
            {
                projectMean<double>(s, x64_.get());
            }
            permuteRows(F64, x64_.get(), ioSol_.get(), true, s);
            NSFEM_HIP(hipStreamSynchronize(s));
            return;
        }
        const bool pcgZero = zeroInitial && opt_.method == SolverMethod::Pcg && !(opt_.precision == SolverPrecision::Mixed && opt_.restart);
        if (zeroInitial && !pcgZero)
        {
            fill(opt_.precision == SolverPrecision::F32 ? F32 : F64, static_cast<long>(n_) * nch, 0.0, ioSol_.get(), s);
        }
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
        if (opt_.precision == SolverPrecision::Mixed && opt_.restart)
        {
            runMixedRestart(s);
            return;
        }
        const double tol2 = opt_.target * opt_.target;
        SolveState st = inner_.view(K, tol2);
        if (opt_.precision == SolverPrecision::F32)
        {
            permuteRows(F32, ioRhs_.get(), b32_.get(), false, s);
            if (!pcgZero)
            {
                permuteRows(F32, ioSol_.get(), x32_.get(), false, s);
            }
            runPcg<float, float>(s, operands<float, float>(), st, pcgZero, nullptr, opt_.maxIters, true, true);
            if (!opt_.projectMean)
            {
                // queued before the statistics' synchronisation: one host
                // round trip fewer with the GPU idle
                permuteRows(F32, x32_.get(), ioSol_.get(), true, s);
            }
            readStats(s, inner_, false, 0);
            lastSolveIters_ = *std::max_element(lastIters_.begin(), lastIters_.end());
            if (opt_.projectMean)
            {
                projectMean<float>(s, x32_.get());
                permuteRows(F32, x32_.get(), ioSol_.get(), true, s);
            }
        }
        else
        {
            // FP64 Krylov vectors and products; the preconditioner in FP64
            // (plain) or FP32 (mixed)
            permuteRows(F64, ioRhs_.get(), b64_.get(), false, s);
            if (!pcgZero)
            {
                permuteRows(F64, ioSol_.get(), x64_.get(), false, s);
            }
            if (opt_.precision == SolverPrecision::F64)
            {
                runPcg<double, double>(s, operands<double, double>(), st, pcgZero, nullptr, opt_.maxIters, true, true);
            }
            else
            {
                runPcg<double, float>(s, operands<double, float>(), st, pcgZero, nullptr, opt_.maxIters, true, true);
            }
            if (!opt_.projectMean)
            {
                // queued before the statistics' synchronisation: one host
                // round trip fewer with the GPU idle
                permuteRows(F64, x64_.get(), ioSol_.get(), true, s);
            }
            readStats(s, inner_, false, 0);
            lastSolveIters_ = *std::max_element(lastIters_.begin(), lastIters_.end());
            if (opt_.projectMean)
            {
                projectMean<double>(s, x64_.get());
                permuteRows(F64, x64_.get(), ioSol_.get(), true, s);
            }
        }
        NSFEM_HIP(hipStreamSynchronize(s));
    }

    void McgsSolver::permuteRows(int dt, const void * src, void * dst, bool inverse, hipStream_t s)
    {
        const int nch = opt_.numCh;
        if (nch > 2 || n_ <= 0 || !(inverse || bmDir_.rowTable))
        {
            permute(dt, dt, n_, nch, perm_.get(), src, dst, inverse, s);
            return;
        }
        const int inv = inverse ? 1 : 0;
        auto go = [&]<typename T>()
        {
            const T * sp = static_cast<const T *>(src);
            T * dp = static_cast<T *>(dst);
            if (inverse)
            {
                const int g = (n_ + kBlock - 1) / kBlock;
                if (nch == 1)
                {
                    unpermuteGatherKernel<T, T, 1><<<g, kBlock, 0, s>>>(n_, iperm_.get(), sp, dp);
                }
                else
                {
                    unpermuteGatherKernel<T, T, 2><<<g, kBlock, 0, s>>>(n_, iperm_.get(), sp, dp);
                }
            }
            else if (nch == 1)
            {
                permuteMappedKernel<T, T, 1><<<nbDir_, kBlock, 0, s>>>(bmDir_, n_, perm_.get(), sp, dp, inv);
            }
            else
            {
                permuteMappedKernel<T, T, 2><<<nbDir_, kBlock, 0, s>>>(bmDir_, n_, perm_.get(), sp, dp, inv);
            }
        };
        if (dt == F64)
        {
            go.template operator()<double>();
        }
        else
        {
            go.template operator()<float>();
        }
        NSFEM_HIP(hipGetLastError());
    }

    template <typename T>
    void McgsSolver::projectMean(hipStream_t s, T * v)
    {
        // reuse the inner state's reduction scratch (no solve is in flight)
        dispatchC(opt_.numCh, [&]<int C>()
                  {
                      meanKernel<T, C><<<numBlocks_, kBlock, 0, s>>>(n_, bm_, v, inner_.partials.get());
                      meanFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bm_, inner_.partials.get(), inner_.totals.get(), means_.get());
                      subtractMeanKernel<T, C><<<numBlocks_, kBlock, 0, s>>>(n_, bm_, v, means_.get());
                  });
        NSFEM_HIP(hipGetLastError());
    }

    void McgsSolver::solveAsync(hipStream_t) { solve(); }

    template <typename T, typename P>
    void McgsSolver::enqueuePcg(hipStream_t s, const PcgOperands<T, P> & v, SolveState st, bool zeroGuess, int budget)
    {
        const int nch = opt_.numCh;
        double * totals = inner_.totals.get();
        fillInt(4, 0, st.flags, s);
        dispatchLC(directionLanes(), nch, [&]<int L, int C>()
                   {
                       pcgInitKernel<T, P, L, C><<<nbDir_, kBlock, 0, s>>>(v.A, bmDir_, st, v.b, v.x, v.r, v.rP, v.z, v.M.invDiag,
                                                                          zColour0End(), zeroGuess ? 1 : 0);
                       pcgInitFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bmDir_, st, nullptr, totals);
                   });
        launchPrecond<T, P>(s, v, st, totals, false);
        for (int i = 0; i < budget; i++)
        {
            pcgIteration<T, P>(s, v, st);
        }
        if (fused())
        {
            // the last update's convergence test (no direction kernel follows)
            dispatchC(nch, [&]<int C>() { pcgUpdateCheckKernel<C><<<1, kFinBlock, 0, s>>>(bmUpd_, st, totals); });
        }
        NSFEM_HIP(hipGetLastError());
    }

    void McgsSolver::enqueueSolve(hipStream_t s, bool zeroInitial, int budget)
    {
        if (amg_)
        {
            if (budget < 1)
            {
                throw std::invalid_argument("captured solve: the iteration budget must be positive");
            }
            const int nch = opt_.numCh;
            permuteRows(F64, ioRhs_.get(), b64_.get(), false, s);
            if (!zeroInitial)
            {
                permuteRows(F64, ioSol_.get(), x64_.get(), false, s);
            }
            amg_->enqueueSolve(s, b64_.get(), x64_.get(), zeroInitial, opt_.target, std::min(budget, opt_.maxIters));
            permuteRows(F64, x64_.get(), ioSol_.get(), true, s);
            return;
        }
        if (opt_.method != SolverMethod::Pcg || (opt_.precision == SolverPrecision::Mixed && opt_.restart))
        {
            throw std::invalid_argument("captured solves need a pcg solver (pcg or pcg_mixed without mixedRestart, amg_pcg, amg_pcg_mixed)");
        }
        if (budget < 1)
        {
            throw std::invalid_argument("captured solve: the iteration budget must be positive");
        }
        budget = std::min(budget, opt_.maxIters);
        const int nch = opt_.numCh;
        const int K = numSystems_ * nch;
        SolveState st = inner_.view(K, opt_.target * opt_.target);
        if (opt_.precision == SolverPrecision::F32)
        {
            permuteRows(F32, ioRhs_.get(), b32_.get(), false, s);
            if (!zeroInitial)
            {
                permuteRows(F32, ioSol_.get(), x32_.get(), false, s);
            }
            enqueuePcg<float, float>(s, operands<float, float>(), st, zeroInitial, budget);
            if (opt_.projectMean)
            {
                projectMean<float>(s, x32_.get());
            }
            permuteRows(F32, x32_.get(), ioSol_.get(), true, s);
        }
        else
        {
            permuteRows(F64, ioRhs_.get(), b64_.get(), false, s);
            if (!zeroInitial)
            {
                permuteRows(F64, ioSol_.get(), x64_.get(), false, s);
            }
            if (opt_.precision == SolverPrecision::F64)
            {
                enqueuePcg<double, double>(s, operands<double, double>(), st, zeroInitial, budget);
            }
            else
            {
                enqueuePcg<double, float>(s, operands<double, float>(), st, zeroInitial, budget);
            }
            if (opt_.projectMean)
            {
                projectMean<double>(s, x64_.get());
            }
            permuteRows(F64, x64_.get(), ioSol_.get(), true, s);
        }
    }

    void McgsSolver::snapshotStats(hipStream_t s, int slot)
    {
        if (slot < 0 || slot >= kStatSlots)
        {
            throw std::invalid_argument("statistics slot out of range");
        }
        const int K = numSystems_ * opt_.numCh;
        const size_t scalBytes = 8 * static_cast<size_t>(K) * sizeof(double);
        const size_t intBytes = (3 * static_cast<size_t>(K) + 4) * sizeof(int);
        const size_t slotBytes = scalBytes + intBytes;
        char * base = statSlots_.get() + slot * slotBytes;
        if (amg_)
        {
            // its per-channel scalars (4 K doubles) and 4 ints fit the slot
            NSFEM_HIP(hipMemcpyAsync(base, amg_->deviceScalars(), 4 * static_cast<size_t>(K) * sizeof(double), hipMemcpyDeviceToDevice, s));
            NSFEM_HIP(hipMemcpyAsync(base + scalBytes, amg_->deviceInts(), 4 * sizeof(int), hipMemcpyDeviceToDevice, s));
            return;
        }
        NSFEM_HIP(hipMemcpyAsync(base, inner_.scal.get(), scalBytes, hipMemcpyDeviceToDevice, s));
        NSFEM_HIP(hipMemcpyAsync(base + scalBytes, inner_.ints.get(), intBytes, hipMemcpyDeviceToDevice, s));
    }

    bool McgsSolver::collectStats(int slot)
    {
        const int K = numSystems_ * opt_.numCh;
        const size_t scalBytes = 8 * static_cast<size_t>(K) * sizeof(double);
        const size_t intBytes = (3 * static_cast<size_t>(K) + 4) * sizeof(int);
        const size_t slotBytes = scalBytes + intBytes;
        if (slot < 0 || slot >= kStatSlots || statSlots_.size() < kStatSlots * slotBytes)
        {
            throw std::invalid_argument("no statistics in that slot");
        }
        const char * base = statSlots_.get() + slot * slotBytes;
        if (amg_)
        {
            NSFEM_HIP(hipMemcpyAsync(hostScal_, base, 4 * K * sizeof(double), hipMemcpyDeviceToHost, stream_));
            NSFEM_HIP(hipMemcpyAsync(hostInts_, base + scalBytes, 4 * sizeof(int), hipMemcpyDeviceToHost, stream_));
            NSFEM_HIP(hipStreamSynchronize(stream_));
            const int stop = hostInts_[0], iters = hostInts_[1], zero = hostInts_[2];
            for (int c = 0; c < K; c++)
            {
                const double bb = hostScal_[2 * K + c], rr = hostScal_[3 * K + c];
                lastMse_[c] = std::sqrt(rr / n_);
                lastRel_[c] = bb > 0 ? std::sqrt(rr / bb) : 0.0;
            }
            lastIters_[0] = iters;
            lastNonFinite_ = stop == 2;
            lastOuter_ = 0;
            // a zero right-hand side from a nonzero guess is zeroed by the
            // regular solve only: rerun
            return zero == 0 && (stop != 0 || iters >= opt_.maxIters);
        }
        NSFEM_HIP(hipMemcpyAsync(hostScal_, base, 7 * K * sizeof(double), hipMemcpyDeviceToHost, stream_));
        NSFEM_HIP(hipMemcpyAsync(hostInts_, base + scalBytes, (2 * K + 4) * sizeof(int), hipMemcpyDeviceToHost, stream_));
        NSFEM_HIP(hipStreamSynchronize(stream_));
        const double * rr = hostScal_ + 3 * K;
        const double * ref = hostScal_ + 5 * K;
        const int * active = hostInts_;
        const int * iters = hostInts_ + K;
        lastNonFinite_ = hostInts_[2 * K + 2] != 0;
        lastOuter_ = 0;
        bool complete = true;
        for (int sys = 0; sys < numSystems_; sys++)
        {
            int it = 0;
            for (int c = 0; c < opt_.numCh; c++)
            {
                const int k = sys * opt_.numCh + c;
                lastMse_[k] = std::sqrt(rr[k] / sysRows_[sys]);
                lastRel_[k] = ref[k] > 0 ? std::sqrt(rr[k] / ref[k]) : 0.0;
                it = std::max(it, iters[k]);
                if (active[k] != 0 && iters[k] < opt_.maxIters)
                {
                    complete = false;
                }
            }
            lastIters_[sys] = it;
        }
        return complete;
    }

    double McgsSolver::timeKernel(const std::string & which, int reps, double * bytesPerLaunch, double * launchesPerIteration)
    {
        if (reps < 1)
        {
            throw std::invalid_argument("reps must be positive");
        }
        if (amg_)
        {
            return amg_->timeKernel(which, reps, bytesPerLaunch, launchesPerIteration, x64_.get());
        }
        if (opt_.method != SolverMethod::Pcg)
        {
            throw std::invalid_argument("kernel timing is available for pcg and amg_pcg solvers");
        }
        hipStream_t s = stream_;
        const int nch = opt_.numCh;
        const int K = numSystems_ * nch;
        // a negative tolerance: the convergence test of a timed iteration
        // never stops it, whatever the work vectors hold
        SolveState st = inner_.view(K, -1.0);
        double * totals = inner_.totals.get();
        double nnzOff = 0;
        for (const auto & cs : colorStats_)
        {
            nnzOff += cs.fullNnz;
        }
        const double n = n_;
        const double n0 = colorStats_.empty() ? 0.0 : colorStats_[0].rows;
        double bytes = 0;
        double launches = 1;
        bool graphTiming = false;
        std::function<void()> launch;
        auto run = [&](auto tagT, auto tagP)
        {
            using T = decltype(tagT);
            using P = decltype(tagP);
            constexpr bool same = std::is_same_v<T, P>;
            const double tsT = sizeof(T), tsP = sizeof(P);
            const PcgOperands<T, P> v = operands<T, P>();
            const int fz = fused() ? 1 : 0;
            // finite work vectors (a fresh solver's memory is arbitrary)
            const long nv = static_cast<long>(n_) * nch;
            const int dtT = std::is_same_v<T, double> ? F64 : F32;
            const int dtP = std::is_same_v<P, double> ? F64 : F32;
            for (T * ptr : {v.x, v.r, v.p, v.q})
            {
                fill(dtT, nv, 0.0, ptr, s);
            }
            for (P * ptr : {v.rP, v.z})
            {
                fill(dtP, nv, 0.0, ptr, s);
            }
            // Algorithmic bytes of one launch, x read once (ideal reuse).
            auto sweepBytes = [&](int c, bool lower)
            {
                const ColorStats & cs = colorStats_[c];
                const double nnz = lower ? cs.lowerNnz : cs.fullNnz;
                const double cols = lower ? cs.lowerCols : cs.fullCols;
                return nnz * (tsP + 4.0) + (cs.rows + 1) * 4.0 + (lower ? cs.rows * 4.0 : 0.0) + cs.rows * tsP +
                       nch * (cs.rows * tsP + cs.rows * tsP + cols * tsP);
            };
            if (which == "direction")
            {
                // matrix and diagonal, z gathered once, p and q read and written
                bytes = nnzOff * (tsT + 4.0) + (n + 1) * 4.0 + n * tsT + nch * (n * tsP + 4.0 * n * tsT);
                launch = [=, this]()
                {
                    dispatchLC(directionLanes(), nch, [&]<int L, int C>()
                               {
                                   pcgDirectionKernel<T, P, L, C><<<nbDir_, kBlock, 0, s>>>(v.A, bmDir_, st, v.z, v.p, v.q, fz, rhoParts(),
                                                                                            nbUpd_);
                               });
                };
            }
            else if (which == "sweep")
            {
                // the colour sweeps of one preconditioner application:
                // forward colours 1 .. C-1 (lower part; from 2 when colour 1
                // is swept with the update), backward C-2 .. 1
                const int firstForward = mergedUpdate() ? 2 : 1;
                double total = 0;
                int count = 0;
                for (int c = firstForward; c < numColors_; c++)
                {
                    if (colorSweepBlockStart_[c + 1] > colorSweepBlockStart_[c])
                    {
                        total += sweepBytes(c, true);
                        count++;
                    }
                }
                for (int c = numColors_ - 2; c >= 1; c--)
                {
                    if (colorSweepBlockStart_[c + 1] > colorSweepBlockStart_[c])
                    {
                        total += sweepBytes(c, false);
                        count++;
                    }
                }
                if (numColors_ < 2)
                {
                    throw std::invalid_argument("the system has a single colour: no sweep launches");
                }
                launches = count;
                bytes = count > 0 ? total / count : 0.0;
                launch = [=, this]() { sgsSweeps<P>(s, v.M, st, v.rP, v.z, firstForward); };
            }
            else if (which == "rho")
            {
                // backward colour 0 (full rows) fused with r.z over all rows
                // (after the fused update and sweep: also r -= alpha q on
                // colour 0, q read and r written there)
                const int fix = mergedUpdate() ? 1 : 0;
                bytes = (numColors_ > 0 ? sweepBytes(0, false) : 0.0) + nch * n * tsT + nch * (n - n0) * tsP -
                        (same ? nch * n0 * tsP : 0.0) + fix * 2.0 * nch * n0 * tsT;
                launch = [=, this]()
                {
                    dispatchLC(lanes_, nch, [&]<int L, int C>()
                               {
                                   pcgRhoKernel<P, T, L, C><<<nbRho_, kBlock, 0, s>>>(v.M, bmRho_, st, v.rP, v.r, v.z, color0End_,
                                                                                  color0BlocksRho_, v.q, fix);
                               });
                };
            }
            else if (which == "fsai")
            {
                // the factorized approximate inverse's two launches (G r over
                // the lower parts, G^T t over the higher parts with r.z): every
                // off-diagonal entry once over both, row bounds and the
                // factor's diagonal in each, r, t, t, z in P and r in T once
                // (gathered columns read once)
                if (precond_ != 1)
                {
                    throw std::invalid_argument("fsai: the solver's preconditioner is not the approximate inverse");
                }
                double full = 0;
                for (const ColorStats & c : colorStats_)
                {
                    full += c.fullNnz;
                }
                launches = 2;
                bytes = (full * (tsP + 4.0) + 2.0 * ((n + 1) * 4.0 + n * 4.0 + n * tsP) + nch * n * (4.0 * tsP + tsT)) / 2.0;
                launch = [=, this]()
                {
                    dispatchLC(directionLanes(), nch, [&]<int L, int C>()
                               {
                                   fsaiLowerKernel<P, L, C><<<nbPre_, kBlock, 0, s>>>(v.G, bmPre_, st, v.rP, v.t);
                                   fsaiUpperKernel<P, T, L, C><<<nbPre_, kBlock, 0, s>>>(v.G, bmPre_, st, v.t, v.z, v.r);
                               });
                };
            }
            else if (which == "update")
            {
                // x, r read and written, p, q read; rP written (P narrower
                // than T); z = r / d on colour 0
                bytes = 6.0 * nch * n * tsT + (same ? 0.0 : nch * n * tsP) + n0 * tsP + nch * n0 * tsP;
                if (mergedUpdate())
                {
                    // with the forward sweep of colour 1: its lower part, its
                    // diagonal and z written; the colour-0 columns it
                    // gathers read r and q and their diagonal once; colour 0's
                    // r is written by the rho kernel instead
                    const ColorStats & c1 = colorStats_[1];
                    bytes += c1.lowerNnz * (tsP + 4.0) + (c1.rows + 1) * 4.0 + c1.rows * 4.0 + c1.rows * tsP +
                             nch * c1.rows * tsP + c1.lowerCols * (2.0 * nch * tsT + tsP) - nch * n0 * tsT;
                    launch = [=, this]() { launchUpdateSweep<T, P>(s, v, st); };
                }
                else
                {
                    launch = [=, this]()
                    {
                        dispatchC(nch, [&]<int C>()
                                  {
                                      pcgUpdateKernel<T, P, C><<<nbUpd_, kBlock, 0, s>>>(n_, bmUpd_, st, v.p, v.q, v.x, v.r, v.rP,
                                                                                        v.z, v.M.invDiag, color0End_, fz, nbDir_);
                                  });
                    };
                }
            }
            else if (which == "finalize")
            {
                // one of the per-iteration one-block finalize kernels (two;
                // none with fused reductions)
                launches = fz ? 0 : 2;
                bytes = static_cast<double>(numBlocks_) * nch * 8.0;
                launch = [=, this]()
                { dispatchC(nch, [&]<int C>() { pcgDirectionFinalizeKernel<C><<<1, kFinBlock, 0, s>>>(bm_, st, totals); }); };
            }
            else if (which == "iteration" || which == "iteration_graph")
            {
                // one whole PCG iteration (all its launches); "iteration_graph"
                // replays the repetitions as one captured graph, as the solver
                // does, "iteration" queues them eagerly behind the delay
                graphTiming = which == "iteration_graph";
                bytes = 0;
                launch = [=, this]() { pcgIteration<T, P>(s, v, st); };
            }
            else
            {
                throw std::invalid_argument("unknown kernel '" + which +
                                            "' (direction, sweep, rho, update, finalize, iteration, iteration_graph)");
            }
        };
        if (opt_.precision == SolverPrecision::F64)
        {
            run(double{}, double{});
        }
        else if (opt_.precision == SolverPrecision::F32 || opt_.restart)
        {
            run(float{}, float{});
        }
        else
        {
            run(double{}, float{});
        }
        // A state in which every kernel does its full work: all channels
        // active, not the first iteration, zero step lengths.
        fillInt(K, 1, st.active, s);
        fillInt(K, 1, st.activeNext, s);
        fillInt(4, 0, st.flags, s);
        // finite partial sums: a fused direction kernel reads the rho and
        // update partials before any launch of this replay has written them
        for (DeviceArray<double> * a : {&inner_.partials, &inner_.partials2, &inner_.partials3})
        {
            fill(F64, static_cast<long>(a->size()), 0.0, a->get(), s);
        }
        fill(F64, 3L * K, 0.0, st.rho, s); // rho, alpha, beta
        fill(F64, 1L * K, 0.0, st.rhoNext, s);
        fill(F64, 1L * K, 1.0, st.ref, s);
        launch();
        NSFEM_HIP(hipGetLastError());
        // The stream idles on a 20 ms delay kernel while the host queues the
        // repetitions, so they run back to back and the events bracket GPU
        // execution only (the figures match a kernel trace of the same run).
        hipGraphExec_t exec = nullptr;
        if (graphTiming)
        {
            hipGraph_t graph = nullptr;
            NSFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < reps; i++)
            {
                launch();
            }
            NSFEM_HIP(hipStreamEndCapture(s, &graph));
            NSFEM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            NSFEM_HIP(hipGraphDestroy(graph));
            NSFEM_HIP(hipGraphLaunch(exec, s));
        }
        NSFEM_HIP(hipStreamSynchronize(s));
        delayKernel<<<1, 64, 0, s>>>(2000000);
        hipEvent_t e0, e1;
        NSFEM_HIP(hipEventCreate(&e0));
        NSFEM_HIP(hipEventCreate(&e1));
        NSFEM_HIP(hipEventRecord(e0, s));
        if (graphTiming)
        {
            NSFEM_HIP(hipGraphLaunch(exec, s));
        }
        else
        {
            for (int i = 0; i < reps; i++)
            {
                launch();
            }
        }
        NSFEM_HIP(hipEventRecord(e1, s));
        NSFEM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        NSFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        NSFEM_HIP(hipEventDestroy(e0));
        NSFEM_HIP(hipEventDestroy(e1));
        if (exec)
        {
            NSFEM_HIP(hipGraphExecDestroy(exec));
        }
        if (bytesPerLaunch)
        {
            *bytesPerLaunch = bytes;
        }
        if (launchesPerIteration)
        {
            *launchesPerIteration = launches;
        }
        // "sweep" and "fsai" launch several kernels per repetition
        const double perRep = which == "sweep" || which == "fsai" ? launches : 1.0;
        return static_cast<double>(ms) / (reps * perRep);
    }

    template void McgsSolver::runGs<double>(hipStream_t);
    template void McgsSolver::runGs<float>(hipStream_t);
} // namespace nsfem
