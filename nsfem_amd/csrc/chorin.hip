#include "chorin.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>

#include "kernels.hpp"

namespace nsfem
{
    // Reference-element convection tensors (see ConvectionTensors).
    __constant__ double gConvTx[216];
    __constant__ double gConvTy[216];

    // Local 6x6 convection matrix of every P2 element from the current
    // velocity, stored structure-of-arrays: elemC[(r*6 + c) * E + e].
    template <typename T>
    __global__ void __launch_bounds__(256) convectionElementKernel(int E, const int * __restrict__ nodes, const double * __restrict__ geom,
                                                                   const T * __restrict__ u, int N, T * __restrict__ elemC)
    {
        const int e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= E)
        {
            return;
        }
        const double * g = geom + static_cast<size_t>(e) * 5;
        const double b00 = g[0], b01 = g[1], b10 = g[2], b11 = g[3], sgn = g[4];
        double a[6], c[6];
#pragma unroll
        for (int k = 0; k < 6; k++)
        {
            const int id = nodes[static_cast<size_t>(e) * 6 + k];
            const double ux = static_cast<double>(u[id]);
            const double uy = static_cast<double>(u[id + N]);
            a[k] = b00 * ux + b10 * uy; // reference-frame flow, x component
            c[k] = b01 * ux + b11 * uy; // y component
        }
#pragma unroll
        for (int rc = 0; rc < 36; rc++)
        {
            double s = 0;
#pragma unroll
            for (int k = 0; k < 6; k++)
            {
                s = fma(a[k], gConvTx[rc * 6 + k], s);
                s = fma(c[k], gConvTy[rc * 6 + k], s);
            }
            elemC[static_cast<size_t>(rc) * E + e] = static_cast<T>(sgn * s);
        }
    }

    // A[j] = nuK[j] + sum of the element contributions that land on entry j.
    template <typename T>
    __global__ void __launch_bounds__(256) convectionGatherKernel(int nnz, const int * __restrict__ cs, const int * __restrict__ ci,
                                                                  const T * __restrict__ elemC, const T * __restrict__ nuK,
                                                                  T * __restrict__ A, T * __restrict__ C)
    {
        const int j = blockIdx.x * blockDim.x + threadIdx.x;
        if (j >= nnz)
        {
            return;
        }
        // two contributions per step, loaded unconditionally (a missing
        // second one reloads the first and adds -0.0, which leaves every sum
        // unchanged), so both gathers go out together; the sum keeps the
        // list order.  A use under a condition would let the compiler move
        // the loads behind a branch that waits for the first pair.
        const int t1 = cs[j + 1];
        const T nk = nuK[j];
        double s = 0;
        for (int t = cs[j]; t < t1; t += 2)
        {
            const bool h = t + 1 < t1;
            const T e0 = elemC[ci[t]];
            const T e1 = elemC[ci[h ? t + 1 : t]];
            s += static_cast<double>(e0);
            s += h ? static_cast<double>(e1) : -0.0;
        }
        if (C)
        {
            C[j] = static_cast<T>(s);
        }
        A[j] = static_cast<T>(static_cast<double>(nk) + s);
    }

    void ChorinGpu::upload(DevCsr & d, const CsrD & m, const std::vector<double> * scaled)
    {
        d.rows = m.rows;
        d.cols = m.cols;
        d.nnz = m.nnz();
        d.rowStart.upload(m.rowStart);
        d.column.upload(m.column);
        const std::vector<double> & v = scaled ? *scaled : m.values;
        if (cfg_.dtype == F64)
        {
            d.values.upload(reinterpret_cast<const char *>(v.data()), v.size() * 8);
        }
        else
        {
            std::vector<float> f(v.begin(), v.end());
            d.values.upload(reinterpret_cast<const char *>(f.data()), f.size() * 4);
        }
    }

    ChorinGpu::ChorinGpu(const ChorinContext & ctx, const ChorinConfig & cfg) : cfg_(cfg)
    {
        checkDtype(cfg_.dtype);
        tsize_ = dtypeSize(cfg_.dtype);
        N_ = ctx.numVelocityNodes;
        V_ = ctx.numPressureNodes;
        E_ = ctx.convection.numElements;
        nu_ = ctx.viscosity;
        B_ = cfg_.instances;
        if (B_ != 1 && B_ != 2 && B_ != 4 && B_ != 8 && B_ != 16)
        {
            throw std::invalid_argument("ensemble instances must be 1, 2, 4, 8 or 16");
        }
        if (B_ > 1 && !ctx.massBlocks.empty())
        {
            throw std::invalid_argument("an ensemble runs on a single-mesh context, not a batch");
        }
        if (!cfg_.inflowScale.empty() && static_cast<int>(cfg_.inflowScale.size()) != B_)
        {
            throw std::invalid_argument("inflowScale needs one value per instance");
        }
        const hipStream_t s = libraryStream();

        std::vector<double> nuK(ctx.velocityStiffness.values);
        for (auto & v : nuK)
        {
            v *= nu_;
        }
        upload(nuK_, ctx.velocityStiffness, &nuK);
        upload(A_, ctx.velocityStiffness, &nuK);
        if (B_ > 1)
        {
            A_.values.resize(static_cast<size_t>(A_.nnz) * B_ * tsize_); // each instance's A (set every step)
        }
        upload(vpd_, ctx.velocityPressureDiv);
        upload(pvd_, ctx.pressureVelocityDiv);

        // Element-wise convection data; contributions re-indexed for the
        // structure-of-arrays element buffer.
        const auto & cd = ctx.convection;
        elemNodes_.upload(cd.elementNodes);
        elemGeom_.upload(cd.geometry);
        contribStart_.upload(cd.contribStart);
        {
            std::vector<int> soa(cd.contribIndex.size());
            for (size_t t = 0; t < soa.size(); t++)
            {
                const int k = cd.contribIndex[t];
                soa[t] = (k % 36) * E_ + k / 36;
            }
            contribIndex_.upload(soa);
        }
        elemC_.resize(static_cast<size_t>(E_) * 36 * tsize_);
        const auto & T = convectionTensors();
        NSFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(gConvTx), &T.tx[0][0][0], sizeof(double) * 216));
        NSFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(gConvTy), &T.ty[0][0][0], sizeof(double) * 216));

        // Dirichlet velocity: X ids, then Y ids offset by N (chorinCuda.cu:41-54),
        // for every instance (offset 2N b, values scaled by its inflow scale)
        const int perInstance = static_cast<int>(ctx.dirichletVx.size() + ctx.dirichletVy.size());
        nDir_ = perInstance * B_;
        std::vector<int> ids;
        std::vector<double> vals;
        for (int b = 0; b < B_; b++)
        {
            const double scale = cfg_.inflowScale.empty() ? 1.0 : cfg_.inflowScale[b];
            const int off = 2 * N_ * b;
            for (const auto & d : ctx.dirichletVx)
            {
                ids.push_back(d.id + off);
                vals.push_back(d.value * scale);
            }
            for (const auto & d : ctx.dirichletVy)
            {
                ids.push_back(d.id + N_ + off);
                vals.push_back(d.value * scale);
            }
        }
        dirIds_.upload(ids.empty() ? std::vector<int>{0} : ids);
        if (cfg_.dtype == F64)
        {
            dirValues_.upload(reinterpret_cast<const char *>(vals.data()), std::max<size_t>(1, vals.size()) * 8);
        }
        else
        {
            std::vector<float> f(vals.begin(), vals.end());
            f.resize(std::max<size_t>(1, f.size()));
            dirValues_.upload(reinterpret_cast<const char *>(f.data()), f.size() * 4);
        }
        nIntP_ = static_cast<int>(ctx.internalPressureNodes.size());
        internalP_.upload(ctx.internalPressureNodes);
        if (B_ > 1)
        {
            std::vector<int> all;
            all.reserve(static_cast<size_t>(nIntP_) * B_);
            for (int b = 0; b < B_; b++)
            {
                for (int id : ctx.internalPressureNodes)
                {
                    all.push_back(id + b * V_);
                }
            }
            pScatterIds_.upload(all);
        }

        u_.resize(static_cast<size_t>(2) * N_ * B_ * tsize_);
        p_.resize(static_cast<size_t>(V_) * B_ * tsize_);

        auto solverOpts = [&](const ChorinSolverConfig & sc, int nch)
        {
            SolverOptions o = parseSolverName(sc.method);
            if (o.precision == SolverPrecision::Mixed)
            {
                if (cfg_.dtype != F64)
                {
                    throw std::invalid_argument("pcg_mixed needs an FP64 state");
                }
            }
            else
            {
                o.precision = cfg_.dtype == F64 ? SolverPrecision::F64 : SolverPrecision::F32;
            }
            o.numCh = nch;
            o.maxIters = sc.maxIterations;
            o.target = sc.target;
            o.mseCheckInterval = sc.checkInterval;
            o.innerTol = cfg_.innerTol;
            o.lanes = sc.lanes;
            o.rowsPerBlock = sc.rowsPerBlock;
            o.sweepLanes = sc.sweepLanes;
            o.sweepRowsPerBlock = sc.sweepRowsPerBlock;
            return o;
        };
        // One solver system per instance (per-instance colouring and
        // convergence) for batches; the whole matrix otherwise.
        std::vector<const CsrD *> vsys, psys;
        if (ctx.massBlocks.empty())
        {
            vsys.push_back(&ctx.velocityMass);
            psys.push_back(&ctx.pressureStiffnessInternal);
        }
        else
        {
            for (const auto & m : ctx.massBlocks) vsys.push_back(&m);
            for (const auto & m : ctx.pressureBlocks) psys.push_back(&m);
        }
        vSolver_ = std::make_unique<McgsSolver>(vsys, solverOpts(cfg_.velocity, 2 * B_));
        pSolver_ = std::make_unique<McgsSolver>(psys, solverOpts(cfg_.pressure, B_));
        vSolver_->setUseGraphs(cfg_.graphs);
        pSolver_->setUseGraphs(cfg_.graphs);
        for (auto & e : ev_)
        {
            NSFEM_HIP(hipEventCreate(&e));
        }
        resetState();
        (void)s;
    }

    ChorinGpu::~ChorinGpu()
    {
        destroyStepGraph();
        if (dlPending_)
        {
            (void)hipEventSynchronize(copyDone_);
        }
        for (auto & e : ev_)
        {
            (void)hipEventDestroy(e);
        }
        for (hipEvent_t e : {snapDone_, copyDone_})
        {
            if (e)
            {
                (void)hipEventDestroy(e);
            }
        }
        if (copyStream_)
        {
            (void)hipStreamDestroy(copyStream_);
        }
        for (void * h : {hostU_, hostP_})
        {
            if (h)
            {
                (void)hipHostFree(h);
            }
        }
    }

    void ChorinGpu::downloadAsync()
    {
        if (dlPending_)
        {
            throw std::invalid_argument("a state download is still pending (wait for it first)");
        }
        const size_t bu = 2 * tsize_ * static_cast<size_t>(N_) * B_;
        const size_t bp = tsize_ * static_cast<size_t>(V_) * B_;
        if (!copyStream_)
        {
            // set up on first use: snapshot buffers, pinned staging, a copy
            // stream (default priority) and its events
            snapU_.resize(bu);
            snapP_.resize(bp);
            NSFEM_HIP(hipHostMalloc(&hostU_, bu, hipHostMallocDefault));
            NSFEM_HIP(hipHostMalloc(&hostP_, bp, hipHostMallocDefault));
            NSFEM_HIP(hipStreamCreateWithFlags(&copyStream_, hipStreamNonBlocking));
            NSFEM_HIP(hipEventCreateWithFlags(&snapDone_, hipEventDisableTiming));
            NSFEM_HIP(hipEventCreateWithFlags(&copyDone_, hipEventDisableTiming));
        }
        const hipStream_t s = libraryStream();
        // the snapshot is ordered after the last step and before the next
        NSFEM_HIP(hipMemcpyAsync(snapU_.get(), u_.get(), bu, hipMemcpyDeviceToDevice, s));
        NSFEM_HIP(hipMemcpyAsync(snapP_.get(), p_.get(), bp, hipMemcpyDeviceToDevice, s));
        NSFEM_HIP(hipEventRecord(snapDone_, s));
        NSFEM_HIP(hipStreamWaitEvent(copyStream_, snapDone_, 0));
        NSFEM_HIP(hipMemcpyAsync(hostU_, snapU_.get(), bu, hipMemcpyDeviceToHost, copyStream_));
        NSFEM_HIP(hipMemcpyAsync(hostP_, snapP_.get(), bp, hipMemcpyDeviceToHost, copyStream_));
        NSFEM_HIP(hipEventRecord(copyDone_, copyStream_));
        dlPending_ = true;
    }

    void ChorinGpu::downloadWait(void * velocity, void * pressure)
    {
        if (!dlPending_)
        {
            throw std::invalid_argument("no state download is pending");
        }
        NSFEM_HIP(hipEventSynchronize(copyDone_));
        dlPending_ = false;
        if (velocity)
        {
            std::memcpy(velocity, hostU_, 2 * tsize_ * static_cast<size_t>(N_) * B_);
        }
        if (pressure)
        {
            std::memcpy(pressure, hostP_, tsize_ * static_cast<size_t>(V_) * B_);
        }
    }

    double ChorinGpu::diffusionLambda()
    {
        if (lambdaNu_ > 0)
        {
            return lambdaNu_;
        }
        const hipStream_t s = libraryStream();
        const int dt = cfg_.dtype;
        const int nch = vSolver_->numCh();
        const size_t n2 = static_cast<size_t>(nch) * N_;
        // independent seeded starts, one per solver channel
        std::vector<double> h(n2);
        uint64_t st = 0x9E3779B97F4A7C15ull;
        for (auto & v : h)
        {
            st ^= st << 13;
            st ^= st >> 7;
            st ^= st << 17;
            v = static_cast<double>(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        }
        DeviceArray<char> x(n2 * tsize_), w(n2 * tsize_);
        if (dt == F64)
        {
            NSFEM_HIP(hipMemcpyAsync(x.get(), h.data(), n2 * 8, hipMemcpyHostToDevice, s));
        }
        else
        {
            std::vector<float> f(h.begin(), h.end());
            NSFEM_HIP(hipMemcpyAsync(x.get(), f.data(), n2 * 4, hipMemcpyHostToDevice, s));
        }
        const CsrView K = nuK_.view(dt);
        void * rhs = vSolver_->ioRhs();
        void * y = vSolver_->ioSol();
        // Each instance of a batch is normalised and measured on its own rows
        // (K and M are block diagonal): normalised over the whole batch, the
        // instances with the smaller lambda would decay geometrically towards
        // the FP32 underflow of a mixed or FP32 preconditioner.
        const int S = vSolver_->numSystems();
        std::vector<int> off(S + 1);
        for (int k = 0; k < S; k++)
        {
            off[k] = vSolver_->systemOffset(k);
        }
        off[S] = N_;
        auto slice = [&](void * v, int k, int c) { return static_cast<char *>(v) + tsize_ * (static_cast<size_t>(c) * N_ + off[k]); };
        double lam = 0;
        for (int it = 0; it < std::max(1, cfg_.lambdaIterations); it++)
        {
            // x scaled to unit RMS per instance and channel, so absolute
            // ("gs") and relative ("pcg") solver targets both apply
            for (int k = 0; k < S; k++)
            {
                const long n = off[k + 1] - off[k];
                for (int c = 0; c < nch; c++)
                {
                    double nrm = 0;
                    dot(dt, n, 1, slice(x.get(), k, c), n, slice(x.get(), k, c), n, &nrm, s);
                    if (!(nrm > 0) || !std::isfinite(nrm))
                    {
                        throw NonFiniteError("lambda_max estimate: the power iteration vector vanished");
                    }
                    scale(dt, n, std::sqrt(n / nrm), slice(x.get(), k, c), s);
                }
            }
            csrMul(K, nch, x.get(), N_, rhs, N_, 1.0, 0.0, nullptr, N_, 0, s);
            vSolver_->solve(true);
            csrMul(K, nch, y, N_, w.get(), N_, 1.0, 0.0, nullptr, N_, 0, s);
            lam = 0;
            for (int k = 0; k < S; k++)
            {
                const long n = off[k + 1] - off[k];
                for (int c = 0; c < nch; c++)
                {
                    double yKy = 0, yMy = 0;
                    dot(dt, n, 1, slice(y, k, c), n, slice(w.get(), k, c), n, &yKy, s);
                    dot(dt, n, 1, slice(y, k, c), n, slice(rhs, k, c), n, &yMy, s); // M y = nu K x
                    if (yMy > 0)
                    {
                        lam = std::max(lam, yKy / yMy);
                    }
                }
            }
            NSFEM_HIP(hipMemcpyAsync(x.get(), y, n2 * tsize_, hipMemcpyDeviceToDevice, s));
        }
        syncLibraryStream();
        if (!(lam > 0) || !std::isfinite(lam))
        {
            throw NonFiniteError("lambda_max estimate failed (non-positive or non-finite)");
        }
        lambdaNu_ = lam;
        return lambdaNu_;
    }

    void ChorinGpu::setTimeStep(double tau)
    {
        if (!(tau > 0) || !std::isfinite(tau))
        {
            throw std::invalid_argument("time step must be positive");
        }
        double used = tau;
        if (cfg_.timeStepPolicy != TimeStepPolicy::Unchecked)
        {
            const double tmax = stableTimeStep();
            if (tau > tmax)
            {
                if (cfg_.timeStepPolicy != TimeStepPolicy::Clamp)
                {
                    // rejected: the previous step stays in force
                    char msg[320];
                    std::snprintf(msg, sizeof msg,
                                  "time step %.6g exceeds the explicit scheme's stability limit %.6g (tau nu lambda_max(M^-1 K) = "
                                  "%.3f > 1); use a smaller step or the clamp policy",
                                  tau, tmax, tau * diffusionLambda());
                    throw std::invalid_argument(msg);
                }
                used = tmax;
            }
        }
        tauRequested_ = tau;
        tau_ = used;
        tauChecked_ = true;
    }

    void ChorinGpu::imposeDirichlet(hipStream_t s) { scatter(cfg_.dtype, nDir_, dirIds_.get(), dirValues_.get(), u_.get(), s); }

    void ChorinGpu::resetState()
    {
        const hipStream_t s = libraryStream();
        fill(cfg_.dtype, 2L * N_ * B_, 0.0, u_.get(), s);
        fill(cfg_.dtype, static_cast<long>(V_) * B_, 0.0, p_.get(), s);
        imposeDirichlet(s);
        fill(pSolver_->ioDtype(), static_cast<long>(pSolver_->n()) * pSolver_->numCh(), 0.0, pSolver_->ioSol(), s);
        if (cfg_.velocityStart != 0)
        {
            const long n = static_cast<long>(vSolver_->n()) * vSolver_->numCh();
            const size_t bytes = static_cast<size_t>(n) * dtypeSize(vSolver_->ioDtype());
            for (auto & kind : vHist_)
            {
                for (auto & h : kind)
                {
                    if (h.size() < bytes)
                    {
                        h.resize(bytes);
                    }
                    fill(vSolver_->ioDtype(), n, 0.0, h.get(), s);
                }
            }
        }
        if (cfg_.pressureExtrapolation)
        {
            const long n = static_cast<long>(pSolver_->n()) * pSolver_->numCh();
            if (pPrev_.size() < static_cast<size_t>(n) * dtypeSize(pSolver_->ioDtype()))
            {
                pPrev_.resize(std::max<size_t>(8, static_cast<size_t>(n) * dtypeSize(pSolver_->ioDtype())));
            }
            fill(pSolver_->ioDtype(), n, 0.0, pPrev_.get(), s);
        }
        syncLibraryStream();
    }

    void ChorinGpu::assembleConvection(const void * u, void * aOut, void * cOut, hipStream_t s)
    {
        const int gE = ceilDiv(E_, 256);
        const int gN = ceilDiv(A_.nnz, 256);
        if (cfg_.dtype == F64)
        {
            convectionElementKernel<double><<<gE, 256, 0, s>>>(E_, elemNodes_.get(), elemGeom_.get(), static_cast<const double *>(u), N_,
                                                               reinterpret_cast<double *>(elemC_.get()));
            convectionGatherKernel<double><<<gN, 256, 0, s>>>(A_.nnz, contribStart_.get(), contribIndex_.get(),
                                                              reinterpret_cast<const double *>(elemC_.get()),
                                                              reinterpret_cast<const double *>(nuK_.values.get()), static_cast<double *>(aOut),
                                                              static_cast<double *>(cOut));
        }
        else
        {
            convectionElementKernel<float><<<gE, 256, 0, s>>>(E_, elemNodes_.get(), elemGeom_.get(), static_cast<const float *>(u), N_,
                                                              reinterpret_cast<float *>(elemC_.get()));
            convectionGatherKernel<float><<<gN, 256, 0, s>>>(A_.nnz, contribStart_.get(), contribIndex_.get(),
                                                             reinterpret_cast<const float *>(elemC_.get()),
                                                             reinterpret_cast<const float *>(nuK_.values.get()), static_cast<float *>(aOut),
                                                             static_cast<float *>(cOut));
        }
        NSFEM_HIP(hipGetLastError());
    }

    double ChorinGpu::timeKernel(const std::string & which, int reps, double * bytesPerLaunch, double * launchesPerIteration)
    {
        if (which.rfind("velocity_", 0) == 0)
        {
            return vSolver_->timeKernel(which.substr(9), reps, bytesPerLaunch, launchesPerIteration);
        }
        if (which.rfind("pressure_", 0) == 0)
        {
            return pSolver_->timeKernel(which.substr(9), reps, bytesPerLaunch, launchesPerIteration);
        }
        if (reps < 1)
        {
            throw std::invalid_argument("reps must be positive");
        }
        const hipStream_t s = libraryStream();
        const double ts = static_cast<double>(tsize_);
        const double nnz = A_.nnz, E = E_, N = N_;
        double bytes = 0;
        std::function<void()> launch;
        if (which == "convection")
        {
            bytes = E * (6 * 4 + 5 * 8 + 12 * ts + 36 * ts) + nnz * (4 + 2 * ts) + 36 * E * (4 + ts);
            launch = [&]() { assembleConvection(u_.get(), A_.values.get(), nullptr, s); };
        }
        else if (which == "spmm")
        {
            bytes = nnz * (ts + 4) + (N + 1) * 4 + 4 * N * ts;
            launch = [&]()
            { csrMul(A_.view(cfg_.dtype), 2, u_.get(), N_, vSolver_->ioRhs(), N_, 1.0, 0.0, nullptr, N_, 0, s); };
        }
        else
        {
            throw std::invalid_argument("unknown kernel '" + which + "'");
        }
        launch();
        NSFEM_HIP(hipStreamSynchronize(s));
        delayKernel<<<1, 64, 0, s>>>(2000000); // the repetitions queue behind it and run back to back
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
        if (launchesPerIteration)
        {
            *launchesPerIteration = 1; // once per step
        }
        return static_cast<double>(ms) / reps;
    }

    namespace
    {
        // Worst instance of a (possibly batched) solve over the channels
        // c0, c0 + stride, ... (an ensemble's channels of one component).
        double maxMse(const McgsSolver & s, int c0, int stride = 1)
        {
            double m = s.lastMse(0, c0);
            for (int c = c0; c < s.numCh(); c += stride)
            {
                for (int k = 0; k < s.numSystems(); k++)
                {
                    const double v = s.lastMse(k, c);
                    m = (v != v || v > m) ? v : m;
                }
            }
            return m;
        }
        double maxRel(const McgsSolver & s)
        {
            double m = -1;
            for (int k = 0; k < s.numSystems(); k++)
            {
                for (int c = 0; c < s.numCh(); c++)
                {
                    m = std::max(m, s.lastRelRes(k, c));
                }
            }
            return m;
        }
        int maxIters(const McgsSolver & s)
        {
            int m = 0;
            for (int k = 0; k < s.numSystems(); k++)
            {
                m = std::max(m, s.lastIterations(k));
            }
            return m;
        }
    } // namespace

    void ChorinGpu::launchConvectionRhs(hipStream_t s)
    {
        // per instance: A_b = nu K + C(u_b), then rhs_b = A_b u_b (the
        // product's lanes per row follow its mean row length)
        const int dt = cfg_.dtype;
        const size_t uStride = 2 * static_cast<size_t>(N_) * tsize_;
        const size_t aStride = static_cast<size_t>(A_.nnz) * tsize_;
        for (int b = 0; b < B_; b++)
        {
            const char * u = static_cast<const char *>(u_.get()) + b * uStride;
            char * a = static_cast<char *>(A_.values.get()) + b * aStride;
            char * rhs = static_cast<char *>(vSolver_->ioRhs()) + b * uStride;
            assembleConvection(u, a, nullptr, s);
            CsrView Ab = A_.view(dt);
            Ab.values = a;
            csrMul(Ab, 2, u, N_, rhs, N_, 1.0, 0.0, nullptr, N_, 0, s);
        }
    }

    void ChorinGpu::launchPressureRhs(hipStream_t s)
    {
        // rhs_b = -(1/tau) (VPD u_b)[internal]: instance b is pressure channel b
        csrMul(vpd_.view(cfg_.dtype), B_, u_.get(), 2L * N_, pSolver_->ioRhs(), nIntP_, -1.0 / tau_, 0.0, internalP_.get(), nIntP_, 0, s);
    }

    void ChorinGpu::startPressure(hipStream_t s)
    {
        if (cfg_.pressureExtrapolation)
        {
            extrapolate(pSolver_->ioDtype(), static_cast<long>(pSolver_->n()) * pSolver_->numCh(), pSolver_->ioSol(), pPrev_.get(), s);
        }
    }

    bool ChorinGpu::startVelocity(int kind, hipStream_t s)
    {
        const long n = static_cast<long>(vSolver_->n()) * vSolver_->numCh();
        const int dt = vSolver_->ioDtype();
        if (cfg_.velocityStart == 1)
        {
            NSFEM_HIP(hipMemcpyAsync(vSolver_->ioSol(), vHist_[kind][0].get(), static_cast<size_t>(n) * dtypeSize(dt),
                                     hipMemcpyDeviceToDevice, s));
        }
        else if (cfg_.velocityStart == 2)
        {
            extrapolateTo(dt, n, vHist_[kind][0].get(), vHist_[kind][1].get(), vSolver_->ioSol(), s);
        }
        return cfg_.velocityStart == 0;
    }

    void ChorinGpu::keepVelocity(int kind, hipStream_t s)
    {
        if (cfg_.velocityStart != 0)
        {
            const long n = static_cast<long>(vSolver_->n()) * vSolver_->numCh();
            NSFEM_HIP(hipMemcpyAsync(vHist_[kind][0].get(), vSolver_->ioSol(), static_cast<size_t>(n) * dtypeSize(vSolver_->ioDtype()),
                                     hipMemcpyDeviceToDevice, s));
        }
    }

    void ChorinGpu::launchScatterGradient(hipStream_t s)
    {
        // p_b[internal] = solution channel b, then rhs_b = PVD p_b
        const int dt = cfg_.dtype;
        scatter(dt, nIntP_ * B_, B_ > 1 ? pScatterIds_.get() : internalP_.get(), pSolver_->ioSol(), p_.get(), s);
        csrMul(pvd_.view(dt), B_, p_.get(), V_, vSolver_->ioRhs(), 2L * N_, 1.0, 0.0, nullptr, 2 * N_, 0, s);
    }

    void ChorinGpu::stepRegular(ChorinStepStats * stats)
    {
        const hipStream_t s = libraryStream();
        const int dt = cfg_.dtype;
        const long twoN = 2L * N_;
        ChorinStepStats st{};
        NSFEM_HIP(hipEventRecord(ev_[0], s));

        // Convection: A = nu K + C(u) and the tentative right-hand side A u
        launchConvectionRhs(s);
        NSFEM_HIP(hipEventRecord(ev_[1], s));

        // Tentative velocity: M a = A u;  u -= tau a
        vSolver_->solve(startVelocity(0, s));
        keepVelocity(0, s);
        st.mseTentX = maxMse(*vSolver_, 0, 2);
        st.mseTentY = maxMse(*vSolver_, 1, 2);
        st.relTent = maxRel(*vSolver_);
        st.itersTent = maxIters(*vSolver_);
        st.outerTent = vSolver_->lastOuterIterations();
        if (!std::isfinite(st.mseTentX) || !std::isfinite(st.mseTentY) || vSolver_->lastNonFinite())
        {
            throw NonFiniteError("Tentative acceleration is bad");
        }
        axpy(dt, twoN * B_, -tau_, vSolver_->ioSol(), u_.get(), s);
        imposeDirichlet(s);
        NSFEM_HIP(hipEventRecord(ev_[2], s));

        // Pressure: Kp_int p = -(1/tau) (VPD u)[internal], warm start
        launchPressureRhs(s);
        startPressure(s);
        pSolver_->solve();
        st.msePressure = maxMse(*pSolver_, 0);
        st.relPressure = maxRel(*pSolver_);
        st.itersPressure = maxIters(*pSolver_);
        st.outerPressure = pSolver_->lastOuterIterations();
        NSFEM_HIP(hipEventRecord(ev_[3], s));

        // Projection: M a = PVD p;  u -= tau a
        launchScatterGradient(s);
        vSolver_->solve(startVelocity(1, s));
        keepVelocity(1, s);
        st.mseFinalX = maxMse(*vSolver_, 0, 2);
        st.mseFinalY = maxMse(*vSolver_, 1, 2);
        st.relFinal = maxRel(*vSolver_);
        st.itersFinal = maxIters(*vSolver_);
        st.outerFinal = vSolver_->lastOuterIterations();
        if (!std::isfinite(st.mseFinalX) || !std::isfinite(st.mseFinalY) || vSolver_->lastNonFinite())
        {
            throw NonFiniteError("Final acceleration is bad");
        }
        axpy(dt, twoN * B_, -tau_, vSolver_->ioSol(), u_.get(), s);
        imposeDirichlet(s);
        NSFEM_HIP(hipEventRecord(ev_[4], s));
        NSFEM_HIP(hipEventSynchronize(ev_[4]));
        float ms[4];
        for (int i = 0; i < 4; i++)
        {
            NSFEM_HIP(hipEventElapsedTime(&ms[i], ev_[i], ev_[i + 1]));
        }
        st.msConvection = ms[0];
        st.msTentative = ms[1];
        st.msPressure = ms[2];
        st.msFinal = ms[3];
        st.msTotal = ms[0] + ms[1] + ms[2] + ms[3];
        st.tauRequested = tauRequested_;
        st.tauUsed = tau_;
        st.capturedStep = 0;
        if (stats)
        {
            *stats = st;
        }
    }

    void ChorinGpu::setStepGraph(bool on)
    {
        cfg_.stepGraph = on;
        destroyStepGraph();
        budget_[0] = budget_[1] = budget_[2] = 0;
    }

    void ChorinGpu::setStepGraphBudgets(int tent, int pres, int fin)
    {
        if (tent < 1 || pres < 1 || fin < 1)
        {
            throw std::invalid_argument("iteration budgets must be positive");
        }
        budget_[0] = tent;
        budget_[1] = pres;
        budget_[2] = fin;
        destroyStepGraph();
    }

    void ChorinGpu::destroyStepGraph()
    {
        if (stepExec_)
        {
            (void)hipGraphExecDestroy(stepExec_);
            stepExec_ = nullptr;
        }
    }

    void ChorinGpu::setBudgets(int tent, int pres, int fin)
    {
        // A measured count c gets c + m iterations, m = max(3, c / 4) for the
        // short velocity solves and max(16, c / 24) for a one-level pressure
        // solve (~2,000 iterations at 500k; per-instance counts of a batch
        // drift by several), max(3, c / 24) for the V-cycle's ~20.  The graph
        // is rebuilt when less than half a margin is left or more than two
        // margins (and a slack of 8 iterations, 2 for the V-cycle, whose idle
        // iterations are ~50 launches each) are idle.
        const int counts[3] = {tent, pres, fin};
        const bool amgPressure = pSolver_->method() == SolverMethod::AmgPcg;
        const int minMargin[3] = {3, amgPressure ? 3 : 16, 3};
        const int slack[3] = {8, amgPressure ? 2 : 8, 8};
        const int frac[3] = {4, 24, 4};
        bool rebuild = false;
        for (int k = 0; k < 3; k++)
        {
            const int c = counts[k];
            const int m = std::max(minMargin[k], c / frac[k]);
            if (c + m / 2 > budget_[k] || budget_[k] > c + 2 * m + slack[k])
            {
                budget_[k] = c + m;
                rebuild = true;
            }
        }
        if (rebuild)
        {
            destroyStepGraph();
        }
    }

    void ChorinGpu::buildStepGraph()
    {
        const hipStream_t s = libraryStream();
        const int dt = cfg_.dtype;
        const long twoN = 2L * N_;
        const size_t pBytes = static_cast<size_t>(pSolver_->n()) * pSolver_->numCh() * dtypeSize(pSolver_->ioDtype());
        const size_t uBytes = static_cast<size_t>(twoN) * B_ * tsize_;
        if (stepSnapU_.size() < uBytes)
        {
            stepSnapU_.resize(uBytes);
        }
        if (stepSnapP_.size() < pBytes)
        {
            stepSnapP_.resize(std::max<size_t>(pBytes, 8));
        }
        syncLibraryStream();
        const auto t0 = std::chrono::steady_clock::now();
        hipGraph_t graph = nullptr;
        NSFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try
        {
            // snapshot for a rollback
            NSFEM_HIP(hipMemcpyAsync(stepSnapU_.get(), u_.get(), uBytes, hipMemcpyDeviceToDevice, s));
            NSFEM_HIP(hipMemcpyAsync(stepSnapP_.get(), pSolver_->ioSol(), pBytes, hipMemcpyDeviceToDevice, s));
            launchConvectionRhs(s);
            vSolver_->enqueueSolve(s, startVelocity(0, s), budget_[0]);
            keepVelocity(0, s);
            vSolver_->snapshotStats(s, 0);
            axpy(dt, twoN * B_, -tau_, vSolver_->ioSol(), u_.get(), s);
            imposeDirichlet(s);
            launchPressureRhs(s);
            startPressure(s);
            pSolver_->enqueueSolve(s, false, budget_[1]);
            pSolver_->snapshotStats(s, 1);
            launchScatterGradient(s);
            vSolver_->enqueueSolve(s, startVelocity(1, s), budget_[2]);
            keepVelocity(1, s);
            vSolver_->snapshotStats(s, 2);
            axpy(dt, twoN * B_, -tau_, vSolver_->ioSol(), u_.get(), s);
            imposeDirichlet(s);
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g)
            {
                (void)hipGraphDestroy(g);
            }
            throw;
        }
        NSFEM_HIP(hipStreamEndCapture(s, &graph));
        const hipError_t rc = hipGraphInstantiate(&stepExec_, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        NSFEM_HIP(rc);
        graphTau_ = tau_;
        graphBuilds_++;
        graphBuildSeconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    bool ChorinGpu::stepCaptured(ChorinStepStats * stats)
    {
        const hipStream_t s = libraryStream();
        if (stepExec_ && graphTau_ != tau_)
        {
            destroyStepGraph(); // tau is a parameter of the captured kernels
        }
        if (!stepExec_)
        {
            buildStepGraph();
        }
        NSFEM_HIP(hipEventRecord(ev_[0], s));
        NSFEM_HIP(hipGraphLaunch(stepExec_, s));
        NSFEM_HIP(hipEventRecord(ev_[4], s));
        NSFEM_HIP(hipEventSynchronize(ev_[4]));
        ChorinStepStats st{};
        const bool okT = vSolver_->collectStats(0);
        st.mseTentX = maxMse(*vSolver_, 0, 2);
        st.mseTentY = maxMse(*vSolver_, 1, 2);
        st.relTent = maxRel(*vSolver_);
        st.itersTent = maxIters(*vSolver_);
        const bool badT = !std::isfinite(st.mseTentX) || !std::isfinite(st.mseTentY) || vSolver_->lastNonFinite();
        const bool okP = pSolver_->collectStats(1);
        st.msePressure = maxMse(*pSolver_, 0);
        st.relPressure = maxRel(*pSolver_);
        st.itersPressure = maxIters(*pSolver_);
        const bool okF = vSolver_->collectStats(2);
        st.mseFinalX = maxMse(*vSolver_, 0, 2);
        st.mseFinalY = maxMse(*vSolver_, 1, 2);
        st.relFinal = maxRel(*vSolver_);
        st.itersFinal = maxIters(*vSolver_);
        const bool badF = !std::isfinite(st.mseFinalX) || !std::isfinite(st.mseFinalY) || vSolver_->lastNonFinite();
        if (badT || badF || !(okT && okP && okF))
        {
            // back to the state before the step (the regular path's state at
            // its error, or the start of the rerun)
            const size_t pBytes = static_cast<size_t>(pSolver_->n()) * pSolver_->numCh() * dtypeSize(pSolver_->ioDtype());
            NSFEM_HIP(hipMemcpyAsync(u_.get(), stepSnapU_.get(), 2 * static_cast<size_t>(N_) * B_ * tsize_, hipMemcpyDeviceToDevice, s));
            NSFEM_HIP(hipMemcpyAsync(pSolver_->ioSol(), stepSnapP_.get(), pBytes, hipMemcpyDeviceToDevice, s));
            syncLibraryStream();
            if (badT)
            {
                throw NonFiniteError("Tentative acceleration is bad");
            }
            if (badF && okT && okP)
            {
                throw NonFiniteError("Final acceleration is bad");
            }
            graphRollbacks_++;
            return false;
        }
        float ms = 0;
        NSFEM_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[4]));
        st.msTotal = ms;
        st.tauRequested = tauRequested_;
        st.tauUsed = tau_;
        st.capturedStep = 1;
        setBudgets(st.itersTent, st.itersPressure, st.itersFinal);
        if (stats)
        {
            *stats = st;
        }
        return true;
    }

    void ChorinGpu::step(ChorinStepStats * stats)
    {
        if (!tauChecked_)
        {
            setTimeStep(tau_); // the default step goes through the policy too
        }
        if (cfg_.stepGraph && budget_[1] > 0 && stepCaptured(stats))
        {
            return;
        }
        ChorinStepStats st{};
        stepRegular(&st);
        if (cfg_.stepGraph)
        {
            setBudgets(st.itersTent, st.itersPressure, st.itersFinal);
        }
        if (stats)
        {
            *stats = st;
        }
    }
} // namespace nsfem
