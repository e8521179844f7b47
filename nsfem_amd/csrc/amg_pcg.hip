#include "amg_pcg.hpp"

#include <chrono>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <type_traits>

#include "ops.hpp"

namespace nsfem
{
    namespace
    {
        struct AmgState
        {
            double * scal; // rho, rho of the iteration before, |b|^2, |r|^2
            int * ints;    // stop (0 running, 1 converged, 2 non-finite or breakdown), iterations, zero right-hand side
            double * partPq;
            double * partRr;
            double * partRho;
            double * partBb;
            int parts;
            double tol2;
        };

        // ---------------- V-cycle ----------------

        // One colour of a multicolour Gauss-Seidel sweep: rows [rowBegin,
        // rowEnd) (one colour, independent of each other) take
        // x_i = (b_i - sum_j a_ij x_j) / a_ii over the lower-colour part of
        // the row (lowerOnly: the first sweep from x = 0, where the higher
        // colours are still zero) or the whole row.  L lanes per row.
        template <typename T, int L>
        __global__ void __launch_bounds__(kBlock) amgSweepKernel(McgsMatrix<T> A, int rowBegin, int rowEnd, const T * __restrict__ b,
                                                                 T * __restrict__ x, int lowerOnly, const int * __restrict__ stop)
        {
            if (*stop)
            {
                return;
            }
            const int t = blockIdx.x * kBlock + threadIdx.x;
            const int row = rowBegin + t / L;
            const int lane = t % L;
            T s = 0;
            if (row < rowEnd)
            {
                const int j1 = lowerOnly ? A.rowSplit[row] : A.rowStart[row + 1];
                for (int j = A.rowStart[row] + lane; j < j1; j += L)
                {
                    s = fmaT(A.values[j], x[A.column[j]], s);
                }
            }
            s = laneGroupSum<L>(s);
            if (row < rowEnd && lane == 0)
            {
                x[row] = (b[row] - s) * A.invDiag[row];
            }
        }

        // r = b - A x (A = stripped + diagonal).
        template <typename T, int L>
        __global__ void __launch_bounds__(kBlock) amgResidualKernel(McgsMatrix<T> A, const T * __restrict__ b, const T * __restrict__ x,
                                                                    T * __restrict__ r, const int * __restrict__ stop)
        {
            if (*stop)
            {
                return;
            }
            const int t = blockIdx.x * kBlock + threadIdx.x;
            const int row = t / L;
            const int lane = t % L;
            T s = 0;
            if (row < A.n)
            {
                for (int j = A.rowStart[row] + lane; j < A.rowStart[row + 1]; j += L)
                {
                    s = fmaT(A.values[j], x[A.column[j]], s);
                }
            }
            s = laneGroupSum<L>(s);
            if (row < A.n && lane == 0)
            {
                r[row] = b[row] - fmaT(A.diag[row], x[row], s);
            }
        }

        // y = M x (accumulate: y += M x) for the restriction and prolongation.
        template <typename T, int L>
        __global__ void __launch_bounds__(kBlock) amgSpmvKernel(int rows, const int * __restrict__ rowStart, const int * __restrict__ column,
                                                                const T * __restrict__ values, const T * __restrict__ x, T * __restrict__ y,
                                                                int accumulate, const int * __restrict__ stop)
        {
            if (*stop)
            {
                return;
            }
            const int t = blockIdx.x * kBlock + threadIdx.x;
            const int row = t / L;
            const int lane = t % L;
            T s = 0;
            if (row < rows)
            {
                for (int j = rowStart[row] + lane; j < rowStart[row + 1]; j += L)
                {
                    s = fmaT(values[j], x[column[j]], s);
                }
            }
            s = laneGroupSum<L>(s);
            if (row < rows && lane == 0)
            {
                y[row] = accumulate ? y[row] + s : s;
            }
        }

        // x = Ainv b on the coarsest level (dense, row-major): one wavefront per row.
        template <typename T>
        __global__ void __launch_bounds__(kBlock) amgDenseKernel(int n, const T * __restrict__ inv, const T * __restrict__ b,
                                                                 T * __restrict__ x, const int * __restrict__ stop)
        {
            if (*stop)
            {
                return;
            }
            const int row = (blockIdx.x * kBlock + threadIdx.x) >> 6;
            const int lane = threadIdx.x & 63;
            T s = 0;
            if (row < n)
            {
                const T * a = inv + static_cast<size_t>(row) * n;
                for (int j = lane; j < n; j += 64)
                {
                    s = fmaT(a[j], b[j], s);
                }
            }
            s = laneGroupSum<64>(s);
            if (row < n && lane == 0)
            {
                x[row] = s;
            }
        }

        // ---------------- CG around it ----------------

        // Sum of parts partials in a fixed order (the same in every block).
        __device__ __forceinline__ double sumParts(const double * __restrict__ parts, int count, double * sm)
        {
            double v[1] = {0};
            for (int k = threadIdx.x; k < count; k += kBlock)
            {
                v[0] += parts[k];
            }
            blockSum<1>(v, sm);
            return v[0]; // thread 0
        }

        // r = b - A x (x = 0 and r = b for a zero guess), the V-cycle's copy
        // of r, partials of |b|^2 and |r|^2.
        template <typename P, int L>
        __global__ void __launch_bounds__(kBlock) amgInitKernel(McgsMatrix<double> A, const double * __restrict__ b, double * __restrict__ x,
                                                                double * __restrict__ r, P * __restrict__ rP, int zeroGuess, AmgState st)
        {
            __shared__ double sm[2 * kWaves];
            constexpr int G = kBlock / L;
            const int lane = threadIdx.x % L;
            double acc[2] = {0, 0};
            for (int base = blockIdx.x * G; base < A.n; base += gridDim.x * G)
            {
                const int row = base + static_cast<int>(threadIdx.x) / L;
                double s = 0;
                if (row < A.n && !zeroGuess)
                {
                    for (int j = A.rowStart[row] + lane; j < A.rowStart[row + 1]; j += L)
                    {
                        s = fmaT(A.values[j], x[A.column[j]], s);
                    }
                }
                s = laneGroupSum<L>(s);
                if (row < A.n && lane == 0)
                {
                    const double bi = b[row];
                    double ri = bi;
                    if (zeroGuess)
                    {
                        x[row] = 0;
                    }
                    else
                    {
                        ri = bi - fmaT(A.diag[row], x[row], s);
                    }
                    r[row] = ri;
                    if constexpr (!std::is_same_v<P, double>)
                    {
                        rP[row] = static_cast<P>(ri);
                    }
                    acc[0] = fmaT(bi, bi, acc[0]);
                    acc[1] = fmaT(ri, ri, acc[1]);
                }
            }
            blockSum<2>(acc, sm);
            if (threadIdx.x == 0)
            {
                st.partBb[blockIdx.x] = acc[0];
                st.partRr[blockIdx.x] = acc[1];
            }
        }

        __global__ void __launch_bounds__(kBlock) amgInitFinalizeKernel(AmgState st)
        {
            __shared__ double sm[kWaves];
            const double bb = sumParts(st.partBb, st.parts, sm);
            const double rr = sumParts(st.partRr, st.parts, sm);
            if (threadIdx.x == 0)
            {
                st.scal[0] = 0;
                st.scal[1] = 0;
                st.scal[2] = bb;
                st.scal[3] = rr;
                st.ints[1] = 0;
                st.ints[2] = bb == 0 ? 1 : 0;
                if (!std::isfinite(bb) || !std::isfinite(rr))
                {
                    st.ints[0] = 2;
                }
                else
                {
                    st.ints[0] = (bb == 0 || rr <= st.tol2 * bb) ? 1 : 0;
                }
            }
        }

        // beta = rho / rho_before (0 on the first direction); p = z + beta p,
        // q = A z + beta q (= A p, as A p_before = q), partials of p.q.
        template <typename P, int L>
        __global__ void __launch_bounds__(kBlock) amgDirectionKernel(McgsMatrix<double> A, const P * __restrict__ z, double * __restrict__ p,
                                                                     double * __restrict__ q, AmgState st)
        {
            if (st.ints[0])
            {
                return;
            }
            __shared__ double sm[kWaves];
            const double rho = st.scal[0], before = st.scal[1];
            const double beta = before != 0 ? rho / before : 0.0;
            constexpr int G = kBlock / L;
            const int lane = threadIdx.x % L;
            double acc[1] = {0};
            for (int base = blockIdx.x * G; base < A.n; base += gridDim.x * G)
            {
                const int row = base + static_cast<int>(threadIdx.x) / L;
                double s = 0;
                if (row < A.n)
                {
                    for (int j = A.rowStart[row] + lane; j < A.rowStart[row + 1]; j += L)
                    {
                        s = fmaT(A.values[j], static_cast<double>(z[A.column[j]]), s);
                    }
                }
                s = laneGroupSum<L>(s);
                if (row < A.n && lane == 0)
                {
                    const double zi = static_cast<double>(z[row]);
                    const double pi = fmaT(beta, p[row], zi);
                    const double qi = fmaT(beta, q[row], fmaT(A.diag[row], zi, s));
                    p[row] = pi;
                    q[row] = qi;
                    acc[0] = fmaT(pi, qi, acc[0]);
                }
            }
            blockSum<1>(acc, sm);
            if (threadIdx.x == 0)
            {
                st.partPq[blockIdx.x] = acc[0];
            }
        }

        // alpha = rho / p.q (every block sums the partials itself, in the same
        // order); x += alpha p, r -= alpha q, the V-cycle's copy of r,
        // partials of |r|^2.
        template <typename P>
        __global__ void __launch_bounds__(kBlock) amgUpdateKernel(int n, const double * __restrict__ p, const double * __restrict__ q,
                                                                  double * __restrict__ x, double * __restrict__ r, P * __restrict__ rP,
                                                                  AmgState st)
        {
            if (st.ints[0])
            {
                return;
            }
            __shared__ double sm[kWaves];
            __shared__ double shAlpha;
            const double pq = sumParts(st.partPq, st.parts, sm);
            if (threadIdx.x == 0)
            {
                shAlpha = st.scal[0] / pq;
            }
            __syncthreads();
            const double alpha = shAlpha;
            double acc[1] = {0};
            for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
            {
                x[i] = fmaT(alpha, p[i], x[i]);
                const double ri = fmaT(-alpha, q[i], r[i]);
                r[i] = ri;
                if constexpr (!std::is_same_v<P, double>)
                {
                    rP[i] = static_cast<P>(ri);
                }
                acc[0] = fmaT(ri, ri, acc[0]);
            }
            blockSum<1>(acc, sm);
            if (threadIdx.x == 0)
            {
                st.partRr[blockIdx.x] = acc[0];
            }
        }

        // partials of r.z
        template <typename P>
        __global__ void __launch_bounds__(kBlock) amgRhoKernel(int n, const double * __restrict__ r, const P * __restrict__ z, AmgState st)
        {
            if (st.ints[0])
            {
                return;
            }
            __shared__ double sm[kWaves];
            double acc[1] = {0};
            for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
            {
                acc[0] = fmaT(r[i], static_cast<double>(z[i]), acc[0]);
            }
            blockSum<1>(acc, sm);
            if (threadIdx.x == 0)
            {
                st.partRho[blockIdx.x] = acc[0];
            }
        }

        // rho = r.z; after an update (init = 0) also |r|^2, the iteration
        // count and the stopping test.  A preconditioned residual with
        // r.z <= 0 (a breakdown) stops the solve as non-finite.
        __global__ void __launch_bounds__(kBlock) amgFinalizeKernel(AmgState st, int init)
        {
            if (st.ints[0])
            {
                return;
            }
            __shared__ double sm[kWaves];
            const double rho = sumParts(st.partRho, st.parts, sm);
            const double rr = init ? 0.0 : sumParts(st.partRr, st.parts, sm);
            if (threadIdx.x == 0)
            {
                if (st.tol2 < 0)
                {
                    // kernel timing: never stops, whatever the vectors hold
                    st.scal[1] = st.scal[0];
                    st.scal[0] = std::isfinite(rho) && rho > 0 ? rho : 1.0;
                    return;
                }
                if (!init)
                {
                    st.scal[3] = rr;
                    st.ints[1] += 1;
                    if (!std::isfinite(rr))
                    {
                        st.ints[0] = 2;
                        return;
                    }
                    if (rr <= st.tol2 * st.scal[2])
                    {
                        st.ints[0] = 1;
                        return;
                    }
                }
                if (!std::isfinite(rho) || !(rho > 0))
                {
                    st.ints[0] = 2;
                    return;
                }
                st.scal[1] = init ? 0.0 : st.scal[0];
                st.scal[0] = rho;
            }
        }

        template <typename F>
        void byLanes(int lanes, F && f)
        {
            switch (lanes)
            {
            case 2: f.template operator()<2>(); break;
            case 4: f.template operator()<4>(); break;
            case 8: f.template operator()<8>(); break;
            case 16: f.template operator()<16>(); break;
            default: throw std::invalid_argument("amg: lanes must be 2, 4, 8 or 16");
            }
        }

        int lanesFor(double meanRow) { return std::max(2, std::min(16, lanesForMeanRow(meanRow))); }
        double meanRow(const std::vector<int> & rowStart)
        {
            const int rows = static_cast<int>(rowStart.size()) - 1;
            return rows > 0 ? static_cast<double>(rowStart.back()) / rows : 1.0;
        }
    } // namespace

    AmgPcg::AmgPcg(const CsrD & stripped, const std::vector<int> & rowSplit, const std::vector<double> & diag,
                   const std::vector<int> & colorStart, bool mixed, McgsMatrix<double> a64, McgsMatrix<float> level0F32, hipStream_t s)
        : mixed_(mixed), n_(stripped.rows), a64_(a64), a32_(level0F32), stream_(s)
    {
        const auto t0 = std::chrono::steady_clock::now();
        const AmgHierarchy h = buildAmgHierarchy(stripped, rowSplit, diag, colorStart);
        complexity_ = h.operatorComplexity;
        levels_.resize(h.levels.size());
        const size_t ts = mixed_ ? 4 : 8;
        for (size_t l = 0; l < h.levels.size(); l++)
        {
            const AmgLevel & H = h.levels[l];
            Level & L = levels_[l];
            L.n = H.n;
            L.colorStart = H.colorStart;
            L.lanes = lanesFor(meanRow(H.stripped.rowStart));
            L.nnz = H.stripped.nnz();
            for (int i = 0; i < H.n; i++)
            {
                L.nnzLower += H.rowSplit[i] - H.stripped.rowStart[i];
            }
            L.nnzP = H.P.nnz();
            L.nnzR = H.R.nnz();
            if (l > 0)
            {
                L.rowStart.upload(H.stripped.rowStart);
                L.rowSplit.upload(H.rowSplit);
                L.column.upload(H.stripped.column.empty() ? std::vector<int>{0} : H.stripped.column);
                const std::vector<double> & v = H.stripped.values.empty() ? std::vector<double>{0.0} : H.stripped.values;
                if (mixed_)
                {
                    std::vector<float> f(v.begin(), v.end()), d(H.diag.begin(), H.diag.end()), inv(H.n);
                    for (int i = 0; i < H.n; i++)
                    {
                        inv[i] = 1.0f / d[i];
                    }
                    L.values32.upload(f);
                    L.diag32.upload(d);
                    L.invDiag32.upload(inv);
                }
                else
                {
                    L.values64.upload(v);
                    L.diag64.upload(H.diag);
                    L.invDiag64.upload(H.invDiag);
                }
                L.b.resize(H.n * ts);
                L.x.resize(H.n * ts);
            }
            if (l + 1 < h.levels.size())
            {
                L.r.resize(H.n * ts);
                L.pRows = H.P.rows;
                L.rRows = H.R.rows;
                L.pLanes = lanesFor(meanRow(H.P.rowStart));
                L.rLanes = lanesFor(meanRow(H.R.rowStart));
                L.pRowStart.upload(H.P.rowStart);
                L.pColumn.upload(H.P.column);
                L.rRowStart.upload(H.R.rowStart);
                L.rColumn.upload(H.R.column);
                if (mixed_)
                {
                    L.pValues32.upload(std::vector<float>(H.P.values.begin(), H.P.values.end()));
                    L.rValues32.upload(std::vector<float>(H.R.values.begin(), H.R.values.end()));
                }
                else
                {
                    L.pValues64.upload(H.P.values);
                    L.rValues64.upload(H.R.values);
                }
            }
        }
        coarseN_ = h.levels.back().n;
        if (mixed_)
        {
            coarse32_.upload(std::vector<float>(h.coarseInverse.begin(), h.coarseInverse.end()));
        }
        else
        {
            coarse64_.upload(h.coarseInverse);
        }
        r_.resize(n_);
        p_.resize(n_);
        q_.resize(n_);
        z_.resize(n_ * ts);
        if (mixed_)
        {
            rP_.resize(n_ * ts);
        }
        parts_ = std::max(1, std::min(1024, (n_ + kBlock - 1) / kBlock));
        scal_.resize(4);
        ints_.resize(4);
        for (auto * a : {&partPq_, &partRr_, &partRho_, &partBb_})
        {
            a->resize(parts_);
        }
        ints_.zero();
        scal_.zero();
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&hostPoll_), 8 * sizeof(int), hipHostMallocDefault));
        for (hipEvent_t & e : pollEvents_)
        {
            NSFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        setupSeconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    AmgPcg::~AmgPcg()
    {
        destroyGraphs();
        for (hipEvent_t e : pollEvents_)
        {
            if (e)
            {
                (void)hipEventDestroy(e);
            }
        }
        if (hostPoll_)
        {
            (void)hipHostFree(hostPoll_);
        }
    }

    void AmgPcg::destroyGraphs()
    {
        if (iterExec_)
        {
            (void)hipGraphExecDestroy(iterExec_);
            iterExec_ = nullptr;
        }
    }

    template <typename T>
    McgsMatrix<T> AmgPcg::levelMatrix(int l) const
    {
        if (l == 0)
        {
            if constexpr (std::is_same_v<T, double>)
            {
                return a64_;
            }
            else
            {
                return a32_;
            }
        }
        const Level & L = levels_[l];
        if constexpr (std::is_same_v<T, double>)
        {
            return McgsMatrix<double>{L.n, L.rowStart.get(), L.rowSplit.get(), L.column.get(), L.values64.get(), L.diag64.get(),
                                      L.invDiag64.get()};
        }
        else
        {
            return McgsMatrix<float>{L.n, L.rowStart.get(), L.rowSplit.get(), L.column.get(), L.values32.get(), L.diag32.get(),
                                     L.invDiag32.get()};
        }
    }

    template <typename T>
    void AmgPcg::enqueueVcycle(hipStream_t s, const T * b0, T * x0)
    {
        const int * stop = ints_.get();
        const int last = numLevels() - 1;
        auto bOf = [&](int l) { return l == 0 ? b0 : reinterpret_cast<const T *>(levels_[l].b.get()); };
        auto xOf = [&](int l) { return l == 0 ? x0 : reinterpret_cast<T *>(levels_[l].x.get()); };
        auto pVals = [&](const Level & L) -> const T *
        {
            if constexpr (std::is_same_v<T, double>)
            {
                return L.pValues64.get();
            }
            else
            {
                return L.pValues32.get();
            }
        };
        auto rVals = [&](const Level & L) -> const T *
        {
            if constexpr (std::is_same_v<T, double>)
            {
                return L.rValues64.get();
            }
            else
            {
                return L.rValues32.get();
            }
        };
        auto sweep = [&](int l, int c, int lowerOnly)
        {
            const Level & L = levels_[l];
            const int r0 = L.colorStart[c], r1 = L.colorStart[c + 1];
            if (r1 <= r0)
            {
                return;
            }
            const McgsMatrix<T> A = levelMatrix<T>(l);
            byLanes(L.lanes, [&]<int LN>()
                    {
                        const int grid = static_cast<int>((static_cast<long>(r1 - r0) * LN + kBlock - 1) / kBlock);
                        amgSweepKernel<T, LN><<<grid, kBlock, 0, s>>>(A, r0, r1, bOf(l), xOf(l), lowerOnly, stop);
                    });
        };
        for (int l = 0; l < last; l++)
        {
            Level & L = levels_[l];
            for (int c = 0; c < static_cast<int>(L.colorStart.size()) - 1; c++)
            {
                sweep(l, c, 1);
            }
            T * r = reinterpret_cast<T *>(L.r.get());
            const McgsMatrix<T> A = levelMatrix<T>(l);
            byLanes(L.lanes, [&]<int LN>()
                    {
                        const int grid = static_cast<int>((static_cast<long>(L.n) * LN + kBlock - 1) / kBlock);
                        amgResidualKernel<T, LN><<<grid, kBlock, 0, s>>>(A, bOf(l), xOf(l), r, stop);
                    });
            byLanes(L.rLanes, [&]<int LN>()
                    {
                        const int grid = static_cast<int>((static_cast<long>(L.rRows) * LN + kBlock - 1) / kBlock);
                        amgSpmvKernel<T, LN><<<grid, kBlock, 0, s>>>(L.rRows, L.rRowStart.get(), L.rColumn.get(), rVals(L), r,
                                                                     reinterpret_cast<T *>(levels_[l + 1].b.get()), 0, stop);
                    });
        }
        {
            const T * inv = nullptr;
            if constexpr (std::is_same_v<T, double>)
            {
                inv = coarse64_.get();
            }
            else
            {
                inv = coarse32_.get();
            }
            const int grid = (coarseN_ * 64 + kBlock - 1) / kBlock;
            amgDenseKernel<T><<<grid, kBlock, 0, s>>>(coarseN_, inv, bOf(last), xOf(last), stop);
        }
        for (int l = last - 1; l >= 0; l--)
        {
            const Level & L = levels_[l];
            byLanes(L.pLanes, [&]<int LN>()
                    {
                        const int grid = static_cast<int>((static_cast<long>(L.pRows) * LN + kBlock - 1) / kBlock);
                        amgSpmvKernel<T, LN><<<grid, kBlock, 0, s>>>(L.pRows, L.pRowStart.get(), L.pColumn.get(), pVals(L), xOf(l + 1),
                                                                     xOf(l), 1, stop);
                    });
            for (int c = static_cast<int>(L.colorStart.size()) - 2; c >= 0; c--)
            {
                sweep(l, c, 0);
            }
        }
        NSFEM_HIP(hipGetLastError());
    }

    template <typename P>
    void AmgPcg::enqueueInit(hipStream_t s, const double * b, double * x, bool zeroGuess, double tol2)
    {
        const AmgState st{scal_.get(), ints_.get(), partPq_.get(), partRr_.get(), partRho_.get(), partBb_.get(), parts_, tol2};
        P * rP = mixed_ ? reinterpret_cast<P *>(rP_.get()) : reinterpret_cast<P *>(r_.get());
        P * z = reinterpret_cast<P *>(z_.get());
        byLanes(levels_[0].lanes, [&]<int LN>()
                { amgInitKernel<P, LN><<<parts_, kBlock, 0, s>>>(a64_, b, x, r_.get(), rP, zeroGuess ? 1 : 0, st); });
        amgInitFinalizeKernel<<<1, kBlock, 0, s>>>(st);
        enqueueVcycle<P>(s, rP, z);
        amgRhoKernel<P><<<parts_, kBlock, 0, s>>>(n_, r_.get(), z, st);
        amgFinalizeKernel<<<1, kBlock, 0, s>>>(st, 1);
        NSFEM_HIP(hipGetLastError());
    }

    template <typename P>
    void AmgPcg::enqueueIteration(hipStream_t s, double tol2)
    {
        const AmgState st{scal_.get(), ints_.get(), partPq_.get(), partRr_.get(), partRho_.get(), partBb_.get(), parts_, tol2};
        P * rP = mixed_ ? reinterpret_cast<P *>(rP_.get()) : reinterpret_cast<P *>(r_.get());
        P * z = reinterpret_cast<P *>(z_.get());
        byLanes(levels_[0].lanes, [&]<int LN>()
                { amgDirectionKernel<P, LN><<<parts_, kBlock, 0, s>>>(a64_, z, p_.get(), q_.get(), st); });
        amgUpdateKernel<P><<<parts_, kBlock, 0, s>>>(n_, p_.get(), q_.get(), xDevice_, r_.get(), rP, st);
        enqueueVcycle<P>(s, rP, z);
        amgRhoKernel<P><<<parts_, kBlock, 0, s>>>(n_, r_.get(), z, st);
        amgFinalizeKernel<<<1, kBlock, 0, s>>>(st, 0);
        NSFEM_HIP(hipGetLastError());
    }

    double AmgPcg::timeKernel(const std::string & which, int reps, double * bytesOut, double * launchesOut, double * x)
    {
        if (reps < 1)
        {
            throw std::invalid_argument("reps must be positive");
        }
        const hipStream_t s = stream_;
        const double vs = mixed_ ? 4 : 8, n = n_;
        const double nnz0 = levels_[0].nnz;
        const double bDir = nnz0 * 12 + n * (4 + 8 + vs + 32);
        const double bUpd = n * (8 + 8 + 16 + 16 + (mixed_ ? 4 : 0));
        const double bRho = n * (8 + vs);
        double bV = static_cast<double>(coarseN_) * coarseN_ * vs;
        for (int l = 0; l + 1 < numLevels(); l++)
        {
            const Level & L = levels_[l];
            // forward lower parts, residual and backward full rows; the diagonal
            // or its inverse three times; b, x and r about eight passes; P and R
            bV += (L.nnzLower + 2 * L.nnz) * (vs + 4) + 3 * 4.0 * (L.n + 1) + 3 * vs * L.n + 8 * vs * L.n;
            bV += (L.nnzP + L.nnzR) * (vs + 4) + 4.0 * (L.pRows + L.rRows + 2);
        }
        double bytes = 0;
        std::function<void()> launch;
        // a stop flag that never rises: every kernel does its work; finite
        // step scalars
        fillInt(4, 0, ints_.get(), s);
        fill(F64, 4, 1.0, scal_.get(), s);
        fill(F64, parts_, 1.0, partPq_.get(), s);
        const double tol2 = -1.0;
        const AmgState st{scal_.get(), ints_.get(), partPq_.get(), partRr_.get(), partRho_.get(), partBb_.get(), parts_, tol2};
        if (x != xDevice_)
        {
            xDevice_ = x;
            destroyGraphs();
        }
        auto typed = [&](auto tag)
        {
            using P = decltype(tag);
            P * rP = mixed_ ? reinterpret_cast<P *>(rP_.get()) : reinterpret_cast<P *>(r_.get());
            P * z = reinterpret_cast<P *>(z_.get());
            if (which == "direction")
            {
                bytes = bDir;
                launch = [&, rP, z]()
                {
                    byLanes(levels_[0].lanes, [&]<int LN>()
                            { amgDirectionKernel<P, LN><<<parts_, kBlock, 0, s>>>(a64_, z, p_.get(), q_.get(), st); });
                };
            }
            else if (which == "update")
            {
                bytes = bUpd;
                launch = [&, rP]() { amgUpdateKernel<P><<<parts_, kBlock, 0, s>>>(n_, p_.get(), q_.get(), x, r_.get(), rP, st); };
            }
            else if (which == "rho")
            {
                bytes = bRho;
                launch = [&, z]() { amgRhoKernel<P><<<parts_, kBlock, 0, s>>>(n_, r_.get(), z, st); };
            }
            else if (which == "vcycle")
            {
                bytes = bV;
                launch = [&, rP, z]() { enqueueVcycle<P>(s, rP, z); };
            }
            else if (which == "iteration")
            {
                bytes = bDir + bUpd + bRho + bV;
                if (!iterExec_ || graphTol2_ != tol2)
                {
                    destroyGraphs();
                    hipGraph_t g = nullptr;
                    NSFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    enqueueIteration<P>(s, tol2);
                    NSFEM_HIP(hipStreamEndCapture(s, &g));
                    const hipError_t rc = hipGraphInstantiate(&iterExec_, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    NSFEM_HIP(rc);
                    graphTol2_ = tol2;
                }
                launch = [&]() { NSFEM_HIP(hipGraphLaunch(iterExec_, s)); };
            }
            else
            {
                throw std::invalid_argument("unknown amg_pcg kernel '" + which + "' (direction, update, rho, vcycle, iteration)");
            }
        };
        if (mixed_)
        {
            typed(float{});
        }
        else
        {
            typed(double{});
        }
        launch(); // warm
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
        NSFEM_HIP(hipGetLastError());
        destroyGraphs(); // the timing graph holds a negative tolerance
        if (bytesOut)
        {
            *bytesOut = bytes;
        }
        if (launchesOut)
        {
            *launchesOut = 1;
        }
        return static_cast<double>(ms) / reps;
    }

    AmgPcg::Result AmgPcg::solve(const double * b, double * x, bool zeroGuess, double target, int maxIters, bool useGraphs)
    {
        const hipStream_t s = stream_;
        const double tol2 = target * target;
        if (x != xDevice_)
        {
            xDevice_ = x; // a kernel argument of the captured iteration
            destroyGraphs();
        }
        if (mixed_)
        {
            enqueueInit<float>(s, b, x, zeroGuess, tol2);
        }
        else
        {
            enqueueInit<double>(s, b, x, zeroGuess, tol2);
        }
        auto iterate = [&]()
        {
            if (!useGraphs)
            {
                if (mixed_)
                {
                    enqueueIteration<float>(s, tol2);
                }
                else
                {
                    enqueueIteration<double>(s, tol2);
                }
                return;
            }
            if (!iterExec_ || graphTol2_ != tol2)
            {
                destroyGraphs();
                hipGraph_t g = nullptr;
                NSFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try
                {
                    if (mixed_)
                    {
                        enqueueIteration<float>(s, tol2);
                    }
                    else
                    {
                        enqueueIteration<double>(s, tol2);
                    }
                }
                catch (...)
                {
                    hipGraph_t gg = nullptr;
                    (void)hipStreamEndCapture(s, &gg);
                    if (gg)
                    {
                        (void)hipGraphDestroy(gg);
                    }
                    throw;
                }
                NSFEM_HIP(hipStreamEndCapture(s, &g));
                const hipError_t rc = hipGraphInstantiate(&iterExec_, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                NSFEM_HIP(rc);
                graphTol2_ = tol2;
            }
            NSFEM_HIP(hipGraphLaunch(iterExec_, s));
        };
        // a burst up to a little before the last solve's count, then one
        // iteration at a time with the stop flag polled one iteration behind
        int launched = 0;
        const int burst = std::max(0, std::min(maxIters, predicted_ - 2));
        for (; launched < burst; launched++)
        {
            iterate();
        }
        int slot = 0;
        bool pending[2] = {false, false};
        while (launched < maxIters)
        {
            iterate();
            launched++;
            NSFEM_HIP(hipMemcpyAsync(hostPoll_ + 4 * slot, ints_.get(), 4 * sizeof(int), hipMemcpyDeviceToHost, s));
            NSFEM_HIP(hipEventRecord(pollEvents_[slot], s));
            pending[slot] = true;
            const int prev = slot ^ 1;
            if (pending[prev])
            {
                NSFEM_HIP(hipEventSynchronize(pollEvents_[prev]));
                pending[prev] = false;
                if (hostPoll_[4 * prev] != 0)
                {
                    break;
                }
            }
            slot = prev;
        }
        double sc[4];
        int in[4];
        NSFEM_HIP(hipMemcpyAsync(sc, scal_.get(), sizeof sc, hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipMemcpyAsync(in, ints_.get(), sizeof in, hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipStreamSynchronize(s));
        Result res;
        res.iterations = in[1];
        res.bb = sc[2];
        res.rr = sc[3];
        res.nonFinite = in[0] == 2;
        res.converged = in[0] == 1;
        if (in[2])
        {
            fill(F64, n_, 0.0, x, s); // b = 0: x = 0
            NSFEM_HIP(hipStreamSynchronize(s));
            res.rr = 0;
        }
        predicted_ = res.iterations;
        return res;
    }
} // namespace nsfem
