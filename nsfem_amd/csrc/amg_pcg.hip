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
            double * scal; // per channel, blocks of NCH: rho, rho of the iteration before, |b|^2, |r|^2
            int * ints;    // stop (0 running, 1 converged, 2 non-finite or breakdown), iterations, zero right-hand sides
            int * active;  // per channel
            double * partPq;
            double * partRr;
            double * partRho;
            double * partBb;
            int parts;
            double tol2;
        };

        // ---------------- V-cycle (channels interleaved per row) ----------------

        // One colour of a multicolour Gauss-Seidel sweep: rows [rowBegin,
        // rowEnd) (one colour, independent of each other) take
        // x_i = (b_i - sum_j a_ij x_j) / a_ii over the lower-colour part of
        // the row (lowerOnly: the first sweep from x = 0, where the higher
        // colours are still zero) or the whole row.  L lanes per row.
        template <typename T, int L, int NCH>
        __global__ void __launch_bounds__(kBlock) amgSweepKernel(McgsMatrix<T> A, int rowBegin, int rowEnd, const T * __restrict__ b,
                                                                 T * __restrict__ x, int lowerOnly, const int * __restrict__ stop)
        {
            // The stop flag, the row bounds and the row's own operands load
            // together, with no condition around a load (one would end in a
            // wait for every load in flight): lanes past the colour read its
            // first row and take no entries.
            const int t = blockIdx.x * kBlock + threadIdx.x;
            const int row = rowBegin + t / L;
            const int lane = t % L;
            const bool valid = row < rowEnd;
            const int rs = valid ? row : rowBegin;
            const int stopped = *stop;
            const int * end = lowerOnly ? A.rowSplit : A.rowStart + 1;
            const int j0 = A.rowStart[rs];
            const int j1 = end[rs];
            const T d = A.invDiag[rs];
            T bi[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                bi[c] = b[static_cast<long>(rs) * NCH + c];
            }
            // (j1 < j0 never holds; testing it keeps the loads ahead of the
            // branch, where the compiler would otherwise sink them)
            if (stopped | (j1 < j0))
            {
                return;
            }
            T s[NCH];
            rowDot<T, L, NCH>(A, j0, valid ? j1 : j0, x, s);
            if (valid && lane == 0)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    x[static_cast<long>(row) * NCH + c] = (bi[c] - s[c]) * d;
                }
            }
        }

        // r = b - A x (A = stripped + diagonal).  Loads as in amgSweepKernel.
        template <typename T, int L, int NCH>
        __global__ void __launch_bounds__(kBlock) amgResidualKernel(McgsMatrix<T> A, const T * __restrict__ b, const T * __restrict__ x,
                                                                    T * __restrict__ r, const int * __restrict__ stop)
        {
            const int t = blockIdx.x * kBlock + threadIdx.x;
            const int row = t / L;
            const int lane = t % L;
            const bool valid = row < A.n;
            const int rs = valid ? row : 0;
            const int stopped = *stop;
            const int j0 = A.rowStart[rs];
            const int j1 = A.rowStart[rs + 1];
            const T d = A.diag[rs];
            T bi[NCH], xi[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                const long i = static_cast<long>(rs) * NCH + c;
                bi[c] = b[i];
                xi[c] = x[i];
            }
            // (j1 < j0 never holds; testing it keeps the loads ahead of the
            // branch, where the compiler would otherwise sink them)
            if (stopped | (j1 < j0))
            {
                return;
            }
            T s[NCH];
            rowDot<T, L, NCH>(A, j0, valid ? j1 : j0, x, s);
            if (valid && lane == 0)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    r[static_cast<long>(row) * NCH + c] = bi[c] - fmaT(d, xi[c], s[c]);
                }
            }
        }

        // y = M x (accumulate: y += M x) for the restriction and prolongation.
        // xInit != nullptr (a restriction onto a level smoothed by the
        // Chebyshev iteration): also xInit = initScale D^-1 y, the level's
        // first iterate from zero (initInv = D^-1 of that level).  Loads as in
        // amgSweepKernel.
        template <typename T, int L, int NCH>
        __global__ void __launch_bounds__(kBlock) amgSpmvKernel(int rows, const int * __restrict__ rowStart, const int * __restrict__ column,
                                                                const T * __restrict__ values, const T * __restrict__ x, T * __restrict__ y,
                                                                int accumulate, const int * __restrict__ stop, const T * __restrict__ initInv,
                                                                T * __restrict__ xInit, T initScale)
        {
            const int t = blockIdx.x * kBlock + threadIdx.x;
            const int row = t / L;
            const int lane = t % L;
            const bool valid = row < rows;
            const int rs = valid ? row : 0;
            const int stopped = *stop;
            const int j0 = rowStart[rs];
            const int j1 = rowStart[rs + 1];
            // the old y only where it is added to (element 0 otherwise)
            T yi[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                yi[c] = y[accumulate ? static_cast<long>(rs) * NCH + c : 0];
            }
            if (stopped)
            {
                return;
            }
            const McgsMatrix<T> M{rows, rowStart, nullptr, column, values, nullptr, nullptr};
            T s[NCH];
            rowDot<T, L, NCH>(M, j0, valid ? j1 : j0, x, s);
            if (valid && lane == 0)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    y[static_cast<long>(row) * NCH + c] = accumulate ? yi[c] + s[c] : s[c];
                }
                if (xInit)
                {
                    const T scale = initScale * initInv[row];
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        xInit[static_cast<long>(row) * NCH + c] = scale * s[c];
                    }
                }
            }
        }

        // One step of the Chebyshev iteration on D^-1 A (Saad, Iterative
        // Methods for Sparse Linear Systems, 2nd ed., Algorithm 12.1, with
        // the diagonal as preconditioner): r = b - A xIn,
        // d = c1 dIn + c2 D^-1 r, xOut = xIn + d, dOut = d when given.  xIn
        // and xOut are different vectors (other rows gather xIn); d is
        // row-local (dIn may be dOut, or xIn when the previous direction is
        // the first iterate itself).  Loads as in amgSweepKernel.
        template <typename T, int L, int NCH>
        __global__ void __launch_bounds__(kBlock) amgChebKernel(McgsMatrix<T> A, const T * __restrict__ b, const T * xIn, const T * dIn,
                                                                T * __restrict__ xOut, T * dOut, T c1, T c2, const int * __restrict__ stop)
        {
            const int t = blockIdx.x * kBlock + threadIdx.x;
            const int row = t / L;
            const int lane = t % L;
            const bool valid = row < A.n;
            const int rs = valid ? row : 0;
            const int stopped = *stop;
            const int j0 = A.rowStart[rs];
            const int j1 = A.rowStart[rs + 1];
            const T dg = A.diag[rs];
            const T inv = A.invDiag[rs];
            T bi[NCH], xi[NCH], di[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                const long i = static_cast<long>(rs) * NCH + c;
                bi[c] = b[i];
                xi[c] = xIn[i];
                di[c] = dIn[i];
            }
            if (stopped | (j1 < j0))
            {
                return; // j1 < j0 never holds (see amgSweepKernel)
            }
            T s[NCH];
            rowDot<T, L, NCH>(A, j0, valid ? j1 : j0, xIn, s);
            if (valid && lane == 0)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    const long i = static_cast<long>(row) * NCH + c;
                    const T r = bi[c] - fmaT(dg, xi[c], s[c]);
                    const T d = fmaT(c1, di[c], c2 * (inv * r));
                    xOut[i] = xi[c] + d;
                    if (dOut)
                    {
                        dOut[i] = d;
                    }
                }
            }
        }

        // x = scale D^-1 b: the first Chebyshev iterate from zero on the
        // finest level (coarser levels get it from their restriction).
        template <typename T, int NCH>
        __global__ void __launch_bounds__(kBlock) amgScaleKernel(int n, const T * __restrict__ invDiag, const T * __restrict__ b,
                                                                 T * __restrict__ x, T scale, const int * __restrict__ stop)
        {
            const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
            const int row = static_cast<int>(t / NCH);
            const bool valid = row < n;
            const int rs = valid ? row : 0;
            const int stopped = *stop;
            const T v = scale * invDiag[rs] * b[valid ? t : 0];
            if (!stopped && valid)
            {
                x[t] = v;
            }
        }

        // x = Ainv b on the coarsest level (dense, row-major): one wavefront per row.
        template <typename T, int NCH>
        __global__ void __launch_bounds__(kBlock) amgDenseKernel(int n, const T * __restrict__ inv, const T * __restrict__ b,
                                                                 T * __restrict__ x, const int * __restrict__ stop)
        {
            if (*stop)
            {
                return;
            }
            const int row = (blockIdx.x * kBlock + threadIdx.x) >> 6;
            const int lane = threadIdx.x & 63;
            T s[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                s[c] = 0;
            }
            if (row < n)
            {
                const T * a = inv + static_cast<size_t>(row) * n;
                for (int j = lane; j < n; j += 64)
                {
                    const T v = a[j];
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        s[c] = fmaT(v, b[static_cast<long>(j) * NCH + c], s[c]);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                s[c] = laneGroupSum<64>(s[c]);
            }
            if (row < n && lane == 0)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    x[static_cast<long>(row) * NCH + c] = s[c];
                }
            }
        }

        // ---------------- CG around it ----------------

        // Per-channel sums of count partials (parts[k * K + c]) in a fixed
        // order, the same in every block; valid in thread 0.
        template <int K>
        __device__ __forceinline__ void sumParts(const double * __restrict__ parts, int count, double (&v)[K], double * sm)
        {
#pragma unroll
            for (int c = 0; c < K; c++)
            {
                v[c] = 0;
            }
            for (int k = threadIdx.x; k < count; k += kBlock)
            {
#pragma unroll
                for (int c = 0; c < K; c++)
                {
                    v[c] += parts[static_cast<size_t>(k) * K + c];
                }
            }
            blockSum<K>(v, sm);
        }

        template <int K>
        __device__ __forceinline__ void storeParts(const double (&v)[K], double * parts)
        {
            if (threadIdx.x == 0)
            {
#pragma unroll
                for (int c = 0; c < K; c++)
                {
                    parts[static_cast<size_t>(blockIdx.x) * K + c] = v[c];
                }
            }
        }

        // r = b - A x (x = 0 and r = b for a zero guess), the V-cycle's copy
        // of r, per-channel partials of |b|^2 and |r|^2.
        template <typename P, int L, int NCH>
        __global__ void __launch_bounds__(kBlock) amgInitKernel(McgsMatrix<double> A, const double * __restrict__ b, double * __restrict__ x,
                                                                double * __restrict__ r, P * __restrict__ rP, int zeroGuess, AmgState st)
        {
            __shared__ double sm[2 * NCH * kWaves];
            constexpr int G = kBlock / L;
            const int lane = threadIdx.x % L;
            double bb[NCH], rr[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                bb[c] = rr[c] = 0;
            }
            for (int base = blockIdx.x * G; base < A.n; base += gridDim.x * G)
            {
                const int row = base + static_cast<int>(threadIdx.x) / L;
                double s[NCH];
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    s[c] = 0;
                }
                if (row < A.n && !zeroGuess)
                {
                    for (int j = A.rowStart[row] + lane; j < A.rowStart[row + 1]; j += L)
                    {
                        const double a = A.values[j];
                        const long col = static_cast<long>(A.column[j]) * NCH;
#pragma unroll
                        for (int c = 0; c < NCH; c++)
                        {
                            s[c] = fmaT(a, x[col + c], s[c]);
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    s[c] = laneGroupSum<L>(s[c]);
                }
                if (row < A.n && lane == 0)
                {
                    const double d = A.diag[row];
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        const long i = static_cast<long>(row) * NCH + c;
                        const double bi = b[i];
                        double ri = bi;
                        if (zeroGuess)
                        {
                            x[i] = 0;
                        }
                        else
                        {
                            ri = bi - fmaT(d, x[i], s[c]);
                        }
                        r[i] = ri;
                        if constexpr (!std::is_same_v<P, double>)
                        {
                            rP[i] = static_cast<P>(ri);
                        }
                        bb[c] = fmaT(bi, bi, bb[c]);
                        rr[c] = fmaT(ri, ri, rr[c]);
                    }
                }
            }
            blockSum<NCH>(bb, sm);
            blockSum<NCH>(rr, sm);
            storeParts<NCH>(bb, st.partBb);
            storeParts<NCH>(rr, st.partRr);
        }

        template <int NCH>
        __global__ void __launch_bounds__(kBlock) amgInitFinalizeKernel(AmgState st)
        {
            __shared__ double sm[NCH * kWaves];
            double bb[NCH], rr[NCH];
            sumParts<NCH>(st.partBb, st.parts, bb, sm);
            sumParts<NCH>(st.partRr, st.parts, rr, sm);
            if (threadIdx.x == 0)
            {
                int any = 0, bad = 0, zero = 0;
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    st.scal[c] = 0;
                    st.scal[NCH + c] = 0;
                    st.scal[2 * NCH + c] = bb[c];
                    st.scal[3 * NCH + c] = rr[c];
                    bad |= !std::isfinite(bb[c]) || !std::isfinite(rr[c]);
                    const int act = bb[c] > 0 && rr[c] > st.tol2 * bb[c];
                    st.active[c] = act;
                    any |= act;
                    if (bb[c] == 0 && rr[c] > 0)
                    {
                        zero |= 1 << c; // b = 0 from a nonzero guess: the solution is 0
                    }
                }
                st.ints[1] = 0;
                st.ints[2] = zero;
                st.ints[0] = bad ? 2 : (any ? 0 : 1);
            }
        }

        // beta = rho / rho_before (0 on the first direction and for stopped
        // channels); p = z + beta p, q = A z + beta q (= A p, as
        // A p_before = q), per-channel partials of p.q.
        template <typename P, int L, int NCH>
        __global__ void __launch_bounds__(kBlock) amgDirectionKernel(McgsMatrix<double> A, const P * __restrict__ z, double * __restrict__ p,
                                                                     double * __restrict__ q, AmgState st)
        {
            if (st.ints[0])
            {
                return;
            }
            __shared__ double sm[NCH * kWaves];
            double beta[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                const double before = st.scal[NCH + c];
                beta[c] = (st.active[c] && before != 0) ? st.scal[c] / before : 0.0;
            }
            constexpr int G = kBlock / L;
            const int lane = threadIdx.x % L;
            double acc[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                acc[c] = 0;
            }
            for (int base = blockIdx.x * G; base < A.n; base += gridDim.x * G)
            {
                // the row's own operands load on its first lane after the
                // row sum: measured faster here than loading them on every
                // lane beside the gathers (short P1 rows, several slabs per
                // block keep enough loads in flight)
                const int row = base + static_cast<int>(threadIdx.x) / L;
                const bool valid = row < A.n;
                const int rs = valid ? row : base;
                const int j0 = A.rowStart[rs];
                const int j1 = A.rowStart[rs + 1];
                double s[NCH];
                rowDot<double, L, NCH>(A, j0, valid ? j1 : j0, z, s);
                if (valid && lane == 0)
                {
                    const double d = A.diag[row];
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        const long i = static_cast<long>(row) * NCH + c;
                        const double zi = static_cast<double>(z[i]);
                        const double pi = fmaT(beta[c], p[i], zi);
                        const double qi = fmaT(beta[c], q[i], fmaT(d, zi, s[c]));
                        p[i] = pi;
                        q[i] = qi;
                        acc[c] = fmaT(pi, qi, acc[c]);
                    }
                }
            }
            blockSum<NCH>(acc, sm);
            storeParts<NCH>(acc, st.partPq);
        }

        // alpha = rho / p.q for the channels still iterating, 0 for the
        // stopped ones (every block sums the partials itself, in the same
        // order); x += alpha p, r -= alpha q, the V-cycle's copy of r,
        // per-channel partials of |r|^2.
        template <typename P, int NCH>
        __global__ void __launch_bounds__(kBlock) amgUpdateKernel(int n, const double * __restrict__ p, const double * __restrict__ q,
                                                                  double * __restrict__ x, double * __restrict__ r, P * __restrict__ rP,
                                                                  AmgState st)
        {
            if (st.ints[0])
            {
                return;
            }
            __shared__ double sm[NCH * kWaves];
            __shared__ double shAlpha[NCH];
            double pq[NCH];
            sumParts<NCH>(st.partPq, st.parts, pq, sm);
            if (threadIdx.x == 0)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    shAlpha[c] = st.active[c] ? st.scal[c] / pq[c] : 0.0;
                }
            }
            __syncthreads();
            double alpha[NCH], acc[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                alpha[c] = shAlpha[c];
                acc[c] = 0;
            }
            for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    const long i = static_cast<long>(row) * NCH + c;
                    x[i] = fmaT(alpha[c], p[i], x[i]);
                    const double ri = fmaT(-alpha[c], q[i], r[i]);
                    r[i] = ri;
                    if constexpr (!std::is_same_v<P, double>)
                    {
                        rP[i] = static_cast<P>(ri);
                    }
                    acc[c] = fmaT(ri, ri, acc[c]);
                }
            }
            blockSum<NCH>(acc, sm);
            storeParts<NCH>(acc, st.partRr);
        }

        // per-channel partials of r.z
        template <typename P, int NCH>
        __global__ void __launch_bounds__(kBlock) amgRhoKernel(int n, const double * __restrict__ r, const P * __restrict__ z, AmgState st)
        {
            if (st.ints[0])
            {
                return;
            }
            __shared__ double sm[NCH * kWaves];
            double acc[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                acc[c] = 0;
            }
            for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    const long i = static_cast<long>(row) * NCH + c;
                    acc[c] = fmaT(r[i], static_cast<double>(z[i]), acc[c]);
                }
            }
            blockSum<NCH>(acc, sm);
            storeParts<NCH>(acc, st.partRho);
        }

        // rho = r.z per channel; after an update (init = 0) also |r|^2, the
        // iteration count and each channel's stopping test.  A channel still
        // iterating with r.z <= 0 (a breakdown) or a non-finite value stops
        // the solve as non-finite; the solve stops when no channel iterates.
        template <int NCH>
        __global__ void __launch_bounds__(kBlock) amgFinalizeKernel(AmgState st, int init)
        {
            if (st.ints[0])
            {
                return;
            }
            __shared__ double sm[NCH * kWaves];
            double rho[NCH], rr[NCH];
            sumParts<NCH>(st.partRho, st.parts, rho, sm);
            if (!init)
            {
                sumParts<NCH>(st.partRr, st.parts, rr, sm);
            }
            if (threadIdx.x == 0)
            {
                if (st.tol2 < 0)
                {
                    // kernel timing: never stops, whatever the vectors hold
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        st.scal[NCH + c] = st.scal[c];
                        st.scal[c] = std::isfinite(rho[c]) && rho[c] > 0 ? rho[c] : 1.0;
                    }
                    return;
                }
                int any = 0, bad = 0;
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    int act = st.active[c];
                    if (!init)
                    {
                        st.scal[3 * NCH + c] = rr[c];
                        if (act)
                        {
                            if (!std::isfinite(rr[c]))
                            {
                                bad = 1;
                            }
                            else if (rr[c] <= st.tol2 * st.scal[2 * NCH + c])
                            {
                                act = 0;
                            }
                        }
                    }
                    if (act && (!std::isfinite(rho[c]) || !(rho[c] > 0)))
                    {
                        bad = 1;
                    }
                    st.active[c] = act;
                    any |= act;
                    st.scal[NCH + c] = init ? 0.0 : st.scal[c];
                    st.scal[c] = rho[c];
                }
                if (!init)
                {
                    st.ints[1] += 1;
                }
                st.ints[0] = bad ? 2 : (any ? 0 : 1);
            }
        }

        // Channels of x (interleaved) whose bit is set in mask become 0.
        __global__ void amgZeroChannelsKernel(int n, int nch, int mask, double * __restrict__ x)
        {
            const long total = static_cast<long>(n) * nch;
            for (long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; t < total; t += static_cast<long>(gridDim.x) * blockDim.x)
            {
                if ((mask >> static_cast<int>(t % nch)) & 1)
                {
                    x[t] = 0;
                }
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
                   const std::vector<int> & colorStart, bool mixed, int numCh, McgsMatrix<double> a64, McgsMatrix<float> level0F32,
                   hipStream_t s)
        : mixed_(mixed), n_(stripped.rows), nch_(numCh), a64_(a64), a32_(level0F32), stream_(s)
    {
        if (nch_ != 1 && nch_ != 2 && nch_ != 4 && nch_ != 8 && nch_ != 16)
        {
            throw std::invalid_argument("amg_pcg solves 1, 2, 4, 8 or 16 right-hand sides");
        }
        const auto t0 = std::chrono::steady_clock::now();
        const AmgHierarchy h = buildAmgHierarchy(stripped, rowSplit, diag, colorStart);
        complexity_ = h.operatorComplexity;
        levels_.resize(h.levels.size());
        const size_t ts = mixed_ ? 4 : 8;
        const size_t vb = ts * nch_; // bytes per row of a V-cycle vector
        for (size_t l = 0; l < h.levels.size(); l++)
        {
            const AmgLevel & H = h.levels[l];
            Level & L = levels_[l];
            L.n = H.n;
            L.colorStart = H.colorStart;
            L.lambdaMax = H.lambdaMax;
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
                L.b.resize(H.n * vb);
                L.x.resize(H.n * vb);
            }
            if (l + 1 < h.levels.size())
            {
                L.r.resize(H.n * vb);
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
        const size_t nv = static_cast<size_t>(n_) * nch_;
        r_.resize(nv);
        p_.resize(nv);
        q_.resize(nv);
        z_.resize(nv * ts);
        if (mixed_)
        {
            rP_.resize(nv * ts);
        }
        parts_ = std::max(1, std::min(1024, (n_ + kBlock - 1) / kBlock));
        scal_.resize(4 * nch_);
        ints_.resize(4);
        active_.resize(nch_);
        for (auto * a : {&partPq_, &partRr_, &partRho_, &partBb_})
        {
            a->resize(static_cast<size_t>(parts_) * nch_);
        }
        ints_.zero();
        scal_.zero();
        active_.zero();
        NSFEM_HIP(hipHostMalloc(reinterpret_cast<void **>(&hostPoll_), 8 * sizeof(int), hipHostMallocDefault));
        for (hipEvent_t & e : pollEvents_)
        {
            NSFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        // Gauss-Seidel on the finest level, two Chebyshev steps over
        // [0.2, 1] x lambda_max on the coarser ones (measured at 500k rows:
        // 52 -> 24 launches per V-cycle, 18.3 -> 15.2 CG iterations)
        setSmoother(1, 2, 0.2);
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

    void AmgPcg::setSmoother(int chebyshevFrom, int degree, double fraction)
    {
        if (chebyshevFrom < 0)
        {
            throw std::invalid_argument("amg: the first Chebyshev level must be >= 0");
        }
        if (degree < 2 || degree > 8)
        {
            // (degree 1, damped Jacobi, measured far weaker: about twice the
            // CG iterations of degree 2 at 500k rows)
            throw std::invalid_argument("amg: the Chebyshev degree must be 2 to 8");
        }
        if (!(fraction > 0 && fraction < 1))
        {
            throw std::invalid_argument("amg: the Chebyshev interval fraction must be in (0, 1)");
        }
        const size_t vb = (mixed_ ? 4 : 8) * static_cast<size_t>(nch_);
        const int last = numLevels() - 1;
        for (int l = 0; l < last; l++)
        {
            Level & L = levels_[l];
            if (l < chebyshevFrom || !(L.lambdaMax > 0))
            {
                L.chebDegree = 0;
                L.chebC1.clear();
                L.chebC2.clear();
                continue;
            }
            // interval [fraction, 1] x lambda_max: theta its centre, delta
            // its half width; step 0 from the residual alone, then the
            // three-term recurrence rho_k = 1 / (2 sigma - rho_k-1)
            const double hi = L.lambdaMax, lo = fraction * hi;
            const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
            L.chebC1.assign(degree, 0.0);
            L.chebC2.assign(degree, 0.0);
            L.chebC2[0] = 1.0 / theta;
            double rho = 1.0 / sigma;
            for (int k = 1; k < degree; k++)
            {
                const double rhoNext = 1.0 / (2.0 * sigma - rho);
                L.chebC1[k] = rhoNext * rho;
                L.chebC2[k] = 2.0 * rhoNext / delta;
                rho = rhoNext;
            }
            L.chebDegree = degree;
            const size_t bytes = static_cast<size_t>(L.n) * vb;
            if (L.xa.size() < bytes)
            {
                L.xa.resize(bytes);
            }
            if (degree > 1 && L.d.size() < bytes)
            {
                L.d.resize(bytes);
            }
        }
        chebFrom_ = chebyshevFrom;
        chebDegree_ = degree;
        chebFraction_ = fraction;
        destroyGraphs(); // the captured iteration holds the old V-cycle
    }

    void AmgPcg::destroyGraphs()
    {
        if (iterExec_)
        {
            (void)hipGraphExecDestroy(iterExec_);
            iterExec_ = nullptr;
        }
    }

    template <typename F>
    void AmgPcg::typed(F && f)
    {
        auto byCh = [&]<typename P>()
        {
            switch (nch_)
            {
            case 1: f.template operator()<P, 1>(); break;
            case 2: f.template operator()<P, 2>(); break;
            case 4: f.template operator()<P, 4>(); break;
            case 8: f.template operator()<P, 8>(); break;
            case 16: f.template operator()<P, 16>(); break;
            default: throw std::invalid_argument("amg_pcg: unsupported channel count");
            }
        };
        if (mixed_)
        {
            byCh.template operator()<float>();
        }
        else
        {
            byCh.template operator()<double>();
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

    template <typename T, int NCH>
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
        auto grid = [](long rows, int lanes) { return static_cast<int>((rows * lanes + kBlock - 1) / kBlock); };
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
                    { amgSweepKernel<T, LN, NCH><<<grid(r1 - r0, LN), kBlock, 0, s>>>(A, r0, r1, bOf(l), xOf(l), lowerOnly, stop); });
        };
        // Chebyshev levels: the steps alternate between x and xa (other rows
        // gather the step's input); the first iterate goes into xa, so after
        // the k - 1 steps down and the k steps up the result is in x again
        auto xaOf = [&](int l) { return reinterpret_cast<T *>(levels_[l].xa.get()); };
        auto dOf = [&](int l) { return reinterpret_cast<T *>(levels_[l].d.get()); };
        auto other = [&](int l, T * v) { return v == xOf(l) ? xaOf(l) : xOf(l); };
        auto chebStep = [&](int l, int k, const T * xIn, const T * dIn, T * xOut, T * dOut)
        {
            const Level & L = levels_[l];
            const McgsMatrix<T> A = levelMatrix<T>(l);
            const T c1 = static_cast<T>(L.chebC1[k]), c2 = static_cast<T>(L.chebC2[k]);
            byLanes(L.lanes, [&]<int LN>()
                    { amgChebKernel<T, LN, NCH><<<grid(L.n, LN), kBlock, 0, s>>>(A, bOf(l), xIn, dIn, xOut, dOut, c1, c2, stop); });
        };
        // the current iterate of every level
        std::vector<T *> cur(levels_.size(), nullptr);
        for (int l = 0; l < last; l++)
        {
            Level & L = levels_[l];
            const int deg = L.chebDegree;
            if (deg == 0)
            {
                for (int c = 0; c < static_cast<int>(L.colorStart.size()) - 1; c++)
                {
                    sweep(l, c, 1);
                }
                cur[l] = xOf(l);
            }
            else
            {
                if (l == 0)
                {
                    const McgsMatrix<T> A = levelMatrix<T>(0);
                    amgScaleKernel<T, NCH><<<grid(static_cast<long>(L.n) * NCH, 1), kBlock, 0, s>>>(L.n, A.invDiag, b0, xaOf(0),
                                                                                                   static_cast<T>(L.chebC2[0]), stop);
                }
                // steps 1..deg-1 from the first iterate x1 (whose direction is x1)
                T * in = xaOf(l);
                const T * dIn = in;
                for (int k = 1; k < deg; k++)
                {
                    T * out = other(l, in);
                    chebStep(l, k, in, dIn, out, k + 1 < deg ? dOf(l) : nullptr);
                    dIn = dOf(l);
                    in = out;
                }
                cur[l] = in;
            }
            T * r = reinterpret_cast<T *>(L.r.get());
            const McgsMatrix<T> A = levelMatrix<T>(l);
            byLanes(L.lanes, [&]<int LN>() { amgResidualKernel<T, LN, NCH><<<grid(L.n, LN), kBlock, 0, s>>>(A, bOf(l), cur[l], r, stop); });
            // the coarser level's first Chebyshev iterate with its right-hand side
            const Level & C = levels_[l + 1];
            const bool init = l + 1 < last && C.chebDegree > 0;
            const T * cInv = nullptr;
            if (init)
            {
                cInv = levelMatrix<T>(l + 1).invDiag;
            }
            T * cInit = init ? xaOf(l + 1) : nullptr;
            const T cScale = init ? static_cast<T>(C.chebC2[0]) : T(0);
            byLanes(L.rLanes, [&]<int LN>()
                    {
                        amgSpmvKernel<T, LN, NCH><<<grid(L.rRows, LN), kBlock, 0, s>>>(L.rRows, L.rRowStart.get(), L.rColumn.get(), rVals(L),
                                                                                      r, reinterpret_cast<T *>(levels_[l + 1].b.get()), 0,
                                                                                      stop, cInv, cInit, cScale);
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
            amgDenseKernel<T, NCH><<<grid(coarseN_, 64), kBlock, 0, s>>>(coarseN_, inv, bOf(last), xOf(last), stop);
            cur[last] = xOf(last);
        }
        for (int l = last - 1; l >= 0; l--)
        {
            const Level & L = levels_[l];
            byLanes(L.pLanes, [&]<int LN>()
                    {
                        amgSpmvKernel<T, LN, NCH><<<grid(L.pRows, LN), kBlock, 0, s>>>(L.pRows, L.pRowStart.get(), L.pColumn.get(), pVals(L),
                                                                                      cur[l + 1], cur[l], 1, stop, nullptr, nullptr, T(0));
                    });
            const int deg = L.chebDegree;
            if (deg == 0)
            {
                for (int c = static_cast<int>(L.colorStart.size()) - 2; c >= 0; c--)
                {
                    sweep(l, c, 0);
                }
            }
            else
            {
                // steps 0..deg-1 from the corrected iterate (step 0 from the
                // residual alone: c1 = 0, its direction input unused)
                T * in = cur[l];
                for (int k = 0; k < deg; k++)
                {
                    T * out = other(l, in);
                    chebStep(l, k, in, k == 0 ? in : dOf(l), out, k + 1 < deg ? dOf(l) : nullptr);
                    in = out;
                }
                if (in != xOf(l))
                {
                    throw std::logic_error("amg: the Chebyshev steps must end in the level's x");
                }
            }
        }
        NSFEM_HIP(hipGetLastError());
    }

    template <typename P, int NCH>
    void AmgPcg::enqueueInit(hipStream_t s, const double * b, double * x, bool zeroGuess, double tol2)
    {
        const AmgState st{scal_.get(), ints_.get(), active_.get(), partPq_.get(), partRr_.get(), partRho_.get(), partBb_.get(), parts_, tol2};
        P * rP = mixed_ ? reinterpret_cast<P *>(rP_.get()) : reinterpret_cast<P *>(r_.get());
        P * z = reinterpret_cast<P *>(z_.get());
        byLanes(levels_[0].lanes, [&]<int LN>()
                { amgInitKernel<P, LN, NCH><<<parts_, kBlock, 0, s>>>(a64_, b, x, r_.get(), rP, zeroGuess ? 1 : 0, st); });
        amgInitFinalizeKernel<NCH><<<1, kBlock, 0, s>>>(st);
        enqueueVcycle<P, NCH>(s, rP, z);
        amgRhoKernel<P, NCH><<<parts_, kBlock, 0, s>>>(n_, r_.get(), z, st);
        amgFinalizeKernel<NCH><<<1, kBlock, 0, s>>>(st, 1);
        NSFEM_HIP(hipGetLastError());
    }

    template <typename P, int NCH>
    void AmgPcg::enqueueIteration(hipStream_t s, double tol2)
    {
        const AmgState st{scal_.get(), ints_.get(), active_.get(), partPq_.get(), partRr_.get(), partRho_.get(), partBb_.get(), parts_, tol2};
        P * rP = mixed_ ? reinterpret_cast<P *>(rP_.get()) : reinterpret_cast<P *>(r_.get());
        P * z = reinterpret_cast<P *>(z_.get());
        byLanes(levels_[0].lanes, [&]<int LN>()
                { amgDirectionKernel<P, LN, NCH><<<parts_, kBlock, 0, s>>>(a64_, z, p_.get(), q_.get(), st); });
        amgUpdateKernel<P, NCH><<<parts_, kBlock, 0, s>>>(n_, p_.get(), q_.get(), xDevice_, r_.get(), rP, st);
        enqueueVcycle<P, NCH>(s, rP, z);
        amgRhoKernel<P, NCH><<<parts_, kBlock, 0, s>>>(n_, r_.get(), z, st);
        amgFinalizeKernel<NCH><<<1, kBlock, 0, s>>>(st, 0);
        NSFEM_HIP(hipGetLastError());
    }

    double AmgPcg::timeKernel(const std::string & which, int reps, double * bytesOut, double * launchesOut, double * x)
    {
        if (reps < 1)
        {
            throw std::invalid_argument("reps must be positive");
        }
        const hipStream_t s = stream_;
        const double vs = mixed_ ? 4 : 8, n = n_, k = nch_;
        const double nnz0 = levels_[0].nnz;
        // matrix once, diagonal, z gathered once, p and q read and written
        const double bDir = nnz0 * 12 + n * (4 + 8) + k * n * (vs + 32);
        const double bUpd = k * n * (8 + 8 + 16 + 16 + (mixed_ ? 4 : 0));
        const double bRho = k * n * (8 + vs);
        double bV = static_cast<double>(coarseN_) * coarseN_ * vs + k * coarseN_ * 2 * vs;
        for (int l = 0; l + 1 < numLevels(); l++)
        {
            const Level & L = levels_[l];
            if (L.chebDegree == 0)
            {
                // forward lower parts, residual and backward full rows; the
                // diagonal or its inverse three times; b, x and r about eight
                // passes
                bV += (L.nnzLower + 2 * L.nnz) * (vs + 4) + 3 * 4.0 * (L.n + 1) + 3 * vs * L.n + 8 * k * vs * L.n;
            }
            else
            {
                // 2 deg - 1 steps (full rows, row bounds, diagonal and its
                // inverse; b, x in, d in, x out, d out), the residual (rows,
                // diagonal, b, x, r), the first iterate (written by the
                // restriction; a scaling of b on the finest level)
                const double steps = 2.0 * L.chebDegree - 1;
                bV += steps * (L.nnz * (vs + 4) + 4.0 * (L.n + 1) + 2 * vs * L.n + 5 * k * vs * L.n);
                bV += L.nnz * (vs + 4) + 4.0 * (L.n + 1) + vs * L.n + 3 * k * vs * L.n;
                bV += k * vs * L.n + (l == 0 ? vs * L.n + k * vs * L.n : 0.0);
            }
            // P and R
            bV += (L.nnzP + L.nnzR) * (vs + 4) + 4.0 * (L.pRows + L.rRows + 2);
        }
        double bytes = 0;
        std::function<void()> launch;
        hipGraphExec_t vcycleExec = nullptr;
        bool composite = false; // a graph of many launches per repetition
        // a stop flag that never rises: every kernel does its work; finite
        // step scalars, every channel active
        const double tol2 = -1.0;
        fillInt(4, 0, ints_.get(), s);
        fillInt(nch_, 1, active_.get(), s);
        fill(F64, 4L * nch_, 1.0, scal_.get(), s);
        fill(F64, static_cast<long>(parts_) * nch_, 1.0, partPq_.get(), s);
        const AmgState st{scal_.get(), ints_.get(), active_.get(), partPq_.get(), partRr_.get(), partRho_.get(), partBb_.get(), parts_, tol2};
        if (x != xDevice_)
        {
            xDevice_ = x;
            destroyGraphs();
        }
        typed([&]<typename P, int NCH>()
              {
                  P * rP = mixed_ ? reinterpret_cast<P *>(rP_.get()) : reinterpret_cast<P *>(r_.get());
                  P * z = reinterpret_cast<P *>(z_.get());
                  if (which == "direction")
                  {
                      bytes = bDir;
                      launch = [=, this]()
                      {
                          byLanes(levels_[0].lanes, [&]<int LN>()
                                  { amgDirectionKernel<P, LN, NCH><<<parts_, kBlock, 0, s>>>(a64_, z, p_.get(), q_.get(), st); });
                      };
                  }
                  else if (which == "update")
                  {
                      bytes = bUpd;
                      launch = [=, this]() { amgUpdateKernel<P, NCH><<<parts_, kBlock, 0, s>>>(n_, p_.get(), q_.get(), x, r_.get(), rP, st); };
                  }
                  else if (which == "rho")
                  {
                      bytes = bRho;
                      launch = [=, this]() { amgRhoKernel<P, NCH><<<parts_, kBlock, 0, s>>>(n_, r_.get(), z, st); };
                  }
                  else if (which == "vcycle")
                  {
                      // one V-cycle replayed from its own captured graph (its 50-odd
                      // short launches queued from the host would outlast them)
                      bytes = bV;
                      hipGraph_t g = nullptr;
                      NSFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                      enqueueVcycle<P, NCH>(s, rP, z);
                      NSFEM_HIP(hipStreamEndCapture(s, &g));
                      const hipError_t rc = hipGraphInstantiate(&vcycleExec, g, nullptr, nullptr, 0);
                      (void)hipGraphDestroy(g);
                      NSFEM_HIP(rc);
                      composite = true;
                      launch = [&vcycleExec, s]() { NSFEM_HIP(hipGraphLaunch(vcycleExec, s)); };
                  }
                  else if (which == "iteration")
                  {
                      bytes = bDir + bUpd + bRho + bV;
                      if (!iterExec_ || graphTol2_ != tol2)
                      {
                          destroyGraphs();
                          hipGraph_t g = nullptr;
                          NSFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                          enqueueIteration<P, NCH>(s, tol2);
                          NSFEM_HIP(hipStreamEndCapture(s, &g));
                          const hipError_t rc = hipGraphInstantiate(&iterExec_, g, nullptr, nullptr, 0);
                          (void)hipGraphDestroy(g);
                          NSFEM_HIP(rc);
                          graphTol2_ = tol2;
                      }
                      composite = true;
                      launch = [this, s]() { NSFEM_HIP(hipGraphLaunch(iterExec_, s)); };
                  }
                  else
                  {
                      throw std::invalid_argument("unknown amg_pcg kernel '" + which + "' (direction, update, rho, vcycle, iteration)");
                  }
              });
        launch(); // warm
        NSFEM_HIP(hipStreamSynchronize(s));
        if (!composite)
        {
            // single kernels: the stream idles on a 20 ms delay kernel while the
            // host queues the repetitions, so they run back to back and the
            // events bracket GPU execution only.  Graph replays (a V-cycle, an
            // iteration) outlast their host launch and run without it, as in a
            // solve (after 20 ms of one idle wave the clocks start lower).
            delayKernel<<<1, 64, 0, s>>>(2000000);
        }
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
        if (vcycleExec)
        {
            (void)hipGraphExecDestroy(vcycleExec);
        }
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

    void AmgPcg::enqueueSolve(hipStream_t s, const double * b, double * x, bool zeroGuess, double target, int budget)
    {
        const double tol2 = target * target;
        xDevice_ = x;
        destroyGraphs(); // the chunk graph may hold another solution vector
        typed([&]<typename P, int NCH>()
              {
                  enqueueInit<P, NCH>(s, b, x, zeroGuess, tol2);
                  for (int i = 0; i < budget; i++)
                  {
                      enqueueIteration<P, NCH>(s, tol2);
                  }
              });
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
        typed([&]<typename P, int NCH>() { enqueueInit<P, NCH>(s, b, x, zeroGuess, tol2); });
        auto iterate = [&]()
        {
            if (!useGraphs)
            {
                typed([&]<typename P, int NCH>() { enqueueIteration<P, NCH>(s, tol2); });
                return;
            }
            if (!iterExec_ || graphTol2_ != tol2)
            {
                destroyGraphs();
                hipGraph_t g = nullptr;
                NSFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try
                {
                    typed([&]<typename P, int NCH>() { enqueueIteration<P, NCH>(s, tol2); });
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
        // a burst up to a little before the last solve's count (at most 64
        // iterations: after a solve that took many more, the no-op
        // iterations of a long burst would cost more than the polls it
        // saves), then one iteration at a time with the stop flag polled one
        // iteration behind
        int launched = 0;
        const int burst = std::max(0, std::min({maxIters, predicted_ - 2, 64}));
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
        std::vector<double> sc(4 * nch_);
        int in[4];
        NSFEM_HIP(hipMemcpyAsync(sc.data(), scal_.get(), sc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipMemcpyAsync(in, ints_.get(), sizeof in, hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipStreamSynchronize(s));
        Result res;
        res.iterations = in[1];
        res.bb.assign(sc.begin() + 2 * nch_, sc.begin() + 3 * nch_);
        res.rr.assign(sc.begin() + 3 * nch_, sc.begin() + 4 * nch_);
        res.nonFinite = in[0] == 2;
        res.converged = in[0] == 1;
        if (in[2])
        {
            // b = 0 from a nonzero guess: x = 0 in those channels
            amgZeroChannelsKernel<<<std::max(1, std::min(4096, (n_ * nch_ + kBlock - 1) / kBlock)), kBlock, 0, s>>>(n_, nch_, in[2], x);
            NSFEM_HIP(hipGetLastError());
            NSFEM_HIP(hipStreamSynchronize(s));
            for (int c = 0; c < nch_; c++)
            {
                if ((in[2] >> c) & 1)
                {
                    res.rr[c] = 0;
                }
            }
        }
        predicted_ = res.iterations;
        return res;
    }
} // namespace nsfem
