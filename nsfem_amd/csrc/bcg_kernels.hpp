// Block PCG kernels: one Krylov space for a tile of NCH right-hand sides
// (NCH = 4, 8, 16, 32) with NCH x NCH step matrices (O'Leary's block CG),
// preconditioned by the multicolour symmetric GS sweeps of the single-RHS
// solver.  One system; vectors are colour-permuted and channel-interleaved,
// v[row * NCH + c], and one lane per channel works on a row (lane c of a
// group of NCH consecutive lanes), so a row's NCH values are one contiguous
// access and the other lanes' values of the same row come by shuffle.
//
// Per iteration, with P the search directions, Q = A P, R the residuals,
// Z = M^-1 R and rho = Z^T R:
//   product   Q = A P and the partial Gram matrices of P^T Q
//   alpha     gamma = P^T Q (summed in block order); alpha = gamma^-1 rho
//   update    X += P alpha, R -= Q alpha, column norms |R_c|^2, colour 0 of Z
//   sweeps    the forward and backward colour sweeps (shared kernels)
//   rho       backward colour 0 of Z and the partial Gram matrices of Z^T R
//   beta      rho' = Z^T R; convergence (every column |R_c| <= tol |B_c|);
//             beta = rho^-1 rho'
//   direction P = Z + P beta
// The small solves are Cholesky factorisations in one workgroup with a
// relative pivot threshold (1e-13 for P^T A P; for rho, which is
// symmetrised, one that matches the preconditioner's precision),
// deterministic like every reduction here (fixed summation order).
//
// Breakdown and verification.  Near the FP64 floor a block's directions
// become linearly dependent and the block recurrences lose accuracy (the
// residual stalls or grows).  Right-hand sides that are themselves dependent
// give a rank-deficient block from the start, which the dropped pivots
// handle.  The beta kernel therefore stops the iteration (flags[0] = 1) and
// asks the host for a restart (flags[3] = 1) when, after the cycle has cut
// its largest relative residual 10-fold, a factorisation keeps fewer
// directions than it did before (rank lost on the way; before such progress
// the lower rank becomes the cycle's reference), or when the largest
// relative residual grows 1000-fold over the best of the cycle; and it asks
// for a verification (flags[3] = 2) when the recursive residuals have
// converged.  Either way the host restarts from the
// current X with the true residual B - A X; the solve ends when the true
// residual of every column meets the tolerance (or at maxIters).
#pragma once

#include "kernels.hpp"

namespace nsfem
{
    struct BcgState
    {
        double * gramParts; // blocks * NCH * NCH
        double * normParts; // blocks * 2 * NCH: [bb | rr] per block
        double * gramTotal; // kBcgSlices * NCH * NCH: slice sums of gramParts
        double * normTotal; // kBcgSlices * 2 * NCH: slice sums of normParts
        double * rho;       // NCH * NCH
        double * step;      // NCH * NCH: alpha, then beta
        double * bb;        // NCH: |b_c|^2
        double * rr;        // NCH: |r_c|^2 after the last update
        double * best;      // [0]: smallest max_c |r_c|^2 / |b_c|^2 of the cycle, [1]: its value at the cycle's start
        int * flags;        // the solve flags ([0] stop, [3] restart 1 / verify 2), shared with the sweeps' state
        int * iters;        // [0]: updates done; [1]: unused; [2], [3]: the cycle's reference rank of
                            // the alpha and the beta factorisation (0: unset); [4]: alpha's last rank
        double tol2;
        // relative pivot threshold of the rho factorisation: rho = Z^T R
        // carries the preconditioner's rounding (about 1e-7 relative with an
        // FP32 preconditioner, where 1e-6 is used; 1e-13 with FP64)
        double rhoThreshold;
        int n;
        int blocks;         // workgroups of the row-wise kernels
        int rowsPerBlock;   // a multiple of kBlock / NCH
    };

    constexpr int kBcgSmallBlock = 1024;
    // The Gram partials of the row blocks are summed in kBcgSlices slices of
    // blocks (b = slice mod kBcgSlices) by as many workgroups per group of 64
    // entries; the small kernels add the slice sums in slice order.
    constexpr int kBcgSlices = 16;

    // The value of lane `a` of this thread's NCH-lane group.
    template <int NCH, typename V>
    __device__ __forceinline__ V groupLane(V v, int a)
    {
        const int lane = static_cast<int>(threadIdx.x) & 63;
        return __shfl(v, (lane & ~(NCH - 1)) + a, 64);
    }

    // acc[a] of thread (group g, lane c) holds the group's sum of u_ra * v_rc;
    // the workgroup's total goes to parts[blk * NCH * NCH + a * NCH + c].
    template <int NCH>
    __device__ __forceinline__ void gramStore(double (&acc)[NCH], double * parts, int blk, double * sm /* kWaves * NCH * NCH */)
    {
#pragma unroll
        for (int a = 0; a < NCH; a++)
        {
#pragma unroll
            for (int off = 32; off >= NCH; off >>= 1)
            {
                acc[a] += __shfl_xor(acc[a], off, 64);
            }
        }
        const int wave = static_cast<int>(threadIdx.x) >> 6;
        const int lane = static_cast<int>(threadIdx.x) & 63;
        if (lane < NCH)
        {
#pragma unroll
            for (int a = 0; a < NCH; a++)
            {
                sm[(wave * NCH + a) * NCH + lane] = acc[a];
            }
        }
        __syncthreads();
        for (int e = static_cast<int>(threadIdx.x); e < NCH * NCH; e += kBlock)
        {
            double s = 0;
#pragma unroll
            for (int w = 0; w < kWaves; w++)
            {
                s += sm[w * NCH * NCH + e];
            }
            parts[static_cast<size_t>(blk) * NCH * NCH + e] = s;
        }
    }

    // Q = A P (A: the permuted off-diagonal CSR plus its diagonal) and the
    // workgroup's part of P^T Q.
    template <typename T, int NCH>
    __global__ void __launch_bounds__(kBlock) bcgProductKernel(McgsMatrix<T> A, BcgState st, const T * __restrict__ p, T * __restrict__ q)
    {
        __shared__ double sm[kWaves * NCH * NCH];
        if (st.flags[0])
        {
            return;
        }
        const int blk = blockIdx.x;
        const int r0 = blk * st.rowsPerBlock;
        const int r1 = min(st.n, r0 + st.rowsPerBlock);
        const int c = static_cast<int>(threadIdx.x) % NCH;
        const int g = static_cast<int>(threadIdx.x) / NCH;
        double acc[NCH];
#pragma unroll
        for (int a = 0; a < NCH; a++)
        {
            acc[a] = 0;
        }
        for (int base = r0; base < r1; base += kBlock / NCH)
        {
            const int row = base + g;
            double pv = 0, qv = 0;
            if (row < r1)
            {
                const long i = static_cast<long>(row) * NCH + c;
                const T own = p[i];
                const T s = rowDotWide<T, NCH>(A, A.rowStart[row], A.rowStart[row + 1], p, c);
                const T qq = fmaT(A.diag[row], own, s);
                q[i] = qq;
                pv = static_cast<double>(own);
                qv = static_cast<double>(qq);
            }
#pragma unroll
            for (int a = 0; a < NCH; a++)
            {
                acc[a] = fma(groupLane<NCH>(pv, a), qv, acc[a]);
            }
        }
        gramStore<NCH>(acc, st.gramParts, blk, sm);
    }

    // Cholesky of the symmetric matrix G (LDS, NCH x NCH, lower triangle
    // overwritten by L) with a relative pivot threshold, then X = G^-1 B for
    // the NCH columns of B (LDS); a dropped pivot's row of X is zero.
    // Returns (in every thread) the number of dropped pivots.
    template <int NCH>
    __device__ int cholSolve(double * G, const double * B, double * X, double * piv, int * flag, double relThreshold)
    {
        const int t = static_cast<int>(threadIdx.x);
        if (t == 0)
        {
            double mx = 0;
            for (int j = 0; j < NCH; j++)
            {
                mx = fmax(mx, G[j * NCH + j]);
            }
            piv[NCH] = mx;
            *flag = 0;
        }
        __syncthreads();
        const double thr = relThreshold * piv[NCH];
        for (int j = 0; j < NCH; j++)
        {
            if (t == 0)
            {
                const double d = G[j * NCH + j];
                const bool keep = d > thr && d > 0;
                piv[j] = keep ? sqrt(d) : 0.0;
                if (!keep)
                {
                    *flag += 1;
                }
            }
            __syncthreads();
            if (t > j && t < NCH)
            {
                G[t * NCH + j] = piv[j] > 0 ? G[t * NCH + j] / piv[j] : 0.0;
            }
            __syncthreads();
            for (int e = t; e < NCH * NCH; e += blockDim.x)
            {
                const int i = e / NCH, k = e % NCH;
                if (i > j && k > j && k <= i)
                {
                    G[i * NCH + k] -= G[i * NCH + j] * G[k * NCH + j];
                }
            }
            __syncthreads();
        }
        if (t < NCH)
        {
            // L y = b, then L^T x = y, for column t
            double y[NCH];
            for (int j = 0; j < NCH; j++)
            {
                double s = B[j * NCH + t];
                for (int i = 0; i < j; i++)
                {
                    s -= G[j * NCH + i] * y[i];
                }
                y[j] = piv[j] > 0 ? s / piv[j] : 0.0;
            }
            for (int j = NCH - 1; j >= 0; j--)
            {
                double s = y[j];
                for (int i = j + 1; i < NCH; i++)
                {
                    s -= G[i * NCH + j] * y[i];
                }
                y[j] = piv[j] > 0 ? s / piv[j] : 0.0;
                X[j * NCH + t] = y[j];
            }
        }
        __syncthreads();
        return *flag;
    }

    // Slice sums of the Gram partials (and, with norms, of the norm
    // partials): workgroup (w, slice) sums entries [64 w, 64 w + 64) over the
    // row blocks b = slice (mod kBcgSlices), four interleaved sub-slices per
    // entry combined in a fixed order; the workgroups after the Gram ones sum
    // the 2 NCH norm entries.
    template <int NCH>
    __global__ void __launch_bounds__(kBlock) bcgGramReduceKernel(BcgState st, int norms)
    {
        __shared__ double sm[kBlock];
        if (st.flags[0])
        {
            return;
        }
        constexpr int kGramGroups = (NCH * NCH + 63) / 64;
        const int w = blockIdx.x / kBcgSlices;
        const int slice = blockIdx.x % kBcgSlices;
        const int t = static_cast<int>(threadIdx.x);
        const int el = t & 63, sub = t >> 6;
        const bool gram = w < kGramGroups;
        if (!gram && !norms)
        {
            return;
        }
        const int e = gram ? w * 64 + el : el;
        const int width = gram ? NCH * NCH : 2 * NCH;
        const double * parts = gram ? st.gramParts : st.normParts;
        double acc = 0;
        if (e < width)
        {
            constexpr int stride = kBcgSlices * (kBlock / 64);
            for (int b = slice + kBcgSlices * sub; b < st.blocks; b += stride)
            {
                acc += parts[static_cast<size_t>(b) * width + e];
            }
        }
        sm[t] = acc;
        __syncthreads();
        if (sub == 0 && e < width)
        {
            double s = 0;
#pragma unroll
            for (int k = 0; k < kBlock / 64; k++)
            {
                s += sm[k * 64 + el];
            }
            (gram ? st.gramTotal : st.normTotal)[static_cast<size_t>(slice) * width + e] = s;
        }
    }

    // Sum of the slice totals of entry e (width entries per slice).
    __device__ __forceinline__ double sliceSum(const double * totals, int width, int e)
    {
        double s = 0;
#pragma unroll
        for (int k = 0; k < kBcgSlices; k++)
        {
            s += totals[static_cast<size_t>(k) * width + e];
        }
        return s;
    }

    // alpha = (P^T Q)^-1 rho
    template <int NCH>
    __global__ void __launch_bounds__(kBcgSmallBlock) bcgAlphaKernel(BcgState st)
    {
        __shared__ double G[NCH * NCH], B[NCH * NCH], X[NCH * NCH], piv[NCH + 1];
        __shared__ int dropped;
        if (st.flags[0])
        {
            return;
        }
        for (int e = static_cast<int>(threadIdx.x); e < NCH * NCH; e += blockDim.x)
        {
            G[e] = sliceSum(st.gramTotal, NCH * NCH, e);
            B[e] = st.rho[e];
        }
        __syncthreads();
        const int kept = NCH - cholSolve<NCH>(G, B, X, piv, &dropped, 1e-13);
        for (int e = static_cast<int>(threadIdx.x); e < NCH * NCH; e += blockDim.x)
        {
            st.step[e] = X[e];
        }
        if (threadIdx.x == 0)
        {
            st.iters[4] = kept; // judged by the beta kernel after this update
            if (st.iters[2] == 0)
            {
                st.iters[2] = kept; // the cycle's reference rank
            }
        }
    }

    // X += P alpha, R -= Q alpha; the preconditioner's copy of R and its
    // colour-0 values (forward sweep from zero: r / d); column norms.
    template <typename T, typename P, int NCH>
    __global__ void __launch_bounds__(kBlock) bcgUpdateKernel(BcgState st, const P * __restrict__ invDiagP, const T * __restrict__ p,
                                                              const T * __restrict__ q, T * __restrict__ x, T * r, P * rP,
                                                              P * __restrict__ z, int color0End)
    {
        __shared__ double al[NCH * NCH];
        __shared__ double sm[NCH * kWaves];
        if (st.flags[0])
        {
            return;
        }
        for (int e = static_cast<int>(threadIdx.x); e < NCH * NCH; e += kBlock)
        {
            al[e] = st.step[e];
        }
        __syncthreads();
        const int blk = blockIdx.x;
        const int r0 = blk * st.rowsPerBlock;
        const int r1 = min(st.n, r0 + st.rowsPerBlock);
        const int c = static_cast<int>(threadIdx.x) % NCH;
        const int g = static_cast<int>(threadIdx.x) / NCH;
        double norm = 0;
        for (int base = r0; base < r1; base += kBlock / NCH)
        {
            const int row = base + g;
            const bool valid = row < r1;
            const long i = static_cast<long>(row) * NCH + c;
            const double pv = valid ? static_cast<double>(p[i]) : 0.0;
            const double qv = valid ? static_cast<double>(q[i]) : 0.0;
            double dx = 0, dr = 0;
#pragma unroll
            for (int a = 0; a < NCH; a++)
            {
                const double w = al[a * NCH + c];
                dx = fma(groupLane<NCH>(pv, a), w, dx);
                dr = fma(groupLane<NCH>(qv, a), w, dr);
            }
            if (valid)
            {
                x[i] = static_cast<T>(static_cast<double>(x[i]) + dx);
                const T rn = static_cast<T>(static_cast<double>(r[i]) - dr);
                r[i] = rn;
                const P rp = static_cast<P>(rn);
                if constexpr (!std::is_same_v<T, P>)
                {
                    rP[i] = rp;
                }
                if (row < color0End)
                {
                    z[i] = rp * invDiagP[row];
                }
                norm = fma(static_cast<double>(rn), static_cast<double>(rn), norm);
            }
        }
        const double total = channelBlockSum<NCH>(norm, sm);
        storeChannelPartial<NCH, 2 * NCH>(total, st.normParts, blk, NCH);
    }

    // Backward colour 0 of the symmetric sweep (full rows) and the
    // workgroup's part of Z^T R.
    template <typename P, typename T, int NCH>
    __global__ void __launch_bounds__(kBlock) bcgRhoKernel(McgsMatrix<P> M, BcgState st, const P * rP, const T * __restrict__ r,
                                                           P * __restrict__ z, int color0End)
    {
        __shared__ double sm[kWaves * NCH * NCH];
        if (st.flags[0])
        {
            return;
        }
        const int blk = blockIdx.x;
        const int r0 = blk * st.rowsPerBlock;
        const int r1 = min(st.n, r0 + st.rowsPerBlock);
        const int c = static_cast<int>(threadIdx.x) % NCH;
        const int g = static_cast<int>(threadIdx.x) / NCH;
        double acc[NCH];
#pragma unroll
        for (int a = 0; a < NCH; a++)
        {
            acc[a] = 0;
        }
        for (int base = r0; base < r1; base += kBlock / NCH)
        {
            const int row = base + g;
            double zv = 0, rv = 0;
            if (row < r1)
            {
                const long i = static_cast<long>(row) * NCH + c;
                P zi;
                if (row < color0End)
                {
                    const P s = rowDotWide<P, NCH>(M, M.rowStart[row], M.rowStart[row + 1], z, c);
                    zi = (rP[i] - s) * M.invDiag[row];
                    z[i] = zi;
                }
                else
                {
                    zi = z[i];
                }
                zv = static_cast<double>(zi);
                rv = static_cast<double>(r[i]);
            }
#pragma unroll
            for (int a = 0; a < NCH; a++)
            {
                acc[a] = fma(groupLane<NCH>(zv, a), rv, acc[a]);
            }
        }
        gramStore<NCH>(acc, st.gramParts, blk, sm);
    }

    // Initial residual R = B - A X (or B for a zero guess), its preconditioner
    // copy and colour-0 values, and the column norms of B and R.
    template <typename T, typename P, int NCH>
    __global__ void __launch_bounds__(kBlock) bcgInitKernel(McgsMatrix<T> A, BcgState st, const T * __restrict__ b, T * __restrict__ x,
                                                            T * r, P * rP, P * __restrict__ z, const P * __restrict__ invDiagP,
                                                            int color0End, int zeroGuess)
    {
        __shared__ double sm[2 * NCH * kWaves];
        const int blk = blockIdx.x;
        const int r0 = blk * st.rowsPerBlock;
        const int r1 = min(st.n, r0 + st.rowsPerBlock);
        const int c = static_cast<int>(threadIdx.x) % NCH;
        const int g = static_cast<int>(threadIdx.x) / NCH;
        double bb = 0, rr = 0;
        for (int base = r0; base < r1; base += kBlock / NCH)
        {
            const int row = base + g;
            if (row < r1)
            {
                const long i = static_cast<long>(row) * NCH + c;
                const T bi = b[i];
                T ri = bi;
                if (zeroGuess)
                {
                    x[i] = T(0);
                }
                else
                {
                    const T s = rowDotWide<T, NCH>(A, A.rowStart[row], A.rowStart[row + 1], x, c);
                    ri = bi - fmaT(A.diag[row], x[i], s);
                }
                r[i] = ri;
                const P rp = static_cast<P>(ri);
                if constexpr (!std::is_same_v<T, P>)
                {
                    rP[i] = rp;
                }
                if (row < color0End)
                {
                    z[i] = rp * invDiagP[row];
                }
                bb = fma(static_cast<double>(bi), static_cast<double>(bi), bb);
                rr = fma(static_cast<double>(ri), static_cast<double>(ri), rr);
            }
        }
        channelBlockSum2<NCH>(bb, rr, sm);
        storeChannelPartial<NCH, 2 * NCH>(bb, st.normParts, blk, 0);
        storeChannelPartial<NCH, 2 * NCH>(rr, st.normParts, blk, NCH);
    }

    // rho' = Z^T R and the column norms.  init (after a (re)start on the
    // true residual): the first rho of the cycle and, the first time, the
    // norms of B; converged columns all meeting the tolerance end the solve.
    // Otherwise: converged recursive residuals ask for a verification, a
    // dropped pivot or a 1000-fold growth of the residual for a restart (both
    // stop the iteration: flags[0] = 1, flags[3] = 2 or 1); else beta =
    // rho^-1 rho' and rho = rho'.
    template <int NCH>
    __global__ void __launch_bounds__(kBcgSmallBlock) bcgBetaKernel(BcgState st, int init)
    {
        __shared__ double G[NCH * NCH], Rn[NCH * NCH], X[NCH * NCH], piv[NCH + 1];
        __shared__ double worst;
        __shared__ int dropped, allDone, bad;
        if (st.flags[0])
        {
            return;
        }
        const int t = static_cast<int>(threadIdx.x);
        for (int e = t; e < NCH * NCH; e += blockDim.x)
        {
            X[e] = sliceSum(st.gramTotal, NCH * NCH, e);
        }
        __syncthreads();
        for (int e = t; e < NCH * NCH; e += blockDim.x)
        {
            const int i = e / NCH, k = e % NCH;
            Rn[e] = 0.5 * (X[e] + X[k * NCH + i]); // Z^T R is symmetric up to the preconditioner's rounding
        }
        if (t == 0)
        {
            allDone = 1;
            bad = 0;
            worst = 0;
        }
        __syncthreads();
        if (t == 0)
        {
            for (int c = 0; c < NCH; c++)
            {
                const double rr = sliceSum(st.normTotal, 2 * NCH, NCH + c);
                double bb = 0;
                if (init == 2)
                {
                    bb = sliceSum(st.normTotal, 2 * NCH, c);
                }
                if (init == 2)
                {
                    st.bb[c] = bb;
                }
                else
                {
                    bb = st.bb[c];
                }
                st.rr[c] = rr;
                const double ref = bb > 0 ? bb : 1.0;
                if (!(rr <= st.tol2 * ref))
                {
                    allDone = 0;
                }
                if (!isfinite(rr))
                {
                    bad = 1;
                }
                worst = fmax(worst, rr / ref);
            }
        }
        __syncthreads();
        if (bad)
        {
            if (t == 0)
            {
                st.flags[2] = 1; // non-finite residual: stop
                st.flags[0] = 1;
            }
            return;
        }
        if (!init && t == 0)
        {
            st.iters[0] += 1;
        }
        if (allDone)
        {
            if (t == 0)
            {
                st.flags[3] = init ? 0 : 2; // converged on the true residual, or verify
                st.flags[0] = 1;
            }
            return;
        }
        if (init)
        {
            for (int e = t; e < NCH * NCH; e += blockDim.x)
            {
                st.rho[e] = Rn[e];
            }
            if (t == 0)
            {
                st.best[0] = worst;
                st.best[1] = worst;
            }
            return;
        }
        const double best = fmin(st.best[0], worst);
        if (worst > 1e6 * best)
        {
            if (t == 0)
            {
                st.flags[3] = 1; // the residual grew 1000-fold: restart
                st.flags[0] = 1;
            }
            return;
        }
        for (int e = t; e < NCH * NCH; e += blockDim.x)
        {
            G[e] = st.rho[e];
        }
        __syncthreads();
        const int kept = NCH - cholSolve<NCH>(G, Rn, X, piv, &dropped, st.rhoThreshold);
        // rank lost after progress: restart; before it: the new reference
        const bool progressed = best < 1e-2 * st.best[1];
        const int keptA = st.iters[4];
        const bool lost = keptA < st.iters[2] || (st.iters[3] != 0 && kept < st.iters[3]);
        if (lost && progressed)
        {
            if (t == 0)
            {
                st.flags[3] = 1;
                st.flags[0] = 1;
            }
            return;
        }
        __syncthreads();
        if (t == 0)
        {
            st.iters[2] = min(st.iters[2], keptA);
            st.iters[3] = st.iters[3] == 0 ? kept : min(st.iters[3], kept);
        }
        for (int e = t; e < NCH * NCH; e += blockDim.x)
        {
            st.step[e] = X[e];
            st.rho[e] = Rn[e];
        }
        if (t == 0)
        {
            st.best[0] = best;
        }
    }

    // P = Z + P beta (first: P = Z).
    template <typename T, typename P, int NCH>
    __global__ void __launch_bounds__(kBlock) bcgDirectionKernel(BcgState st, const P * __restrict__ z, T * p, int first)
    {
        __shared__ double be[NCH * NCH];
        if (st.flags[0])
        {
            return;
        }
        if (!first)
        {
            for (int e = static_cast<int>(threadIdx.x); e < NCH * NCH; e += kBlock)
            {
                be[e] = st.step[e];
            }
        }
        __syncthreads();
        const int blk = blockIdx.x;
        const int r0 = blk * st.rowsPerBlock;
        const int r1 = min(st.n, r0 + st.rowsPerBlock);
        const int c = static_cast<int>(threadIdx.x) % NCH;
        const int g = static_cast<int>(threadIdx.x) / NCH;
        for (int base = r0; base < r1; base += kBlock / NCH)
        {
            const int row = base + g;
            const bool valid = row < r1;
            const long i = static_cast<long>(row) * NCH + c;
            const double pv = (valid && !first) ? static_cast<double>(p[i]) : 0.0;
            double s = 0;
            if (!first)
            {
#pragma unroll
                for (int a = 0; a < NCH; a++)
                {
                    s = fma(groupLane<NCH>(pv, a), be[a * NCH + c], s);
                }
            }
            if (valid)
            {
                p[i] = static_cast<T>(static_cast<double>(z[i]) + s);
            }
        }
    }
} // namespace nsfem
