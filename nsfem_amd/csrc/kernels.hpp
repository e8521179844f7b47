// Device kernels for the sparse hot path on gfx950 (wave64).
//
// Layout conventions
//   * CSR: int32 rowStart/column, values of type T.  Rows are handled by a
//     group of LANES consecutive lanes of one wavefront (LANES | 64): the
//     lanes stride over the row, so one wavefront reads 64/LANES adjacent row
//     segments as one coalesced stream, and the row sum is finished with
//     xor-shuffles inside the lane group.
//   * Multi-channel vectors are planar: channel c of vector v is v[c*ld + i]
//     (the reference's [ux; uy] layout, chorinCuda.cu:189-196).
//   * Reductions are deterministic and fence-free: each block writes its
//     partial, a one-block finalize kernel sums them in block-map order.
//   * Solver kernels walk a block map: every block owns one row range that
//     lies inside one (colour, system) chunk of the colour-major permuted
//     system, so colour sweeps launch only their colour's blocks and every
//     block knows which system's scalars it updates.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace nsfem
{
    constexpr int kBlock = 256;
    constexpr int kWaves = kBlock / 64;
    // One-block finalize kernels run 1024 threads: more partial sums in
    // flight per round trip, more systems reduced side by side.
    constexpr int kFinBlock = 1024;
    constexpr int kFinWaves = kFinBlock / 64;

    template <typename T>
    __device__ __forceinline__ T fmaT(T a, T b, T c);
    template <>
    __device__ __forceinline__ double fmaT<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
    template <>
    __device__ __forceinline__ float fmaT<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

    template <int LANES, typename T>
    __device__ __forceinline__ T laneGroupSum(T v)
    {
#pragma unroll
        for (int off = LANES / 2; off > 0; off >>= 1)
        {
            v += __shfl_xor(v, off, 64);
        }
        return v;
    }

    // Workgroups are dispatched round-robin over the 8 XCDs (each with its own
    // L2).  Map the dispatch index to a logical block so that every XCD works
    // on one contiguous range of logical blocks (neighbouring rows share
    // their x entries); a bijection on [0, G) for any G.
    __device__ __forceinline__ int logicalBlock(int p, int G)
    {
        const int q = G >> 3, r = G & 7, x = p & 7, i = p >> 3;
        return x * q + (x < r ? x : r) + i;
    }

    // Bijection on [0, B) placing the first B0 indices evenly among the rest:
    // slot l holds index floor(l B0 / B) where that value steps up, else the
    // next of the remaining B - B0 indices, in order.
    __device__ __forceinline__ int interleaveFirst(int l, int B, int B0)
    {
        const long c = static_cast<long>(l) * B0 / B;
        const long c1 = static_cast<long>(l + 1) * B0 / B;
        return c1 > c ? static_cast<int>(c) : B0 + (l - static_cast<int>(c));
    }

    // Sum of v over the block (W wavefronts); result valid in thread 0.
    template <int K, int W = kWaves>
    __device__ __forceinline__ void blockSum(double (&v)[K], double * sm /* K * W */)
    {
#pragma unroll
        for (int k = 0; k < K; k++)
        {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
            {
                v[k] += __shfl_xor(v[k], off, 64);
            }
        }
        const int wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0)
        {
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                sm[k * W + wave] = v[k];
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                double s = 0;
                for (int w = 0; w < W; w++)
                {
                    s += sm[k * W + w];
                }
                v[k] = s;
            }
        }
        __syncthreads();
    }

    // Block partial -> partials[slot*K + k] (thread 0, after blockSum).  The
    // partials are summed by a one-block finalize kernel launched next on the
    // same stream: the kernel boundary makes them visible across XCDs, which
    // a device-scope fence inside the producer would do per block at the
    // price of an L2 write-back and invalidate on every block.
    template <int K>
    __device__ __forceinline__ void storePartial(const double (&v)[K], double * partials, int slot)
    {
        if (threadIdx.x == 0)
        {
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                partials[static_cast<size_t>(slot) * K + k] = v[k];
            }
        }
    }

    // ------------------------------------------------------------------
    // Plain CSR products: y = alpha * A x (+ beta y), NCH planar channels.
    // rowIds != nullptr: output row i is A's row rowIds[i] (fused gather).
    // ------------------------------------------------------------------
    template <typename T, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) csrMulKernel(int rows, const int * __restrict__ rowStart,
                                                            const int * __restrict__ column, const T * __restrict__ values,
                                                            const T * __restrict__ x, long ldx, T * __restrict__ y, long ldy,
                                                            T alpha, T beta, const int * __restrict__ rowIds)
    {
        const long gid = static_cast<long>(logicalBlock(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
        const int row = static_cast<int>(gid / LANES);
        const int lane = threadIdx.x & (LANES - 1);
        T acc[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            acc[c] = T(0);
        }
        if (row < rows)
        {
            const int src = rowIds ? rowIds[row] : row;
            const int j1 = rowStart[src + 1];
            // U strided entries per step, loaded unconditionally (an entry
            // past the row's end reloads entry j and its term is dropped), so
            // their loads, then their gathers, go out together; each lane
            // adds its terms in the same order as one by one.  Lane groups
            // of 8 or more take one entry per step: their lanes see about one
            // entry of a row each, and the reloads measured slower (P2 rows,
            // 8 lanes: 155 -> 171 us at 2M rows).
            constexpr int U = LANES <= 2 ? 4 : (LANES <= 4 ? 2 : 1);
            for (int j = rowStart[src] + lane; j < j1; j += U * LANES)
            {
                T v[U];
                int col[U];
                bool has[U];
#pragma unroll
                for (int u = 0; u < U; u++)
                {
                    const int ju = j + u * LANES;
                    has[u] = ju < j1;
                    const int js = has[u] ? ju : j;
                    v[u] = values[js];
                    col[u] = column[js];
                }
                T xv[U][NCH];
#pragma unroll
                for (int u = 0; u < U; u++)
                {
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        xv[u][c] = x[col[u] + c * ldx];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++)
                {
                    if (has[u])
                    {
#pragma unroll
                        for (int c = 0; c < NCH; c++)
                        {
                            acc[c] = fmaT(v[u], xv[u][c], acc[c]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            acc[c] = laneGroupSum<LANES>(acc[c]);
        }
        if (row < rows && lane == 0)
        {
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                T out = alpha * acc[c];
                if (beta != T(0))
                {
                    out = fmaT(beta, y[row + c * ldy], out);
                }
                y[row + c * ldy] = out;
            }
        }
    }

    // ------------------------------------------------------------------
    // BLAS-1 style kernels (grid-stride, n up to 2^31 * channels)
    // ------------------------------------------------------------------
    template <typename T>
    __global__ void axpyKernel(long n, T alpha, const T * __restrict__ src, T * __restrict__ dst)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            dst[i] = fmaT(alpha, src[i], dst[i]);
        }
    }

    template <typename T>
    __global__ void scaleKernel(long n, T alpha, T * __restrict__ dst)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            dst[i] *= alpha;
        }
    }

    template <typename T>
    __global__ void geamKernel(long n, T alpha, const T * __restrict__ a, T beta, const T * __restrict__ b, T * __restrict__ c)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            c[i] = alpha * a[i] + beta * b[i];
        }
    }

    // x <- 2 x - prev, prev <- x (old): a linear extrapolation in time of a
    // solver's warm start, keeping the history for the next step
    template <typename T>
    __global__ void extrapolateKernel(long n, T * __restrict__ x, T * __restrict__ prev)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            const T a = x[i];
            x[i] = 2 * a - prev[i];
            prev[i] = a;
        }
    }

    // out <- 2 cur - prev, prev <- cur: an extrapolated initial guess into a
    // solver's vector, keeping the history
    template <typename T>
    __global__ void extrapolateToKernel(long n, const T * __restrict__ cur, T * __restrict__ prev, T * __restrict__ out)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            const T a = cur[i];
            out[i] = 2 * a - prev[i];
            prev[i] = a;
        }
    }

    // One wave that idles for the given number of ticks of the constant
    // 100 MHz wall clock, then exits (holds a stream while the host queues
    // work behind it, so that work runs back to back).
    static __global__ void delayKernel(long ticks)
    {
        const long t0 = static_cast<long>(wall_clock64());
        while (static_cast<long>(wall_clock64()) - t0 < ticks)
        {
            __builtin_amdgcn_s_sleep(127);
        }
    }

    // dst = src over n 16-byte elements, one per thread and grid-stride
    // (the achievable streaming bandwidth of the device).
    static __global__ void streamCopyKernel(long n, const double2 * __restrict__ src, double2 * __restrict__ dst)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            dst[i] = src[i];
        }
    }

    template <typename T>
    __global__ void fillKernel(long n, T value, T * __restrict__ dst)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            dst[i] = value;
        }
    }

    template <typename S, typename D>
    __global__ void convertKernel(long n, const S * __restrict__ src, D * __restrict__ dst)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            dst[i] = static_cast<D>(src[i]);
        }
    }

    template <typename T>
    __global__ void scatterKernel(int nnz, const int * __restrict__ idx, const T * __restrict__ vals, T * __restrict__ dense)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < nnz)
        {
            dense[idx[i]] = vals[i];
        }
    }

    template <typename T>
    __global__ void gatherKernel(int nnz, const T * __restrict__ dense, const int * __restrict__ idx, T * __restrict__ vals)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < nnz)
        {
            vals[i] = dense[idx[i]];
        }
    }

    // Two-level dot products: K pairs (x_k, y_k) of length n each.
    template <typename T, int K>
    __global__ void __launch_bounds__(kBlock) dotKernel(long n, const T * __restrict__ x, long ldx, const T * __restrict__ y,
                                                         long ldy, double * partials)
    {
        __shared__ double sm[K * kWaves];
        double acc[K];
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            acc[k] = 0;
        }
        for (long i = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * kBlock)
        {
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                acc[k] = __builtin_fma(static_cast<double>(x[i + k * ldx]), static_cast<double>(y[i + k * ldy]), acc[k]);
            }
        }
        blockSum<K>(acc, sm);
        storePartial<K>(acc, partials, blockIdx.x);
    }

    // One block: out[k] = sum over the producer's blocks, in block order.
    template <int K>
    __global__ void __launch_bounds__(kBlock) dotFinalizeKernel(int blocks, const double * __restrict__ partials, double * out)
    {
        __shared__ double sm[K * kWaves];
        double tot[K];
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            tot[k] = 0;
        }
        for (int b = threadIdx.x; b < blocks; b += kBlock)
        {
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                tot[k] += partials[static_cast<size_t>(b) * K + k];
            }
        }
        blockSum<K>(tot, sm);
        if (threadIdx.x == 0)
        {
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                out[k] = tot[k];
            }
        }
    }

    // ------------------------------------------------------------------
    // Colour permutation between planar io vectors (channel c of row r at
    // c*n + r) and the solvers' interleaved internal vectors (row i of the
    // colour order, channel c at i*nch + c).
    // fwd: dst[i*nch + c] = src[c*n + perm[i]];   inv: dst[c*n + perm[i]] = src[i*nch + c]
    // ------------------------------------------------------------------
    template <typename S, typename D>
    __global__ void permuteKernel(int n, int nch, const int * __restrict__ perm, const S * __restrict__ src,
                                  D * __restrict__ dst, int inverse)
    {
        const long total = static_cast<long>(n) * nch;
        for (long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; t < total; t += static_cast<long>(gridDim.x) * blockDim.x)
        {
            const int i = static_cast<int>(t / nch);
            const int c = static_cast<int>(t - static_cast<long>(i) * nch);
            const long p = static_cast<long>(c) * n + perm[i];
            if (inverse)
            {
                dst[p] = static_cast<D>(src[t]);
            }
            else
            {
                dst[t] = static_cast<D>(src[p]);
            }
        }
    }

    // ------------------------------------------------------------------
    // Multicolour Gauss-Seidel / PCG engine
    // ------------------------------------------------------------------
    template <typename T>
    struct McgsMatrix
    {
        int n;
        const int * rowStart; // stripped (off-diagonal) CSR
        const int * rowSplit; // end of the lower-colour part of each row
        const int * column;
        const T * values;
        const T * diag;
        const T * invDiag;
    };

    struct BlockMap
    {
        const int4 * blocks;       // per block: rowBegin, rowEnd, system, partial slot (colour-major)
        const int * sysBlockStart; // numSystems + 1
        const int * sysBlocks;     // blocks grouped by system
        const int * systemRows;    // rows per system (for RMS norms)
        int numSystems;
        int numBlocks;
        // flatRows > 0 (one system): block b is rows [b*rowsPerBlock, ...)
        // of [0, flatRows) (of the launch's colour for the sweeps), its slot
        // is b; the kernels compute the range instead of loading it.
        int flatRows;
        int rowsPerBlock;
        // flat maps only, optional: block b is rows rowTable[b] instead
        // (same slot b); the region-major direction map, where the blocks
        // running on one XCD take the same stretch of every colour
        const int2 * rowTable;
    };

    // The rows, system and partial slot of block lb.
    __device__ __forceinline__ int4 blockRange(const BlockMap & bm, int lb)
    {
        if (bm.flatRows > 0)
        {
            if (bm.rowTable)
            {
                const int2 t = bm.rowTable[lb];
                return make_int4(t.x, t.y, 0, lb);
            }
            const int r0 = lb * bm.rowsPerBlock;
            return make_int4(r0, min(bm.flatRows, r0 + bm.rowsPerBlock), 0, lb);
        }
        return bm.blocks[lb];
    }

    // The permutation of permuteKernel over the rows of a block map: block lb
    // moves rows blockRange(bm, lb).  On the region-major direction map the
    // workgroups of one XCD take stretch k of every colour, whose rows come
    // from one stretch of the original order, so the gathered (forward) or
    // scattered (inverse) side stays in that XCD's L2 instead of every
    // colour sweeping the whole vector.
    template <typename S, typename D, int NCH>
    __global__ void __launch_bounds__(kBlock) permuteMappedKernel(BlockMap bm, int n, const int * __restrict__ perm,
                                                                  const S * __restrict__ src, D * __restrict__ dst, int inverse)
    {
        const int4 range = blockRange(bm, logicalBlock(blockIdx.x, gridDim.x));
        for (int row = range.x + static_cast<int>(threadIdx.x); row < range.y; row += kBlock)
        {
            const long p = perm[row];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                const long i = static_cast<long>(row) * NCH + c;
                const long o = static_cast<long>(c) * n + p;
                if (inverse)
                {
                    dst[o] = static_cast<D>(src[i]);
                }
                else
                {
                    dst[i] = static_cast<D>(src[o]);
                }
            }
        }
    }

    // The inverse direction of the permutation as a gather: original row r
    // takes colour-major row iperm[r], dst[c*n + r] = src[iperm[r]*NCH + c]
    // (iperm the inverse of perm).  The writes are whole lines in order, and
    // the contiguous range of original rows one XCD takes (logicalBlock)
    // reads from the matching stretch of each colour; scattering the same
    // permutation's 8-byte writes takes twice as long.
    template <typename S, typename D, int NCH>
    __global__ void __launch_bounds__(kBlock) unpermuteGatherKernel(int n, const int * __restrict__ iperm,
                                                                    const S * __restrict__ src, D * __restrict__ dst)
    {
        const long r = static_cast<long>(logicalBlock(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
        if (r >= n)
        {
            return;
        }
        const long q = static_cast<long>(iperm[r]) * NCH;
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            dst[static_cast<long>(c) * n + r] = static_cast<D>(src[q + c]);
        }
    }

    // Block lb of a colour sweep launch: rows [rowBegin, rowEnd) of the colour
    // (flat map) or table entry blockOffset + lb.
    __device__ __forceinline__ int4 sweepRange(const BlockMap & bm, int blockOffset, int lb, int rowBegin, int rowEnd)
    {
        if (bm.flatRows > 0)
        {
            const int r0 = rowBegin + lb * bm.rowsPerBlock;
            return make_int4(r0, min(rowEnd, r0 + bm.rowsPerBlock), 0, 0);
        }
        return bm.blocks[blockOffset + lb];
    }

    // Finalize kernels stage the per-channel solver scalars in LDS while the
    // partial sums stream in, for up to this many (system, channel) pairs.
    constexpr int kStaged = 64;

    // Per (system, channel) scalars, index s*NCH + c, plus global flags.
    struct SolveState
    {
        double * rho;
        double * alpha;
        double * beta;
        double * rr;    // |r|^2 after the last update
        double * bb;    // |b|^2
        double * ref;   // reference |.|^2 the tolerance is relative to
        double * mse;   // RMS residual (gs mode)
        int * active;   // 1 while the channel iterates
        int * iters;    // iterations done / iteration index at convergence
        int * flags;    // [0] done, [1] first (p = z), [2] non-finite seen, [3] unused
        double * partials;  // blocks * 2 * NCH
        double * partials2; // blocks * NCH: |r|^2 of the update, read with rho's
        double * partials3; // blocks * NCH: p.q of the direction kernel (fused reductions)
        double * rhoNext;   // fused reductions: rho and activity decided by the
        int * activeNext;   //   direction kernel, taken over by the update kernel
        double tol2;        // relative tolerance squared (pcg) or target RMS (gs)
    };


    // Fused reductions: the partial slots of a block's system (one system: all
    // of them; a batch: the system's contiguous slot range) and whether the
    // block records the system's decision (its first block).
    struct FusedSlice
    {
        int s;        // system
        int base;     // first partial slot of the system
        int rhoParts; // rho kernel partials of the system
        int updParts; // update kernel partials of the system
        bool leader;
    };

    __device__ __forceinline__ FusedSlice fusedSlice(const BlockMap & bm, int4 range, int lb, int rhoParts, int updParts)
    {
        FusedSlice f;
        f.s = range.z;
        if (bm.flatRows > 0)
        {
            f.base = 0;
            f.rhoParts = rhoParts;
            f.updParts = updParts;
            f.leader = lb == 0;
        }
        else
        {
            f.base = bm.sysBlockStart[f.s];
            f.rhoParts = f.updParts = bm.sysBlockStart[f.s + 1] - f.base;
            f.leader = range.w == f.base;
        }
        return f;
    }

    // Registers of the fused reductions' loads (fusedRhoLoad / fusedRhoFinish).
    constexpr int kFusedPre = 8;
    template <int NCH>
    struct FusedRhoLoads
    {
        int first, was;
        double ref, rhoOld;
        double rho[kFusedPre], rr[kFusedPre];
    };

    // Finishes a per-system reduction in a one-block finalize kernel of
    // kFinBlock threads: out[s*K + k] is the sum over the blocks of system s
    // of their partials.  Producers store a block's partial at its slot (the
    // block's position in the system-grouped order, block table entry .w), so
    // each system's partials are contiguous: [sysBlockStart[s], sysBlockStart[s+1]).
    // The summation order depends only on the block map: results repeat.
    template <int K>
    __device__ __forceinline__ void systemTotals(const BlockMap & bm, const double * partials, double * out, double * sm)
    {
        // at most 8 sums at a time keep the accumulators in registers
        constexpr int G = K < 8 ? K : 8;
        static_assert(K % G == 0, "sum groups must divide K");
        for (int k0 = 0; k0 < K; k0 += G)
        {
            if (bm.numSystems == 1)
            {
                double acc[G];
#pragma unroll
                for (int k = 0; k < G; k++)
                {
                    acc[k] = 0;
                }
                // independent loads: keep many in flight
                const int nb = bm.numBlocks;
#pragma unroll 8
                for (int b = threadIdx.x; b < nb; b += kFinBlock)
                {
#pragma unroll
                    for (int k = 0; k < G; k++)
                    {
                        acc[k] += partials[static_cast<size_t>(b) * K + k0 + k];
                    }
                }
                blockSum<G, kFinWaves>(acc, sm);
                if (threadIdx.x == 0)
                {
#pragma unroll
                    for (int k = 0; k < G; k++)
                    {
                        out[k0 + k] = acc[k];
                    }
                }
                __syncthreads();
                continue;
            }
            // One wavefront per system, lanes stride the system's partials.
            const int wave = threadIdx.x >> 6;
            const int lane = threadIdx.x & 63;
            for (int s = wave; s < bm.numSystems; s += kFinWaves)
            {
                double acc[G];
#pragma unroll
                for (int k = 0; k < G; k++)
                {
                    acc[k] = 0;
                }
                const int i1 = bm.sysBlockStart[s + 1];
#pragma unroll 4
                for (int i = bm.sysBlockStart[s] + lane; i < i1; i += 64)
                {
#pragma unroll
                    for (int k = 0; k < G; k++)
                    {
                        acc[k] += partials[static_cast<size_t>(i) * K + k0 + k];
                    }
                }
#pragma unroll
                for (int k = 0; k < G; k++)
                {
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1)
                    {
                        acc[k] += __shfl_xor(acc[k], off, 64);
                    }
                }
                if (lane == 0)
                {
#pragma unroll
                    for (int k = 0; k < G; k++)
                    {
                        out[s * K + k0 + k] = acc[k];
                    }
                }
            }
            __syncthreads();
        }
    }

    template <int NCH>
    __device__ __forceinline__ bool systemActive(const SolveState & st, int s)
    {
        int a = 0;
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            a |= st.active[s * NCH + c];
        }
        return a != 0;
    }

    // Lane-group sum over j in [j0, j1) of val[j] * gather(col[j], c), in the
    // matrix type T.  rowDot gathers v[col*NCH + c] (v may be stored
    // narrower, e.g. an FP32 preconditioned vector under an FP64 matrix;
    // solver vectors interleave their channels, so one gathered column is one
    // contiguous access); the fused update-and-sweep kernel computes the
    // gathered values instead.
    template <typename T, int LANES, int NCH, typename G>
    __device__ __forceinline__ void rowDotG(const McgsMatrix<T> & A, int j0, int j1, G gather, T (&acc)[NCH])
    {
        const int lane = threadIdx.x & (LANES - 1);
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            acc[c] = T(0);
        }
        // U strided entries per step: their loads, then their gathers, go out
        // together (the terms are added in the same order as one by one).
        // Four for short rows on 1-2 lanes, two where a lane sees few entries.
        // The loads are unconditional: an entry past the row's end reloads
        // the step's first entry j (in the row by the loop test) and only its
        // term is dropped.  Conditional loads compile to branches that each
        // end in a wait for every load in flight, which serialises the step.
        constexpr int U = LANES <= 2 ? 4 : 2;
        for (int j = j0 + lane; j < j1; j += U * LANES)
        {
            T a[U];
            int col[U];
            bool has[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const int ju = j + u * LANES;
                has[u] = ju < j1;
                const int js = has[u] ? ju : j;
                a[u] = A.values[js];
                col[u] = A.column[js];
            }
            T xv[U][NCH];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    xv[u][c] = gather(col[u], c);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                if (has[u])
                {
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        acc[c] = fmaT(a[u], xv[u][c], acc[c]);
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            acc[c] = laneGroupSum<LANES>(acc[c]);
        }
    }

    template <typename T, int LANES, int NCH, typename V>
    __device__ __forceinline__ void rowDot(const McgsMatrix<T> & A, int j0, int j1, const V * __restrict__ v, T (&acc)[NCH])
    {
        rowDotG<T, LANES, NCH>(A, j0, j1, [v](int col, int c) { return static_cast<T>(v[static_cast<long>(col) * NCH + c]); },
                               acc);
    }

    template <int LANES>
    __device__ __forceinline__ bool groupLeader()
    {
        return (threadIdx.x & (LANES - 1)) == 0;
    }

    // Gauss-Seidel update of one colour.  lowerOnly = 0: full row (plain GS
    // sweep, backward half of symmetric GS); lowerOnly = 1: lower-colour part
    // only (forward half of symmetric GS started from zero).
    template <typename T, int LANES, int NCH>
    __device__ __forceinline__ void mcgsColorNarrow(McgsMatrix<T> A, BlockMap bm, int blockOffset, int rowBegin, int rowEnd, SolveState st,
                                                              const T * __restrict__ b, T * __restrict__ x, int lowerOnly)
    {
        // Latency: the block's rows, its system and the done flag are loaded
        // together, and the first slab's row bounds are in flight before the
        // early-exit test resolves (the test reads them, so they are issued
        // first); the chain to the first gather is three loads deep.
        const int4 range = sweepRange(bm, blockOffset, logicalBlock(blockIdx.x, gridDim.x), rowBegin, rowEnd);
        const int sys = range.z;
        const int done = st.flags[0];
        int row = range.x + static_cast<int>(threadIdx.x) / LANES;
        bool valid = row < range.y;
        int j0 = valid ? A.rowStart[row] : 0;
        int j1 = valid ? (lowerOnly ? A.rowSplit[row] : A.rowStart[row + 1]) : 0;
        const bool act = systemActive<NCH>(st, sys);
        if (done | !act | (j1 < j0))
        {
            return;
        }
        for (int base = range.x; base < range.y; base += kBlock / LANES)
        {
            if (base != range.x)
            {
                row = base + static_cast<int>(threadIdx.x) / LANES;
                valid = row < range.y;
                j0 = valid ? A.rowStart[row] : 0;
                j1 = valid ? (lowerOnly ? A.rowSplit[row] : A.rowStart[row + 1]) : 0;
            }
            // the row's own operands load alongside the gathers, on every
            // lane (lanes past the block's rows read the slab's first row):
            // loads under a condition end in a wait for every load in flight
            const bool lead = valid && groupLeader<LANES>();
            const int rs = valid ? row : base;
            const T inv = A.invDiag[rs];
            T bi[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                bi[c] = b[static_cast<long>(rs) * NCH + c];
            }
            T acc[NCH];
            rowDot<T, LANES, NCH>(A, j0, j1, x, acc);
            if (lead)
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    x[static_cast<long>(row) * NCH + c] = (bi[c] - acc[c]) * inv;
                }
            }
        }
    }

    // Residual check of the reference GS solver: mse = |b - A x| / sqrt(rows)
    // per system and channel; a system stops when all its channels are below
    // the target (gaussSeidel.cu:267-290, 354-380).  st.tol2 holds the target.
    template <typename T, int LANES, int NCH>
    __device__ __forceinline__ void gsResidualNarrow(McgsMatrix<T> A, BlockMap bm, SolveState st, const T * __restrict__ b,
                                                               const T * __restrict__ x)
    {
        __shared__ double sm[NCH * kWaves];
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        if (st.flags[0])
        {
            return; // the whole grid and its finalize skip
        }
        double part[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            part[c] = 0;
        }
        const int4 range = blockRange(bm, lb);
        if (systemActive<NCH>(st, range.z))
        {
            for (int base = range.x; base < range.y; base += kBlock / LANES)
            {
                const int row = base + static_cast<int>(threadIdx.x) / LANES;
                const bool valid = row < range.y;
                const int j0 = valid ? A.rowStart[row] : 0;
                const int j1 = valid ? A.rowStart[row + 1] : 0;
                T acc[NCH];
                rowDot<T, LANES, NCH>(A, j0, j1, x, acc);
                if (valid && groupLeader<LANES>())
                {
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        const long i = static_cast<long>(row) * NCH + c;
                        const T ax = fmaT(A.diag[row], x[i], acc[c]);
                        const double r = static_cast<double>(ax - b[i]);
                        part[c] = fma(r, r, part[c]);
                    }
                }
            }
        }
        blockSum<NCH>(part, sm);
        storePartial<NCH>(part, st.partials, range.w);
    }

    // PCG set-up: r = b - A x (or r = b with x = 0), z = r / d on colour 0,
    // per-channel |b|^2 and |r|^2 (then pcgInitFinalizeKernel).
    // Two precisions: T for the Krylov vectors and the matrix products, P for
    // the preconditioner (the symmetric GS sweeps).  P narrower than T: the
    // residual is also stored as rP = (P) r for the sweeps; P == T: rP == r.
    template <typename T, typename P, int LANES, int NCH>
    __device__ __forceinline__ void pcgInitNarrow(McgsMatrix<T> A, BlockMap bm, SolveState st, const T * __restrict__ b,
                                                            T * __restrict__ x, T * r, P * rP, P * __restrict__ z,
                                                            const P * __restrict__ invDiagP, int color0End, int zeroGuess)
    {
        __shared__ double sm[2 * NCH * kWaves];
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        double part[2 * NCH];
#pragma unroll
        for (int k = 0; k < 2 * NCH; k++)
        {
            part[k] = 0;
        }
        const int4 range = blockRange(bm, lb);
        for (int base = range.x; base < range.y; base += kBlock / LANES)
        {
            const int row = base + static_cast<int>(threadIdx.x) / LANES;
            const bool valid = row < range.y;
            // The row's own operands load before its entries, on every lane
            // and with no condition around a load (one would end in a wait
            // for every load in flight): lanes past the block's rows read the
            // slab's first row; x is read but not used for a zero guess, the
            // colour-0 scale at element 0 for the other colours.
            const int rs = valid ? row : base;
            const T dg = A.diag[rs];
            const P dInv = invDiagP[rs < color0End ? rs : 0];
            T bv[NCH], xv[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                const long i = static_cast<long>(rs) * NCH + c;
                bv[c] = b[i];
                xv[c] = x[i];
            }
            T acc[NCH];
            if (!zeroGuess)
            {
                const int j0 = A.rowStart[rs];
                const int j1 = A.rowStart[rs + 1];
                rowDot<T, LANES, NCH>(A, j0, valid ? j1 : j0, x, acc);
            }
            if (valid && groupLeader<LANES>())
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    const long i = static_cast<long>(row) * NCH + c;
                    const T bi = bv[c];
                    T ri;
                    if (zeroGuess)
                    {
                        x[i] = T(0);
                        ri = bi;
                    }
                    else
                    {
                        ri = bi - fmaT(dg, xv[c], acc[c]);
                    }
                    r[i] = ri;
                    const P rp = static_cast<P>(ri);
                    if constexpr (!std::is_same_v<T, P>)
                    {
                        rP[i] = rp;
                    }
                    if (row < color0End)
                    {
                        z[i] = rp * dInv;
                    }
                    part[c] = fma(static_cast<double>(bi), static_cast<double>(bi), part[c]);
                    part[NCH + c] = fma(static_cast<double>(ri), static_cast<double>(ri), part[NCH + c]);
                }
            }
        }
        blockSum<2 * NCH>(part, sm);
        storePartial<2 * NCH>(part, st.partials, range.w);
    }

    // Backward colour 0 of symmetric GS (on the preconditioner matrix M, type
    // P) fused with the partial sums of rho = r.z (r of the Krylov type T).
    // Colour-0 blocks carry the sweep work; they are interleaved evenly among
    // the others (block map order: colour 0 first, color0Blocks of them) so
    // that every XCD's range of logical blocks gets its share.
    template <typename P, typename T, int LANES, int NCH>
    __device__ __forceinline__ void pcgRhoNarrow(McgsMatrix<P> M, BlockMap bm, SolveState st, const P * rP,
                                                           T * r, P * __restrict__ z, int color0End, int color0Blocks,
                                                           const T * __restrict__ q, int r0Fix)
    {
        __shared__ double sm[NCH * kWaves];
        const int lb = interleaveFirst(logicalBlock(blockIdx.x, gridDim.x), gridDim.x, color0Blocks);
        // independent loads together; the first slab's row bounds in flight
        // before the early-exit test (which reads them)
        const int4 range = blockRange(bm, lb);
        const int sys = range.z;
        const int done = st.flags[0];
        // colour-0 rows (a flat block may straddle the end of colour 0)
        const bool anyC0 = range.x < color0End;
        int row = range.x + static_cast<int>(threadIdx.x) / LANES;
        bool valid = row < range.y;
        bool c0 = valid && row < color0End;
        int j0 = c0 ? M.rowStart[row] : 0;
        int j1 = c0 ? M.rowStart[row + 1] : 0;
        // the system's activity loads with the flag and the row bounds
        const bool act = systemActive<NCH>(st, sys);
        // r0Fix: r -= alpha q on colour 0 is still to be stored (the fused
        // update-and-sweep kernel leaves it), per active channel
        bool fixC[NCH];
        T alphaC[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            fixC[c] = r0Fix && st.active[sys * NCH + c];
            alphaC[c] = fixC[c] ? static_cast<T>(st.alpha[sys * NCH + c]) : T(0);
        }
        if (done || j1 < j0)
        {
            return;
        }
        double part[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            part[c] = 0;
        }
        if (act)
        {
            for (int base = range.x; base < range.y; base += kBlock / LANES)
            {
                if (base != range.x)
                {
                    row = base + static_cast<int>(threadIdx.x) / LANES;
                    valid = row < range.y;
                    c0 = valid && row < color0End;
                    j0 = c0 ? M.rowStart[row] : 0;
                    j1 = c0 ? M.rowStart[row + 1] : 0;
                }
                // The row's own operands load alongside the gathers, on every
                // lane and with no condition around a load (one would end in
                // a wait for every load in flight): lanes past the block's
                // rows read the slab's first row, a row's own value is read
                // through a selected pointer, and a value a row does not use
                // is read at element 0 (one shared line for all such lanes).
                const bool lead = valid && groupLeader<LANES>();
                const int rs = valid ? row : base;
                const P inv = M.invDiag[c0 ? rs : 0];
                // colour 0 starts from its residual (with P == T, rP is r),
                // the other colours keep the preconditioned value
                const P * ownSrc = c0 ? rP : z;
                P own[NCH];
                T rFix[NCH];
                double rr[NCH];
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    const long i = static_cast<long>(rs) * NCH + c;
                    const bool fix = c0 && fixC[c];
                    const T ri = r[i];
                    const T qi = q[fix ? i : 0];
                    const P oi = ownSrc[i];
                    rFix[c] = fix ? fmaT(-alphaC[c], qi, ri) : ri;
                    if constexpr (std::is_same_v<T, P>)
                    {
                        own[c] = c0 ? rFix[c] : oi;
                    }
                    else
                    {
                        own[c] = oi;
                    }
                    rr[c] = static_cast<double>(rFix[c]);
                }
                P acc[NCH];
                if (anyC0)
                {
                    // rows past colour 0 have j0 = j1: no entries
                    rowDot<P, LANES, NCH>(M, j0, j1, z, acc);
                }
                if (lead)
                {
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        const long i = static_cast<long>(row) * NCH + c;
                        P zi;
                        if (c0 && fixC[c])
                        {
                            r[i] = rFix[c];
                        }
                        if (c0)
                        {
                            zi = (own[c] - acc[c]) * inv;
                            z[i] = zi;
                        }
                        else
                        {
                            zi = own[c];
                        }
                        part[c] = fma(rr[c], static_cast<double>(zi), part[c]);
                    }
                }
            }
        }
        blockSum<NCH>(part, sm);
        storePartial<NCH>(part, st.partials, range.w);
    }

    // p = z + beta p;  q = A z + beta q  (= A p);  partial sums of p.q.
    // z is the preconditioned vector (type P), p and q are of the Krylov type T.
    template <typename T, typename P, int LANES, int NCH>
    __device__ __forceinline__ void pcgDirectionNarrow(McgsMatrix<T> A, BlockMap bm, SolveState st, const P * __restrict__ z,
                                                       T * __restrict__ p, T * __restrict__ q, int fused, int rhoParts, int updParts)
    {
        __shared__ double sm[2 * NCH * kWaves];
        __shared__ double shBeta[NCH];
        __shared__ int shAct[NCH];
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        // independent loads together; the first slab's row bounds in flight
        // before the early-exit test (which reads them)
        const int4 range = blockRange(bm, lb);
        const int s = range.z;
        const int done = st.flags[0];
        int row = range.x + static_cast<int>(threadIdx.x) / LANES;
        bool valid = row < range.y;
        int j0 = valid ? A.rowStart[row] : 0;
        int j1 = valid ? A.rowStart[row + 1] : 0;
        if (done || j1 < j0)
        {
            return;
        }
        double part[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            part[c] = 0;
        }
        // fused: the partial sums load now, with the first slab's rows, and
        // are finished after its gathers (beta is needed only to combine)
        FusedRhoLoads<NCH> fl;
        FusedSlice sl;
        bool pending = fused != 0;
        if (pending)
        {
            sl = fusedSlice(bm, range, lb, rhoParts, updParts);
            fusedRhoLoad<NCH>(st, sl, fl);
        }
        if (pending || systemActive<NCH>(st, s))
        {
            const bool first = st.flags[1] != 0;
            T beta[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                beta[c] = pending ? T(0) : static_cast<T>(st.beta[s * NCH + c]);
            }
            for (int base = range.x; base < range.y; base += kBlock / LANES)
            {
                if (base != range.x)
                {
                    row = base + static_cast<int>(threadIdx.x) / LANES;
                    valid = row < range.y;
                    j0 = valid ? A.rowStart[row] : 0;
                    j1 = valid ? A.rowStart[row + 1] : 0;
                }
                // The row's own operands load alongside the gathers, on every
                // lane and unconverted (lanes past the block's rows read the
                // slab's first row; p and q are read but not used on the
                // first iteration): loads under a condition, or converted
                // where they load, make the compiler wait for every load in
                // flight before the row's entries are even requested.
                const bool lead = valid && groupLeader<LANES>();
                const int rs = valid ? row : base;
                const T dg = A.diag[rs];
                P zr[NCH];
                T pr[NCH], qr[NCH];
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    const long i = static_cast<long>(rs) * NCH + c;
                    zr[c] = z[i];
                    pr[c] = p[i];
                    qr[c] = q[i];
                }
                T acc[NCH];
                rowDot<T, LANES, NCH>(A, j0, j1, z, acc);
                if (pending)
                {
                    // uniform: every thread is in the first slab here
                    pending = false;
                    if (!fusedRhoFinish<NCH>(st, fl, sl, sm, shBeta, shAct))
                    {
                        break;
                    }
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        beta[c] = static_cast<T>(shBeta[c]);
                    }
                }
                if (lead)
                {
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        const long i = static_cast<long>(row) * NCH + c;
                        const T zi = static_cast<T>(zr[c]);
                        const T az = fmaT(dg, zi, acc[c]);
                        T pi, qi;
                        if (first)
                        {
                            pi = zi;
                            qi = az;
                        }
                        else
                        {
                            pi = fmaT(beta[c], pr[c], zi);
                            qi = fmaT(beta[c], qr[c], az);
                        }
                        p[i] = pi;
                        q[i] = qi;
                        part[c] = fma(static_cast<double>(pi), static_cast<double>(qi), part[c]);
                    }
                }
            }
        }
        blockSum<NCH>(part, sm);
        storePartial<NCH>(part, fused ? st.partials3 : st.partials, range.w);
    }

    // ------------------------------------------------------------------
    // Wide channel tiles: 4 to 32 right-hand sides sharing one matrix.  NCH
    // lanes per row, lane c on channel c walks the whole row in order, so a
    // gathered column is NCH contiguous values across the lanes, the row's
    // entries are one broadcast load per lane group, and no lane reduction
    // is needed.  1 and 2 channels keep the lane-split rows above.
    // ------------------------------------------------------------------
    template <int NCH>
    constexpr bool kWide = NCH > 2;

    // sum over j in [j0, j1) of val[j] * v[col[j]*NCH + c], in entry order;
    // four entries' loads, then their gathers, go out together.
    template <typename T, int NCH, typename V>
    __device__ __forceinline__ T rowDotWide(const McgsMatrix<T> & A, int j0, int j1, const V * __restrict__ v, int c)
    {
        T acc = T(0);
        int j = j0;
        for (; j + 4 <= j1; j += 4)
        {
            T a[4], xv[4];
            int col[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
            {
                a[k] = A.values[j + k];
                col[k] = A.column[j + k];
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
            {
                xv[k] = static_cast<T>(v[static_cast<long>(col[k]) * NCH + c]);
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
            {
                acc = fmaT(a[k], xv[k], acc);
            }
        }
        for (; j < j1; j++)
        {
            acc = fmaT(A.values[j], static_cast<T>(v[static_cast<long>(A.column[j]) * NCH + c]), acc);
        }
        return acc;
    }

    // Block sum per channel when every thread holds one channel's value,
    // channel = threadIdx.x % NCH (lanes of one channel are NCH apart in a
    // wavefront).  The total of channel c is returned in thread c (c < NCH).
    template <int NCH>
    __device__ __forceinline__ double channelBlockSum(double v, double * sm /* NCH * kWaves */)
    {
#pragma unroll
        for (int off = 32; off >= NCH; off >>= 1)
        {
            v += __shfl_xor(v, off, 64);
        }
        const int wave = threadIdx.x >> 6;
        const int lane = threadIdx.x & 63;
        if (lane < NCH)
        {
            sm[lane * kWaves + wave] = v;
        }
        __syncthreads();
        double s = 0;
        if (static_cast<int>(threadIdx.x) < NCH)
        {
#pragma unroll
            for (int w = 0; w < kWaves; w++)
            {
                s += sm[threadIdx.x * kWaves + w];
            }
        }
        __syncthreads();
        return s;
    }

    // channelBlockSum of two values at once (one pair of barriers); sm holds
    // 2 * NCH * kWaves doubles.  Totals in threads c < NCH.
    template <int NCH>
    __device__ __forceinline__ void channelBlockSum2(double & a, double & b, double * sm)
    {
#pragma unroll
        for (int off = 32; off >= NCH; off >>= 1)
        {
            a += __shfl_xor(a, off, 64);
            b += __shfl_xor(b, off, 64);
        }
        const int wave = threadIdx.x >> 6;
        const int lane = threadIdx.x & 63;
        if (lane < NCH)
        {
            sm[lane * kWaves + wave] = a;
            sm[(NCH + lane) * kWaves + wave] = b;
        }
        __syncthreads();
        double sa = 0, sb = 0;
        if (static_cast<int>(threadIdx.x) < NCH)
        {
#pragma unroll
            for (int w = 0; w < kWaves; w++)
            {
                sa += sm[threadIdx.x * kWaves + w];
                sb += sm[(NCH + threadIdx.x) * kWaves + w];
            }
        }
        __syncthreads();
        a = sa;
        b = sb;
    }

    // partials[slot*K + offset + c] = total of channel c (threads c < NCH).
    template <int NCH, int K>
    __device__ __forceinline__ void storeChannelPartial(double total, double * partials, int slot, int offset)
    {
        if (static_cast<int>(threadIdx.x) < NCH)
        {
            partials[static_cast<size_t>(slot) * K + offset + threadIdx.x] = total;
        }
    }

    // Sum over count partial slots of channel threadIdx.x % NCH (slot i,
    // channel c at parts[i*NCH + c]); the total of channel c is returned in
    // thread c (c < NCH).  Same order in every block.
    template <int NCH>
    __device__ __forceinline__ double channelPartialSum(const double * __restrict__ parts, int count, double * sm)
    {
        constexpr int G = kBlock / NCH; // threads per channel
        const int c = static_cast<int>(threadIdx.x) % NCH;
        double acc = 0;
#pragma unroll 4
        for (int i = static_cast<int>(threadIdx.x) / NCH; i < count; i += G)
        {
            acc += parts[static_cast<size_t>(i) * NCH + c];
        }
        return channelBlockSum<NCH>(acc, sm);
    }

    // Fused reductions, first half: the flags, the system's per-channel state
    // and up to kFusedPre partial sums per thread of each set load into
    // registers, so that they are in flight with the kernel's first row loads.
    template <int NCH>
    __device__ __forceinline__ void fusedRhoLoad(const SolveState & st, const FusedSlice & sl, FusedRhoLoads<NCH> & f)
    {
        const int t = static_cast<int>(threadIdx.x);
        f.first = st.flags[1];
        f.was = 0;
        f.ref = 0;
        f.rhoOld = 0;
        if (t < NCH)
        {
            const int k = sl.s * NCH + t;
            f.was = st.active[k];
            f.ref = st.ref[k];
            f.rhoOld = st.rho[k];
        }
        constexpr int G = kBlock / NCH; // threads per channel
        const int c = t % NCH;
        const int i0 = t / NCH;
        const double * pr = st.partials + static_cast<size_t>(sl.base) * NCH;
        const double * pu = st.partials2 + static_cast<size_t>(sl.base) * NCH;
#pragma unroll
        for (int k = 0; k < kFusedPre; k++)
        {
            const int i = i0 + k * G;
            f.rho[k] = i < sl.rhoParts ? pr[static_cast<size_t>(i) * NCH + c] : 0.0;
            f.rr[k] = i < sl.updParts ? pu[static_cast<size_t>(i) * NCH + c] : 0.0;
        }
    }

    // Second half: finish the previous iteration's rho = r.z (rho kernel
    // partials) and |r|^2 (update kernel partials, unused in the first
    // iteration) of the block's system.  Every block of the system sums them
    // in the same order, so all hold the same convergence decision and the
    // same beta, as pcgRhoFinalizeKernel would compute them.  The system's
    // first block records rho and the decision for the update kernel
    // (rhoNext, activeNext) and rr, the iteration count and non-finite
    // results for the host; no block of this launch reads those.  Returns
    // whether any channel of the system goes on; beta and activity per
    // channel in LDS.
    template <int NCH>
    __device__ __forceinline__ bool fusedRhoFinish(const SolveState & st, const FusedRhoLoads<NCH> & f, const FusedSlice & sl,
                                                   double * sm, double * shBeta, int * shAct)
    {
        const int t = static_cast<int>(threadIdx.x);
        constexpr int G = kBlock / NCH;
        const int c = t % NCH;
        double rho = 0, rr = 0;
#pragma unroll
        for (int k = 0; k < kFusedPre; k++)
        {
            rho += f.rho[k];
            rr += f.rr[k];
        }
        // beyond the preloaded slots (more than kFusedPre * G partials)
        const double * pr = st.partials + static_cast<size_t>(sl.base) * NCH;
        const double * pu = st.partials2 + static_cast<size_t>(sl.base) * NCH;
        for (int i = t / NCH + kFusedPre * G; i < sl.rhoParts; i += G)
        {
            rho += pr[static_cast<size_t>(i) * NCH + c];
        }
        for (int i = t / NCH + kFusedPre * G; i < sl.updParts; i += G)
        {
            rr += pu[static_cast<size_t>(i) * NCH + c];
        }
        channelBlockSum2<NCH>(rho, rr, sm);
        const int first = f.first;
        if (t < NCH)
        {
            const int was = f.was;
            int act = was;
            int bad = 0;
            if (was && !first)
            {
                if (!(rr == rr) || rr > 1e300)
                {
                    act = 0;
                    bad = 1;
                }
                else if (rr <= st.tol2 * f.ref)
                {
                    act = 0;
                }
            }
            shBeta[t] = (first || f.rhoOld == 0) ? 0.0 : rho / f.rhoOld;
            shAct[t] = act;
            if (act && !(rho == rho))
            {
                bad = 1;
            }
            if (sl.leader)
            {
                const int k = sl.s * NCH + t;
                st.rhoNext[k] = rho;
                st.activeNext[k] = act;
                if (was && !first)
                {
                    st.rr[k] = rr;
                    st.iters[k] = st.iters[k] + 1;
                }
                if (bad)
                {
                    st.flags[2] = 1;
                }
            }
        }
        __syncthreads();
        int any = 0;
#pragma unroll
        for (int k = 0; k < NCH; k++)
        {
            any |= shAct[k];
        }
        return any != 0;
    }

    template <int NCH>
    __device__ __forceinline__ bool fusedRhoStep(const SolveState & st, const FusedSlice & sl, double * sm, double * shBeta,
                                                 int * shAct)
    {
        FusedRhoLoads<NCH> f;
        fusedRhoLoad<NCH>(st, sl, f);
        return fusedRhoFinish<NCH>(st, f, sl, sm, shBeta, shAct);
    }

    template <typename T, int NCH>
    __device__ __forceinline__ void mcgsColorWide(McgsMatrix<T> A, BlockMap bm, int blockOffset, int rowBegin, int rowEnd, SolveState st, const T * __restrict__ b,
                                                  T * __restrict__ x, int lowerOnly)
    {
        constexpr int slab = kBlock / NCH;
        const int4 range = sweepRange(bm, blockOffset, logicalBlock(blockIdx.x, gridDim.x), rowBegin, rowEnd);
        const int sys = range.z;
        const int done = st.flags[0];
        const int c = static_cast<int>(threadIdx.x) % NCH;
        int row = range.x + static_cast<int>(threadIdx.x) / NCH;
        bool valid = row < range.y;
        int j0 = valid ? A.rowStart[row] : 0;
        int j1 = valid ? (lowerOnly ? A.rowSplit[row] : A.rowStart[row + 1]) : 0;
        const bool act = systemActive<NCH>(st, sys);
        if (done | !act | (j1 < j0))
        {
            return;
        }
        for (int base = range.x; base < range.y; base += slab)
        {
            if (base != range.x)
            {
                row = base + static_cast<int>(threadIdx.x) / NCH;
                valid = row < range.y;
                j0 = valid ? A.rowStart[row] : 0;
                j1 = valid ? (lowerOnly ? A.rowSplit[row] : A.rowStart[row + 1]) : 0;
            }
            if (valid)
            {
                const long i = static_cast<long>(row) * NCH + c;
                const T inv = A.invDiag[row];
                const T bi = b[i];
                const T acc = rowDotWide<T, NCH>(A, j0, j1, x, c);
                x[i] = (bi - acc) * inv;
            }
        }
    }

    template <typename T, int NCH>
    __device__ __forceinline__ void gsResidualWide(McgsMatrix<T> A, BlockMap bm, SolveState st, const T * __restrict__ b,
                                                   const T * __restrict__ x)
    {
        __shared__ double sm[NCH * kWaves];
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        if (st.flags[0])
        {
            return; // the whole grid and its finalize skip
        }
        const int c = static_cast<int>(threadIdx.x) % NCH;
        double part = 0;
        const int4 range = blockRange(bm, lb);
        if (systemActive<NCH>(st, range.z))
        {
            for (int base = range.x; base < range.y; base += kBlock / NCH)
            {
                const int row = base + static_cast<int>(threadIdx.x) / NCH;
                if (row < range.y)
                {
                    const long i = static_cast<long>(row) * NCH + c;
                    const T acc = rowDotWide<T, NCH>(A, A.rowStart[row], A.rowStart[row + 1], x, c);
                    const T ax = fmaT(A.diag[row], x[i], acc);
                    const double r = static_cast<double>(ax - b[i]);
                    part = fma(r, r, part);
                }
            }
        }
        const double total = channelBlockSum<NCH>(part, sm);
        storeChannelPartial<NCH, NCH>(total, st.partials, range.w, 0);
    }

    template <typename T, typename P, int NCH>
    __device__ __forceinline__ void pcgInitWide(McgsMatrix<T> A, BlockMap bm, SolveState st, const T * __restrict__ b, T * __restrict__ x,
                                                T * r, P * rP, P * __restrict__ z, const P * __restrict__ invDiagP, int color0End,
                                                int zeroGuess)
    {
        __shared__ double sm[NCH * kWaves];
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        const int c = static_cast<int>(threadIdx.x) % NCH;
        double bb = 0, rr = 0;
        const int4 range = blockRange(bm, lb);
        for (int base = range.x; base < range.y; base += kBlock / NCH)
        {
            const int row = base + static_cast<int>(threadIdx.x) / NCH;
            if (row < range.y)
            {
                const long i = static_cast<long>(row) * NCH + c;
                const T bi = b[i];
                T ri;
                if (zeroGuess)
                {
                    x[i] = T(0);
                    ri = bi;
                }
                else
                {
                    const T acc = rowDotWide<T, NCH>(A, A.rowStart[row], A.rowStart[row + 1], x, c);
                    ri = bi - fmaT(A.diag[row], x[i], acc);
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
        const double tb = channelBlockSum<NCH>(bb, sm);
        const double tr = channelBlockSum<NCH>(rr, sm);
        storeChannelPartial<NCH, 2 * NCH>(tb, st.partials, range.w, 0);
        storeChannelPartial<NCH, 2 * NCH>(tr, st.partials, range.w, NCH);
    }

    template <typename P, typename T, int NCH>
    __device__ __forceinline__ void pcgRhoWide(McgsMatrix<P> M, BlockMap bm, SolveState st, const P * rP, const T * r,
                                               P * __restrict__ z, int color0End, int color0Blocks)
    {
        // (wide tiles never take the fused update and sweep: no r0Fix here)
        __shared__ double sm[NCH * kWaves];
        const int lb = interleaveFirst(logicalBlock(blockIdx.x, gridDim.x), gridDim.x, color0Blocks);
        const int4 range = blockRange(bm, lb);
        const int sys = range.z;
        const int done = st.flags[0];
        const int c = static_cast<int>(threadIdx.x) % NCH;
        int row = range.x + static_cast<int>(threadIdx.x) / NCH;
        bool valid = row < range.y;
        // colour-0 rows (a flat block may straddle the end of colour 0)
        bool c0 = valid && row < color0End;
        int j0 = c0 ? M.rowStart[row] : 0;
        int j1 = c0 ? M.rowStart[row + 1] : 0;
        const bool act = systemActive<NCH>(st, sys);
        if (done || j1 < j0)
        {
            return;
        }
        double part = 0;
        if (act)
        {
            for (int base = range.x; base < range.y; base += kBlock / NCH)
            {
                if (base != range.x)
                {
                    row = base + static_cast<int>(threadIdx.x) / NCH;
                    valid = row < range.y;
                    c0 = valid && row < color0End;
                    j0 = c0 ? M.rowStart[row] : 0;
                    j1 = c0 ? M.rowStart[row + 1] : 0;
                }
                if (valid)
                {
                    const long i = static_cast<long>(row) * NCH + c;
                    const double ri = static_cast<double>(r[i]);
                    P zi;
                    if (c0)
                    {
                        const P inv = M.invDiag[row];
                        const P own = rP[i];
                        const P acc = rowDotWide<P, NCH>(M, j0, j1, z, c);
                        zi = (own - acc) * inv;
                        z[i] = zi;
                    }
                    else
                    {
                        zi = z[i];
                    }
                    part = fma(ri, static_cast<double>(zi), part);
                }
            }
        }
        const double total = channelBlockSum<NCH>(part, sm);
        storeChannelPartial<NCH, NCH>(total, st.partials, range.w, 0);
    }

    template <typename T, typename P, int NCH>
    __device__ __forceinline__ void pcgDirectionWide(McgsMatrix<T> A, BlockMap bm, SolveState st, const P * __restrict__ z,
                                                     T * __restrict__ p, T * __restrict__ q, int fused, int rhoParts, int updParts)
    {
        __shared__ double sm[2 * NCH * kWaves];
        __shared__ double shBeta[NCH];
        __shared__ int shAct[NCH];
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        const int4 range = blockRange(bm, lb);
        const int s = range.z;
        const int done = st.flags[0];
        const int c = static_cast<int>(threadIdx.x) % NCH;
        int row = range.x + static_cast<int>(threadIdx.x) / NCH;
        bool valid = row < range.y;
        int j0 = valid ? A.rowStart[row] : 0;
        int j1 = valid ? A.rowStart[row + 1] : 0;
        if (done || j1 < j0)
        {
            return;
        }
        double part = 0;
        const bool go = fused ? fusedRhoStep<NCH>(st, fusedSlice(bm, range, lb, rhoParts, updParts), sm, shBeta, shAct)
                              : systemActive<NCH>(st, s);
        if (go)
        {
            const bool first = st.flags[1] != 0;
            const T beta = static_cast<T>(fused ? shBeta[c] : st.beta[s * NCH + c]);
            for (int base = range.x; base < range.y; base += kBlock / NCH)
            {
                if (base != range.x)
                {
                    row = base + static_cast<int>(threadIdx.x) / NCH;
                    valid = row < range.y;
                    j0 = valid ? A.rowStart[row] : 0;
                    j1 = valid ? A.rowStart[row + 1] : 0;
                }
                if (valid)
                {
                    const long i = static_cast<long>(row) * NCH + c;
                    const T dg = A.diag[row];
                    const T zi = static_cast<T>(z[i]);
                    T pr = T(0), qr = T(0);
                    if (!first)
                    {
                        pr = p[i];
                        qr = q[i];
                    }
                    const T acc = rowDotWide<T, NCH>(A, j0, j1, z, c);
                    const T az = fmaT(dg, zi, acc);
                    const T pi = first ? zi : fmaT(beta, pr, zi);
                    const T qi = first ? az : fmaT(beta, qr, az);
                    p[i] = pi;
                    q[i] = qi;
                    part = fma(static_cast<double>(pi), static_cast<double>(qi), part);
                }
            }
        }
        const double total = channelBlockSum<NCH>(part, sm);
        storeChannelPartial<NCH, NCH>(total, fused ? st.partials3 : st.partials, range.w, 0);
    }
    template <typename T, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) mcgsColorKernel(McgsMatrix<T> A, BlockMap bm, int blockOffset, int rowBegin, int rowEnd, SolveState st, const T * __restrict__ b, T * __restrict__ x, int lowerOnly)
    {
        if constexpr (kWide<NCH>)
        {
            mcgsColorWide<T, NCH>(A, bm, blockOffset, rowBegin, rowEnd, st, b, x, lowerOnly);
        }
        else
        {
            mcgsColorNarrow<T, LANES, NCH>(A, bm, blockOffset, rowBegin, rowEnd, st, b, x, lowerOnly);
        }
    }

    template <typename T, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) gsResidualKernel(McgsMatrix<T> A, BlockMap bm, SolveState st, const T * __restrict__ b, const T * __restrict__ x)
    {
        if constexpr (kWide<NCH>)
        {
            gsResidualWide<T, NCH>(A, bm, st, b, x);
        }
        else
        {
            gsResidualNarrow<T, LANES, NCH>(A, bm, st, b, x);
        }
    }

    template <typename T, typename P, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) pcgInitKernel(McgsMatrix<T> A, BlockMap bm, SolveState st, const T * __restrict__ b, T * __restrict__ x, T * r, P * rP, P * __restrict__ z, const P * __restrict__ invDiagP, int color0End, int zeroGuess)
    {
        if constexpr (kWide<NCH>)
        {
            pcgInitWide<T, P, NCH>(A, bm, st, b, x, r, rP, z, invDiagP, color0End, zeroGuess);
        }
        else
        {
            pcgInitNarrow<T, P, LANES, NCH>(A, bm, st, b, x, r, rP, z, invDiagP, color0End, zeroGuess);
        }
    }

    // r0Fix = 1 (1-2 channels, after pcgUpdateSweepKernel): the colour-0
    // residuals r -= alpha q are stored here first (alpha and the channels'
    // activity from the solve state).
    template <typename P, typename T, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) pcgRhoKernel(McgsMatrix<P> M, BlockMap bm, SolveState st, const P * rP, T * r, P * __restrict__ z, int color0End, int color0Blocks,
                                                           const T * __restrict__ q, int r0Fix)
    {
        if constexpr (kWide<NCH>)
        {
            pcgRhoWide<P, T, NCH>(M, bm, st, rP, r, z, color0End, color0Blocks);
        }
        else
        {
            pcgRhoNarrow<P, T, LANES, NCH>(M, bm, st, rP, r, z, color0End, color0Blocks, q, r0Fix);
        }
    }

    // Factorized sparse approximate inverse preconditioner z = G^T G r (the
    // factor built on the host: G's values share the matrix's column
    // structure, G(i, j) for the lower-colour entries of row i, G(j, i) for
    // its higher-colour entries, diag the diagonal of G; rowSplit splits the
    // two).  Two launches replace the multicolour sweeps: t = G r over the
    // lower parts, then z = G^T t over the higher parts with the partial
    // sums of rho = r.z, stored as the rho kernel stores them.  Same block
    // map as the direction kernel.  1-2 channels split a row over lanes;
    // wide tiles run one lane per channel (below).
    template <typename P, int LANES, int NCH>
    __device__ __forceinline__ void fsaiLowerNarrow(McgsMatrix<P> G, BlockMap bm, SolveState st, const P * __restrict__ r,
                                                    P * __restrict__ t)
    {
        const int4 range = blockRange(bm, logicalBlock(blockIdx.x, gridDim.x));
        const int done = st.flags[0];
        int row = range.x + static_cast<int>(threadIdx.x) / LANES;
        bool valid = row < range.y;
        int j0 = valid ? G.rowStart[row] : 0;
        int j1 = valid ? G.rowSplit[row] : 0;
        const bool act = systemActive<NCH>(st, range.z);
        if (done | !act | (j1 < j0))
        {
            return;
        }
        for (int base = range.x; base < range.y; base += kBlock / LANES)
        {
            if (base != range.x)
            {
                row = base + static_cast<int>(threadIdx.x) / LANES;
                valid = row < range.y;
                j0 = valid ? G.rowStart[row] : 0;
                j1 = valid ? G.rowSplit[row] : 0;
            }
            // the row's own operands on every lane, unconditionally (lanes
            // past the block's rows read the slab's first row)
            const int rs = valid ? row : base;
            const P gd = G.diag[rs];
            P rv[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                rv[c] = r[static_cast<long>(rs) * NCH + c];
            }
            P acc[NCH];
            rowDot<P, LANES, NCH>(G, j0, j1, r, acc);
            if (valid && groupLeader<LANES>())
            {
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    t[static_cast<long>(row) * NCH + c] = fmaT(gd, rv[c], acc[c]);
                }
            }
        }
    }

    template <typename P, typename T, int LANES, int NCH>
    __device__ __forceinline__ void fsaiUpperNarrow(McgsMatrix<P> G, BlockMap bm, SolveState st, const P * __restrict__ t,
                                                    P * __restrict__ z, const T * __restrict__ r)
    {
        __shared__ double sm[NCH * kWaves];
        const int4 range = blockRange(bm, logicalBlock(blockIdx.x, gridDim.x));
        const int done = st.flags[0];
        int row = range.x + static_cast<int>(threadIdx.x) / LANES;
        bool valid = row < range.y;
        int j0 = valid ? G.rowSplit[row] : 0;
        int j1 = valid ? G.rowStart[row + 1] : 0;
        const bool act = systemActive<NCH>(st, range.z);
        if (done || j1 < j0)
        {
            return;
        }
        double part[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            part[c] = 0;
        }
        if (act)
        {
            for (int base = range.x; base < range.y; base += kBlock / LANES)
            {
                if (base != range.x)
                {
                    row = base + static_cast<int>(threadIdx.x) / LANES;
                    valid = row < range.y;
                    j0 = valid ? G.rowSplit[row] : 0;
                    j1 = valid ? G.rowStart[row + 1] : 0;
                }
                const int rs = valid ? row : base;
                const P gd = G.diag[rs];
                P tv[NCH];
                T rv[NCH];
#pragma unroll
                for (int c = 0; c < NCH; c++)
                {
                    const long i = static_cast<long>(rs) * NCH + c;
                    tv[c] = t[i];
                    rv[c] = r[i];
                }
                P acc[NCH];
                rowDot<P, LANES, NCH>(G, j0, j1, t, acc);
                if (valid && groupLeader<LANES>())
                {
#pragma unroll
                    for (int c = 0; c < NCH; c++)
                    {
                        const P zi = fmaT(gd, tv[c], acc[c]);
                        z[static_cast<long>(row) * NCH + c] = zi;
                        part[c] = fma(static_cast<double>(rv[c]), static_cast<double>(zi), part[c]);
                    }
                }
            }
        }
        blockSum<NCH>(part, sm);
        storePartial<NCH>(part, st.partials, range.w);
    }

    // Wide tiles (4-32 channels): one lane per channel walks the whole
    // (lower or higher) part of its row, as the other wide kernels.
    template <typename P, int NCH>
    __device__ __forceinline__ void fsaiLowerWide(McgsMatrix<P> G, BlockMap bm, SolveState st, const P * __restrict__ r,
                                                  P * __restrict__ t)
    {
        const int4 range = blockRange(bm, logicalBlock(blockIdx.x, gridDim.x));
        const int done = st.flags[0];
        const int c = static_cast<int>(threadIdx.x) % NCH;
        int row = range.x + static_cast<int>(threadIdx.x) / NCH;
        bool valid = row < range.y;
        int j0 = valid ? G.rowStart[row] : 0;
        int j1 = valid ? G.rowSplit[row] : 0;
        const bool act = systemActive<NCH>(st, range.z);
        if (done | !act | (j1 < j0))
        {
            return;
        }
        for (int base = range.x; base < range.y; base += kBlock / NCH)
        {
            if (base != range.x)
            {
                row = base + static_cast<int>(threadIdx.x) / NCH;
                valid = row < range.y;
                j0 = valid ? G.rowStart[row] : 0;
                j1 = valid ? G.rowSplit[row] : 0;
            }
            const int rs = valid ? row : base;
            const P gd = G.diag[rs];
            const P rv = r[static_cast<long>(rs) * NCH + c];
            const P acc = rowDotWide<P, NCH>(G, j0, j1, r, c);
            if (valid)
            {
                t[static_cast<long>(row) * NCH + c] = fmaT(gd, rv, acc);
            }
        }
    }

    template <typename P, typename T, int NCH>
    __device__ __forceinline__ void fsaiUpperWide(McgsMatrix<P> G, BlockMap bm, SolveState st, const P * __restrict__ t,
                                                  P * __restrict__ z, const T * __restrict__ r)
    {
        __shared__ double sm[NCH * kWaves];
        const int4 range = blockRange(bm, logicalBlock(blockIdx.x, gridDim.x));
        const int done = st.flags[0];
        const int c = static_cast<int>(threadIdx.x) % NCH;
        int row = range.x + static_cast<int>(threadIdx.x) / NCH;
        bool valid = row < range.y;
        int j0 = valid ? G.rowSplit[row] : 0;
        int j1 = valid ? G.rowStart[row + 1] : 0;
        const bool act = systemActive<NCH>(st, range.z);
        if (done || j1 < j0)
        {
            return;
        }
        double part = 0;
        if (act)
        {
            for (int base = range.x; base < range.y; base += kBlock / NCH)
            {
                if (base != range.x)
                {
                    row = base + static_cast<int>(threadIdx.x) / NCH;
                    valid = row < range.y;
                    j0 = valid ? G.rowSplit[row] : 0;
                    j1 = valid ? G.rowStart[row + 1] : 0;
                }
                const int rs = valid ? row : base;
                const long is = static_cast<long>(rs) * NCH + c;
                const P gd = G.diag[rs];
                const P tv = t[is];
                const T rv = r[is];
                const P acc = rowDotWide<P, NCH>(G, j0, j1, t, c);
                if (valid)
                {
                    const P zi = fmaT(gd, tv, acc);
                    z[static_cast<long>(row) * NCH + c] = zi;
                    part = fma(static_cast<double>(rv), static_cast<double>(zi), part);
                }
            }
        }
        const double total = channelBlockSum<NCH>(part, sm);
        storeChannelPartial<NCH, NCH>(total, st.partials, range.w, 0);
    }

    template <typename P, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) fsaiLowerKernel(McgsMatrix<P> G, BlockMap bm, SolveState st, const P * __restrict__ r,
                                                              P * __restrict__ t)
    {
        if constexpr (kWide<NCH>)
        {
            fsaiLowerWide<P, NCH>(G, bm, st, r, t);
        }
        else
        {
            fsaiLowerNarrow<P, LANES, NCH>(G, bm, st, r, t);
        }
    }

    template <typename P, typename T, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) fsaiUpperKernel(McgsMatrix<P> G, BlockMap bm, SolveState st, const P * __restrict__ t,
                                                              P * __restrict__ z, const T * __restrict__ r)
    {
        if constexpr (kWide<NCH>)
        {
            fsaiUpperWide<P, T, NCH>(G, bm, st, t, z, r);
        }
        else
        {
            fsaiUpperNarrow<P, T, LANES, NCH>(G, bm, st, t, z, r);
        }
    }

    template <typename T, typename P, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) pcgDirectionKernel(McgsMatrix<T> A, BlockMap bm, SolveState st, const P * __restrict__ z, T * __restrict__ p, T * __restrict__ q,
                                                                 int fused, int rhoParts, int updParts)
    {
        if constexpr (kWide<NCH>)
        {
            pcgDirectionWide<T, P, NCH>(A, bm, st, z, p, q, fused, rhoParts, updParts);
        }
        else
        {
            pcgDirectionNarrow<T, P, LANES, NCH>(A, bm, st, z, p, q, fused, rhoParts, updParts);
        }
    }

    // x += alpha p;  r -= alpha q;  z = r / d on colour 0 (preconditioner
    // type P; rP = (P) r when P is narrower than T); partial sums of |r|^2.
    // fused = 0: alpha from pcgDirectionFinalizeKernel, activity from the
    // solve state.  fused = 1 (fused reductions): every block finishes p.q of
    // its system from the direction kernel's partials (one system: dirParts of
    // them) itself (same order in every block of the system, so all hold the
    // same alpha = rho / p.q) while its first elements' operands are already
    // loading, with rho and activity as the direction kernel decided them;
    // block 0 makes those decisions the solve state for the kernels that follow.
    template <typename T, typename P, int NCH>
    __global__ void __launch_bounds__(kBlock) pcgUpdateKernel(int n, BlockMap bm, SolveState st, const T * __restrict__ p,
                                                              const T * __restrict__ q, T * __restrict__ x, T * r, P * rP,
                                                              P * __restrict__ z, const P * __restrict__ invDiagP, int color0End,
                                                              int fused, int dirParts)
    {
        __shared__ double sm[NCH * kWaves];
        __shared__ double shAlpha[NCH];
        __shared__ int shAct[NCH];
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        // the block's rows, its system and the done flag load together
        const int4 range = blockRange(bm, lb);
        const int s = range.z;
        // flat over the block's interleaved elements, one per thread: the
        // stride (kBlock) is a multiple of NCH, so a thread keeps one channel
        static_assert(kBlock % NCH == 0, "channels must divide the block");
        const int c = static_cast<int>(threadIdx.x) % NCH;
        const long i0 = static_cast<long>(range.x) * NCH + threadIdx.x;
        const long end = static_cast<long>(range.y) * NCH;
        const int row0 = static_cast<int>(i0 / NCH);
        bool go;
        int act;
        T alpha;
        // the first element's operands go out before alpha is known
        T r0 = T(0), x0 = T(0), p0 = T(0), q0 = T(0);
        P d0 = P(0);
        if (fused)
        {
            // every load goes out first: the system's next activity and rho,
            // the first element's operands and the direction kernel's partials
            // of the system (the stop test comes after)
            const int t = static_cast<int>(threadIdx.x);
            int sysAny = 0;
#pragma unroll
            for (int k = 0; k < NCH; k++)
            {
                sysAny |= st.activeNext[s * NCH + k];
            }
            int aNext = 0;
            double rhoNext = 0;
            if (t < NCH)
            {
                aNext = st.activeNext[s * NCH + t];
                rhoNext = st.rhoNext[s * NCH + t];
            }
            if (i0 < end)
            {
                r0 = r[i0];
                x0 = x[i0];
                p0 = p[i0];
                q0 = q[i0];
                if (row0 < color0End)
                {
                    d0 = invDiagP[row0];
                }
            }
            const FusedSlice sl = fusedSlice(bm, range, lb, dirParts, dirParts);
            const double pq = channelPartialSum<NCH>(st.partials3 + static_cast<size_t>(sl.base) * NCH, sl.rhoParts, sm);
            // block 0 makes the direction kernel's decisions (every system)
            // the solve state, also when they stop the solve; no block of
            // this launch reads what it writes: the blocks test their own
            // system's activeNext and read rhoNext
            if (lb == 0)
            {
                const int K = bm.numSystems * NCH;
                int any = 0;
                for (int k = t; k < K; k += kBlock)
                {
                    const int a = st.activeNext[k];
                    st.active[k] = a;
                    st.rho[k] = st.rhoNext[k];
                    any |= a;
                }
                any = __syncthreads_or(any);
                if (t == 0)
                {
                    st.flags[0] = any ? 0 : 1;
                    st.flags[1] = 0; // the direction kernel before has used p = z
                }
            }
            if (!sysAny)
            {
                return;
            }
            if (t < NCH)
            {
                shAct[t] = aNext;
                shAlpha[t] = (aNext && pq != 0) ? rhoNext / pq : 0.0;
            }
            __syncthreads();
            go = true;
            act = shAct[c];
            alpha = static_cast<T>(shAlpha[c]);
        }
        else
        {
            if (st.flags[0])
            {
                return;
            }
            if (lb == 0 && threadIdx.x == 0)
            {
                st.flags[1] = 0; // the direction kernel before has used p = z
            }
            go = systemActive<NCH>(st, s);
            act = st.active[s * NCH + c];
            alpha = static_cast<T>(st.alpha[s * NCH + c]);
            if (go && i0 < end)
            {
                r0 = r[i0];
                if (act)
                {
                    x0 = x[i0];
                    p0 = p[i0];
                    q0 = q[i0];
                }
                if (row0 < color0End)
                {
                    d0 = invDiagP[row0];
                }
            }
        }
        double sum = 0;
        if (go)
        {
            auto finish = [&](long i, int row, T ri, T xi, T pi, T qi, P di)
            {
                if (act)
                {
                    x[i] = fmaT(alpha, pi, xi);
                    ri = fmaT(-alpha, qi, ri);
                    r[i] = ri;
                }
                const P rp = static_cast<P>(ri);
                if constexpr (!std::is_same_v<T, P>)
                {
                    if (act)
                    {
                        rP[i] = rp;
                    }
                }
                if (row < color0End)
                {
                    z[i] = rp * di;
                }
                sum = fma(static_cast<double>(ri), static_cast<double>(ri), sum);
            };
            if (i0 < end)
            {
                finish(i0, row0, r0, x0, p0, q0, d0);
            }
            for (long i = i0 + kBlock; i < end; i += kBlock)
            {
                const int row = static_cast<int>(i / NCH);
                const T ri = r[i];
                T xi = T(0), pi = T(0), qi = T(0);
                if (act)
                {
                    xi = x[i];
                    pi = p[i];
                    qi = q[i];
                }
                const P di = row < color0End ? invDiagP[row] : P(0);
                finish(i, row, ri, xi, pi, qi, di);
            }
        }
        if constexpr (kWide<NCH>)
        {
            const double total = channelBlockSum<NCH>(sum, sm);
            storeChannelPartial<NCH, NCH>(total, st.partials2, range.w, 0);
        }
        else
        {
            double part[NCH];
#pragma unroll
            for (int cc = 0; cc < NCH; cc++)
            {
                part[cc] = cc == c ? sum : 0.0;
            }
            blockSum<NCH>(part, sm);
            storePartial<NCH>(part, st.partials2, range.w);
        }
    }

    // The update kernel fused with the first forward sweep (one system, flat
    // map, fused reductions, 1-2 channels): the update of every row and, on
    // the rows of colour 1, their forward symmetric-GS sweep.  That sweep's
    // lower part gathers colour-0 rows only; their z = (P)(r - alpha q) / d is
    // computed from r and q as this launch finds them, with the arithmetic of
    // the stored z, so the sweep equals update-then-sweep bit for bit.  This
    // launch therefore leaves the colour-0 residuals (r, and rP when it is the
    // same array) unstored; the rho kernel of the same iteration stores them
    // (pcgRhoKernel r0Fix).  Workgroups holding colour-1 rows are interleaved
    // evenly among the others, so every XCD gets its share of sweep work.
    template <typename T, typename P, int LANES, int NCH>
    __global__ void __launch_bounds__(kBlock) pcgUpdateSweepKernel(McgsMatrix<P> M, BlockMap bm, SolveState st, const T * __restrict__ p,
                                                                   const T * __restrict__ q, T * __restrict__ x, T * r, P * rP,
                                                                   P * __restrict__ z, int color0End, int c1Begin, int c1End,
                                                                   int specialBegin, int specialCount, int dirParts)
    {
        static_assert(!kWide<NCH>, "the fused update and sweep takes 1 or 2 channels");
        static_assert(kBlock % NCH == 0, "channels must divide the block");
        __shared__ double sm[NCH * kWaves];
        __shared__ double shAlpha[NCH];
        __shared__ int shAct[NCH];
        const int k = interleaveFirst(logicalBlock(blockIdx.x, gridDim.x), gridDim.x, specialCount);
        const int lb = k < specialCount ? specialBegin + k : (k - specialCount < specialBegin ? k - specialCount : k);
        const int4 range = blockRange(bm, lb);
        const int s = range.z;
        const int t = static_cast<int>(threadIdx.x);
        // the block's colour-1 rows; the first slab's row bounds go out with
        // the prologue's loads
        const int s0 = max(range.x, c1Begin), s1 = min(range.y, c1End);
        int srow = s0 + t / LANES;
        bool svalid = srow < s1;
        int j0 = svalid ? M.rowStart[srow] : 0;
        int j1 = svalid ? M.rowSplit[srow] : 0;
        int sysAny = 0;
#pragma unroll
        for (int kk = 0; kk < NCH; kk++)
        {
            sysAny |= st.activeNext[s * NCH + kk];
        }
        int aNext = 0;
        double rhoNext = 0;
        if (t < NCH)
        {
            aNext = st.activeNext[s * NCH + t];
            rhoNext = st.rhoNext[s * NCH + t];
        }
        const FusedSlice sl = fusedSlice(bm, range, lb, dirParts, dirParts);
        const double pq = channelPartialSum<NCH>(st.partials3 + static_cast<size_t>(sl.base) * NCH, sl.rhoParts, sm);
        // block 0 makes the direction kernel's decisions the solve state (as
        // pcgUpdateKernel); no block of this launch reads what it writes
        if (lb == 0)
        {
            const int K = bm.numSystems * NCH;
            int any = 0;
            for (int kk = t; kk < K; kk += kBlock)
            {
                const int a = st.activeNext[kk];
                st.active[kk] = a;
                st.rho[kk] = st.rhoNext[kk];
                any |= a;
            }
            any = __syncthreads_or(any);
            if (t == 0)
            {
                st.flags[0] = any ? 0 : 1;
                st.flags[1] = 0;
            }
        }
        if (!sysAny | (j1 < j0))
        {
            return;
        }
        if (t < NCH)
        {
            shAct[t] = aNext;
            const double a = (aNext && pq != 0) ? rhoNext / pq : 0.0;
            shAlpha[t] = a;
            if (sl.leader)
            {
                st.alpha[s * NCH + t] = a; // for the rho kernel's colour-0 residuals
            }
        }
        __syncthreads();
        double part[NCH];
#pragma unroll
        for (int cc = 0; cc < NCH; cc++)
        {
            part[cc] = 0;
        }
        {
            // 1. the update of the rows outside colour 1, one element per thread
            const int c = t % NCH;
            const int act = shAct[c];
            const T alpha = static_cast<T>(shAlpha[c]);
            double sum = 0;
            const long end = static_cast<long>(range.y) * NCH;
            for (long i = static_cast<long>(range.x) * NCH + t; i < end; i += kBlock)
            {
                const int row = static_cast<int>(i / NCH);
                if (row >= c1Begin && row < c1End)
                {
                    continue;
                }
                T ri = r[i];
                if (act)
                {
                    const T xi = x[i], pi = p[i], qi = q[i];
                    x[i] = fmaT(alpha, pi, xi);
                    ri = fmaT(-alpha, qi, ri);
                }
                const P rp = static_cast<P>(ri);
                if (row < color0End)
                {
                    z[i] = rp * M.invDiag[row];
                    if constexpr (!std::is_same_v<T, P>)
                    {
                        if (act)
                        {
                            rP[i] = rp; // a separate array: not read by this launch
                        }
                    }
                }
                else if (act)
                {
                    r[i] = ri;
                    if constexpr (!std::is_same_v<T, P>)
                    {
                        rP[i] = rp;
                    }
                }
                sum = fma(static_cast<double>(ri), static_cast<double>(ri), sum);
            }
            part[c] = sum;
        }
        if (s0 < s1)
        {
            // 2. colour-1 rows: the update and the forward sweep, LANES lanes
            // per row
            int actC[NCH];
            T alphaC[NCH];
#pragma unroll
            for (int cc = 0; cc < NCH; cc++)
            {
                actC[cc] = shAct[cc];
                alphaC[cc] = static_cast<T>(shAlpha[cc]);
            }
            auto zColour0 = [&](int col, int cc) -> P
            {
                const long jj = static_cast<long>(col) * NCH + cc;
                T rj = r[jj];
                if (actC[cc])
                {
                    rj = fmaT(-alphaC[cc], q[jj], rj);
                }
                return static_cast<P>(rj) * M.invDiag[col];
            };
            for (int base = s0; base < s1; base += kBlock / LANES)
            {
                if (base != s0)
                {
                    srow = base + t / LANES;
                    svalid = srow < s1;
                    j0 = svalid ? M.rowStart[srow] : 0;
                    j1 = svalid ? M.rowSplit[srow] : 0;
                }
                const bool lead = svalid && groupLeader<LANES>();
                P inv = P(0);
                T ri[NCH], xi[NCH], pi[NCH], qi[NCH];
#pragma unroll
                for (int cc = 0; cc < NCH; cc++)
                {
                    ri[cc] = xi[cc] = pi[cc] = qi[cc] = T(0);
                }
                if (lead)
                {
                    inv = M.invDiag[srow];
#pragma unroll
                    for (int cc = 0; cc < NCH; cc++)
                    {
                        const long i = static_cast<long>(srow) * NCH + cc;
                        ri[cc] = r[i];
                        if (actC[cc])
                        {
                            xi[cc] = x[i];
                            pi[cc] = p[i];
                            qi[cc] = q[i];
                        }
                    }
                }
                P acc[NCH];
                rowDotG<P, LANES, NCH>(M, j0, j1, zColour0, acc);
                if (lead)
                {
#pragma unroll
                    for (int cc = 0; cc < NCH; cc++)
                    {
                        const long i = static_cast<long>(srow) * NCH + cc;
                        T rn = ri[cc];
                        if (actC[cc])
                        {
                            x[i] = fmaT(alphaC[cc], pi[cc], xi[cc]);
                            rn = fmaT(-alphaC[cc], qi[cc], rn);
                            r[i] = rn;
                        }
                        const P rp = static_cast<P>(rn);
                        if constexpr (!std::is_same_v<T, P>)
                        {
                            if (actC[cc])
                            {
                                rP[i] = rp;
                            }
                        }
                        z[i] = (rp - acc[cc]) * inv;
                        part[cc] = fma(static_cast<double>(rn), static_cast<double>(rn), part[cc]);
                    }
                }
            }
        }
        blockSum<NCH>(part, sm);
        storePartial<NCH>(part, st.partials2, range.w);
    }

    // Per-system RMS residuals of gsResidualKernel, stop test and flags.
    template <int NCH>
    __global__ void __launch_bounds__(kFinBlock) gsResidualFinalizeKernel(BlockMap bm, SolveState st, int iter, double * totals)
    {
        __shared__ double sm[NCH * kFinWaves];
        if (st.flags[0])
        {
            return; // the producer skipped as well
        }
        systemTotals<NCH>(bm, st.partials, totals, sm);
        if (threadIdx.x == 0)
        {
            int anyActive = 0;
            for (int sys = 0; sys < bm.numSystems; sys++)
            {
                if (!systemActive<NCH>(st, sys))
                {
                    continue;
                }
                bool all = true;
                for (int c = 0; c < NCH; c++)
                {
                    const double m = sqrt(totals[sys * NCH + c] / bm.systemRows[sys]);
                    st.mse[sys * NCH + c] = m;
                    if (!(m < st.tol2))
                    {
                        all = false;
                    }
                    if (!(m == m) || m > 1e300)
                    {
                        st.flags[2] = 1;
                    }
                }
                if (all)
                {
                    for (int c = 0; c < NCH; c++)
                    {
                        st.active[sys * NCH + c] = 0;
                        st.iters[sys * NCH + c] = iter;
                    }
                }
                else
                {
                    anyActive = 1;
                }
            }
            if (!anyActive)
            {
                st.flags[0] = 1;
            }
        }
    }

    // |b|^2, |r|^2 of pcgInitKernel; resets the iteration scalars.  gate
    // (optional) switches off channels whose outer refinement has converged.
    template <int NCH>
    __global__ void __launch_bounds__(kFinBlock) pcgInitFinalizeKernel(BlockMap bm, SolveState st, const int * __restrict__ gate, double * totals)
    {
        __shared__ double sm[2 * NCH * kFinWaves];
        systemTotals<2 * NCH>(bm, st.partials, totals, sm);
        if (threadIdx.x == 0)
        {
            int anyActive = 0;
            for (int s = 0; s < bm.numSystems; s++)
            {
                for (int c = 0; c < NCH; c++)
                {
                    const int k = s * NCH + c;
                    const double bb = totals[s * 2 * NCH + c];
                    const double rr = totals[s * 2 * NCH + NCH + c];
                    st.bb[k] = bb;
                    st.rr[k] = rr;
                    st.ref[k] = bb > 0 ? bb : rr;
                    st.iters[k] = 0;
                    st.alpha[k] = 0;
                    st.beta[k] = 0;
                    st.rho[k] = 0;
                    if (!(rr == rr) || rr > 1e300)
                    {
                        st.flags[2] = 1;
                    }
                    int act = rr > st.tol2 * st.ref[k] ? 1 : 0;
                    if (gate && !gate[k])
                    {
                        act = 0;
                    }
                    st.active[k] = act;
                    anyActive |= act;
                }
            }
            st.flags[0] = anyActive ? 0 : 1;
            st.flags[1] = 1;
        }
    }

    // rho = r.z and beta = rho / rho_old.
    template <int NCH>
    // withUpdate: also finish the preceding update's |r|^2 (partials2): the
    // iteration count and the convergence test of the update come first,
    // then rho and beta for the channels that go on.  (The sweeps and the
    // rho kernel in between ran on the update's result either way.)
    __global__ void __launch_bounds__(kFinBlock) pcgRhoFinalizeKernel(BlockMap bm, SolveState st, double * totals, int withUpdate)
    {
        __shared__ double sm[NCH * kFinWaves];
        __shared__ double shRho[kStaged], shRef[kStaged];
        __shared__ int shAct[kStaged], shIters[kStaged];
        const int K = bm.numSystems * NCH;
        const bool staged = K <= kStaged;
        const int done = st.flags[0];
        const int firstFlag = st.flags[1];
        if (staged && static_cast<int>(threadIdx.x) < K)
        {
            shRho[threadIdx.x] = st.rho[threadIdx.x];
            shRef[threadIdx.x] = st.ref[threadIdx.x];
            shAct[threadIdx.x] = st.active[threadIdx.x];
            shIters[threadIdx.x] = st.iters[threadIdx.x];
        }
        double * rrTotals = totals + K;
        if (withUpdate)
        {
            systemTotals<NCH>(bm, st.partials2, rrTotals, sm);
        }
        systemTotals<NCH>(bm, st.partials, totals, sm); // ends with a barrier
        if (done)
        {
            return; // the producers skipped as well
        }
        if (threadIdx.x == 0)
        {
            const bool first = firstFlag != 0;
            int anyActive = 0;
            for (int k = 0; k < K; k++)
            {
                if (!(staged ? shAct[k] : st.active[k]))
                {
                    continue;
                }
                if (withUpdate)
                {
                    const double rr = rrTotals[k];
                    st.rr[k] = rr;
                    st.iters[k] = (staged ? shIters[k] : st.iters[k]) + 1;
                    if (!(rr == rr) || rr > 1e300)
                    {
                        st.flags[2] = 1;
                        st.active[k] = 0;
                        continue;
                    }
                    if (rr <= st.tol2 * (staged ? shRef[k] : st.ref[k]))
                    {
                        st.active[k] = 0;
                        continue;
                    }
                }
                anyActive = 1;
                const double rhoNew = totals[k];
                const double rhoOld = staged ? shRho[k] : st.rho[k];
                st.beta[k] = (first || rhoOld == 0) ? 0.0 : rhoNew / rhoOld;
                st.rho[k] = rhoNew;
                if (!(rhoNew == rhoNew))
                {
                    st.flags[2] = 1;
                }
            }
            if (withUpdate && !anyActive)
            {
                st.flags[0] = 1;
            }
        }
    }

    // Fused reductions: the convergence test of a solve's last update (no
    // direction kernel follows to make it): rr, the iteration count and the
    // stop flag as pcgRhoFinalizeKernel records them.
    template <int NCH>
    __global__ void __launch_bounds__(kFinBlock) pcgUpdateCheckKernel(BlockMap bm, SolveState st, double * totals)
    {
        __shared__ double sm[NCH * kFinWaves];
        if (st.flags[0])
        {
            return;
        }
        systemTotals<NCH>(bm, st.partials2, totals, sm);
        if (threadIdx.x == 0)
        {
            int any = 0;
            for (int k = 0; k < bm.numSystems * NCH; k++)
            {
                if (!st.active[k])
                {
                    continue;
                }
                const double rr = totals[k];
                st.rr[k] = rr;
                st.iters[k] = st.iters[k] + 1;
                if (!(rr == rr) || rr > 1e300)
                {
                    st.flags[2] = 1;
                    st.active[k] = 0;
                    continue;
                }
                if (rr <= st.tol2 * st.ref[k])
                {
                    st.active[k] = 0;
                    continue;
                }
                any = 1;
            }
            if (!any)
            {
                st.flags[0] = 1;
            }
        }
    }

    // alpha = rho / (p.q).
    template <int NCH>
    __global__ void __launch_bounds__(kFinBlock) pcgDirectionFinalizeKernel(BlockMap bm, SolveState st, double * totals)
    {
        __shared__ double sm[NCH * kFinWaves];
        __shared__ double shRho[kStaged];
        __shared__ int shAct[kStaged];
        const int K = bm.numSystems * NCH;
        const bool staged = K <= kStaged;
        const int done = st.flags[0];
        if (staged && static_cast<int>(threadIdx.x) < K)
        {
            shRho[threadIdx.x] = st.rho[threadIdx.x];
            shAct[threadIdx.x] = st.active[threadIdx.x];
        }
        systemTotals<NCH>(bm, st.partials, totals, sm); // ends with a barrier
        if (done)
        {
            return; // the producer skipped as well
        }
        if (threadIdx.x == 0)
        {
            for (int k = 0; k < K; k++)
            {
                const double pq = totals[k];
                const double rho = staged ? shRho[k] : st.rho[k];
                const int act = staged ? shAct[k] : st.active[k];
                st.alpha[k] = (act && pq != 0) ? rho / pq : 0.0;
            }
            st.flags[1] = 0;
        }
    }

    // x64 += (double) e  over all rows and channels (outer refinement update).
    template <typename S>
    __global__ void addConvertKernel(long n, const S * __restrict__ e, double * __restrict__ x)
    {
        for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * blockDim.x)
        {
            x[i] += static_cast<double>(e[i]);
        }
    }

    // Mean of each (system, channel) over its rows: mean[s*NCH + c] (FP64).
    template <typename T, int NCH>
    __global__ void __launch_bounds__(kBlock) meanKernel(int n, BlockMap bm, const T * __restrict__ v, double * partials)
    {
        __shared__ double sm[NCH * kWaves];
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        double part[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++)
        {
            part[c] = 0;
        }
        const int4 range = blockRange(bm, lb);
        for (int row = range.x + static_cast<int>(threadIdx.x); row < range.y; row += kBlock)
        {
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                part[c] += static_cast<double>(v[static_cast<long>(row) * NCH + c]);
            }
        }
        blockSum<NCH>(part, sm);
        storePartial<NCH>(part, partials, range.w);
    }

    // mean[s*NCH + c] from the partial sums of meanKernel.
    template <int NCH>
    __global__ void __launch_bounds__(kFinBlock) meanFinalizeKernel(BlockMap bm, const double * __restrict__ partials, double * totals, double * mean)
    {
        __shared__ double sm[NCH * kFinWaves];
        systemTotals<NCH>(bm, partials, totals, sm);
        if (threadIdx.x == 0)
        {
            for (int s = 0; s < bm.numSystems; s++)
            {
                for (int c = 0; c < NCH; c++)
                {
                    mean[s * NCH + c] = totals[s * NCH + c] / bm.systemRows[s];
                }
            }
        }
    }

    // v -= mean of its (system, channel).
    template <typename T, int NCH>
    __global__ void __launch_bounds__(kBlock) subtractMeanKernel(int n, BlockMap bm, T * __restrict__ v, const double * __restrict__ mean)
    {
        const int lb = logicalBlock(blockIdx.x, gridDim.x);
        const int4 range = blockRange(bm, lb);
        const int s = range.z;
        for (int row = range.x + static_cast<int>(threadIdx.x); row < range.y; row += kBlock)
        {
#pragma unroll
            for (int c = 0; c < NCH; c++)
            {
                const long i = static_cast<long>(row) * NCH + c;
                v[i] = static_cast<T>(static_cast<double>(v[i]) - mean[s * NCH + c]);
            }
        }
    }
} // namespace nsfem
