// Host launchers for the generic sparse / vector kernels.  Everything is
// enqueued on the given stream; nothing here synchronises except dot().
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace nsfem
{
    enum DType : int
    {
        F32 = 0, // ids match the linalg .bin format (LinAlg/src/io.cpp:11-21)
        F64 = 1
    };

    inline size_t dtypeSize(int dt) { return dt == F64 ? 8 : 4; }
    void checkDtype(int dt);

    // Lanes per row for a CSR product, from the mean row length.
    int lanesForMeanRow(double meanNnzPerRow);

    struct CsrView
    {
        int rows, cols, nnz;
        const int * rowStart;
        const int * column;
        const void * values;
        int dtype;
    };

    // y[c*ldy + i] = alpha * sum_j A[i,j] x[c*ldx + j] + beta * y[...]
    // rowIds (optional, device): output row i is A's row rowIds[i]; outRows rows.
    void csrMul(const CsrView & A, int numCh, const void * x, long ldx, void * y, long ldy, double alpha, double beta,
                const int * rowIds, int outRows, int lanes, hipStream_t s);

    void axpy(int dt, long n, double alpha, const void * src, void * dst, hipStream_t s);
    void scale(int dt, long n, double alpha, void * dst, hipStream_t s);
    void geam(int dt, long n, double alpha, const void * a, double beta, const void * b, void * c, hipStream_t s);
    void fill(int dt, long n, double value, void * dst, hipStream_t s);
    // x <- 2 x - prev and prev <- the old x, elementwise
    void extrapolate(int dt, long n, void * x, void * prev, hipStream_t s);
    // out <- 2 cur - prev and prev <- cur, elementwise
    void extrapolateTo(int dt, long n, const void * cur, void * prev, void * out, hipStream_t s);
    void fillInt(long n, int value, int * dst, hipStream_t s);
    void convert(int srcDt, int dstDt, long n, const void * src, void * dst, hipStream_t s);
    void scatter(int dt, int nnz, const int * idx, const void * vals, void * dense, hipStream_t s);
    void gather(int dt, int nnz, const void * dense, const int * idx, void * vals, hipStream_t s);
    void permute(int srcDt, int dstDt, int n, int nch, const int * perm, const void * src, void * dst, bool inverse, hipStream_t s);

    // Blocking FP64-accumulated dot product(s): out[k] = x_k . y_k, k < K (1 or 2).
    // Partial sums of at most kDotMaxBlocks workgroups, summed in workgroup
    // order (deterministic); the scratch is a library buffer allocated on
    // first use.
    constexpr int kDotMaxBlocks = 2048;
    void dot(int dt, long n, int K, const void * x, long ldx, const void * y, long ldy, double * out, hipStream_t s);
    // The same single dot product (the same kernels, bit-identical result)
    // with its scratch in a fixed file-scope device array instead of
    // allocated memory: no allocation, ever.
    constexpr int kNoAllocDotScratchDoubles = 2 * kDotMaxBlocks + 2;
    void dotNoAlloc(int dt, long n, const void * x, const void * y, double * out, hipStream_t s);

    // Grid size for grid-stride element-wise kernels.
    int elementwiseGrid(long n);
    // Achievable streaming bandwidth: a device-to-device copy of bytes
    // (rounded down to 16) timed with HIP events on s over reps back-to-back
    // launches; returns read + write bytes per second.
    double copyBandwidth(size_t bytes, int reps, hipStream_t s);
} // namespace nsfem
