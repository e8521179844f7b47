#include "ops.hpp"

#include <algorithm>
#include <mutex>

#include "kernels.hpp"

namespace nsfem
{
    hipStream_t libraryStream()
    {
        static hipStream_t stream = []
        {
            hipStream_t s = nullptr;
            NSFEM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            return s;
        }();
        return stream;
    }

    void syncLibraryStream() { NSFEM_HIP(hipStreamSynchronize(libraryStream())); }

    void checkDtype(int dt)
    {
        if (dt != F32 && dt != F64)
        {
            throw std::invalid_argument("Unknown dtype " + std::to_string(dt) + " (0 = f32, 1 = f64)");
        }
    }

    int lanesForMeanRow(double m)
    {
        if (m < 5)
        {
            return 4;
        }
        if (m < 12)
        {
            return 8;
        }
        if (m < 24)
        {
            return 16;
        }
        if (m < 48)
        {
            return 32;
        }
        return 64;
    }

    int elementwiseGrid(long n)
    {
        const long blocks = (n + 255) / 256;
        return static_cast<int>(std::max(1L, std::min(blocks, 8192L)));
    }

    namespace
    {
        template <typename T>
        void csrMulT(const CsrView & A, int numCh, const void * x, long ldx, void * y, long ldy, double alpha, double beta,
                     const int * rowIds, int outRows, int lanes, hipStream_t s)
        {
            const T * vals = static_cast<const T *>(A.values);
            auto launch = [&]<int L, int C>(const T * xs, T * ys)
            {
                const long threads = static_cast<long>(outRows) * L;
                const int grid = ceilDiv(threads, kBlock);
                csrMulKernel<T, L, C><<<grid, kBlock, 0, s>>>(outRows, A.rowStart, A.column, vals, xs, ldx, ys, ldy,
                                                              static_cast<T>(alpha), static_cast<T>(beta), rowIds);
            };
            auto byLanes = [&]<int C>(const T * xs, T * ys)
            {
                switch (lanes)
                {
                case 4: launch.template operator()<4, C>(xs, ys); break;
                case 8: launch.template operator()<8, C>(xs, ys); break;
                case 16: launch.template operator()<16, C>(xs, ys); break;
                case 32: launch.template operator()<32, C>(xs, ys); break;
                case 64: launch.template operator()<64, C>(xs, ys); break;
                default: throw std::invalid_argument("lanes must be 4, 8, 16, 32 or 64");
                }
            };
            // Channels are processed two at a time (the matrix is read once per pair).
            for (int c0 = 0; c0 < numCh; c0 += 2)
            {
                const T * xs = static_cast<const T *>(x) + static_cast<long>(c0) * ldx;
                T * ys = static_cast<T *>(y) + static_cast<long>(c0) * ldy;
                if (numCh - c0 >= 2)
                {
                    byLanes.template operator()<2>(xs, ys);
                }
                else
                {
                    byLanes.template operator()<1>(xs, ys);
                }
            }
            NSFEM_HIP(hipGetLastError());
        }
    } // namespace

    void csrMul(const CsrView & A, int numCh, const void * x, long ldx, void * y, long ldy, double alpha, double beta,
                const int * rowIds, int outRows, int lanes, hipStream_t s)
    {
        if (numCh < 1)
        {
            throw std::invalid_argument("SPMM channels must be at least 1");
        }
        if (outRows == 0)
        {
            return;
        }
        if (lanes <= 0)
        {
            lanes = lanesForMeanRow(A.rows > 0 ? static_cast<double>(A.nnz) / A.rows : 1.0);
        }
        if (A.dtype == F64)
        {
            csrMulT<double>(A, numCh, x, ldx, y, ldy, alpha, beta, rowIds, outRows, lanes, s);
        }
        else
        {
            csrMulT<float>(A, numCh, x, ldx, y, ldy, alpha, beta, rowIds, outRows, lanes, s);
        }
    }

#define NSFEM_TYPED(dt, CALL_F32, CALL_F64) \
    do                                      \
    {                                       \
        checkDtype(dt);                     \
        if ((dt) == F64)                    \
        {                                   \
            CALL_F64;                       \
        }                                   \
        else                                \
        {                                   \
            CALL_F32;                       \
        }                                   \
        NSFEM_HIP(hipGetLastError());       \
    } while (0)

    void axpy(int dt, long n, double alpha, const void * src, void * dst, hipStream_t s)
    {
        if (n <= 0)
        {
            return;
        }
        const int g = elementwiseGrid(n);
        NSFEM_TYPED(dt,
                    (axpyKernel<float><<<g, 256, 0, s>>>(n, static_cast<float>(alpha), static_cast<const float *>(src), static_cast<float *>(dst))),
                    (axpyKernel<double><<<g, 256, 0, s>>>(n, alpha, static_cast<const double *>(src), static_cast<double *>(dst))));
    }

    void scale(int dt, long n, double alpha, void * dst, hipStream_t s)
    {
        if (n <= 0)
        {
            return;
        }
        const int g = elementwiseGrid(n);
        NSFEM_TYPED(dt, (scaleKernel<float><<<g, 256, 0, s>>>(n, static_cast<float>(alpha), static_cast<float *>(dst))),
                    (scaleKernel<double><<<g, 256, 0, s>>>(n, alpha, static_cast<double *>(dst))));
    }

    double copyBandwidth(size_t bytes, int reps, hipStream_t s)
    {
        const long n = static_cast<long>(bytes / 16);
        if (n < 1 || reps < 1)
        {
            throw std::invalid_argument("copy bandwidth: at least 16 bytes and one repetition");
        }
        DeviceArray<double2> src(n), dst(n);
        NSFEM_HIP(hipMemsetAsync(src.get(), 0, n * sizeof(double2), s));
        const long blocks = std::min<long>((n + 255) / 256, 1L << 30);
        const int g = static_cast<int>(blocks);
        streamCopyKernel<<<g, 256, 0, s>>>(n, src.get(), dst.get());
        NSFEM_HIP(hipGetLastError());
        NSFEM_HIP(hipStreamSynchronize(s));
        // the repetitions queue behind a 20 ms delay kernel and run back to back
        delayKernel<<<1, 64, 0, s>>>(2000000);
        hipEvent_t e0, e1;
        NSFEM_HIP(hipEventCreate(&e0));
        NSFEM_HIP(hipEventCreate(&e1));
        NSFEM_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < reps; i++)
        {
            streamCopyKernel<<<g, 256, 0, s>>>(n, src.get(), dst.get());
        }
        NSFEM_HIP(hipEventRecord(e1, s));
        NSFEM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        NSFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        NSFEM_HIP(hipEventDestroy(e0));
        NSFEM_HIP(hipEventDestroy(e1));
        return 2.0 * static_cast<double>(n) * sizeof(double2) * reps / (static_cast<double>(ms) * 1e-3);
    }

    void geam(int dt, long n, double alpha, const void * a, double beta, const void * b, void * c, hipStream_t s)
    {
        if (n <= 0)
        {
            return;
        }
        const int g = elementwiseGrid(n);
        NSFEM_TYPED(dt,
                    (geamKernel<float><<<g, 256, 0, s>>>(n, static_cast<float>(alpha), static_cast<const float *>(a), static_cast<float>(beta),
                                                         static_cast<const float *>(b), static_cast<float *>(c))),
                    (geamKernel<double><<<g, 256, 0, s>>>(n, alpha, static_cast<const double *>(a), beta, static_cast<const double *>(b),
                                                          static_cast<double *>(c))));
    }

    void extrapolate(int dt, long n, void * x, void * prev, hipStream_t s)
    {
        if (n <= 0)
        {
            return;
        }
        const int g = elementwiseGrid(n);
        NSFEM_TYPED(dt, (extrapolateKernel<float><<<g, 256, 0, s>>>(n, static_cast<float *>(x), static_cast<float *>(prev))),
                    (extrapolateKernel<double><<<g, 256, 0, s>>>(n, static_cast<double *>(x), static_cast<double *>(prev))));
    }

    void extrapolateTo(int dt, long n, const void * cur, void * prev, void * out, hipStream_t s)
    {
        if (n <= 0)
        {
            return;
        }
        const int g = elementwiseGrid(n);
        NSFEM_TYPED(dt,
                    (extrapolateToKernel<float><<<g, 256, 0, s>>>(n, static_cast<const float *>(cur), static_cast<float *>(prev),
                                                                  static_cast<float *>(out))),
                    (extrapolateToKernel<double><<<g, 256, 0, s>>>(n, static_cast<const double *>(cur), static_cast<double *>(prev),
                                                                   static_cast<double *>(out))));
    }

    void fill(int dt, long n, double value, void * dst, hipStream_t s)
    {
        if (n <= 0)
        {
            return;
        }
        const int g = elementwiseGrid(n);
        NSFEM_TYPED(dt, (fillKernel<float><<<g, 256, 0, s>>>(n, static_cast<float>(value), static_cast<float *>(dst))),
                    (fillKernel<double><<<g, 256, 0, s>>>(n, value, static_cast<double *>(dst))));
    }

    void fillInt(long n, int value, int * dst, hipStream_t s)
    {
        if (n <= 0)
        {
            return;
        }
        fillKernel<int><<<elementwiseGrid(n), 256, 0, s>>>(n, value, dst);
        NSFEM_HIP(hipGetLastError());
    }

    void convert(int srcDt, int dstDt, long n, const void * src, void * dst, hipStream_t s)
    {
        checkDtype(srcDt);
        checkDtype(dstDt);
        if (n <= 0)
        {
            return;
        }
        const int g = elementwiseGrid(n);
        if (srcDt == F64 && dstDt == F32)
        {
            convertKernel<double, float><<<g, 256, 0, s>>>(n, static_cast<const double *>(src), static_cast<float *>(dst));
        }
        else if (srcDt == F32 && dstDt == F64)
        {
            convertKernel<float, double><<<g, 256, 0, s>>>(n, static_cast<const float *>(src), static_cast<double *>(dst));
        }
        else
        {
            NSFEM_HIP(hipMemcpyAsync(dst, src, n * dtypeSize(srcDt), hipMemcpyDeviceToDevice, s));
        }
        NSFEM_HIP(hipGetLastError());
    }

    void scatter(int dt, int nnz, const int * idx, const void * vals, void * dense, hipStream_t s)
    {
        if (nnz <= 0)
        {
            return;
        }
        const int g = ceilDiv(nnz, 256);
        NSFEM_TYPED(dt, (scatterKernel<float><<<g, 256, 0, s>>>(nnz, idx, static_cast<const float *>(vals), static_cast<float *>(dense))),
                    (scatterKernel<double><<<g, 256, 0, s>>>(nnz, idx, static_cast<const double *>(vals), static_cast<double *>(dense))));
    }

    void gather(int dt, int nnz, const void * dense, const int * idx, void * vals, hipStream_t s)
    {
        if (nnz <= 0)
        {
            return;
        }
        const int g = ceilDiv(nnz, 256);
        NSFEM_TYPED(dt, (gatherKernel<float><<<g, 256, 0, s>>>(nnz, static_cast<const float *>(dense), idx, static_cast<float *>(vals))),
                    (gatherKernel<double><<<g, 256, 0, s>>>(nnz, static_cast<const double *>(dense), idx, static_cast<double *>(vals))));
    }

    void permute(int srcDt, int dstDt, int n, int nch, const int * perm, const void * src, void * dst, bool inverse, hipStream_t s)
    {
        checkDtype(srcDt);
        checkDtype(dstDt);
        const long total = static_cast<long>(n) * nch;
        if (total <= 0)
        {
            return;
        }
        // one element per thread (each is a dependent index load then a
        // gather or scatter; a capped grid's second round would wait for the
        // first): the full grid
        const int g = static_cast<int>(std::min<long>((total + 255) / 256, 1L << 30));
        const int inv = inverse ? 1 : 0;
        if (srcDt == F64 && dstDt == F64)
        {
            permuteKernel<double, double><<<g, 256, 0, s>>>(n, nch, perm, static_cast<const double *>(src), static_cast<double *>(dst), inv);
        }
        else if (srcDt == F32 && dstDt == F32)
        {
            permuteKernel<float, float><<<g, 256, 0, s>>>(n, nch, perm, static_cast<const float *>(src), static_cast<float *>(dst), inv);
        }
        else if (srcDt == F64 && dstDt == F32)
        {
            permuteKernel<double, float><<<g, 256, 0, s>>>(n, nch, perm, static_cast<const double *>(src), static_cast<float *>(dst), inv);
        }
        else
        {
            permuteKernel<float, double><<<g, 256, 0, s>>>(n, nch, perm, static_cast<const float *>(src), static_cast<double *>(dst), inv);
        }
        NSFEM_HIP(hipGetLastError());
    }

    namespace
    {
        struct DotScratch
        {
            DeviceArray<double> partials;
            DeviceArray<double> out;
            std::mutex mutex;
        };
        DotScratch & dotScratch()
        {
            static DotScratch * d = []
            {
                auto * p = new DotScratch;
                p->partials.resize(2 * kDotMaxBlocks);
                p->out.resize(2);
                return p;
            }();
            return *d;
        }
    } // namespace

    namespace
    {
        // Launches the two dot-product kernels: partials[g * K], then out[K]
        // (device); returns g.
        void dotLaunch(int dt, long n, int K, const void * x, long ldx, const void * y, long ldy, double * partials, double * devOut,
                       hipStream_t s)
        {
            const int g = static_cast<int>(std::max(1L, std::min((n + kBlock - 1) / kBlock, static_cast<long>(kDotMaxBlocks))));
            if (dt == F64)
            {
                if (K == 1)
                {
                    dotKernel<double, 1><<<g, kBlock, 0, s>>>(n, static_cast<const double *>(x), ldx, static_cast<const double *>(y), ldy,
                                                             partials);
                }
                else
                {
                    dotKernel<double, 2><<<g, kBlock, 0, s>>>(n, static_cast<const double *>(x), ldx, static_cast<const double *>(y), ldy,
                                                             partials);
                }
            }
            else
            {
                if (K == 1)
                {
                    dotKernel<float, 1><<<g, kBlock, 0, s>>>(n, static_cast<const float *>(x), ldx, static_cast<const float *>(y), ldy,
                                                            partials);
                }
                else
                {
                    dotKernel<float, 2><<<g, kBlock, 0, s>>>(n, static_cast<const float *>(x), ldx, static_cast<const float *>(y), ldy,
                                                            partials);
                }
            }
            if (K == 1)
            {
                dotFinalizeKernel<1><<<1, kBlock, 0, s>>>(g, partials, devOut);
            }
            else
            {
                dotFinalizeKernel<2><<<1, kBlock, 0, s>>>(g, partials, devOut);
            }
            NSFEM_HIP(hipGetLastError());
        }
    } // namespace

    void dot(int dt, long n, int K, const void * x, long ldx, const void * y, long ldy, double * out, hipStream_t s)
    {
        checkDtype(dt);
        if (K < 1 || K > 2)
        {
            throw std::invalid_argument("dot: K must be 1 or 2");
        }
        if (n <= 0)
        {
            for (int k = 0; k < K; k++)
            {
                out[k] = 0;
            }
            return;
        }
        auto & d = dotScratch();
        std::lock_guard<std::mutex> lock(d.mutex);
        dotLaunch(dt, n, K, x, ldx, y, ldy, d.partials.get(), d.out.get(), s);
        NSFEM_HIP(hipMemcpyAsync(out, d.out.get(), K * sizeof(double), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipStreamSynchronize(s));
    }

    // ------------------------------------------------------------------
    // Allocation-free dot product.  Its scratch (the first kernel's
    // per-workgroup partial sums and the result slot) is this file-scope
    // device array: one copy per GPU, loaded with the library's code object
    // and held until the process ends.  In host code the array's name is
    // not a device address; the address comes from hipGetSymbolAddress, once
    // per device with that device current, and only that address is used.
    // Nothing in it needs clearing: every partial the finalize kernel reads
    // is written by the same call's first kernel, and the result slot by the
    // finalize kernel.
    // ------------------------------------------------------------------
    static __device__ __align__(256) double scratch[kNoAllocDotScratchDoubles];

    namespace
    {
        constexpr int kMaxDevices = 64;
        std::mutex & noAllocMutex()
        {
            static std::mutex m;
            return m;
        }

        double * noAllocScratch()
        {
            int dev = -1;
            NSFEM_HIP(hipGetDevice(&dev));
            if (dev < 0 || dev >= kMaxDevices)
            {
                throw std::runtime_error("allocation-free dot: device index " + std::to_string(dev) + " out of range");
            }
            static double * ptrs[kMaxDevices] = {};
            if (!ptrs[dev])
            {
                void * p = nullptr;
                const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
                if (rc != hipSuccess || p == nullptr)
                {
                    throw DeviceError(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorName(rc));
                }
                ptrs[dev] = static_cast<double *>(p);
            }
            return ptrs[dev];
        }
    } // namespace

    void dotNoAlloc(int dt, long n, const void * x, const void * y, double * out, hipStream_t s)
    {
        checkDtype(dt);
        if (n < 0)
        {
            throw std::invalid_argument("dot: n must be non-negative");
        }
        if (n == 0)
        {
            *out = 0;
            return;
        }
        // calls from several host threads would share the array: one at a time
        std::lock_guard<std::mutex> lock(noAllocMutex());
        double * base = noAllocScratch();
        double * partials = base;
        double * result = base + 2 * kDotMaxBlocks;
        dotLaunch(dt, n, 1, x, n, y, n, partials, result, s);
        NSFEM_HIP(hipMemcpyAsync(out, result, sizeof(double), hipMemcpyDeviceToHost, s));
        NSFEM_HIP(hipStreamSynchronize(s));
    }
} // namespace nsfem
