// C++ host API over the C ABI (include/nsfem_hip.h), keeping the names,
// signatures, argument meaning and exception types of the reference's
// CudaUtils library so that solver code written against it builds
// unchanged:
//   cu::vec<T>           cpp/libs/CudaUtils/include/cu/vec.h:16-387
//   cu::csrF             cpp/libs/CudaUtils/include/cu/csrF.h:12-30
//   cu::spmv / cu::spmm  cpp/libs/CudaUtils/include/cu/spmv.h:13-46, spmm.h:14-47
//   cu::Blas, saxpy, scale     include/cu/blas.h:8-28
//   cu::Sparse                 include/cu/sparse.h
//   cu::AbstractSolver         include/cu/solvers/abstractSolver.h:8-32
//   cu::createSolver           include/cu/solvers/solverFactory.h:13-14
// Host matrices are the caller's own type: anything with the data members of
// the reference's linalg::CsrMatrix<F> (cpp/libs/LinAlg/include/linalg/
// csrMatrix.h:13-23: rows, cols, values, column, rowStart) binds, so the
// header sits next to the reference's linalg headers without redefining them.
// Handles (Blas, Sparse) carry no state: the library owns one stream.
#pragma once

#include <cmath>
#include <concepts>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../nsfem_hip.h"

namespace nsfem
{
    // A host CSR matrix: int32 rows, cols, 0-based column indices and
    // rowStart[rows + 1], contiguous values.
    template <typename M>
    concept CsrMatrixLike = requires(const M & m) {
        { m.rows } -> std::convertible_to<int>;
        { m.cols } -> std::convertible_to<int>;
        { m.values.data() };
        { m.values.size() } -> std::convertible_to<size_t>;
        { m.column.data() } -> std::convertible_to<const int *>;
        { m.rowStart.data() } -> std::convertible_to<const int *>;
    };

    template <CsrMatrixLike M>
    using CsrValueType = std::remove_cvref_t<decltype(*std::declval<const M &>().values.data())>;

    // The reference's binary CSR file (cpp/libs/LinAlg/src/io.cpp:49-84) as
    // read by include/nsfem/linalg_io.hpp; a CsrMatrixLike.
    template <typename F>
    struct HostCsr
    {
        int rows = 0, cols = 0;
        std::vector<F> values;
        std::vector<int> column;
        std::vector<int> rowStart; // rows + 1
    };
} // namespace nsfem

namespace cu
{
    inline void check(int rc)
    {
        if (rc == NSFEM_OK)
        {
            return;
        }
        const std::string msg = nsfem_last_error();
        if (rc == NSFEM_EINVAL)
        {
            throw std::invalid_argument(msg);
        }
        throw std::runtime_error(msg);
    }

    template <typename T>
    constexpr int dtypeOf();
    template <>
    constexpr int dtypeOf<float>() { return NSFEM_F32; }
    template <>
    constexpr int dtypeOf<double>() { return NSFEM_F64; }

    template <typename T>
    class vec
    {
        T * devicePtr = nullptr;
        size_t length = 0;
        bool externallyOwned = false;

    public:
        vec() = default;
        explicit vec(size_t n) : length(n)
        {
            if (n == 0)
            {
                throw std::invalid_argument("Size can't be zero");
            }
            void * p = nullptr;
            check(nsfem_malloc(n * sizeof(T), &p));
            devicePtr = static_cast<T *>(p);
        }
        // Wrapper for externally managed memory (never freed)
        vec(T * extData, size_t n) : devicePtr(extData), length(n), externallyOwned(true) {}
        explicit vec(const std::vector<T> & host) : vec(host.size()) { upload(host); }
        vec(const vec & o) : vec(o.length) { check(nsfem_copy(devicePtr, o.devicePtr, length * sizeof(T))); }
        vec(vec && o) noexcept : devicePtr(o.devicePtr), length(o.length), externallyOwned(o.externallyOwned)
        {
            o.devicePtr = nullptr;
            o.length = 0;
        }
        vec & operator=(vec && o) noexcept
        {
            if (this != &o)
            {
                reset();
                devicePtr = o.devicePtr;
                length = o.length;
                externallyOwned = o.externallyOwned;
                o.devicePtr = nullptr;
                o.length = 0;
                o.externallyOwned = false;
            }
            return *this;
        }
        ~vec() { reset(); }

        void reset() noexcept
        {
            if (devicePtr && !externallyOwned)
            {
                nsfem_free(devicePtr);
            }
            devicePtr = nullptr;
            length = 0;
            externallyOwned = false;
        }
        T * get() { return devicePtr; }
        const T * get() const { return devicePtr; }
        size_t size() const { return length; }
        explicit operator bool() const noexcept { return devicePtr != nullptr; }

        void copyTo(vec & dst) const
        {
            if (dst.size() != size())
            {
                throw std::invalid_argument("vec copy: size mismatch");
            }
            check(nsfem_copy(dst.get(), get(), size() * sizeof(T)));
        }
        void memsetZero()
        {
            if (!devicePtr)
            {
                throw std::runtime_error("memsetZero called on an empty vector");
            }
            check(nsfem_memset_zero(devicePtr, length * sizeof(T)));
        }
        void upload(const T * src) { check(nsfem_upload(devicePtr, src, length * sizeof(T))); }
        void download(T * dst) const { check(nsfem_download(dst, devicePtr, length * sizeof(T))); }
        void upload(const std::vector<T> & src)
        {
            if (src.size() != length)
            {
                throw std::invalid_argument("Size mismatch - got " + std::to_string(src.size()) + ", expected " + std::to_string(length));
            }
            upload(src.data());
        }
        void download(std::vector<T> & dst) const
        {
            if (dst.size() != length)
            {
                throw std::invalid_argument("Size mismatch - got " + std::to_string(dst.size()) + ", expected " + std::to_string(length));
            }
            download(dst.data());
        }
        void overwriteUpload(const std::vector<T> & src)
        {
            if (src.empty())
            {
                throw std::invalid_argument("Size can't be 0");
            }
            *this = vec(src.size());
            upload(src);
        }
    };

    // Stateless handles kept for source compatibility.
    struct Blas
    {
        void * handle = nullptr;
    };
    struct Sparse
    {
        void * handle() const { return nullptr; }
    };

    template <typename F>
    struct csrT
    {
        int rows = 0, cols = 0;
        nsfem_csr * h = nullptr;
        vec<F> values; // view of the device values, mutable in place

        template <nsfem::CsrMatrixLike M>
            requires std::same_as<nsfem::CsrValueType<M>, F>
        explicit csrT(const M & m) : rows(m.rows), cols(m.cols)
        {
            if (m.rowStart.size() != static_cast<size_t>(m.rows) + 1 || m.column.size() != m.values.size())
            {
                throw std::invalid_argument("CSR matrix: rowStart must have rows + 1 entries and column one per value");
            }
            check(nsfem_csr_create(m.rows, m.cols, static_cast<int>(m.values.size()), m.rowStart.data(), m.column.data(), m.values.data(),
                                   dtypeOf<F>(), &h));
            void * p = nullptr;
            check(nsfem_csr_values(h, &p));
            values = vec<F>(static_cast<F *>(p), m.values.size());
        }
        csrT(const csrT &) = delete;
        csrT & operator=(const csrT &) = delete;
        ~csrT() { nsfem_csr_destroy(h); }
    };
    using csrF = csrT<float>;
    using csrD = csrT<double>;

    template <typename F>
    struct spmvT
    {
        csrT<F> & mat;
        vec<F> x, b;
        spmvT(void *, csrT<F> & m) : mat(m), x(m.cols), b(m.rows) {}
        explicit spmvT(csrT<F> & m) : spmvT(nullptr, m) {}
        void compute() { compute(x, b); }
        void compute(vec<F> & otherX, vec<F> & otherB) { check(nsfem_spmv(mat.h, otherX.get(), otherB.get())); }
    };
    using spmv = spmvT<float>;

    template <typename F>
    struct spmmT
    {
        csrT<F> & mat;
        int numCh;
        vec<F> x, b;
        spmmT(void *, csrT<F> & m, int numCh) : mat(m), numCh(numCh), x(static_cast<size_t>(m.cols) * numCh), b(static_cast<size_t>(m.rows) * numCh)
        {
            if (numCh < 1)
            {
                throw std::invalid_argument("SPMM channels must be at least 1");
            }
        }
        void compute() { compute(x, b); }
        void compute(vec<F> & otherX, vec<F> & otherB) { check(nsfem_spmm(mat.h, numCh, otherX.get(), otherB.get())); }
    };
    using spmm = spmmT<float>;

    inline void saxpy(Blas &, const int n, float * src, float * dst, float alpha) { check(nsfem_axpy(NSFEM_F32, n, alpha, src, dst)); }
    inline void saxpy(Blas &, const int n, double * src, double * dst, double alpha) { check(nsfem_axpy(NSFEM_F64, n, alpha, src, dst)); }
    inline void scale(Blas &, const int n, float * dst, float alpha) { check(nsfem_scale(NSFEM_F32, n, alpha, dst)); }
    inline void scale(Blas &, const int n, double * dst, double alpha) { check(nsfem_scale(NSFEM_F64, n, alpha, dst)); }

    // cu::AbstractSolver with the io vector type of the solver's precision.
    template <typename T>
    class AbstractSolverT
    {
    public:
        virtual int getNumCh() const = 0;
        virtual float getLastMse(const int ch = -1) const = 0;
        virtual int getLastIterations() const = 0;
        virtual vec<T> & getRhs() = 0;
        virtual vec<T> & getSol() = 0;
        virtual void solve() = 0;
        virtual void setMaxIters(const int n) = 0;
        virtual void setTargetMse(const float mse) = 0;
        virtual void setMseCheckInterval(const int mseMod) = 0;
        virtual ~AbstractSolverT() {}
    };
    using AbstractSolver = AbstractSolverT<float>;

    template <typename T>
    class HipSolver final : public AbstractSolverT<T>
    {
        nsfem_solver * h = nullptr;
        int numCh;
        vec<T> rhs, sol;

    public:
        HipSolver(nsfem_solver * handle, int numCh, size_t n) : h(handle), numCh(numCh)
        {
            void * r = nullptr;
            void * s = nullptr;
            check(nsfem_solver_rhs(h, &r));
            check(nsfem_solver_sol(h, &s));
            rhs = vec<T>(static_cast<T *>(r), n * numCh);
            sol = vec<T>(static_cast<T *>(s), n * numCh);
        }
        ~HipSolver() override { nsfem_solver_destroy(h); }
        int getNumCh() const override { return numCh; }
        float getLastMse(const int ch = -1) const override
        {
            double m = -1;
            check(nsfem_solver_last_mse(h, ch, &m));
            return static_cast<float>(m);
        }
        int getLastIterations() const override
        {
            int it = -1;
            check(nsfem_solver_last_iterations(h, &it));
            return it;
        }
        vec<T> & getRhs() override { return rhs; }
        vec<T> & getSol() override { return sol; }
        void solve() override { check(nsfem_solver_solve(h)); }
        void setMaxIters(const int n) override { check(nsfem_solver_set_max_iters(h, n)); }
        void setTargetMse(const float mse) override { check(nsfem_solver_set_target_mse(h, mse)); }
        void setMseCheckInterval(const int k) override { check(nsfem_solver_set_mse_check_interval(h, k)); }
    };

    // name: "gs", "pcg" ("dss" is rejected, as is an unknown name) for FP32
    // matrices (the reference's), "gs", "pcg", "pcg_mixed" for FP64 ones; the
    // solver's io vectors have the matrix's value type.
    template <nsfem::CsrMatrixLike M>
        requires std::same_as<nsfem::CsrValueType<M>, float> || std::same_as<nsfem::CsrValueType<M>, double>
    std::unique_ptr<AbstractSolverT<nsfem::CsrValueType<M>>> createSolver(const std::string & name, const int numCh, const M & m,
                                                                          const int maxIters, const nsfem::CsrValueType<M> targetMse,
                                                                          const int mseCheckInterval)
    {
        using F = nsfem::CsrValueType<M>;
        if (m.rowStart.size() != static_cast<size_t>(m.rows) + 1 || m.column.size() != m.values.size())
        {
            throw std::invalid_argument("CSR matrix: rowStart must have rows + 1 entries and column one per value");
        }
        nsfem_solver * h = nullptr;
        check(nsfem_solver_create(name.c_str(), numCh, m.rows, static_cast<int>(m.values.size()), m.rowStart.data(), m.column.data(),
                                  m.values.data(), dtypeOf<F>(), maxIters, targetMse, mseCheckInterval, &h));
        return std::make_unique<HipSolver<F>>(h, numCh, static_cast<size_t>(m.rows));
    }
} // namespace cu
