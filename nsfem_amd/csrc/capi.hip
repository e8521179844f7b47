// extern "C" entry points: exceptions -> status codes + thread-local message.
#include "../../include/nsfem_hip.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "chorin.hpp"
#include "common.hpp"
#include "context.hpp"
#include "csr.hpp"
#include "fem.hpp"
#include "mesh.hpp"
#include "ops.hpp"
#include "solver.hpp"

using namespace nsfem;

namespace
{
    thread_local std::string gLastError;

    template <typename F>
    int guarded(F && f)
    {
        try
        {
            f();
            gLastError.clear();
            return NSFEM_OK;
        }
        catch (const std::invalid_argument & e)
        {
            gLastError = e.what();
            return NSFEM_EINVAL;
        }
        catch (const NonFiniteError & e)
        {
            gLastError = e.what();
            return NSFEM_ENONFINITE;
        }
        catch (const DeviceError & e)
        {
            gLastError = e.what();
            return NSFEM_EDEVICE;
        }
        catch (const std::exception & e)
        {
            gLastError = e.what();
            return NSFEM_ERUNTIME;
        }
        catch (...)
        {
            gLastError = "unknown error";
            return NSFEM_ERUNTIME;
        }
    }

    void need(const void * p, const char * what)
    {
        if (!p)
        {
            throw std::invalid_argument(std::string(what) + " is NULL");
        }
    }

    void applySolverOption(McgsSolver & solver, const std::string & k, double value)
    {
        if (k == "innerTol")
        {
            if (!(value > 0 && value < 1)) throw std::invalid_argument("innerTol must be in (0, 1)");
            solver.setInnerTol(value);
        }
        else if (k == "projectMean")
        {
            solver.setProjectMean(value != 0);
        }
        else if (k == "blockPivot")
        {
            if (!(value > 0 && value < 1)) throw std::invalid_argument("blockPivot must be in (0, 1)");
            solver.setBlockPivot(value);
        }
        else if (k == "mixedRestart")
        {
            solver.setMixedRestart(value != 0);
        }
        else if (k == "graphs")
        {
            solver.setUseGraphs(value != 0);
        }
        else if (k == "fusedReductions")
        {
            if (value != -1 && value != 0 && value != 1) throw std::invalid_argument("fusedReductions must be -1, 0 or 1");
            solver.setFusedReductions(static_cast<int>(value));
        }
        else if (k == "fuseUpdateSweep")
        {
            if (value != 0 && value != 1) throw std::invalid_argument("fuseUpdateSweep must be 0 or 1");
            solver.setFusedUpdateSweep(static_cast<int>(value));
        }
        else if (k == "regionDirection")
        {
            if (value != 0 && value != 1) throw std::invalid_argument("regionDirection must be 0 or 1");
            solver.setRegionedDirection(static_cast<int>(value));
        }
        else if (k == "lanesDirection")
        {
            if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8 && value != 16)
                throw std::invalid_argument("lanesDirection must be 0, 1, 2, 4, 8 or 16");
            solver.setDirectionLanes(static_cast<int>(value));
        }
        else if (k == "preconditioner")
        {
            if (value != 0 && value != 1) throw std::invalid_argument("preconditioner must be 0 or 1");
            solver.setPreconditioner(static_cast<int>(value));
        }
        else if (k == "amgChebyshevFrom" || k == "amgChebyshevDegree" || k == "amgChebyshevFraction")
        {
            AmgPcg * amg = solver.amg();
            if (!amg)
            {
                throw std::invalid_argument(k + " applies to the amg_pcg solvers only");
            }
            if (k == "amgChebyshevFrom")
            {
                if (!(value >= 0 && value <= 1 << 20)) throw std::invalid_argument("amgChebyshevFrom must be in [0, 2^20]");
                amg->setSmoother(static_cast<int>(value), amg->chebyshevDegree(), amg->chebyshevFraction());
            }
            else if (k == "amgChebyshevDegree")
            {
                if (value != static_cast<int>(value)) throw std::invalid_argument("amgChebyshevDegree must be an integer");
                amg->setSmoother(amg->chebyshevFrom(), static_cast<int>(value), amg->chebyshevFraction());
            }
            else
            {
                amg->setSmoother(amg->chebyshevFrom(), amg->chebyshevDegree(), value);
            }
        }
        else if (k == "rowsPreconditioner")
        {
            if (!(value >= 0 && value <= 1 << 20) || value != static_cast<int>(value))
                throw std::invalid_argument("rowsPreconditioner must be an integer in [0, 2^20]");
            solver.setPreconditionerRows(static_cast<int>(value));
        }
        else if (k == "rowsDirection" || k == "rowsUpdate" || k == "rowsRho")
        {
            if (!(value >= 0 && value <= 1 << 20)) throw std::invalid_argument(k + " must be in [0, 2^20]");
            const int v = static_cast<int>(value);
            solver.setKindRows(k == "rowsDirection" ? v : solver.kindRows(0), k == "rowsUpdate" ? v : solver.kindRows(1),
                               k == "rowsRho" ? v : solver.kindRows(2));
        }
        else
        {
            throw std::invalid_argument("Unknown solver option: " + k);
        }
    }

    CsrD hostCsr(int rows, int cols, int nnz, const int * rowStart, const int * column, const void * values, int dtype)
    {
        checkDtype(dtype);
        if (rows < 1 || cols < 1)
        {
            throw std::invalid_argument("CSR: rows and cols must be positive");
        }
        if (nnz < 0)
        {
            throw std::invalid_argument("CSR: negative nnz");
        }
        need(rowStart, "rowStart");
        if (nnz > 0)
        {
            need(column, "column");
            need(values, "values");
        }
        CsrD m;
        m.rows = rows;
        m.cols = cols;
        m.rowStart.assign(rowStart, rowStart + rows + 1);
        if (m.rowStart[0] != 0 || m.rowStart[rows] != nnz)
        {
            throw std::invalid_argument("CSR: rowStart must start at 0 and end at nnz");
        }
        for (int r = 0; r < rows; r++)
        {
            if (m.rowStart[r + 1] < m.rowStart[r])
            {
                throw std::invalid_argument("CSR: rowStart is not monotone");
            }
        }
        m.column.assign(column, column + nnz);
        for (int c : m.column)
        {
            if (c < 0 || c >= cols)
            {
                throw std::invalid_argument("CSR: column index out of range");
            }
        }
        m.values.resize(nnz);
        for (int j = 0; j < nnz; j++)
        {
            m.values[j] = dtype == F64 ? static_cast<const double *>(values)[j] : static_cast<const float *>(values)[j];
        }
        return m;
    }
} // namespace

struct nsfem_csr
{
    int rows, cols, nnz, dtype, lanes;
    DeviceArray<int> rowStart, column;
    DeviceArray<char> values;
    CsrView view() const { return CsrView{rows, cols, nnz, rowStart.get(), column.get(), values.get(), dtype}; }
};

struct nsfem_solver
{
    std::unique_ptr<McgsSolver> core;
};

struct nsfem_mesh
{
    TriangleMesh tri;
};

struct nsfem_context
{
    ConcreteMesh velocityMesh, pressureMesh;
    ChorinContext ctx;
};

struct nsfem_chorin
{
    std::unique_ptr<ChorinGpu> core;
};

extern "C"
{
    const char * nsfem_last_error(void) { return gLastError.c_str(); }

    int nsfem_version(int * major, int * minor)
    {
        return guarded([&]
                       {
            need(major, "major");
            need(minor, "minor");
            *major = 1;
            *minor = 0; });
    }

    int nsfem_synchronize(void)
    {
        return guarded([&]
                       {
            syncLibraryStream();
            NSFEM_HIP(hipDeviceSynchronize()); });
    }

    int nsfem_malloc(size_t bytes, void ** dptr)
    {
        return guarded([&]
                       {
            need(dptr, "dptr");
            if (bytes == 0) throw std::invalid_argument("Size can't be zero");
            NSFEM_HIP(hipMalloc(dptr, bytes)); });
    }

    int nsfem_free(void * dptr)
    {
        return guarded([&]
                       {
            if (dptr)
            {
                syncLibraryStream();
                NSFEM_HIP(hipFree(dptr));
            } });
    }

    int nsfem_upload(void * dst, const void * src, size_t bytes)
    {
        return guarded([&]
                       {
            if (!bytes) return;
            need(dst, "dst");
            need(src, "src");
            NSFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, libraryStream()));
            syncLibraryStream(); });
    }

    int nsfem_download(void * dst, const void * src, size_t bytes)
    {
        return guarded([&]
                       {
            if (!bytes) return;
            need(dst, "dst");
            need(src, "src");
            NSFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, libraryStream()));
            syncLibraryStream(); });
    }

    int nsfem_copy(void * dst, const void * src, size_t bytes)
    {
        return guarded([&]
                       {
            if (!bytes) return;
            need(dst, "dst");
            need(src, "src");
            NSFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, libraryStream()));
            syncLibraryStream(); });
    }

    int nsfem_memset_zero(void * dptr, size_t bytes)
    {
        return guarded([&]
                       {
            need(dptr, "dptr");
            NSFEM_HIP(hipMemsetAsync(dptr, 0, bytes, libraryStream()));
            syncLibraryStream(); });
    }

    int nsfem_csr_create(int rows, int cols, int nnz, const int * rowStart, const int * column, const void * values, int dtype,
                         nsfem_csr ** out)
    {
        return guarded([&]
                       {
            need(out, "out");
            CsrD h = hostCsr(rows, cols, nnz, rowStart, column, values, dtype);
            auto m = std::make_unique<nsfem_csr>();
            m->rows = rows;
            m->cols = cols;
            m->nnz = nnz;
            m->dtype = dtype;
            m->lanes = lanesForMeanRow(static_cast<double>(nnz) / rows);
            m->rowStart.upload(h.rowStart);
            m->column.upload(nnz ? h.column : std::vector<int>{0});
            const size_t vb = std::max<size_t>(1, nnz) * dtypeSize(dtype);
            m->values.resize(vb);
            if (nnz)
            {
                NSFEM_HIP(hipMemcpyAsync(m->values.get(), values, nnz * dtypeSize(dtype), hipMemcpyHostToDevice, libraryStream()));
            }
            syncLibraryStream();
            *out = m.release(); });
    }

    int nsfem_csr_destroy(nsfem_csr * m)
    {
        return guarded([&]
                       {
            syncLibraryStream();
            delete m; });
    }

    int nsfem_csr_values(nsfem_csr * m, void ** dvalues)
    {
        return guarded([&]
                       {
            need(m, "matrix");
            need(dvalues, "dvalues");
            *dvalues = m->values.get(); });
    }

    int nsfem_csr_shape(const nsfem_csr * m, int * rows, int * cols, int * nnz, int * dtype)
    {
        return guarded([&]
                       {
            need(m, "matrix");
            if (rows) *rows = m->rows;
            if (cols) *cols = m->cols;
            if (nnz) *nnz = m->nnz;
            if (dtype) *dtype = m->dtype; });
    }

    int nsfem_csr_mul(nsfem_csr * m, int numCh, const void * x, long ldx, void * y, long ldy, double alpha, double beta, int lanes)
    {
        return guarded([&]
                       {
            need(m, "matrix");
            need(x, "x");
            need(y, "y");
            csrMul(m->view(), numCh, x, ldx, y, ldy, alpha, beta, nullptr, m->rows, lanes > 0 ? lanes : m->lanes, libraryStream());
            syncLibraryStream(); });
    }

    int nsfem_spmv(nsfem_csr * m, const void * x, void * y)
    {
        if (!m)
        {
            gLastError = "matrix is NULL";
            return NSFEM_EINVAL;
        }
        return nsfem_csr_mul(m, 1, x, m->cols, y, m->rows, 1.0, 0.0, 0);
    }

    int nsfem_spmm(nsfem_csr * m, int numCh, const void * x, void * y)
    {
        if (!m)
        {
            gLastError = "matrix is NULL";
            return NSFEM_EINVAL;
        }
        return nsfem_csr_mul(m, numCh, x, m->cols, y, m->rows, 1.0, 0.0, 0);
    }

    int nsfem_axpy(int dtype, long n, double alpha, const void * src, void * dst)
    {
        return guarded([&]
                       {
            axpy(dtype, n, alpha, src, dst, libraryStream());
            syncLibraryStream(); });
    }

    int nsfem_scale(int dtype, long n, double alpha, void * dst)
    {
        return guarded([&]
                       {
            scale(dtype, n, alpha, dst, libraryStream());
            syncLibraryStream(); });
    }

    int nsfem_copy_bandwidth(size_t bytes, int reps, double * bytesPerSecond)
    {
        return guarded([&]
                       {
            need(bytesPerSecond, "bytesPerSecond");
            *bytesPerSecond = copyBandwidth(bytes, reps, libraryStream());
            syncLibraryStream(); });
    }

    int nsfem_geam(int dtype, long n, double alpha, const void * a, double beta, const void * b, void * c)
    {
        return guarded([&]
                       {
            geam(dtype, n, alpha, a, beta, b, c, libraryStream());
            syncLibraryStream(); });
    }

    int nsfem_dot(int dtype, long n, const void * x, const void * y, double * out)
    {
        return guarded([&]
                       {
            need(out, "out");
            dot(dtype, n, 1, x, n, y, n, out, libraryStream()); });
    }

    int nsfem_dot_noalloc(int dtype, long n, const void * x, const void * y, double * out)
    {
        return guarded([&]
                       {
            need(out, "out");
            if (n > 0)
            {
                need(x, "x");
                need(y, "y");
            }
            dotNoAlloc(dtype, n, x, y, out, libraryStream()); });
    }

    int nsfem_nrm2(int dtype, long n, const void * x, double * out)
    {
        return guarded([&]
                       {
            need(out, "out");
            double d = 0;
            dot(dtype, n, 1, x, n, x, n, &d, libraryStream());
            *out = std::sqrt(d); });
    }

    int nsfem_scatter(int dtype, int nnz, const int * idx, const void * values, void * dense)
    {
        return guarded([&]
                       {
            scatter(dtype, nnz, idx, values, dense, libraryStream());
            syncLibraryStream(); });
    }

    int nsfem_gather(int dtype, int nnz, const void * dense, const int * idx, void * values)
    {
        return guarded([&]
                       {
            gather(dtype, nnz, dense, idx, values, libraryStream());
            syncLibraryStream(); });
    }

    int nsfem_convert(int srcDtype, int dstDtype, long n, const void * src, void * dst)
    {
        return guarded([&]
                       {
            convert(srcDtype, dstDtype, n, src, dst, libraryStream());
            syncLibraryStream(); });
    }

    static void applyName(SolverOptions & o, const char * name, int numCh, int dtype, int maxIters, double target, int interval)
    {
        need(name, "name");
        checkDtype(dtype);
        if (numCh < 1)
        {
            throw std::invalid_argument("Invalid number of channels for solver: " + std::to_string(numCh));
        }
        if (interval < 1)
        {
            throw std::invalid_argument("Invalid MSE check interval for solver: " + std::to_string(interval));
        }
        o = parseSolverName(name);
        if (o.precision == SolverPrecision::Mixed)
        {
            if (dtype != F64)
            {
                throw std::invalid_argument("pcg_mixed needs FP64 values");
            }
        }
        else
        {
            o.precision = dtype == F64 ? SolverPrecision::F64 : SolverPrecision::F32;
        }
        o.numCh = numCh;
        o.maxIters = maxIters;
        o.target = target;
        o.mseCheckInterval = interval;
    }

    int nsfem_solver_create(const char * name, int numCh, int rows, int nnz, const int * rowStart, const int * column, const void * values,
                            int valueDtype, int maxIters, double targetMse, int mseCheckInterval, nsfem_solver ** out)
    {
        return guarded([&]
                       {
            need(out, "out");
            SolverOptions o;
            applyName(o, name, numCh, valueDtype, maxIters, targetMse, mseCheckInterval);
            CsrD m = hostCsr(rows, rows, nnz, rowStart, column, values, valueDtype);
            auto s = std::make_unique<nsfem_solver>();
            s->core = std::make_unique<McgsSolver>(std::vector<const CsrD *>{&m}, o);
            *out = s.release(); });
    }

    int nsfem_solver_create_batch(const char * name, int numCh, int numSystems, const int * rows, const int * nnz,
                                  const int * const * rowStart, const int * const * column, const double * const * values, int dtype,
                                  int maxIters, double targetMse, int mseCheckInterval, nsfem_solver ** out)
    {
        return guarded([&]
                       {
            need(out, "out");
            if (numSystems < 1) throw std::invalid_argument("numSystems must be positive");
            need(rows, "rows");
            need(nnz, "nnz");
            need(rowStart, "rowStart");
            need(column, "column");
            need(values, "values");
            SolverOptions o;
            applyName(o, name, numCh, dtype, maxIters, targetMse, mseCheckInterval);
            std::vector<CsrD> mats;
            mats.reserve(numSystems);
            for (int s = 0; s < numSystems; s++)
            {
                mats.push_back(hostCsr(rows[s], rows[s], nnz[s], rowStart[s], column[s], values[s], F64));
            }
            std::vector<const CsrD *> ptrs;
            for (auto & m : mats) ptrs.push_back(&m);
            auto h = std::make_unique<nsfem_solver>();
            h->core = std::make_unique<McgsSolver>(ptrs, o);
            *out = h.release(); });
    }

    int nsfem_solver_destroy(nsfem_solver * s)
    {
        return guarded([&]
                       {
            syncLibraryStream();
            delete s; });
    }

    int nsfem_solver_num_ch(const nsfem_solver * s, int * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            *out = s->core->numCh(); });
    }

    int nsfem_solver_rows(const nsfem_solver * s, int * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            *out = s->core->n(); });
    }

    int nsfem_solver_io_dtype(const nsfem_solver * s, int * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            *out = s->core->ioDtype(); });
    }

    int nsfem_solver_rhs(nsfem_solver * s, void ** dptr)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(dptr, "dptr");
            *dptr = s->core->ioRhs(); });
    }

    int nsfem_solver_sol(nsfem_solver * s, void ** dptr)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(dptr, "dptr");
            *dptr = s->core->ioSol(); });
    }

    int nsfem_solver_solve(nsfem_solver * s)
    {
        return guarded([&]
                       {
            need(s, "solver");
            s->core->solve(); });
    }

    int nsfem_solver_last_mse(const nsfem_solver * s, int ch, double * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            const int nch = s->core->numCh();
            if (ch < 0)
            {
                double sum = 0;
                for (int c = 0; c < nch; c++)
                {
                    const double m = s->core->lastMse(0, c);
                    sum += m * m;
                }
                *out = nch == 1 ? s->core->lastMse(0, 0) : std::sqrt(sum);
            }
            else if (ch < nch)
            {
                *out = s->core->lastMse(0, ch);
            }
            else
            {
                *out = -1;
            } });
    }

    int nsfem_solver_last_relres(const nsfem_solver * s, int system, int ch, double * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            if (system < 0 || system >= s->core->numSystems() || ch < 0 || ch >= s->core->numCh())
                throw std::invalid_argument("system or channel out of range");
            *out = s->core->lastRelRes(system, ch); });
    }

    int nsfem_solver_last_iterations(const nsfem_solver * s, int * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            int it = 0;
            for (int k = 0; k < s->core->numSystems(); k++) it = std::max(it, s->core->lastIterations(k));
            *out = it; });
    }

    int nsfem_solver_last_iterations_system(const nsfem_solver * s, int system, int * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            if (system < 0 || system >= s->core->numSystems()) throw std::invalid_argument("system out of range");
            *out = s->core->lastIterations(system); });
    }

    int nsfem_solver_last_outer_iterations(const nsfem_solver * s, int * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            *out = s->core->lastOuterIterations(); });
    }

    int nsfem_solver_set_max_iters(nsfem_solver * s, int n)
    {
        return guarded([&]
                       {
            need(s, "solver");
            if (n < 0) throw std::invalid_argument("max iterations must be >= 0");
            s->core->setMaxIters(n); });
    }

    int nsfem_solver_set_target_mse(nsfem_solver * s, double t)
    {
        return guarded([&]
                       {
            need(s, "solver");
            s->core->setTarget(t); });
    }

    int nsfem_solver_set_mse_check_interval(nsfem_solver * s, int k)
    {
        return guarded([&]
                       {
            need(s, "solver");
            s->core->setMseCheckInterval(k); });
    }

    int nsfem_solver_set_option(nsfem_solver * s, const char * key, double value)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(key, "key");
            applySolverOption(*s->core, key, value); });
    }

    int nsfem_solver_time_kernel(nsfem_solver * s, const char * which, int reps, double * avgMs, double * bytes,
                                 double * launchesPerIteration)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(which, "which");
            need(avgMs, "avgMs");
            *avgMs = s->core->timeKernel(which, reps, bytes, launchesPerIteration); });
    }

    int nsfem_solver_reshape(nsfem_solver * s, int lanes, int rowsPerBlock)
    {
        return guarded([&]
                       {
            need(s, "solver");
            s->core->reshape(lanes, rowsPerBlock); });
    }

    int nsfem_solver_shape(const nsfem_solver * s, int * lanes, int * rowsPerBlock)
    {
        return guarded([&]
                       {
            need(s, "solver");
            if (lanes) *lanes = s->core->lanes();
            if (rowsPerBlock) *rowsPerBlock = s->core->rowsPerBlock(); });
    }

    int nsfem_solver_reshape_sweeps(nsfem_solver * s, int lanes, int rowsPerBlock)
    {
        return guarded([&]
                       {
            need(s, "solver");
            s->core->reshapeSweeps(lanes, rowsPerBlock); });
    }

    int nsfem_solver_sweep_shape(const nsfem_solver * s, int * lanes, int * rowsPerBlock)
    {
        return guarded([&]
                       {
            need(s, "solver");
            if (lanes) *lanes = s->core->sweepLanes();
            if (rowsPerBlock) *rowsPerBlock = s->core->sweepRowsPerBlock(); });
    }

    int nsfem_solver_kind_rows(const nsfem_solver * s, int * direction, int * update, int * rho)
    {
        return guarded([&]
                       {
            need(s, "solver");
            if (direction) *direction = s->core->kindRows(0);
            if (update) *update = s->core->kindRows(1);
            if (rho) *rho = s->core->kindRows(2); });
    }

    int nsfem_solver_autotune(nsfem_solver * s, int reps, int * lanes, int * rowsPerBlock)
    {
        return guarded([&]
                       {
            need(s, "solver");
            if (reps < 1) throw std::invalid_argument("reps must be positive");
            const auto shape = s->core->autotune(reps);
            if (lanes) *lanes = shape.first;
            if (rowsPerBlock) *rowsPerBlock = shape.second; });
    }

    int nsfem_solver_num_colors(const nsfem_solver * s, int system, int * out)
    {
        return guarded([&]
                       {
            need(s, "solver");
            need(out, "out");
            if (system < 0 || system >= s->core->numSystems()) throw std::invalid_argument("system out of range");
            *out = static_cast<int>(s->core->partitionStart(system).size()) - 1; });
    }

    int nsfem_solver_coloring(const nsfem_solver * s, int system, int * coloring, int * partitionStart)
    {
        return guarded([&]
                       {
            need(s, "solver");
            if (system < 0 || system >= s->core->numSystems()) throw std::invalid_argument("system out of range");
            const auto & c = s->core->coloring(system);
            const auto & p = s->core->partitionStart(system);
            if (coloring) std::copy(c.begin(), c.end(), coloring);
            if (partitionStart) std::copy(p.begin(), p.end(), partitionStart); });
    }

    int nsfem_mcgs_coloring(int rows, const int * rowStart, const int * column, int * coloring, int * partitionStart, int * numColors)
    {
        return guarded([&]
                       {
            if (rows < 1) throw std::invalid_argument("rows must be positive");
            need(rowStart, "rowStart");
            need(coloring, "coloring");
            need(partitionStart, "partitionStart");
            need(numColors, "numColors");
            std::vector<int> rs(rowStart, rowStart + rows + 1);
            std::vector<int> col(column, column + rs[rows]);
            auto graph = buildCsrGraph(rs, col, rows);
            auto parts = partitionGraphGreedy(graph, smallestLastOrdering(graph));
            int pos = 0;
            for (size_t c = 0; c < parts.size(); c++)
            {
                std::sort(parts[c].begin(), parts[c].end());
                partitionStart[c] = pos;
                std::copy(parts[c].begin(), parts[c].end(), coloring + pos);
                pos += static_cast<int>(parts[c].size());
            }
            partitionStart[parts.size()] = rows;
            *numColors = static_cast<int>(parts.size()); });
    }

    int nsfem_mesh_load_gmsh(const char * path, nsfem_mesh ** out)
    {
        return guarded([&]
                       {
            need(path, "path");
            need(out, "out");
            auto m = std::make_unique<nsfem_mesh>();
            m->tri = parseTriangleGmsh(path);
            *out = m.release(); });
    }

    int nsfem_mesh_from_arrays(int numNodes, const double * xy, int numTriangles, const int * triangles, int numBorderEdges,
                               const int * borderEdges, int numGroups, const char * const * groupNames, nsfem_mesh ** out)
    {
        return guarded([&]
                       {
            need(out, "out");
            if (numNodes < 3 || numTriangles < 1) throw std::invalid_argument("mesh needs nodes and triangles");
            need(xy, "xy");
            need(triangles, "triangles");
            std::vector<Point> nodes(numNodes);
            for (int i = 0; i < numNodes; i++) nodes[i] = Point{xy[2 * i], xy[2 * i + 1]};
            std::vector<std::array<int, 3>> tris(numTriangles);
            for (int i = 0; i < numTriangles; i++) tris[i] = {triangles[3 * i], triangles[3 * i + 1], triangles[3 * i + 2]};
            std::vector<std::array<int, 3>> edges(numBorderEdges);
            for (int i = 0; i < numBorderEdges; i++) edges[i] = {borderEdges[3 * i], borderEdges[3 * i + 1], borderEdges[3 * i + 2]};
            std::vector<std::string> groups;
            for (int g = 0; g < numGroups; g++)
            {
                need(groupNames[g], "group name");
                groups.emplace_back(groupNames[g]);
            }
            auto m = std::make_unique<nsfem_mesh>();
            m->tri = buildTriangleMesh(nodes, tris, edges, groups);
            *out = m.release(); });
    }

    int nsfem_mesh_destroy(nsfem_mesh * m)
    {
        return guarded([&]
                       { delete m; });
    }

    int nsfem_mesh_info(const nsfem_mesh * m, int * numNodes, int * numTriangles, int * numBorderElements, int * numGroups)
    {
        return guarded([&]
                       {
            need(m, "mesh");
            if (numNodes) *numNodes = static_cast<int>(m->tri.nodes.size());
            if (numTriangles) *numTriangles = static_cast<int>(m->tri.elements.size());
            if (numBorderElements) *numBorderElements = static_cast<int>(m->tri.borderElements.size());
            if (numGroups) *numGroups = static_cast<int>(m->tri.groups.size()); });
    }

    int nsfem_mesh_arrays(const nsfem_mesh * m, double * xy, int * triangles, int * border)
    {
        return guarded([&]
                       {
            need(m, "mesh");
            const auto & t = m->tri;
            if (xy)
                for (size_t i = 0; i < t.nodes.size(); i++) { xy[2 * i] = t.nodes[i].x; xy[2 * i + 1] = t.nodes[i].y; }
            if (triangles)
                for (size_t i = 0; i < t.elements.size(); i++)
                    for (int k = 0; k < 3; k++) triangles[3 * i + k] = t.elements[i][k];
            if (border)
                for (size_t i = 0; i < t.borderElements.size(); i++)
                {
                    border[3 * i] = t.borderElements[i].element;
                    border[3 * i + 1] = t.borderElements[i].side;
                    border[3 * i + 2] = t.borderElements[i].group;
                } });
    }

    int nsfem_mesh_group_name(const nsfem_mesh * m, int group, char * buf, int bufSize)
    {
        return guarded([&]
                       {
            need(m, "mesh");
            need(buf, "buf");
            if (group < 0 || group >= static_cast<int>(m->tri.groups.size())) throw std::invalid_argument("group out of range");
            const std::string & n = m->tri.groups[group];
            if (bufSize < static_cast<int>(n.size()) + 1) throw std::invalid_argument("buffer too small");
            std::memcpy(buf, n.c_str(), n.size() + 1); });
    }

    int nsfem_context_build(const nsfem_mesh * mesh, const nsfem_dfg_conditions * cond, int numThreads, nsfem_context ** out)
    {
        return nsfem_context_build_ex(mesh, cond, numThreads, 0, out);
    }

    int nsfem_context_build_ex(const nsfem_mesh * mesh, const nsfem_dfg_conditions * cond, int numThreads, int flags,
                               nsfem_context ** out)
    {
        return guarded([&]
                       {
            if (flags & ~NSFEM_CONTEXT_GPU_ASSEMBLY) throw std::invalid_argument("unknown context flags");
            need(mesh, "mesh");
            need(cond, "cond");
            need(out, "out");
            if (!(cond->channelHeight > 0)) throw std::invalid_argument("channelHeight must be positive");
            DfgConditions c;
            c.viscosity = cond->viscosity;
            c.peakVelocity = cond->peakVelocity;
            c.channelHeight = cond->channelHeight;
            c.profileScale = cond->profileScale;
            auto h = std::make_unique<nsfem_context>();
            h->velocityMesh = createMesh(mesh->tri, ElementType::P2);
            h->pressureMesh = createMesh(mesh->tri, ElementType::P1);
            h->ctx = buildChorinContext(h->velocityMesh, h->pressureMesh, c, numThreads > 0 ? numThreads : defaultThreads(),
                                        (flags & NSFEM_CONTEXT_GPU_ASSEMBLY) != 0);
            *out = h.release(); });
    }

    int nsfem_context_build_batch(const nsfem_mesh * const * meshes, const nsfem_dfg_conditions * conds, int count, int numThreads,
                                  int flags, nsfem_context ** out)
    {
        return guarded([&]
                       {
            if (flags & ~NSFEM_CONTEXT_GPU_ASSEMBLY) throw std::invalid_argument("unknown context flags");
            need(meshes, "meshes");
            need(conds, "conds");
            need(out, "out");
            if (count < 1) throw std::invalid_argument("count must be positive");
            std::vector<ChorinContext> parts;
            parts.reserve(count);
            const int nt = numThreads > 0 ? numThreads : defaultThreads();
            for (int k = 0; k < count; k++)
            {
                need(meshes[k], "mesh");
                if (!(conds[k].channelHeight > 0)) throw std::invalid_argument("channelHeight must be positive");
                DfgConditions c{conds[k].viscosity, conds[k].peakVelocity, conds[k].channelHeight, conds[k].profileScale};
                ConcreteMesh vm = createMesh(meshes[k]->tri, ElementType::P2);
                ConcreteMesh pm = createMesh(meshes[k]->tri, ElementType::P1);
                parts.push_back(buildChorinContext(vm, pm, c, nt, (flags & NSFEM_CONTEXT_GPU_ASSEMBLY) != 0));
            }
            auto h = std::make_unique<nsfem_context>();
            h->ctx = concatenateContexts(parts);
            h->velocityMesh.numElements = h->ctx.convection.numElements;
            *out = h.release(); });
    }

    int nsfem_context_num_systems(const nsfem_context * c, int * out)
    {
        return guarded([&]
                       {
            need(c, "context");
            need(out, "out");
            *out = c->ctx.numSystems(); });
    }

    int nsfem_context_system_offsets(const nsfem_context * c, int * velocityOffsets, int * pressureOffsets)
    {
        return guarded([&]
                       {
            need(c, "context");
            const int S = c->ctx.numSystems();
            for (int k = 0; k < S; k++)
            {
                if (velocityOffsets) velocityOffsets[k] = c->ctx.velocityOffsets.empty() ? 0 : c->ctx.velocityOffsets[k];
                if (pressureOffsets) pressureOffsets[k] = c->ctx.pressureOffsets.empty() ? 0 : c->ctx.pressureOffsets[k];
            } });
    }

    int nsfem_context_destroy(nsfem_context * c)
    {
        return guarded([&]
                       { delete c; });
    }

    int nsfem_context_sizes(const nsfem_context * c, int * N, int * V, int * nDir, int * nIntP, int * E)
    {
        return guarded([&]
                       {
            need(c, "context");
            if (N) *N = c->ctx.numVelocityNodes;
            if (V) *V = c->ctx.numPressureNodes;
            if (nDir) *nDir = static_cast<int>(c->ctx.dirichletVx.size());
            if (nIntP) *nIntP = static_cast<int>(c->ctx.internalPressureNodes.size());
            if (E) *E = c->velocityMesh.numElements; });
    }

    int nsfem_context_matrix(const nsfem_context * c, int which, int * rows, int * cols, int * nnz, const int ** rowStart,
                             const int ** column, const double ** values)
    {
        return guarded([&]
                       {
            need(c, "context");
            const CsrD * m = nullptr;
            switch (which)
            {
            case NSFEM_MAT_VELOCITY_MASS: m = &c->ctx.velocityMass; break;
            case NSFEM_MAT_VELOCITY_STIFFNESS: m = &c->ctx.velocityStiffness; break;
            case NSFEM_MAT_PRESSURE_STIFFNESS: m = &c->ctx.pressureStiffness; break;
            case NSFEM_MAT_PRESSURE_STIFFNESS_INTERNAL: m = &c->ctx.pressureStiffnessInternal; break;
            case NSFEM_MAT_VELOCITY_PRESSURE_DIV: m = &c->ctx.velocityPressureDiv; break;
            case NSFEM_MAT_PRESSURE_VELOCITY_DIV: m = &c->ctx.pressureVelocityDiv; break;
            default: throw std::invalid_argument("unknown matrix id");
            }
            if (rows) *rows = m->rows;
            if (cols) *cols = m->cols;
            if (nnz) *nnz = m->nnz();
            if (rowStart) *rowStart = m->rowStart.data();
            if (column) *column = m->column.data();
            if (values) *values = m->values.data(); });
    }

    int nsfem_context_dirichlet(const nsfem_context * c, int which, int * count, int * ids, double * values)
    {
        return guarded([&]
                       {
            need(c, "context");
            const std::vector<DirichletNode> * d = nullptr;
            if (which == 0) d = &c->ctx.dirichletVx;
            else if (which == 1) d = &c->ctx.dirichletVy;
            else if (which == 2) d = &c->ctx.dirichletPressure;
            else throw std::invalid_argument("which must be 0, 1 or 2");
            if (count) *count = static_cast<int>(d->size());
            for (size_t i = 0; i < d->size(); i++)
            {
                if (ids) ids[i] = (*d)[i].id;
                if (values) values[i] = (*d)[i].value;
            } });
    }

    int nsfem_context_internal_pressure(const nsfem_context * c, int * ids)
    {
        return guarded([&]
                       {
            need(c, "context");
            need(ids, "ids");
            std::copy(c->ctx.internalPressureNodes.begin(), c->ctx.internalPressureNodes.end(), ids); });
    }

    static void meshOut(const ConcreteMesh & m, double * xy, int * elements)
    {
        if (xy)
        {
            for (size_t i = 0; i < m.nodes.size(); i++)
            {
                xy[2 * i] = m.nodes[i].x;
                xy[2 * i + 1] = m.nodes[i].y;
            }
        }
        if (elements)
        {
            std::copy(m.elements.begin(), m.elements.end(), elements);
        }
    }

    int nsfem_context_velocity_mesh(const nsfem_context * c, double * xy, int * elements)
    {
        return guarded([&]
                       {
            need(c, "context");
            if (c->ctx.numSystems() > 1) throw std::invalid_argument("batch contexts have no single mesh");
            meshOut(c->velocityMesh, xy, elements); });
    }

    int nsfem_context_pressure_mesh(const nsfem_context * c, double * xy, int * elements)
    {
        return guarded([&]
                       {
            need(c, "context");
            if (c->ctx.numSystems() > 1) throw std::invalid_argument("batch contexts have no single mesh");
            meshOut(c->pressureMesh, xy, elements); });
    }

    int nsfem_fast_convection_map(const nsfem_context * c, int numThreads, int * rows, int * cols, int * nnz, int ** rowStart, int ** column,
                                  double ** values)
    {
        return guarded([&]
                       {
            need(c, "context");
            need(rows, "rows");
            need(cols, "cols");
            need(nnz, "nnz");
            need(rowStart, "rowStart");
            need(column, "column");
            need(values, "values");
            if (c->ctx.numSystems() > 1) throw std::invalid_argument("batch contexts have no single mesh");
            CsrD m = buildFastConvectionMap(c->velocityMesh, c->ctx.velocityStiffness, numThreads > 0 ? numThreads : defaultThreads());
            *rows = m.rows;
            *cols = m.cols;
            *nnz = m.nnz();
            *rowStart = static_cast<int *>(std::malloc(sizeof(int) * m.rowStart.size()));
            *column = static_cast<int *>(std::malloc(sizeof(int) * std::max<size_t>(1, m.column.size())));
            *values = static_cast<double *>(std::malloc(sizeof(double) * std::max<size_t>(1, m.values.size())));
            if (!*rowStart || !*column || !*values) throw std::runtime_error("out of host memory");
            std::copy(m.rowStart.begin(), m.rowStart.end(), *rowStart);
            std::copy(m.column.begin(), m.column.end(), *column);
            std::copy(m.values.begin(), m.values.end(), *values); });
    }

    void nsfem_host_free(void * p) { std::free(p); }

    int nsfem_chorin_create(const nsfem_context * c, const nsfem_chorin_config * cfg, nsfem_chorin ** out)
    {
        return nsfem_chorin_create_ensemble(c, cfg, 1, nullptr, out);
    }

    int nsfem_chorin_create_ensemble(const nsfem_context * c, const nsfem_chorin_config * cfg, int instances, const double * inflowScale,
                                     nsfem_chorin ** out)
    {
        return guarded([&]
                       {
            need(c, "context");
            need(cfg, "config");
            need(out, "out");
            ChorinConfig cc;
            auto conv = [](const nsfem_solver_config & s, ChorinSolverConfig & d)
            {
                need(s.method, "solver method");
                d.method = s.method;
                d.maxIterations = s.maxIterations;
                d.target = s.targetMse;
                d.checkInterval = s.mseCheckInterval;
                d.lanes = s.lanes;
                d.rowsPerBlock = s.rowsPerBlock;
                d.sweepLanes = s.sweepLanes;
                d.sweepRowsPerBlock = s.sweepRowsPerBlock;
                if (d.rowsPerBlock < 0 || d.sweepRowsPerBlock < 0) throw std::invalid_argument("rowsPerBlock must be >= 0");
                if (d.checkInterval < 1) throw std::invalid_argument("Invalid MSE check interval for solver: " + std::to_string(d.checkInterval));
            };
            conv(cfg->velocitySolver, cc.velocity);
            conv(cfg->pressureSolver, cc.pressure);
            cc.dtype = cfg->dtype;
            cc.innerTol = cfg->innerTol > 0 ? cfg->innerTol : 1e-4;
            cc.graphs = cfg->eagerLaunches == 0;
            if (cfg->timeStepPolicy < NSFEM_TAU_REJECT || cfg->timeStepPolicy > NSFEM_TAU_UNCHECKED)
            {
                throw std::invalid_argument("timeStepPolicy must be NSFEM_TAU_REJECT, NSFEM_TAU_CLAMP or NSFEM_TAU_UNCHECKED");
            }
            cc.timeStepPolicy = static_cast<TimeStepPolicy>(cfg->timeStepPolicy);
            if (cfg->stepGraph != 0 && cfg->stepGraph != 1)
            {
                throw std::invalid_argument("stepGraph must be 0 or 1");
            }
            cc.stepGraph = cfg->stepGraph == 1;
            if (cfg->pressureStart != NSFEM_PRESSURE_START_LAST && cfg->pressureStart != NSFEM_PRESSURE_START_EXTRAPOLATE)
            {
                throw std::invalid_argument("pressureStart must be NSFEM_PRESSURE_START_LAST or NSFEM_PRESSURE_START_EXTRAPOLATE");
            }
            cc.pressureExtrapolation = cfg->pressureStart == NSFEM_PRESSURE_START_EXTRAPOLATE;
            if (cfg->velocityStart < NSFEM_VELOCITY_START_ZERO || cfg->velocityStart > NSFEM_VELOCITY_START_EXTRAPOLATE)
            {
                throw std::invalid_argument(
                    "velocityStart must be NSFEM_VELOCITY_START_ZERO, NSFEM_VELOCITY_START_LAST or NSFEM_VELOCITY_START_EXTRAPOLATE");
            }
            cc.velocityStart = cfg->velocityStart;
            if (cc.stepGraph)
            {
                for (const auto * sc : {&cc.velocity, &cc.pressure})
                {
                    const SolverOptions o = parseSolverName(sc->method);
                    if (o.method != SolverMethod::Pcg && o.method != SolverMethod::AmgPcg)
                    {
                        throw std::invalid_argument("stepGraph needs pcg, pcg_mixed, amg_pcg or amg_pcg_mixed solvers");
                    }
                }
            }
            cc.instances = instances;
            if (inflowScale && instances > 0)
            {
                cc.inflowScale.assign(inflowScale, inflowScale + instances);
                for (double v : cc.inflowScale)
                {
                    if (!std::isfinite(v)) throw std::invalid_argument("inflowScale values must be finite");
                }
            }
            auto h = std::make_unique<nsfem_chorin>();
            h->core = std::make_unique<ChorinGpu>(c->ctx, cc);
            *out = h.release(); });
    }

    int nsfem_chorin_destroy(nsfem_chorin * ch)
    {
        return guarded([&]
                       {
            syncLibraryStream();
            delete ch; });
    }

    int nsfem_chorin_set_time_step(nsfem_chorin * ch, double tau)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            if (!(tau > 0)) throw std::invalid_argument("time step must be positive");
            ch->core->setTimeStep(tau); });
    }

    int nsfem_chorin_step_graph_info(const nsfem_chorin * ch, int * builds, double * buildSeconds, int * rollbacks, int * budgets)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            if (builds) *builds = ch->core->stepGraphBuilds();
            if (buildSeconds) *buildSeconds = ch->core->stepGraphBuildSeconds();
            if (rollbacks) *rollbacks = ch->core->stepGraphRollbacks();
            if (budgets)
            {
                for (int k = 0; k < 3; k++) budgets[k] = ch->core->stepGraphBudget(k);
            } });
    }

    int nsfem_chorin_set_step_graph_budgets(nsfem_chorin * ch, int tentative, int pressure, int final)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            ch->core->setStepGraphBudgets(tentative, pressure, final); });
    }

    int nsfem_chorin_time_step_info(nsfem_chorin * ch, double * tauRequested, double * tauUsed, double * lambdaNu, double * tauMax)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            if (tauRequested) *tauRequested = ch->core->requestedTimeStep();
            if (tauUsed) *tauUsed = ch->core->timeStep();
            if (lambdaNu) *lambdaNu = ch->core->diffusionLambda();
            if (tauMax) *tauMax = ch->core->stableTimeStep(); });
    }

    int nsfem_chorin_step(nsfem_chorin * ch, nsfem_step_stats * stats)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            ChorinStepStats s{};
            ch->core->step(&s);
            if (stats)
            {
                stats->mseTentX = s.mseTentX;
                stats->mseTentY = s.mseTentY;
                stats->msePressure = s.msePressure;
                stats->mseFinalX = s.mseFinalX;
                stats->mseFinalY = s.mseFinalY;
                stats->relTent = s.relTent;
                stats->relPressure = s.relPressure;
                stats->relFinal = s.relFinal;
                stats->msConvection = s.msConvection;
                stats->msTentative = s.msTentative;
                stats->msPressure = s.msPressure;
                stats->msFinal = s.msFinal;
                stats->msTotal = s.msTotal;
                stats->itersTent = s.itersTent;
                stats->itersPressure = s.itersPressure;
                stats->itersFinal = s.itersFinal;
                stats->outerTent = s.outerTent;
                stats->outerPressure = s.outerPressure;
                stats->outerFinal = s.outerFinal;
                stats->tauRequested = s.tauRequested;
                stats->tauUsed = s.tauUsed;
                stats->capturedStep = s.capturedStep;
            } });
    }

    int nsfem_chorin_state(nsfem_chorin * ch, void ** velocity, void ** pressure)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            if (velocity) *velocity = ch->core->velocity();
            if (pressure) *pressure = ch->core->pressure(); });
    }

    int nsfem_chorin_download(nsfem_chorin * ch, void * velocity, void * pressure)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            const size_t ts = dtypeSize(ch->core->dtype());
            hipStream_t s = libraryStream();
            if (velocity)
                NSFEM_HIP(hipMemcpyAsync(velocity, ch->core->velocity(), 2 * ts * ch->core->numVelocityNodes() * ch->core->instances(),
                                         hipMemcpyDeviceToHost, s));
            if (pressure)
                NSFEM_HIP(hipMemcpyAsync(pressure, ch->core->pressure(), ts * ch->core->numPressureNodes() * ch->core->instances(),
                                         hipMemcpyDeviceToHost, s));
            syncLibraryStream(); });
    }

    int nsfem_chorin_download_async(nsfem_chorin * ch)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            ch->core->downloadAsync(); });
    }

    int nsfem_chorin_download_wait(nsfem_chorin * ch, void * velocity, void * pressure)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            ch->core->downloadWait(velocity, pressure); });
    }

    int nsfem_chorin_upload_velocity(nsfem_chorin * ch, const void * velocity)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            need(velocity, "velocity");
            const size_t ts = dtypeSize(ch->core->dtype());
            NSFEM_HIP(hipMemcpyAsync(ch->core->velocity(), velocity, 2 * ts * ch->core->numVelocityNodes() * ch->core->instances(),
                                     hipMemcpyHostToDevice, libraryStream()));
            syncLibraryStream(); });
    }

    int nsfem_chorin_instances(const nsfem_chorin * ch, int * instances)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            need(instances, "instances");
            *instances = ch->core->instances(); });
    }

    int nsfem_chorin_reset(nsfem_chorin * ch)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            ch->core->resetState(); });
    }

    void chorinShape(const nsfem_chorin & ch, int * shape)
    {
        const McgsSolver & v = ch.core->velocitySolver();
        const McgsSolver & p = ch.core->pressureSolver();
        const int s[18] = {v.lanes(), v.rowsPerBlock(), p.lanes(), p.rowsPerBlock(),
                           v.sweepLanes(), v.sweepRowsPerBlock(), p.sweepLanes(), p.sweepRowsPerBlock(),
                           v.kindRows(0), v.kindRows(1), v.kindRows(2), p.kindRows(0), p.kindRows(1), p.kindRows(2),
                           v.directionLanes(), p.directionLanes(), v.fused() ? 1 : 0, p.fused() ? 1 : 0};
        std::copy(s, s + 18, shape);
    }

    int nsfem_chorin_shape(const nsfem_chorin * ch, int * shape)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            need(shape, "shape");
            chorinShape(*ch, shape); });
    }

    int nsfem_chorin_set_solver_option(nsfem_chorin * ch, int which, const char * key, double value)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            need(key, "key");
            if (which < 0 || which > 2) throw std::invalid_argument("which must be 0 (velocity), 1 (pressure) or 2 (both)");
            if (which != 1) applySolverOption(ch->core->velocitySolver(), key, value);
            if (which != 0) applySolverOption(ch->core->pressureSolver(), key, value);
            ch->core->invalidateStepGraph(); });
    }

    int nsfem_chorin_autotune(nsfem_chorin * ch, int reps, int * shape)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            if (reps < 1) throw std::invalid_argument("reps must be positive");
            // block PCG keeps its fixed shapes
            for (McgsSolver * sv : {&ch->core->velocitySolver(), &ch->core->pressureSolver()})
            {
                if (sv->method() == SolverMethod::Pcg) sv->autotune(reps);
            }
            if (shape)
            {
                chorinShape(*ch, shape);
            } });
    }

    int nsfem_chorin_time_kernel(nsfem_chorin * ch, const char * which, int reps, double * avgMs, double * bytes,
                                 double * launchesPerIteration)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            need(which, "which");
            need(avgMs, "avgMs");
            *avgMs = ch->core->timeKernel(which, reps, bytes, launchesPerIteration); });
    }

    int nsfem_chorin_assemble_convection(nsfem_chorin * ch, const void * velocity, void * aValues, void * cValues)
    {
        return guarded([&]
                       {
            need(ch, "chorin");
            need(velocity, "velocity");
            need(aValues, "aValues");
            ch->core->assembleConvection(velocity, aValues, cValues, libraryStream());
            syncLibraryStream(); });
    }
}
