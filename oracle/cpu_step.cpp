// ORACLE - test infrastructure only: the CPU baseline.
//
// A multithreaded FP64 C++ restatement of the reference's Chorin projection
// step, the data flow of the reference's CPU path (chorinEigen.cpp:272-484:
// convection update, A = A0 + C, A [ux uy], mass solve, tentative update,
// divergence, pressure solve on the internal nodes, gradient, mass solve,
// final update) with the GPU path's step semantics (chorinCuda.cu:236-376:
// u -= tau a on every node, then the Dirichlet values re-imposed, X ids then
// Y ids + N; the pressure solve warm-started from the previous step, the
// velocity solves from zero).  Every linear solve is CG preconditioned by one
// symmetric multicolour Gauss-Seidel sweep (smallest-last greedy colouring,
// forward colours 0..C-1 then backward C-2..0, from zero), stopped at
// |r| <= rtol |b|.  Threads: OpenMP, static schedules (deterministic for a
// fixed thread count).  Only tests and the bench's CPU-baseline leg load it.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" int oracle_sl_greedy(int n, const int * rowStart, const int * column, int * order, int * color);

namespace
{
    // The published 6-point degree-4 triangle rule (x, y, weight including
    // the reference area 1/2), the table the reference integrates with
    // (triangleIntegrator.cpp:27-35).
    const double kQuad[6][3] = {
        {9.15762135097710067155318824916321e-02, 9.15762135097710067155318824916321e-02, 5.49758718276610602870846378209535e-02},
        {9.15762135097710067155318824916321e-02, 8.16847572980457958813360619387822e-01, 5.49758718276610602870846378209535e-02},
        {8.16847572980457958813360619387822e-01, 9.15762135097710067155318824916321e-02, 5.49758718276610602870846378209535e-02},
        {4.45948490915965001235576892213430e-01, 4.45948490915965001235576892213430e-01, 1.11690794839005624883299105931656e-01},
        {4.45948490915965001235576892213430e-01, 1.08103018168069997528846215573139e-01, 1.11690794839005624883299105931656e-01},
        {1.08103018168069997528846215573139e-01, 4.45948490915965001235576892213430e-01, 1.11690794839005624883299105931656e-01},
    };

    // P2 on the reference triangle, nodes: corner 0, mid 01, corner 1, mid
    // 12, corner 2, mid 20.
    void p2Shape(double x, double y, double * s)
    {
        const double l = 1 - x - y;
        s[0] = l * (1 - 2 * x - 2 * y);
        s[1] = 4 * x * l;
        s[2] = x * (2 * x - 1);
        s[3] = 4 * x * y;
        s[4] = y * (2 * y - 1);
        s[5] = 4 * y * l;
    }
    void p2Grad(double x, double y, double * gx, double * gy)
    {
        gx[0] = 4 * x + 4 * y - 3;
        gx[1] = 4 - 8 * x - 4 * y;
        gx[2] = 4 * x - 1;
        gx[3] = 4 * y;
        gx[4] = 0;
        gx[5] = -4 * y;
        gy[0] = 4 * x + 4 * y - 3;
        gy[1] = -4 * x;
        gy[2] = 0;
        gy[3] = 4 * x;
        gy[4] = 4 * y - 1;
        gy[5] = 4 - 4 * x - 8 * y;
    }

    struct Csr
    {
        int rows = 0, cols = 0;
        std::vector<int> rowStart, column;
        std::vector<double> values;
        Csr() = default;
        Csr(int r, int c, const int * rs, const int * col, const double * val) : rows(r), cols(c), rowStart(rs, rs + r + 1)
        {
            const int nnz = rs[r];
            column.assign(col, col + nnz);
            values.assign(val, val + nnz);
        }
        // y = alpha A x (+ y if accumulate)
        void mul(const double * x, double * y, double alpha = 1.0) const
        {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < rows; i++)
            {
                double s = 0;
                for (int k = rowStart[i]; k < rowStart[i + 1]; k++)
                {
                    s += values[k] * x[column[k]];
                }
                y[i] = alpha * s;
            }
        }
    };

    double dot(const double * a, const double * b, long n)
    {
        double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
        for (long i = 0; i < n; i++)
        {
            s += a[i] * b[i];
        }
        return s;
    }

    // CG with a symmetric multicolour GS preconditioner on the colour-
    // permuted matrix.
    struct Pcg
    {
        int n = 0;
        std::vector<int> perm;       // new -> old
        std::vector<int> colorStart; // numColors + 1
        Csr P;                       // permuted, full rows
        std::vector<double> diag;
        std::vector<double> bp, xp, r, z, p, q;

        explicit Pcg(const Csr & A)
        {
            n = A.rows;
            std::vector<int> order(n), color(n);
            const int k = oracle_sl_greedy(n, A.rowStart.data(), A.column.data(), order.data(), color.data());
            if (k < 0)
            {
                throw std::runtime_error("colouring failed");
            }
            colorStart.assign(k + 1, 0);
            for (int i = 0; i < n; i++)
            {
                colorStart[color[i] + 1]++;
            }
            for (int c = 0; c < k; c++)
            {
                colorStart[c + 1] += colorStart[c];
            }
            perm.resize(n);
            std::vector<int> fill(colorStart.begin(), colorStart.end() - 1);
            for (int i = 0; i < n; i++)
            {
                perm[fill[color[i]]++] = i; // ascending old index within each colour
            }
            std::vector<int> inv(n);
            for (int i = 0; i < n; i++)
            {
                inv[perm[i]] = i;
            }
            P.rows = P.cols = n;
            P.rowStart.assign(n + 1, 0);
            for (int i = 0; i < n; i++)
            {
                const int o = perm[i];
                P.rowStart[i + 1] = P.rowStart[i] + (A.rowStart[o + 1] - A.rowStart[o]);
            }
            P.column.resize(P.rowStart[n]);
            P.values.resize(P.rowStart[n]);
            diag.assign(n, 0.0);
            for (int i = 0; i < n; i++)
            {
                const int o = perm[i];
                std::vector<std::pair<int, double>> row;
                for (int t = A.rowStart[o]; t < A.rowStart[o + 1]; t++)
                {
                    row.emplace_back(inv[A.column[t]], A.values[t]);
                    if (A.column[t] == o)
                    {
                        diag[i] = A.values[t];
                    }
                }
                std::sort(row.begin(), row.end());
                for (size_t t = 0; t < row.size(); t++)
                {
                    P.column[P.rowStart[i] + t] = row[t].first;
                    P.values[P.rowStart[i] + t] = row[t].second;
                }
                if (diag[i] == 0)
                {
                    throw std::runtime_error("zero on the diagonal");
                }
            }
            bp.resize(n);
            xp.resize(n);
            r.resize(n);
            z.resize(n);
            p.resize(n);
            q.resize(n);
        }

        void sweepColor(int c, const double * b, double * x) const
        {
#pragma omp parallel for schedule(static)
            for (int i = colorStart[c]; i < colorStart[c + 1]; i++)
            {
                double s = b[i];
                for (int t = P.rowStart[i]; t < P.rowStart[i + 1]; t++)
                {
                    const int j = P.column[t];
                    if (j != i)
                    {
                        s -= P.values[t] * x[j];
                    }
                }
                x[i] = s / diag[i];
            }
        }

        void sgs(const double * rr, double * zz) const
        {
            std::memset(zz, 0, sizeof(double) * n);
            const int k = static_cast<int>(colorStart.size()) - 1;
            for (int c = 0; c < k; c++)
            {
                sweepColor(c, rr, zz);
            }
            for (int c = k - 2; c >= 0; c--)
            {
                sweepColor(c, rr, zz);
            }
        }

        // x (natural order) on entry is the initial guess; returns iterations.
        int solve(const double * b, double * x, double rtol, int maxIters)
        {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++)
            {
                bp[i] = b[perm[i]];
                xp[i] = x[perm[i]];
            }
            P.mul(xp.data(), q.data());
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++)
            {
                r[i] = bp[i] - q[i];
            }
            const double bb = dot(bp.data(), bp.data(), n);
            double rr = dot(r.data(), r.data(), n);
            const double ref = bb > 0 ? bb : rr;
            int it = 0;
            if (rr > rtol * rtol * ref)
            {
                sgs(r.data(), z.data());
                std::copy(z.begin(), z.end(), p.begin());
                double rho = dot(r.data(), z.data(), n);
                while (it < maxIters)
                {
                    P.mul(p.data(), q.data());
                    const double alpha = rho / dot(p.data(), q.data(), n);
#pragma omp parallel for schedule(static)
                    for (int i = 0; i < n; i++)
                    {
                        xp[i] += alpha * p[i];
                        r[i] -= alpha * q[i];
                    }
                    it++;
                    rr = dot(r.data(), r.data(), n);
                    if (rr <= rtol * rtol * ref)
                    {
                        break;
                    }
                    sgs(r.data(), z.data());
                    const double rhoNew = dot(r.data(), z.data(), n);
                    const double beta = rhoNew / rho;
                    rho = rhoNew;
#pragma omp parallel for schedule(static)
                    for (int i = 0; i < n; i++)
                    {
                        p[i] = z[i] + beta * p[i];
                    }
                }
            }
            if (!std::isfinite(rr))
            {
                throw std::runtime_error("non-finite residual");
            }
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++)
            {
                x[perm[i]] = xp[i];
            }
            return it;
        }
    };
} // namespace

struct cpu_step
{
    int N = 0, V = 0, E = 0, Vi = 0;
    double nu = 0, tau = 0, rtol = 1e-10;
    int maxVel = 5000, maxPres = 100000;
    Csr M, K, Kp, VPD, PVD;
    std::vector<int> elements;   // E x 6
    std::vector<double> geom;    // E x 5: B00 B01 B10 B11 sgn
    std::vector<int> slotStart;  // nnz + 1: contributions of each CSR entry
    std::vector<int> slotList;   // element * 36 + local (r * 6 + c)
    std::vector<int> dirIds;     // X ids then Y ids + N
    std::vector<double> dirVals;
    std::vector<int> internalP;
    std::unique_ptr<Pcg> vSolver, pSolver;
    std::vector<double> u, p, pInt, elemC, A, rhs, a, d, dInt, g;
    int lastIters[3] = {0, 0, 0};
    std::string error;
};

namespace
{
    void impose(cpu_step & s)
    {
        for (size_t k = 0; k < s.dirIds.size(); k++)
        {
            s.u[s.dirIds[k]] = s.dirVals[k];
        }
    }

    void assembleA(cpu_step & s)
    {
        const int E = s.E, N = s.N;
        double shp[6][6], gx[6][6], gy[6][6];
        for (int q = 0; q < 6; q++)
        {
            p2Shape(kQuad[q][0], kQuad[q][1], shp[q]);
            p2Grad(kQuad[q][0], kQuad[q][1], gx[q], gy[q]);
        }
#pragma omp parallel for schedule(static)
        for (int e = 0; e < E; e++)
        {
            const int * nd = &s.elements[6 * static_cast<size_t>(e)];
            const double * gm = &s.geom[5 * static_cast<size_t>(e)];
            double ux[6], uy[6];
            for (int k = 0; k < 6; k++)
            {
                ux[k] = s.u[nd[k]];
                uy[k] = s.u[nd[k] + N];
            }
            double loc[36] = {0};
            for (int q = 0; q < 6; q++)
            {
                double fx = 0, fy = 0;
                for (int k = 0; k < 6; k++)
                {
                    fx += ux[k] * shp[q][k];
                    fy += uy[k] * shp[q][k];
                }
                const double lx = gm[0] * fx + gm[2] * fy;
                const double ly = gm[1] * fx + gm[3] * fy;
                for (int r = 0; r < 6; r++)
                {
                    const double wr = kQuad[q][2] * shp[q][r];
                    for (int c = 0; c < 6; c++)
                    {
                        loc[r * 6 + c] += wr * (lx * gx[q][c] + ly * gy[q][c]);
                    }
                }
            }
            double * out = &s.elemC[36 * static_cast<size_t>(e)];
            for (int k = 0; k < 36; k++)
            {
                out[k] = gm[4] * loc[k];
            }
        }
        const int nnz = static_cast<int>(s.K.values.size());
#pragma omp parallel for schedule(static)
        for (int j = 0; j < nnz; j++)
        {
            double c = 0;
            for (int t = s.slotStart[j]; t < s.slotStart[j + 1]; t++)
            {
                c += s.elemC[s.slotList[t]];
            }
            s.A[j] = s.nu * s.K.values[j] + c;
        }
    }

    int velocitySolve(cpu_step & s, const double * b, double * x)
    {
        int it = 0;
        for (int c = 0; c < 2; c++)
        {
            std::fill(x + static_cast<size_t>(c) * s.N, x + static_cast<size_t>(c + 1) * s.N, 0.0);
            it = std::max(it, s.vSolver->solve(b + static_cast<size_t>(c) * s.N, x + static_cast<size_t>(c) * s.N, s.rtol, s.maxVel));
        }
        return it;
    }

    void step(cpu_step & s)
    {
        const int N = s.N;
        assembleA(s);
        // rhs = A [ux uy]
        {
            const Csr & K = s.K;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < N; i++)
            {
                double sx = 0, sy = 0;
                for (int k = K.rowStart[i]; k < K.rowStart[i + 1]; k++)
                {
                    const int j = K.column[k];
                    sx += s.A[k] * s.u[j];
                    sy += s.A[k] * s.u[j + N];
                }
                s.rhs[i] = sx;
                s.rhs[i + N] = sy;
            }
        }
        s.lastIters[0] = velocitySolve(s, s.rhs.data(), s.a.data());
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2L * N; i++)
        {
            s.u[i] -= s.tau * s.a[i];
        }
        impose(s);
        // pressure: Kp_int p = -(1/tau) (VPD u)[internal], warm start
        s.VPD.mul(s.u.data(), s.d.data(), -1.0 / s.tau);
        for (int k = 0; k < s.Vi; k++)
        {
            s.dInt[k] = s.d[s.internalP[k]];
        }
        s.lastIters[1] = s.pSolver->solve(s.dInt.data(), s.pInt.data(), s.rtol, s.maxPres);
        std::fill(s.p.begin(), s.p.end(), 0.0);
        for (int k = 0; k < s.Vi; k++)
        {
            s.p[s.internalP[k]] = s.pInt[k];
        }
        // projection: M a = PVD p; u -= tau a
        s.PVD.mul(s.p.data(), s.g.data());
        s.lastIters[2] = velocitySolve(s, s.g.data(), s.a.data());
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2L * N; i++)
        {
            s.u[i] -= s.tau * s.a[i];
        }
        impose(s);
    }

    int fail(cpu_step * s, const std::exception & e)
    {
        if (s)
        {
            s->error = e.what();
        }
        return 1;
    }
} // namespace

extern "C" {

// Matrices as CSR arrays (int32, 0-based): the velocity mass M and stiffness
// K (N x N, one pattern, which the convection shares), the internal pressure
// Laplacian Kp_int (Vi x Vi), VPD (V x 2N), PVD (2N x V); P2 elements (E x 6,
// corners at 0, 2, 4) and P2 node coordinates (N x 2); Dirichlet ids/values
// per component; the internal pressure nodes.  threads <= 0: OpenMP's
// default.
int cpu_step_create(int N, int V, int E, const int * mRs, const int * mCol, const double * mVal, const int * kRs, const int * kCol,
                    const double * kVal, double viscosity, int Vi, const int * pRs, const int * pCol, const double * pVal,
                    const int * vpdRs, const int * vpdCol, const double * vpdVal, const int * pvdRs, const int * pvdCol,
                    const double * pvdVal, const int * elements, const double * xy, int nDirX, const int * dirXIds,
                    const double * dirXVals, int nDirY, const int * dirYIds, const double * dirYVals, const int * internalP,
                    double tau, double rtol, int threads, cpu_step ** out)
{
    try
    {
        if (threads > 0)
        {
            omp_set_num_threads(threads);
        }
        auto s = std::make_unique<cpu_step>();
        s->N = N;
        s->V = V;
        s->E = E;
        s->Vi = Vi;
        s->nu = viscosity;
        s->tau = tau;
        s->rtol = rtol;
        s->M = Csr(N, N, mRs, mCol, mVal);
        s->K = Csr(N, N, kRs, kCol, kVal);
        s->Kp = Csr(Vi, Vi, pRs, pCol, pVal);
        s->VPD = Csr(V, 2 * N, vpdRs, vpdCol, vpdVal);
        s->PVD = Csr(2 * N, V, pvdRs, pvdCol, pvdVal);
        s->elements.assign(elements, elements + 6 * static_cast<size_t>(E));
        s->geom.resize(5 * static_cast<size_t>(E));
        for (int e = 0; e < E; e++)
        {
            const int * nd = &s->elements[6 * static_cast<size_t>(e)];
            const double * p0 = xy + 2 * static_cast<size_t>(nd[0]);
            const double * p1 = xy + 2 * static_cast<size_t>(nd[2]);
            const double * p2 = xy + 2 * static_cast<size_t>(nd[4]);
            const double m00 = p1[0] - p0[0], m01 = p2[0] - p0[0];
            const double m10 = p1[1] - p0[1], m11 = p2[1] - p0[1];
            const double det = m00 * m11 - m01 * m10;
            double * g = &s->geom[5 * static_cast<size_t>(e)];
            g[0] = m11;  // B = [[m11, -m10], [-m01, m00]]
            g[1] = -m10;
            g[2] = -m01;
            g[3] = m00;
            g[4] = det < 0 ? -1.0 : 1.0;
        }
        // CSR slot of every element-local entry, then the contributions of
        // each CSR entry in ascending (element, local) order
        const int nnz = kRs[N];
        std::vector<int> slot(36 * static_cast<size_t>(E));
#pragma omp parallel for schedule(static)
        for (int e = 0; e < E; e++)
        {
            const int * nd = &s->elements[6 * static_cast<size_t>(e)];
            for (int r = 0; r < 6; r++)
            {
                const int * b = kCol + kRs[nd[r]];
                const int * en = kCol + kRs[nd[r] + 1];
                for (int c = 0; c < 6; c++)
                {
                    const int * f = std::lower_bound(b, en, nd[c]);
                    slot[36 * static_cast<size_t>(e) + r * 6 + c] = (f != en && *f == nd[c]) ? static_cast<int>(f - kCol) : -1;
                }
            }
        }
        s->slotStart.assign(nnz + 1, 0);
        for (int v : slot)
        {
            if (v < 0)
            {
                throw std::runtime_error("element entry outside the stiffness pattern");
            }
            s->slotStart[v + 1]++;
        }
        for (int j = 0; j < nnz; j++)
        {
            s->slotStart[j + 1] += s->slotStart[j];
        }
        s->slotList.resize(slot.size());
        {
            std::vector<int> fill(s->slotStart.begin(), s->slotStart.end() - 1);
            for (size_t k = 0; k < slot.size(); k++)
            {
                s->slotList[fill[slot[k]]++] = static_cast<int>(k);
            }
        }
        for (int k = 0; k < nDirX; k++)
        {
            s->dirIds.push_back(dirXIds[k]);
            s->dirVals.push_back(dirXVals[k]);
        }
        for (int k = 0; k < nDirY; k++)
        {
            s->dirIds.push_back(dirYIds[k] + N);
            s->dirVals.push_back(dirYVals[k]);
        }
        s->internalP.assign(internalP, internalP + Vi);
        s->vSolver = std::make_unique<Pcg>(s->M);
        s->pSolver = std::make_unique<Pcg>(s->Kp);
        s->u.assign(2 * static_cast<size_t>(N), 0.0);
        s->p.assign(V, 0.0);
        s->pInt.assign(Vi, 0.0);
        s->elemC.resize(36 * static_cast<size_t>(E));
        s->A.resize(nnz);
        s->rhs.resize(2 * static_cast<size_t>(N));
        s->a.resize(2 * static_cast<size_t>(N));
        s->d.resize(V);
        s->dInt.resize(Vi);
        s->g.resize(2 * static_cast<size_t>(N));
        impose(*s);
        *out = s.release();
        return 0;
    }
    catch (const std::exception & e)
    {
        return 1;
    }
}

int cpu_step_set_velocity(cpu_step * s, const double * u)
{
    std::copy(u, u + 2 * static_cast<size_t>(s->N), s->u.begin());
    return 0;
}

// Runs steps; iters (optional) receives 3 iteration counts per step.
int cpu_step_run(cpu_step * s, int steps, int * iters)
{
    try
    {
        for (int k = 0; k < steps; k++)
        {
            step(*s);
            if (iters)
            {
                std::copy(s->lastIters, s->lastIters + 3, iters + 3 * k);
            }
        }
        return 0;
    }
    catch (const std::exception & e)
    {
        return fail(s, e);
    }
}

int cpu_step_state(const cpu_step * s, double * u, double * p)
{
    std::copy(s->u.begin(), s->u.end(), u);
    std::copy(s->p.begin(), s->p.end(), p);
    return 0;
}

int cpu_step_threads(void) { return omp_get_max_threads(); }

const char * cpu_step_error(const cpu_step * s) { return s ? s->error.c_str() : ""; }

void cpu_step_destroy(cpu_step * s) { delete s; }
}
