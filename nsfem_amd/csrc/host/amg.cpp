#include "amg.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace nsfem
{
    namespace
    {
        // Full matrix of a level (stripped + diagonal), ascending columns.
        CsrD assemble(const AmgLevel & L)
        {
            CsrD a;
            a.rows = a.cols = L.n;
            a.rowStart.resize(L.n + 1);
            a.column.reserve(L.stripped.nnz() + L.n);
            a.values.reserve(L.stripped.nnz() + L.n);
            std::vector<std::pair<int, double>> row;
            for (int i = 0; i < L.n; i++)
            {
                a.rowStart[i] = static_cast<int>(a.column.size());
                row.clear();
                for (int j = L.stripped.rowStart[i]; j < L.stripped.rowStart[i + 1]; j++)
                {
                    row.emplace_back(L.stripped.column[j], L.stripped.values[j]);
                }
                row.emplace_back(i, L.diag[i]);
                std::sort(row.begin(), row.end(), [](const auto & x, const auto & y) { return x.first < y.first; });
                for (const auto & [c, v] : row)
                {
                    a.column.push_back(c);
                    a.values.push_back(v);
                }
            }
            a.rowStart[L.n] = static_cast<int>(a.column.size());
            return a;
        }

        CsrD transpose(const CsrD & m)
        {
            CsrD t;
            t.rows = m.cols;
            t.cols = m.rows;
            t.rowStart.assign(m.cols + 1, 0);
            for (int c : m.column)
            {
                t.rowStart[c + 1]++;
            }
            for (int i = 0; i < m.cols; i++)
            {
                t.rowStart[i + 1] += t.rowStart[i];
            }
            t.column.resize(m.nnz());
            t.values.resize(m.nnz());
            std::vector<int> at(t.rowStart.begin(), t.rowStart.end() - 1);
            for (int i = 0; i < m.rows; i++) // rows in order: columns of t stay ascending
            {
                for (int j = m.rowStart[i]; j < m.rowStart[i + 1]; j++)
                {
                    const int k = at[m.column[j]]++;
                    t.column[k] = i;
                    t.values[k] = m.values[j];
                }
            }
            return t;
        }

        // C = A B, rows with ascending columns (dense accumulator per row).
        CsrD multiply(const CsrD & a, const CsrD & b)
        {
            if (a.cols != b.rows)
            {
                throw std::invalid_argument("amg: product shapes differ");
            }
            CsrD c;
            c.rows = a.rows;
            c.cols = b.cols;
            c.rowStart.resize(a.rows + 1);
            std::vector<double> acc(b.cols, 0.0);
            std::vector<int> mark(b.cols, -1), cols;
            for (int i = 0; i < a.rows; i++)
            {
                c.rowStart[i] = static_cast<int>(c.column.size());
                cols.clear();
                for (int j = a.rowStart[i]; j < a.rowStart[i + 1]; j++)
                {
                    const int k = a.column[j];
                    const double v = a.values[j];
                    for (int l = b.rowStart[k]; l < b.rowStart[k + 1]; l++)
                    {
                        const int col = b.column[l];
                        if (mark[col] != i)
                        {
                            mark[col] = i;
                            acc[col] = 0.0;
                            cols.push_back(col);
                        }
                        acc[col] += v * b.values[l];
                    }
                }
                std::sort(cols.begin(), cols.end());
                for (int col : cols)
                {
                    c.column.push_back(col);
                    c.values.push_back(acc[col]);
                }
            }
            c.rowStart[a.rows] = static_cast<int>(c.column.size());
            return c;
        }

        // (A + A^T) / 2 for a structurally symmetric A.
        CsrD symmetrise(const CsrD & a)
        {
            const CsrD t = transpose(a);
            CsrD s = a;
            for (int i = 0; i < a.rows; i++)
            {
                if (a.rowStart[i + 1] - a.rowStart[i] != t.rowStart[i + 1] - t.rowStart[i])
                {
                    throw std::invalid_argument("amg: coarse operator is not structurally symmetric");
                }
                for (int j = a.rowStart[i], k = t.rowStart[i]; j < a.rowStart[i + 1]; j++, k++)
                {
                    s.values[j] = 0.5 * (a.values[j] + t.values[k]);
                }
            }
            return s;
        }

        // Largest eigenvalue of D^-1 A (similar to the symmetric D^-1/2 A D^-1/2)
        // by power iteration from a fixed start.
        double lambdaMaxDinvA(const CsrD & a, const std::vector<double> & diag)
        {
            const int n = a.rows;
            std::vector<double> x(n), y(n);
            uint32_t s = 12345u;
            for (auto & v : x)
            {
                s = s * 1664525u + 1013904223u;
                v = 0.5 + static_cast<double>(s >> 8) / 16777216.0;
            }
            double lam = 0;
            for (int it = 0; it < 30; it++)
            {
                double nrm = 0;
                for (double v : x)
                {
                    nrm += v * v;
                }
                nrm = std::sqrt(nrm);
                if (!(nrm > 0))
                {
                    break;
                }
                for (auto & v : x)
                {
                    v /= nrm;
                }
                double num = 0, den = 0;
                for (int i = 0; i < n; i++)
                {
                    double t = 0;
                    for (int j = a.rowStart[i]; j < a.rowStart[i + 1]; j++)
                    {
                        t += a.values[j] * x[a.column[j]];
                    }
                    y[i] = t / diag[i];
                    // Rayleigh quotient in the D inner product: x^T A x / x^T D x
                    num += x[i] * t;
                    den += x[i] * x[i] * diag[i];
                }
                lam = num / den;
                x.swap(y);
            }
            return lam;
        }

        // Aggregates of the strong-connection graph (three greedy passes).
        std::vector<int> aggregate(const CsrD & a, const std::vector<double> & diag, double theta, int & numAgg)
        {
            const int n = a.rows;
            std::vector<int> strongStart(n + 1, 0), strong;
            strong.reserve(a.nnz());
            for (int i = 0; i < n; i++)
            {
                strongStart[i] = static_cast<int>(strong.size());
                for (int j = a.rowStart[i]; j < a.rowStart[i + 1]; j++)
                {
                    const int c = a.column[j];
                    if (c != i && std::abs(a.values[j]) >= theta * std::sqrt(std::abs(diag[i] * diag[c])))
                    {
                        strong.push_back(c);
                    }
                }
            }
            strongStart[n] = static_cast<int>(strong.size());
            std::vector<int> agg(n, -1);
            numAgg = 0;
            // 1. roots whose strong neighbourhood is wholly unaggregated
            for (int i = 0; i < n; i++)
            {
                if (agg[i] >= 0)
                {
                    continue;
                }
                bool free = true;
                for (int j = strongStart[i]; j < strongStart[i + 1] && free; j++)
                {
                    free = agg[strong[j]] < 0;
                }
                if (!free)
                {
                    continue;
                }
                agg[i] = numAgg;
                for (int j = strongStart[i]; j < strongStart[i + 1]; j++)
                {
                    agg[strong[j]] = numAgg;
                }
                numAgg++;
            }
            // 2. the rest join the pass-1 aggregate they are most strongly tied to
            std::vector<int> first = agg;
            for (int i = 0; i < n; i++)
            {
                if (first[i] >= 0)
                {
                    continue;
                }
                double best = -1;
                for (int j = a.rowStart[i]; j < a.rowStart[i + 1]; j++)
                {
                    const int c = a.column[j];
                    if (c == i || first[c] < 0)
                    {
                        continue;
                    }
                    const double w = std::abs(a.values[j]) / std::sqrt(std::abs(diag[i] * diag[c]));
                    if (w >= theta && w > best)
                    {
                        best = w;
                        agg[i] = first[c];
                    }
                }
            }
            // 3. leftovers: new aggregates of themselves and their free strong neighbours
            for (int i = 0; i < n; i++)
            {
                if (agg[i] >= 0)
                {
                    continue;
                }
                agg[i] = numAgg;
                for (int j = strongStart[i]; j < strongStart[i + 1]; j++)
                {
                    if (agg[strong[j]] < 0)
                    {
                        agg[strong[j]] = numAgg;
                    }
                }
                numAgg++;
            }
            return agg;
        }

        // Dense inverse of an SPD matrix (Cholesky, then L^-T L^-1); singular:
        // of A + s 1 1^T, s = mean diagonal / n (A positive semidefinite with
        // the constants as its null space).
        std::vector<double> denseInverse(const CsrD & a, bool singular)
        {
            const int n = a.rows;
            std::vector<double> m(static_cast<size_t>(n) * n, 0.0);
            double diagSum = 0;
            for (int i = 0; i < n; i++)
            {
                for (int j = a.rowStart[i]; j < a.rowStart[i + 1]; j++)
                {
                    m[static_cast<size_t>(i) * n + a.column[j]] = a.values[j];
                    if (a.column[j] == i)
                    {
                        diagSum += a.values[j];
                    }
                }
            }
            if (singular)
            {
                const double shift = diagSum / n / n;
                for (auto & v : m)
                {
                    v += shift;
                }
            }
            // lower Cholesky factor in place
            for (int j = 0; j < n; j++)
            {
                const double ajj = m[static_cast<size_t>(j) * n + j];
                double d = ajj;
                for (int k = 0; k < j; k++)
                {
                    d -= m[static_cast<size_t>(j) * n + k] * m[static_cast<size_t>(j) * n + k];
                }
                // a pivot that has lost all but rounding noise of its
                // diagonal: singular (e.g. the constant null space of a
                // pure-Neumann Laplacian)
                if (!(d > 1e-10 * std::abs(ajj)))
                {
                    throw std::invalid_argument(
                        "amg: the coarsest operator is not positive definite (a singular system without zero row sums?)");
                }
                d = std::sqrt(d);
                m[static_cast<size_t>(j) * n + j] = d;
                for (int i = j + 1; i < n; i++)
                {
                    double v = m[static_cast<size_t>(i) * n + j];
                    for (int k = 0; k < j; k++)
                    {
                        v -= m[static_cast<size_t>(i) * n + k] * m[static_cast<size_t>(j) * n + k];
                    }
                    m[static_cast<size_t>(i) * n + j] = v / d;
                }
            }
            // columns of the inverse: solve L L^T x = e_c
            std::vector<double> inv(static_cast<size_t>(n) * n), y(n);
            for (int c = 0; c < n; c++)
            {
                for (int i = 0; i < n; i++)
                {
                    double v = i == c ? 1.0 : 0.0;
                    for (int k = 0; k < i; k++)
                    {
                        v -= m[static_cast<size_t>(i) * n + k] * y[k];
                    }
                    y[i] = v / m[static_cast<size_t>(i) * n + i];
                }
                for (int i = n - 1; i >= 0; i--)
                {
                    double v = y[i];
                    for (int k = i + 1; k < n; k++)
                    {
                        v -= m[static_cast<size_t>(k) * n + i] * inv[static_cast<size_t>(k) * n + c];
                    }
                    inv[static_cast<size_t>(i) * n + c] = v / m[static_cast<size_t>(i) * n + i];
                }
            }
            // exactly symmetric
            for (int i = 0; i < n; i++)
            {
                for (int j = 0; j < i; j++)
                {
                    const double v = 0.5 * (inv[static_cast<size_t>(i) * n + j] + inv[static_cast<size_t>(j) * n + i]);
                    inv[static_cast<size_t>(i) * n + j] = v;
                    inv[static_cast<size_t>(j) * n + i] = v;
                }
            }
            return inv;
        }
    } // namespace

    AmgHierarchy buildAmgHierarchy(const CsrD & stripped, const std::vector<int> & rowSplit, const std::vector<double> & diag,
                                   const std::vector<int> & colorStart, const AmgOptions & opt)
    {
        const int n0 = stripped.rows;
        if (n0 < 1 || static_cast<int>(diag.size()) != n0 || static_cast<int>(rowSplit.size()) != n0 || colorStart.empty() ||
            colorStart.back() != n0)
        {
            throw std::invalid_argument("amg: inconsistent finest level");
        }
        AmgHierarchy h;
        {
            // zero row sums (to rounding): a pure-Neumann operator
            double worst = 0, dmax = 0;
            for (int i = 0; i < n0; i++)
            {
                double sum = diag[i];
                for (int j = stripped.rowStart[i]; j < stripped.rowStart[i + 1]; j++)
                {
                    sum += stripped.values[j];
                }
                worst = std::max(worst, std::abs(sum));
                dmax = std::max(dmax, std::abs(diag[i]));
            }
            h.singular = worst <= 1e-10 * dmax;
        }
        {
            AmgLevel L0;
            L0.n = n0;
            L0.stripped = stripped;
            L0.rowSplit = rowSplit;
            L0.diag = diag;
            L0.invDiag.resize(n0);
            for (int i = 0; i < n0; i++)
            {
                L0.invDiag[i] = 1.0 / diag[i];
            }
            L0.colorStart = colorStart;
            h.levels.push_back(std::move(L0));
        }
        double nnz0 = stripped.nnz() + n0, nnzAll = nnz0;
        double theta = opt.theta;
        for (;;)
        {
            AmgLevel & L = h.levels.back();
            const CsrD A = assemble(L);
            if (L.n <= opt.maxCoarse || static_cast<int>(h.levels.size()) >= opt.maxLevels)
            {
                h.coarseInverse = denseInverse(A, h.singular);
                break;
            }
            int nc = 0;
            const std::vector<int> agg = aggregate(A, L.diag, theta, nc);
            if (nc >= L.n || nc < 1)
            {
                h.coarseInverse = denseInverse(A, h.singular); // no coarsening left
                break;
            }
            // smoothed prolongation P = (I - omega D^-1 A) P_tent
            const double lambda = lambdaMaxDinvA(A, L.diag);
            const double omega = opt.omegaScale / lambda;
            {
                // upper bound of the spectrum of D^-1 A for a polynomial
                // smoother on this level: the power iteration's estimate
                // (from below) raised by 10 %, but never above the
                // Gershgorin bound max_i sum_j |a_ij| / a_ii
                double gersh = 0;
                for (int i = 0; i < A.rows; i++)
                {
                    double row = 0;
                    for (int j = A.rowStart[i]; j < A.rowStart[i + 1]; j++)
                    {
                        row += std::abs(A.values[j]);
                    }
                    gersh = std::max(gersh, row / L.diag[i]);
                }
                L.lambdaMax = std::min(1.1 * lambda, gersh);
            }
            CsrD P;
            P.rows = L.n;
            P.cols = nc;
            P.rowStart.resize(L.n + 1);
            {
                std::vector<double> acc(nc, 0.0);
                std::vector<int> mark(nc, -1), cols;
                for (int i = 0; i < L.n; i++)
                {
                    P.rowStart[i] = static_cast<int>(P.column.size());
                    cols.clear();
                    auto add = [&](int c, double v)
                    {
                        if (mark[c] != i)
                        {
                            mark[c] = i;
                            acc[c] = 0.0;
                            cols.push_back(c);
                        }
                        acc[c] += v;
                    };
                    add(agg[i], 1.0);
                    const double s = omega / L.diag[i];
                    for (int j = A.rowStart[i]; j < A.rowStart[i + 1]; j++)
                    {
                        add(agg[A.column[j]], -s * A.values[j]);
                    }
                    std::sort(cols.begin(), cols.end());
                    for (int c : cols)
                    {
                        if (acc[c] != 0.0)
                        {
                            P.column.push_back(c);
                            P.values.push_back(acc[c]);
                        }
                    }
                }
                P.rowStart[L.n] = static_cast<int>(P.column.size());
            }
            const CsrD R = transpose(P);
            const CsrD Ac = symmetrise(multiply(R, multiply(A, P)));
            // the coarse level in its own multicolour order
            McgsLayout lay = buildMcgsLayout(Ac);
            std::vector<int> pos(nc);
            for (int k = 0; k < nc; k++)
            {
                pos[lay.coloring[k]] = k;
            }
            for (auto & c : P.column)
            {
                c = pos[c];
            }
            for (int i = 0; i < L.n; i++) // columns ascending again
            {
                std::vector<std::pair<int, double>> row;
                for (int j = P.rowStart[i]; j < P.rowStart[i + 1]; j++)
                {
                    row.emplace_back(P.column[j], P.values[j]);
                }
                std::sort(row.begin(), row.end(), [](const auto & x, const auto & y) { return x.first < y.first; });
                for (int j = P.rowStart[i], k = 0; j < P.rowStart[i + 1]; j++, k++)
                {
                    P.column[j] = row[k].first;
                    P.values[j] = row[k].second;
                }
            }
            L.P = P;
            L.R = transpose(P);
            AmgLevel C;
            C.n = nc;
            C.stripped = std::move(lay.stripped);
            C.rowSplit = std::move(lay.rowSplit);
            C.diag = std::move(lay.diag);
            C.invDiag = std::move(lay.invDiag);
            C.colorStart = std::move(lay.partitionStart);
            nnzAll += C.stripped.nnz() + nc;
            h.levels.push_back(std::move(C));
            theta *= 0.5;
        }
        h.operatorComplexity = nnzAll / nnz0;
        return h;
    }
} // namespace nsfem
