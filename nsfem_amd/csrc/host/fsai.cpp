#include "fsai.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace nsfem
{
    namespace
    {
        // M(p, q) of the split matrix (q in row p, or the diagonal)
        double entry(int p, int q, const std::vector<int> & rowStart, const std::vector<int> & column, const std::vector<double> & values,
                     const std::vector<double> & diag)
        {
            if (p == q)
            {
                return diag[p];
            }
            for (int j = rowStart[p]; j < rowStart[p + 1]; j++)
            {
                if (column[j] == q)
                {
                    return values[j];
                }
            }
            return 0.0;
        }

        // Runs f(begin, end) over [0, n) on the host's cores.
        template <typename F>
        void parallelRows(int n, F && f)
        {
            const int threads = std::max(1, std::min(32, static_cast<int>(std::thread::hardware_concurrency())));
            if (threads == 1 || n < 4096)
            {
                f(0, n);
                return;
            }
            std::vector<std::thread> pool;
            std::vector<std::exception_ptr> errors(threads);
            for (int t = 0; t < threads; t++)
            {
                const int b = static_cast<int>(static_cast<long>(n) * t / threads);
                const int e = static_cast<int>(static_cast<long>(n) * (t + 1) / threads);
                pool.emplace_back(
                    [&, b, e, t]()
                    {
                        try
                        {
                            f(b, e);
                        }
                        catch (...)
                        {
                            errors[t] = std::current_exception();
                        }
                    });
            }
            for (auto & th : pool)
            {
                th.join();
            }
            for (auto & e : errors)
            {
                if (e)
                {
                    std::rethrow_exception(e);
                }
            }
        }
    } // namespace

    void buildFsai(int n, const std::vector<int> & rowStart, const std::vector<int> & rowSplit, const std::vector<int> & column,
                   const std::vector<double> & values, const std::vector<double> & diag, std::vector<double> & gValues,
                   std::vector<double> & gDiag)
    {
        if (n < 1 || static_cast<int>(rowStart.size()) != n + 1 || static_cast<int>(rowSplit.size()) != n ||
            static_cast<int>(diag.size()) != n || column.size() != values.size() || rowStart[n] != static_cast<int>(column.size()))
        {
            throw std::invalid_argument("fsai: inconsistent matrix");
        }
        gValues.assign(values.size(), 0.0);
        gDiag.assign(n, 0.0);
        // rows of G: the local system on P_i = lower-colour columns + i,
        // solved by Cholesky (m <= the row length + 1)
        parallelRows(n,
                     [&](int begin, int end)
                     {
                         std::vector<int> P;
                         std::vector<double> a, y;
                         for (int i = begin; i < end; i++)
                         {
                             P.assign(column.begin() + rowStart[i], column.begin() + rowSplit[i]);
                             P.push_back(i);
                             const int m = static_cast<int>(P.size());
                             a.assign(static_cast<size_t>(m) * m, 0.0);
                             for (int r = 0; r < m; r++)
                             {
                                 for (int c = 0; c <= r; c++)
                                 {
                                     a[r * m + c] = entry(P[r], P[c], rowStart, column, values, diag);
                                 }
                             }
                             // lower Cholesky factor in place
                             for (int c = 0; c < m; c++)
                             {
                                 double d = a[c * m + c];
                                 for (int k = 0; k < c; k++)
                                 {
                                     d -= a[c * m + k] * a[c * m + k];
                                 }
                                 if (!(d > 0))
                                 {
                                     throw std::invalid_argument("fsai: a local system is not positive definite");
                                 }
                                 d = std::sqrt(d);
                                 a[c * m + c] = d;
                                 for (int r = c + 1; r < m; r++)
                                 {
                                     double s = a[r * m + c];
                                     for (int k = 0; k < c; k++)
                                     {
                                         s -= a[r * m + k] * a[c * m + k];
                                     }
                                     a[r * m + c] = s / d;
                                 }
                             }
                             // L L^T y = e_last: L w = e_last gives w = e_last / L_last,last,
                             // then L^T y = w from the last row up
                             const double lmm = a[(m - 1) * m + (m - 1)];
                             y.assign(m, 0.0);
                             y[m - 1] = 1.0 / (lmm * lmm);
                             for (int r = m - 2; r >= 0; r--)
                             {
                                 double s = 0.0;
                                 for (int k = r + 1; k < m; k++)
                                 {
                                     s -= a[k * m + r] * y[k];
                                 }
                                 y[r] = s / a[r * m + r];
                             }
                             if (!(y[m - 1] > 0))
                             {
                                 throw std::invalid_argument("fsai: a local system is not positive definite");
                             }
                             const double scale = 1.0 / std::sqrt(y[m - 1]);
                             for (int r = 0; r < m - 1; r++)
                             {
                                 gValues[rowStart[i] + r] = y[r] * scale;
                             }
                             gDiag[i] = y[m - 1] * scale;
                         }
                     });
        // G^T: G(i, c) for c lower of i goes to row c's higher part, column i
        parallelRows(n,
                     [&](int begin, int end)
                     {
                         for (int c = begin; c < end; c++)
                         {
                             for (int k = rowSplit[c]; k < rowStart[c + 1]; k++)
                             {
                                 const int i = column[k];
                                 int found = -1;
                                 for (int j = rowStart[i]; j < rowSplit[i]; j++)
                                 {
                                     if (column[j] == c)
                                     {
                                         found = j;
                                         break;
                                     }
                                 }
                                 if (found < 0)
                                 {
                                     throw std::invalid_argument("fsai: the sparsity pattern is not symmetric");
                                 }
                                 gValues[k] = gValues[found];
                             }
                         }
                     });
    }
} // namespace nsfem
