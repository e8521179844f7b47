#include "csr.hpp"

#include <algorithm>
#include <queue>
#include <stdexcept>
#include <string>

namespace nsfem
{
    template <typename T>
    Csr<T> Csr<T>::slice(const std::vector<int> & rowIds, const std::vector<int> & colIds) const
    {
        Csr<T> result;
        if (rowIds.empty() || colIds.empty())
        {
            result.rowStart.assign(1, 0);
            return result;
        }
        for (int r : rowIds)
        {
            if (r < 0 || r >= rows)
            {
                throw std::invalid_argument("slice: Bad row index " + std::to_string(r));
            }
        }
        std::vector<int> colMap(cols, -1);
        for (size_t i = 0; i < colIds.size(); i++)
        {
            const int c = colIds[i];
            if (c < 0 || c >= cols)
            {
                throw std::invalid_argument("slice: Bad col index " + std::to_string(c));
            }
            colMap[c] = static_cast<int>(i);
        }
        result.rows = static_cast<int>(rowIds.size());
        result.cols = static_cast<int>(colIds.size());
        result.rowStart.resize(result.rows + 1);
        std::vector<std::pair<int, T>> tmp;
        for (int i = 0; i < result.rows; i++)
        {
            result.rowStart[i] = static_cast<int>(result.column.size());
            tmp.clear();
            const int src = rowIds[i];
            for (int j = rowStart[src]; j < rowStart[src + 1]; j++)
            {
                const int dc = colMap[column[j]];
                if (dc >= 0)
                {
                    tmp.emplace_back(dc, values[j]);
                }
            }
            // New column ids within a row are unique, so the order is total.
            std::sort(tmp.begin(), tmp.end(), [](const auto & a, const auto & b) { return a.first < b.first; });
            for (const auto & p : tmp)
            {
                result.column.push_back(p.first);
                result.values.push_back(p.second);
            }
        }
        result.rowStart[result.rows] = static_cast<int>(result.column.size());
        return result;
    }

    template <typename T>
    int Csr<T>::find(int row, int col) const
    {
        if (row < 0 || row >= rows)
        {
            return -1;
        }
        const int * b = column.data() + rowStart[row];
        const int * e = column.data() + rowStart[row + 1];
        const int * it = std::lower_bound(b, e, col);
        if (it == e || *it != col)
        {
            return -1;
        }
        return static_cast<int>(it - column.data());
    }

    template struct Csr<double>;
    template struct Csr<float>;

    std::vector<std::vector<int>> buildCsrGraph(const std::vector<int> & rowStart, const std::vector<int> & column, int n)
    {
        std::vector<std::vector<int>> graph(n);
        for (int r = 0; r < n; r++)
        {
            for (int j = rowStart[r]; j < rowStart[r + 1]; j++)
            {
                if (column[j] != r)
                {
                    graph[r].push_back(column[j]);
                }
            }
        }
        return graph;
    }

    std::vector<int> smallestLastOrdering(const std::vector<std::vector<int>> & graph)
    {
        // Lazy-deletion min-heap on degree.  The queue, its element type and
        // its comparator are the ones the reference uses, so ties between equal
        // degrees are broken identically by the standard library's heap.
        const int n = static_cast<int>(graph.size());
        struct E
        {
            int vertex;
            int degree;
            bool operator<(const E & other) const { return degree > other.degree; }
        };
        std::priority_queue<E> q;
        std::vector<int> degree(n);
        std::vector<char> removed(n, 0);
        for (int i = 0; i < n; i++)
        {
            degree[i] = static_cast<int>(graph[i].size());
            q.push({i, degree[i]});
        }
        std::vector<int> result(n);
        for (int i = 0; i < n; i++)
        {
            int v = -1;
            while (!q.empty())
            {
                const E top = q.top();
                q.pop();
                if (!removed[top.vertex])
                {
                    v = top.vertex;
                    break;
                }
            }
            if (v < 0)
            {
                throw std::runtime_error("smallestLastOrdering: heap exhausted");
            }
            result[i] = v;
            removed[v] = 1;
            for (int w : graph[v])
            {
                if (removed[w])
                {
                    continue;
                }
                degree[w]--;
                q.push({w, degree[w]});
            }
        }
        std::reverse(result.begin(), result.end());
        return result;
    }

    std::vector<std::vector<int>> partitionGraphGreedy(const std::vector<std::vector<int>> & graph, const std::vector<int> & order)
    {
        const int n = static_cast<int>(graph.size());
        std::vector<std::vector<int>> parts;
        if (n == 0)
        {
            return parts;
        }
        std::vector<int> color(n, -1);
        std::vector<char> used;
        color[order[0]] = 0;
        parts.push_back({order[0]});
        used.assign(1, 0);
        for (int k = 1; k < n; k++)
        {
            const int v = order[k];
            std::fill(used.begin(), used.end(), 0);
            for (int w : graph[v])
            {
                if (color[w] >= 0)
                {
                    used[color[w]] = 1;
                }
            }
            int lowest = 0;
            while (lowest < static_cast<int>(used.size()) && used[lowest])
            {
                lowest++;
            }
            color[v] = lowest;
            if (lowest == static_cast<int>(parts.size()))
            {
                parts.push_back({v});
                used.push_back(0);
            }
            else
            {
                parts[lowest].push_back(v);
            }
        }
        return parts;
    }

    bool verifyColoring(const std::vector<std::vector<int>> & graph, const std::vector<std::vector<int>> & parts)
    {
        const int n = static_cast<int>(graph.size());
        std::vector<int> color(n, -1);
        for (size_t c = 0; c < parts.size(); c++)
        {
            for (int v : parts[c])
            {
                if (v < 0 || v >= n || color[v] != -1)
                {
                    return false;
                }
                color[v] = static_cast<int>(c);
            }
        }
        for (int v = 0; v < n; v++)
        {
            if (color[v] < 0)
            {
                return false;
            }
            for (int w : graph[v])
            {
                if (color[w] == color[v])
                {
                    return false;
                }
            }
        }
        return true;
    }

    McgsLayout buildMcgsLayout(const CsrD & m)
    {
        if (m.rows != m.cols)
        {
            throw std::invalid_argument("Gauss-Seidel needs a square matrix");
        }
        const int n = m.rows;
        if (n < 1)
        {
            throw std::invalid_argument("Gauss-Seidel needs a non-empty matrix");
        }
        auto graph = buildCsrGraph(m.rowStart, m.column, n);
        auto order = smallestLastOrdering(graph);
        auto parts = partitionGraphGreedy(graph, order);

        McgsLayout L;
        L.n = n;
        L.coloring.resize(n);
        L.partitionStart.resize(parts.size() + 1);
        int pos = 0;
        for (size_t c = 0; c < parts.size(); c++)
        {
            auto & p = parts[c];
            std::sort(p.begin(), p.end());
            L.partitionStart[c] = pos;
            std::copy(p.begin(), p.end(), L.coloring.begin() + pos);
            pos += static_cast<int>(p.size());
        }
        L.partitionStart.back() = n;

        const CsrD reordered = m.slice(L.coloring, L.coloring);

        // Strip the diagonal (gaussSeidel.cpp:26-58): zero or missing diagonal
        // entries are errors.
        L.stripped.rows = n;
        L.stripped.cols = n;
        L.stripped.rowStart.resize(n + 1);
        L.diag.assign(n, 0.0);
        L.invDiag.assign(n, 0.0);
        L.rowSplit.resize(n);
        std::vector<int> colorOf(n);
        for (int c = 0; c < L.numColors(); c++)
        {
            for (int i = L.partitionStart[c]; i < L.partitionStart[c + 1]; i++)
            {
                colorOf[i] = c;
            }
        }
        for (int r = 0; r < n; r++)
        {
            L.stripped.rowStart[r] = static_cast<int>(L.stripped.column.size());
            bool found = false;
            const int lowerEnd = L.partitionStart[colorOf[r]];
            int split = -1;
            for (int j = reordered.rowStart[r]; j < reordered.rowStart[r + 1]; j++)
            {
                const int c = reordered.column[j];
                const double v = reordered.values[j];
                if (c == r)
                {
                    if (v == 0.0)
                    {
                        throw std::invalid_argument("Matrix contains zeros on the diagonal");
                    }
                    L.diag[r] = v;
                    L.invDiag[r] = 1.0 / v;
                    found = true;
                }
                else
                {
                    if (split < 0 && c >= lowerEnd)
                    {
                        split = static_cast<int>(L.stripped.column.size());
                    }
                    L.stripped.column.push_back(c);
                    L.stripped.values.push_back(v);
                }
            }
            if (!found)
            {
                throw std::invalid_argument("Matrix has missing diagonal elements!");
            }
            L.rowSplit[r] = split < 0 ? static_cast<int>(L.stripped.column.size()) : split;
        }
        L.stripped.rowStart[n] = static_cast<int>(L.stripped.column.size());
        return L;
    }
    std::vector<int> localityRank(const std::vector<int> & rowStart, const std::vector<int> & column, int n)
    {
        std::vector<int> degree(n);
        for (int i = 0; i < n; i++)
        {
            degree[i] = rowStart[i + 1] - rowStart[i];
        }
        std::vector<int> rank(n, -1), queue;
        queue.reserve(n);
        std::vector<int> mark(n, -1);
        std::vector<int> nbrs;
        // breadth-first from `start`, marking with `tag`; appends to `out`
        auto bfs = [&](int start, int tag, std::vector<int> & out)
        {
            size_t head = out.size();
            out.push_back(start);
            mark[start] = tag;
            while (head < out.size())
            {
                const int u = out[head++];
                nbrs.clear();
                for (int j = rowStart[u]; j < rowStart[u + 1]; j++)
                {
                    const int v = column[j];
                    if (v != u && mark[v] != tag)
                    {
                        mark[v] = tag;
                        nbrs.push_back(v);
                    }
                }
                std::sort(nbrs.begin(), nbrs.end(), [&](int a, int b) { return degree[a] != degree[b] ? degree[a] < degree[b] : a < b; });
                out.insert(out.end(), nbrs.begin(), nbrs.end());
            }
        };
        int next = 0;
        std::vector<int> probe;
        for (int seed = 0; seed < n; seed++)
        {
            if (rank[seed] >= 0)
            {
                continue;
            }
            // a far node of this component: the last one reached from the
            // component's minimum-degree node
            probe.clear();
            bfs(seed, -2 - seed, probe);
            int start = seed;
            for (int v : probe)
            {
                if (degree[v] < degree[start])
                {
                    start = v;
                }
            }
            probe.clear();
            bfs(start, -2 - n - seed, probe);
            start = probe.back();
            const size_t before = queue.size();
            bfs(start, n + seed, queue);
            for (size_t k = before; k < queue.size(); k++)
            {
                rank[queue[k]] = next++;
            }
        }
        return rank;
    }
} // namespace nsfem
