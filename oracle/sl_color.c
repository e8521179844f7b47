/*
 * ORACLE - test infrastructure only.  Not part of the product; only tests/,
 * the smoke check and the CPU baseline of bench.py load it.
 *
 * Smallest-last vertex ordering and greedy colouring, restated in C from
 *   cpp/libs/LinAlg/src/graphs.cpp:64-121  (buildSmallestLastOrdering)
 *   cpp/libs/LinAlg/src/graphs.cpp:123-176 (partitionGraphGreedy)
 *   cpp/libs/LinAlg/src/graphs.cpp:264-284 (buildCsrGraph)
 *
 * The reference keeps its lazy-deletion queue in std::priority_queue<E> with
 * operator< returning (degree > other.degree), so ties between equal degrees
 * are broken by the binary-heap algorithm of the C++ standard library it was
 * built with (libstdc++, GCC 11: std::push_heap / std::pop_heap in
 * <bits/stl_heap.h>).  That published algorithm is restated below
 * (sift-up on push; on pop, move the last element to a hole sifted from the
 * root down to a leaf along the larger child, then sift it up), so this file
 * reproduces the reference's order without using the library itself.
 */
#include <stdlib.h>
#include <string.h>

typedef struct
{
    int vertex;
    int degree;
} Entry;

/* std::less<E> with the reference's E::operator< */
static int entryLess(Entry a, Entry b) { return a.degree > b.degree; }

/* __push_heap: move the hole up while the parent is "less" than value */
static void siftUp(Entry * h, long hole, long top, Entry value)
{
    long parent = (hole - 1) / 2;
    while (hole > top && entryLess(h[parent], value))
    {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = value;
}

/* __adjust_heap over [0, len) with a hole at 'hole' to be filled by value */
static void adjustHeap(Entry * h, long hole, long len, Entry value)
{
    const long top = hole;
    long second = hole;
    while (second < (len - 1) / 2)
    {
        second = 2 * (second + 1);
        if (entryLess(h[second], h[second - 1]))
        {
            second--;
        }
        h[hole] = h[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2)
    {
        second = 2 * (second + 1);
        h[hole] = h[second - 1];
        hole = second - 1;
    }
    siftUp(h, hole, top, value);
}

static void heapPush(Entry * h, long * size, Entry e)
{
    h[*size] = e;
    (*size)++;
    siftUp(h, *size - 1, 0, h[*size - 1]);
}

/* top() then pop(): pop_heap moves front to the back and re-heapifies */
static Entry heapPop(Entry * h, long * size)
{
    Entry topEntry = h[0];
    if (*size > 1)
    {
        Entry value = h[*size - 1];
        h[*size - 1] = h[0];
        adjustHeap(h, 0, *size - 1, value);
    }
    (*size)--;
    return topEntry;
}

/*
 * rows, rowStart, column: square CSR pattern (diagonal entries are skipped).
 * order: out, the smallest-last order (after the final reversal).
 * color: out, colour of each vertex.  Returns the number of colours, or -1
 * on allocation failure.
 */
int oracle_sl_greedy(int n, const int * rowStart, const int * column, int * order, int * color)
{
    if (n <= 0)
    {
        return 0;
    }
    const long nnz = rowStart[n];
    Entry * heap = (Entry *)malloc(sizeof(Entry) * (size_t)(n + nnz + 1));
    int * degree = (int *)malloc(sizeof(int) * (size_t)n);
    char * removed = (char *)calloc((size_t)n, 1);
    char * used = (char *)calloc((size_t)n + 1, 1);
    if (!heap || !degree || !removed || !used)
    {
        free(heap);
        free(degree);
        free(removed);
        free(used);
        return -1;
    }
    long size = 0;
    for (int i = 0; i < n; i++)
    {
        int d = 0;
        for (int j = rowStart[i]; j < rowStart[i + 1]; j++)
        {
            if (column[j] != i)
            {
                d++;
            }
        }
        degree[i] = d;
        Entry e = {i, d};
        heapPush(heap, &size, e);
    }
    for (int k = 0; k < n; k++)
    {
        int v = -1;
        while (size > 0)
        {
            Entry e = heapPop(heap, &size);
            if (!removed[e.vertex])
            {
                v = e.vertex;
                break;
            }
        }
        order[k] = v;
        removed[v] = 1;
        for (int j = rowStart[v]; j < rowStart[v + 1]; j++)
        {
            const int w = column[j];
            if (w == v || removed[w])
            {
                continue;
            }
            degree[w]--;
            Entry e = {w, degree[w]};
            heapPush(heap, &size, e);
        }
    }
    /* reverse */
    for (int a = 0, b = n - 1; a < b; a++, b--)
    {
        int t = order[a];
        order[a] = order[b];
        order[b] = t;
    }
    /* greedy: lowest colour not used by an already coloured neighbour */
    for (int i = 0; i < n; i++)
    {
        color[i] = -1;
    }
    int numColors = 0;
    for (int k = 0; k < n; k++)
    {
        const int v = order[k];
        memset(used, 0, (size_t)numColors + 1);
        for (int j = rowStart[v]; j < rowStart[v + 1]; j++)
        {
            const int w = column[j];
            if (w != v && color[w] >= 0)
            {
                used[color[w]] = 1;
            }
        }
        int c = 0;
        while (c < numColors && used[c])
        {
            c++;
        }
        color[v] = c;
        if (c == numColors)
        {
            numColors++;
        }
    }
    free(heap);
    free(degree);
    free(removed);
    free(used);
    return numColors;
}

/* Sequential Gauss-Seidel sweeps in a given row order (the reference's
 * gaussSeidelCustomOrder, cpp/libs/LinAlg/src/gaussSeidel.cpp:67-92, in
 * FP64 with x_i = (b_i - sum_{j != i} a_ij x_j) / a_ii). */
void oracle_gs_custom_order(int n, const int * rowStart, const int * column, const double * values, double * x, const double * b,
                            const int * order, int sweeps)
{
    for (int s = 0; s < sweeps; s++)
    {
        for (int k = 0; k < n; k++)
        {
            const int i = order[k];
            double diag = 0, negSum = 0;
            for (int j = rowStart[i]; j < rowStart[i + 1]; j++)
            {
                if (column[j] == i)
                {
                    diag = values[j];
                }
                else
                {
                    negSum += values[j] * x[column[j]];
                }
            }
            x[i] = (b[i] - negSum) / diag;
        }
    }
}
