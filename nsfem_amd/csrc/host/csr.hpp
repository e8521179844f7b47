// Host CSR matrix, symmetric slicing and multicolour Gauss-Seidel set-up.
//
// Mirrors the parts of the reference's linalg library that feed the GPU path:
//   CsrMatrix::slice            cpp/libs/LinAlg/src/csrMatrix.cpp:12-91
//   findOffsetsUnsorted         cpp/libs/LinAlg/src/csrMatrix.cpp:133-159
//   buildCsrGraph               cpp/libs/LinAlg/src/graphs.cpp:264-284
//   buildSmallestLastOrdering   cpp/libs/LinAlg/src/graphs.cpp:64-121
//   partitionGraphGreedy        cpp/libs/LinAlg/src/graphs.cpp:123-176
//   verifyColoring              cpp/libs/LinAlg/src/graphs.cpp:11-62
//   GaussSeidel colouring set-up cpp/libs/CudaUtils/src/gaussSeidel.cu:140-182
//   buildGaussSeidelContext     cpp/libs/LinAlg/src/gaussSeidel.cpp:9-65
#pragma once

#include <cstdint>
#include <vector>

namespace nsfem
{
    template <typename T>
    struct Csr
    {
        int rows = 0;
        int cols = 0;
        std::vector<int> rowStart; // rows + 1
        std::vector<int> column;   // nnz, sorted within each row
        std::vector<T> values;     // nnz

        int nnz() const { return static_cast<int>(column.size()); }

        // Row i of the result is row rowIds[i] of this matrix; columns are
        // renumbered through colIds (unlisted columns dropped) and re-sorted.
        Csr<T> slice(const std::vector<int> & rowIds, const std::vector<int> & colIds) const;

        // Position of (row, col) in values, or -1.  Columns must be sorted.
        int find(int row, int col) const;
    };

    using CsrD = Csr<double>;

    std::vector<std::vector<int>> buildCsrGraph(const std::vector<int> & rowStart, const std::vector<int> & column, int n);

    std::vector<int> smallestLastOrdering(const std::vector<std::vector<int>> & graph);

    std::vector<std::vector<int>> partitionGraphGreedy(const std::vector<std::vector<int>> & graph, const std::vector<int> & order);

    bool verifyColoring(const std::vector<std::vector<int>> & graph, const std::vector<std::vector<int>> & parts);

    // The multicolour layout of one square matrix: the colour permutation
    // (coloring[newIndex] = oldIndex), colour boundaries, and the permuted
    // matrix split into strictly off-diagonal part + diagonal.
    struct McgsLayout
    {
        int n = 0;
        std::vector<int> coloring;       // n
        std::vector<int> partitionStart; // numColors + 1
        CsrD stripped;                   // permuted, diagonal removed
        std::vector<double> diag;        // permuted diagonal
        std::vector<double> invDiag;     // 1 / diag
        // First position in each stripped row whose column is not of a lower colour.
        std::vector<int> rowSplit;
        int numColors() const { return static_cast<int>(partitionStart.size()) - 1; }
    };

    McgsLayout buildMcgsLayout(const CsrD & m);

    // Locality rank of every row of a square pattern: its position in a
    // breadth-first (Cuthill-McKee) order of the row graph, each component
    // started from a far node (the last one reached from a minimum-degree
    // node) and neighbours taken by increasing degree.  Rows that are close
    // in the graph get close ranks.
    std::vector<int> localityRank(const std::vector<int> & rowStart, const std::vector<int> & column, int n);
} // namespace nsfem
