// Test driver: reads a split colour-major matrix (n, nnz, rowStart, rowSplit,
// column, values, diag as raw little-endian arrays), runs buildFsai and
// writes the factor's values and diagonal; exit code 3 on a rejected matrix.
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "fsai.hpp"

template <typename T>
static bool readAll(FILE * f, std::vector<T> & v)
{
    return fread(v.data(), sizeof(T), v.size(), f) == v.size();
}

int main(int argc, char ** argv)
{
    if (argc != 3)
    {
        return 2;
    }
    FILE * f = fopen(argv[1], "rb");
    int hdr[2];
    if (!f || fread(hdr, 4, 2, f) != 2)
    {
        return 2;
    }
    const int n = hdr[0], nnz = hdr[1];
    std::vector<int> rowStart(n + 1), rowSplit(n), column(nnz);
    std::vector<double> values(nnz), diag(n);
    if (!readAll(f, rowStart) || !readAll(f, rowSplit) || !readAll(f, column) || !readAll(f, values) || !readAll(f, diag))
    {
        return 2;
    }
    fclose(f);
    std::vector<double> g, gd;
    try
    {
        nsfem::buildFsai(n, rowStart, rowSplit, column, values, diag, g, gd);
    }
    catch (const std::invalid_argument & e)
    {
        fprintf(stderr, "%s\n", e.what());
        return 3;
    }
    FILE * o = fopen(argv[2], "wb");
    if (!o || fwrite(g.data(), 8, g.size(), o) != g.size() || fwrite(gd.data(), 8, gd.size(), o) != gd.size())
    {
        return 2;
    }
    fclose(o);
    return 0;
}
