// The reference's binary vector / CSR format (cpp/libs/LinAlg/src/io.cpp:49-148):
// int32 dtypeId (0 = f32, 1 = f64), then for a vector int32 n and the data,
// for a CSR int32 rows, cols, nnz, then values, column, rowStart.  In
// namespace nsfem, so a translation unit may also include the reference's own
// linalg/io.h.
#pragma once

#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cu.hpp"

namespace nsfem
{
    template <typename F>
    constexpr int32_t dtypeId() { return sizeof(F) == 8 ? 1 : 0; }

    template <typename F>
    void writeVec(const std::string & fileName, const std::vector<F> & v)
    {
        std::ofstream f(fileName, std::ios::binary);
        if (!f.is_open())
        {
            throw std::runtime_error("Failed to open output file " + fileName);
        }
        const int32_t id = dtypeId<F>();
        const int32_t n = static_cast<int32_t>(v.size());
        f.write(reinterpret_cast<const char *>(&id), 4);
        f.write(reinterpret_cast<const char *>(&n), 4);
        f.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(F)));
    }

    template <typename F>
    std::vector<F> readVec(const std::string & fileName)
    {
        std::ifstream f(fileName, std::ios::binary);
        if (!f.is_open())
        {
            throw std::runtime_error("Failed to open input file " + fileName);
        }
        int32_t id = -1, n = 0;
        if (!f.read(reinterpret_cast<char *>(&id), 4) || id != dtypeId<F>())
        {
            throw std::runtime_error("Failed to read vector: Bad dtype");
        }
        if (!f.read(reinterpret_cast<char *>(&n), 4) || n < 1)
        {
            throw std::runtime_error("Failed to read vector: Bad size");
        }
        std::vector<F> v(n);
        if (!f.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(n * sizeof(F))))
        {
            throw std::runtime_error("Failed to read vector: Couldn't read data vector");
        }
        return v;
    }
    template <typename F>
    HostCsr<F> readCsr(const std::string & fileName)
    {
        std::ifstream f(fileName, std::ios::binary);
        if (!f.is_open())
        {
            throw std::runtime_error("Failed to open input file " + fileName);
        }
        int32_t id = -1, nnz = 0;
        HostCsr<F> m;
        if (!f.read(reinterpret_cast<char *>(&id), 4) || id != dtypeId<F>())
        {
            throw std::runtime_error("Failed to read CSR matrix: Bad dtype");
        }
        if (!f.read(reinterpret_cast<char *>(&m.rows), 4) || !f.read(reinterpret_cast<char *>(&m.cols), 4) ||
            !f.read(reinterpret_cast<char *>(&nnz), 4))
        {
            throw std::runtime_error("Failed to read CSR matrix: Couldn't read header");
        }
        if (m.rows < 1 || m.cols < 1 || nnz < 1)
        {
            throw std::runtime_error("Failed to read CSR matrix: Bad header values");
        }
        m.values.resize(nnz);
        m.column.resize(nnz);
        m.rowStart.resize(m.rows + 1);
        if (!f.read(reinterpret_cast<char *>(m.values.data()), static_cast<std::streamsize>(nnz * sizeof(F))) ||
            !f.read(reinterpret_cast<char *>(m.column.data()), static_cast<std::streamsize>(nnz * 4)) ||
            !f.read(reinterpret_cast<char *>(m.rowStart.data()), static_cast<std::streamsize>((m.rows + 1) * 4)))
        {
            throw std::runtime_error("Failed to read CSR matrix: Couldn't read data vectors");
        }
        if (m.rowStart.back() != nnz)
        {
            throw std::runtime_error("Failed to read CSR matrix: Bad end of rowStart");
        }
        return m;
    }
} // namespace nsfem
