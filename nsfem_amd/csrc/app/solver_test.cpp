// ./solver_test <algo> <matrix.bin> <rhs.bin> [maxIters] [target]
//
// The reference's SolverTest harness shape (cpp/tests/SolverTest/src/solverTest.cpp:254-340)
// over the cu:: host API: reads a linalg .bin CSR (float) and RHS, splits the
// RHS into channels of the matrix size, solves, prints the final RMS
// residual per channel ("Final MSE = ...") and writes the solution next to
// the RHS as <rhs>.sol.bin.  algo: gaussSeidelCuda (gs, 1 ch),
// gaussSeidelCuda2 (gs, 2 ch), pcgCuda (pcg, 1 ch).
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

#include "../../../include/nsfem/cu.hpp"
#include "../../../include/nsfem/linalg_io.hpp"

int main(int argc, char ** argv)
{
    if (argc < 4)
    {
        std::cerr << "./solver_test <algo> <matrix> <rhs> [maxIters] [target]\n";
        return 1;
    }
    try
    {
        const std::string algo = argv[1];
        const nsfem::HostCsr<float> m = nsfem::readCsr<float>(argv[2]);
        const auto rhs = nsfem::readVec<float>(argv[3]);
        const int maxIters = argc > 4 ? std::stoi(argv[4]) : 100;
        const float target = argc > 5 ? std::stof(argv[5]) : 1e-9f;
        if (m.rows != m.cols)
        {
            std::cerr << "Only square matrices supported for now\n";
            return 1;
        }
        const int n = m.rows;
        const int numCh = static_cast<int>(rhs.size()) / n;
        if (numCh * n != static_cast<int>(rhs.size()))
        {
            std::cerr << "Failed to determine number of channels\n";
            return 1;
        }
        std::string name;
        int solverCh = 1;
        if (algo == "gaussSeidelCuda")
        {
            name = "gs";
        }
        else if (algo == "gaussSeidelCuda2")
        {
            name = "gs";
            solverCh = 2;
        }
        else if (algo == "pcgCuda")
        {
            name = "pcg";
        }
        else
        {
            std::cerr << "No algorithm matching " << algo << "\n";
            return 1;
        }
        if (solverCh == 2 && numCh != 2)
        {
            std::cerr << "Impossible - input doesn't have 2 channels\n";
            return 1;
        }
        std::vector<float> x(rhs.size(), 0.0f);
        const int passes = solverCh == 2 ? 1 : numCh;
        auto solver = cu::createSolver(name, solverCh, m, maxIters, target, 1);
        for (int c = 0; c < passes; c++)
        {
            const size_t len = static_cast<size_t>(n) * solverCh;
            std::vector<float> b(rhs.begin() + c * len, rhs.begin() + (c + 1) * len), xc(len, 0.0f);
            solver->getRhs().upload(b);
            solver->getSol().upload(xc);
            solver->solve();
            solver->getSol().download(xc);
            std::copy(xc.begin(), xc.end(), x.begin() + c * len);
            for (int k = 0; k < solverCh; k++)
            {
                double err = 0;
                for (int r = 0; r < n; r++)
                {
                    double s = 0;
                    for (int j = m.rowStart[r]; j < m.rowStart[r + 1]; j++)
                    {
                        s += static_cast<double>(m.values[j]) * xc[k * n + m.column[j]];
                    }
                    const double d = s - b[k * n + r];
                    err += d * d;
                }
                std::cout << "Final MSE = " << std::sqrt(err / n) << " iterations = " << solver->getLastIterations() << "\n";
            }
        }
        nsfem::writeVec(std::string(argv[3]) + ".sol.bin", x);
    }
    catch (const std::exception & e)
    {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
