// Smoothed-aggregation multigrid hierarchy (host set-up) for the symmetric
// positive definite systems of the pressure solve (P1 Laplacians with
// Dirichlet rows removed).
//
// No counterpart in the reference, which solves the pressure Poisson system
// with multicolour Gauss-Seidel (cpp/libs/CudaUtils/src/gaussSeidel.cu) or a
// direct cuDSS factorisation (cpp/execs/NavierStokes/src/chorinCuda.cu).
// Here the multicolour Gauss-Seidel sweep becomes the smoother of every level
// of a V-cycle that preconditions CG.  Construction (Vanek, Mandel, Brezina,
// "Algebraic multigrid by smoothed aggregation", Computing 56, 1996):
//   strength     |a_ij| >= theta_l sqrt(|a_ii a_jj|), theta_l = theta 2^-l
//   aggregation  the three greedy passes of the paper (roots whose strong
//                neighbourhood is free; attach the rest to a neighbouring
//                aggregate; leftovers form aggregates of their own)
//   prolongation P = (I - omega D^-1 A) P_tent, omega = (4/3) / lambda_max(D^-1 A)
//   coarse op    A_c = P^T A P (symmetrised against rounding)
// Every level is kept in its own multicolour (colour-major) order from the
// smallest-last greedy colouring, so the GPU smooths each level with one
// multicolour sweep per colour; the coarsest is inverted densely.
#pragma once

#include <vector>

#include "csr.hpp"

namespace nsfem
{
    struct AmgOptions
    {
        double theta = 0.08;     // strength threshold of the finest level (halved per level)
        int maxCoarse = 512;     // the coarsest level has at most this many rows
        int maxLevels = 12;
        double omegaScale = 4.0 / 3.0;
    };

    struct AmgLevel
    {
        int n = 0;
        // colour-major order: strictly off-diagonal part with the lower-
        // colour entries of each row first (rowSplit ends them), diagonal
        CsrD stripped;
        std::vector<int> rowSplit;
        std::vector<double> diag, invDiag;
        std::vector<int> colorStart; // numColors + 1
        // upper bound of the spectrum of D^-1 A for a polynomial smoother:
        // min(1.1 x the power-iteration estimate, Gershgorin bound); set on
        // every level but the coarsest
        double lambdaMax = 0;
        // to / from the next coarser level (empty on the coarsest)
        CsrD P; // n x nNext
        CsrD R; // nNext x n, P^T
        int numColors() const { return static_cast<int>(colorStart.size()) - 1; }
    };

    struct AmgHierarchy
    {
        std::vector<AmgLevel> levels;
        std::vector<double> coarseInverse; // coarsest n x n, row-major
        double operatorComplexity = 0;     // sum of nnz over the levels / nnz of the finest
        // Zero row sums on the finest level (a pure-Neumann Laplacian): the
        // constants span the null space on every level (P maps constants to
        // constants), and the coarsest "inverse" is that of A_c + s 1 1^T
        // (s = mean diagonal / n_c), which solves A_c x = b with 1^T x = 0 for
        // every b with 1^T b = 0.
        bool singular = false;
    };

    // Level 0 is given in the caller's colour-major order (stripped
    // off-diagonal rows with rowSplit, the diagonal, colour boundaries) and
    // kept as it is.  Throws std::invalid_argument when a level is not
    // positive (semi)definite: the coarsest factorisation meets a pivot that
    // has lost all but rounding noise of its diagonal (for a singular matrix
    // without zero row sums).
    AmgHierarchy buildAmgHierarchy(const CsrD & stripped, const std::vector<int> & rowSplit, const std::vector<double> & diag,
                                   const std::vector<int> & colorStart, const AmgOptions & opt = AmgOptions{});
} // namespace nsfem
