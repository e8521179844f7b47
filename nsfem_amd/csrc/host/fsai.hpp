// Factorized sparse approximate inverse (FSAI) of a symmetric positive
// definite matrix for the PCG solvers: M^-1 ~ G^T G with G lower triangular
// (Kolotilina and Yeremin, "Factorized sparse approximate inverse
// preconditionings I. Theory", SIAM J. Matrix Anal. Appl. 14 (1993)).
//
// No counterpart in the reference (its PCG preconditioner is multicolour
// symmetric Gauss-Seidel, cpp/libs/CudaUtils/src/gaussSeidel.cu).  The matrix
// is given in the solvers' colour-major order, split into its strictly
// off-diagonal rows (the lower-colour entries of row i in [rowStart[i],
// rowSplit[i]), the higher-colour ones after them) and the diagonal.  Row i
// of G lives on P_i = the lower-colour columns of row i and i itself: the
// vector g solving M[P_i, P_i] g = e_i, scaled by 1 / sqrt(g_i) so that
// (G M G^T)_ii = 1.  Applying G needs the lower parts of the rows and G^T
// the higher-colour parts, so both share the matrix's column structure:
//   values[j], j in [rowStart[i], rowSplit[i])   = G(i, column[j])
//   values[j], j in [rowSplit[i], rowStart[i+1]) = G(column[j], i)
//   diag[i]                                      = G(i, i)
#pragma once

#include <vector>

namespace nsfem
{
    // The sparsity pattern must be symmetric (every j in row i's higher
    // part has i in its lower part); throws std::invalid_argument otherwise
    // or when a local system is not positive definite.
    void buildFsai(int n, const std::vector<int> & rowStart, const std::vector<int> & rowSplit, const std::vector<int> & column,
                   const std::vector<double> & values, const std::vector<double> & diag, std::vector<double> & gValues,
                   std::vector<double> & gDiag);
} // namespace nsfem
