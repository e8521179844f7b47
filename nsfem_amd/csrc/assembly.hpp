// GPU assembly of the Chorin matrices (mass, stiffness, pressure stiffness,
// both divergence blocks) into host-built CSR patterns.
//
// Replaces the values part of buildChorinCsrMatrices
// (cpp/libs/Fem/src/chorinCsr.cpp:296-502); the patterns and the per-entry
// slot lists stay on the host (buildAssemblyTopology).  One thread per
// element evaluates every local matrix; one thread per CSR entry sums its
// slots in ascending element order, as the host assembly does, so with the
// element integrals shared and no FP contraction the values are
// bit-identical to the host's.
#pragma once

#include "fem.hpp"

namespace nsfem
{
    // Fills the values of out (patterns copied from topo).  Blocking.
    ChorinMatrices assembleChorinMatricesGpu(const ConcreteMesh & velocityMesh, const ConcreteMesh & pressureMesh,
                                             const AssemblyTopology & topo);
} // namespace nsfem
