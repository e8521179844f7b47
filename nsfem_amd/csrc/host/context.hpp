// Everything one Chorin time loop needs, built once on the host.
//
// Follows buildChorinContext (cpp/execs/NavierStokes/src/buildContext.cpp:17-142),
// extractDirichletNodes (include/NavierStokes/borders.h:12-47),
// extractInternalNodes (src/borders.cpp:19-52) and the DFG inflow profile
// (include/NavierStokes/dfgCondtions.h:9-12).
#pragma once

#include <string>
#include <vector>

#include "csr.hpp"
#include "fem.hpp"
#include "mesh.hpp"

namespace nsfem
{
    struct DirichletNode
    {
        int id;
        double value;
    };

    // Inflow u_x(y) = profileScale * peak * y (H - y) / H^2 on the "Left" border.
    // H = 0.41 and profileScale = 1 is the reference's DFG profile.
    struct DfgConditions
    {
        double viscosity = 1e-3;
        double peakVelocity = 1.0;
        double channelHeight = 0.41;
        double profileScale = 1.0;

        double leftVelocity(double y) const
        {
            return profileScale * peakVelocity * y * (channelHeight - y) / (channelHeight * channelHeight);
        }
    };

    struct ChorinContext
    {
        int numVelocityNodes = 0; // N, one component
        int numPressureNodes = 0; // V
        std::vector<DirichletNode> dirichletVx, dirichletVy;
        std::vector<int> internalVelocityNodes;
        std::vector<DirichletNode> dirichletPressure;
        std::vector<int> internalPressureNodes;

        CsrD velocityMass;
        CsrD velocityStiffness;
        CsrD pressureStiffness;
        CsrD pressureStiffnessInternal;
        CsrD velocityPressureDiv;
        CsrD pressureVelocityDiv;

        ConvectionElementData convection;
        double viscosity = 0;

        // Block-diagonal batches of independent instances: the solvers are
        // built per instance from these blocks (empty for one instance).
        std::vector<CsrD> massBlocks;
        std::vector<CsrD> pressureBlocks;
        std::vector<int> velocityOffsets; // per instance, into one velocity component
        std::vector<int> pressureOffsets;
        int numSystems() const { return massBlocks.empty() ? 1 : static_cast<int>(massBlocks.size()); }
    };

    std::vector<DirichletNode> extractDirichletNodes(const ConcreteMesh & mesh, const std::vector<int> & groupIds,
                                                     const DfgConditions * cond, bool inflowX);

    std::vector<int> extractInternalNodes(int numNodes, const std::vector<DirichletNode> & sorted);

    // gpuAssembly: the matrix values (and the convection slot lists) come
    // from assembleChorinMatricesGpu over host-built patterns; identical
    // values to the host assembly.
    ChorinContext buildChorinContext(const ConcreteMesh & velocityMesh, const ConcreteMesh & pressureMesh,
                                     const DfgConditions & cond, int nThreads, bool gpuAssembly = false);

    // Concatenates independent instances (different meshes and conditions)
    // into one block-diagonal context: velocity planar over the whole batch
    // ([ux of all instances; uy of all instances]), K pre-scaled by each
    // instance's viscosity (the batch viscosity is 1).
    ChorinContext concatenateContexts(const std::vector<ChorinContext> & parts);
} // namespace nsfem
