#include "context.hpp"

#include <algorithm>
#include <stdexcept>

namespace nsfem
{
    std::vector<DirichletNode> extractDirichletNodes(const ConcreteMesh & mesh, const std::vector<int> & groupIds,
                                                     const DfgConditions * cond, bool inflowX)
    {
        // Border elements are visited in order and the first group that
        // claims a node decides its value (borders.h:23-45).
        std::vector<DirichletNode> result;
        const int stride = mesh.borderElementStride();
        const int elSize = mesh.ptsPerSide;
        const int leftId = mesh.findGroupId("Left");
        std::vector<char> seen(mesh.nodes.size(), 0);
        for (int i = 0; i < mesh.numBorderElements; i++)
        {
            const int * be = mesh.borderElements.data() + static_cast<size_t>(i) * stride;
            const int group = be[2];
            if (std::find(groupIds.begin(), groupIds.end(), group) == groupIds.end())
            {
                continue;
            }
            for (int k = 0; k < elSize; k++)
            {
                const int id = be[3 + k];
                if (seen[id])
                {
                    continue;
                }
                seen[id] = 1;
                double v = 0.0;
                if (inflowX && cond && group == leftId)
                {
                    v = cond->leftVelocity(mesh.nodes[id].y);
                }
                result.push_back(DirichletNode{id, v});
            }
        }
        std::sort(result.begin(), result.end(), [](const auto & a, const auto & b) { return a.id < b.id; });
        return result;
    }

    std::vector<int> extractInternalNodes(int numNodes, const std::vector<DirichletNode> & sorted)
    {
        std::vector<int> result;
        size_t j = 0;
        for (int i = 0; i < numNodes; i++)
        {
            while (j < sorted.size() && sorted[j].id < i)
            {
                j++;
            }
            if (j < sorted.size() && sorted[j].id == i)
            {
                continue;
            }
            result.push_back(i);
        }
        return result;
    }

    ChorinContext buildChorinContext(const ConcreteMesh & vm, const ConcreteMesh & pm, const DfgConditions & cond, int nThreads)
    {
        if (vm.numElements != pm.numElements || vm.groups.size() != pm.groups.size())
        {
            throw std::invalid_argument("Velocity and pressure meshes do not match");
        }
        auto need = [&](const char * name)
        {
            const int id = vm.findGroupId(name);
            if (id < 0)
            {
                throw std::invalid_argument(std::string("No ") + name + " border!");
            }
            return id;
        };
        const int idLeft = need("Left");
        const int idRight = need("Right");
        const int idTop = need("Top");
        const int idBottom = need("Bottom");
        const int idCircle = need("Circle");

        ChorinContext ctx;
        ctx.numVelocityNodes = static_cast<int>(vm.nodes.size());
        ctx.numPressureNodes = static_cast<int>(pm.nodes.size());
        ctx.viscosity = cond.viscosity;

        const std::vector<int> velocityBorders = {idLeft, idTop, idBottom, idCircle};
        ctx.dirichletVx = extractDirichletNodes(vm, velocityBorders, &cond, true);
        ctx.dirichletVy = extractDirichletNodes(vm, velocityBorders, nullptr, false);
        ctx.internalVelocityNodes = extractInternalNodes(ctx.numVelocityNodes, ctx.dirichletVx);
        ctx.dirichletPressure = extractDirichletNodes(pm, {idRight}, nullptr, false);
        ctx.internalPressureNodes = extractInternalNodes(ctx.numPressureNodes, ctx.dirichletPressure);
        if (ctx.internalPressureNodes.empty())
        {
            throw std::invalid_argument("No internal pressure nodes");
        }

        auto mats = buildChorinMatrices(vm, pm, nThreads);
        ctx.velocityMass = std::move(mats.velocityMass);
        ctx.velocityStiffness = std::move(mats.velocityStiffness);
        ctx.pressureStiffness = std::move(mats.pressureStiffness);
        ctx.velocityPressureDiv = std::move(mats.velocityPressureDiv);
        ctx.pressureVelocityDiv = std::move(mats.pressureVelocityDiv);
        ctx.pressureStiffnessInternal = ctx.pressureStiffness.slice(ctx.internalPressureNodes, ctx.internalPressureNodes);
        ctx.convection = buildConvectionElementData(vm, ctx.velocityStiffness);
        return ctx;
    }
} // namespace nsfem
