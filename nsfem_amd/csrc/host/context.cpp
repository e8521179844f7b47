#include "context.hpp"

#include "../assembly.hpp"

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

    ChorinContext buildChorinContext(const ConcreteMesh & vm, const ConcreteMesh & pm, const DfgConditions & cond, int nThreads,
                                     bool gpuAssembly)
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

        if (gpuAssembly)
        {
            AssemblyTopology topo = buildAssemblyTopology(vm, pm, nThreads);
            auto mats = assembleChorinMatricesGpu(vm, pm, topo);
            ctx.velocityMass = std::move(mats.velocityMass);
            ctx.velocityStiffness = std::move(mats.velocityStiffness);
            ctx.pressureStiffness = std::move(mats.pressureStiffness);
            ctx.velocityPressureDiv = std::move(mats.velocityPressureDiv);
            ctx.pressureVelocityDiv = std::move(mats.pressureVelocityDiv);
            ctx.convection = buildConvectionElementData(vm, ctx.velocityStiffness, &topo.velocitySlots);
        }
        else
        {
            auto mats = buildChorinMatrices(vm, pm, nThreads);
            ctx.velocityMass = std::move(mats.velocityMass);
            ctx.velocityStiffness = std::move(mats.velocityStiffness);
            ctx.pressureStiffness = std::move(mats.pressureStiffness);
            ctx.velocityPressureDiv = std::move(mats.velocityPressureDiv);
            ctx.pressureVelocityDiv = std::move(mats.pressureVelocityDiv);
            ctx.convection = buildConvectionElementData(vm, ctx.velocityStiffness);
        }
        ctx.pressureStiffnessInternal = ctx.pressureStiffness.slice(ctx.internalPressureNodes, ctx.internalPressureNodes);
        return ctx;
    }

    namespace
    {
        // Appends rows of m to out with row / column index maps.
        template <typename RowMap, typename ColMap>
        void appendRows(CsrD & out, const CsrD & m, RowMap rowOf, ColMap colOf, std::vector<std::vector<std::pair<int, double>>> & rows)
        {
            for (int r = 0; r < m.rows; r++)
            {
                auto & dst = rows[rowOf(r)];
                for (int j = m.rowStart[r]; j < m.rowStart[r + 1]; j++)
                {
                    dst.emplace_back(colOf(m.column[j]), m.values[j]);
                }
            }
            (void)out;
        }

        CsrD fromRows(int rows, int cols, std::vector<std::vector<std::pair<int, double>>> & lists)
        {
            CsrD m;
            m.rows = rows;
            m.cols = cols;
            m.rowStart.assign(rows + 1, 0);
            size_t nnz = 0;
            for (auto & l : lists)
            {
                nnz += l.size();
            }
            m.column.reserve(nnz);
            m.values.reserve(nnz);
            for (int r = 0; r < rows; r++)
            {
                for (const auto & e : lists[r])
                {
                    m.column.push_back(e.first);
                    m.values.push_back(e.second);
                }
                m.rowStart[r + 1] = static_cast<int>(m.column.size());
                lists[r].clear();
                lists[r].shrink_to_fit();
            }
            return m;
        }
    } // namespace

    ChorinContext concatenateContexts(const std::vector<ChorinContext> & parts)
    {
        if (parts.empty())
        {
            throw std::invalid_argument("empty batch");
        }
        ChorinContext b;
        const int S = static_cast<int>(parts.size());
        std::vector<int> no(S), vo(S), io(S), eo(S);
        long long nTot = 0, vTot = 0, iTot = 0, eTot = 0;
        for (int s = 0; s < S; s++)
        {
            no[s] = static_cast<int>(nTot);
            vo[s] = static_cast<int>(vTot);
            io[s] = static_cast<int>(iTot);
            eo[s] = static_cast<int>(eTot);
            nTot += parts[s].numVelocityNodes;
            vTot += parts[s].numPressureNodes;
            iTot += static_cast<long long>(parts[s].internalPressureNodes.size());
            eTot += parts[s].convection.numElements;
        }
        if (2 * nTot > 0x7fffffffLL || 36 * eTot > 0x7fffffffLL)
        {
            throw std::invalid_argument("batch too large for 32-bit indices");
        }
        const int N = static_cast<int>(nTot), V = static_cast<int>(vTot), I = static_cast<int>(iTot);
        b.numVelocityNodes = N;
        b.numPressureNodes = V;
        b.viscosity = 1.0;
        b.velocityOffsets = no;
        b.pressureOffsets = vo;
        for (int s = 0; s < S; s++)
        {
            const auto & p = parts[s];
            for (const auto & d : p.dirichletVx) b.dirichletVx.push_back({d.id + no[s], d.value});
            for (const auto & d : p.dirichletVy) b.dirichletVy.push_back({d.id + no[s], d.value});
            for (const auto & d : p.dirichletPressure) b.dirichletPressure.push_back({d.id + vo[s], d.value});
            for (int id : p.internalVelocityNodes) b.internalVelocityNodes.push_back(id + no[s]);
            for (int id : p.internalPressureNodes) b.internalPressureNodes.push_back(id + vo[s]);
            b.massBlocks.push_back(p.velocityMass);
            b.pressureBlocks.push_back(p.pressureStiffnessInternal);
        }
        {
            std::vector<std::vector<std::pair<int, double>>> rows(N);
            CsrD dummy;
            for (int s = 0; s < S; s++)
            {
                appendRows(dummy, parts[s].velocityMass, [&](int r) { return r + no[s]; }, [&](int c) { return c + no[s]; }, rows);
            }
            b.velocityMass = fromRows(N, N, rows);
            rows.assign(N, {});
            for (int s = 0; s < S; s++)
            {
                CsrD k = parts[s].velocityStiffness;
                for (auto & v : k.values) v *= parts[s].viscosity;
                appendRows(dummy, k, [&](int r) { return r + no[s]; }, [&](int c) { return c + no[s]; }, rows);
            }
            b.velocityStiffness = fromRows(N, N, rows);
            rows.assign(V, {});
            for (int s = 0; s < S; s++)
            {
                appendRows(dummy, parts[s].pressureStiffness, [&](int r) { return r + vo[s]; }, [&](int c) { return c + vo[s]; }, rows);
            }
            b.pressureStiffness = fromRows(V, V, rows);
            rows.assign(I, {});
            for (int s = 0; s < S; s++)
            {
                appendRows(dummy, parts[s].pressureStiffnessInternal, [&](int r) { return r + io[s]; }, [&](int c) { return c + io[s]; }, rows);
            }
            b.pressureStiffnessInternal = fromRows(I, I, rows);
            rows.assign(V, {});
            for (int s = 0; s < S; s++)
            {
                const int n = parts[s].numVelocityNodes;
                appendRows(dummy, parts[s].velocityPressureDiv, [&](int r) { return r + vo[s]; },
                           [&](int c) { return c < n ? c + no[s] : N + no[s] + (c - n); }, rows);
            }
            b.velocityPressureDiv = fromRows(V, 2 * N, rows);
            rows.assign(2 * static_cast<size_t>(N), {});
            for (int s = 0; s < S; s++)
            {
                const int n = parts[s].numVelocityNodes;
                appendRows(dummy, parts[s].pressureVelocityDiv, [&](int r) { return r < n ? r + no[s] : N + no[s] + (r - n); },
                           [&](int c) { return c + vo[s]; }, rows);
            }
            b.pressureVelocityDiv = fromRows(2 * N, V, rows);
        }
        // Convection: node ids and element slots shifted per instance.
        auto & c = b.convection;
        c.numElements = static_cast<int>(eTot);
        c.contribStart.assign(1, 0);
        for (int s = 0; s < S; s++)
        {
            const auto & pc = parts[s].convection;
            for (int id : pc.elementNodes) c.elementNodes.push_back(id + no[s]);
            c.geometry.insert(c.geometry.end(), pc.geometry.begin(), pc.geometry.end());
            const int base = c.contribStart.back();
            for (size_t j = 1; j < pc.contribStart.size(); j++) c.contribStart.push_back(base + pc.contribStart[j]);
            for (int k : pc.contribIndex) c.contribIndex.push_back(k + 36 * eo[s]);
        }
        return b;
    }
} // namespace nsfem
