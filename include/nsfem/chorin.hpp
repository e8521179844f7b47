// The reference's top entry, solveNsChorinCuda
// (cpp/execs/NavierStokes/include/NavierStokes/chorinCuda.h:10-12, the loop of
// src/chorinCuda.cu:154-379), over the C ABI (include/nsfem_hip.h).
//
// nsfem::solveNsChorin<Solution>(velocityMesh, pressureMesh, cond, tau0, maxT,
// cfg) takes the caller's own types, so the reference's chorinCuda.cu can be
// replaced by a file that forwards to it unchanged:
//
//   Solution solveNsChorinCuda(const mesh::ConcreteMesh & v, const mesh::ConcreteMesh & p,
//                              const DfgConditions & cond, const float tau0, const float maxT,
//                              const ChorinCudaConfig & cfg)
//   {
//       return nsfem::solveNsChorin<Solution>(v, p, cond, tau0, maxT, cfg);
//   }
//
// Requirements on the types (the data members the reference's own types have):
//   mesh:      nodes[i].x / .y, elements (nodes per element x numElements
//              ints), numElements, borderElements ((3 + points per side) x
//              numBorderElements ints: triangle, side, group, points),
//              numBorderElements, groups (names) - mesh::ConcreteMesh,
//              cpp/libs/Mesh/include/mesh/concreteMesh.h:11-61;
//   cond:      viscosity, peakVelocity - DfgConditions, dfgCondtions.h:4-13;
//   cfg:       velocitySolver / pressureSolver with method, maxIterations,
//              targetMse, mseCheckInterval - chorinCudaConfig.h:6-19;
//   Solution:  steps, a vector of {time, velocity [ux; uy], pressure} with
//              float (or double) vectors - solution.h:6-17.
//
// The pressure (P1) mesh gives the triangle mesh; the library numbers the P2
// velocity nodes itself with the reference's createMesh rules and checks that
// the result equals velocityMesh (std::runtime_error otherwise).  Steps:
// tau = maxT / ceil(maxT / tau0), numTimeSteps + 1 steps from u = 0 plus the
// Dirichlet values, as the reference; the state after each step is kept.
// Options beyond the reference: the state precision (FP64 by default; the
// reference is FP32), the time-step policy (the explicit scheme is stable
// only for tau below a limit: rejected by default, or clamped - tau0 is then
// lowered to the limit first so the run still ends at maxT - or unchecked,
// the reference's behaviour) and which states to keep.
#pragma once

#include <cmath>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cu.hpp"

namespace nsfem
{
    template <typename M>
    concept ConcreteMeshLike = requires(const M & m) {
        { m.nodes[0].x } -> std::convertible_to<double>;
        { m.nodes[0].y } -> std::convertible_to<double>;
        { m.nodes.size() } -> std::convertible_to<size_t>;
        { m.elements.data() } -> std::convertible_to<const int *>;
        { m.numElements } -> std::convertible_to<int>;
        { m.borderElements.data() } -> std::convertible_to<const int *>;
        { m.numBorderElements } -> std::convertible_to<int>;
        { m.groups.size() } -> std::convertible_to<size_t>;
        { std::string(m.groups[0]) };
    };

    struct ChorinOptions
    {
        int dtype = NSFEM_F64;                    // NSFEM_F64 (plain) or NSFEM_F32 (the reference's precision)
        int timeStepPolicy = NSFEM_TAU_REJECT;    // NSFEM_TAU_REJECT, NSFEM_TAU_CLAMP, NSFEM_TAU_UNCHECKED
        int frameStep = 1;                        // keep every frameStep-th state (1: all, as the reference)
        double channelHeight = 0.41;              // inflow profile scale * peak * y (H - y) / H^2
        double profileScale = 1.0;                // (dfgCondtions.h:9-12 has H = 0.41, scale 1)
        int autotuneReps = 10;                    // nsfem_chorin_autotune before the first step (0: default shapes)
    };

    namespace detail
    {
        template <auto Destroy>
        struct HandleDeleter
        {
            template <typename H>
            void operator()(H * h) const noexcept
            {
                Destroy(h);
            }
        };
    } // namespace detail

    template <typename SolutionT, ConcreteMeshLike MeshT, typename CondT, typename CfgT>
    SolutionT solveNsChorin(const MeshT & velocityMesh, const MeshT & pressureMesh, const CondT & cond, const float timeStep0,
                            const float maxT, const CfgT & cfg, const ChorinOptions & opt = {})
    {
        using cu::check;
        if (!(timeStep0 > 0) || !(maxT > 0))
        {
            throw std::invalid_argument("time step and maxT must be positive");
        }
        if (opt.frameStep < 1)
        {
            throw std::invalid_argument("frameStep must be at least 1");
        }
        // Triangle mesh from the P1 mesh: vertices, triangles, border edges
        // (from, to, group) in the orientation of their side.
        const int V = static_cast<int>(pressureMesh.nodes.size());
        const int E = pressureMesh.numElements;
        const int nb = pressureMesh.numBorderElements;
        if (static_cast<size_t>(E) * 3 != pressureMesh.elements.size())
        {
            throw std::invalid_argument("pressure mesh must be a P1 mesh (3 nodes per element)");
        }
        const size_t bstep = nb > 0 ? pressureMesh.borderElements.size() / static_cast<size_t>(nb) : 5;
        if (bstep != 5)
        {
            throw std::invalid_argument("pressure mesh border elements must have 2 points per side");
        }
        std::vector<double> xy(2 * static_cast<size_t>(V));
        for (int i = 0; i < V; i++)
        {
            xy[2 * i] = pressureMesh.nodes[i].x;
            xy[2 * i + 1] = pressureMesh.nodes[i].y;
        }
        std::vector<int> edges(3 * static_cast<size_t>(nb));
        for (int b = 0; b < nb; b++)
        {
            const int * be = pressureMesh.borderElements.data() + 5 * static_cast<size_t>(b);
            edges[3 * b] = be[3];
            edges[3 * b + 1] = be[4];
            edges[3 * b + 2] = be[2];
        }
        std::vector<std::string> names;
        std::vector<const char *> namePtrs;
        for (size_t g = 0; g < pressureMesh.groups.size(); g++)
        {
            names.emplace_back(pressureMesh.groups[g]);
        }
        for (const auto & n : names)
        {
            namePtrs.push_back(n.c_str());
        }

        std::unique_ptr<nsfem_mesh, detail::HandleDeleter<nsfem_mesh_destroy>> mesh;
        {
            nsfem_mesh * h = nullptr;
            check(nsfem_mesh_from_arrays(V, xy.data(), E, pressureMesh.elements.data(), nb, edges.data(), static_cast<int>(names.size()),
                                         namePtrs.data(), &h));
            mesh.reset(h);
        }
        const nsfem_dfg_conditions c{static_cast<double>(cond.viscosity), static_cast<double>(cond.peakVelocity), opt.channelHeight,
                                     opt.profileScale};
        std::unique_ptr<nsfem_context, detail::HandleDeleter<nsfem_context_destroy>> ctx;
        {
            nsfem_context * h = nullptr;
            check(nsfem_context_build(mesh.get(), &c, 0, &h));
            ctx.reset(h);
        }
        int N = 0, Vc = 0, numDir = 0, numInt = 0, numEl = 0;
        check(nsfem_context_sizes(ctx.get(), &N, &Vc, &numDir, &numInt, &numEl));
        if (static_cast<size_t>(N) != velocityMesh.nodes.size() || Vc != V ||
            velocityMesh.elements.size() != 6 * static_cast<size_t>(numEl))
        {
            throw std::runtime_error("velocity mesh does not match the P2 numbering of the pressure mesh");
        }
        {
            std::vector<double> vxy(2 * static_cast<size_t>(N));
            std::vector<int> vel(6 * static_cast<size_t>(numEl));
            check(nsfem_context_velocity_mesh(ctx.get(), vxy.data(), vel.data()));
            for (size_t k = 0; k < vel.size(); k++)
            {
                if (vel[k] != velocityMesh.elements[k])
                {
                    throw std::runtime_error("velocity mesh does not match the P2 numbering of the pressure mesh");
                }
            }
        }

        const std::string vm = cfg.velocitySolver.method, pm = cfg.pressureSolver.method;
        nsfem_chorin_config cc{};
        cc.velocitySolver.method = vm.c_str();
        cc.velocitySolver.maxIterations = cfg.velocitySolver.maxIterations;
        cc.velocitySolver.targetMse = cfg.velocitySolver.targetMse;
        cc.velocitySolver.mseCheckInterval = cfg.velocitySolver.mseCheckInterval;
        cc.pressureSolver.method = pm.c_str();
        cc.pressureSolver.maxIterations = cfg.pressureSolver.maxIterations;
        cc.pressureSolver.targetMse = cfg.pressureSolver.targetMse;
        cc.pressureSolver.mseCheckInterval = cfg.pressureSolver.mseCheckInterval;
        cc.dtype = opt.dtype;
        cc.timeStepPolicy = opt.timeStepPolicy;
        std::unique_ptr<nsfem_chorin, detail::HandleDeleter<nsfem_chorin_destroy>> ch;
        {
            nsfem_chorin * h = nullptr;
            check(nsfem_chorin_create(ctx.get(), &cc, &h));
            ch.reset(h);
        }
        if (opt.autotuneReps > 0)
        {
            check(nsfem_chorin_autotune(ch.get(), opt.autotuneReps, nullptr));
        }

        double tau0 = timeStep0;
        if (opt.timeStepPolicy == NSFEM_TAU_CLAMP)
        {
            double tauMax = 0;
            check(nsfem_chorin_time_step_info(ch.get(), nullptr, nullptr, nullptr, &tauMax));
            tau0 = tau0 < tauMax ? tau0 : tauMax;
        }
        const int numTimeSteps = static_cast<int>(std::ceil(static_cast<double>(maxT) / tau0));
        const double tau = static_cast<double>(maxT) / numTimeSteps;
        check(nsfem_chorin_set_time_step(ch.get(), tau));

        using StepT = typename decltype(SolutionT{}.steps)::value_type;
        using VelT = typename decltype(StepT{}.velocity)::value_type;
        using PresT = typename decltype(StepT{}.pressure)::value_type;
        SolutionT result;
        std::vector<double> u64(2 * static_cast<size_t>(N)), p64(V);
        std::vector<float> u32(opt.dtype == NSFEM_F32 ? u64.size() : 0), p32(opt.dtype == NSFEM_F32 ? p64.size() : 0);
        for (int iT = 0; iT <= numTimeSteps; iT++)
        {
            check(nsfem_chorin_step(ch.get(), nullptr));
            if (iT % opt.frameStep != 0)
            {
                continue;
            }
            StepT s{};
            s.time = static_cast<decltype(s.time)>((iT + 1) * tau);
            s.velocity.resize(u64.size());
            s.pressure.resize(p64.size());
            if (opt.dtype == NSFEM_F32)
            {
                check(nsfem_chorin_download(ch.get(), u32.data(), p32.data()));
                for (size_t k = 0; k < u32.size(); k++) s.velocity[k] = static_cast<VelT>(u32[k]);
                for (size_t k = 0; k < p32.size(); k++) s.pressure[k] = static_cast<PresT>(p32[k]);
            }
            else
            {
                check(nsfem_chorin_download(ch.get(), u64.data(), p64.data()));
                for (size_t k = 0; k < u64.size(); k++) s.velocity[k] = static_cast<VelT>(u64[k]);
                for (size_t k = 0; k < p64.size(); k++) s.pressure[k] = static_cast<PresT>(p64[k]);
            }
            result.steps.push_back(std::move(s));
        }
        return result;
    }
} // namespace nsfem
