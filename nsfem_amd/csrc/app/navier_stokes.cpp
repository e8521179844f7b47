// ./navier_stokes <config> <msh file> <output dir>
//
// Host program of the GPU Chorin path with the reference's command line and
// config keys (cpp/execs/NavierStokes/src/navierStokes.cpp:20-83,
// src/nsConfig.cpp:8-45, utils ConfigParser key = value with '#' comments).
// Frames are written as linalg .bin vectors (velocity_<j>.bin = [ux; uy],
// pressure_<j>.bin) instead of rendered images.  Extra keys: hip.dtype
// (f64 | f32), hip.innerTol (pcg_mixed), hip.timeStepPolicy, hip.pressureStart,
// hip.velocityStart, hip.velocityPreconditioner (gs | fsai), hip.autotune.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "../../../include/nsfem/cu.hpp"
#include "../../../include/nsfem/linalg_io.hpp"

namespace
{
    std::string trim(const std::string & s)
    {
        const auto a = s.find_first_not_of(" \t\r\n");
        if (a == std::string::npos)
        {
            return "";
        }
        const auto b = s.find_last_not_of(" \t\r\n");
        return s.substr(a, b - a + 1);
    }

    std::map<std::string, std::string> parseConfig(const std::string & fileName)
    {
        std::ifstream f(fileName);
        if (!f.is_open())
        {
            throw std::runtime_error("Failed to open config file");
        }
        std::map<std::string, std::string> m;
        std::string line;
        while (std::getline(f, line))
        {
            if (line.empty() || line[0] == '#')
            {
                continue;
            }
            const auto eq = line.find('=');
            if (eq == std::string::npos || line.find('=', eq + 1) != std::string::npos)
            {
                continue;
            }
            const std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
            if (!k.empty() && !v.empty())
            {
                m[k] = v;
            }
        }
        return m;
    }

    struct Cfg
    {
        std::map<std::string, std::string> m;
        std::string str(const std::string & k, const std::string & d) const { auto it = m.find(k); return it == m.end() ? d : it->second; }
        double num(const std::string & k, double d) const { auto it = m.find(k); return it == m.end() ? d : std::stod(it->second); }
    };
} // namespace

int main(int argc, char ** argv)
{
    if (argc != 4)
    {
        std::cerr << "./navier_stokes <config> <msh file> <output dir>\n";
        return 1;
    }
    try
    {
        Cfg cfg{parseConfig(argv[1])};
        const std::string algo = cfg.str("algo", "chorinEigen");
        if (algo != "chorinCuda" && algo != "chorinHip")
        {
            std::cerr << "algo = " << algo << " is not provided by this build (GPU path: chorinCuda / chorinHip)\n";
            return 1;
        }
        nsfem_mesh * mesh = nullptr;
        cu::check(nsfem_mesh_load_gmsh(argv[2], &mesh));
        nsfem_dfg_conditions cond{cfg.num("viscosity", 0.001), cfg.num("peakVelocity", 1.0), 0.41, 1.0};
        nsfem_context * ctx = nullptr;
        cu::check(nsfem_context_build(mesh, &cond, 0, &ctx));
        int N = 0, V = 0;
        cu::check(nsfem_context_sizes(ctx, &N, &V, nullptr, nullptr, nullptr));

        const std::string vm = cfg.str("chorinCuda.velocitySolver.method", "gs");
        const std::string pmth = cfg.str("chorinCuda.pressureSolver.method", "gs");
        const bool f32 = cfg.str("hip.dtype", "f64") == "f32";
        nsfem_chorin_config cc{};
        cc.velocitySolver = {vm.c_str(), static_cast<int>(cfg.num("chorinCuda.velocitySolver.maxIterations", 200)),
                             cfg.num("chorinCuda.velocitySolver.targetMse", 1e-6),
                             static_cast<int>(cfg.num("chorinCuda.velocitySolver.mseCheckInterval", 1))};
        cc.pressureSolver = {pmth.c_str(), static_cast<int>(cfg.num("chorinCuda.pressureSolver.maxIterations", 200)),
                             cfg.num("chorinCuda.pressureSolver.targetMse", 1e-6),
                             static_cast<int>(cfg.num("chorinCuda.pressureSolver.mseCheckInterval", 1))};
        cc.dtype = f32 ? NSFEM_F32 : NSFEM_F64;
        cc.innerTol = cfg.num("hip.innerTol", 1e-4);
        // what a time step past the explicit scheme's stability limit does
        const std::string policy = cfg.str("hip.timeStepPolicy", "reject");
        if (policy == "reject")
        {
            cc.timeStepPolicy = NSFEM_TAU_REJECT;
        }
        else if (policy == "clamp")
        {
            cc.timeStepPolicy = NSFEM_TAU_CLAMP;
        }
        else if (policy == "unchecked")
        {
            cc.timeStepPolicy = NSFEM_TAU_UNCHECKED;
        }
        else
        {
            std::cerr << "hip.timeStepPolicy must be reject, clamp or unchecked\n";
            return 1;
        }
        // the pressure solve's initial guess: the last pressure or 2 p_n - p_n-1
        const std::string start = cfg.str("hip.pressureStart", "last");
        if (start != "last" && start != "extrapolate")
        {
            std::cerr << "hip.pressureStart must be last or extrapolate\n";
            return 1;
        }
        cc.pressureStart = start == "last" ? NSFEM_PRESSURE_START_LAST : NSFEM_PRESSURE_START_EXTRAPOLATE;
        // the velocity solves' initial guess: zero (as the reference), last or extrapolate
        const std::string vstart = cfg.str("hip.velocityStart", "zero");
        if (vstart == "zero")
        {
            cc.velocityStart = NSFEM_VELOCITY_START_ZERO;
        }
        else if (vstart == "last")
        {
            cc.velocityStart = NSFEM_VELOCITY_START_LAST;
        }
        else if (vstart == "extrapolate")
        {
            cc.velocityStart = NSFEM_VELOCITY_START_EXTRAPOLATE;
        }
        else
        {
            std::cerr << "hip.velocityStart must be zero, last or extrapolate\n";
            return 1;
        }
        nsfem_chorin * ch = nullptr;
        cu::check(nsfem_chorin_create(ctx, &cc, &ch));
        // the velocity solves' preconditioner (pcg methods): the reference's
        // symmetric GS sweep or the factorized approximate inverse
        const std::string vprec = cfg.str("hip.velocityPreconditioner", "gs");
        if (vprec == "fsai")
        {
            cu::check(nsfem_chorin_set_solver_option(ch, 0, "preconditioner", 1.0));
        }
        else if (vprec != "gs")
        {
            std::cerr << "hip.velocityPreconditioner must be gs or fsai\n";
            return 1;
        }
        // launch shapes tuned once before the first step (0: the defaults)
        const int tuneReps = static_cast<int>(cfg.num("hip.autotune", 10));
        if (tuneReps > 0)
        {
            cu::check(nsfem_chorin_autotune(ch, tuneReps, nullptr));
        }

        // tau = maxT / ceil(maxT / tau0) (chorinCuda.cu:221-222); under the
        // clamp policy tau0 is first lowered to the stability limit, so the
        // run still ends at maxT
        const double maxT = cfg.num("maxT", 1.0);
        double tau0 = cfg.num("tau", 1e-4);
        if (cc.timeStepPolicy == NSFEM_TAU_CLAMP)
        {
            double tauMax = 0;
            cu::check(nsfem_chorin_time_step_info(ch, nullptr, nullptr, nullptr, &tauMax));
            if (tau0 > tauMax)
            {
                std::printf("time step %g is past the stability limit %g: clamped\n", tau0, tauMax);
                tau0 = tauMax;
            }
        }
        const int numTimeSteps = static_cast<int>(std::ceil(maxT / tau0));
        const double tau = maxT / numTimeSteps;
        cu::check(nsfem_chorin_set_time_step(ch, tau));
        const int frameStep = std::max(1, static_cast<int>(cfg.num("output.frameStep", 1)));
        std::filesystem::create_directories(argv[3]);
        std::vector<double> u(2 * static_cast<size_t>(N)), p(V);
        std::vector<float> uf(u.size()), pf(p.size());
        int frame = 0;
        for (int iT = 0; iT <= numTimeSteps; iT++)
        {
            nsfem_step_stats st{};
            cu::check(nsfem_chorin_step(ch, &st));
            std::printf("%d / %d: %.3f ms\n", iT, numTimeSteps, st.msTotal);
            std::printf("\tconvection = %.3f, tentative = %.3f, pressure = %.3f, final = %.3f\n", st.msConvection, st.msTentative,
                        st.msPressure, st.msFinal);
            std::printf("\tMSEs: tent (X / Y / iters) = %g / %g / %d, pressure = %g / %d, final (X / Y / iters) = %g / %g / %d\n",
                        st.mseTentX, st.mseTentY, st.itersTent, st.msePressure, st.itersPressure, st.mseFinalX, st.mseFinalY,
                        st.itersFinal);
            if (iT % frameStep == 0)
            {
                const std::string dir = argv[3];
                if (f32)
                {
                    cu::check(nsfem_chorin_download(ch, uf.data(), pf.data()));
                    nsfem::writeVec(dir + "/velocity_" + std::to_string(frame) + ".bin", uf);
                    nsfem::writeVec(dir + "/pressure_" + std::to_string(frame) + ".bin", pf);
                }
                else
                {
                    cu::check(nsfem_chorin_download(ch, u.data(), p.data()));
                    nsfem::writeVec(dir + "/velocity_" + std::to_string(frame) + ".bin", u);
                    nsfem::writeVec(dir + "/pressure_" + std::to_string(frame) + ".bin", p);
                }
                frame++;
            }
        }
        nsfem_chorin_destroy(ch);
        nsfem_context_destroy(ctx);
        nsfem_mesh_destroy(mesh);
    }
    catch (const std::exception & e)
    {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
