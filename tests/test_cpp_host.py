"""The C++ host tools built over the C ABI: the cu:: facade header compiles
with the reference's call shapes, the CLIs validate their arguments (CPU),
and run end to end on the GPU."""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, mesh_path

APP = os.path.join(ROOT, "nsfem_amd")


def _write_vec(path, v):
    v = np.asarray(v)
    with open(path, "wb") as f:
        np.array([0 if v.dtype == np.float32 else 1, v.size], np.int32).tofile(f)
        v.tofile(f)


def _write_csr(path, m):
    m = m.tocsr()
    with open(path, "wb") as f:
        np.array([0, m.shape[0], m.shape[1], m.nnz], np.int32).tofile(f)
        m.data.astype(np.float32).tofile(f)
        m.indices.astype(np.int32).tofile(f)
        m.indptr.astype(np.int32).tofile(f)


def _read_vec(path):
    with open(path, "rb") as f:
        dt, n = np.fromfile(f, np.int32, 2)
        return np.fromfile(f, np.float32 if dt == 0 else np.float64, n)


REFERENCE_CPP = "/root/reference/cpp"


def _compile(tmp_path, code, extra=()):
    src = tmp_path / "tu.cpp"
    src.write_text(code)
    return subprocess.run(["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-c", str(src), "-I", os.path.join(ROOT, "include"),
                           *extra, "-o", str(tmp_path / "tu.o")], capture_output=True, text=True)


# The data members of the reference's linalg::CsrMatrix<F> (csrMatrix.h:13-23)
# and a selection of its methods, declared as the reference declares them.
REFERENCE_SHAPED_CSR = r"""
#include <span>
#include <vector>
namespace linalg
{
    template <typename F>
    struct CsrMatrix
    {
        int rows, cols;
        std::vector<F> values;
        std::vector<int> column;
        std::vector<int> rowStart;
        CsrMatrix<F> slice(std::span<const int> rowIds, std::span<const int> colIds) const;
        void findOffsets(const int row, std::span<const int> columnIds, std::span<int> dstOffsets) const;
        bool compareLayout(const CsrMatrix<F> & other) const;
        bool operator==(const CsrMatrix<F> & other) const;
        void rMult(const F * src, F * dst) const;
        std::vector<F> rMult(const std::vector<F> & src) const;
    };
}
"""

FACADE_SHAPES = r"""
#include "nsfem/cu.hpp"
#include "nsfem/linalg_io.hpp"
void shapes(const linalg::CsrMatrix<float> & m, const linalg::CsrMatrix<double> & md)
{
    cu::Blas blas;
    cu::Sparse sparse;
    cu::csrF a(m);
    cu::spmv s1(sparse.handle(), a);
    cu::spmm s2(sparse.handle(), a, 2);
    s1.compute(s1.x, s1.b);
    auto solver = cu::createSolver("gs", 2, m, 200, 1e-6f, 1);
    cu::vec<float> & rhs = solver->getRhs();
    s2.compute(s2.x, rhs);
    solver->getSol().memsetZero();
    solver->solve();
    cu::saxpy(blas, 10, solver->getSol().get(), s2.x.get(), -1e-3f);
    cu::scale(blas, 10, s1.b.get(), -2.0f);
    (void)solver->getLastMse(0);
    (void)solver->getLastIterations();
    std::unique_ptr<cu::AbstractSolver> keep = std::move(solver);
    auto s64 = cu::createSolver("pcg_mixed", 1, md, 5000, 1e-10, 1);
    cu::vec<double> & r64 = s64->getRhs();
    (void)r64;
    cu::csrD b(md);
}
"""


def test_facade_compiles_with_reference_call_shapes(tmp_path):
    """The cu:: facade in a translation unit that already defines the
    reference's linalg::CsrMatrix (members and methods), with the call shapes
    of chorinCuda.cu:158-214 and solverFactory.h:13-14."""
    out = _compile(tmp_path, REFERENCE_SHAPED_CSR + FACADE_SHAPES)
    assert out.returncode == 0, out.stderr


REPLACEMENT_TU = r"""
#include <NavierStokes/chorinCuda.h>
#include "nsfem/chorin.hpp"
#include "nsfem/cu.hpp"

Solution solveNsChorinCuda(const mesh::ConcreteMesh & velocityMesh, const mesh::ConcreteMesh & pressureMesh,
                           const DfgConditions & cond, const float timeStep0, const float maxT,
                           const ChorinCudaConfig & cfg)
{
    return nsfem::solveNsChorin<Solution>(velocityMesh, pressureMesh, cond, timeStep0, maxT, cfg);
}
"""


def test_solve_ns_chorin_compiles_against_reference_headers(tmp_path):
    """The replacement for chorinCuda.cu: the reference's own headers
    (chorinCuda.h with mesh::ConcreteMesh, DfgConditions, ChorinCudaConfig,
    Solution) and the reference-shaped CsrMatrix in one translation unit with
    the facade; solveNsChorinCuda keeps its exact signature."""
    if not os.path.isdir(REFERENCE_CPP):
        pytest.skip("reference sources absent")
    inc = []
    for d in ("libs/Mesh/include", "libs/Element/include", "execs/NavierStokes/include"):
        inc += ["-I", os.path.join(REFERENCE_CPP, d)]
    out = _compile(tmp_path, REFERENCE_SHAPED_CSR + REPLACEMENT_TU, inc)
    assert out.returncode == 0, out.stderr


def test_cli_usage_and_algo_check(tmp_path):
    r = subprocess.run([os.path.join(APP, "navier_stokes")], capture_output=True, text=True)
    assert r.returncode == 1 and "<config>" in r.stderr
    cfg = tmp_path / "cfg.txt"
    cfg.write_text("algo = chorinEigen\n")
    r = subprocess.run([os.path.join(APP, "navier_stokes"), str(cfg), mesh_path("c3"), str(tmp_path / "out")], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "not provided" in r.stderr
    r = subprocess.run([os.path.join(APP, "solver_test")], capture_output=True, text=True)
    assert r.returncode == 1


@pytest.mark.gpu
def test_navier_stokes_cli_runs(tmp_path, oracle, octx_c3):
    cfg = tmp_path / "cfg.txt"
    cfg.write_text("\n".join([
        "# reference config keys", "algo = chorinCuda", "viscosity = 0.001", "peakVelocity = 1", "maxT = 0.002", "tau = 0.001",
        "chorinCuda.velocitySolver.method = pcg", "chorinCuda.velocitySolver.targetMse = 1e-13",
        "chorinCuda.velocitySolver.maxIterations = 2000", "chorinCuda.pressureSolver.method = pcg",
        "chorinCuda.pressureSolver.targetMse = 1e-13", "chorinCuda.pressureSolver.maxIterations = 8000", "output.frameStep = 1", ""]))
    out = tmp_path / "frames"
    r = subprocess.run([os.path.join(APP, "navier_stokes"), str(cfg), mesh_path("c3"), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "2 / 2:" in r.stdout
    orc = oracle.ChorinOracle(octx_c3, 0.001, vel=("pcg", 2000, 1e-13, 1), pres=("pcg", 8000, 1e-13, 1))
    for j in range(3):
        uo, po = orc.step()
        u = _read_vec(out / f"velocity_{j}.bin")
        p = _read_vec(out / f"pressure_{j}.bin")
        assert np.max(np.abs(u - uo)) <= 1e-10 * np.max(np.abs(uo))
        assert np.max(np.abs(p - po)) <= 1e-10 * np.max(np.abs(po))


@pytest.mark.gpu
def test_solver_test_harness(tmp_path, oracle, octx_c3):
    m = octx_c3["Kp_int"]
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, m.shape[0]).astype(np.float32)
    _write_csr(tmp_path / "m.bin", m)
    _write_vec(tmp_path / "b.bin", b)
    r = subprocess.run([os.path.join(APP, "solver_test"), "gaussSeidelCuda", str(tmp_path / "m.bin"), str(tmp_path / "b.bin"), "30",
                        "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    x = _read_vec(str(tmp_path / "b.bin") + ".sol.bin")
    xo, _, _ = oracle.gs_solve(m.astype(np.float32).astype(np.float64), b.astype(np.float64), np.zeros(m.shape[0]), 30, 0.0, 1)
    assert np.max(np.abs(x - xo)) <= 1e-4 * np.max(np.abs(xo))
    assert "Final MSE" in r.stdout


# Types with the data members of the reference's mesh::ConcreteMesh,
# DfgConditions, ChorinCudaConfig and Solution, and a driver that fills them
# from the library's own gmsh reader (P1) and P2 numbering, then runs
# nsfem::solveNsChorin.
SOLVE_DRIVER = r"""
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>
#include "nsfem/chorin.hpp"

namespace el { struct Point { float x, y; }; }
namespace mesh
{
    struct ConcreteMesh
    {
        std::vector<el::Point> nodes;
        int numElements = 0;
        std::vector<int> elements;
        std::vector<std::string> groups;
        int numBorderElements = 0;
        std::vector<int> borderElements;
    };
}
struct DfgConditions { float viscosity; float peakVelocity; };
struct SolverConfig { std::string method = "gs"; int maxIterations = 200; float targetMse = 1e-6f; int mseCheckInterval = 1; };
struct ChorinCudaConfig { SolverConfig velocitySolver; SolverConfig pressureSolver; };
struct TimeStepSolution { float time; std::vector<float> velocity; std::vector<float> pressure; };
struct Solution { std::vector<TimeStepSolution> steps; };

static const int kSides[6][2] = {{0, 1}, {1, 0}, {1, 2}, {2, 1}, {2, 0}, {0, 2}};

int main(int argc, char ** argv)
{
    if (argc != 3) return 2;
    try
    {
        nsfem_mesh * m = nullptr;
        cu::check(nsfem_mesh_load_gmsh(argv[1], &m));
        int nn, nt, nb, ng;
        cu::check(nsfem_mesh_info(m, &nn, &nt, &nb, &ng));
        std::vector<double> xy(2 * nn);
        std::vector<int> tri(3 * nt), border(3 * nb);
        cu::check(nsfem_mesh_arrays(m, xy.data(), tri.data(), border.data()));
        mesh::ConcreteMesh pm;
        for (int i = 0; i < nn; i++) pm.nodes.push_back({static_cast<float>(xy[2 * i]), static_cast<float>(xy[2 * i + 1])});
        pm.numElements = nt;
        pm.elements = tri;
        for (int g = 0; g < ng; g++) { char buf[64]; cu::check(nsfem_mesh_group_name(m, g, buf, 64)); pm.groups.push_back(buf); }
        pm.numBorderElements = nb;
        for (int b = 0; b < nb; b++)
        {
            const int e = border[3 * b], side = border[3 * b + 1], g = border[3 * b + 2];
            pm.borderElements.insert(pm.borderElements.end(), {e, side, g, tri[3 * e + kSides[side][0]], tri[3 * e + kSides[side][1]]});
        }
        nsfem_dfg_conditions c{1e-3, 1.0, 0.41, 1.0};
        nsfem_context * ctx = nullptr;
        cu::check(nsfem_context_build(m, &c, 0, &ctx));
        int N, V, d0, d1, E;
        cu::check(nsfem_context_sizes(ctx, &N, &V, &d0, &d1, &E));
        std::vector<double> vxy(2 * N);
        mesh::ConcreteMesh vm;
        vm.elements.resize(6 * E);
        cu::check(nsfem_context_velocity_mesh(ctx, vxy.data(), vm.elements.data()));
        for (int i = 0; i < N; i++) vm.nodes.push_back({static_cast<float>(vxy[2 * i]), static_cast<float>(vxy[2 * i + 1])});
        vm.numElements = E;
        vm.groups = pm.groups;
        nsfem_context_destroy(ctx);
        nsfem_mesh_destroy(m);

        ChorinCudaConfig cfg;
        cfg.velocitySolver = {"pcg", 2000, 1e-13f, 1};
        cfg.pressureSolver = {"pcg", 8000, 1e-13f, 1};
        const DfgConditions cond{1e-3f, 1.0f};
        const Solution sol = nsfem::solveNsChorin<Solution>(vm, pm, cond, 1e-3f, 3e-3f, cfg);
        std::printf("steps %zu\n", sol.steps.size());
        for (size_t k = 0; k < sol.steps.size(); k++)
        {
            std::ofstream f(std::string(argv[2]) + "/step_" + std::to_string(k) + ".bin", std::ios::binary);
            f.write(reinterpret_cast<const char *>(sol.steps[k].velocity.data()), sol.steps[k].velocity.size() * 4);
            f.write(reinterpret_cast<const char *>(sol.steps[k].pressure.data()), sol.steps[k].pressure.size() * 4);
        }
        // a velocity mesh that is not the P2 numbering of the pressure mesh is refused
        std::swap(vm.elements[0], vm.elements[1]);
        try
        {
            (void)nsfem::solveNsChorin<Solution>(vm, pm, cond, 1e-3f, 3e-3f, cfg);
            std::printf("mismatch accepted\n");
        }
        catch (const std::runtime_error & e)
        {
            std::printf("mismatch refused: %s\n", e.what());
        }
        std::swap(vm.elements[0], vm.elements[1]);
        // a step past the stability limit is refused by default
        try
        {
            (void)nsfem::solveNsChorin<Solution>(vm, pm, cond, 1.0f, 3.0f, cfg);
            std::printf("unstable accepted\n");
        }
        catch (const std::invalid_argument & e)
        {
            std::printf("unstable refused: %s\n", e.what());
        }
    }
    catch (const std::exception & e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
"""


@pytest.mark.gpu
def test_solve_ns_chorin_drop_in_runs(tmp_path, oracle, octx_c3):
    """nsfem::solveNsChorin with reference-shaped types: numTimeSteps + 1
    steps of tau = maxT / ceil(maxT / tau0), each kept state equal to the
    FP64 oracle's (the FP64 state rounded to the float Solution vectors); a
    mismatched velocity mesh and a step past the stability limit are
    refused."""
    src = tmp_path / "drv.cpp"
    src.write_text(SOLVE_DRIVER)
    exe = tmp_path / "drv"
    b = subprocess.run(["g++", "-std=c++20", "-O1", "-o", str(exe), str(src), "-I", os.path.join(ROOT, "include"), "-L", APP,
                        "-lnsfem_hip", "-Wl,-rpath," + APP], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([str(exe), mesh_path("c3"), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "steps 4" in r.stdout
    assert "mismatch refused" in r.stdout and "unstable refused" in r.stdout, r.stdout
    N, V = octx_c3["N"], octx_c3["V"]
    orc = oracle.ChorinOracle(octx_c3, 1e-3, vel=("pcg", 2000, 1e-13, 1), pres=("pcg", 8000, 1e-13, 1))
    for k in range(4):
        uo, po = orc.step()
        v = np.fromfile(tmp_path / f"step_{k}.bin", np.float32)
        u, p = v[:2 * N].astype(np.float64), v[2 * N:].astype(np.float64)
        assert p.size == V
        # FP64 results rounded once to float: within half a float ulp plus
        # the 1e-10 step tolerance
        assert np.max(np.abs(u - uo)) <= 2 ** -24 * np.max(np.abs(uo)) + 1e-10 * np.max(np.abs(uo)), k
        assert np.max(np.abs(p - po)) <= 2 ** -24 * np.max(np.abs(po)) + 1e-10 * np.max(np.abs(po)), k
