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


def test_facade_compiles_with_reference_call_shapes(tmp_path):
    src = tmp_path / "shape.cpp"
    src.write_text(r'''
#include "nsfem/cu.hpp"
// Call shapes of chorinCuda.cu:158-214 and solverFactory.h:13-14
void shapes(const linalg::CsrMatrix<float> & m)
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
}
''')
    out = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-c", str(src), "-I", os.path.join(ROOT, "include"),
                          "-o", str(tmp_path / "shape.o")], capture_output=True, text=True)
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
