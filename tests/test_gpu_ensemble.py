"""Ensembles: B independent flows on one mesh sharing every matrix.

Instance b with inflow scale s_b is the flow whose Dirichlet velocity is
s_b times the DFG profile, i.e. the oracle's context with peak velocity s_b;
each instance of the GPU ensemble is held to the same bound against its own
FP64 oracle run as a single flow (1e-10 relative for the FP64 solvers at a
1e-13 residual, the mixed bounds of the single-flow tests otherwise)."""

import numpy as np
import pytest

from conftest import mesh_path

pytestmark = pytest.mark.gpu

SCALES = [1.0, 0.5, 1.5, 0.8, 1.2, 0.7, 0.9, 1.1]


def _ctx(name="c3", visc=1e-3):
    from nsfem_amd import fem
    return fem.ChorinContext(fem.TriangleMesh.load_gmsh(mesh_path(name)), fem.DfgConditions(visc, 1.0))


def _cfg(vel, pres, dtype=1, step_graph=0):
    from nsfem_amd import fem
    return fem.ChorinConfig(fem.SolverConfig(*vel), fem.SolverConfig(*pres), dtype=dtype, stepGraph=step_graph)


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _oracles(oracle, scales, tau, name="c3"):
    mesh = oracle.parse_gmsh(mesh_path(name))
    return [oracle.ChorinOracle(oracle.build_context(mesh, 1e-3, s), tau, vel=("pcg", 2000, 1e-13, 1),
                                pres=("pcg", 8000, 1e-13, 1)) for s in scales]


def _split(ch, u, p):
    B, N, V = ch.instances, ch.N, ch.V
    return u.reshape(B, 2 * N), p.reshape(B, V)


@pytest.mark.parametrize("pressure", ["pcg", "bpcg"])
def test_fp64_ensemble_instances_match_oracle(oracle, pressure):
    from nsfem_amd import fem
    tau, B = 1e-3, 4
    ch = fem.Chorin(_ctx(), _cfg(("pcg", 2000, 1e-13, 1), (pressure, 8000, 1e-13, 1)), instances=B, inflow_scale=SCALES[:B])
    assert ch.instances == B
    ch.set_time_step(tau)
    orcs = _oracles(oracle, SCALES[:B], tau)
    for k in range(3):
        st = ch.step()
        assert st["relTent"] <= 1e-13 and st["relFinal"] <= 1e-13
        assert st["relPressure"] <= 1e-12
        u, p = _split(ch, *ch.download())
        for b, orc in enumerate(orcs):
            uo, po = orc.step()
            assert _rel(u[b], uo) <= 1e-10, (k, b)
            assert _rel(p[b], po) <= 1e-10, (k, b)


def test_mixed_ensemble_with_block_pressure_matches_oracle(oracle):
    """The ensemble's intended shape: velocity pcg_mixed over 2B columns,
    pressure bpcg_mixed over B columns sharing one Krylov space."""
    from nsfem_amd import fem
    tau, B = 1e-3, 8
    ch = fem.Chorin(_ctx(), _cfg(("pcg_mixed", 2000, 1e-10, 1), ("bpcg_mixed", 8000, 1e-10, 1)), instances=B,
                    inflow_scale=SCALES)
    ch.set_time_step(tau)
    orcs = _oracles(oracle, SCALES, tau)
    for _ in range(3):
        st = ch.step()
        assert st["relTent"] <= 1e-10 and st["relFinal"] <= 1e-10
        assert st["relPressure"] <= 1e-9
        outs = [orc.step() for orc in orcs]
    u, p = _split(ch, *ch.download())
    for b, (uo, po) in enumerate(outs):
        eu, ep = _rel(u[b], uo), _rel(p[b], po)
        assert eu < 1e-8 and ep < 1e-7, (b, eu, ep)


@pytest.mark.parametrize("pressure,B,target,bounds", [("amg_pcg", 4, 1e-13, (1e-10, 1e-10)),
                                                     ("amg_pcg_mixed", 8, 1e-10, (1e-8, 1e-7))])
def test_ensemble_with_amg_pressure_matches_oracle(oracle, pressure, B, target, bounds):
    """The pressure solve of an ensemble as one AMG-preconditioned solve of B
    right-hand sides."""
    from nsfem_amd import fem
    tau = 1e-3
    vel = "pcg" if target < 1e-12 else "pcg_mixed"
    ch = fem.Chorin(_ctx(), _cfg((vel, 2000, target, 1), (pressure, 500, target, 1)), instances=B, inflow_scale=SCALES[:B])
    ch.set_time_step(tau)
    orcs = _oracles(oracle, SCALES[:B], tau)
    for _ in range(3):
        st = ch.step()
        assert st["relPressure"] <= target and st["itersPressure"] < 60
        outs = [orc.step() for orc in orcs]
    u, p = _split(ch, *ch.download())
    for b, (uo, po) in enumerate(outs):
        assert _rel(u[b], uo) <= bounds[0], b
        assert _rel(p[b], po) <= bounds[1], b


def test_ensemble_of_one_equals_the_single_flow():
    from nsfem_amd import fem
    cfg = _cfg(("pcg", 2000, 1e-13, 1), ("pcg", 8000, 1e-13, 1))
    a = fem.Chorin(_ctx(), cfg)
    b = fem.Chorin(_ctx(), cfg, instances=1)
    for ch in (a, b):
        ch.set_time_step(1e-3)
        ch.step()
        ch.step()
    ua, pa = a.download()
    ub, pb = b.download()
    assert np.array_equal(ua, ub) and np.array_equal(pa, pb)


def test_captured_ensemble_step_matches_regular_steps():
    from nsfem_amd import fem
    B = 4
    regular = fem.Chorin(_ctx(), _cfg(("pcg", 2000, 1e-13, 1), ("pcg", 8000, 1e-13, 1)), instances=B,
                         inflow_scale=SCALES[:B])
    captured = fem.Chorin(_ctx(), _cfg(("pcg", 2000, 1e-13, 1), ("pcg", 8000, 1e-13, 1), step_graph=1), instances=B,
                          inflow_scale=SCALES[:B])
    for ch in (regular, captured):
        ch.set_time_step(1e-3)
    kinds = []
    for _ in range(5):
        regular.step()
        kinds.append(captured.step()["capturedStep"])
    assert kinds[0] == 0 and sum(kinds[1:]) >= 3
    ur, pr = regular.download()
    uc, pc = captured.download()
    assert _rel(uc, ur) <= 1e-11 and _rel(pc, pr) <= 1e-10


def test_ensemble_async_download_and_upload_cover_every_instance():
    from nsfem_amd import fem
    B = 2
    ch = fem.Chorin(_ctx(), _cfg(("pcg", 2000, 1e-13, 1), ("pcg", 8000, 1e-13, 1)), instances=B, inflow_scale=[1.0, 2.0])
    ch.set_time_step(1e-3)
    ch.step()
    u0, p0 = ch.download()
    ch.download_async()
    u1, p1 = ch.download_wait()
    assert np.array_equal(u0, u1) and np.array_equal(p0, p1)
    u, _ = _split(ch, u0, p0)
    # the two instances differ by their inflow (scale 2 vs 1)
    assert _rel(u[1], u[0]) > 0.5
    ch.upload_velocity(np.zeros(2 * ch.N * B))
    z, _ = ch.download()
    assert not z.any()


def test_ensemble_rejects_bad_shapes():
    from nsfem_amd import fem
    from nsfem_amd._lib import NsfemError
    cfg = _cfg(("pcg", 100, 1e-10, 1), ("pcg", 100, 1e-10, 1))
    with pytest.raises(NsfemError, match="instances must be"):
        fem.Chorin(_ctx(), cfg, instances=3)
    with pytest.raises(ValueError, match="one value per instance"):
        fem.Chorin(_ctx(), cfg, instances=2, inflow_scale=[1.0])
    with pytest.raises(NsfemError, match="finite"):
        fem.Chorin(_ctx(), cfg, instances=2, inflow_scale=[1.0, float("nan")])


@pytest.mark.parametrize("velocity,pressure,B,target,bounds", [("pcg", "amg_pcg", 4, 1e-13, (1e-10, 1e-10)),
                                                              ("pcg_mixed", "amg_pcg_mixed", 8, 1e-10, (1e-8, 1e-7))])
def test_ensemble_with_fsai_velocity_matches_oracle(oracle, velocity, pressure, B, target, bounds):
    """The velocity solves of an ensemble (2B columns, one lane per column)
    preconditioned by the factorized approximate inverse: every instance
    against its own oracle run."""
    from nsfem_amd import fem
    tau = 1e-3
    ch = fem.Chorin(_ctx(), _cfg((velocity, 2000, target, 1), (pressure, 500, target, 1)), instances=B,
                    inflow_scale=SCALES[:B])
    ch.set_solver_option("preconditioner", 1, "velocity")
    ch.set_time_step(tau)
    orcs = _oracles(oracle, SCALES[:B], tau)
    for _ in range(3):
        st = ch.step()
        assert st["relTent"] <= target and st["relFinal"] <= target
        outs = [orc.step() for orc in orcs]
    u, p = _split(ch, *ch.download())
    for b, (uo, po) in enumerate(outs):
        eu, ep = _rel(u[b], uo), _rel(p[b], po)
        assert eu <= bounds[0] and ep <= bounds[1], (b, eu, ep)
