import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: large workload")
    # Build what is missing (a checkout without built artefacts); a present
    # library is used as is.
    lib = os.path.join(ROOT, "nsfem_amd", "libnsfem_hip.so")
    if not os.path.exists(lib):
        import importlib.util
        spec = importlib.util.spec_from_file_location("nsfem_build", os.path.join(ROOT, "nsfem_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    if not os.path.exists(os.path.join(ROOT, "oracle", "_build", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def mesh_path(name):
    return os.path.join(GOLDEN, name + ".msh")


@pytest.fixture(scope="session")
def oracle():
    from oracle import nsfem_oracle
    return nsfem_oracle


@pytest.fixture(scope="session")
def octx_c3(oracle):
    return oracle.build_context(oracle.parse_gmsh(mesh_path("c3")), 1e-3, 1.0)


@pytest.fixture(scope="session")
def octx_h025(oracle):
    return oracle.build_context(oracle.parse_gmsh(mesh_path("h_0.025")), 1e-3, 1.0)
