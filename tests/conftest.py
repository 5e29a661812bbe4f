import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
EPS = float(np.finfo(float).eps)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.fixture(scope="session")
def gx():
    import gauxc_amd
    return gauxc_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu(gx):
    """Device presence is required, not optional: the HIP path has no fallback."""
    n = gx.device_count()
    assert n > 0, "GPU tests need a HIP device; none is visible"
    return 0


def benzene_setup(tag, grid="UltraFineGrid"):
    """Molecule, basis (fixture-normalised, tolerance = eps as the GauXC test)
    and load balancer of a benzene golden file (tests/xc_integrator.cxx:163-172)."""
    import gauxc_amd as gx
    d = load_golden(f"benzene_{tag}.npz")
    mol = gx.Molecule(d["Z"], d["xyz"])
    bs = gx.BasisSet.from_arrays(d["sh_nprim"], d["sh_l"], d["sh_pure"], d["sh_alpha"],
                                 d["sh_coeff"], d["sh_origin"], normalize=False,
                                 shell_tolerance=EPS)
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles", grid)
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    return d, mol, bs, lb
