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


def unpack_sym(packed, n):
    """Symmetric n x n matrix from its packed upper triangle (row-major
    triu order)."""
    m = np.zeros((n, n))
    m[np.triu_indices(n)] = packed
    return m + np.triu(m, 1).T


def cytosine_uks_setup():
    """Doublet cytosine / SVWN5 / cc-pVDZ on the Robust-pruned ultra-fine grid
    (tests/xc_integrator.cxx:455-459): fixture dict with unpacked densities
    and potentials, molecule, basis, load balancer."""
    import gauxc_amd as gx
    d = dict(load_golden("cytosine_svwn5_uks.npz"))
    n = int(d["nbf"][0])
    for k in ("density_s", "density_z", "vxc_s", "vxc_z"):
        d[k] = unpack_sym(d[k + "_packed"], n)
    mol = gx.Molecule(d["Z"], d["xyz"])
    bs = gx.BasisSet.from_arrays(d["sh_nprim"], d["sh_l"], d["sh_pure"], d["sh_alpha"],
                                 d["sh_coeff"], d["sh_origin"], normalize=False,
                                 shell_tolerance=EPS)
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Robust", 512, "MuraKnowles",
                                                  "UltraFineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    return d, mol, bs, lb


CALLER_SRC = os.path.join(ROOT, "tests", "cxx", "caller.cxx")
CALLER_BIN = os.path.join(ROOT, "tests", "cxx", "_build", "caller")


def build_caller():
    """The C++ caller program (tests/cxx/caller.cxx) against include/ and
    libgauxc_amd.so; rebuilt when the source or a header is newer than the
    binary (the classes' layout is part of the ABI it uses)."""
    import subprocess
    lib = os.path.join(ROOT, "gauxc_amd", "lib")
    import glob
    deps = [CALLER_SRC] + glob.glob(os.path.join(ROOT, "include", "gauxc", "**", "*.h*"),
                                    recursive=True)
    if (not os.path.exists(CALLER_BIN)
            or os.path.getmtime(CALLER_BIN) < max(os.path.getmtime(d) for d in deps)):
        os.makedirs(os.path.dirname(CALLER_BIN), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                        CALLER_SRC, "-o", CALLER_BIN, "-L" + lib, "-lgauxc_amd",
                        "-Wl,-rpath," + lib, "-ldl"], check=True)
    return CALLER_BIN


def write_caller_input(path, Z, xyz, l, pure, alpha, coeff, origin, P):
    """Input file of tests/cxx/caller.cxx (uncontracted shells)."""
    with open(path, "wb") as f:
        f.write(np.int64(len(Z)).tobytes())
        for z, r in zip(Z, xyz):
            f.write(np.int64(z).tobytes())
            f.write(np.asarray(r, dtype=np.float64).tobytes())
        f.write(np.int64(len(l)).tobytes())
        for i in range(len(l)):
            f.write(np.int32(l[i]).tobytes())
            f.write(np.int32(int(pure[i])).tobytes())
            f.write(np.float64(alpha[i][0]).tobytes())
            f.write(np.float64(coeff[i][0]).tobytes())
            f.write(np.asarray(origin[i], dtype=np.float64).tobytes())
        f.write(np.int64(P.shape[0]).tobytes())
        f.write(np.asfortranarray(P, dtype=np.float64).tobytes(order="F"))


def run_caller(args, timeout=120):
    import subprocess
    r = subprocess.run([build_caller(), *args], capture_output=True, text=True,
                       timeout=timeout)
    return r.returncode, r.stdout + r.stderr
