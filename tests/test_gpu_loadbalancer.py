"""Load balancer with the batch screening on the device (GauXC's device load
balancer, src/load_balancer/device/hip/replicated_hip_load_balancer.cxx:72-220):
the task list is index work, so it must equal the host load balancer's and
GauXC's benzene task fixture exactly (tests/load_balancer_test.cxx:40-82),
on the fixture and on the 24-water workload."""
import time

import numpy as np
import pytest

from conftest import EPS, load_golden

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _tasks(gx, mol, bs, grid, ex):
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles", grid)
    lb = gx.LoadBalancerFactory(ex, "Default").get_instance(gx.RuntimeEnvironment(), mol, mg, bs)
    tic = time.perf_counter()
    t = lb.get_tasks()
    return t, time.perf_counter() - tic


def _assert_same(t, r):
    assert len(t["npts"]) == len(r["npts"])
    for k in ("iparent", "npts", "nshells", "nbe", "shell_list"):
        np.testing.assert_array_equal(t[k], r[k], err_msg=k)
    for k in ("points", "weights", "dist_nearest"):
        np.testing.assert_array_equal(t[k], r[k], err_msg=k)


def test_device_load_balancer_equals_gauxc_fixture(gx, gpu):
    d = load_golden("benzene_svwn5.npz")
    ref = load_golden("benzene_tasks_ufg.npz")
    mol = gx.Molecule(d["Z"], d["xyz"])
    bs = gx.BasisSet.from_arrays(d["sh_nprim"], d["sh_l"], d["sh_pure"], d["sh_alpha"],
                                 d["sh_coeff"], d["sh_origin"], normalize=False,
                                 shell_tolerance=EPS)
    _flush("device load balancer, benzene UltraFineGrid")
    t, _ = _tasks(gx, mol, bs, "UltraFineGrid", "Device")
    assert len(t["npts"]) == 664
    np.testing.assert_array_equal(t["iparent"], ref["iParent"])
    np.testing.assert_array_equal(t["npts"], ref["npts"])
    np.testing.assert_array_equal(t["nbe"], ref["nbe"])
    np.testing.assert_array_equal(t["shell_list"], ref["shell_list"])
    h, _ = _tasks(gx, mol, bs, "UltraFineGrid", "Host")
    _assert_same(t, h)


def test_device_load_balancer_equals_host_24_waters(gx, gpu):
    mol, bs, P, cfg = gx.workloads.make_config("c1", seed=0)
    _flush("device load balancer, 24 waters")
    t, td = _tasks(gx, mol, bs, cfg["grid"], "Device")
    h, th = _tasks(gx, mol, bs, cfg["grid"], "Host")
    _flush(f"{len(t['npts'])} tasks; task build host {th:.3f} s, device screening {td:.3f} s")
    _assert_same(t, h)
