"""The drop-in boundary on the device: the runtime environment's fill
fraction and caller-owned device arena reach the integrator
(GauXC include/gauxc/runtime_environment/decl.hpp:71-73; the arena-too-small
error of xc_device_stack_data.cxx:720-721), and a C++ caller built against
include/gauxc/*.hpp gets the same result as the C API."""
import ctypes
import re

import numpy as np
import pytest

from conftest import run_caller

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def _device_buffer(nbytes):
    hip = _hip()
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
    return p


def _free(p):
    _hip().hipFree(p)


@pytest.fixture(scope="module")
def system(gx, gpu):
    Z, xyz = gx.workloads.water_cluster(3, 6.0, 0)
    l, pure, alpha, coeff, origin = gx.workloads.even_tempered_basis(Z, xyz,
                                                                     gx.workloads.BASIS_C0)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=1e-10)
    P = gx.workloads.density(bs.nbf(), 15, 0)
    mol = gx.Molecule(Z, xyz)
    return (Z, xyz, l, pure, alpha, coeff, origin), mol, bs, P


def _eval(gx, mol, bs, P, rt):
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles", "FineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(rt, mol, mg, bs)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional("PBE"), lb)
    return integ.eval_exc_vxc(P)


def test_fill_fraction_and_caller_arena(gx, system):
    _, mol, bs, P = system
    _flush("boundary: default environment")
    e0, v0 = _eval(gx, mol, bs, P, gx.RuntimeEnvironment())
    _flush("boundary: device environment, fill fraction 0.5")
    e1, v1 = _eval(gx, mol, bs, P, gx.DeviceRuntimeEnvironment(0.5))
    assert abs(e1 - e0) <= 1e-12 * abs(e0) and np.abs(v1 - v0).max() <= 1e-13
    _flush("boundary: fill fraction too small for one task")
    with pytest.raises(gx.GauXCError, match="Insufficient memory to even start"):
        _eval(gx, mol, bs, P, gx.DeviceRuntimeEnvironment(1e-12))
    _flush("boundary: caller arena of 4 KiB")
    small = _device_buffer(4096)
    try:
        with pytest.raises(gx.GauXCError, match="Insufficient memory to even start"):
            _eval(gx, mol, bs, P, gx.DeviceRuntimeEnvironment(mem=small.value, mem_size=4096))
    finally:
        _free(small)
    _flush("boundary: caller arena of 64 MiB")
    big = _device_buffer(64 << 20)
    try:
        e2, v2 = _eval(gx, mol, bs, P, gx.DeviceRuntimeEnvironment(mem=big.value,
                                                                   mem_size=64 << 20))
    finally:
        _free(big)
    assert abs(e2 - e0) <= 1e-12 * abs(e0) and np.abs(v2 - v0).max() <= 1e-13


def test_device_context_caller_arena_chunks(gx, system):
    """xcq_context_create_mem: an arena that holds a few tasks at a time
    splits the evaluation into many chunks, with the same result."""
    _, mol, bs, P = system
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles", "FineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol, mg,
                                                                 bs)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    t = lb.get_tasks()
    ref = gx.DeviceContext(0)
    ref.upload_basis(bs.xcq_shells())
    ref.upload_tasks(t)
    ref.upload_density(P)
    ref.exc_vxc("PBE", "fp64", True)
    e0, n0, v0 = ref.results()
    # Scratch of one task: five FP64 planes of nbe x npts (npts rounded up to
    # 64).  An arena of twice the largest task holds a few tasks at a time.
    task_bytes = 5 * 8 * t["nbe"].astype(np.int64) * ((t["npts"].astype(np.int64) + 63) // 64 * 64)
    nbytes = int(2 * task_bytes.max() + 255) // 256 * 256
    assert task_bytes.sum() > 4 * nbytes
    buf = _device_buffer(nbytes)
    try:
        lib = ref.lib
        h = ctypes.c_void_p()
        assert lib.xcq_context_create_mem(0, buf, nbytes, ctypes.byref(h)) == 0
        ctx = gx.DeviceContext.__new__(gx.DeviceContext)
        ctx.lib, ctx.h, ctx.nbf, ctx.tasks = lib, h, None, None
        ctx.upload_basis(bs.xcq_shells())
        ctx.upload_tasks(t)
        ctx.upload_density(P)
        for prec in ("fp64", "mixed"):
            _flush(f"caller arena {nbytes >> 20} MiB {prec}")
            ctx.exc_vxc("PBE", prec, True)
            e, n, v = ctx.results()
            tol = 1e-12 if prec == "fp64" else 1e-6
            assert abs(e - e0) <= tol * max(1.0, abs(e0))
            assert np.abs(v - v0).max() <= (1e-13 if prec == "fp64" else 1e-7)
        ctx.close()
        misaligned = ctypes.c_void_p(buf.value + 8)
        assert lib.xcq_context_create_mem(0, misaligned, 4096, ctypes.byref(h)) != 0
        assert b"256-byte aligned" in lib.xcq_last_error()
    finally:
        _free(buf)


def test_cxx_caller_on_device(gx, system, tmp_path):
    """The C++ caller (GauXC's call sequence against include/gauxc/*.hpp)
    on the device: same EXC as the C API; a caller arena too small for one
    task fails with GauXC's message; a large one gives the same EXC."""
    from conftest import write_caller_input
    arrays, mol, bs, P = system
    Z, xyz, l, pure, alpha, coeff, origin = arrays
    path = str(tmp_path / "caller_input.bin")
    write_caller_input(path, Z, xyz, l, pure, alpha, coeff, origin, P)
    e0, v0 = _eval(gx, mol, bs, P, gx.RuntimeEnvironment())
    _flush("C++ caller: device")
    rc, out = run_caller([path, "device"])
    assert rc == 0, out
    exc = float(re.search(r"EXC (\S+)", out).group(1))
    vsum = float(re.search(r"VXC_abs_sum (\S+)", out).group(1))
    assert abs(exc - e0) <= 1e-12 * abs(e0)
    assert abs(vsum - np.abs(v0).sum()) <= 1e-10 * np.abs(v0).sum()
    _flush("C++ caller: caller arena of 1 KiB")
    rc, out = run_caller([path, "device-mem", "1024"])
    assert rc == 1 and "Insufficient memory to even start!" in out, out
    _flush("C++ caller: caller arena of 256 MiB")
    rc, out = run_caller([path, "device-mem", str(256 << 20)])
    assert rc == 0, out
    assert abs(float(re.search(r"EXC (\S+)", out).group(1)) - e0) <= 1e-12 * abs(e0)
