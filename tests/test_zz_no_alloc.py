"""xcq_ssf_weights_noalloc: the allocation-free SSF entry point (static
device scratch, caller-owned device buffers) against xcq_ssf_weights on the
same inputs, on two calls in a row (the second one smaller, so it runs over
the first call's leftover scratch), plus its size limit and its parent-index
check."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    return ctypes.CDLL("libamdhip64.so")


class _DevBuf:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.p = ctypes.c_void_p()
        hip = _hip()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(self.host.nbytes)) == 0
        assert hip.hipMemcpy(self.p, self.host.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(self.host.nbytes), H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert _hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p,
                                ctypes.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        _hip().hipFree(self.p)


def _system(gx, nwater, seed):
    Z, xyz = gx.workloads.water_cluster(nwater, 6.0 * (nwater / 3.0) ** (1 / 3), seed)
    l, pure, alpha, coeff, origin = gx.workloads.even_tempered_basis(Z, xyz,
                                                                     gx.workloads.BASIS_C0)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=1e-10)
    lb = gx.workloads.build_tasks(gx.Molecule(Z, xyz), bs, grid="GM3")
    t = lb.get_tasks()
    parent = np.repeat(t["iparent"], t["npts"]).astype(np.int32)
    return np.ascontiguousarray(xyz, dtype=np.float64), t["points"].copy(), parent, \
        t["weights"].copy()


def _reference(lib, xyz, pts, parent, w):
    out = w.copy()
    rc = lib.xcq_ssf_weights(0, len(xyz), xyz.ctypes.data, len(w), pts.ctypes.data,
                             parent.ctypes.data, out.ctypes.data)
    assert rc == 0, lib.xcq_last_error()
    return out


def _noalloc(lib, xyz, pts, parent, w):
    bufs = [_DevBuf(pts), _DevBuf(parent), _DevBuf(w)]
    try:
        rc = lib.xcq_ssf_weights_noalloc(0, len(xyz), xyz.ctypes.data, len(w), bufs[0].p,
                                         bufs[1].p, bufs[2].p)
        return rc, bufs[2].get()
    finally:
        for b in bufs:
            b.free()


def test_ssf_noalloc_matches_allocating_entry_point(gx, gpu):
    from gauxc_amd import _native
    lib = _native.device_lib()
    for nwater, seed in ((12, 3), (3, 0)):  # second call smaller than the first
        print(f"noalloc SSF: {nwater} waters", flush=True)
        xyz, pts, parent, w = _system(gx, nwater, seed)
        ref = _reference(lib, xyz, pts, parent, w)
        rc, got = _noalloc(lib, xyz, pts, parent, w)
        assert rc == 0, lib.xcq_last_error()
        assert np.array_equal(got != w, ref != w)
        # Same partition factor P_parent / sum_A P_A to 1e-13 of each point's
        # unpartitioned weight (the kernels are the same; the distance matrix
        # is formed on the device here, on the host there, and may differ in
        # the last bit, which moves factors near the switching edge of the
        # Frisch polynomial by a few ulps).
        assert np.all(np.abs(got - ref) <= 1e-13 * np.abs(w))


def test_ssf_noalloc_limits(gx, gpu):
    from gauxc_amd import _native
    lib = _native.device_lib()
    xyz, pts, parent, w = _system(gx, 3, 0)
    n = _native.device_lib()
    big = np.zeros((1025, 3))
    rc = n.xcq_ssf_weights_noalloc(0, 1025, big.ctypes.data, 0, None, None, None)
    assert rc != 0 and b"XCQ_SSF_NOALLOC_MAX_ATOMS" in lib.xcq_last_error()
    bad = parent.copy()
    bad[5] = len(xyz)  # out of range: the kernel skips the point and flags it
    print("noalloc SSF: out-of-range parent", flush=True)
    rc, got = _noalloc(lib, xyz, pts, bad, w)
    assert rc != 0 and b"parent atom out of range" in lib.xcq_last_error()
    assert got[5] == w[5]
