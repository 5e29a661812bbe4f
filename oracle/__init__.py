"""Python access to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (oracle.cpp) restates GauXC's host reference integrator; it is
the checker of the HIP path in tests/, in __graft_entry__.smoke(), and the
CPU baseline timed by bench.py.  Product code never imports this package.
``gg_ref()`` exposes GauXC's own gau2grid collocation compiled from the
reference sources into oracle/_ref (present only where it was built).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
GG_REF = os.path.join(HERE, "_ref", "libgg_ref.so")

FUNC = {"SVWN5": 1, "PBE": 2, "PBE0": 3}


class OracleShell(ctypes.Structure):
    _fields_ = [("nprim", ctypes.c_int32), ("l", ctypes.c_int32), ("pure", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("alpha", ctypes.c_double * 32),
                ("coeff", ctypes.c_double * 32), ("origin", ctypes.c_double * 3),
                ("cutoff_radius", ctypes.c_double)]


class OracleTask(ctypes.Structure):
    _fields_ = [("iparent", ctypes.c_int32), ("npts", ctypes.c_int32),
                ("nshells", ctypes.c_int32), ("nbe", ctypes.c_int32),
                ("point_offset", ctypes.c_int64), ("shell_offset", ctypes.c_int64),
                ("dist_nearest", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.oracle_ssf_weights.argtypes = [ctypes.c_int32, vp, ctypes.c_int32, vp, vp, vp,
                                         ctypes.c_int]
        L.oracle_ssf_weights.restype = None
        L.oracle_collocation.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, vp, vp, vp, vp,
                                         vp]
        L.oracle_collocation.restype = None
        L.oracle_functional.argtypes = [ctypes.c_int, ctypes.c_int64, vp, vp, vp, vp, vp]
        L.oracle_functional.restype = ctypes.c_int
        L.oracle_exc_vxc.argtypes = [ctypes.c_int, ctypes.c_int32, vp, vp, ctypes.c_int64,
                                     ctypes.c_int32, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int64,
                                     ctypes.c_int]
        L.oracle_exc_vxc.restype = ctypes.c_int
        L.oracle_functional_pol.argtypes = [ctypes.c_int, ctypes.c_int64, vp, vp, vp, vp, vp]
        L.oracle_functional_pol.restype = ctypes.c_int
        L.oracle_exc_vxc_uks.argtypes = [ctypes.c_int, ctypes.c_int32, vp, vp, ctypes.c_int64,
                                         vp, ctypes.c_int64, ctypes.c_int32, vp, vp, vp, vp,
                                         vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int]
        L.oracle_exc_vxc_uks.restype = ctypes.c_int
        L.oracle_functional_fxc.argtypes = [ctypes.c_int, ctypes.c_int64] + [vp] * 8
        L.oracle_functional_fxc.restype = ctypes.c_int
        L.oracle_fxc_contraction.argtypes = [ctypes.c_int, ctypes.c_int32, vp, vp,
                                             ctypes.c_int64, vp, ctypes.c_int64,
                                             ctypes.c_int32, vp, vp, vp, vp, vp, vp,
                                             ctypes.c_int64, ctypes.c_int]
        L.oracle_fxc_contraction.restype = ctypes.c_int
        L.oracle_fxc_contraction_uks.argtypes = [ctypes.c_int, ctypes.c_int32, vp, vp, vp, vp,
                                                 vp, ctypes.c_int64, ctypes.c_int32, vp, vp,
                                                 vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int]
        L.oracle_fxc_contraction_uks.restype = ctypes.c_int
        L.oracle_exc_grad_uks.argtypes = [ctypes.c_int, ctypes.c_int32, vp, ctypes.c_int32, vp,
                                          vp, vp, ctypes.c_int64, ctypes.c_int32, vp, vp, vp,
                                          vp, ctypes.c_int, vp, ctypes.c_int]
        L.oracle_exc_grad_uks.restype = ctypes.c_int
        L.oracle_collocation_hessian.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, vp, vp]
        L.oracle_collocation_hessian.restype = None
        L.oracle_exc_grad.argtypes = [ctypes.c_int, ctypes.c_int32, vp, ctypes.c_int32, vp, vp,
                                      ctypes.c_int64, ctypes.c_int32, vp, vp, vp, vp,
                                      ctypes.c_int, vp, ctypes.c_int]
        L.oracle_exc_grad.restype = ctypes.c_int
        L.oracle_use_blas.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.oracle_use_blas.restype = ctypes.c_int
        _lib = L
    return _lib


def shells_from(struct_array):
    """Copy any shell struct array with nprim/l/pure/alpha/coeff/origin fields."""
    n = len(struct_array)
    out = (OracleShell * n)()
    for i, s in enumerate(struct_array):
        d = out[i]
        d.nprim, d.l, d.pure = s.nprim, s.l, int(s.pure)
        for k in range(s.nprim):
            d.alpha[k] = s.alpha[k]
            d.coeff[k] = s.coeff[k]
        for k in range(3):
            d.origin[k] = s.origin[k]
    return out


_TASK_DTYPE = np.dtype([("iparent", "<i4"), ("npts", "<i4"), ("nshells", "<i4"),
                        ("nbe", "<i4"), ("point_offset", "<i8"), ("shell_offset", "<i8"),
                        ("dist_nearest", "<f8")], align=True)
assert _TASK_DTYPE.itemsize == ctypes.sizeof(OracleTask)


def _task_array(t):
    """oracle_task records of a task dict, as a numpy structured array."""
    arr = np.zeros(len(t["npts"]), dtype=_TASK_DTYPE)
    for k in _TASK_DTYPE.names:
        arr[k] = t[k]
    return arr


def ssf_weights(xyz, t, nthreads=0):
    """SSF-modified copy of t['weights'] (t: task dict of LoadBalancer.get_tasks())."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    pts = np.ascontiguousarray(t["points"], dtype=np.float64)
    w = np.array(t["weights"], dtype=np.float64, copy=True)
    tasks = _task_array(t)
    lib().oracle_ssf_weights(len(xyz), xyz.ctypes.data, len(t["npts"]), tasks.ctypes.data,
                             pts.ctypes.data, w.ctypes.data, nthreads)
    return w


def collocation(shells, shell_list, points, deriv=True):
    """[comp][npts][nbe] values (comp = value, d/dx, d/dy, d/dz)."""
    sl = np.ascontiguousarray(shell_list, dtype=np.int32)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    nbe = 0
    for s in sl:
        sh = shells[int(s)]
        nbe += 2 * sh.l + 1 if sh.pure else (sh.l + 1) * (sh.l + 2) // 2
    out = np.zeros((4 if deriv else 1, len(pts), nbe))
    ptrs = [out[k].ctypes.data for k in range(out.shape[0])] + [None] * (4 - out.shape[0])
    lib().oracle_collocation(ctypes.addressof(shells), len(sl), sl.ctypes.data, len(pts),
                             pts.ctypes.data, *ptrs)
    return out


def functional(func, rho, sigma=None):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = len(rho)
    eps, vr, vs = np.zeros(n), np.zeros(n), np.zeros(n)
    sig = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
    rc = lib().oracle_functional(FUNC[func], n, rho.ctypes.data,
                                 None if sig is None else sig.ctypes.data, eps.ctypes.data,
                                 vr.ctypes.data, None if sig is None else vs.ctypes.data)
    assert rc == 0
    return eps, vr, (vs if sig is not None else None)


def exc_vxc(func, nbf, shells, P, t, weights=None, nthreads=0):
    """(EXC, NEL, VXC) of the restated host integrator over task dict t."""
    P = np.asfortranarray(P, dtype=np.float64)
    pts = np.ascontiguousarray(t["points"], dtype=np.float64)
    w = np.ascontiguousarray(t["weights"] if weights is None else weights, dtype=np.float64)
    sl = np.ascontiguousarray(t["shell_list"], dtype=np.int32)
    vxc = np.zeros((nbf, nbf), order="F")
    exc = ctypes.c_double(0.)
    nel = ctypes.c_double(0.)
    tasks = _task_array(t)
    rc = lib().oracle_exc_vxc(FUNC[func], nbf, ctypes.addressof(shells), P.ctypes.data, nbf,
                              len(t["npts"]), tasks.ctypes.data, pts.ctypes.data,
                              w.ctypes.data, sl.ctypes.data, ctypes.addressof(exc),
                              ctypes.addressof(nel), vxc.ctypes.data, nbf, nthreads)
    assert rc == 0
    return exc.value, nel.value, vxc


def functional_pol(func, rho, sigma=None):
    """Spin-polarised functional: rho (n, 2) = (rho_a, rho_b), sigma (n, 3) =
    (aa, ab, bb); returns eps (n), vrho (n, 2), vsigma (n, 3) or None."""
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, 2)
    n = len(rho)
    eps, vr, vs = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 3))
    sig = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1, 3)
    rc = lib().oracle_functional_pol(FUNC[func], n, rho.ctypes.data,
                                     None if sig is None else sig.ctypes.data, eps.ctypes.data,
                                     vr.ctypes.data, None if sig is None else vs.ctypes.data)
    assert rc == 0
    return eps, vr, (vs if sig is not None else None)


def exc_vxc_uks(func, nbf, shells, Ps, Pz, t, weights=None, nthreads=0):
    """(EXC, NEL, VXC_s, VXC_z) of the restated UKS host integrator."""
    Ps = np.asfortranarray(Ps, dtype=np.float64)
    Pz = np.asfortranarray(Pz, dtype=np.float64)
    pts = np.ascontiguousarray(t["points"], dtype=np.float64)
    w = np.ascontiguousarray(t["weights"] if weights is None else weights, dtype=np.float64)
    sl = np.ascontiguousarray(t["shell_list"], dtype=np.int32)
    vs, vz = np.zeros((nbf, nbf), order="F"), np.zeros((nbf, nbf), order="F")
    exc = ctypes.c_double(0.)
    nel = ctypes.c_double(0.)
    tasks = _task_array(t)
    rc = lib().oracle_exc_vxc_uks(FUNC[func], nbf, ctypes.addressof(shells), Ps.ctypes.data,
                                  nbf, Pz.ctypes.data, nbf, len(t["npts"]), tasks.ctypes.data,
                                  pts.ctypes.data, w.ctypes.data, sl.ctypes.data,
                                  ctypes.addressof(exc), ctypes.addressof(nel), vs.ctypes.data,
                                  vz.ctypes.data, nbf, nthreads)
    assert rc == 0
    return exc.value, nel.value, vs, vz


def functional_fxc(func, rho, sigma, drho, dsigma):
    """Unpolarised functional with directional second derivatives: returns
    vrho, vsigma, a = v2rho2 drho + v2rhosigma dsigma and
    b = v2rhosigma drho + v2sigma2 dsigma (vsigma, b None for LDA)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    drho = np.ascontiguousarray(drho, dtype=np.float64)
    n = len(rho)
    gga = func != "SVWN5"
    sig = np.ascontiguousarray(sigma if gga else np.zeros(n), dtype=np.float64)
    dsig = np.ascontiguousarray(dsigma if gga else np.zeros(n), dtype=np.float64)
    vr, vs, a, b = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    rc = lib().oracle_functional_fxc(FUNC[func], n, rho.ctypes.data, sig.ctypes.data,
                                     drho.ctypes.data, dsig.ctypes.data, vr.ctypes.data,
                                     vs.ctypes.data, a.ctypes.data, b.ctypes.data)
    assert rc == 0
    return vr, (vs if gga else None), a, (b if gga else None)


def fxc_contraction(func, nbf, shells, P, tP, t, weights=None, nthreads=0):
    """(NEL, FXC) of the restated RKS FXC contraction over task dict t."""
    P = np.asfortranarray(P, dtype=np.float64)
    tP = np.asfortranarray(tP, dtype=np.float64)
    pts = np.ascontiguousarray(t["points"], dtype=np.float64)
    w = np.ascontiguousarray(t["weights"] if weights is None else weights, dtype=np.float64)
    sl = np.ascontiguousarray(t["shell_list"], dtype=np.int32)
    fxc = np.zeros((nbf, nbf), order="F")
    nel = ctypes.c_double(0.)
    tasks = _task_array(t)
    rc = lib().oracle_fxc_contraction(FUNC[func], nbf, ctypes.addressof(shells), P.ctypes.data,
                                      nbf, tP.ctypes.data, nbf, len(t["npts"]),
                                      tasks.ctypes.data, pts.ctypes.data, w.ctypes.data,
                                      sl.ctypes.data, ctypes.addressof(nel), fxc.ctypes.data,
                                      nbf, nthreads)
    assert rc == 0
    return nel.value, fxc


def exc_grad_uks(func, nbf, shells, xyz, Ps, Pz, t, weights=None, nthreads=0,
                 weight_derivs=True):
    """UKS EXC gradient (natoms, 3) of the restated host integrator."""
    Ps = np.asfortranarray(Ps, dtype=np.float64)
    Pz = np.asfortranarray(Pz, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    pts = np.ascontiguousarray(t["points"], dtype=np.float64)
    w = np.ascontiguousarray(t["weights"] if weights is None else weights, dtype=np.float64)
    sl = np.ascontiguousarray(t["shell_list"], dtype=np.int32)
    g = np.zeros((len(xyz), 3))
    tasks = _task_array(t)
    rc = lib().oracle_exc_grad_uks(FUNC[func], nbf, ctypes.addressof(shells), len(xyz),
                                   xyz.ctypes.data, Ps.ctypes.data, Pz.ctypes.data, nbf,
                                   len(t["npts"]), tasks.ctypes.data, pts.ctypes.data,
                                   w.ctypes.data, sl.ctypes.data, int(bool(weight_derivs)),
                                   g.ctypes.data, nthreads)
    assert rc == 0
    return g


def fxc_contraction_uks(func, nbf, shells, Ps, Pz, tPs, tPz, t, weights=None, nthreads=0):
    """(FXC_s, FXC_z) of the restated UKS FXC contraction over task dict t."""
    mats = [np.asfortranarray(m, dtype=np.float64) for m in (Ps, Pz, tPs, tPz)]
    pts = np.ascontiguousarray(t["points"], dtype=np.float64)
    w = np.ascontiguousarray(t["weights"] if weights is None else weights, dtype=np.float64)
    sl = np.ascontiguousarray(t["shell_list"], dtype=np.int32)
    fs, fz = np.zeros((nbf, nbf), order="F"), np.zeros((nbf, nbf), order="F")
    tasks = _task_array(t)
    rc = lib().oracle_fxc_contraction_uks(FUNC[func], nbf, ctypes.addressof(shells),
                                          *[m.ctypes.data for m in mats], nbf, len(t["npts"]),
                                          tasks.ctypes.data, pts.ctypes.data, w.ctypes.data,
                                          sl.ctypes.data, fs.ctypes.data, fz.ctypes.data, nbf,
                                          nthreads)
    assert rc == 0
    return fs, fz


def collocation_hessian(shells, shell_list, points):
    """(10, npts, nbe): B, its gradient (x, y, z) and Hessian (xx, xy, xz,
    yy, yz, zz) of the shell list at the points."""
    sl = np.ascontiguousarray(shell_list, dtype=np.int32)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    nbe = 0
    for s in sl:
        sh = shells[int(s)]
        nbe += 2 * sh.l + 1 if sh.pure else (sh.l + 1) * (sh.l + 2) // 2
    out = np.zeros((10, len(pts), nbe))
    ptrs = (ctypes.c_void_p * 10)(*[out[q].ctypes.data for q in range(10)])
    lib().oracle_collocation_hessian(ctypes.addressof(shells), len(sl), sl.ctypes.data, len(pts),
                                     pts.ctypes.data, ptrs)
    return out


def exc_grad(func, nbf, shells, xyz, P, t, weights=None, nthreads=0, weight_derivs=True):
    """RKS EXC gradient (natoms, 3), with the grid-weight derivatives unless
    weight_derivs is False (Hellmann-Feynman form)."""
    P = np.asfortranarray(P, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    pts = np.ascontiguousarray(t["points"], dtype=np.float64)
    w = np.ascontiguousarray(t["weights"] if weights is None else weights, dtype=np.float64)
    sl = np.ascontiguousarray(t["shell_list"], dtype=np.int32)
    g = np.zeros((len(xyz), 3))
    tasks = _task_array(t)
    rc = lib().oracle_exc_grad(FUNC[func], nbf, ctypes.addressof(shells), len(xyz),
                               xyz.ctypes.data, P.ctypes.data, nbf, len(t["npts"]),
                               tasks.ctypes.data, pts.ctypes.data, w.ctypes.data, sl.ctypes.data,
                               int(bool(weight_derivs)), g.ctypes.data, nthreads)
    assert rc == 0
    return g


# Host BLAS libraries tried, in order, for the CPU baseline (MKL's single
# dynamic library first, then OpenBLAS and the reference BLAS of the system).
BLAS_CANDIDATES = ["/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1",
                   "libmkl_rt.so", "libopenblas.so.0", "libopenblas.so", "libblas.so.3"]


def use_blas(path="auto"):
    """Route the oracle's two per-task products through a host BLAS.

    path="auto" tries BLAS_CANDIDATES; None returns to the plain loops.
    Returns the loaded library's path, or None when the loops are in use."""
    L = lib()
    buf = ctypes.create_string_buffer(512)
    if path is None:
        L.oracle_use_blas(None, buf, len(buf))
        return None
    for cand in (BLAS_CANDIDATES if path == "auto" else [path]):
        if L.oracle_use_blas(cand.encode(), buf, len(buf)) == 0:
            return buf.value.decode()
    return None


def gg_ref():
    """GauXC's own gau2grid built from the reference sources, or None."""
    if not os.path.exists(GG_REF):
        return None
    L = ctypes.CDLL(GG_REF)
    vp = ctypes.c_void_p
    L.gg_collocation_deriv1.argtypes = [ctypes.c_int, ctypes.c_ulong, vp, ctypes.c_ulong,
                                        ctypes.c_int, vp, vp, vp, ctypes.c_int, vp, vp, vp, vp]
    L.gg_collocation_deriv1.restype = None
    L.gg_collocation_deriv2.argtypes = [ctypes.c_int, ctypes.c_ulong, vp, ctypes.c_ulong,
                                        ctypes.c_int, vp, vp, vp, ctypes.c_int] + [vp] * 10
    L.gg_collocation_deriv2.restype = None
    return L


GG_SPHERICAL_CCA, GG_CARTESIAN_CCA = 300, 400  # gau2grid.h orderings
