"""Parity of the HIP path with the oracle and with GauXC's golden data.

All calls go through the C ABIs (GauXC C API in libgauxc_amd.so, device C
ABI in libgauxc_amd_device.so).  FP64 bars (the path's parity tolerances):
|dEXC| <= 1e-10 Eh, max|dVXC_ij| <= 1e-9 against the oracle, and GauXC's
own ||VXC - VXC_ref||_F / nbf < 1e-10 against its fixtures.
"""
import numpy as np
import pytest

from conftest import EPS, benzene_setup, load_golden

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_device_functional_matches_oracle(gpu, gx, orc):
    rng = np.random.default_rng(11)
    rho = np.concatenate([10 ** rng.uniform(-18, 2, 3000), [0.0, 1e-15, 1e-12, 2e-12]])
    sigma = rho ** (8 / 3) * 10 ** rng.uniform(-4, 3, len(rho))
    for func in ("SVWN5", "PBE", "PBE0"):
        _flush(f"functional {func}")
        s = None if func == "SVWN5" else sigma
        e_d, v_d, s_d = gx.eval_functional_device(func, rho, s)
        e_o, v_o, s_o = orc.functional(func, rho, s)
        # Device and host libm round cbrt/log/atan differently in the last bits.
        np.testing.assert_allclose(e_d, e_o, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(v_d, v_o, rtol=1e-12, atol=1e-300)
        if s is not None:
            # vsigma = exchange (< 0) + correlation (> 0) cancels at extreme
            # sigma/rho; its natural scale is |rho eps| / sigma.
            scale = np.abs(e_o * rho) / np.maximum(s, 1e-300)
            assert np.all(np.abs(s_d - s_o) <= 1e-12 * np.abs(s_o) + 1e-11 * scale)


def test_device_ssf_matches_fixture_and_oracle(gpu, gx, orc):
    _flush("ssf weights fixture")
    d = load_golden("benzene_weights_ssf.npz")
    parent = np.repeat(d["iParent"], d["npts"])
    w = gx.ssf_weights_device(d["xyz"], d["points"], parent, d["weights_unm"])
    np.testing.assert_allclose(w, d["weights_mod"], rtol=1e-12, atol=0)
    _flush("ssf weights benzene FineGrid vs oracle")
    dd, mol, bs, lb = benzene_setup("svwn5", grid="FineGrid")
    t = lb.get_tasks()
    w_o = orc.ssf_weights(dd["xyz"], t)
    w_d = gx.ssf_weights_device(dd["xyz"], t["points"], np.repeat(t["iparent"], t["npts"]),
                                t["weights"])
    # Same partition; the device forms each P_A as a product over all atoms,
    # GauXC's pairwise loop stops refining factors below 1e-13.
    assert np.abs(w_d - w_o).max() <= 1e-12 * np.abs(w_o).max()


def test_device_ssf_water_cluster_matches_oracle(gpu, gx, orc):
    """SSF weights on an 8-water cluster (24 atoms, FineGrid): most atom
    pairs are far apart relative to the point, the case where the kernel
    skips factors that are exactly 1."""
    _flush("ssf weights 8 waters FineGrid vs oracle")
    mol, bs, P, cfg = gx.workloads.make_config("c1", seed=1, n_water=8, grid="FineGrid")
    lb = gx.workloads.build_tasks(mol, bs, grid="FineGrid")
    t = lb.get_tasks()
    w_o = orc.ssf_weights(cfg["xyz"], t)
    w_d = gx.ssf_weights_device(cfg["xyz"], t["points"], np.repeat(t["iparent"], t["npts"]),
                                t["weights"])
    assert np.abs(w_d - w_o).max() <= 1e-12 * np.abs(w_o).max()


def _context_for(gx, bs, t, P, arena=0):
    ctx = gx.DeviceContext(0, arena_bytes=arena)
    ctx.upload_basis(bs.xcq_shells())
    ctx.upload_tasks(t)
    ctx.upload_density(P)
    return ctx


@pytest.mark.parametrize("precision", ["fp64"])
def test_device_collocation_matches_oracle(gpu, gx, orc, precision):
    _flush(f"collocation {precision}")
    mol, bs, P, cfg = gx.workloads.make_config("c0", seed=0)
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    t = lb.get_tasks()
    ctx = _context_for(gx, bs, t, P)
    sh = orc.shells_from(bs.xcq_shells())
    for task in sorted(set([0, 1, len(t["npts"]) // 2, len(t["npts"]) - 1])):
        s0, n = t["shell_offset"][task], t["nshells"][task]
        p0, m = t["point_offset"][task], t["npts"][task]
        ref = orc.collocation(sh, t["shell_list"][s0:s0 + n], t["points"][p0:p0 + m], deriv=True)
        out = ctx.collocation(task, deriv=True, precision=precision)
        for k in range(4):
            np.testing.assert_allclose(out[k], ref[k].T, rtol=1e-12,
                                       atol=1e-14 * max(1.0, np.abs(ref[k]).max()))


@pytest.mark.parametrize("tag,func", [("svwn5", "SVWN5"), ("pbe0", "PBE0")])
def test_integrator_matches_gauxc_fixture(gpu, gx, tag, func):
    """GauXC's end-to-end test (tests/xc_integrator.cxx:163-216) on the
    device path through the C API."""
    _flush(f"benzene {func} through the C API")
    d, mol, bs, lb = benzene_setup(tag)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated", "Default", "Default",
                                   "Default").get_instance(gx.Functional(func), lb)
    exc, vxc = integ.eval_exc_vxc(d["density"])
    assert abs(exc - d["exc"][0]) <= 1e-10 * abs(d["exc"][0])
    assert np.linalg.norm(vxc - d["vxc"]) / bs.nbf() < 1e-10
    # repeat call propagates state correctly
    exc1, vxc1 = integ.eval_exc_vxc(d["density"])
    assert abs(exc1 - exc) <= 1e-13 * abs(exc)
    assert np.linalg.norm(vxc1 - d["vxc"]) / bs.nbf() < 1e-10
    # EXC-only path equals EXC
    assert abs(integ.eval_exc(d["density"]) - exc) <= 1e-13 * abs(exc)
    # integrate_den: P is the alpha density -> N/2
    assert abs(integ.integrate_den(d["density"]) - 21.0) <= 21.0 * 1e-6


def test_integrator_errors(gpu, gx):
    d, mol, bs, lb = benzene_setup("svwn5", grid="GM3")
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional("SVWN5"),
                                                                       lb)
    with pytest.raises(gx.GauXCError, match="Weights Have Not Been Modified"):
        integ.eval_exc_vxc(d["density"])
    mw = gx.MolecularWeightsFactory("Device", "Default").get_instance()
    mw.modify_weights(lb)
    with pytest.raises(gx.GauXCError, match="Attempting to Overwrite Modified Weights"):
        mw.modify_weights(lb)
    with pytest.raises(gx.GauXCError, match="Same Dimension as Basis"):
        integ.eval_exc_vxc(np.eye(10))


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_device_fp64_matches_oracle_config0(gpu, gx, orc, func):
    """Config 0 molecule/basis/density (3 waters, 84 bf); FineGrid."""
    _flush(f"config0 {func} fp64 vs oracle")
    mol, bs, P, cfg = gx.workloads.make_config("c0", seed=0)
    lb = gx.workloads.build_tasks(mol, bs, grid="FineGrid")
    t = lb.get_tasks()
    w = orc.ssf_weights(cfg["xyz"], t)
    exc_o, nel_o, vxc_o = orc.exc_vxc(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), P, t,
                                      weights=w)
    t2 = dict(t)
    t2["weights"] = w
    ctx = _context_for(gx, bs, t2, P)
    ctx.exc_vxc(func, "fp64", True)
    exc, nel, vxc = ctx.results()
    assert abs(exc - exc_o) <= 1e-10
    assert abs(nel - nel_o) <= 1e-10
    assert np.abs(vxc - vxc_o).max() <= 1e-9
    # deterministic: a second evaluation is bit-identical in EXC
    ctx.exc_vxc(func, "fp64", True)
    exc2, _, vxc2 = ctx.results()
    assert exc2 == exc
    assert np.abs(vxc2 - vxc).max() <= 1e-13


def test_device_chunking_matches_single_chunk(gpu, gx, orc):
    """A small arena splits the tasks into many chunks: same result."""
    _flush("chunking")
    mol, bs, P, cfg = gx.workloads.make_config("c0", seed=1)
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    t = lb.get_tasks()
    big = _context_for(gx, bs, t, P)
    big.exc_vxc("PBE", "fp64", True)
    e1, n1, v1 = big.results()
    small = _context_for(gx, bs, t, P, arena=4 << 20)
    small.exc_vxc("PBE", "fp64", True)
    e2, n2, v2 = small.results()
    assert abs(e1 - e2) <= 1e-12 * abs(e1)
    assert np.abs(v1 - v2).max() <= 1e-12
    with pytest.raises(gx.GauXCError, match="Insufficient memory to even start"):
        _context_for(gx, bs, t, P, arena=1024)


def test_device_collocation_fp32_storage(gpu, gx, orc):
    """Mixed path: collocation evaluated in FP64, stored as FP32."""
    _flush("collocation mixed")
    mol, bs, P, cfg = gx.workloads.make_config("c0", seed=0)
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    t = lb.get_tasks()
    ctx = _context_for(gx, bs, t, P)
    sh = orc.shells_from(bs.xcq_shells())
    for task in (0, len(t["npts"]) - 1):
        s0, n = t["shell_offset"][task], t["nshells"][task]
        p0, m = t["point_offset"][task], t["npts"][task]
        ref = orc.collocation(sh, t["shell_list"][s0:s0 + n], t["points"][p0:p0 + m], deriv=True)
        out = ctx.collocation(task, deriv=True, precision="mixed")
        for k in range(4):
            # FP32 storage: relative rounding 2^-24 of each element, plus the
            # FP64 evaluation's own cancellation noise near zeros of a
            # gradient component (as in the FP64 collocation test).
            r = ref[k].T
            err = np.abs(out[k] - r)
            bound = 6e-8 * np.abs(r) + 1e-14 * max(1.0, np.abs(r).max()) + 1e-38
            worst = np.unravel_index(np.argmax(err - bound), err.shape)
            assert np.all(err <= bound), (k, worst, out[k][worst], r[worst])


@pytest.mark.parametrize("cfgname,n_water,grid,func",
                         [("c0", None, "FineGrid", "SVWN5"), ("c0", None, "FineGrid", "PBE"),
                          ("c1", 4, "UltraFineGrid", "PBE")])
def test_device_mixed_matches_fp64(gpu, gx, orc, cfgname, n_water, grid, func):
    """Mixed precision against the FP64 device path and the oracle at the
    mixed-precision bars: |dEXC| <= 1e-6 Eh, max|dVXC_ij| <= 1e-7."""
    _flush(f"mixed {cfgname} n_water={n_water} {func}")
    mol, bs, P, cfg = gx.workloads.make_config(cfgname, seed=0, n_water=n_water)
    lb = gx.workloads.build_tasks(mol, bs, grid=grid)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    t = lb.get_tasks()
    ctx = _context_for(gx, bs, t, P)
    ctx.exc_vxc(func, "fp64", True)
    e64, n64, v64 = ctx.results()
    ctx.exc_vxc(func, "mixed", True)
    emx, nmx, vmx = ctx.results()
    print(f"dEXC={emx - e64:.3e} max|dVXC|={np.abs(vmx - v64).max():.3e}", flush=True)
    assert abs(emx - e64) <= 1e-6
    assert np.abs(vmx - v64).max() <= 1e-7
    assert np.allclose(vmx, vmx.T, atol=0)
    if cfgname == "c0":
        exc_o, _, vxc_o = orc.exc_vxc(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), P, t)
        assert abs(emx - exc_o) <= 1e-6
        assert np.abs(vmx - vxc_o).max() <= 1e-7


def test_mixed_p32_correction(gpu, gx):
    """The mixed path multiplies with an FP32 copy P32 of P and corrects EXC
    to first order, EXC(P) = EXC(P32) + 2 sum (P - P32) . VXC.  P32 + d with
    |d| a quarter ulp of P32 has the same FP32 copy, so the mixed EXC
    difference between P32 + d and P32 comes from the correction alone; it
    must match the FP64 path's difference (a first-order quantity, second
    order ~1e-15 here).  EXC-only requests give the same corrected EXC."""
    _flush("mixed P32 correction c1 n_water=4 PBE")
    mol, bs, P, cfg = gx.workloads.make_config("c1", seed=0, n_water=4)
    lb = gx.workloads.build_tasks(mol, bs, grid="UltraFineGrid")
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    t = lb.get_tasks()
    P32 = np.asfortranarray(P.astype(np.float32).astype(np.float64))
    Pd = np.asfortranarray(P32 * (1.0 + 2.0 ** -26))
    assert np.array_equal(Pd.astype(np.float32), P32.astype(np.float32))
    ctx = _context_for(gx, bs, t, P32)
    e = {}
    for name, mat in (("base", P32), ("pert", Pd)):
        ctx.upload_density(mat)
        for prec in ("fp64", "mixed"):
            ctx.exc_vxc("PBE", prec, True)
            e[name, prec] = ctx.results(want_vxc=False)[0]
        ctx.exc_vxc("PBE", "mixed", False)
        assert ctx.results(want_vxc=False)[0] == e[name, "mixed"]
    d64 = e["pert", "fp64"] - e["base", "fp64"]
    dmx = e["pert", "mixed"] - e["base", "mixed"]
    print(f"dEXC fp64 {d64:.6e} mixed {dmx:.6e}", flush=True)
    assert abs(d64) > 1e-8
    assert abs(dmx - d64) <= 1e-3 * abs(d64) + 1e-12


def test_mixed_through_gauxc_api(gpu, gx):
    """The "Scheme1-Mixed" local work driver through the C API on GauXC's
    benzene PBE0 fixture."""
    d, mol, bs, lb = benzene_setup("pbe0")
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated", "Default", "Scheme1-Mixed",
                                   "Default").get_instance(gx.Functional("PBE0"), lb)
    exc, vxc = integ.eval_exc_vxc(d["density"])
    assert abs(exc - d["exc"][0]) <= 1e-6
    assert np.abs(vxc - d["vxc"]).max() <= 1e-7


def _high_l_system(gx, lmax):
    """Two atoms carrying every angular momentum up to lmax, spherical and
    Cartesian from l = 2 on, some shells contracted (two primitives, one
    with six: more than the kernel stages per shell in LDS)."""
    Z = [8, 1]
    xyz = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 1.8]])
    l, pure, alpha, coeff, origin = [], [], [], [], []
    for a, r in zip(Z, xyz):
        for am in range(lmax + 1):
            for p in ([False] if am < 2 else [True, False]):
                l.append(am)
                pure.append(p)
                alpha.append([1.7 / (am + 1), 0.45])
                coeff.append([0.6, 0.8])
                origin.append(r)
        l.append(0)
        pure.append(False)
        alpha.append([30.0, 9.0, 3.0, 1.0, 0.35, 0.12])
        coeff.append([0.05, 0.15, 0.3, 0.4, 0.3, 0.1])
        origin.append(r)
    mol = gx.Molecule(Z, xyz)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=1e-10)
    P = gx.workloads.density(bs.nbf(), 4, 7)
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    return mol, bs, P, lb.get_tasks()


@pytest.mark.parametrize("lmax", [3, 4])
def test_device_collocation_high_l(gpu, gx, orc, lmax):
    """Collocation kernels instantiated for the basis' highest angular
    momentum (3 and 4) against the oracle (itself pinned to GauXC's compiled
    gau2grid for l = 0..4), FP64 and FP32 storage."""
    _flush(f"collocation lmax={lmax}")
    mol, bs, P, t = _high_l_system(gx, lmax)
    ctx = _context_for(gx, bs, t, P)
    sh = orc.shells_from(bs.xcq_shells())
    for task in sorted(set([0, len(t["npts"]) // 2, len(t["npts"]) - 1])):
        s0, n = t["shell_offset"][task], t["nshells"][task]
        p0, m = t["point_offset"][task], t["npts"][task]
        ref = orc.collocation(sh, t["shell_list"][s0:s0 + n], t["points"][p0:p0 + m], deriv=True)
        out = ctx.collocation(task, deriv=True, precision="fp64")
        for k in range(4):
            np.testing.assert_allclose(out[k], ref[k].T, rtol=1e-12,
                                       atol=1e-14 * max(1.0, np.abs(ref[k]).max()))
        out = ctx.collocation(task, deriv=True, precision="mixed")
        for k in range(4):
            r = ref[k].T
            bound = 6e-8 * np.abs(r) + 1e-14 * max(1.0, np.abs(r).max()) + 1e-38
            assert np.all(np.abs(out[k] - r) <= bound), (lmax, task, k)


@pytest.mark.parametrize("lmax", [3, 4])
def test_device_fp64_matches_oracle_high_l(gpu, gx, orc, lmax):
    """The whole FP64 pipeline on a basis with g functions and contracted
    shells against the oracle (PBE), and mixed against FP64."""
    _flush(f"pipeline lmax={lmax} fp64 vs oracle")
    mol, bs, P, t = _high_l_system(gx, lmax)
    xyz = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 1.8]])
    w = orc.ssf_weights(xyz, t)
    exc_o, nel_o, vxc_o = orc.exc_vxc("PBE", bs.nbf(), orc.shells_from(bs.xcq_shells()), P, t,
                                      weights=w)
    t2 = dict(t)
    t2["weights"] = w
    ctx = _context_for(gx, bs, t2, P)
    ctx.exc_vxc("PBE", "fp64", True)
    exc, nel, vxc = ctx.results()
    assert abs(exc - exc_o) <= 1e-10
    assert abs(nel - nel_o) <= 1e-10
    assert np.abs(vxc - vxc_o).max() <= 1e-9
    ctx.exc_vxc("PBE", "mixed", True)
    emx, nmx, vmx = ctx.results()
    assert abs(emx - exc) <= 1e-6
    assert np.abs(vmx - vxc).max() <= 1e-7
