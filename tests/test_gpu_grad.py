"""EXC nuclear gradient on the device (GauXC eval_exc_grad(P) and
eval_exc_grad(Ps, Pz)): GauXC's benzene SVWN5 / PBE0 EXC_GRAD and
EXC_GRAD_HELLFEY fixtures through the GauXC API at GauXC's bar
(tests/xc_integrator.cxx:278-300) and tighter; water clusters and a
two-atom basis with g functions against the oracle, RKS and UKS, with and
without the weight derivatives; translational invariance; UKS with P_z = 0
against RKS; the boundary's checks.  (The cytosine UKS fixture is checked in
the UKS tests.)"""
import numpy as np
import pytest

from conftest import benzene_setup

pytestmark = pytest.mark.gpu

GRAD_BAR = 1e-9


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("tag,func", [("svwn5", "SVWN5"), ("pbe0", "PBE0")])
def test_exc_grad_benzene_matches_gauxc_fixture(gx, gpu, tag, func):
    d, mol, bs, lb = benzene_setup(tag)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional(func), lb)
    _flush(f"EXC gradient benzene {func} through the GauXC API")
    g = integ.eval_exc_grad(d["density"])
    ref = d["exc_grad"].reshape(-1, 3)
    err = np.linalg.norm(g - ref) / np.sqrt(g.size)
    _flush(f"|dG|/sqrt(3N)={err:.3e} max|dG|={np.abs(g - ref).max():.3e}")
    assert err < 1e-11  # GauXC's bar is 1e-8
    assert np.abs(g.sum(axis=0)).max() < 1e-11
    gh = integ.eval_exc_grad(d["density"], include_weight_derivatives=False)
    err_h = np.linalg.norm(gh - d["exc_grad_hf"].reshape(-1, 3)) / np.sqrt(g.size)
    _flush(f"Hellmann-Feynman |dG|/sqrt(3N)={err_h:.3e}")
    assert err_h < 1e-11
    # An EXC/VXC evaluation afterwards is unaffected.
    exc, vxc = integ.eval_exc_vxc(d["density"])
    assert abs(exc - d["exc"][0]) <= 1e-10 * abs(d["exc"][0])


def _cluster(gx, nw=4):
    Z, xyz = gx.workloads.water_cluster(nw, 6.0, 7)
    xyz = np.asarray(xyz, dtype=np.float64)
    l, pure, alpha, coeff, origin = gx.workloads.even_tempered_basis(Z, xyz,
                                                                     gx.workloads.BASIS_C0)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=1e-10)
    mol = gx.Molecule(Z, xyz)
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    P = gx.workloads.density(bs.nbf(), 5 * nw, 31)
    return xyz, mol, bs, lb, P


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_exc_grad_water_cluster_matches_oracle(gx, orc, gpu, func):
    xyz, mol, bs, lb, P = _cluster(gx)
    t = lb.get_tasks()
    w = orc.ssf_weights(xyz, t)
    sh = orc.shells_from(bs.xcq_shells())
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional(func), lb)
    for wd in (True, False):
        g_o = orc.exc_grad(func, bs.nbf(), sh, xyz, P, t, weights=w, weight_derivs=wd)
        _flush(f"EXC gradient 4 waters {func} weight derivatives {wd}")
        g = integ.eval_exc_grad(P, include_weight_derivatives=wd)
        _flush(f"max|dG|={np.abs(g - g_o).max():.3e} (max|G|={np.abs(g_o).max():.3e})")
        assert np.abs(g - g_o).max() <= GRAD_BAR
    # The device C ABI directly: EXC and NEL of the same evaluation.
    exc_o, nel_o, _ = orc.exc_vxc(func, bs.nbf(), sh, P, t, weights=w)
    ctx = gx.DeviceContext(0)
    ctx.upload_basis(bs.xcq_shells())
    t2 = dict(t)
    t2["weights"] = w
    ctx.upload_tasks(t2)
    ctx.upload_density(P)
    ctx.exc_grad(func, xyz, weight_derivs=False)
    exc, nel, g3 = ctx.results_grad()
    assert abs(exc - exc_o) <= 1e-10 and abs(nel - nel_o) <= 1e-10
    assert np.abs(g3 - g).max() <= GRAD_BAR  # g: the Hellmann-Feynman form above
    # The weight part's all-atom loops (used above 8,192 atoms) agree with
    # the bounded ones.
    ctx.exc_grad(func, xyz, weight_derivs=True)
    g_l = ctx.results_grad()[2]
    ctx.set_ssf_neighbour_lists(False)
    ctx.exc_grad(func, xyz, weight_derivs=True)
    g_a = ctx.results_grad()[2]
    ctx.set_ssf_neighbour_lists(True)
    _flush(f"lists vs all atoms max|dG|={np.abs(g_l - g_a).max():.3e}")
    assert np.abs(g_l - g_a).max() <= 1e-13
    ctx.exc_vxc(func, "fp64", True)
    with pytest.raises(gx.GauXCError, match="not an EXC gradient"):
        ctx.results_grad()


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_uks_exc_grad_water_cluster_matches_oracle(gx, orc, gpu, func):
    xyz, mol, bs, lb, P = _cluster(gx)
    nbf = bs.nbf()
    pa = P
    pb = gx.workloads.density(nbf, 5 * 4 - 1, 32)
    Ps, Pz = pa + pb, pa - pb
    t = lb.get_tasks()
    w = orc.ssf_weights(xyz, t)
    sh = orc.shells_from(bs.xcq_shells())
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional(func, polarized=True), lb)
    for wd in (True, False):
        g_o = orc.exc_grad_uks(func, nbf, sh, xyz, Ps, Pz, t, weights=w, weight_derivs=wd)
        _flush(f"UKS EXC gradient 4 waters {func} weight derivatives {wd}")
        g = integ.eval_exc_grad(Ps, Pz, include_weight_derivatives=wd)
        _flush(f"max|dG|={np.abs(g - g_o).max():.3e} (max|G|={np.abs(g_o).max():.3e})")
        assert np.abs(g - g_o).max() <= GRAD_BAR
        if wd:
            assert np.abs(g.sum(axis=0)).max() < 1e-11
    # P_z = 0 gives the RKS gradient of P_s / 2.
    rks = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional(func), lb)
    g_r = rks.eval_exc_grad(pa)
    g_u = integ.eval_exc_grad(2 * pa, np.zeros_like(pa))
    assert np.abs(g_u - g_r).max() <= 1e-12 * np.abs(g_r).max()
    # The device C ABI directly.
    ctx = gx.DeviceContext(0)
    ctx.upload_basis(bs.xcq_shells())
    t2 = dict(t)
    t2["weights"] = w
    ctx.upload_tasks(t2)
    ctx.upload_density(Ps)
    with pytest.raises(gx.GauXCError, match="needs P_z"):
        ctx.exc_grad(func, xyz, uks=True)
    ctx.upload_density_z(Pz)
    ctx.exc_grad(func, xyz, weight_derivs=True, uks=True)
    exc, nel, g3 = ctx.results_grad()
    exc_o, nel_o = orc.exc_vxc_uks(func, nbf, sh, Ps, Pz, t, weights=w)[:2]
    assert abs(exc - exc_o) <= 1e-10 and abs(nel - nel_o) <= 1e-10
    assert np.abs(g3 - orc.exc_grad_uks(func, nbf, sh, xyz, Ps, Pz, t, weights=w)).max() <= GRAD_BAR


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
@pytest.mark.parametrize("lmax", [3, 4])
def test_exc_grad_high_l_matches_oracle(gx, orc, gpu, lmax, func):
    """Basis with f and g functions (spherical and Cartesian) and a
    six-primitive shell: the Hessian-vector products of every l."""
    Z = [8, 1]
    xyz = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 1.8]])
    l, pure, alpha, coeff, origin = [], [], [], [], []
    for r in xyz:
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
    t = lb.get_tasks()
    w = orc.ssf_weights(xyz, t)
    g_o = orc.exc_grad(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), xyz, P, t, weights=w)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional(func), lb)
    _flush(f"EXC gradient lmax={lmax} {func}")
    g = integ.eval_exc_grad(P)
    _flush(f"max|dG|={np.abs(g - g_o).max():.3e} (max|G|={np.abs(g_o).max():.3e})")
    assert np.abs(g - g_o).max() <= GRAD_BAR
    # UKS: the full Hessian per function (two coefficient vectors).
    Pb = gx.workloads.density(bs.nbf(), 3, 8)
    gu_o = orc.exc_grad_uks(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), xyz, P + Pb,
                            P - Pb, t, weights=w)
    integ_u = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional(func, polarized=True), lb)
    _flush(f"UKS EXC gradient lmax={lmax}")
    gu = integ_u.eval_exc_grad(P + Pb, P - Pb)
    _flush(f"max|dG|={np.abs(gu - gu_o).max():.3e}")
    assert np.abs(gu - gu_o).max() <= GRAD_BAR


def test_exc_grad_boundary_errors(gx, gpu):
    d, mol, bs, lb = benzene_setup("svwn5")
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    P = d["density"]
    pol = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional("SVWN5", polarized=True), lb)
    with pytest.raises(gx.GauXCError, match="unpolarized"):
        pol.eval_exc_grad(P)
    mixed = gx.XCIntegratorFactory("Device", "Replicated", "Default", "Scheme1-Mixed").get_instance(
        gx.Functional("SVWN5"), lb)
    with pytest.raises(gx.GauXCError, match="FP64"):
        mixed.eval_exc_grad(P)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional("SVWN5"), lb)
    with pytest.raises(gx.GauXCError, match="Same Dimension"):
        integ.eval_exc_grad(P[:-1, :-1])
    with pytest.raises(gx.GauXCError, match="polarized"):
        integ.eval_exc_grad(P, 0 * P)
    mixed_u = gx.XCIntegratorFactory("Device", "Replicated", "Default", "Scheme1-Mixed").get_instance(
        gx.Functional("SVWN5", polarized=True), lb)
    with pytest.raises(gx.GauXCError, match="FP64"):
        mixed_u.eval_exc_grad(P, 0 * P)
