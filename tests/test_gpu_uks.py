"""UKS EXC/VXC on the device (GauXC eval_exc_vxc(Ps, Pz), the generic driver
with spin_dim 2): GauXC's Li / SVWN5 / STO-3G UKS fixture through the GauXC
API at GauXC's bar and tighter (tests/xc_integrator.cxx:218-237, 448-453), a
spin-polarised PBE water dimer against the oracle at the FP64 bars, the
equal-spin case against RKS, the EXC-only path, and the spin checks of the
boundary."""
import numpy as np
import pytest

from conftest import EPS, load_golden

pytestmark = pytest.mark.gpu

FP64_EXC, FP64_VXC = 1e-10, 1e-9


def _flush(msg):
    print(msg, flush=True)


def test_uks_li_svwn5_matches_gauxc_fixture(gx, gpu):
    d = load_golden("li_svwn5_uks.npz")
    mol = gx.Molecule(d["Z"], d["xyz"])
    bs = gx.BasisSet.from_arrays(d["sh_nprim"], d["sh_l"], d["sh_pure"], d["sh_alpha"],
                                 d["sh_coeff"], d["sh_origin"], normalize=False,
                                 shell_tolerance=EPS)
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles",
                                                  "UltraFineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional("SVWN5", polarized=True), lb)
    _flush("UKS Li SVWN5 through the GauXC API")
    exc, vs, vz = integ.eval_exc_vxc(d["density_s"], d["density_z"])
    _flush(f"dEXC={exc - d['exc'][0]:.3e} "
           f"|dVs|/nbf={np.linalg.norm(vs - d['vxc_s']) / bs.nbf():.3e} "
           f"|dVz|/nbf={np.linalg.norm(vz - d['vxc_z']) / bs.nbf():.3e}")
    assert abs(exc - d["exc"][0]) <= 1e-10 * abs(d["exc"][0])
    assert np.linalg.norm(vs - d["vxc_s"]) / bs.nbf() < 1e-10
    assert np.linalg.norm(vz - d["vxc_z"]) / bs.nbf() < 1e-10
    # State propagates: a second call gives the same result; EXC-only agrees.
    exc1, vs1, vz1 = integ.eval_exc_vxc(d["density_s"], d["density_z"])
    assert exc1 == exc and np.array_equal(vs1, vs) and np.array_equal(vz1, vz)
    assert abs(integ.eval_exc(d["density_s"], d["density_z"]) - exc) <= 1e-14 * abs(exc)


def _dimer(gx):
    Z, xyz = gx.workloads.water_cluster(2, 5.0, 4)
    l, pure, alpha, coeff, origin = gx.workloads.even_tempered_basis(Z, xyz,
                                                                     gx.workloads.BASIS_C0)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=1e-10)
    mol = gx.Molecule(Z, xyz)
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    nbf = bs.nbf()
    pa = gx.workloads.density(nbf, 11, 21) / 2
    pb = gx.workloads.density(nbf, 9, 22) / 2
    return xyz, mol, bs, lb, pa, pb


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_uks_dimer_matches_oracle(gx, orc, gpu, func):
    xyz, mol, bs, lb, pa, pb = _dimer(gx)
    t = lb.get_tasks()
    w = orc.ssf_weights(xyz, t)
    ps, pz = pa + pb, pa - pb
    exc_o, nel_o, vs_o, vz_o = orc.exc_vxc_uks(func, bs.nbf(), orc.shells_from(bs.xcq_shells()),
                                               ps, pz, t, weights=w)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional(func, polarized=True), lb)
    _flush(f"UKS water dimer {func}")
    exc, vs, vz = integ.eval_exc_vxc(ps, pz)
    _flush(f"dEXC={exc - exc_o:.3e} max|dVs|={np.abs(vs - vs_o).max():.3e} "
           f"max|dVz|={np.abs(vz - vz_o).max():.3e}")
    assert abs(exc - exc_o) <= FP64_EXC
    assert np.abs(vs - vs_o).max() <= FP64_VXC
    assert np.abs(vz - vz_o).max() <= FP64_VXC
    assert np.array_equal(vs, vs.T) and np.array_equal(vz, vz.T)
    # Device context: NEL of the UKS evaluation against the oracle's.
    ctx = gx.DeviceContext(0)
    ctx.upload_basis(bs.xcq_shells())
    ctx.upload_tasks(lb.get_tasks())
    ctx.upload_density(ps)
    ctx.upload_density_z(pz)
    ctx.exc_vxc_uks(func, True)
    e2, n2, _, _ = ctx.results_uks()
    assert abs(e2 - exc) <= 1e-12 * abs(exc) and abs(n2 - nel_o) <= 1e-10


def test_uks_equal_spins_equals_rks(gx, gpu):
    xyz, mol, bs, lb, pa, pb = _dimer(gx)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    rks = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional("PBE"), lb)
    uks = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional("PBE", polarized=True), lb)
    e_r, v_r = rks.eval_exc_vxc(pa)
    e_u, vs, vz = uks.eval_exc_vxc(2 * pa, np.zeros_like(pa))
    assert abs(e_u - e_r) <= 1e-12 * abs(e_r)
    assert np.abs(vs - v_r).max() <= 1e-12 * np.abs(v_r).max()
    assert np.abs(vz).max() <= 1e-13 * np.abs(v_r).max()


def test_uks_boundary_checks(gx, gpu):
    xyz, mol, bs, lb, pa, pb = _dimer(gx)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    pol = gx.Functional("PBE", polarized=True)
    unpol = gx.Functional("PBE")
    with pytest.raises(gx.GauXCError, match="needs a polarized functional"):
        gx.XCIntegratorFactory("Device", "Replicated").get_instance(unpol, lb).eval_exc_vxc(
            pa + pb, pa - pb)
    with pytest.raises(gx.GauXCError, match="needs an unpolarized functional"):
        gx.XCIntegratorFactory("Device", "Replicated").get_instance(pol, lb).eval_exc_vxc(pa)
    with pytest.raises(gx.GauXCError, match="FP64"):
        gx.XCIntegratorFactory("Device", "Replicated", "Default", "Scheme1-Mixed",
                               "Default").get_instance(pol, lb).eval_exc_vxc(pa + pb, pa - pb)


def test_uks_cytosine_matches_gauxc_fixture_and_oracle(gx, orc, gpu):
    """Doublet cytosine / SVWN5 / cc-pVDZ on the Robust-pruned ultra-fine grid
    (tests/xc_integrator.cxx:218-237,278-300,455-459) through the GauXC API:
    EXC / VXC against GauXC's values at GauXC's bars and against the oracle
    at the FP64 bars; the full UKS gradient against the fixture (1e-12) and
    the oracle; the Hellmann-Feynman gradient against the oracle (the
    fixture's EXC, VXC_s and Hellmann-Feynman gradient sit 2.6e-11
    relative, 6.5e-12 and 2.8e-9 from the oracle, which its own test
    records)."""
    from conftest import cytosine_uks_setup
    d, mol, bs, lb = cytosine_uks_setup()
    t = lb.get_tasks()
    xyz = np.asarray(d["xyz"], dtype=np.float64)
    n = bs.nbf()
    w = orc.ssf_weights(xyz, t)
    sh = orc.shells_from(bs.xcq_shells())
    Ps, Pz = d["density_s"], d["density_z"]
    exc_o, _, vs_o, vz_o = orc.exc_vxc_uks("SVWN5", n, sh, Ps, Pz, t, weights=w)[:4]
    g_o = orc.exc_grad_uks("SVWN5", n, sh, xyz, Ps, Pz, t, weights=w)
    gh_o = orc.exc_grad_uks("SVWN5", n, sh, xyz, Ps, Pz, t, weights=w, weight_derivs=False)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional("SVWN5", polarized=True), lb)
    _flush("UKS cytosine SVWN5 (Robust grid) through the GauXC API")
    exc, vs, vz = integ.eval_exc_vxc(Ps, Pz)
    _flush(f"dEXC(fixture)={exc - d['exc'][0]:.3e} dEXC(oracle)={exc - exc_o:.3e} "
           f"max|dVs|={np.abs(vs - vs_o).max():.3e} max|dVz|={np.abs(vz - vz_o).max():.3e}")
    assert abs(exc - d["exc"][0]) <= 1e-10 * abs(exc)
    assert np.linalg.norm(vs - d["vxc_s"]) / n < 1e-10
    assert np.linalg.norm(vz - d["vxc_z"]) / n < 1e-10
    assert abs(exc - exc_o) <= FP64_EXC
    assert np.abs(vs - vs_o).max() <= FP64_VXC and np.abs(vz - vz_o).max() <= FP64_VXC
    _flush("UKS cytosine EXC gradient")
    g = integ.eval_exc_grad(Ps, Pz)
    gh = integ.eval_exc_grad(Ps, Pz, include_weight_derivatives=False)
    _flush(f"max|dG|(fixture)={np.abs(g.ravel() - d['exc_grad']).max():.3e} "
           f"max|dG|(oracle)={np.abs(g - g_o).max():.3e} "
           f"max|dG_HF|(oracle)={np.abs(gh - gh_o).max():.3e}")
    assert np.abs(g.ravel() - d["exc_grad"]).max() < 1e-12
    assert np.abs(g - g_o).max() < 1e-12
    assert np.abs(g.sum(axis=0)).max() < 1e-12
    assert np.abs(gh - gh_o).max() < 1e-12
    assert np.linalg.norm(gh.ravel() - d["exc_grad_hf"]) / np.sqrt(gh.size) < 1e-8
