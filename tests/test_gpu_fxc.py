"""RKS FXC contraction on the device (GauXC eval_fxc_contraction(P, tP)):
GauXC's benzene SVWN5 / PBE0 FXC fixtures through the GauXC API at GauXC's
bar (tests/2nd_derivative_test.cxx:121-126) and tighter, a water cluster
against the oracle at the FP64 VXC bar, linearity in the trial density, and
the boundary's argument and state checks."""
import numpy as np
import pytest

from conftest import benzene_setup

pytestmark = pytest.mark.gpu

FP64_VXC = 1e-9


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("tag,func", [("svwn5", "SVWN5"), ("pbe0", "PBE0")])
def test_fxc_benzene_matches_gauxc_fixture(gx, gpu, tag, func):
    d, mol, bs, lb = benzene_setup(tag)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional(func), lb)
    _flush(f"FXC benzene {func} through the GauXC API")
    fxc = integ.eval_fxc_contraction(d["density"], d["trial_density"])
    err = np.linalg.norm(fxc - d["fxc"]) / bs.nbf()
    _flush(f"|dFXC|/nbf={err:.3e} max|dFXC|={np.abs(fxc - d['fxc']).max():.3e}")
    assert err < 1e-12  # GauXC's bar is 1e-10
    assert np.array_equal(fxc, fxc.T)
    # State propagates: the same call again gives the same matrix (to the
    # rounding of the scatter's atomic adds, whose order varies); an EXC/VXC
    # evaluation afterwards is unaffected.
    fxc1 = integ.eval_fxc_contraction(d["density"], d["trial_density"])
    assert np.abs(fxc1 - fxc).max() <= 1e-14 * np.abs(fxc).max()
    exc, vxc = integ.eval_exc_vxc(d["density"])
    assert abs(exc - d["exc"][0]) <= 1e-10 * abs(d["exc"][0])
    assert np.linalg.norm(vxc - d["vxc"]) / bs.nbf() < 1e-10


def _cluster(gx, nw=4):
    Z, xyz = gx.workloads.water_cluster(nw, 6.0, 7)
    l, pure, alpha, coeff, origin = gx.workloads.even_tempered_basis(Z, xyz,
                                                                     gx.workloads.BASIS_C0)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=1e-10)
    mol = gx.Molecule(Z, xyz)
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    nbf = bs.nbf()
    P = gx.workloads.density(nbf, 5 * nw, 31)
    rng = np.random.default_rng(32)
    tP = rng.standard_normal((nbf, nbf)) * 0.01
    return xyz, mol, bs, lb, P, tP + tP.T


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_fxc_water_cluster_matches_oracle(gx, orc, gpu, func):
    xyz, mol, bs, lb, P, tP = _cluster(gx)
    t = lb.get_tasks()
    w = orc.ssf_weights(xyz, t)
    nel_o, fxc_o = orc.fxc_contraction(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), P, tP,
                                       t, weights=w)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional(func), lb)
    _flush(f"FXC 4 waters {func}")
    fxc = integ.eval_fxc_contraction(P, tP)
    _flush(f"max|dFXC|={np.abs(fxc - fxc_o).max():.3e} (max|FXC|={np.abs(fxc_o).max():.3e})")
    assert np.abs(fxc - fxc_o).max() <= FP64_VXC
    assert np.array_equal(fxc, fxc.T)
    # Linear in the trial density (a size-independent property).
    fxc2 = integ.eval_fxc_contraction(P, -2.0 * tP)
    assert np.abs(fxc2 + 2.0 * fxc).max() <= 1e-12 * np.abs(fxc).max()
    # The device C ABI directly: same FXC, and the integrated density of P.
    ctx = gx.DeviceContext(0)
    ctx.upload_basis(bs.xcq_shells())
    ctx.upload_tasks(lb.get_tasks())
    ctx.upload_density(P)
    ctx.upload_trial_density(tP)
    ctx.fxc_contraction(func)
    nel, fxc3 = ctx.results_fxc()
    assert abs(nel - nel_o) <= 1e-10 * abs(nel_o)
    assert np.abs(fxc3 - fxc).max() <= 1e-13 * np.abs(fxc).max()
    # After an EXC/VXC evaluation the context holds no FXC to download.
    ctx.exc_vxc(func, "fp64", True)
    with pytest.raises(gx.GauXCError, match="not an FXC"):
        ctx.results_fxc()


def test_fxc_boundary_errors(gx, gpu):
    d, mol, bs, lb = benzene_setup("svwn5")
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    P, tP = d["density"], d["trial_density"]
    pol = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional("SVWN5", polarized=True), lb)
    with pytest.raises(gx.GauXCError, match="unpolarized"):
        pol.eval_fxc_contraction(P, tP)
    mixed = gx.XCIntegratorFactory("Device", "Replicated", "Default", "Scheme1-Mixed").get_instance(
        gx.Functional("SVWN5"), lb)
    with pytest.raises(gx.GauXCError, match="FP64"):
        mixed.eval_fxc_contraction(P, tP)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional("SVWN5"), lb)
    with pytest.raises(gx.GauXCError, match="Same Dimension"):
        integ.eval_fxc_contraction(P[:-1, :-1], tP[:-1, :-1])
