"""FXC contraction on the device (GauXC eval_fxc_contraction): RKS against
GauXC's benzene SVWN5 / PBE0 FXC fixtures through the GauXC API at GauXC's
bar (tests/2nd_derivative_test.cxx:121-126) and tighter, a water cluster
against the oracle at the FP64 VXC bar, linearity in the trial density, and
the boundary's argument and state checks; UKS against GauXC's Li fixture
(:128-134) and a spin-polarised water dimer against the oracle; both
contractions and the UKS EXC/VXC with f and g shells (in-stage and
stored-plane basis evaluation) against the oracle."""
import numpy as np
import pytest

from conftest import EPS, benzene_setup, load_golden

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


def _li(gx):
    d = load_golden("li_svwn5_uks.npz")
    mol = gx.Molecule(d["Z"], d["xyz"])
    bs = gx.BasisSet.from_arrays(d["sh_nprim"], d["sh_l"], d["sh_pure"], d["sh_alpha"],
                                 d["sh_coeff"], d["sh_origin"], normalize=False,
                                 shell_tolerance=EPS)
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles",
                                                  "UltraFineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    return d, bs, lb


def test_fxc_uks_li_matches_gauxc_fixture(gx, gpu):
    """GauXC's UKS FXC check (tests/2nd_derivative_test.cxx:128-134): Li /
    SVWN5 / STO-3G, ||FXC_s - ref|| / nbf and ||FXC_z - ref|| / nbf < 1e-10."""
    d, bs, lb = _li(gx)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional("SVWN5", polarized=True), lb)
    _flush("UKS FXC Li SVWN5 through the GauXC API")
    fs, fz = integ.eval_fxc_contraction(d["density_s"], d["trial_density_s"], d["density_z"],
                                        d["trial_density_z"])
    es = np.linalg.norm(fs - d["fxc_s"]) / bs.nbf()
    ez = np.linalg.norm(fz - d["fxc_z"]) / bs.nbf()
    _flush(f"|dFXC_s|/nbf={es:.3e} |dFXC_z|/nbf={ez:.3e}")
    assert es < 1e-12 and ez < 1e-12  # GauXC's bar is 1e-10


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_fxc_uks_dimer_matches_oracle(gx, orc, gpu, func):
    Z, xyz = gx.workloads.water_cluster(2, 5.0, 4)
    l, pure, alpha, coeff, origin = gx.workloads.even_tempered_basis(Z, xyz,
                                                                     gx.workloads.BASIS_C0)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=1e-10)
    mol = gx.Molecule(Z, xyz)
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    nbf = bs.nbf()
    pa = gx.workloads.density(nbf, 11, 21) / 2
    pb = gx.workloads.density(nbf, 9, 22) / 2
    rng = np.random.default_rng(6)
    ta, tb = rng.standard_normal((2, nbf, nbf)) * 0.01
    ta, tb = ta + ta.T, tb + tb.T
    Ps, Pz, tPs, tPz = pa + pb, pa - pb, ta + tb, ta - tb
    t = lb.get_tasks()
    w = orc.ssf_weights(xyz, t)
    fs_o, fz_o = orc.fxc_contraction_uks(func, nbf, orc.shells_from(bs.xcq_shells()), Ps, Pz,
                                         tPs, tPz, t, weights=w)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional(func, polarized=True), lb)
    _flush(f"UKS FXC water dimer {func}")
    fs, fz = integ.eval_fxc_contraction(Ps, tPs, Pz, tPz)
    _flush(f"max|dFXC_s|={np.abs(fs - fs_o).max():.3e} max|dFXC_z|={np.abs(fz - fz_o).max():.3e}")
    assert np.abs(fs - fs_o).max() <= FP64_VXC
    assert np.abs(fz - fz_o).max() <= FP64_VXC
    assert np.array_equal(fs, fs.T) and np.array_equal(fz, fz.T)
    # An unpolarized functional is refused for the UKS form.
    rks = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional(func), lb)
    with pytest.raises(gx.GauXCError, match="polarized"):
        rks.eval_fxc_contraction(Ps, tPs, Pz, tPz)


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
@pytest.mark.parametrize("lmax", [3, 4])
def test_derivative_paths_high_l_match_oracle(gx, orc, gpu, lmax, func):
    """RKS and UKS FXC contractions and the UKS EXC/VXC on a two-atom basis
    with f (and g) shells, spherical and Cartesian, and a six-primitive s
    shell: up to f the density stage evaluates the basis itself and Z comes
    from the per-point coefficients (in-stage mode); with g shells both read
    the stored collocation planes."""
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
    nbf = bs.nbf()
    pa = gx.workloads.density(nbf, 4, 7)
    pb = gx.workloads.density(nbf, 3, 8)
    rng = np.random.default_rng(lmax)
    ta, tb = rng.standard_normal((2, nbf, nbf)) * 0.01
    ta, tb = ta + ta.T, tb + tb.T
    lb = gx.workloads.build_tasks(mol, bs, grid="GM3")
    t = lb.get_tasks()
    w = orc.ssf_weights(xyz, t)
    sh = orc.shells_from(bs.xcq_shells())
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    rks = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional(func), lb)
    uks = gx.XCIntegratorFactory("Device", "Replicated").get_instance(
        gx.Functional(func, polarized=True), lb)
    _flush(f"FXC lmax={lmax}")
    f = rks.eval_fxc_contraction(pa, ta)
    f_o = orc.fxc_contraction(func, nbf, sh, pa, ta, t, weights=w)[1]
    _flush(f"max|dFXC|={np.abs(f - f_o).max():.3e}")
    assert np.abs(f - f_o).max() <= FP64_VXC
    _flush(f"UKS FXC lmax={lmax}")
    fs, fz = uks.eval_fxc_contraction(pa + pb, ta + tb, pa - pb, ta - tb)
    fs_o, fz_o = orc.fxc_contraction_uks(func, nbf, sh, pa + pb, pa - pb, ta + tb, ta - tb, t,
                                         weights=w)
    _flush(f"max|dFXC_s|={np.abs(fs - fs_o).max():.3e} max|dFXC_z|={np.abs(fz - fz_o).max():.3e}")
    assert np.abs(fs - fs_o).max() <= FP64_VXC and np.abs(fz - fz_o).max() <= FP64_VXC
    _flush(f"UKS EXC/VXC lmax={lmax}")
    exc, vs, vz = uks.eval_exc_vxc(pa + pb, pa - pb)
    exc_o, _, vs_o, vz_o = orc.exc_vxc_uks(func, nbf, sh, pa + pb, pa - pb, t, weights=w)[:4]
    _flush(f"dEXC={exc - exc_o:.3e} max|dVs|={np.abs(vs - vs_o).max():.3e} "
           f"max|dVz|={np.abs(vz - vz_o).max():.3e}")
    assert abs(exc - exc_o) <= 1e-10
    assert np.abs(vs - vs_o).max() <= FP64_VXC and np.abs(vz - vz_o).max() <= FP64_VXC
