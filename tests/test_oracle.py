"""Pins the CPU oracle against GauXC's own golden data and against
mathematics, so that it can serve as the checker of the HIP path.

Each check is chosen so that a plausible mistake fails it:
  * benzene SVWN5 / PBE0 EXC+VXC (GauXC tests/xc_integrator.cxx:199-212):
    grid, SSF, collocation, X = 2BP, uvvar, functional (incl. the VWN
    variant and the PBE constants), Z, SYR2K, scatter and symmetrisation;
  * SSF weights of 15 tasks (tests/weights_host.hpp:22-57);
  * gau2grid collocation + gradient (tests/collocation_host.hpp:111-180),
    and GauXC's gau2grid itself compiled under oracle/_ref for l = 0..4;
  * functional derivatives by central differences, and closed-form limits;
  * UKS: GauXC's Li / SVWN5 UKS fixture; the polarised functionals' unpolarised
    limit and finite-difference derivatives; UKS with equal spins equals RKS;
    VXC_s / VXC_z are the finite-difference derivatives of the UKS EXC;
  * FXC contraction: GauXC's benzene SVWN5 / PBE0 FXC fixtures
    (tests/2nd_derivative_test.cxx:121-126) and the Li UKS FXC fixture;
    the second derivatives by differences of the first; FXC as the
    derivative of VXC along tP (RKS and UKS); UKS with P_z = 0 equals RKS;
  * EXC gradient: Hessian collocation against gau2grid's deriv2 and central
    differences; GauXC's benzene EXC_GRAD (with grid-weight derivatives) and
    EXC_GRAD_HELLFEY fixtures; translational invariance; dEXC/dR by central
    differences (RKS and UKS); UKS with P_z = 0 equals RKS; GauXC's doublet
    cytosine UKS fixture (EXC, VXC_s, VXC_z, both gradients);
  * pruned grids: GauXC's benzene Robust and Treutler fixtures (point
    counts, EXC, VXC, FXC, both gradients).
"""
import ctypes

import numpy as np
import pytest

from conftest import EPS, benzene_setup, load_golden


@pytest.mark.parametrize("tag,func", [("svwn5", "SVWN5"), ("pbe0", "PBE0")])
def test_oracle_matches_gauxc_benzene_fixture(orc, tag, func):
    d, mol, bs, lb = benzene_setup(tag)
    t = lb.get_tasks()
    w = orc.ssf_weights(d["xyz"], t)
    exc, nel, vxc = orc.exc_vxc(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), d["density"],
                                t, weights=w)
    # GauXC checks EXC to Catch's default relative epsilon and
    # ||VXC - VXC_ref||_F / nbf < 1e-10; the oracle meets far tighter bounds.
    assert abs(exc - d["exc"][0]) <= 1e-10 * abs(d["exc"][0])
    assert np.linalg.norm(vxc - d["vxc"]) / bs.nbf() < 1e-12
    assert abs(nel - 42.0) < 1e-4  # 2 x alpha electrons of benzene
    assert np.allclose(vxc, vxc.T, atol=0)


def test_oracle_ssf_matches_gauxc_weights_fixture(orc):
    d = load_golden("benzene_weights_ssf.npz")
    npts = d["npts"]
    t = dict(iparent=d["iParent"], npts=npts, nshells=np.zeros_like(npts),
             nbe=np.zeros_like(npts), dist_nearest=d["dist_nearest"],
             point_offset=np.concatenate([[0], np.cumsum(npts)[:-1]]),
             shell_offset=np.zeros_like(npts), points=d["points"], weights=d["weights_unm"])
    w = orc.ssf_weights(d["xyz"], t)
    ref = d["weights_mod"]
    assert np.all(np.abs(w - ref) <= 1e-12 * np.abs(ref) + 1e-300)
    # GauXC saved the 15 tasks with the smallest weights: points inside the
    # 0.5(1-a) R_nearest sphere, whose weight SSF leaves unchanged.  This
    # fixture therefore pins the inner-sphere rule; the Frisch-polynomial
    # part is pinned by the benzene EXC/VXC fixtures and by
    # test_oracle_ssf_partition_of_unity.
    assert np.array_equal(ref, d["weights_unm"])


def test_oracle_ssf_partition_of_unity(gx, orc):
    """SSF weights must partition unity: the molecular quadrature then
    integrates a normalised Gaussian on every atom to 1 (FineGrid accuracy).
    A wrong Frisch polynomial, sign of mu or normalisation breaks this."""
    d = load_golden("benzene_svwn5.npz")
    mol = gx.Molecule(d["Z"], d["xyz"])
    bs = gx.BasisSet.from_arrays(d["sh_nprim"], d["sh_l"], d["sh_pure"], d["sh_alpha"],
                                 d["sh_coeff"], d["sh_origin"], normalize=False)
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles",
                                                  "FineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    t = lb.get_tasks()
    w = orc.ssf_weights(d["xyz"], t)
    assert np.sum(w != t["weights"]) > 10000
    pts = t["points"]
    for a in (0.3, 1.0):
        f = np.zeros(len(w))
        for R in d["xyz"]:
            f += np.exp(-a * np.sum((pts - R) ** 2, axis=1))
        integral = np.dot(w, f) / (len(d["xyz"]) * (np.pi / a) ** 1.5)
        assert abs(integral - 1.0) < 1e-4  # FineGrid quadrature error ~1e-5
        # Without the partition (raw atomic weights) the atoms overlap.
        raw = np.dot(t["weights"], f) / (len(d["xyz"]) * (np.pi / a) ** 1.5)
        assert abs(raw - 1.0) > 1e-2


def _water_shells(gx, d):
    bs = gx.BasisSet.from_arrays(d["raw_nprim"], d["raw_l"], d["raw_pure"], d["raw_alpha"],
                                 d["raw_coeff"], d["raw_origin"], normalize=True,
                                 shell_tolerance=1e-6)
    return bs


def test_oracle_collocation_matches_gau2grid_fixture(gx, orc):
    d = load_golden("water_collocation.npz")
    bs = _water_shells(gx, d)
    sh = orc.shells_from(bs.xcq_shells())
    for e in range(len(d["pts_off"]) - 1):
        pts = d["pts"][d["pts_off"][e]:d["pts_off"][e + 1]]
        mask = d["mask"][d["mask_off"][e]:d["mask_off"][e + 1]]
        out = orc.collocation(sh, mask, pts, deriv=True)
        sl = slice(d["eval_off"][e], d["eval_off"][e + 1])
        for k, key in enumerate(["eval", "deval_x", "deval_y", "deval_z"]):
            ref = d[key][sl].reshape(len(pts), -1)
            assert np.abs(out[k] - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def _random_shells(orc, rng, lmax=4, pure=True):
    shells = (orc.OracleShell * (lmax + 1))()
    for l in range(lmax + 1):
        s = shells[l]
        s.nprim, s.l, s.pure = 3, l, int(pure)
        for k, a in enumerate([3.1, 0.9, 0.21]):
            s.alpha[k] = a
            s.coeff[k] = rng.uniform(0.2, 1.0)
        for k in range(3):
            s.origin[k] = rng.uniform(-1, 1)
    return shells


@pytest.mark.parametrize("pure", [True, False])
def test_oracle_collocation_matches_compiled_gau2grid(orc, pure):
    gg = orc.gg_ref()
    if gg is None:
        pytest.skip("oracle/_ref not built (needs the GauXC source tree)")
    rng = np.random.default_rng(7)
    shells = _random_shells(orc, rng, 4, pure)
    pts = rng.uniform(-2.5, 2.5, (57, 3))
    for l in range(5):
        s = shells[l]
        out = orc.collocation(shells, [l], pts, deriv=True)
        nc = 2 * l + 1 if pure else (l + 1) * (l + 2) // 2
        ref = np.zeros((4, nc, len(pts)))
        coeff = np.array(s.coeff[:3])
        alpha = np.array(s.alpha[:3])
        center = np.array(s.origin[:3])
        P = ctypes.c_void_p
        gg.gg_collocation_deriv1(l, len(pts), np.ascontiguousarray(pts).ctypes.data, 3, 3,
                                 coeff.ctypes.data, alpha.ctypes.data, center.ctypes.data,
                                 orc.GG_SPHERICAL_CCA if pure else orc.GG_CARTESIAN_CCA,
                                 *[P(ref[k].ctypes.data) for k in range(4)])
        for k in range(4):
            np.testing.assert_allclose(out[k], ref[k].T, rtol=1e-12, atol=1e-14)


def _fd(func, orc, rho, sigma, h_rel=1e-5):
    e = lambda r, s: orc.functional(func, r, s)[0] * r  # noqa: E731
    hr = h_rel * rho
    d_rho = (e(rho + hr, sigma) - e(rho - hr, sigma)) / (2 * hr)
    d_sig = None
    if sigma is not None:
        hs = h_rel * np.maximum(sigma, 1e-8)
        d_sig = (e(rho, sigma + hs) - e(rho, sigma - hs)) / (2 * hs)
    return d_rho, d_sig


@pytest.mark.parametrize("func", ["SVWN5", "PBE", "PBE0"])
def test_oracle_functional_derivatives_by_finite_difference(orc, func):
    rng = np.random.default_rng(3)
    rho = 10 ** rng.uniform(-4, 1.5, 400)
    sigma = None if func == "SVWN5" else rho ** (8 / 3) * 10 ** rng.uniform(-3, 1.5, 400)
    eps, vr, vs = orc.functional(func, rho, sigma)
    d_rho, d_sig = _fd(func, orc, rho, sigma)
    e = np.abs(eps * rho)
    # Central differences of e = rho*eps: truncation ~h^2, round-off ~u|e|/h.
    assert np.all(np.abs(vr - d_rho) <= 1e-7 * np.abs(vr) + 1e-8 * e / rho)
    if sigma is not None:
        assert np.all(np.abs(vs - d_sig) <= 1e-6 * np.abs(vs) + 1e-8 * e / sigma)


def test_oracle_functional_limits(orc):
    rho = np.array([1e-3, 0.05, 0.7, 12.0])
    # PBE exchange at zero gradient is Slater exchange; PBE0 XC at zero
    # gradient is 0.75 Slater + PBE correlation at t = 0 (= PW92).
    e_lda, v_lda, _ = orc.functional("SVWN5", rho)
    e_pbe, v_pbe, vs = orc.functional("PBE", rho, np.zeros_like(rho))
    e_pbe0, _, _ = orc.functional("PBE0", rho, np.zeros_like(rho))
    slater = -0.75 * (3 / np.pi) ** (1 / 3) * rho ** (1 / 3)
    ec_pw = e_pbe - slater
    np.testing.assert_allclose(e_pbe0, 0.75 * slater + ec_pw, rtol=1e-14)
    # PW92 paramagnetic correlation at rs = 1 (Perdew & Wang 1992, Table):
    # -0.0598 Ha; with the modified parameters -0.05975 to 4 digits.
    rs1 = 3 / (4 * np.pi)
    e1, _, _ = orc.functional("PBE", np.array([rs1]), np.zeros(1))
    ec1 = e1[0] + 0.75 * (3 / np.pi) ** (1 / 3) * rs1 ** (1 / 3)
    assert abs(ec1 + 0.05975) < 5e-5
    # Large gradient: PBE exchange enhancement saturates at 1 + kappa and the
    # correlation vanishes (H -> -eps_c).
    big = rho ** (8 / 3) * 1e12
    e_big, _, _ = orc.functional("PBE", rho, big)
    np.testing.assert_allclose(e_big, 1.804 * slater, rtol=1e-8)
    # Below the density threshold everything is zero.
    e0, v0, s0 = orc.functional("PBE", np.array([1e-16]), np.array([1e-20]))
    assert e0[0] == 0 and v0[0] == 0 and s0[0] == 0
    # VWN(RPA) at rs = 1: -0.0790 Ha (the RPA fit lies below VWN5's -0.0600).
    e_v, _, _ = orc.functional("SVWN5", np.array([rs1]))
    ecv = e_v[0] + 0.75 * (3 / np.pi) ** (1 / 3) * rs1 ** (1 / 3)
    assert -0.0800 < ecv < -0.0780
    assert np.all(v_lda < 0) and np.all(v_pbe < 0)


def test_oracle_uks_matches_gauxc_li_fixture(gx, orc):
    """UKS: Li / SVWN5 / STO-3G on the unpruned ultra-fine grid (GauXC
    tests/xc_integrator.cxx:218-237, 448-453): pins the spin-resolved
    densities, the polarised Slater + VWN(RPA) with its spin interpolation,
    the weight fold, Z_s / Z_z and both potentials."""
    d = load_golden("li_svwn5_uks.npz")
    mol = gx.Molecule(d["Z"], d["xyz"])
    bs = gx.BasisSet.from_arrays(d["sh_nprim"], d["sh_l"], d["sh_pure"], d["sh_alpha"],
                                 d["sh_coeff"], d["sh_origin"], normalize=False,
                                 shell_tolerance=EPS)
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles",
                                                  "UltraFineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    t = lb.get_tasks()
    exc, nel, vs, vz = orc.exc_vxc_uks("SVWN5", bs.nbf(), orc.shells_from(bs.xcq_shells()),
                                       d["density_s"], d["density_z"], t)
    assert abs(exc - d["exc"][0]) <= 1e-10 * abs(d["exc"][0])
    assert np.linalg.norm(vs - d["vxc_s"]) / bs.nbf() < 1e-11
    assert np.linalg.norm(vz - d["vxc_z"]) / bs.nbf() < 1e-11
    assert abs(nel - 3.0) < 1e-9


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_oracle_polarised_functional_unpolarised_limit(orc, func):
    """rho_a = rho_b = rho/2, every sigma = sigma/4: the polarised functional
    is the unpolarised one (eps, vrho_a = vrho_b = vrho, and the sigma
    derivatives sum to 4 vsigma along that path)."""
    rng = np.random.default_rng(11)
    rho = 10 ** rng.uniform(-4, 1.5, 300)
    sigma = None if func == "SVWN5" else rho ** (8 / 3) * 10 ** rng.uniform(-3, 1.5, 300)
    e_u, v_u, s_u = orc.functional(func, rho, sigma)
    rp = np.stack([rho / 2, rho / 2], axis=1)
    sp = None if sigma is None else np.stack([sigma / 4] * 3, axis=1)
    e_p, v_p, s_p = orc.functional_pol(func, rp, sp)
    np.testing.assert_allclose(e_p, e_u, rtol=1e-12)
    np.testing.assert_allclose(v_p[:, 0], v_u, rtol=1e-11)
    np.testing.assert_allclose(v_p[:, 1], v_u, rtol=1e-11)
    if sigma is not None:
        np.testing.assert_allclose(s_p.sum(axis=1) / 4, s_u, rtol=1e-10,
                                   atol=1e-12 * np.abs(s_u).max())


def test_oracle_polarised_pbe_derivatives_by_finite_difference(orc):
    """Every spin derivative of the polarised PBE energy density against
    central differences, at strongly polarised points (zeta up to +-0.9)."""
    rng = np.random.default_rng(12)
    n = 200
    rt = 10 ** rng.uniform(-3, 1, n)
    zeta = rng.uniform(-0.9, 0.9, n)
    rho = np.stack([rt * (1 + zeta) / 2, rt * (1 - zeta) / 2], axis=1)
    saa = rho[:, 0] ** (8 / 3) * 10 ** rng.uniform(-2, 1, n)
    sbb = rho[:, 1] ** (8 / 3) * 10 ** rng.uniform(-2, 1, n)
    sab = np.sqrt(saa * sbb) * rng.uniform(-0.9, 0.9, n)
    sig = np.stack([saa, sab, sbb], axis=1)
    eps, vr, vs = orc.functional_pol("PBE", rho, sig)

    def energy(r, s):
        e, _, _ = orc.functional_pol("PBE", r, s)
        return e * r.sum(axis=1)

    for k in range(2):
        h = 1e-6 * rho[:, k]
        dr = np.zeros_like(rho)
        dr[:, k] = h
        fd = (energy(rho + dr, sig) - energy(rho - dr, sig)) / (2 * h)
        np.testing.assert_allclose(vr[:, k], fd, rtol=2e-6, atol=1e-9)
    for k in range(3):
        # Step on each sigma's own scale (sigma_ab's is sqrt(saa sbb)).
        h = 1e-3 * (saa, np.sqrt(saa * sbb), sbb)[k]
        ds = np.zeros_like(sig)
        ds[:, k] = h
        fd = (energy(rho, sig + ds) - energy(rho, sig - ds)) / (2 * h)
        np.testing.assert_allclose(vs[:, k], fd, rtol=2e-5, atol=1e-9)


def _uks_system(gx):
    Z, xyz = gx.workloads.water_cluster(2, 5.0, 4)
    l, pure, alpha, coeff, origin = gx.workloads.even_tempered_basis(Z, xyz,
                                                                     gx.workloads.BASIS_C0)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=1e-10)
    lb = gx.workloads.build_tasks(gx.Molecule(Z, xyz), bs, grid="GM3")
    nbf = bs.nbf()
    pa = gx.workloads.density(nbf, 11, 21) / 2
    pb = gx.workloads.density(nbf, 9, 22) / 2
    return xyz, bs, lb.get_tasks(), pa, pb


def test_oracle_uks_reduces_to_rks_and_is_its_own_derivative(gx, orc):
    """UKS with P_a = P_b = P equals RKS with P (PBE: EXC, VXC_s = VXC,
    VXC_z = 0); and for a spin-polarised density VXC_s / VXC_z are the
    derivatives of the UKS EXC itself (d EXC / d P_s,ij = 2 VXC_s,ij for a
    symmetric perturbation of an off-diagonal pair), which pins the Z_s / Z_z
    and sigma combinations of the GGA spin terms."""
    xyz, bs, t, pa, pb = _uks_system(gx)
    nbf = bs.nbf()
    sh = orc.shells_from(bs.xcq_shells())
    w = orc.ssf_weights(xyz, t)
    e_r, n_r, v_r = orc.exc_vxc("PBE", nbf, sh, pa, t, weights=w)
    e_u, n_u, vs, vz = orc.exc_vxc_uks("PBE", nbf, sh, 2 * pa, np.zeros_like(pa), t, weights=w)
    assert abs(e_u - e_r) <= 1e-12 * abs(e_r) and abs(n_u - n_r) <= 1e-12 * n_r
    np.testing.assert_allclose(vs, v_r, rtol=0, atol=1e-13 * np.abs(v_r).max())
    assert np.abs(vz).max() <= 1e-13 * np.abs(v_r).max()
    ps, pz = pa + pb, pa - pb
    e0, _, vs, vz = orc.exc_vxc_uks("PBE", nbf, sh, ps, pz, t, weights=w)
    for (i, j) in ((3, 7), (0, nbf - 1), (12, 13)):
        for which, vref in ((0, vs), (1, vz)):
            h = 1e-5
            d = np.zeros((nbf, nbf))
            d[i, j] = d[j, i] = h
            args = [ps, pz]
            args[which] = args[which] + d
            ep = orc.exc_vxc_uks("PBE", nbf, sh, *args, t, weights=w)[0]
            args[which] = args[which] - 2 * d
            em = orc.exc_vxc_uks("PBE", nbf, sh, *args, t, weights=w)[0]
            fd = (ep - em) / (2 * h)
            assert abs(fd - 2 * vref[i, j]) <= 1e-7 * max(1.0, abs(fd)), (which, i, j)


@pytest.mark.parametrize("tag,func", [("svwn5", "SVWN5"), ("pbe0", "PBE0")])
def test_oracle_fxc_matches_gauxc_benzene_fixture(orc, tag, func):
    """GauXC's FXC contraction check (tests/2nd_derivative_test.cxx:121-126):
    ||FXC - FXC_ref||_F / nbf < 1e-10 for the fixture's trial density."""
    d, mol, bs, lb = benzene_setup(tag)
    t = lb.get_tasks()
    w = orc.ssf_weights(d["xyz"], t)
    nel, fxc = orc.fxc_contraction(func, bs.nbf(), orc.shells_from(bs.xcq_shells()),
                                   d["density"], d["trial_density"], t, weights=w)
    assert np.linalg.norm(fxc - d["fxc"]) / bs.nbf() < 1e-12
    assert abs(nel - 42.0) < 1e-4
    assert np.array_equal(fxc, fxc.T)


@pytest.mark.parametrize("func", ["SVWN5", "PBE", "PBE0"])
def test_oracle_fxc_functional_by_finite_difference(orc, func):
    """a and b are the derivatives of vrho and vsigma along (drho, dsigma):
    central differences of the first-derivative functional (itself pinned by
    the EXC/VXC fixtures and by test_oracle_functional_derivatives_by_finite_difference)."""
    rng = np.random.default_rng(11)
    n = 300
    rho = 10 ** rng.uniform(-3, 1.0, n)
    gga = func != "SVWN5"
    sigma = rho ** (8 / 3) * 10 ** rng.uniform(-2, 1.0, n) if gga else None
    drho = rho * rng.uniform(-1, 1, n)
    dsig = sigma * rng.uniform(-1, 1, n) if gga else None
    vr, vs, a, b = orc.functional_fxc(func, rho, sigma, drho, dsig)
    eps0, vr0, vs0 = orc.functional(func, rho, sigma)
    # Same derivatives by two routes (closed forms / duals); the PBE
    # correlation's cancellations at low density leave ~1e-10 relative.
    assert np.allclose(vr, vr0, rtol=1e-11, atol=0)
    if gga:
        assert np.allclose(vs, vs0, rtol=1e-9, atol=0)
    h = 1e-4
    _, vrp, vsp = orc.functional(func, rho + h * drho, None if not gga else sigma + h * dsig)
    _, vrm, vsm = orc.functional(func, rho - h * drho, None if not gga else sigma - h * dsig)
    fa = (vrp - vrm) / (2 * h)
    # truncation ~ h^2 |third derivative| (relative 1e-7 here), round-off ~ u |vrho| / h
    assert np.all(np.abs(a - fa) <= 1e-6 * np.abs(fa) + 1e-11 * np.abs(vr0) / h)
    if gga:
        fb = (vsp - vsm) / (2 * h)
        assert np.all(np.abs(b - fb) <= 1e-6 * np.abs(fb) + 1e-11 * np.abs(vs0) / h)


def test_oracle_fxc_is_the_derivative_of_vxc(orc):
    """FXC = d VXC[P + e tP] / de at e = 0 (the contraction's definition):
    central differences of the oracle's VXC along the fixture's trial density
    (PBE0, the GGA terms included)."""
    d, mol, bs, lb = benzene_setup("pbe0")
    t = lb.get_tasks()
    w = orc.ssf_weights(d["xyz"], t)
    sh = orc.shells_from(bs.xcq_shells())
    P, tP = d["density"], d["trial_density"]
    _, fxc = orc.fxc_contraction("PBE0", bs.nbf(), sh, P, tP, t, weights=w)
    h = 1e-4
    _, _, vp = orc.exc_vxc("PBE0", bs.nbf(), sh, P + h * tP, t, weights=w)
    _, _, vm = orc.exc_vxc("PBE0", bs.nbf(), sh, P - h * tP, t, weights=w)
    fd = (vp - vm) / (2 * h)
    assert np.abs(fxc - fd).max() <= 1e-6 * np.abs(fxc).max()


def test_oracle_uks_fxc_matches_gauxc_li_fixture(gx, orc):
    """UKS FXC: GauXC's Li / SVWN5 check (tests/2nd_derivative_test.cxx:128-134):
    ||FXC_s - ref|| / nbf and ||FXC_z - ref|| / nbf < 1e-10."""
    d = load_golden("li_svwn5_uks.npz")
    mol = gx.Molecule(d["Z"], d["xyz"])
    bs = gx.BasisSet.from_arrays(d["sh_nprim"], d["sh_l"], d["sh_pure"], d["sh_alpha"],
                                 d["sh_coeff"], d["sh_origin"], normalize=False,
                                 shell_tolerance=EPS)
    mg = gx.MolGridFactory.create_default_molgrid(mol, "Unpruned", 512, "MuraKnowles",
                                                  "UltraFineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    t = lb.get_tasks()
    fs, fz = orc.fxc_contraction_uks("SVWN5", bs.nbf(), orc.shells_from(bs.xcq_shells()),
                                     d["density_s"], d["density_z"], d["trial_density_s"],
                                     d["trial_density_z"], t)
    assert np.linalg.norm(fs - d["fxc_s"]) / bs.nbf() < 1e-13
    assert np.linalg.norm(fz - d["fxc_z"]) / bs.nbf() < 1e-13


def test_oracle_uks_fxc_is_the_derivative_of_vxc_and_reduces_to_rks(gx, orc):
    """UKS FXC (PBE, spin-polarised water dimer) equals central differences
    of the UKS VXC_s / VXC_z along the trial direction; with P_z = tP_z = 0
    and P_s = 2P, tP_s = 2tP it equals the RKS FXC of (P, tP)."""
    xyz, bs, t, pa, pb = _uks_system(gx)
    w = orc.ssf_weights(xyz, t)
    sh = orc.shells_from(bs.xcq_shells())
    nbf = bs.nbf()
    rng = np.random.default_rng(5)
    ta = rng.standard_normal((nbf, nbf)) * 0.01
    tb = rng.standard_normal((nbf, nbf)) * 0.01
    ta, tb = ta + ta.T, tb + tb.T
    Ps, Pz, tPs, tPz = pa + pb, pa - pb, ta + tb, ta - tb
    fs, fz = orc.fxc_contraction_uks("PBE", nbf, sh, Ps, Pz, tPs, tPz, t, weights=w)
    h = 1e-4
    _, _, vsp, vzp = orc.exc_vxc_uks("PBE", nbf, sh, Ps + h * tPs, Pz + h * tPz, t, weights=w)
    _, _, vsm, vzm = orc.exc_vxc_uks("PBE", nbf, sh, Ps - h * tPs, Pz - h * tPz, t, weights=w)
    assert np.abs(fs - (vsp - vsm) / (2 * h)).max() <= 1e-6 * np.abs(fs).max()
    assert np.abs(fz - (vzp - vzm) / (2 * h)).max() <= 1e-6 * np.abs(fz).max()
    _, fr = orc.fxc_contraction("PBE", nbf, sh, pa, ta, t, weights=w)
    fs0, fz0 = orc.fxc_contraction_uks("PBE", nbf, sh, 2 * pa, 0 * pa, 2 * ta, 0 * ta, t,
                                       weights=w)
    assert np.abs(fs0 - fr).max() <= 1e-11 * np.abs(fr).max()
    assert np.abs(fz0).max() <= 1e-14 * np.abs(fr).max()


@pytest.mark.parametrize("pure", [True, False])
def test_oracle_collocation_hessian(orc, pure):
    """Hessian collocation: against GauXC's own gau2grid (deriv2, compiled
    under oracle/_ref) where built, and always against central differences
    of the pinned gradient collocation; value and gradient equal the
    gradient collocation's."""
    rng = np.random.default_rng(8)
    shells = _random_shells(orc, rng, 4, pure)
    pts = rng.uniform(-2.0, 2.0, (41, 3))
    gg = orc.gg_ref()
    h = 1e-5
    for l in range(5):
        out = orc.collocation_hessian(shells, [l], pts)
        base = orc.collocation(shells, [l], pts, deriv=True)
        for k in range(4):
            np.testing.assert_allclose(out[k], base[k], rtol=1e-13, atol=1e-15)
        # d/dx_j of the gradient: column (i, j) of the Hessian.
        comp = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            gp = orc.collocation(shells, [l], pts + e, deriv=True)
            gm = orc.collocation(shells, [l], pts - e, deriv=True)
            for i in range(3):
                fd = (gp[1 + i] - gm[1 + i]) / (2 * h)
                hij = out[comp[(min(i, j), max(i, j))]]
                assert np.abs(hij - fd).max() <= 1e-7 * max(1.0, np.abs(hij).max())
        if gg is not None:
            s = shells[l]
            nc = 2 * l + 1 if pure else (l + 1) * (l + 2) // 2
            ref = np.zeros((10, nc, len(pts)))
            coeff = np.array(s.coeff[:3])
            alpha = np.array(s.alpha[:3])
            center = np.array(s.origin[:3])
            P = ctypes.c_void_p
            gg.gg_collocation_deriv2(l, len(pts), np.ascontiguousarray(pts).ctypes.data, 3, 3,
                                     coeff.ctypes.data, alpha.ctypes.data, center.ctypes.data,
                                     orc.GG_SPHERICAL_CCA if pure else orc.GG_CARTESIAN_CCA,
                                     *[P(ref[k].ctypes.data) for k in range(10)])
            for k in range(10):
                assert np.abs(out[k] - ref[k].T).max() <= 1e-12 * max(1.0, np.abs(ref[k]).max())


@pytest.mark.parametrize("tag,func", [("svwn5", "SVWN5"), ("pbe0", "PBE0")])
def test_oracle_exc_grad_matches_gauxc_benzene_fixture(orc, tag, func):
    """GauXC's gradient check (tests/xc_integrator.cxx:278-289): the full
    gradient with grid-weight derivatives (the fixture's EXC_GRAD carries
    includes_weight_derivatives = 1), ||dG||_F / sqrt(3 natoms) < 1e-8."""
    d, mol, bs, lb = benzene_setup(tag)
    t = lb.get_tasks()
    w = orc.ssf_weights(d["xyz"], t)
    g = orc.exc_grad(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), d["xyz"], d["density"],
                     t, weights=w)
    ref = d["exc_grad"].reshape(-1, 3)
    assert np.linalg.norm(g - ref) / np.sqrt(g.size) < 1e-12
    # Translational invariance: the rows sum to zero.
    assert np.abs(g.sum(axis=0)).max() < 1e-12
    # The Hellmann-Feynman form (EXC_GRAD_HELLFEY, tests/xc_integrator.cxx:291-300).
    gh = orc.exc_grad(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), d["xyz"], d["density"],
                      t, weights=w, weight_derivs=False)
    assert np.linalg.norm(gh - d["exc_grad_hf"].reshape(-1, 3)) / np.sqrt(g.size) < 1e-11


@pytest.mark.parametrize("scheme", ["Robust", "Treutler"])
def test_oracle_pruned_grid_matches_gauxc_benzene_fixture(gx, orc, scheme):
    """Benzene / SVWN5 on the pruned ultra-fine grids (tests/xc_integrator.cxx:
    410-420, 2nd_derivative_test.cxx:198-208): the point count follows the
    scheme's regions (C and H: 99 radial nodes, base rule 590 of order 41;
    Robust: 25 nodes on 26 points, 25 on 434 (order 35), 49 on 590;
    Treutler: 34 on 26, 16 on 50, 49 on 590), and EXC, VXC, FXC and both
    gradients match GauXC's values."""
    from conftest import unpack_sym
    d, mol, bs, _ = benzene_setup("svwn5")
    pr = load_golden("benzene_svwn5_pruned.npz")
    key = scheme.lower()
    mg = gx.MolGridFactory.create_default_molgrid(mol, scheme, 512, "MuraKnowles",
                                                  "UltraFineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    t = lb.get_tasks()
    per_atom = 25 * 26 + 25 * 434 + 49 * 590 if scheme == "Robust" else 34 * 26 + 16 * 50 + 49 * 590
    assert int(np.sum(t["npts"])) == 12 * per_atom
    n = bs.nbf()
    w = orc.ssf_weights(d["xyz"], t)
    sh = orc.shells_from(bs.xcq_shells())
    exc, nel, vxc = orc.exc_vxc("SVWN5", n, sh, d["density"], t, weights=w)
    assert abs(exc - pr[f"{key}_exc"][0]) <= 1e-12 * abs(exc)
    assert np.abs(vxc - unpack_sym(pr[f"{key}_vxc_packed"], n)).max() < 1e-11
    fxc = orc.fxc_contraction("SVWN5", n, sh, d["density"], d["trial_density"], t, weights=w)[-1]
    assert np.abs(fxc - unpack_sym(pr[f"{key}_fxc_packed"], n)).max() < 1e-11
    for wd, ref in ((True, pr[f"{key}_exc_grad"]), (False, pr[f"{key}_exc_grad_hf"])):
        g = orc.exc_grad("SVWN5", n, sh, d["xyz"], d["density"], t, weights=w, weight_derivs=wd)
        assert np.abs(g.ravel() - ref).max() < 1e-11, wd


def test_oracle_uks_matches_gauxc_cytosine_fixture(orc):
    """Doublet cytosine / SVWN5 / cc-pVDZ, Robust-pruned ultra-fine grid
    (tests/xc_integrator.cxx:218-237,278-300,455-459).  The full gradient
    (EXC_GRAD, with weight derivatives) matches to 1e-12.  EXC, VXC_s and
    the Hellmann-Feynman gradient match at GauXC's own bars (EXC relative
    Catch epsilon, ||dVXC||_F / nbf < 1e-10, ||dG|| / sqrt(3 natoms) < 1e-8)
    but not tighter: measured 2.6e-11 relative, 6.5e-12 and 2.8e-9, with
    VXC_z at 1e-14.  The same restatement reproduces every benzene fixture
    (unpruned, Robust, Treutler; EXC, VXC, FXC, both gradients) at 1e-12 and
    ties UKS to RKS at 1e-11, and the fixture's own full gradient agrees at
    1e-13, so the residual is taken to be in how those three datasets were
    produced; the bounds below are GauXC's, with the measured values noted."""
    from conftest import cytosine_uks_setup
    d, mol, bs, lb = cytosine_uks_setup()
    t = lb.get_tasks()
    xyz = np.asarray(d["xyz"], dtype=np.float64)
    n = bs.nbf()
    w = orc.ssf_weights(xyz, t)
    sh = orc.shells_from(bs.xcq_shells())
    exc, nel, vs, vz = orc.exc_vxc_uks("SVWN5", n, sh, d["density_s"], d["density_z"], t,
                                       weights=w)[:4]
    assert abs(exc - d["exc"][0]) <= 1e-10 * abs(exc)          # measured 2.6e-11
    assert np.linalg.norm(vs - d["vxc_s"]) / n < 1e-10         # measured 6.5e-12
    assert np.linalg.norm(vz - d["vxc_z"]) / n < 1e-13         # measured 1.0e-14
    assert abs(nel - 57.0) < 1e-4                               # 29 + 28 electrons
    g = orc.exc_grad_uks("SVWN5", n, sh, xyz, d["density_s"], d["density_z"], t, weights=w)
    assert np.abs(g.ravel() - d["exc_grad"]).max() < 1e-12     # measured 1.3e-13
    assert np.abs(g.sum(axis=0)).max() < 1e-12
    gh = orc.exc_grad_uks("SVWN5", n, sh, xyz, d["density_s"], d["density_z"], t, weights=w,
                          weight_derivs=False)
    assert np.linalg.norm(gh.ravel() - d["exc_grad_hf"]) / np.sqrt(gh.size) < 1e-8


def _dimer_system(gx, orc, xyz, Z):
    l, pure, alpha, coeff, origin = gx.workloads.even_tempered_basis(Z, xyz,
                                                                     gx.workloads.BASIS_C0)
    bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=EPS)
    lb = gx.workloads.build_tasks(gx.Molecule(Z, xyz), bs, grid="GM3")
    t = lb.get_tasks()
    return bs, t, orc.ssf_weights(xyz, t)


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_oracle_exc_grad_is_the_derivative_of_exc(gx, orc, func):
    """dEXC/dR_A by central differences of the whole evaluation (basis, grid
    and SSF weights moving with the atoms) on a water dimer."""
    Z, xyz = gx.workloads.water_cluster(2, 5.0, 4)
    xyz = np.asarray(xyz, dtype=np.float64)
    bs, t, w = _dimer_system(gx, orc, xyz, Z)
    P = gx.workloads.density(bs.nbf(), 10, 21)
    g = orc.exc_grad(func, bs.nbf(), orc.shells_from(bs.xcq_shells()), xyz, P, t, weights=w)
    assert np.abs(g.sum(axis=0)).max() < 1e-10
    h = 1e-4
    for atom, comp in ((0, 0), (4, 2)):
        e = []
        for sgn in (1, -1):
            x2 = xyz.copy()
            x2[atom, comp] += sgn * h
            b2, t2, w2 = _dimer_system(gx, orc, x2, Z)
            e.append(orc.exc_vxc(func, b2.nbf(), orc.shells_from(b2.xcq_shells()), P, t2,
                                 weights=w2)[0])
        fd = (e[0] - e[1]) / (2 * h)
        assert abs(g[atom, comp] - fd) <= 1e-6 * max(abs(fd), 1e-3), (atom, comp, g[atom, comp], fd)


@pytest.mark.parametrize("func", ["SVWN5", "PBE"])
def test_oracle_uks_exc_grad(gx, orc, func):
    """UKS gradient: dEXC/dR by central differences of the whole UKS
    evaluation on a spin-polarised water dimer, translational invariance,
    and the RKS limit (P_s = 2P, P_z = 0 gives the RKS gradient of P)."""
    Z, xyz = gx.workloads.water_cluster(2, 5.0, 4)
    xyz = np.asarray(xyz, dtype=np.float64)
    bs, t, w = _dimer_system(gx, orc, xyz, Z)
    nbf = bs.nbf()
    pa = gx.workloads.density(nbf, 11, 21) / 2
    pb = gx.workloads.density(nbf, 9, 22) / 2
    Ps, Pz = pa + pb, pa - pb
    sh = orc.shells_from(bs.xcq_shells())
    g = orc.exc_grad_uks(func, nbf, sh, xyz, Ps, Pz, t, weights=w)
    assert np.abs(g.sum(axis=0)).max() < 1e-10
    h = 1e-4
    for atom, comp in ((1, 1), (3, 0)):
        e = []
        for sgn in (1, -1):
            x2 = xyz.copy()
            x2[atom, comp] += sgn * h
            b2, t2, w2 = _dimer_system(gx, orc, x2, Z)
            e.append(orc.exc_vxc_uks(func, b2.nbf(), orc.shells_from(b2.xcq_shells()), Ps, Pz,
                                     t2, weights=w2)[0])
        fd = (e[0] - e[1]) / (2 * h)
        assert abs(g[atom, comp] - fd) <= 1e-6 * max(abs(fd), 1e-3), (atom, comp, g[atom, comp], fd)
    gr = orc.exc_grad(func, nbf, sh, xyz, pa, t, weights=w)
    g0 = orc.exc_grad_uks(func, nbf, sh, xyz, 2 * pa, 0 * pa, t, weights=w)
    assert np.abs(g0 - gr).max() <= 1e-11 * np.abs(gr).max()
    for wd in (False,):
        gh = orc.exc_grad_uks(func, nbf, sh, xyz, 2 * pa, 0 * pa, t, weights=w, weight_derivs=wd)
        ghr = orc.exc_grad(func, nbf, sh, xyz, pa, t, weights=w, weight_derivs=wd)
        assert np.abs(gh - ghr).max() <= 1e-11 * np.abs(ghr).max()


def test_oracle_mlp_gradient_matches_finite_differences():
    """The cost-model network's CPU oracle: analytic gradient against
    central differences of its own loss (pins forward and backward)."""
    import oracle.mlp as M
    rng = np.random.default_rng(7)
    widths, shapes, n = M.layout(6, 24, 3, 2)
    p = rng.normal(size=n) * 0.4
    x = rng.normal(size=(32, 6))
    t = rng.normal(size=(32, 2))
    loss, g = M.loss_and_grad(p, shapes, x, t)
    acts, _ = M.forward(p, shapes, x)
    assert np.isclose(loss, np.mean((acts[-1] - t) ** 2))
    h = 1e-6
    for k in rng.choice(n, 60, replace=False):
        e = np.zeros(n)
        e[k] = h
        fd = (M.loss_and_grad(p + e, shapes, x, t)[0] - M.loss_and_grad(p - e, shapes, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(g[k]), (k, fd, g[k])
    # a planted mistake (bias gradient of the last layer doubled) is caught
    off, o, i = shapes[-1]
    g_bad = g.copy()
    g_bad[off + o * i:off + o * i + o] *= 2
    k = off + o * i
    e = np.zeros(n)
    e[k] = h
    fd = (M.loss_and_grad(p + e, shapes, x, t)[0] - M.loss_and_grad(p - e, shapes, x, t)[0]) / (2 * h)
    assert abs(fd - g_bad[k]) > 1e-6
