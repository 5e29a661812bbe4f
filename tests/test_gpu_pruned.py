"""Pruned grids on the device: benzene / SVWN5 / cc-pVDZ on the Robust and
Treutler pruned ultra-fine grids (tests/xc_integrator.cxx:410-420,
2nd_derivative_test.cxx:198-208) through the GauXC API, with the device
weights: EXC, VXC, FXC and both gradients against GauXC's values, at GauXC's
bars and the FP64 bars."""
import numpy as np
import pytest

from conftest import benzene_setup, load_golden, unpack_sym

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("scheme", ["Robust", "Treutler"])
def test_pruned_benzene_matches_gauxc_fixture(gx, gpu, scheme):
    d, mol, bs, _ = benzene_setup("svwn5")
    pr = load_golden("benzene_svwn5_pruned.npz")
    key = scheme.lower()
    n = bs.nbf()
    mg = gx.MolGridFactory.create_default_molgrid(mol, scheme, 512, "MuraKnowles",
                                                  "UltraFineGrid")
    lb = gx.LoadBalancerFactory("Host", "Default").get_instance(gx.RuntimeEnvironment(), mol,
                                                                 mg, bs)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = gx.XCIntegratorFactory("Device", "Replicated").get_instance(gx.Functional("SVWN5"),
                                                                         lb)
    _flush(f"benzene SVWN5 {scheme}-pruned grid: EXC/VXC")
    exc, vxc = integ.eval_exc_vxc(d["density"])
    vref = unpack_sym(pr[f"{key}_vxc_packed"], n)
    _flush(f"dEXC={exc - pr[f'{key}_exc'][0]:.3e} max|dVXC|={np.abs(vxc - vref).max():.3e}")
    assert abs(exc - pr[f"{key}_exc"][0]) <= 1e-10
    assert np.abs(vxc - vref).max() <= 1e-9
    _flush("FXC contraction")
    fxc = integ.eval_fxc_contraction(d["density"], d["trial_density"])
    fref = unpack_sym(pr[f"{key}_fxc_packed"], n)
    _flush(f"max|dFXC|={np.abs(fxc - fref).max():.3e}")
    assert np.abs(fxc - fref).max() <= 1e-9
    _flush("EXC gradients")
    g = integ.eval_exc_grad(d["density"])
    gh = integ.eval_exc_grad(d["density"], include_weight_derivatives=False)
    _flush(f"max|dG|={np.abs(g.ravel() - pr[f'{key}_exc_grad']).max():.3e} "
           f"max|dG_HF|={np.abs(gh.ravel() - pr[f'{key}_exc_grad_hf']).max():.3e}")
    assert np.abs(g.ravel() - pr[f"{key}_exc_grad"]).max() <= 1e-10
    assert np.abs(gh.ravel() - pr[f"{key}_exc_grad_hf"]).max() <= 1e-10
