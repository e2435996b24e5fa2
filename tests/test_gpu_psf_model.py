"""PSF models on the gfx950 path against the float64 oracle.

* Moffat_PSF sampled on a PSF_Image window with Simpson sampling, threshold
  integration and unit-sum normalisation (psf_model_object.py:180-264,
  moffat_model.py:71-119), and its Jacobian (normalisation included);
* a galaxy convolved with an auxiliary Moffat_PSF model (model_object.py:
  133-147, 305-310): image, Jacobian with the PSF parameters' columns first,
  hand-written adjoint, for both sub-pixel shift modes;
* a group holding such a galaxy (tests/test_group_models.py:112-165 in the
  reference: 16 parameters) saved and loaded;
* LM recovers a galaxy and its PSF width together from a noiseless image.

Tolerances as in test_gpu_parity: image 1e-5, Jacobian columns 2e-4 of the
oracle's max.
"""

import math

import numpy as np
import pytest
import torch

import astrophot_amd as ap
from oracle import hotpath as H
from tests.pairs import rel_err, target

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
JAC_TOL = 2e-4
SERSIC = {"center": [20.3, 17.6], "q": 0.6, "PA": 0.7, "n": 2.5, "Re": 6.0, "Ie": 1.0}


def psf_pair(n, Rd, side, P=1.0, free_I0=False, **knobs):
    """(oracle Component, Moffat_PSF) on a side x side PSF window."""
    g = H.Grid(side, side, P, ref_ij=((side - 1) / 2, (side - 1) / 2), ref_xy=(0.0, 0.0))
    locked = ("center",) if free_I0 else ("center", "I0")
    oc = H.Component("moffat_psf", g, {"n": n, "Rd": Rd, "I0": 0.0}, locked=locked,
                     sampling_mode=knobs.get("sampling_mode", "simpsons"),
                     sampling_tolerance=knobs.get("sampling_tolerance", 1e-3),
                     integrate_mode=knobs.get("integrate_mode", "threshold"))
    params = {"n": n, "Rd": Rd}
    if free_I0:
        params["I0"] = {"value": 0.0, "locked": False}
    m = ap.models.AstroPhot_Model(name=f"psf{side}", model_type="moffat psf model",
                                  target=ap.image.PSF_Image(data=np.zeros((side, side)), pixelscale=P),
                                  parameters=params, **knobs)
    return oc, m


@pytest.mark.parametrize("n,Rd,side", [(2.5, 2.2, 15), (3.0, 2.942, 51), (1.3, 0.8, 9), (4.5, 5.0, 25)])
@pytest.mark.parametrize("integrate", ["none", "threshold"])
def test_psf_model_sample(n, Rd, side, integrate):
    oc, m = psf_pair(n, Rd, side, integrate_mode=integrate)
    om = H.Model([oc])
    ref = om.forward(om.vector_representation())
    got = m().data
    assert got.is_cuda and tuple(got.shape) == (side, side)
    assert rel_err(got, ref) < IMG_TOL
    assert abs(float(got.double().sum()) - 1.0) < 1e-5


def test_psf_model_quad_sampling_and_pixelscale():
    oc, m = psf_pair(2.0, 1.5, 13, P=0.7, sampling_mode="quad:3")
    om = H.Model([oc])
    assert rel_err(m().data, om.forward(om.vector_representation())) < IMG_TOL


@pytest.mark.parametrize("free_I0", [False, True])
def test_psf_model_jacobian(free_I0):
    oc, m = psf_pair(2.5, 2.2, 15, free_I0=free_I0)
    om = H.Model([oc])
    rep = om.vector_representation()
    assert np.allclose(rep.numpy(), m.parameters.vector_representation().numpy())
    Jref = om.jacobian(rep)
    J = m.jacobian(parameters=rep.clone(), as_representation=True).data
    assert J.shape == Jref.shape
    for k in range(J.shape[-1]):
        if Jref[..., k].abs().max() < 1e-12:  # I0 cancels in the normalisation
            assert float(J[..., k].abs().max()) < 1e-6
        else:
            assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k


def galaxy_pair(shift="bilinear", psf_side=11, kind="sersic", params=None, nx=40, ny=36):
    opsf, mpsf = psf_pair(2.2, 1.8, psf_side)
    params = dict(SERSIC if params is None else params)
    og = H.Component(kind, H.Grid(nx, ny), dict(params), psf=opsf, psf_mode="full")
    tgt = target(nx, ny)
    mg = ap.models.AstroPhot_Model(name="gal", model_type=f"{kind} galaxy model", target=tgt, psf=mpsf,
                                   psf_mode="full", psf_subpixel_shift=shift, parameters=params)
    return H.Model([og]), mg, mpsf


def test_galaxy_with_psf_model():
    om, mg, mpsf = galaxy_pair("bilinear")
    rep = om.vector_representation()
    assert list(mg.parameters.vector_names())[:2] == ["n", "Rd"]
    assert np.allclose(rep.numpy(), mg.parameters.vector_representation().numpy())
    assert rel_err(mg().data, om.forward(rep)) < IMG_TOL
    Jref = om.jacobian(rep)
    J = mg.jacobian(parameters=rep.clone(), as_representation=True).data
    assert J.shape == Jref.shape == (36, 40, 9)
    for k in range(J.shape[-1]):
        assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k


def test_psf_model_equals_its_image():
    """Shift "none" (no oracle restatement): a galaxy convolved with a PSF
    model renders as with that model's sampled image as a fixed PSF, and its
    PSF-parameter columns match central differences of the render."""
    _, mg, mpsf = galaxy_pair("none")
    img = mpsf().data.clone()
    tgt = target(40, 36)
    fixed = ap.models.AstroPhot_Model(name="fixed", model_type="sersic galaxy model", target=tgt,
                                      psf=ap.image.PSF_Image(data=img, pixelscale=1.0), psf_mode="full",
                                      psf_subpixel_shift="none", parameters=dict(SERSIC))
    assert rel_err(mg().data, fixed().data) < 1e-6
    rep = mg.parameters.vector_representation().clone()
    J = mg.jacobian(parameters=rep.clone(), as_representation=True).data
    for k in range(2):
        h = 1e-3
        r1, r2 = rep.clone(), rep.clone()
        r1[k] += h
        r2[k] -= h
        fd = (mg(parameters=r1, as_representation=True).data.double()
              - mg(parameters=r2, as_representation=True).data.double()) / (2 * h)
        assert rel_err(J[..., k], fd) < 5e-3, k


@pytest.mark.parametrize("shift", ["bilinear", "none"])
def test_galaxy_with_psf_model_vjp(shift):
    _, mg, _ = galaxy_pair(shift)
    rep = mg.parameters.vector_representation().clone()
    rng = np.random.default_rng(7)
    G = torch.tensor(rng.normal(size=(36, 40)))
    got = mg.vjp(G.float(), parameters=rep.clone(), as_representation=True).cpu()
    J = mg.jacobian(parameters=rep.clone(), as_representation=True).data.double().cpu()
    jtg = (J * G[..., None]).sum(dim=(0, 1))
    assert ((got - jtg).abs() / jtg.abs().max()).max() < JAC_TOL, (got, jtg)


def test_group_with_psf_model_save_load(tmp_path):
    """The reference's group save/load test layout: the first model carries an
    auxiliary Moffat PSF (n, Rd free), 16 parameters in all."""
    tar = target(51, 51)
    _, mpsf = psf_pair(2.0, 3.0, 11)
    m1 = ap.models.AstroPhot_Model(name="base model 1", model_type="sersic galaxy model", target=tar, psf=mpsf,
                                   psf_mode="full", parameters={"center": [25.2, 24.6], "q": 0.7, "PA": 0.4,
                                                                "n": 1.8, "Re": 6.0, "Ie": 0.5})
    m2 = ap.models.AstroPhot_Model(name="base model 2", model_type="sersic galaxy model", target=tar,
                                   parameters={"center": [30.1, 20.3], "q": 0.5, "PA": 2.0, "n": 3.0,
                                               "Re": 4.0, "Ie": 0.2})
    grp = ap.models.AstroPhot_Model(name="group model", model_type="group model", models=[m1, m2], target=tar)
    assert grp.parameters.size == 16
    fn = str(tmp_path / "g.yaml")
    grp.save(fn)
    new = ap.models.AstroPhot_Model(name="group model", filename=fn, target=tar)
    assert new.parameters.size == 16
    assert torch.equal(new.parameters.vector_values(), grp.parameters.vector_values())
    assert type(new.models["base model 1"].psf).__name__ == "Moffat_PSF"
    assert torch.equal(new().data, grp().data)
    assert torch.equal(new.jacobian().data, grp.jacobian().data)


def test_lm_fits_galaxy_and_psf_width():
    """Noiseless image of a Sersic galaxy through a Moffat PSF (Rd 1.8): LM
    from a perturbed start (Rd 2.4, Re 7.5) recovers both."""
    om, mg, mpsf = galaxy_pair()
    data = mg().data.clone()
    tgt = target(40, 36, data=data.cpu().numpy(), variance=np.full((36, 40), 1e-4))
    _, mpsf2 = psf_pair(2.2, 2.4, 11)
    start = dict(SERSIC, Re=7.5, center=[20.0, 17.9])
    fit_m = ap.models.AstroPhot_Model(name="fitgal", model_type="sersic galaxy model", target=tgt, psf=mpsf2,
                                      psf_mode="full", parameters=start)
    res = ap.fit.LM(fit_m, max_iter=60).fit()
    assert "success" in res.message, res.message
    assert abs(float(mpsf2["Rd"].value) - 1.8) < 1e-3
    assert abs(float(fit_m["Re"].value) - 6.0) < 1e-3
    assert np.allclose(fit_m["center"].value.numpy(), [20.3, 17.6], atol=1e-4)


def test_galaxy_with_upscaled_psf_model():
    """An auxiliary Moffat_PSF on a PSF window twice as fine as the image
    (model_object.py:313-314 take the upscale from the PSF's pixel length):
    image and the 9-column Jacobian (n, Rd first) against the oracle."""
    opsf, mpsf = psf_pair(2.2, 1.8, 21, P=0.5)
    og = H.Component("sersic", H.Grid(40, 36), dict(SERSIC), psf=opsf, psf_mode="full")
    om = H.Model([og])
    mg = ap.models.AstroPhot_Model(name="gal_up", model_type="sersic galaxy model", target=target(40, 36), psf=mpsf,
                                   psf_mode="full", parameters=dict(SERSIC))
    rep = om.vector_representation()
    assert np.allclose(rep.numpy(), mg.parameters.vector_representation().numpy())
    assert rel_err(mg().data, om.forward(rep)) < IMG_TOL
    Jref = om.jacobian(rep)
    J = mg.jacobian(parameters=rep.clone(), as_representation=True).data
    assert J.shape == Jref.shape == (36, 40, 9)
    for k in range(J.shape[-1]):
        assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k
