"""Parity of the gfx950 path with the float64 oracle on identical inputs.

Tolerances (max-norm relative to the oracle's max):
  model image  1e-5   (float32 render vs float64 oracle)
  Jacobian     2e-4 per column (float32 analytic partials vs float64 forward AD)
"""

import math

import numpy as np
import pytest
import torch

import astrophot_amd as ap
from oracle import hotpath as H
from tests.pairs import GAL, component, rel_err, single, target

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
JAC_TOL = 2e-4

PSF11 = H.gaussian_psf(1.5, 11, 1.0)


@pytest.mark.parametrize("kind", ["sersic", "exponential", "gaussian", "moffat", "flat_sky"])
@pytest.mark.parametrize("integrate", ["none", "threshold"])
def test_render_psf_none(kind, integrate):
    om, m = single(kind, GAL[kind], integrate_mode=integrate)
    ref = om.forward(om.vector_representation())
    got = m().data
    assert got.is_cuda
    assert rel_err(got, ref) < IMG_TOL


@pytest.mark.parametrize("kind", ["sersic", "exponential", "gaussian", "moffat"])
@pytest.mark.parametrize("integrate", ["none", "threshold"])
def test_render_psf_full(kind, integrate):
    om, m = single(kind, GAL[kind], psf=PSF11, psf_mode="full", integrate_mode=integrate)
    ref = om.forward(om.vector_representation())
    got = m().data
    assert rel_err(got, ref) < IMG_TOL


@pytest.mark.parametrize("kind", ["sersic", "exponential", "gaussian", "moffat", "flat_sky"])
@pytest.mark.parametrize("psf_mode", ["none", "full"])
def test_jacobian(kind, psf_mode):
    if kind == "flat_sky" and psf_mode == "full":
        pytest.skip("sky models never convolve (sky_model_object.py:21-27)")
    kw = dict(psf=PSF11, psf_mode="full") if psf_mode == "full" else {}
    om, m = single(kind, GAL[kind], **kw)
    rep = om.vector_representation()
    Jref = om.jacobian(rep)
    J = m.jacobian(parameters=rep.clone(), as_representation=True).data
    assert J.shape == Jref.shape
    for k in range(J.shape[-1]):
        assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k


def test_sersic_quad_sampling():
    """sampling_mode quad:5 (the reference fixtures' truth render, tests/utils.py:64)."""
    om, m = single("sersic", GAL["sersic"], psf=PSF11, psf_mode="full", sampling_mode="quad:5")
    ref = om.forward(om.vector_representation())
    assert rel_err(m().data, ref) < IMG_TOL


@pytest.mark.parametrize("kind", ["sersic", "exponential", "gaussian", "moffat"])
@pytest.mark.parametrize("integrate", ["none", "threshold"])
def test_trapezoid_sampling(kind, integrate):
    """sampling_mode "trapezoid" (_model_methods.py:121-139): the mean of the
    four pixel corners, selection by its Laplacian.  Image, Jacobian columns
    and the adjoint against the oracle, with the PSF convolved."""
    om, m = single(kind, GAL[kind], psf=PSF11, psf_mode="full", sampling_mode="trapezoid",
                   integrate_mode=integrate)
    rep = om.vector_representation()
    assert rel_err(m().data, om.forward(rep)) < IMG_TOL
    Jref = om.jacobian(rep)
    J = m.jacobian(parameters=rep.clone(), as_representation=True).data
    for k in range(J.shape[-1]):
        assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k
    G = torch.tensor(np.random.default_rng(3).normal(size=(m.window.ny, m.window.nx)))
    r = rep.clone().requires_grad_(True)
    ref, = torch.autograd.grad((om.forward(r) * G).sum(), r)
    got = m.vjp(G.float(), parameters=rep.clone(), as_representation=True).cpu()
    assert ((got - ref).abs() / ref.abs().max()).max() < JAC_TOL, (got, ref)


def test_trapezoid_sampling_psf_none():
    om, m = single("sersic", GAL["sersic"], sampling_mode="trapezoid")
    assert rel_err(m().data, om.forward(om.vector_representation())) < IMG_TOL


def upscaled_pair(kind, U, nx=30, ny=26):
    """A model whose PSF is U times finer than the image (target_image.py:
    303-324), in the oracle and on the GPU."""
    psf = H.gaussian_psf(1.3, 13, 1.0 / U)
    params = dict(GAL[kind], center=[15.3, 12.6])
    oc = H.Component(kind, H.Grid(nx, ny), dict(params), psf=torch.tensor(psf), psf_mode="full", psf_upscale=U)
    tgt = ap.image.Target_Image(data=np.zeros((ny, nx)), pixelscale=1.0, psf=psf, psf_upscale=U)
    m = ap.models.AstroPhot_Model(name=f"up{kind}{U}", model_type=f"{kind} galaxy model", target=tgt,
                                  psf_mode="full", parameters=params)
    return H.Model([oc]), m


@pytest.mark.parametrize("U", [2, 3])
@pytest.mark.parametrize("kind", ["sersic", "moffat"])
def test_psf_upscale(kind, U):
    """model_object.py:313-349: sampled and convolved on the U-times finer
    grid, U x U blocks summed back (with U = 3 the reference's float32 1/U
    leaves the last row and column empty).  Image, Jacobian and adjoint
    against the oracle."""
    om, m = upscaled_pair(kind, U)
    rep = om.vector_representation()
    ref = om.forward(rep)
    got = m().data
    assert rel_err(got, ref) < IMG_TOL
    if U == 3:
        assert float(got[-1].abs().max()) == 0.0 and float(got[:, -1].abs().max()) == 0.0
    Jref = om.jacobian(rep)
    J = m.jacobian(parameters=rep.clone(), as_representation=True).data
    assert J.shape == Jref.shape
    for k in range(J.shape[-1]):
        assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k
    G = torch.tensor(np.random.default_rng(5).normal(size=tuple(ref.shape)))
    r = rep.clone().requires_grad_(True)
    gref, = torch.autograd.grad((om.forward(r) * G).sum(), r)
    gv = m.vjp(G.float(), parameters=rep.clone(), as_representation=True).cpu()
    assert ((gv - gref).abs() / gref.abs().max()).max() < JAC_TOL, (gv, gref)


def test_lm_with_upscaled_psf_matches_oracle():
    """fit.LM through the upscaled render and Jacobian (PSF upscale 2) against
    the oracle's float64 LM on the same seeded data (variance = model/5 +
    0.01) from the same +-3 % start: parameters 1e-4, chi^2 1e-6.  (The
    bilinear PSF shift makes the model piecewise linear in the centre with a
    kink at every working-grid pixel; on the twice finer grid the kinks are
    twice as dense, and from +-10 % the reference LM's curvature test damps
    both fits into immobility.)"""
    om, m = upscaled_pair("sersic", 2)
    img = om.forward(om.vector_representation()).numpy()
    rng = np.random.default_rng(9)
    var = img / 5.0 + 0.01
    data = img + rng.normal(size=img.shape) * np.sqrt(var)
    oc = om.components[0]
    start_vals = om.vector_values().numpy() * (1 + rng.uniform(-0.03, 0.03, 7))
    o_start = oc.val_to_rep(torch.tensor(start_vals))
    lm_o = H.LM(om, torch.tensor(data), weight=torch.tensor(1 / var), initial_state=o_start).fit()
    psf = H.gaussian_psf(1.3, 13, 0.5)
    tgt = ap.image.Target_Image(data=data, pixelscale=1.0, variance=var, psf=psf, psf_upscale=2)
    fit_m = ap.models.AstroPhot_Model(name="upfit", model_type="sersic galaxy model", target=tgt,
                                      psf_mode="full", parameters=dict(GAL["sersic"], center=[15.3, 12.6]))
    fit_m.parameters.vector_set_values(torch.tensor(start_vals))
    start = fit_m.parameters.vector_representation()
    assert torch.allclose(start, o_start, rtol=1e-12, atol=1e-12)
    lm_g = ap.fit.LM(fit_m, initial_state=start).fit()
    assert lm_o.message.startswith("success"), lm_o.message
    assert lm_g.message.startswith("success"), lm_g.message
    vo = oc.rep_to_values(torch.as_tensor(lm_o.res()))
    v_o = torch.cat([vo[k].reshape(-1) for k in ("center", "q", "PA", "n", "Re", "Ie")])
    v_g = fit_m.parameters.vector_values()
    err = ((v_g - v_o).abs() / v_o.abs().clamp_min(1.0)).max().item()
    assert err < 1e-4, (err, v_g, v_o)
    chi_o, chi_g = min(lm_o.loss_history), min(lm_g.loss_history)
    assert abs(chi_g - chi_o) / chi_o < 1e-6, (chi_g, chi_o)


def test_group_c0_like():
    """Sersic + flat sky group with a 25x25 Gaussian PSF (the 100x100 C0 setup at 64x64)."""
    nx = ny = 64
    psf = H.gaussian_psf(1.5, 25, 1.0)
    tgt = target(nx, ny, psf=psf)
    p = {"center": [32.3, 31.7], "q": 0.7, "PA": 0.6, "n": 2.0, "Re": 8.0, "Ie": 1.0}
    oc, mc = component("sersic", p, tgt, nx, ny, psf=psf, psf_mode="full")
    os_, ms = component("flat_sky", {"F": math.log10(0.5)}, tgt, nx, ny)
    om = H.Model([oc, os_], window=H.Grid(nx, ny))
    g = ap.models.AstroPhot_Model(name="grp", model_type="group model", models=[mc, ms], target=tgt)
    rep = om.vector_representation()
    assert np.allclose(rep.numpy(), g.parameters.vector_representation().numpy())
    assert rel_err(g().data, om.forward(rep)) < IMG_TOL
    J = g.jacobian(parameters=rep.clone(), as_representation=True).data
    Jref = om.jacobian(rep)
    for k in range(J.shape[-1]):
        assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k


def test_group_subwindows_union_and_jacobian_scatter():
    """Two Sersic components on overlapping sub-windows of a 72 x 64 image plus
    a sky on all of it: the group renders over the union of the windows with
    each component confined to its own window (group_model_object.py:92-129,
    183-231), and each component's Jacobian columns land in the group
    Jacobian at its own pixels, zero elsewhere (group_model_object.py:233-283,
    jacobian_image.py:79-118) -- against the oracle."""
    nx, ny = 72, 64
    psf = H.gaussian_psf(1.2, 11, 1.0)
    tgt = target(nx, ny, psf=psf)
    wins = [((4, 44), (6, 50)), ((30, 70), (14, 60))]  # [xmin, xmax), [ymin, ymax) in pixels
    pars = [{"center": [24.6, 27.2], "q": 0.6, "PA": 0.9, "n": 2.2, "Re": 5.0, "Ie": 1.1},
            {"center": [49.3, 37.8], "q": 0.8, "PA": 2.1, "n": 1.1, "Re": 7.0, "Ie": 0.6}]
    ocs, ms = [], []
    for k, (((x0, x1), (y0, y1)), p) in enumerate(zip(wins, pars)):
        g = H.Grid(x1 - x0, y1 - y0, ref_ij=(-0.5 - x0, -0.5 - y0))
        ocs.append(H.Component("sersic", g, dict(p), psf=torch.as_tensor(psf, dtype=torch.float64), psf_mode="full"))
        ms.append(ap.models.AstroPhot_Model(name=f"sub{k}", model_type="sersic galaxy model", target=tgt,
                                            parameters=dict(p), psf_mode="full", window=[[x0, x1], [y0, y1]]))
    osky, msky = component("flat_sky", {"F": math.log10(0.3)}, tgt, nx, ny)
    om = H.Model(ocs + [osky], window=H.Grid(nx, ny))
    grp = ap.models.AstroPhot_Model(name="subgrp", model_type="group model", models=ms + [msky], target=tgt)
    rep = om.vector_representation()
    assert np.allclose(rep.numpy(), grp.parameters.vector_representation().numpy())
    assert rel_err(grp().data, om.forward(rep)) < IMG_TOL
    J = grp.jacobian(parameters=rep.clone(), as_representation=True).data
    Jref = om.jacobian(rep)
    assert J.shape == Jref.shape
    for k in range(J.shape[-1]):
        assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k
    # the first component's columns vanish outside its window
    (x0, x1), (y0, y1) = wins[0]
    outside = torch.ones(ny, nx, dtype=torch.bool)
    outside[y0:y1, x0:x1] = False
    assert float(J[..., :7].cpu()[outside].abs().max()) == 0.0


def test_saved_group_renders_the_same(tmp_path):
    """A group saved to YAML and loaded by filename (core_model.py:386-449,
    group_model_object.py:325-365) renders, and differentiates, bit for bit
    as the group it was saved from."""
    nx, ny = 60, 52
    tgt = target(nx, ny, psf=H.gaussian_psf(1.3, 11, 1.0))
    a = ap.models.AstroPhot_Model(name="a", model_type="sersic galaxy model", target=tgt, psf_mode="full",
                                  parameters={"center": [24.3, 22.8], "q": 0.7, "PA": 0.4, "n": 1.7,
                                              "Re": 6.0, "Ie": 0.9}, window=[[2, 50], [0, 48]])
    b = ap.models.AstroPhot_Model(name="b", model_type="moffat galaxy model", target=tgt,
                                  parameters={"center": [40.1, 30.6], "q": 0.8, "PA": 2.0, "n": 2.5,
                                              "Rd": 3.0, "I0": 0.5})
    grp = ap.models.AstroPhot_Model(name="g", model_type="group model", models=[a, b], target=tgt)
    fn = str(tmp_path / "g.yaml")
    grp.save(fn)
    new = ap.models.AstroPhot_Model(name="g", filename=fn, target=tgt)
    assert torch.equal(new().data, grp().data)
    assert torch.equal(new.jacobian().data, grp.jacobian().data)


def test_chi2_and_normal_equations():
    lib = ap._lib.load()
    torch.manual_seed(0)
    n, npar = 5000, 7
    dev = "cuda"
    J = torch.randn(n, npar, device=dev, dtype=torch.float32)
    m = torch.rand(n, device=dev)
    y = torch.rand(n, device=dev)
    w = torch.rand(n, device=dev) + 0.5
    mask = (torch.rand(n, device=dev) < 0.1).to(torch.uint8)
    wsb = lib.apx_reduce_workspace_bytes(n, npar)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    out = torch.zeros(1, dtype=torch.float64, device=dev)
    st = ap._lib.stream()
    ap._lib.check(lib.apx_chi2(m.data_ptr(), y.data_ptr(), w.data_ptr(), mask.data_ptr(), n, out.data_ptr(),
                               ws.data_ptr(), wsb, st))
    keep = mask == 0
    ref = (w.double() * (y.double() - m.double()) ** 2)[keep].sum()
    assert abs(out.item() - ref.item()) <= 1e-9 * abs(ref.item())
    hess = torch.zeros(npar, npar, dtype=torch.float64, device=dev)
    grad = torch.zeros(npar, dtype=torch.float64, device=dev)
    ap._lib.check(lib.apx_normal_equations(J.data_ptr(), m.data_ptr(), y.data_ptr(), w.data_ptr(), mask.data_ptr(),
                                           n, npar, hess.data_ptr(), grad.data_ptr(), ws.data_ptr(), wsb, st))
    Jd, wd = J.double()[keep], w.double()[keep]
    Href = Jd.T @ (wd[:, None] * Jd)
    gref = Jd.T @ (wd * (y.double() - m.double())[keep])
    assert torch.allclose(hess, Href, rtol=1e-5, atol=1e-5 * Href.abs().max().item())
    assert torch.allclose(grad, gref, rtol=1e-5, atol=1e-5 * gref.abs().max().item())


def test_lm_solve_matches_reference_formula():
    lib = ap._lib.load()
    torch.manual_seed(1)
    npar = 8
    A = torch.randn(40, npar, dtype=torch.float64)
    hess = (A.T @ A).cuda()
    grad = torch.randn(npar, dtype=torch.float64).cuda()
    for Lval in (1e-9, 1e-3, 1.0, 1e4):
        h = torch.empty_like(grad)
        st = torch.zeros(1, dtype=torch.int32, device="cuda")
        ap._lib.check(lib.apx_lm_solve(hess.data_ptr(), grad.data_ptr(), npar, Lval, h.data_ptr(), st.data_ptr(),
                                       ap._lib.stream()))
        ref = H.lm_damped_solve(Lval, grad.cpu(), hess.cpu())
        assert st.item() == 0
        assert torch.allclose(h.cpu(), ref, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("kind", ["sersic", "exponential", "gaussian", "moffat", "flat_sky"])
@pytest.mark.parametrize("psf_mode", ["none", "full"])
def test_adjoint_vjp(kind, psf_mode):
    """Hand-written adjoint vs the oracle's reverse-mode autograd gradient and
    vs J^T G from the forward-mode Jacobian (tolerance 2e-4 of the max)."""
    if kind == "flat_sky" and psf_mode == "full":
        pytest.skip("sky models never convolve")
    kw = dict(psf=PSF11, psf_mode="full") if psf_mode == "full" else {}
    om, m = single(kind, GAL[kind], **kw)
    rep = om.vector_representation()
    rng = np.random.default_rng(11)
    G = torch.tensor(rng.normal(size=(m.window.ny, m.window.nx)))
    r = rep.clone().requires_grad_(True)
    ref, = torch.autograd.grad((om.forward(r) * G).sum(), r)
    got = m.vjp(G.float(), parameters=rep.clone(), as_representation=True).cpu()
    scale = ref.abs().max()
    assert ((got - ref).abs() / scale).max() < JAC_TOL, (got, ref)
    J = m.jacobian(parameters=rep.clone(), as_representation=True).data.double().cpu()
    jtg = (J * G[..., None]).sum(dim=(0, 1))
    assert ((got - jtg).abs() / scale).max() < JAC_TOL


def test_autograd_function_group():
    """torch.autograd through differentiable_sample on the C0-like group."""
    nx = ny = 48
    psf = H.gaussian_psf(1.5, 11, 1.0)
    tgt = target(nx, ny, psf=psf)
    p = {"center": [24.3, 23.6], "q": 0.7, "PA": 0.6, "n": 2.0, "Re": 6.0, "Ie": 1.0}
    oc, mc = component("sersic", p, tgt, nx, ny, psf=psf, psf_mode="full")
    os_, ms = component("flat_sky", {"F": math.log10(0.5)}, tgt, nx, ny)
    om = H.Model([oc, os_], window=H.Grid(nx, ny))
    g = ap.models.AstroPhot_Model(name="grp_ad", model_type="group model", models=[mc, ms], target=tgt)
    rep = om.vector_representation()
    W = torch.tensor(np.random.default_rng(3).uniform(0.5, 1.5, (ny, nx)))
    r = rep.clone().requires_grad_(True)
    ref, = torch.autograd.grad(((om.forward(r) - 1.0) ** 2 * W).sum(), r)
    r2 = rep.clone().cuda().requires_grad_(True)
    img = ap.models.differentiable_sample(g, r2)
    loss = ((img.double() - 1.0) ** 2 * W.cuda()).sum()
    loss.backward()
    assert ((r2.grad.cpu() - ref).abs() / ref.abs().max()).max() < JAC_TOL
    # central finite differences of the oracle (h = 1e-3 in representation space)
    h = 1e-3
    for k in range(len(rep)):
        a, b = rep.clone(), rep.clone()
        a[k] += h
        b[k] -= h
        fd = ((((om.forward(a) - 1.0) ** 2 * W).sum() - ((om.forward(b) - 1.0) ** 2 * W).sum()) / (2 * h)).item()
        assert abs(r2.grad[k].item() - fd) <= 2e-3 * max(abs(fd), ref.abs().max().item() * 1e-3), (k, fd)


def _masked_pair():
    nx, ny = 40, 36
    mask = torch.zeros((ny, nx))
    mask[10, 13] = 1
    mask[17:20, 19:23] = 1
    tgt = target(nx, ny, psf=PSF11)
    p = dict(GAL["sersic"])
    plain = ap.models.AstroPhot_Model(name="plain", model_type="sersic galaxy model", target=tgt, parameters=dict(p),
                                      psf_mode="full")
    masked = ap.models.AstroPhot_Model(name="masked", model_type="sersic galaxy model", target=tgt,
                                       parameters=dict(p), psf_mode="full", mask=mask)
    return tgt, mask, plain, masked


def test_model_mask():
    """model_object.py:370-371 (reference test_model.py:139-161): masked
    samples are zero, the others are the unmasked model's; the Jacobian is the
    derivative of the masked sample and the adjoint its transpose."""
    tgt, mask, plain, masked = _masked_pair()
    keep = torch.logical_not(mask).float().cuda()
    img = masked().data
    assert img[10, 13].item() == 0.0 and img[11, 12].item() != 0.0
    assert torch.equal(img, plain().data * keep)
    J = masked.jacobian().data
    Jp = plain.jacobian().data
    assert torch.equal(J, Jp * keep[..., None])
    G = torch.tensor(np.random.default_rng(4).normal(size=(36, 40)), dtype=torch.float32)
    rep = masked.parameters.vector_representation().clone()
    got = masked.vjp(G, parameters=rep.clone(), as_representation=True).cpu()
    Jr = masked.jacobian(parameters=rep.clone(), as_representation=True).data.double().cpu()
    jtg = (Jr * G.double()[..., None]).sum(dim=(0, 1))
    assert ((got - jtg).abs() / jtg.abs().max()).max() < JAC_TOL


def test_masked_member_keeps_other_light():
    """A group sampled into an image it is handed calls each member on the
    shared working image (group_model_object.py:217-229): a member's mask
    zeroes only its own samples (model_object.py:370-373), not the light the
    other members put there."""
    tgt, mask, plain, masked = _masked_pair()
    sky = ap.models.AstroPhot_Model(name="sky", model_type="flat sky model", target=tgt,
                                    parameters={"F": math.log10(0.5)})
    grp = ap.models.AstroPhot_Model(name="g", model_type="group model", models=[sky, masked], target=tgt)
    img = grp.make_model_image()
    grp.sample(image=img)
    keep = torch.logical_not(mask).float().cuda()
    want = sky().data + plain().data * keep
    assert float((img.data - want).abs().max()) <= 1e-6 * float(want.abs().max())
    assert img.data[10, 13].item() == pytest.approx(sky().data[10, 13].item(), rel=1e-6)
    own = grp()
    assert float((own.data - want).abs().max()) <= 1e-6 * float(want.abs().max())
    # group Jacobian: the masked member's columns vanish at its masked pixels
    J = grp.jacobian().data
    assert float(J[10, 13, 1:].abs().max()) == 0.0 and J[10, 13, 0].item() != 0.0


def test_group_with_upscaled_psf():
    """Sersic + flat sky group on a target whose PSF is 3 times finer
    (target_image.py:303-324): the group image and Jacobian against the
    oracle; the sky ignores the PSF."""
    nx, ny, U = 42, 38, 3
    psf = H.gaussian_psf(1.2, 15, 1.0 / U)
    tgt = ap.image.Target_Image(data=np.zeros((ny, nx)), pixelscale=1.0, psf=psf, psf_upscale=U)
    p = {"center": [21.3, 18.6], "q": 0.7, "PA": 0.6, "n": 2.0, "Re": 6.0, "Ie": 1.0}
    oc = H.Component("sersic", H.Grid(nx, ny), dict(p), psf=torch.tensor(psf), psf_mode="full", psf_upscale=U)
    osky = H.Component("flat_sky", H.Grid(nx, ny), {"F": math.log10(0.5)})
    om = H.Model([oc, osky], window=H.Grid(nx, ny))
    ms = ap.models.AstroPhot_Model(name="s_up", model_type="sersic galaxy model", target=tgt, parameters=dict(p),
                                   psf_mode="full")
    mk = ap.models.AstroPhot_Model(name="sky_up", model_type="flat sky model", target=tgt,
                                   parameters={"F": math.log10(0.5)})
    g = ap.models.AstroPhot_Model(name="g_up", model_type="group model", models=[ms, mk], target=tgt)
    rep = om.vector_representation()
    assert np.allclose(rep.numpy(), g.parameters.vector_representation().numpy())
    assert rel_err(g().data, om.forward(rep)) < IMG_TOL
    J = g.jacobian(parameters=rep.clone(), as_representation=True).data
    Jref = om.jacobian(rep)
    for k in range(J.shape[-1]):
        assert rel_err(J[..., k], Jref[..., k]) < JAC_TOL, k
