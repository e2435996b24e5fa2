"""Moment-based initialiser (apx_batch_moment_init).

* the device estimator equals a numpy restatement of the same estimator on
  the same cutouts (centre, PA, q, Re, n, Ie to float32 accuracy of the
  sums, away from the discrete decisions);
* on seeded C2-like cutouts it lands near the truth (centre within 1 px,
  Re within a factor 2), and LM started from it ends where LM started from
  the perturbed truth for the well-constrained fits.
"""

import math

import numpy as np
import pytest
import torch

from astrophot_amd.batch import BatchFit, Cutout, rep_to_values, values_to_rep
from astrophot_amd.population import perturbed_start, sample_population

pytestmark = pytest.mark.gpu


def numpy_moment_init(img, psf):
    """The estimator of csrc/init.hip restated with numpy (float64)."""
    ny, nx = img.shape
    edge_px = np.concatenate([img[0, :], img[-1, :], img[1:-1, 0], img[1:-1, -1]])
    edge, scatter = edge_px.mean(), edge_px.std()
    ph, pw = psf.shape
    yy, xx = np.mgrid[0:ph, 0:pw]
    psf2 = 0.5 * (psf * ((xx - (pw - 1) // 2) ** 2 + (yy - (ph - 1) // 2) ** 2)).sum() / psf.sum()
    win = max(min(min(nx, ny) // 10, 30), 6)
    win += win % 2
    cen = np.array([0.5 * (ny - 1), 0.5 * (nx - 1)])
    init = cen.copy()
    for _ in range(100):
        r0, c0 = int(np.rint(cen[0])) - win // 2, int(np.rint(cen[1])) - win // 2
        if r0 < 0 or c0 < 0 or r0 + win >= ny or c0 + win >= nx:
            cen = init
            break
        sub = img[r0:r0 + win, c0:c0 + win]
        rr, cc = np.mgrid[0:win, 0:win]
        new = np.array([r0 + (sub * rr).sum() / sub.sum(), c0 + (sub * cc).sum() / sub.sum()])
        if np.abs(new - cen).sum() < 0.1:
            break
        cen = new
    pr, pc = cen
    J, I = np.mgrid[0:ny, 0:nx]
    X, Y = I - pc, J - pr
    w = img - edge
    r2 = X**2 + Y**2
    ok = r2 > 0
    cs = (w[ok] * (X[ok]**2 - Y[ok]**2) / r2[ok]).sum()
    sn = (w[ok] * 2 * X[ok] * Y[ok] / r2[ok]).sum()
    ang = 0.5 * math.atan2(sn, cs)
    m = w > 3 * scatter
    u = math.cos(ang) * X + math.sin(ang) * Y
    v = -math.sin(ang) * X + math.cos(ang) * Y
    mw = w[m].sum()
    q = 0.7
    if mw > 0:
        a2 = max((w[m] * u[m] ** 2).sum() / mw - psf2, 0.25)
        b2 = max((w[m] * v[m] ** 2).sum() / mw - psf2, 0.01)
        q = math.sqrt(b2 / a2)
    if q > 1:
        q, ang = 1 / q, ang + math.pi / 2
    q = min(max(q, 0.1), 0.95)
    u = math.cos(ang) * X + math.sin(ang) * Y
    v = (-math.sin(ang) * X + math.cos(ang) * Y) / q
    R = np.sqrt(u**2 + v**2)
    return dict(centre=(pc, pr), ang=ang % math.pi, q=q, R=R, w=w, psf2=psf2)


@pytest.fixture(scope="module")
def population():
    pop = sample_population(48, seed=11, side_range=(40, 140), snr_range=(80, 1000), q_range=(0.3, 0.8))
    cuts = [Cutout(np.zeros((s, s)), np.ones((s, s)), p) for s, p in zip(pop.sides, pop.psfs)]
    bf0 = BatchFit(cuts)
    imgs = [t.double().cpu().numpy() for t in bf0.split(bf0.render(values_to_rep(pop.truth)))]
    rng = np.random.default_rng(12)
    data = [im + rng.normal(size=im.shape) * sig for im, sig in zip(imgs, pop.sky_sigma)]
    cuts = [Cutout(d, np.full(d.shape, sig**2), p) for d, sig, p in zip(data, pop.sky_sigma, pop.psfs)]
    return pop, cuts, data


def test_moment_init_matches_numpy_restatement(population):
    pop, cuts, data = population
    bf = BatchFit(cuts)
    vals = rep_to_values(bf.moment_init().cpu().numpy())
    for k, (img, psf) in enumerate(zip(data, pop.psfs)):
        ref = numpy_moment_init(img, psf)
        assert abs(vals[k, 0] - (ref["centre"][0] + 0.5)) < 1e-3, k
        assert abs(vals[k, 1] - (ref["centre"][1] + 0.5)) < 1e-3, k
        dpa = (vals[k, 3] - (ref["ang"] + math.pi / 2)) % math.pi
        assert min(dpa, math.pi - dpa) < 1e-3, k
        assert abs(vals[k, 2] - ref["q"]) < 1e-3, k


def test_moment_init_near_truth_and_lm_agrees(population):
    pop, cuts, data = population
    bf = BatchFit(cuts)
    rep_m = bf.moment_init()
    vm = rep_to_values(rep_m.cpu().numpy())
    t = pop.truth
    assert np.all(np.abs(vm[:, :2] - t[:, :2]) < 1.0)
    ratio = vm[:, 5] / t[:, 5]
    assert np.all((ratio > 0.5) & (ratio < 2.0)), ratio
    res_m = bf.fit(rep_m)
    res_t = bf.fit(torch.tensor(values_to_rep(perturbed_start(t, 13)), dtype=torch.float64, device="cuda"))
    both = (res_m["status"] == 1) & (res_t["status"] == 1)
    assert both.mean() > 0.8, (res_m["status"], res_t["status"])
    rel = np.abs(res_m["chi2"][both] - res_t["chi2"][both]) / res_t["chi2"][both]
    assert np.median(rel) < 1e-6 and np.mean(rel < 1e-4) > 0.9, rel
