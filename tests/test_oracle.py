"""Pins of the float64 oracle (CPU only).

Each test checks an oracle function against something that does not restate
it: the reference's own known-answer tests, golden vectors written by running
the reference's utils/operations.py, library routines (scipy, numpy), closed
forms, invariants and brute force on tiny inputs.
"""

import math
import os

import numpy as np
import pytest
import torch
from scipy.signal import fftconvolve
from scipy.special import erf, gamma, roots_legendre

from oracle import hotpath as H

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "operations_vectors.npz"))
F64 = torch.float64


# -- reference KATs: test_utils.py:344-434 (Sersic conversions) ---------------
def test_sersic_flux_kat():
    b1 = H.sersic_n_to_b(1.0)
    got = H.sersic_Ie_to_flux(torch.tensor(1.0, dtype=F64), torch.tensor(1.0, dtype=F64), 1.0, 1.0)
    assert got.item() == pytest.approx(2 * np.pi * gamma(2) * np.exp(b1) * b1 ** (-2), rel=1e-12)


@pytest.mark.parametrize("n", [0.7, 1.0, 2.5, 4.0])
def test_sersic_flux_brute_force(n):
    """Integral of the profile over a fine polar grid equals the total flux."""
    Re, q = 2.0, 0.6
    r = np.geomspace(1e-6, 400 * Re, 400001)
    I = H.sersic(torch.tensor(r), torch.tensor(n), torch.tensor(Re), torch.tensor(1.0)).numpy()
    flux = 2 * np.pi * q * np.trapezoid(I * r, r)  # elliptical: area element scales by q
    F = H.sersic_Ie_to_flux(torch.tensor(1.0, dtype=F64), torch.tensor(n, dtype=F64), Re, q).item()
    assert flux == pytest.approx(F, rel=2e-4)


def test_b_n_half_light():
    """b(n) makes Re the half-light radius: incomplete gamma P(2n, b) = 1/2."""
    from scipy.special import gammainc
    for n in (0.5, 1.0, 2.0, 4.0, 8.0):
        assert gammainc(2 * n, H.sersic_n_to_b(n)) == pytest.approx(0.5, abs=2e-4)


def test_profile_closed_forms():
    r = np.geomspace(1e-8, 2000.0, 400001)
    # gaussian with the 1-D normalisation quirk: integral = I0 * sqrt(2 pi) * sigma
    s = 1.7
    I = H.gaussian(torch.tensor(r), torch.tensor(s), torch.tensor(1.0)).numpy()
    assert 2 * np.pi * np.trapezoid(I * r, r) == pytest.approx(np.sqrt(2 * np.pi) * s, rel=1e-6)
    # moffat: integral = I0 pi Rd^2 / (n - 1)   (conversions/functions.py moffat_I0_to_flux)
    n, Rd = 3.0, 2.2
    I = H.moffat(torch.tensor(r), torch.tensor(n), torch.tensor(Rd), torch.tensor(1.0)).numpy()
    assert 2 * np.pi * np.trapezoid(I * r, r) == pytest.approx(np.pi * Rd**2 / (n - 1), rel=1e-5)
    # exponential is the Sersic n = 1 profile
    R = torch.linspace(0.01, 30, 200, dtype=F64)
    e = H.exponential(R, torch.tensor(3.0, dtype=F64), torch.tensor(2.0, dtype=F64))
    s1 = H.sersic(R, torch.tensor(1.0, dtype=F64), torch.tensor(3.0, dtype=F64), torch.tensor(2.0, dtype=F64))
    assert torch.allclose(e, s1, rtol=1e-14)


# -- parameter maps: test_parameter.py:117-161, 379-418 ------------------------
def test_boundaries_round_trip_and_kats():
    v = torch.tensor([0.2, 0.5, 0.93], dtype=F64)
    assert torch.allclose(H.inv_boundaries(H.boundaries(v, 0.0, 1.0), 0.0, 1.0), v, rtol=1e-14)
    v = torch.tensor([1.5, 4.0, 100.0], dtype=F64)
    assert torch.allclose(H.inv_boundaries(H.boundaries(v, 1.0, None), 1.0, None), v, rtol=1e-13)
    v = torch.tensor([-100.0, -3.0, 0.5], dtype=F64)
    assert torch.allclose(H.inv_boundaries(H.boundaries(v, None, 1.0), None, 1.0), v, rtol=1e-13)
    # cyclic wrap of (-1, 1): 2 -> 0 and -2 -> 0 (test_parameter.py:146-161)
    assert H.cyclic_boundaries(torch.tensor(2.0), -1.0, 1.0).item() == 0.0
    assert H.cyclic_boundaries(torch.tensor(-2.0), -1.0, 1.0).item() == 0.0


def test_boundaries_derivatives_closed_form():
    """optimization.py:40-47 closed forms vs autograd of inv_boundaries."""
    r = torch.linspace(-5, 5, 41, dtype=F64, requires_grad=True)
    for lo, hi, f in [(0.0, 1.0, lambda r: (1.0 - 0.0) / (np.pi * (r**2 + 1))),
                      (0.0, None, lambda r: 0.5 + 0.5 * (r - 0.0) / torch.sqrt((r - 0.0) ** 2 + 4)),
                      (None, 2.0, lambda r: 0.5 - 0.5 * (r - 2.0) / torch.sqrt((r - 2.0) ** 2 + 4))]:
        g, = torch.autograd.grad(H.inv_boundaries(r, lo, hi).sum(), r)
        assert torch.allclose(g, f(r.detach()), rtol=1e-12)


# -- quadrature: operations.py:94-247 -----------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7])
def test_quad_table_golden(n):
    P = torch.tensor(GOLD["P"])
    x, y, w = H.quad_table(n, P)
    assert np.allclose(x.numpy(), GOLD[f"qt{n}_x"], rtol=0, atol=1e-15)
    assert np.allclose(y.numpy(), GOLD[f"qt{n}_y"], rtol=0, atol=1e-15)
    assert np.allclose(w.numpy(), GOLD[f"qt{n}_w"], rtol=0, atol=1e-15)
    # numpy's leggauss (what the HIP path's host side passes to the kernels)
    # gives the table scipy's roots_legendre gives the reference
    a, wt = np.polynomial.legendre.leggauss(n)
    sa, sw = roots_legendre(n)
    assert np.allclose(a, sa, atol=1e-15) and np.allclose(wt, sw, atol=1e-15)
    assert np.isclose(w.sum().item(), 1.0, atol=1e-14)


def test_quad_exact_on_polynomials():
    """3x3 Gauss-Legendre integrates x^a y^b (a, b <= 5) over a unit pixel exactly."""
    P = torch.eye(2, dtype=F64)
    for a in range(6):
        for b in range(6):
            f = lambda X, Y, P: X**a * Y**b
            res, _ = H.single_quad_integrate(torch.zeros(1, dtype=F64), torch.zeros(1, dtype=F64), P, f, 3)
            exact = ((0.5 ** (a + 1) - (-0.5) ** (a + 1)) / (a + 1)) * ((0.5 ** (b + 1) - (-0.5) ** (b + 1)) / (b + 1))
            assert res.item() == pytest.approx(exact, abs=1e-15)


def _gold_brightness(X, Y, P):
    n, Re, Ie, q, PA, soft = GOLD["params"]
    th = -(PA - math.pi / 2)
    c, s = math.cos(th), math.sin(th)
    Xr, Yr = c * X - s * Y, (s * X + c * Y) / q
    R = torch.sqrt(Xr**2 + Yr**2 + soft**2)
    area = torch.linalg.det(P).abs()
    return area * 10**Ie * torch.exp(-H.sersic_n_to_b(n) * (torch.pow(R / Re, 1 / n) - 1))


@pytest.mark.parametrize("depth", [1, 2, 3])
@pytest.mark.parametrize("ref", [1e-3, 1e-2])
def test_grid_integrate_golden(depth, ref):
    """Vectors written by the reference's own grid_integrate."""
    X, Y, P = (torch.tensor(GOLD[k]) for k in ("X", "Y", "P"))
    got = H.grid_integrate(X, Y, P, _gold_brightness, 3, 5, 1, depth, ref)
    assert np.allclose(got.numpy(), GOLD[f"grid_d{depth}_r{ref:g}"], rtol=1e-13, atol=0)


def test_single_quad_golden():
    X, Y, P = (torch.tensor(GOLD[k]) for k in ("X", "Y", "P"))
    res, refc = H.single_quad_integrate(X, Y, P, _gold_brightness, 5)
    assert np.allclose(res.numpy(), GOLD["quad5_res"], rtol=1e-13)
    assert np.allclose(refc.numpy(), GOLD["quad5_ref"], rtol=1e-13)


def test_grid_integrate_gaussian_erf():
    """Deep adaptive integration of a narrow round Gaussian matches the erf
    closed form of its pixel integral."""
    s = 0.6
    P = torch.eye(2, dtype=F64)
    f = lambda X, Y, P: torch.linalg.det(P).abs() * torch.exp(-0.5 * (X**2 + Y**2) / s**2) / (2 * np.pi * s**2)
    xs = torch.tensor([0.0, 0.3, 0.8, -1.2], dtype=F64)
    ys = torch.tensor([0.0, -0.4, 0.5, 0.9], dtype=F64)
    got = H.grid_integrate(xs, ys, P, f, 3, 5, 1, 3, 1e-9)
    ex = [0.25 * (erf((x + 0.5) / (s * np.sqrt(2))) - erf((x - 0.5) / (s * np.sqrt(2))))
          * (erf((y + 0.5) / (s * np.sqrt(2))) - erf((y - 0.5) / (s * np.sqrt(2)))) for x, y in zip(xs.numpy(), ys.numpy())]
    assert np.allclose(got.numpy(), ex, rtol=2e-7, atol=1e-12)


# -- curvature, bilinear shift, convolution ------------------------------------
def test_curvature_brute_force():
    rng = np.random.default_rng(3)
    m = rng.normal(size=(7, 9))
    got = H.curvature(torch.tensor(m)).numpy()
    for j in range(7):
        for i in range(9):
            jj, ii = min(max(j, 1), 5), min(max(i, 1), 7)
            want = m[jj - 1, ii] + m[jj + 1, ii] + m[jj, ii - 1] + m[jj, ii + 1] - 4 * m[jj, ii]
            assert got[j, i] == pytest.approx(want, abs=1e-12)


@pytest.mark.parametrize("shift", [(0.3, -0.2), (-0.5, 0.5), (0.0, 0.0), (0.49, 0.01)])
def test_shift_psf_brute_force(shift):
    """interp2d with clamped corners, written out pixel by pixel."""
    rng = np.random.default_rng(4)
    psf = rng.uniform(0.1, 1.0, (5, 3))
    got = H.shift_psf(torch.tensor(psf), torch.tensor(shift, dtype=F64)).numpy()
    p = np.pad(psf, 1)
    h, w = p.shape
    for r in range(h):
        for c in range(w):
            x, y = c - shift[0], r - shift[1]
            x0f, y0f = math.floor(x), math.floor(y)
            x0, x1 = min(max(x0f, 0), w - 2), min(max(x0f + 1, 1), w - 1)
            y0, y1 = min(max(y0f, 0), h - 2), min(max(y0f + 1, 1), h - 1)
            want = (p[y0, x0] * (x1 - x) * (y1 - y) + p[y1, x0] * (x1 - x) * (y - y0)
                    + p[y0, x1] * (x - x0) * (y1 - y) + p[y1, x1] * (x - x0) * (y - y0))
            assert got[r, c] == pytest.approx(want, abs=1e-14)
    if shift == (0.0, 0.0):
        assert np.allclose(got[1:-1, 1:-1], psf)


def test_fft_convolve_golden_and_scipy():
    img, psf = torch.tensor(GOLD["fft_img"]), torch.tensor(GOLD["fft_psf"])
    assert np.allclose(H.fft_convolve(img, psf).numpy(), GOLD["fft_prepadded"], rtol=1e-12, atol=1e-12)
    # reference KAT (test_utils.py:15-38): interior equals scipy fftconvolve mode="same"
    sp = fftconvolve(img.numpy(), psf.numpy(), mode="same")
    got = H.fft_convolve(img, psf).numpy()
    assert np.allclose(got[4:-4, 3:-3], sp[4:-4, 3:-3], rtol=1e-10, atol=1e-12)
    # and the direct mode agrees on the interior (_model_methods.py:247-255)
    d = H.direct_convolve(img, psf).numpy()
    assert np.allclose(got[4:-4, 3:-3], d[4:-4, 3:-3], rtol=1e-10, atol=1e-12)


# -- whole-component invariants -----------------------------------------------
def _comp(kind="gaussian", center=(20.0, 18.0), psf=None, **kw):
    p = {"gaussian": {"q": 0.999999, "PA": 0.4, "sigma": 2.0, "flux": 1.0},
         "sersic": {"q": 0.6, "PA": 0.4, "n": 2.0, "Re": 4.0, "Ie": 0.0}}[kind]
    p = dict(p, center=list(center))
    return H.Component(kind, H.Grid(40, 36), p, psf=psf, **kw)


def test_flux_conservation_under_psf():
    """A unit-sum PSF moves flux but does not change its total (centre on a
    pixel centre, so the sampling grids coincide)."""
    psf = torch.tensor(H.gaussian_psf(1.2, 9, 1.0))
    a = _comp(center=(20.5, 18.5), psf=psf, psf_mode="full", integrate_mode="none")
    b = _comp(center=(20.5, 18.5), integrate_mode="none")
    ia = H.Model([a]).forward(H.Model([a]).vector_representation())
    ib = H.Model([b]).forward(H.Model([b]).vector_representation())
    assert ia.sum().item() == pytest.approx(ib.sum().item(), rel=1e-9)


def test_whole_pixel_translation():
    psf = torch.tensor(H.gaussian_psf(1.2, 9, 1.0))
    a = _comp("sersic", (18.3, 16.7), psf=psf, psf_mode="full")
    b = _comp("sersic", (21.3, 18.7), psf=psf, psf_mode="full")
    ia = H.Model([a]).forward(H.Model([a]).vector_representation())
    ib = H.Model([b]).forward(H.Model([b]).vector_representation())
    # selection threshold uses the same reference (same numel, same flux): images shift exactly
    assert torch.allclose(ib[2 + 5:-5, 3 + 5:-5], ia[5:-5 - 2, 5:-5 - 3], rtol=1e-12, atol=1e-14)


def test_round_profile_pa_independent_and_symmetric():
    a = H.Component("sersic", H.Grid(31, 31), {"center": [15.5, 15.5], "q": 0.99999999, "PA": 0.3,
                                               "n": 1.5, "Re": 4.0, "Ie": 0.0})
    b = H.Component("sersic", H.Grid(31, 31), {"center": [15.5, 15.5], "q": 0.99999999, "PA": 2.1,
                                               "n": 1.5, "Re": 4.0, "Ie": 0.0})
    ia = H.Model([a]).forward(H.Model([a]).vector_representation())
    ib = H.Model([b]).forward(H.Model([b]).vector_representation())
    assert torch.allclose(ia, ib, rtol=1e-7)
    assert torch.allclose(ia, ia.flip(0), rtol=1e-7)
    assert torch.allclose(ia, ia.T, rtol=1e-7)


def test_jacobian_vs_central_fd_and_reverse_mode():
    psf = torch.tensor(H.gaussian_psf(1.2, 7, 1.0))
    c = _comp("sersic", (20.3, 17.6), psf=psf, psf_mode="full")
    m = H.Model([c])
    rep = m.vector_representation()
    J = m.jacobian(rep)
    h = 1e-5
    for k in range(len(rep)):
        r1, r2 = rep.clone(), rep.clone()
        r1[k] += h
        r2[k] -= h
        fd = (m.forward(r1) - m.forward(r2)) / (2 * h)
        assert (fd - J[..., k]).abs().max() <= 1e-6 * J[..., k].abs().max() + 1e-12
    # reverse mode on a projection
    r = rep.clone().requires_grad_(True)
    wgt = torch.linspace(0, 1, J.shape[0] * J.shape[1], dtype=F64).reshape(J.shape[:2])
    g, = torch.autograd.grad((m.forward(r) * wgt).sum(), r)
    assert torch.allclose(g, (J * wgt[..., None]).sum(dim=(0, 1)), rtol=1e-10)


# -- PSF models: psf_model_object.py:180-264, _model_methods.py:97-109 ---------
class _PolyComp:
    """Stand-in component whose brightness is a polynomial in (x, y)."""

    sampling_mode = "simpsons"

    def __init__(self, a, b):
        self.a, self.b = a, b

    def evaluate(self, X, Y, P, v):
        return (X + 0.3) ** self.a * (Y - 0.2) ** self.b + 0.5 * X * Y


@pytest.mark.parametrize("P", [0.8, [[0.7, 0.2], [-0.1, 0.9]]])
def test_simpsons_sampling_exact_on_cubics(P):
    """Simpson's 3 x 3 rule on each (parallelogram) pixel integrates a cubic
    exactly: checked against 4-point Gauss-Legendre quadrature of the same
    pixel; the reference sample is the pixel-centre value."""
    grid = H.Grid(5, 4, P, ref_ij=(1.3, 0.4), ref_xy=(0.2, -0.7))
    xg, wg = roots_legendre(4)
    # degree <= 3 in each pixel coordinate: any total degree <= 3, and x^a y^b
    # with a, b <= 3 when the pixel axes are the plane axes
    cases = [(0, 0), (1, 2), (3, 0), (2, 1), (0, 3)] + ([(2, 3), (3, 3)] if np.ndim(P) == 0 else [])
    for a, b in cases:
        comp = _PolyComp(a, b)
        mid, simps = H.sample_init(comp, grid, {}, torch.zeros(2, dtype=F64))
        C = grid.meshgrid()
        Pm = grid.P
        exact = torch.zeros_like(mid)
        for ux, wx in zip(xg / 2, wg / 2):
            for uy, wy in zip(xg / 2, wg / 2):
                dx = Pm[0, 0] * ux + Pm[0, 1] * uy
                dy = Pm[1, 0] * ux + Pm[1, 1] * uy
                exact += wx * wy * comp.evaluate(C[0] + dx, C[1] + dy, Pm, None)
        assert torch.allclose(simps, exact, rtol=1e-12, atol=1e-12), (a, b)
        assert torch.allclose(mid, comp.evaluate(C[0], C[1], Pm, None), rtol=1e-14, atol=1e-14)


def _laplacian_replicate(a):
    """5-point Laplacian on the interior, border rows/columns copied outward."""
    ny, nx = a.shape
    L = torch.zeros_like(a)
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            L[j, i] = a[j, i - 1] + a[j, i + 1] + a[j - 1, i] + a[j + 1, i] - 4 * a[j, i]
    L[0, :], L[-1, :] = L[1, :], L[-2, :]
    L[:, 0], L[:, -1] = L[:, 1], L[:, -2]
    return L


def test_trapezoid_sampling_round_gaussian_closed_form():
    """_model_methods.py:121-139.  A round Gaussian is separable in the plane
    axes, so with a diagonal pixelscale the mean of a pixel's four corner
    values is the product of the means of the 1-D profile at the pixel's two
    x edges and two y edges.  The returned reference is that mean plus its
    Laplacian (replicate border)."""
    P = [[0.8, 0.0], [0.0, 1.1]]
    grid = H.Grid(9, 7, P, ref_ij=(0.7, -0.2), ref_xy=(0.4, 0.1))
    c = torch.tensor([3.3, 2.1], dtype=F64)
    sigma, flux = 1.7, 0.3
    comp = H.Component("gaussian", grid, {"center": c.tolist(), "q": 1.0, "PA": 0.4, "sigma": sigma,
                                          "flux": flux}, sampling_mode="trapezoid")
    v = {"center": c, "q": torch.tensor(1.0, dtype=F64), "PA": torch.tensor(0.4, dtype=F64),
         "sigma": torch.tensor(sigma, dtype=F64), "flux": torch.tensor(flux, dtype=F64)}
    ref, trapz = H.sample_init(comp, grid, v, c)
    xe = 0.8 * (np.arange(10) - 0.5 - 0.7) + 0.4 - 3.3
    ye = 1.1 * (np.arange(8) - 0.5 + 0.2) + 0.1 - 2.1
    ex, ey = np.exp(-xe ** 2 / (2 * sigma ** 2)), np.exp(-ye ** 2 / (2 * sigma ** 2))
    C = 0.8 * 1.1 * 10 ** flux / math.sqrt(2 * math.pi * sigma ** 2) * math.exp(-1e-6 / (2 * sigma ** 2))
    closed = C * np.outer(0.5 * (ey[:-1] + ey[1:]), 0.5 * (ex[:-1] + ex[1:]))
    assert np.allclose(trapz.numpy(), closed, rtol=1e-13, atol=0)
    assert torch.allclose(ref, trapz + _laplacian_replicate(trapz), rtol=1e-13, atol=1e-16)


class _PixelBilinear:
    """Stand-in component whose brightness is (a + b u)(c + d v) in the pixel
    coordinates (u, v) of a sheared grid."""

    sampling_mode = "trapezoid"

    def __init__(self, grid):
        self.Pinv = torch.linalg.inv(grid.P)

    def evaluate(self, X, Y, P, v):
        u = self.Pinv[0, 0] * X + self.Pinv[0, 1] * Y
        w = self.Pinv[1, 0] * X + self.Pinv[1, 1] * Y
        return (0.7 + 0.3 * u) * (1.2 - 0.4 * w)


def test_trapezoid_sampling_exact_on_pixel_bilinear():
    """The corner mean is the pixel average of a function bilinear in the
    pixel coordinates: for (a + b u)(c + d v) that average is its value at the
    pixel centre, whatever the shear of the pixelscale."""
    grid = H.Grid(6, 5, [[0.7, 0.2], [-0.1, 0.9]], ref_ij=(1.3, 0.4), ref_xy=(0.2, -0.7))
    comp = _PixelBilinear(grid)
    _, trapz = H.sample_init(comp, grid, {}, torch.zeros(2, dtype=F64))
    Cm = grid.meshgrid()
    assert torch.allclose(trapz, comp.evaluate(Cm[0], Cm[1], grid.P, None), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("U", [2, 3])
def test_psf_upscale_block_sums(U):
    """model_object.py:313-349 with a PSF U times finer than the image
    (target_image.py:303-324): the model is sampled on a grid of pixelscale
    P * float32(1/U) and shape floor(n / float32(1/U)), convolved there, and
    U x U blocks are summed back onto the image grid.  With U = 3 the float32
    scale leaves the fine grid one pixel short (3n - 1), so the last row and
    column of the image receive nothing.  Against a direct restatement:
    profile values at the fine pixel centres, scipy convolve2d, explicit block
    sums.  The centre sits on a fine pixel centre, so the sub-pixel shift is
    zero and the shifted kernel is the PSF itself."""
    from scipy.signal import convolve2d
    nx, ny, sigma, flux = 12, 10, 2.0, 0.5
    s = float(np.float32(1.0 / U))
    nfx, nfy = int(nx // s), int(ny // s)
    assert (nfx, nfy) == ((U * nx - 1, U * ny - 1) if U == 3 else (U * nx, U * ny))
    rng = np.random.default_rng(U)
    psf = rng.uniform(0.1, 1.0, size=(5, 5))
    psf /= psf.sum()
    kx, ky = 17, 13
    c = [s * (kx + 0.5), s * (ky + 0.5)]
    comp = H.Component("gaussian", H.Grid(nx, ny), {"center": c, "q": 1.0, "PA": 0.4, "sigma": sigma, "flux": flux},
                       psf=torch.tensor(psf), psf_mode="full", psf_upscale=U, integrate_mode="none")
    m = H.Model([comp])
    got = m.forward(m.vector_representation()).numpy()
    b = 3
    xf = s * (np.arange(nfx + 2 * b) - b + 0.5) - c[0]
    yf = s * (np.arange(nfy + 2 * b) - b + 0.5) - c[1]
    r2 = xf[None, :] ** 2 + yf[:, None] ** 2 + 1e-6
    f = s * s * 10 ** flux / math.sqrt(2 * math.pi * sigma ** 2) * np.exp(-r2 / (2 * sigma ** 2))
    conv = convolve2d(f, psf, mode="same")[b:-b, b:-b]
    want = np.zeros((ny, nx))
    for j in range(nfy // U):
        for i in range(nfx // U):
            want[j, i] = conv[U * j:U * j + U, U * i:U * i + U].sum()
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15 * want.max())
    if U == 3:
        assert np.all(got[-1, :] == 0) and np.all(got[:, -1] == 0)


def _psf_comp(n=2.5, Rd=2.2, side=15, P=1.0, **kw):
    g = H.Grid(side, side, P, ref_ij=((side - 1) / 2, (side - 1) / 2), ref_xy=(0.0, 0.0))
    kw.setdefault("sampling_mode", "simpsons")
    kw.setdefault("sampling_tolerance", 1e-3)
    return H.Component("moffat_psf", g, {"n": n, "Rd": Rd}, locked=("center", "I0"), **kw)


@pytest.mark.parametrize("n,Rd", [(2.5, 2.2), (1.2, 0.9), (4.0, 5.0)])
def test_psf_sample_brute_force(n, Rd):
    """A sampled Moffat PSF model is the normalised pixel integral of
    (1 + (r/Rd)^2)^-n: against 60 x 60 midpoint sub-sampling of every pixel,
    unit sum, and the symmetries of a centred round profile."""
    c = _psf_comp(n, Rd)
    m = H.Model([c])
    img = m.forward(m.vector_representation())
    assert img.sum().item() == pytest.approx(1.0, abs=1e-13)
    assert torch.allclose(img, img.T, rtol=1e-12) and torch.allclose(img, img.flip(0), rtol=1e-12)
    k = 60
    u = (torch.arange(k, dtype=F64) + 0.5) / k - 0.5
    xs = (torch.arange(15, dtype=F64) - 7)[:, None] + u[None, :]
    X, Y = xs.reshape(-1)[None, :], xs.reshape(-1)[:, None]
    f = (1 + (X**2 + Y**2 + 1e-6) / Rd**2) ** -n
    brute = f.reshape(15, k, 15, k).mean(dim=(1, 3))
    brute = brute / brute.sum()
    assert (img - brute).abs().max().item() < 2e-4 * brute.max().item()


def test_aux_psf_model_matches_image_psf_and_fd():
    """A galaxy convolved with an auxiliary PSF model renders exactly as with
    that model's sampled image as a fixed PSF, and its Jacobian, the PSF
    parameters' columns included (first in the vector), matches central
    finite differences."""
    aux = _psf_comp(2.2, 1.8, side=9)
    psf_img, _ = H.psf_sample(aux, aux.rep_to_values(aux.val_to_rep(aux.vector_values())))
    a = _comp("sersic", (20.3, 17.6), psf=aux, psf_mode="full")
    b = _comp("sersic", (20.3, 17.6), psf=psf_img, psf_mode="full")
    ma, mb = H.Model([a]), H.Model([b])
    ra = ma.vector_representation()
    assert len(ra) == len(mb.vector_representation()) + 2
    assert torch.equal(ra[2:], mb.vector_representation())
    assert torch.equal(ma.forward(ra), mb.forward(mb.vector_representation()))
    J = ma.jacobian(ra)
    h = 1e-5
    for k in range(len(ra)):
        r1, r2 = ra.clone(), ra.clone()
        r1[k] += h
        r2[k] -= h
        fd = (ma.forward(r1) - ma.forward(r2)) / (2 * h)
        assert (fd - J[..., k]).abs().max() <= 1e-6 * J[..., k].abs().max() + 1e-12, k
    assert J[..., 0].abs().max() > 0 and J[..., 1].abs().max() > 0


# -- chi^2 KATs: test_utils.py:55-121 ------------------------------------------
def test_chi2_kats():
    ones, zeros = torch.ones(10, dtype=F64), torch.zeros(10, dtype=F64)
    mask = torch.zeros(10, dtype=torch.bool)
    mask[2] = True
    var = 2 * torch.ones(10, dtype=F64)
    assert H.chi_squared(ones, zeros, mask=mask, variance=var).item() == 4.5
    assert H.reduced_chi_squared(ones, zeros, 3, mask=mask, variance=var).item() == 0.75
    assert H.chi_squared(ones, zeros, variance=var).item() == 5
    assert H.reduced_chi_squared(ones, zeros, 3, variance=var).item() == pytest.approx(5 / 7)
    assert H.chi_squared(ones, zeros, mask=mask).item() == 9
    assert H.reduced_chi_squared(ones, zeros, 3, mask=mask).item() == 1.5
    assert H.chi_squared(ones, zeros).item() == 10
    assert H.reduced_chi_squared(ones, zeros, 3).item() == pytest.approx(10 / 7)


# -- LM --------------------------------------------------------------------------
def test_lm_damped_solve_numpy():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(30, 6))
    Hm, g = A.T @ A, rng.normal(size=6)
    for L in (1e-9, 0.1, 1.0, 1e5):
        M = Hm / (1 + L)
        np.fill_diagonal(M, np.diag(Hm) + L * (1 + np.diag(Hm)))
        want = np.linalg.solve(M, g)
        got = H.lm_damped_solve(L, torch.tensor(g), torch.tensor(Hm)).numpy()
        assert np.allclose(got, want, rtol=1e-10)


def test_lm_recovers_truth_noiseless():
    psf = torch.tensor(H.gaussian_psf(1.2, 7, 1.0))
    truth = {"center": [16.2, 15.7], "q": 0.7, "PA": 0.6, "n": 1.8, "Re": 4.0, "Ie": 1.0}
    c = H.Component("sersic", H.Grid(32, 32), truth, psf=psf, psf_mode="full")
    m = H.Model([c])
    rep0 = m.vector_representation()
    data = m.forward(rep0)
    start = rep0.clone()
    start[2:] *= 1.05
    start[:2] += 0.2
    lm = H.LM(m, data, weight=torch.ones_like(data) * 100.0, initial_state=start).fit()
    assert lm.message.startswith("success")
    assert torch.allclose(torch.as_tensor(lm.res()), rep0, rtol=1e-5, atol=1e-6)
