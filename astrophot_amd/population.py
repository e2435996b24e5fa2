"""Synthetic Sersic cutout populations for batched fitting and for training the
learned initialiser.

Priors (seeded numpy PCG64):
  cutout side  s = round(exp U(ln 32, ln 256)) px (square), pixelscale 1
  centre       s/2 + U(-1, 1) per axis (plane coordinates, corner at 0)
  n ~ U(0.5, 6), Re = exp U(ln 2, ln 40) px, q ~ U(0.2, 1), PA ~ U(0, pi)
  Ie (log10 flux/px^2) ~ U(-0.5, 1.5)
  PSF: Gaussian sigma ~ U(1, 3) px in a (2 ceil(4 sigma) + 1)^2 box
       (utils/initialize/construct_psf.py:8-35)
  S/N = exp U(ln 10, ln 1000), defined as F_tot / (sigma_sky sqrt(pi Re^2 q)),
       i.e. the total flux over the sky noise inside the half-light ellipse;
       variance = sigma_sky^2 everywhere.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

from ._engine import sersic_total_flux
from .utils import gaussian_psf


@dataclass
class Population:
    truth: np.ndarray       # (n, 7) values [cx, cy, q, PA, n, Re, Ie]
    sides: np.ndarray       # (n,) cutout side
    psf_sigma: np.ndarray   # (n,)
    psfs: List[np.ndarray]
    sky_sigma: np.ndarray   # (n,)
    snr: np.ndarray         # (n,)


def sample_population(n: int, seed: int, side_range=(32, 256), re_range=(2.0, 40.0), n_range=(0.5, 6.0),
                      q_range=(0.2, 1.0), snr_range=(10.0, 1000.0), sigma_range=(1.0, 3.0),
                      ie_range=(-0.5, 1.5), make_psfs=True) -> Population:
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(math.log(side_range[0]), math.log(side_range[1]), n))).astype(int)
    cx = sides / 2 + rng.uniform(-1, 1, n)
    cy = sides / 2 + rng.uniform(-1, 1, n)
    ns = rng.uniform(*n_range, n)
    re = np.exp(rng.uniform(math.log(re_range[0]), math.log(re_range[1]), n))
    q = rng.uniform(*q_range, n)
    q = np.minimum(q, 0.999)
    pa = rng.uniform(0, math.pi, n)
    ie = rng.uniform(*ie_range, n)
    snr = np.exp(rng.uniform(math.log(snr_range[0]), math.log(snr_range[1]), n))
    sig = rng.uniform(*sigma_range, n)
    psfs, sky = [], np.empty(n)
    for k in range(n):
        if make_psfs:
            psfs.append(gaussian_psf(sig[k], 2 * math.ceil(4 * sig[k]) + 1, 1.0))
        F = sersic_total_flux(10 ** ie[k], ns[k], re[k], q[k])
        sky[k] = F / (snr[k] * math.sqrt(math.pi * re[k] ** 2 * q[k]))
    truth = np.stack([cx, cy, q, pa, ns, re, ie], axis=1)
    return Population(truth, sides, sig, psfs, sky, snr)


def perturbed_start(truth: np.ndarray, seed: int, frac: float = 0.1, center_px=None) -> np.ndarray:
    """Every parameter times (1 + U(-frac, frac)), kept inside its limits
    (the C0 start rule).  With ``center_px`` the centres move by
    U(-center_px, center_px) pixels instead, which keeps sources far from the
    plane origin inside their windows."""
    rng = np.random.default_rng(seed)
    s = truth * (1 + rng.uniform(-frac, frac, truth.shape))
    if center_px is not None:
        s[:, :2] = truth[:, :2] + rng.uniform(-center_px, center_px, (truth.shape[0], 2))
    s[:, 2] = np.clip(s[:, 2], 0.02, 0.999)
    s[:, 4] = np.clip(s[:, 4], 0.37, 7.99)
    s[:, 5] = np.maximum(s[:, 5], 0.05)
    return s


@dataclass
class GroupScene:
    """One image holding many Sersic sources on their own windows plus a flat
    sky, all seen through one Moffat PSF (the joint-group configuration)."""

    side: int
    truth: np.ndarray        # (n, 7) values [cx, cy, q, PA, n, Re, Ie], image corner at the plane origin
    windows: np.ndarray      # (n, 4) pixel boxes x0, x1, y0, y1 (half-open), inside the image
    sky_F: float             # log10 sky brightness per unit area
    sky_sigma: float         # noise sigma per pixel
    psf: np.ndarray          # (psf_side, psf_side) unit sum
    beta: float
    Rd: float
    snr: np.ndarray          # (n,)


def group_scene(n: int, seed: int, side: int = 2048, psf_side: int = 65, fwhm: float = 3.0,
                sky_sigma: float = 1.0, sky_F: float = 0.0, re_range=(2.0, 40.0), n_range=(0.5, 6.0),
                q_range=(0.2, 1.0), snr_range=(10.0, 1000.0), min_half=33) -> GroupScene:
    """n sources at U(0, side)^2 with the batched-fit priors; each source's
    S/N = exp U(ln 10, ln 1000) against sky_sigma (definition as in
    sample_population) sets its Ie; windows of half side max(5 Re, min_half)
    px clipped to the image; Moffat beta ~ U(2.5, 4.5) with
    Rd = fwhm / (2 sqrt(2^(1/beta) - 1)), sampled by moffat_psf on an odd box."""
    from .utils import moffat_psf
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0, side, n)
    cy = rng.uniform(0, side, n)
    ns = rng.uniform(*n_range, n)
    re = np.exp(rng.uniform(math.log(re_range[0]), math.log(re_range[1]), n))
    q = np.minimum(rng.uniform(*q_range, n), 0.999)
    pa = rng.uniform(0, math.pi, n)
    snr = np.exp(rng.uniform(math.log(snr_range[0]), math.log(snr_range[1]), n))
    ie = np.empty(n)
    for k in range(n):
        F = snr[k] * sky_sigma * math.sqrt(math.pi * re[k] ** 2 * q[k])
        ie[k] = math.log10(F / sersic_total_flux(1.0, ns[k], re[k], q[k]))
    half = np.maximum(5 * re, min_half)
    wins = np.stack([np.clip(np.floor(cx - half), 0, side), np.clip(np.ceil(cx + half), 0, side),
                     np.clip(np.floor(cy - half), 0, side), np.clip(np.ceil(cy + half), 0, side)], axis=1)
    beta = float(rng.uniform(2.5, 4.5))
    Rd = fwhm / (2 * math.sqrt(2 ** (1 / beta) - 1))
    if psf_side % 2 == 0:
        psf_side += 1   # PSF images must have odd sides (psf_image.py:55-61)
    psf = moffat_psf(beta, Rd, psf_side, 1.0)
    truth = np.stack([cx, cy, q, pa, ns, re, ie], axis=1)
    return GroupScene(side, truth, wins.astype(np.int64), sky_F, sky_sigma, psf, beta, Rd, snr)
