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


def perturbed_start(truth: np.ndarray, seed: int, frac: float = 0.1) -> np.ndarray:
    """Every parameter times (1 + U(-frac, frac)), kept inside its limits
    (the C0 start rule)."""
    rng = np.random.default_rng(seed)
    s = truth * (1 + rng.uniform(-frac, frac, truth.shape))
    s[:, 2] = np.clip(s[:, 2], 0.02, 0.999)
    s[:, 4] = np.clip(s[:, 4], 0.37, 7.99)
    s[:, 5] = np.maximum(s[:, 5], 0.05)
    return s
