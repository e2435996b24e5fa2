"""Utility functions of the reference needed on the hot path:
fixture PSFs (utils/initialize/construct_psf.py:8-65), Sersic conversions
(utils/conversions/functions.py) and chi^2 helpers (utils/optimization.py)."""

from __future__ import annotations

import math

import numpy as np
import torch

from ._engine import sersic_b as sersic_n_to_b, sersic_total_flux

__all__ = ["gaussian_psf", "moffat_psf", "sersic_n_to_b", "sersic_Ie_to_flux", "chi_squared",
           "reduced_chi_squared"]


def _psf_grid(img_width, pixelscale, upsample):
    assert img_width % 2 == 1, "psf images should have an odd shape"
    N = img_width * upsample
    g = np.linspace(-(N - 1) * pixelscale / (2 * upsample), (N - 1) * pixelscale / (2 * upsample), N)
    return np.meshgrid(g, g)


def gaussian_psf(sigma, img_width, pixelscale, upsample=4, normalize=True):
    """construct_psf.py:8-35: 4x supersampled, block-averaged Gaussian."""
    XX, YY = _psf_grid(img_width, pixelscale, upsample)
    ZZ = np.exp(-0.5 * (XX**2 + YY**2) / sigma**2)
    ZZ = ZZ.reshape(img_width, upsample, img_width, upsample).sum(axis=(1, 3)) / (upsample**2)
    return ZZ / np.sum(ZZ) if normalize else ZZ


def moffat_psf(n, Rd, img_width, pixelscale, upsample=4, normalize=True):
    """construct_psf.py:38-65: 4x supersampled, block-averaged Moffat."""
    XX, YY = _psf_grid(img_width, pixelscale, upsample)
    ZZ = 1.0 / (1.0 + (XX**2 + YY**2) / (Rd**2)) ** n
    ZZ = ZZ.reshape(img_width, upsample, img_width, upsample).sum(axis=(1, 3)) / (upsample**2)
    return ZZ / np.sum(ZZ) if normalize else ZZ


def sersic_Ie_to_flux(Ie, n, R, q):
    """conversions/functions.py:168-190 (Ie linear)."""
    return sersic_total_flux(Ie, n, R, q)


def chi_squared(target, model, mask=None, variance=None):
    """utils/optimization.py:6-17."""
    if mask is None:
        r = (target - model) ** 2
        return torch.sum(r if variance is None else r / variance)
    keep = torch.logical_not(mask)
    r = (target[keep] - model[keep]) ** 2
    return torch.sum(r if variance is None else r / variance[keep])


def reduced_chi_squared(target, model, params, mask=None, variance=None):
    """utils/optimization.py:20-30."""
    ndf = (len(target) if mask is None else torch.sum(torch.logical_not(mask))) - params
    return chi_squared(target, model, mask, variance) / ndf
