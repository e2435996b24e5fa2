"""Build the same model twice: once in the oracle (float64 CPU restatement)
and once through astrophot_amd (gfx950 kernels), from one specification."""

from __future__ import annotations

import math

import numpy as np
import torch

import astrophot_amd as ap
from oracle import hotpath as H

MODEL_TYPE = {"sersic": "sersic galaxy model", "exponential": "exponential galaxy model",
              "gaussian": "gaussian galaxy model", "moffat": "moffat galaxy model",
              "flat_sky": "flat sky model"}

_count = [0]


def target(nx, ny, pixelscale=1.0, psf=None, data=None, variance=None, mask=None):
    d = np.zeros((ny, nx)) if data is None else data
    return ap.image.Target_Image(data=d, pixelscale=pixelscale,
                                 psf=None if psf is None else ap.image.PSF_Image(data=psf, pixelscale=pixelscale),
                                 variance=variance, mask=mask)


def component(kind, params, tgt, nx, ny, pixelscale=1.0, psf=None, **knobs):
    """(oracle Component, astrophot_amd model) with identical settings."""
    _count[0] += 1
    g = H.Grid(nx, ny, pixelscale)
    oc = H.Component(kind, g, dict(params), psf=None if psf is None else torch.as_tensor(psf, dtype=torch.float64),
                     **knobs)
    m = ap.models.AstroPhot_Model(name=f"{kind}{_count[0]}", model_type=MODEL_TYPE[kind], target=tgt,
                                  parameters=dict(params), **knobs)
    return oc, m


def single(kind, params, nx=40, ny=36, pixelscale=1.0, psf=None, **knobs):
    tgt = target(nx, ny, pixelscale, psf)
    oc, m = component(kind, params, tgt, nx, ny, pixelscale, psf, **knobs)
    return H.Model([oc]), m


def rel_err(a, b):
    """max |a - b| / max |b|"""
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


GAL = {
    "sersic": {"center": [20.3, 17.6], "q": 0.6, "PA": 0.7, "n": 2.5, "Re": 6.0, "Ie": 1.0},
    "exponential": {"center": [19.8, 18.4], "q": 0.5, "PA": 2.0, "Re": 7.0, "Ie": 0.5},
    "gaussian": {"center": [20.1, 17.9], "q": 0.8, "PA": 1.1, "sigma": 4.0, "flux": 2.0},
    "moffat": {"center": [19.6, 18.2], "q": 0.7, "PA": 0.3, "n": 2.0, "Rd": 3.0, "I0": 1.2},
    "flat_sky": {"F": math.log10(0.5)},
}
