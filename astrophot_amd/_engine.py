"""Host side of one component render: window geometry in float64, then the
gfx950 launches render -> psf_shift -> convolve through the C ABI.

This is Component_Model.sample (models/model_object.py:258-375) and the
per-component forward-mode Jacobian (models/_model_methods.py:260-347)
re-expressed as kernel launches:

* psf_mode "full": the working window is the sampling window padded by the
  PSF border (psf_image.py:71-93); the sub-pixel centre shift
  s = pc - round(pc) moves into a bilinearly shifted PSF and the grid is
  sampled at X = P (i - round(pc)) (model_object.py:320-325).  Centre
  derivatives come from d(shifted PSF)/ds, every other column from the render
  channels convolved with the shifted PSF.
* psf_mode "none": X = P (i - pc); every column, the centre included, comes
  from the render channels; a 1x1 unit kernel crops and accumulates.

The integration threshold reference is the Sersic total flux over the
working-grid pixel count (sersic_model.py:78-89) or the mean initial sample
(_model_methods.py:151-152).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import AP_config
from . import _lib as L
from .errors import SpecificationConflict

PROFILE_CODE = {"sersic": L.APX_SERSIC, "exponential": L.APX_EXPONENTIAL, "gaussian": L.APX_GAUSSIAN,
                "moffat": L.APX_MOFFAT, "flat_sky": L.APX_FLAT_SKY}
# profile parameter name -> slot
PROFILE_SLOTS = {
    "sersic": {"n": L.APX_P_A, "Re": L.APX_P_B, "Ie": L.APX_P_C},
    "exponential": {"Re": L.APX_P_A, "Ie": L.APX_P_B},
    "gaussian": {"sigma": L.APX_P_A, "flux": L.APX_P_B},
    "moffat": {"n": L.APX_P_A, "Rd": L.APX_P_B, "I0": L.APX_P_C},
    "flat_sky": {"F": L.APX_P_A},
}
GEOM_SLOTS = {"q": L.APX_P_Q, "PA": L.APX_P_PA}


def gauss_legendre(n: int):
    """Abscissae and weights on [-1, 1] (the table scipy.special.roots_legendre
    gives the reference, utils/operations.py:105-120)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def sersic_b(n):
    """utils/conversions/functions.py:7-22."""
    return (2 * n - 1 / 3 + 4 / (405 * n) + 46 / (25515 * n**2) + 131 / (1148175 * n**3)
            - 2194697 / (30690717750 * n**4))


def sersic_total_flux(Ie_lin, n, Re, q):
    """utils/conversions/functions.py:168-190."""
    b = sersic_b(n)
    return 2 * math.pi * Ie_lin * Re**2 * q * n * math.exp(b) * b ** (-2 * n) * math.exp(math.lgamma(2 * n))


@dataclass
class RenderRequest:
    """What one component render has to produce."""

    window: object                  # sampling (working) window before padding
    image: Optional[object] = None  # Model_Image to accumulate the value into
    jac: Optional[object] = None    # Jacobian_Image to accumulate columns into
    slot_cols: Optional[Dict[int, int]] = None  # slot -> Jacobian column
    dvdr: Optional[Dict[int, float]] = None     # slot -> d value / d rep


_UNIT_KERNEL = {}


def _unit_kernel(device):
    k = _UNIT_KERNEL.get(device)
    if k is None:
        k = torch.ones((1, 1, 1), dtype=torch.float32, device=device)
        _UNIT_KERNEL[device] = k
    return k


def _box(dst_window, src_window):
    """Rows/columns of ``src`` covering ``dst`` and where they land in ``dst``
    (image_object.py:450-457 via window_object.py:241-269)."""
    dr, dc = src_window.get_other_indices(dst_window)  # into dst
    sr, sc = dst_window.get_other_indices(src_window)  # into src
    return (dr.start, dc.start), (sr.start, sc.start), (sr.stop - sr.start, sc.stop - sc.start)


def render(model, values: Dict[str, torch.Tensor], req: RenderRequest):
    """Render ``model`` with parameter ``values`` (host float64) on
    ``req.window``; accumulate into ``req.image`` / ``req.jac``."""
    lib = L.load()
    kind = model._profile
    psf_full = "full" in model.psf_mode and kind != "flat_sky"
    window = req.window.copy()
    center = values["center"]
    P = window.pixelscale
    cp = L.Component()
    cp.profile = PROFILE_CODE[kind]
    # sampling mode (_model_methods.py:82-148)
    sm = model.sampling_mode
    if sm == "midpoint":
        cp.sampling = L.APX_SAMPLE_MIDPOINT
    elif sm.startswith("quad"):
        cp.sampling = L.APX_SAMPLE_QUAD
        cp.sample_quad = int(sm[sm.find(":") + 1:])
        if not 1 <= cp.sample_quad <= L.APX_MAX_QUAD_LEVEL:
            raise SpecificationConflict(f"sampling quad level must be 1..{L.APX_MAX_QUAD_LEVEL}")
        x, w = gauss_legendre(cp.sample_quad)
        for k in range(cp.sample_quad):
            cp.sq_node[k], cp.sq_weight[k] = x[k], w[k]
    else:
        raise SpecificationConflict(
            f"{model.name} has unknown sampling mode: {sm}. Should be one of: midpoint, quad:level")
    im = model.integrate_mode if kind != "flat_sky" else "none"
    if im not in ("none", "threshold"):
        raise SpecificationConflict(
            f"{model.name} has unknown integration mode: {im}. Should be one of: none, threshold")
    cp.integrate = 1 if im == "threshold" else 0
    if cp.integrate:
        cp.quad_level = int(model.integrate_quad_level)
        cp.gridding = int(model.integrate_gridding)
        cp.max_depth = int(model.integrate_max_depth)
        x, w = gauss_legendre(cp.quad_level)
        for k in range(cp.quad_level):
            cp.gl_node[k], cp.gl_weight[k] = x[k], w[k]

    shift = None
    if psf_full:
        psf = model.psf
        if psf is None:
            raise SpecificationConflict(f"{model.name} has psf_mode 'full' but no PSF")
        upscale = int(torch.round(window.pixel_length / psf.pixel_length))
        if upscale != 1:
            raise SpecificationConflict("PSF pixelscale must equal the image pixelscale (upscale 1)")
        bx, by = psf.psf_border_int
        window.pad_pixel((bx, by))
        pc = window.plane_to_pixel(center)
        if model.psf_subpixel_shift == "none":
            cij = torch.round(pc)
            frac = P @ (cij - pc)
        elif model.psf_subpixel_shift == "bilinear":
            cij = torch.round(pc)
            shift = pc - cij
            frac = torch.zeros(2, dtype=torch.float64)
        else:
            raise SpecificationConflict(f"unrecognized subpixel shift method: {model.psf_subpixel_shift}")
    else:
        bx = by = 0
        pc = window.plane_to_pixel(center)
        cij = torch.round(pc)
        frac = P @ (cij - pc)
    nx, ny = window.nx, window.ny
    cp.nx, cp.ny = nx, ny
    cp.ci, cp.cj = int(cij[0]), int(cij[1])
    for k in range(4):
        cp.P[k] = float(P.reshape(-1)[k])
    cp.frac[0], cp.frac[1] = float(frac[0]), float(frac[1])
    cp.pixel_area = float(window.pixel_area)
    cp.softening = float(model.softening)
    cp.tolerance = float(model.sampling_tolerance)
    if kind == "flat_sky":
        cp.q, cp.pa = 1.0, 0.0
    else:
        cp.q, cp.pa = float(values["q"]), float(values["PA"])
    for name, slot in PROFILE_SLOTS[kind].items():
        cp.prm[slot - L.APX_P_A] = float(values[name])
    if kind == "sersic":
        cp.mean_reference = 0
        cp.reference = sersic_total_flux(10 ** float(values["Ie"]), float(values["n"]),
                                         float(values["Re"]), float(values["q"])) / (nx * ny)
    else:
        cp.mean_reference = 1

    # Jacobian channels
    slot_cols = req.slot_cols or {}
    dvdr = req.dvdr or {}
    chan_slots = [s for s in sorted(slot_cols)
                  if not (psf_full and shift is not None and s in (L.APX_P_CX, L.APX_P_CY))]
    cp.nchan = 1 + (len(chan_slots) if req.jac is not None else 0)
    if req.jac is not None:
        for c, s in enumerate(chan_slots):
            cp.chan_slot[c + 1] = s
    for s, v in dvdr.items():
        cp.dvdr[s] = float(v)

    dev = AP_config.ap_device
    planes = torch.empty((cp.nchan, ny, nx), dtype=torch.float32, device=dev)
    wsb = lib.apx_render_workspace_bytes(nx * ny)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    st = L.stream()
    L.check(lib.apx_render_component(ctypes.byref(cp), L.ptr(planes), L.ptr(ws), wsb, st))

    # shifted PSF (or unit kernel)
    if psf_full:
        pdata = model.psf.data.to(torch.float32).contiguous()
        ph, pw = int(pdata.shape[0]), int(pdata.shape[1])
        if shift is not None:
            kernels = torch.empty((3, ph + 2, pw + 2), dtype=torch.float32, device=dev)
            pinv = (ctypes.c_double * 4)(*[float(v) for v in window.pixelscale_inv.reshape(-1)])
            L.check(lib.apx_psf_shift(L.ptr(pdata), ph, pw, float(shift[0]), float(shift[1]), pinv,
                                      float(dvdr.get(L.APX_P_CX, 1.0)), float(dvdr.get(L.APX_P_CY, 1.0)),
                                      L.ptr(kernels), st))
        else:
            kernels = torch.flip(pdata / pdata.sum(), dims=(0, 1)).reshape(1, ph, pw).contiguous()
        kh, kw = int(kernels.shape[1]), int(kernels.shape[2])
    else:
        kernels = _unit_kernel(dev)
        kh = kw = 1

    # crop back to the sampling window, then window-aware accumulation
    sample_window = req.window
    job = L.ConvJob()
    job.nx, job.ny, job.nchan, job.kh, job.kw = nx, ny, cp.nchan, kh, kw
    for c in range(L.APX_MAX_CHAN):
        job.chan_col[c] = -1
    job.dk_col[0] = job.dk_col[1] = -1
    if req.image is not None:
        (dy, dx), (sy, sx), (h, w) = _box(req.image.window, sample_window)
        j1 = _job_copy(job)
        j1.nchan = 1
        j1.x0, j1.y0, j1.ow, j1.oh = bx + sx, by + sy, w, h
        img = req.image.data
        if img.dtype != torch.float32 or not img.is_contiguous():
            raise SpecificationConflict("model image must be contiguous float32")
        j1.img, j1.img_row, j1.img_x0, j1.img_y0 = L.ptr(img), img.shape[1], dx, dy
        j1.img_w, j1.img_h = img.shape[1], img.shape[0]
        if h > 0 and w > 0:
            L.check(lib.apx_convolve(ctypes.byref(j1), L.ptr(planes), L.ptr(kernels), st))
    if req.jac is not None and slot_cols:
        (dy, dx), (sy, sx), (h, w) = _box(req.jac.window, sample_window)
        j2 = _job_copy(job)
        j2.x0, j2.y0, j2.ow, j2.oh = bx + sx, by + sy, w, h
        jd = req.jac.data
        npj = int(jd.shape[2])
        if not jd.is_contiguous() or jd.dtype != torch.float32:
            raise SpecificationConflict("Jacobian must be contiguous float32")
        j2.jac, j2.jac_row, j2.jac_np, j2.jac_x0, j2.jac_y0 = L.ptr(jd), jd.shape[1] * npj, npj, dx, dy
        j2.jac_w, j2.jac_h = jd.shape[1], jd.shape[0]
        for c, s in enumerate(chan_slots):
            j2.chan_col[c + 1] = slot_cols[s]
        if psf_full and shift is not None:
            j2.dk_col[0] = slot_cols.get(L.APX_P_CX, -1)
            j2.dk_col[1] = slot_cols.get(L.APX_P_CY, -1)
        if h > 0 and w > 0:
            L.check(lib.apx_convolve(ctypes.byref(j2), L.ptr(planes), L.ptr(kernels), st))
    return ws


def _job_copy(job):
    j = L.ConvJob()
    ctypes.pointer(j)[0] = job
    return j
