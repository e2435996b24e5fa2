"""Windows and images with the reference's API (image/window_object.py,
image/wcs.py PPCS, image/image_object.py, image/target_image.py,
image/psf_image.py, image/model_image.py, image/jacobian_image.py).

Window geometry is host-side float64, exactly as the reference computes it;
pixel data live on the GPU in float32 (``AP_config.ap_image_dtype``).  World
(RA/Dec) coordinates and FITS I/O are not part of this package.
"""

from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np
import torch

from . import AP_config
from .errors import InvalidData, SpecificationConflict

__all__ = ["Window", "Image", "Target_Image", "Model_Image", "PSF_Image", "Jacobian_Image"]

F64 = torch.float64


def _h(v):
    """host float64 tensor"""
    return torch.as_tensor(v, dtype=F64, device="cpu")


def _dev(v, dtype=None):
    dtype = AP_config.ap_image_dtype if dtype is None else dtype
    if isinstance(v, torch.Tensor):
        return v.detach().to(device=AP_config.ap_device, dtype=dtype)
    return torch.as_tensor(np.asarray(v), dtype=dtype, device=AP_config.ap_device)


def _pair(v):
    """two host floats from a number pair (tensor, ndarray, sequence)"""
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().reshape(-1).tolist()
    elif isinstance(v, np.ndarray):
        v = v.reshape(-1).tolist()
    return (float(v[0]), float(v[1]))


def _round_half_even(x):
    """torch.round / numpy round semantics (ties to even), on a float."""
    return float(round(x))


class Window:
    """A window on the tangent plane: ``pixel_shape`` (nx, ny) pixels of
    parallelogram ``pixelscale``; pixel ``reference_imageij`` is pinned to plane
    point ``reference_imagexy`` (image/wcs.py:561-584).

    The geometry is kept in Python floats (IEEE double, the arithmetic the
    reference does in float64 torch) so the per-call host cost stays small; the
    properties return float64 tensors as the reference's do.
    """

    def __init__(self, *, pixel_shape=None, origin=None, center=None, pixelscale=None,
                 reference_imageij=(-0.5, -0.5), reference_imagexy=(0.0, 0.0), state=None, **kwargs):
        # world-coordinate fields (image/wcs.py:36-43): carried through save/load;
        # every computation here stays on the tangent plane
        self.projection = kwargs.get("projection", "gnomonic")
        self.reference_radec = _pair(kwargs.get("reference_radec", (0.0, 0.0)))
        self.reference_planexy = _pair(kwargs.get("reference_planexy", (0.0, 0.0)))
        if state is not None:
            self.set_state(state)
            return
        self.pixelscale = 1.0 if pixelscale is None else pixelscale
        self.pixel_shape = pixel_shape
        if origin is not None and center is not None:
            raise SpecificationConflict("Please provide only one reference position for the window")
        if origin is not None:  # window_object.py:143-150
            self._rij = (-0.5, -0.5)
            self._rxy = _pair(origin)
        elif center is not None:  # window_object.py:151-160
            self._rij = (self._nx / 2 - 0.5, self._ny / 2 - 0.5)
            self._rxy = _pair(center)
        else:
            self._rij = _pair(reference_imageij)
            self._rxy = _pair(reference_imagexy)

    # -- PPCS (wcs.py:492-639) ----------------------------------------------
    def _set_P(self, P):
        self._P = P
        a, b, c, d = P
        det = a * d - b * c
        self._area = abs(det)
        self._Pi = (d / det, -b / det, -c / det, a / det)

    @property
    def pixelscale(self):
        return torch.tensor(self._P, dtype=F64).reshape(2, 2)

    @pixelscale.setter
    def pixelscale(self, pix):
        p = _h(pix).reshape(-1)
        if p.numel() == 1:
            v = float(p[0])
            self._set_P((v, 0.0, 0.0, v))
        else:
            self._set_P(tuple(float(x) for x in p.tolist()))

    @property
    def pixel_area(self):
        return torch.tensor(self._area, dtype=F64)

    @property
    def pixel_length(self):
        return torch.tensor(math.sqrt(self._area), dtype=F64)

    @property
    def pixelscale_inv(self):
        return torch.tensor(self._Pi, dtype=F64).reshape(2, 2)

    @property
    def reference_imageij(self):
        return torch.tensor(self._rij, dtype=F64)

    @reference_imageij.setter
    def reference_imageij(self, v):
        self._rij = _pair(v)

    @property
    def reference_imagexy(self):
        return torch.tensor(self._rxy, dtype=F64)

    @reference_imagexy.setter
    def reference_imagexy(self, v):
        self._rxy = _pair(v)

    @property
    def pixel_shape(self):
        return torch.tensor([self._nx, self._ny], dtype=torch.int32)

    @pixel_shape.setter
    def pixel_shape(self, shape):
        if shape is None:
            self._nx = self._ny = None
            return
        x, y = _pair(shape)
        self._nx, self._ny = int(_round_half_even(x)), int(_round_half_even(y))

    # float-level transforms
    def _p2x(self, i, j):
        P, r, o = self._P, self._rij, self._rxy
        di, dj = i - r[0], j - r[1]
        return (P[0] * di + P[1] * dj + o[0], P[2] * di + P[3] * dj + o[1])

    def _x2p(self, x, y):
        Pi, r, o = self._Pi, self._rij, self._rxy
        dx, dy = x - o[0], y - o[1]
        return (Pi[0] * dx + Pi[1] * dy + r[0], Pi[2] * dx + Pi[3] * dy + r[1])

    def pixel_to_plane(self, ij):
        return torch.tensor(self._p2x(*_pair(ij)), dtype=F64)

    def plane_to_pixel(self, xy):
        return torch.tensor(self._x2p(*_pair(xy)), dtype=F64)

    def pixel_to_plane_delta(self, d):
        return self.pixelscale @ _h(d)

    def plane_to_pixel_delta(self, d):
        return self.pixelscale_inv @ _h(d)

    # -- window properties (window_object.py:165-239) ---------------------
    @property
    def nx(self):
        return self._nx

    @property
    def ny(self):
        return self._ny

    @property
    def size(self):
        return self._nx * self._ny

    def _end(self):
        P = self._P
        return (P[0] * self._nx + P[1] * self._ny, P[2] * self._nx + P[3] * self._ny)

    def _origin(self):
        return self._p2x(-0.5, -0.5)

    @property
    def end(self):
        return torch.tensor(self._end(), dtype=F64)

    @property
    def origin(self):
        return torch.tensor(self._origin(), dtype=F64)

    @property
    def center(self):
        o, e = self._origin(), self._end()
        return torch.tensor((o[0] + e[0] / 2, o[1] + e[1] / 2), dtype=F64)

    @property
    def shape(self):
        P = self._P
        return torch.tensor((math.hypot(P[0] * self._nx, P[2] * self._nx),
                             math.hypot(P[1] * self._ny, P[3] * self._ny)), dtype=F64)

    def copy(self, **kwargs):
        w = Window.__new__(Window)
        w._P, w._Pi, w._area = self._P, self._Pi, self._area
        w._nx, w._ny = self._nx, self._ny
        w._rij, w._rxy = self._rij, self._rxy
        w.projection, w.reference_radec, w.reference_planexy = (self.projection, self.reference_radec,
                                                                self.reference_planexy)
        if "pixelscale" in kwargs:
            w.pixelscale = kwargs["pixelscale"]
        if "pixel_shape" in kwargs:
            w.pixel_shape = kwargs["pixel_shape"]
        if "origin" in kwargs:
            w._rij = (-0.5, -0.5)
            w._rxy = _pair(kwargs["origin"])
        if "reference_imageij" in kwargs:
            w._rij = _pair(kwargs["reference_imageij"])
        if "reference_imagexy" in kwargs:
            w._rxy = _pair(kwargs["reference_imagexy"])
        return w

    def rescale_pixel(self, scale):
        """window_object.py:233-239."""
        scale = float(scale)
        return self.copy(pixelscale=self.pixelscale * scale,
                         pixel_shape=(self._nx // scale, self._ny // scale),
                         reference_imageij=((self._rij[0] + 0.5) / scale - 0.5, (self._rij[1] + 0.5) / scale - 0.5))

    def pixel_shift(self, shift):
        """window_object.py:284-290."""
        s = _pair(shift)
        self._rij = (self._rij[0] - s[0], self._rij[1] - s[1])
        return self

    def shift(self, shift):
        s = _pair(shift)
        self._rxy = (self._rxy[0] + s[0], self._rxy[1] + s[1])
        return self

    @staticmethod
    def _pix(pixels):
        p = [float(v) for v in (pixels.tolist() if isinstance(pixels, torch.Tensor) else pixels)]
        if len(p) == 1:
            p = p * 2
        if len(p) not in (2, 4):
            raise ValueError(f"Unrecognized pixel pad/crop format: {pixels}")
        return p

    def pad_pixel(self, pixels):
        """window_object.py:367-388."""
        p = self._pix(pixels)
        lo = p[:2]
        hi = p[:2] if len(p) == 2 else p[2:]
        self.pixel_shape = (self._nx + lo[0] + hi[0], self._ny + lo[1] + hi[1])
        self._rij = (self._rij[0] + lo[0], self._rij[1] + lo[1])
        return self

    def crop_pixel(self, pixels):
        """window_object.py:335-356."""
        p = self._pix(pixels)
        lo = p[:2]
        hi = p[:2] if len(p) == 2 else p[2:]
        self.pixel_shape = (self._nx - lo[0] - hi[0], self._ny - lo[1] - hi[1])
        self._rij = (self._rij[0] - lo[0], self._rij[1] - lo[1])
        return self

    def crop_to_pixel(self, pixels):
        """window_object.py:358-365; pixels = [[xmin, xmax], [ymin, ymax]]."""
        p = [[float(a), float(b)] for a, b in (pixels.tolist() if isinstance(pixels, torch.Tensor) else pixels)]
        self._rij = (self._rij[0] - p[0][0], self._rij[1] - p[1][0])
        self.pixel_shape = (p[0][1] - p[0][0], p[1][1] - p[1][0])
        return self

    @staticmethod
    def _get_indices(ref_window, obj_window):
        """window_object.py:241-253: (row slice, column slice) of ref's pixels
        covered by obj."""
        ox, oy = obj_window._origin()
        ex, ey = obj_window._end()
        a = ref_window._x2p(ox, oy)
        b = ref_window._x2p(ox + ex, oy + ey)
        x0 = max(0, int(_round_half_even(a[0] + 0.5)))
        y0 = max(0, int(_round_half_even(a[1] + 0.5)))
        x1 = min(ref_window._nx, int(_round_half_even(b[0] + 0.5)))
        y1 = min(ref_window._ny, int(_round_half_even(b[1] + 0.5)))
        return slice(y0, y1), slice(x0, x1)

    def get_self_indices(self, obj):
        return self._get_indices(self, obj if isinstance(obj, Window) else obj.window)

    def get_other_indices(self, obj):
        return self._get_indices(obj if isinstance(obj, Window) else obj.window, self)

    def _box(self, other, op):
        ox, oy = other._origin()
        ex, ey = other._end()
        o = self._x2p(ox, oy)
        e = self._x2p(ox + ex, oy + ey)
        if op == "and":  # window_object.py:511-523
            lo = (max(-0.5, o[0]), max(-0.5, o[1]))
            hi = (min(self._nx - 0.5, e[0]), min(self._ny - 0.5, e[1]))
        else:  # window_object.py:483-495
            lo = (min(-0.5, o[0]), min(-0.5, o[1]))
            hi = (max(float(self._nx), e[0]), max(float(self._ny), e[1]))
        return self.copy(origin=self._p2x(*lo), pixel_shape=(hi[0] - lo[0], hi[1] - lo[1]))

    def __and__(self, other):
        return self._box(other, "and")

    def __or__(self, other):
        return self._box(other, "or")

    def overlap_frac(self, other):
        overlap = self & other
        oa = torch.prod(overlap.shape)
        return oa / (torch.prod(self.shape) + torch.prod(other.shape) - oa)

    def get_coordinate_meshgrid(self, device="cpu"):
        """window_object.py:390-405: plane coordinates of every pixel centre (2, ny, nx)."""
        i = torch.arange(self._nx, dtype=F64)
        j = torch.arange(self._ny, dtype=F64)
        mi, mj = torch.meshgrid(i, j, indexing="xy")
        c = self.pixelscale @ (torch.stack((mi.reshape(-1), mj.reshape(-1))) - self.reference_imageij.view(2, 1))
        return (c + self.reference_imagexy.view(2, 1)).reshape(2, self._ny, self._nx).to(device)

    def __eq__(self, other):
        return (self._nx == other._nx and self._ny == other._ny and self._P == other._P
                and self._p2x(0.0, 0.0) == other._p2x(0.0, 0.0))

    def get_state(self):
        """The keys of image/wcs.py:361-370, 656-661 and window_object.py:311-314."""
        return {"projection": self.projection, "reference_radec": list(self.reference_radec),
                "reference_planexy": list(self.reference_planexy), "pixelscale": self.pixelscale.tolist(),
                "reference_imageij": list(self._rij), "reference_imagexy": list(self._rxy),
                "pixel_shape": [self._nx, self._ny]}

    def set_state(self, state):
        self.projection = state.get("projection", "gnomonic")
        self.reference_radec = _pair(state.get("reference_radec", (0.0, 0.0)))
        self.reference_planexy = _pair(state.get("reference_planexy", (0.0, 0.0)))
        self.pixelscale = state.get("pixelscale", 1.0)
        self.pixel_shape = state["pixel_shape"]
        self._rij = _pair(state.get("reference_imageij", (-0.5, -0.5)))
        self._rxy = _pair(state.get("reference_imagexy", (0.0, 0.0)))

    def __repr__(self):
        return (f"window pixel_shape: {[self._nx, self._ny]}, origin: {list(self._origin())}, "
                f"pixelscale: {list(self._P)}")


class Image:
    """Pixel data on the GPU plus a window (image/image_object.py:17-479)."""

    north = math.pi / 2  # image/image_header.py:36

    def __init__(self, *, data=None, pixelscale=None, window: Optional[Window] = None, zeropoint=None,
                 origin=None, center=None, identity=None, state=None, **kwargs):
        if state is not None:  # image_object.py:397-399
            header = state.get("header", {})
            window = Window(state=header["window"])
            zeropoint = header.get("zeropoint", zeropoint)
            data = np.asarray(state["data"], dtype=np.float64)
        if data is None and window is None:
            raise InvalidData("Image must have either data or a window to construct itself.")
        if window is None:
            shape = tuple(np.shape(data)) if not isinstance(data, torch.Tensor) else tuple(data.shape)
            window = Window(pixel_shape=(shape[1], shape[0]), pixelscale=pixelscale, origin=origin,
                            center=center)
        self.window = window
        self.zeropoint = None if zeropoint is None else _h(zeropoint)
        self.identity = str(id(self)) if identity is None else identity
        if data is None:
            self._data = torch.zeros((window.ny, window.nx), dtype=AP_config.ap_image_dtype,
                                     device=AP_config.ap_device)
        else:
            self.set_data(data, require_shape=False)

    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, d):
        self.set_data(d)

    def set_data(self, data, require_shape=True):
        d = _dev(data)
        if require_shape and getattr(self, "_data", None) is not None and d.shape != self._data.shape:
            raise SpecificationConflict(
                f"Attempting to change image data with tensor that has a different shape! "
                f"({tuple(d.shape)} vs {tuple(self._data.shape)})")
        self._data = d

    @property
    def pixelscale(self):
        return self.window.pixelscale

    @property
    def pixel_area(self):
        return self.window.pixel_area

    @property
    def pixel_length(self):
        return self.window.pixel_length

    @property
    def origin(self):
        return self.window.origin

    @property
    def center(self):
        return self.window.center

    @property
    def shape(self):
        return self.window.shape

    def plane_to_pixel(self, xy):
        return self.window.plane_to_pixel(xy)

    def pixel_to_plane(self, ij):
        return self.window.pixel_to_plane(ij)

    def get_coordinate_meshgrid(self):
        return self.window.get_coordinate_meshgrid()

    def _new(self, data, window):
        return self.__class__(data=data, window=window, zeropoint=self.zeropoint, identity=self.identity)

    def get_state(self):
        """image_object.py:390-395 with the header keys of image_header.py:252-263."""
        header = {"window": self.window.get_state()}
        if self.zeropoint is not None:
            header["zeropoint"] = float(self.zeropoint)
        return {"type": self.__class__.__name__, "data": self.data.detach().cpu().tolist(), "header": header}

    def copy(self):
        return self._new(self.data.clone(), self.window.copy())

    def blank_copy(self):
        return self._new(torch.zeros_like(self.data), self.window.copy())

    def get_window(self, window: Window):
        """image_object.py:294-300 (sub-image covering ``window``)."""
        return self._new(self.data[self.window.get_self_indices(window)], self.window & window)

    def __getitem__(self, key):
        if isinstance(key, Window):
            return self.get_window(key)
        if isinstance(key, Image):
            return self.get_window(key.window)
        raise ValueError("Unrecognized Image getitem request!")

    def flatten(self, attribute="data"):
        return getattr(self, attribute).reshape(-1)

    def crop(self, pixels):
        """image_object.py:307-334."""
        p = [int(v) for v in pixels]
        if len(p) == 1:
            self.set_data(self.data[p[0]:self.data.shape[0] - p[0], p[0]:self.data.shape[1] - p[0]],
                          require_shape=False)
        elif len(p) == 2:
            self.set_data(self.data[p[1]:self.data.shape[0] - p[1], p[0]:self.data.shape[1] - p[0]],
                          require_shape=False)
        elif len(p) == 4:
            self.set_data(self.data[p[2]:self.data.shape[0] - p[3], p[0]:self.data.shape[1] - p[1]],
                          require_shape=False)
        self.window.crop_pixel(p)
        return self

    def reduce(self, scale: int):
        """image_object.py:348-376: sum scale x scale blocks."""
        if not isinstance(scale, int):
            raise SpecificationConflict(f"Reduce scale must be an integer! not {type(scale)}")
        if scale == 1:
            return self
        MS, NS = self.data.shape[0] // scale, self.data.shape[1] // scale
        d = self.data[: MS * scale, : NS * scale].reshape(MS, scale, NS, scale).sum(dim=(1, 3))
        return self._new(d, self.window.rescale_pixel(scale))

    def __iadd__(self, other):
        """image_object.py:450-457."""
        if isinstance(other, Image):
            self.data[other.window.get_other_indices(self)] += other.data[self.window.get_other_indices(other)]
        else:
            self.data += other
        return self

    def __isub__(self, other):
        if isinstance(other, Image):
            self.data[other.window.get_other_indices(self)] -= other.data[self.window.get_other_indices(other)]
        else:
            self.data -= other
        return self

    def __repr__(self):
        return f"{self.__class__.__name__} {tuple(self.data.shape)} {self.window!r}"


class Model_Image(Image):
    """image/model_image.py:13-91."""

    def clear_image(self):
        self.data.zero_()


class PSF_Image(Image):
    """image/psf_image.py: odd-shaped PSF pinned at its centre pixel."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.window.reference_imageij = np.flip(np.array(self.data.shape, dtype=float) - 1.0).copy() / 2
        self.window.reference_imagexy = (0.0, 0.0)

    def set_data(self, data, require_shape=True):
        super().set_data(data, require_shape)
        if any(s % 2 != 1 for s in self.data.shape):
            raise SpecificationConflict(f"psf must have odd shape, not {tuple(self.data.shape)}")
        if torch.any(self.data < 0):
            AP_config.ap_logger.warning("psf data should be non-negative")

    def normalize(self):
        self.data /= torch.sum(self.data)

    def jacobian_image(self, parameters=None, data=None):
        """psf_image.py:106-130."""
        return Jacobian_Image(parameters=[] if parameters is None else parameters, data=data,
                              window=self.window.copy())

    def model_image(self, window=None):
        """psf_image.py:132-141: a blank image on this PSF's window."""
        w = self.window.copy() if window is None else window.copy()
        return Model_Image(window=w)

    @property
    def psf_border_int(self):
        """psf_image.py:71-93: ceil((1 + shape)/2) in (x, y) order."""
        return (int(math.ceil((1 + self.data.shape[1]) / 2)), int(math.ceil((1 + self.data.shape[0]) / 2)))

    @property
    def mask(self):
        return torch.zeros_like(self.data, dtype=torch.bool)


class Jacobian_Image(Image):
    """image/jacobian_image.py: (ny, nx, Np) float32, pixel-major; parameter
    identities label the last axis."""

    def __init__(self, *, parameters, data=None, window: Window, **kwargs):
        self.parameters = list(parameters)
        if data is None:
            data = torch.zeros((window.ny, window.nx, len(self.parameters)),
                               dtype=AP_config.ap_image_dtype, device=AP_config.ap_device)
        super().__init__(data=data, window=window, **kwargs)

    def set_data(self, data, require_shape=True):
        self._data = _dev(data)

    def flatten(self, attribute="data"):
        """jacobian_image.py:36-37: (ny*nx, Np)."""
        d = getattr(self, attribute)
        return d.reshape(-1, d.shape[-1])

    def _new(self, data, window):
        return Jacobian_Image(parameters=self.parameters, data=data, window=window)

    def __iadd__(self, other):
        """jacobian_image.py:79-118: add columns matched by parameter identity."""
        if not isinstance(other, Jacobian_Image):
            raise ValueError("Jacobian images can only add Jacobian images")
        si = self.window.get_other_indices(other)
        oi = other.window.get_other_indices(self)
        for k, pid in enumerate(other.parameters):
            if pid in self.parameters:
                j = self.parameters.index(pid)
                self.data[oi + (j,)] += other.data[si + (k,)]
            else:
                self.parameters.append(pid)
                col = torch.zeros(self.data.shape[:2] + (1,), dtype=self.data.dtype, device=self.data.device)
                col[oi + (0,)] = other.data[si + (k,)]
                self._data = torch.cat((self.data, col), dim=-1)
        return self


class Target_Image(Image):
    """image/target_image.py:17-532: data, variance/weight, mask and PSF."""

    def __init__(self, *args, mask=None, variance=None, weight=None, psf=None, psf_upscale=1, **kwargs):
        super().__init__(*args, **kwargs)
        self._weight = None
        self._mask = None
        self._psf = None
        if weight is not None:
            self.weight = weight
        elif variance is not None:
            self.variance = variance
        if mask is not None:
            self.mask = mask
        if psf is not None:
            self.set_psf(psf, psf_upscale)

    # variance / weight (target_image.py:117-184, 271-281)
    @property
    def has_variance(self):
        return self._weight is not None

    @property
    def weight(self):
        if self._weight is None:
            return torch.ones_like(self.data)
        return self._weight

    @weight.setter
    def weight(self, w):
        self._weight = None if w is None else _dev(w)

    @property
    def variance(self):
        if self._weight is None:
            return torch.ones_like(self.data)
        return 1.0 / self._weight

    @variance.setter
    def variance(self, v):
        self._weight = None if v is None else 1.0 / _dev(v, torch.float64).to(AP_config.ap_image_dtype)

    @property
    def has_mask(self):
        return self._mask is not None

    @property
    def mask(self):
        if self._mask is None:
            return torch.zeros_like(self.data, dtype=torch.bool)
        return self._mask

    @mask.setter
    def mask(self, m):
        self._mask = None if m is None else _dev(m, torch.bool)

    @property
    def has_psf(self):
        return self._psf is not None

    @property
    def psf(self):
        return self._psf

    @psf.setter
    def psf(self, psf):
        self.set_psf(psf)

    def set_psf(self, psf, psf_upscale=1):
        """target_image.py:303-325: a PSF given as data gets the pixelscale
        of this image divided by ``psf_upscale``; models convolved with it are
        sampled on a grid that much finer and summed back in blocks."""
        if psf is None or isinstance(psf, PSF_Image):
            self._psf = psf
        else:
            self._psf = PSF_Image(data=psf, pixelscale=self.pixelscale / psf_upscale)

    def _new(self, data, window):
        return Target_Image(data=data, window=window, zeropoint=self.zeropoint, identity=self.identity)

    def get_window(self, window):
        idx = self.window.get_self_indices(window)
        new = Target_Image(data=self.data[idx], window=self.window & window, zeropoint=self.zeropoint,
                           identity=self.identity)
        if self._weight is not None:
            new._weight = self._weight[idx]
        if self._mask is not None:
            new._mask = self._mask[idx]
        new._psf = self._psf
        return new

    def jacobian_image(self, parameters=None, data=None):
        """target_image.py:406-430."""
        return Jacobian_Image(parameters=[] if parameters is None else parameters, data=data,
                              window=self.window.copy())

    def model_image(self, window=None):
        """target_image.py:432-441."""
        w = self.window.copy() if window is None else window.copy()
        return Model_Image(window=w)
