"""Write golden input/output vectors for the reference's quadrature and FFT
convolution (astrophot/utils/operations.py) into operations_vectors.npz.

Run in a container that holds the reference at /root/reference:
    python tests/golden/make_operations_vectors.py

Only utils/operations.py is loaded (it imports torch, scipy and numpy, nothing
from the astrophot package).  Its functions take the image header and the
brightness function as arguments; the header passed here is a plain object
with the two members grid_integrate uses (``pixelscale`` and
``rescale_pixel``), and the brightness function is a Sersic profile written
out below, so the vectors exercise the reference's own recursion, node
placement and per-level threshold (operations.py:150-247).
"""

import importlib.util
import math
import os
import sys

import numpy as np
import torch

REF = "/root/reference/astrophot/utils/operations.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "operations_vectors.npz")


def load_operations():
    spec = importlib.util.spec_from_file_location("ref_operations", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class Header:
    """The two members of an image header that grid_integrate reads."""

    def __init__(self, pixelscale):
        self.pixelscale = torch.as_tensor(pixelscale, dtype=torch.float64)

    @property
    def pixel_area(self):
        return torch.linalg.det(self.pixelscale).abs()

    def rescale_pixel(self, scale):
        return Header(self.pixelscale * scale)


def sersic_b(n):
    return 2 * n - 1 / 3 + 4 / (405 * n) + 46 / (25515 * n**2) + 131 / (1148175 * n**3) - 2194697 / (
        30690717750 * n**4)


PARAMS = dict(n=2.5, Re=3.0, Ie=1.0, q=0.6, PA=0.7, soft=1e-3)


def brightness(X, Y, image=None, parameters=None):
    p = PARAMS
    th = -(p["PA"] - math.pi / 2)
    c, s = math.cos(th), math.sin(th)
    Xr, Yr = c * X - s * Y, (s * X + c * Y) / p["q"]
    R = torch.sqrt(Xr**2 + Yr**2 + p["soft"] ** 2)
    b = sersic_b(p["n"])
    return image.pixel_area * 10 ** p["Ie"] * torch.exp(-b * (torch.pow(R / p["Re"], 1 / p["n"]) - 1))


def main():
    ops = load_operations()
    rng = np.random.default_rng(7)
    P = np.array([[0.9, 0.1], [-0.05, 1.1]])
    X = torch.tensor(rng.uniform(-4, 4, 64))
    Y = torch.tensor(rng.uniform(-4, 4, 64))
    out = {"X": X.numpy(), "Y": Y.numpy(), "P": P, "params": np.array([PARAMS[k] for k in
                                                                       ("n", "Re", "Ie", "q", "PA", "soft")])}
    for depth in (1, 2, 3):
        for ref in (1e-3, 1e-2):
            r = ops.grid_integrate(X.clone(), Y.clone(), Header(P), brightness, None, torch.float64, "cpu",
                                   quad_level=3, gridding=5, max_depth=depth, reference=ref)
            out[f"grid_d{depth}_r{ref:g}"] = r.numpy()
    res, refc = ops.single_quad_integrate(X.clone(), Y.clone(), Header(P), brightness, None, torch.float64,
                                          "cpu", quad_level=5)
    out["quad5_res"], out["quad5_ref"] = res.numpy(), refc.numpy()
    for n in (1, 2, 3, 4, 5, 7):
        qx, qy, qw = ops.quad_table(n, torch.tensor(P), torch.float64, "cpu")
        out[f"qt{n}_x"], out[f"qt{n}_y"], out[f"qt{n}_w"] = qx.numpy(), qy.numpy(), qw.numpy()
    img = torch.tensor(rng.uniform(0, 1, (31, 27)))
    psf = torch.tensor(rng.uniform(0, 1, (9, 7)))
    out["fft_img"], out["fft_psf"] = img.numpy(), psf.numpy()
    out["fft_prepadded"] = ops.fft_convolve_torch(img, psf, img_prepadded=True).numpy()
    np.savez(OUT, **out)
    print(f"wrote {OUT} ({len(out)} arrays)")


if __name__ == "__main__":
    sys.exit(main())
