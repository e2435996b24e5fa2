"""Global configuration (AP_config.py:7-21 of the reference).

Parameters and host-side geometry are float64 (``ap_dtype``); images, model
renders and Jacobians live on ``ap_device`` in float32, the arithmetic type of
the gfx950 kernels.
"""

import logging
import sys

import torch

__all__ = ["ap_dtype", "ap_device", "ap_image_dtype", "ap_logger", "set_logging_output"]

ap_dtype = torch.float64
ap_image_dtype = torch.float32
ap_device = "cuda:0" if torch.cuda.is_available() else "cpu"
ap_verbose = 0

ap_logger = logging.getLogger("astrophot_amd")
if not ap_logger.handlers:
    _h = logging.StreamHandler(sys.stdout)
    _h.setFormatter(logging.Formatter("%(message)s"))
    ap_logger.addHandler(_h)
    ap_logger.setLevel(logging.INFO)


def set_logging_output(stdout=True, filename=None, **kwargs):
    """AP_config.py:24-66: route log output to stdout and/or a file."""
    for h in list(ap_logger.handlers):
        ap_logger.removeHandler(h)
    if stdout:
        h = logging.StreamHandler(sys.stdout)
        h.setFormatter(logging.Formatter(kwargs.get("stdout_format", "%(message)s")))
        ap_logger.addHandler(h)
    if filename is not None:
        h = logging.FileHandler(filename)
        h.setFormatter(logging.Formatter(kwargs.get("filename_format", "%(asctime)s:%(levelname)s: %(message)s")))
        ap_logger.addHandler(h)
