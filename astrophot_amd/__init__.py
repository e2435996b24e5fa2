"""astrophot_amd: AstroPhot's image-model hot path on MI355X (gfx950).

``model.sample`` of Sersic / exponential / Gaussian / Moffat galaxies and a
flat sky (adaptive supersampled centres, PSF convolution), the parameter
Jacobian and ``fit.LM`` keep AstroPhot's Model / Image / Window / fit API and
run on hand-written HIP kernels behind the C ABI of include/astrophot_hip.h.
"""

from . import AP_config, errors, fit, image, models, param, utils
from . import batch, cnn

__all__ = ["AP_config", "errors", "fit", "image", "models", "param", "utils", "batch", "cnn"]
__version__ = "0.1.0"
