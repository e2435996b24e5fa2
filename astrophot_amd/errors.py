"""Exception hierarchy of the reference's ``astrophot.errors`` (errors/*.py),
plus ``APXError`` for failures of the HIP library itself."""

__all__ = (
    "AstroPhotError", "NameNotAllowed", "SpecificationConflict", "OptimizeStop", "InvalidWindow",
    "InvalidData", "InvalidImage", "InvalidModel", "InvalidTarget", "UnrecognizedModel",
    "InvalidParameter", "APXError",
)


class AstroPhotError(Exception):
    """Base exception (errors/base.py)."""


class NameNotAllowed(AstroPhotError):
    """Invalid name of an object (errors/base.py)."""


class SpecificationConflict(AstroPhotError):
    """Conflicting or unsupported inputs (errors/base.py)."""


class OptimizeStop(AstroPhotError):
    """Stops an optimisation (errors/fit.py)."""


class InvalidWindow(AstroPhotError):
    """Misspecified window (errors/image.py)."""


class InvalidData(AstroPhotError):
    """Image cannot determine its data (errors/image.py)."""


class InvalidImage(AstroPhotError):
    """Image cannot be used as given (errors/image.py)."""


class InvalidModel(AstroPhotError):
    """Inappropriate model object (errors/models.py)."""


class InvalidTarget(AstroPhotError):
    """Target assigned incorrectly (errors/models.py)."""


class UnrecognizedModel(AstroPhotError):
    """Unknown model_type (errors/models.py)."""


class InvalidParameter(AstroPhotError):
    """Parameter value violates its limits (errors/param.py)."""


class APXError(AstroPhotError):
    """The gfx950 library is missing, no GPU is visible, or a library call failed."""
