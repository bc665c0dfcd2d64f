"""corrfunc_amd -- MI355X-native theory pair counting with Corrfunc's API.

`corrfunc_amd.theory` mirrors `Corrfunc.theory` (DD, DDrppi, DDsmu, xi, wp);
the C ABI of the same engine is declared in include/corrfunc_theory.h.
"""
__version__ = "0.1.0"

from . import _abi  # noqa: F401

__all__ = ("theory", "utils")
