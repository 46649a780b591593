"""contact_amd -- MI355X (gfx950) HIP engine for CONTACT's inner contact solve.

The library libcntc_amd.so exports the module-3 subset of CONTACT's cntc_* C interface
(include/cntc_amd.h).  `contact_amd.cntc` mirrors the reference's ctypes wrapper
(python_intfc) function by function.
"""
from . import cntc  # noqa: F401
from .cntc import lib, library_path  # noqa: F401
