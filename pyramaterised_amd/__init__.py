"""MI355X (gfx950) statevector engine behind pyramaterised's PQC / gate /
Measurements API.

``import pyramaterised_amd as pyqc`` gives the reference's top-level names
(pyramaterised/__init__.py:1-4 re-exports circuit, and through it every gate,
plus the measure, gates and templates modules).  Kets, overlaps, partial
traces, the QFIM and the sampling measures run in the HIP library
``lib/libpqs.so``; there is no CPU fallback.
"""
from .circuit import *  # noqa: F401,F403
from . import autograd, gates, measure, serve, templates, qobj  # noqa: F401
from .qobj import Ket, PauliSum, Qobj, basis, bell_state, expect, qeye, sigmax, sigmay, sigmaz, tensor  # noqa: F401

__version__ = "0.1.0"
