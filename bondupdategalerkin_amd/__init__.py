"""MI355X (gfx950) backend for the hot path of the BUG time integrator.

The step ``cbe_bug_step`` of madhavmen/BondUpdateGalerkin (src/BUG/TimeEvolution/bug_step.jl)
on device-resident block-sparse U(1) matrix product states; every contraction, Krylov recurrence,
QR and SVD runs in ``libbug.so`` (hand-written HIP for gfx950) behind the C ABI of
``include/bug.h``.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
