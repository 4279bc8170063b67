"""Host-side probe generation of the DD-projected sketches (drawn on the host, uploaded).

The probe stream is counter based: number ``i`` of the stream is a complex Gaussian built by the
Box-Muller transform from two 53-bit uniforms, splitmix64(key + 2i) and splitmix64(key + 2i + 1),
scaled to E|z|^2 = 1 as Julia's ``randn(ComplexF64)``.  Being counter based, a probe block can be
"drawn at full width and trimmed" (``_randomize_like`` + ``_trim_per_sector``,
src/BUG/Randomised/cbe.jl:156-178 and truncation.jl:283-288) by skipping the unused numbers
instead of generating them.
"""
from __future__ import annotations

import os
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_MUL1 = np.uint64(0xBF58476D1CE4E5B9)
_MUL2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _MUL1
        z = (z ^ (z >> np.uint64(27))) * _MUL2
        return z ^ (z >> np.uint64(31))


def _values(key: np.uint64, i: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        two_i = np.uint64(2) * i + key
        a = _mix(two_i)
        b = _mix(two_i + np.uint64(1))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0
    u2 = (b >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    rad = np.sqrt(-np.log(u1))
    ang = 2.0 * np.pi * u2
    return rad * np.cos(ang) + 1j * rad * np.sin(ang)


# Large draws are cut into chunks evaluated by a small thread pool (numpy's element-wise loops
# release the GIL); every number depends only on its own position, so the chunking does not
# change any value.
_CHUNK = 1 << 15
_pool = None
_pool_lock = threading.Lock()


def _executor() -> ThreadPoolExecutor:
    global _pool
    with _pool_lock:
        if _pool is None:
            _pool = ThreadPoolExecutor(max_workers=max(1, min(8, (os.cpu_count() or 1) // 2)),
                                       thread_name_prefix="probes")
        return _pool


class ProbeStream:
    """Seeded, skippable stream of complex Gaussian numbers."""

    def __init__(self, seed: int):
        self._key = _mix(np.array([seed], dtype=np.uint64))[0]
        self.position = 0

    def at(self, i: np.ndarray) -> np.ndarray:
        """The numbers at stream positions ``i`` (uint64), without moving the stream."""
        i = np.asarray(i, dtype=np.uint64)
        n = i.size
        if n <= 2 * _CHUNK:
            return _values(self._key, i)
        flat = i.reshape(-1)
        out = np.empty(n, dtype=np.complex128)

        def fill(lo):
            out[lo: lo + _CHUNK] = _values(self._key, flat[lo: lo + _CHUNK])

        list(_executor().map(fill, range(0, n, _CHUNK)))
        return out.reshape(i.shape)

    def draw(self, n: int) -> np.ndarray:
        vals = self.at(np.arange(self.position, self.position + n, dtype=np.uint64))
        self.position += n
        return vals

    def skip(self, n: int):
        self.position += n

    def split(self) -> "ProbeStream":
        """A new stream seeded by this stream's next 64-bit word (the stream advances by one
        number): the parallel right half-sweep's generator, ``MersenneTwister(rand(rng, UInt64))``
        (src/BUG/TimeEvolution/bug_step.jl:238)."""
        with np.errstate(over="ignore"):
            word = _mix(np.array([np.uint64(2) * np.uint64(self.position) + self._key],
                                 dtype=np.uint64))[0]
        self.position += 1
        return ProbeStream(int(word))
