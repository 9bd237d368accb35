"""Counter-based SplitMix64 generator for seeded synthetic inputs.

The reference seeds its initial field from the wall clock
(src/m_FlowCase.f90:69-71), so it cannot reproduce a case.  Here every seeded
input is a pure function of (seed, stream, index):

    x_q = mix(key + (q + 1) * 0x9E3779B97F4A7C15),  key = seed + stream * 2**32

with ``mix`` the SplitMix64 finaliser (Steele, Lea and Flood, 2014).  For
stream 0 this is exactly the sequential SplitMix64 generator started from
``seed``.  Uniform doubles take the top 53 bits; normals use Box-Muller on the
pair (u_{2q}, u_{2q+1}).
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, n: int, stream: int = 0, start: int = 0) -> np.ndarray:
    """Raw 64-bit outputs q = start .. start+n-1."""
    key = np.uint64((seed + (stream << 32)) & 0xFFFFFFFFFFFFFFFF)
    q = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(key + q * GOLDEN)


def uniform(seed: int, n: int, stream: int = 0, start: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1)."""
    x = splitmix64(seed, n, stream, start)
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller; value q uses uniforms 2q and 2q+1."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, stream)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    t = 2.0 * np.pi * u2
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(t)
    out[1::2] = r * np.sin(t)
    return out[:n]
