"""Deterministic synthetic inputs (seeded splitmix64), identical on every host.

Fields are generated once on the host and copied to the GPU, so the CPU checker
and the GPU engine read the same bits.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """The first n outputs of splitmix64 started from state `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + _GAMMA * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, lo=0.0, hi=1.0) -> np.ndarray:
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def normal(seed: int, n: int) -> np.ndarray:
    """Box-Muller pairs from consecutive uniforms."""
    m = (n + 1) // 2
    raw = (splitmix64(seed, 2 * m) >> np.uint64(11)).astype(np.float64)
    u1 = (raw[0::2] + 1.0) * (2.0 ** -53)
    u2 = raw[1::2] * (2.0 ** -53)
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:n]


def periodic_fill(a: np.ndarray, g: int = 1) -> np.ndarray:
    """Copy-fill g ghost layers of a (z, y, x) block periodically (x, y, z sweeps)."""
    for ax in (2, 1, 0):
        n = a.shape[ax] - 2 * g
        src_lo = [slice(None)] * 3
        dst_lo = [slice(None)] * 3
        src_lo[ax] = slice(n, n + g)
        dst_lo[ax] = slice(0, g)
        a[tuple(dst_lo)] = a[tuple(src_lo)]
        src_hi = [slice(None)] * 3
        dst_hi = [slice(None)] * 3
        src_hi[ax] = slice(g, 2 * g)
        dst_hi[ax] = slice(n + g, n + 2 * g)
        a[tuple(dst_hi)] = a[tuple(src_hi)]
    return a


def cell_centres(n: int, g: int = 1, L: float = 1.0) -> np.ndarray:
    """x = (i + 0.5) * dx for window index i - g (ghosts included), dx = L/n."""
    dx = L / n
    return (np.arange(-g, n + g, dtype=np.float64) + 0.5) * dx


def fourier_field(n: int, seed: int, modes: int = 8, g: int = 1) -> np.ndarray:
    """phi = sum_m a_m sin(2 pi (k_m . x) + theta_m), k_m in {1..8}^3, a_m ~ N(0,1),
    theta_m ~ U[0, 2 pi); periodic on [0,1)^3; returned as an (n+2g)^3 (z,y,x) block."""
    kk = (splitmix64(seed, 3 * modes) % np.uint64(8)).astype(np.int64).reshape(modes, 3) + 1
    amp = normal(seed + 1, modes)
    theta = uniform(seed + 2, modes, 0.0, 2.0 * np.pi)
    x = cell_centres(n, g)
    out = np.zeros((n + 2 * g,) * 3)
    for m in range(modes):
        kx, ky, kz = kk[m]
        # separable phase: sin(a+b+c) built once per mode
        ph = (2.0 * np.pi * (kz * x[:, None, None] + ky * x[None, :, None] + kx * x[None, None, :])
              + theta[m])
        out += amp[m] * np.sin(ph)
    return out


def coefficient_field(n: int, seed: int, lo=0.5, hi=2.0, g: int = 1) -> np.ndarray:
    """k ~ U[lo, hi) on the interior, periodic ghosts."""
    a = np.zeros((n + 2 * g,) * 3)
    a[g:n + g, g:n + g, g:n + g] = uniform(seed, n ** 3, lo, hi).reshape(n, n, n)
    return periodic_fill(a, g)
