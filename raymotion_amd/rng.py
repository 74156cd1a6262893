"""Vectorised numpy form of the per-pixel random stream of the renderer.

PCG RXS-M-XS 32 (O'Neill 2014) seeded through splitmix64, as the kernels and
the reference shader use it (src/engine/pathtrace.metal:93-101). Used to
synthesise test rays from the same generator family the GPU consumes.
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_state(seeds) -> np.ndarray:
    """32-bit PCG states from 64-bit seeds (Rng::seed)."""
    z = np.asarray(seeds, np.uint64) + GOLDEN
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return (z ^ (z >> np.uint64(31))).astype(np.uint32)


def draws(states: np.ndarray, count: int) -> tuple[np.ndarray, np.ndarray]:
    """`count` uniforms in [0,1) from each stream; returns (values (n, count), new states)."""
    st = np.array(states, np.uint32, copy=True)
    out = np.empty((st.size, count), np.float32)
    with np.errstate(over="ignore"):
        for k in range(count):
            st = st * np.uint32(747796405) + np.uint32(2891336453)
            word = ((st >> ((st >> np.uint32(28)) + np.uint32(4))) ^ st) * np.uint32(277803737)
            out[:, k] = (((word >> np.uint32(22)) ^ word) >> np.uint32(8)).astype(np.float32) * np.float32(2.0 ** -24)
    return out, st


def random_rays(n: int, seed: int = 0, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """n rays with origins uniform in the box [lo, hi] and directions uniform on
    the sphere (z = 2u - 1, phi = 2 pi v); ray i uses the stream seeded with
    seed ^ (i + 1) * golden and takes five draws."""
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        seeds = np.uint64(seed) ^ ((i + np.uint64(1)) * GOLDEN)
    u, _ = draws(splitmix_state(seeds), 5)
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    origins = lo + (hi - lo) * u[:, 0:3]
    z = 2.0 * u[:, 3].astype(np.float64) - 1.0
    phi = 2.0 * np.pi * u[:, 4].astype(np.float64)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)
    return np.ascontiguousarray(origins, np.float32), np.ascontiguousarray(dirs)
