"""Seeded SplitMix64 streams, vectorised with numpy.

Every synthetic scene is drawn from these streams so that the GPU path and the
CPU oracle consume byte-identical input buffers. Output k (k = 0, 1, ...) of
stream `seed` is SplitMix64 applied to state seed + (k + 1) * golden_gamma.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

DEFAULT_SEED = 0x6B617A616E  # "kazan"


def splitmix64(seed, n, offset=0):
    """n consecutive 64-bit outputs of stream `seed`, starting at output `offset`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform_f32(seed, n, offset=0):
    """n floats in [0, 1): (u64 >> 40) * 2**-24, exact in f32."""
    u = splitmix64(seed, n, offset) >> np.uint64(40)
    return (u.astype(np.float64) * (1.0 / 16777216.0)).astype(np.float32)


def uniform_u8(seed, n, offset=0):
    return (splitmix64(seed, n, offset) >> np.uint64(56)).astype(np.uint8)


def uniform_u32(seed, n, offset=0):
    return (splitmix64(seed, n, offset) >> np.uint64(32)).astype(np.uint32)
