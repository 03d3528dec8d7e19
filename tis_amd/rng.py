"""Host side of the counter-based random streams (Philox4x32-10, Salmon et
al., SC'11) that the kernels use: the uniform numbers of the replica-exchange
acceptance test are drawn here with the same counter layout as on the device,
(lane, step_lo, system, step_hi | tag << 16) under key (seed_lo, seed_hi)."""
from __future__ import annotations

_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32(ctr, key):
    c0, c1, c2, c3 = (int(v) & _MASK for v in ctr)
    k0, k1 = (int(v) & _MASK for v in key)
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & _MASK, p1 & _MASK, ((p0 >> 32) ^ c3 ^ k1) & _MASK, p0 & _MASK
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return [c0, c1, c2, c3]


def uniform(seed: int, sys: int, step: int, stream: int, idx: int) -> float:
    """U(0,1) with 53 random bits for (seed, system, step, stream, index)."""
    ctr = [idx, step & _MASK, sys, ((step >> 32) & 0xFFFF) | ((2 + stream) << 16)]
    r = philox4x32(ctr, [seed & _MASK, (seed >> 32) & _MASK])
    m = (r[0] << 21) ^ (r[1] >> 11)
    return (m + 0.5) * 2.0 ** -53
