"""Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) in numpy, and the
per-particle uniform streams the device sources draw from it: a stream is
keyed by a 64-bit key and counted by (c0, c1, tag, block); each block yields
four 32-bit words, two words make one uniform double with 53 random bits.

Test infrastructure: it lets a test recompute the exact draws of a device
source and drive the oracle's restatement with them."""
from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = 0x9E3779B9
W1 = 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, key: int):
    """Vectorised Philox4x32-10; counters are uint32 arrays (broadcast)."""
    c = [np.asarray(v, np.uint64) & MASK for v in (c0, c1, c2, c3)]
    c = list(np.broadcast_arrays(*c))
    c = [v.copy() for v in c]
    k0, k1 = key & 0xFFFFFFFF, (key >> 32) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        n0 = hi1 ^ c[1] ^ np.uint64(k0)
        n2 = hi0 ^ c[3] ^ np.uint64(k1)
        c = [n0, lo1, n2, lo0]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return c


def u01(a, b):
    """Uniform double in [0, 1) from words a, b: (a >> 5) * 2^26 + (b >> 6), * 2^-53."""
    m = ((np.asarray(a, np.uint64) >> np.uint64(5)) << np.uint64(26)) | (
        np.asarray(b, np.uint64) >> np.uint64(6))
    return m.astype(np.float64) * (1.0 / 9007199254740992.0)


def uniforms(key: int, c0, c1, tag: int, count: int):
    """The first `count` uniforms of the stream (key; c0, c1, tag), for arrays
    of (c0, c1): shape [..., count]."""
    c0 = np.asarray(c0, np.uint64)
    c1 = np.asarray(c1, np.uint64)
    out = []
    nblk = (count + 1) // 2
    for blk in range(nblk):
        w = philox4x32(c0, c1, np.uint64(tag), np.uint64(blk), key)
        out.append(u01(w[0], w[1]))
        out.append(u01(w[2], w[3]))
    return np.stack(out[:count], axis=-1)
