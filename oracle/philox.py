"""Philox4x32-10 and the Box-Muller normals of the device noise source.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

The reference draws its noise from numpy's legacy MT19937 ``RandomState``
(tvb_library/tvb/simulator/noise.py:102-105, :195-198), which cannot be
reproduced per instance on a GPU.  The device's batched noise source is the
counter-based Philox4x32-10 generator of Salmon, Moraes, Dror and Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11, restated here from the
paper's description in numpy uint32/uint64 arithmetic.  It is pinned by the
paper's published known-answer vectors (tests/test_oracle_philox.py).

Stream layout (the device uses the same):
    key     = (seed & 0xffffffff, seed >> 32)
    counter = (node | (call << 20), stream, step & 0xffffffff, step >> 32)
one call yields four uint32 words -> two Box-Muller pairs -> four normals for
state variables ``4*call .. 4*call+3``.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32(ctr, key, rounds=10):
    """Philox4x32 with ``rounds`` rounds.

    ``ctr``: uint32 array (..., 4); ``key``: uint32 array (..., 2).
    Returns uint32 array (..., 4)."""
    c = [np.asarray(ctr[..., i], np.uint32) for i in range(4)]
    k0 = np.asarray(key[..., 0], np.uint32).copy()
    k1 = np.asarray(key[..., 1], np.uint32).copy()
    with np.errstate(over="ignore"):
        for r in range(rounds):
            if r:
                k0 = (k0 + W0).astype(np.uint32)
                k1 = (k1 + W1).astype(np.uint32)
            p0 = M0 * c[0].astype(np.uint64)
            p1 = M1 * c[2].astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & MASK32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & MASK32).astype(np.uint32)
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack(c, axis=-1)


def counters(node, call, stream, step):
    """Counter words of the device noise stream (broadcasting)."""
    node = np.asarray(node, np.uint64)
    call = np.asarray(call, np.uint64)
    stream = np.asarray(stream, np.uint64)
    step = np.asarray(step, np.uint64)
    shape = np.broadcast(node, call, stream, step).shape
    out = np.empty(shape + (4,), np.uint32)
    out[..., 0] = ((node | (call << np.uint64(20))) & MASK32).astype(np.uint32)
    out[..., 1] = (stream & MASK32).astype(np.uint32)
    out[..., 2] = (step & MASK32).astype(np.uint32)
    out[..., 3] = ((step >> np.uint64(32)) & MASK32).astype(np.uint32)
    return out


def seed_key(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32)


def box_muller(words, precision="f64"):
    """Four uint32 words -> four standard normals, as the device computes them.

    f64: u1 = (w0 + 1) 2^-32 in (0, 1], u2 = w1 2^-32 in [0, 1),
         z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = ... sin(2 pi u2); same for (w2, w3).
    f32: the same with 24-bit uniforms u1 = ((w0 >> 8) + 1) 2^-24,
         u2 = (w1 >> 8) 2^-24; returned in float64 (the device evaluates it in
         float32 with the hardware log2, sqrt, sin and cos instructions, a few
         float32 ulp from these values)."""
    w = words.astype(np.float64) if precision == "f64" else (words >> np.uint32(8)).astype(np.float64)
    scale = 2.0 ** -32 if precision == "f64" else 2.0 ** -24
    out = np.empty(words.shape, np.float64)
    for a, b in ((0, 1), (2, 3)):
        u1 = (w[..., a] + 1.0) * scale
        u2 = w[..., b] * scale
        rad = np.sqrt(-2.0 * np.log(u1))
        th = 2.0 * np.pi * u2
        out[..., a] = rad * np.cos(th)
        out[..., b] = rad * np.sin(th)
    return out


def normals(seed, stream, n_steps, nsvar, n_node, step0=1, precision="f64"):
    """Standard normals z[k, sv, node] the device draws for steps
    ``step0 .. step0 + n_steps - 1`` of one instance (Philox mode)."""
    ncall = (nsvar + 3) // 4
    steps = np.arange(step0, step0 + n_steps, dtype=np.uint64)[:, None, None]
    calls = np.arange(ncall, dtype=np.uint64)[None, :, None]
    nodes = np.arange(n_node, dtype=np.uint64)[None, None, :]
    ctr = counters(nodes, calls, np.uint64(stream), steps)
    key = np.broadcast_to(seed_key(seed), ctr.shape[:-1] + (2,))
    words = philox4x32(ctr, key)                     # (n_steps, ncall, N, 4)
    z = box_muller(words, precision)                 # (n_steps, ncall, N, 4)
    z = np.moveaxis(z, -1, 2).reshape(n_steps, ncall * 4, n_node)
    return z[:, :nsvar]
