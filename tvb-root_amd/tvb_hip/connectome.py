"""Seeded synthetic connectomes and the benchmark workloads built on them.

Generator: choose round(density * N (N-1) / 2) unordered node pairs i < j,
weights lognormal(0, 1) made symmetric and scaled to max 1, zero diagonal;
tract lengths U[lo, hi] for every pair, symmetric, zero diagonal.  The
horizon therefore depends on the full length matrix, as the reference's
does (connectivity.py:323).
"""
import numpy as np


def synthetic(n, density, length_lo, length_hi, seed):
    """Returns (weights, tract_lengths), both (n, n) float64."""
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    npairs = iu[0].size
    k = int(round(density * npairs))
    chosen = rng.choice(npairs, size=k, replace=False)
    w = np.zeros((n, n))
    w[iu[0][chosen], iu[1][chosen]] = rng.lognormal(0.0, 1.0, size=k)
    w = w + w.T
    if k:
        w /= w.max()
    lengths = np.zeros((n, n))
    lengths[iu] = rng.uniform(length_lo, length_hi, size=npairs)
    lengths = lengths + lengths.T
    return w, lengths


def idelays(lengths, speed, dt):
    """Integer delays and horizon as Connectivity.set_idelays computes them."""
    d = np.rint((lengths / speed) / dt).astype(np.int32)
    return d, int(d.max()) + 1


# Workloads (sizes of the five benchmark configurations).
WORKLOADS = {
    "c0_g2d_76": dict(n=76, density=0.30, lengths=(10.0, 200.0), speed=3.0, dt=0.1, model="G2D",
                      coupling_a=0.0126, scheme="HeunDeterministic", duration=1000.0, tavg=1.0),
    "c1_mpr_192": dict(n=192, density=0.20, lengths=(5.0, 250.0), speed=4.0, dt=0.05, model="MPR",
                       coupling_a=0.5, scheme="HeunStochastic", nsig=(0.01, 0.02), duration=10000.0,
                       tavg=1.0, bold=720.0),
    "c2_mpr_sweep_68x8192": dict(n=68, density=0.25, lengths=(10.0, 180.0), speed=5.0, dt=0.05, model="MPR",
                                 grid=(128, 64), G=(0.0, 2.0), sigma=(1e-4, 1e-2),
                                 scheme="HeunStochastic", duration=20000.0, tavg=1.0, bold=2000.0),
    "c3_jr_uneven_1024": dict(n_choices=(68, 96, 192, 400, 998), n_inst=1024, density=(0.1, 0.3),
                              max_delay=800, dt=0.1, model="JR", scheme="EulerStochastic",
                              duration=(2000.0, 30000.0), tavg=2.0, bold=2000.0),
}


def sweep_grid(n_g=128, n_s=64, g_range=(0.0, 2.0), s_range=(1e-4, 1e-2)):
    """The C2 parameter grid: coupling gain G (Linear a) linearly spaced and
    noise sigma (Additive nsig of both state variables) log spaced.
    Returns (G, sigma) arrays of length n_g * n_s, G-major."""
    G = np.linspace(g_range[0], g_range[1], n_g)
    s = np.geomspace(s_range[0], s_range[1], n_s)
    GG, SS = np.meshgrid(G, s, indexing="ij")
    return GG.reshape(-1), SS.reshape(-1)
