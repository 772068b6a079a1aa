"""Synthetic tridiagonal workloads (numpy PCG64, fixed seeds).

Sizes and spectra follow the benchmark definitions:

* ``uniform(n)``: D, E i.i.d. U(-1, 1);
* ``toeplitz_pert(n)``: the 1-2-1 Toeplitz matrix (D = 2, E = 1) plus a 1e-6
  uniform perturbation of every entry;
* ``glued_wilkinson(n, glue)``: copies of W21+ (D = |10 - i|, i = 0..20, E = 1)
  glued by off-diagonals glue * U(0.5, 1), padded to n with 1-2-1 rows;
* ``normal(n)``: D, E ~ N(0, 1).

Each generator returns (D, E) as float64 arrays of length n (E[n-1] = 0).
"""
import numpy as np

SEED_BASE = 0x4D5233_0000


def rng_for(stream, extra=0):
    return np.random.default_rng([SEED_BASE + stream, extra])


def uniform(n, rng):
    D = rng.uniform(-1.0, 1.0, n)
    E = rng.uniform(-1.0, 1.0, n)
    E[n - 1] = 0.0
    return D, E


def normal(n, rng):
    D = rng.standard_normal(n)
    E = rng.standard_normal(n)
    E[n - 1] = 0.0
    return D, E


def toeplitz_pert(n, rng, pert=1e-6):
    D = 2.0 + pert * rng.uniform(-1.0, 1.0, n)
    E = 1.0 + pert * rng.uniform(-1.0, 1.0, n)
    E[n - 1] = 0.0
    return D, E


def glued_wilkinson(n, rng, glue=1e-8):
    w = np.abs(10.0 - np.arange(21.0))
    nb = n // 21
    D = np.full(n, 2.0)
    E = np.ones(n)
    for b in range(nb):
        D[21 * b:21 * b + 21] = w
        E[21 * b:21 * b + 20] = 1.0
        if 21 * b + 20 < n - 1:
            E[21 * b + 20] = glue * rng.uniform(0.5, 1.0)
    E[n - 1] = 0.0
    return D, E


def single_uniform_1000():
    """n = 1000, U(-1, 1), jobz V, range A."""
    return uniform(1000, rng_for(1))


def single_glued_20000():
    """n = 20000 glued Wilkinson (952 W21+ blocks, glue 1e-8), jobz V, range A."""
    return glued_wilkinson(20000, rng_for(2), 1e-8)


def mixed_sizes(count, lo, hi, rng):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)


def mixed_matrix(n, rng, glue=1e-10):
    """The spectrum mix of the batched workload: 50% uniform, 25% perturbed
    Toeplitz, 25% glued Wilkinson."""
    u = rng.uniform()
    if u < 0.5:
        return uniform(n, rng), "uniform"
    if u < 0.75:
        return toeplitz_pert(n, rng), "toeplitz"
    return glued_wilkinson(n, rng, glue), "wilkinson"


def batched_mixed(count=4096, seed_extra=0):
    """Batched: `count` matrices, n log-uniform in [32, 2048], the 50/25/25 mix,
    jobz V, range A.  Returns a list of (D, E, kind)."""
    rng = rng_for(3, seed_extra)
    sizes = mixed_sizes(count, 32, 2048, rng)
    out = []
    for n in sizes:
        (D, E), kind = mixed_matrix(int(n), rng)
        out.append((D, E, kind))
    return out


def batched_range_i(count=256, seed_extra=0):
    """Batched subset: n log-uniform in [4096, 50000], N(0,1), jobz V, range I over a
    contiguous window of ceil(0.1 n) indices.  Returns (D, E, il, iu)."""
    rng = rng_for(4, seed_extra)
    sizes = mixed_sizes(count, 4096, 50000, rng)
    out = []
    for n in sizes:
        n = int(n)
        D, E = normal(n, rng)
        w = int(np.ceil(0.1 * n))
        il = int(rng.integers(1, n - w + 2))
        out.append((D, E, il, il + w - 1))
    return out


def served_stream(count, rate, seed_extra=0):
    """Served stream: Poisson arrivals at `rate` per second, n log-uniform in
    [64, 16384], jobz 70% V / 30% N, range 60% A / 25% I (5-20% window) / 15% V
    (10% of the Gershgorin span).  Returns a list of dicts with arrival times."""
    rng = rng_for(5, seed_extra)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    sizes = mixed_sizes(count, 64, 16384, rng)
    out = []
    for i in range(count):
        n = int(sizes[i])
        (D, E), kind = mixed_matrix(n, rng)
        jobz = "V" if rng.uniform() < 0.7 else "N"
        u = rng.uniform()
        req = dict(t=float(t[i]), n=n, D=D, E=E, kind=kind, jobz=jobz)
        if u < 0.6:
            req["range"] = "A"
        elif u < 0.85:
            frac = rng.uniform(0.05, 0.20)
            w = max(1, int(np.ceil(frac * n)))
            il = int(rng.integers(1, n - w + 2))
            req.update(range="I", il=il, iu=il + w - 1)
        else:
            ae = np.abs(E[: n - 1])
            off = np.concatenate([[0.0], ae]) + np.concatenate([ae, [0.0]])
            gl, gu = float(np.min(D - off)), float(np.max(D + off))
            span = 0.1 * (gu - gl)
            lo = float(rng.uniform(gl, gu - span))
            req.update(range="V", vl=lo, vu=lo + span)
        out.append(req)
    return out
