"""The tree form of dlaebz's IJOB=3 loop that k_prep runs for the range-I bracket
(mrrr_val.c:293-319, odebz.f:380-536 with one interval) gives the sequential loop's
points and results.  Both are restated here in Python on small matrices: the tree form
evaluates all 2^L - 1 points of the next L steps from the current interval, then walks
the sequential loop's path through them with its updates, stopping test and iteration
cap.  The device code follows the tree form step for step."""
import numpy as np


def count2(d, e2, x, pivmin):
    """Sturm count, dlaebz IJOB=2/3 convention (odebz.f:395-415)."""
    t = d[0] - x
    c = 0
    if t <= pivmin:
        c = 1
        t = min(t, -pivmin)
    for j in range(1, len(d)):
        t = d[j] - e2[j - 1] / t - x
        if t <= pivmin:
            c += 1
            t = min(t, -pivmin)
    return c


def stop(a1, a2, n1, n2, abstol, reltol, pivmin):
    return abs(a2 - a1) < max(max(abstol, pivmin), reltol * max(abs(a2), abs(a1))) or n1 >= n2


def sequential(d, e2, pivmin, abstol, reltol, nitmax, nval, c, a1, a2, n1, n2):
    for _ in range(nitmax):
        k = count2(d, e2, c, pivmin)
        if k <= nval:
            a1, n1 = c, k
        if k >= nval:
            a2, n2 = c, k
        if stop(a1, a2, n1, n2, abstol, reltol, pivmin):
            break
        c = 0.5 * (a1 + a2)
    return a1, a2, n1, n2


def tree(d, e2, pivmin, abstol, reltol, nitmax, nval, c, a1, a2, n1, n2, L=7):
    it, done, rounds = 0, nitmax <= 0, 0
    while not done:
        cnt = {}
        for t in range(1, 1 << L):
            dt = t.bit_length() - 1
            if it + dt < nitmax:
                b1, b2, cc = a1, a2, c
                for b in range(dt - 1, -1, -1):
                    if (t >> b) & 1:
                        b1 = cc
                    else:
                        b2 = cc
                    cc = 0.5 * (b1 + b2)
                cnt[t] = count2(d, e2, cc, pivmin)
        rounds += 1
        t = 1
        for _ in range(L):
            if it >= nitmax:
                done = True
                break
            k = cnt[t]
            if k <= nval:
                a1, n1 = c, k
            if k >= nval:
                a2, n2 = c, k
            it += 1
            if stop(a1, a2, n1, n2, abstol, reltol, pivmin):
                done = True
                break
            c = 0.5 * (a1 + a2)
            t = 2 * t + (1 if k < nval else 0)
        if it >= nitmax:
            done = True
    return (a1, a2, n1, n2), rounds


def test_tree_form_equals_sequential_loop():
    rng = np.random.default_rng(7)
    eps = np.finfo(np.float64).eps
    checked = 0
    for trial in range(36):
        n = int(rng.integers(3, 48))
        kind = trial % 3
        if kind == 0:
            d, e = rng.normal(size=n), rng.normal(size=n)
        elif kind == 1:
            d, e = np.full(n, 2.0), np.ones(n)
        else:   # near-splits and repeated diagonal values
            d = np.round(rng.normal(size=n))
            e = np.where(rng.random(n) < 0.3, 0.0, 1e-9 * rng.random(n))
        e2 = e[: n - 1] ** 2
        ae = np.abs(e[: n - 1])
        off = np.concatenate([[0.0], ae]) + np.concatenate([ae, [0.0]])
        pivmin = np.finfo(np.float64).tiny * max(1.0, float(ae.max()) ** 2 if n > 1 else 1.0)
        gl, gu = float(np.min(d - off)), float(np.max(d + off))
        tnorm = max(abs(gl), abs(gu))
        gl -= 2 * tnorm * eps * n + 4 * pivmin
        gu += 2 * tnorm * eps * n + 4 * pivmin
        itmax = int((np.log(tnorm + pivmin) - np.log(pivmin)) / np.log(2.0) + 2)
        atoli = 2 * (2 * np.finfo(np.float64).tiny + 2 * pivmin)
        il = int(rng.integers(1, n + 1))
        iu = int(rng.integers(il, n + 1))
        for nval, c in ((il - 1, gl), (iu, gu)):
            for cap in (itmax, 5, 9):
                args = (d, e2, pivmin, atoli, 2 * eps, cap, nval, c, gl, gu, -1, n + 1)
                got, rounds = tree(*args)
                assert got == sequential(*args), (trial, nval, cap)
                checked += 1
    assert checked == 36 * 6
