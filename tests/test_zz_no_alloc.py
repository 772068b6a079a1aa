"""mr3_solve_small_noalloc (the small-matrix route with every work array in a static
device buffer) against mr3_solve_batch, through the C ABI: the same return code, m,
info and scalar outputs, and W, Z, Zsupp bit for bit, on two calls in a row (the second
with fewer and smaller problems, so that it runs over what the first left in the
buffer), and its rejections (n >= 500, too many problems, too much work), which are
decided on the host before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import mr3smp_amd
    return mr3smp_amd


def _problems(mats, dev, seed):
    import torch
    rng = np.random.default_rng(seed)
    out = []
    for i, (D, E, _) in enumerate(mats):
        n = len(D)
        il = int(rng.integers(1, n + 1))
        iu = int(rng.integers(il, n + 1))
        med = float(np.median(D))
        kw = [dict(jobz="V", range="A"), dict(jobz="N", range="A"),
              dict(jobz="V", range="I", il=il, iu=iu, mcols=iu - il + 1),
              dict(jobz="V", range="V", vl=med - 0.3, vu=med + 0.4),
              dict(jobz="C", range="V", vl=med - 0.3, vu=med + 0.4),
              dict(jobz="V", range="A", tryrac=0)][i % 6]
        out.append(dict(D=torch.tensor(D, device=dev), E=torch.tensor(E, device=dev), **kw))
    return out


def _fresh(problems):
    return [{k: v for k, v in p.items() if k not in ("W", "Z", "Zsupp")} for p in problems]


def _compare(M, problems, label):
    import torch
    ref = _fresh(problems)
    rc0, a0, k0 = M.solve_batch(ref)
    got = _fresh(problems)
    a1, k1 = M.prepare_batch(got)
    print(f"{label}: {len(problems)} problems", flush=True)
    rc1 = M.run_batch_noalloc(a1)
    torch.cuda.synchronize()
    assert rc1 == rc0 == 0, (label, rc0, rc1)
    for i in range(len(problems)):
        p, q = a0[i], a1[i]
        assert (q.m, q.info, q.il_out, q.iu_out, q.tryrac_out) == (p.m, p.info, p.il_out, p.iu_out,
                                                                    p.tryrac_out), (label, i)
        assert (q.vl_out, q.vu_out) == (p.vl_out, p.vu_out), (label, i)
        m = p.m
        if problems[i]["jobz"] == "C":
            continue
        assert torch.equal(k1[i]["W"][:m], k0[i]["W"][:m]), (label, i, "W")
        if problems[i]["jobz"] == "V":
            n = problems[i]["D"].numel()
            assert torch.equal(k1[i]["Z"][:m, :n], k0[i]["Z"][:m, :n]), (label, i, "Z")
            assert torch.equal(k1[i]["Zsupp"][:2 * m], k0[i]["Zsupp"][:2 * m]), (label, i, "Zsupp")


def test_noalloc_equals_batch_on_two_calls(M):
    import torch
    from mr3smp_amd import workloads as WL
    dev = torch.device("cuda:0")
    mats = [m for m in WL.batched_mixed(4096) if len(m[0]) < 500]
    first = sorted(mats[::23], key=lambda m: -len(m[0]))[:36]   # up to n = 499
    second = [m for m in mats[5::61] if len(m[0]) < 200][:20]   # fewer and smaller
    _compare(M, _problems(first, dev, 1), "first call")
    _compare(M, _problems(second, dev, 2), "second call")


def test_noalloc_rejects_what_it_cannot_hold(M):
    import torch
    from mr3smp_amd import workloads as WL
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(3)
    big = WL.uniform(600, rng)
    probs = [dict(D=torch.tensor(big[0], device=dev), E=torch.tensor(big[1], device=dev))]
    arr, _ = M.prepare_batch(probs)
    assert M.run_batch_noalloc(arr) == 3 and arr[0].info == 3 and arr[0].m == 0
    tiny = WL.uniform(5, rng)
    Dt, Et = torch.tensor(tiny[0], device=dev), torch.tensor(tiny[1], device=dev)
    arr, _ = M.prepare_batch([dict(D=Dt, E=Et) for _ in range(1025)])
    assert M.run_batch_noalloc(arr) == 3 and all(a.info == 3 for a in arr)
    n499 = WL.uniform(499, rng)
    D4, E4 = torch.tensor(n499[0], device=dev), torch.tensor(n499[1], device=dev)
    arr, _ = M.prepare_batch([dict(D=D4, E=E4) for _ in range(48)])   # about 78 MB of work
    assert M.run_batch_noalloc(arr) == 3 and all(a.info == 3 for a in arr)
