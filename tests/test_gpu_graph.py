"""GPU tests of the captured batched solve (mr3_graph_create / mr3_graph_launch): a
HIP graph that replays the whole batch -- the small-matrix route and the task-parallel
pipeline with a fixed number of tree levels -- must give the bits of the eager
mr3_solve_batch on every replay, with the program's own eager work and
synchronisations between the replays, and must report (not hide) a tree that outgrows
the capture."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import mr3smp_amd
    return mr3smp_amd


def _batch(WL, torch, count, seed):
    dev = torch.device("cuda:0")
    mats = WL.batched_mixed(count, seed_extra=seed)
    return mats, [dict(D=torch.tensor(D, device=dev), E=torch.tensor(E, device=dev)) for D, E, _ in mats]


def _same(a, b):
    return bool((a == b).all().item()) if a.numel() else True


def _check(arr, keep, ref_arr, ref_keep, tag):
    import torch
    for i in range(len(arr)):
        assert arr[i].info == ref_arr[i].info == 0, (tag, i, arr[i].info)
        m = arr[i].m
        assert m == ref_arr[i].m, (tag, i)
        n = keep[i]["D"].numel()
        for k, sl in (("W", slice(0, m)), ("Zsupp", slice(0, 2 * m))):
            assert torch.equal(keep[i][k][sl], ref_keep[i][k][sl]), (tag, i, k)
        assert torch.equal(keep[i]["Z"][:m, :n], ref_keep[i]["Z"][:m, :n]), (tag, i, "Z")


def test_graph_replays_equal_eager_solve(M):
    import torch
    from mr3smp_amd import workloads as WL
    mats, probs = _batch(WL, torch, 96, 11)
    assert any(len(D) >= 500 for D, _, _ in mats) and any(len(D) < 500 for D, _, _ in mats)
    print("graph: eager reference solve", flush=True)
    rc, ref_arr, ref_keep = M.solve_batch([dict(D=p["D"], E=p["E"]) for p in probs])
    torch.cuda.synchronize()
    assert rc == 0
    arr, keep = M.prepare_batch(probs)
    print("graph: capture", flush=True)
    g = M.BatchGraph(arr)
    assert g.rc == 0
    _check(arr, keep, ref_arr, ref_keep, "after create")
    other = [dict(D=p["D"].clone(), E=p["E"].clone()) for p in probs[:8]]
    for r in range(3):
        for k in keep:   # spoil the outputs: each replay must write them all again
            k["W"].fill_(float("nan"))
            k["Z"].fill_(float("nan"))
            k["Zsupp"].fill_(-7)
        print(f"graph: replay {r}", flush=True)
        assert g.launch() == 0
        _check(arr, keep, ref_arr, ref_keep, f"replay {r}")
        # the program's own eager work between replays
        rc2, _, _ = M.solve_batch(other)
        x = torch.randn(1000, 1000, device="cuda:0", dtype=torch.float64)
        (x @ x).sum().item()
        torch.cuda.synchronize()
        assert rc2 == 0
    g.close()


def test_graph_new_inputs_same_bits_or_reported(M):
    """New D and E in the same arrays: the replay equals an eager solve of the new
    inputs bit for bit, or -- when their tree outgrows the capture -- returns 4 with
    info 2 on the problems of the pipeline."""
    import torch
    from mr3smp_amd import workloads as WL
    mats, probs = _batch(WL, torch, 24, 12)
    arr, keep = M.prepare_batch(probs)
    g = M.BatchGraph(arr)
    assert g.rc == 0
    gen = torch.Generator(device="cpu").manual_seed(5)
    for p in probs:   # new values in the same arrays: D scaled by 1 + O(1e-3) per entry
        n = p["D"].numel()
        f = 1.0 + 1e-3 * torch.rand(n, generator=gen, dtype=torch.float64)
        p["D"].mul_(f.to(p["D"].device))
    torch.cuda.synchronize()
    rc = g.launch()
    assert rc in (0, 4), rc
    rc_e, ref_arr, ref_keep = M.solve_batch([dict(D=p["D"], E=p["E"]) for p in probs])
    assert rc_e == 0
    if rc == 0:
        _check(arr, keep, ref_arr, ref_keep, "new inputs")
    else:
        assert any(arr[i].info == 2 for i in range(len(arr)) if probs[i]["D"].numel() >= 500)
        assert all(arr[i].info == 0 for i in range(len(arr)) if probs[i]["D"].numel() < 500)
    g.close()
