"""GPU tests of the boundary's failure and concurrency semantics: a host-forced allocation refusal
(the library's device-byte cap) seen by the status form, the void slot answering through the
registered CPU fallback, no device memory leaked by failing calls, concurrent host callers, the
erase flag, and batch plans split into several groups."""
import random
import threading

import numpy as np
import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")
pytestmark = pytest.mark.gpu


def _inputs(c, n, seed):
    sc = O.gen_scalars(c, seed, n, edge_frac=0.05)
    pb = O.gen_sswu_bases(c, seed, n, dup_frac=0.05, id_frac=0.02)
    return sc, pb, GH.to_jac_array(c, pb, n, rnd=random.Random(seed), lazy=True)


@pytest.mark.parametrize("c", [0, 1])
def test_allocation_refusal_status_fallback_and_recovery(c):
    n = 300
    sc, pb, pts = _inputs(c, n, 11 + c)
    want = O.msm(c, sc, pb, n)
    s_np = GH.scalars_np(sc, n)
    calls = []
    hs.synchronize()
    hs.trim()
    held = hs.device_bytes()
    hs.set_device_limit(held + 4096)  # nothing new fits
    try:
        with pytest.raises(hs.HsGpuError, match="-3"):
            hs.msm(c, s_np, pts)
        with pytest.raises(hs.HsGpuError, match="-3"):
            hs.msm_batch([(c, s_np, pts), (1 - c, s_np, GH.to_jac_array(1 - c, O.gen_sswu_bases(1 - c, 2, n), n))])
        assert hs.device_bytes() <= held + 4096  # failed calls keep nothing
        hs.set_cpu_fallback(c, GH.oracle_fallback(c, calls))
        before = hs.fallback_count()
        jac = hs.msm_slot(c, s_np, pts)
        assert calls == [n] and hs.fallback_count() == before + 1
        assert GH.jac_to_bytes(c, jac) == want
    finally:
        hs.set_device_limit(0)
    # the device path works again, without the fallback
    jac = hs.msm_slot(c, s_np, pts)
    assert GH.jac_to_bytes(c, jac) == want and calls == [n]
    hs.set_cpu_fallback(c)


def test_concurrent_host_callers():
    """Eight threads issue host-pointer MSMs at once (each takes a staging slot of its own); every
    result equals the sequential one."""
    jobs = []
    for k in range(16):
        c = k % 2
        n = [100, 1500, 4000, 20000][k % 4]
        sc, pb, pts = _inputs(c, n, 100 + k)
        jobs.append((c, GH.scalars_np(sc, n), pts))
    seq = [hs.msm(c, s, p)[0] for c, s, p in jobs]
    got = [None] * len(jobs)

    def worker(idx):
        for i in idx:
            c, s, p = jobs[i]
            got[i] = hs.msm(c, s, p)[0]

    th = [threading.Thread(target=worker, args=(list(range(j, len(jobs), 8)),)) for j in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert got == seq


@pytest.mark.parametrize("c", [0, 1])
def test_erase_flag_keeps_results(c):
    n = 5000
    sc, pb, pts = _inputs(c, n, 40 + c)
    s_np = GH.scalars_np(sc, n)
    a, _ = hs.msm(c, s_np, pts)
    b, _ = hs.msm(c, s_np, pts, flags=hs.ERASE)
    e, _ = hs.msm(c, s_np, pts, flags=hs.ERASE | hs.EXACT)
    assert a == b == O.msm(c, sc, pb, n)
    assert e == O.msm_naive(c, sc, pb, n)


def test_batch_groups_split_at_the_limit(monkeypatch):
    """With the group limit lowered to a few thousand entry slots, a batch becomes many groups of the
    same (curve, width); results equal the one-group plan's."""
    rng = np.random.default_rng(9)
    items = []
    for k in range(24):
        c = k % 2
        n = int(rng.integers(200, 3000))
        sc, pb, pts = _inputs(c, n, 300 + k)
        items.append((c, GH.scalars_np(sc, n), pts))
    one = hs.msm_batch(items)
    monkeypatch.setenv("HS_GPU_BATCH_GROUP_LIMIT", "40000")
    many = hs.msm_batch(items)
    assert [o for o, _ in one] == [o for o, _ in many]
    assert [o for o, _ in one] == [hs.msm(c, s, p)[0] for c, s, p in items]
