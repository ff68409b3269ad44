"""GPU tests of the serving loop (hs_gpu_server_*): continuous batching of requests submitted
from several threads, admission by device footprint, and the property the serving design must
keep: every request's result is the same alone (plain MSM) and in the mix."""
import threading
import time

import numpy as np
import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")
pytestmark = pytest.mark.gpu


def _requests(count, seed, lo=32, hi=6000):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        c = int(rng.integers(0, 2))
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        if k == 3:
            n = 0
        if k == 5:
            n = 1500  # the reference's window-8 band
        sc = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
        sc[:, 31] &= 0x7F
        pts = hs.map_to_curve(c, rng.integers(0, 256, size=n * 32, dtype=np.uint8).tobytes())[0] if n else \
            np.zeros((0, 15), np.uint64)
        out.append((c, sc, pts))
    return out


def test_results_equal_alone_and_in_mix():
    reqs = _requests(48, 1)
    alone = [hs.msm(c, sc, pts)[0] if pts.shape[0] else bytes(32) for c, sc, pts in reqs]
    srv = hs.Server(lanes=3, max_batch=16)
    results = [None] * len(reqs)

    def client(idx):
        rng = np.random.default_rng(100 + idx[0])
        tickets = []
        for i in idx:
            c, sc, pts = reqs[i]
            tickets.append((i, srv.submit(c, sc, pts)))
            time.sleep(float(rng.uniform(0, 0.002)))
        for i, t in tickets:
            results[i] = srv.wait(t)

    th = [threading.Thread(target=client, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    st = srv.stats()
    srv.close()
    for i, (r, a) in enumerate(zip(results, alone)):
        assert r is not None and r[0] == a, i
    assert st["completed"] == len(reqs)
    assert st["p50_us"] <= st["p99_us"] <= st["max_us"]
    assert st["batches"] >= 1


def test_results_match_oracle_small():
    srv = hs.Server(lanes=2)
    cases = []
    for k, n in enumerate((1, 17, 300, 1000)):
        c = k % 2
        sc = O.gen_scalars(c, 500 + k, n)
        pb = O.gen_sswu_bases(c, 600 + k, n, dup_frac=0.05, id_frac=0.02)
        cases.append((c, sc, pb, n, srv.submit(c, GH.scalars_np(sc, n), GH.to_jac_array(c, pb, n))))
    for c, sc, pb, n, t in cases:
        assert srv.wait(t)[0] == O.msm(c, sc, pb, n), n
    srv.close()


def test_admission_by_footprint():
    """A small device budget: requests still complete (in smaller batches); one whose footprint
    exceeds a lane fails with ENOMEM without blocking the others."""
    reqs = _requests(12, 2, lo=64, hi=3000)
    srv = hs.Server(device_bytes=48 << 20, lanes=2, max_batch=64)
    ts = [srv.submit(c, sc, pts) for c, sc, pts in reqs]
    rng = np.random.default_rng(3)
    n_big = 1 << 18
    big_sc = rng.integers(0, 256, size=(n_big, 32), dtype=np.uint8)
    big_sc[:, 31] &= 0x7F
    big_pts = np.tile(reqs[0][2][:1], (n_big, 1)) if reqs[0][2].shape[0] else None
    big = srv.submit(0, big_sc, big_pts)
    ts2 = [srv.submit(c, sc, pts) for c, sc, pts in reqs[:4]]
    with pytest.raises(hs.HsGpuError):
        srv.wait(big)
    got = [srv.wait(t)[0] for t in ts] + [srv.wait(t)[0] for t in ts2]
    st = srv.stats()
    srv.close()
    want = [hs.msm(c, sc, pts)[0] if pts.shape[0] else bytes(32) for c, sc, pts in reqs]
    assert got == want + want[:4]
    assert st["batches"] >= 2


def test_device_inputs():
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    reqs = _requests(6, 4, lo=100, hi=4000)
    srv = hs.Server(lanes=2)
    bufs, ts = [], []
    for c, sc, pts in reqs:
        n = pts.shape[0]
        if n == 0:
            continue
        ds = torch.from_numpy(sc).to(dev)
        dp = torch.from_numpy(pts.view(np.int64)).to(dev)
        bufs.append((ds, dp))
        torch.cuda.synchronize()
        ts.append((hs.msm(c, sc, pts)[0], srv.submit(c, d_scalars=ds.data_ptr(), d_points=dp.data_ptr(), n=n)))
    for want, t in ts:
        assert srv.wait(t)[0] == want
    srv.close()


def test_full_size_range_staggered():
    """Requests log-uniform over the whole served range [32, 2^18] (the plain workload's
    generator), submitted from two threads with staggered gaps while others run: each result
    equals the plain MSM of the same request, and the three fixture requests equal the compiled
    reference's bytes (tests/golden/workload_runs.json)."""
    import json
    import os

    from helioselene_amd import workloads as W

    runs = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workload_runs.json")))
    fixed = [r for r in runs if r["kind"] == "plain" and r["seed"] in (7, 8, 9)]
    rng = np.random.default_rng(21)
    reqs = []
    for r in fixed:
        c = O.CURVES.index(r["curve"])
        w = W.plain_inputs(c, r["n"], r["seed"])
        pts = W.apply_dups_and_identities(hs.map_to_curve(c, w["u"].tobytes())[0], w)
        reqs.append((c, w["scalars"], pts, r["ref_result"]))
    for k in range(21):
        c = int(rng.integers(0, 2))
        n = int(np.exp(rng.uniform(np.log(32), np.log(1 << 18))))
        w = W.plain_inputs(c, n, 1000 + k)
        pts = W.apply_dups_and_identities(hs.map_to_curve(c, w["u"].tobytes())[0], w)
        reqs.append((c, w["scalars"], pts, None))
    order = rng.permutation(len(reqs))
    alone = [hs.msm(c, sc, pts)[0] for c, sc, pts, _ in reqs]
    for (_, _, _, ref), a in zip(reqs, alone):
        if ref is not None:
            assert a.hex() == ref
    srv = hs.Server(device_bytes=12 << 30, lanes=4)
    results = [None] * len(reqs)

    def client(part, seed):
        g = np.random.default_rng(seed)
        tickets = []
        for i in part:
            c, sc, pts, _ = reqs[i]
            tickets.append((i, srv.submit(c, sc, pts)))
            time.sleep(float(g.uniform(0, 0.004)))
        for i, t in tickets:
            results[i] = srv.wait(t)[0]

    th = [threading.Thread(target=client, args=(list(order[j::2]), 50 + j)) for j in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    st = srv.stats()
    srv.close()
    assert results == alone
    assert st["completed"] == len(reqs)


def test_tuned_groups_with_mixed_run_lengths_fit_their_lane(tmp_path):
    """A tune table whose entries for one window width differ in run length L and segment size G
    (sizes 2^10 and 2^11 both at c = 8): a lane filled close to its budget with both sizes still
    plans its batch inside the admitted footprint, and every result equals the plain MSM's."""
    f = tmp_path / "tune.txt"
    f.write_text("1 0 10 8 4 4 0.1\n1 0 11 8 64 16 0.1\n1 1 10 8 4 4 0.1\n1 1 11 8 64 16 0.1\n")
    hs.tune_load(str(f))
    try:
        rng = np.random.default_rng(31)
        reqs = []
        for k in range(24):
            c = k % 2
            n = 1024 if k % 3 else 2048
            sc = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
            sc[:, 31] &= 0x7F
            pts = hs.map_to_curve(c, rng.integers(0, 256, size=n * 32, dtype=np.uint8).tobytes())[0]
            reqs.append((c, sc, pts))
        srv = hs.Server(device_bytes=2 * (10 << 20), lanes=2, max_batch=64)
        ts = [srv.submit(c, sc, pts) for c, sc, pts in reqs]
        got = [srv.wait(t)[0] for t in ts]
        srv.close()
        assert got == [hs.msm(c, sc, pts)[0] for c, sc, pts in reqs]
    finally:
        hs.tune_clear()
