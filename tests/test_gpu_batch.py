"""GPU parity tests for batched MSMs (many uneven MSMs of both curves in one launch sequence):
every element must equal the oracle's result for that MSM alone, and the plain single-MSM path."""
import random

import numpy as np
import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")
pytestmark = pytest.mark.gpu


def _case(c, n, seed, edge=0.05, dup=0.05, ident=0.02):
    sc = O.gen_scalars(c, seed, n, edge_frac=edge) if n else b""
    pb = O.gen_sswu_bases(c, seed, n, dup_frac=dup, id_frac=ident) if n else b""
    pts = GH.to_jac_array(c, pb, n) if n else np.zeros((0, 15), np.uint64)
    return sc, pb, pts


def test_batch_mixed_sizes_and_curves():
    rnd = random.Random(5)
    sizes = [0, 1, 2, 17, 33, 64, 95, 200, 700, 900, 1000, 1500, 3000, 5000, 9000]
    items, want = [], []
    for k, n in enumerate(sizes * 2):
        c = k % 2
        sc, pb, pts = _case(c, n, 300 + k)
        items.append((c, GH.scalars_np(sc, n) if n else np.zeros((0, 32), np.uint8), pts))
        want.append(O.msm(c, sc, pb, n) if n else bytes(32))
    got = hs.msm_batch(items)
    for k, ((out, jac), w) in enumerate(zip(got, want)):
        assert out == w, (k, items[k][0], items[k][2].shape[0])
        assert GH.jac_to_bytes(items[k][0], jac) == w


def test_batch_random_z_points():
    """Batch ingest with random Z (lazy limbs) mixed with Z = 1 and identities across MSMs."""
    rnd = random.Random(3)
    items, want = [], []
    for k, n in enumerate((70, 3000, 5000, 1)):
        c = k % 2
        sc, pb, _ = _case(c, n, 400 + k)
        pts = GH.to_jac_array(c, pb, n, rnd=rnd, lazy=True) if k % 3 != 2 else GH.to_jac_array(c, pb, n)
        items.append((c, GH.scalars_np(sc, n), pts))
        want.append(O.msm(c, sc, pb, n))
    got = hs.msm_batch(items)
    assert [g[0] for g in got] == want


def test_batch_exact_mode_window8():
    """Exact mode in a batch: elements in the reference window-8 band equal the naive sum."""
    items, want = [], []
    for k, n in enumerate((1000, 1200, 2000)):
        c = k % 2
        sc, pb, pts = _case(c, n, 77 + k, edge=0.0)
        items.append((c, GH.scalars_np(sc, n), pts))
        want.append(O.msm_naive(c, sc, pb, n))
    got = hs.msm_batch(items, flags=hs.EXACT)
    assert [g[0] for g in got] == want


def test_batch_mix_vs_compiled_reference():
    """The batched-mix shape (sizes log-uniform in [64, 2^14], both curves interleaved, distinct
    SSWU bases per MSM, uniform scalars), every element against the compiled reference's bytes."""
    if O.ref() is None:
        pytest.skip("compiled reference (oracle/_ref) not built")
    rng = np.random.default_rng(33)
    items, want = [], []
    for k in range(24):
        c = k % 2
        n = int(np.exp(rng.uniform(np.log(64), np.log(1 << 14))))
        if k == 5:
            n = 1700  # the reference's window-8 band
        sc = O.gen_scalars(c, 3300 + k, n)
        pb = GH.ref_sswu_bases(c, 3400 + k, n)
        items.append((c, GH.scalars_np(sc, n), GH.to_jac_array(c, pb, n)))
        want.append(O.ref_msm(c, sc, pb, n))
    got = hs.msm_batch(items)
    assert [g[0] for g in got] == want


@pytest.mark.parametrize("flags", [0, hs.ERASE])
def test_batch_matches_plain_path(flags):
    """Host-buffer batch against the plain path; the 150000-point element's upload (4.8 MB of
    scalars, 18 MB of points) goes through the pinned staging slots, the points in two pieces."""
    items = []
    for k, n in enumerate((100, 4096, 12000, 150000)):
        c = k % 2
        rng = np.random.default_rng(k)
        sc = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
        sc[:, 31] &= 0x7F
        pts, _ = hs.map_to_curve(c, rng.integers(0, 256, size=n * 32, dtype=np.uint8).tobytes())
        items.append((c, sc, pts))
    got = hs.msm_batch(items, flags=flags)
    for (c, sc, pts), (out, _) in zip(items, got):
        assert out == hs.msm(c, sc, pts)[0]


def test_batch_plan_graph_replays():
    """Plan over device buffers: eager launch, then hipGraph capture and several replays with
    unrelated device work in between; every replay must reproduce the eager results."""
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(9)
    sizes = [64, 300, 1000, 2500, 8000, 16384]
    bufs, items, want = [], [], []
    for k, n in enumerate(sizes):
        c = k % 2
        sc = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
        sc[:, 31] &= 0x7F
        pts, _ = hs.map_to_curve(c, rng.integers(0, 256, size=n * 32, dtype=np.uint8).tobytes())
        ds = torch.from_numpy(sc).to(dev)
        dp = torch.from_numpy(pts.view(np.int64)).to(dev)
        do = torch.zeros(8, dtype=torch.int32, device=dev)
        bufs += [ds, dp, do]
        items.append((c, n, ds.data_ptr(), dp.data_ptr(), do.data_ptr(), None))
        want.append(hs.msm(c, sc, pts)[0])
    b = hs.Batch(items)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        b.launch(stream.cuda_stream)
        torch.cuda.synchronize()
        eager = [bufs[3 * k + 2].cpu().numpy().tobytes() for k in range(len(sizes))]
        assert eager == want
        b.capture()
        for rep in range(3):
            for k in range(len(sizes)):
                bufs[3 * k + 2].zero_()
            x = torch.randn(1 << 20, device=dev)
            (x * x).sum().item()
            b.launch(stream.cuda_stream)
            torch.cuda.synchronize()
            got = [bufs[3 * k + 2].cpu().numpy().tobytes() for k in range(len(sizes))]
            assert got == eager, rep
    b.close()
