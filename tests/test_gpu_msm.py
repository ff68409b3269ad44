"""GPU parity tests for the plain MSM (calls through the C ABI of include/hs_gpu_msm.h).

Bar: bit-exact 32-byte encodings.  Checked against the reference's golden vectors, the compiled
reference's results on seeded inputs (committed fixture), the oracle at mid sizes, and at the
full 2^16..2^20 sizes against a closed form (bases on an arithmetic progression of multiples of
G, so the MSM equals (sum s_i k_i mod order) * G).
"""
import random

import numpy as np
import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")

pytestmark = pytest.mark.gpu
H = bytes.fromhex


def run(c, sc, pts_jac, flags=0):
    out, jac = hs.msm(c, GH.scalars_np(sc, pts_jac.shape[0]) if not isinstance(sc, np.ndarray) else sc, pts_jac, flags)
    return out, jac


@pytest.mark.parametrize("c", [0, 1])
def test_golden_msm_vectors(golden, c):
    v = golden[0][O.CURVES[c] + "_point"]
    for t in v["msm"]:
        n = t["n"]
        sc = b"".join(H(x) for x in t["scalars"])
        pts = GH.to_jac_array(c, b"".join(H(x) for x in t["points"]), n)
        out, jac = run(c, sc, pts)
        print(O.CURVES[c], t["label"], flush=True)
        assert out.hex() == t["result"], t["label"]
        assert GH.jac_to_bytes(c, jac).hex() == t["result"]


@pytest.mark.parametrize("c", [0, 1])
def test_golden_pedersen_as_msm(golden, c):
    """pedersen_commit vectors = one MSM of n+1 terms (helios_pedersen.h:55-86)."""
    v = golden[0][O.CURVES[c] + "_point"]
    for t in v["pedersen_commit"]:
        n = t["n"]
        sc = H(t["blinding"]) + b"".join(H(x) for x in t["values"])
        pts = GH.to_jac_array(c, H(t["H"]) + b"".join(H(x) for x in t["generators"]), n + 1)
        out, _ = run(c, sc, pts)
        assert out.hex() == t["result"], t["label"]


def test_reference_runs_fixture(golden):
    """Compiled-reference results on seeded SSWU bases (1% duplicates, 0.5% identities) in every
    window band, including the window-8 band where the reference drops a recode carry."""
    for r in golden[1]:
        c = O.CURVES.index(r["curve"])
        n = r["n"]
        if r["bases"] != "sswu":
            continue
        sc = O.gen_scalars(c, r["seed"], n)
        pb = O.gen_sswu_bases(c, r["seed"], n, r["dup_frac"], r["id_frac"])
        pts = GH.to_jac_array(c, pb, n)
        print(r["curve"], n, flush=True)
        out, _ = run(c, sc, pts)
        assert out.hex() == r["ref_result"], (r["curve"], n)
        if r["naive_result"] is not None:
            out_exact, _ = run(c, sc, pts, flags=hs.EXACT)
            assert out_exact.hex() == r["naive_result"], (r["curve"], n, "exact")


@pytest.mark.slow
def test_reference_run_4096_kG(golden):
    """The n = 4096 Helios workload on k*G points (every Z != 1 after re-randomisation)."""
    r = [x for x in golden[1] if x["bases"] == "kG"][0]
    n = r["n"]
    sc = O.gen_scalars(0, r["seed"], n)
    pb = O.gen_kG_points(0, r["seed"], n)
    pts = GH.to_jac_array(0, pb, n, rnd=random.Random(3), lazy=True)
    out, _ = run(0, sc, pts)
    assert out.hex() == r["ref_result"]


@pytest.mark.parametrize("c", [0, 1])
def test_mixed_z_ingest(c):
    """Several ingest blocks with every kind of Z mixed within one thread's points (random Z with
    lazy limbs, Z = 1, identities): the shared batched inversion must give each point its own
    affine form."""
    n = 6000
    rnd = random.Random(77 + c)
    sc = O.gen_scalars(c, 17, n, edge_frac=0.01)
    pb = O.gen_sswu_bases(c, 19, n, dup_frac=0.01, id_frac=0.02)
    rz = GH.to_jac_array(c, pb, n, rnd=rnd, lazy=True)
    z1 = GH.to_jac_array(c, pb, n)
    pick = np.array([rnd.random() < 0.4 for _ in range(n)])
    pts = np.where(pick[:, None], z1, rz)
    out, _ = run(c, sc, pts)
    assert out == O.msm(c, sc, pb, n)
    allz1, _ = run(c, sc, z1)
    assert allz1 == out


@pytest.mark.parametrize("c", [0, 1])
@pytest.mark.parametrize("n", [1, 2, 3, 16, 17, 31, 100, 700])
def test_random_vs_oracle(c, n):
    """Seeded scalars with edge values (0, 1, order-1, 8-bit), bases with duplicates and
    identities, random Z and unnormalised limbs; compared with the oracle's reference-algorithm
    MSM and, in exact mode, with the naive sum."""
    rnd = random.Random(1000 * c + n)
    sc = O.gen_scalars(c, 7 + n, n, edge_frac=0.1)
    pb = O.gen_sswu_bases(c, 7 + n, n, dup_frac=0.1, id_frac=0.05)
    pts = GH.to_jac_array(c, pb, n, rnd=rnd, lazy=True)
    out, jac = run(c, sc, pts)
    assert out == O.msm(c, sc, pb, n)
    assert GH.jac_to_bytes(c, jac) == out
    out2, _ = run(c, sc, pts, flags=hs.EXACT)
    assert out2 == O.msm_naive(c, sc, pb, n)


@pytest.mark.parametrize("c", [0, 1])
def test_edge_cases(c):
    """All-zero scalars, all-identity points, P and -P cancelling, equal points in one bucket,
    non-canonical scalars (>= order, >= 2^255)."""
    g = O.generator(c)
    order = O.ORDER[c]
    n = 40
    pb = g * n
    pts = GH.to_jac_array(c, pb, n)
    out, _ = run(c, bytes(32 * n), pts)
    assert out == bytes(32)
    pts_id = GH.to_jac_array(c, bytes(32 * n), n)
    out, _ = run(c, O.gen_scalars(c, 5, n), pts_id)
    assert out == bytes(32)
    # s*G + (order - s)*G = O, all in the same buckets
    s = 123456789
    sc = (s.to_bytes(32, "little") + (order - s).to_bytes(32, "little")) * (n // 2)
    out, _ = run(c, sc, pts)
    assert out == bytes(32)
    # 40 equal (point, scalar) pairs -> doubling inside buckets
    sc = (77).to_bytes(32, "little") * n
    out, _ = run(c, sc, pts)
    assert out == O.scalar_mul(c, (77 * n).to_bytes(32, "little"), g)
    # the same with negative digits: the doubling path gets a lazily negated base
    sc = (order - 77).to_bytes(32, "little") * n
    out, _ = run(c, sc, pts)
    assert out == O.scalar_mul(c, ((order - 77) * n % order).to_bytes(32, "little"), g)
    # non-canonical scalars: order + k and 2^256 - 1
    big = [(order + 5), (1 << 256) - 1, (1 << 255) + 3, order]
    sc = b"".join(v.to_bytes(32, "little") for v in big) * (n // 4)
    out, _ = run(c, sc, pts, flags=hs.EXACT)
    assert out == O.msm_naive(c, sc, pb, n)
    assert out == O.msm(c, sc, pb, n)  # n = 40: window 5, no carry drop in the reference


@pytest.mark.parametrize("c", [0, 1])
def test_window8_band_compat_and_exact(c):
    """n = 1500 (reference window 8): default mode reproduces the reference's bytes, exact mode
    the naive sum; scalars are chosen so that several drop the final carry."""
    n = 1500
    rnd = random.Random(99 + c)
    order = O.ORDER[c]
    vals = [rnd.randrange(order) for _ in range(n)]
    for i in range(0, n, 50):
        vals[i] = (0x7F << 248) | (0xFF << 240) | rnd.getrandbits(240)  # top digit overflows
    sc = b"".join((v % order).to_bytes(32, "little") for v in vals)
    pb = O.gen_sswu_bases(c, 5, n)
    pts = GH.to_jac_array(c, pb, n)
    assert any(O.drops_carry(sc[32 * i:32 * i + 32], 8) for i in range(n))
    out, _ = run(c, sc, pts)
    assert out == O.msm(c, sc, pb, n)
    ref = O.ref_msm(c, sc, pb, n)
    if ref is not None:
        assert out == ref
    out2, _ = run(c, sc, pts, flags=hs.EXACT)
    assert out2 == O.msm_naive(c, sc, pb, n)
    assert out2 != out


@pytest.mark.parametrize("c", [0, 1])
@pytest.mark.parametrize("logn", [16, 18, 20])
def test_full_size_closed_form(c, logn):
    """Full-size MSMs (2^16, 2^18, 2^20) against the closed form over progression bases, with 1%
    edge scalars."""
    n = 1 << logn
    pts_xy, a, d = GH.progression_bases(c, n, seed=logn * 10 + c)
    pts = hs.jacobian_array(pts_xy)
    sc = O.gen_scalars(c, logn, n, edge_frac=0.01) if logn <= 16 else _fast_scalars(c, n, logn)
    print(O.CURVES[c], n, flush=True)
    out, _ = run(c, sc, pts)
    assert out == GH.progression_expected(c, sc, n, a, d)


def _fast_scalars(c, n, seed):
    rnd = np.random.default_rng(seed)
    raw = rnd.integers(0, 256, size=(n, 32), dtype=np.uint8)
    raw[:, 31] &= 0x7F  # < 2^255, not necessarily < order: the library reduces
    k = max(1, n // 100)
    idx = rnd.choice(n, size=k, replace=False)
    for j, i in enumerate(idx):
        v = (0, 1, O.ORDER[c] - 1, int(raw[i, 0]) | 1)[j % 4]
        raw[i] = np.frombuffer(v.to_bytes(32, "little"), dtype=np.uint8)
    return raw.tobytes()


@pytest.mark.parametrize("c,logn", [(0, 16), (1, 16), (1, 18)])
def test_config2_shape_vs_compiled_reference(c, logn):
    """The plain-MSM workload shape (SSWU bases with 0.5% duplicates and 0.1% identities, 1% edge
    scalars) at n = 2^16 and 2^18, against the compiled reference's own MSM bytes on the same
    inputs (bytes ABI: the points are decoded on the device)."""
    if O.ref() is None:
        pytest.skip("compiled reference (oracle/_ref) not built")
    n = 1 << logn
    sc = O.gen_scalars(c, 2018 + c, n, edge_frac=0.01)
    pb = GH.ref_sswu_bases(c, 2018 + c, n, dup_frac=0.005, id_frac=0.001)
    want = O.ref_msm(c, sc, pb, n)
    assert hs.msm_bytes(c, sc, pb) == want
    if logn >= 17:
        # 8 MB inputs: the pinned staging path, wiped afterwards with ERASE
        assert hs.msm_bytes(c, sc, pb, flags=hs.ERASE) == want


@pytest.mark.parametrize("c", [0, 1])
def test_chunked_msm(c, monkeypatch):
    """MSMs of 2^24 points or more run as chunks whose sums are added on the device; the chunk
    size is lowered here so the path runs at test sizes.  n = 3000 in chunks of 1024 would hit the
    reference's window-8 band per chunk but not as a whole: the band follows the whole MSM."""
    monkeypatch.setenv("HS_GPU_CHUNK", "1024")
    for n in (3000, 2000, 1025):
        sc = O.gen_scalars(c, 60 + n, n, edge_frac=0.02)
        pb = O.gen_sswu_bases(c, 61 + n, n, dup_frac=0.02, id_frac=0.01)
        pts = GH.to_jac_array(c, pb, n)
        out, jac = run(c, sc, pts)
        assert out == O.msm(c, sc, pb, n), n
        assert GH.jac_to_bytes(c, jac) == out
        out2, _ = run(c, sc, pts, flags=hs.EXACT)
        assert out2 == O.msm_naive(c, sc, pb, n), n


def test_drop_in_slot_and_n0():
    g = O.generator(0)
    pts = GH.to_jac_array(0, g * 3, 3)
    sc = (5).to_bytes(32, "little") * 3
    jac = hs.msm_slot(0, GH.scalars_np(sc, 3), pts)
    assert GH.jac_to_bytes(0, jac) == O.scalar_mul(0, (15).to_bytes(32, "little"), g)
    out, jac = hs.msm(0, np.zeros((0, 32), np.uint8), np.zeros((0, 15), np.uint64))
    assert out == bytes(32) and GH.jac_to_bytes(0, jac) == bytes(32)


@pytest.mark.parametrize("K", [1, 2, 4])
@pytest.mark.parametrize("c", [0, 1])
def test_affine_bucket_method(c, K, monkeypatch):
    """The batched-affine bucket kernel (K chains per lane, one wavefront-wide batch inversion per
    step), forced for the plain path: duplicated points (doubling in a bucket), P and -P in one
    bucket (cancellation), identities, edge scalars, random Z; the window-8 band both ways; and the
    2^16 plain workload against the compiled reference's bytes."""
    import json
    import os

    from helioselene_amd import workloads as W

    monkeypatch.setenv("HS_GPU_AFFINE_K", str(K))
    g = O.generator(c)
    order = O.ORDER[c]
    n = 40
    pts = GH.to_jac_array(c, g * n, n)
    sc = (77).to_bytes(32, "little") * n
    assert run(c, sc, pts)[0] == O.scalar_mul(c, (77 * n).to_bytes(32, "little"), g)
    s = 123456789
    sc = (s.to_bytes(32, "little") + (order - s).to_bytes(32, "little")) * (n // 2)
    assert run(c, sc, pts)[0] == bytes(32)
    for n in (3, 100, 700, 1500, 5000):
        rnd = random.Random(300 * c + n)
        sc = O.gen_scalars(c, 90 + n, n, edge_frac=0.1)
        pb = O.gen_sswu_bases(c, 91 + n, n, dup_frac=0.1, id_frac=0.05)
        pts = GH.to_jac_array(c, pb, n, rnd=rnd, lazy=True)
        assert run(c, sc, pts)[0] == O.msm(c, sc, pb, n), n
        assert run(c, sc, pts, flags=hs.EXACT)[0] == O.msm_naive(c, sc, pb, n), n
    runs = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workload_runs.json")))
    r = [x for x in runs if x["kind"] == "plain" and x["curve"] == O.CURVES[c] and x["n"] == 1 << 16][0]
    w = W.plain_inputs(c, r["n"], r["seed"])
    pts = W.apply_dups_and_identities(hs.map_to_curve(c, w["u"].tobytes())[0], w)
    assert hs.msm(c, w["scalars"], pts)[0].hex() == r["ref_result"]
