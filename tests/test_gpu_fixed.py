"""GPU parity tests for Pedersen commitments and the fixed-base MSM (16-tooth comb and windowed tables).

Checked against the reference's golden pedersen_commit vectors, the oracle's restatements of
helios_pedersen_commit / helios_msm_fixed (oracle/hs_oracle.c), the oracle's scalar multiple for
every table entry, and the plain GPU MSM (itself pinned against the oracle and the reference).
"""
import random

import numpy as np
import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")
pytestmark = pytest.mark.gpu
H = bytes.fromhex


def _affine_entry_bytes(c, e):
    """reference affine struct (x[5], y[5]) -> compressed bytes; (0, 0) -> 0^32."""
    if not e.any():
        return bytes(32)
    jac = np.zeros(15, dtype=np.uint64)
    jac[:10] = e
    jac[10] = 1
    return GH.jac_to_bytes(c, jac)


@pytest.mark.parametrize("c", [0, 1])
def test_precompute_tables_are_multiples(c):
    n = 24
    pb = bytearray(O.gen_sswu_bases(c, 11 + c, n))
    pb[32 * 5:32 * 6] = bytes(32)  # one identity base
    pts = GH.to_jac_array(c, bytes(pb), n, rnd=random.Random(c))
    tab = hs.fixed_precompute(c, pts)
    assert tab.shape == (n, 16, 10)
    for i in range(n):
        P = bytes(pb[32 * i:32 * i + 32])
        for k in range(16):
            got = _affine_entry_bytes(c, tab[i, k])
            want = bytes(32) if P == bytes(32) else O.scalar_mul(c, (k + 1).to_bytes(32, "little"), P)
            assert got == want, (i, k)


@pytest.mark.parametrize("c", [0, 1])
@pytest.mark.parametrize("n", [1, 2, 7, 100, 1000])
def test_msm_fixed_slot_matches_oracle(c, n):
    sc = bytearray(O.gen_scalars(c, 40 + n, n, edge_frac=0.1 if n > 10 else 0.0))
    pb = bytearray(O.gen_sswu_bases(c, 50 + n, n, dup_frac=0.05 if n > 20 else 0.0))
    if n > 3:
        pb[32:64] = bytes(32)
    pts = GH.to_jac_array(c, bytes(pb), n)
    tab = hs.fixed_precompute(c, pts)
    res = hs.msm_fixed(c, GH.scalars_np(bytes(sc), n), tab)
    assert GH.jac_to_bytes(c, res) == O.msm_naive(c, bytes(sc), bytes(pb), n)


def test_msm_fixed_slot_empty():
    res = hs.msm_fixed(0, np.zeros((0, 32), np.uint8), np.zeros((0, 16, 10), np.uint64))
    assert GH.jac_to_bytes(0, res) == bytes(32)


@pytest.mark.parametrize("c", [0, 1])
def test_pedersen_slot_golden(golden, c):
    v = golden[0][O.CURVES[c] + "_point"]
    for t in v["pedersen_commit"]:
        n = t["n"]
        Hp = GH.to_jac_array(c, H(t["H"]), 1)[0]
        gens = GH.to_jac_array(c, b"".join(H(x) for x in t["generators"]), n) if n else np.zeros((0, 15), np.uint64)
        vals = GH.scalars_np(b"".join(H(x) for x in t["values"]), n) if n else np.zeros((0, 32), np.uint8)
        res = hs.pedersen_commit(c, H(t["blinding"]), Hp, vals, gens)
        assert GH.jac_to_bytes(c, res).hex() == t["result"], t["label"]


@pytest.mark.parametrize("c", [0, 1])
@pytest.mark.parametrize("n", [0, 5, 999])
def test_pedersen_slot_matches_oracle(c, n):
    """n = 999 puts the n + 1 term MSM in the reference's window-8 band."""
    sc = O.gen_scalars(c, 70 + n, n + 1)
    pb = O.gen_sswu_bases(c, 80 + n, n + 1)
    pts = GH.to_jac_array(c, pb, n + 1)
    res = hs.pedersen_commit(c, sc[:32], pts[0], GH.scalars_np(sc[32:], n) if n else np.zeros((0, 32), np.uint8),
                             pts[1:])
    want = O.pedersen(c, sc[:32], pb[:32], sc[32:], pb[32:], n)
    assert GH.jac_to_bytes(c, res) == want


@pytest.mark.parametrize("c", [0, 1])
@pytest.mark.parametrize("method", [hs.FIXED_COMB, hs.FIXED_PIPPENGER])
def test_fixed_set_rows_match_oracle(c, method):
    n, K = 300, 5
    pb = bytearray(O.gen_sswu_bases(c, 90 + c, n, dup_frac=0.02))
    pb[32 * 7:32 * 8] = bytes(32)
    pts = GH.to_jac_array(c, bytes(pb), n)
    rows = [O.gen_scalars(c, 100 + k, n, edge_frac=0.05) for k in range(K)]
    rows[1] = bytes(32 * n)  # all-zero row: identity
    fs = hs.FixedSet(c, pts)
    out, jac = fs.msm(np.frombuffer(b"".join(rows), np.uint8).reshape(K, n, 32), method=method)
    fs.close()
    for k in range(K):
        want = O.msm(c, rows[k], bytes(pb), n)
        assert out[k] == want, k
        assert GH.jac_to_bytes(c, jac[k]) == want


@pytest.mark.parametrize("method", [hs.FIXED_COMB, hs.FIXED_PIPPENGER])
def test_fixed_set_window8_band(method):
    """A 1000-base set: default rows reproduce the reference's MSM bytes, exact rows the sum."""
    c, n, K = 0, 1000, 3
    pb = O.gen_sswu_bases(c, 123, n)
    pts = GH.to_jac_array(c, pb, n)
    rows = [O.gen_scalars(c, 200 + k, n) for k in range(K)]
    S = np.frombuffer(b"".join(rows), np.uint8).reshape(K, n, 32)
    fs = hs.FixedSet(c, pts)
    compat, _ = fs.msm(S, method=method)
    exact, _ = fs.msm(S, flags=hs.EXACT, method=method)
    fs.close()
    for k in range(K):
        assert compat[k] == O.msm(c, rows[k], pb, n)
        assert exact[k] == O.msm_naive(c, rows[k], pb, n)


@pytest.mark.parametrize("c", [0, 1])
def test_pedersen_batch_4096_generators(c):
    """The commitment shape of the Pedersen workload (H + 2^12 generators), both methods against
    the plain MSM of each row."""
    m, K = 4096, 3
    rng = np.random.default_rng(c)
    pts, _ = hs.map_to_curve(c, rng.integers(0, 256, size=(m + 1) * 32, dtype=np.uint8).tobytes())
    blind = rng.integers(0, 256, size=(K, 32), dtype=np.uint8)
    vals = rng.integers(0, 256, size=(K, m, 32), dtype=np.uint8)
    blind[:, 31] &= 0x7F
    vals[:, :, 31] &= 0x7F
    fs = hs.FixedSet(c, pts)
    comb, _ = fs.pedersen(blind, vals, method=hs.FIXED_COMB)
    pip, _ = fs.pedersen(blind, vals, method=hs.FIXED_PIPPENGER)
    fs.close()
    for k in range(K):
        row = np.concatenate([blind[k:k + 1], vals[k]], axis=0)
        want = hs.msm(c, row, pts)[0]
        assert comb[k] == want and pip[k] == want, k


@pytest.mark.parametrize("flags", [0, hs.ERASE])
def test_fixed_set_large_host_upload(flags):
    """160 rows over a set of 4097 bases: 21 MB of host scalars, uploaded in two 16 MiB pieces
    through the pinned staging slots (with ERASE wiped afterwards).  Every row is checked against
    the Pippenger method over the same set, three rows against the plain MSM, and the Pedersen form
    (which always erases) against the rows."""
    c, m, K = 1, 4097, 160
    rng = np.random.default_rng(77 + flags)
    pts, _ = hs.map_to_curve(c, rng.integers(0, 256, size=m * 32, dtype=np.uint8).tobytes())
    S = rng.integers(0, 256, size=(K, m, 32), dtype=np.uint8)
    S[:, :, 31] &= 0x7F
    fs = hs.FixedSet(c, pts)
    comb, _ = fs.msm(S, flags=flags, method=hs.FIXED_COMB)
    pip, _ = fs.msm(S, flags=flags, method=hs.FIXED_PIPPENGER)
    ped, _ = fs.pedersen(S[:, 0], S[:, 1:], method=hs.FIXED_COMB)
    fs.close()
    assert comb == pip and ped == comb
    # rows 0, 127 (it straddles the two upload pieces) and 159
    for k in (0, 127, K - 1):
        assert comb[k] == hs.msm(c, S[k], pts)[0], k


@pytest.mark.parametrize("c", [0, 1])
def test_pedersen_batch_vs_compiled_reference(c):
    """The Pedersen-batch shape (H + 2^12 SSWU generators held resident, uniform values and
    blindings) against the compiled reference's helios_pedersen_commit / selene_pedersen_commit."""
    import ctypes

    if O.ref() is None:
        pytest.skip("compiled reference (oracle/_ref) not built")
    m, K = 4096, 2
    pb = GH.ref_sswu_bases(c, 4400 + c, m + 1)
    pts = GH.to_jac_array(c, pb, m + 1)
    rows = [O.gen_scalars(c, 4500 + 10 * c + k, m + 1) for k in range(K)]
    fs = hs.FixedSet(c, pts)
    S = np.frombuffer(b"".join(rows), np.uint8).reshape(K, m + 1, 32)
    comb, _ = fs.pedersen(S[:, 0], S[:, 1:], method=hs.FIXED_COMB)
    pip, _ = fs.pedersen(S[:, 0], S[:, 1:], method=hs.FIXED_PIPPENGER)
    fs.close()
    f = getattr(O.ref(), f"hsref_{O.CURVES[c]}_pedersen_bytes")
    for k in range(K):
        out = ctypes.create_string_buffer(32)
        assert f(out, rows[k][:32], pb[:32], rows[k][32:], pb[32:], m) == 0
        assert comb[k] == out.raw and pip[k] == out.raw, k


@pytest.mark.parametrize("c", [0, 1])
def test_fixed_set_windowed_tables(c):
    """A set too large for the comb's 2 MiB per base under the 9 GiB budget (5000 bases: 14-bit
    windows), with an identity base and edge scalars, against the oracle and the Pippenger method."""
    n, K = 5000, 2
    pb = bytearray(O.gen_sswu_bases(c, 700 + c, n, dup_frac=0.01))
    pb[32 * 11:32 * 12] = bytes(32)
    pts = GH.to_jac_array(c, bytes(pb), n)
    rows = [O.gen_scalars(c, 710 + 10 * c + k, n, edge_frac=0.05) for k in range(K)]
    S = np.frombuffer(b"".join(rows), np.uint8).reshape(K, n, 32)
    fs = hs.FixedSet(c, pts)
    comb, jac = fs.msm(S, flags=hs.EXACT, method=hs.FIXED_COMB)
    pip, _ = fs.msm(S, flags=hs.EXACT, method=hs.FIXED_PIPPENGER)
    fs.close()
    for k in range(K):
        want = O.msm_naive(c, rows[k], bytes(pb), n)
        assert comb[k] == want and pip[k] == want, k
        assert GH.jac_to_bytes(c, jac[k]) == want


@pytest.mark.parametrize("c", [0, 1])
@pytest.mark.parametrize("width", [16, 14, 13, 12, 10, 8, 5])
def test_fixed_set_every_table_width(c, width, monkeypatch):
    """Every resident-set layout (comb = 16, windows of 14..5 bits), forced at a small n: an identity
    base, a duplicated base and edge scalars, default and exact modes, against the oracle and the
    Pippenger method."""
    monkeypatch.setenv("HS_GPU_FIXED_WIDTH", str(width))
    n, K = 70, 3
    pb = bytearray(O.gen_sswu_bases(c, 800 + width, n))
    pb[32 * 3:32 * 4] = bytes(32)
    pb[32 * 9:32 * 10] = pb[32 * 4:32 * 5]
    pts = GH.to_jac_array(c, bytes(pb), n, rnd=random.Random(width))
    rows = [O.gen_scalars(c, 810 + k + width, n, edge_frac=0.2) for k in range(K)]
    S = np.frombuffer(b"".join(rows), np.uint8).reshape(K, n, 32)
    fs = hs.FixedSet(c, pts)
    comb, _ = fs.msm(S, method=hs.FIXED_COMB)
    comb_x, _ = fs.msm(S, flags=hs.EXACT, method=hs.FIXED_COMB)
    pip, _ = fs.msm(S, flags=hs.EXACT, method=hs.FIXED_PIPPENGER)
    fs.close()
    for k in range(K):
        want = O.msm_naive(c, rows[k], bytes(pb), n)
        assert comb[k] == O.msm(c, rows[k], bytes(pb), n) == want
        assert comb_x[k] == want and pip[k] == want, k


@pytest.mark.parametrize("c", [0, 1])
@pytest.mark.parametrize("width", [14, 13, 12, 10, 8])
def test_fixed_set_windowed_tables_every_entry(c, width, monkeypatch):
    """Every entry of the windowed tables, 1..2^(w-1) times the base: row m puts scalar m on a
    random base (entry m, or entry 2^w - m negated for m >= 2^(w-1), so m = 1..2^(w-1) reads all of
    them) and a random scalar on an identity base; each row against the oracle's m P."""
    monkeypatch.setenv("HS_GPU_FIXED_WIDTH", str(width))
    E = 1 << (width - 1)
    P = O.gen_sswu_bases(c, 900 + width, 1)
    pts = GH.to_jac_array(c, P + bytes(32), 2, rnd=random.Random(width))
    rnd = np.random.default_rng(width)
    S = np.zeros((E, 2, 32), np.uint8)
    for m in range(1, E + 1):
        S[m - 1, 0] = np.frombuffer(m.to_bytes(32, "little"), np.uint8)
    S[:, 1] = rnd.integers(0, 256, size=(E, 32), dtype=np.uint8)
    S[:, 1, 31] &= 0x7F
    fs = hs.FixedSet(c, pts)
    got, _ = fs.msm(S, flags=hs.EXACT, method=hs.FIXED_COMB)
    fs.close()
    for m in range(1, E + 1):
        assert got[m - 1] == O.scalar_mul(c, m.to_bytes(32, "little"), P), m


def test_windowed_tables_without_room_for_half_tables(monkeypatch):
    """Under a device cap that leaves room for a set's 8-bit-window tables but not for the
    temporary half tables, the tables are built entry by entry instead: same results."""
    monkeypatch.setenv("HS_GPU_FIXED_WIDTH", "8")
    c, n, K = 0, 3000, 2
    pb = O.gen_sswu_bases(c, 950, n)
    pts = GH.to_jac_array(c, pb, n)
    rows = [O.gen_scalars(c, 960 + k, n, edge_frac=0.05) for k in range(K)]
    S = np.frombuffer(b"".join(rows), np.uint8).reshape(K, n, 32)
    hs.trim()
    tables = n * 128 * 64
    hs.set_device_limit(hs.device_bytes() + tables + (2 << 20))  # the half tables need 7.1 MB
    try:
        fs = hs.FixedSet(c, pts)
    finally:
        hs.set_device_limit(0)
    got, _ = fs.msm(S, flags=hs.EXACT, method=hs.FIXED_COMB)
    fs.close()
    for k in range(K):
        assert got[k] == O.msm_naive(c, rows[k], pb, n), k


@pytest.mark.parametrize("c", [0, 1])
def test_msm_fixed_degenerate_case_is_exact(c):
    """Repeated (base, scalar) pairs, where the reference's msm_fixed meets an accumulator equal to
    the table entry and drops the pair: the drop-in returns the exact sum (the documented
    departure); the reference's own bytes are in the fixture."""
    import json
    import os

    runs = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workload_runs.json")))
    for r in [x for x in runs if x["kind"] == "fixed_degenerate" and x["curve"] == O.CURVES[c]]:
        pb = b"".join(H(p) for p in r["points"])
        sb = b"".join(H(s) for s in r["scalars"])
        n = len(r["points"])
        tab = hs.fixed_precompute(c, GH.to_jac_array(c, pb, n))
        got = GH.jac_to_bytes(c, hs.msm_fixed(c, GH.scalars_np(sb, n), tab))
        assert got.hex() == r["exact_result"] and got.hex() != r["ref_result"], r["label"]
