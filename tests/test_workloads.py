"""CPU checks of the synthetic workload generator (helioselene_amd/workloads.py) and of the
compiled-reference fixtures written from it (tests/golden/workload_runs.json)."""
import json
import os

import numpy as np
import pytest

import _oracle as O
from helioselene_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def _runs(kind):
    with open(os.path.join(HERE, "golden", "workload_runs.json")) as f:
        return [r for r in json.load(f) if r["kind"] == kind]


def _affine_bytes(c, row):
    """reference Jacobian struct (uint64[15], any Z) -> compressed encoding, in Python integers"""
    m = O.FIELD[c]
    X, Y, Z = (sum(int(row[5 * k + j]) << (51 * j) for j in range(5)) % m for k in range(3))
    if Z == 0:
        return bytes(32)
    zi = pow(Z, -1, m)
    x, y = X * zi * zi % m, Y * zi * zi * zi % m
    b = bytearray(x.to_bytes(32, "little"))
    b[31] |= (y & 1) << 7
    return bytes(b)


@pytest.mark.parametrize("c", [0, 1])
def test_generators_match_vectors(golden, c):
    par = golden[0]["parameters"]
    name = O.CURVES[c]
    assert W.GEN[c][0] == int.from_bytes(bytes.fromhex(par[name + "_gx"]), "little")
    assert W.GEN[c][1] == int.from_bytes(bytes.fromhex(par[name + "_gy"]), "little")
    assert W.ORDER[c] == O.ORDER[c] and W.FIELD[c] == O.FIELD[c]


@pytest.mark.parametrize("c", [0, 1])
def test_doubling_chain_is_powers_of_two_times_g(c):
    """Point i of the chain is 2^i G (checked against the oracle's scalar multiplication), with
    Z != 1 from the second point on."""
    n = 300
    pts = W.doubling_chain(c, n)
    g = O.generator(c)
    for i in (0, 1, 2, 7, 64, 255, 299):
        k = pow(2, i, O.ORDER[c])
        assert _affine_bytes(c, pts[i]) == O.scalar_mul(c, k.to_bytes(32, "little"), g), i
    assert all(int(pts[i][10]) != 1 or int(pts[i][11:15].sum()) != 0 for i in range(1, n))


def test_doubling_chain_equals_reference_chain():
    """Same projective representatives as the reference's own chain (helios_dbl / selene_dbl)."""
    if O.ref() is None:
        pytest.skip("compiled reference (oracle/_ref) not built")
    for c in (0, 1):
        a, b = O.ref_dbl_chain(c, 400), W.doubling_chain(c, 400)
        m = O.FIELD[c]
        for i in range(400):
            for k in range(3):
                va = sum(int(a[i][5 * k + j]) << (51 * j) for j in range(5)) % m
                vb = sum(int(b[i][5 * k + j]) << (51 * j) for j in range(5)) % m
                assert va == vb, (c, i, k)


def test_fcmpp_fixture_matches_oracle():
    """The committed reference bytes of the FCMP++ shapes against the oracle's MSM over the same
    points (encoded) and scalars."""
    runs = _runs("fcmpp")
    assert len(runs) == 8
    for r in runs:
        c, n = O.CURVES.index(r["curve"]), r["n"]
        pts = W.doubling_chain(c, n)
        pb = b"".join(_affine_bytes(c, pts[i]) for i in range(n))
        sc = W.fcmpp_scalars(n).tobytes()
        assert O.msm(c, sc, pb, n).hex() == r["ref_result"], (r["curve"], n)
        assert O.msm_naive(c, sc, pb, n).hex() == r["ref_result"]


def test_tree_hash_fixture_matches_oracle():
    """The committed reference bytes of the FCMP++ tree-hash nodes (Selene 228 and 38, Helios 18
    doubling-chain generators, seeded uniform scalars) against the oracle's MSM and naive sum."""
    runs = _runs("tree_hash")
    assert sorted({(r["curve"], r["n"]) for r in runs}) == [("helios", 18), ("selene", 38), ("selene", 228)]
    for c, n, name in W.TREE_HASH_SHAPES:
        rs = sorted([r for r in runs if r["curve"] == O.CURVES[c] and r["n"] == n], key=lambda r: r["index"])
        assert [r["index"] for r in rs] == list(range(len(rs))) and rs[0]["layer"] == name
        pts = W.doubling_chain(c, n)
        pb = b"".join(_affine_bytes(c, pts[i]) for i in range(n))
        sc = W.tree_layer_scalars(n, len(rs), rs[0]["seed"])
        for h, r in enumerate(rs):
            assert O.msm(c, sc[h].tobytes(), pb, n).hex() == r["ref_result"], (name, h)
            assert O.msm_naive(c, sc[h].tobytes(), pb, n).hex() == r["ref_result"], (name, h)


@pytest.mark.parametrize("c", [0, 1])
def test_plain_inputs_shape(c):
    n = 1 << 14
    w = W.plain_inputs(c, n, 5)
    w2 = W.plain_inputs(c, n, 5)
    assert all((w[k] == w2[k]).all() for k in w)
    assert w["u"].shape == (n, 32) and w["scalars"].shape == (n, 32)
    assert len(w["dup_dst"]) == n // 200 and len(w["id_dst"]) == n // 1000
    assert len(set(w["dup_dst"].tolist()) | set(w["id_dst"].tolist())) == n // 200 + n // 1000
    vals = [int.from_bytes(w["scalars"][i].tobytes(), "little") for i in range(n)]
    assert max(vals) < O.ORDER[c]
    # 1% edge scalars: 0, 1, order - 1 and small values, a quarter each
    k = n // 100
    per = [len(range(r, k, 4)) for r in range(4)]
    assert vals.count(0) >= per[0] and vals.count(1) >= per[1] and vals.count(O.ORDER[c] - 1) == per[2]
    assert sum(1 for v in vals if 2 <= v < 256) >= per[3]
    pts = np.zeros((n, 15), np.uint64)
    pts[:, 0] = np.arange(n)
    W.apply_dups_and_identities(pts, w)
    assert (pts[w["dup_dst"], 0] == w["dup_src"]).all() or set(w["dup_src"]) & set(w["id_dst"])
    assert (pts[w["id_dst"]] == W.IDENTITY_STRUCT).all()


@pytest.mark.slow
@pytest.mark.parametrize("c", [0, 1])
def test_plain_fixture_2e16_matches_oracle(c):
    """The reference's 2^16 plain-workload bytes against the oracle's own MSM over the oracle's
    SSWU images of the same seeds."""
    r = [x for x in _runs("plain") if x["curve"] == O.CURVES[c] and x["n"] == 1 << 16][0]
    n = r["n"]
    w = W.plain_inputs(c, n, r["seed"])
    pb = bytearray(O.map_to_curve(c, w["u"].tobytes()))
    orig = bytes(pb)
    for d, s in zip(w["dup_dst"], w["dup_src"]):
        pb[32 * d:32 * d + 32] = orig[32 * s:32 * s + 32]
    for d in w["id_dst"]:
        pb[32 * d:32 * d + 32] = bytes(32)
    assert O.msm(c, w["scalars"].tobytes(), bytes(pb), n).hex() == r["ref_result"]


def test_fixed_degenerate_fixture():
    """The reference's msm_fixed on repeated (base, scalar) pairs differs from the exact sum (its
    madd-2007-bl meets an accumulator equal to the table entry); the exact side is the oracle's."""
    runs = _runs("fixed_degenerate")
    assert len(runs) == 4
    for r in runs:
        c = O.CURVES.index(r["curve"])
        pb = b"".join(bytes.fromhex(p) for p in r["points"])
        sb = b"".join(bytes.fromhex(s) for s in r["scalars"])
        n = len(r["points"])
        assert O.msm_naive(c, sb, pb, n).hex() == r["exact_result"]
        assert r["ref_result"] != r["exact_result"]
