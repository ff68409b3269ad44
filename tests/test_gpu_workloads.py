"""GPU parity on the benchmark's own workloads, against the compiled reference's bytes.

* The plain-MSM workload (SSWU bases with 0.5% copies and 0.1% identities, 1% edge scalars) at
  n = 2^16, 2^18 and 2^20 on both curves, generated by helioselene_amd/workloads.py exactly as
  bench.py generates it, with the bases mapped by the library's own SSWU kernel: compared with the
  reference's helios_msm_vartime / selene_msm_vartime bytes on the reference's own SSWU images of
  the same seeds (tests/golden/workload_runs.json, written by make_workload_fixtures.py).
* The library's SSWU map itself: the reference's map_to_curve vectors and, on 2^15 seeded inputs
  per curve plus edge inputs, the compiled reference's map (helios/src/x64/helios_map_to_curve.cpp:98-198,
  selene/src/x64/selene_map_to_curve.cpp).
* The FCMP++ verification shapes (src/benchmark_fcmpp.cpp:347-365) on doubling-chain bases (Z != 1),
  one MSM at a time and as one batch.
"""
import json
import os

import numpy as np
import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")
from helioselene_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _runs(kind):
    with open(os.path.join(HERE, "golden", "workload_runs.json")) as f:
        return [r for r in json.load(f) if r["kind"] == kind]


def plain_points(c, w):
    """The workload's bases as reference Jacobian structs: the library's SSWU map, then the copies
    and identities."""
    pts, _ = hs.map_to_curve(c, w["u"].tobytes())
    return W.apply_dups_and_identities(pts, w)


@pytest.mark.parametrize("case", _runs("plain")[:6], ids=lambda r: f"{r['curve']}-n{r['n']}")
def test_plain_workload_vs_reference(case):
    c, n = O.CURVES.index(case["curve"]), case["n"]
    print(case["curve"], n, flush=True)
    w = W.plain_inputs(c, n, case["seed"])
    out, jac = hs.msm(c, w["scalars"], plain_points(c, w))
    assert out.hex() == case["ref_result"]
    assert GH.jac_to_bytes(c, jac) == out


@pytest.mark.parametrize("c", [0, 1])
def test_sswu_golden_vectors(golden, c):
    v = golden[0][O.CURVES[c] + "_point"]
    for t in v["map_to_curve_single"]:
        _, enc = hs.map_to_curve(c, bytes.fromhex(t["u"]), want_points=False, want_bytes=True)
        assert enc.hex() == t["result"], t["label"]
    for t in v["map_to_curve_double"]:
        # hash_to_curve = map(u0) + map(u1): both maps on the GPU, the sum by the GPU MSM
        pts, _ = hs.map_to_curve(c, bytes.fromhex(t["u0"]) + bytes.fromhex(t["u1"]))
        out, _ = hs.msm(c, hs.scalars_array([1, 1]), pts)
        assert out.hex() == t["result"], t["label"]


@pytest.mark.parametrize("c", [0, 1])
def test_sswu_vs_compiled_reference(c):
    """2^15 seeded inputs plus edge inputs (0, 1, m - 1, m, 2^255 - 1, bit 255 set, all ones):
    the encoded images and the Jacobian structs (Z = 1) against the reference's own map."""
    if O.ref() is None:
        pytest.skip("compiled reference (oracle/_ref) not built")
    m = O.FIELD[c]
    n = 1 << 15
    u = bytearray(np.random.default_rng(4242 + c).integers(0, 256, size=n * 32, dtype=np.uint8).tobytes())
    edge = [0, 1, 2, m - 1, m, m + 1, (1 << 255) - 1, (1 << 256) - 1, (1 << 255) | 5, m - 2]
    for k, v in enumerate(edge):
        u[32 * k:32 * k + 32] = v.to_bytes(32, "little")
    u = bytes(u)
    pts, enc = hs.map_to_curve(c, u, want_points=True, want_bytes=True)
    want = O.ref_map_to_curve_bytes(c, u)
    assert enc == want
    # the struct output denotes the same points (Z = 1, canonical coordinates)
    for i in list(range(len(edge))) + list(range(n - 64, n)):
        assert GH.jac_to_bytes(c, pts[i]) == want[32 * i:32 * i + 32], i


def test_fcmpp_shapes_single_and_batched():
    """Selene 522 + 80 b and Helios 265 + 80 b, b = 1, 2, 4, 10, on doubling-chain points whose
    Z != 1 (every point goes through the ingest's inversion)."""
    runs = {(r["curve"], r["n"]): r["ref_result"] for r in _runs("fcmpp")}
    items, want = [], []
    for c, n, b in W.fcmpp_shapes():
        pts = W.doubling_chain(c, n)
        sc = W.fcmpp_scalars(n)
        ref = runs[(O.CURVES[c], n)]
        out, _ = hs.msm(c, sc, pts)
        assert out.hex() == ref, (O.CURVES[c], n)
        items.append((c, sc, pts))
        want.append(ref)
    res = hs.msm_batch(items)
    assert [o.hex() for o, _ in res] == want


def test_fcmpp_tree_hash_layers():
    """FCMP++ tree hashes (src/benchmark_fcmpp.cpp:133-198): nodes of a layer over its fixed
    generators (Selene 228 and 38, Helios 18, doubling chains with Z != 1) as plain MSMs, as one
    batch, and over a resident fixed-base set of the generators by both of its methods, against the
    compiled reference's bytes."""
    runs = _runs("tree_hash")
    for c, n, name in W.TREE_HASH_SHAPES:
        rs = sorted([r for r in runs if r["curve"] == O.CURVES[c] and r["n"] == n], key=lambda r: r["index"])
        want = [r["ref_result"] for r in rs]
        pts = W.doubling_chain(c, n)
        sc = W.tree_layer_scalars(n, len(rs), rs[0]["seed"])
        print(name, "plain", flush=True)
        assert [hs.msm(c, sc[h], pts)[0].hex() for h in range(len(rs))] == want, name
        print(name, "batch", flush=True)
        assert [o.hex() for o, _ in hs.msm_batch([(c, sc[h], pts) for h in range(len(rs))])] == want, name
        fs = hs.FixedSet(c, pts)
        try:
            for method in (hs.FIXED_COMB, hs.FIXED_PIPPENGER):
                print(name, "fixed", method, flush=True)
                out, _ = fs.msm(sc, method=method)
                assert [o.hex() for o in out] == want, (name, method)
        finally:
            fs.close()


def test_fcmpp_reference_chain_structs():
    """The reference's own doubling-chain structs (its limbs) give the same bytes as the
    workload generator's canonical structs."""
    if O.ref() is None:
        pytest.skip("compiled reference (oracle/_ref) not built")
    for c, n in ((1, 602), (0, 345)):
        pts = O.ref_dbl_chain(c, n)
        sc = W.fcmpp_scalars(n)
        out, _ = hs.msm(c, sc, pts)
        assert out == O.ref_msm_jac(c, sc, pts)
