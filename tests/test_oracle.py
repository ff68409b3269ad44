"""Pins the CPU oracle (oracle/hs_oracle.c) before it is trusted as the checker.

Each oracle function is checked against (a) the reference's own golden vectors and known answers,
(b) closed-form big-integer arithmetic, and (c) the reference library compiled from its sources
(oracle/_ref, when built), so that a dropped term, a wrong sign or a transposed operand fails.
"""
import os
import random

import pytest

import _oracle as O

H = bytes.fromhex


def le(x):
    return int.from_bytes(x, "little")


def b32(x):
    return int(x).to_bytes(32, "little")


# ---------------------------------------------------------------- field arithmetic
def test_tests_cpp_field_kats(golden):
    """src/tests.cpp:102-139: a*b, a^2, a^-1 in F_p and F_q, sqrt(4) in F_q."""
    k = golden[0]["tests_cpp_kat"]
    a, b = H(k["test_a"]), H(k["test_b"])
    for fk, name in ((0, "fp"), (1, "fq")):
        assert O.fe_op("mul", fk, a, b)[0].hex() == k[name + "_ab"]
        assert O.fe_op("sq", fk, a)[0].hex() == k[name + "_asq"]
        assert O.fe_op("inv", fk, a)[0].hex() == k[name + "_ainv"]
    r, rc = O.fe_op("sqrt", 1, b32(4))
    assert rc == 0 and le(r) in (2, O.FIELD[1] - 2)


@pytest.mark.parametrize("sec,fk", [("selene_scalar", 0), ("helios_scalar", 1)])
def test_scalar_vectors_are_field_ops(golden, sec, fk):
    """Helios scalars live in F_q and Selene scalars in F_p, so the reference's scalar vectors
    (test_vectors json: *_scalar/add,sub,mul,sq,invert,negate) pin the field arithmetic."""
    v = golden[0][sec]
    for t in v["mul"]:
        assert O.fe_op("mul", fk, H(t["a"]), H(t["b"]))[0].hex() == t["result"], t["label"]
    for t in v["sq"]:
        assert O.fe_op("sq", fk, H(t["a"]))[0].hex() == t["result"], t["label"]
    for t in v["add"]:
        assert O.fe_op("add", fk, H(t["a"]), H(t["b"]))[0].hex() == t["result"], t["label"]
    for t in v["sub"]:
        assert O.fe_op("sub", fk, H(t["a"]), H(t["b"]))[0].hex() == t["result"], t["label"]
    for t in v["invert"]:
        if t.get("result") is None:
            continue
        assert O.fe_op("inv", fk, H(t["a"]))[0].hex() == t["result"], t["label"]
    for t in v["negate"]:
        assert O.fe_op("sub", fk, bytes(32), H(t["a"]))[0].hex() == t["result"], t["label"]


@pytest.mark.parametrize("fk", [0, 1])
def test_field_closed_form(fk):
    """mul/sq/add/sub/inv/sqrt against Python big integers, including values near the modulus
    and unreduced inputs (the oracle reduces what it loads, like the lazy reference limbs)."""
    m = O.FIELD[fk]
    rnd = random.Random(7 + fk)
    vals = [0, 1, 2, m - 1, m - 2, (1 << 255) - 1, (1 << 256) - 1, m, m + 5, 2 ** 128, 2 ** 255 - 2 ** 128]
    vals += [rnd.getrandbits(256) for _ in range(40)]
    for a in vals:
        for bb in vals[:12] + [rnd.getrandbits(256)]:
            ab, ib = b32(a), b32(bb)
            assert le(O.fe_op("mul", fk, ab, ib)[0]) == (a * bb) % m
            assert le(O.fe_op("add", fk, ab, ib)[0]) == (a + bb) % m
            assert le(O.fe_op("sub", fk, ab, ib)[0]) == (a - bb) % m
        assert le(O.fe_op("inv", fk, b32(a))[0]) == (pow(a, m - 2, m) if a % m else 0)
        r, rc = O.fe_op("sqrt", fk, b32(a))
        is_qr = pow(a % m, (m - 1) // 2, m) in (0, 1)
        assert (rc == 0) == is_qr
        if rc == 0:
            assert le(r) ** 2 % m == a % m


def test_batch_invert_vectors(golden):
    bi = golden[0]["batch_invert"]
    import ctypes
    for fk, key in ((0, "fp"), (1, "fq")):
        for t in bi[key]:
            n = t["n"]
            out = ctypes.create_string_buffer(32 * n)
            O.oracle().hso_batch_invert(fk, out, b"".join(H(x) for x in t["inputs"]), n)
            assert [out.raw[32 * i:32 * i + 32].hex() for i in range(n)] == t["results"], t["label"]


# ---------------------------------------------------------------- points
@pytest.mark.parametrize("c", [0, 1])
def test_point_vectors(golden, c):
    v = golden[0][O.CURVES[c] + "_point"]
    assert O.generator(c).hex() == v["generator"]
    for t in v["from_bytes"]:
        if t["label"] == "identity_encoding":
            continue  # the reference decoder rejects 0^32; the oracle loader maps it to identity
        ok = O.oracle().hso_point_valid(c, H(t["input"]))
        assert bool(ok) == (t["result"] is not None), t["label"]
    for t in v["add"]:
        assert O.point_op("add", c, H(t["a"]), H(t["b"])).hex() == t["result"], t["label"]
    for t in v["dbl"]:
        assert O.point_op("dbl", c, H(t["a"])).hex() == t["result"], t["label"]
    for t in v["negate"]:
        assert O.point_op("neg", c, H(t["a"])).hex() == t["result"], t["label"]
    for t in v["scalar_mul"]:
        assert O.scalar_mul(c, H(t["scalar"]), H(t["point"])).hex() == t["result"], t["label"]
    for t in v["map_to_curve_single"]:
        assert O.map_to_curve(c, H(t["u"])).hex() == t["result"], t["label"]
    import ctypes
    for t in v["map_to_curve_double"]:
        out = ctypes.create_string_buffer(32)
        O.oracle().hso_map_to_curve2(c, out, H(t["u0"]), H(t["u1"]))
        assert out.raw.hex() == t["result"], t["label"]


def test_tests_cpp_point_kats(golden):
    """src/tests.cpp:141-165: G, 2G, 7G compressed encodings for both curves."""
    k = golden[0]["tests_cpp_kat"]
    for c, name in ((0, "helios"), (1, "selene")):
        g = O.generator(c)
        assert g.hex() == k[name + "_g"]
        assert O.point_op("dbl", c, g).hex() == k[name + "_2g"]
        assert O.scalar_mul(c, b32(7), g).hex() == k[name + "_7g"]


@pytest.mark.parametrize("c", [0, 1])
def test_group_order_and_identity(c):
    """Closed forms: order*G = O, (order-1)*G = -G, P + (-P) = O, 2P = P + P."""
    g = O.generator(c)
    assert O.scalar_mul(c, b32(O.ORDER[c]), g) == bytes(32)
    neg = O.point_op("neg", c, g)
    assert O.scalar_mul(c, b32(O.ORDER[c] - 1), g) == neg
    assert O.point_op("add", c, g, neg) == bytes(32)
    p = O.scalar_mul(c, b32(123456789), g)
    assert O.point_op("add", c, p, p) == O.point_op("dbl", c, p)


# ---------------------------------------------------------------- MSM
@pytest.mark.parametrize("c", [0, 1])
def test_msm_vectors(golden, c):
    v = golden[0][O.CURVES[c] + "_point"]
    for t in v["msm"]:
        n = t["n"]
        s = b"".join(H(x) for x in t["scalars"])
        p = b"".join(H(x) for x in t["points"])
        assert O.msm(c, s, p, n).hex() == t["result"], t["label"]
        assert O.msm_naive(c, s, p, n).hex() == t["result"], t["label"]
    for t in v["pedersen_commit"]:
        n = t["n"]
        vals = b"".join(H(x) for x in t["values"])
        gens = b"".join(H(x) for x in t["generators"])
        assert O.pedersen(c, H(t["blinding"]), H(t["H"]), vals, gens, n).hex() == t["result"], t["label"]


def test_msm_reference_runs(golden):
    """The oracle's MSM reproduces the compiled reference's bytes on seeded inputs in every
    window band, including the window-8 band (864 <= n < 2592) where the reference drops the
    final recode carry; outside that band it equals the exact naive sum."""
    for r in golden[1]:
        c = O.CURVES.index(r["curve"])
        n = r["n"]
        if r["bases"] == "kG":
            continue  # covered by test_msm_reference_run_kG (slow)
        s = O.gen_scalars(c, r["seed"], n)
        p = O.gen_sswu_bases(c, r["seed"], n, r["dup_frac"], r["id_frac"])
        got = O.msm(c, s, p, n)
        assert got.hex() == r["ref_result"], (r["curve"], n)
        if r["naive_result"] is not None:
            w8 = O.ref_window(n) == 8
            dropped = any(O.drops_carry(s[32 * i:32 * i + 32], 8) for i in range(n))
            assert (r["naive_result"] == r["ref_result"]) == (not (w8 and dropped)), (r["curve"], n)


def test_window8_defect_model(golden):
    """Pins the model the GPU compat mode uses: in the window-8 band the reference result equals
    the exact sum with every scalar whose top digit overflows replaced by s - 2^256 (mod order)."""
    for r in golden[1]:
        n = r["n"]
        if r["bases"] != "sswu" or O.ref_window(n) != 8:
            continue
        c = O.CURVES.index(r["curve"])
        s = bytearray(O.gen_scalars(c, r["seed"], n))
        p = O.gen_sswu_bases(c, r["seed"], n, r["dup_frac"], r["id_frac"])
        for i in range(n):
            si = s[32 * i:32 * i + 32]
            if O.drops_carry(si, 8):
                s[32 * i:32 * i + 32] = b32((le(si) - (1 << 256)) % O.ORDER[c])
        assert O.msm_naive(c, bytes(s), p, n).hex() == r["ref_result"]


@pytest.mark.slow
def test_msm_reference_run_kG(golden):
    """n = 4096 Helios on k*G points (the reference fuzz recipe)."""
    r = [x for x in golden[1] if x["bases"] == "kG"][0]
    s = O.gen_scalars(0, r["seed"], r["n"])
    p = O.gen_kG_points(0, r["seed"], r["n"])
    assert O.msm(0, s, p, r["n"]).hex() == r["ref_result"]


@pytest.mark.parametrize("c", [0, 1])
def test_recode_invariant(c):
    """sum d_i 2^(w i) = s - carry*2^(w*count) and |d_i| <= 2^(w-1) for every width."""
    import ctypes
    rnd = random.Random(11)
    vals = [0, 1, O.ORDER[c] - 1, (1 << 255) - 1, (1 << 256) - 1, 0x7F << 248, (0x7F << 248) | (0x80 << 240)]
    vals += [rnd.getrandbits(256) for _ in range(50)]
    for w in range(4, 17):
        for v in vals:
            d = (ctypes.c_int16 * 80)()
            cnt = O.oracle().hso_recode(d, b32(v), w)
            dropped = O.drops_carry(b32(v), w)
            total = sum(d[i] << (w * i) for i in range(cnt))
            assert total == v - (dropped << (w * cnt))
            assert all(-(1 << (w - 1)) <= d[i] < (1 << (w - 1)) for i in range(cnt))


@pytest.mark.parametrize("c", [0, 1])
def test_msm_fixed_and_pedersen_consistency(c):
    """Fixed-base MSM and Pedersen agree with the plain MSM on random inputs."""
    n = 9
    s = O.gen_scalars(c, 77, n + 1)
    p = O.gen_sswu_bases(c, 77, n + 1)
    assert O.msm_fixed(c, s, p, n) == O.msm_naive(c, s, p, n)
    assert O.pedersen(c, s[:32], p[:32], s[32:], p[32:], n) == O.msm_naive(c, s, p, n + 1)


@pytest.mark.skipif(not os.path.exists(O.REF_SO), reason="reference library not built")
@pytest.mark.parametrize("c", [0, 1])
def test_oracle_matches_compiled_reference(c):
    """Cross-check the restatement against the compiled reference on fresh seeded inputs:
    msm, naive sum, Pedersen, fixed-base MSM and SSWU."""
    import ctypes
    r = O.ref()
    name = O.CURVES[c]
    for n in (1, 5, 16, 17, 40, 300, 900):
        s = O.gen_scalars(c, 4242 + n, n, edge_frac=0.05)
        p = O.gen_sswu_bases(c, 4242 + n, n, dup_frac=0.05, id_frac=0.02)
        assert O.msm(c, s, p, n) == O.ref_msm(c, s, p, n), n
        out = ctypes.create_string_buffer(32)
        getattr(r, f"hsref_{name}_msm_fixed_bytes")(out, s, p, n)
        assert O.msm_fixed(c, s, p, n) == out.raw, n
    u = O.xoshiro(99, 32 * 64)
    out = ctypes.create_string_buffer(32 * 64)
    getattr(r, f"hsref_{name}_map_to_curve_bytes")(out, u, 64)
    assert O.map_to_curve(c, u) == out.raw
