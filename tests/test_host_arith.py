"""CPU checks of the product's limb arithmetic (helioselene_amd/csrc/field.h, curve.h) compiled
for the host by tests/native/Makefile: the same source the GPU kernels use, checked against
Python big integers (field) and the oracle (points, including P == Q and P == -Q)."""
import ctypes
import os
import random
import subprocess

import pytest

import _oracle as O
import gpu_helpers as GH

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "native", "_build", "libfield_host.so")


@pytest.fixture(scope="module")
def L():
    subprocess.run(["make", "-C", os.path.join(HERE, "native")], check=True, capture_output=True)
    return ctypes.CDLL(SO)


def call(L, k, o, a, b=0):
    out = ctypes.create_string_buffer(32)
    L.hst_fe_op(k, o, out, a.to_bytes(32, "little"), b.to_bytes(32, "little"))
    return int.from_bytes(out.raw, "little")


@pytest.mark.parametrize("k", [0, 1])
def test_field_ops_vs_bigint(L, k):
    m = O.FIELD[k]
    rnd = random.Random(5 + k)
    special = [0, 1, 2, m - 1, m, m + 1, 2 * m - 1, 2 * m, (1 << 256) - 1, (1 << 256) - 2, 1 << 255,
               (1 << 255) - 1, 1 << 128, (1 << 256) - m, (1 << 256) - 38]
    vals = special + [rnd.getrandbits(256) for _ in range(2000)]
    for i, a in enumerate(vals):
        b = vals[(i * 7 + 3) % len(vals)] if i >= len(special) else rnd.getrandbits(256)
        for o, want in ((0, a * b), (1, a * a), (2, a + b), (3, a - b), (6, -a)):
            r = call(L, k, o, a, b)
            assert r < 1 << 256 and r % m == want % m, (o, hex(a), hex(b))
        assert call(L, k, 5, a) == a % m
        sm = rnd.getrandbits(14)
        assert call(L, k, 7, a, sm) % m == a * sm % m
    for a in special[:6] + [rnd.getrandbits(256) for _ in range(10)]:
        assert call(L, k, 4, a) % m == pow(a, m - 2, m)


@pytest.mark.parametrize("k", [0, 1])
def test_divstep_inverse(L, k):
    """fe_inv (divsteps) against Python's modular inverse and the Fermat form: 0, small values,
    m - 1, non-canonical inputs (>= m, up to 2^256 - 1), powers of two, values just below m, and
    random values; canonical outputs."""
    m = O.FIELD[k]
    rnd = random.Random(77 + k)
    vals = [0, 1, 2, 3, m - 1, m - 2, m, m + 1, 2 * m - 1, (1 << 256) - 1, (1 << 255) - 1, (1 << 254),
            (m + 1) // 2, m // 3]
    vals += [1 << i for i in range(0, 256, 7)] + [m - (1 << i) for i in range(0, 250, 9)]
    vals += [rnd.getrandbits(256) for _ in range(3000)] + [rnd.randrange(m) for _ in range(3000)]
    for a in vals:
        r = call(L, k, 4, a)
        want = 0 if a % m == 0 else pow(a, -1, m)
        assert r == want, hex(a)
    for a in vals[:200]:
        assert call(L, k, 8, a) == call(L, k, 4, a)


@pytest.mark.parametrize("k", [0, 1])
def test_r51_conversion(L, k):
    """Reference radix-2^51 limbs (lazy: limbs up to 64 bits) in and canonical limbs out."""
    m = O.FIELD[k]
    rnd = random.Random(9)
    for _ in range(2000):
        limbs = [rnd.getrandbits(rnd.choice([51, 52, 54, 60, 64])) for _ in range(5)]
        arr = (ctypes.c_uint64 * 5)(*limbs)
        out = ctypes.create_string_buffer(32)
        L.hst_fe_from_r51(k, out, arr)
        v = sum(x << (51 * i) for i, x in enumerate(limbs))
        r = int.from_bytes(out.raw, "little")
        assert r < 1 << 256 and r % m == v % m
        L.hst_fe_to_r51(k, arr, out.raw)
        assert sum(x << (51 * i) for i, x in enumerate(arr)) == v % m
        assert all(x < (1 << 51) for x in arr)


def _enc(blob):
    if not blob[64]:
        return bytes(32)
    b = bytearray(blob[:32])
    b[31] |= (blob[32] & 1) << 7
    return bytes(b)


@pytest.mark.parametrize("c", [0, 1])
def test_point_ops_vs_oracle(L, c):
    g = O.generator(c)
    pts = [O.scalar_mul(c, (i * 1234567 + 5).to_bytes(32, "little"), g) for i in range(12)]
    out = ctypes.create_string_buffer(65)

    def aff(p):
        x, y = GH.decompress(c, p)
        return x.to_bytes(32, "little") + y.to_bytes(32, "little")

    for i, P in enumerate(pts):
        Q = pts[(i * 3 + 1) % len(pts)]
        pa, qa = aff(P), aff(Q)
        for o, want in ((0, O.point_op("add", c, P, Q)), (1, O.point_op("add", c, P, Q)),
                        (2, O.point_op("dbl", c, P)), (3, O.point_op("dbl", c, P)),
                        (5, O.point_op("add", c, O.point_op("dbl", c, P), Q))):
            L.hst_pt_op(c, o, out, pa, qa, 0)
            assert _enc(out.raw) == want, o
        for o in (0, 1):  # P + P and P + (-P)
            L.hst_pt_op(c, o, out, pa, pa, 0)
            assert _enc(out.raw) == O.point_op("dbl", c, P)
            L.hst_pt_op(c, o, out, pa, aff(O.point_op("neg", c, P)), 0)
            assert _enc(out.raw) == bytes(32)
        k = random.Random(i).getrandbits(20)
        L.hst_pt_op(c, 4, out, pa, qa, k)
        assert _enc(out.raw) == O.scalar_mul(c, k.to_bytes(32, "little"), P)


@pytest.mark.parametrize("fn", ["hst_madd_chain", "hst_walk_chain"])
@pytest.mark.parametrize("c", [0, 1])
def test_madd_chain_vs_oracle(L, c, fn):
    """Long xyzz_madd chains with repeated points, negated points, identities and cancellations,
    against the oracle's sum.  hst_walk_chain is the bucket walk's form: points arrive with their
    sign applied, the run starts from its first point and continues with xyzz_madd_hot (trial 2
    starts with a negated point, trial 3 cancels the first point with the second)."""
    rnd = random.Random(40 + c)
    g = O.generator(c)
    base = [O.scalar_mul(c, rnd.getrandbits(250).to_bytes(32, "little"), g) for _ in range(40)]
    for trial in range(6):
        n = 300
        idx = [rnd.randrange(len(base)) for _ in range(n)]
        neg = [rnd.random() < 0.5 for _ in range(n)]
        if trial % 2:
            # force doublings (acc == P) and cancellations (acc == -P) mid-chain
            idx[10:14] = [idx[9]] * 4
            neg[10:14] = [neg[9]] * 4
            idx[50], neg[50] = idx[49], not neg[49]
        if trial == 2:
            neg[0] = True
        if trial == 3:
            # the first point cancelled by the second: the run continues from the identity
            idx[1], neg[0], neg[1] = idx[0], True, False
        pts, want_sc = bytearray(), bytearray()
        for i in range(n):
            P = base[idx[i]]
            x, y = GH.decompress(c, P)
            pts += x.to_bytes(32, "little") + y.to_bytes(32, "little")
        pts[64 * 7:64 * 8] = bytes(64)  # one identity entry
        sc = [0] * len(base)
        for i in range(n):
            if i != 7:
                sc[idx[i]] += -1 if neg[i] else 1
        order = O.ORDER[c]
        want = O.msm_naive(c, b"".join((v % order).to_bytes(32, "little") for v in sc), b"".join(base), len(base))
        out = ctypes.create_string_buffer(65)
        f = getattr(L, fn)
        f.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
        f(c, out, bytes(pts), bytes(int(v) for v in neg), n)
        assert _enc(out.raw) == want, trial


@pytest.mark.parametrize("c", [0, 1])
def test_decompress_and_sswu_vs_oracle(L, c, golden):
    """Device decode / hash-to-curve code on the host vs the oracle and the reference vectors."""
    v = golden[0][O.CURVES[c] + "_point"]
    out = ctypes.create_string_buffer(64)
    for t in v["from_bytes"]:
        rc = L.hst_decompress(c, out, bytes.fromhex(t["input"]))
        if t["label"] == "identity_encoding":
            assert rc == 1
            continue
        assert (rc == 0) == (t["result"] is not None), t["label"]
    for t in v["map_to_curve_single"]:
        L.hst_sswu(c, out, bytes.fromhex(t["u"]))
        assert _enc(out.raw + b"\x01").hex() == t["result"], t["label"]
    u = O.xoshiro(31 + c, 32 * 200)
    want = O.map_to_curve(c, u)
    for i in range(200):
        L.hst_sswu(c, out, u[32 * i:32 * i + 32])
        got = _enc(out.raw + b"\x01")
        assert got == want[32 * i:32 * i + 32]
        rc = L.hst_decompress(c, out, got)
        assert rc == 0 and GH.decompress(c, got) == (int.from_bytes(out.raw[:32], "little"),
                                                     int.from_bytes(out.raw[32:], "little"))


@pytest.mark.parametrize("c", [0, 1])
def test_scalar_prepare_and_recode(L, c):
    """The device recode: prepared scalar = s mod order (minus 2^256 when the reference's window-8
    recode drops its carry, in compat mode); digits sum back to it, lie in
    [-(2^(w-1)) + 1, 2^(w-1)], and no digit is a signed zero (v == 2^w with a carry)."""
    order = O.ORDER[c]
    rnd = random.Random(21 + c)
    vals = [0, 1, order - 1, order, order + 1, (1 << 256) - 1, (1 << 255) - 1, (0x7F << 248) | (0xFF << 240),
            int("f" * 64, 16) >> 3]
    vals += [rnd.getrandbits(256) for _ in range(300)]
    vals += [int("".join(rnd.choice("0f") for _ in range(64)), 16) for _ in range(100)]  # all-ones windows
    prep = ctypes.create_string_buffer(32)
    dig = (ctypes.c_int32 * 80)()
    for w in range(4, 17):
        for v in vals:
            for compat in (0, 1):
                W = L.hst_recode(c, w, compat, v.to_bytes(32, "little"), prep, dig)
                assert W == (256 + w - 1) // w
                want = v % order
                if compat and O.drops_carry(v.to_bytes(32, "little"), 8):
                    want = (v - (1 << 256)) % order
                assert int.from_bytes(prep.raw, "little") == want
                assert sum(dig[i] << (w * i) for i in range(W)) == want
                assert all(-(1 << (w - 1)) < dig[i] <= (1 << (w - 1)) for i in range(W))


@pytest.mark.parametrize("k", [0, 1])
def test_field_output_bounds(L, k):
    """Every result keeps the representation invariant: limbs <= 2^29, value < 2^262."""
    m = O.FIELD[k]
    rnd = random.Random(77 + k)
    vals = [0, 1, m - 1, (1 << 256) - 1, (1 << 255), (1 << 256) - 38] + [rnd.getrandbits(256) for _ in range(3000)]
    out = ctypes.create_string_buffer(40)
    for i, a in enumerate(vals):
        b = vals[(i * 13 + 1) % len(vals)]
        for o, want in ((0, a * b), (1, a * a), (2, a + b), (3, a - b)):
            L.hst_fe_op_raw(k, o, out, a.to_bytes(32, "little"), b.to_bytes(32, "little"))
            assert out.raw[39] == 0, ("limb bound", o)
            v = int.from_bytes(out.raw[:39], "little")
            assert v < (1 << 262) and v % m == want % m, (o, hex(a), hex(b))


def _limbs_value(l):
    return sum(x << (29 * i) for i, x in enumerate(l))


@pytest.mark.parametrize("k", [0, 1])
def test_field_extreme_limbs(L, k):
    """Inputs at the representation's limits (every limb 2^29, or 2^29 - 1, mixed with zeros):
    products and squares stay exact and normalised; the zero test is exact on multiples of m
    written with non-canonical limbs."""
    m = O.FIELD[k]
    rnd = random.Random(31 + k)
    top = 1 << 29
    pats = [[top] * 9, [top - 1] * 9, [top, 0] * 4 + [top], [0] * 8 + [top], [top] + [0] * 8]
    pats += [[rnd.choice([0, top, top - 1, rnd.getrandbits(29)]) for _ in range(9)] for _ in range(400)]
    out = ctypes.create_string_buffer(40)
    for i, la in enumerate(pats):
        lb = pats[(i * 7 + 1) % len(pats)]
        a, b = _limbs_value(la), _limbs_value(lb)
        A, B = (ctypes.c_uint32 * 9)(*la), (ctypes.c_uint32 * 9)(*lb)
        for o, want in ((0, a * b), (1, a * a), (2, a + b), (3, a - b)):
            L.hst_fe_op_limbs(k, o, out, A, B)
            assert out.raw[39] == 0, ("limb bound", o, la, lb)
            v = int.from_bytes(out.raw[:39], "little")
            assert v < (1 << 262) and v % m == want % m, (o, la, lb)
    # products with one lazily subtracted operand (limbs up to 2^29 + 2^31) and one normalised
    lazy_top = top + (1 << 31) - 1
    lazy = [[lazy_top] * 9, [lazy_top, top] * 4 + [lazy_top]] + \
        [[rnd.randrange(0, lazy_top + 1) for _ in range(9)] for _ in range(200)]
    for i, la in enumerate(lazy):
        lb = pats[i % len(pats)]
        a, b = _limbs_value(la), _limbs_value(lb)
        for x, y in ((la, lb), (lb, la)):
            L.hst_fe_op_limbs(k, 0, out, (ctypes.c_uint32 * 9)(*x), (ctypes.c_uint32 * 9)(*y))
            assert out.raw[39] == 0, ("limb bound", la, lb)
            v = int.from_bytes(out.raw[:39], "little")
            assert v < (1 << 262) and v % m == (a * b) % m, (la, lb)
    # zero test: j*m for j = 0..64 in strict digits, and with one digit borrowed into the next
    for j in list(range(65)):
        v = j * m
        d = [(v >> (29 * i)) & (top - 1) for i in range(9)]
        assert L.hst_fe_op_limbs(k, 4, out, (ctypes.c_uint32 * 9)(*d), (ctypes.c_uint32 * 9)()) == 1, j
        for pos in range(8):
            if d[pos + 1] > 0:
                e = list(d)
                e[pos + 1] -= 1
                e[pos] += top  # same value, limb pos = digit + 2^29 (<= 2^30)
                if e[pos] <= top:
                    assert L.hst_fe_op_limbs(k, 4, out, (ctypes.c_uint32 * 9)(*e), (ctypes.c_uint32 * 9)()) == 1
    for _ in range(2000):
        la = [rnd.getrandbits(29) for _ in range(9)]
        want = _limbs_value(la) % m == 0
        assert L.hst_fe_op_limbs(k, 4, out, (ctypes.c_uint32 * 9)(*la), (ctypes.c_uint32 * 9)()) == int(want)
    for v in (1, m - 1, m + 1, 2 * m - 1, (1 << 255), 19, 1 << 145):
        d = [(v >> (29 * i)) & (top - 1) for i in range(9)]
        assert L.hst_fe_op_limbs(k, 4, out, (ctypes.c_uint32 * 9)(*d), (ctypes.c_uint32 * 9)()) == 0, hex(v)


@pytest.mark.parametrize("k", [0, 1])
def test_zero_test_of_normalised_values(L, k):
    """fe_is_zero_norm (no carry pass) on the values normalisation can produce: limbs 3..8 strict
    29-bit digits, limbs 0..2 up to 2^29.  Every multiple j m (j <= 64) in every such form is
    zero, and random / near-multiple values agree with the exact test."""
    m = O.FIELD[k]
    rnd = random.Random(91 + k)
    top = 1 << 29
    out = ctypes.create_string_buffer(40)
    U = lambda l: (ctypes.c_uint32 * 9)(*l)

    def forms(v):
        d = [(v >> (29 * i)) & (top - 1) for i in range(9)]
        if v >> 261:
            return
        yield d
        # borrow one from digit p + 1 into digit p (p <= 2), while the digit stays <= 2^29
        for p in range(3):
            if d[p + 1] > 0 and d[p] == 0:
                e = list(d)
                e[p + 1] -= 1
                e[p] += top
                yield e

    for j in range(65):
        for d in forms(j * m):
            assert L.hst_fe_op_limbs(k, 5, out, U(d), U([0] * 9)) == 1, (j, d)
    for v in [1, m - 1, m + 1, 2 * m - 1, 1 << 255, 1 << 145, (1 << 261) - 1] + \
             [j * m + rnd.choice([-1, 1]) * rnd.getrandbits(rnd.choice([3, 40, 140])) for j in range(1, 64)] + \
             [rnd.getrandbits(261) for _ in range(2000)]:
        for d in forms(v % (1 << 261)):
            want = int(_limbs_value(d) % m == 0)
            assert L.hst_fe_op_limbs(k, 5, out, U(d), U([0] * 9)) == want, hex(v)


@pytest.mark.parametrize("k", [0, 1])
def test_mul_sum2_extreme_limbs(L, k):
    """a b + c d with one reduction, as the point formulas form it: b = x - y with the small-limb
    bias, c = -z, every input at the representation's limits (limbs 0, 2^29 - 1 or 2^29); also
    with the lazy operands passed raw at their largest limbs (2^29 + 2^30 - 2 and 2^30 - 2)."""
    m = O.FIELD[k]
    rnd = random.Random(77 + k)
    top = 1 << 29
    pats = [[top] * 9, [top - 1] * 9, [0] * 9, [top, 0] * 4 + [top], [0, top] * 4 + [0]]
    pats += [[rnd.choice([0, top, top - 1, rnd.getrandbits(29)]) for _ in range(9)] for _ in range(300)]
    out = ctypes.create_string_buffer(40)
    U = lambda l: (ctypes.c_uint32 * 9)(*l)
    for i, la in enumerate(pats):
        x, y, z, d = (pats[(i * s + 1) % len(pats)] for s in (3, 5, 7, 11))
        A, X, Y, Z, D = (_limbs_value(v) for v in (la, x, y, z, d))
        L.hst_fe_mul_sum2(k, out, U(la), U(x), U(y), U(z), U(d), 1)
        assert out.raw[39] == 0, ("limb bound", i)
        v = int.from_bytes(out.raw[:39], "little")
        assert v < (1 << 262) and v % m == (A * (X - Y) - Z * D) % m, i
    hi_b, hi_c = top + (1 << 30) - 2, (1 << 30) - 2
    for b, c in (([hi_b] * 9, [hi_c] * 9), ([hi_b, 0] * 4 + [hi_b], [0, hi_c] * 4 + [0])):
        for a, d in (([top] * 9, [top] * 9), ([top - 1] * 9, [top] * 9)):
            L.hst_fe_mul_sum2(k, out, U(a), U(b), U([0] * 9), U(c), U(d), 0)
            assert out.raw[39] == 0
            v = int.from_bytes(out.raw[:39], "little")
            want = _limbs_value(a) * _limbs_value(b) + _limbs_value(c) * _limbs_value(d)
            assert v < (1 << 262) and v % m == want % m


@pytest.mark.parametrize("k", [0, 1])
def test_sub_small_mul_extreme_limbs(L, k):
    """a - j b with one carry pass (j = 1..15) for normalised inputs at the limb limits."""
    m = O.FIELD[k]
    rnd = random.Random(66 + k)
    top = 1 << 29
    pats = [[top] * 9, [top - 1] * 9, [0] * 9, [top, 0] * 4 + [top], [0, top] * 4 + [0]]
    pats += [[rnd.choice([0, top, top - 1, rnd.getrandbits(29)]) for _ in range(9)] for _ in range(120)]
    out = ctypes.create_string_buffer(40)
    U = lambda l: (ctypes.c_uint32 * 9)(*l)
    for i, la in enumerate(pats):
        lb = pats[(i * 7 + 3) % len(pats)]
        A, B = _limbs_value(la), _limbs_value(lb)
        for j in (1, 2, 3, 4, 8, 15):
            L.hst_fe_op_limbs(k, 16 + j, out, U(la), U(lb))
            v = int.from_bytes(out.raw[:39], "little")
            assert out.raw[39] == 0 and v < (1 << 262) and v % m == (A - j * B) % m, (i, j)


@pytest.mark.parametrize("k", [0, 1])
def test_mul_add_extreme_limbs(L, k):
    """a b - c and a^2 - c - d - e with one reduction (the lazy negation as the low columns' start
    value), every input at the representation's limits, and one product operand a lazily negated
    base coordinate (limbs up to 2^30 - 2)."""
    m = O.FIELD[k]
    rnd = random.Random(55 + k)
    top = 1 << 29
    pats = [[top] * 9, [top - 1] * 9, [0] * 9, [top, 0] * 4 + [top]]
    pats += [[rnd.choice([0, top, top - 1, rnd.getrandbits(29)]) for _ in range(9)] for _ in range(300)]
    out = ctypes.create_string_buffer(40)
    U = lambda l: (ctypes.c_uint32 * 9)(*l)
    for i, la in enumerate(pats):
        b, c, d, e = (pats[(i * s + 1) % len(pats)] for s in (3, 5, 7, 11))
        A, B, C, D, E = (_limbs_value(v) for v in (la, b, c, d, e))
        L.hst_fe_mul_add(k, out, U(la), U(b), U(c), U(d), U(e), 0)
        v = int.from_bytes(out.raw[:39], "little")
        assert out.raw[39] == 0 and v < (1 << 262) and v % m == (A * B - C) % m, i
        L.hst_fe_mul_add(k, out, U(la), U(b), U(c), U(d), U(e), 1)
        v = int.from_bytes(out.raw[:39], "little")
        assert out.raw[39] == 0 and v < (1 << 262) and v % m == (A * A - C - D - E) % m, i
    lazy = [(1 << 30) - 2] * 9  # 65 m - y for a base with zero limbs: the largest lazy operand
    for b in pats[:4]:
        L.hst_fe_mul_add(k, out, U(lazy), U(b), U(pats[0]), U(pats[0]), U(pats[0]), 0)
        v = int.from_bytes(out.raw[:39], "little")
        assert out.raw[39] == 0 and v % m == (_limbs_value(lazy) * _limbs_value(b) - _limbs_value(pats[0])) % m


@pytest.mark.parametrize("c", [0, 1])
def test_jacobian_chain_vs_oracle(L, c):
    """xyzz -> Jacobian conversion, dbl-2001-b chains and add-2007-bl (including P + P and
    P + (-P)) against the oracle."""
    g = O.generator(c)

    def aff(p):
        x, y = GH.decompress(c, p)
        return x.to_bytes(32, "little") + y.to_bytes(32, "little")

    out = ctypes.create_string_buffer(65)
    for a, b, k in ((12345, 999, 0), (12345, 999, 1), (777, 5, 13), (3, 3, 0), (3, 12, 0), (5, O.ORDER[c] - 10, 1)):
        P = O.scalar_mul(c, a.to_bytes(32, "little"), g)
        Q = O.scalar_mul(c, b.to_bytes(32, "little"), g)
        L.hst_jac_chain(c, out, aff(P), aff(Q), k)
        want = O.scalar_mul(c, ((a * 2 ** (k + 1) + 2 * b) % O.ORDER[c]).to_bytes(32, "little"), g)
        assert _enc(out.raw) == want, (a, b, k)
