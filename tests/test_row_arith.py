"""CPU checks of the row-distributed field and curve arithmetic (helioselene_amd/csrc/rowfield.h,
rowcurve.h) that the Horner kernels use: the same source compiled for the host over a lane-by-lane
wavefront model (tests/native/row_emu.h) that flags every wrapped intermediate.

* field operations on four rows at once against Python big integers, at the operand bounds the
  header states (largest row-normalised limbs, lazy sums and differences, raw 32-bit limbs);
* results row-normalised (limbs < 2^29 + 2^18) with zeros in lanes 9..15 and no wrap anywhere;
* doubling chains, additions (generic, P == Q, P == -Q, identities) and Horner combines against the
  single-lane curve.h formulas they restate: the same residues in every coordinate.
"""
import ctypes
import os
import random
import subprocess

import pytest

import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "native", "_build", "librow_host.so")
RN = (1 << 29) + (1 << 18)
U32 = ctypes.c_uint32


@pytest.fixture(scope="module")
def L():
    subprocess.run(["make", "-C", os.path.join(HERE, "native")], check=True, capture_output=True)
    return ctypes.CDLL(SO)


def limbs(v):
    return [(v >> (29 * i)) & ((1 << 29) - 1) for i in range(9)]


def val(ls):
    return sum(int(x) << (29 * i) for i, x in enumerate(ls))


def arr(rows):
    flat = [x for r in rows for x in r]
    return (U32 * len(flat))(*flat)


def row_op(L, k, o, a, b=None, c=None, d=None, s=0):
    z = [[0] * 9] * 4
    out = (U32 * 36)()
    st = (U32 * 3)()
    L.hst_row_op(k, o, out, st, arr(a), arr(b or z), arr(c or z), arr(d or z), U32(s))
    return [list(out[9 * r:9 * r + 9]) for r in range(4)], list(st)


def rand_rn(rnd, hi=RN):
    """limbs anywhere below the row-normalised bound (value may exceed m: a lazy residue)"""
    return [rnd.randrange(hi) for _ in range(9)]


def pick(rnd, m):
    kind = rnd.randrange(6)
    if kind == 0:
        return [RN - 1] * 9
    if kind == 1:
        return rand_rn(rnd)
    if kind == 2:
        return limbs(rnd.randrange(m))
    if kind == 3:
        return limbs(rnd.choice([0, 1, m - 1, m, 2 * m, (1 << 255) - 1]) % (1 << 261))
    if kind == 4:
        return [rnd.choice([0, (1 << 29) - 1, RN - 1]) for _ in range(9)]
    return [rnd.randrange(1 << 29) for _ in range(9)]


@pytest.mark.parametrize("k", [0, 1])
def test_row_field_ops_vs_bigint(L, k):
    m = O.FIELD[k]
    rnd = random.Random(90 + k)
    for it in range(300):
        a, b, c, d = ([pick(rnd, m) for _ in range(4)] for _ in range(4))
        cases = [
            (0, lambda r: val(a[r]) * val(b[r])),
            (1, lambda r: val(a[r]) ** 2),
            (2, lambda r: val(a[r]) - val(b[r])),
            (3, lambda r: val(a[r]) + val(b[r])),
            (4, lambda r: val(a[r]) - val(b[r]) - val(c[r])),
            (5, lambda r: val(a[r]) - val(b[r]) - val(c[r]) - val(d[r])),
            (8, lambda r: val(a[r]) * (val(b[r]) - val(c[r]))),
            (9, lambda r: (val(a[r]) + val(b[r])) * (val(c[r]) + val(d[r]))),
            (11, lambda r: (val(a[r]) + val(b[r])) ** 2),
        ]
        for o, f in cases:
            out, st = row_op(L, k, o, a, b, c, d)
            assert st == [0, st[1], 0] and st[1] < RN, (o, st)
            for r in range(4):
                assert val(out[r]) % m == f(r) % m, (o, r, a[r], b[r])
        s = rnd.randrange(8)
        out, st = row_op(L, k, 6, a, s=s)
        assert st[0] == 0 and st[1] < RN and st[2] == 0
        assert all(val(out[r]) % m == val(a[r]) * s % m for r in range(4))
        s = rnd.randrange(17)
        out, st = row_op(L, k, 7, a, b, s=s)
        assert st[0] == 0 and st[1] < RN and st[2] == 0
        assert all(val(out[r]) % m == (val(a[r]) - s * val(b[r])) % m for r in range(4))
        raw = [[rnd.choice([rnd.getrandbits(32), 0xffffffff, 0]) for _ in range(9)] for _ in range(4)]
        out, st = row_op(L, k, 10, raw)
        assert st[0] == 0 and st[1] < RN and st[2] == 0
        assert all(val(out[r]) % m == val(raw[r]) % m for r in range(4))


def jac_arr(pts):
    flat = [x for p in pts for c in p for x in limbs(c)]
    return (U32 * max(27, len(flat)))(*flat)


def row_curve(L, k, o, p, q, n):
    out = (U32 * 58)()
    L.hst_row_curve(k, o, out, jac_arr(p), jac_arr(q), n)
    return out


def lane_curve(L, k, o, p, q, n):
    out = (U32 * 27)()
    L.hst_lane_curve(k, o, out, jac_arr(p), jac_arr(q), n)
    return [val(out[9 * c:9 * c + 9]) for c in range(3)]


def check_same(L, k, o, p, q, n):
    m = O.FIELD[k]
    out = row_curve(L, k, o, p, q, n)
    wraps, mx, dirty, disagree = out[27], out[28], out[29], out[30]
    assert wraps == 0 and mx < RN and dirty == 0 and disagree == 0, (wraps, mx, dirty, disagree)
    got = [val(out[9 * c:9 * c + 9]) % m for c in range(3)]
    uni = [val(out[31 + 9 * c:31 + 9 * c + 9]) for c in range(3)]
    assert max(out[31:58]) <= 1 << 29
    want = [x % m for x in lane_curve(L, k, o, p, q, n)]
    assert got == want and [u % m for u in uni] == want
    return got


def rand_pt(rnd, m):
    return [rnd.randrange(m), rnd.randrange(m), rnd.randrange(1, m)]


def scaled(p, lam, m):
    return [p[0] * lam * lam % m, p[1] * lam ** 3 % m, p[2] * lam % m]


@pytest.mark.parametrize("k", [0, 1])
def test_row_curve_vs_lane(L, k):
    m = O.FIELD[k]
    rnd = random.Random(300 + k)
    ident = [1, 1, 0]
    for _ in range(12):
        p, q = rand_pt(rnd, m), rand_pt(rnd, m)
        for n in (1, 2, 7):
            check_same(L, k, 0, [p], [ident], n)
        check_same(L, k, 1, [p], [q], 0)
        lam = rnd.randrange(2, m)
        check_same(L, k, 1, [p], [scaled(p, lam, m)], 0)                        # P + P
        neg = scaled(p, lam, m)
        neg[1] = (m - neg[1]) % m
        assert check_same(L, k, 1, [p], [neg], 0)[2] == 0                       # P + (-P)
        check_same(L, k, 1, [p], [ident], 0)
        check_same(L, k, 1, [ident], [q], 0)
    check_same(L, k, 1, [ident], [ident], 0)
    check_same(L, k, 0, [ident], [ident], 3)


@pytest.mark.parametrize("k", [0, 1])
def test_row_horner_vs_lane(L, k):
    """Horner over 8 window sums (some identities, one equal to the running value) with c = 5."""
    m = O.FIELD[k]
    rnd = random.Random(700 + k)
    ident = [1, 1, 0]
    for trial in range(3):
        wins = [rand_pt(rnd, m) for _ in range(8)]
        wins[2] = ident
        if trial == 1:
            wins[7] = ident  # the top window empty: doublings of the identity skipped
        check_same(L, k, 2, wins, [[5, 0, 0]], 8)
