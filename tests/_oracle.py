"""ctypes access to the checker libraries (test infrastructure only).

* ``oracle/_build/libhs_oracle.so`` -- the plain-C restatement of the reference path (oracle/hs_oracle.c).
* ``oracle/_ref/libhsref.so``      -- the reference itself, compiled from its own sources by
  ``make -C oracle ref`` (present when it was built in the development container; optional).

Also holds the seeded input generators used by the parity tests.  Inputs follow the reference
fuzz recipe (src/tests_fuzz.cpp:103-185): xoshiro256** bytes, scalars = reduce_wide(64 bytes),
SSWU bases from 32 random bytes (helios/src/x64/helios_map_to_curve.cpp:184-189).
"""
import ctypes
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libhs_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libhsref.so")

CURVES = ("helios", "selene")
HELIOS, SELENE = 0, 1
# group orders (helios_constants.h:57-64, selene_constants.h:56-64)
ORDER = (
    0x7FFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFBF7F782CB7656B586EB6D2727927C79F,
    (1 << 255) - 19,
)
# coordinate field moduli: Helios is over F_p, Selene over F_q
FIELD = (ORDER[1], ORDER[0])

_o = None
_r = None


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        _o = ctypes.CDLL(ORACLE_SO)
        sz = ctypes.c_size_t
        _o.hso_xoshiro_fill.argtypes = [ctypes.c_uint64, ctypes.c_char_p, sz]
        for f in ("hso_msm", "hso_msm_naive", "hso_msm_fixed"):
            getattr(_o, f).argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz]
        _o.hso_map_to_curve.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, sz]
        _o.hso_batch_invert.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, sz]
        _o.hso_pedersen.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                    ctypes.c_char_p, ctypes.c_char_p, sz]
        _o.hso_ref_window.argtypes = [sz]
    return _o


def ref():
    """The compiled reference, or None when it was not built."""
    global _r
    if _r is None and os.path.exists(REF_SO):
        _r = ctypes.CDLL(REF_SO)
        sz = ctypes.c_size_t
        _r.hsref_init.restype = ctypes.c_uint
        _r.hsref_init()
        for c in CURVES:
            for f in ("msm_bytes", "naive_bytes", "msm_fixed_bytes"):
                getattr(_r, f"hsref_{c}_{f}").argtypes = [ctypes.c_char_p] * 3 + [sz]
            getattr(_r, f"hsref_{c}_pedersen_bytes").argtypes = [ctypes.c_char_p] * 5 + [sz]
            getattr(_r, f"hsref_{c}_map_to_curve_bytes").argtypes = [ctypes.c_char_p] * 2 + [sz]
            getattr(_r, f"hsref_{c}_to_jac").argtypes = [ctypes.c_char_p] * 2 + [sz]
            getattr(_r, f"hsref_{c}_time_msm").argtypes = [ctypes.c_char_p, ctypes.c_char_p, sz,
                                                           ctypes.c_int, ctypes.c_int]
            getattr(_r, f"hsref_{c}_time_msm").restype = ctypes.c_double
            getattr(_r, f"hsref_{c}_dbl_chain").argtypes = [ctypes.c_void_p, sz]
            getattr(_r, f"hsref_{c}_msm_jac_bytes").argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, sz]
            getattr(_r, f"hsref_{c}_map_to_curve_jac").argtypes = [ctypes.c_void_p, ctypes.c_char_p, sz]
    return _r


def _buf(n):
    return ctypes.create_string_buffer(n)


# ---------------------------------------------------------------- oracle calls
def fe_op(name, k, *args):
    out = _buf(32)
    rc = getattr(oracle(), "hso_fe_" + name)(k, out, *[bytes(a) for a in args])
    return out.raw, rc


def point_op(name, c, *args):
    out = _buf(32)
    rc = getattr(oracle(), "hso_point_" + name)(c, out, *[bytes(a) for a in args])
    return (out.raw if rc == 0 else None)


def scalar_mul(c, s, p):
    out = _buf(32)
    rc = oracle().hso_scalar_mul(c, out, bytes(s), bytes(p))
    return out.raw if rc == 0 else None


def generator(c):
    out = _buf(32)
    oracle().hso_generator(c, out)
    return out.raw


def reduce_wide(c, wide):
    out = _buf(32)
    oracle().hso_reduce_wide(c, out, bytes(wide))
    return out.raw


def map_to_curve(c, u):
    n = len(u) // 32
    out = _buf(32 * n)
    oracle().hso_map_to_curve(c, out, bytes(u), n)
    return out.raw


def msm(c, scalars, points, n):
    out = _buf(32)
    rc = oracle().hso_msm(c, out, bytes(scalars), bytes(points), n)
    assert rc == 0, rc
    return out.raw


def msm_naive(c, scalars, points, n):
    out = _buf(32)
    rc = oracle().hso_msm_naive(c, out, bytes(scalars), bytes(points), n)
    assert rc == 0, rc
    return out.raw


def pedersen(c, blinding, H, values, gens, n):
    out = _buf(32)
    rc = oracle().hso_pedersen(c, out, bytes(blinding), bytes(H), bytes(values), bytes(gens), n)
    assert rc == 0, rc
    return out.raw


def msm_fixed(c, scalars, points, n):
    out = _buf(32)
    rc = oracle().hso_msm_fixed(c, out, bytes(scalars), bytes(points), n)
    assert rc == 0, rc
    return out.raw


def ref_window(n):
    return oracle().hso_ref_window(n)


def drops_carry(s, w):
    return oracle().hso_recode_drops_carry(bytes(s), w)


def xoshiro(seed, nbytes):
    out = _buf(nbytes)
    oracle().hso_xoshiro_fill(seed, out, nbytes)
    return out.raw


# ---------------------------------------------------------------- seeded inputs
def gen_scalars(c, seed, n, edge_frac=0.0):
    """n canonical scalars (reduce_wide of 64 xoshiro bytes).  With edge_frac > 0, that fraction
    (at evenly spread indices) is replaced by 0, 1, order-1 and an 8-bit value in turn."""
    raw = xoshiro(seed, 64 * n)
    s = bytearray()
    for i in range(n):
        s += reduce_wide(c, raw[64 * i:64 * i + 64])
    if edge_frac > 0:
        k = max(1, int(n * edge_frac))
        step = max(1, n // k)
        for j, i in enumerate(range(0, n, step)[:k]):
            kind = j % 4
            v = (0, 1, ORDER[c] - 1, raw[64 * i] | 1)[kind]
            s[32 * i:32 * i + 32] = v.to_bytes(32, "little")
    return bytes(s)


def gen_sswu_bases(c, seed, n, dup_frac=0.0, id_frac=0.0):
    """n SSWU images of 32 xoshiro bytes each; dup_frac of them copy an earlier base and id_frac
    are the identity (0^32)."""
    u = xoshiro(seed ^ 0x5555AAAA5555AAAA, 32 * n)
    p = bytearray(map_to_curve(c, u))
    if dup_frac > 0:
        k = int(n * dup_frac)
        for j in range(k):
            i = (j * 7919 + 13) % n
            src = (j * 2654435761 + 7) % i if i > 0 else 1
            p[32 * i:32 * i + 32] = p[32 * src:32 * src + 32]
    if id_frac > 0:
        k = int(n * id_frac)
        for j in range(k):
            i = (j * 104729 + 101) % n
            p[32 * i:32 * i + 32] = bytes(32)
    return bytes(p)


def gen_kG_points(c, seed, n):
    """Points k_i*G for k_i = reduce_wide(64 xoshiro bytes) (tests_fuzz.cpp:177-185)."""
    ks = gen_scalars(c, seed ^ 0x1234, n)
    g = generator(c)
    out = bytearray()
    for i in range(n):
        out += scalar_mul(c, ks[32 * i:32 * i + 32], g)
    return bytes(out)


def ref_dbl_chain(c, n):
    """The reference's own FCMP++ bases G, 2G, 4G, ... (its Jacobian structs, Z != 1), uint64 (n, 15)."""
    import numpy as np

    out = np.zeros((n, 15), dtype=np.uint64)
    getattr(ref(), f"hsref_{CURVES[c]}_dbl_chain")(out.ctypes.data, n)
    return out


def ref_map_to_curve_jac(c, u):
    """The reference's SSWU map of n 32-byte inputs as its Jacobian structs, uint64 (n, 15)."""
    import numpy as np

    n = len(u) // 32
    out = np.zeros((n, 15), dtype=np.uint64)
    getattr(ref(), f"hsref_{CURVES[c]}_map_to_curve_jac")(out.ctypes.data, bytes(u), n)
    return out


def ref_map_to_curve_bytes(c, u):
    """The reference's SSWU map of n 32-byte inputs, encoded (helios_tobytes)."""
    n = len(u) // 32
    out = _buf(32 * n)
    getattr(ref(), f"hsref_{CURVES[c]}_map_to_curve_bytes")(out, bytes(u), n)
    return out.raw


def ref_msm_jac(c, scalars, pts):
    """Reference MSM bytes over the reference's Jacobian structs (uint64 (n, 15)) and scalars
    (uint8 (n, 32))."""
    import numpy as np

    s = np.ascontiguousarray(scalars, dtype=np.uint8)
    p = np.ascontiguousarray(pts, dtype=np.uint64)
    out = _buf(32)
    getattr(ref(), f"hsref_{CURVES[c]}_msm_jac_bytes")(out, s.ctypes.data, p.ctypes.data, p.shape[0])
    return out.raw


def ref_msm(c, scalars, points, n):
    """Reference MSM bytes via the compiled reference, or None when it is not built."""
    r = ref()
    if r is None:
        return None
    out = _buf(32)
    rc = getattr(r, f"hsref_{CURVES[c]}_msm_bytes")(out, bytes(scalars), bytes(points), n)
    assert rc == 0
    return out.raw
