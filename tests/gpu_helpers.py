"""Shared helpers for the GPU parity tests: input conversion and closed-form workloads."""
import ctypes
import random

import numpy as np

import _oracle as O


def decompress(c, p32):
    """compressed point -> (x, y) ints, or None for 0^32."""
    out = ctypes.create_string_buffer(64)
    rc = O.oracle().hso_decompress(c, out, bytes(p32))
    assert rc in (0, 1), rc
    if rc == 1:
        return None
    return int.from_bytes(out.raw[:32], "little"), int.from_bytes(out.raw[32:], "little")


def to_jac_array(c, points_bytes, n, rnd=None, lazy=False):
    """compressed points -> reference Jacobian structs (n, 15) uint64.  With `rnd`, each point is
    re-expressed with a random Z (X*Z^2, Y*Z^3, Z); with `lazy`, limbs are left unnormalised
    (value + k*m spread over 52-bit limbs), as the reference's lazy arithmetic produces."""
    m = O.FIELD[c]
    out = np.zeros((n, 15), dtype=np.uint64)
    for i in range(n):
        xy = decompress(c, points_bytes[32 * i:32 * i + 32])
        if xy is None:
            out[i, 0] = 1
            out[i, 5] = 1
            if rnd is not None:
                out[i, 0] = rnd.getrandbits(51)  # any X, Y with Z = 0 is the identity
            continue
        x, y = xy
        z = 1
        if rnd is not None:
            z = rnd.randrange(1, m)
            x, y = x * z * z % m, y * z * z * z % m
        vals = (x, y, z)
        for k, v in enumerate(vals):
            if lazy:
                v = v + m * (1 if (i + k) % 3 == 0 else 0)
            limbs = [(v >> (51 * j)) & ((1 << 51) - 1) for j in range(5)]
            top = v >> 255
            limbs[4] += top << 51  # keep the value exact when v >= 2^255
            if lazy and (i + k) % 5 == 1:
                # move one unit of limb j+1 down into limb j (limb j grows past 51 bits)
                if limbs[1] > 0:
                    limbs[1] -= 1
                    limbs[0] += 1 << 51
            out[i, 5 * k:5 * k + 5] = limbs
    return out


def scalars_np(sc_bytes, n):
    return np.frombuffer(bytes(sc_bytes), dtype=np.uint8).reshape(n, 32).copy()


def jac_to_bytes(c, jac):
    """reference Jacobian struct (uint64[15]) with Z = 1 or 0 -> compressed bytes."""
    m = O.FIELD[c]
    z = sum(int(jac[10 + j]) << (51 * j) for j in range(5)) % m
    if z == 0:
        return bytes(32)
    assert z == 1
    x = sum(int(jac[j]) << (51 * j) for j in range(5)) % m
    y = sum(int(jac[5 + j]) << (51 * j) for j in range(5)) % m
    b = bytearray(x.to_bytes(32, "little"))
    b[31] |= (y & 1) << 7
    return bytes(b)


# ---------------------------------------------------------------- arithmetic-progression bases
def _aff_add(c, P, Q):
    m = O.FIELD[c]
    if P is None:
        return Q
    if Q is None:
        return P
    (x1, y1), (x2, y2) = P, Q
    if x1 == x2:
        if (y1 + y2) % m == 0:
            return None
        lam = (3 * x1 * x1 - 3) * pow(2 * y1, -1, m) % m
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, m) % m
    x3 = (lam * lam - x1 - x2) % m
    return x3, (lam * (x1 - x3) - y1) % m


def progression_bases(c, n, seed):
    """Affine points P_i = (a + i*d) G for seeded a, d; returns (points, a, d).  The MSM of any
    scalars over these bases has the closed form (sum s_i (a + i d) mod order) G."""
    rnd = random.Random(seed)
    order = O.ORDER[c]
    a, d = rnd.randrange(1, order), rnd.randrange(1, order)
    g = O.generator(c)
    A = decompress(c, O.scalar_mul(c, a.to_bytes(32, "little"), g))
    D = decompress(c, O.scalar_mul(c, d.to_bytes(32, "little"), g))
    pts = []
    P = A
    for _ in range(n):
        pts.append(P)
        P = _aff_add(c, P, D)
    return pts, a, d


def progression_expected(c, scalars_bytes, n, a, d):
    order = O.ORDER[c]
    tot = 0
    for i in range(n):
        s = int.from_bytes(scalars_bytes[32 * i:32 * i + 32], "little")
        tot += s * (a + i * d)
    k = tot % order
    return O.scalar_mul(c, k.to_bytes(32, "little"), O.generator(c))


def ref_sswu_bases(c, seed, n, dup_frac=0.0, id_frac=0.0):
    """n SSWU images (compressed) of seeded uniform bytes, mapped by the compiled reference's own
    map_to_curve, with dup_frac of them copies of another base and id_frac the identity (0^32).
    None when the compiled reference is not built."""
    import ctypes as _ct

    r = O.ref()
    if r is None:
        return None
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 256, size=n * 32, dtype=np.uint8).tobytes()
    out = _ct.create_string_buffer(32 * n)
    getattr(r, f"hsref_{O.CURVES[c]}_map_to_curve_bytes")(out, u, n)
    pb = bytearray(out.raw)
    for _ in range(int(n * dup_frac)):
        i, src = int(rng.integers(1, n)), int(rng.integers(0, n))
        pb[32 * i:32 * i + 32] = pb[32 * src:32 * src + 32]
    for _ in range(int(n * id_frac)):
        i = int(rng.integers(0, n))
        pb[32 * i:32 * i + 32] = bytes(32)
    return bytes(pb)


# ---------------------------------------------------------------- CPU fallback for the drop-ins
def struct_to_bytes(c, words):
    """reference Jacobian struct (15 uint64, any Z, lazy limbs) -> compressed encoding (Python ints)"""
    m = O.FIELD[c]
    X, Y, Z = (sum(int(words[5 * k + j]) << (51 * j) for j in range(5)) % m for k in range(3))
    if Z == 0:
        return bytes(32)
    zi = pow(Z, -1, m)
    x, y = X * zi * zi % m, Y * zi * zi * zi % m
    b = bytearray(x.to_bytes(32, "little"))
    b[31] |= (y & 1) << 7
    return bytes(b)


def bytes_to_struct(c, p32):
    xy = decompress(c, p32)
    out = np.zeros(15, dtype=np.uint64)
    if xy is None:
        out[0] = out[5] = 1
        return out
    m51 = (1 << 51) - 1
    for k, v in enumerate((xy[0], xy[1], 1)):
        out[5 * k:5 * k + 5] = [(v >> (51 * j)) & m51 for j in range(5)]
    return out


def oracle_fallback(c, calls):
    """A CPU MSM with the drop-in slot's signature, computed by the oracle (test infrastructure):
    stands in for the reference backend a caller registers with hs_gpu_set_cpu_fallback."""
    def fb(res, sc, pts, n):
        calls.append(n)
        s = ctypes.string_at(sc, 32 * n)
        w = np.frombuffer(ctypes.string_at(pts, 120 * n), dtype=np.uint64).reshape(n, 15)
        pb = b"".join(struct_to_bytes(c, w[i]) for i in range(n))
        out = bytes_to_struct(c, O.msm(c, s, pb, n))
        ctypes.memmove(res, out.ctypes.data, 120)
    return fb
