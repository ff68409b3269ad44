"""Seeded synthetic MSM workloads (host side, numpy only).

The benchmark, the parity tests and the fixture script build their inputs from these functions, so
the inputs a test checks are byte for byte the inputs the benchmark times.

* ``plain_inputs`` -- the plain single-MSM workload: 32-byte seeds for SSWU hash-to-curve bases
  (the map itself runs on the device, or in the reference for fixtures), 0.5% of the bases replaced
  by copies of other bases and 0.1% by the identity, scalars uniform below 2^255 and reduced
  modulo the group order, 1% of them forced to 0, 1, order - 1 or a small 8-bit value.
* ``doubling_chain`` / ``fcmpp_scalars`` -- the FCMP++ verification shapes of the reference's
  benchmark (src/benchmark_fcmpp.cpp:51-83, :347-365): points G, 2G, 4G, ... built by successive
  Jacobian doublings (so Z != 1), scalars (i + 1) + 2^24.
"""
import numpy as np

P = (1 << 255) - 19
Q = 0x7FFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFBF7F782CB7656B586EB6D2727927C79F
ORDER = (Q, P)  # Helios order = q, Selene order = p
FIELD = (P, Q)  # Helios coordinates in F_p, Selene in F_q
# generators (reference helios_constants.h / selene_constants.h, test_vectors "generator")
GEN = (
    (3, 0x537b74d97ac0721cbd92668350205f0759003bddc586a5dcd243e639e3183ef4),
    (1, 0x7a19d927b85cca9257c93177455c825f938bb198c8f09b37741e0aa6a1d3fdd2),
)
FCMPP_FIXED = (265, 522)  # Helios, Selene generators shared by every proof
FCMPP_PER_PROOF = 80
FCMPP_BATCHES = (1, 2, 4, 10)


def plain_inputs(curve, n, seed):
    """The plain-MSM workload of one curve.  Returns a dict of numpy arrays:
    u (n, 32) uint8 SSWU inputs; dup_dst, dup_src (base dup_dst[k] becomes a copy of the original
    base dup_src[k]); id_dst (bases replaced by the identity); scalars (n, 32) uint8."""
    rng = np.random.default_rng(seed)
    order = ORDER[curve]
    u = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
    ndup, nid = n // 200, n // 1000
    dst = rng.choice(n, size=ndup + nid, replace=False)
    src = rng.integers(0, n, size=ndup)
    sc = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
    sc[:, 31] &= 0x7F
    # uniform in [0, 2^255); values >= order (probability ~2^-128) are reduced
    top = sc[:, 16:].view(np.uint64)
    for i in np.nonzero((top == np.uint64(0xFFFFFFFFFFFFFFFF)).all(axis=1))[0]:
        v = int.from_bytes(sc[i].tobytes(), "little") % order
        sc[i] = np.frombuffer(v.to_bytes(32, "little"), dtype=np.uint8)
    k = n // 100
    for j, i in enumerate(rng.choice(n, size=k, replace=False)):
        small = int(rng.integers(2, 256))  # drawn for every edge scalar, used by every fourth
        v = (0, 1, order - 1, small)[j % 4]
        sc[i] = np.frombuffer(v.to_bytes(32, "little"), dtype=np.uint8)
    return {"u": u, "dup_dst": dst[:ndup], "dup_src": src, "id_dst": dst[ndup:], "scalars": sc}


IDENTITY_STRUCT = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0], dtype=np.uint64)


def apply_dups_and_identities(points, w):
    """Copies / identities of ``plain_inputs`` applied to (n, 15) Jacobian structs (numpy or a torch
    tensor of int64 on any device; the copies read the bases as they were before any change)."""
    try:
        import torch

        if isinstance(points, torch.Tensor):
            dev = points.device
            dst = torch.from_numpy(w["dup_dst"]).to(dev)
            src = torch.from_numpy(w["dup_src"]).to(dev)
            points[dst] = points[src]
            ident = torch.from_numpy(IDENTITY_STRUCT.view(np.int64)).to(dev)
            points[torch.from_numpy(w["id_dst"]).to(dev)] = ident
            return points
    except ImportError:
        pass
    points[w["dup_dst"]] = points[w["dup_src"]].copy()
    points[w["id_dst"]] = IDENTITY_STRUCT
    return points


def r51(v):
    m = (1 << 51) - 1
    return [(v >> (51 * i)) & m for i in range(5)]


def jacobian_struct(curve, X, Y, Z):
    """Field elements -> the reference's 120-byte Jacobian struct as uint64[15] (canonical limbs)."""
    m = FIELD[curve]
    return np.array(r51(X % m) + r51(Y % m) + r51(Z % m), dtype=np.uint64)


def doubling_chain(curve, n):
    """n Jacobian points G, 2G, 4G, ... by successive dbl-2001-b doublings with a = -3, the formula
    of the reference's helios_dbl / selene_dbl (helios/src/x64/helios_dbl.cpp:32-88), so each point
    is the same projective representative (X : Y : Z) the reference's FCMP++ benchmark builds
    (src/benchmark_fcmpp.cpp:51-69).  Returns uint64 (n, 15)."""
    m = FIELD[curve]
    X, Y, Z = GEN[curve][0], GEN[curve][1], 1
    out = np.zeros((n, 15), dtype=np.uint64)
    for i in range(n):
        out[i] = jacobian_struct(curve, X, Y, Z)
        delta = Z * Z % m
        gamma = Y * Y % m
        beta = X * gamma % m
        alpha = 3 * (X - delta) * (X + delta) % m
        X3 = (alpha * alpha - 8 * beta) % m
        Z3 = ((Y + Z) * (Y + Z) - gamma - delta) % m
        Y3 = (alpha * (4 * beta - X3) - 8 * gamma * gamma) % m
        X, Y, Z = X3, Y3, Z3
    return out


def fcmpp_scalars(n):
    """Scalars of the reference's FCMP++ benchmark (src/benchmark_fcmpp.cpp:72-83): bytes 0..2 =
    i + 1, byte 3 = 1, the rest zero.  Returns uint8 (n, 32)."""
    sc = np.zeros((n, 32), dtype=np.uint8)
    i = np.arange(1, n + 1, dtype=np.uint32)
    sc[:, 0] = i & 0xFF
    sc[:, 1] = (i >> 8) & 0xFF
    sc[:, 2] = (i >> 16) & 0xFF
    sc[:, 3] = 1
    return sc


# FCMP++ tree hashes (src/benchmark_fcmpp.cpp:133-198): a node's hash is an MSM over its layer's
# fixed generators (the benchmark's doubling chains) followed by to_affine; the layers hash 228
# (Selene leaves), 38 (Selene branches) and 18 (Helios branches) children.
TREE_HASH_SHAPES = [(1, 228, "selene_leaf"), (1, 38, "selene_branch"), (0, 18, "helios_branch")]


def tree_layer_scalars(n, count, seed):
    """count nodes of one tree layer with n children each: uniform 255-bit scalars (the children's
    hashes; the library and the reference reduce them mod the group order).  uint8 (count, n, 32)."""
    rng = np.random.default_rng(seed)
    sc = rng.integers(0, 256, size=(count, n, 32), dtype=np.uint8)
    sc[:, :, 31] &= 0x7F
    return sc


def fcmpp_shapes():
    """[(curve, n, batch)] of the FCMP++ verification MSMs: Selene 522 + 80 b and Helios 265 + 80 b
    for b in 1, 2, 4, 10 (src/benchmark_fcmpp.cpp:347-365)."""
    out = []
    for b in FCMPP_BATCHES:
        out.append((1, FCMPP_FIXED[1] + FCMPP_PER_PROOF * b, b))
        out.append((0, FCMPP_FIXED[0] + FCMPP_PER_PROOF * b, b))
    return out
