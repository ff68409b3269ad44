"""helioselene_amd -- MI355X (gfx950) variable-time multi-scalar multiplication for the
Helios / Selene curve cycle, behind the C ABI of include/hs_gpu_msm.h.

This module is a thin ctypes binding of helioselene_amd/_lib/libhs_gpu.so.  It mirrors the
reference's point-level MSM call (HeliosPoint::multi_scalar_mul / helios_msm_vartime,
include/helioselene_point.h:166, helios/include/helios_msm_vartime.h:43-72): inputs are the
reference's byte formats (32-byte scalars, 120-byte radix-2^51 Jacobian points), outputs the
32-byte compressed encoding or the Jacobian struct.  There is no CPU fallback: if the HIP
library is missing or fails to load, every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libhs_gpu.so")

HELIOS, SELENE = 0, 1
CURVES = {"helios": HELIOS, "selene": SELENE, HELIOS: HELIOS, SELENE: SELENE}
EXACT = 1
ERASE = 2  # zero the device copies of scalars and digits after use

# group orders and coordinate moduli (reference helios_constants.h:57-64, selene_constants.h:56-64)
P = (1 << 255) - 19
Q = 0x7FFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFBF7F782CB7656B586EB6D2727927C79F
ORDER = {HELIOS: Q, SELENE: P}
FIELD = {HELIOS: P, SELENE: Q}

_lib = None

# hs_gpu_cpu_msm_fn / hs_gpu_cpu_msm_fixed_fn: (result, scalars, points or tables, n)
CPU_MSM_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)
CPU_MSM_FIXED_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)
_fallbacks = {}


class HsGpuError(RuntimeError):
    pass


def lib():
    """The loaded HIP library (raises when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HsGpuError(f"HIP library not built: {LIB_PATH} (run `python -m helioselene_amd.build`)")
        # PyTorch-ROCm ships its own HIP runtime under the same soname.  Loading torch first makes
        # this library bind to that one copy, so device pointers and stream handles from torch
        # are valid here and torch can still initialise the GPU after this library has.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        sz, vp, cp = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p
        L.hs_gpu_msm_ex.argtypes = [ctypes.c_int, vp, vp, vp, vp, sz, ctypes.c_uint]
        L.hs_gpu_msm_device.argtypes = [ctypes.c_int, vp, vp, sz, ctypes.c_uint, vp, vp, vp]
        L.hs_gpu_helios_msm_vartime.argtypes = [vp, vp, vp, sz]
        L.hs_gpu_selene_msm_vartime.argtypes = [vp, vp, vp, sz]
        L.hs_gpu_last_error.restype = cp
        L.hs_gpu_version.restype = cp
        L.hs_gpu_window_bits.argtypes = [ctypes.c_int, sz]
        L.hs_gpu_reference_window_bits.argtypes = [sz]
        L.hs_gpu_msm_bytes.argtypes = [ctypes.c_int, vp, vp, vp, sz, ctypes.c_uint]
        L.hs_gpu_map_to_curve.argtypes = [ctypes.c_int, vp, sz, vp, vp]
        L.hs_gpu_map_to_curve_device.argtypes = [ctypes.c_int, vp, sz, vp, vp, vp]
        L.hs_gpu_profile.argtypes = [ctypes.c_int]
        L.hs_gpu_phase_times.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int]
        L.hs_gpu_phase_name.restype = cp
        L.hs_gpu_phase_summary.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int)]
        L.hs_gpu_probe_int_mul_peak.argtypes = [ctypes.POINTER(ctypes.c_double)]
        L.hs_gpu_msm_batch.argtypes = [vp, sz, ctypes.c_uint]
        L.hs_gpu_batch_create.argtypes = [vp, sz, ctypes.c_uint, ctypes.POINTER(vp)]
        L.hs_gpu_batch_launch.argtypes = [vp, vp]
        L.hs_gpu_batch_capture.argtypes = [vp]
        L.hs_gpu_batch_points.argtypes = [vp]
        L.hs_gpu_batch_points.restype = sz
        L.hs_gpu_batch_destroy.argtypes = [vp]
        L.hs_gpu_helios_pedersen_commit.argtypes = [vp, vp, vp, vp, vp, sz]
        L.hs_gpu_selene_pedersen_commit.argtypes = [vp, vp, vp, vp, vp, sz]
        L.hs_gpu_helios_msm_fixed.argtypes = [vp, vp, vp, sz]
        L.hs_gpu_selene_msm_fixed.argtypes = [vp, vp, vp, sz]
        L.hs_gpu_fixed_precompute.argtypes = [ctypes.c_int, vp, sz, vp]
        L.hs_gpu_fixed_create.argtypes = [ctypes.c_int, vp, sz, ctypes.POINTER(vp)]
        L.hs_gpu_fixed_create_device.argtypes = [ctypes.c_int, vp, sz, vp, ctypes.POINTER(vp)]
        L.hs_gpu_fixed_size.argtypes = [vp]
        L.hs_gpu_fixed_size.restype = sz
        L.hs_gpu_fixed_msm.argtypes = [vp, vp, sz, ctypes.c_uint, ctypes.c_int, vp, vp]
        L.hs_gpu_fixed_msm_device.argtypes = [vp, vp, sz, ctypes.c_uint, ctypes.c_int, vp, vp, vp]
        L.hs_gpu_pedersen_commit_batch.argtypes = [vp, vp, vp, sz, ctypes.c_uint, ctypes.c_int, vp, vp]
        L.hs_gpu_fixed_destroy.argtypes = [vp]
        L.hs_gpu_server_create.argtypes = [vp, ctypes.POINTER(vp)]
        L.hs_gpu_server_submit.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        L.hs_gpu_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int]
        L.hs_gpu_server_get_stats.argtypes = [vp, vp]
        L.hs_gpu_server_reset_stats.argtypes = [vp]
        L.hs_gpu_server_destroy.argtypes = [vp]
        L.hs_gpu_autotune.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, cp]
        L.hs_gpu_tune_load.argtypes = [cp]
        L.hs_gpu_tune_get.argtypes = [ctypes.c_int, ctypes.c_int, sz, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.hs_gpu_debug_dump.argtypes = [ctypes.c_int, ctypes.c_int, vp, sz, ctypes.POINTER(ctypes.c_size_t)]
        L.hs_gpu_set_cpu_fallback.argtypes = [ctypes.c_int, CPU_MSM_FN, CPU_MSM_FIXED_FN]
        L.hs_gpu_fallback_count.restype = ctypes.c_uint64
        L.hs_gpu_set_device_limit.argtypes = [sz]
        L.hs_gpu_device_info.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), cp, sz]
        L.hs_gpu_device_bytes.restype = sz
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise HsGpuError(f"hs_gpu error {rc}: {lib().hs_gpu_last_error().decode()}")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------- format helpers (host)
def fe_to_r51(v):
    m = (1 << 51) - 1
    return [(v >> (51 * i)) & m for i in range(5)]


def _words_to_r51(w):
    """uint64 words (n, 4) of values < 2^256 -> radix-2^51 limbs (n, 5) (top limb keeps bits 204+)."""
    m = np.uint64((1 << 51) - 1)
    out = np.empty((w.shape[0], 5), dtype=np.uint64)
    s = np.uint64
    out[:, 0] = w[:, 0] & m
    out[:, 1] = ((w[:, 0] >> s(51)) | (w[:, 1] << s(13))) & m
    out[:, 2] = ((w[:, 1] >> s(38)) | (w[:, 2] << s(26))) & m
    out[:, 3] = ((w[:, 2] >> s(25)) | (w[:, 3] << s(39))) & m
    out[:, 4] = w[:, 3] >> s(12)
    return out


def jacobian_array(points_xy):
    """[(x, y) or None] affine points (ints) -> uint64 array (n, 15) in the reference struct
    layout (Z = 1, or X = Y = 1, Z = 0 for the identity)."""
    n = len(points_xy)
    xb = bytearray(32 * n)
    yb = bytearray(32 * n)
    ident = np.zeros(n, dtype=bool)
    for i, p in enumerate(points_xy):
        if p is None:
            ident[i] = True
            continue
        xb[32 * i:32 * i + 32] = p[0].to_bytes(32, "little")
        yb[32 * i:32 * i + 32] = p[1].to_bytes(32, "little")
    return jacobian_from_affine_bytes(bytes(xb), bytes(yb), ident)


def jacobian_from_affine_bytes(xb, yb, ident=None):
    """Affine coordinates as n x 32-byte little-endian strings -> reference Jacobian structs."""
    xw = np.frombuffer(xb, dtype="<u8").reshape(-1, 4)
    yw = np.frombuffer(yb, dtype="<u8").reshape(-1, 4)
    n = xw.shape[0]
    out = np.zeros((n, 15), dtype=np.uint64)
    out[:, 0:5] = _words_to_r51(xw)
    out[:, 5:10] = _words_to_r51(yw)
    out[:, 10] = 1
    if ident is not None and ident.any():
        out[ident] = 0
        out[ident, 0] = 1
        out[ident, 5] = 1
    return out


def scalars_array(scalars):
    """ints or 32-byte strings -> uint8 array (n, 32) little-endian."""
    buf = bytearray()
    for s in scalars:
        buf += s if isinstance(s, (bytes, bytearray)) else int(s).to_bytes(32, "little")
    return np.frombuffer(bytes(buf), dtype=np.uint8).reshape(-1, 32).copy()


FIXED_AUTO, FIXED_COMB, FIXED_PIPPENGER = 0, 1, 2

# ---------------------------------------------------------------- MSM calls
def msm(curve, scalars, points, flags=0):
    """Plain MSM through the status-returning slot.  scalars: uint8 (n, 32); points: uint64
    (n, 15) reference Jacobian structs.  Returns (compressed 32 bytes, jacobian uint64[15])."""
    c = CURVES[curve]
    s = np.ascontiguousarray(scalars, dtype=np.uint8)
    p = np.ascontiguousarray(points, dtype=np.uint64)
    n = p.shape[0]
    assert s.shape == (n, 32) and p.shape == (n, 15)
    out = np.zeros(32, dtype=np.uint8)
    jac = np.zeros(15, dtype=np.uint64)
    _check(lib().hs_gpu_msm_ex(c, _ptr(jac), _ptr(out), _ptr(s), _ptr(p), n, flags))
    return out.tobytes(), jac


def msm_slot(curve, scalars, points):
    """The void drop-in slot (hs_gpu_helios_msm_vartime / hs_gpu_selene_msm_vartime)."""
    c = CURVES[curve]
    s = np.ascontiguousarray(scalars, dtype=np.uint8)
    p = np.ascontiguousarray(points, dtype=np.uint64)
    jac = np.zeros(15, dtype=np.uint64)
    f = lib().hs_gpu_helios_msm_vartime if c == HELIOS else lib().hs_gpu_selene_msm_vartime
    f(_ptr(jac), _ptr(s), _ptr(p), p.shape[0])
    return jac


def msm_device(curve, d_scalars, d_points, n, d_out32, d_out_jac=None, flags=0, stream=None):
    """Device-resident MSM: arguments are device addresses (ints, e.g. torch tensor data_ptr())."""
    _check(lib().hs_gpu_msm_device(CURVES[curve], d_scalars, d_points, n, flags, d_out32, d_out_jac, stream))


def set_cpu_fallback(curve, msm=None, msm_fixed=None):
    """Registers CPU functions the void drop-ins call when the device fails (hs_gpu_set_cpu_fallback).
    msm(result_addr, scalars_addr, points_addr, n) and msm_fixed(result_addr, scalars_addr,
    tables_addr, n) receive raw addresses; None clears.  The callbacks are kept alive here."""
    c = CURVES[curve]
    f1 = CPU_MSM_FN(msm) if msm else CPU_MSM_FN()
    f2 = CPU_MSM_FIXED_FN(msm_fixed) if msm_fixed else CPU_MSM_FIXED_FN()
    _fallbacks[c] = (f1, f2)
    _check(lib().hs_gpu_set_cpu_fallback(c, f1, f2))


def fallback_count():
    return lib().hs_gpu_fallback_count()


def set_device_limit(nbytes):
    """Cap on the device bytes the library may hold (0 = none)."""
    _check(lib().hs_gpu_set_device_limit(nbytes))


def device_bytes():
    return lib().hs_gpu_device_bytes()


def trim():
    """Frees the library's cached workspaces (after waiting for its streams)."""
    _check(lib().hs_gpu_trim())


def synchronize():
    _check(lib().hs_gpu_synchronize())


def window_bits(curve, n):
    return lib().hs_gpu_window_bits(CURVES[curve], n)


def reference_window_bits(n):
    return lib().hs_gpu_reference_window_bits(n)


def msm_bytes(curve, scalars, points32, flags=0):
    """MSM over compressed points (decoded on the device).  scalars: bytes/uint8 (n*32);
    points32: bytes (n*32).  Returns the 32-byte result."""
    s = np.frombuffer(bytes(scalars), dtype=np.uint8)
    p = np.frombuffer(bytes(points32), dtype=np.uint8)
    n = p.size // 32
    out = np.zeros(32, dtype=np.uint8)
    _check(lib().hs_gpu_msm_bytes(CURVES[curve], _ptr(out), _ptr(s), _ptr(p), n, flags))
    return out.tobytes()


def map_to_curve(curve, u, want_points=True, want_bytes=False):
    """SSWU hash-to-curve of n 32-byte inputs on the GPU.  Returns (jacobian uint64 (n,15) or
    None, compressed bytes or None)."""
    ub = np.frombuffer(bytes(u), dtype=np.uint8)
    n = ub.size // 32
    pts = np.zeros((n, 15), dtype=np.uint64) if want_points else None
    enc = np.zeros(n * 32, dtype=np.uint8) if want_bytes else None
    _check(lib().hs_gpu_map_to_curve(CURVES[curve], _ptr(ub), n, _ptr(pts), _ptr(enc)))
    return pts, (enc.tobytes() if enc is not None else None)


def map_to_curve_device(curve, d_u, n, d_points, d_out32=None, stream=None):
    _check(lib().hs_gpu_map_to_curve_device(CURVES[curve], d_u, n, d_points, d_out32, stream))


def profile(on=True):
    _check(lib().hs_gpu_profile(1 if on else 0))


def phase_times():
    """{phase: ms} of the last plain MSM (after profile(True))."""
    buf = (ctypes.c_float * 16)()
    k = lib().hs_gpu_phase_times(buf, 16)
    return {lib().hs_gpu_phase_name(i).decode(): float(buf[i]) for i in range(k)}


def probe_int_mul_peak():
    """Measured 32x32->64 multiply throughput of the device (partial products per second)."""
    v = ctypes.c_double()
    _check(lib().hs_gpu_probe_int_mul_peak(ctypes.byref(v)))
    return v.value


def device_info():
    """(compute units, peak clock kHz, device name) of the library's device."""
    cu, clk = ctypes.c_int(), ctypes.c_int()
    name = ctypes.create_string_buffer(256)
    _check(lib().hs_gpu_device_info(ctypes.byref(cu), ctypes.byref(clk), name, 256))
    return cu.value, clk.value, name.value.decode()


def debug_dump(curve, which, max_items):
    """Intermediate state of the last plain MSM: which 0/1/2 -> list of 32-byte encodings of the
    window / segment / bucket sums; 3/4 -> uint32 array of sorted entries / bucket offsets."""
    item = 32 if which in (0, 1, 2, 5) else 4
    buf = np.zeros(max_items * item, dtype=np.uint8)
    cnt = ctypes.c_size_t()
    _check(lib().hs_gpu_debug_dump(CURVES[curve], which, _ptr(buf), buf.size, ctypes.byref(cnt)))
    if which in (0, 1, 2, 5):
        return [buf[32 * i:32 * i + 32].tobytes() for i in range(cnt.value)]
    return buf.view(np.uint32)[:cnt.value].copy()


def phase_summary(curve=-1):
    """({phase: total ms}, calls) over all MSMs profiled since profile(True) (curve -1: both)."""
    buf = (ctypes.c_float * 16)()
    calls = ctypes.c_int()
    c = -1 if curve == -1 else CURVES[curve]
    k = lib().hs_gpu_phase_summary(c, buf, 16, ctypes.byref(calls))
    if k < 0:
        _check(k)
    return {lib().hs_gpu_phase_name(i).decode(): float(buf[i]) for i in range(k)}, calls.value


class MsmDesc(ctypes.Structure):
    """hs_gpu_msm_desc of include/hs_gpu_msm.h."""
    _fields_ = [("curve", ctypes.c_int), ("n", ctypes.c_size_t), ("scalars", ctypes.c_void_p),
                ("points", ctypes.c_void_p), ("result", ctypes.c_void_p), ("out32", ctypes.c_void_p)]


def msm_batch(items, flags=0):
    """Synchronous batch over host arrays.  items: [(curve, scalars uint8 (n,32), points uint64
    (n,15))].  Returns [(compressed 32 bytes, jacobian uint64[15])]."""
    keep, descs, outs = [], (MsmDesc * max(1, len(items)))(), []
    for k, (curve, sc, pts) in enumerate(items):
        s = np.ascontiguousarray(sc, dtype=np.uint8)
        p = np.ascontiguousarray(pts, dtype=np.uint64)
        o = np.zeros(32, dtype=np.uint8)
        j = np.zeros(15, dtype=np.uint64)
        keep += [s, p]
        outs.append((o, j))
        descs[k] = MsmDesc(CURVES[curve], p.shape[0], s.ctypes.data, p.ctypes.data, j.ctypes.data, o.ctypes.data)
    _check(lib().hs_gpu_msm_batch(ctypes.cast(descs, ctypes.c_void_p), len(items), flags))
    return [(o.tobytes(), j) for o, j in outs]


class Batch:
    """A batch plan over DEVICE buffers (addresses as ints): create once, launch many times,
    optionally replayed from one captured hipGraph."""

    def __init__(self, items, flags=0):
        # items: [(curve, n, d_scalars, d_points, d_out32, d_outjac or None)]
        self._descs = (MsmDesc * max(1, len(items)))()
        for k, (curve, n, ds, dp, do, dj) in enumerate(items):
            self._descs[k] = MsmDesc(CURVES[curve], n, ds, dp, dj, do)
        h = ctypes.c_void_p()
        _check(lib().hs_gpu_batch_create(ctypes.cast(self._descs, ctypes.c_void_p), len(items), flags,
                                         ctypes.byref(h)))
        self.handle = h

    @property
    def points(self):
        return lib().hs_gpu_batch_points(self.handle)

    def launch(self, stream=None):
        _check(lib().hs_gpu_batch_launch(self.handle, stream))

    def capture(self):
        _check(lib().hs_gpu_batch_capture(self.handle))

    def close(self):
        if self.handle:
            lib().hs_gpu_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- Pedersen / fixed-base
def pedersen_commit(curve, blinding, H, values, generators):
    """The void drop-in (hs_gpu_helios_pedersen_commit / selene): blinding 32 bytes, H uint64[15],
    values uint8 (n, 32), generators uint64 (n, 15).  Returns the result struct uint64[15]."""
    c = CURVES[curve]
    b = np.frombuffer(bytes(blinding), dtype=np.uint8).copy()
    h = np.ascontiguousarray(H, dtype=np.uint64).reshape(15)
    v = np.ascontiguousarray(values, dtype=np.uint8).reshape(-1, 32)
    g = np.ascontiguousarray(generators, dtype=np.uint64).reshape(-1, 15)
    res = np.zeros(15, dtype=np.uint64)
    f = lib().hs_gpu_helios_pedersen_commit if c == HELIOS else lib().hs_gpu_selene_pedersen_commit
    f(_ptr(res), _ptr(b), _ptr(h), _ptr(v), _ptr(g), g.shape[0])
    return res


def fixed_precompute(curve, points):
    """Width-5 tables of n bases: uint64 (n, 16, 10), entry k of base i = (k+1) P_i as the
    reference's affine struct (x[5], y[5] radix 2^51)."""
    p = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1, 15)
    t = np.zeros((p.shape[0], 16, 10), dtype=np.uint64)
    _check(lib().hs_gpu_fixed_precompute(CURVES[curve], _ptr(p), p.shape[0], _ptr(t)))
    return t


def msm_fixed(curve, scalars, tables):
    """The void drop-in (hs_gpu_helios_msm_fixed / selene) over per-base tables uint64 (n, 16, 10);
    passes an array of n table pointers like the reference's `const helios_affine *const *`."""
    c = CURVES[curve]
    s = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1, 32)
    t = np.ascontiguousarray(tables, dtype=np.uint64).reshape(-1, 16, 10)
    n = t.shape[0]
    ptrs = (ctypes.c_void_p * max(1, n))(*[t.ctypes.data + i * 16 * 80 for i in range(n)])
    res = np.zeros(15, dtype=np.uint64)
    f = lib().hs_gpu_helios_msm_fixed if c == HELIOS else lib().hs_gpu_selene_msm_fixed
    f(_ptr(res), _ptr(s), ctypes.cast(ptrs, ctypes.c_void_p), n)
    return res


class FixedSet:
    """A resident fixed-base set (bases and their tables kept in HBM: a 16-tooth comb up to 4608
    bases, windows of 14..5 bits above, at most 9 GiB of tables)."""

    def __init__(self, curve, points=None, d_points=None, n=None, stream=None):
        self.curve = CURVES[curve]
        h = ctypes.c_void_p()
        if points is not None:
            p = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1, 15)
            _check(lib().hs_gpu_fixed_create(self.curve, _ptr(p), p.shape[0], ctypes.byref(h)))
        else:
            _check(lib().hs_gpu_fixed_create_device(self.curve, d_points, n, stream, ctypes.byref(h)))
        self.handle = h
        self.n = lib().hs_gpu_fixed_size(h)

    def msm(self, scalars, flags=0, method=FIXED_AUTO):
        """scalars uint8 (count, n, 32) -> (list of 32-byte results, uint64 (count, 15))."""
        s = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1, self.n, 32)
        k = s.shape[0]
        o = np.zeros((k, 32), dtype=np.uint8)
        j = np.zeros((k, 15), dtype=np.uint64)
        _check(lib().hs_gpu_fixed_msm(self.handle, _ptr(s), k, flags, method, _ptr(o), _ptr(j)))
        return [o[i].tobytes() for i in range(k)], j

    def pedersen(self, blindings, values, flags=0, method=FIXED_AUTO):
        """Set = [H, G_0..G_{n-2}]: blindings (count, 32), values (count, n-1, 32)."""
        b = np.ascontiguousarray(blindings, dtype=np.uint8).reshape(-1, 32)
        k = b.shape[0]
        v = np.ascontiguousarray(values, dtype=np.uint8).reshape(k, self.n - 1, 32)
        o = np.zeros((k, 32), dtype=np.uint8)
        j = np.zeros((k, 15), dtype=np.uint64)
        _check(lib().hs_gpu_pedersen_commit_batch(self.handle, _ptr(b), _ptr(v), k, flags, method, _ptr(o), _ptr(j)))
        return [o[i].tobytes() for i in range(k)], j

    def msm_device(self, d_scalars, count, d_out32, d_out_jac=None, flags=0, method=FIXED_AUTO, stream=None):
        _check(lib().hs_gpu_fixed_msm_device(self.handle, d_scalars, count, flags, method, d_out32, d_out_jac, stream))

    def close(self):
        if self.handle:
            lib().hs_gpu_fixed_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- serving loop
class ServerConfig(ctypes.Structure):
    _fields_ = [("device_bytes", ctypes.c_size_t), ("lanes", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("flags", ctypes.c_uint)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("points", ctypes.c_uint64),
                ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double), ("max_us", ctypes.c_double),
                ("mean_us", ctypes.c_double), ("elapsed_s", ctypes.c_double)]


class Server:
    """The native serving loop (hs_gpu_server_*): submit from any thread, wait on tickets."""

    def __init__(self, device_bytes=0, lanes=0, max_batch=0, flags=0):
        cfg = ServerConfig(device_bytes, lanes, max_batch, flags)
        h = ctypes.c_void_p()
        _check(lib().hs_gpu_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._live = {}

    def submit(self, curve, scalars=None, points=None, d_scalars=None, d_points=None, n=None):
        """Host arrays (scalars uint8 (n, 32), points uint64 (n, 15)) or device addresses with n."""
        out = np.zeros(32, dtype=np.uint8)
        jac = np.zeros(15, dtype=np.uint64)
        if d_scalars is not None:
            keep = ()
            d = MsmDesc(CURVES[curve], n, d_scalars, d_points, jac.ctypes.data, out.ctypes.data)
            dev = 1
        else:
            s = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1, 32)
            p = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1, 15)
            keep = (s, p)
            d = MsmDesc(CURVES[curve], p.shape[0], s.ctypes.data, p.ctypes.data, jac.ctypes.data, out.ctypes.data)
            dev = 0
        t = ctypes.c_uint64()
        _check(lib().hs_gpu_server_submit(self.handle, ctypes.byref(d), dev, ctypes.byref(t)))
        self._live[t.value] = (keep, out, jac)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """(32-byte result, jacobian uint64[15]); None on timeout."""
        rc = lib().hs_gpu_server_wait(self.handle, ticket, timeout_ms)
        if rc == 1:
            return None
        _, out, jac = self._live.pop(ticket)
        if rc != 0:
            raise HsGpuError(f"hs_gpu request {ticket} failed with status {rc}")
        return out.tobytes(), jac

    def stats(self):
        st = ServerStats()
        _check(lib().hs_gpu_server_get_stats(self.handle, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(lib().hs_gpu_server_reset_stats(self.handle))

    def close(self):
        if self.handle:
            lib().hs_gpu_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- autotuner
def autotune(curve, mode, lg_lo, lg_hi, path=None):
    """Measure and install the fastest (c, L, G) per size class 2^lg_lo..2^lg_hi (mode 0 plain,
    1 batch groups); optionally append them to a tuning file."""
    _check(lib().hs_gpu_autotune(CURVES[curve], mode, lg_lo, lg_hi, path.encode() if path else None))


def tune_load(path):
    _check(lib().hs_gpu_tune_load(path.encode()))


def tune_get(curve, mode, n):
    c, L, G = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib().hs_gpu_tune_get(CURVES[curve], mode, n, ctypes.byref(c), ctypes.byref(L), ctypes.byref(G)))
    return c.value, L.value, G.value


def tune_clear():
    _check(lib().hs_gpu_tune_clear())
