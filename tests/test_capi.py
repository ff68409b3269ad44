"""The C-ABI library loads and exports every entry point include/hs_gpu_msm.h declares (no
compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR_DIR = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "helioselene_amd", "_lib", "libhs_gpu.so")


def declared():
    names = set()
    for f in os.listdir(HDR_DIR):
        if f.endswith(".h"):
            txt = open(os.path.join(HDR_DIR, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(hs_gpu_\w+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    d = declared()
    for n in ("hs_gpu_helios_msm_vartime", "hs_gpu_selene_msm_vartime", "hs_gpu_msm_ex", "hs_gpu_msm_device"):
        assert n in d


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_exports_all_declared_symbols():
    L = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared()) if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_reference_window_table():
    """hs_gpu_reference_window_bits restates pippenger_window_size (helios_msm_vartime.cpp:248-263)."""
    import _oracle as O
    L = ctypes.CDLL(LIB)
    L.hs_gpu_reference_window_bits.argtypes = [ctypes.c_size_t]
    for n in [0, 1, 16, 17, 95, 96, 287, 288, 863, 864, 2591, 2592, 7775, 7776, 23327, 23328, 1 << 20]:
        assert L.hs_gpu_reference_window_bits(n) == O.ref_window(n)


def test_layout_matches_reference_struct():
    """hs_gpu_jacobian is the reference's 120-byte struct (helios/include/helios.h:41-46)."""
    import _oracle as O
    r = O.ref()
    if r is None:
        pytest.skip("reference library not built")
    assert r.hsref_sizeof_jacobian() == 120
