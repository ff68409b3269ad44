"""GPU test of the no-allocation entry point hs_gpu_msm_device_noalloc, called through the C ABI:
two calls in a row per curve, the second with a smaller input, each against hs_gpu_msm_device on
the same device inputs; no device bytes allocated by the calls; the size limit rejected with a
clear error."""
import ctypes

import numpy as np
import pytest

hs = pytest.importorskip("helioselene_amd")
torch = pytest.importorskip("torch")
from helioselene_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


def _noalloc():
    f = hs.lib().hs_gpu_msm_device_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p]
    return f


def _inputs(c, n, seed, dev):
    w = W.plain_inputs(c, n, seed)
    u = torch.from_numpy(w["u"]).to(dev)
    pts = torch.empty((n, 15), dtype=torch.int64, device=dev)
    hs.map_to_curve_device(c, u.data_ptr(), n, pts.data_ptr(), None, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    W.apply_dups_and_identities(pts, w)
    return torch.from_numpy(w["scalars"]).to(dev), pts


@pytest.mark.parametrize("c", [0, 1])
def test_noalloc_matches_msm_device(c):
    dev = torch.device("cuda:0")
    f = _noalloc()
    assert hs.lib().hs_gpu_init() == 0
    st = torch.cuda.current_stream().cuda_stream
    cases = [(1 << 16, 5 + c), (1500, 6 + c)]  # full size, then a smaller input (the window-8 band)
    ins = [_inputs(c, n, seed, dev) for n, seed in cases]
    got = []
    torch.cuda.synchronize()
    held = hs.device_bytes()
    for (n, _), (sc, pts) in zip(cases, ins):
        for flags in (0, hs.EXACT):
            o = torch.zeros(8, dtype=torch.int32, device=dev)
            j = torch.zeros(15, dtype=torch.int64, device=dev)
            assert f(c, sc.data_ptr(), pts.data_ptr(), n, flags, o.data_ptr(), j.data_ptr(), st) == 0, \
                hs.lib().hs_gpu_last_error()
            torch.cuda.synchronize()
            got.append((o.cpu().numpy().tobytes(), j.cpu().numpy().tobytes()))
    assert hs.device_bytes() == held  # the calls allocated nothing
    k = 0
    for (n, _), (sc, pts) in zip(cases, ins):
        for flags in (0, hs.EXACT):
            o = torch.zeros(8, dtype=torch.int32, device=dev)
            j = torch.zeros(15, dtype=torch.int64, device=dev)
            hs.msm_device(c, sc.data_ptr(), pts.data_ptr(), n, o.data_ptr(), j.data_ptr(), flags, st)
            torch.cuda.synchronize()
            assert got[k] == (o.cpu().numpy().tobytes(), j.cpu().numpy().tobytes()), (n, flags)
            k += 1
    assert got[2][0] != got[3][0]  # n = 1500: the reference's window-8 bytes differ from the exact sum


def test_noalloc_rejects_oversize():
    dev = torch.device("cuda:0")
    f = _noalloc()
    n = (1 << 16) + 1
    sc = torch.zeros((n, 32), dtype=torch.uint8, device=dev)
    pts = torch.zeros((n, 15), dtype=torch.int64, device=dev)
    o = torch.zeros(8, dtype=torch.int32, device=dev)
    assert f(0, sc.data_ptr(), pts.data_ptr(), n, 0, o.data_ptr(), None, None) == -1
    assert b"at most 65536 points" in hs.lib().hs_gpu_last_error()
