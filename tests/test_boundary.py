"""Failure semantics of the void drop-in slots without a device (CPU only; skipped where a GPU is
visible).  The slot never returns a wrong point: it answers through the registered CPU fallback,
and with none registered it stops the process with a message."""
import os
import subprocess
import sys

import numpy as np
import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gpu_visible():
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


pytestmark = pytest.mark.skipif(_gpu_visible() or not os.path.exists(hs.LIB_PATH),
                                reason="checks the no-device path; needs the built library and no GPU")


def test_slot_without_device_and_fallback_aborts():
    code = ("import numpy as np, helioselene_amd as hs\n"
            "hs.msm_slot(0, np.zeros((2, 32), np.uint8), np.zeros((2, 15), np.uint64))\n"
            "print('returned')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == -6, (r.returncode, r.stdout, r.stderr)
    assert "no CPU fallback is registered" in r.stderr and "returned" not in r.stdout


def test_status_form_reports_the_device_error():
    with pytest.raises(hs.HsGpuError, match="-2"):
        hs.msm(0, np.zeros((2, 32), np.uint8), np.zeros((2, 15), np.uint64))


@pytest.mark.parametrize("c", [0, 1])
def test_slot_without_device_answers_through_the_fallback(c):
    n = 20
    sc = O.gen_scalars(c, 3, n, edge_frac=0.1)
    pb = O.gen_sswu_bases(c, 3, n, dup_frac=0.1, id_frac=0.1)
    pts = GH.to_jac_array(c, pb, n, rnd=__import__("random").Random(5), lazy=True)
    calls = []
    hs.set_cpu_fallback(c, GH.oracle_fallback(c, calls))
    try:
        before = hs.fallback_count()
        jac = hs.msm_slot(c, GH.scalars_np(sc, n), pts)
        assert calls == [n] and hs.fallback_count() == before + 1
        assert GH.jac_to_bytes(c, jac) == O.msm(c, sc, pb, n)
    finally:
        hs.set_cpu_fallback(c)
