"""GPU tests of the plan variants the environment selects for measurements: window-group pipelining
(HS_GPU_Q, the combine of the top windows overlapping the bucket sums of the lower ones) and the
quad-of-lanes Horner kernels (HS_GPU_HORNER=quad, read once when the library loads, so that case
runs in a process of its own).  Their results must be the oracle's, like the default path's."""
import os
import subprocess
import sys

import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")


# n = 300 plans c = 4 (64 windows), n = 5000 c = 8 (32 windows): with Q groups the Horner runs in
# Q pieces, each window on its own (no window groups), the last piece finalizing
@pytest.mark.parametrize("Q", [2, 3])
def test_window_group_pipelining(monkeypatch, Q):
    monkeypatch.setenv("HS_GPU_Q", str(Q))
    for c in (0, 1):
        for n in (300, 5000):
            sc = O.gen_scalars(c, 400 + n, n, edge_frac=0.02)
            pb = O.gen_sswu_bases(c, 410 + n, n, dup_frac=0.02, id_frac=0.01)
            print(f"curve {c} n={n} Q={Q}", flush=True)
            assert hs.msm(c, GH.scalars_np(sc, n), GH.to_jac_array(c, pb, n))[0] == O.msm(c, sc, pb, n)


QUAD_CODE = r"""
import sys
sys.path[:0] = [{root!r}, {tests!r}]
import helioselene_amd as hs, _oracle as O, gpu_helpers as GH
items = []
for c in (0, 1):
    for n in (100, 3000):
        sc = O.gen_scalars(c, 500 + n, n, edge_frac=0.02)
        pb = O.gen_sswu_bases(c, 510 + n, n, dup_frac=0.02, id_frac=0.01)
        pts = GH.to_jac_array(c, pb, n)
        want = O.msm(c, sc, pb, n)
        print("plain", c, n, flush=True)
        assert hs.msm(c, GH.scalars_np(sc, n), pts)[0] == want, (c, n)
        items.append((c, GH.scalars_np(sc, n), pts, want))
print("batch", flush=True)
got = hs.msm_batch([(c, s, p) for c, s, p, _ in items])
assert [g[0] for g in got] == [w for *_, w in items]
print("ok", flush=True)
"""


def test_quad_horner_kernels():
    env = dict(os.environ, HS_GPU_HORNER="quad")
    r = subprocess.run([sys.executable, "-c", QUAD_CODE.format(root=ROOT, tests=TESTS)], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
