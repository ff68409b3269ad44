"""GPU tests of the autotuner: results never depend on the plan parameters it may choose (window
bits, run length, segment size), and a tuning run installs usable parameters."""
import numpy as np
import pytest

import _oracle as O
import gpu_helpers as GH

hs = pytest.importorskip("helioselene_amd")
pytestmark = pytest.mark.gpu


# (14, 32, 1) and (16, 32, 1) give the window reduction its largest shape, 8192 segments in 32
# blocks per window (the plan raises the latter's G to 4, so that no window has more segments);
# (5, 1, 1) and (11, 64, 16) give windows of one block, (13, 7, 2) eight blocks.
@pytest.mark.parametrize("params", [(9, 16, 4), (11, 64, 16), (13, 7, 2), (5, 1, 1), (14, 32, 1), (16, 32, 1)])
def test_forced_parameters_keep_parity(tmp_path, params):
    c, L, G = params
    plan_G = max(G, (1 << (c - 1)) // 8192)
    n = 1500  # also in the reference's window-8 band: compat substitution must still apply
    f = tmp_path / "t.txt"
    f.write_text("".join(f"{m} {cv} 10 {c} {L} {G} 1.0\n" for m in (0, 1) for cv in (0, 1)))
    hs.tune_clear()
    hs.tune_load(str(f))
    try:
        for cv in (0, 1):
            assert hs.tune_get(cv, 0, n) == (c, L, plan_G)
            sc = O.gen_scalars(cv, 900 + c, n, edge_frac=0.02)
            pb = O.gen_sswu_bases(cv, 910 + c, n, dup_frac=0.02, id_frac=0.01)
            pts = GH.to_jac_array(cv, pb, n)
            want = O.msm(cv, sc, pb, n)
            assert hs.msm(cv, GH.scalars_np(sc, n), pts)[0] == want
            got = hs.msm_batch([(cv, GH.scalars_np(sc, n), pts), (1 - cv, np.zeros((0, 32), np.uint8),
                                                                  np.zeros((0, 15), np.uint64))])
            assert got[0][0] == want
    finally:
        hs.tune_clear()


# Every window width, plain and batched: the window sums are combined in groups of up to four
# windows (W = ceil(256 / c) windows; c = 6, 7, 9, 10 and 12 leave a ragged last group, c >= 14 one
# window per group), each group's sums scaled by 2^(c j) and added before the Horner over groups.
@pytest.mark.parametrize("c", list(range(4, 17)))
def test_every_window_width_keeps_parity(tmp_path, c):
    n = 700
    f = tmp_path / "t.txt"
    f.write_text("".join(f"{m} {cv} 9 {c} 16 4 1.0\n" for m in (0, 1) for cv in (0, 1)))
    hs.tune_clear()
    hs.tune_load(str(f))
    try:
        for cv in (0, 1):
            assert hs.tune_get(cv, 0, n)[0] == c
            sc = O.gen_scalars(cv, 300 + c, n, edge_frac=0.02)
            pb = O.gen_sswu_bases(cv, 310 + c, n, dup_frac=0.02, id_frac=0.01)
            pts = GH.to_jac_array(cv, pb, n)
            want = O.msm(cv, sc, pb, n)
            assert hs.msm(cv, GH.scalars_np(sc, n), pts)[0] == want
            got = hs.msm_batch([(cv, GH.scalars_np(sc, n), pts), (1 - cv, np.zeros((0, 32), np.uint8),
                                                                  np.zeros((0, 15), np.uint64))])
            assert got[0][0] == want
    finally:
        hs.tune_clear()


def test_autotune_installs_parameters(tmp_path):
    f = tmp_path / "tune.txt"
    hs.tune_clear()
    try:
        hs.autotune(0, 0, 10, 11, str(f))
        hs.autotune(1, 1, 8, 9, str(f))
        lines = f.read_text().split("\n")
        rows = [list(map(float, ln.split())) for ln in lines if ln.strip()]
        assert [(int(r[0]), int(r[1]), int(r[2])) for r in rows] == [(0, 0, 10), (0, 0, 11), (1, 1, 8), (1, 1, 9)]
        for r in rows:
            assert 4 <= r[3] <= 16 and r[4] >= 1 and r[5] >= 1 and r[6] > 0
        c, L, G = hs.tune_get(0, 0, 1 << 10)
        assert (c, L, G) == tuple(int(x) for x in rows[0][3:6])
        n = 1 << 10
        sc = O.gen_scalars(0, 77, n)
        pb = O.gen_sswu_bases(0, 78, n)
        assert hs.msm(0, GH.scalars_np(sc, n), GH.to_jac_array(0, pb, n))[0] == O.msm(0, sc, pb, n)
    finally:
        hs.tune_clear()
