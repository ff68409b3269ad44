"""CPU checks of the autotuner's table (no GPU needed: loading a tuning file and asking which
plan parameters a size would use touch no device)."""
import os

import pytest

hs = pytest.importorskip("helioselene_amd")


@pytest.fixture
def lib_or_skip():
    if not os.path.exists(hs.LIB_PATH):
        pytest.skip("HIP library not built")
    return hs.lib()


def test_tune_file_lookup(lib_or_skip, tmp_path):
    hs.tune_clear()
    c0 = hs.tune_get(0, 0, 1 << 16)
    assert c0 == (12, 32, 8)  # built-in default for 2^16: c = log2 n - 4
    f = tmp_path / "tune.txt"
    f.write_text("0 0 16 11 64 4 1.5\n"
                 "1 1 10 7 16 16 0.9\n"
                 "garbage line\n")
    hs.tune_load(str(f))
    assert hs.tune_get(0, 0, 1 << 16) == (11, 64, 4)
    assert hs.tune_get(0, 0, (1 << 17) - 1) == (11, 64, 4)   # same class
    assert hs.tune_get(0, 0, 1 << 17) == (11, 64, 4)         # neighbouring class
    assert hs.tune_get(0, 0, 1 << 19) == (15, 64, 8)         # too far: default (runs of 64 from 2^19)
    assert hs.tune_get(0, 0, 1 << 10) == (6, 16, 4)          # default below 2^15: short runs and segments
    assert hs.tune_get(1, 0, 1 << 16) == (12, 32, 8)         # other curve: default
    assert hs.tune_get(1, 1, 1 << 10) == (7, 16, 16)
    assert hs.tune_get(0, 1, 1 << 10) == (6, 32, 8)          # batch default for the other curve
    hs.tune_clear()
    assert hs.tune_get(0, 0, 1 << 16) == c0
