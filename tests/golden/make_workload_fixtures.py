"""Writes tests/golden/workload_runs.json: the compiled reference's own results on the benchmark's
workloads (run in the development container, where ``make -C oracle ref`` has built the reference
library from the reference's sources).

Inputs are regenerated from seeds by helioselene_amd/workloads.py, the generator bench.py uses:

* ``plain``: the plain single-MSM workload at n = 2^16, 2^18, 2^20 for both curves with the seeds
  bench.py uses (1 for Helios, 18 for Selene), plus three served-request sized cases.  Bases are
  the reference's own SSWU images (helios/src/x64/helios_map_to_curve.cpp:184-198) of the seeded
  32-byte strings, with the seeded copies and identities applied; the result is the reference's
  helios_msm_vartime / selene_msm_vartime (helios/src/x64/helios_msm_vartime.cpp:396-416) on its
  own Jacobian structs, encoded by helios_tobytes.
* ``fcmpp``: the FCMP++ verification shapes of src/benchmark_fcmpp.cpp:347-365 on the reference's
  own doubling-chain bases (Z != 1) and benchmark scalars.
* ``tree_hash``: FCMP++ tree hashes (src/benchmark_fcmpp.cpp:133-198): nodes of a layer over the
  doubling-chain generators (Selene 228 and 38, Helios 18) with seeded uniform scalars, each the
  reference's MSM of one node's children.
* ``fixed_degenerate``: helios_msm_fixed / selene_msm_fixed (helios/include/helios_msm_fixed.h:60-155)
  on repeated bases with repeated scalars, where its madd-2007-bl meets an accumulator equal to the
  table entry, next to the exact sum.

    python tests/golden/make_workload_fixtures.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import _oracle as O  # noqa: E402
from helioselene_amd import workloads as W  # noqa: E402

PLAIN_CASES = [(0, 1 << 16, 1), (1, 1 << 16, 18), (0, 1 << 18, 1), (1, 1 << 18, 18), (0, 1 << 20, 1),
               (1, 1 << 20, 18),
               # sizes of served requests (log-uniform in [32, 2^18])
               (0, 200000, 7), (1, 150000, 8), (1, 1 << 18, 9)]


def plain_case(c, n, seed):
    w = W.plain_inputs(c, n, seed)
    pts = O.ref_map_to_curve_jac(c, w["u"].tobytes())
    W.apply_dups_and_identities(pts, w)
    t0 = time.perf_counter()
    r = O.ref_msm_jac(c, w["scalars"], pts)
    return {"kind": "plain", "curve": O.CURVES[c], "n": n, "seed": seed, "generator": "workloads.plain_inputs",
            "ref_result": r.hex(), "ref_seconds": round(time.perf_counter() - t0, 2)}


def fcmpp_case(c, n, b):
    pts = O.ref_dbl_chain(c, n)
    sc = W.fcmpp_scalars(n)
    return {"kind": "fcmpp", "curve": O.CURVES[c], "n": n, "batch": b, "generator": "workloads.doubling_chain",
            "ref_result": O.ref_msm_jac(c, sc, pts).hex()}


TREE_NODES = 4


def tree_hash_cases():
    out = []
    for c, n, name in W.TREE_HASH_SHAPES:
        pts = O.ref_dbl_chain(c, n)
        seed = 7000 + n
        sc = W.tree_layer_scalars(n, TREE_NODES, seed)
        for h in range(TREE_NODES):
            out.append({"kind": "tree_hash", "curve": O.CURVES[c], "n": n, "layer": name, "seed": seed, "index": h,
                        "generator": "workloads.doubling_chain, workloads.tree_layer_scalars",
                        "ref_result": O.ref_msm_jac(c, sc[h], pts).hex()})
    return out


def fixed_degenerate_cases():
    out = []
    for c in (0, 1):
        g = O.generator(c)
        P = O.scalar_mul(c, (0x1234567).to_bytes(32, "little"), g)
        Qp = O.scalar_mul(c, (0x89abcdef).to_bytes(32, "little"), g)
        s = (0x0BADC0FFEE12345).to_bytes(32, "little")
        t = (0x777).to_bytes(32, "little")
        for label, pts, sc in (("[P,P] x [s,s]", [P, P], [s, s]), ("[P,Q,P] x [s,t,s]", [P, Qp, P], [s, t, s])):
            pb, sb = b"".join(pts), b"".join(sc)
            n = len(pts)
            ref = _buf_fixed(c, sb, pb, n)
            out.append({"kind": "fixed_degenerate", "curve": O.CURVES[c], "label": label,
                        "points": [p.hex() for p in pts], "scalars": [x.hex() for x in sc],
                        "ref_result": ref.hex(), "exact_result": O.msm_naive(c, sb, pb, n).hex()})
    return out


def _buf_fixed(c, sb, pb, n):
    import ctypes

    out = ctypes.create_string_buffer(32)
    assert getattr(O.ref(), f"hsref_{O.CURVES[c]}_msm_fixed_bytes")(out, sb, pb, n) == 0
    return out.raw


def main():
    assert O.ref() is not None, "build the reference first: make -C oracle ref"
    runs = []
    for c, n, seed in PLAIN_CASES:
        runs.append(plain_case(c, n, seed))
        print(runs[-1], flush=True)
    for c, n, b in W.fcmpp_shapes():
        runs.append(fcmpp_case(c, n, b))
        print(runs[-1], flush=True)
    runs += tree_hash_cases()
    runs += fixed_degenerate_cases()
    with open(os.path.join(HERE, "workload_runs.json"), "w") as f:
        json.dump(runs, f, indent=1)


if __name__ == "__main__":
    main()
