"""Writes the golden fixtures under tests/golden/ (run in the development container, where the
reference tree is mounted and the reference library is built by ``make -C oracle ref``).

1. ``reference_vectors.json``: the MSM-path sections of the reference's own test-vector file
   (test_vectors/helioselene_test_vectors.json) -- field/scalar KATs, point encode/decode,
   add/dbl/neg, scalar_mul, msm (n <= 64), pedersen_commit, SSWU, batch_invert -- plus the
   F_p/F_q known answers held as byte arrays in src/tests.cpp:102-165.  Data only.
2. ``reference_msm_runs.json``: results of the compiled reference MSM on seeded inputs
   regenerated by tests/_oracle.py (seed, n, curve -> 32-byte result, plus the exact naive sum),
   covering every window-width band of helios_msm_vartime.cpp:248-263, the window-8 band where
   the reference drops a final recode carry, and the n = 4096 Helios workload on k*G points.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402

REF_JSON = "/root/reference/test_vectors/helioselene_test_vectors.json"

# src/tests.cpp:102-165 (transcribed byte arrays, hex little-endian)
TESTS_CPP_KAT = {
    "test_a": "efcdab9078563412bebafecaefbeadde00000000000000000000000000000000",
    "test_b": "08070605040302010df0adbacefaedfe00000000000000000000000000000000",
    "fp_ab": "8bf899b681c39d32379183ab63dfe3395abb62cf01db9b0740050f2e7564bf5d",
    "fp_asq": "34a5f2a2095f47a68023116b3872b0ef206511b6cc2e41d218fa928213cdb141",
    "fp_ainv": "3f3a94edeaf400ef5609c094eb9322cb71873d9b459cdef40a2013c1fc616625",
    "fq_ab": "d930723d0ff1e6c3de251ef43667647a5abb62cf01db9b0740050f2e7564bf5d",
    "fq_asq": "82ddca29978c903728b8abb30bfa3030216511b6cc2e41d218fa928213cdb141",
    "fq_ainv": "eee9dcce6d3757f1fd9058f5ffff5fb3303cb4b2814ab84fcfbe50e06b8ee160",
    "fq_sqrt4": "0200000000000000000000000000000000000000000000000000000000000000",
    "helios_g": "0300000000000000000000000000000000000000000000000000000000000000",
    "helios_2g": "262942408090b3c507b8ac94d46fc495fc129fb4d165372411d5e5ea008402f2",
    "helios_7g": "03df58ab3f9099c34d76642b4c99e582e38cf47e1bee444c4817a481ba499826",
    "selene_g": "0100000000000000000000000000000000000000000000000000000000000000",
    "selene_2g": "9dc7277972d2b66e586b65b72c787fbfffffffffffffffffffffffffffffffff",
    "selene_7g": "9930214df235941dba78b61cebf3812c69c0431828f9089e01695d8afd58be2f",
}

# seeded MSM runs: (curve, n, seed, kind) where kind selects the base generator
RUN_CASES = [(c, n, 1000 + i) for c in (0, 1)
             for i, n in enumerate([17, 33, 95, 96, 287, 288, 863, 864, 1000, 2000, 2591, 2592, 7776])]


def vectors():
    with open(REF_JSON) as f:
        d = json.load(f)
    keep = {
        "parameters": d["parameters"],
        "batch_invert": d["batch_invert"],
        "tests_cpp_kat": TESTS_CPP_KAT,
    }
    for s in ("helios_scalar", "selene_scalar"):
        keep[s] = {k: d[s][k] for k in ("add", "sub", "mul", "sq", "invert", "reduce_wide", "negate")}
    for s in ("helios_point", "selene_point"):
        keep[s] = {k: d[s][k] for k in ("generator", "identity", "from_bytes", "add", "dbl", "negate",
                                        "scalar_mul", "msm", "pedersen_commit", "map_to_curve_single",
                                        "map_to_curve_double")}
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(keep, f, indent=1, sort_keys=True)


def runs():
    assert O.ref() is not None, "build the reference first: make -C oracle ref"
    out = []
    for c, n, seed in RUN_CASES:
        s = O.gen_scalars(c, seed, n)
        p = O.gen_sswu_bases(c, seed, n, dup_frac=0.01, id_frac=0.005)
        r = O.ref_msm(c, s, p, n)
        naive = O.msm_naive(c, s, p, n) if n <= 2592 else None
        out.append({"curve": O.CURVES[c], "n": n, "seed": seed, "bases": "sswu", "dup_frac": 0.01,
                    "id_frac": 0.005, "ref_result": r.hex(), "naive_result": naive.hex() if naive else None})
    # the n = 4096 Helios workload: scalars and k*G points from the fuzz recipe
    c, n, seed = 0, 4096, 1
    s = O.gen_scalars(c, seed, n)
    p = O.gen_kG_points(c, seed, n)
    out.append({"curve": "helios", "n": n, "seed": seed, "bases": "kG", "dup_frac": 0, "id_frac": 0,
                "ref_result": O.ref_msm(c, s, p, n).hex(), "naive_result": None})
    with open(os.path.join(HERE, "reference_msm_runs.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    vectors()
    runs()
