#!/usr/bin/env python3
"""Benchmark of the variable-time Pippenger MSM on one MI355X (one JSON line).

Headline (`value`): plain single MSMs.  One step = one Helios MSM and one Selene MSM of n = 2^20
points each, on inputs already resident in HBM:
  * bases: SSWU hash-to-curve images of seeded uniform 32-byte strings (computed on the GPU by
    the library's own map), with 0.5% of them duplicates of another base and 0.1% the identity;
  * scalars: uniform modulo each curve's group order, 1% forced to edge values
    (0, 1, order - 1, a small 8-bit value).
Helios and Selene run on their own streams and consecutive steps alternate between two stream
pairs, so `value` (points per second over the timed steps, 2 * n points per step) is the
throughput of a stream of plain MSMs.  Before measuring, the library's autotuner picks plan
parameters per size class on this device (--no-tune: built-in defaults).

The same line carries:
  * `roofline`: k_bucket_flat (the dominant kernel) against the device's measured 32x32->64
    multiply throughput, from HIP events in an isolated pass (one MSM at a time), plus the whole
    MSM's fraction and the kernel's HBM traffic from the committed counter summary in profiles/;
  * `batched_mix`: 256 Helios + 256 Selene MSMs of log-uniform sizes in [64, 2^14], plain
    per-MSM launches against one batch plan replayed from a hipGraph;
  * `pedersen_batch`: 1024 commitments per curve over H + 2^12 resident generators, both
    fixed-base methods and the width-5 tables, with each table build timed to completion;
  * `serving`: Poisson request stream through the native server against plain per-request
    launches, p50/p99 latency and saturated throughput;
  * `host_slot_points_per_s`: the drop-in slot with host buffers (PCIe upload included);
  * `parity_vs_reference`: the timed steps' own 32-byte results against the compiled reference's
    bytes on the same inputs (tests/golden/workload_runs.json);
  * `single_msm`: one MSM at a time per curve with the scalar upload and result download timed;
  * `gate`: the whole single MSM's fraction of the INT32 roofline, and the same over the steps;
  * `roofline.hbm_phases`: achieved HBM bandwidth of the sort/scatter and gather phases;
  * `config1_helios_4096`: Helios n = 4096 on k*G points, GPU and reference CPU;
  * `fcmpp_verify`: the FCMP++ verification MSM pairs (Z != 1 bases), GPU and reference CPU;
  * `fcmpp_tree_hash`: FCMP++ tree-hash layers (Selene 228 / 38, Helios 18 fixed generators), as a
    batch and over a resident fixed-base set, against the reference per node;
  * `cpu_baseline`: the reference's own CPU MSM (compiled from its sources into
    oracle/_ref/libhsref.so) on a bounded sample of the same inputs, one pinned thread, median
    of three runs, with the CPU model.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MSM points/s at n=2^20 (plain vs batched) + p50/p99 latency, bit-exact"
# algorithmic partial products per field operation (8 x 32-bit limb schoolbook + reduction):
# F_p mul 72, sq 44; F_q mul 112, sq 84; a batched-affine bucket addition is 5M + 1S
PRODUCTS = {0: (72, 44), 1: (112, 84)}


def algorithmic_bucket_ops(curve, n):
    """INT32 multiply operations (2 per 32x32->64 partial product) of the bucket-accumulation
    phase of an n-point MSM under the fixed convention: W(c*) * n bucket additions at 5M + 1S,
    c* minimising the whole-MSM count."""
    m, s = PRODUCTS[curve]
    best = None
    for c in range(4, 21):
        W = -(-256 // c)
        tot = W * n * (5 * m + s) + W * (1 << c) * (11 * m + 5 * s)
        if best is None or tot < best[0]:
            best = (tot, c, W)
    _, c, W = best
    return 2 * W * n * (5 * m + s), c


def msm_int32_ops(curve, n):
    """INT32 multiply ops of a whole n-point MSM under the fixed convention: min over c of
    W(c) n (5m+s) + W(c) 2^c (11m+5s), plus 256 doublings at 3m+5s, 2 ops per product."""
    m, s = PRODUCTS[curve]
    best = min(-(-256 // c) * n * (5 * m + s) + -(-256 // c) * (1 << c) * (11 * m + 5 * s) for c in range(4, 21))
    return 2 * (best + 256 * (3 * m + 5 * s))


def measured_traffic(n):
    """HBM bytes per k_bucket_flat launch from the committed counter summary (two rocprofv3 --pmc
    runs, FETCH_SIZE and WRITE_SIZE, scaled by a calibration kernel of known size), or None."""
    import re

    path = os.path.join(ROOT, "profiles", "plain_msm_n2e20_kernel_trace.md")
    if n != 1 << 20 or not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch: helios=(\d+) selene=(\d+)", open(path).read())
    if not m:
        return None
    return {"helios": int(m.group(1)), "selene": int(m.group(2)), "source": os.path.relpath(path, ROOT)}


def make_inputs(hs, torch, curve, n, seed, dev):
    """The plain workload of helioselene_amd/workloads.py (the generator the parity tests and the
    reference fixtures use) with the bases mapped by the library's SSWU kernel on the device."""
    from helioselene_amd import workloads as W

    w = W.plain_inputs(curve, n, seed)
    u = torch.from_numpy(w["u"]).to(dev)
    pts = torch.empty((n, 15), dtype=torch.int64, device=dev)
    hs.map_to_curve_device(curve, u.data_ptr(), n, pts.data_ptr(), None, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    W.apply_dups_and_identities(pts, w)
    return torch.from_numpy(w["scalars"]).to(dev), pts


def reference_results(kind):
    """{(curve name, n, seed or batch): 32-byte hex} of the compiled reference on these workloads
    (tests/golden/workload_runs.json, written by tests/golden/make_workload_fixtures.py)."""
    path = os.path.join(ROOT, "tests", "golden", "workload_runs.json")
    if not os.path.exists(path):
        return {}
    out = {}
    for r in json.load(open(path)):
        if r["kind"] == kind:
            out[(r["curve"], r["n"], r.get("seed", r.get("batch")))] = r["ref_result"]
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_single(hs, torch, inputs, n, args, st):
    """One MSM at a time per curve, the way a caller holding resident bases uses it: the scalars
    go from pinned host memory to HBM, the MSM runs, the 32-byte result comes back; all inside the
    timed region (bases resident, not timed)."""
    import numpy as np

    res = {}
    for c, cv in enumerate(("helios", "selene")):
        sc, pts = inputs[c]
        host = sc.cpu().pin_memory()
        dsc = torch.empty_like(sc)
        o = torch.zeros(8, dtype=torch.int32, device=sc.device)
        oh = torch.zeros(8, dtype=torch.int32).pin_memory()
        lat = []
        with torch.cuda.stream(st):
            for i in range(2 + args.single_steps):
                t0 = time.perf_counter()
                dsc.copy_(host, non_blocking=True)
                hs.msm_device(c, dsc.data_ptr(), pts.data_ptr(), n, o.data_ptr(), None, 0, st.cuda_stream)
                oh.copy_(o, non_blocking=True)
                st.synchronize()
                lat.append(time.perf_counter() - t0)
        lat = np.array(lat[2:]) * 1e3
        res[cv] = {"ms_p50": float(np.percentile(lat, 50)), "ms_p99": float(np.percentile(lat, 99)),
                   "points_per_s": n / (float(np.median(lat)) * 1e-3), "result": oh.numpy().tobytes().hex()}
    res["workload"] = (f"one plain MSM at a time of n={n} per curve: scalar upload ({32 * n} B from pinned host "
                       f"memory) + MSM + result download timed, bases resident; {args.single_steps} runs per curve")
    return res


def make_mix(hs, torch, dev, seed, per_curve=256, lo=64, hi=1 << 14):
    """The uneven batch: per_curve MSMs per curve, Helios and Selene interleaved, sizes
    log-uniform in [lo, hi], distinct seeded SSWU bases per MSM, scalars uniform mod order."""
    import numpy as np

    rng = np.random.default_rng(seed)
    M = 2 * per_curve
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), size=M)).astype(np.int64).clip(lo, hi)
    curves = [k % 2 for k in range(M)]
    tot = [int(sizes[[k for k in range(M) if curves[k] == c]].sum()) for c in (0, 1)]
    sc, pts = [], []
    for c in (0, 1):
        u = torch.from_numpy(rng.integers(0, 256, size=(tot[c], 32), dtype=np.uint8)).to(dev)
        p = torch.empty((tot[c], 15), dtype=torch.int64, device=dev)
        hs.map_to_curve_device(c, u.data_ptr(), tot[c], p.data_ptr(), None, torch.cuda.current_stream().cuda_stream)
        s = rng.integers(0, 256, size=(tot[c], 32), dtype=np.uint8)
        s[:, 31] &= 0x7F  # uniform 255-bit; each order is within 2^-126 of 2^255, the library reduces
        sc.append(torch.from_numpy(s).to(dev))
        pts.append(p)
    torch.cuda.synchronize()
    out = torch.zeros((M, 8), dtype=torch.int32, device=dev)
    items, off = [], [0, 0]
    for k in range(M):
        c, n = curves[k], int(sizes[k])
        items.append((c, n, sc[c].data_ptr() + 32 * off[c], pts[c].data_ptr() + 120 * off[c],
                      out[k].data_ptr(), None))
        off[c] += n
    return items, out, (sc, pts), sizes


def ref_lib():
    lib = os.path.join(ROOT, "oracle", "_ref", "libhsref.so")
    if not os.path.exists(lib):
        return None
    L = ctypes.CDLL(lib)
    L.hsref_init.restype = ctypes.c_uint
    L.hsref_init()
    for cv in ("helios", "selene"):
        f = getattr(L, f"hsref_{cv}_time_msm")
        f.restype = ctypes.c_double
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
        getattr(L, f"hsref_{cv}_msm_jac_bytes").argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p,
                                                            ctypes.c_size_t]
    return L


def cpu_time_median(L, curve, sc, pts, runs=3):
    """Median wall seconds of `runs` single-threaded reference MSMs (one pinned core)."""
    import numpy as np

    f = getattr(L, ("hsref_helios_time_msm", "hsref_selene_time_msm")[curve])
    s = np.ascontiguousarray(sc, dtype=np.uint8)
    p = np.ascontiguousarray(pts).view(np.uint64)
    return float(np.median([f(s.ctypes.data, p.ctypes.data, p.shape[0], 1, 1) for _ in range(runs)]))


def cpu_ref_bytes(L, curve, sc, pts):
    import numpy as np

    out = ctypes.create_string_buffer(32)
    s = np.ascontiguousarray(sc, dtype=np.uint8)
    p = np.ascontiguousarray(pts).view(np.uint64)
    getattr(L, f"hsref_{('helios', 'selene')[curve]}_msm_jac_bytes")(out, s.ctypes.data, p.ctypes.data, p.shape[0])
    return out.raw.hex()


def run_config1(hs, torch, args, st, L):
    """Helios n = 4096 on points k_i G (k_i, s_i = 64 seeded bytes reduced mod the order, the
    reference fuzz recipe): the GPU MSM (inputs resident) and the reference CPU MSM on the same
    inputs.  The bases are computed on the GPU as 4096 one-term MSMs over the resident set {G}."""
    import numpy as np

    from helioselene_amd import workloads as W

    n, order = 4096, hs.ORDER[0]
    rng = np.random.default_rng(args.seed + 4000)

    def reduced():
        raw = rng.integers(0, 256, size=(n, 64), dtype=np.uint8)
        vals = [int.from_bytes(raw[i].tobytes(), "little") % order for i in range(n)]
        return np.frombuffer(b"".join(v.to_bytes(32, "little") for v in vals), dtype=np.uint8).reshape(n, 32).copy()

    k, sc = reduced(), reduced()
    g = hs.FixedSet(0, W.jacobian_struct(0, *W.GEN[0], 1).reshape(1, 15))
    _, pts = g.msm(k.reshape(n, 1, 32), flags=hs.EXACT)
    g.close()
    dev = torch.device("cuda:0")
    dsc, dpt = torch.from_numpy(sc).to(dev), torch.from_numpy(pts.view(np.int64)).to(dev)
    o = torch.zeros(8, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    lat = []
    with torch.cuda.stream(st):
        for i in range(3 + args.single_steps):
            t0 = time.perf_counter()
            hs.msm_device(0, dsc.data_ptr(), dpt.data_ptr(), n, o.data_ptr(), None, 0, st.cuda_stream)
            st.synchronize()
            lat.append(time.perf_counter() - t0)
    gpu_ms = float(np.median(lat[3:])) * 1e3
    got = o.cpu().numpy().tobytes().hex()
    res = {"workload": "Helios MSM n=4096, points k_i*G, scalars uniform mod order (64 seeded bytes reduced)",
           "gpu_ms": gpu_ms, "gpu_points_per_s": n / (gpu_ms * 1e-3)}
    if L is not None:
        t = cpu_time_median(L, 0, sc, pts)
        res.update({"cpu_ms": t * 1e3, "cpu_points_per_s": n / t, "cpu_cores": 1,
                    "cpu_kind": "reference (oracle/_ref, helios_msm_vartime via helioselene_init dispatch), median of 3",
                    "bit_exact_vs_reference": got == cpu_ref_bytes(L, 0, sc, pts)})
    return res


def run_fcmpp(hs, torch, args, st, L):
    """FCMP++ verification MSMs (src/benchmark_fcmpp.cpp:347-365 of the reference): Selene 522 + 80 b
    and Helios 265 + 80 b for b proofs, on doubling-chain bases (Z != 1: every point goes through the
    ingest's inversion) with the benchmark's scalars.  Per b: the two MSMs as plain launches and as
    one batch (captured graph), latency per verification, against the reference CPU MSMs."""
    import numpy as np

    from helioselene_amd import workloads as W

    dev = torch.device("cuda:0")
    ref = reference_results("fcmpp")
    out = {"workload": "per verification: Selene 522+80b and Helios 265+80b points, doubling-chain bases; "
                       "batches: the reference benchmark's scalars (i + 1 + 2^24), uniform_scalars: uniform "
                       "255-bit scalars",
           "batches": {}, "uniform_scalars": {}}
    rng = np.random.default_rng(args.seed + 6000)
    for b, uniform in [(b, False) for b in W.FCMPP_BATCHES] + [(b, True) for b in W.FCMPP_BATCHES]:
        shapes = [(1, W.FCMPP_FIXED[1] + W.FCMPP_PER_PROOF * b), (0, W.FCMPP_FIXED[0] + W.FCMPP_PER_PROOF * b)]
        host, dev_in, items = [], [], []
        res_t = torch.zeros((2, 8), dtype=torch.int32, device=dev)
        for k, (c, n) in enumerate(shapes):
            pts, sc = W.doubling_chain(c, n), W.fcmpp_scalars(n)
            if uniform:
                # verification scalars are uniform field elements in a real proof
                sc = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
                sc[:, 31] &= 0x7F
            host.append((c, sc, pts))
            ds, dp = torch.from_numpy(sc).to(dev), torch.from_numpy(pts.view(np.int64)).to(dev)
            dev_in.append((ds, dp))
            items.append((c, n, ds.data_ptr(), dp.data_ptr(), res_t[k].data_ptr(), None))
        torch.cuda.synchronize()
        sh = st.cuda_stream

        def plain():
            for (c, n, ds, dp, do, _) in items:
                hs.msm_device(c, ds, dp, n, do, None, 0, sh)

        def timed(fn, reps):
            lat = []
            with torch.cuda.stream(st):
                for i in range(3 + reps):
                    t0 = time.perf_counter()
                    fn()
                    st.synchronize()
                    lat.append(time.perf_counter() - t0)
            return float(np.median(lat[3:])) * 1e3

        t_plain = timed(plain, args.single_steps)
        ok_plain = [res_t[k].cpu().numpy().tobytes().hex() for k in range(2)]
        bt = hs.Batch(items)
        bt.capture()
        res_t.zero_()
        t_batch = timed(lambda: bt.launch(sh), args.single_steps)
        ok_batch = [res_t[k].cpu().numpy().tobytes().hex() for k in range(2)]
        bt.close()
        if uniform:
            want = [cpu_ref_bytes(L, c, sc, pts) for c, sc, pts in host] if L is not None else [None, None]
        else:
            want = [ref.get((("helios", "selene")[c], n, b)) for c, n in shapes]
        npts = sum(n for _, n in shapes)
        rec = {"points": npts, "gpu_plain_ms": t_plain, "gpu_batched_ms": t_batch,
               "gpu_points_per_s": npts / (min(t_plain, t_batch) * 1e-3),
               "bit_exact_vs_reference": (ok_plain == want and ok_batch == want) if None not in want else None}
        if L is not None:
            tc = sum(cpu_time_median(L, c, sc, pts) for c, sc, pts in host)
            rec.update({"cpu_ms": tc * 1e3, "cpu_points_per_s": npts / tc})
        out["uniform_scalars" if uniform else "batches"][str(b)] = rec
    return out


def run_tree_hash(hs, torch, args, st, L):
    """FCMP++ tree hashes (src/benchmark_fcmpp.cpp:133-198 of the reference): a node's hash is an
    MSM over its layer's fixed generators (doubling chains, Z != 1) plus the affine conversion
    (the 32-byte result carries the canonical x).  Per layer shape (Selene 228 / 38, Helios 18
    children): one layer of nodes with uniform 255-bit scalars as one batch (captured graph) and over
    a resident fixed-base set of the generators (16-tooth comb), a single node's latency, and the
    reference's MSM + tobytes per node on one host thread."""
    import numpy as np

    from helioselene_amd import workloads as W

    dev = torch.device("cuda:0")
    ref = {(r["curve"], r["n"], r["index"]): r["ref_result"]
           for r in (json.load(open(os.path.join(ROOT, "tests", "golden", "workload_runs.json"))))
           if r["kind"] == "tree_hash"} if os.path.exists(os.path.join(ROOT, "tests", "golden", "workload_runs.json")) else {}
    out = {"workload": "one tree layer per shape: nodes of n children over the layer's doubling-chain generators, "
                       "uniform 255-bit scalars; a node = MSM + affine conversion (32-byte result)", "layers": {}}
    sh = st.cuda_stream

    def timed(fn, reps):
        lat = []
        with torch.cuda.stream(st):
            for i in range(2 + reps):
                t0 = time.perf_counter()
                fn()
                st.synchronize()
                lat.append(time.perf_counter() - t0)
        return float(np.median(lat[2:]))

    for c, n, name in W.TREE_HASH_SHAPES:
        K = 2048 if n > 100 else 4096
        pts = W.doubling_chain(c, n)
        sc = W.tree_layer_scalars(n, K, 7000 + n)
        with torch.cuda.stream(st):
            dp = torch.from_numpy(pts.view(np.int64)).to(dev)
            ds = torch.from_numpy(sc).to(dev)
            o_b = torch.zeros((K, 8), dtype=torch.int32, device=dev)
            o_f = torch.zeros((K, 8), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        items = [(c, n, ds.data_ptr() + 32 * n * h, dp.data_ptr(), o_b[h].data_ptr(), None) for h in range(K)]
        bt = hs.Batch(items)
        bt.capture()
        t_batch = timed(lambda: bt.launch(sh), 5)
        bt.close()
        t_single = timed(lambda: hs.msm_device(c, items[0][2], items[0][3], n, items[0][4], None, 0, sh), 10)
        t0 = time.perf_counter()
        fs = hs.FixedSet(c, d_points=dp.data_ptr(), n=n, stream=sh)
        torch.cuda.synchronize()
        t_build = time.perf_counter() - t0
        t_fixed = timed(lambda: fs.msm_device(ds.data_ptr(), K, o_f.data_ptr(), method=hs.FIXED_COMB, stream=sh), 5)
        fs.close()
        got_b = [o_b[h].cpu().numpy().tobytes().hex() for h in range(4)]
        got_f = [o_f[h].cpu().numpy().tobytes().hex() for h in range(4)]
        want = [ref.get((("helios", "selene")[c], n, h)) for h in range(4)]
        rec = {"curve": ("helios", "selene")[c], "children": n, "nodes": K,
               "gpu_single_node_ms": t_single * 1e3,
               "gpu_batch_ms": t_batch * 1e3, "gpu_batch_nodes_per_s": K / t_batch,
               "gpu_fixed_comb_ms": t_fixed * 1e3, "gpu_fixed_comb_nodes_per_s": K / t_fixed,
               "fixed_table_build_ms": t_build * 1e3,
               "bit_exact_vs_reference": (got_b == want and got_f == want) if None not in want else None,
               "batch_equals_fixed": bool((o_b.cpu().numpy() == o_f.cpu().numpy()).all())}
        if L is not None:
            m = 256 if n < 100 else 64
            t0 = time.perf_counter()
            for h in range(m):
                cpu_ref_bytes(L, c, sc[h], pts)
            tc = (time.perf_counter() - t0) / m
            rec.update({"cpu_node_ms": tc * 1e3, "cpu_nodes_per_s": 1.0 / tc,
                        "cpu_sample": f"{m} nodes, reference MSM + tobytes, one host thread"})
        out["layers"][name] = rec
    return out


def run_mix(hs, torch, args):
    """Plain per-MSM launches vs one batched plan replayed from a hipGraph on the uneven mix."""
    import numpy as np

    dev = torch.device("cuda:0")
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        items, out, keep, sizes = make_mix(hs, torch, dev, args.seed + 1000)
        npts = int(sizes.sum())
        sh = st.cuda_stream

        def plain_pass():
            for (c, n, ds, dp, do, _) in items:
                hs.msm_device(c, ds, dp, n, do, None, 0, sh)

        plain_pass()
        torch.cuda.synchronize()
        want = out.cpu().numpy().copy()
        t = []
        for _ in range(args.mix_plain_steps):
            t0 = time.perf_counter()
            plain_pass()
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
        plain_s = float(np.median(t))

        b = hs.Batch(items)
        out.zero_()
        b.launch(sh)
        torch.cuda.synchronize()
        eager_ok = bool((out.cpu().numpy() == want).all())
        b.capture()
        for _ in range(args.warmup):
            b.launch(sh)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.mix_steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record(st)
            b.launch(sh)
            e1.record(st)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
        graph_ok = bool((out.cpu().numpy() == want).all())
        b.close()
    batched = args.mix_steps * npts / elapsed
    return {
        "workload": "256 Helios + 256 Selene MSMs, sizes log-uniform in [64, 2^14], interleaved",
        "msms": len(items),
        "points": npts,
        "plain_points_per_s": npts / plain_s,
        "plain_ms_per_mix": plain_s * 1e3,
        "batched_points_per_s": batched,
        "batched_ms_per_mix": elapsed / args.mix_steps * 1e3,
        "batched_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "speedup_vs_plain": batched * plain_s / npts,
        "bit_exact_vs_plain": eager_ok and graph_ok,
    }


def run_pedersen(hs, torch, args, count=1024, m=4096):
    """Batched Pedersen commitments r*H + sum v_i G_i over one resident set per curve of H and
    m = 2^12 generators (seeded SSWU), count commitments per curve, both fixed-base methods."""
    import numpy as np

    dev = torch.device("cuda:0")
    st = torch.cuda.Stream()
    rng = np.random.default_rng(args.seed + 2000)
    res = {"workload": f"{count} Pedersen commitments per curve over H + {m} resident generators",
           "commitments_per_curve": count, "terms_per_commitment": m + 1}
    with torch.cuda.stream(st):
        sh = st.cuda_stream
        outs = {}
        for c in (0, 1):
            u = torch.from_numpy(rng.integers(0, 256, size=(m + 1, 32), dtype=np.uint8)).to(dev)
            pts = torch.empty((m + 1, 15), dtype=torch.int64, device=dev)
            hs.map_to_curve_device(c, u.data_ptr(), m + 1, pts.data_ptr(), None, sh)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fs = hs.FixedSet(c, d_points=pts.data_ptr(), n=m + 1, stream=sh)
            torch.cuda.synchronize()
            build = time.perf_counter() - t0
            res[f"{['helios', 'selene'][c]}_table_build_ms"] = build * 1e3
            s = rng.integers(0, 256, size=(count, m + 1, 32), dtype=np.uint8)
            s[:, :, 31] &= 0x7F
            ds = torch.from_numpy(s).to(dev)
            for method, name in ((hs.FIXED_COMB, "comb"), (hs.FIXED_PIPPENGER, "pippenger")):
                out = torch.zeros((count, 8), dtype=torch.int32, device=dev)
                fs.msm_device(ds.data_ptr(), count, out.data_ptr(), None, 0, method, sh)
                torch.cuda.synchronize()
                t = []
                for _ in range(args.pedersen_steps):
                    t0 = time.perf_counter()
                    fs.msm_device(ds.data_ptr(), count, out.data_ptr(), None, 0, method, sh)
                    torch.cuda.synchronize()
                    t.append(time.perf_counter() - t0)
                tm = float(np.median(t))
                res[f"{['helios', 'selene'][c]}_{name}_ms"] = tm * 1e3
                res[f"{['helios', 'selene'][c]}_{name}_points_per_s"] = count * (m + 1) / tm
                if name == "comb":
                    res[f"{['helios', 'selene'][c]}_comb_points_per_s_incl_build"] = count * (m + 1) / (tm + build)
                outs[(c, name)] = out.cpu().numpy()
            fs.close()
            # the width-5 affine tables of the workload as written ([1P..16P] per generator, 52
            # windows of 5 shared doublings; helios_scalarmult_fixed_precompute / helios_msm_fixed)
            os.environ["HS_GPU_FIXED_WIDTH"] = "5"
            try:
                t0 = time.perf_counter()
                f5 = hs.FixedSet(c, d_points=pts.data_ptr(), n=m + 1, stream=sh)
                torch.cuda.synchronize()
                build5 = time.perf_counter() - t0
            finally:
                del os.environ["HS_GPU_FIXED_WIDTH"]
            out = torch.zeros((count, 8), dtype=torch.int32, device=dev)
            f5.msm_device(ds.data_ptr(), count, out.data_ptr(), None, 0, hs.FIXED_COMB, sh)
            torch.cuda.synchronize()
            t = []
            for _ in range(args.pedersen_steps):
                t0 = time.perf_counter()
                f5.msm_device(ds.data_ptr(), count, out.data_ptr(), None, 0, hs.FIXED_COMB, sh)
                torch.cuda.synchronize()
                t.append(time.perf_counter() - t0)
            tm = float(np.median(t))
            cv = ['helios', 'selene'][c]
            res[f"{cv}_w5_table_build_ms"] = build5 * 1e3
            res[f"{cv}_w5_ms"] = tm * 1e3
            res[f"{cv}_w5_points_per_s"] = count * (m + 1) / tm
            res[f"{cv}_w5_points_per_s_incl_build"] = count * (m + 1) / (tm + build5)
            outs[(c, "w5")] = out.cpu().numpy()
            f5.close()
            res[f"{cv}_methods_agree"] = bool((outs[(c, "comb")] == outs[(c, "pippenger")]).all() and
                                              (outs[(c, "comb")] == outs[(c, "w5")]).all())
            del ds
    best = sum(min(res[f"{cv}_comb_ms"], res[f"{cv}_pippenger_ms"]) for cv in ("helios", "selene"))
    res["points_per_s_best_method"] = 2 * count * (m + 1) / (best * 1e-3)
    w5 = sum(res[f"{cv}_w5_ms"] for cv in ("helios", "selene"))
    res["points_per_s_w5_tables"] = 2 * count * (m + 1) / (w5 * 1e-3)
    res["note"] = ("comb: 16-tooth signed comb tables (8 GiB per set, built once in *_table_build_ms); w5: the "
                   "width-5 affine tables the workload names (1 KiB per generator, built in *_w5_table_build_ms)")
    return res


def run_serving(hs, torch, args):
    """Served request stream: Poisson arrivals at a fixed offered load, sizes log-uniform in
    [32, 2^18], curve by coin flip, seeded SSWU bases and uniform scalars resident in HBM (each
    request reads its own slice of a per-curve pool).  The native server (continuous batching,
    footprint admission) against plain per-request launches on the same arrival schedule, plus
    both at saturation (all requests queued at once)."""
    import numpy as np

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(args.seed + 3000)
    pool = 1 << 18
    R, lam = args.serve_requests, args.serve_rate
    sizes = np.exp(rng.uniform(np.log(32), np.log(pool), size=R)).astype(np.int64).clip(32, pool)
    curves = rng.integers(0, 2, size=R)
    offs = np.array([int(rng.integers(0, pool - n + 1)) for n in sizes])
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=R))
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        dsc, dpt = [], []
        for c in (0, 1):
            u = torch.from_numpy(rng.integers(0, 256, size=(pool, 32), dtype=np.uint8)).to(dev)
            p = torch.empty((pool, 15), dtype=torch.int64, device=dev)
            hs.map_to_curve_device(c, u.data_ptr(), pool, p.data_ptr(), None, st.cuda_stream)
            s = rng.integers(0, 256, size=(pool, 32), dtype=np.uint8)
            s[:, 31] &= 0x7F
            dsc.append(torch.from_numpy(s).to(dev))
            dpt.append(p)
        torch.cuda.synchronize()

    def req(i):
        c, n, o = int(curves[i]), int(sizes[i]), int(offs[i])
        return c, n, dsc[c].data_ptr() + 32 * o, dpt[c].data_ptr() + 120 * o

    def served(schedule):
        srv = hs.Server(device_bytes=args.serve_bytes << 30, lanes=args.serve_lanes)
        out = [None] * R
        for i in range(min(8, R)):  # warm the lanes
            c, n, ds, dp = req(i)
            srv.wait(srv.submit(c, d_scalars=ds, d_points=dp, n=n))
        srv.reset_stats()
        t0 = time.perf_counter()
        tickets = []
        for i in range(R):
            if schedule:
                dt = t0 + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            c, n, ds, dp = req(i)
            tickets.append(srv.submit(c, d_scalars=ds, d_points=dp, n=n))
        for i, t in enumerate(tickets):
            out[i] = srv.wait(t)[0]
        wall = time.perf_counter() - t0
        stats = srv.stats()
        srv.close()
        return out, stats, wall

    def plain(schedule):
        sh = st.cuda_stream
        o = torch.zeros(8, dtype=torch.int32, device=dev)
        lat, out = [], [None] * R
        t0 = time.perf_counter()
        with torch.cuda.stream(st):
            for i in range(R):
                if schedule:
                    dt = t0 + arrivals[i] - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                c, n, ds, dp = req(i)
                hs.msm_device(c, ds, dp, n, o.data_ptr(), None, 0, sh)
                st.synchronize()
                done = time.perf_counter()
                lat.append((done - (t0 + arrivals[i] if schedule else t0)) * 1e6)
                out[i] = o.cpu().numpy().tobytes()
        wall = time.perf_counter() - t0
        lat = np.array(lat)
        return out, {"p50_us": float(np.percentile(lat, 50)), "p99_us": float(np.percentile(lat, 99))}, wall

    npts = int(sizes.sum())
    s_out, s_st, s_wall = served(True)
    p_out, p_st, p_wall = plain(True)
    ss_out, ss_st, ss_wall = served(False)
    ps_out, _, ps_wall = plain(False)
    return {
        "workload": f"{R} requests, Poisson arrivals at {lam:g}/s, sizes log-uniform in [32, 2^18], 50/50 curves, "
                    "bases and scalars resident in HBM",
        "requests": R, "points": npts, "offered_points_per_s": npts / float(arrivals[-1]),
        "served": {"p50_ms": s_st["p50_us"] / 1e3, "p99_ms": s_st["p99_us"] / 1e3, "points_per_s": npts / s_wall,
                   "batches": s_st["batches"], "lanes": args.serve_lanes},
        "plain_per_request": {"p50_ms": p_st["p50_us"] / 1e3, "p99_ms": p_st["p99_us"] / 1e3,
                              "points_per_s": npts / p_wall},
        "saturated_points_per_s": {"served": npts / ss_wall, "plain": npts / ps_wall},
        "saturated_speedup": ps_wall / ss_wall,
        "bit_exact_served_vs_plain": s_out == p_out and ss_out == p_out and ps_out == p_out,
    }


def cpu_baseline(hs_pts_cpu, sc_cpu, n_sample, runs=3):
    """The reference CPU MSM (x64 Pippenger behind its runtime dispatch) on the first n_sample
    inputs of each curve: one thread pinned to one core (as include/helioselene_benchmark.h:119-128
    pins its benchmark thread), median of `runs` runs per curve."""
    L = ref_lib()
    if L is None:
        return None
    feats = L.hsref_init()
    cpus = sorted(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {cpus[len(cpus) // 2]})
    try:
        med = [cpu_time_median(L, c, sc_cpu[c][:n_sample], hs_pts_cpu[c][:n_sample], runs) for c in (0, 1)]
    finally:
        os.sched_setaffinity(0, set(cpus))
    tot = med[0] + med[1]
    return {
        "value": 2 * n_sample / tot,
        "unit": "points/s",
        "cores": 1,
        "kind": "reference",
        "cpu_model": cpu_model(),
        "per_curve_points_per_s": {"helios": n_sample / med[0], "selene": n_sample / med[1]},
        "sample": (f"one Helios and one Selene reference MSM (helios_msm_vartime / selene_msm_vartime via "
                   f"helioselene_init dispatch, cpu feature mask {feats:#x}) at n={n_sample} on the first "
                   f"{n_sample} bases/scalars of this workload, 1 thread pinned to one core, median of {runs} "
                   f"runs per curve ({med[0]:.2f} s + {med[1]:.2f} s)"),
    }


def cpu_many(hs_pts_cpu, sc_cpu, jobs, threads):
    """The reference CPU MSM on many independent MSMs at once: for each (curve, n) in `jobs`,
    `threads` host threads each run one whole reference MSM of that size (the function is
    reentrant), on the first n bases/scalars of the plain workload.  Returns points/s."""
    lib = os.path.join(ROOT, "oracle", "_ref", "libhsref.so")
    if not os.path.exists(lib) or not jobs:
        return None
    L = ctypes.CDLL(lib)
    L.hsref_init.restype = ctypes.c_uint
    L.hsref_init()
    pts, tot = 0, 0.0
    for curve, n in jobs:
        f = getattr(L, ("hsref_helios_time_msm", "hsref_selene_time_msm")[curve])
        f.restype = ctypes.c_double
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
        s = sc_cpu[curve][:n].copy()
        p = hs_pts_cpu[curve][:n].copy()
        tot += f(s.ctypes.data, p.ctypes.data, n, threads, threads)
        pts += threads * n
    return pts / tot, tot


def add_cpu_legs(rec, pts_cpu, sc_cpu, args):
    """N-thread reference baselines for the batched mix, the Pedersen batch and the served stream,
    each on a bounded sample of its own size distribution."""
    import numpy as np

    threads = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(args.seed + 5000)
    legs = {
        "batched_mix": [(k % 2, int(np.exp(rng.uniform(np.log(64), np.log(1 << 14))))) for k in range(24)],
        "pedersen_batch": [(0, 4097), (1, 4097)],
        "serving": [(int(rng.integers(0, 2)), int(np.exp(rng.uniform(np.log(32), np.log(1 << 18)))))
                    for _ in range(8)],
    }
    for leg, jobs in legs.items():
        if not rec.get(leg):
            continue
        r = cpu_many(pts_cpu, sc_cpu, jobs, threads)
        if r is None:
            continue
        rate, secs = r
        rec[leg]["cpu_baseline"] = {
            "value": rate, "unit": "points/s", "cores": threads, "kind": "reference",
            "sample": (f"{len(jobs)} MSM sizes drawn from this leg's distribution; for each, {threads} host threads "
                       f"run one whole reference MSM each at the same time ({secs:.1f} s in total)"),
        }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40,
                    help="timed steps (enough that filling and draining the stream pipeline is a small part)")
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mix-steps", type=int, default=20, help="timed replays of the batched mix")
    ap.add_argument("--mix-plain-steps", type=int, default=3, help="timed plain passes over the mix")
    ap.add_argument("--no-mix", action="store_true")
    ap.add_argument("--pedersen-steps", type=int, default=3)
    ap.add_argument("--no-pedersen", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="offered load, requests per second")
    ap.add_argument("--serve-lanes", type=int, default=4)
    ap.add_argument("--serve-bytes", type=int, default=16, help="server device budget, GiB")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--roofline-steps", type=int, default=3)
    ap.add_argument("--single-steps", type=int, default=10, help="timed runs per curve of the one-at-a-time legs")
    ap.add_argument("--no-fcmpp", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--affine-k", type=int, default=0, choices=(0, 1, 2, 4),
                    help="plain MSMs with the batched-affine bucket kernel, K chains per lane (measurement of that "
                         "kernel; 0 = the default xyzz kernel)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (built-in plan parameters)")
    ap.add_argument("--streams", type=int, default=4, choices=(1, 2, 4),
                    help="1: both MSMs of a step on one stream; 2: Helios and Selene on their own streams; "
                         "4: consecutive steps alternate between two such stream pairs (pipelined steps)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (each MSM fits one "
              f"MI355X; multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    if args.affine_k:
        os.environ["HS_GPU_AFFINE_K"] = str(args.affine_k)
    import numpy as np
    import torch

    import helioselene_amd as hs

    dev = torch.device("cuda:0")
    torch.cuda.init()
    streams = [torch.cuda.Stream() for _ in range(args.streams)]
    # HIP binds a stream to one of the process's hardware queues (4 by default) when the stream is
    # first used. The MSM streams are used first, one after another, so each gets a queue of its
    # own; streams used later (the autotuner's, the input setup's) share queues with them but are
    # idle while the MSMs are timed. With two MSM streams on one queue the two stream pairs no
    # longer overlapped (-10% throughput).
    for s_ in streams:
        with torch.cuda.stream(s_):
            torch.zeros(1, device=dev)
    torch.cuda.synchronize()
    n = args.n
    tuning = None
    if not args.no_tune:
        # the library's autotuner picks (window bits, run length, segment size) per size class on
        # this device: plain MSMs from 2^6 up to the headline size (the mix's per-MSM launches use the
        # small classes), batch groups over the batched / served / Pedersen size range
        t_tune = time.perf_counter()
        lg = max(6, min(22, n.bit_length() - 1))
        for c in (0, 1):
            hs.autotune(c, 0, 6, lg)
            hs.autotune(c, 1, 6, 18)
        tuning = {"seconds": time.perf_counter() - t_tune,
                  "plain": {cv: hs.tune_get(c, 0, n) for c, cv in enumerate(("helios", "selene"))},
                  "plain_by_size": {cv: {f"2^{k}": hs.tune_get(c, 0, 1 << k) for k in range(6, lg + 1)}
                                    for c, cv in enumerate(("helios", "selene"))},
                  "batch": {cv: {f"2^{k}": hs.tune_get(c, 1, 1 << k) for k in range(6, 19)}
                            for c, cv in enumerate(("helios", "selene"))},
                  "fields": "(window bits c, entries per thread L, buckets per segment G)"}
    setup = torch.cuda.Stream()
    with torch.cuda.stream(setup):
        inputs = [make_inputs(hs, torch, c, n, args.seed + 17 * c, dev) for c in (0, 1)]
        out = torch.zeros((4, 8), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        pairs = [[streams[min(c, args.streams - 1)].cuda_stream for c in (0, 1)]] if args.streams < 4 else \
            [[streams[0].cuda_stream, streams[1].cuda_stream], [streams[2].cuda_stream, streams[3].cuda_stream]]
        sts = pairs[0]

        def step(i):
            for c in (0, 1):
                sc, pts = inputs[c]
                k = i % len(pairs)
                hs.msm_device(c, sc.data_ptr(), pts.data_ptr(), n, out[2 * k + c].data_ptr(), None, 0, pairs[k][c])

        for i in range(args.warmup):
            step(i)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            step(i)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # HBM counter calibration: one elementwise kernel that reads 2 GiB and writes 2 GiB (eight
        # times the 256 MB last-level cache, so nothing it reads is cache-resident), so a FETCH_SIZE /
        # WRITE_SIZE counter run of this command can scale the MSM kernels' counts
        cal = torch.zeros(1 << 29, dtype=torch.int32, device=dev)
        cal2 = torch.empty_like(cal)
        torch.cuda.synchronize()
        torch.add(cal, 1, out=cal2)
        torch.cuda.synchronize()
        del cal, cal2
        # roofline pass: the same MSMs one at a time on one stream, HIP events between phases
        hs.profile(True)
        for _ in range(args.roofline_steps):
            for c in (0, 1):
                sc, pts = inputs[c]
                hs.msm_device(c, sc.data_ptr(), pts.data_ptr(), n, out[c].data_ptr(), None, 0, sts[0])
                torch.cuda.synchronize()
        phases = [hs.phase_summary(c) for c in (0, 1)]
        hs.profile(False)
        torch.cuda.synchronize()
        # the timed steps' own results against the compiled reference's bytes on these inputs
        ref = reference_results("plain")
        rows = out.cpu().numpy()
        parity = {}
        for c, cv in enumerate(("helios", "selene")):
            want = ref.get((cv, n, args.seed + 17 * c))
            got = {rows[k].tobytes().hex() for k in range(4) if k % 2 == c and (args.streams == 4 or k < 2)}
            parity[cv] = None if want is None else (got == {want})
        parity["source"] = "tests/golden/workload_runs.json (compiled reference helios_msm_vartime / selene_msm_vartime)"
        single = run_single(hs, torch, inputs, n, args, streams[0])
        for c, cv in enumerate(("helios", "selene")):
            want = ref.get((cv, n, args.seed + 17 * c))
            single[cv]["bit_exact_vs_reference"] = None if want is None else single[cv].pop("result") == want

    elapsed = t1 - t0
    value = args.steps * 2 * n / elapsed
    peak_products = hs.probe_int_mul_peak()
    ops, cstar = zip(*[algorithmic_bucket_ops(c, n) for c in (0, 1)])
    bucket_ms = [phases[c][0].get("bucket", 0.0) / max(1, phases[c][1]) for c in (0, 1)]
    achieved = (ops[0] + ops[1]) / ((bucket_ms[0] + bucket_ms[1]) * 1e-3)
    peak = 2 * peak_products
    per_curve_ms = {["helios", "selene"][c]: {k: v / max(1, phases[c][1]) for k, v in phases[c][0].items()}
                    for c in (0, 1)}
    msm_ms = [sum(per_curve_ms[cv].values()) for cv in ("helios", "selene")]
    msm_ops = [msm_int32_ops(c, n) for c in (0, 1)]
    traffic = measured_traffic(n)
    cus, clk_khz, dev_name = hs.device_info()
    nominal = 2.0 * cus * 4 * 16 * clk_khz * 1e3  # INT32 ops/s: one v_mad_u64_u32 per lane per clock
    # HBM phases (fixed convention, c* of the INT32 formula): sort = scalars read once plus an 8-byte
    # (key, index) record written and read per point and window; gather = one 64-byte affine base per
    # point and window, during the bucket phase
    W_star = -(-256 // cstar[0])
    bytes_sort, bytes_gather = 32 * n + 16 * W_star * n, 64 * W_star * n
    hbm = {}
    for c, cv in enumerate(("helios", "selene")):
        ph = per_curve_ms[cv]
        t_sort = (ph.get("recode", 0.0) + ph.get("sort", 0.0)) * 1e-3
        t_bucket = ph.get("bucket", 0.0) * 1e-3
        hbm[cv] = {"sort_GBps": bytes_sort / t_sort / 1e9 if t_sort else None,
                   "sort_frac": bytes_sort / t_sort / 8e12 if t_sort else None,
                   "gather_GBps": bytes_gather / t_bucket / 1e9 if t_bucket else None,
                   "gather_frac": bytes_gather / t_bucket / 8e12 if t_bucket else None}
    hbm.update({"bytes_sort": bytes_sort, "bytes_gather": bytes_gather, "peak_GBps": 8000,
                "phases": "sort = recode + sort phases, gather = bucket phase (isolated MSMs, HIP events)"})

    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded SSWU bases with 0.5% duplicates and 0.1% identities; scalars uniform mod order "
                "with 1% edge values)",
        "config": {
            "workload": "plain single MSM, one Helios and one Selene MSM of n=2^20 points per step",
            "n": n,
            "curves": ["helios", "selene"],
            "window_bits": [hs.window_bits(c, n) for c in (0, 1)],
            "result_mode": "reference-compatible",
            "bucket_method": f"batched affine, {args.affine_k} chains per lane" if args.affine_k else "xyzz mixed additions",
            "streams": args.streams,
        },
        "roofline": {
            "bound": "valu_int32",
            "kernel": "k_bucket_flat (bucket accumulation)",
            "achieved": achieved / 1e12,
            "peak": peak / 1e12,
            "unit": "TOP/s",
            "frac": achieved / peak,
            "traffic": (None if traffic is None else (traffic["helios"] + traffic["selene"]) / 2),
            "traffic_per_launch": traffic,
            "algorithmic_bytes_per_launch": -(-256 // cstar[0]) * n * (4 + 64),
            "algorithmic_ops_per_launch": {"helios": ops[0], "selene": ops[1], "c_star": list(cstar)},
            "kernel_ms": {"helios": bucket_ms[0], "selene": bucket_ms[1]},
            "kernel_ms_source": "HIP events around k_chunk_start + k_bucket_flat on the MSM's stream, "
                                f"{args.roofline_steps} isolated MSMs per curve after the timed steps",
            "peak_source": "measured v_mad_u64_u32 throughput x 2 INT32 ops per partial product",
            "whole_msm": {cv: {"ms": msm_ms[c], "int32_ops": msm_ops[c],
                               "frac": msm_ops[c] / (msm_ms[c] * 1e-3) / peak,
                               "frac_of_nominal": msm_ops[c] / (msm_ms[c] * 1e-3) / nominal}
                          for c, cv in enumerate(("helios", "selene"))},
            "peak_nominal": nominal / 1e12,
            "peak_nominal_source": f"{cus} CUs x 4 SIMDs x 16 lanes x {clk_khz / 1e6:.2f} GHz x 2 ops per "
                                   f"v_mad_u64_u32 ({dev_name})",
            "frac_of_nominal": achieved / nominal,
            "hbm_phases": hbm,
        },
        "gate": {
            "single_msm_frac_of_int32_peak": {cv: msm_ops[c] / (msm_ms[c] * 1e-3) / peak
                                              for c, cv in enumerate(("helios", "selene"))},
            "stream_frac_of_int32_peak": (msm_ops[0] + msm_ops[1]) * args.steps / elapsed / peak,
            "target": 0.5,
            "definition": "convention INT32 ops of one n=2^20 MSM / its isolated wall time (sum of its phases) / "
                          "measured peak; stream_frac: the same over the timed steps (MSMs overlapping)",
        },
        "parity_vs_reference": parity,
        "single_msm": single,
        "autotune": tuning,
        "phase_ms_isolated": per_curve_ms,
        "per_curve_points_per_s_isolated": {cv: n / (msm_ms[c] * 1e-3) for c, cv in enumerate(("helios", "selene"))},
    }
    L = None if args.no_cpu else ref_lib()
    rec["config1_helios_4096"] = None if args.no_config1 else run_config1(hs, torch, args, streams[0], L)
    rec["fcmpp_verify"] = None if args.no_fcmpp else run_fcmpp(hs, torch, args, streams[0], L)
    rec["fcmpp_tree_hash"] = None if args.no_fcmpp else run_tree_hash(hs, torch, args, streams[0], L)
    rec["batched_mix"] = None if args.no_mix else run_mix(hs, torch, args)
    rec["pedersen_batch"] = None if args.no_pedersen else run_pedersen(hs, torch, args)
    rec["serving"] = None if args.no_serve else run_serving(hs, torch, args)
    pts_cpu = [inputs[c][1].cpu().numpy().view(np.uint64) for c in (0, 1)]
    sc_cpu = [inputs[c][0].cpu().numpy() for c in (0, 1)]
    # the drop-in slot with host buffers: PCIe upload of 152 B per point included
    t_host = 0.0
    for c in (0, 1):
        hs.msm(c, sc_cpu[c], pts_cpu[c])
        t0 = time.perf_counter()
        hs.msm(c, sc_cpu[c], pts_cpu[c])
        t_host += time.perf_counter() - t0
    rec["host_slot_points_per_s"] = 2 * n / t_host
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(pts_cpu, sc_cpu, min(n, args.cpu_sample))
        add_cpu_legs(rec, pts_cpu, sc_cpu, args)
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
