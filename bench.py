"""Benchmark of the pil2-stark commitment hot path on one MI355X.

Default workload (the plain-large configuration): one trace of 2^22 rows x 128
uniform-random Goldilocks columns in HBM (row-major, as the CPU path hands it
over), LDE at blowup 4 to 2^24 rows, then the Poseidon2 (width 12, arity 3)
Merkle commit over the LDE rows. One step = one full commit (LDE + tree + root).

value = commit GElem/s = NExt * nCols / seconds per commit, inputs resident in
HBM when the timed region starts. The dominant kernel (Poseidon2 leaf hashing)
and the LDE passes are timed with HIP events on the commit's own stream.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config plain_large]
       [--no-cpu-baseline]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "pil2-proofman_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "commit GElem/s (LDE+Poseidon2 Merkle), batched vs plain; p50/p99 latency"

CONFIGS = {
    "plain_large": dict(n_bits=22, blowup_bits=2, cols=128, arity=3, seed=0x5EED0002),
    "plain_small": dict(n_bits=16, blowup_bits=1, cols=32, arity=3, seed=0x5EED0001),
    # 64 instances: log2 rows U{12..21}, columns U{8..256}, blowup 2 or 4 (seeded)
    "batched": dict(n_instances=64, arity=3, seed=0x5EED0003),
    # cubic-extension polynomial on a 2^22 domain, fold factor 8 per step down to 2^5
    "fri": dict(n_bits_ext=22, seed=0x5EED0004, arity=3),
    # quotient commit: q of 2^24 rows x 3 (cubic) -> qDeg 4 pieces of 2^22, extended to 2^24, merkelized
    "compute_q": dict(n_bits=22, n_bits_ext=24, q_deg=4, q_dim=3, arity=3, seed=0x5EED0006),
    # Poisson stream of 2000 jobs (log2 rows U{12..22}), HBM budget 200 GB, 32 openings per job
    "serving": dict(n_jobs=2000, arity=3, seed=0x5EED0005, budget=200e9, queries=32, llv=2, load=1.2),
}

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
# VALU lane-operations per second: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz
# (= the published 157.3 TFLOP/s FP32 vector rate / 2 for FMA / 2 for packed).
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
N_SIMD = 256 * 4
CLOCK_HZ = 2.4e9
# Poseidon2 field multiplications per permutation (algorithmic count):
# 8 full rounds x W x 4 (x^7) + 22 partial rounds x (4 + W)
MULS_PER_PERM = {8: 8 * 8 * 4 + 22 * 12, 12: 8 * 12 * 4 + 22 * 16, 16: 8 * 16 * 4 + 22 * 20}


def host_info():
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return info


def cpu_baseline(cfg, target_s=15.0, max_cols=None):
    """Time the reference's own CPU path (oracle/_ref: NTT_Goldilocks::LDE +
    Poseidon2Goldilocks<W>::merkletree, OpenMP over all host cores) on a bounded
    sample of the same workload: full trace height, same blowup and arity, and as
    many columns as fit about `target_s` seconds (sized from a 1-column probe).
    Falls back to our C restatement if the reference library cannot be loaded."""
    import datagen
    from oracle.oracle import Oracle, RefLib
    n, ne, arity = 1 << cfg["n_bits"], 1 << (cfg["n_bits"] + cfg["blowup_bits"]), cfg["arity"]
    o = Oracle()
    try:
        lib, kind = RefLib(), "reference"
        tree_len = o.tree_num_elements(ne, arity)
        lde_f, tree_f = (lambda t: lib.lde(t, ne)), (lambda e: lib.merkletree(e, arity, tree_len))
        threads = lib.num_threads()
    except (OSError, FileNotFoundError):
        lib, kind = o, "port"
        lde_f, tree_f = (lambda t: o.lde(t, ne)), (lambda e: o.merkletree(e, arity))
        threads = o.num_threads()
    split = [0.0, 0.0]  # LDE seconds, merkletree seconds

    def run(t):
        t0 = time.perf_counter()
        e = lde_f(t)
        t1 = time.perf_counter()
        tree_f(e)
        split[0] += t1 - t0
        split[1] += time.perf_counter() - t1

    pc_ = 4
    probe = datagen.uniform(n * pc_, cfg["seed"]).reshape(n, pc_)
    t0 = time.perf_counter()
    run(probe)
    t_probe = time.perf_counter() - t0
    split[0] = split[1] = 0.0
    cols = int(max(1, min(max_cols or cfg["cols"], pc_ * target_s / max(t_probe, 1e-3))))
    trace = datagen.uniform(n * cols, cfg["seed"]).reshape(n, cols)
    reps, dt = 0, 0.0
    while reps == 0 or (dt < target_s / 2 and reps < 1000):  # a short workload runs again up to ~target_s/2
        t0 = time.perf_counter()
        run(trace)
        dt += time.perf_counter() - t0
        reps += 1
    return {"value": reps * ne * cols / dt / 1e9, "unit": "GElem/s", "cores": threads, "kind": kind,
            "lde_s_per_run": round(split[0] / reps, 4), "merkletree_s_per_run": round(split[1] / reps, 4),
            "sample": "2^%d x %d cols -> 2^%d rows (blowup %d), Poseidon2 W%d arity %d; %d run(s), %.2f s"
                      % (cfg["n_bits"], cols, cfg["n_bits"] + cfg["blowup_bits"], 1 << cfg["blowup_bits"],
                         4 * arity, arity, reps, dt),
            **host_info()}


def cpu_baseline_batched(shapes, arity, target_s=15.0, what="batch"):
    """Reference CPU path (oracle/_ref, all host cores) over the workload's own
    instances (the batch's, or the served jobs'), taken in generator order while
    the estimated time stays within about `target_s` seconds (rate estimated from
    the instances already run)."""
    import datagen
    from oracle.oracle import Oracle, RefLib
    o = Oracle()
    try:
        lib, kind = RefLib(), "reference"
        threads = lib.num_threads()
    except (OSError, FileNotFoundError):
        lib, kind = o, "port"
        threads = o.num_threads()
    elems, spent, used, lde_s = 0, 0.0, [], 0.0
    for j in shapes:
        n, ne, c = 1 << j["n_bits"], 1 << j["n_bits_ext"], j["n_cols"]
        if spent > 0 and spent + ne * c / (elems / spent) > target_s:
            continue
        trace = datagen.uniform(n * c, j["seed"]).reshape(n, c)
        t0 = time.perf_counter()
        if kind == "reference":
            e = lib.lde(trace, ne)
            t1 = time.perf_counter()
            lib.merkletree(e, arity, o.tree_num_elements(ne, arity))
        else:
            e = o.lde(trace, ne)
            t1 = time.perf_counter()
            o.merkletree(e, arity)
        spent += time.perf_counter() - t0
        lde_s += t1 - t0
        elems += ne * c
        used.append("2^%d x %d (x%d)" % (j["n_bits"], c, ne // n))
    return {"value": elems / spent / 1e9, "unit": "GElem/s", "cores": threads, "kind": kind,
            "lde_s": round(lde_s, 3), "merkletree_s": round(spent - lde_s, 3),
            "sample": "%d of the %s's instances [%s], LDE + Poseidon2 W%d arity %d Merkle, %.2f s"
                      % (len(used), what, ", ".join(used), 4 * arity, arity, spent),
            **host_info()}


_MARKER = []


def profile_region(args, on):
    """With --profile-region, mark the start (on) and end of a timed region in a
    profile: a one-block k_fill_uniform launch (grid 256, unlike any input fill),
    with the device idle on both sides. The profile summaries keep the dispatches
    between the two markers: the timed steps only, without warm-up, tuning,
    calibration or the per-instance pass."""
    if not getattr(args, "profile_region", False):
        return
    import torch
    import pil2_commit as pc
    if not _MARKER:
        _MARKER.append(pc.empty(1))
    torch.cuda.synchronize()
    pc.fill_uniform(_MARKER[0], 1 if on else 2)
    torch.cuda.synchronize()


_MUL_PEAK = []


def measured_mul_peak(lanes=1 << 21, iters=1024, reps=3):
    """Measured throughput of the Poseidon2 kernels' own reduced product
    (pil2hip_mul_peak: 18 independent chains per lane, the permutation's
    p2_mul3), in T products/s, timed with HIP events on the launch stream.
    Measured once per process, outside every timed region."""
    if not _MUL_PEAK:
        import torch
        import pil2_commit as pc
        c = pc.MUL_PEAK_CHAINS
        inp = pc.empty(2 * c * lanes)
        pc.fill_uniform(inp, 0x5EED00FF)
        out = pc.empty(c * lanes)
        pc.mul_peak(out, inp, lanes, iters)
        st = torch.cuda.current_stream()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(st)
        for _ in range(reps):
            pc.mul_peak(out, inp, lanes, iters)
        t1.record(st)
        t1.synchronize()
        ms = t0.elapsed_time(t1) / reps
        _MUL_PEAK.append((lanes * iters * c / (ms / 1e3) / 1e12, ms))
        del inp, out
    return _MUL_PEAK[0]


def _leaf_roofline(width, perms, leaf_ms, leaf_bytes, traffic, launches):
    """VALU roofline of the Poseidon2 leaf kernel: algorithmic Goldilocks
    multiplications (736 per W12 permutation) against the rate at which the VALU
    can issue the four 32x32->64 partial products each one needs (v_mad_u64_u32
    at the single-issue lane rate), i.e. a products-only ceiling."""
    leaf_s = leaf_ms / 1e3
    muls_per_s = perms * MULS_PER_PERM[width] / leaf_s
    mul_peak = VALU_PEAK_TOPS / 4.0  # T mul/s
    return {
        "bound": "valu", "kernel": "k_p2_leaves<%d>" % width,
        "achieved": round(muls_per_s / 1e12, 4), "peak": round(mul_peak, 4), "unit": "Tmul/s",
        "frac": round(muls_per_s / 1e12 / mul_peak, 4),
        "model": "64x64 field products: %d per W%d permutation, 4 v_mad_u64_u32 each at %.1f T lane-op/s "
                 "(published 157.3 TFLOP/s FP32 vector / 2 FMA / 2 packed)" % (MULS_PER_PERM[width], width,
                                                                             VALU_PEAK_TOPS),
        "perm_per_s": perms / leaf_s,
        "mul_peak_measured": round(measured_mul_peak()[0], 4),
        "frac_of_measured_mul_peak": round(muls_per_s / 1e12 / measured_mul_peak()[0], 4),
        "mul_peak_note": "pil2hip_mul_peak: the kernel's own reduced product (11 single-rate instructions), "
                         "18 independent chains per lane, 2^21 lanes",
        "hbm_achieved_GBs": leaf_bytes / leaf_s / 1e9, "hbm_frac": leaf_bytes / leaf_s / 1e9 / HBM_PEAK_GBS,
        "traffic": traffic,
        "avg_ms": leaf_ms / launches,
    }


def _lde_kernels(nb, nbe, cols):
    """The four kernels of one LDE as the library launches them for this shape
    (blowup 4 with two or more columns takes the two-column block kernel)."""
    block = "k_lde_block2" if nbe - nb == 2 and cols >= 2 else "k_lde_block<false, false, false>"
    return ("k_transpose_rm2cm", "k_ntt_strided_inv<false>", block, "k_ntt_strided_fwd<false, false>")
# single-rate VALU instructions of one radix-2 butterfly at the floor: a reduced
# Goldilocks product (4 v_mad_u64_u32 + a 7-instruction reduction) plus a 6-instruction
# add and a 6-instruction sub
INSTR_PER_BUTTERFLY = 23


def _lde_valu_roofline(nb, nbe, cols, lde_ms, config):
    """VALU roofline of the LDE: the algorithmic radix-2 butterflies (the N-point
    inverse transform, then the NExt-point forward transform of the zero-extended
    coefficients, whose first log2(B) stages only copy) against the rate at which
    the VALU issues them at INSTR_PER_BUTTERFLY single-rate instructions each; plus
    the measured share of VALU issue slots the four kernels fill (committed
    counters, live time)."""
    n, ne = 1 << nb, 1 << nbe
    bfly = cols * (n // 2 * nb + ne // 2 * nb)
    peak = VALU_PEAK_TOPS / INSTR_PER_BUTTERFLY  # T butterflies/s
    achieved = bfly / (lde_ms / 1e3) / 1e12
    out = {"bound": "valu", "achieved": round(achieved, 4), "peak": round(peak, 4), "unit": "Tbutterfly/s",
           "frac": round(achieved / peak, 4), "butterflies": bfly,
           "floor_ms": round(bfly * INSTR_PER_BUTTERFLY / (VALU_PEAK_TOPS * 1e12) * 1e3, 3),
           "model": "radix-2 butterflies of the iNTT (N/2 log N) and the forward NTT after zero extension "
                    "(NExt/2 log N), %d single-rate instructions each at %.1f T lane-op/s"
                    % (INSTR_PER_BUTTERFLY, VALU_PEAK_TOPS)}
    cyc, lanes = 0.0, 0.0
    for k in _lde_kernels(nb, nbe, cols):
        c = _profile_kernel(k, config)
        if "valu_wave_instr_per_launch" not in c:
            return out
        cyc += 4.0 * (c["valu_wave_instr_per_launch"] - c["valu_dual_issue_quad_cycles_per_launch"]) / N_SIMD
        lanes += 64.0 * c["valu_wave_instr_per_launch"]
    out["valu_issue_frac"] = round(cyc / (CLOCK_HZ * lde_ms / 1e3), 4)
    out["lane_instructions"] = lanes
    return out


def _with_issue(roofline, config):
    """Adds the leaf kernel's measured VALU issue share to its roofline object."""
    f = _valu_issue(roofline["kernel"], config, roofline["avg_ms"])
    if f is not None:
        roofline["valu_issue_frac"] = round(f, 4)
        roofline["valu_issue_note"] = ("VALU issue slots filled (rocprof SQ_INSTS_VALU - SQ_ACTIVE_INST_VALU2 per "
                                       "launch from profiles/, 1024 SIMDs at 2.4 GHz); the products-only 'frac' "
                                       "counts 4 of the 11 single-rate instructions a reduced product takes")
    return roofline


def _reference_fixture():
    """The compiled reference's expectations at the configurations' real sizes
    (tests/golden/config_reference.json), or None where the file is absent."""
    try:
        with open(os.path.join(ROOT, "tests", "golden", "config_reference.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def run_batched(args, cfg, embedded=False):
    """Mixed-size batch: all instances resident in HBM, committed (a) one by one
    through the plain per-instance path (commit + root back to the host, the way
    proofman calls commit_witness per instance) and (b) as one hipGraph launch of
    pil2hip_commit_batch. Both produce every instance's root; they must agree, and
    equal the compiled reference's roots where the fixture file is present.
    embedded: return a summary (at most 3 timed graph replays, no CPU baseline)
    for the default benchmark line instead of printing a line of its own."""
    import ctypes
    import torch
    import datagen
    import pil2_commit as pc

    arity = cfg["arity"]
    width = 4 * arity
    shapes = datagen.batched_workload(seed=cfg["seed"], n_instances=cfg["n_instances"])
    pc.set_hash_family(2)
    dev = torch.device("cuda:0")
    need = 0
    for j in shapes:
        n, ne, c = 1 << j["n_bits"], 1 << j["n_bits_ext"], j["n_cols"]
        need += 8 * (2 * n * c + ne * c + pc.get_tree_num_elements(ne, arity))
    free, total = torch.cuda.mem_get_info(dev)
    if need > free - (4 << 30):
        raise SystemExit("batched workload needs %.1f GB of HBM, %.1f GB free" % (need / 1e9, free / 1e9))
    jobs = []
    for j in shapes:
        tr = pc.empty((1 << j["n_bits"]) * j["n_cols"], dev)
        pc.fill_uniform(tr, j["seed"])
        jobs.append(dict(j, trace=tr))
    batch = pc.CommitBatch(jobs, arity=arity)
    torch.cuda.synchronize()
    elems = sum((1 << j["n_bits_ext"]) * j["n_cols"] for j in shapes)

    # (a) plain per-instance path, with phase timing per commit
    def plain_pass(timed):
        roots = []
        for jb in batch.jobs:
            pc.commit(jb["trace"], arity, jb["n_bits"], jb["n_bits_ext"], jb["n_cols"], lde=jb["lde"],
                      nodes=jb["nodes"], scratch=jb["scratch"])
            roots.append(pc.MerkleTreeGL(arity, 0, True, 1 << jb["n_bits_ext"], jb["n_cols"],
                                         nodes=jb["nodes"]).get_root())
        return roots

    plain_pass(False)
    torch.cuda.synchronize()
    pc.lib.pil2hip_timing_read(None, True)
    pc.lib.pil2hip_timing_enable(True)
    t0 = time.perf_counter()
    plain_roots = plain_pass(True)
    torch.cuda.synchronize()
    dt_plain = time.perf_counter() - t0
    pc.lib.pil2hip_timing_enable(False)
    ph = (ctypes.c_double * 6)()
    ncalls = pc.lib.pil2hip_timing_read(ph, True)
    phase_ms = [ph[i] for i in range(6)]  # sums over the pass
    plain_value = elems / dt_plain / 1e9

    # (b) batched: the first call autotunes the launch plan (graph branches, NTT tile
    # and radix, Merkle level schedule) and instantiates the graph; later calls replay it
    # (the same request replays the graph the first call built: an all-0 plan is tuned per shape)
    t_tune = time.perf_counter()
    request = (dict(parse_plan(args.plan), n_streams=args.batch_streams) if args.plan
               else dict(n_streams=args.batch_streams))
    streams = batch.launch(plan=request)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t_tune
    plan = batch.plan
    shape_plans = {}
    for j in shapes:
        p = pc.batch_shape_plan(j["n_bits"], j["n_bits_ext"], j["n_cols"], arity)[1]
        k = "tile 2^%d radix %d coop %d" % (p["ntt_tile_log"], 1 << p["ntt_radix_log"], p["coop_max_parents"])
        shape_plans[k] = shape_plans.get(k, 0) + 1
    b_steps, b_warm = (min(args.steps, 3), min(args.warmup, 1)) if embedded else (args.steps, args.warmup)
    for _ in range(b_warm):
        batch.launch(plan=request)
    torch.cuda.synchronize()
    profile_region(args, True)
    t0 = time.perf_counter()
    for _ in range(b_steps):
        batch.launch(plan=request)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    profile_region(args, False)
    value = elems / (dt / b_steps) / 1e9
    roots = batch.roots_host()
    match = all(list(map(int, roots[k])) == list(map(int, plain_roots[k])) for k in range(len(shapes)))
    fx = _reference_fixture()
    ref_match = None
    if fx is not None and fx["batched"]["seed"] == cfg["seed"] and len(fx["batched"]["roots"]) == len(shapes):
        ref_match = all([int(v) for v in roots[k]] == [int(x, 16) for x in fx["batched"]["roots"][k]]
                        for k in range(len(shapes)))
    del batch, jobs
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if embedded:
        out = {"workload": "64 instances, log2 rows U{12..21}, cols U{8..256}, blowup 2/4 (seed 0x%X), one "
                           "hipGraph launch" % cfg["seed"],
               "value": round(value, 4), "unit": "GElem/s", "ms_per_step": round(dt * 1e3 / b_steps, 3),
               "steps": b_steps, "lde_elements": elems, "graph_branches": streams, "autotune_s": round(t_tune, 3),
               "plain_value": round(plain_value, 4), "plain_ms_per_pass": round(dt_plain * 1e3, 3),
               "batched_over_plain": round(value / plain_value, 4), "roots_match_plain": match,
               "roots_match_reference": ref_match}
        if not match or ref_match is False:
            out["error"] = "batched roots differ from the per-instance or reference roots"
        return out

    rate = width - 4
    perms = sum((1 << j["n_bits_ext"]) * ((j["n_cols"] + rate - 1) // rate) for j in shapes)
    leaf_bytes = sum(8 * (1 << j["n_bits_ext"]) * j["n_cols"] + 32 * (1 << j["n_bits_ext"]) for j in shapes)
    roofline = _leaf_roofline(width, perms, phase_ms[4], leaf_bytes,
                              _profile_traffic("k_p2_leaves<%d>" % width, "batched"), len(shapes))
    roofline["measured_on"] = "per-instance commits of the same batch (HIP events around each leaf launch)"
    lde_bytes = sum(8 * j["n_cols"] * ((1 << j["n_bits"]) + (1 << j["n_bits_ext"])) for j in shapes)
    lde_ms = sum(phase_ms[0:4])
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "GElem/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64 (Goldilocks field, exact integer)",
        "data": "synthetic: seeded uniform Goldilocks elements generated in HBM",
        "config": {"workload": "batched: %d instances, log2 rows U{12..21}, cols U{8..256}, blowup 2/4 "
                               "(seed 0x%X), Poseidon2 W%d arity %d Merkle commit each, one hipGraph launch"
                               % (len(shapes), cfg["seed"], width, arity),
                   "n_instances": len(shapes), "lde_elements": elems, "arity": arity, "seed": cfg["seed"],
                   "hbm_resident_GB": round(need / 1e9, 2), "path": "batched (pil2hip_commit_batch)"},
        "batched": {"graph_branches": streams, "plan": plan, "autotune_s": round(t_tune, 3),
                    "instances_per_shape_plan": shape_plans if not args.plan else None,
                    "roots_match_plain": match, "roots_match_reference": ref_match},
        "plain": {"value": round(plain_value, 4), "unit": "GElem/s", "ms_per_pass": round(dt_plain * 1e3, 3),
                  "path": "per-instance pil2hip_commit + root to host, one instance at a time",
                  "phases_ms_sum": dict(zip(["transpose", "intt_strided", "lde_block", "ntt_strided", "leaves",
                                             "tree_levels"], [round(x, 3) for x in phase_ms])),
                  "timed_commits": ncalls},
        "batched_over_plain": round(value / plain_value, 4),
        "roofline": roofline,
        "roofline_lde": {"bound": "hbm", "achieved": lde_bytes / (lde_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": lde_bytes / (lde_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                         "traffic": None, "avg_ms": lde_ms / len(shapes)},
        "device": torch.cuda.get_device_name(0), "host": platform.node(),
    }
    if not match or ref_match is False:
        out["error"] = "batched roots differ from the per-instance or reference roots"
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batched(shapes, arity, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    if not match or ref_match is False:
        sys.exit(1)


def run_fri(args, cfg):
    """FRI folding with a Merkle commit per layer, in the reference's order
    (gen_proof: step 0 commits the LDE-domain polynomial, every later step folds
    by 2^3 and commits, the last step only folds): FRI::fold (fri.hpp:32-99),
    getTransposed + merkelize (fri.hpp:102-109, 209-226). One step of the bench is
    the whole layer sequence; value = cubic input elements of all folds per second."""
    import numpy as np
    import torch
    import datagen
    import pil2_commit as pc

    nbe, arity = cfg["n_bits_ext"], cfg["arity"]
    steps = pc.FRI.steps(nbe)
    chal = datagen.challenges(len(steps), cfg["seed"] + 1)
    pc.set_hash_family(2)
    dev = torch.device("cuda:0")
    src = pc.empty(3 << nbe, dev)
    pc.fill_uniform(src, cfg["seed"])
    work = pc.empty(3 << nbe, dev)
    trees = []
    for s_ in range(len(steps) - 1):
        cur, nxt = steps[s_], steps[s_ + 1]
        trees.append(pc.MerkleTreeGL(arity, 2, False, 1 << nxt, 3 << (cur - nxt)))
    for t in trees:
        t.source = pc.empty(t.height * t.width, dev)
        t.nodes = pc.empty(t.num_nodes, dev)
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * len(steps) + 1)]

    def layers(record):
        roots = []
        for s_ in range(len(steps)):
            if s_ > 0:
                pc.FRI.fold(s_, src if s_ == 1 else work, chal[s_], nbe, steps[s_ - 1], steps[s_],
                            out=work)
            if record:
                ev[2 * s_].record(st)
            if s_ < len(steps) - 1:
                pol = src if s_ == 0 else work
                t = trees[s_]
                degree = 1 << steps[s_]
                pc._check(pc._lib.pil2hip_fri_transpose(pc._ptr(t.source), pc._ptr(pol), degree, steps[s_ + 1],
                                                        pc._stream()))
                t.row_stride, t.col_stride = t.width, 1
                t.merkelize()
                if record:
                    ev[2 * s_ + 1].record(st)
                roots.append(t)
        return roots

    # timed: the whole loop as one cached hipGraph (pil2hip_fri_layers), challenges in HBM
    d_ch = pc.to_device(np.asarray(chal, dtype=np.uint64).reshape(-1), dev)
    for _ in range(args.warmup):
        pc.FRI.layers(src, work, d_ch, nbe, steps, trees, arity=arity, graph=True)
    torch.cuda.synchronize()  # the warm-up's work ends before the clock starts
    profile_region(args, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pc.FRI.layers(src, work, d_ch, nbe, steps, trees, arity=arity, graph=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    profile_region(args, False)
    graph_roots = [["%016x" % v for v in t.get_root()] for t in trees]
    graph_final = pc.to_host(work)[:3 << steps[-1]]
    # the same loop launched call by call (per-layer entry points), for comparison
    for _ in range(args.warmup):
        layers(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        layers(False)
    torch.cuda.synchronize()
    dt_eager = (time.perf_counter() - t0) / args.steps
    ev_start = torch.cuda.Event(enable_timing=True)
    ev_start.record(st)
    layers(True)
    torch.cuda.synchronize()
    fold_ms, commit_ms, prev = 0.0, 0.0, ev_start
    for s_ in range(len(steps)):
        fold_ms += prev.elapsed_time(ev[2 * s_])
        prev = ev[2 * s_]
        if s_ < len(steps) - 1:
            commit_ms += prev.elapsed_time(ev[2 * s_ + 1])
            prev = ev[2 * s_ + 1]
    roots = [["%016x" % v for v in t.get_root()] for t in trees]
    final = pc.to_host(work)[:3 << steps[-1]]
    elems = sum(1 << steps[s_ - 1] for s_ in range(1, len(steps)))
    perms = sum(t.height * ((t.width + 4 * arity - 5) // (4 * arity - 4)) for t in trees)
    out = {
        "metric": METRIC, "value": round(elems / dt / 1e9, 4), "unit": "G cubic elem/s folded (with per-layer commits)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 (Goldilocks cubic extension, exact integer)",
        "data": "synthetic: seeded uniform cubic evaluations generated in HBM, seeded challenges",
        "config": {"workload": "fri: degree 2^%d cubic polynomial, steps %s, Poseidon2 W%d arity %d commit per layer"
                               % (nbe, steps, 4 * arity, arity), "steps": steps, "seed": cfg["seed"]},
        "path": "pil2hip_fri_layers: fold + layer commits as one cached hipGraph",
        "per_call_path": {"ms_per_step": round(dt_eager * 1e3, 4),
                          "note": "the same loop through the per-layer entry points (one launch per kernel)"},
        "graph_matches_per_call": bool(graph_roots == roots and (graph_final == final).all()),
        "phases_ms": {"fold": round(fold_ms, 4), "transpose_commit": round(commit_ms, 4),
                      "note": "from the per-call path with events between layers"},
        "roofline": {"bound": "valu", "kernel": "k_p2_leaves<%d> + k_p2_nodes (layer trees)" % (4 * arity),
                     "achieved_perm_per_s": perms / (commit_ms / 1e3), "traffic": None,
                     "note": "layer commits dominate; per-layer trees are small (launch-bound below 2^13 rows)"},
        "roots": roots,
        "final_pol_sha256": __import__("hashlib").sha256(final.tobytes()).hexdigest(),
        "device": torch.cuda.get_device_name(0), "host": platform.node(),
    }
    if not args.no_cpu_baseline:
        from oracle.oracle import Oracle
        o = Oracle()
        t0 = time.perf_counter()
        host = datagen.uniform(3 << nbe, cfg["seed"])
        t_gen = time.perf_counter() - t0
        src_host, ok, cpu_dt, reps = host, True, 0.0, 0
        while reps == 0 or (cpu_dt < args.cpu_seconds / 2 and reps < 100):
            host = src_host
            t0 = time.perf_counter()
            for s_ in range(len(steps)):
                if s_ > 0:
                    host = o.fri_fold(host, chal[s_], s_, nbe, steps[s_ - 1], steps[s_])
                if s_ < len(steps) - 1:
                    nxt = steps[s_ + 1]
                    aux = o.fri_transpose(host, 1 << steps[s_], nxt)
                    r = o.merkletree(aux.reshape(1 << nxt, 3 << (steps[s_] - nxt)), arity)[-4:]
                    ok &= ["%016x" % v for v in r] == roots[s_]
            cpu_dt += time.perf_counter() - t0
            reps += 1
        ok &= bool((host[:3 << steps[-1]] == final).all())
        out["cpu_baseline"] = {"value": reps * elems / cpu_dt / 1e9, "unit": out["unit"], "cores": o.num_threads(),
                               "kind": "port", "sample": "the full layer sequence, %d run(s), %.2f s (input "
                                                         "generation %.2f s not counted)" % (reps, cpu_dt, t_gen),
                               **host_info()}
        out["parity_vs_oracle"] = ok
    print(json.dumps(out), flush=True)


def run_serving(args, cfg, embedded=False):
    """Serving: jobs arrive as a Poisson stream and are admitted by the commit
    server (pil2hip_serve_*) under an HBM budget while earlier jobs still run;
    each job commits its trace (synthetic values generated in HBM from its seed)
    and returns its root and 32 group proofs (getGroupProof, llv 2). The arrival
    rate is 1.2x the throughput the same server sustains on a closed burst of
    jobs. Reports aggregate GElem/s and p50/p99 of completion - arrival. The jobs
    of the stream's prefix that the fixture file holds are checked against the
    compiled reference's roots and proof digests as they complete.
    embedded: a bounded stream (the first 400 jobs, then the first 200 at 0.7x)
    summarised for the default benchmark line, no CPU baseline."""
    import hashlib
    import numpy as np
    import torch
    import datagen
    import pil2_commit as pc

    arity, llv, nq = cfg["arity"], cfg["llv"], cfg["queries"]
    n_jobs = args.jobs or (400 if embedded else cfg["n_jobs"])
    shapes = datagen.batched_workload(seed=cfg["seed"], n_instances=n_jobs, min_log=12, max_log=22)
    qwords = datagen.raw_words(n_jobs * nq, cfg["seed"] ^ 0xA5A5A5A5).reshape(n_jobs, nq)
    pc.set_hash_family(2)
    free, _ = torch.cuda.mem_get_info(torch.device("cuda:0"))
    budget = int(min(cfg["budget"], free - (8 << 30)))
    eng = pc.ServeEngine(budget, n_streams=args.serve_streams, max_inflight=args.serve_inflight, arity=arity,
                         llv=llv)
    elems = [(1 << j["n_bits_ext"]) * j["n_cols"] for j in shapes]
    fx = _reference_fixture()
    fx_jobs = fx["serving"]["jobs"] if fx is not None and fx["serving"]["seed"] == cfg["seed"] else []
    checked = {"jobs": 0, "mismatch": 0}

    def check(rec):
        """Root and proof digest of a fixture job against the compiled reference's."""
        k = rec["tag"]
        if rec["status"] != 0 or k >= len(fx_jobs):
            return
        fj = fx_jobs[k]
        got = hashlib.sha256(rec["proofs"][:nq * rec["pw"]].astype("<u8").tobytes()).hexdigest()
        ok = [int(v) for v in rec["root"]] == [int(x, 16) for x in fj["root"]] and got == fj["proofs_sha256"]
        checked["jobs"] += 1
        checked["mismatch"] += 0 if ok else 1

    def submit(k):
        j = shapes[k]
        q = qwords[k] % np.uint64(1 << j["n_bits_ext"])
        return eng.submit(j["n_bits"], j["n_bits_ext"], j["n_cols"], queries=q, seed=j["seed"], tag=k)

    # capacity: a closed burst of the first jobs, all submitted at once
    n_cal = min(48, n_jobs)
    t0 = time.perf_counter()
    for k in range(n_cal):
        submit(k)
    got = 0
    while got < n_cal:
        for rec in eng.poll(max_n=256, timeout_ms=200):
            assert rec["status"] == 0, rec["status"]
            check(rec)
            eng.forget(rec["job_id"])
            got += 1
    cal_s = time.perf_counter() - t0
    capacity = sum(elems[:n_cal]) / cal_s  # elements per second
    mean_elems = sum(elems) / n_jobs

    def open_loop(jobs, load, seed):
        """Poisson arrivals of `jobs` (indices) at `load` x capacity; returns latency
        and queue-delay arrays (ms), the span (s) and the failed-job count."""
        lam_ = load * capacity / mean_elems  # jobs per second
        arr = np.cumsum(np.random.default_rng(seed).exponential(1.0 / lam_, len(jobs)))
        lat_, qd_, done_, bad_, k_ = [], [], 0, 0, 0
        t_first_ = t_last_ = None
        start = time.perf_counter()
        base_us = eng.now_us()
        last_progress = start
        while done_ < len(jobs):
            now = time.perf_counter() - start
            while k_ < len(jobs) and arr[k_] <= now:
                submit(jobs[k_])
                k_ += 1
            wait_ms = 0 if k_ >= len(jobs) else max(0, int((arr[k_] - (time.perf_counter() - start)) * 1e3))
            for rec in eng.poll(max_n=256, timeout_ms=min(wait_ms, 50) if k_ < len(jobs) else 100):
                if rec["status"] != 0:
                    bad_ += 1
                check(rec)
                lat_.append((rec["t_done_us"] - rec["t_submit_us"]) / 1e3)
                qd_.append((rec["t_admit_us"] - rec["t_submit_us"]) / 1e3)
                t_first_ = rec["t_submit_us"] if t_first_ is None else min(t_first_, rec["t_submit_us"])
                t_last_ = rec["t_done_us"] if t_last_ is None else max(t_last_, rec["t_done_us"])
                eng.forget(rec["job_id"])
                done_ += 1
            if time.perf_counter() - last_progress > 20:
                last_progress = time.perf_counter()
                print("serving: %d/%d submitted, %d done" % (k_, len(jobs), done_), file=sys.stderr, flush=True)
        return np.array(lat_), np.array(qd_), (t_last_ - min(t_first_, base_us)) / 1e6, bad_, lam_

    # the measured stream: every job at the configured offered load
    lat, qdelay, span_s, bad, lam = open_loop(list(range(n_jobs)), cfg["load"], cfg["seed"])
    # the same server below capacity: the first jobs of the stream again at 0.7x, where
    # latency is the service time plus little queueing
    n_low = min(200 if embedded else 400, n_jobs)
    lat_lo, qd_lo, span_lo, bad_lo, lam_lo = open_loop(list(range(n_low)), 0.7, cfg["seed"] + 1)
    bad += bad_lo
    st = eng.stats()
    eng.close()
    value = sum(elems) / span_s / 1e9
    bad += checked["mismatch"]
    lat_ms = {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)}
    lat_lo_ms = {"p50": round(float(np.percentile(lat_lo, 50)), 3), "p99": round(float(np.percentile(lat_lo, 99)), 3)}
    if embedded:
        return {"workload": "first %d jobs of the serving stream (log2 rows U{12..22}, cols U{8..256}, blowup 2/4, "
                            "seed 0x%X), HBM budget %.0f GB, %d group proofs per job (llv %d)"
                            % (n_jobs, cfg["seed"], budget / 1e9, nq, llv),
                "offered_load": cfg["load"], "value": round(value, 4), "unit": "GElem/s",
                "latency_ms": lat_ms, "arrival_rate_jobs_s": round(lam, 3),
                "below_capacity": {"offered_load": 0.7, "jobs": n_low, "latency_ms": lat_lo_ms,
                                   "GElem_s": round(sum(elems[:n_low]) / span_lo / 1e9, 4)},
                "capacity_GElem_s": round(capacity / 1e9, 4), "failed_jobs": bad,
                "reference_checked_jobs": checked["jobs"], "reference_mismatches": checked["mismatch"],
                "peak_running": st["peak_running"]}
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "GElem/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(span_s * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 (Goldilocks field, exact integer)",
        "data": "synthetic: each job's uniform trace generated in HBM from its seed at admission",
        "config": {"workload": "serving: %d Poisson jobs (log2 rows U{12..22}, cols U{8..256}, blowup 2/4, "
                               "seed 0x%X), HBM budget %.0f GB, %d group proofs per job (llv %d), "
                               "offered load %.1fx capacity" % (n_jobs, cfg["seed"], budget / 1e9, nq, llv,
                                                                cfg["load"]),
                   "n_jobs": n_jobs, "arity": arity, "streams": args.serve_streams,
                   "max_inflight": args.serve_inflight, "path": "served (pil2hip_serve_*)"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "mean": round(float(lat.mean()), 3), "max": round(float(lat.max()), 3)},
        "queue_delay_ms": {"p50": round(float(np.percentile(qdelay, 50)), 3),
                           "p99": round(float(np.percentile(qdelay, 99)), 3)},
        "below_capacity": {"offered_load": 0.7, "jobs": n_low, "arrival_rate_jobs_s": round(lam_lo, 3),
                           "latency_ms": {"p50": round(float(np.percentile(lat_lo, 50)), 3),
                                          "p99": round(float(np.percentile(lat_lo, 99)), 3)},
                           "queue_delay_ms": {"p50": round(float(np.percentile(qd_lo, 50)), 3),
                                              "p99": round(float(np.percentile(qd_lo, 99)), 3)},
                           "GElem_s": round(sum(elems[:n_low]) / span_lo / 1e9, 4)},
        "capacity_GElem_s": round(capacity / 1e9, 4), "arrival_rate_jobs_s": round(lam, 3),
        "lde_elements": int(sum(elems)), "failed_jobs": bad,
        "reference_checked_jobs": checked["jobs"], "reference_mismatches": checked["mismatch"],
        "server": {k2: st[k2] for k2 in ("admitted", "completed", "failed", "peak_bytes", "peak_running")},
        "device": torch.cuda.get_device_name(0), "host": platform.node(),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batched(shapes, arity, args.cpu_seconds, what="served job stream")
        out["cpu_baseline"]["sample"] += " (commits only; the 32 openings per job are not timed on the CPU)"
    print(json.dumps(out), flush=True)
    if bad:
        sys.exit(1)


def run_compute_q(args, cfg):
    """Quotient commit: Starks::computeQ (starks.hpp:187-239) + the stage's
    merkelize. q (row-major NExt x qDim, the prover's layout) is resident in HBM;
    one step = INTT over NExt, regrouping by shiftIn^p, zero-extended NTT over NExt
    of qDeg*qDim columns, Merkle tree. value = NExt * qDeg*qDim / s (the columns
    committed, as for the trace commit)."""
    import numpy as np
    import torch
    import datagen
    import pil2_commit as pc

    nb, nbe, qdeg, qdim, arity = cfg["n_bits"], cfg["n_bits_ext"], cfg["q_deg"], cfg["q_dim"], cfg["arity"]
    n, ne, width, nc = 1 << nb, 1 << nbe, 4 * arity, qdeg * qdim
    pc.set_hash_family(2)
    dev = torch.device("cuda:0")
    q = pc.empty(ne * qdim, dev)
    pc.fill_uniform(q, cfg["seed"])
    cmq = pc.empty(ne * nc, dev)
    nodes = pc.empty(pc.get_tree_num_elements(ne, arity), dev)
    scratch = pc.empty(n * nc, dev)
    step = lambda: pc.compute_q(q, nb, nbe, qdeg, q_dim=qdim, arity=arity, cmq=cmq, nodes=nodes, scratch=scratch)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()  # the warm-up's work ends before the clock starts
    profile_region(args, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    profile_region(args, False)
    # phases on the same stream: the transform alone, then the tree alone
    st = torch.cuda.current_stream()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record(st)
    pc.compute_q(q, nb, nbe, qdeg, q_dim=qdim, arity=arity, cmq=cmq, scratch=scratch, merkelize=False)
    e1.record(st)
    pc._check(pc.lib.pil2hip_merkelize(pc._ptr(nodes), pc._ptr(cmq), ne, nc, 1, ne, arity, pc._stream()))
    e2.record(st)
    torch.cuda.synchronize()
    t_ms, m_ms = e0.elapsed_time(e1), e1.elapsed_time(e2)
    root = pc.to_host(nodes[-4:])
    rate = width - 4
    perms = ne * ((nc + rate - 1) // rate)
    # transform: algorithmic bytes = read q once, write cmQ once
    q_bytes = 8 * ne * qdim + 8 * ne * nc
    out = {
        "metric": METRIC, "value": round(ne * nc / dt / 1e9, 4), "unit": "GElem/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64 (Goldilocks field, exact integer)",
        "data": "synthetic: seeded uniform Goldilocks elements generated in HBM",
        "config": {"workload": "compute_q: q 2^%d x %d (cubic) -> %d pieces of 2^%d extended to 2^%d, Poseidon2 W%d "
                               "arity %d Merkle commit" % (nbe, qdim, qdeg, nb, nbe, width, arity),
                   "n_bits": nb, "n_bits_ext": nbe, "q_deg": qdeg, "q_dim": qdim, "arity": arity,
                   "seed": cfg["seed"]},
        "phases_ms": {"transform": round(t_ms, 4), "merkelize": round(m_ms, 4)},
        "roofline": _leaf_roofline(width, perms, m_ms, 8 * ne * nc + 32 * ne,
                                   _profile_traffic("k_p2_leaves<%d>" % width, "compute_q"), 1),
        "roofline_transform": {"bound": "hbm", "achieved": q_bytes / (t_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": q_bytes / (t_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                               "traffic": None, "avg_ms": t_ms},
        "root": ["%016x" % v for v in root],
        "device": torch.cuda.get_device_name(0), "host": platform.node(),
    }
    out["roofline"]["note"] = "merkelize phase (leaves + levels) timed as one"
    if not args.no_cpu_baseline:
        from oracle.oracle import Oracle, RefLib
        o = Oracle()
        try:
            lib, kind, threads = RefLib(), "reference", None
            threads = lib.num_threads()
        except (OSError, FileNotFoundError):
            lib, kind, threads = o, "port", o.num_threads()
        def cpu_run(sb, sbe):
            sn, sne = 1 << sb, 1 << sbe
            hq = datagen.uniform(sne * qdim, cfg["seed"]).reshape(sne, qdim)
            t0 = time.perf_counter()
            cm = lib.compute_q(hq, sn, qdeg)
            if kind == "reference":
                lib.merkletree(cm, arity, o.tree_num_elements(sne, arity))
            else:
                o.merkletree(cm, arity)
            return time.perf_counter() - t0

        # a quarter of the rows first; the whole commit (repeated up to ~cpu_seconds/2) when that was quick
        sb, sbe = nb - 2, nbe - 2
        cdt, reps = cpu_run(sb, sbe), 1
        if cdt < args.cpu_seconds / 5:
            sb, sbe = nb, nbe
            cdt, reps = 0.0, 0
            while reps == 0 or (cdt < args.cpu_seconds / 2 and reps < 100):
                cdt += cpu_run(sb, sbe)
                reps += 1
        sne = 1 << sbe
        out["cpu_baseline"] = {"value": reps * sne * nc / cdt / 1e9, "unit": "GElem/s", "cores": threads,
                               "kind": kind,
                               "sample": "computeQ 2^%d x %d -> %d x 2^%d extended to 2^%d + W%d Merkle; %d run(s), "
                                         "%.2f s" % (sbe, qdim, qdeg, sb, sbe, width, reps, cdt), **host_info()}
    print(json.dumps(out), flush=True)


def parse_plan(text):
    """'S,tile,radix,coop' -> Pil2CommitPlan fields (0 = default / tuned branch count)."""
    f = [int(x) for x in text.split(",")]
    if len(f) != 4:
        raise SystemExit("--plan takes S,tile,radix,coop")
    return dict(n_streams=f[0], ntt_tile_log=f[1], ntt_radix_log=f[2], coop_max_parents=f[3])


def run_plain_plan(args, cfg):
    """The plain commit under an explicit launch plan: one job through
    pil2hip_commit_batch_plan (a one-job graph), timed like the plain path."""
    import torch
    import pil2_commit as pc

    nb, nbe, cols, arity = cfg["n_bits"], cfg["n_bits"] + cfg["blowup_bits"], cfg["cols"], cfg["arity"]
    ne = 1 << nbe
    pc.set_hash_family(2)
    dev = torch.device("cuda:0")
    trace = pc.empty((1 << nb) * cols, dev)
    pc.fill_uniform(trace, cfg["seed"])
    plan = dict(parse_plan(args.plan), n_streams=1)
    b = pc.CommitBatch([{"trace": trace, "n_bits": nb, "n_bits_ext": nbe, "n_cols": cols}], arity=arity)
    b.launch(plan=plan)
    for _ in range(args.warmup):
        b.launch(plan=plan)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch(plan=plan)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    out = {"metric": METRIC, "value": round(ne * cols / dt / 1e9, 4), "unit": "GElem/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u64 (Goldilocks field, exact integer)",
           "data": "synthetic: seeded uniform Goldilocks elements generated in HBM",
           "config": {"workload": "%s under launch plan %s (one-job graph)" % (args.config, b.plan),
                      "plan": b.plan, "n_bits": nb, "n_bits_ext": nbe, "n_cols": cols, "arity": arity},
           "root": ["%016x" % v for v in b.roots_host()[0]],
           "device": torch.cuda.get_device_name(0), "host": platform.node()}
    print(json.dumps(out), flush=True)


def _plain_from_host(args, pc, trace, lde, nodes, scratch, root, nb, nbe, cols, arity):
    """The plain commit with the trace in pinned host memory: copy-then-commit
    and the streamed pil2hip_commit_host, each step ending with the root on the
    host (PCIe-inclusive figures; `value` is the HBM-resident rate)."""
    import torch
    n, ne, dev = 1 << nb, 1 << nbe, trace.device
    host_trace = torch.empty(n * cols, dtype=trace.dtype, pin_memory=True)
    host_trace.copy_(trace)
    torch.cuda.synchronize()
    h2d_steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(h2d_steps):
        trace.copy_(host_trace, non_blocking=True)
        pc.commit(trace, arity, nb, nbe, cols, lde=lde, nodes=nodes, scratch=scratch)
        root_h2d = pc.MerkleTreeGL(arity, 0, True, ne, cols, nodes=nodes).get_root()
    h2d_dt = (time.perf_counter() - t0) / h2d_steps
    assert np.array_equal(root_h2d, root), "commit from host memory gave another root"
    t0 = time.perf_counter()
    trace.copy_(host_trace, non_blocking=True)
    torch.cuda.synchronize()
    copy_dt = time.perf_counter() - t0
    # the same from host memory through pil2hip_commit_host: column groups copied
    # on the library's copy stream while the previous group is extended and hashed
    stage = pc.empty(pc.lib.pil2hip_commit_host_stage_words(nb, cols, arity, 0), dev)
    pc.commit_host(host_trace, arity, nb, nbe, cols, lde=lde, nodes=nodes, stage=stage)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):  # as many steps as the HBM-resident timing
        pc.commit_host(host_trace, arity, nb, nbe, cols, lde=lde, nodes=nodes, stage=stage)
        root_s = pc.MerkleTreeGL(arity, 0, True, ne, cols, nodes=nodes).get_root()
    streamed_dt = (time.perf_counter() - t0) / args.steps
    assert np.array_equal(root_s, root), "streamed commit from host memory gave another root"
    del host_trace, stage
    with_h2d = {"value": round(ne * cols / h2d_dt / 1e9, 4), "unit": "GElem/s", "ms_per_step": round(h2d_dt * 1e3, 3),
                "steps": h2d_steps, "h2d_ms": round(copy_dt * 1e3, 3),
                "h2d_GBs": round(8 * n * cols / copy_dt / 1e9, 2),
                "note": "trace in pinned host memory: H2D copy + commit + root read per step",
                "streamed": {"value": round(ne * cols / streamed_dt / 1e9, 4), "ms_per_step": round(streamed_dt * 1e3, 3),
                             "steps": args.steps,
                             "note": "pil2hip_commit_host: from the first host-to-device copy to the root on the "
                                     "host; column groups copied on a copy stream, overlapped with the previous "
                                     "group's LDE and leaf absorb"}}
    return with_h2d


def _profile_kernel(kernel, config):
    """Per-launch counters of `kernel` from the committed rocprofv3 summary, or {}."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % config)
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return {}


def _profile_traffic(kernel, config):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), or None."""
    return _profile_kernel(kernel, config).get("hbm_bytes_per_launch")


def _valu_issue(kernel, config, ms_per_launch):
    """Share of the VALU issue slots (1024 SIMDs at the 2.4 GHz peak clock, one wave
    instruction per 4 cycles, a dual-issued pair per 4 cycles) that the kernel's
    instructions fill, from the committed SQ_INSTS_VALU / SQ_ACTIVE_INST_VALU2
    counters per launch and the live launch time; None without a profile."""
    c = _profile_kernel(kernel, config)
    if "valu_wave_instr_per_launch" not in c or not ms_per_launch:
        return None
    cycles = 4.0 * (c["valu_wave_instr_per_launch"] - c["valu_dual_issue_quad_cycles_per_launch"]) / N_SIMD
    return cycles / (CLOCK_HZ * ms_per_launch / 1e3)


def copy_calibration(nbytes):
    """Plain device-to-device copy of a known size, for scaling FETCH_SIZE/WRITE_SIZE."""
    import torch
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 8, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration) and exit")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="plain_large", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-h2d", action="store_true",
                    help="plain configs: skip the passes from host memory (the with_h2d object)")
    ap.add_argument("--profile-region", action="store_true",
                    help="mark the timed steps in a profile with one-block fill launches before and after")
    ap.add_argument("--plain-only", action="store_true",
                    help="plain_large: leave out the embedded batched and serving runs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate duration of the bounded CPU-baseline sample")
    ap.add_argument("--batch-streams", type=int, default=0,
                    help="graph branches for the batched config (0 = autotune)")
    ap.add_argument("--jobs", type=int, default=0, help="serving: number of jobs (default 2000)")
    ap.add_argument("--serve-streams", type=int, default=4, help="serving: device streams")
    ap.add_argument("--serve-inflight", type=int, default=32, help="serving: admitted jobs in flight")
    ap.add_argument("--plan", default="", metavar="S,TILE,RADIX,COOP",
                    help="explicit launch plan (plain configs: run as a one-job graph; batched: the batch plan)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU only; runs on several GPUs are not part of this benchmark "
              "(batched and served jobs shard per job, one large commit per column group)", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import torch
    import pil2_commit as pc

    cfg = CONFIGS[args.config]
    if args.config == "batched":
        run_batched(args, cfg)
        return
    if args.config == "serving":
        run_serving(args, cfg)
        return
    if args.config == "fri":
        run_fri(args, cfg)
        return
    if args.config == "compute_q":
        run_compute_q(args, cfg)
        return
    if args.plan:
        run_plain_plan(args, cfg)
        return
    out = run_plain(args, cfg)
    if args.config == "plain_large" and not args.plain_only:
        # the rest of the metric ("batched vs plain; p50/p99 latency") in the same line
        torch.cuda.empty_cache()
        b = run_batched(args, CONFIGS["batched"], embedded=True)
        out["batched"] = b
        out["batched_over_plain"] = b["batched_over_plain"]
        print("bench: batched %.3f GElem/s (%.2fx the per-instance path)" % (b["value"], b["batched_over_plain"]),
              file=sys.stderr, flush=True)
        sv = run_serving(args, CONFIGS["serving"], embedded=True)
        out["serving"] = sv
        out["latency_ms"] = sv["latency_ms"]
        print("bench: serving p50 %.1f ms p99 %.1f ms" % (sv["latency_ms"]["p50"], sv["latency_ms"]["p99"]),
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    if out.get("batched", {}).get("error") or out.get("serving", {}).get("failed_jobs"):
        sys.exit(1)


def run_plain(args, cfg):
    """The plain commit (one instance per call) of `cfg`, HBM-resident inputs; the
    same commit from pinned host memory; rooflines; the CPU baseline. Returns the
    benchmark line as a dict, with every device buffer released."""
    import ctypes
    import torch
    import pil2_commit as pc

    nb, nbe, cols, arity = cfg["n_bits"], cfg["n_bits"] + cfg["blowup_bits"], cfg["cols"], cfg["arity"]
    n, ne, width = 1 << nb, 1 << nbe, 4 * arity
    pc.set_hash_family(2)
    dev = torch.device("cuda:0")
    trace = pc.empty(n * cols, dev)
    pc.fill_uniform(trace, cfg["seed"])
    lde = pc.empty(ne * cols, dev)
    nodes = pc.empty(pc.get_tree_num_elements(ne, arity), dev)
    scratch = pc.empty(n * cols, dev)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        pc.commit(trace, arity, nb, nbe, cols, lde=lde, nodes=nodes, scratch=scratch)
    torch.cuda.synchronize()

    # timed region: K commits; phase events on the commit stream
    pc.lib.pil2hip_timing_read(None, True)
    pc.lib.pil2hip_timing_enable(True)
    profile_region(args, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pc.commit(trace, arity, nb, nbe, cols, lde=lde, nodes=nodes, scratch=scratch)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    profile_region(args, False)
    pc.lib.pil2hip_timing_enable(False)
    ph = (ctypes.c_double * 6)()
    ncalls = pc.lib.pil2hip_timing_read(ph, True)
    phase_ms = [ph[i] / max(ncalls, 1) for i in range(6)]
    root = pc.MerkleTreeGL(arity, 0, True, ne, cols, nodes=nodes).get_root()

    # The same commit with the trace in pinned host memory: each step copies it
    # to HBM, commits and reads the root back (PCIe-inclusive; `value` is not).
    with_h2d = None
    if not args.no_h2d:
        with_h2d = _plain_from_host(args, pc, trace, lde, nodes, scratch, root, nb, nbe, cols, arity)

    ms_per_step = dt * 1e3 / args.steps
    value = ne * cols / (dt / args.steps) / 1e9

    # dominant kernel: Poseidon2 leaf hashing (k_p2_leaves<W>)
    rate = width - 4
    perms = ne * ((cols + rate - 1) // rate)
    leaf_bytes = 8 * ne * cols + 32 * ne  # reads every LDE word once, writes one digest per row
    roofline = _with_issue(_leaf_roofline(width, perms, phase_ms[4], leaf_bytes,
                                          _profile_traffic("k_p2_leaves<%d>" % width, args.config), 1),
                           args.config)
    # LDE passes: algorithmic bytes 8*C*(N + NExt) over the four LDE kernels
    lde_ms = sum(phase_ms[0:4])
    lde_bytes = 8 * cols * (n + ne)
    pass_bytes = 16 * cols * n * 2 + 8 * cols * (n + ne) + 16 * cols * ne  # each pass reads+writes its array
    lde_traffic = None
    tr_k = [_profile_traffic(k, args.config) for k in _lde_kernels(nb, nbe, cols)]
    if all(t is not None for t in tr_k):
        lde_traffic = sum(tr_k)
    roofline_lde = {
        "bound": "hbm", "kernels": "transpose + strided iNTT + fused block + strided NTT",
        "achieved": lde_bytes / (lde_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": lde_bytes / (lde_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "per_pass_achieved": pass_bytes / (lde_ms / 1e3) / 1e9,
        "per_pass_frac": pass_bytes / (lde_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "traffic": lde_traffic, "avg_ms": lde_ms,
        "limited_by": "valu: the field arithmetic's instruction count, not HBM (the 'valu' object)",
        "valu": _lde_valu_roofline(nb, nbe, cols, lde_ms, args.config),
    }
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "GElem/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64 (Goldilocks field, exact integer)",
        "data": "synthetic: seeded uniform Goldilocks elements generated in HBM",
        "config": {"workload": "%s: trace 2^%d x %d -> LDE 2^%d (blowup %d), Poseidon2 W%d arity %d Merkle commit"
                               % (args.config, nb, cols, nbe, 1 << cfg["blowup_bits"], width, arity),
                   "n_bits": nb, "n_bits_ext": nbe, "n_cols": cols, "arity": arity, "seed": cfg["seed"],
                   "path": "plain (one instance per call)"},
        "phases_ms": dict(zip(["transpose", "intt_strided", "lde_block", "ntt_strided", "leaves", "tree_levels"],
                              [round(x, 4) for x in phase_ms])),
        "roofline": roofline, "roofline_lde": roofline_lde,
        "root": ["%016x" % v for v in root],
        "device": torch.cuda.get_device_name(0), "host": platform.node(),
    }
    if with_h2d is not None:
        out["with_h2d"] = with_h2d
        out["value_from_host"] = with_h2d["streamed"]["value"]
    fx = _reference_fixture()
    if fx is not None and args.config == "plain_large" and fx["plain_large"]["seed"] == cfg["seed"]:
        out["root_matches_reference"] = [int(v) for v in root] == [int(x, 16) for x in fx["plain_large"]["root"]]
    del trace, lde, nodes, scratch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    return out


if __name__ == "__main__":
    main()
