"""Benchmark: Groth16/BN254 proofs/s and p50/p99 latency at 2^20 constraints,
batched vs plain, on one MI355X.

Sections (all inputs resident in HBM before any timed region; every proof has
fresh blinding r, s; batched/served proofs are compared bit for bit with plain
proofs of the same inputs):

  headline (2^20): the seeded synthetic R1CS with m = 2^20 constraints and 4096
      public inputs (3 terms per A/B row, C_i = aux_i; domain n = 2^21),
      trapdoor CRS resident in HBM with the window tables of its bases
      (--tables on|off|auto).
        plain   : one proof at a time (zk_prove_device, stream-synchronised)
        batched : the same proofs with several in flight, as streams
                  (zk_prove_batch_device); the captured-graph variant
                  (zk_batch_graph_*) is tried and reported as not run while
                  the library keeps it switched off. `value` is the faster
                  of the batched variants that ran.
      Kernel probes (alone on the device): the G1 and G2 witness MSMs over the
      resident bases (roofline of the bucket-accumulation kernel), NTTs of
      2^22 (forward, inverse, coset) and a G1 MSM of 2^22 points.
  mixed: 64 proofs, m log-uniform in [2^14, 2^20], 1024 public inputs; half
      over 4 shared circuits (shared CRS), half over 32 circuits of their own;
      windows from the autotuner; plain vs streams (vs one captured graph,
      when the library enables it).
  served: Poisson arrivals against 4 resident circuits (m = 2^18, 2^20, 2^22,
      2^24, 1024 public inputs), continuous batching (zk_server_*), at three
      arrival rates set from the measured plain capacity; proofs/s, p50/p99.

Prints one JSON line.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_MAC_TS = 256 * 64 * 2.4e9 / 1e12  # VALU lane-ops: 256 CUs x 64 lanes/clk x 2.4 GHz; v_mad_u64_u32 issues at full rate
PEAK_HBM_GBS = 8000.0
MACS_PER_FPMUL = 136                    # 8x8 products + 8x8 reduction + 8 quotient words (32-bit Montgomery)
MADD_FPMUL = 10                         # XYZZ mixed add (madd-2008-s): 8M + 2S
MADD_FPMUL_G2 = 30                      # the same in Fp2: 10 Fp2 products x 3 Fp products (Karatsuba)
CALIB_BYTES = 384 << 20                 # device copy before the probes (profiler counter calibration)
# HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE runs of the probes
# (`bench.py --sections none --no-cpu --no-tune --probe-reps 1`), scaled by the
# calibration copy; profiles/probe_kernels_hbm_counters.md
# G1 / G2 accumulation over the window tables of the 2^20 circuit (7 windows
# of 18-19 bits over the GLV halves sharing one bucket set), per entry read:
# of the sort's 14,737,409 entries the A' probe reads the 11,054,444 of wires
# with a finite A' base, the B2' probe the 11,052,135 with finite B' bases

HBM_BYTES_PER_ENTRY_G1 = (1.9644e9 + 0.1069e9) / 11054444
HBM_BYTES_PER_ENTRY_G2 = (1.8688e9 + 0.0671e9) / 11052135
ALGO_BYTES_PER_ENTRY_G1 = 68            # y and x (or beta*x) of the base, 32 B each packed, + the 4 B entry
ALGO_BYTES_PER_ENTRY_G2 = 132           # the same over Fp2, + the 4 B entry
# NTT of 2^22, per transform (three pass pairs + the last pass)
HBM_BYTES_NTT_2P22 = {"forward": 1.2250e9, "inverse": 1.3579e9, "coset": 1.3621e9}
# SpMV of the 2^20 circuit (k_spmv_abc)
HBM_BYTES_SPMV_2P20 = 1.3459e9 + 0.2017e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def f_g1(n):
    return min(math.ceil(254 / c) * (11 * n + 16 * 2 ** c) for c in range(11, 19))


def f_g2(n):
    return min(math.ceil(254 / c) * (29 * n + 43 * 2 ** c) for c in range(11, 19))


def proof_work(m, n_pub):
    """Algorithmic Fp-mul equivalents of one proof's MSMs (A, B1, C, H in G1, B2 in G2)."""
    nv = 1 + n_pub + m
    n = 1 << math.ceil(math.log2(m + n_pub + 1))
    return 2 * f_g1(nv + 1) + f_g1(m) + f_g1(n) + f_g2(nv + 1)


def pct(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def r3(x):
    return None if x is None else round(float(x), 3)


def cpu_info():
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model,
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_baseline(m_cpu, n_pub_cpu, m, n_pub, threads):
    """Oracle (GMP) proves on the host: the small circuit on one thread and on
    all threads, then the metric's own circuit on all threads (measured, not
    scaled). CRS setup (parallel) is untimed."""
    import oracle as O

    info = cpu_info()
    all_thr = O.set_threads(threads if threads > 0 else (info["affinity"] or info["nproc"] or 1))
    t0 = time.time()
    c = O.Circuit(99, m_cpu, n_pub_cpu, with_crs=True)
    setup_small = time.time() - t0
    r, s = c.blinding()
    O.set_threads(1)
    t0 = time.perf_counter()
    c.prove(r, s)
    t_small_1 = time.perf_counter() - t0
    O.set_threads(all_thr)
    t0 = time.perf_counter()
    c.prove(r, s)
    t_small_all = time.perf_counter() - t0
    del c
    log(f"[bench] cpu: m=2^{int(math.log2(m_cpu))}: 1 thread {t_small_1:.2f} s, {all_thr} threads {t_small_all:.2f} s")
    t0 = time.time()
    c = O.Circuit(98, m, n_pub, with_crs=True)
    setup_big = time.time() - t0
    r, s = c.blinding()
    t0 = time.perf_counter()
    c.prove(r, s)
    t_big_all = time.perf_counter() - t0
    del c
    log(f"[bench] cpu: m=2^{int(math.log2(m))}: {all_thr} threads {t_big_all:.2f} s (setup {setup_big:.1f} s)")
    ratio = proof_work(m_cpu, n_pub_cpu) / proof_work(m, n_pub)
    return {
        "value": round(1.0 / t_big_all, 6),
        "unit": "proofs/s",
        "cores": all_thr,
        "kind": "port",
        "sample": (f"one oracle prove (GMP, OpenMP over MSM windows, NTT butterflies, SpMV rows) of the "
                   f"m=2^{int(math.log2(m))}, {n_pub}-input synthetic circuit on {all_thr} threads: "
                   f"{t_big_all:.2f} s (CRS setup {setup_big:.1f} s untimed)"),
        "host": info,
        "single_thread": {"m": m_cpu, "n_public": n_pub_cpu, "seconds": round(t_small_1, 3),
                          "proofs_per_s": round(1.0 / t_small_1, 6),
                          "scaled_to_metric_proofs_per_s": round(ratio / t_small_1, 6),
                          "scale_note": f"MSM work ratio {ratio:.5f} (F_G1/F_G2 formulas)"},
        "all_threads_small": {"m": m_cpu, "seconds": round(t_small_all, 3), "threads": all_thr},
        "setup_s": {"small": round(setup_small, 1), "metric": round(setup_big, 1)},
    }


def upload_all(ctx, arrays):
    out = []
    for a in arrays:
        p = ctx.dev_alloc(a.nbytes)
        ctx.upload(p, a)
        out.append(p)
    return out


# ---------------------------------------------------------------- headline section
def tuned_table(ctx):
    """The context's tuning table as rows (kind, ceil log2 n, c, G1 split, G2 split), via zk_tune_save."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "windows.txt")
        ctx.tune_save(path)
        rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    kinds = {"0": "witness", "1": "h", "2": "witness (window tables)", "3": "h (window tables)"}
    return [{"kind": kinds.get(r[1], r[1]), "log2_n": int(r[2]), "c": int(r[3]),
             "split_g1": int(r[4]), "split_g2": int(r[5])} for r in rows]


def setup_headline(args, zk, ctx):
    m, n_pub = 1 << args.log_m, args.n_pub
    t0 = time.time()
    circ = ctx.synth_circuit(args.seed, m, n_pub)
    log(f"[bench] circuit m=2^{args.log_m} n_pub={n_pub} n_vars={circ.n_vars} domain=2^{circ.log_n} "
        f"resident {circ.device_bytes / 2**30:.2f} GiB, setup {time.time() - t0:.1f} s")
    tuned = None
    if not args.no_tune:
        t0 = time.time()
        tuned = ctx.autotune(circ)
        log(f"[bench] autotuned windows (witness, h) = {tuned} in {time.time() - t0:.1f} s")
    return circ, tuned


def section_headline(args, zk, ctx, circ, tuned):
    K, W = args.steps, args.warmup
    # satisfying witnesses (distinct public inputs) resident in HBM, cycled over
    # the proofs; every proof gets fresh blinding r, s
    total = W + K
    t0 = time.time()
    d_uniq = upload_all(ctx, [circ.witness(args.seed * 1000 + j + 1) for j in range(args.witnesses)])
    d_w = [d_uniq[i % len(d_uniq)] for i in range(total)]
    bl = [zk.synth_blinding(args.seed * 1000 + i + 1) for i in range(total)]
    R = np.stack([b[0] for b in bl])
    S = np.stack([b[1] for b in bl])
    log(f"[bench] {args.witnesses} witnesses generated and uploaded in {time.time() - t0:.1f} s")

    # ---- plain: one proof at a time, profiled ----
    ctx.set_profiling(True)
    for i in range(W):
        ctx.prove_device(circ, d_w[i], R[i], S[i])
    lat, stats, plain_proofs = [], [], []
    t_start = time.perf_counter()
    for i in range(W, total):
        t0 = time.perf_counter()
        pr = ctx.prove_device(circ, d_w[i], R[i], S[i])
        lat.append((time.perf_counter() - t0) * 1e3)
        stats.append(ctx.prove_stats())
        plain_proofs.append(pr.words())
    plain_ps = K / (time.perf_counter() - t_start)
    ctx.set_profiling(False)
    log(f"[bench] plain: {plain_ps:.2f} proofs/s, p50 {pct(lat, 50):.1f} ms, p99 {pct(lat, 99):.1f} ms")

    # ---- batched as streams: all K proofs through the slot pipeline ----
    ctx.prove_batch_device([circ] * min(W, K), d_w[:min(W, K)], R[:min(W, K)], S[:min(W, K)])
    t0 = time.perf_counter()
    out = ctx.prove_batch_device([circ] * K, d_w[W:], R[W:], S[W:])
    streams_ps = K / (time.perf_counter() - t0)
    same_streams = all(np.array_equal(out[j].words(), plain_proofs[j]) for j in range(K))
    log(f"[bench] batched (streams, {args.slots} slots): {streams_ps:.2f} proofs/s; identical to plain: {same_streams}")

    # ---- batched as one captured graph, replayed (when the library enables it) ----
    graph_ps, same_graph, n_nodes, build_s, graph_note = 0.0, True, 0, 0.0, None
    try:
        t0 = time.perf_counter()
        g = ctx.batch_graph([circ] * K, d_w[W:], slots=args.slots)
        build_s = time.perf_counter() - t0
        g.run(R[:K][::-1], S[:K][::-1])                  # warm replay, other blinding
        t0 = time.perf_counter()
        out = g.run(R[W:], S[W:])
        graph_ps = K / (time.perf_counter() - t0)
        same_graph = all(np.array_equal(out[j].words(), plain_proofs[j]) for j in range(K))
        n_nodes = g.n_nodes
        g.close()
        log(f"[bench] batched (graph, {n_nodes} nodes, built in {build_s:.2f} s): {graph_ps:.2f} proofs/s; "
            f"identical to plain: {same_graph}")
    except zk.ZkError as e:
        graph_note = str(e)
        log(f"[bench] batched (graph): not run: {graph_note}")

    mean = {k: float(np.mean([s[k] for s in stats])) for k in stats[0]}
    best_mode = "graph" if graph_ps >= streams_ps else "streams"
    batch_ps = max(graph_ps, streams_ps)
    for p in d_uniq:
        ctx.dev_free(p)
    return {
        "c_w": int(mean["c_w"]), "c_h": int(mean["c_h"]),
        "plain": {"proofs_per_s": round(plain_ps, 4), "p50_ms": r3(pct(lat, 50)), "p99_ms": r3(pct(lat, 99))},
        "batched": {"proofs_per_s": round(batch_ps, 4), "mode": best_mode,
                    "streams_proofs_per_s": round(streams_ps, 4), "graph_proofs_per_s": round(graph_ps, 4) if graph_ps else None,
                    "identical_to_plain": bool(same_streams and same_graph), "slots": args.slots,
                    "graph_nodes": int(n_nodes), "graph_build_s": round(build_s, 3), "graph_note": graph_note,
                    "latency_note": "all proofs submitted at t=0; per-proof latency = batch position x period"},
        "speedup_batched_vs_plain": round(batch_ps / plain_ps, 4),
        "phases_ms": {k: round(v, 4) for k, v in mean.items() if k.endswith("_ms")},
        "tuned_windows": tuned,
        "tuned_table": tuned_table(ctx),
    }


# ---------------------------------------------------------------- kernel probes
def probe_msm(args, zk, ctx, circ, which, c_w, g2):
    """Witness MSM over resident bases through the arrays a proof reads (the
    window tables when the circuit has them), alone on the device:
    (total, sort, accumulation ms, entries, reduction ms)."""
    d_bases, n_ext = circ.bases(which)
    w0 = circ.witness(args.seed * 1000 + 1)
    ext = np.ascontiguousarray(np.concatenate([w0, zk.synth_uniform(args.seed + 7, 4)]), dtype=np.uint64)
    assert ext.shape[0] == n_ext
    d_sc = ctx.dev_alloc(ext.nbytes)
    ctx.upload(d_sc, ext)
    rows = []
    for i in range(2 + args.probe_reps):
        ctx.circuit_msm_device(circ, which, d_sc, c_w)
        if i >= 2:
            rows.append(ctx.last_timing())
    ctx.dev_free(d_sc)
    return n_ext, np.array(rows).mean(axis=0)


def probe_ntt(args, zk, ctx, log_n):
    n = 1 << log_n
    x = zk.synth_uniform(args.seed + 11, n)
    a, b = ctx.dev_alloc(x.nbytes), ctx.dev_alloc(x.nbytes)
    res = {}
    for kind, name in ((0, "forward"), (1, "inverse"), (2, "coset")):
        ts = []
        for i in range(2 + args.probe_reps):
            ctx.upload(a, x)
            ctx.ntt_device([a], [b], log_n, kind)
            if i >= 2:
                ts.append(ctx.last_timing()[0])
        res[name] = float(np.mean(ts))
    ctx.dev_free(a)
    ctx.dev_free(b)
    return res


def section_probes(args, zk, ctx, circ):
    """Kernels alone on the device, before any proof. A device copy of a known
    size first gives the HBM-counter calibration point for profiler runs."""
    tabs = ctx.circuit_window_tables(circ)
    c_w = tabs[0] or ctx.tuned_window(0, circ.n_vars + 4)
    nb = CALIB_BYTES
    a, b = ctx.dev_alloc(nb), ctx.dev_alloc(nb)
    ctx.dev_copy(b, a, nb)
    ctx.dev_free(a)
    ctx.dev_free(b)
    n_ext, t1 = probe_msm(args, zk, ctx, circ, 0, c_w, False)
    _, t2 = probe_msm(args, zk, ctx, circ, 2, c_w, True)
    t_launch = t1[2] / 1e3
    units = float(t1[3])
    achieved = units * MADD_FPMUL * MACS_PER_FPMUL / t_launch / 1e12
    g2_achieved = float(t2[3]) * MADD_FPMUL_G2 * MACS_PER_FPMUL / (t2[2] / 1e3) / 1e12
    log(f"[bench] isolated G1 witness MSM n={n_ext} c={c_w}: total {t1[0]:.3f} ms, sort {t1[1]:.3f} ms, "
        f"accumulation {t1[2]:.3f} ms over {int(units)} entries -> {achieved:.2f} TMAC/s, reduction {t1[4]:.3f} ms")
    log(f"[bench] isolated G2 witness MSM: total {t2[0]:.3f} ms, accumulation {t2[2]:.3f} ms -> {g2_achieved:.2f} TMAC/s, "
        f"reduction {t2[4]:.3f} ms")
    roofline = {
        "bound": "valu", "kernel": "k_msm_accum29 (G1 bucket accumulation, 9 x 29-bit limbs)",
        "achieved": round(achieved, 3), "peak": round(PEAK_MAC_TS, 2), "unit": "TMAC/s",
        "frac": round(achieved / PEAK_MAC_TS, 4), "traffic": int(units * HBM_BYTES_PER_ENTRY_G1),
        "traffic_note": ("HBM bytes per launch: per-entry figure from the profiler's FETCH_SIZE + WRITE_SIZE runs "
                         "of this probe (calibrated by a 384 MiB copy) x entries; algorithmic bytes 68 per entry "
                         "(64 B of packed base coordinates + 4 B entry)"),
        "algorithmic_bytes": int(units * ALGO_BYTES_PER_ENTRY_G1),
        "avg_launch_ms": round(t_launch * 1e3, 4), "units_per_launch": int(units),
        "unit_def": ("one signed-digit (point, bucket) entry = one XYZZ mixed add = 10 Fp products x 136 32x32 "
                     "MACs (the algorithmic count of a 256-bit Montgomery product, whatever the limb size)"),
        "probe": (f"G1 MSM over the resident A' bases, n={n_ext}, c={c_w}"
                  f"{' over window tables (one bucket set)' if tabs[0] else ''}, as a proof runs it (the "
                  f"entries of the wires whose A' base is finite), alone on the device, "
                  f"{args.probe_reps} launches, HIP events on the MSM stream"),
        "msm_total_ms": r3(t1[0]), "sort_ms": r3(t1[1]), "reduction_ms": r3(t1[4]),
        "g2": {"kernel": "k_msm_accum29q (G2, 29-bit limbs, accumulator in LDS)", "achieved": round(g2_achieved, 3),
               "frac": round(g2_achieved / PEAK_MAC_TS, 4), "avg_launch_ms": r3(t2[2]),
               "units_per_launch": int(t2[3]),
               "traffic": int(float(t2[3]) * HBM_BYTES_PER_ENTRY_G2),
               "algorithmic_bytes": int(float(t2[3]) * ALGO_BYTES_PER_ENTRY_G2),
               "msm_total_ms": r3(t2[0]), "reduction_ms": r3(t2[4]), "unit_def": "entry = Fp2 XYZZ mixed add = 30 Fp-mul x 136 MACs"},
    }
    ntt = probe_ntt(args, zk, ctx, 22)
    n = 1 << 22
    gbs = {k: 64 * n / (v / 1e3) / 1e9 for k, v in ntt.items()}
    log(f"[bench] NTT 2^22 ms {ntt} -> GB/s {({k: round(v) for k, v in gbs.items()})}")
    roofline_ntt = {
        "bound": "hbm", "kernel": "NTT 2^22 (three k_ntt_pass2 launches + k_ntt_final<4>), forward",
        "achieved": round(gbs["forward"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": round(gbs["forward"] / PEAK_HBM_GBS, 4), "traffic": int(HBM_BYTES_NTT_2P22["forward"]),
        "traffic_by_kind": {k: int(v) for k, v in HBM_BYTES_NTT_2P22.items()},
        "traffic_GBps": {k: round(HBM_BYTES_NTT_2P22[k] / (v / 1e3) / 1e9, 1) for k, v in ntt.items()},
        "note": ("the passes are mostly instruction-bound (29-bit Fr products), the rest memory latency at 2 waves "
                 "per SIMD, not HBM bandwidth: the profiler's traffic runs at about a quarter of the HBM peak"),
        "avg_launch_ms": r3(ntt["forward"]), "algorithmic_bytes": 64 * n,
        "ms": {k: r3(v) for k, v in ntt.items()},
        "GBps": {k: round(v, 1) for k, v in gbs.items()},
    }
    msm22 = probe_msm_2p22(args, zk, ctx, circ)
    roofline_spmv = probe_spmv(args, zk, ctx, circ)
    # the VALU issue ceiling the MSM roofline is priced against, measured here
    mad = ctx.probe_valu(0) / 1e12
    add = ctx.probe_valu(1) / 1e12
    log(f"[bench] VALU probe: v_mad_u64_u32 {mad:.2f} T lane-ops/s, v_add_u32 {add:.2f} T lane-ops/s "
        f"(nominal {PEAK_MAC_TS:.2f})")
    roofline["peak_measured"] = {"v_mad_u64_u32": round(mad, 2), "v_add_u32": round(add, 2), "unit": "T lane-ops/s",
                                 "probe": "8 independent chains per thread, 2048 x 16 x 8 ops, 8 waves per SIMD"}
    # what one to eight waves per SIMD can issue: instruction streams, then the
    # accumulation's and the reduction's point additions as dependent chains
    waves = (1, 2, 4, 8)
    by_waves = {}
    for op, name, scale in ((0, "v_mad_u64_u32 (8 chains)", 1e12), (3, "v_mad_u64_u32 (1 chain)", 1e12),
                            (2, "v_fma_f64 (8 chains)", 1e12), (1, "v_add_u32", 1e12),
                            (4, "29-bit mixed XYZZ add (accumulation)", 1e9),
                            (5, "29-bit general XYZZ add (reduction)", 1e9)):
        vals = [ctx.probe_valu(op + 16 * w) / scale for w in waves]
        by_waves[name] = [round(v, 2) for v in vals]
        unit = "T lane-ops/s" if scale == 1e12 else "G adds/s"
        log(f"[bench] issue probe {name}, 1/2/4/8 waves per SIMD: " + " / ".join(f"{v:.2f}" for v in vals) + f" {unit}")
    roofline["peak_measured"]["by_waves_per_simd"] = {
        "waves": list(waves), "rates": by_waves,
        "units": "T lane-ops/s for instructions, G point additions/s for the XYZZ chains (1024 SIMDs)"}
    return roofline, roofline_ntt, msm22, roofline_spmv


def probe_spmv(args, zk, ctx, circ):
    """The fused A*w, B*w, (A*w)o(B*w) kernel of the headline circuit alone on the
    device, against HBM. Algorithmic bytes per matrix M: 36 nnz_M (value + column)
    + 4 (rows + 1) (row pointers) + 32 rows (output) + 32 n_vars (witness), plus
    32 rows for the product output."""
    w = circ.witness(args.seed * 1000 + 1)
    d_w = ctx.dev_alloc(w.nbytes)
    ctx.upload(d_w, w)
    outs = [ctx.dev_alloc(circ.n * 32) for _ in range(3)]
    ts = []
    for i in range(2 + args.probe_reps):
        t = ctx.spmv_device(circ, d_w, *outs)
        if i >= 2:
            ts.append(t)
    for p in outs + [d_w]:
        ctx.dev_free(p)
    ms = float(np.mean(ts))
    rows = circ.n
    algo = sum(36 * nnz + 4 * (rows + 1) + 32 * rows + 32 * circ.n_vars for nnz in (circ.nnz_a, circ.nnz_b)) + 32 * rows
    gbs = algo / (ms / 1e3) / 1e9
    log(f"[bench] SpMV (A, B, AoB) alone: {ms:.3f} ms, {algo / 1e6:.1f} MB algorithmic -> {gbs:.0f} GB/s")
    return {"bound": "hbm", "kernel": "k_spmv_abc (A*w, B*w and their product, fused)", "achieved": round(gbs, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
            "traffic": int(HBM_BYTES_SPMV_2P20), "traffic_GBps": round(HBM_BYTES_SPMV_2P20 / (ms / 1e3) / 1e9, 1),
            "avg_launch_ms": r3(ms), "algorithmic_bytes": int(algo),
            "unit_def": "36 B per nonzero, 4 B per row pointer, 32 B per output row and per witness value"}


def probe_msm_2p22(args, zk, ctx, circ):
    """The G1 MSM of N = 2^22 points alone on the device: the circuit's resident
    H, A' and B1' bases end to end (distinct points, already in HBM) with uniform
    scalars. Priced by the implementation-independent Fp-mul count F_G1(N) x 136."""
    N = 1 << 22
    d_pts = ctx.dev_alloc(N * 64)
    off = 0
    for which in (4, 0, 1):
        p, cnt = circ.bases(which)
        take = min(cnt, N - off)
        ctx.dev_copy(ctypes.c_void_p(d_pts.value + off * 64), p, take * 64)
        off += take
        if off == N:
            break
    assert off == N
    sc = zk.synth_uniform(args.seed + 22, N)
    d_sc = ctx.dev_alloc(sc.nbytes)
    ctx.upload(d_sc, sc)
    rows = []
    for i in range(2 + args.probe_reps):
        ctx.msm_g1_device(d_pts, d_sc, N, 0)
        if i >= 2:
            rows.append(ctx.last_timing())
    ctx.dev_free(d_sc)
    ctx.dev_free(d_pts)
    t = np.array(rows).mean(axis=0)
    macs = f_g1(N) * MACS_PER_FPMUL
    achieved = macs / (t[0] / 1e3) / 1e12
    log(f"[bench] G1 MSM 2^22 alone: {t[0]:.3f} ms (sort {t[1]:.3f}, accumulation {t[2]:.3f}, reduction {t[4]:.3f} ms, "
        f"{int(t[3])} entries) -> {achieved:.2f} T MAC/s algorithmic")
    return {
        "n": N, "ms": r3(t[0]), "sort_ms": r3(t[1]), "accumulation_ms": r3(t[2]), "reduction_ms": r3(t[4]),
        "entries": int(t[3]),
        "msm_per_s": round(1e3 / t[0], 2), "achieved": round(achieved, 3), "peak": round(PEAK_MAC_TS, 2),
        "unit": "TMAC/s", "frac": round(achieved / PEAK_MAC_TS, 4),
        "work_def": "F_G1(N) = min_c ceil(254/c)(11N + 16*2^c) Fp-mul equivalents x 136 MACs, device time "
                    "(GLV split + sort + accumulation + reduction), host finish excluded",
        "points": "resident H, A', B1' bases of the 2^20 circuit, end to end; uniform scalars",
    }


# ---------------------------------------------------------------- mixed batch
def section_mixed(args, zk, ctx):
    rng = np.random.default_rng(args.seed + 3)
    n_pub = 1024
    lo, hi = 14, 20
    n_distinct = args.mixed_proofs // 2
    shared_m = [int(round(2 ** rng.uniform(lo, hi))) for _ in range(4)]
    distinct_m = [int(round(2 ** rng.uniform(lo, hi))) for _ in range(n_distinct)]
    t0 = time.time()
    shared = [ctx.synth_circuit(30000 + i, m, n_pub) for i, m in enumerate(shared_m)]
    distinct = [ctx.synth_circuit(40000 + j, m, n_pub) for j, m in enumerate(distinct_m)]
    setup_s = time.time() - t0
    plan = [shared[int(rng.integers(0, 4))] for _ in range(args.mixed_proofs - n_distinct)] + distinct
    plan = [plan[i] for i in rng.permutation(len(plan))]
    t0 = time.time()
    if not args.no_tune:
        for c in shared + distinct:
            ctx.autotune(c)
    tune_s = time.time() - t0
    d_ws = upload_all(ctx, [c.witness(50000 + k) for k, c in enumerate(plan)])
    N = len(plan)
    bl = [zk.synth_blinding(60000 + k) for k in range(N)]
    R = np.stack([b[0] for b in bl])
    S = np.stack([b[1] for b in bl])
    log(f"[bench] mixed: {N} proofs, n_vars in [{min(c.n_vars for c in plan)}, {max(c.n_vars for c in plan)}], "
        f"{len(shared) + len(distinct)} circuits (setup {setup_s:.1f} s, tuning {tune_s:.1f} s)")
    # plain (one untimed proof per circuit first)
    seen = set()
    for k, c in enumerate(plan):
        if id(c) not in seen:
            seen.add(id(c))
            ctx.prove_device(c, d_ws[k], R[k], S[k])
    t0 = time.perf_counter()
    plain = [ctx.prove_device(c, d_ws[k], R[k], S[k]).words() for k, c in enumerate(plan)]
    plain_s = time.perf_counter() - t0
    # one captured graph over the whole batch (when the library enables it)
    graph_s, same_graph, nodes, build_s, graph_note = None, True, 0, 0.0, None
    try:
        t0 = time.perf_counter()
        g = ctx.batch_graph(plan, d_ws, slots=args.mixed_slots)
        build_s = time.perf_counter() - t0
        g.run(R[::-1], S[::-1])
        t0 = time.perf_counter()
        out = g.run(R, S)
        graph_s = time.perf_counter() - t0
        same_graph = all(np.array_equal(out[k].words(), plain[k]) for k in range(N))
        nodes = g.n_nodes
        g.close()
    except zk.ZkError as e:
        graph_note = str(e)
    # streams
    ctx.set_batch_slots(args.mixed_slots)
    ctx.prove_batch_device(plan, d_ws, R[::-1], S[::-1])     # warm pass: slots sized for every circuit
    t0 = time.perf_counter()
    out = ctx.prove_batch_device(plan, d_ws, R, S)
    streams_s = time.perf_counter() - t0
    ctx.set_batch_slots(args.slots)
    same_streams = all(np.array_equal(out[k].words(), plain[k]) for k in range(N))
    gtxt = f"graph {N / graph_s:.2f} ({nodes} nodes, built {build_s:.2f} s)" if graph_s else "graph not run"
    log(f"[bench] mixed: plain {N / plain_s:.2f} proofs/s, {gtxt}, streams {N / streams_s:.2f}; "
        f"identical: graph {same_graph}, streams {same_streams}")
    windows = sorted({(c.n_vars + 4, ctx.tuned_window(0, c.n_vars + 4), c.n, ctx.tuned_window(1, c.n))
                      for c in shared + distinct})
    res = {
        "proofs": N, "n_public": n_pub, "m_range_log2": [lo, hi],
        "circuits": {"shared": len(shared), "shared_proofs": N - n_distinct, "distinct": n_distinct},
        "plain_proofs_per_s": round(N / plain_s, 4),
        "graph_proofs_per_s": round(N / graph_s, 4) if graph_s else None,
        "streams_proofs_per_s": round(N / streams_s, 4),
        "speedup_graph_vs_plain": round(plain_s / graph_s, 4) if graph_s else None,
        "speedup_streams_vs_plain": round(plain_s / streams_s, 4),
        "identical_to_plain": bool(same_graph and same_streams),
        "graph_nodes": int(nodes), "graph_build_s": round(build_s, 3), "graph_note": graph_note,
        "slots": args.mixed_slots,
        "windows_by_size": [{"n_witness": a, "c_w": b, "n_h": c, "c_h": d} for a, b, c, d in windows],
    }
    for p in d_ws:
        ctx.dev_free(p)
    for c in shared + distinct:
        c.free()
    return res


# ---------------------------------------------------------------- served
def section_served(args, zk, ctx):
    """Poisson arrivals against the resident circuits through zk_server_*. Every
    request carries a fresh witness (public inputs from a fresh seed, the aux
    wires computed on the device, zk_circuit_witness_device) and fresh r, s.
    Slots are sized for every circuit up front (zk_server_prepare), so no
    request allocates device memory while others run."""
    n_pub = 1024
    sizes = args.serve_log_m
    circs = {}
    for lg in sizes:
        t0 = time.time()
        circs[lg] = ctx.synth_circuit(70000 + lg, 1 << lg, n_pub)
        log(f"[bench] served: circuit 2^{lg} resident {circs[lg].device_bytes / 2**30:.2f} GiB, "
            f"setup {time.time() - t0:.1f} s")
        if not args.no_tune:
            t0 = time.time()
            tw = ctx.autotune(circs[lg])
            log(f"[bench] served: 2^{lg} windows {tw} tuned in {time.time() - t0:.1f} s")
    # per circuit, one device witness buffer for the plain service times and
    # the final checks
    ring = {lg: [ctx.dev_alloc(circs[lg].n_vars * 32)] for lg in sizes}
    gen_ms = {}
    for lg in sizes:
        t0 = time.perf_counter()
        circs[lg].witness_device(71000 + lg, ring[lg][0])
        gen_ms[lg] = (time.perf_counter() - t0) * 1e3
    log(f"[bench] served: device witness generation ms {({k: round(v, 1) for k, v in gen_ms.items()})}")
    # plain capacity: one proof at a time per circuit
    svc = {}
    for lg in sizes:
        r, s = zk.synth_blinding(80000 + lg)
        ctx.prove_device(circs[lg], ring[lg][0], r, s)
        ts = []
        for i in range(args.serve_svc_reps):
            t0 = time.perf_counter()
            ctx.prove_device(circs[lg], ring[lg][0], r, s)
            ts.append(time.perf_counter() - t0)
        svc[lg] = float(np.mean(ts))
    capacity = len(sizes) / sum(svc.values())          # requests/s, uniform circuit choice, one at a time
    log(f"[bench] served: plain service times (ms) {({k: round(v * 1e3, 1) for k, v in svc.items()})}, "
        f"capacity {capacity:.2f} req/s")
    srv = zk.Server(ctx, max_slots=args.serve_slots, reserve_bytes=args.serve_reserve_gib << 30)
    t0 = time.time()
    n_slots = srv.prepare([circs[lg] for lg in sizes])
    log(f"[bench] served: {n_slots} slots prepared for all circuits in {time.time() - t0:.1f} s")
    rng = np.random.default_rng(args.seed + 5)
    rates = []
    check = {}
    for frac in args.serve_loads:
        rate = frac * capacity
        N = args.serve_requests
        arrivals = np.cumsum(rng.exponential(1.0 / rate, N))
        which = rng.integers(0, len(sizes), N)
        picks = [(sizes[int(q)], int(rng.integers(0, 1 << 30))) for q in which]
        results = {}
        # every request's witness is made on the device before the timed loop
        # (inputs resident in HBM, a fresh witness per request), so that the
        # submitter keeps to the arrival times
        t0 = time.time()
        bufs = []
        for lg, seed in picks:
            d = ctx.dev_alloc(circs[lg].n_vars * 32)
            circs[lg].witness_device(seed, d)
            bufs.append(d)
        gen_s = time.time() - t0
        tickets, late = [], 0
        t_base = time.perf_counter()
        for k in range(N):
            t_sched = t_base + arrivals[k]
            while True:
                dt = t_sched - time.perf_counter()
                if dt <= 0:
                    break
                time.sleep(min(dt, 0.002))
            lg, seed = picks[k]
            r, s = zk.synth_blinding(seed)
            t_sub = time.perf_counter()
            t = srv.submit_device(circs[lg], bufs[k], r, s)
            lag = max(0.0, t_sub - t_sched)
            late += int(lag > 0.005)
            tickets.append((t, lg, seed, lag))
        # latency counts from the scheduled arrival: a late submission adds its lag
        lat, queued, by, lags = [], [], {lg: [] for lg in sizes}, []
        for t, lg, seed, lag in tickets:
            if t not in results:
                results[t] = srv.wait(t)
            pr, tm = results[t]
            lat.append(tm[2] + lag * 1e3)
            queued.append(tm[0])
            lags.append(lag * 1e3)
            by[lg].append(tm[2] + lag * 1e3)
            check.setdefault(lg, (pr, seed))
        wall = time.perf_counter() - t_base
        for d in bufs:
            ctx.dev_free(d)
        st = srv.stats()
        rates.append({
            "load": frac, "offered_req_per_s": round(rate, 3), "requests": N,
            "proofs_per_s": round(N / wall, 3),
            "p50_ms": r3(pct(lat, 50)), "p99_ms": r3(pct(lat, 99)), "mean_queued_ms": r3(np.mean(queued)),
            "p50_ms_by_log_m": {str(lg): r3(pct(v, 50)) for lg, v in by.items() if v},
            "p99_ms_by_log_m": {str(lg): r3(pct(v, 99)) for lg, v in by.items() if v},
            "peak_in_flight": int(st["peak_in_flight"]), "slots": int(st["slots"]),
            "late_arrivals": late, "mean_submit_lag_ms": r3(np.mean(lags)),
            "witnesses_made_s": round(gen_s, 2),
        })
        log(f"[bench] served load {frac:.2f} ({rate:.2f} req/s): {N / wall:.2f} proofs/s, p50 {pct(lat, 50):.1f} ms, "
            f"p99 {pct(lat, 99):.1f} ms, peak in flight {int(st['peak_in_flight'])}, {late} late arrivals")
    failed = int(srv.stats()["failed"])
    srv.close()
    # one served result per circuit against the plain proof of the same inputs,
    # and against the pairing check
    same = verified = True
    for lg, (pr, seed) in check.items():
        d = ring[lg][0]
        circs[lg].witness_device(seed, d)
        r, s = zk.synth_blinding(seed)
        same &= bool(np.array_equal(pr.words(), ctx.prove_device(circs[lg], d, r, s).words()))
        w_pub = ctx.download(d, circs[lg].n_vars * 32).reshape(-1, 4)[1:n_pub + 1]
        verified &= zk.verify(circs[lg], w_pub, pr)
    for lg in sizes:
        for p in ring[lg]:
            ctx.dev_free(p)
        circs[lg].free()
    return {
        "circuits_log_m": sizes, "n_public": n_pub,
        "witnesses": "fresh per request: public inputs from the request's seed, aux wires on the device",
        "witness_generation_ms": {str(lg): r3(v) for lg, v in gen_ms.items()},
        "plain_service_ms": {str(lg): r3(v * 1e3) for lg, v in svc.items()},
        "plain_capacity_req_per_s": round(capacity, 3),
        "rates": rates, "identical_to_plain": same, "pairing_verified": verified, "failed": failed,
        "max_slots": args.serve_slots, "slots_prepared": n_slots, "reserve_gib": args.serve_reserve_gib,
        "arrivals": "Poisson, circuit uniform over the sizes, fresh witness and r, s per request",
        "latency": ("from the scheduled arrival to the result: the server's submit-to-done time plus any lag of "
                    "the submission behind its arrival (every request's witness is made on the device before the "
                    "timed loop)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-m", type=int, default=20)
    ap.add_argument("--n-pub", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2020)
    ap.add_argument("--slots", type=int, default=3)
    ap.add_argument("--witnesses", type=int, default=4)
    ap.add_argument("--cpu-log-m", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: every core this process may use)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tables", choices=("auto", "on", "off"), default="auto",
                    help="window tables of the resident circuits (default: on when they fit)")
    ap.add_argument("--probe-reps", type=int, default=5)
    ap.add_argument("--sections", default="headline,mixed,served",
                    help="comma list of: headline, mixed, served (the kernel probes always run)")
    ap.add_argument("--mixed-proofs", type=int, default=64)
    ap.add_argument("--mixed-slots", type=int, default=4)
    ap.add_argument("--serve-log-m", type=lambda s: [int(x) for x in s.split(",")], default=[18, 20, 22, 24])
    ap.add_argument("--serve-loads", type=lambda s: [float(x) for x in s.split(",")], default=[0.4, 0.7, 1.0])
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-reserve-gib", type=int, default=8)
    ap.add_argument("--serve-svc-reps", type=int, default=2)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); every box here has a single "
              f"MI355X and multi-GPU sharding is not part of this build", file=sys.stderr)
        sys.exit(2)

    import zyklop_amd as zk

    sections = set(args.sections.split(","))

    ctx = zk.Context(0)
    ctx.set_batch_slots(args.slots)
    ctx.set_window_tables({"auto": -1, "on": 1, "off": 0}[args.tables])
    circ, tuned = setup_headline(args, zk, ctx)
    roofline, roofline_ntt, msm22, roofline_spmv = section_probes(args, zk, ctx, circ)
    m, n_pub, K = 1 << args.log_m, args.n_pub, args.steps
    result = {
        "metric": "Groth16 proofs/s + p50/p99 latency at 2^20 constraints, batched vs plain",
        "value": None,
        "unit": "proofs/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"groth16_bn254_prove_m2^{args.log_m}_pub{n_pub}", "constraints": m,
                   "n_public": n_pub, "n_vars": circ.n_vars, "domain": circ.n, "batched_slots": args.slots,
                   "blinding": "fresh r,s", "window_tables": list(ctx.circuit_window_tables(circ))},
    }
    if "headline" in sections:
        head = section_headline(args, zk, ctx, circ, tuned)
        result["value"] = head["batched"]["proofs_per_s"]
        result["ms_per_step"] = round(1e3 / head["batched"]["proofs_per_s"], 3)
        result["config"]["window_witness"] = head.pop("c_w")
        result["config"]["window_h"] = head.pop("c_h")
        result.update(head)
    result["roofline"] = roofline
    result["roofline_ntt"] = roofline_ntt
    result["roofline_spmv"] = roofline_spmv
    result["msm_g1_2p22"] = msm22
    circ.free()
    if "mixed" in sections:
        result["mixed_batch"] = section_mixed(args, zk, ctx)
    if "served" in sections:
        result["served"] = section_served(args, zk, ctx)
    if not args.no_cpu:
        log(f"[bench] cpu baseline: oracle proves at m=2^{args.cpu_log_m} and 2^{args.log_m} ...")
        result["cpu_baseline"] = cpu_baseline(1 << args.cpu_log_m, 1024, m, n_pub, args.cpu_threads)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result, default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)


if __name__ == "__main__":
    main()
