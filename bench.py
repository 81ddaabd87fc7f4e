#!/usr/bin/env python3
"""Benchmark of the network-likelihood hot path on one MI355X.

Metric: lnL + full gradient evaluations per second at 50 taxa, 4
reticulations (16 displayed trees), 50,000 DNA sites in two partitions of
20k / 30k sites, GTR+G4 -- one evaluation is a full (non-incremental)
forward pass over all partitions and displayed trees plus dlnL/d(branch
length) for every edge and dlnL/d(reticulation probability) for every
reticulation.  Each step moves every branch length (as an optimiser step
would), so every step recomputes all P matrices and all CLVs.

Prints one JSON line (see --help).  `--gpus N` with N > 1 exits: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "lnL+grad evals/s, 50 taxa/4 retic/50k sites, vs plain FP64 at |dlnL|<=1e-6"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_engine(w, precision, timing=True, use_graph=False):
    from netrax_amd import AnnotatedNetwork, NetraxOptions, Precision

    opts = NetraxOptions(precision=Precision(precision), timing=timing, use_graph=use_graph)
    return AnnotatedNetwork(w.network.copy(), w.partitions, opts)


SHAPE = None  # --shape G x TS: fixed FP32 kernel shape instead of autotuning


def run_form(w, precision, steps, warmup, use_graph=False, tune=False):
    """Times `steps` lnL+gradient evaluations; returns a result dict.

    Throughput (`evals_per_s`) is measured on an engine without timing
    events; kernel durations for the roofline come from a second engine of
    the same shape whose every pruning / adjoint launch is bracketed by HIP
    events on its stream (events between launches would otherwise add their
    own gaps to the throughput).  With `tune`
    (FP32 form) the kernel shape is autotuned first, outside the timed
    region (or set from --shape)."""
    ann = make_engine(w, precision, timing=False, use_graph=use_graph)
    tim = make_engine(w, precision, timing=True)
    tuning = None
    if tune and SHAPE:
        shape = SHAPE
        tuning = {"group_width": SHAPE[0], "tile_sites": SHAPE[1], "row_lanes": SHAPE[2], "fixed": True}
    elif tune:
        best, cands = tim.autotune(reps=3)
        shape = best[:3]
        tuning = {"group_width": best[0], "tile_sites": best[1], "row_lanes": best[2],
                  "candidates_ms": {f"{g}x{t}x{h}": round(ms, 4) for g, t, h, ms in cands}}
        log(f"autotune: {best} of {cands}")
    if tune:
        ann.set_tuning(*shape)
        tim.set_tuning(*shape)
    b0 = w.network.brlen.copy()
    b1 = b0 * (1.0 + 1e-7)
    for i in range(warmup):
        for e in (ann, tim):
            e.set_branch_lengths(b1 if i % 2 else b0)
            e.gradient()
    t0 = time.perf_counter()
    for i in range(steps):
        ann.set_branch_lengths(b1 if (warmup + i) % 2 else b0)
        lnl, g, gp = ann.gradient()
    wall = time.perf_counter() - t0
    fwd_ms = adj_ms = tot_ms = 0.0
    fwd_l = adj_l = 0
    for i in range(steps):
        tim.set_branch_lengths(b1 if (warmup + i) % 2 else b0)
        tim.gradient()
        t = tim.last_timing()
        fwd_ms += t.fwd_ms
        adj_ms += t.adj_ms
        tot_ms += t.total_ms
        fwd_l += t.fwd_launches
        adj_l += t.adj_launches
    st = ann.stats()
    ann.set_branch_lengths(b0)
    lnl, g, gp = ann.gradient()
    res = {
        "evals_per_s": steps / wall,
        "ms_per_eval": 1e3 * wall / steps,
        "device_ms_per_eval": tot_ms / steps,
        "fwd_ms_per_eval": fwd_ms / steps,
        "adj_ms_per_eval": adj_ms / steps,
        "fwd_launches_per_eval": fwd_l / steps,
        "adj_launches_per_eval": adj_l / steps,
        "lnl": lnl, "grad": g, "dprob": gp, "stats": st, "tuning": tuning,
    }
    ann.close()
    tim.close()
    return res


PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_traffic_50taxa_4retic_50k.md")


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by the copy-kernel calibration),
    or None when that kernel shape was not profiled."""
    try:
        text = open(PMC_SUMMARY).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        k = json.loads(block)["kernels"].get(kernel)
        return None if k is None else k["traffic_bytes"]
    except (OSError, IndexError, ValueError, KeyError):
        return None


def roofline(res, ncat=4):
    st = res["stats"]
    kern = [("forward pruning", st.fwd_op_bytes, res["fwd_ms_per_eval"], res["fwd_launches_per_eval"]),
            ("adjoint pruning", st.adj_bytes, res["adj_ms_per_eval"], res["adj_launches_per_eval"])]
    name, by, ms, launches = max(kern, key=lambda k: k[2])
    achieved = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    t = res.get("tuning")
    step = "forward" if name.startswith("forward") else "adjoint"
    if t:  # persistent form: one launch per evaluation
        split = step == "adjoint" and t.get("row_lanes", 1) == 2
        kname = (f"k_adjoint_split<{ncat}, {t['group_width']}, {t['tile_sites']}>" if split else
                 f"k_{step}_persistent<{ncat}, {t['group_width']}, {t['tile_sites']}>")
    else:  # plain form: one forward and one reverse-sweep launch walking every level
        kname = f"k_adjoint_level<4, {ncat}>" if step == "adjoint" else f"k_forward_walk<4, {ncat}, 4>"
    traffic = pmc_traffic(kname)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic, "kernel": kname,
            "algorithmic_bytes_per_eval": by, "algorithmic_bytes_per_launch": round(by / max(launches, 1)),
            "kernel_ms_per_eval": round(ms, 4), "launches_per_eval": launches,
            "traffic_source": "profiles/pmc_traffic_50taxa_4retic_50k.md (rocprofv3 --pmc FETCH_SIZE / "
                              "WRITE_SIZE, copy-kernel calibrated), bytes per launch" if traffic else None}


MFMA_FP32_PEAK_TFLOPS = 157.3
MFMA_FLOPS_32x32x2 = 32 * 32 * 2 * 2  # one v_mfma_f32_32x32x2f32


def bench_protein(steps, warmup):
    """40 taxa / 2 reticulations / 20,000 amino-acid sites, GTR-like 20-state
    model + G4: plain FP64 form vs the FP32 matrix-core form; matrix-core
    utilisation of the FP32 kernels from the instruction count the op list
    implies (10 instructions per (op, child, category, 32-site tile) forward;
    20 + 10 per inner child in the adjoint)."""
    from netrax_amd.simulate import workload_protein_20k

    w = workload_protein_20k()
    out = {"workload": "40 taxa / 2 reticulations / 20,000 amino-acid sites (1 partition), "
                       "20-state model + G4, lnL + full gradient",
           "patterns": int(sum(p.n_sites for p in w.partitions))}
    forms = {}
    for name, prec in (("plain_fp64", 0), ("fp32_mfma", 1)):
        # the FP32 form replays one captured graph per evaluation, as the headline does
        r = run_form(w, prec, steps, warmup, use_graph=prec == 1, tune=prec == 1)
        forms[name] = r
        st = r["stats"]
        d = {"evals_per_s": round(r["evals_per_s"], 3),
             "device_ms_per_eval": round(r["device_ms_per_eval"], 4),
             "fwd_ms": round(r["fwd_ms_per_eval"], 4), "adj_ms": round(r["adj_ms_per_eval"], 4),
             "fwd_hbm_gbps": round(st.fwd_op_bytes / (r["fwd_ms_per_eval"] * 1e-3) / 1e9, 1),
             "adj_hbm_gbps": round(st.adj_bytes / (r["adj_ms_per_eval"] * 1e-3) / 1e9, 1)}
        if prec == 1:
            C = w.partitions[0].model.ncat
            blocks = st.n_sites_padded // 32 * C
            fwd_fl = st.n_uses * blocks * 10 * MFMA_FLOPS_32x32x2
            adj_fl = (2 * st.n_uses + st.n_inner_uses) * blocks * 10 * MFMA_FLOPS_32x32x2
            fu = fwd_fl / (r["fwd_ms_per_eval"] * 1e-3) / 1e12
            au = adj_fl / (r["adj_ms_per_eval"] * 1e-3) / 1e12
            d["mfma"] = {"peak_tflops": MFMA_FP32_PEAK_TFLOPS,
                         "fwd_tflops": round(fu, 2), "fwd_util": round(fu / MFMA_FP32_PEAK_TFLOPS, 4),
                         "adj_tflops": round(au, 2), "adj_util": round(au / MFMA_FP32_PEAK_TFLOPS, 4),
                         "useful_fraction": round(400 / 640, 4),
                         "note": "issued matrix-instruction flops (20 states padded to 32 rows)"}
            d["tuning"] = r["tuning"]
        out[name] = d
        log(f"protein {name}: {r['evals_per_s']:.1f} evals/s, fwd {r['fwd_ms_per_eval']:.3f} ms, "
            f"adj {r['adj_ms_per_eval']:.3f} ms")
    p64, a32 = forms["plain_fp64"], forms["fp32_mfma"]
    out["fp32_vs_fp64"] = {"speedup": round(a32["evals_per_s"] / p64["evals_per_s"], 3),
                           "rel_lnl_diff": abs(a32["lnl"] - p64["lnl"]) / abs(p64["lnl"]),
                           "max_grad_diff_rel_to_norm": float(np.abs(a32["grad"] - p64["grad"]).max()
                                                              / max(1.0, np.abs(p64["grad"]).max()))}
    return out


def bench_candidates(n_cand, steps, warmup):
    """A search step's candidates: n_cand networks (8 seeded bases, 20-60
    taxa, 0-5 reticulations, 2k-50k DNA sites log-uniform, GTR+G4; each
    candidate one seeded rSPR / arc insertion / arc removal) scored -- lnL +
    full gradient -- by one hipGraph launch of the network batch, against
    the same engines evaluated one at a time."""
    from netrax_amd import AnnotatedNetwork, NetraxOptions, NetworkBatch, Precision
    from netrax_amd.simulate import candidate_batch

    t0 = time.perf_counter()
    per_base = max(1, n_cand // 8)
    cands = candidate_batch(8, per_base, seed=5)
    anns = [AnnotatedNetwork(n, p, NetraxOptions(precision=Precision.FP32, use_graph=True, timing=True))
            for n, p in cands]
    stats = [a.stats() for a in anns]
    log(f"candidates: {len(anns)} engines in {time.perf_counter() - t0:.1f}s, "
        f"{sum(s.clv_bytes for s in stats) / 1e9:.1f} GB of CLVs/adjoints")
    big = max(range(len(anns)), key=lambda i: stats[i].grad_bytes)
    best, _ = anns[big].autotune(reps=2)
    for a in anns:
        a.set_tuning(*best[:3])
    batch = NetworkBatch(anns)
    for _ in range(warmup):
        batch.evaluate(0)
    dev = 0.0
    t1 = time.perf_counter()
    for _ in range(steps):
        lnl, dbs, dps, ms = batch.evaluate(0)
        dev += ms
    wall = time.perf_counter() - t1
    # the same engines one at a time (each evaluation one graph launch).  The
    # batch left every CLV valid, so each engine first gets its branch lengths
    # moved by 1e-7 (relative): every CLV is then recomputed, as in the batch
    moved = [a.branch_lengths() * (1.0 + 1e-7) for a in anns]
    seq_dev = 0.0
    t2 = time.perf_counter()
    for a, b in zip(anns, moved):
        a.set_branch_lengths(b)
        a.gradient()
        seq_dev += a.last_timing().total_ms
    seq_wall = time.perf_counter() - t2
    # spot check against the FP64 form
    errs = []
    for i in range(0, len(anns), max(1, len(anns) // 4)):
        n, p = cands[i]
        a64 = AnnotatedNetwork(n, p, NetraxOptions())
        l64 = a64.gradient()[0]
        errs.append(abs(lnl[i] - l64) / abs(l64))
        a64.close()
    gb = sum(s.grad_bytes for s in stats)
    out = {"workload": f"{len(anns)} candidate networks (8 bases x {per_base} seeded moves), "
                       "20-60 taxa, 0-5 reticulations, 2k-50k DNA sites, GTR+G4, lnL + full gradient",
           "networks_per_s": round(len(anns) * steps / wall, 2),
           "batch_device_ms": round(dev / steps, 4),
           "batch_wall_ms": round(1e3 * wall / steps, 4),
           "one_at_a_time_device_ms": round(seq_dev, 4),
           "one_at_a_time_wall_ms": round(1e3 * seq_wall, 4),
           "batch_speedup_device": round(seq_dev / (dev / steps), 3),
           "algorithmic_gbps": round(gb / (dev / steps * 1e-3) / 1e9, 1),
           "patterns_total": int(sum(s.n_sites_padded for s in stats)),
           "tuning": {"group_width": best[0], "tile_sites": best[1], "row_lanes": best[2]},
           "max_rel_lnl_diff_vs_fp64_sampled": max(errs)}
    batch.close()
    for a in anns:
        a.close()
    log(f"candidates: {out['networks_per_s']} networks/s, batch {out['batch_device_ms']} ms vs "
        f"{out['one_at_a_time_device_ms']} ms one at a time")
    return out


def bench_newton_sweep(w):
    """Newton-Raphson sweep over every branch of the headline network (plain
    FP64 engine, reference-normalised derivatives,
    BranchLengthOptimization.cpp:422-477) from the same perturbed lengths:
    the device iteration (nrx_optimize_branches: per branch one captured
    graph of Newton rounds and one host wait) against the host-driven loop
    (netrax_amd.optimization.optimize_branches: one C-ABI call and one wait
    per Newton round)."""
    from netrax_amd import AnnotatedNetwork, NetraxOptions
    from netrax_amd.optimization import optimize_branches

    rng = np.random.default_rng(11)
    start = w.network.brlen * rng.uniform(0.5, 2.0, w.network.n_edges)
    res = {}
    for name in ("device", "host"):
        net = w.network.copy()
        net.brlen = start.copy()
        ann = AnnotatedNetwork(net, w.partitions, NetraxOptions())
        l0 = ann.loglikelihood()
        t0 = time.perf_counter()
        l1 = ann.optimize_branches() if name == "device" else optimize_branches(ann)
        res[name] = (time.perf_counter() - t0, l0, l1, ann.network.brlen.copy())
        ann.close()
    dv, ho = res["device"], res["host"]
    out = {"workload": "every branch of the 50-taxon network (110), plain FP64 engine, lengths perturbed x U(0.5, 2)",
           "device_sweep_s": round(dv[0], 4), "host_sweep_s": round(ho[0], 4),
           "lnl_start": dv[1], "lnl_device": dv[2], "lnl_host": ho[2],
           "max_rel_brlen_diff": float(np.max(np.abs(dv[3] - ho[3]) / np.maximum(np.abs(ho[3]), 1e-12)))}
    log(f"newton sweep: device {dv[0]:.3f}s, host {ho[0]:.3f}s, lnL {dv[1]:.4f} -> {dv[2]:.4f}")
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def _median_ms(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(1e3 * (time.perf_counter() - t0))
    return statistics.median(ts)


def bench_widened(w):
    """The calls around the hot path on the headline workload (plain FP64
    engine unless noted), host-inclusive wall times:

    * model gradient (nrx_model_gradient: dlnL/d exchangeabilities,
      frequencies, Gamma shape for both partitions) against the central
      differences it replaces (2 full lnL evaluations per parameter);
    * reticulation probabilities: d lnL/dp and d2 lnL/dp2 with the CLVs
      reused, and the Newton optimiser from perturbed probabilities;
    * lnL + gradient under a memory budget (the smallest of 50-90 % of the
      resident CLV rows that a segment plan fits) against the resident
      evaluation (bit-identical results);
    * per-candidate optimisation of 32 candidate networks
      (nrx_batch_optimize: branches around the move and every reticulation,
      then one batched scoring) against the same steps one engine at a time."""
    from netrax_amd import AnnotatedNetwork, NetraxOptions, NetworkBatch, Precision
    from netrax_amd.simulate import candidate_batch

    out = {}
    ann = AnnotatedNetwork(w.network.copy(), w.partitions, NetraxOptions())
    b0 = w.network.brlen.copy()
    flip = [0]

    def full_grad():
        flip[0] ^= 1
        ann.set_branch_lengths(b0 * (1.0 + 1e-7 * flip[0]))
        return ann.gradient()

    def full_lnl():
        flip[0] ^= 1
        ann.set_branch_lengths(b0 * (1.0 + 1e-7 * flip[0]))
        return ann.loglikelihood(incremental=0)

    full_grad()
    grad_ms = _median_ms(full_grad, 7)
    lnl_ms = _median_ms(full_lnl, 7)
    mg_ms = _median_ms(ann.model_gradient, 7)
    n_par = sum(len(p.model.subst_rates) - 1 + len(p.model.freqs) - 1 + 1 for p in w.partitions)
    out["model_gradient"] = {"ms": round(mg_ms, 3), "parameters": n_par,
                             "central_difference_ms": round(2 * n_par * lnl_ms, 3),
                             "full_lnl_ms": round(lnl_ms, 3), "lnl_grad_ms": round(grad_ms, 3)}
    pd_ms = _median_ms(ann.prob_derivatives, 7)
    rng = np.random.default_rng(12)
    p0 = np.clip(w.network.retic_prob * rng.uniform(0.5, 1.5, w.network.n_retic), 0.05, 0.95)
    ann.set_reticulation_probs(p0)
    l0 = ann.loglikelihood()
    t0 = time.perf_counter()
    l1 = ann.optimize_reticulations(10)
    opt_ms = 1e3 * (time.perf_counter() - t0)
    out["reticulation_probs"] = {"derivatives_ms": round(pd_ms, 3), "optimise_ms": round(opt_ms, 3),
                                 "lnl_start": l0, "lnl_end": l1, "reticulations": int(w.network.n_retic)}
    ann.set_reticulation_probs(w.network.retic_prob)
    # memory budget: the smallest of 50-90 % of the resident rows that a
    # segment plan fits
    from netrax_amd import _lib
    ann.set_branch_lengths(b0)
    ref = ann.gradient()
    plan0 = ann.memory_plan()
    frac = None
    for f in (0.5, 0.6, 0.7, 0.8, 0.9):
        try:
            ann.set_memory_budget(int(f * plan0["resident_rows"] * plan0["row_bytes"]))
            frac = f
            break
        except _lib.NrxError:
            continue
    plan = ann.memory_plan()
    ann.set_branch_lengths(b0)
    got = ann.gradient()
    same = bool(got[0] == ref[0] and np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2]))
    bud_ms = _median_ms(full_grad, 7)
    ann.set_memory_budget(0)
    out["memory_budget"] = {"budget_fraction_of_resident": frac,
                            "lnl_grad_ms": round(bud_ms, 3), "resident_lnl_grad_ms": round(grad_ms, 3),
                            "rows": plan["rows"], "resident_rows": plan0["resident_rows"],
                            "segments": plan["segments"], "row_bytes": plan["row_bytes"],
                            "bit_identical": same}
    ann.close()
    # per-candidate optimisation inside a batch
    cands = candidate_batch(8, 4, seed=5)

    def move_edges(net):
        es = {0}
        for r in range(net.n_retic):
            es.add(int(net.retic_first_edge[r]))
            es.add(int(net.retic_second_edge[r]))
        return sorted(es)

    edges = [move_edges(n) for n, _p in cands]
    retics = [list(range(n.n_retic)) for n, _p in cands]
    opts = NetraxOptions(precision=Precision.FP32, use_graph=True)
    anns = [AnnotatedNetwork(n.copy(), p, opts) for n, p in cands]
    for a in anns:
        a.loglikelihood()
    batch = NetworkBatch(anns)
    t0 = time.perf_counter()
    lnl_b, st = batch.optimize(edges, retics, brlen_iters=32, prob_iters=10, n_threads=8)
    batch_ms = 1e3 * (time.perf_counter() - t0)
    batch.close()
    for a in anns:
        a.close()
    twins = [AnnotatedNetwork(n.copy(), p, opts) for n, p in cands]
    for a in twins:
        a.loglikelihood()
    t0 = time.perf_counter()
    lnl_s = []
    for i, a in enumerate(twins):
        a.optimize_branches(edges[i], max_iters=32)
        lnl_s.append(a.optimize_reticulations(10) if cands[i][0].n_retic else a.loglikelihood())
    seq_ms = 1e3 * (time.perf_counter() - t0)
    for a in twins:
        a.close()
    out["candidate_optimisation"] = {
        "candidates": len(cands), "batch_ms": round(batch_ms, 2), "one_at_a_time_ms": round(seq_ms, 2),
        "host_workers": 8, "statuses_ok": bool(np.all(st == 0)),
        "max_rel_lnl_diff_vs_one_at_a_time": float(np.max(np.abs(lnl_b - np.array(lnl_s)) / np.abs(lnl_s)))}
    log(f"widened: model gradient {mg_ms:.2f} ms (central differences {2 * n_par * lnl_ms:.1f} ms), "
        f"prob optimiser {opt_ms:.1f} ms, budget {bud_ms:.2f} vs {grad_ms:.2f} ms, "
        f"candidate optimisation {batch_ms:.1f} vs {seq_ms:.1f} ms")
    return out


def bench_tile_sweep(tile_counts, steps, warmup, protein=False):
    """Forward and reverse-sweep time of both forms against the number of
    64-site tiles: the headline network (50 taxa, 4 reticulations, seed 3)
    with two uncompressed partitions holding 2/5 and 3/5 of the tiles, the
    batched form at a fixed shape (3 ops per group, 64-site tiles, or
    --shape).  A time that stays nearly flat while the tiles per CU grow,
    then jumps when the tiles outnumber the resident workgroups, marks a
    walk bound by its per-tile dependency chain, not by HBM.  With
    `protein`: the 40-taxon, 2-reticulation protein network (seed 4), one
    uncompressed partition, plain FP64 form and the FP32 matrix-core form
    (4 ops per group; two 32-site workgroups per tile)."""
    global SHAPE
    from netrax_amd.simulate import make_workload

    shape = SHAPE or ((4, 32, 1) if protein else (3, 64, 1))
    SHAPE = shape  # run_form reads it for the FP32 form
    forms = (("plain", 0), ("mfma", 1)) if protein else (("batched", 1), ("plain", 0))
    rows = []
    for tiles in tile_counts:
        if protein:
            w = make_workload("tile_sweep_aa", 40, 2, [64 * tiles], 4, protein=True, compress=False)
        else:
            k1 = (2 * tiles) // 5
            w = make_workload("tile_sweep", 50, 4, [64 * k1, 64 * (tiles - k1)], 3, compress=False)
        row = {"tiles": tiles}
        for name, prec in forms:
            r = run_form(w, prec, steps, warmup, use_graph=prec == 1, tune=prec == 1)
            row[f"{name}_fwd_ms"] = round(r["fwd_ms_per_eval"], 4)
            row[f"{name}_adj_ms"] = round(r["adj_ms_per_eval"], 4)
        log(f"tile sweep: {row}")
        rows.append(row)
    return {"shape": list(shape), "states": 20 if protein else 4, "rows": rows}


def cpu_baseline(w, sample_patterns, threads, budget_s):
    """The oracle (C FP64 restatement, OpenMP over sites) timed on the host
    cores: lnL + full gradient of the whole workload (every pattern), with
    all `threads` threads (median of at least 3 runs) and with 1 thread (at
    least 1 run, more while the time budget lasts).  `sample_patterns` > 0
    times only the first patterns instead and scales linearly (a quick
    look, not the reported baseline)."""
    from netrax_amd.simulate import Partition
    from oracle import oracle as O

    full = sum(p.n_sites for p in w.partitions)
    parts = w.partitions
    if sample_patterns:
        parts = []
        for p in w.partitions:
            k = max(1, int(round(sample_patterns * p.n_sites / full)))
            parts.append(Partition(p.model, np.ascontiguousarray(p.masks[:, :k]),
                                   np.ascontiguousarray(p.weights[:k])))
    got = sum(p.n_sites for p in parts)
    scale = full / got

    def timed(nthreads, min_runs, budget):
        O.set_threads(nthreads)
        times = []
        t_start = time.perf_counter()
        while len(times) < min_runs or (time.perf_counter() - t_start < budget and len(times) < 20):
            t0 = time.perf_counter()
            O.gradient(w.network, parts)
            times.append(time.perf_counter() - t0)
        return statistics.median(times) * scale, len(times)

    O.set_threads(threads)
    O.gradient(w.network, parts)  # warm (pages, thread pool)
    t_all, n_all = timed(threads, 3, budget_s / 4)
    t_one, n_one = timed(1, 1, budget_s / 2)
    O.set_threads(threads)
    return {"value": round(1.0 / t_all, 5), "unit": "evals/s", "cores": threads,
            "kind": "port",
            "threads_all": {"threads": threads, "evals_per_s": round(1.0 / t_all, 5),
                            "s_per_eval": round(t_all, 4), "runs": n_all},
            "threads_1": {"threads": 1, "evals_per_s": round(1.0 / t_one, 5),
                          "s_per_eval": round(t_one, 4), "runs": n_one},
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "gradient_method": ("per-tree outside pass: one forward and one reverse pass per displayed "
                                "tree, not the reference's per-branch reroot + sumtable evaluation "
                                "(E x (reroot + sumtable) per gradient), so this is faster than the "
                                "reference's own CPU path would be"),
            "sample": (f"oracle (C FP64 restatement, OpenMP over sites) lnL + full gradient on "
                       + (f"all {full} patterns" if got == full else
                          f"the first {got} of {full} patterns, scaled linearly")
                       + f"; median of {n_all} runs at {threads} threads, {n_one} at 1 thread")}


def oracle_check(w, lnl, g, gp, threads):
    from oracle import oracle as O

    O.set_threads(threads)
    olnl, og, ogp = O.gradient(w.network, w.partitions)
    return {
        "rel_lnl_err_vs_oracle": abs(lnl - olnl) / abs(olnl),
        "max_abs_grad_err_vs_oracle": float(np.abs(g - og).max()),
        "max_grad_vs_oracle_rel_to_norm": float(np.abs(g - og).max() / max(1.0, np.abs(og).max())),
        "max_abs_dprob_err_vs_oracle": float(np.abs(gp - ogp).max()) if len(ogp) else 0.0,
        "oracle_lnl": olnl,
    }


def stream_reference(nbytes=1 << 30):
    """Practical HBM rates on this GPU for plain streams (torch elementwise
    kernels over 1 GiB float64 buffers, best of 5, HIP events): a read-only
    sum, a copy (1 read + 1 write) and an add (2 reads + 1 write).  Reported
    beside the published 8 TB/s peak that `roofline.frac` uses."""
    import torch

    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.ones_like(a)
    c = torch.empty_like(a)
    out = {}
    for name, moved, fn in (("read_sum", 1, lambda: a.sum()),
                            ("copy", 2, lambda: c.copy_(a)),
                            ("add_2r1w", 3, lambda: torch.add(a, b, out=c))):
        fn()
        best = float("inf")
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[name + "_gbps"] = round(moved * nbytes / (best * 1e-3) / 1e9, 1)
    del a, b, c
    torch.cuda.empty_cache()
    out["note"] = ("torch elementwise kernels, 1 GiB float64 buffers: what plain streams reach on this "
                   "GPU, beside the 8000 GB/s peak used for roofline.frac")
    return out


def copy_calibration(nbytes):
    """A copy kernel of known size (reads and writes nbytes each), three
    times, for scaling the PMC byte counters of a profiled run."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "kernel": "torch copy_ (elementwise)"}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--form", choices=["batched", "plain"], default="batched",
                    help="headline form: batched = FP32 CLVs, persistent site-tile kernels; "
                         "plain = FP64 CLVs, one launch per pass walking every pruning level")
    ap.add_argument("--only", choices=["batched", "plain", "protein", "candidates", "sweep", "widened"],
                    default=None,
                    help="time only this form or section (profiling runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true")
    ap.add_argument("--cpu-sample-patterns", type=int, default=0, help="0 = every pattern")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cores)")
    ap.add_argument("--cpu-budget-s", type=float, default=24.0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the protein and candidate-batch sections")
    ap.add_argument("--candidates", type=int, default=256)
    ap.add_argument("--stream-reference", action="store_true",
                    help="print the plain-stream HBM rates of this GPU (torch kernels) and exit")
    ap.add_argument("--shape", default=None, metavar="GxTSxH",
                    help="fixed FP32 kernel shape (ops per group x sites per workgroup x lanes "
                         "per row in the reverse sweep) instead of autotuning, e.g. 3x64x2 "
                         "(profiling runs)")
    ap.add_argument("--tile-sweep", default=None, metavar="T1,T2,...",
                    help="only time both forms on the headline network at these numbers of "
                         "64-site tiles (uncompressed synthetic partitions)")
    ap.add_argument("--tile-sweep-protein", action="store_true",
                    help="with --tile-sweep: the 20-state protein network instead")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (PMC counter calibration)")
    args = ap.parse_args()
    global SHAPE
    if args.shape:
        SHAPE = tuple(int(v) for v in args.shape.lower().split("x"))
        if len(SHAPE) == 2:
            SHAPE = SHAPE + (1,)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.stream_reference:
        print(json.dumps(stream_reference()), flush=True)
        return 0
    if args.tile_sweep:
        counts = [int(v) for v in args.tile_sweep.split(",")]
        print(json.dumps({"tile_sweep": bench_tile_sweep(counts, args.steps, args.warmup,
                                                         protein=args.tile_sweep_protein)}), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the site-sharded multi-GPU form is "
              "a design only (partitions split by site tiles, one all-reduce of the per-root site sums). "
              "Run with --gpus 1.", flush=True)
        return 1
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    # torch's HIP runtime and this library's do not share one process, so the
    # stream reference runs in a child process, before this one touches the GPU
    stream_ref = None
    if not args.only:
        try:
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--stream-reference"],
                               capture_output=True, text=True, timeout=300)
            stream_ref = json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
        except (subprocess.TimeoutExpired, ValueError, IndexError):
            stream_ref = None
    if args.only == "protein":
        print(json.dumps({"protein": bench_protein(args.steps, args.warmup)}), flush=True)
        return 0
    if args.only == "candidates":
        print(json.dumps({"candidate_batch": bench_candidates(args.candidates, args.steps, args.warmup)}),
              flush=True)
        return 0

    from netrax_amd.simulate import workload_50taxa_50k

    t0 = time.perf_counter()
    w = workload_50taxa_50k(seed=3)
    n_pat = sum(p.n_sites for p in w.partitions)
    log(f"workload ready in {time.perf_counter() - t0:.1f}s: {n_pat} patterns")
    if args.only == "sweep":
        print(json.dumps({"newton_sweep": bench_newton_sweep(w)}), flush=True)
        return 0
    if args.only == "widened":
        print(json.dumps({"widened": bench_widened(w)}), flush=True)
        return 0

    forms = {}
    for name, prec in (("plain", 0), ("batched", 1)):
        if args.only and args.only != name:
            continue
        forms[name] = run_form(w, prec, args.steps, args.warmup, use_graph=prec == 1, tune=prec == 1)
        r = forms[name]
        log(f"{name}: {r['evals_per_s']:.1f} evals/s, device {r['device_ms_per_eval']:.3f} ms/eval "
            f"(fwd {r['fwd_ms_per_eval']:.3f} ms, adj {r['adj_ms_per_eval']:.3f} ms)")
    head_name = args.only or args.form
    head = forms[head_name]
    st = head["stats"]
    desc = {"plain": "plain FP64 (FP64 CLVs and adjoint; one forward and one reverse-sweep launch, a "
                     "workgroup per site tile walking every pruning level)",
            "batched": "batched FP32 (FP32 CLVs + per-site power-of-two exponents, one persistent "
                       "site-tile launch for all ops and trees, FP32 adjoint with per-site "
                       "exponents, FP64 site sums / mixture / gradient reduction; each evaluation "
                       "replays one captured HIP graph)"}
    out = {
        "metric": METRIC,
        "value": round(head["evals_per_s"], 3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_eval"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head_name == "batched" else "f64",
        "data": "synthetic: seeded Yule network + arc insertions, GTR+G4 alignment simulated on displayed trees",
        "config": {
            "workload": "50 taxa / 4 reticulations (16 displayed trees) / 50,000 DNA sites in 20k+30k "
                        "partitions, GTR+G4, lnL + full gradient (all branch lengths + reticulation probs)",
            "form": desc[head_name],
            "patterns": n_pat, "sites": 50000, "edges": int(w.network.n_edges),
            "clv_ops": st.n_ops_needed, "levels": st.n_levels, "displayed_trees": st.n_classes,
            "adjoint_uses": st.n_uses, "adjoint_lds_handoffs": st.n_lds_handoffs, "parallelism": "single GPU",
            "tuning": head["tuning"],
        },
        "roofline": roofline(head),
        "stream_reference": stream_ref,
        "device_ms_per_eval": round(head["device_ms_per_eval"], 4),
        "timing_note": "value/ms_per_step: engine without timing events; device_ms and roofline "
                       "kernel durations: same-shape engine with HIP events around every launch",
    }
    for name, r in forms.items():
        out[name] = {"evals_per_s": round(r["evals_per_s"], 3),
                     "ms_per_eval": round(r["ms_per_eval"], 4),
                     "device_ms_per_eval": round(r["device_ms_per_eval"], 4),
                     "fwd_ms": round(r["fwd_ms_per_eval"], 4),
                     "adj_ms": round(r["adj_ms_per_eval"], 4),
                     "grad_bytes_per_eval": r["stats"].grad_bytes,
                     "roofline": roofline(r)}
    if "plain" in forms and "batched" in forms:
        p64, b32 = forms["plain"], forms["batched"]
        out["batched_vs_plain"] = {
            "speedup": round(b32["evals_per_s"] / p64["evals_per_s"], 3),
            "rel_lnl_diff": abs(b32["lnl"] - p64["lnl"]) / abs(p64["lnl"]),
            "max_grad_diff_rel_to_norm": float(np.abs(b32["grad"] - p64["grad"]).max()
                                               / max(1.0, np.abs(p64["grad"]).max())),
        }
    if not args.no_extras and not args.only:
        out["newton_sweep"] = bench_newton_sweep(w)
        out["widened"] = bench_widened(w)
        out["protein"] = bench_protein(max(5, args.steps // 2), args.warmup)
        out["candidate_batch"] = bench_candidates(args.candidates, 3, 1)
    if not args.no_oracle_check:
        out["accuracy"] = oracle_check(w, head["lnl"], head["grad"], head["dprob"], threads)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_sample_patterns, threads, args.cpu_budget_s)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
