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


def run_form(w, precision, steps, warmup, use_graph=False):
    """Times `steps` lnL+gradient evaluations; returns a result dict."""
    ann = make_engine(w, precision, use_graph=use_graph)
    b0 = w.network.brlen.copy()
    b1 = b0 * (1.0 + 1e-7)
    fwd_ms = adj_ms = tot_ms = 0.0
    fwd_l = adj_l = 0
    for i in range(warmup):
        ann.set_branch_lengths(b1 if i % 2 else b0)
        ann.gradient()
    t0 = time.perf_counter()
    for i in range(steps):
        ann.set_branch_lengths(b1 if (warmup + i) % 2 else b0)
        lnl, g, gp = ann.gradient()
        t = ann.last_timing()
        fwd_ms += t.fwd_ms
        adj_ms += t.adj_ms
        tot_ms += t.total_ms
        fwd_l += t.fwd_launches
        adj_l += t.adj_launches
    wall = time.perf_counter() - t0
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
        "lnl": lnl, "grad": g, "dprob": gp, "stats": st,
    }
    ann.close()
    return res


def roofline(res):
    st = res["stats"]
    kern = [("forward pruning", st.fwd_op_bytes, res["fwd_ms_per_eval"], res["fwd_launches_per_eval"]),
            ("adjoint pruning", st.adj_bytes, res["adj_ms_per_eval"], res["adj_launches_per_eval"])]
    name, by, ms, launches = max(kern, key=lambda k: k[2])
    achieved = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None, "kernel": name,
            "algorithmic_bytes_per_eval": by, "kernel_ms_per_eval": round(ms, 4),
            "launches_per_eval": launches}


def cpu_baseline(w, sample_patterns, threads, budget_s):
    from netrax_amd.simulate import Partition
    from oracle import oracle as O

    O.set_threads(threads)
    full = sum(p.n_sites for p in w.partitions)
    parts = []
    for p in w.partitions:
        k = max(1, int(round(sample_patterns * p.n_sites / full)))
        parts.append(Partition(p.model, np.ascontiguousarray(p.masks[:, :k]),
                               np.ascontiguousarray(p.weights[:k])))
    got = sum(p.n_sites for p in parts)
    O.gradient(w.network, parts)  # warm
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < budget_s and len(times) < 50):
        t0 = time.perf_counter()
        O.gradient(w.network, parts)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    per_eval_full = med * full / got  # CPU cost is linear in patterns
    return {"value": round(1.0 / per_eval_full, 5), "unit": "evals/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle (C FP64 restatement, OpenMP) lnL+gradient on the first {got} of "
                       f"{full} patterns, median of {len(times)} runs = {med:.4f} s, "
                       f"scaled linearly to all patterns")}


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


def copy_calibration(nbytes):
    """Device-to-device copy of a known size, for scaling PMC byte counters."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"copy_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--form", choices=["batched", "plain"], default="batched",
                    help="headline form: batched = FP32 CLVs, persistent site-tile kernels; "
                         "plain = FP64 CLVs, one launch per pruning level")
    ap.add_argument("--only", choices=["batched", "plain"], default=None,
                    help="time only this form (profiling runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true")
    ap.add_argument("--cpu-sample-patterns", type=int, default=4000)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cores)")
    ap.add_argument("--cpu-budget-s", type=float, default=12.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (PMC counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the site-sharded multi-GPU form is "
              "design only (DESIGN.md). Run with --gpus 1.", flush=True)
        return 1
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)

    from netrax_amd.simulate import workload_50taxa_50k

    t0 = time.perf_counter()
    w = workload_50taxa_50k(seed=3)
    n_pat = sum(p.n_sites for p in w.partitions)
    log(f"workload ready in {time.perf_counter() - t0:.1f}s: {n_pat} patterns")

    forms = {}
    for name, prec in (("plain", 0), ("batched", 1)):
        if args.only and args.only != name:
            continue
        forms[name] = run_form(w, prec, args.steps, args.warmup)
        r = forms[name]
        log(f"{name}: {r['evals_per_s']:.1f} evals/s, device {r['device_ms_per_eval']:.3f} ms/eval "
            f"(fwd {r['fwd_ms_per_eval']:.3f} ms, adj {r['adj_ms_per_eval']:.3f} ms)")
    head_name = args.only or args.form
    head = forms[head_name]
    st = head["stats"]
    desc = {"plain": "plain FP64 (FP64 CLVs, one launch per pruning level, FP64 adjoint)",
            "batched": "batched FP32 (FP32 CLVs + per-site power-of-two exponents, one persistent "
                       "site-tile launch for all ops and trees, FP32 adjoint with per-site "
                       "exponents, FP64 site sums / mixture / gradient reduction)"}
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
            "adjoint_uses": st.n_uses, "parallelism": "single GPU",
        },
        "roofline": roofline(head),
        "device_ms_per_eval": round(head["device_ms_per_eval"], 4),
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
    if not args.no_oracle_check:
        out["accuracy"] = oracle_check(w, head["lnl"], head["grad"], head["dprob"], threads)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_sample_patterns, threads, args.cpu_budget_s)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
