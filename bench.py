#!/usr/bin/env python3
"""Benchmark: subvolume-leaps/s and firings/s of the RDME tau-leap step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload brusselator|schlogl|batch]

A "step" is one iteration of the device-resident step loop over the whole
workload: one tau pass (propensities, Cao tau', tau'', global min) and one
leap pass (Poisson firings, gather-form diffusion, critical event, negativity
check, commit or leap-back).  With plain tau the iteration runs the leap pass,
then the tau pass on the counts it wrote, then one kernel that commits the
step and decides the next (commit_decide_kernel): its time is reported as
commit_ms_avg and decide_ms_avg is the empty slot between them.  The headline workload is the 1,048,576-
subvolume Brusselator random mesh (128x128x64, degree 4..12, weighted edges,
5 species, 8 reactions) with plain Cao tau; inputs are resident in HBM
before timing starts.  value = accepted subvolume-leaps per second
(a subvolume advanced by one accepted global step).

The CPU baseline is the oracle (same algorithm and Philox streams) on the
host cores over a bounded sample of the same workload; its final state is
compared with the GPU's after the same number of steps (full-size parity).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PMC_TRAFFIC = os.path.join(ROOT, "profiles", "pmc_traffic.json")
PMC_VALU = os.path.join(ROOT, "profiles", "pmc_valu.json")
MLP_FLOP = 2 * (12 * 64 + 64 * 64 + 64 * 2)  # per subvolume and inference
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense bf16 MFMA


def algorithmic_bytes(S, topos, part):
    """Compulsory HBM bytes of one launch: what the kernel must read and
    write once.  part: "tau" (phase_a_kernel: counts, CSR row pointers,
    neighbour ids and weights), "leap" (the whole leap pass: counts in and
    out plus the topology once), "leap_kernel" (phase_b_kernel of the
    edge-parallel pass: counts, the forward CSR -- each undirected edge once
    -- and the 64-bit balances it produces), "finish" (phase_b_finish_kernel:
    counts and balances in, new counts out), "step" (one whole step:
    N (12 S + 10) + 2 topology)."""
    tot = 0
    for t in topos:
        N = int(t["N"])
        E = int(len(t["col"]))
        wt = t.get("w") is not None
        if part == "tau":
            tot += 4 * S * N + 4 * (N + 1) + 4 * E + (4 * E if wt else 0)
        elif part == "leap":
            tot += 8 * S * N + 4 * (N + 1) + 4 * E + (4 * E if wt else 0)
        elif part == "leap_kernel":
            tot += 4 * S * N + 4 * (N + 1) + 2 * E + (2 * E if wt else 0) + 8 * S * N
        elif part == "finish":
            tot += 4 * S * N + 8 * S * N + 4 * S * N
        elif part == "step":
            # the minimal two-pass step: counts read twice and written once,
            # a 10-byte tau / flag per subvolume, the topology read twice
            tot += N * (12 * S + 10) + 2 * (4 * (N + 1) + 4 * E + (4 * E if wt else 0))
        else:
            raise ValueError(part)
    return tot


def build_workload(name):
    from pliscuda_amd import workloads as W
    if name == "brusselator":
        return W.brusselator_mesh()
    if name == "schlogl":
        return W.schlogl()
    if name == "batch":
        return W.schlogl_batch()
    if name == "bruss_batch":
        return W.brusselator_batch()
    raise SystemExit(f"unknown workload {name}")


def tuned_shape(wl, seed=2024):
    """(waves per block, blocks per CU) the autotuner picks for a workload,
    found on a throwaway context."""
    import torch

    from pliscuda_amd import capi
    ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], eps=wl["eps"], nc=wl["nc"], seed=seed)
    ctx.arm(t_end=float("inf"))
    ctx.autotune(3)
    sh = ctx.launch_shape()
    ctx.close()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    return sh["waves_per_block"], -(-sh["blocks"] // cus)


def gpu_measure(wl, steps, warmup, seed=2024, autotune=True, shape=None):
    import torch

    from pliscuda_amd import capi
    wpb, bpc = shape if shape else (0, 0)
    if autotune and not shape:
        # the autotuner times real iterations, which advance the trajectory:
        # tune on a context of its own so that the timed run starts from the
        # initial state (as the learned leg does)
        wpb, bpc = tuned_shape(wl, seed)
    ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], eps=wl["eps"], nc=wl["nc"], seed=seed,
                       waves_per_block=wpb, blocks_per_cu=bpc)
    ctx.arm(t_end=float("inf"))
    shape = ctx.launch_shape()
    ctx.iterate(warmup)
    ctx.sync()
    s0 = [ctx.stats(r) for r in range(ctx.n)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.iterate(steps)
    ctx.sync()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s1 = [ctx.stats(r) for r in range(ctx.n)]
    acc = sum(b["steps"] - a["steps"] for a, b in zip(s0, s1))
    leaps = sum(b["subvolume_leaps"] - a["subvolume_leaps"] for a, b in zip(s0, s1))
    fir = sum(b["firings"] - a["firings"] for a, b in zip(s0, s1))
    lbs = sum(b["leapbacks"] - a["leapbacks"] for a, b in zip(s0, s1))
    crit = sum(b["crit_events"] - a["crit_events"] for a, b in zip(s0, s1))
    simt = sum((b["t"] - a["t"]) * n for a, b, n in zip(s0, s1, ctx.N))
    # per-kernel durations: HIP events around every launch on the context's stream
    # (the leap pass by kernel in separate iterations: the extra events
    # between its kernels cost a few microseconds each)
    n_prof = min(steps, 50)
    prof = ctx.profile(n_prof)
    prof += ctx.profile_kernels(n_prof)[5:]
    ctx.close()
    return {"seconds": dt, "accepted_steps": acc, "subvolume_leaps": leaps, "firings": fir,
            "leapbacks": lbs, "crit_events": crit, "subvolume_simtime": simt, "shape": shape, "tuned": (wpb, bpc),
            "tau_ms": prof[1] / n_prof, "decide_ms": prof[2] / n_prof, "leap_ms": prof[3] / n_prof,
            "commit_ms": prof[4] / n_prof, "iter_ms_events": prof[0] / n_prof,
            "leap_kernel_ms": prof[5] / n_prof, "redo_ms": prof[6] / n_prof, "finish_ms": prof[7] / n_prof}


def learned_measure(wl, steps, warmup, train_steps, batch, seed=2024, kappa=1.0, snapshots=16,
                    per_snapshot=16384, shape=None, act=1, records=None, net=None):
    """Records from plain steps of the same workload, GPU training, then the
    learned step timed like the plain one.  act: the network's activation
    (1 GELU, 0 ReLU); records: (F, Y, L, plain_stats) to reuse; net: a
    trained network {"w", "norm", "vmax"} to use instead of records and
    training."""
    import torch

    from pliscuda_amd import capi, learned
    t_rec = t_train = 0.0
    info = {}
    if net is None:
        rec = capi.Context(wl["net"], wl["topos"], wl["x0"], eps=wl["eps"], nc=wl["nc"], seed=seed + 1)
        t0 = time.perf_counter()
        if records is None:
            F, Y, Cc = learned.collect_records(rec, snapshots=snapshots, stride=4, per_snapshot=per_snapshot)
            plain_stats = [rec.stats(r) for r in range(rec.n)]
        else:
            F, Y, Cc, plain_stats = records
        t_rec = time.perf_counter() - t0
        vmax = learned.calibrate_viol_rate(plain_stats, [int(t["N"]) for t in wl["topos"]])
        rec.mlp_activation(act)
        t0 = time.perf_counter()
        norm, losses = learned.train(rec, F, Y, Cc, steps=train_steps, batch=batch, lr=1e-3)
        rec.sync()
        t_train = time.perf_counter() - t0
        w = rec.mlp_get()
        rec.close()
        info = {"records": int(F.shape[0]), "reject_label_rate": float(Cc.mean()), "_records": (F, Y, Cc, plain_stats),
                "train_steps": train_steps, "train_batch": batch,
                "loss_first": float(losses[:10].mean()), "loss_last": float(losses[-100:].mean()),
                "_net": {"w": w, "norm": norm, "vmax": vmax}}
    else:
        w, norm, vmax = net["w"], net["norm"], float(net["vmax"])
        info = {"network": "loaded"}
    wpb, bpc = shape if shape else (0, 0)
    ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], eps=wl["eps"], nc=wl["nc"], seed=seed,
                       waves_per_block=wpb, blocks_per_cu=bpc)
    ctx.mlp_set(w, norm)
    ctx.mlp_activation(act)
    ctx.set_tau_mode(True, kappa=kappa, theta=learned.THETA, viol_rate_max=vmax)
    ctx.arm(t_end=float("inf"))
    ctx.iterate(warmup)
    ctx.sync()
    s0 = [ctx.stats(r) for r in range(ctx.n)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.iterate(steps)
    ctx.sync()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s1 = [ctx.stats(r) for r in range(ctx.n)]
    n_prof = min(steps, 20)
    prof = ctx.profile_kernels(n_prof)
    ctx.close()
    d = lambda f: sum(b[f] - a[f] for a, b in zip(s0, s1))
    props = d("learned_steps") + d("learned_rejects")
    return {"seconds": dt, "subvolume_leaps": d("subvolume_leaps"), "firings": d("firings"),
            "accepted_steps": d("steps"), "learned_steps": d("learned_steps"),
            "learned_rejects": d("learned_rejects"), "leapbacks": d("leapbacks"),
            "crit_events": d("crit_events"),
            "rejection_rate": d("learned_rejects") / props if props else None,
            "subvolume_simtime": sum((b["t"] - a["t"]) * n for a, b, n in
                                     zip(s0, s1, [int(t["N"]) for t in wl["topos"]])),
            "activation": "gelu" if act == 1 else "relu",
            "tau_ms": prof[1] / n_prof, "decide_ms": prof[2] / n_prof, "leap_ms": prof[3] / n_prof,
            "commit_ms": prof[4] / n_prof, "leap_kernel_ms": prof[5] / n_prof, "redo_ms": prof[6] / n_prof,
            "finish_with_proposals_ms": prof[7] / n_prof, "iteration_ms_events": prof[0] / n_prof,
            "records_seconds": t_rec, "train_seconds": t_train,
            "viol_rate_max": vmax, "kappa": kappa, **info}


def cpu_baseline(wl, seed, sample_steps, gpu_check=True):
    """Oracle on the host cores for sample_steps accepted steps (bounded)."""
    from oracle import ffi
    threads = ffi.num_threads()
    t0 = time.perf_counter()
    states, leaps, fires = [], 0, 0
    nreal = len(wl["x0"])
    use = min(nreal, 4)
    for r in range(use):
        m = ffi.Model(wl["net"], wl["topos"][r])
        x = np.array(wl["x0"][r], dtype=np.uint32).copy()
        st = ffi.run(m, x, wl["eps"], wl["nc"], seed, float("inf"), step_limit=sample_steps,
                     realization=r)
        states.append(x)
        leaps += st.subvolume_leaps
        fires += st.firings
    dt = time.perf_counter() - t0
    parity = None
    if gpu_check:
        from pliscuda_amd import capi
        ctx = capi.Context(wl["net"], wl["topos"][:use], wl["x0"][:use], eps=wl["eps"], nc=wl["nc"],
                           seed=seed)
        ctx.run(float("inf"), step_limit=sample_steps)
        parity = all(np.array_equal(ctx.get_state(r), states[r]) for r in range(use))
        ctx.close()
    return {"value": leaps / dt, "unit": "subvolume-leaps/s", "cores": threads, "kind": "port",
            "firings_per_s": fires / dt,
            "sample": f"{use} realization(s) x {sample_steps} accepted steps of the same workload "
                      f"from the same initial state and seed ({dt:.1f} s of CPU work)",
            "seconds": dt, "gpu_state_bit_identical": parity}


def abc_to_t_end(seed=2024):
    """The reference's CPU-runnable case (64 realizations of the 8^3 A-B-C
    reflecting grid to t = 10): GPU wall time to t_end, the oracle on the host
    cores for the same work, and whether their final states agree bit for bit."""
    import torch

    from oracle import ffi
    from pliscuda_amd import capi
    from pliscuda_amd import workloads as W
    wl = W.abc_reflecting()
    ctx = capi.Context(wl["net"], wl["topos"], wl["x0"], eps=wl["eps"], nc=wl["nc"], seed=seed)
    warm = capi.Context(wl["net"], wl["topos"], wl["x0"], eps=wl["eps"], nc=wl["nc"], seed=seed + 1)
    warm.run(wl["t_end"])
    warm.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.run(wl["t_end"])
    ctx.sync()
    dt = time.perf_counter() - t0
    st = [ctx.stats(r) for r in range(ctx.n)]
    leaps = sum(s["subvolume_leaps"] for s in st)
    gpu_states = [ctx.get_state(r) for r in range(ctx.n)]
    ctx.close()
    m = ffi.Model(wl["net"], wl["topos"][0])
    xs = [np.array(x, dtype=np.uint32).copy() for x in wl["x0"]]
    t0 = time.perf_counter()
    ffi.run_batch(m, wl["topos"], xs, wl["eps"], wl["nc"], seed, wl["t_end"])
    cdt = time.perf_counter() - t0
    return {"subvolume_leaps_per_s": leaps / dt, "seconds_to_t_end": dt, "steps_per_realization":
            float(np.mean([s["steps"] for s in st])), "cpu_seconds_to_t_end": cdt,
            "cpu_subvolume_leaps_per_s": leaps / cdt, "cpu_threads": ffi.num_threads(),
            "gpu_state_bit_identical": all(np.array_equal(g, x) for g, x in zip(gpu_states, xs))}


def reference_benchmarks(reps=2, seed=2024):
    """The reference's benchmark loop (PLISCuda/benchmarks/run.sh: each case
    run NUM times through the binary, steps/second and elapsed time parsed
    from stdout) over its four benchmark inputs, with the driver's extra
    keys: subvolume-leaps/second and firings."""
    import re
    import shutil
    import subprocess
    import tempfile
    root = os.path.join(ROOT, "tests", "golden", "ref_benchmarks")
    driver = os.path.join(ROOT, "pliscuda_amd", "bin", "pliscuda")
    res = {}
    for name in ("chemotaxis", "diff-5125", "diff-64", "react-64"):
        runs = []
        with tempfile.TemporaryDirectory() as tmp:
            for f in os.listdir(os.path.join(root, name)):
                shutil.copyfile(os.path.join(root, name, f), os.path.join(tmp, f))
            with open(os.path.join(tmp, "conf.txt"), "a") as f:
                f.write(f"\nseed={seed}\n")
            for _ in range(reps):
                r = subprocess.run([driver, "conf.txt"], cwd=tmp, capture_output=True, text=True, timeout=120)
                if r.returncode != 0:
                    raise RuntimeError(f"driver failed on {name}: {r.stderr[-500:]}")
                def num(key):
                    m = re.search(re.escape(key) + r"\s*([0-9.eE+-]+)", r.stdout)
                    return float(m.group(1)) if m else None
                el = re.search(r"elapsed time:\s*(\d+)h\s*(\d+)m\s*([0-9.]+)s", r.stdout)
                secs = int(el.group(1)) * 3600 + int(el.group(2)) * 60 + float(el.group(3)) if el else None
                runs.append({"steps_per_s": num("steps/second:"), "elapsed_s": secs,
                             "subvolume_leaps_per_s": num("subvolume-leaps/second:"),
                             "firings": num("firings:"), "steps": num("simulation steps:")})
        mean = lambda k: float(np.mean([q[k] for q in runs])) if all(q[k] is not None for q in runs) else None
        res[name] = {"runs": reps, "steps_per_s": mean("steps_per_s"), "elapsed_s": mean("elapsed_s"),
                     "subvolume_leaps_per_s": mean("subvolume_leaps_per_s"),
                     "firings_per_s": (mean("firings") / mean("elapsed_s")) if mean("elapsed_s") else None,
                     "steps": mean("steps")}
    return res


def copy_probe(nbytes=1 << 28, reps=5):
    """A plain device-to-device copy of known size (torch elementwise copy),
    used to calibrate the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def load_traffic(kernel, path=None):
    try:
        with open(path or PMC_TRAFFIC) as f:
            d = json.load(f)
        return d.get(kernel)
    except (OSError, ValueError):
        return None


# Instruction-issue bound: a wave64 VALU instruction holds one SIMD for 4
# cycles (16 lanes per cycle); MI355X has 256 CUs x 4 SIMDs at up to 2.4 GHz.
VALU_SIMDS, VALU_CYCLES_PER_INST, CLOCK_HZ = 1024, 4, 2.4e9


def valu_issue(workload, kname, ms):
    """VALU instructions per launch (rocprofv3 SQ_INSTS_VALU, committed
    under profiles/) against the launch's measured duration: the fraction
    of the chip's VALU issue slots the kernel fills."""
    n = load_traffic(f"{workload}:{kname}", PMC_VALU)
    if not n or not ms:
        return None
    t_min = n * VALU_CYCLES_PER_INST / VALU_SIMDS / CLOCK_HZ
    return {"kernel": kname, "valu_instructions": n, "min_us": t_min * 1e6, "measured_us": ms * 1e3,
            "frac": t_min / (ms * 1e-3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="brusselator")
    ap.add_argument("--cpu-sample-steps", type=int, default=70)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--shape", default=None,
                    help="WPB,BPC: fixed waves per block and blocks per CU instead of autotuning")
    ap.add_argument("--extra", action="store_true", help="also measure the Schlogl grid and batch")
    ap.add_argument("--copy-probe", action="store_true", help="only run the counter-calibration copy")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-ref-benchmarks", action="store_true",
                    help="skip the loop over the reference's benchmark inputs")
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--train-batch", type=int, default=65536)
    ap.add_argument("--kappa", type=float, default=1.0)
    ap.add_argument("--net-file", default=None,
                    help="trained networks of the learned legs (.npz): loaded when the file exists "
                         "(no records, no training), else written after training")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(realizations shard across GPUs by construction, see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    wl = build_workload(args.workload)
    S = int(wl["net"]["S"])
    shape = tuple(int(v) for v in args.shape.split(",")) if args.shape else None
    res = gpu_measure(wl, args.steps, args.warmup, autotune=not args.no_autotune, shape=shape)
    value = res["subvolume_leaps"] / res["seconds"]
    bytes_leap = algorithmic_bytes(S, wl["topos"], "leap")
    bytes_tau = algorithmic_bytes(S, wl["topos"], "tau")
    split = res["finish_ms"] > 0.0  # edge-parallel leap pass: three kernels
    # the roofline is taken on the single kernel with the longest launches
    cand = {"phase_a_kernel": (res["tau_ms"], bytes_tau),
            "phase_b_kernel": (res["leap_kernel_ms"],
                               algorithmic_bytes(S, wl["topos"], "leap_kernel" if split else "leap"))}
    if split:
        cand["phase_b_finish_kernel"] = (res["finish_ms"], algorithmic_bytes(S, wl["topos"], "finish"))
    kname = max(cand, key=lambda k: cand[k][0])
    dom_ms, dom_bytes = cand[kname]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = load_traffic(f"{args.workload}:{kname}")
    # whole-step compulsory bytes x accepted steps / wall time
    b_step = algorithmic_bytes(S, wl["topos"], "step")
    b_step_gbs = b_step * res["accepted_steps"] / len(wl["topos"]) / res["seconds"] / 1e9
    out = {
        "metric": "subvolume-leaps/s (plain Cao tau), 1 MI355X",
        "value": value,
        "unit": "subvolume-leaps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["seconds"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["name"], "subvolumes": int(sum(t["N"] for t in wl["topos"])),
                   "directed_edges": int(sum(len(t["col"]) for t in wl["topos"])),
                   "species": S, "reactions": int(wl["net"]["R"]), "realizations": len(wl["topos"]),
                   "epsilon": wl["eps"], "n_critical": wl["nc"], "tau": "plain Cao",
                   "launch": res["shape"]},
        "firings_per_s": res["firings"] / res["seconds"],
        "accepted_steps": res["accepted_steps"],
        "leapbacks": res["leapbacks"],
        "crit_events": res["crit_events"],  # steps that ended at tau'' (a critical event fired)
        "subvolume_simtime_per_s": res["subvolume_simtime"] / res["seconds"],
        "kernels": {"step_order": "leap pass, tau pass on its counts, commit_decide_kernel",
                    "tau_ms_avg": res["tau_ms"], "leap_ms_avg": res["leap_ms"],
                    "decide_ms_avg": res["decide_ms"], "commit_ms_avg": res["commit_ms"],
                    "iteration_ms_events": res["iter_ms_events"],
                    "tau_alg_bytes": bytes_tau, "leap_alg_bytes": bytes_leap,
                    "tau_GBs": bytes_tau / (res["tau_ms"] * 1e-3) / 1e9,
                    "leap_GBs": bytes_leap / (res["leap_ms"] * 1e-3) / 1e9,
                    "leap_pass_split_ms_avg": {"phase_b_kernel": res["leap_kernel_ms"],
                                               "phase_b_redo_kernel": res["redo_ms"],
                                               "phase_b_finish_kernel": res["finish_ms"]},
                    "per_kernel_GBs": {k: b / (ms * 1e-3) / 1e9 for k, (ms, b) in cand.items() if ms > 0}},
        "step_roofline": {"B_step": b_step, "achieved": b_step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": b_step_gbs / HBM_PEAK_GBS},
        "valu_issue": [v for v in (valu_issue(args.workload, k, ms) for k, (ms, _) in cand.items()) if v],
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic},
    }
    # the bound that applies: VALU issue when the kernel fills a larger
    # fraction of the chip's VALU issue slots than of HBM bandwidth (the
    # GB/s figures stay those of its algorithmic bytes)
    vi = [v for v in out["valu_issue"] if v["kernel"] == kname]
    if vi and vi[0]["frac"] > out["roofline"]["frac"]:
        out["roofline"]["bound"] = "valu"
        out["roofline"]["valu_issue_frac"] = vi[0]["frac"]
    if not args.no_learned and args.workload == "brusselator":
        simt_plain = res["subvolume_simtime"] / res["seconds"]
        recs = None
        nets = {}
        if args.net_file and os.path.exists(args.net_file):
            with np.load(args.net_file) as z:
                nets = {a: {"w": z[f"w{a}"], "norm": z[f"norm{a}"], "vmax": z[f"vmax{a}"]} for a in (0, 1)}
        # the network as specified (GELU), then the same records with ReLU
        for key, act in (("learned", 1), ("learned_relu", 0)):
            lr_ = learned_measure(wl, args.steps, args.warmup, args.train_steps, args.train_batch,
                                  kappa=args.kappa, shape=res["tuned"], act=act, records=recs,
                                  net=nets.get(act))
            recs = lr_.pop("_records", None)
            if "_net" in lr_:
                nets_new = lr_.pop("_net")
                if args.net_file:
                    nets[act] = nets_new
            if args.net_file and len(nets) == 2 and not os.path.exists(args.net_file):
                np.savez(args.net_file, **{f"{k}{a}": np.asarray(v) for a, d in nets.items() for k, v in d.items()})
            lv = lr_["subvolume_leaps"] / lr_["seconds"]
            simt_learned = lr_["subvolume_simtime"] / lr_["seconds"]
            out[key] = {"subvolume_leaps_per_s": lv, "firings_per_s": lr_["firings"] / lr_["seconds"],
                        "ms_per_step": lr_["seconds"] * 1e3 / args.steps,
                        "speedup_per_step": lv / value,
                        "subvolume_simtime_per_s": simt_learned,
                        "speedup_simulated_time": simt_learned / simt_plain if simt_plain else None,
                        **{k: v for k, v in lr_.items() if k not in ("seconds",)}}
            # MFMA accounting: 2 (12x64 + 64x64 + 64x2) = 9,984 FLOP per subvolume per
            # inference, training about 3x that per record
            n_sub = int(sum(t["N"] for t in wl["topos"]))
            inf_flop = MLP_FLOP * n_sub
            mf = {"inference_flop_per_step": inf_flop, "peak_bf16_tflops": PEAK_BF16_TFLOPS,
                  "inference_tflops_over_finish": inf_flop / (lr_["finish_with_proposals_ms"] * 1e-3) / 1e12}
            mf["inference_frac_over_finish"] = mf["inference_tflops_over_finish"] / PEAK_BF16_TFLOPS
            if lr_.get("train_steps") and lr_.get("train_seconds"):
                tr = 3 * MLP_FLOP * lr_["train_batch"] * lr_["train_steps"] / lr_["train_seconds"] / 1e12
                mf["train_tflops"] = tr
                mf["train_frac"] = tr / PEAK_BF16_TFLOPS
            out[key]["mfma"] = mf
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, 2024, args.cpu_sample_steps)
    if not args.no_ref_benchmarks:
        out["reference_benchmarks"] = reference_benchmarks()
    if args.extra:
        extra = {}
        for name in ("schlogl", "batch"):
            w2 = build_workload(name)
            r2 = gpu_measure(w2, 200, 20)
            extra[w2["name"]] = {"subvolume_leaps_per_s": r2["subvolume_leaps"] / r2["seconds"],
                                 "firings_per_s": r2["firings"] / r2["seconds"],
                                 "ms_per_step": r2["seconds"] * 1e3 / 200,
                                 "tau_ms_avg": r2["tau_ms"], "leap_ms_avg": r2["leap_ms"],
                                 "leap_GBs": algorithmic_bytes(1, w2["topos"], "leap") / (r2["leap_ms"] * 1e-3) / 1e9}
        extra["abc_reflecting_8^3_x64_to_t10"] = abc_to_t_end()
        wb = build_workload("bruss_batch")
        rp = gpu_measure(wb, 100, 10)
        rl = learned_measure(wb, 100, 10, 5000, 65536, snapshots=8, per_snapshot=512, shape=rp["tuned"])
        rl.pop("_records")
        vp, vl = rp["subvolume_leaps"] / rp["seconds"], rl["subvolume_leaps"] / rl["seconds"]
        sp, sl = rp["subvolume_simtime"] / rp["seconds"], rl["subvolume_simtime"] / rl["seconds"]
        extra[wb["name"] + "_learned"] = {
            "plain_subvolume_leaps_per_s": vp, "learned_subvolume_leaps_per_s": vl,
            "speedup_per_step": vl / vp, "speedup_simulated_time": sl / sp,
            "plain_ms_per_step": rp["seconds"] * 10.0, "learned_ms_per_step": rl["seconds"] * 10.0,
            "rejection_rate": rl["rejection_rate"], "leapbacks": rl["leapbacks"], "records": rl["records"],
            "train_seconds": rl["train_seconds"], "loss_last": rl["loss_last"]}
        out["extra"] = extra
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
