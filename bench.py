#!/usr/bin/env python3
"""Throughput of the batched FP64 GRPP matrix build on one MI355X.

One step = one full-GRPP matrix build (type-1 + semilocal type-2 + spin-orbit + outercore, all
ordered shell pairs x all GRPP centres) of the ``ragged_molecules`` workload: 200 independent
molecules of 2-30 atoms in one plan, 1-6 GRPP centres each (semilocal L <= 5 + local, SO, outercore
s/p), Cartesian shells l = 0..6 drawn with probability proportional to 1/(l+1) and 1-8 primitives
(11,051 shells, 666 centres, 2.70e6 shell-pair x centre tasks), synthetic, seed 0.
``--workload heavy_cluster`` times the 8-atom cluster instead (100,352 tasks).  Inputs are packed and
resident in HBM before timing; a step runs the kernels.  The build is timed twice: with the plain
adaptive method, then with radial tables ordered by the learned level predictor (trained
beforehand, untimed, on seeds 1-4); ``value`` is the latter and ``learned_grid`` holds the
comparison.  ``parity`` compares the matrices with the genuine reference's blocks for the sampled
ordered pairs of tests/golden (per operator: max abs / relative error and the largest error over
the reference's own noise floor of the block).  ``percall`` times the reference's own usage
pattern through this library (host threads calling libgrpp_full_grpp_integrals per task).

Prints one JSON line (see the keys below).  --gpus N with N > 1 is not supported: the build is
single-GPU (tasks would shard by shell-pair rows without communication).
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

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix), AMD published


def workload_config(name, sysm, ntask, nint):
    if isinstance(sysm, list):
        return {"workload": f"{name}: {len(sysm)} independent molecules of 2-30 atoms in one plan, 1-6 GRPP "
                            f"centres each (L<=5 + local + SO + OC s/p), Cartesian shells l=0..6 with P ~ 1/(l+1), "
                            f"1-8 primitives",
                "molecules": len(sysm), "shells": sum(len(m.shells) for m in sysm),
                "centres": sum(len(m.centers) for m in sysm), "tasks": ntask, "integrals_per_step": nint, "seed": 0}
    return {"workload": f"{name}: 8 atoms, 12 bohr cube, GRPP L<=4 + local + SO L<=4 + OC s/p, "
                        f"basis 4s4p3d2f1g x 6 primitives",
            "shells": len(sysm.shells), "centres": len(sysm.centers), "dim": sysm.dim, "tasks": ntask,
            "integrals_per_step": nint, "seed": 0}


def traffic_bytes(kernel: str, workload: str):
    """HBM bytes per launch of ``kernel`` on ``workload``, from the committed rocprofv3 counter
    summary of that workload (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a
    plain copy of known size); None when no summary of that workload is present."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 2 and f[0] == kernel + ":" and f[1].isdigit():
            return int(f[1])
    return None


def fp64_counters(workload: str, step_s: float):
    """Issued FP64 work per build from the committed rocprofv3 counter summary of ``workload``
    (profiles/counters_<workload>.txt: wave instructions per launch): FLOP = 64 lanes x (2 FMA + MUL
    + ADD + TRANS) + 512 x MFMA_MOPS_F64, an upper bound (every lane counted active).  Returns the
    per-kernel figures and the whole build's fraction of the FP64 peak at the measured step time;
    None when no summary of that workload is present."""
    rows, path = _counter_rows(workload)
    if not rows:
        return None
    per = {}
    for k, r in rows.items():
        if "VALU_FMA_F64" not in r:
            continue
        fl = 64.0 * (2 * r["VALU_FMA_F64"] + r["VALU_MUL_F64"] + r["VALU_ADD_F64"] + r["VALU_TRANS_F64"]) + \
            512.0 * r.get("VALU_MFMA_MOPS_F64", 0.0)
        per[k] = fl
    if not per:
        return None
    total = sum(per.values())
    return {"source": os.path.relpath(path, ROOT), "issued_flops_per_build": total,
            "per_kernel": per, "build_achieved_tflops": total / step_s / 1e12,
            "build_frac": total / step_s / 1e12 / FP64_PEAK_TFLOPS,
            "note": "upper bound: 64 active lanes per wave instruction"}


def _counter_rows(workload: str):
    """per-kernel rows of the committed rocprofv3 counter summary of ``workload`` (both tables)"""
    path = os.path.join(ROOT, "profiles", f"counters_{workload}.txt")
    if not os.path.exists(path):
        return None, path
    rows, cols = {}, None
    for line in open(path):
        f = line.split()
        if len(f) >= 2 and f[0] == "kernel":
            cols = f[1:]
        elif cols and len(f) == len(cols) + 1 and f[0].startswith("k_"):
            try:
                rows.setdefault(f[0], {}).update({c: float(v) for c, v in zip(cols, f[1:])})
            except ValueError:
                pass
    return rows, path


def valu_issue(workload: str, kernel: str, seconds: float):
    """VALU issue load of ``kernel``: its wave instructions per launch (SQ_INSTS_VALU from the
    committed counter summary) at 4 cycles each (a wave64 instruction on a 16-lane SIMD; FP64 FMA,
    MUL and ADD issue at that rate, transcendentals slower, some 32-bit operations faster) over the
    issue capacity of the measured duration: 1024 SIMDs at the 2.4 GHz peak engine clock.  None when
    no summary of that workload is present."""
    rows, path = _counter_rows(workload)
    if not rows or kernel not in rows or "VALU" not in rows[kernel] or seconds <= 0:
        return None
    v = rows[kernel]["VALU"]
    return {"source": os.path.relpath(path, ROOT), "valu_wave_insts": v, "cycles_per_inst": 4,
            "frac": v * 4.0 / (seconds * 1024 * 2.4e9)}


def host_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def _threads():
    """host threads for the CPU legs: the process's CPU affinity, at most 16 (one GPU's share)"""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, n))


LSUM_MAX = 5   # CPU-side samples: ordered pairs with l_A + l_B <= 5 only


def cpu_sample(sysm, rng_seed=1):
    """the fixed sample both CPU-side legs use: (molecule index, sorted ordered-pair indices) in a
    fixed random order; for one system, molecule index 0.  Only pairs with l_A + l_B <= LSUM_MAX
    are drawn: the reference spends minutes on one g-g pair of a six-centre molecule, which would
    make a bounded CPU leg unbounded.  The restriction favours the CPU (cheap pairs)."""
    rng = np.random.default_rng(rng_seed)
    mols = sysm if isinstance(sysm, list) else [sysm]
    order = rng.permutation(len(mols)) if isinstance(sysm, list) else np.zeros(64, int)
    for m in order:
        mol = mols[int(m)]
        ns = len(mol.shells)
        ok = [p for p in rng.permutation(ns * ns) if mol.shells[p // ns].L + mol.shells[p % ns].L <= LSUM_MAX]
        yield int(m), np.sort(np.array(ok[:8 if isinstance(sysm, list) else 16], dtype=np.int64))


def cpu_baseline(sysm, budget_s: float = 25.0):
    """Genuine libgrpp (compiled from its own sources into oracle/_ref) on a bounded sample of the
    same workload, OpenMP dynamic over ordered pairs on the host threads: chunks of the fixed
    sample are added (untimed warm-up) until a pass takes about budget/6, then that sample is timed
    5 times and the median reported.  Falls back to the oracle port when the reference build is
    absent."""
    import oracle
    lib, kind = oracle.reference(), "reference"
    if lib is None:
        lib, kind = oracle.oracle(), "port"
    cores = _threads()
    mols = sysm if isinstance(sysm, list) else [sysm]
    chunks, t_warm = [], 0.0
    for m, pairs in cpu_sample(sysm):
        t_warm += lib.matrices(mols[m], nthreads=cores, pairs=pairs)[4]
        chunks.append((m, pairs))
        if t_warm >= budget_s / 6:
            break
    reps = []
    for _ in range(5):
        reps.append(sum(lib.matrices(mols[m], nthreads=cores, pairs=p)[4] for m, p in chunks))
    t = float(np.median(reps))
    ntask = sum(len(p) * len(mols[m].centers) for m, p in chunks)
    out = {"value": ntask / t, "unit": "tasks/s", "cores": cores, "kind": kind,
           "sample": f"{sum(len(p) for _, p in chunks)} random ordered shell pairs x all centres "
                     f"({len(chunks)} {'molecules' if len(mols) > 1 else 'chunks'}) = {ntask} tasks; median of 5 "
                     f"passes after a warm-up pass, {t:.2f} s each",
           "repetitions_s": reps}
    out.update(host_info())
    return out, chunks


def percall_leg(sysm, chunks, batched, reps: int = 3):
    """The reference's own usage pattern through this library's drop-in API: host threads (as
    OpenMP threads in eval_integrals.c:78-123) each calling libgrpp_full_grpp_integrals for one
    (shell A, shell B, centre) task of the CPU sample, outputs summed per pair.  Median of ``reps``
    passes after a warm-up pass; max_abs_vs_batched compares the summed blocks with the batched build."""
    from concurrent.futures import ThreadPoolExecutor
    from libgrpp_amd import lib as G
    from libgrpp_amd.abi import as_dptr, darr
    L, fac = G.lib(), G.factory()
    mols = sysm if isinstance(sysm, list) else [sysm]
    tasks, objs = [], []
    for m, pairs in chunks:
        mol = mols[m]
        sh = [fac.shell(x) for x in mol.shells]
        gr = [(darr(c), fac.grpp(g)) for c, g in mol.centers]
        objs.append((sh, [h for _, h in gr]))
        ns = len(mol.shells)
        for p in pairs:
            i, j = divmod(int(p), ns)
            n = mol.shells[i].cart_size * mol.shells[j].cart_size
            for cc, h in gr:
                tasks.append((m, i, j, sh[i], sh[j], h, cc, [np.zeros(n) for _ in range(4)]))

    def run(t):
        L.libgrpp_full_grpp_integrals(t[3], t[4], t[5], t[6], *[as_dptr(x) for x in t[7]])

    threads = _threads()
    times = []
    with ThreadPoolExecutor(threads) as ex:
        for r in range(reps + 1):
            t0 = time.perf_counter()
            list(ex.map(run, tasks))
            if r:
                times.append(time.perf_counter() - t0)
    err = 0.0
    sums = {}
    for m, i, j, *_, outs in tasks:
        sums.setdefault((m, i, j), [0.0, 0.0, 0.0, 0.0])
        for k in range(4):
            sums[(m, i, j)][k] = sums[(m, i, j)][k] + outs[k]
    for (m, i, j), blk in sums.items():
        mol = mols[m]
        off = mol.offsets()
        A, B = mol.shells[i], mol.shells[j]
        M = batched[m] if isinstance(sysm, list) else batched
        for k in range(4):
            ref = M[k][off[i]:off[i] + A.cart_size, off[j]:off[j] + B.cart_size]
            err = max(err, float(np.abs(blk[k].reshape(A.cart_size, B.cart_size) - ref).max()))
    for sh, gr in objs:
        for x in sh:
            fac.f["delete_shell"](x)
        for h in gr:
            fac.free_grpp(h)
    t = float(np.median(times))
    return {"value": len(tasks) / t, "unit": "tasks/s", "threads": threads, "tasks": len(tasks),
            "seconds": t, "max_abs_vs_batched": err,
            "pattern": "host threads each calling libgrpp_full_grpp_integrals per (A, B, centre) task"}


def parity_fields(workload, sysm, out):
    """per-operator error of the build's matrices against the genuine reference's blocks for the
    sampled ordered pairs of tests/golden/<workload>_reference.npz (None when absent): max abs,
    max relative (|x| > 1), largest error / noise floor of the block among elements outside the
    parity rule (|d| <= 1e-10 or relative 1e-9 where |x| > 1), and the number of block operators
    outside the rule and beyond 3 floors"""
    name = {"ragged_molecules": "ragged", "heavy_cluster": "heavy_cluster"}.get(workload)
    path = os.path.join(ROOT, "tests", "golden", f"{name}_reference.npz") if name else ""
    if not os.path.exists(path):
        return None
    d = np.load(path)
    mols = sysm if isinstance(sysm, list) else [sysm]
    res = {}
    for k, nm in enumerate(["arep", "so_x", "so_y", "so_z"]):
        res[nm] = {"max_abs": 0.0, "max_rel": 0.0, "worst_error_over_floor": 0.0, "blocks_outside_rule": 0,
                   "blocks_beyond_3_floors": 0}
    for b, (m, p) in enumerate(zip(d["molecule"], d["pair"])):
        mol = mols[int(m)]
        M = out[int(m)] if isinstance(sysm, list) else out
        ns = len(mol.shells)
        i, j = divmod(int(p), ns)
        off = mol.offsets()
        o0, o1 = d["offset"][b], d["offset"][b + 1]
        n = (o1 - o0) // 4
        for k, nm in enumerate(["arep", "so_x", "so_y", "so_z"]):
            want = d["ref"][o0 + k * n:o0 + (k + 1) * n]
            got = M[k][off[i]:off[i] + mol.shells[i].cart_size, off[j]:off[j] + mol.shells[j].cart_size].ravel()
            dd = np.abs(got - want)
            big = np.abs(want) > 1.0
            rule = (dd <= 1e-10) | (big & (dd <= 1e-9 * np.abs(want)))
            fl = float(d["floor"][b][k])
            r = res[nm]
            r["max_abs"] = max(r["max_abs"], float(dd.max()))
            r["max_rel"] = max(r["max_rel"], float((dd[big] / np.abs(want[big])).max(initial=0.0)))
            if not rule.all():
                r["blocks_outside_rule"] += 1
                q = float((dd[~rule].max() - 1e-10) / max(fl, 1e-300))
                r["worst_error_over_floor"] = max(r["worst_error_over_floor"], q)
                if q > 3.0:
                    r["blocks_beyond_3_floors"] += 1
    return {"reference": os.path.relpath(path, ROOT), "sampled_blocks": int(len(d["pair"])),
            "molecules": int(len(np.unique(d["molecule"]))), "per_matrix": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ragged_molecules")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="time the plain method only")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shapes screen_probe,contract_threads,type1_threads[,stream_layout[,screen_split]] "
                         "(defaults 0 and 4) instead of autotuning")
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--no-percall", action="store_true", help="skip the per-call drop-in leg")
    ap.add_argument("--graph", action="store_true",
                    help="also time the plain build captured once as a hipGraph and replayed (graph_leg)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device-to-device copy (3 times): the known-size kernel "
                         "against which the FETCH_SIZE / WRITE_SIZE counters are scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: --gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)

    import torch  # device plumbing only: synchronisation around the timed region
    if args.copy_calibration:
        x = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": x.numel() * 8}))
        return
    from libgrpp_amd import configs, lib as G

    from libgrpp_amd import learned as LG
    sysm = configs.WORKLOADS[args.workload]()
    molecules = isinstance(sysm, list)
    phases = ["tables_seconds", "screen_seconds", "radial_seconds", "canon_seconds", "proj_seconds",
              "type1_seconds", "contract_seconds", "device_seconds"]

    def timed(plan):
        """warm-up, then exactly --steps timed builds bracketed by device synchronisation"""
        for _ in range(args.warmup):
            plan.execute()
        torch.cuda.synchronize()
        times = {k: [] for k in phases}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute()
            st = plan.stats()   # per-phase HIP-event times of this step, on the plan's stream
            for k in phases:
                times[k].append(st[k])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, plan.stats(), {k: float(np.mean(v)) for k, v in times.items()}

    def tune(plan):
        """the plan's launch shapes picked by timing its own builds (untimed; results unchanged)"""
        if args.tuning:
            v = [int(x) for x in args.tuning.split(",")]
            probe, ct, t1, lay, split = (v + [0, 4][len(v) - 3:])[:5]
            plan.set_tuning(probe, ct, t1, lay, split)
            return {"screen_probe": probe, "contract_threads": ct, "type1_threads": t1, "stream_layout": lay,
                    "screen_split": split, "fixed": True}
        if args.no_tune:
            return None
        plan.execute()
        reps = 2 if plan.stats()["device_seconds"] < 0.2 else 1
        return plan.autotune(repeats=reps)

    # plain adaptive method
    plan = LG.make_plan(sysm, 0)
    tuning_plain = tune(plan)
    wall_plain, st_plain, mean_plain = timed(plan)
    plain_out = plan.download()
    plan.close()
    graph_info = None
    if args.graph:
        # the same plain build captured into one hipGraph on its first execute, then replayed
        gplan = LG.make_plan(sysm, G.BATCH_GRAPH)
        if tuning_plain:
            gplan.set_tuning(tuning_plain["screen_probe"], tuning_plain["contract_threads"],
                             tuning_plain["type1_threads"], tuning_plain["stream_layout"],
                             tuning_plain["screen_split"])
        for _ in range(max(1, args.warmup)):
            gplan.execute()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            gplan.execute()
        torch.cuda.synchronize()
        gwall = time.perf_counter() - t0
        gout = gplan.download()
        gplan.close()
        flat = lambda ms: [m for x in ms for m in (x if isinstance(x, list) else [x])]
        graph_info = {"graph_ms_per_step": 1e3 * gwall / args.steps,
                      "eager_ms_per_step": 1e3 * wall_plain / args.steps,
                      "bitwise_equal_to_eager": all(np.array_equal(x, y)
                                                    for x, y in zip(flat(gout), flat(plain_out)))}
    # learned grid levels: the predictor is trained (untimed) on heavy_cluster seeds 1-4, then its
    # BF16 inference orders the radial tables of the seed-0 workload
    learned_info = None
    wall, st, mean = wall_plain, st_plain, mean_plain
    if not args.no_learned:
        t0 = time.perf_counter()
        X, T = LG.training_set(LG.default_training_systems(4, args.workload))
        net, loss = LG.train_predictor(X, T, epochs=30)
        t_train = time.perf_counter() - t0
        plan = LG.make_plan(sysm, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        LG.apply(plan, net, bf16=True)
        torch.cuda.synchronize()
        t_pred = time.perf_counter() - t0
        # the plain plan's tuned launch shapes, so that the two timings differ only in the row order
        tuning = tuning_plain
        if tuning_plain:
            plan.set_tuning(tuning_plain["screen_probe"], tuning_plain["contract_threads"],
                            tuning_plain["type1_threads"], tuning_plain["stream_layout"],
                            tuning_plain["screen_split"])
        wall, st, mean = timed(plan)
        out = plan.download()
        plan.close()
        net.close()
        if molecules:
            diff = max(float(np.abs(a - b).max()) for ma, mb in zip(out, plain_out) for a, b in zip(ma, mb))
        else:
            diff = max(float(np.abs(a - b).max()) for a, b in zip(out, plain_out))
        learned_info = {"plain_tasks_per_s": st_plain["n_tasks"] / (wall_plain / args.steps),
                        "plain_ms_per_step": 1e3 * wall_plain / args.steps,
                        "speedup_vs_plain": wall_plain / wall,
                        "radial_points_plain": st_plain["n_radial_points"], "radial_points_learned": st["n_radial_points"],
                        "radial_ms_plain": 1e3 * mean_plain["radial_seconds"],
                        "radial_ms_learned": 1e3 * mean["radial_seconds"],
                        "max_abs_diff_vs_plain": diff, "training_rows": int(X.shape[0]),
                        "training_loss": loss, "training_s": t_train, "inference_and_reorder_ms": 1e3 * t_pred,
                        "launch_shapes": "the plain plan's tuned shapes, used for both plans"}

    ntask = st["n_tasks"]
    nint = st["n_integrals"]
    assert ntask == (sum(configs.n_tasks(m) for m in sysm) if molecules else configs.n_tasks(sysm))
    ms = 1e3 * wall / args.steps
    tasks_per_s = ntask / (wall / args.steps)
    scr = mean["screen_seconds"]
    achieved = st["screen_flops"] / scr / 1e12
    rad = mean["radial_seconds"]
    rad_achieved = st["fp64_flops"] / rad / 1e12
    line = {
        "metric": ("FP64 GRPP shell-pair x centre tasks/s: learned-grid vs plain GPU (full GRPP matrix build)"
                   if learned_info else "FP64 GRPP shell-pair x centre tasks/s: plain GPU (full GRPP matrix build)"),
        "value": tasks_per_s,
        "unit": "tasks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": workload_config(args.workload, sysm, ntask, nint),
        "integrals_per_s": nint / (wall / args.steps),
        "device_ms_per_step": 1e3 * mean["device_seconds"],
        "kernel_ms": {k.replace("_seconds", ""): 1e3 * v for k, v in mean.items() if k != "device_seconds"},
        # dominant kernel: the Shaw-Hill envelope estimates (FP64 VALU; FLOP equivalents per triple and
        # per fixed-point evaluation as counted by the engine)
        "roofline": {"bound": "valu", "kernel": "k_screen", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic_bytes("k_screen", args.workload),
                     "screen_flops": st["screen_flops"], "screen_triples": st["screen_triples"],
                     "screen_steps": st["screen_steps"], "screen_ms": 1e3 * scr,
                     "radial_product": {"kernel": "k_radial_mfma", "bound": "mfma", "achieved": rad_achieved,
                                        "frac": rad_achieved / FP64_PEAK_TFLOPS, "ms": 1e3 * rad,
                                        "flops": st["fp64_flops"], "points": st["n_radial_points"],
                                        "tiles": st["n_tiles"], "traffic": traffic_bytes("k_radial_mfma", args.workload)}},
    }
    if learned_info:
        line["learned_grid"] = learned_info
    final = out if learned_info else plain_out
    line["parity"] = parity_fields(args.workload, sysm, final)
    line["fp64_counters"] = fp64_counters(args.workload, ms / 1e3)
    line["roofline"]["valu_issue"] = valu_issue(args.workload, "k_screen", scr)
    line["autotune"] = tuning_plain
    if graph_info:
        line["graph_leg"] = graph_info
    if not args.no_cpu:
        line["cpu_baseline"], _ = cpu_baseline(sysm, args.cpu_budget)
    if not args.no_percall:
        chunks, ntask = [], 0
        mols = sysm if molecules else [sysm]
        for m, pairs in cpu_sample(sysm):
            chunks.append((m, pairs))
            ntask += len(pairs) * len(mols[m].centers)
            if ntask >= 1024:
                break
        line["percall"] = percall_leg(sysm, chunks, final)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
