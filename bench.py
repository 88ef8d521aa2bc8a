"""Benchmark: transient time steps of the 65k-gate CMOS DAG on one MI355X.

Workload (one "step" = one accepted transient time step, each running the full
Newton loop: device eval, stamping, sparse LU refactorisation, triangular
solves): the seeded gate DAG with 65 000 two-input NAND/NOR gates (260 000
MOSFETs, ~130 000 unknowns), adaptive BDF2 from the power-up initial state.
One run takes ``--warmup W`` accepted steps untimed -- through the 100 ps
supply ramp into the switching of the primary inputs -- and the next
``--steps K`` accepted steps are timed (the driver's measurement window,
between stream synchronisations).  A short run before it performs the
one-time symbolic analysis and the launch-plan autotuner, which are excluded
like the reference excludes JIT and setup and reported beside the result.

Prints one JSON line with ms/step and Newton iterations/s (``value``) over the
timed window, the same window in mixed precision with its waveform error
against the FP64 run on the FP64 run's time grid, a roofline estimate for the
dominant kernel class from HIP events recorded on the library's own stream in
a second, eager (non-graph) pass over the same steps, and a CPU baseline: the
C++ restatement of the reference algorithm (oracle/ctran.cpp) timed on the
same circuit's first steps on all host threads and on one.  Per step the two
sides run different windows (different Newton iteration counts), so the
speed-up is stated per Newton iteration as well.

``--summarize-trace kernel_trace.csv`` instead prints a markdown table of the
kernels a rocprofv3 kernel trace holds between the two window markers
(k_window_mark) of the timed run; ``--summarize-pmc`` turns FETCH_SIZE and
WRITE_SIZE counter passes into the HBM bytes per dispatch the roofline's
``traffic`` reads.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("ms/timestep + Newton iters/s at ~150k unknowns; batched instance·steps/s")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_circuit(gates: int, seed: int):
    from vajax_amd import generators as G

    return G.gate_dag(gates, seed=seed)


def run_gpu(eng, steps: int, precision: str, use_graph: bool, band_width=None, autotune=True,
            nd_leaf=None, window_from=0, replay=None, t_stop=1e-9):
    from vajax_amd.engine import AdaptiveConfig

    eng.prepare(t_stop=t_stop, dt=1e-12, adaptive_config=AdaptiveConfig(integration_method="gear2"),
                precision=precision, max_steps=steps + 1, use_graph=use_graph,
                band_width=band_width, autotune=autotune, nd_leaf=nd_leaf,
                window_from_step=window_from, replay_times=replay)
    return eng.run_transient()


def window_stats(res):
    st = res.stats
    k = max(st["window_steps"], 1)
    el = st["window_time"]
    return {"accepted_steps": st["window_steps"], "rejected_attempts": st["window_rejected"],
            "newton_iterations": st["window_nr_iterations"],
            "newton_iters_per_step": round(st["window_nr_iterations"] / k, 3),
            "ms_per_step": round(el / k * 1e3, 4),
            "ms_per_newton_iter": round(el / max(st["window_nr_iterations"], 1) * 1e3, 4),
            "newton_iters_per_s": round(st["window_nr_iterations"] / el, 3) if el > 0 else None,
            "loop_s": round(el, 6),
            "t_window_ps": [round(float(res.times[-1 - st["window_steps"]]) * 1e12, 3),
                            round(float(res.times[-1]) * 1e12, 3)]}


# model parameter columns read by k_eval (the MOSFETs' include their init cache)
N_PAR = {"resistor": 1, "capacitor": 1, "mosfet": 11, "ekv26": 52}
# FP64 VALU operations of one EKV 2.6 instance eval (value + three partials),
# from rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 over the eval
# kernel of the 65k-gate EKV 2.6 DAG (64 lanes per instruction, FMA = 2),
# divided by its instances per launch (profiles/eval_valu_ekv26_gate_dag_65000.md)
EKV26_F64_OPS_PER_INSTANCE = 5595
FP64_VALU_PEAK_TFLOPS = 78.6


def mixed_error(eng, ref_res, use_graph):
    """Largest node-voltage difference between the mixed (FP32 factor + FP64
    refinement) solution and the plain FP64 run over every step of the FP64
    run's time grid (the mixed run replays that grid, so the waveforms
    compare point by point)."""
    import numpy as np

    times = ref_res.times
    res = run_gpu(eng, len(times) - 1, "f32", use_graph, replay=times)
    dv = np.abs(res.stats["V_out"] - ref_res.stats["V_out"])
    return {"max_abs_dV_vs_fp64": float(np.max(dv)), "steps": len(times) - 1,
            "nonconverged_steps": int(res.stats["rejected_steps"]),
            "grid": "the FP64 run's accepted grid (gear2 adaptive), replayed"}


def kernel_bytes(topo, info, precision: str):
    """Algorithmic HBM bytes per Newton iteration of each kernel class (the
    data each kernel must move at least once: indices, gathered operands,
    results), from the circuit topology and the LU analysis."""
    vb = 4 if precision == "f32" else 8
    n = topo.n_augmented
    nse = topo.nse
    ev = 0
    n_jc = 0
    n_rc = 0
    for name, ms in topo.models.items():
        N = ms.count
        n_v = ms.v_node1.shape[1]
        jv = int((ms.csr_pos >= 0).sum())
        rv = int((ms.res_idx >= 0).sum())
        n_jc += jv
        n_rc += rv
        # params, node indices + voltage gathers, destination indices, writes
        ev += N * (8 * N_PAR[name] + n_v * (2 * 4 + 2 * 8) + 4 * (ms.csr_pos.shape[1] + ms.res_idx.shape[1]))
        ev += 8 * jv + 16 * rv
    # stamping: contributions + segment pointers + per-slot constants + csr write;
    # residual: rf/rq + segments + history vectors + f/Q/mrc writes
    stamp = 8 * n_jc + 4 * (nse + 1) + 12 * nse + 8 * nse + 16 * n_rc + 4 * n + 8 * 8 * n
    nnz_lu = info["nnz_l"] + info["nnz_u"]
    # LU band tasks (run from LDS): each item's 16-byte record and each
    # update's 8-byte record (packed local L / U indices, pivot index) are
    # streamed once, and the task's slots -- targets, operands and its
    # pivots' L columns, scaled in LDS -- are gathered (index + value) and
    # the written ones stored back: at most the LU's entries once each.
    # Schur steps: the L and U slots of each update (L already scaled) once,
    # per target the item, slot, segment pair and the value's
    # read-modify-write.  Operand values gathered many times are counted once.
    lu_tasks = (16 * info["fact_task_targets"] + 8 * info["fact_task_updates"]
                + (4 + 2 * vb) * nnz_lu)
    lu_schur = 8 * info["fact_schur_updates"] + (4 + 4 + 8 + 2 * vb) * info["fact_schur_targets"]
    lu_gather = (4 + 8 + vb) * nnz_lu
    lu_scale = 0  # folded into the tasks
    offd = nnz_lu - n            # off-diagonal entries, each read once per solve
    # solves: each off-diagonal entry's column index and value once (the y
    # operands it gathers are counted once per row), per row the item, row
    # pointers, the right-hand side / result and the pivot
    solves = (4 + vb) * offd + (4 + 8 + 3 * vb) * 2 * n
    if precision == "f32":       # second solve of the refinement step + FP64 SpMV
        solves = 2 * solves
    # dense trailing block: its slots gathered and written back, the
    # contributing L / U slots gathered, and D (K x K FP64) read and written
    # once per panel step it takes part in (counted once)
    K, m = info.get("dense_k", 0), info.get("dense_m", 0)
    lu_dense = 8 * (3 * K * K + 2 * K * m) if K else 0
    return {"eval": ev, "stamp": stamp, "lu_gather": lu_gather, "lu_tasks": lu_tasks,
            "lu_schur": lu_schur, "lu_scale": lu_scale, "solve_tasks": solves, "solve_hub": 0,
            "lu_dense": lu_dense}


KERNEL_NAMES = {"eval": "k_eval<MOSFET> (+R/C)", "stamp": "k_chunk_sum/k_reduce_jac/k_chunk_res/k_residual",
                "lu_gather": "k_lu_gather", "lu_tasks": "k_lu_band_lds (band tasks)",
                "lu_schur": "k_lu_wide (Schur step)",
                "lu_scale": "k_lu_scale (level programs only)", "solve_tasks": "k_tri_band_lde (forward + backward)",
                "solve_hub": "k_tri_hub_part/k_tri_hub_finish",
                "lu_dense": "k_dgemm_sub/k_dense_panel (dense trailing block, FP64 MFMA)"}


# kernel class -> kernel-name prefixes in the rocprofv3 traffic summaries
TRAFFIC_KERNELS = {
    "eval": ["k_eval<"], "stamp": ["k_reduce_jac", "k_residual", "k_chunk_sum", "k_chunk_res"],
    "lu_gather": ["k_lu_gather<double>"], "lu_tasks": ["k_lu_band_lds<double", "k_lu_band<double>"],
    "lu_schur": ["k_lu_wide<double"], "lu_scale": ["k_lu_scale<double>"],
    "solve_tasks": ["k_tri_band_lde<double", "k_tri_band<double"],
    "lu_dense": ["k_dgemm_sub", "k_dense_panel", "k_dense_diag"]}


def measured_traffic(args, cls):
    """HBM bytes per dispatch of a kernel class (dispatch-weighted mean over
    its kernels) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    of this workload, scaled by the copy-kernel calibration; None when no
    summary covers it."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_gate_dag_{args.gates}.json")
    if args.precision != "f64" or args.seed != 0 or not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)["kernels"]
    ks = [v for k, v in data.items() if any(k.startswith(p) for p in TRAFFIC_KERNELS.get(cls, []))]
    n = sum(k["dispatches"] for k in ks)
    if n == 0:
        return None
    return int(sum(k["traffic_bytes"] * k["dispatches"] for k in ks) / n)


def _kernel_name(raw):
    """Kernel name without the return type, the namespace and the argument
    list."""
    name = raw.replace("vjx::(anonymous namespace)::", "").replace("vjx::", "").split("(")[0].strip()
    return name[5:] if name.startswith("void ") else name


def _counter_means(path):
    """kernel -> (dispatches, mean counter value per dispatch) of a rocprofv3
    --pmc CSV with one counter."""
    import csv

    per = {}
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            d = per.setdefault(_kernel_name(r["Kernel_Name"]), {})
            d[r["Dispatch_Id"]] = d.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    return {k: (len(v), sum(v.values()) / len(v)) for k, v in per.items()}


VALU_F64 = ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64",
            "SQ_INSTS_VALU_TRANS_F64", "SQ_INSTS_VALU_FLOPS_FP64", "SQ_INSTS_VALU_FLOPS_FP64_TRANS")


def summarize_valu(path, kernel, grid):
    """FP64 VALU work of one kernel from a rocprofv3 --pmc CSV holding the
    VALU_F64 counters: the mean per dispatch over the dispatches of
    ``kernel`` with ``grid`` work-items, and the operations per work-item
    (64 lanes per instruction, an FMA counting two)."""
    import csv

    per = {}
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            if _kernel_name(r["Kernel_Name"]) != kernel or int(r["Grid_Size"]) != grid:
                continue
            d = per.setdefault(r["Counter_Name"], {})
            d[r["Dispatch_Id"]] = d.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    mean = {c: sum(v.values()) / len(v) for c, v in per.items()}
    n = {c: len(v) for c, v in per.items()}
    ops = 64.0 * (mean["SQ_INSTS_VALU_ADD_F64"] + mean["SQ_INSTS_VALU_MUL_F64"]
                  + 2.0 * mean["SQ_INSTS_VALU_FMA_F64"] + mean["SQ_INSTS_VALU_TRANS_F64"])
    print(f"* kernel `{kernel}`, grid {grid} work-items, {min(n.values())} dispatches\n")
    print("| counter | mean per dispatch | per work-item |")
    print("|---|---:|---:|")
    for c in VALU_F64:
        if c in mean:
            print(f"| {c} | {mean[c]:.4g} | {mean[c] / grid:.2f} |")
    print(f"\nFP64 operations per work-item, 64 (ADD + MUL + 2 FMA + TRANS) / grid: "
          f"{ops / grid:.1f}")


def summarize_pmc(paths, json_out):
    """HBM bytes per dispatch from FETCH_SIZE and WRITE_SIZE passes (KiB per
    dispatch), each scaled by what the same counter reads for the 1 GiB
    device-to-device copies of ``--copy-calibration``; writes the JSON that
    the roofline's ``traffic`` reads and prints a markdown table."""
    fetch, write, cal_f, cal_w = (_counter_means(p) for p in paths)
    copy = "__amd_rocclr_copyBuffer"
    gib_kib = float(1 << 20)
    ff = gib_kib / cal_f[copy][1]
    wf = gib_kib / cal_w[copy][1]
    out = {"fetch_factor": ff, "write_factor": wf, "kernels": {}}
    for k in sorted(set(fetch) | set(write), key=lambda k: -fetch.get(k, (0, 0.0))[0]):
        nf, f = fetch.get(k, (0, 0.0))
        nw, w = write.get(k, (0, 0.0))
        out["kernels"][k] = {"dispatches": max(nf, nw), "fetch_kib": f, "write_kib": w,
                             "traffic_bytes": (f * ff + w * wf) * 1024.0}
    with open(json_out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"* FETCH_SIZE of the copy: {cal_f[copy][1]:.1f} KiB per 1 GiB read -> factor {ff:.4f}")
    print(f"* WRITE_SIZE of the copy: {cal_w[copy][1]:.1f} KiB per 1 GiB written -> factor "
          f"{wf:.4f}\n")
    print("| kernel | dispatches | FETCH KiB | WRITE KiB | traffic MB/dispatch (scaled) |")
    print("|---|---:|---:|---:|---:|")
    for k, v in out["kernels"].items():
        print(f"| `{k}` | {v['dispatches']} | {v['fetch_kib']:.1f} | {v['write_kib']:.1f} | "
              f"{v['traffic_bytes'] / 1e6:.2f} |")


def host_threads() -> int:
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(env)) if env and env.isdigit() and int(env) > 0 else n


def cpu_baseline(args):
    """The C++ CPU restatement of the same algorithm (OpenMP over device
    instances, serial KLU-style sparse LU) on the same circuit and the same
    first steps, on all host threads and on one, in a child process so its
    OpenMP runtime stays apart from this one's."""
    import subprocess

    threads = host_threads()
    cmd = [sys.executable, "-m", "oracle.cpu_bench", "--gates", str(args.gates), "--seed",
           str(args.seed), "--steps", str(args.cpu_steps), "--threads", str(threads), "1"]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed ({r.returncode}): {r.stderr[-2000:]}")
    data = json.loads(r.stdout.strip().splitlines()[-1])
    full, one = data["runs"][0], data["runs"][1]
    out = {
        "value": round(full["newton_iters_per_s"], 3),
        "unit": "Newton iters/s",
        "ms_per_step": round(full["ms_per_step"], 3),
        "ms_per_newton_iter": round(full["wall_s"] / max(full["newton_iterations"], 1) * 1e3, 4),
        "cores": full["threads"],
        "kind": "port",
        "sample": (f"C++ restatement of the reference algorithm (oracle/ctran.cpp: OpenMP device "
                   f"eval and stamping, serial Gilbert-Peierls LU with threshold pivoting 0.1 "
                   f"then refactorisation on the pivot order), same {data['workload']} circuit, "
                   f"first {full['accepted_steps']} accepted steps ({full['newton_iterations']} "
                   f"Newton iterations) in {full['wall_s']:.2f} s on {full['threads']} host "
                   f"threads; setup and ordering excluded, first factorisation included"),
        "phase_s": {k: round(v, 4) for k, v in full["phase_s"].items()},
        "single_thread": {"value": round(one["newton_iters_per_s"], 3),
                          "ms_per_step": round(one["ms_per_step"], 3),
                          "phase_s": {k: round(v, 4) for k, v in one["phase_s"].items()}},
    }
    if args.cpu_numpy:
        out["python_oracle"] = numpy_oracle_baseline(args)
    return out


def numpy_oracle_baseline(args):
    """Secondary figure: the numpy/SciPy-SuperLU oracle on one thread."""
    from threadpoolctl import threadpool_limits

    from oracle import mna
    from oracle import transient as T

    net = build_circuit(args.gates, args.seed)
    steps = min(args.cpu_steps, 3)
    with threadpool_limits(limits=1):
        cir = mna.setup(net)
        cfg = T.AdaptiveConfig(integration_method="gear2")
        out = T.run_adaptive(cir, 1e-9, 1e-12, cfg, T.Options(icmode="uic"), max_steps=steps + 1)
    st = out.stats
    acc = max(st["n_steps"] - 1, 1)
    return {"value": round(st["total_nr_iterations"] / st["wall_time"], 3),
            "ms_per_step": round(st["wall_time"] / acc * 1e3, 3), "cores": 1,
            "sample": f"first {acc} accepted steps, numpy/SciPy oracle, one thread"}


def run_batched(args, model="mosfet"):
    """Batched throughput: Monte Carlo instances of a 2k-gate DAG plus ragged
    topologies packed into one block-diagonal system (shared analysis per
    distinct block pattern, one launch per level for every instance), all
    instances advancing on the packed system's common time grid."""
    import torch

    from vajax_amd import generators as G
    from vajax_amd.engine import CircuitEngine

    t0 = time.time()
    ragged = tuple(int(g) for g in args.batch_ragged.split(",")) if args.batch_ragged else ()
    net = G.batch_mc(args.batch, args.batch_gates, ragged=ragged, seed=args.seed, model=model)
    n_inst = args.batch + len(ragged)
    eng = CircuitEngine(net)
    eng.parse()
    topo = eng.topology
    log(f"batch setup {time.time() - t0:.1f}s: {n_inst} instances, "
        f"n_unknowns={topo.n_augmented} nse={topo.nse}")
    out = {"workload": net.title, "instances": n_inst, "unknowns": topo.n_augmented,
           "mosfets": int(net.groups[model].count), "unit": "instance·steps/s"}
    wb = args.warmup
    for prec in ("f32", "f64"):
        t0 = time.time()
        run_gpu(eng, 3, prec, True)  # analysis, launch-plan autotuner, graph capture
        log(f"batch set-up run {prec} {time.time() - t0:.1f}s")
        # the same measurement window as the single circuit: untimed steps
        # through the supply ramp, then the timed accepted steps
        res = run_gpu(eng, wb + args.batch_steps, prec, True, window_from=wb + 1)
        w = window_stats(res)
        out[prec] = {"value": round(n_inst * w["accepted_steps"] / w["loop_s"], 3), **w}
        log(f"batch {prec}: {w['accepted_steps']} steps in {w['loop_s']:.3f}s")
    # the plain form: one instance at a time, FP64
    single = CircuitEngine(G.gate_dag(args.batch_gates, seed=args.seed, model=model))
    single.parse()
    run_gpu(single, 3, "f64", True)
    res = run_gpu(single, wb + args.batch_steps, "f64", True, window_from=wb + 1)
    w = window_stats(res)
    out["single_f64"] = {"value": round(w["accepted_steps"] / w["loop_s"], 3), **w,
                         "workload": f"gate_dag_{args.batch_gates}_s{args.seed}"}
    info = eng.device.lu.info()
    out["lu"] = {k: info[k] for k in ("n_blocks", "n_block_patterns", "nnz_l", "nnz_u", "n_updates",
                                      "fact_levels")}
    out["value"] = out["f32"]["value"]
    out["precision"] = "f32-lu+f64"
    out["mixed_vs_fp64_batched"] = round(out["f32"]["value"] / max(out["f64"]["value"], 1e-30), 3)
    out["mixed_batched_vs_fp64_single"] = round(out["f32"]["value"] / max(out["single_f64"]["value"], 1e-30), 3)
    out["time_grid"] = "common to all instances (dt = smallest LTE step over the batch)"
    return out


MFMA_F64_PEAK_TFLOPS = 78.6
MUL64_REFERENCE_MS_PER_STEP = 648.0  # JAX GPU, cuDSS f32 + refinement, Tesla T4, PSP103


# the reference's published small-circuit step times (JAX GPU path, its CI GPU
# runner; README.md:88-89): a 47-node PSP103 ring and a 4-node RC circuit
REFERENCE_SMALL_MS_PER_STEP = {"ring": 1.49, "rc": 0.24}


def run_small(args):
    """The two small configurations, whose step time is launch- and
    round-trip-bound: the 200-section RC ladder (fixed-step BDF2, 10 000
    steps, the configuration's own length) and the 101-stage CMOS ring
    (adaptive Gear2, 2 000 timed steps after 200), one HIP-graph Newton
    iteration per launch."""
    from vajax_amd import generators as G
    from vajax_amd.engine import AdaptiveConfig, CircuitEngine

    out = {}
    eng = CircuitEngine(G.rc_ladder(200, seed=args.seed))
    eng.parse()
    eng.options.fixed_step_transient = True
    cfg = AdaptiveConfig(integration_method="gear2")
    cfg.tran_fs = 1.0
    eng.prepare(t_stop=10e-9, dt=1e-12, adaptive_config=cfg, use_graph=True, max_steps=10002)
    eng.run_transient()
    eng.prepare(t_stop=10e-9, dt=1e-12, adaptive_config=cfg, use_graph=True, max_steps=10002,
                window_from_step=1)
    out["rc_ladder_200"] = {"unknowns": eng.topology.n_augmented,
                            "integration": "gear2 fixed 1 ps", **window_stats(eng.run_transient())}
    e = CircuitEngine(G.ring_oscillator(101, seed=args.seed))
    e.parse()
    run_gpu(e, 3, "f64", True, t_stop=50e-9)
    r = run_gpu(e, 2200, "f64", True, window_from=201, t_stop=50e-9)
    out["ring_101"] = {"unknowns": e.topology.n_augmented, "integration": "gear2 adaptive",
                       **window_stats(r)}
    out["reference_ms_per_step"] = REFERENCE_SMALL_MS_PER_STEP
    out["reference_note"] = ("JAX GPU path on the reference's CI GPU runner: a 47-node PSP103 "
                             "ring and a 4-node RC circuit -- other circuits, model and hardware")
    for k in ("rc_ladder_200", "ring_101"):
        log(f"small {k}: {out[k]['ms_per_step']} ms/step")
    return out


def run_mul64(args):
    """The reference's own benchmark circuit, mul64
    (vajax/benchmarks/data/mul64/runme.sim: the 64x64 array multiplier of its
    generate block, every input driven high through 1 ohm, a uic transient to
    2 ns with 2 ps steps and tran_lteratio = 1.5), built by this build's
    generator of the same topology, with the EKV substitute for PSP103 (no
    internal nodes: 133 592 unknowns where the reference's PSP103 circuit has
    666 409).  A set-up run (analysis, autotuner, graph), then the whole
    transient timed on the step loop's clock, then a short eager pass for the
    kernel split, including the dense trailing block on the FP64 matrix
    cores."""
    from vajax_amd import _lib as L
    from vajax_amd import generators as G
    from vajax_amd import vacask
    from vajax_amd.engine import CircuitEngine

    t0 = time.time()
    net = vacask.load("", text=G.array_multiplier_testbench(64))
    eng = CircuitEngine(net)
    eng.parse()
    topo = eng.topology
    log(f"mul64 setup {time.time() - t0:.1f}s: n_unknowns={topo.n_augmented} nse={topo.nse}")

    def run(use_graph, max_steps=None, window_from=0):
        eng.prepare(t_stop=2e-9, dt=2e-12, precision="f64", use_graph=use_graph, autotune=True,
                    max_steps=max_steps, window_from_step=window_from)
        return eng.run_transient()

    t0 = time.time()
    run(True, max_steps=3)
    setup_s = time.time() - t0
    log(f"mul64 analysis + autotune {setup_s:.1f}s")
    res = run(True, window_from=1)
    w = window_stats(res)
    log(f"mul64: {w['accepted_steps']} steps in {w['loop_s']:.2f}s")
    L.lib.vjx_profile_enable(1)
    L.lib.vjx_profile_reset()
    run(False, max_steps=60)
    pt = L.PhaseTimes()
    L.lib.vjx_profile_read(C.byref(pt))
    L.lib.vjx_profile_enable(0)
    ph = pt.as_dict()
    iters = max(ph["n_iterations"], 1)
    info = eng.device.lu.info()
    out = {"workload": "mul64 (64x64 array multiplier, reference benchmark topology)",
           "mosfets": int(net.groups["mosfet"].count), "unknowns": topo.n_augmented,
           "csr_nnz": topo.nse, "t_stop_ns": 2.0, "dt_ps": 2.0, **w,
           "end_time_ns": round(float(res.times[-1]) * 1e9, 4), "analysis_s": round(setup_s, 2),
           "lu": {k: info[k] for k in ("nnz_l", "nnz_u", "n_updates", "fact_levels", "dense_k",
                                       "dense_m")},
           "kernel_ms_per_newton_iter": {name: round(kt["ms"] / iters, 4)
                                         for name, kt in ph["kernels"].items() if kt["launches"]},
           "reference_ms_per_step": MUL64_REFERENCE_MS_PER_STEP,
           "reference_note": "JAX GPU (cuDSS FP32 factor + FP64 refinement) on a Tesla T4 with "
                             "PSP103 (666 409 unknowns); this run: EKV substitute, 133 592 "
                             "unknowns, MI355X -- different model and hardware"}
    # the same circuit with every MOSFET the EKV 2.6 model (charges, junctions,
    # short-channel effects): the whole transient at a PSP-class eval cost
    eng_e = CircuitEngine(G.with_ekv26(net))
    eng_e.parse()
    eng_e.prepare(t_stop=2e-9, dt=2e-12, precision="f64", use_graph=True, autotune=True,
                  max_steps=3)
    eng_e.run_transient()
    eng_e.prepare(t_stop=2e-9, dt=2e-12, precision="f64", use_graph=True, autotune=True,
                  window_from_step=1)
    we = window_stats(eng_e.run_transient())
    out["ekv26"] = {"workload": "mul64 with EKV 2.6 MOSFETs", **we}
    log(f"mul64 ekv26: {we['accepted_steps']} steps, {we['ms_per_step']} ms/step")
    del eng_e
    dense = ph["kernels"].get("lu_dense")
    if dense and dense["launches"]:
        ms = dense["ms"] / iters  # one refactorisation per Newton iteration
        tf = info["dense_flops"] / (ms * 1e-3) / 1e12
        out["dense_block"] = {
            "kernels": KERNEL_NAMES["lu_dense"], "k": info["dense_k"], "m": info["dense_m"],
            "flops_per_refactor": info["dense_flops"], "ms_per_refactor": round(ms, 4),
            "launches_per_refactor": round(dense["launches"] / iters, 1),
            "achieved_TFLOPs": round(tf, 4), "peak_TFLOPs": MFMA_F64_PEAK_TFLOPS,
            "frac": round(tf / MFMA_F64_PEAK_TFLOPS, 6)}
    return out


def run_ekv26(args):
    """The 65k-gate DAG with every MOSFET the EKV 2.6 model of
    vajax/devices/models/ekv/ekv.va (charges, junctions, short-channel
    effects; 52-column instance cache, dual-number Jacobian): the step loop
    at a PSP-class eval cost, timed like the main leg, then an eager pass for
    the eval kernel's time, FP64 rate and HBM fraction."""
    from vajax_amd import _lib as L
    from vajax_amd import generators as G
    from vajax_amd.engine import CircuitEngine

    t0 = time.time()
    net = G.gate_dag(args.gates, seed=args.seed, model="ekv26")
    eng = CircuitEngine(net)
    eng.parse()
    topo = eng.topology
    log(f"ekv26 setup {time.time() - t0:.1f}s: n_unknowns={topo.n_augmented}")
    run_gpu(eng, 3, "f64", True)
    total = args.warmup + args.steps
    res = run_gpu(eng, total, "f64", True, window_from=args.warmup + 1)
    w = window_stats(res)
    log(f"ekv26: {w['ms_per_step']} ms/step, {w['newton_iters_per_s']} Newton iters/s")
    mixed = None
    if not args.no_mixed:
        # FP32 LU + one FP64 refinement over the same window, and on the FP64
        # run's grid for the voltage difference
        r32 = run_gpu(eng, total, "f32", True, window_from=args.warmup + 1)
        mixed = {"window": window_stats(r32), **mixed_error(eng, res, True)}
        mixed["mixed_vs_fp64_ms_per_step"] = round(
            w["ms_per_step"] / mixed["window"]["ms_per_step"], 3)
        log(f"ekv26 mixed: {mixed['window']['ms_per_step']} ms/step, max |dV| vs FP64 "
            f"{mixed['max_abs_dV_vs_fp64']:.2e} V")
    L.lib.vjx_profile_enable(1)
    L.lib.vjx_profile_reset()
    run_gpu(eng, min(total, 60), "f64", False)
    pt = L.PhaseTimes()
    L.lib.vjx_profile_read(C.byref(pt))
    L.lib.vjx_profile_enable(0)
    ph = pt.as_dict()
    iters = max(ph["n_iterations"], 1)
    n_dev = int(net.groups["ekv26"].count)
    ev = ph["kernels"]["eval"]
    avg_ms = ev["ms"] / max(ev["launches"], 1)
    alg = kernel_bytes(topo, eng.device.lu.info(), "f64")["eval"]
    gbs = alg / (avg_ms * 1e-3) / 1e9
    out = {"workload": f"gate_dag_{args.gates}_seed{args.seed} with EKV 2.6 MOSFETs",
           "mosfets": n_dev, "unknowns": topo.n_augmented, **w,
           "analysis_ms": [round(x, 1) for x in eng.device.lu.info()["analyze_ms"]],
           "eval_blocks": eng.device.eval_blocks(),
           "mixed_precision": mixed,
           "phase_ms_per_newton_iter": {"build": round(ph["build_ms"] / iters, 4),
                                        "lu": round(ph["lu_ms"] / iters, 4),
                                        "solve": round(ph["solve_ms"] / iters, 4)},
           "eval": {"kernel": "k_eval<EKV26>", "avg_launch_us": round(avg_ms * 1e3, 2),
                    "ms_per_newton_iter": round(ev["ms"] / iters, 4),
                    "alg_bytes_per_launch": int(alg),
                    "achieved_GBs": round(gbs, 2), "hbm_frac": round(gbs / HBM_PEAK_GBS, 5)}}
    if EKV26_F64_OPS_PER_INSTANCE:
        tf = EKV26_F64_OPS_PER_INSTANCE * n_dev / (avg_ms * 1e-3) / 1e12
        out["eval"].update({"f64_ops_per_instance": EKV26_F64_OPS_PER_INSTANCE,
                            "achieved_f64_TFLOPs": round(tf, 3),
                            "f64_valu_frac": round(tf / FP64_VALU_PEAK_TFLOPS, 5)})
    del eng
    # the other synthetic workloads with the same model
    small = {}
    for name, netf, warm, steps, t_stop in (
            ("dag20k", lambda: G.gate_dag(20000, seed=args.seed, model="ekv26"), args.warmup,
             args.steps, 1e-9),
            ("ring101", lambda: G.ring_oscillator(101, seed=args.seed, model="ekv26"), 200, 2000,
             50e-9)):
        e = CircuitEngine(netf())
        e.parse()
        run_gpu(e, 3, "f64", True, t_stop=t_stop)
        r = run_gpu(e, warm + steps, "f64", True, window_from=warm + 1, t_stop=t_stop)
        small[name] = {"unknowns": e.topology.n_augmented, **window_stats(r)}
        log(f"ekv26 {name}: {small[name]['ms_per_step']} ms/step")
        del e
    out.update(small)
    if args.batch > 0:
        out["batch"] = run_batched(args, model="ekv26")
    return out


def summarize_trace(path, window=0):
    """Markdown table of the kernels in a rocprofv3 kernel-trace CSV inside
    one timed window: between the start and end k_window_mark dispatches of
    the window-th timed run of the process (0: the first)."""
    import csv

    with open(path, newline="") as fh:
        rows = list(csv.DictReader(fh))
    key = "Start_Timestamp" if "Start_Timestamp" in rows[0] else "Start_Time"
    endk = "End_Timestamp" if "End_Timestamp" in rows[0] else "End_Time"
    rows.sort(key=lambda r: int(r[key]))
    marks = [i for i, r in enumerate(rows) if _kernel_name(r["Kernel_Name"]).startswith("k_window_mark")]
    if len(marks) < 2 * window + 2:
        raise SystemExit(f"the trace holds {len(marks) // 2} windows, not {window + 1}")
    a, b = marks[2 * window], marks[2 * window + 1]
    win = rows[a + 1:b]
    t0, t1 = int(rows[a][endk]), int(rows[b][key])
    agg = {}
    for r in win:
        name = _kernel_name(r["Kernel_Name"])
        d = int(r[endk]) - int(r[key])
        a = agg.setdefault(name, [0, 0])
        a[0] += 1
        a[1] += d
    total = sum(v[1] for v in agg.values())
    print(f"Timed window: {len(win)} dispatches, {(t1 - t0) / 1e6:.3f} ms between the markers, "
          f"{total / 1e6:.3f} ms of kernel time\n")
    print("| kernel | calls | total ms | avg us | % of kernel time |")
    print("|---|---:|---:|---:|---:|")
    for name, (n, d) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{name}` | {n} | {d / 1e6:.3f} | {d / n / 1e3:.2f} | {100.0 * d / total:.1f} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--gates", type=int, default=65000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", choices=("f64", "f32"), default="f64")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--band-width", type=int, default=None,
                    help="levels per triangular-solve band (0: one launch per level)")
    ap.add_argument("--nd-leaf", type=int, default=None,
                    help="nested-dissection leaf size of the LU ordering")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-numpy", action="store_true",
                    help="also time the numpy oracle (slow) as a secondary CPU figure")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision runs (profiling passes)")
    ap.add_argument("--no-eager", action="store_true",
                    help="skip the eager per-kernel timing pass")
    ap.add_argument("--batch", type=int, default=256, help="Monte Carlo instances (0: skip)")
    ap.add_argument("--batch-gates", type=int, default=2000)
    ap.add_argument("--batch-ragged", default="500,1000,2000,4000")
    ap.add_argument("--batch-steps", type=int, default=50,
                    help="timed accepted steps of the batch (after --warmup untimed ones)")
    ap.add_argument("--no-mul64", action="store_true",
                    help="skip the mul64 leg (the reference's 64x64 multiplier benchmark)")
    ap.add_argument("--mul64-only", action="store_true",
                    help="run only the mul64 leg and print its JSON")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the small-circuit leg (RC ladder, ring oscillator)")
    ap.add_argument("--small-only", action="store_true",
                    help="run only the small-circuit leg and print its JSON")
    ap.add_argument("--no-ekv26", action="store_true",
                    help="skip the EKV 2.6 leg (the DAG with the full EKV 2.6 MOSFET model)")
    ap.add_argument("--ekv26-only", action="store_true",
                    help="run only the EKV 2.6 leg and print its JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device buffers (reference kernel for HBM counters)")
    ap.add_argument("--summarize-trace", default=None,
                    help="print the timed window of a rocprofv3 kernel_trace.csv and exit")
    ap.add_argument("--window", type=int, default=0,
                    help="with --summarize-trace: which timed window of the process (0: first)")
    ap.add_argument("--summarize-pmc", default=None,
                    help="FETCH,WRITE,CAL_FETCH,CAL_WRITE: four rocprofv3 --pmc CSVs (the bench "
                         "and --copy-calibration); write --pmc-json, print a table and exit")
    ap.add_argument("--pmc-json", default="hbm_traffic.json")
    ap.add_argument("--summarize-valu", default=None,
                    help="rocprofv3 --pmc CSV of the FP64 VALU counters: print the FP64 "
                         "operations per work-item of --valu-kernel and exit")
    ap.add_argument("--valu-kernel", default="k_eval<4>")
    ap.add_argument("--valu-grid", type=int, default=260096,
                    help="work-items of the dispatches to average (260096: 260 000 devices)")
    args = ap.parse_args()
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.window)
        return
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc.split(","), args.pmc_json)
        return
    if args.summarize_valu:
        summarize_valu(args.summarize_valu, args.valu_kernel, args.valu_grid)
        return
    if args.copy_calibration:
        import torch

        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": a.numel() * 8, "copies": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only: instances and circuits are not "
                                   "sharded across GPUs", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    if args.mul64_only:
        print(json.dumps(run_mul64(args)), flush=True)
        return
    if args.ekv26_only:
        print(json.dumps(run_ekv26(args)), flush=True)
        return
    if args.small_only:
        print(json.dumps(run_small(args)), flush=True)
        return
    from vajax_amd import _lib as L
    from vajax_amd.engine import CircuitEngine

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    t0 = time.time()
    net = build_circuit(args.gates, args.seed)
    eng = CircuitEngine(net)
    eng.parse()
    topo = eng.topology
    setup_s = time.time() - t0
    log(f"setup {setup_s:.1f}s: n_unknowns={topo.n_augmented} nse={topo.nse}")

    use_graph = not args.no_graph
    t0 = time.time()
    run_gpu(eng, 3, args.precision, use_graph, args.band_width, nd_leaf=args.nd_leaf)
    analysis_s = time.time() - t0
    log(f"analysis + autotune run {analysis_s:.1f}s")
    lu_info = eng.device.lu.info()

    total = args.warmup + args.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = run_gpu(eng, total, args.precision, use_graph, args.band_width, nd_leaf=args.nd_leaf,
                  window_from=args.warmup + 1)
    torch.cuda.synchronize()
    call_s = time.perf_counter() - t0
    st = res.stats
    win = window_stats(res)
    steps_done = win["accepted_steps"]
    nr = win["newton_iterations"]
    elapsed = st["window_time"]
    ms_per_step = win["ms_per_step"]
    log(f"timed window: {steps_done} steps, {nr} Newton iterations, {elapsed:.3f}s "
        f"(whole run {st['n_steps'] - 1} steps in {st['wall_time']:.3f}s, call {call_s:.3f}s)")

    mixed = None
    if args.precision == "f64" and not args.no_mixed:
        r32 = run_gpu(eng, total, "f32", use_graph, args.band_width, nd_leaf=args.nd_leaf,
                      window_from=args.warmup + 1)
        mixed = {"window": window_stats(r32)}
        mixed.update(mixed_error(eng, res, use_graph))
        mixed["mixed_vs_fp64_ms_per_step"] = round(ms_per_step / mixed["window"]["ms_per_step"], 3)
        mixed["mixed_vs_fp64_ms_per_newton_iter"] = round(
            win["ms_per_newton_iter"] / mixed["window"]["ms_per_newton_iter"], 3)
        log(f"mixed: {mixed['window']['ms_per_step']} ms/step, max |dV| vs FP64 "
            f"{mixed['max_abs_dV_vs_fp64']:.2e} V")

    ph = None
    kernels = {}
    if not args.no_eager:
        # eager pass with per-launch events on the library stream
        L.lib.vjx_profile_enable(1)
        L.lib.vjx_profile_reset()
        run_gpu(eng, total, args.precision, False, args.band_width, nd_leaf=args.nd_leaf)
        pt = L.PhaseTimes()
        L.lib.vjx_profile_read(C.byref(pt))
        L.lib.vjx_profile_enable(0)
        ph = pt.as_dict()
    iters = max(ph["n_iterations"], 1) if ph else 1
    per_iter = ({"build": ph["build_ms"] / iters, "lu": ph["lu_ms"] / iters,
                 "solve": ph["solve_ms"] / iters} if ph else {})
    kb = kernel_bytes(topo, lu_info, args.precision)
    for name, kt in (ph["kernels"].items() if ph else ()):
        if kt["launches"] == 0:
            continue
        avg_ms = kt["ms"] / kt["launches"]
        per_launch = kb[name] / (kt["launches"] / iters)
        kernels[name] = {"kernel": KERNEL_NAMES[name], "ms_per_newton_iter": round(kt["ms"] / iters, 4),
                         "launches_per_newton_iter": round(kt["launches"] / iters, 2),
                         "avg_launch_us": round(avg_ms * 1e3, 2),
                         "alg_bytes_per_launch": int(per_launch),
                         "achieved_GBs": round(per_launch / (avg_ms * 1e-3) / 1e9, 2)}
    roofline = None
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k]["ms_per_newton_iter"])
        d = kernels[dom]
        traffic = measured_traffic(args, dom)
        roofline = {
            "bound": "hbm",
            "kernel": d["kernel"],
            "achieved": d["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(d["achieved_GBs"] / HBM_PEAK_GBS, 6),
            "traffic": traffic,
            "avg_launch_us": d["avg_launch_us"],
            "alg_bytes_per_launch": d["alg_bytes_per_launch"],
            "phase_ms_per_newton_iter": {k: round(v, 4) for k, v in per_iter.items()},
            "kernels": kernels,
        }
        lt = kernels.get("lu_tasks")
        if lt:
            # the band tasks are bound by their chain of dependent pivot
            # levels, not by bandwidth: time per level of the elimination
            levels = lu_info["fact_levels"]
            roofline["level_latency"] = {
                "kernel": lt["kernel"], "pivot_levels_per_refactor": levels,
                "launches_per_refactor": lt["launches_per_newton_iter"],
                "us_per_level": round(lt["ms_per_newton_iter"] * 1e3 / max(levels, 1), 3),
                "note": "each level waits on one LDS round trip and a workgroup barrier"}

    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = cpu_baseline(args)
        log(f"cpu baseline {time.time() - t0:.1f}s: {cpu['ms_per_step']:.1f} ms/step on "
            f"{cpu['cores']} threads, {cpu['single_thread']['ms_per_step']:.1f} on one")

    batched = None
    if args.batch > 0:
        batched = run_batched(args)

    mul64 = None
    if not args.no_mul64:
        mul64 = run_mul64(args)

    ekv26 = None
    if not args.no_ekv26:
        ekv26 = run_ekv26(args)

    small = None
    if not args.no_small:
        small = run_small(args)

    line = {
        "metric": METRIC,
        "value": win["newton_iters_per_s"],
        "unit": "Newton iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "call_ms_per_step": round(call_s / max(total, 1) * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "f64" else "f32-lu+f64",
        "data": "synthetic",
        "config": {
            "workload": f"gate_dag_{args.gates}_seed{args.seed}",
            "gates": args.gates,
            "mosfets": 4 * args.gates,
            "unknowns": topo.n_augmented,
            "csr_nnz": topo.nse,
            "lu_nnz": lu_info["nnz_l"] + lu_info["nnz_u"],
            "lu_levels": lu_info["fact_levels"],
            "solve_bands": [lu_info["fwd_bands"], lu_info["bwd_bands"]],
            "band_width": lu_info["band_width"],
            "lu_pipeline": lu_info["lu_pipeline"],
            "eval_blocks": eng.device.eval_blocks(),
            "solve_band_width": lu_info["solve_band_width"],
            "integration": "gear2 adaptive",
            "precision": args.precision,
            "hip_graph": use_graph,
            "window": win,
            "whole_run": {"accepted_steps": st["n_steps"] - 1, "rejected": st["rejected_steps"],
                          "newton_iterations": st["total_nr_iterations"],
                          "loop_s": round(st["wall_time"], 4)},
            "setup_s": round(setup_s, 2),
            "analysis_s": round(analysis_s, 2),
            "analysis_ms": [round(x, 1) for x in lu_info["analyze_ms"]],
            "parallelism": "single-gpu",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "mul64": mul64,
        "ekv26": ekv26,
        "small": small,
        "mixed_precision": mixed,
    }
    if cpu:
        line["speedup_vs_cpu_ms_per_newton_iter"] = round(
            cpu["ms_per_newton_iter"] / win["ms_per_newton_iter"], 2)
        line["speedup_vs_cpu_ms_per_step"] = round(cpu["ms_per_step"] / ms_per_step, 2)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
