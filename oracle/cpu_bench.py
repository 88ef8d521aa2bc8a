"""CPU baseline run: the C++ restatement (``oracle/ctran.cpp``) on the gate-DAG
transient, timed on host cores.  Test / benchmark infrastructure only.

    python -m oracle.cpu_bench --gates 65000 --steps 20 --threads 16 1

prints one JSON line with, per thread count, wall time, accepted steps,
Newton iterations and the phase split (build, LU, solves).  Setup (circuit
arrays, the fill-reducing order) is excluded like the reference excludes JIT
and symbolic analysis; the first pivoting factorisation is inside the timed
loop, as UMFPACK factorises every call.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gates", type=int, default=65000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--threads", type=int, nargs="+", default=[0])
    args = ap.parse_args()

    from oracle import ctran, mna
    from oracle import transient as T
    from vajax_amd import generators as G

    t0 = time.time()
    net = G.gate_dag(args.gates, seed=args.seed)
    cir = mna.setup(net)
    ct = ctran.CTran(cir)
    setup_s = time.time() - t0
    cfg = T.AdaptiveConfig(integration_method="gear2")
    runs = []
    for th in args.threads:
        out = ct.run_adaptive(1e-9, 1e-12, cfg, T.Options(icmode="uic"), max_steps=args.steps + 1,
                              threads=th)
        st = out.stats
        acc = max(int(st["n_steps"]) - 1, 1)
        runs.append({
            "threads": int(st["threads"]),
            "wall_s": st["wall"],
            "accepted_steps": acc,
            "newton_iterations": int(st["total_nr"]),
            "ms_per_step": st["wall"] / acc * 1e3,
            "newton_iters_per_s": st["total_nr"] / st["wall"],
            "phase_s": {"build": st["t_build"], "lu": st["t_lu"], "solve": st["t_solve"]},
            "lu_factorisations": int(st["lu_factors"]),
            "lu_refactorisations": int(st["lu_refactors"]),
        })
    print(json.dumps({"workload": net.title, "unknowns": cir.n_aug, "csr_nnz": cir.nse,
                      "setup_s": setup_s, "runs": runs}))


if __name__ == "__main__":
    main()
