"""Benchmark: sigma-vector throughput of the gfx950 engine on one GPU.

One step = one application of the two-site effective Hamiltonian (sigma = H_eff psi) at the
middle site of the chain, inputs resident in HBM.  value = sigma TFLOP/s with the reference's
own flop model (solver/tasks.cpp:388-398, 777-796).  The roofline entry is the dominant kernel
of the step, timed with HIP events on the engine's stream; the CPU baseline is the oracle (the
CPU restatement of the reference's CPU branch) on the same inputs on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hubbard-L64-m1024]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "maquis-dmrg_amd"))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # name: (model, L, m)
    "hubbard-L64-m1024": ("hubbard", 64, 1024),
    "heisenberg-L32-m128": ("heisenberg", 32, 128),
    "qc-L16-m256": ("qc", 16, 256),
    "qc-L24-m1000": ("qc", 24, 1000),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="hubbard-L64-m1024", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=7)
    return ap.parse_args()


def cpu_baseline(ws, Lcoh, Rcoh, psi, flops, budget_s):
    """Time the oracle's CPU sigma (OpenMP over MPS blocks, all host cores) on the same inputs."""
    from oracle import oracle as O
    blas = O.use_numpy_blas()
    Lb = O.HostBoundary(ws.left_index, [c if c is not None else __import__("numpy").zeros(0) for c in Lcoh])
    Rb = O.HostBoundary(ws.right_index, [c if c is not None else __import__("numpy").zeros(0) for c in Rcoh])
    so = O.SigmaOracle(ws.site_desc, ws.mpo_desc, Lb, Rb)
    t0 = time.perf_counter()
    out = so.apply(psi)
    n, t = 1, time.perf_counter() - t0
    while t < budget_s and n < 50:
        so.apply(psi)
        n += 1
        t = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return out, {"value": flops * n / t / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
                 "sample": "%d full sigma applications of the same workload (%.1f s), dgemm: %s"
                           % (n, t, "numpy OpenBLAS, 1 thread per call" if blas else "oracle blocked loop")}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from mqdmrg import engine as E
    from mqdmrg.workloads import SyntheticSite

    model, L, m = WORKLOADS[args.workload]
    site = L // 2 - 1
    ws = SyntheticSite(model, L, site, m, seed=args.seed)
    eng = E.Engine(0)
    Lcoh = ws.random_cohorts(ws.left_index)
    Rcoh = ws.random_cohorts(ws.right_index)
    Lb = E.DeviceBoundary.upload(eng, ws.left_index, Lcoh)
    Rb = E.DeviceBoundary.upload(eng, ws.right_index, Rcoh)
    t_plan = time.perf_counter()
    plan = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb)
    t_plan = time.perf_counter() - t_plan
    psi_h = ws.random_psi()
    psi = torch.from_numpy(psi_h).to(eng.torch_device)
    sigma = torch.zeros_like(psi)

    for _ in range(args.warmup):
        plan.apply(psi, sigma)
    eng.sync()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(eng.stream)
    for _ in range(args.steps):
        plan.apply(psi, sigma)
    ev1.record(eng.stream)
    eng.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    value = plan.flops / (ms_step * 1e-3) / 1e12

    # per-stage timing of the same kernels (HIP events between stages on the engine stream)
    stage_ms, stage_flops, gather_bytes = plan.profile(psi, sigma, reps=max(3, min(args.steps, 20)))
    names = ["fill", "gemm_T", "gather_S", "gemm_final"]
    dom = max(range(1, 4), key=lambda k: stage_ms[k])
    if dom == 2:
        achieved = gather_bytes / (stage_ms[2] * 1e-3) / 1e9
        roof = {"kernel": "gather_kernel (S stage)", "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": None}
    else:
        fl = stage_flops[0] if dom == 1 else stage_flops[2]
        achieved = fl / (stage_ms[dom] * 1e-3) / 1e12
        roof = {"kernel": "gemm_f64_kernel (%s)" % names[dom], "bound": "mfma", "achieved": achieved,
                "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_MFMA_TFLOPS,
                "traffic": None}
    roof["stage_ms"] = dict(zip(names, [round(x, 4) for x in stage_ms]))

    cpu = None
    if not args.no_cpu:
        ref, cpu = cpu_baseline(ws, Lcoh, Rcoh, psi_h, plan.flops, args.cpu_seconds)
        plan.apply(psi, sigma)
        got = sigma.cpu().numpy()
        cpu["parity_rel_err"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))

    out = {
        "metric": "sweep wall-time & σ TFLOP/s at L=30 m=2048, mixed vs FP64 at |ΔE|≤1e-6 Eh",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "sigma (H_eff psi) at the middle site, %s chain L=%d, bond dimension m=%d, FP64 plain path"
                        % (model, L, m),
            "model": model, "L": L, "m": m, "site": site,
            "bond_sectors": len(ws.bonds[site + 1].sizes),
            "flops_per_sigma": plan.flops,
            "psi_elements": plan.n,
            "plan_seconds": round(t_plan, 3),
            "workspace_bytes": plan.workspace_bytes,
            "wall_seconds_timed": round(wall, 4),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
