"""Benchmark: sigma-vector throughput of the gfx950 engine on one GPU, FP64 against the
reduced-precision paths.

One step = one application of the two-site effective Hamiltonian (sigma = H_eff psi) at the
middle site of the chain, inputs resident in HBM.  value = sigma TFLOP/s with the reference's own
flop model (solver/tasks.cpp:388-398, 777-796) in the reported precision (default: the fastest
reduced-precision path whose relative sigma error stays within 1e-4).  The FP64 plain path and
every reduced precision are timed in the same run.  A two-site DMRG run of the same ab initio
Hamiltonian at a smaller size is then swept with the FP64 sigma and with the reported precision
(refined Jacobi-Davidson: reduced-precision products, FP64 Rayleigh quotient and residual) to
report the energy difference in Eh and the sweep wall times.  The roofline entry is the dominant kernel of the step, timed
with HIP events on the engine's stream; the CPU baseline is the oracle (the CPU restatement of
the reference's CPU branch) on a bounded sample of the same sigma (a subset of its MPS blocks)
on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload qc_su2-L30-m2048]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "maquis-dmrg_amd"))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0 / 3}   # dense MFMA; bf16x3 = 3 bf16 products
PEAK_HBM_GBS = 8000.0
MIXED_RTOL = 1e-4

WORKLOADS = {
    # name: (model, L, m); qc_su2: spin-adapted (SU2U1) ab initio chain, N = L, singlet, m multiplets
    "qc_su2-L30-m2048": ("qc_su2", 30, 2048),
    "qc_su2-L24-m1000": ("qc_su2", 24, 1000),
    "qc_su2-L16-m256": ("qc_su2", 16, 256),
    "qc-L30-m2048": ("qc", 30, 2048),
    "qc-L24-m1000": ("qc", 24, 1000),
    "qc-L16-m256": ("qc", 16, 256),
    "hubbard-L64-m1024": ("hubbard", 64, 1024),
    "heisenberg-L32-m128": ("heisenberg", 32, 128),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="qc_su2-L30-m2048", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="auto", choices=["auto", "f64", "f32", "bf16x3"],
                    help="precision reported as value (auto: fastest reduced precision within the sigma bar)")
    ap.add_argument("--only", default=None, choices=["f64", "f32", "bf16x3"],
                    help="time this precision alone (profiling runs; it is then also the reported one)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-L", type=int, default=16, help="chain length of the sweep comparison")
    ap.add_argument("--sweep-m", type=int, default=256, help="bond dimension of the sweep comparison")
    ap.add_argument("--sweep-n", type=int, default=6, help="sweeps of the sweep comparison (converged energies)")
    ap.add_argument("--batch", type=int, default=0,
                    help="also time N independent ab initio instances (L uniform in [12, 24], m=256) "
                         "applied one by one and as one batch")
    ap.add_argument("--no-headline-sweep", action="store_true",
                    help="skip the wall-time of full two-site sweeps at the workload's shape in FP64 and in the "
                         "reported precision")
    ap.add_argument("--headline-sweeps", type=int, default=1, help="sweeps of the headline sweep comparison")
    ap.add_argument("--headline-which", default="both", choices=["both", "f64", "reduced"],
                    help="run the headline sweeps in FP64, in the reported precision, or both")
    ap.add_argument("--headline-verbose", action="store_true",
                    help="print every site pair of the headline (and differentiation) sweeps: energy, "
                         "iterations, bond dimension")
    ap.add_argument("--adjoint", action="store_true",
                    help="also run the differentiation workload: two-site DMRG of the ab initio chain "
                         "(--adjoint-L orbitals, N = L, bond dimension --adjoint-m), the sigma-adjoint sweep "
                         "giving dE/dh and dE/d(pq|rs), its identities, and central finite differences of the "
                         "re-optimised energy for a few integrals")
    ap.add_argument("--adjoint-L", type=int, default=16)
    ap.add_argument("--adjoint-m", type=int, default=256)
    ap.add_argument("--adjoint-sweeps", type=int, default=6)
    ap.add_argument("--adjoint-fd", type=int, default=3, help="integrals checked by finite differences")
    ap.add_argument("--adjoint-cache", action="store_true",
                    help="run the differentiation workload with the boundary cache: only the boundaries of the "
                         "current site pair on the device, the others in pinned host memory, and during the "
                         "adjoint sweep the left boundaries recomputed as the orthogonality center moves")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile configurations (no timing of candidates before the run)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="GB",
                    help="only copy a device buffer of this many GB five times (a kernel of known HBM traffic: "
                         "the reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(ws, Lcoh, Rcoh, psi, budget_s, threads):
    """Time the oracle's CPU sigma (OpenMP over MPS blocks, `threads` host threads) on a sample of
    the MPS blocks of the same sigma: `threads` blocks of about the median size (so every thread
    works; smaller blocks if one application of those exceeds budget_s / 3), applied repeatedly for
    about budget_s; the GPU's sigma restricted to the same blocks is the parity check."""
    import numpy as np
    from oracle import oracle as O
    blas = O.use_numpy_blas()
    Lb = O.HostBoundary(ws.left_index, [c if c is not None else np.zeros(0) for c in Lcoh])
    Rb = O.HostBoundary(ws.right_index, [c if c is not None else np.zeros(0) for c in Rcoh])
    lay = ws.layout
    nblk = len(lay.left.sizes)
    by_size = sorted(range(nblk), key=lambda b: lay.left.sizes[b] * lay.rp_cols[b])
    count = min(threads, nblk)
    start = max(0, min(nblk - count, nblk // 2 - count // 2))
    while True:
        blocks = sorted(by_size[start:start + count])
        so = O.SigmaOracle(ws.site_desc, ws.mpo_desc, Lb, Rb, only_blocks=blocks)
        t0 = time.perf_counter()
        out = so.apply(psi)
        t1 = time.perf_counter() - t0
        if t1 <= budget_s / 3 or start == 0:
            break
        start = max(0, start - max(1, count // 2))
    n, t = 1, t1
    while t < budget_s and n < 20:
        t0 = time.perf_counter()
        so.apply(psi)
        t += time.perf_counter() - t0
        n += 1
    res = {"value": so.flops * n / t / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
           "sample": "%d applications of sigma restricted to %d of %d MPS blocks (ranks %d-%d by size; %.3e of the "
                     "reference-model flops per application), %.1f s, OpenMP over the %d blocks; dgemm: %s"
                     % (n, len(blocks), nblk, start, start + len(blocks) - 1, so.flops, t, threads,
                        "numpy OpenBLAS, 1 thread per call" if blas else "oracle blocked loop")}
    return res, blocks, out


def headline_sweep(eng, args, report, prec_code, chain, model, L, m):
    """Wall time of n_sweeps full two-site sweeps (both directions) of the ab initio chain (L
    orbitals, N = L, singlet for SU2U1) at bond dimension m, from the same random state of total
    bond dimension m, in FP64 and in the reported precision (refined Jacobi-Davidson: reduced-
    precision products, one FP64 product per site), with the per-sweep energies of both runs and
    their difference.  chain is the (basis, mpo, target) of the timed workload."""
    import torch
    from mqdmrg.dmrg import TwoSiteDMRG
    from mqdmrg import engine as E
    sym = "SU2U1, singlet" if model == "qc_su2" else "TwoU1"
    n = args.headline_sweeps
    out = {"model": "ab initio (synthetic integrals), %s, L=%d, N=%d, m=%d, %d sweep(s) from a random state of "
                    "total bond dimension %d, JD tol 1e-8, 8 iterations per site" % (sym, L, L, m, n, m)}
    basis, mpo, target = chain
    runs = [("f64", E.PREC_F64), (report, prec_code)]
    if args.headline_which != "both":
        runs = runs[:1] if args.headline_which == "f64" else runs[1:]
    for name, pc in runs:
        t0 = time.perf_counter()
        d = TwoSiteDMRG(eng, model, L, m=m, cutoff=1e-10, init_m=m, seed=3, jd_tol=1e-8, jd_maxiter=8,
                        mpo=mpo, basis=basis, target=target, precision=pc, init_bonds="total",
                        verbose=args.headline_verbose)
        d.init_boundaries()
        t_setup = time.perf_counter() - t0
        torch.cuda.synchronize()
        secs = []
        for _ in range(n):
            t0 = time.perf_counter()
            d.sweep()
            torch.cuda.synchronize()
            secs.append(round(time.perf_counter() - t0, 2))
            log("headline sweep %s %d: %.1f s E=%.10f" % (name, len(secs), secs[-1], d.energies[-1]))
        out[name] = {"sweep_seconds": secs, "setup_seconds": round(t_setup, 2), "energies_Eh": list(d.energies),
                     "breakdown_seconds": {k: round(v, 2) for k, v in d.timings.items() if k != "sweep"},
                     "sigma_calls": d.sigma_calls}
        log("headline sweeps %s: %s s %s" % (name, secs, out[name]["breakdown_seconds"]))
        del d
        torch.cuda.empty_cache()
    if report != "f64" and "f64" in out and report in out:
        f, r = out["f64"], out[report]
        out["speedup_vs_f64"] = sum(f["sweep_seconds"]) / sum(r["sweep_seconds"])
        out["abs_delta_e_per_sweep_Eh"] = [abs(a - b) for a, b in zip(f["energies_Eh"], r["energies_Eh"])]
    return out


def adjoint_workload(eng, args, su2_model):
    """dE/dh_pq and dE/d(pq|rs) of the DMRG ground state from the sigma adjoint (one-site adjoint at
    every site, Hellmann-Feynman), with size-independent identities and finite differences of the
    re-optimised energy (warm start from the converged state, step 1e-5).  su2_model: the
    spin-adapted Hamiltonian (singlet) instead of the TwoU1 one (n_up = n_down)."""
    import numpy as np
    import torch
    from mqdmrg import qc, su2
    from mqdmrg.dmrg import TwoSiteDMRG, SiteTensor
    L, m = args.adjoint_L, args.adjoint_m
    N = L
    h, eri = qc.synthetic_integrals(L, seed=0)
    kw = dict(m=m, cutoff=1e-12, init_m=16, seed=3, jd_tol=1e-10, jd_maxiter=12, boundary_cache=args.adjoint_cache,
              verbose=args.headline_verbose)
    model = "qc_su2" if su2_model else "qc"
    target = (N, 0) if su2_model else (N // 2, N - N // 2)

    def make(hh, ee):
        if su2_model:
            basis, mpo, _, elements, src = su2.qc_su2_mpo(hh, ee, N, sources=True)
            return basis, mpo, (elements, src)
        basis, mpo, _, src = qc.qc_mpo(hh, ee, sources=True)
        return basis, mpo, src

    def to_tensors(mpo, src, grads):
        if su2_model:
            elements, esrc = src
            g = su2.mpo_gradient_to_elements(mpo, esrc, grads, len(elements))
            return su2.element_gradient_to_tensors(elements, g, L)
        return qc.mpo_gradient_to_integrals(mpo, src, grads, L)
    basis, mpo, src = make(h, eri)
    d = TwoSiteDMRG(eng, model, L, mpo=mpo, basis=basis, target=target, **kw)
    t0 = time.perf_counter()
    e = d.run(args.adjoint_sweeps)
    t_dmrg = time.perf_counter() - t0
    log("adjoint workload: L=%d m=%d E=%.12f after %d sweeps (%.1f s)" % (L, m, e, args.adjoint_sweeps, t_dmrg))
    state = [SiteTensor(eng, s.sym, s.left, s.phys, s.right, s.data.clone()) for s in d.sites]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grads, energies = d.mpo_gradients()
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    dh, deri, _ = to_tensors(mpo, src, grads)
    t_map = time.perf_counter() - t0 - t_adj
    euler = float(np.sum(h * dh) + np.sum(eri * deri))
    contr = np.einsum("pqrr->pq", deri)
    res = {
        "model": "ab initio (synthetic integrals), %s, L=%d, N=%d, m=%d, %d sweeps, FP64%s"
                 % ("SU2U1 singlet" if su2_model else "TwoU1", L, N, m, args.adjoint_sweeps,
                    ", boundary cache (left boundaries recomputed in the adjoint sweep, right boundaries "
                    "fetched from host memory)" if args.adjoint_cache else ""),
        "energy_Eh": e, "dmrg_seconds": round(t_dmrg, 2),
        "adjoint_sweep_seconds": round(t_adj, 3), "integral_map_seconds": round(t_map, 3),
        "site_energy_spread_Eh": float(max(energies) - min(energies)),
        "checks": {
            "euler_sum_x_dEdx_minus_E": euler - e,
            "trace_1rdm_minus_N": float(np.trace(dh) - N),
            "1rdm_asymmetry": float(np.linalg.norm(dh - dh.T)),
            "2rdm_partial_trace_residual": float(np.linalg.norm(contr - 0.5 * (N - 1) * dh) / np.linalg.norm(dh)),
        },
    }
    log("adjoint sweep %.2f s; checks %s" % (t_adj, res["checks"]))
    # finite differences of the re-optimised energy, warm-started from the converged state
    fd = []
    picks = [("h", (0, 0)), ("h", (L // 2, L // 2 + 1)), ("eri", (1, 1, 2, 2)), ("eri", (0, 1, 1, 0))][:args.adjoint_fd]
    def perm_class(kind, idx):
        # the spin-adapted generator reads one element per permutation class: perturb the class.
        # TwoU1: the entry with its Hermitian partner (a single off-diagonal entry alone makes H
        # non-Hermitian); the derivatives of the two are equal
        if not su2_model:
            return sorted({idx, qc.hermitian_partner(kind, idx)})
        if kind == "h":
            return sorted({idx, idx[::-1]})
        p, q, r, s_ = idx
        return sorted({(p, q, r, s_), (q, p, r, s_), (p, q, s_, r), (q, p, s_, r), (r, s_, p, q), (s_, r, p, q),
                       (r, s_, q, p), (s_, r, q, p)})
    for kind, idx in picks:
        cls = perm_class(kind, idx)
        vals = []
        for sgn in (+1, -1):
            hh, ee = h.copy(), eri.copy()
            for x in cls:
                (hh if kind == "h" else ee)[x] += sgn * 1e-5
            b2, mpo2, _ = make(hh, ee)
            d2 = TwoSiteDMRG(eng, model, L, mpo=mpo2, basis=b2, target=target, **kw)
            d2.sites = [SiteTensor(eng, s.sym, s.left, s.phys, s.right, s.data.clone()) for s in state]
            vals.append(d2.run(2))
        num = (vals[0] - vals[1]) / 2e-5
        ana = float(sum((dh if kind == "h" else deri)[x] for x in cls))
        fd.append({"integral": kind, "index": list(idx), "adjoint": ana, "finite_difference": num,
                   "rel_err": abs(ana - num) / max(abs(num), 1e-300)})
        log("  d E / d %s%s: adjoint %.10f  FD %.10f" % (kind, idx, ana, num))
    res["finite_differences"] = fd
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (the engine has no multi-GPU path); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if args.calibrate_copy > 0:
        n = int(args.calibrate_copy * 1e9) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 8 * n, "calibrate_copy_bytes_written": 8 * n, "copies": 5}))
        return
    from mqdmrg import engine as E
    from mqdmrg.workloads import SyntheticSite

    model, L, m = WORKLOADS[args.workload]
    site = L // 2 - 1
    t_site = time.perf_counter()
    ws = SyntheticSite(model, L, site, m, seed=args.seed)
    t_site = time.perf_counter() - t_site
    log("site built in %.1f s: %d psi elements" % (t_site, ws.layout.rp_total))
    eng = E.Engine(0)
    Lcoh = ws.random_cohorts(ws.left_index)
    Rcoh = ws.random_cohorts(ws.right_index)
    Lb = E.DeviceBoundary.upload(eng, ws.left_index, Lcoh)
    Rb = E.DeviceBoundary.upload(eng, ws.right_index, Rcoh)
    t_plan = time.perf_counter()
    plan = E.SigmaPlan(eng, ws.site_desc, ws.mpo_desc, Lb, Rb, algorithm=E.SIGMA_GROUPED)
    t_plan = time.perf_counter() - t_plan
    log("plan built in %.1f s: %.3e flops per sigma, %.2f GB workspace" % (t_plan, plan.flops,
                                                                         plan.workspace_bytes / 1e9))
    psi_h = ws.random_psi()
    psi = torch.from_numpy(psi_h).to(eng.torch_device)
    sigma = torch.zeros_like(psi)

    # ---- timed sigma applications per precision
    per = {}
    ref = None
    precs = (("f64", E.PREC_F64), ("f32", E.PREC_F32), ("bf16x3", E.PREC_BF16X3))
    if args.only:
        precs = [x for x in precs if x[0] == args.only]
    for name, prec in precs:
        tune = None
        if not args.no_autotune:
            # tile configurations of the T and final stages: stage times before and after tuning
            untuned, _, _ = plan.profile(psi, sigma, reps=3, precision=prec)
            t0 = time.perf_counter()
            n_timed = plan.autotune(prec)
            tune = {"timings": n_timed, "seconds": round(time.perf_counter() - t0, 3),
                    "untuned_stage_ms": dict(zip(["fill", "gemm_T", "gemm_S", "gemm_final"],
                                                 [round(x, 4) for x in untuned]))}
            log("%s autotune: %d stage timings in %.2f s" % (name, n_timed, tune["seconds"]))
        for _ in range(args.warmup):
            plan.apply(psi, sigma, precision=prec)
        eng.sync()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(eng.stream)
        for _ in range(args.steps):
            plan.apply(psi, sigma, precision=prec)
        ev1.record(eng.stream)
        eng.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = ev0.elapsed_time(ev1) / args.steps
        out = sigma.clone()
        if name == "f64":
            ref = out
        err = float(torch.linalg.norm(out - ref) / torch.linalg.norm(ref)) if ref is not None else None
        stage_ms, stage_flops, _ = plan.profile(psi, sigma, reps=max(2, min(args.steps, 5)), precision=prec)
        per[name] = {"ms_per_sigma": ms, "tflops": plan.flops / (ms * 1e-3) / 1e12, "sigma_rel_err_vs_f64": err,
                     "wall_seconds_timed": round(wall, 4),
                     "stage_ms": dict(zip(["fill", "gemm_T", "gemm_S", "gemm_final"], [round(x, 4) for x in stage_ms])),
                     "stage_gemm_flops": dict(zip(["T", "S", "final"], stage_flops)), "autotune": tune}
        log("%s: %.3f ms/sigma, %.2f TFLOP/s, rel err %s, stages %s" % (
            name, ms, per[name]["tflops"], "n/a" if err is None else "%.2e" % err, per[name]["stage_ms"]))

    if args.only:
        report = args.only
    elif args.precision == "auto":
        ok = [p for p in ("f32", "bf16x3") if per[p]["sigma_rel_err_vs_f64"] <= MIXED_RTOL]
        report = min(ok, key=lambda p: per[p]["ms_per_sigma"]) if ok else "f64"
    else:
        report = args.precision
    prec_code = E.PRECISIONS[report]

    # ---- sweeps: the same two-site DMRG run (ab initio Hamiltonian from the benchmark integral
    # generator) with the FP64 sigma and with the reported precision; energies in Eh and wall times
    sweep = None
    if not args.no_sweep and report != "f64":
        from mqdmrg.dmrg import TwoSiteDMRG
        sL, sm, nsw = args.sweep_L, args.sweep_m, args.sweep_n
        kw = dict(m=sm, cutoff=1e-10, init_m=16, seed=3, jd_tol=1e-8, jd_maxiter=8, integral_seed=0, ecore=0.0,
                  N=sL)
        sweep_model = "qc_su2" if model == "qc_su2" else "qc"
        runs = {}
        for name, pc in (("f64", E.PREC_F64), (report, prec_code)):
            d = TwoSiteDMRG(eng, sweep_model, sL, precision=pc, **kw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e = d.run(nsw)
            torch.cuda.synchronize()
            runs[name] = {"energy_Eh": e, "energies_per_sweep": d.energies, "seconds": time.perf_counter() - t0,
                          "sweep_seconds": d.timings["sweep"], "jd_seconds": d.timings["jd"],
                          "sigma_calls": d.sigma_calls}
            log("sweeps %s: E=%.10f  %s s  calls %s" % (name, e, [round(x, 2) for x in d.timings["sweep"]],
                                                       d.sigma_calls))
        sweep = {"model": "ab initio (synthetic integrals), %s, L=%d, N=%d, m=%d, %d sweeps, JD tol 1e-8, "
                          "max_iter 8; reduced precision: one FP64 refinement step per site"
                          % ("SU2U1 singlet" if sweep_model == "qc_su2" else "TwoU1", sL, sL, sm, nsw),
                 "runs": runs, "abs_delta_e_Eh": abs(runs[report]["energy_Eh"] - runs["f64"]["energy_Eh"]),
                 "speedup_vs_f64": runs["f64"]["seconds"] / runs[report]["seconds"]}

    # ---- roofline of the dominant kernel in the reported precision.  achieved = the stage's flops in
    # the reference's model (solver/tasks.cpp:388-398, 777-796) per second: for the T and final stages
    # these are the block products the kernel executes; for the S stage the model counts one alpha
    # per nonzero MPO coupling, i.e. the plan's model flops minus those of the T and final stages
    # (the executed MFMA flops, which include the zeros inside the visited W_D tiles, are reported
    # beside it).
    pr = per[report]
    st = pr["stage_ms"]
    dom = max(["gemm_T", "gemm_S", "gemm_final"], key=lambda k: st[k])
    sf = pr["stage_gemm_flops"]
    executed = sf[{"gemm_T": "T", "gemm_S": "S", "gemm_final": "final"}[dom]]
    fl = max(0, plan.flops - sf["T"] - sf["final"]) if dom == "gemm_S" else executed
    achieved = fl / (st[dom] * 1e-3) / 1e12
    if dom == "gemm_S":
        kernel = "mq::s_bsr_kernel<%s> (block-sparse S stage over the nonzero 16-key tiles of W_D, %s MFMA)" % (
            {"f64": "MODE_F64", "f32": "MODE_F32", "bf16x3": "MODE_BF16X3"}[report],
            {"f64": "FP64", "f32": "FP32", "bf16x3": "bf16 split-product, FP32 accumulate"}[report])
        peak_key = report
    else:
        # T and final stages: the lean MFMA kernels (and the generic grouped GEMM for other tile
        # shapes), FP32 accumulation in the reduced precisions
        kernel = ("mq::t_stack_kernel (T stage)" if dom == "gemm_T" else "mq::f_unit_kernel (final stage)")
        peak_key = "f64" if report == "f64" else "f32"
    roof = {"kernel": kernel, "bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[peak_key],
            "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[peak_key], "traffic": None, "stage_ms": st,
            "flops_per_launch_set": fl, "flops_model": "reference (alpha per nonzero MPO coupling)",
            "executed_mfma_flops": executed, "executed_tflops": executed / (st[dom] * 1e-3) / 1e12}
    # HBM bytes of the same launch set, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes of
    # this workload (profiles/), when they cover the reported stage and precision
    tf = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (args.workload.replace("-", "_"), report))
    if dom == "gemm_S" and os.path.exists(tf):
        with open(tf) as fh:
            t = json.load(fh)
        roof["traffic"] = t["traffic_bytes"]
        roof["traffic_source"] = os.path.relpath(tf, ROOT)

    # ---- CPU baseline on a bounded sample of the same sigma
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        cpu, blocks, ref_cpu = cpu_baseline(ws, Lcoh, Rcoh, psi_h, args.cpu_seconds, threads)
        # GPU sigma of psi restricted to the sampled ket blocks (others zeroed): same contributions
        masked = np.zeros_like(psi_h)
        lay = ws.layout
        for b in blocks:
            o = lay.rp_offset[b]
            n = lay.left.sizes[b] * lay.rp_cols[b]
            masked[o:o + n] = psi_h[o:o + n]
        plan.apply(torch.from_numpy(masked).to(eng.torch_device), sigma)
        eng.sync()
        got = sigma.cpu().numpy()
        cpu["parity_rel_err"] = float(np.linalg.norm(got - ref_cpu) / np.linalg.norm(ref_cpu))

    # ---- batch of independent instances: sequential applications vs one batched launch
    batch = None
    if args.batch > 0:
        rng = np.random.default_rng(args.seed)
        plans, psis, sigs, keep = [], [], [], []
        flops = 0
        batch_model = "qc_su2" if model == "qc_su2" else "qc"
        for i in range(args.batch):
            bL = int(rng.integers(12, 25))
            inst = SyntheticSite(batch_model, bL, bL // 2 - 1, 256, seed=100 + i)
            lb = E.DeviceBoundary.upload(eng, inst.left_index, inst.random_cohorts(inst.left_index))
            rb = E.DeviceBoundary.upload(eng, inst.right_index, inst.random_cohorts(inst.right_index))
            pl = E.SigmaPlan(eng, inst.site_desc, inst.mpo_desc, lb, rb, algorithm=E.SIGMA_GROUPED)
            x = torch.from_numpy(inst.random_psi()).to(eng.torch_device)
            plans.append(pl)
            psis.append(x)
            sigs.append(torch.zeros_like(x))
            keep += [lb, rb, inst]
            flops += pl.flops
        res = {}
        for name, pc in (("f64", E.PREC_F64), (report, prec_code)):
            for mode in ("sequential", "batched"):
                def once():
                    if mode == "batched":
                        E.apply_batched(eng, plans, psis, sigs, precision=pc)
                    else:
                        for pl, x, y in zip(plans, psis, sigs):
                            pl.apply(x, y, precision=pc)
                for _ in range(3):   # eager, capture, warm replay
                    once()
                eng.sync()
                ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ev0.record(eng.stream)
                for _ in range(args.steps):
                    once()
                ev1.record(eng.stream)
                eng.sync()
                ms = ev0.elapsed_time(ev1) / args.steps
                res["%s_%s" % (name, mode)] = {"ms_per_batch_step": ms, "tflops": flops / (ms * 1e-3) / 1e12}
                log("batch %s %s: %.3f ms per step of %d instances, %.2f TFLOP/s" % (name, mode, ms, args.batch,
                                                                                   res["%s_%s" % (name, mode)]["tflops"]))
        batch = {"instances": args.batch, "model": "ab initio (synthetic integrals, %s), L uniform in [12, 24], "
                 "m=256, middle site" % ("SU2U1" if batch_model == "qc_su2" else "TwoU1"),
                 "flops_per_step": flops, "runs": res}

    adjoint = adjoint_workload(eng, args, model == "qc_su2") if args.adjoint else None
    hsweep = None
    if not args.no_headline_sweep and model in ("qc", "qc_su2") and args.headline_sweeps > 0:
        plan.close()   # the sweep builds its own plans; release the timed plan's workspace first
        hsweep = headline_sweep(eng, args, report, prec_code, ws.chain, model, L, m)

    out = {
        "metric": "sweep wall-time & σ TFLOP/s at L=30 m=2048, mixed vs FP64 at |ΔE|≤1e-6 Eh",
        "value": pr["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pr["ms_per_sigma"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "bf16x3": "bf16x3 (fp32 accumulate)"}[report],
        "data": "synthetic",
        "config": {
            "workload": ("sigma (two-site H_eff psi) at the middle site of the ab initio chain, SU2U1, L=%d, N=%d, "
                         "singlet, m=%d multiplets, grouped MFMA path" % (L, L, m)) if model == "qc_su2" else
                        ("sigma (two-site H_eff psi) at the middle site of a %s chain, L=%d, bond dimension m=%d "
                         "(%s symmetry), grouped MFMA path" % (model, L, m, "TwoU1" if model in ("qc", "hubbard")
                                                               else "U1")),
            "model": model, "L": L, "m": m, "site": site,
            "bond_sectors": len(ws.bonds[site + 1].sizes),
            "mpo_bond_dim": int(ws.mpo[site].col_dim),
            "flops_per_sigma": plan.flops,
            "psi_elements": plan.n,
            "site_build_seconds": round(t_site, 2),
            "plan_seconds": round(t_plan, 3),
            "workspace_bytes": plan.workspace_bytes,
            "reported_precision": report,
        },
        "precisions": per,
        "speedup_vs_f64": per["f64"]["ms_per_sigma"] / pr["ms_per_sigma"] if "f64" in per else None,
        "sweeps": sweep,
        "headline_sweep": hsweep,
        "batch": batch,
        "adjoint": adjoint,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
