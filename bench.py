"""Benchmark: 3D vector coupled dipoles, N=4096 points in the unit ball
(system 12288 x 12288 complex), 128 Fibonacci directions x 2 polarisations
(T=256), fields at 2048 points on a sphere of radius 2.

A step is one solve: assemble M, factor, solve all 256 columns (+ GMRES-IR
for the mixed-precision methods), propagate to the 2048 measurement points.
Inputs (positions and incident fields) are resident in HBM before timing.
Prints one JSON line (rank 0, one GPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D N=4096 solves/s: mixed-prec+IR vs FP64 @≤1e-9 err; serving p50/p99 latency"
PEAK = {"fp64": 78.6, "c64": 157.3, "bf16x3": 2500.0 / 3.0, "bf16x6": 2500.0 / 6.0}   # TFLOP/s per product form
# The FP64 kernels multiply complex numbers with three real MFMA products (Gauss) where the algorithmic
# count is four (8 m n k flops): the form's peak in algorithmic flops is 4/3 of the FP64 MFMA peak.
PEAK_Z3 = 78.6 * 4.0 / 3.0
# peak of each trailing-update kernel class's product form (the FP64 64x64 register-staged tile multiplies
# with four real products, every other FP64 / FP32 kernel here with three)
PEAK_CLASS = {("fp64", "dma"): PEAK_Z3, ("fp64", "reg"): 78.6, ("c64", "reg"): 157.3 * 4.0 / 3.0,
              ("bf16x3", "dma"): PEAK["bf16x3"], ("bf16x3", "reg"): PEAK["bf16x3"], ("bf16x6", "reg"): PEAK["bf16x6"]}
# the factorisation's trailing-update launches (timed with HIP events on their stream), per kernel class
KERNEL = {"fp64": {"dma": "k_zgemm_dma_trail<64, true> (pivoted LU trailing update; 128x64 tiles, operands DMA'd to "
                          "LDS through three stages, C DMA'd late into the freed stages, 3-product complex MFMA)",
                   "reg": "k_cgemm_trail<double, 64, 64> (the updates too small for a tile per CU)"},
          "c64": {"reg": "k_cgemm_trail<float> (block-LDU trailing update; 128x64 tiles, 3-product complex MFMA)"},
          "bf16x3": {"dma": "k_ldu_update_dma<1|2, LATE> (block-LDU trailing update over pre-split panels, one or "
                            "two steps per launch; BF16x3 MFMA, 128x128 tiles, operands DMA'd to LDS, C read late "
                            "by DMA)",
                     "reg": "sdma::k_ldu_update_small<64> (the chain-side updates too small for a 128x128 tile "
                            "per CU: next diagonal blocks, block rows and columns of the next super-step; 64x32 "
                            "tiles, operands DMA'd to LDS two k-tiles ahead)"},
          "bf16x6": {"reg": "k_ldu_update<3, 64, 64, 2, 2> (block-LDU trailing update over pre-split panels, one or "
                            "two steps per launch; BF16x6 MFMA, 64x64 tiles)"}}
PRODUCT_KERNEL = ("k_zgemm_dma<64> + k_splitk_reduce<double> (the refinement's FP64 Krylov product M Z, "
                  "12288 x 256 x 12288 split over K in two; 128x64 tiles, operands DMA'd to LDS, 3-product "
                  "complex MFMA)")
# HBM bytes per launch (average over one solve of this workload), measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE on this script with --copy-probe 2 and scaled by the counters' reading for
# that 2 GiB copy (fetch x2.000, write x1.000), beside the algorithmic bytes of the same launches: C read
# and written once plus each operand panel read once (pivoted LU: 16 B per complex128 element each way;
# block LDU: 8 B per complex64 element each way, 8 B per pre-split BF16x3 panel element and k); the
# product: M and Z read once, the two K-split partials written and read once, the result written once.
TRAFFIC = {("fp64", "dma"): 2.181e9, ("fp64", "reg"): 6.44e7, ("c64", "reg"): 4.410e8,
           ("bf16x3", "dma"): 1.189e9, ("bf16x3", "reg"): 3.95e7, ("bf16x6", "reg"): 2.709e8, "product": 4.179e9}
ALGO_BYTES = {("fp64", "dma"): 1.345e9, ("fp64", "reg"): 4.93e7, ("bf16x3", "dma"): 6.7e8, "product": 2.718e9}


def colerr(x, ref):
    x = x.reshape(-1, x.shape[-1])
    ref = ref.reshape(-1, ref.shape[-1])
    return (torch.linalg.norm(x - ref, dim=0) / torch.linalg.norm(ref, dim=0)).max().item()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--methods", default="fp64,c64,bf16x3,bf16x6,bf16x3-norelax",
                    help="fp64, c64, bf16x3, bf16x6; a '-norelax' suffix takes every refinement product in FP64")
    ap.add_argument("--n-points", type=int, default=4096)
    ap.add_argument("--n-dirs", type=int, default=128)
    ap.add_argument("--n-measure", type=int, default=2048)
    ap.add_argument("--tol", type=float, default=1e-9,
                    help="relative FP64 residual per column at which the refinement stops (the error of every "
                         "method against the FP64 path is measured and reported beside it)")
    ap.add_argument("--batch-tol", type=float, default=1e-10,
                    help="refinement tolerance of the ensemble and serving measurements (their smaller, "
                         "differently conditioned systems reach a 1e-9 error only at a tighter residual)")
    ap.add_argument("--no-relax", action="store_true",
                    help="every refinement product in FP64 (default: the late products on split-bf16 MFMA once "
                         "the residual allows their error, refine.RELAX_LEVELS)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--ensemble", type=int, default=256,
                    help="configurations of the batched-ensemble measurement (0: skip)")
    ap.add_argument("--ensemble-runs", default="fp64_loop,fp64_streams4,bf16x3_streams4,fp64_grouped,bf16x3_grouped",
                    help="which ensemble measurements to run (the first one is the reference of the errors)")
    ap.add_argument("--serving", type=int, default=40, help="requests per serving load level (0: skip)")
    ap.add_argument("--ldos-sites", type=int, default=20000,
                    help="2D TE perturbed lattice at about this many sites, for each k0 a in --k0a: LDOS at every "
                         "site and 4096 probes, and the 256 incidence fields in mixed precision (0: skip)")
    ap.add_argument("--k0a", default="0.6,0.9,1.2,1.5", help="k0 a values of the lattice measurement")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="GiB copied device to device once before the workload (untimed): a transfer of known "
                         "size that calibrates profiler memory counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from magreete_amd import engine as E
    from magreete_amd import workloads as W
    from magreete_amd.pipeline import Problem, Solver

    dev = torch.device("cuda", 0)
    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    w = W.cfg_3d_vector(n_points=args.n_points, n_dirs=args.n_dirs, n_measure=args.n_measure, seed=args.seed)
    pts = w.points.to(dev)
    meas = w.measure.to(dev)
    # incident plane waves at the scatterers, generated on the device before timing (inputs); at the
    # measurement points the propagation evaluates them in its final pass
    from magreete_amd.Transmission3D import _pvec
    pol = _pvec(w.p)
    E0s = E.sources(3, False, pts, w.u, pol, w.k0, 0.2).reshape(-1, w.T).contiguous()
    src = (False, torch.as_tensor(w.u, dtype=torch.float64).to(dev).contiguous(), pol.to(dev).contiguous(), 0.2)
    prob = Problem("3d_vector", pts, w.k0, w.alpha, E0s, meas, None, source=src)
    n = prob.n
    print("# workload 3d_vector N=%d n=%d T=%d measure=%d k0=%.4f" % (args.n_points, n, w.T, args.n_measure, w.k0),
          file=sys.stderr, flush=True)

    results = {}
    outputs = {}
    for method in [m for m in args.methods.split(",") if m]:
        try:
            base = method.split("-")[0]
            relax = not args.no_relax and not method.endswith("-norelax")
            s = Solver(prob, method=base, tol=args.tol, relax=relax)
            for _ in range(args.warmup):
                s.run()
            torch.cuda.synchronize()
            E.profile_updates(dev, True)
            E.read_update_profile(dev)   # clear
            t0 = time.perf_counter()
            its = []
            for _ in range(args.steps):
                s.run()
                its.append(dict(s.stats))
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            cls = E.read_update_profile_classes(dev)
            E.profile_updates(dev, False)
            E.check_barriers(dev)
            info = int(s.info.item())
            launches = sum(v[0] for v in cls.values())
            flops = sum(v[1] for v in cls.values())
            ms = sum(v[2] for v in cls.values())
            results[method] = {
                "solves_per_s": args.steps / dt, "ms_per_step": 1e3 * dt / args.steps, "info": info,
                "update_launches": launches, "update_tflops": (flops / (ms * 1e-3) / 1e12) if ms > 0 else None,
                "update_ms": ms / args.steps, "refine": its[-1],
                "updates": {k: {"launches_per_step": v[0] / args.steps, "ms_per_step": v[2] / args.steps,
                                "tflops": v[1] / (v[2] * 1e-3) / 1e12} for k, v in cls.items() if v[0] > 0},
            }
            outputs[method] = (s.X.clone(), s.out.clone())
            if method == best_candidate(args):
                results[method]["phases"] = phase_rooflines(s, prob, args)
            print("# %s: %.3f solves/s, %.1f ms/step, trailing GEMM %.1f TF/s, stats %s" % (
                method, args.steps / dt, 1e3 * dt / args.steps, results[method]["update_tflops"] or 0, its[-1]),
                file=sys.stderr, flush=True)
            del s
            torch.cuda.empty_cache()
        except Exception as e:  # a method that is not built yet is reported, not fatal
            results[method] = {"error": str(e)[:300]}
            print("# %s failed: %s" % (method, e), file=sys.stderr, flush=True)

    # errors of every method against the FP64 GPU path, and of the FP64 path against the oracle
    errors = {}
    if "fp64" in outputs:
        X64, F64 = outputs["fp64"]
        for m, (X, F) in outputs.items():
            if m != "fp64":
                errors[m + "_vs_fp64"] = {"Ek": colerr(X, X64), "E": colerr(F, F64),
                                         "I": colerr(F.abs() ** 2, F64.abs() ** 2)}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import transmission as O
        torch.set_num_threads(os.cpu_count() or 1)
        # warm-up on a quarter-size problem, then the median of three full solves
        ww = W.cfg_3d_vector(n_points=max(64, args.n_points // 4), n_dirs=max(2, args.n_dirs // 4),
                             n_measure=max(16, args.n_measure // 4), seed=args.seed)
        ow = O.Transmission3D_vector(ww.points, source="plane")
        ow.propagate(ww.measure, ow.solve(ww.k0, ww.alpha, 0.0, ow.generate_source(ww.points, ww.k0, ww.u, ww.p, 0.2),
                                          self_interaction=False), ww.k0, ww.alpha,
                     ow.generate_source(ww.measure, ww.k0, ww.u, ww.p, 0.2))
        o = O.Transmission3D_vector(w.points, source="plane")
        cE0s = o.generate_source(w.points, w.k0, w.u, w.p, 0.2)
        cE0m = o.generate_source(w.measure, w.k0, w.u, w.p, 0.2)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            cEk = o.solve(w.k0, w.alpha, 0.0, cE0s, self_interaction=False)
            cE = o.propagate(w.measure, cEk, w.k0, w.alpha, cE0m)
            times.append(time.perf_counter() - t0)
        tc = float(np.median(times))
        cpu = {"value": 1.0 / tc, "unit": "solves/s", "cores": torch.get_num_threads(), "kind": "port",
               "sample": "full solves of this workload (assemble + torch.linalg.solve + propagate, complex128) by "
                         "the oracle: median of 3 after a quarter-size warm-up; %s s" % ", ".join(
                             "%.2f" % t for t in times),
               "host": host_info()}
        for m, (X, F) in outputs.items():
            errors[m + "_vs_oracle"] = {"Ek": colerr(X.cpu(), cEk), "E": colerr(F.cpu().reshape(cE.shape), cE),
                                        "I": colerr((F.cpu().abs() ** 2).reshape(cE.shape).sum(1),
                                                    (cE.abs() ** 2).sum(1))}
        print("# cpu oracle %.1f s, errors %s" % (tc, errors), file=sys.stderr, flush=True)

    # headline: fastest method whose error vs the FP64 path is <= 1e-9 (FP64 itself otherwise)
    best = "fp64" if "solves_per_s" in results.get("fp64", {}) else None
    for m in ("bf16x6", "bf16x3", "c64", "bf16x3-norelax"):
        r = results.get(m, {})
        e = errors.get(m + "_vs_fp64")
        if "solves_per_s" in r and e and max(e.values()) <= 1e-9 and not r["refine"].get("fallback"):
            if best is None or r["solves_per_s"] > results[best]["solves_per_s"]:
                best = m
    if best is None:
        print(json.dumps({"metric": METRIC, "value": None, "error": results}))
        sys.exit(1)
    rb = results[best]

    def kernel_rooflines(method):
        """Per-kernel rooflines of a method: its trailing-update classes and, for the mixed methods, the
        FP64 Krylov product (timed in phase_rooflines), each with its time per solve."""
        r = results.get(method, {})
        base = method.split("-")[0]
        out = {}
        for k, u in (r.get("updates") or {}).items():
            pk = PEAK_CLASS[(base, k)]
            out["update_" + k] = {
                "bound": "mfma", "achieved": u["tflops"], "peak": pk, "unit": "TFLOP/s",
                "frac": u["tflops"] / pk, "traffic": TRAFFIC.get((base, k)),
                "algorithmic_bytes": ALGO_BYTES.get((base, k)), "kernel": KERNEL[base][k],
                "launches_per_step": u["launches_per_step"], "ms_per_step": u["ms_per_step"]}
        ph = r.get("phases") or (results.get(best_candidate(args), {}).get("phases") or {})
        prod = ph.get("fp64_product")
        nprod = sum(1 for p in (r.get("refine") or {}).get("products", []) if p.startswith("fp64"))
        if prod and nprod and base != "fp64":
            out["fp64_product"] = dict(prod, launches_per_step=nprod, ms_per_step=nprod * prod["ms"])
        return out

    def roofline(method):
        """The dominant kernel of the method: the one with the most time per solve."""
        ks = kernel_rooflines(method)
        if not ks:
            return None
        name = max(ks, key=lambda k: ks[k]["ms_per_step"])
        return dict(ks[name], which=name)

    roof = roofline(best)
    speedup = {}
    if "solves_per_s" in results.get("fp64", {}):
        for m, r in results.items():
            if m != "fp64" and "solves_per_s" in r:
                speedup[m] = r["solves_per_s"] / results["fp64"]["solves_per_s"]
    ldos = run_ldos(args, dev) if args.ldos_sites > 0 else None
    ens = run_ensemble(args, dev) if args.ensemble > 0 else None
    srv = run_serving(args, dev) if args.serving > 0 else None
    line = {
        "metric": METRIC, "value": rb["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": rb["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {"fp64": "complex128", "c64": "complex64 LU + complex128 GMRES-IR",
                                       "bf16x3": "complex64 (bf16x3 products) LU + complex128 GMRES-IR",
                                       "bf16x6": "complex64 (bf16x6 products) LU + complex128 GMRES-IR"}[best.split("-")[0]]
        + ("" if best == "fp64" or args.no_relax or best.endswith("-norelax")
           else " (late Krylov products split-bf16 where the residual allows, FP64 residual check)"),
        "data": "synthetic (seeded uniform ball, Fibonacci directions)",
        "config": {"workload": "3d_vector N=%d (n=%d) T=%d (%d dirs x 2 pol) measure=%d" % (
            args.n_points, n, w.T, args.n_dirs, args.n_measure), "method": best, "k0": w.k0,
            "seed": args.seed, "tol": args.tol,
            "relaxed_products": not args.no_relax and not best.endswith("-norelax")},
        "roofline": roof, "roofline_fp64": roofline("fp64") if best != "fp64" else None,
        "rooflines_kernels": kernel_rooflines(best),
        "speedup_vs_fp64": speedup,
        "rooflines_other": dict(method=best_candidate(args),
                                **(results.get(best_candidate(args), {}).get("phases") or {})),
        "cpu_baseline": cpu, "methods": results, "errors": errors,
        "ensemble": ens, "serving": srv, "ldos_lattice": ldos,
    }
    print(json.dumps(line))


def host_info():
    """Host cores and CPU model of the machine the CPU baseline ran on."""
    info = {"nproc": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def best_candidate(args):
    """The method whose other kernels get rooflines (the expected headline method)."""
    ms = [m for m in args.methods.split(",") if m]
    for m in ("bf16x3-norelax", "bf16x3"):
        if m in ms:
            return m
    return ms[0] if ms else None


def phase_rooflines(s, prob, args):
    """One more solve with HIP events around its phases (on the solve's
    stream), and one preconditioner application of a fresh split factor:
    assembly (HBM write: the complex128 matrix and its complex64 copy),
    propagation (FP64 MFMA, 8 (d M)(d N) T flops, the Green tile evaluated
    in-kernel), and the factor's two triangular sweeps plus the
    block-diagonal product (4 n^2 T algorithmic flops per sweep)."""
    from magreete_amd import engine as E
    ph = {}
    s.run(phases=ph)
    n, T = prob.n, prob.T
    dM = prob.d * prob.measure.shape[0]
    out = {}
    wbytes = n * n * (16 + (8 if s.M32 is not None or s.method != "fp64" else 0))
    if ph.get("assemble"):
        t = ph["assemble"] * 1e-3
        out["assembly"] = {"bound": "hbm", "achieved": wbytes / t / 1e9, "peak": 8000.0, "unit": "GB/s",
                           "frac": wbytes / t / 8e12, "ms": ph["assemble"],
                           "bytes": wbytes, "kernel": "k_assemble<3D_VECTOR> (complex128 + complex64 outputs)"}
    if ph.get("propagate"):
        fl = 8.0 * dM * n * T
        t = ph["propagate"] * 1e-3
        out["propagation"] = {"bound": "mfma", "achieved": fl / t / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                              "frac": fl / t / 78.6e12, "ms": ph["propagate"], "flops": fl,
                              "kernel": "k_propagate<3D_VECTOR> (+ k_sum_parts_src: partial sums plus the incident field)"}
    if s.method != "fp64":
        # the refinement's FP64 Krylov product M Z (the same call as refine.py's), on the current stream
        st = torch.cuda.current_stream()
        Z = torch.randn((n, T), dtype=torch.complex128, device=s.M.device)
        Wz = torch.empty_like(Z)
        E.cgemm(s.M, Z, Wz, split=True)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record(st)
        for _ in range(reps):
            E.cgemm(s.M, Z, Wz, split=True)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        fl = 8.0 * n * n * T
        out["fp64_product"] = {"bound": "mfma", "achieved": fl / (ms * 1e-3) / 1e12, "peak": PEAK_Z3,
                               "unit": "TFLOP/s", "frac": fl / (ms * 1e-3) / (PEAK_Z3 * 1e12), "ms": ms,
                               "flops": fl, "traffic": TRAFFIC.get("product"),
                               "algorithmic_bytes": 16.0 * n * n + 16.0 * n * T * 6,
                               "kernel": PRODUCT_KERNEL}
        del Z, Wz
    if s.method in ("bf16x3", "bf16x6"):
        M32 = torch.empty((n, n), dtype=torch.complex64, device=s.M.device)
        E.convert(s.M, M32)
        F = E.Factor(M32, s.method, pivot=False)
        B = torch.randn((n, T), dtype=torch.complex64, device=s.M.device)
        X = B.clone()
        F.solve(X)
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record(st)
        for _ in range(reps):
            X.copy_(B)
            F.solve(X)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        fl = 8.0 * n * n * T
        pk = PEAK[s.method]
        out["ldu_solve"] = {"bound": "mfma", "achieved": fl / (ms * 1e-3) / 1e12, "peak": pk, "unit": "TFLOP/s",
                            "frac": fl / (ms * 1e-3) / (pk * 1e12), "ms": ms, "flops": fl,
                            "kernel": ("sdma::k_ldu_sweep_dma forward + k_d_product + sdma::k_ldu_sweep_dma "
                                       "backward (one preconditioner application, latency-bound chain of 256-row "
                                       "blocks)") if s.method == "bf16x3" else
                                      ("k_ldu_sweep forward + k_d_product + k_ldu_sweep backward (one "
                                       "preconditioner application, latency-bound chain of 256-row blocks)")}
        del F, M32
        torch.cuda.empty_cache()
    return out


def run_ldos(args, dev):
    """The 2D TE perturbed triangular lattice (unit spacing, sigma 0.1, ~args.ldos_sites sites in a disk)
    at every k0 a of args.k0a:
      * LDOS at every site and at 4096 probes through the drop-in's batch_LDOS_measurements (batches of
        1000 points, self_interaction=False: alpha is already dressed): the sites from the diagonal blocks
        of M^-1 (complex128 LU, in-place triangular inverses), the probes by solves against G_p^T;
      * the 256 incidence fields by the BF16x3 factor with FP64 GMRES refinement (relative residual
        1e-10), their error measured against a complex128 LU solve of the same system (untimed).
    With the CPU baseline on, the oracle's LDOS and field solve at 1500 sites, scaled by n^3."""
    from magreete_amd import workloads as W
    from magreete_amd.Transmission2D import Transmission2D_vector
    from magreete_amd.pipeline import Solver, problem
    out = {"workload": "2d_te perturbed triangular lattice, ~%d sites, LDOS at sites + 4096 probes, "
                       "256 incidence fields (BF16x3 LU + FP64 GMRES-IR, tol 1e-10)" % args.ldos_sites,
           "levels": {}}
    for k0a in [float(x) for x in args.k0a.split(",") if x]:
        w = W.cfg_2d_te_lattice(n_sites=args.ldos_sites, k0a=k0a, n_angles=256, n_probes=4096, seed=args.seed)
        n = 2 * w.points.shape[0]
        pts = torch.cat([w.points, w.probes]).to(dev)
        cls = Transmission2D_vector(w.points.to(dev), "plane")
        batches = torch.split(pts, 1000)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ld = cls.batch_LDOS_measurements(batches, w.k0, w.alpha, 0.1, self_interaction=False)
        torch.cuda.synchronize()
        t_ldos = time.perf_counter() - t0
        del cls, ld
        torch.cuda.empty_cache()
        prob = problem(w, dev, measure=False)
        s = Solver(prob, "bf16x3", tol=1e-10, relax=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run()
        torch.cuda.synchronize()
        t_fields = time.perf_counter() - t0
        X = s.X.clone()
        st = dict(s.stats)
        del s
        torch.cuda.empty_cache()
        s64 = Solver(prob, "fp64")
        s64.run()
        err = colerr(X, s64.X)
        del s64, X, prob
        torch.cuda.empty_cache()
        # LU (8/3) n^3, triangular inverses (8/3) n^3, probe solves 8 n^2 (2 x 4096)
        fl = (16.0 / 3.0) * n ** 3 + 8.0 * n * n * 2 * 4096
        lvl = {"n": n, "sites": int(w.points.shape[0]), "ldos_seconds": t_ldos, "ldos_tflops": fl / t_ldos / 1e12,
               "fields_seconds": t_fields, "fields_err_vs_fp64": err, "fields_products": st.get("products"),
               "fields_residual": st.get("residual"), "fields_tier": st.get("tier")}
        out["levels"]["k0a=%.1f" % k0a] = lvl
        print("# lattice k0a=%.1f n=%d: LDOS %.2f s (%.1f TF/s), fields %.2f s, err %.2e" % (
            k0a, n, t_ldos, lvl["ldos_tflops"], t_fields, err), file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        from oracle import transmission as O
        ws = W.cfg_2d_te_lattice(n_sites=1500, k0a=0.6, n_angles=256, n_probes=4096, seed=args.seed)
        o = O.Transmission2D_vector(ws.points, "plane")
        t0 = time.perf_counter()
        o.batch_LDOS_measurements(torch.split(torch.cat([ws.points, ws.probes]), 1000), ws.k0, ws.alpha, 0.1,
                                  self_interaction=False)
        E0 = o.generate_source(ws.points, ws.k0, ws.thetas, 0.2)
        o.solve(ws.k0, ws.alpha, 0.0, E0, self_interaction=False)
        tc = time.perf_counter() - t0
        ns = 2 * ws.points.shape[0]
        nfull = max(l["n"] for l in out["levels"].values()) if out["levels"] else ns
        out["cpu_oracle"] = {"seconds_at_n": tc, "n": ns, "seconds_scaled_n3": tc * (nfull / ns) ** 3,
                             "note": "one k0: LDOS (full inverse) + 256 fields; the reference form at full size "
                                     "needs M, eye and W (~77 GB) in host memory"}
    return out


def run_ensemble(args, dev):
    """Batched ensemble: args.ensemble 3D vector configurations (N log-uniform
    in [500, 3000], 16 directions x 2 polarisations), each assembled and
    solved: one at a time by the FP64 path (the loop), concurrently on 4
    streams by the FP64 and BF16x3 paths, and by the grouped lockstep LU
    (FP64 pivoted; BF16x3 factor + FP64 iterative refinement to
    args.batch_tol).  Errors of every batched result against the loop.  The
    grouped runs are preceded by one untimed run of the same ensemble, which
    autotunes the group's tunables (cached per size class)."""
    from magreete_amd import workloads as W
    from magreete_amd.ensemble import Ensemble
    ens = W.cfg_ensemble(n_configs=args.ensemble, seed=args.seed)
    out = {"workload": "%d x 3d_vector, N log-uniform [500, 3000], T=32" % len(ens), "unit": "configs/s"}
    ref = None
    runs = (("fp64_loop", "fp64", 1, False), ("fp64_streams4", "fp64", 4, False),
            ("bf16x3_streams4", "bf16x3", 4, False), ("fp64_grouped", "fp64", 1, True),
            ("bf16x3_grouped", "bf16x3", 1, True))
    wanted = set(args.ensemble_runs.split(","))
    for name, method, streams, grouped in runs:
        if name not in wanted:
            continue
        e = Ensemble(dev, streams=streams, method=method, tol=args.batch_tol, grouped=grouped)
        res = e.run(ens if grouped else ens[:4])
        del res
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = e.run(ens)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[name] = len(ens) / dt
        if ref is None:
            ref = [r["X"] for r in res]
        else:
            out[name + "_max_err_vs_loop"] = max(colerr(r["X"], x) for r, x in zip(res, ref))
        if grouped:
            out[name + "_groups"] = len(set(id(r["X"].untyped_storage()) for r in res))
            if method == "bf16x3":
                out[name + "_fallbacks"] = sum(bool(r["stats"]["fallback"]) for r in res)
                out[name + "_max_refine_iters"] = max(r["stats"]["iters"] for r in res)
        print("# ensemble %s: %.1f configs/s" % (name, out[name]), file=sys.stderr, flush=True)
        del res
        torch.cuda.empty_cache()
    if "fp64_loop" in out:
        out["grouped_vs_loop"] = {m: out[m + "_grouped"] / out["fp64_loop"] for m in ("fp64", "bf16x3")
                                  if m + "_grouped" in out}
    return out


def run_serving(args, dev):
    """Serving: capacity from a burst of args.serving requests, then Poisson
    arrivals at 0.5x, 0.8x and 1.1x of it (mix 50% 2D TM / 25% 2D TE / 25% 3D
    vector, N log-uniform in [256, 16384], T uniform in [1, 256], 30% reuse,
    HBM budget 200 GB, 4 workers, BF16x3 mixed path)."""
    import math
    from magreete_amd import serving as S
    srv = S.Server(dev, workers=4, method="bf16x3", hbm_budget=200e9, tol=args.batch_tol)
    srv.run(S.make_trace(4, math.inf, seed=args.seed + 99, n_range=(256, 2048)))
    burst = S.make_trace(args.serving, math.inf, seed=args.seed + 7)
    t0 = time.perf_counter()
    srv.run(burst)
    cap = args.serving / (time.perf_counter() - t0)
    out = {"requests_per_level": args.serving, "capacity_req_per_s": cap, "workers": 4, "hbm_budget_gb": 200}
    for f in (0.5, 0.8, 1.1):
        t0 = time.perf_counter()
        res = srv.run(S.make_trace(args.serving, cap * f, seed=args.seed + 8))
        dt = time.perf_counter() - t0
        lvl = S.percentiles(res)
        # achieved throughput over the level's wall time, and every result's FP64 residual b - M x
        # (relative, largest column; the refinement's tolerance is --batch-tol)
        resid = [r["stats"]["residual"] for r in res if "residual" in r["stats"]]
        lvl.update({"offered_req_per_s": cap * f, "achieved_req_per_s": len(res) / dt,
                    "achieved_solves_per_s": sum(r["T"] for r in res) / dt, "hits": sum(r["hit"] for r in res),
                    "max_residual": max(resid) if resid else None,
                    "fallbacks": sum(1 for r in res if r["stats"].get("fallback"))})
        out["load_%.1fx" % f] = lvl
        print("# serving %.1fx: %s" % (f, lvl), file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    main()
