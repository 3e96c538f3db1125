"""Ensembles of independent coupled-dipole solves on one GPU.

An ensemble is many (configuration, k0) problems, each a full solve:
assemble M, factor, solve its T incidence columns (and propagate when the
workloads carry measurement points).  magreete runs such ensembles one
problem at a time (one `solve` per configuration in its driver loop,
magreete.py:346-452); small systems then leave most of the GPU idle, since
a factorisation of n = 1500 keeps only a few CUs busy between its
synchronisations.  Here `streams` worker threads, each with its own torch
stream and its own library handle (look-ahead stream, leaf exchange
buffers), take the problems largest first (longest-processing-time order)
and run side by side.

The pivoted LU's leaf kernel synchronises its workgroups through a grid
barrier, so every handle caps its leaf at (CUs / streams) workgroups: the
leaves that run at the same time then always fit on the device together.
Every kernel is deterministic, so a problem's result does not depend on
what runs beside it; the leaf cap changes the factorisation's blocking,
hence its rounding, so results match a lone solve with the default cap to
rounding, not bit for bit.
"""
import threading
import time

import torch

from . import _lib as L
from . import engine as E
from .pipeline import Solver, problem


def cost(w):
    """Relative cost of a workload: the factorisation's (d N)^3."""
    return float(w.d * w.points.shape[0]) ** 3


class Ensemble:
    """Run many workloads concurrently on one device.

    method: Solver method ('fp64', 'c64', 'bf16x3'); streams: concurrent
    solves; measure: propagate to the workloads' measurement points.

    grouped: instead of concurrent per-problem solves, the workloads go in
    groups (largest first, at most 256 problems and group_bytes of matrices
    each) through the grouped LU, which factors all problems of a group in
    lockstep, one launch per step of the blocked algorithm (mg_group_*):
    'fp64' solves with the complex128 pivoted factors directly; 'bf16x3'
    factors complex64 copies (trailing updates on BF16x3 products) and
    refines every problem with FP64 residuals b - M x (classic iterative
    refinement, the group's solve replayed from a captured graph) until
    max_t ||b_t - M x_t|| / ||b_t|| <= tol; a problem that stops improving
    is solved by the FP64 pivoted LU instead.  tune: autotune the group's
    tunables (cached per size class)."""

    def __init__(self, device="cuda:0", streams=4, method="fp64", tol=1e-12, measure=False, grouped=False,
                 group_bytes=None, tune=True, max_refine=12):
        # group_bytes: matrix bytes per group (default 0.6 of the device's memory)
        self.dev = torch.device(device)
        if self.dev.index is None:
            self.dev = torch.device("cuda", torch.cuda.current_device())
        self.streams = max(1, int(streams))
        self.method = method
        self.tol = tol
        self.measure = measure
        self.grouped = grouped
        self.group_bytes = group_bytes
        self.tune = tune
        self.max_refine = max_refine
        if grouped and method not in ("fp64", "bf16x3"):
            raise ValueError("grouped ensembles run 'fp64' or 'bf16x3'")

    def run(self, workloads):
        """Solve every workload.  Returns, in input order, dicts with 'X'
        (d N x T complex128 on the device), 'out' (measurement fields or None),
        'stats', 'info' (device int) and 'ms' (time on its worker's stream;
        grouped: the group's time)."""
        if self.grouped:
            return self._run_grouped(workloads)
        order = sorted(range(len(workloads)), key=lambda i: -cost(workloads[i]))
        results = [None] * len(workloads)
        lock = threading.Lock()
        cursor = [0]
        errors = []
        cus = torch.cuda.get_device_properties(self.dev).multi_processor_count
        cap = max(1, cus // self.streams) if self.streams > 1 else 0

        def worker():
            torch.cuda.set_device(self.dev)
            stream = torch.cuda.Stream(device=self.dev)
            ws = E.Workspace(self.dev, leaf_blocks=cap)
            try:
                with ws, torch.cuda.stream(stream):
                    while True:
                        with lock:
                            if cursor[0] >= len(order) or errors:
                                break
                            i = order[cursor[0]]
                            cursor[0] += 1
                        w = workloads[i]
                        t0 = torch.cuda.Event(enable_timing=True)
                        t1 = torch.cuda.Event(enable_timing=True)
                        t0.record(stream)
                        s = Solver(problem(w, self.dev, measure=self.measure), self.method, tol=self.tol)
                        out = s.run()
                        t1.record(stream)
                        results[i] = {"X": s.X, "out": out, "stats": dict(s.stats), "info": s.info,
                                      "events": (t0, t1)}
                        del s
                    E.check_barriers(self.dev)         # this worker's handle; synchronises its stream
                stream.synchronize()
            except BaseException as e:   # reported to the caller after every worker stopped
                errors.append(e)
                stream.synchronize()
            finally:
                ws.close()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.streams)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        for r in results:
            t0, t1 = r.pop("events")
            r["ms"] = t0.elapsed_time(t1)
            E.check_info(r["info"], pivoted=(self.method == "fp64"))
        return results

    # -- grouped ------------------------------------------------------------
    def _groups(self, workloads):
        """Largest first, at most GroupLU.MAX problems and group_bytes of
        matrices (complex128, plus for bf16x3 the complex64 copies and the
        pre-split update operands, three buffers of n x nb x 8 bytes at panel
        widths nb up to 384) per group."""
        order = sorted(range(len(workloads)), key=lambda i: -cost(workloads[i]))
        bf = self.method == "bf16x3"
        per = 16 + (8 if bf else 0)
        cap = self.group_bytes
        if cap is None:
            cap = 0.6 * torch.cuda.get_device_properties(self.dev).total_memory
        groups, cur, used = [], [], 0
        for i in order:
            w = workloads[i]
            n = w.d * w.points.shape[0]
            b = per * n * n + (3 * 384 * 8 * ((n + 127) // 128 * 128) if bf else 0)
            if cur and (len(cur) >= E.GroupLU.MAX or used + b > cap):
                groups.append(cur)
                cur, used = [], 0
            cur.append(i)
            used += b
        if cur:
            groups.append(cur)
        return groups

    def _run_grouped(self, workloads):
        results = [None] * len(workloads)
        for idx in self._groups(workloads):
            self._group(workloads, idx, results)
        for r in results:
            E.check_info(r["info"], pivoted=True)
        return results

    def _group(self, workloads, idx, results):
        dev = self.dev
        marks = []                       # (phase, event) at the end of each phase

        def mark(name):
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            marks.append((name, e))

        mark("start")
        probs = [problem(workloads[i], dev, measure=self.measure) for i in idx]
        Ms, M32s = [], []
        for p in probs:
            M = torch.empty((p.n, p.n), dtype=torch.complex128, device=dev)
            M32 = torch.empty((p.n, p.n), dtype=torch.complex64, device=dev) if self.method == "bf16x3" else None
            flags = L.MG_F_ZERO_SELF | L.MG_F_DIAG | (L.MG_F_BESSEL_EXACT if p.bessel_exact else 0)
            E.assemble(p.kind, p.points, p.points, p.k0, scale=-p.alpha * p.k0 * p.k0, flags=flags, diag=1.0,
                       out=M, out64=M32)
            Ms.append(M)
            M32s.append(M32)
        mark("assemble")
        T = probs[0].T
        if any(p.T != T for p in probs):
            raise ValueError("a group's problems share the number of incidence columns")
        offs = [0]
        for p in probs:
            offs.append(offs[-1] + p.n * T)
        view = lambda flat, i: flat[offs[i]:offs[i + 1]].view(probs[i].n, T)
        X = torch.empty(offs[-1], dtype=torch.complex128, device=dev)
        stats = [{"iters": 0, "fallback": False} for _ in probs]
        if self.method == "fp64":
            for i, p in enumerate(probs):
                view(X, i).copy_(p.E0s)
            g = E.GroupLU(Ms, [view(X, i) for i in range(len(probs))], prec="c128")
            if self.tune:
                g.tune()
            mark("tune")
            g.factor()
            mark("factor")
            g.solve()
            mark("solve")
            info = g.info
        else:
            B = torch.empty_like(X)
            for i, p in enumerate(probs):
                view(B, i).copy_(p.E0s)
            X.zero_()
            R = B.clone()
            D = torch.empty(offs[-1], dtype=torch.complex64, device=dev)
            g = E.GroupLU(M32s, [view(D, i) for i in range(len(probs))], prec="bf16x3")
            if self.tune:
                g.tune()
            mark("tune")
            g.factor()
            mark("factor")
            # the FP64 residual group: R_i -= M_i X_i and its column norms, one launch each
            gr = E.GroupLU(Ms, [view(R, i) for i in range(len(probs))], prec="c128")
            gr.set_aux([view(X, i) for i in range(len(probs))])
            bnorm = gr.colnorms()
            bnorm = torch.where(bnorm == 0, torch.ones_like(bnorm), bnorm)
            live = list(range(len(probs)))
            last = [float("inf")] * len(probs)
            fall = []
            for it in range(self.max_refine):
                # every problem takes the step (one launch serves them all);
                # a converged one only moves by its own residual's correction
                D.copy_(R)
                g.solve(graph=True)
                X.add_(D)
                R.copy_(B)
                gr.residual()
                worst = (gr.colnorms() / bnorm).max(dim=1).values.cpu().tolist()
                nxt = []
                for i in live:
                    stats[i]["iters"] = it + 1
                    stats[i]["residual"] = worst[i]
                    if worst[i] <= self.tol:
                        continue
                    if worst[i] > 0.5 * last[i]:       # stopped improving
                        fall.append(i)
                        continue
                    last[i] = worst[i]
                    nxt.append(i)
                live = nxt
                if not live:
                    break
            for i in range(len(probs)):
                stats[i]["residual"] = worst[i]      # after the last step every problem took
            gr.close()
            fall += live
            mark("refine")
            info = g.info
            infos = [info[j:j + 1] for j in range(len(probs))]
            # these problems get the FP64 pivoted LU of their complex128 matrix
            for i in fall:
                F = E.Factor(Ms[i])
                view(X, i).copy_(probs[i].E0s)
                F.solve(view(X, i))
                infos[i] = F.info
                stats[i]["fallback"] = True
            mark("fallback")
        if self.method == "fp64":
            infos = [info[j:j + 1] for j in range(len(probs))]
        g.close()
        outs = [None] * len(probs)
        for i, p in enumerate(probs):
            if p.measure is not None:
                out = torch.empty((p.d * p.measure.shape[0], T), dtype=torch.complex128, device=dev)
                out.copy_(p.E0m)
                flags = L.MG_F_BESSEL_EXACT if p.bessel_exact else 0
                E.propagate(p.kind, p.measure, p.points, p.k0, p.alpha * p.k0 * p.k0, view(X, i), out, flags=flags)
                outs[i] = out
        mark("propagate")
        marks[-1][1].synchronize()
        ms = marks[0][1].elapsed_time(marks[-1][1])
        phases = {name: marks[q - 1][1].elapsed_time(e) for q, (name, e) in enumerate(marks) if q > 0}
        for j, i in enumerate(idx):
            st = dict(stats[j])
            st["group"] = len(idx)
            st["group_phases_ms"] = phases
            results[i] = {"X": view(X, j), "out": outs[j], "stats": st, "info": infos[j], "ms": ms}
