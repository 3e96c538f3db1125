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

import torch

from . import engine as E
from .pipeline import Solver, problem


def cost(w):
    """Relative cost of a workload: the factorisation's (d N)^3."""
    return float(w.d * w.points.shape[0]) ** 3


class Ensemble:
    """Run many workloads concurrently on one device.

    method: Solver method ('fp64', 'c64', 'bf16x3'); streams: concurrent
    solves; measure: propagate to the workloads' measurement points."""

    def __init__(self, device="cuda:0", streams=4, method="fp64", tol=1e-12, measure=False):
        self.dev = torch.device(device)
        if self.dev.index is None:
            self.dev = torch.device("cuda", torch.cuda.current_device())
        self.streams = max(1, int(streams))
        self.method = method
        self.tol = tol
        self.measure = measure

    def run(self, workloads):
        """Solve every workload.  Returns, in input order, dicts with 'X'
        (d N x T complex128 on the device), 'out' (measurement fields or None),
        'stats', 'info' (device int) and 'ms' (time on its worker's stream)."""
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
