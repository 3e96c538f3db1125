"""Serving loop: a stream of coupled-dipole solve requests on one GPU.

A request asks for one solve (assemble M, factor, solve its T incidence
columns, propagate to its measurement points) of a configuration at a k0.
Requests arrive as a Poisson process and mix 2D TM, 2D TE and 3D vector
problems of very different sizes.  The server

  * admits a request when its HBM footprint fits the budget next to the
    requests in flight and the resident factors (first fit in arrival
    order; resident factors not in use are evicted least recently used
    first to make room);
  * runs up to `workers` requests at once, each worker thread with its own
    torch stream and library handle (the pivoted LU's leaf is capped at
    CUs / workers workgroups so concurrent leaves fit the device together);
  * keeps M and its factorisation resident after a request, keyed by
    (configuration, k0), so that a later request for the same key with new
    incidences skips assembly and factorisation.

Every kernel is deterministic and a request's result depends only on its
own inputs (one request uses a resident factor at a time), so a request
served in a busy mix returns the same bits as the same request served
alone.  Latency is measured from arrival to the end of the request's work
on its stream.
"""
import math
import threading
import time
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from . import engine as E
from . import refine
from . import workloads as W
from .pipeline import Problem, incident, source_spec

KINDS = ("2d_tm", "2d_te", "3d_vector")


@dataclass
class Request:
    rid: int
    kind: str
    n_points: int
    T: int
    key: int            # configuration and k0: requests with equal keys share M and its factorisation
    arrival: float      # seconds after the start of the run
    n_measure: int = 64

    @property
    def d(self):
        return E.DIM[self.kind]

    @property
    def n(self):
        return self.d * self.n_points


def make_trace(n_requests, rate, seed=0, mix=(0.5, 0.25, 0.25), n_range=(256, 16384), t_range=(1, 256),
               reuse=0.3, n_measure=64):
    """Poisson arrivals at `rate` per second; kinds drawn from `mix`
    (2D TM, 2D TE, 3D vector), N log-uniform in n_range, T uniform in
    t_range; a fraction `reuse` repeats the configuration and k0 of an
    earlier request (with its own T).  rate = inf puts every arrival at 0."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        if math.isfinite(rate) and i > 0:
            t += rng.exponential(1.0 / rate)
        T = int(rng.integers(t_range[0], t_range[1] + 1))
        if out and rng.uniform() < reuse:
            prev = out[int(rng.integers(0, len(out)))]
            out.append(Request(i, prev.kind, prev.n_points, T, prev.key, t, n_measure))
            continue
        kind = KINDS[int(rng.choice(len(KINDS), p=np.asarray(mix) / np.sum(mix)))]
        N = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        out.append(Request(i, kind, N, T, i + seed * 1000003, t, n_measure))
    return out


def workload(req):
    """The configuration of a request: N points uniform in a disk (2D) or
    ball (3D) at the density of 1000 points per unit disk / 4096 per unit
    ball, k0 * (mean nearest-neighbour spacing) uniform in [0.5, 1.5], and
    alpha from the radiative-damping rule of its kind; all drawn from the
    key, so equal keys give equal problems.  Incidences: T plane waves
    (2D: theta_j = 2 pi j / T; 3D: Fibonacci directions x 2 polarisations),
    fields at n_measure points on a circle / sphere of twice the radius."""
    rng = np.random.default_rng(req.key)
    N = req.n_points
    if req.kind in ("2d_tm", "2d_te"):
        rad = math.sqrt(N / 1000.0)
        pts = rad * W.unit_disk(N, rng)
    else:
        rad = (N / 4096.0) ** (1.0 / 3.0)
        pts = rad * W.unit_ball(N, rng)
    k0 = rng.uniform(0.5, 1.5) / W.mean_nn_spacing(pts)
    T = req.T
    kw = {}
    if req.kind in ("2d_tm", "2d_te"):
        th = 2 * np.pi * np.arange(T) / T
        tm = 2 * np.pi * np.arange(req.n_measure) / req.n_measure
        meas = 2 * rad * np.stack([np.cos(tm), np.sin(tm)], axis=1)
        kw["thetas"] = th
        alpha = W.alpha_2d_tm(k0) if req.kind == "2d_tm" else W.alpha_2d_te(k0)
    else:
        u = W.fibonacci_sphere(max(2, (T + 1) // 2))
        p1, p2 = W.polarisation_pairs(u)
        uu = np.repeat(u, 2, axis=0)[:T]
        pp = np.empty((2 * u.shape[0], 3))
        pp[0::2] = p1
        pp[1::2] = p2
        kw["u"] = torch.from_numpy(uu)
        kw["p"] = torch.from_numpy(pp[:T])
        meas = 2 * rad * W.fibonacci_sphere(max(2, req.n_measure))
        alpha = W.alpha_3d(k0)
    return W.Workload(req.kind, torch.from_numpy(pts), k0, alpha, "plane", measure=torch.from_numpy(meas), **kw)


def resident_bytes(n, method):
    """Bytes that stay resident for a factorised system of order n: M
    (complex128) and the factors: the complex128 LU with its diagonal-block
    inverses; the complex64 block LDU with its D inverses; or, for the split
    precisions, the D inverses and the panels stored split into bf16 parts
    (4 bytes per part and complex element, Lh and Uh), the complex64 working
    copy being released after the factorisation."""
    nblk = (n + 255) // 256
    if method == "fp64":
        return 16 * n * n + 16 * 256 * 256 * 2 * nblk
    parts = {"bf16x3": 2, "bf16x6": 3}.get(method, 0)
    if parts:
        return 16 * n * n + 8 * 256 * 256 * nblk + 2 * 4 * parts * 256 * 256 * nblk * (nblk - 1) // 2
    return 16 * n * n + 8 * n * n + 8 * 256 * 256 * nblk


def factor_bytes(n, method):
    """Bytes a factorisation needs while it runs, beyond its resident result:
    the split precisions' complex64 working copy, and the LDU workspace."""
    if method == "fp64":
        return 8 * 256 * n * 4
    np_ = 256 * ((n + 255) // 256)
    work = 8 * (4 * 256 * n + 2 * 256 * 256 + 4 * np_ * 32 * 6 * 2)
    return work + (8 * n * n if method in ("bf16x3", "bf16x6") else 0)


def transient_bytes(n, T, n_meas, method, restart=12):
    """Bytes a request needs while it runs, beyond its resident system."""
    rhs = 16 * n * T * 4 + 16 * n_meas * T * 2
    if method == "fp64":
        return rhs + 8 * 256 * n * 4
    # split-K partials of the FP64 Krylov products: at most two workgroups'
    # worth of 128 x 64 tiles per CU (1024 on 256 CUs), or 16 splits
    splitk = 16 * min(16 * n * T, 1024 * 128 * 64)
    return rhs + 16 * n * T * (2 * restart + 2) + 8 * n * T * 2 + 8 * 256 * n * 4 + splitk


class _Entry:
    """M and its factorisation for one key, resident between requests."""

    def __init__(self, prob, method, tol, stream):
        self.prob = prob
        self.method = method
        n = prob.n
        dev = prob.points.device
        self.M = torch.empty((n, n), dtype=torch.complex128, device=dev)
        M32 = None if method == "fp64" else torch.empty((n, n), dtype=torch.complex64, device=dev)
        E.assemble(prob.kind, prob.points, prob.points, prob.k0, scale=-prob.alpha * prob.k0 * prob.k0,
                   flags=L.MG_F_ZERO_SELF | L.MG_F_DIAG, diag=1.0, out=self.M, out64=M32)
        if method == "fp64":
            self.solver = E.Factor(self.M)                  # LU in place of M
        else:
            self.solver = refine.MixedSolver(self.M, M32, method, tol=tol)
        self.bytes = resident_bytes(n, method)
        self.users = 0
        self.done = torch.cuda.Event()
        self.done.record(stream)

    def solve(self, B, X):
        if self.method == "fp64":
            X.copy_(B)
            self.solver.solve(X)
            return {"tier": "fp64", "iters": 0}
        return self.solver.solve(B, X)

    def held_bytes(self):
        """Bytes this entry holds now: a mixed-precision solver that fell
        back to the complex128 LU holds that factor (a copy of M) instead of
        its split factors."""
        n = self.prob.n
        if self.method != "fp64" and getattr(self.solver, "tier", None) == "fp64":
            return 16 * n * n + resident_bytes(n, "fp64")
        return resident_bytes(n, self.method)

    def check(self):
        """Raises if the entry's factorisation failed (after the stream synchronised)."""
        info = self.solver.info
        E.check_info(info, pivoted=(self.method == "fp64"))


class Server:
    """Serve requests on one device (see the module docstring)."""

    def __init__(self, device="cuda:0", workers=4, method="bf16x3", hbm_budget=200e9, tol=1e-12, cache=True,
                 keep_outputs=False):
        self.dev = torch.device(device)
        if self.dev.index is None:
            self.dev = torch.device("cuda", torch.cuda.current_device())
        self.workers = max(1, int(workers))
        self.method = method
        self.budget = float(hbm_budget)
        self.tol = tol
        self.cache = cache
        self.keep_outputs = keep_outputs

    def run(self, trace, time_scale=1.0):
        """Serve the trace (arrival times scaled by time_scale).  Returns one
        dict per request (in trace order): latency_ms, wait_ms, service_ms,
        hit (resident factor reused), stats, and X / E (on the device) when
        keep_outputs."""
        self.peak_reserved = 0.0
        cond = threading.Condition()
        pending = []                      # arrived, not started (arrival order)
        entries = OrderedDict()           # key -> _Entry (LRU order)
        state = {"reserved": 0.0, "finished": 0, "error": None}
        building = set()                  # keys whose system is being assembled and factorised
        results = [None] * len(trace)
        order = sorted(range(len(trace)), key=lambda i: trace[i].arrival)
        cus = torch.cuda.get_device_properties(self.dev).multi_processor_count
        cap = max(1, cus // self.workers) if self.workers > 1 else 0
        t_start = time.perf_counter()

        def now():
            return time.perf_counter() - t_start

        def admit():
            """Under cond: the first pending request that fits, with its
            reservation made; None if nothing can start now."""
            for j, i in enumerate(pending):
                r = trace[i]
                ent = entries.get(r.key)
                if (ent is not None and ent.users > 0) or r.key in building:
                    continue                                   # its factor is in use or on its way
                need = transient_bytes(r.n, r.T, r.d * r.n_measure, self.method)
                if ent is None:
                    need += resident_bytes(r.n, self.method) + factor_bytes(r.n, self.method)
                if state["reserved"] + need > self.budget:
                    evicted = False
                    for k in list(entries.keys()):           # evict idle resident factors, oldest first
                        if state["reserved"] + need <= self.budget:
                            break
                        e = entries[k]
                        if e.users == 0 and k != r.key:
                            e.done.synchronize()
                            state["reserved"] -= e.bytes
                            del entries[k]
                            evicted = True
                    if evicted:
                        # hand the evicted factors' memory back to the device, so that
                        # allocations outside torch's cache (the runtime's) find room
                        torch.cuda.empty_cache()
                if state["reserved"] + need > self.budget:
                    if not state["reserved"] and not entries:
                        raise MemoryError("request %d needs %.1f GB, over the %.1f GB budget" % (
                            r.rid, need / 1e9, self.budget / 1e9))
                    continue
                state["reserved"] += need
                self.peak_reserved = max(self.peak_reserved, state["reserved"])
                del pending[j]
                if ent is not None:
                    ent.users += 1
                    entries.move_to_end(r.key)
                else:
                    building.add(r.key)
                return i, need, ent
            return None

        def worker():
            torch.cuda.set_device(self.dev)
            stream = torch.cuda.Stream(device=self.dev)
            ws = E.Workspace(self.dev, leaf_blocks=cap)
            try:
                with ws, torch.cuda.stream(stream):
                    while True:
                        with cond:
                            while True:
                                if state["error"] is not None or state["finished"] == len(trace):
                                    return
                                got = admit()
                                if got is not None:
                                    break
                                cond.wait(timeout=0.05)
                        i, need, ent = got
                        r = trace[i]
                        t0 = now()
                        hit = ent is not None
                        if hit:
                            stream.wait_event(ent.done)
                            prob = ent.prob
                            w = workload(r)
                            E0s = incident(r.kind, prob.points, w)
                        else:
                            w = workload(r)
                            pts = w.points.to(self.dev)
                            meas = w.measure.to(self.dev)
                            prob = Problem(r.kind, pts, w.k0, w.alpha, incident(r.kind, pts, w), meas)
                            E0s = prob.E0s
                            ent = _Entry(prob, self.method, self.tol, stream)
                        X = torch.empty_like(E0s)
                        st = ent.solve(E0s, X)
                        # fields at the measurement points, the incident part evaluated in the final pass
                        _, beam, dirs, pol = source_spec(r.kind, w, self.dev)
                        F = torch.empty((prob.d * prob.measure.shape[0], X.shape[1]), dtype=X.dtype, device=self.dev)
                        E.propagate_src(r.kind, prob.measure, prob.points, prob.k0, prob.alpha * prob.k0 * prob.k0,
                                        X, F, beam, dirs, pol, w.beam_waist)
                        ent.done.record(stream)
                        stream.synchronize()
                        t1 = now()
                        E.check_barriers(self.dev)          # this worker's handle (already synchronised)
                        if not hit:
                            ent.check()
                        res = {"rid": r.rid, "kind": r.kind, "n": r.n, "T": r.T, "hit": hit,
                               "wait_ms": 1e3 * (t0 - r.arrival * time_scale),
                               "service_ms": 1e3 * (t1 - t0),
                               "latency_ms": 1e3 * (t1 - r.arrival * time_scale), "stats": st}
                        if self.keep_outputs:
                            res["X"], res["E"] = X, F
                        with cond:
                            results[i] = res
                            state["reserved"] -= need
                            held = ent.held_bytes()
                            if held != ent.bytes:
                                # a fallback to the complex128 LU changed what the entry holds:
                                # account for it from now on (the fallback's transient copy
                                # during the solve was not reserved in advance)
                                if hit:
                                    state["reserved"] += held - ent.bytes
                                ent.bytes = held
                            if hit:
                                ent.users -= 1
                            else:
                                building.discard(r.key)
                                if self.cache:
                                    entries[r.key] = ent
                                    state["reserved"] += ent.bytes
                                    self.peak_reserved = max(self.peak_reserved, state["reserved"])
                            state["finished"] += 1
                            cond.notify_all()
                        del X, F, ent
            except BaseException as e:
                with cond:
                    if state["error"] is None:
                        state["error"] = e
                    cond.notify_all()
            finally:
                stream.synchronize()
                ws.close()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.workers)]
        for t in threads:
            t.start()
        # release arrivals on time
        for i in order:
            delay = trace[i].arrival * time_scale - now()
            if delay > 0:
                time.sleep(delay)
            with cond:
                if state["error"] is not None:
                    break
                pending.append(i)
                cond.notify_all()
        for t in threads:
            t.join()
        if state["error"] is not None:
            raise state["error"]
        with cond:
            for e in entries.values():
                e.done.synchronize()
            entries.clear()
        return results


def percentiles(results, qs=(50, 99)):
    lat = np.array([r["latency_ms"] for r in results])
    return {"p%d_ms" % q: float(np.percentile(lat, q)) for q in qs}
