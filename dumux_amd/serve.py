"""Serving loop: requests (one porous-medium solve each) arrive over time and are continuously
batched on one GPU.

A request is admitted when its device footprint (dmx_estimate_bytes, batched form) fits the HBM
budget next to the requests already resident; admitted requests are prepared on host threads
(problem arrays, BCRS pattern, device upload) while the GPU iterates the batch, and join the
batch between two Newton phases. Finished requests leave the batch at once and free their
memory, so the next waiting request can be admitted. Per request the state (solution, previous
time level, Jacobian, Krylov vectors) stays on the device from admission to completion.

The reference runs one problem per process and has no serving layer; per request the
numerical control is that of NewtonSolver::solve(vars, timeLoop) (newtonsolver.hh:309-355).
"""
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import PRECOND_AMG, Batch, FVAssembler, Problem, estimate_bytes


class Request:
    """One solve: a `kind` (1: OneP, 2: TwoP) problem on an n-cell box with a seeded lognormal K,
    `nsteps` implicit-Euler steps of `dt` (linear_pde: one stationary step)."""

    def __init__(self, rid, kind, dim, n, seed, nsteps=5, dt=10.0, linear_pde=False, arrival=0.0, keep=False, **build):
        self.rid = rid
        self.kind, self.dim, self.n = kind, dim, tuple(n)
        self.seed = seed
        self.nsteps, self.dt, self.linear_pde = nsteps, dt, linear_pde
        self.arrival = arrival       # seconds after the start of the stream
        self.keep = keep             # keep the final solution on the host
        self.build = build
        self.cells = int(np.prod(self.n))
        self.bytes = 0
        self.t_admit = self.t_start = self.t_done = None
        self.finished = 0
        self.info = None
        self.solution = None

    def problem(self):
        return Problem(kind=self.kind, dim=self.dim, n=self.n, seed=self.seed, **self.build)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


class Server:
    def __init__(self, hbm_budget, device=0, linear=None, prep_threads=4, **newton_kw):
        self.budget = int(hbm_budget)
        self.device = device
        self.batch = Batch(device=device, linear=linear, **newton_kw)
        self.multigrid = bool(linear) and linear.get("precond") == PRECOND_AMG
        self.pool = ThreadPoolExecutor(max_workers=prep_threads)
        self.phases = 0

    def close(self):
        self.pool.shutdown(wait=True)
        self.batch.close()

    def _prepare(self, req):
        prob = req.problem()
        asm = FVAssembler(prob, device=self.device)
        asm.setSolution(prob.u0)
        return prob, asm

    def run(self, requests, realtime=True, on_phase=None):
        """Serve `requests` (arrival times relative to the call). realtime=False admits every
        request at once (closed batch). Returns the requests with their timings."""
        waiting = deque(sorted(requests, key=lambda r: r.arrival))
        for r in waiting:
            r.bytes = estimate_bytes(r.kind, r.dim, r.n, batched=True, multigrid=self.multigrid)
        preparing = []          # (req, future)
        active = {}             # req -> (prob, asm)
        resident = 0
        done = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while waiting or preparing or active:
            t = now()
            # admission in arrival order (no overtaking), gated by the HBM budget
            while waiting and (not realtime or waiting[0].arrival <= t):
                r = waiting[0]
                if resident + r.bytes > self.budget and (resident > 0 or preparing):
                    break
                waiting.popleft()
                resident += r.bytes
                r.t_admit = now()
                preparing.append((r, self.pool.submit(self._prepare, r)))
            # prepared requests join the batch
            still = []
            for r, fut in preparing:
                if fut.done():
                    prob, asm = fut.result()
                    self.batch.add(asm, nsteps=r.nsteps, dt=r.dt, linear_pde=r.linear_pde)
                    r.t_start = now()
                    active[r] = (prob, asm)
                else:
                    still.append((r, fut))
            preparing = still
            if not active:
                # nothing to iterate: wait for a preparation or the next arrival
                if preparing:
                    preparing[0][1].result()
                elif waiting:
                    time.sleep(max(0.0, min(waiting[0].arrival - now(), 0.05)))
                continue
            self.batch.step()
            self.phases += 1
            if on_phase:
                on_phase(self)
            for r in list(active):
                prob, asm = active[r]
                f, info = self.batch.status(asm)
                if f != 0:
                    r.finished, r.info = f, info
                    if r.keep:
                        r.solution = asm.solution()
                    self.batch.remove(asm)
                    asm.close()
                    del active[r]
                    resident -= r.bytes
                    r.t_done = now()
                    done.append(r)
        return done


def summarize(done, t_span=None):
    """Throughput (Newton cell-steps/s over the span from the first arrival to the last
    completion) and latency percentiles of served requests."""
    ok = [r for r in done if r.finished == 1]
    lat = np.array([r.latency for r in ok]) if ok else np.array([np.nan])
    if t_span is None:
        t_span = max(r.t_done for r in done) - min(r.arrival for r in done) if done else float("nan")
    cell_steps = sum(r.cells * r.info.newton_steps for r in ok)
    return {
        "requests": len(done), "failed": len(done) - len(ok),
        "fp64_fallback": sum(r.info.fp64_fallback for r in done if r.info is not None),
        "cell_steps_per_s": cell_steps / t_span if t_span > 0 else float("nan"),
        "requests_per_s": len(ok) / t_span if t_span > 0 else float("nan"),
        "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
        "span_s": t_span,
    }


def poisson_stream(rate, count, rng, cells_lo=1e4, cells_hi=2e6, nsteps=5, dt=10.0, seed0=0, **build):
    """`count` TwoP 3D requests with exponential inter-arrival times of mean 1/rate and cell
    counts log-uniform in [cells_lo, cells_hi] (cubes of the nearest integer side)."""
    t = 0.0
    reqs = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        cells = np.exp(rng.uniform(np.log(cells_lo), np.log(cells_hi)))
        side = max(2, int(round(cells ** (1.0 / 3.0))))
        reqs.append(Request(i, 2, 3, (side, side, side), seed=seed0 + i, nsteps=nsteps, dt=dt, arrival=t, **build))
    return reqs
