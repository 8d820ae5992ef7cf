"""Serving loop: continuous batching of solver cases on one GPU.

The reference runs one case per process (``main.f90``); a parameter study
runs many such processes one after another.  This module serves such cases
as requests instead: every admitted case lives in one :class:`lbmib.Batch`,
so all running cases advance together, every kernel of a step covering all
of them in one launch.

* Requests queue in arrival order (FIFO).  Before each window of
  ``window`` steps the server admits queued requests while the device memory
  of the running cases plus the new one stays within ``hbm_budget`` bytes
  (and at most ``max_cases`` run at once).  A case's footprint is estimated
  from its lattice before it is created and taken from the library
  (``lbmib_device_bytes``) once it exists.
* An admitted case is created on the device, joins the batch at its step 0
  and stays resident until its last step; finished cases leave the batch at
  the end of the window in which they finish, and their force history is
  the result.  Cases of one batch join at multiples of
  ``lbmib.GEOMETRY_BATCH`` steps (the geometry batch), which a window of a
  multiple of that many steps guarantees.
* A case's result does not depend on what else runs: the per-case
  arithmetic is the single-case path's (checked bit for bit by the GPU tests).

The scheduling logic talks to a backend (``GPUBackend`` here; the CPU tests
use a stand-in) so that it can be tested without a device.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Callable, Dict, List, Optional

import numpy as np

from . import lbmib

# device bytes per lattice node: populations (two state planes of 9 and the
# middle plane of the two-step passes) + rho/u/v
# planes of the last two recorded steps
BYTES_PER_NODE = {lbmib.PREC_FP64: 3 * 9 * 8 + 2 * 3 * 8, lbmib.PREC_FP32DEV: 3 * 9 * 4 + 2 * 3 * 8}
# the middle plane of the two-step passes, allocated on a case's first pass
MID_PLANE_BYTES = {lbmib.PREC_FP64: 9 * 8, lbmib.PREC_FP32DEV: 9 * 4}


@dataclasses.dataclass
class Result:
    tag: object
    history: np.ndarray      # force rows, one per step (lbmib_get_history)
    steps: int
    nodes: int
    arrival: float           # seconds on the server clock
    admitted: float
    finished: float

    @property
    def time_to_result(self) -> float:
        return self.finished - self.arrival


@dataclasses.dataclass
class _Job:
    tag: object
    case: object
    arrival: float
    est_bytes: int
    sim: object = None
    admitted: float = 0.0
    bytes: int = 0


class GPUBackend:
    """The device side: cases as :class:`lbmib.LBMIB` handles in one batch."""

    def __init__(self, device: int = 0, precision: int = lbmib.PREC_FP32DEV, shape: Optional[dict] = None):
        self.device = device
        self.precision = precision
        self.batch = lbmib.Batch(device)
        if shape:
            self.batch.set_shape(**{k: v for k, v in shape.items() if k != "ms_per_step"})

    def estimate_bytes(self, case) -> int:
        p = case.params
        n = p.nx * p.ny
        return int(BYTES_PER_NODE[self.precision] * n * 1.1) + (64 << 20)

    def open(self, case):
        return lbmib.LBMIB(case, precision=self.precision, device=self.device)

    def device_bytes(self, sim) -> int:
        # with the middle plane its first window will allocate
        p = sim.case.params
        return sim.device_bytes() + MID_PLANE_BYTES[self.precision] * p.nx * p.ny

    def add(self, sim):
        self.batch.add(sim)

    def remove(self, sim):
        self.batch.remove(sim)

    def step(self, n: int):
        self.batch.step(n)

    def steps_done(self, sim) -> int:
        return sim.steps_done

    def max_steps(self, sim) -> int:
        return sim.max_steps()

    def result(self, sim) -> np.ndarray:
        return sim.history()

    def close_case(self, sim):
        sim.close()

    def close(self):
        self.batch.close()


class Server:
    """Continuous-batching scheduler over a backend (see the module notes)."""

    def __init__(self, backend, hbm_budget: int, max_cases: int = 256, window: int = 32,
                 clock: Callable[[], float] = time.perf_counter):
        if window <= 0 or window % lbmib.GEOMETRY_BATCH:
            raise ValueError("window must be a positive multiple of %d steps" % lbmib.GEOMETRY_BATCH)
        self.be = backend
        self.budget = int(hbm_budget)
        self.max_cases = int(max_cases)
        self.window = int(window)
        self.clock = clock
        self.t0 = clock()
        self.queue: collections.deque = collections.deque()
        self.running: List[_Job] = []
        self.results: List[Result] = []
        self.busy_s = 0.0          # time spent stepping the batch
        self.windows = 0
        self.node_steps = 0        # lattice-node updates performed
        self.peak_bytes = 0
        self.peak_cases = 0
        self.case_windows = 0      # sum over windows of running cases

    def now(self) -> float:
        return self.clock() - self.t0

    # ---- requests ----
    def submit(self, case, tag=None, arrival: Optional[float] = None):
        """Queue a case; ``arrival`` defaults to now (server clock)."""
        est = self.be.estimate_bytes(case)
        if est > self.budget:
            raise ValueError("case %s needs about %d bytes, more than the whole budget of %d"
                             % (getattr(case, "name", tag), est, self.budget))
        self.queue.append(_Job(tag, case, self.now() if arrival is None else arrival, est))

    def resident_bytes(self) -> int:
        return sum(j.bytes for j in self.running)

    def admit(self) -> int:
        """Admit queued requests in arrival order while memory and slots allow."""
        n = 0
        while self.queue and len(self.running) < self.max_cases:
            job = self.queue[0]
            if self.resident_bytes() + job.est_bytes > self.budget:
                break
            self.queue.popleft()
            job.sim = self.be.open(job.case)
            job.bytes = self.be.device_bytes(job.sim)
            self.be.add(job.sim)
            job.admitted = self.now()
            self.running.append(job)
            n += 1
        self.peak_bytes = max(self.peak_bytes, self.resident_bytes())
        self.peak_cases = max(self.peak_cases, len(self.running))
        return n

    def step_window(self) -> List[Result]:
        """Admit, advance every running case one window, retire finished ones."""
        self.admit()
        if not self.running:
            return []
        before = {id(j): self.be.steps_done(j.sim) for j in self.running}
        t = self.clock()
        self.be.step(self.window)
        self.busy_s += self.clock() - t
        self.windows += 1
        self.case_windows += len(self.running)
        done, keep = [], []
        for j in self.running:
            s = self.be.steps_done(j.sim)
            p = j.case.params
            self.node_steps += (s - before[id(j)]) * p.nx * p.ny
            (done if s >= self.be.max_steps(j.sim) else keep).append(j)
        out = []
        for j in done:
            hist = self.be.result(j.sim)
            self.be.remove(j.sim)
            self.be.close_case(j.sim)
            p = j.case.params
            r = Result(j.tag, hist, len(hist), p.nx * p.ny, j.arrival, j.admitted, self.now())
            self.results.append(r)
            out.append(r)
        self.running = keep
        return out

    def idle(self) -> bool:
        return not self.queue and not self.running

    def run(self, requests, sleep: Callable[[float], None] = time.sleep,
            progress: Optional[Callable[["Server"], None]] = None) -> List[Result]:
        """Serve ``requests`` (objects with .case, .arrival in seconds from
        now, .index) as they arrive on the server clock; returns the results
        in completion order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.t0 = self.clock()
        k = 0
        while k < len(pending) or not self.idle():
            now = self.now()
            while k < len(pending) and pending[k].arrival <= now:
                r = pending[k]
                self.submit(r.case, tag=r.index, arrival=r.arrival)
                k += 1
            if self.running or self.queue:
                self.step_window()
                if progress:
                    progress(self)
            elif k < len(pending):
                sleep(max(0.0, pending[k].arrival - self.now()))
        return self.results

    def stats(self) -> Dict[str, float]:
        ttr = np.array([r.time_to_result for r in self.results]) if self.results else np.zeros(1)
        wait = np.array([r.admitted - r.arrival for r in self.results]) if self.results else np.zeros(1)
        span = max((r.finished for r in self.results), default=0.0) - min((r.arrival for r in self.results),
                                                                          default=0.0)
        return {
            "requests": len(self.results),
            "p50_time_to_result_s": float(np.percentile(ttr, 50)),
            "p99_time_to_result_s": float(np.percentile(ttr, 99)),
            "mean_time_to_result_s": float(ttr.mean()),
            "p99_queue_wait_s": float(np.percentile(wait, 99)),
            "span_s": float(span),
            "aggregate_MLUPS": float(self.node_steps / max(span, 1e-9) / 1e6),
            "busy_fraction": float(self.busy_s / max(span, 1e-9)),
            "MLUPS_while_busy": float(self.node_steps / max(self.busy_s, 1e-9) / 1e6),
            "windows": self.windows,
            "mean_running_cases": float(self.case_windows / max(self.windows, 1)),
            "peak_running_cases": self.peak_cases,
            "peak_resident_GB": float(self.peak_bytes / 1e9),
        }

    def close(self):
        for j in self.running:
            self.be.remove(j.sim)
            self.be.close_case(j.sim)
        self.running = []
        self.be.close()
