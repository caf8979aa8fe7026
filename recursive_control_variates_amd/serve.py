"""Continuous-batching optimisation service.

Step requests ("advance job j by one iteration") arrive over time; jobs are
admitted while their device footprint fits the HBM budget; every engine
iteration advances, in one batched device iteration (librcv rcv_batch_*),
every admitted job that has a pending request. Jobs that arrive while others
run join at the next engine iteration (the batch is rebuilt with each job's
CV, Adam and two-state parameter state carried over on the device), finished
jobs leave and free their memory, and each job's CV state stays resident for
its whole life. A job's random streams depend only on its seed, iteration,
pixel and sample, so its images are the same alone and in any mix; its
gradients differ only by float-atomic summation order.

``run_simulation`` drives the service with a Poisson arrival process of step
requests (rate ``rate`` per second) over a job mix of surfaces and volumes
and reports p50/p99 step latency (arrival to completion) and Msamples/s.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from . import load_dict, traverse
from .batch import BatchOptimizer, JobSpec, _JobBuffers


@dataclass
class ServedJob:
    jid: int
    buffers: _JobBuffers
    iterations: int
    meta: Dict
    done: int = 0
    pending: collections.deque = field(default_factory=collections.deque)
    latencies: List[float] = field(default_factory=list)
    admitted_at: float = 0.0
    finished_at: Optional[float] = None
    footprint: int = 0


def make_job(seed: int, kind: str, method: str = "cv_ps", device="cuda", spp_ref_factor: int = 4,
             lr: float = 0.01):
    """Builds a JobSpec for one job of the served mix (scene, reference image
    of the target, initial value) and its description."""
    from . import scenes
    if kind == "volume":
        d, m = scenes.volume_job_dict(seed, twostate=(method == "cv_ps"))
        integ = {"type": "prbvolpath", "max_depth": m["max_depth"], "rr_depth": m["max_depth"]}
        clamp = (0.0, 1e9)
    else:
        d, m = scenes.surface_job_dict(seed, twostate=(method == "cv_ps"))
        integ = {"type": "prb", "max_depth": m["max_depth"], "rr_depth": 5}
        clamp = (0.0, 1.0)
    sc = load_dict(d, device=device)
    p = traverse(sc)
    key = m["key"]
    p[key] = torch.from_numpy(m["target"]).to(sc.device)
    p.update()
    p.update()
    plain = load_dict(dict(integ))
    ref = plain.render(sc, 0, 17843 + seed, min(256, spp_ref_factor * m["spp"]))
    p[key] = torch.from_numpy(m["init"]).to(sc.device)
    p.update()
    p.update()
    spec = JobSpec(scene=sc, key=key, init=m["init"], img_ref=ref, integrator=integ, method=method,
                   loss="L1", spp=m["spp"], spp_grad=m["spp"], seed=(seed * 1000003) & 0x7FFFFFFF, lr=lr,
                   clamp=clamp)
    return spec, m


def footprint_estimate(meta) -> int:
    """Device bytes a job will hold: parameter ring x3 + Adam m, v + gradient
    (6 parameter-sized arrays) and 13 image-sized arrays (reference, H, F,
    image, image gradient, 8 CV arrays), plus a fixed 1 MiB for the scene."""
    n_param = meta["grid"] ** 3 if meta["kind"] == "volume" else 64 * 64 * 3
    n_img = meta["side"] * meta["side"] * 3
    return 4 * (6 * n_param + 13 * n_img) + (1 << 20)


class Server:
    def __init__(self, budget_bytes: float = 200e9, device="cuda", use_graph: bool = True, block_size: int = 0,
                 queue_grab: int = 0, max_jobs: int = 1024):
        self.budget = int(budget_bytes)
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.use_graph = use_graph
        self.launch = (block_size, queue_grab)
        self.max_jobs = max_jobs
        self.jobs: List[ServedJob] = []          # admitted, in batch order
        self.waiting: collections.deque = collections.deque()  # (jid, builder, iterations, meta)
        self.finished: List[ServedJob] = []
        self.used = 0
        self.batch: Optional[BatchOptimizer] = None
        self.dirty = False
        self.stream = torch.cuda.Stream(device=self.device)
        self.samples = 0
        self.engine_iterations = 0
        self._next_jid = 0

    # ----------------------------------------------------------- clients
    def submit_job(self, builder: Callable[[], tuple], iterations: int, meta: Dict) -> int:
        """Queues a job for admission; ``builder()`` returns (JobSpec, meta)
        and is called when the job is admitted."""
        jid = self._next_jid
        self._next_jid += 1
        self.waiting.append((jid, builder, iterations, meta))
        return jid

    def submit_step(self, jid: int, t_arrival: float) -> bool:
        for j in self.jobs:
            if j.jid == jid:
                if j.done + len(j.pending) < j.iterations:
                    j.pending.append(t_arrival)
                    return True
                return False
        return False

    # ------------------------------------------------------------ engine
    def _admit(self, now: float):
        while self.waiting and len(self.jobs) < self.max_jobs:
            jid, builder, iters, meta = self.waiting[0]
            fp = footprint_estimate(meta)
            if self.used + fp > self.budget:
                break
            self.waiting.popleft()
            spec, m = builder()
            buf = _JobBuffers(spec, self.device)
            self.jobs.append(ServedJob(jid, buf, iters, m, admitted_at=now, footprint=fp))
            self.used += fp
            self.dirty = True

    def _retire(self, now: float):
        keep = []
        for j in self.jobs:
            if j.done >= j.iterations and not j.pending:
                j.finished_at = now
                j.final_param = None
                j.buffers = None          # device memory returns once the batch is rebuilt
                self.finished.append(j)
                self.used -= j.footprint
                self.dirty = True
            else:
                keep.append(j)
        self.jobs = keep

    def _rebuild(self):
        resume = None
        if self.batch is not None:
            old = {id(b): k for k, b in enumerate(self.batch.jobs)}
            resume = []
            states = {}
            for k in range(len(self.batch.jobs)):
                states[k] = self.batch.read_job(k)
            for j in self.jobs:
                k = old.get(id(j.buffers))
                resume.append(states[k] if k is not None else None)
            self.batch.close()
            self.batch = None
        if self.jobs:
            self.batch = BatchOptimizer([j.buffers for j in self.jobs], device=self.device, use_graph=self.use_graph,
                                        block_size=self.launch[0], queue_grab=self.launch[1], resume=resume,
                                        stream=self.stream)
        self.dirty = False

    def engine_step(self, now: float) -> int:
        """One continuous-batching iteration; returns the number of steps served."""
        self._retire(now)
        self._admit(now)
        if self.dirty:
            self._rebuild()
        if self.batch is None:
            return 0
        mask = [len(j.pending) > 0 for j in self.jobs]
        if not any(mask):
            return 0
        self.batch.set_active(mask)
        self.batch.step(1)
        self.stream.synchronize()
        t_done = time.perf_counter()
        served = 0
        for j, on in zip(self.jobs, mask):
            if on:
                t_arr = j.pending.popleft()
                j.latencies.append(t_done - t_arr)
                j.done += 1
                s = j.buffers.spec
                n_pix = j.buffers.img.numel() // 3
                self.samples += n_pix * (s.spp + (s.spp_grad or s.spp))
                served += 1
        self.engine_iterations += 1
        return served

    def close(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None


def run_simulation(duration: float = 20.0, rate: float = 200.0, surface_fraction: float = 0.7,
                   target_jobs: int = 16, iterations=(100, 500), budget_bytes: float = 200e9, seed: int = 0,
                   method: str = "cv_ps", device="cuda", use_graph: bool = True, block_size: int = 0,
                   queue_grab: int = 0, log=None, skip_steps: int = 17):
    """Poisson step requests at ``rate``/s, each for a uniformly chosen open
    job; ``target_jobs`` jobs are kept submitted (a finished job is replaced
    by a new one drawn from the mix: surface with probability
    ``surface_fraction``, else volume; iterations U[iterations]). Step
    latency percentiles leave out each job's first ``skip_steps`` steps, as
    the reference's per-iteration timings leave out the first 17
    (optimize.py:105); the ``*_all_steps`` figures keep them."""
    rng = np.random.default_rng(seed)
    srv = Server(budget_bytes, device, use_graph, block_size, queue_grab)
    next_seed = [100 + seed * 10000]

    def new_job():
        from . import scenes
        s = next_seed[0]
        next_seed[0] += 1
        kind = "surface" if rng.random() < surface_fraction else "volume"
        meta = scenes.surface_job_params(s) if kind == "surface" else scenes.volume_job_params(s)
        its = int(rng.integers(iterations[0], iterations[1] + 1))
        srv.submit_job(lambda s=s, kind=kind: make_job(s, kind, method, device), its, meta)

    for _ in range(target_jobs):
        new_job()
    t0 = time.perf_counter()
    next_arrival = t0 + rng.exponential(1.0 / rate)
    dropped = 0
    arrivals = 0
    peak = 0
    while True:
        now = time.perf_counter()
        if now - t0 > duration:
            break
        while next_arrival <= now:
            open_jobs = [j for j in srv.jobs if j.done + len(j.pending) < j.iterations]
            arrivals += 1
            if open_jobs:
                j = open_jobs[int(rng.integers(len(open_jobs)))]
                srv.submit_step(j.jid, next_arrival)
            else:
                dropped += 1
            next_arrival += rng.exponential(1.0 / rate)
        n_open = len(srv.jobs) + len(srv.waiting)
        for _ in range(max(0, target_jobs - n_open + len([j for j in srv.jobs if j.done >= j.iterations]))):
            new_job()
        srv.engine_step(time.perf_counter())
        peak = max(peak, len(srv.jobs))
        if log and srv.engine_iterations % 50 == 0 and srv.engine_iterations:
            log(f"[serve] t={now - t0:.1f}s engine its={srv.engine_iterations} jobs={len(srv.jobs)} "
                f"finished={len(srv.finished)}")
    elapsed = time.perf_counter() - t0
    lat_all = np.array([x for j in srv.jobs + srv.finished for x in j.latencies])
    lat = np.array([x for j in srv.jobs + srv.finished for x in j.latencies[skip_steps:]])
    out = {
        "duration_s": round(elapsed, 3), "rate_per_s": rate, "arrivals": arrivals, "dropped": dropped,
        "served_steps": int(lat_all.size), "engine_iterations": srv.engine_iterations,
        "jobs_admitted": len(srv.jobs) + len(srv.finished), "jobs_finished": len(srv.finished),
        "p50_step_ms": round(float(np.percentile(lat, 50)) * 1e3, 3) if lat.size else None,
        "p99_step_ms": round(float(np.percentile(lat, 99)) * 1e3, 3) if lat.size else None,
        "skipped_first_steps_per_job": skip_steps,
        "p50_step_ms_all_steps": round(float(np.percentile(lat_all, 50)) * 1e3, 3) if lat_all.size else None,
        "p99_step_ms_all_steps": round(float(np.percentile(lat_all, 99)) * 1e3, 3) if lat_all.size else None,
        "msamples_per_s": round(srv.samples / elapsed / 1e6, 3),
        "hbm_budget_gb": budget_bytes / 1e9, "peak_resident_jobs": peak,
    }
    srv.close()
    return out
