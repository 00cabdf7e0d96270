"""Serving imaging requests on one GPU: continuous batching of measurement-operator jobs.

A request is one imaging job: an operator description (image size, visibilities, optional w-terms)
and the vectors to apply it to.  The server

  * admits pending requests in arrival order while their estimated HBM footprint fits the budget and
    a slot is free (``max_inflight``); a request that arrives while others run is admitted as soon as
    memory and a slot free up,
  * builds each admitted request's operator on a worker thread (the C library releases the GIL, so
    builds of different requests proceed in parallel on the host),
  * runs its applies on the worker's own stream (default priority), so the kernels of concurrent
    requests overlap on the device,
  * records the latency from arrival to completion.

Every kernel of the operator is deterministic, so a request's result does not depend on what runs
beside it: it is bit-identical alone and in the mix.

The reference has no serving layer (its applications build one operator per run, cpp/main.cc:85-111);
this module is the batching/serving half of the workload the operator is measured on.
"""
from __future__ import annotations

import heapq
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .operators import measurementoperator


@dataclass
class ImagingRequest:
    """One job.  kind: "forward" (y = Phi x), "adjoint" (x = Phi^dagger y), "pair" (both) or "padmm" (a
    reconstruction from y: the operator normalised by `power_iters` power iterations, a DB4 prior with
    `levels` levels, the l2-ball radius from `sigma`, `iters` PADMM iterations)."""
    u: np.ndarray
    v: np.ndarray
    imsizey: int
    imsizex: int
    kind: str = "pair"
    x: np.ndarray = None  # image (complex128), for forward / pair
    y: np.ndarray = None  # visibilities (complex128), for adjoint / pair
    w: np.ndarray = None
    weights: np.ndarray = None
    kernel: str = "kb"
    J: int = 4
    w_term: bool = False
    Jw: int = 6
    cell: float = 1.0
    oversample: float = 2.0
    precision: str = "c64"
    arrival: float = 0.0  # seconds from the server's start; set by submit() when 0
    rid: int = -1
    iters: int = 100       # padmm
    sigma: float = 0.0     # padmm: noise standard deviation of y
    levels: int = 4        # padmm: DB4 levels
    power_iters: int = 50  # padmm: normalisation of Phi


@dataclass
class RequestResult:
    rid: int
    y: np.ndarray = None
    x: np.ndarray = None
    arrival: float = 0.0
    admitted: float = 0.0
    finished: float = 0.0
    device_bytes: int = 0

    @property
    def latency(self):
        return self.finished - self.arrival


def estimate_bytes(req: ImagingRequest) -> int:
    """Upper estimate of the HBM a request holds: the operator's grids, visibility data, tile lists and
    projection rows, its staging buffers and the request's own vectors."""
    e = 8 if req.precision == "c64" else 16
    Nu = int(np.floor(req.oversample * req.imsizex))
    Nv = int(np.floor(req.oversample * req.imsizey))
    M = int(np.asarray(req.u).size)
    n2 = req.imsizex * req.imsizey
    grid = 2 * e * Nu * Nv
    vis = M * (16 + 8 + 8 + 2 * e + 4 * 4)  # coordinates, weights, permutation plan, sorted and staged values, 2 list sets
    rows = M * (req.J + req.Jw - 1) ** 2 * e if req.w_term else 0
    stage = 16 * max(n2, M) + e * (n2 + M)
    io = 2 * e * (n2 + M)
    solver = (e * (8 * n2 + 6 * M) + 16 * M) if req.kind == "padmm" else 0  # PADMM state, DB4 buffers
    return int(1.15 * (grid + vis + rows + stage + io + solver)) + (8 << 20)


class Server:
    """Continuous-batching server over one GPU (see the module docstring)."""

    def __init__(self, hbm_budget_bytes: int, max_inflight: int = 4, device: int = 0):
        import torch
        self._torch = torch
        self.budget = int(hbm_budget_bytes)
        self.max_inflight = int(max_inflight)
        self.device = torch.device("cuda", device)
        self._pool = ThreadPoolExecutor(max_workers=self.max_inflight)
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(self.max_inflight)]
        self._free_streams = list(range(self.max_inflight))
        self._lock = threading.Condition()
        self._pending: list = []  # heap of (arrival, seq, request, future)
        self._seq = 0
        self._in_use = 0
        self._inflight = 0
        self._t0 = time.perf_counter()
        self._closed = False
        self.peak_in_use = 0
        self.peak_inflight = 0
        self._sched = threading.Thread(target=self._schedule, daemon=True)
        self._sched.start()

    def now(self) -> float:
        return time.perf_counter() - self._t0

    def submit(self, req: ImagingRequest) -> Future:
        if estimate_bytes(req) > self.budget:
            raise ValueError("request does not fit the HBM budget on its own")
        fut: Future = Future()
        with self._lock:
            if req.arrival <= 0:
                req.arrival = self.now()
            if req.rid < 0:
                req.rid = self._seq
            heapq.heappush(self._pending, (req.arrival, self._seq, req, fut))
            self._seq += 1
            self._lock.notify_all()
        return fut

    def close(self):
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        self._sched.join()
        self._pool.shutdown(wait=True)

    # -- scheduler: admit in arrival order while memory and a slot are free
    def _schedule(self):
        while True:
            with self._lock:
                while True:
                    if self._closed and not self._pending and self._inflight == 0:
                        return
                    if self._pending and self._free_streams:
                        arrival, _, req, fut = self._pending[0]
                        need = estimate_bytes(req)
                        if arrival <= self.now() and self._in_use + need <= self.budget:
                            heapq.heappop(self._pending)
                            self._in_use += need
                            self.peak_in_use = max(self.peak_in_use, self._in_use)
                            self._inflight += 1
                            self.peak_inflight = max(self.peak_inflight, self._inflight)
                            slot = self._free_streams.pop()
                            break
                        # not yet arrived: wait until it does (or a completion frees memory)
                        wait = max(0.0, arrival - self.now()) if arrival > self.now() else None
                        self._lock.wait(timeout=wait if wait else 0.05)
                        continue
                    self._lock.wait(timeout=0.05)
            self._pool.submit(self._run, req, fut, slot, need, self.now())

    @staticmethod
    def _padmm(req: ImagingRequest, op):
        """The reconstruction: normalise Phi, then `iters` PADMM iterations on the solver's own stream."""
        from . import algorithms, utilities, wavelets
        if req.power_iters > 0:
            op.set_norm(op.power_method(req.power_iters, 1e-4))
        lv = req.levels
        while lv > 0 and (req.imsizey % (1 << lv) or req.imsizex % (1 << lv)):
            lv -= 1
        W = wavelets.WaveletOperator([("DB4", lv)], req.imsizey, req.imsizex, req.precision)
        sigma = req.sigma if req.sigma > 0 else 1e-3 * float(np.linalg.norm(req.y)) / np.sqrt(max(req.y.size, 1))
        eps = utilities.calculate_l2_radius(req.y.size, sigma)
        sol = algorithms.ImagingProximalADMM(op, W, req.y, eps, itermax=max(req.iters, 1))
        # launched eagerly on the solver's stream: a graph capture would make every other request's
        # legacy-stream call (operator builds, power methods) fail while it lasts
        sol.iterate(req.iters, use_graph=False)
        x = sol.image()
        sol.close()
        W.close()
        return x

    def _run(self, req: ImagingRequest, fut: Future, slot: int, need: int, admitted: float):
        torch = self._torch
        res = RequestResult(rid=req.rid, arrival=req.arrival, admitted=admitted)
        try:
            lt = measurementoperator.init_degrid_operator_2d(
                req.u, req.v, req.w, req.weights, req.imsizey, req.imsizex, req.oversample, 0, 1e-4, req.kernel,
                req.J, req.J, req.w_term, req.Jw, req.cell, req.cell, precision=req.precision)
            op = lt.op
            res.device_bytes = op.info().device_bytes
            if req.kind == "padmm":
                res.x = self._padmm(req, op)
                op.close()
                res.finished = self.now()
                fut.set_result(res)
                return
            st = self._streams[slot]
            dt = torch.complex64 if req.precision == "c64" else torch.complex128
            with torch.cuda.stream(st):
                if req.kind in ("forward", "pair"):
                    xd = torch.from_numpy(np.asarray(req.x, np.complex128)).to(dt).to(self.device, non_blocking=True)
                    yd = op.forward(xd, stream=st.cuda_stream)
                    res.y = yd
                if req.kind in ("adjoint", "pair"):
                    vd = torch.from_numpy(np.asarray(req.y, np.complex128)).to(dt).to(self.device, non_blocking=True)
                    xo = op.adjoint_apply(vd, stream=st.cuda_stream)
                    res.x = xo
                st.synchronize()
                if res.y is not None:
                    res.y = res.y.cpu().numpy()
                if res.x is not None:
                    res.x = res.x.cpu().numpy()
            op.close()
            res.finished = self.now()
            fut.set_result(res)
        except BaseException as exc:  # surfaced through the future
            fut.set_exception(exc)
        finally:
            with self._lock:
                self._in_use -= need
                self._inflight -= 1
                self._free_streams.append(slot)
                self._lock.notify_all()


def latency_stats(results) -> dict:
    lat = np.array([r.latency for r in results])
    if lat.size == 0:
        return {}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "requests": int(lat.size)}


def poisson_requests(seed: int, n_requests: int, rate_per_s: float, sizes=(256, 512, 1024, 2048),
                     M_range=(1e4, 4e6), w_fraction=0.3, wmax=200.0, precision="c64", kind="pair",
                     padmm_fraction=0.0, padmm_iters=100):
    """The serving workload: Poisson arrivals; image side uniform over `sizes`; M log-uniform over
    M_range; a fraction of the jobs with w-terms (w uniform in [0, wmax], 1 degree field of view); a
    fraction `padmm_fraction` of the jobs are `padmm_iters`-iteration reconstructions instead of `kind`.
    The shared generator draws every request's parameters in order; each request's coordinates, image and
    visibilities come from generators seeded by its index, on a thread pool."""
    from . import utilities
    rng = np.random.default_rng(seed)
    t = 0.0
    params = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate_per_s)
        n = int(rng.choice(sizes))
        M = int(round(10 ** rng.uniform(np.log10(M_range[0]), np.log10(M_range[1]))))
        w_term = bool(rng.uniform() < w_fraction)
        w = rng.uniform(0.0, wmax, M) if w_term else None
        kd = "padmm" if rng.uniform() < padmm_fraction else kind
        params.append((k, t, n, M, w_term, w, kd))

    def make(p):
        k, t, n, M, w_term, w, kd = p
        u, v = utilities.synth_uv(seed * 1000 + k, M, np.pi / 3)
        u = utilities.radians_to_pixels(u, n)
        v = utilities.radians_to_pixels(v, n)
        x = utilities.synth_image(seed * 1000 + k, n, n, 20)
        y = utilities.synth_cnormal(seed * 1000 + k + 7, M)
        return ImagingRequest(u=u, v=v, imsizey=n, imsizex=n, kind=kd, x=x, y=y, w=w, J=4, w_term=w_term, Jw=6,
                              cell=3600.0 / n, precision=precision, arrival=t, rid=k, iters=padmm_iters,
                              sigma=0.1 * float(np.linalg.norm(y)) / np.sqrt(M))

    with ThreadPoolExecutor(max_workers=8) as ex:
        reqs = list(ex.map(make, params))
    return reqs
