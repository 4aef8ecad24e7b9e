"""Serving loop: continuous batching of mixed-size nonlinear solve requests.

Requests (a cube mesh, a material, a load) arrive over time.  The server
admits them while batch slots and the HBM budget allow, keeps each admitted
request's mesh, constraints and Newton state resident on the GPU, and runs
the inner FP32 Jacobi-PCG solves of all admitted requests together through a
``tatva_batch``: every PCG iteration of every request in flight shares the
same few kernel launches, whatever the request sizes.  A request that
finishes its linear solve takes its FP64 Newton step (residual, update,
convergence test, next load step) between batch launches and re-enters the
batch with its next linear system; a request arriving meanwhile joins at the
next launch.  Arriving requests are prepared (mesh, tiling, constraints,
device buffers) on a worker thread while the GPU iterates the admitted ones.
The arithmetic of a request does not depend on what else is in the batch, so
its result is the same served alone or in any mix.

There is no counterpart in the reference; the request workload follows the
mixed Hex8 linear-elastic / Neo-Hookean cube stream: n = round(exp(U(ln 8,
ln 48))) cells per side, E ~ LogNormal(0, 0.5), nu ~ U(0.2, 0.45), materials
50/50, Poisson arrivals at a fixed rate.
"""

from __future__ import annotations

import ctypes
import math
import threading
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor
from concurrent.futures import wait as futures_wait
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat
from .element import Hexahedron8
from .energies import LinearElasticity, NeoHookean
from .lifter import Fixed, Lifter, RuntimeValue
from .mesh import Mesh
from .operator import Operator


def traction_loads(coords, n, t=(1e-3, 0.0, 0.0)):
    """Consistent nodal loads of a uniform traction on the x = 1 face of a
    structured n^3 unit-cube Hex8 mesh: t h^2 {1/4 corner, 1/2 edge, 1 interior}."""
    h = 1.0 / n
    f = np.zeros(coords.shape[0] * 3)
    face = np.where(np.isclose(coords[:, 0], 1.0))[0]
    j = np.rint(coords[face, 1] / h).astype(int)
    k = np.rint(coords[face, 2] / h).astype(int)
    w = np.where((j == 0) | (j == n), 0.5, 1.0) * np.where((k == 0) | (k == n), 0.5, 1.0)
    for c in range(3):
        f[face * 3 + c] = t[c] * h * h * w
    return f


def lame(E: float, nu: float):
    """(mu, lambda) from Young's modulus and Poisson's ratio."""
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


@dataclass
class SolveRequest:
    """One Hex8 cube solve: x = 0 clamped.  "linear_elastic": traction
    ``traction`` on x = 1 (one step).  "neo_hookean": u_x prescribed on x = 1,
    raised by ``stretch`` per load step over ``load_steps`` steps."""

    rid: int
    n: int
    material: str
    E: float
    nu: float
    arrival: float = 0.0
    load_steps: int = 1
    stretch: float = 0.03
    traction: float = 1e-3
    max_cg: int | None = None   # inner PCG iteration limit (None: the server's default)

    def __post_init__(self):
        if self.material not in ("linear_elastic", "neo_hookean"):
            raise ValueError("material must be 'linear_elastic' or 'neo_hookean'")
        if self.n < 1 or self.load_steps < 1:
            raise ValueError("n and load_steps must be >= 1")
        if self.max_cg is not None and self.max_cg < 1:
            raise ValueError("max_cg must be >= 1")

    @property
    def n_elements(self) -> int:
        return self.n ** 3

    @property
    def n_dofs(self) -> int:
        return 3 * (self.n + 1) ** 3

    def hbm_bytes(self) -> int:
        """Upper bound of the device memory the request holds while admitted:
        mesh (coords, connectivity, incidence, tiles, tile partials) and the
        Newton/PCG vectors."""
        E, N, d = self.n_elements, (self.n + 1) ** 3, self.n_dofs
        mesh = 24 * N + 32 * E + 4 * (N + 1) + 32 * E + 4 * E + 16 * E + 8 * (N + 1) \
            + 2 * 8 * E * 4 + 2 * 8 * E * 3 * 8
        vectors = d * (8 + 8 + 4 + 4 + 8 + 4 + 4) + d * 4 * 8
        return int(mesh + vectors + (1 << 20))


def make_stream(count: int, qps: float, seed: int = 0, n_range=(8, 48),
                nh_load_steps: int = 2, nh_stretch: float = 0.03):
    """The mixed request stream: Poisson arrivals at rate qps, n log-uniform
    in n_range, E ~ LogNormal(0, 0.5), nu ~ U(0.2, 0.45), materials 50/50."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    for i in range(count):
        t += rng.exponential(1.0 / qps) if qps > 0 else 0.0
        n = int(round(math.exp(rng.uniform(lo, hi))))
        E = float(rng.lognormal(0.0, 0.5))
        nu = float(rng.uniform(0.2, 0.45))
        nh = bool(rng.uniform() < 0.5)
        out.append(SolveRequest(rid=i, n=n, material="neo_hookean" if nh else "linear_elastic",
                                E=E, nu=nu, arrival=t,
                                load_steps=nh_load_steps if nh else 1, stretch=nh_stretch))
    return out


@dataclass
class RequestResult:
    rid: int
    n: int
    material: str
    n_free: int
    arrival: float
    admitted: float
    finished: float
    newton_iterations: int = 0
    cg_iterations: int = 0
    load_steps: int = 0
    converged: bool = False
    rel_residual: float = float("nan")
    # why a request did not converge: "non-finite residual", "inner PCG iteration
    # limit", "CG breakdown", "Newton iteration limit", "exceeds the HBM budget"
    failure: str | None = None
    u: object = None  # full FP64 solution on the device (kept if requested)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def queue_wait(self) -> float:
        return self.admitted - self.arrival


class CubeProblem:
    """The problem of a request: Hex8 unit cube of req.n cells per side, x = 0
    clamped; linear elastic with the x = 1 traction as external force, or
    Neo-Hookean with u_x on x = 1 prescribed per load step."""

    def __init__(self, req: SolveRequest):
        t = nat.torch()
        self.req = req
        n = req.n
        mesh = Mesh.unit_cube(n)
        self.mesh = mesh
        self.op = Operator(mesh, Hexahedron8())
        mu, lam = lame(req.E, req.nu)
        x0 = np.where(np.isclose(mesh.coords[:, 0], 0.0))[0]
        clamp = (x0[:, None] * 3 + np.arange(3)).ravel()
        self.size = mesh.n_nodes * 3
        self.ramp = None
        if req.material == "linear_elastic":
            f = traction_loads(mesh.coords, n, (req.traction, 0.0, 0.0))
            self.energy = self.op.energy(LinearElasticity(mu, lam), external_force=f)
            self.lifter0 = Lifter(self.size, Fixed(clamp, 0.0))
            self._x1 = None
        else:
            self.energy = self.op.energy(NeoHookean(mu, lam))
            x1 = np.where(np.isclose(mesh.coords[:, 0], 1.0))[0]
            self._x1 = x1
            # the prescribed u_x is one scalar per load step, written by the
            # device kernel without a host-to-device copy of a value array
            self.lifter0 = Lifter(self.size, Fixed(clamp, 0.0),
                                  Fixed(x1 * 3, RuntimeValue("load")))
            self.ramp = t.zeros(self.size, dtype=t.float64, device=nat.device())
            self.ramp[0::3] = t.as_tensor(mesh.coords[:, 0], device=nat.device())

    def lifter(self, step: int) -> Lifter:
        """Constraints of load step ``step`` (1-based)."""
        if self._x1 is None:
            return self.lifter0
        return self.lifter0.at["load"].set(float(self.req.stretch * step))

    def predict(self, u) -> None:
        """Start of a load step: homogeneous-stretch increment u_x += stretch * x."""
        if self.ramp is not None:
            u.add_(self.ramp, alpha=self.req.stretch)


class _Job:
    """A request resident on the GPU: mesh, constraints, FP64 state, the FP32
    linear system of its current Newton step and its batch slot.  Built
    (mesh, tiling, constraints, device buffers) before it has a slot, on the
    server's preparation thread; ``attach`` gives it the slot."""

    def __init__(self, req: SolveRequest, server: "Server"):
        t = nat.torch()
        self.req = req
        self.slot = -1
        self.server = server
        self.prob = CubeProblem(req)
        self.op, self.energy = self.prob.op, self.prob.energy
        self.lifter = self.prob.lifter(1)
        self.red = self.energy.reduced(self.lifter)
        self.n = self.red.size
        dev = nat.device()
        size = self.prob.size
        self.u = t.zeros(size, dtype=t.float64, device=dev)
        self.u32 = t.empty(size, dtype=t.float32, device=dev)
        self.b32 = t.empty(self.n, dtype=t.float32, device=dev)
        self.x32 = t.empty(self.n, dtype=t.float32, device=dev)
        L = nat.lib()
        nbytes = L.tatva_batch_slot_workspace(self.op.handle, 3, self.n)
        if nbytes < 0:
            nat.check(nat.EINVAL)
        self.ws = t.empty(int(nbytes), dtype=t.uint8, device=dev)
        self.result = RequestResult(rid=req.rid, n=req.n, material=req.material, n_free=self.n,
                                    arrival=req.arrival, admitted=float("nan"),
                                    finished=float("nan"))
        self.step = 0
        self.r0 = None
        self.newton_in_step = 0
        self.done = False
        self.solving = False
        self._r = None
        self._ev = None
        self._rr_host = t.empty(1, dtype=t.float64, pin_memory=True)  # r . r

    def attach(self, slot: int, admitted: float) -> None:
        self.slot = slot
        self.result.admitted = admitted

    # ------------------------------------------------------------ phases --

    def _set_load_step(self, s: int):
        self.step = s
        self.r0 = None
        self.newton_in_step = 0
        self.prob.predict(self.u)
        self.lifter = self.prob.lifter(s)
        self.red.lifter = self.lifter
        self.lifter.apply_constraints(self.u)

    # A job is either solving (its linear system iterates in the batch) or
    # waiting for the norm of its FP64 residual: the residual and its dot are
    # queued on the current stream with a copy of the dot to pinned host
    # memory and an event, and the server takes the Newton decision
    # (on_norm) once the event has completed, without blocking on it.

    def begin(self):
        """Queue the first residual (load step 1)."""
        self._set_load_step(1)
        self._start_residual()

    def _start_residual(self):
        from .solvers import dot  # noqa: PLC0415

        t = nat.torch()
        self._r = self.red.residual_full_state(self.u)
        self._rr_host.copy_(dot(self._r, self._r).reshape(1), non_blocking=True)
        self._ev = t.cuda.Event()
        self._ev.record()
        self.solving = False

    def norm_ready(self) -> bool:
        return not self.solving and not self.done and self._ev.query()

    def on_norm(self, batch_stream=None) -> bool:
        """The Newton decision at the residual norm: next load step (another
        residual queued), end, or the next linear solve (its set-up queued on
        the current stream, the solve started in the batch on batch_stream,
        after the current stream's work).  Returns False when the job is done."""
        srv = self.server
        L = nat.lib()
        rn = math.sqrt(max(float(self._rr_host[0]), 0.0))
        r = self._r
        if self.r0 is None:
            self.r0 = rn
        rel = rn / self.r0 if self.r0 > 0 else 0.0
        self.result.rel_residual = rel
        if not math.isfinite(rn):
            return self._finish(False, "non-finite residual")
        if self.newton_in_step >= srv.max_newton:
            return self._finish(False, "Newton iteration limit")
        if rn <= srv.tol * self.r0 or rn == 0.0:
            if self.step < self.req.load_steps:
                self._set_load_step(self.step + 1)
                self._start_residual()
                return True
            return self._finish(True)
        self.newton_in_step += 1
        self.result.newton_iterations += 1
        s = nat.stream()
        nat.check(L.tatva_cast_f64_f32(self.u.numel(), 1.0, nat.dptr(self.u), nat.dptr(self.u32), s))
        nat.check(L.tatva_cast_f64_f32(self.n, -1.0, nat.dptr(r), nat.dptr(self.b32), s))
        self._r = None
        dmap = self.lifter.dof_map_device()
        diag = self.energy.hessian_diag(self.u32, dof_map=dmap, n_out=self.n,
                                        dtype=nat.torch().float32)
        self._diag = diag
        self.x32.zero_()
        t = nat.torch()
        cur = t.cuda.current_stream()
        bs = cur if batch_stream is None else batch_stream
        if bs != cur:
            # the solve's set-up reads what this stream wrote (and diag, made here)
            bs.wait_stream(cur)
            diag.record_stream(bs)
        nat.check(L.tatva_batch_start(srv._batch, self.slot, self.op.handle,
                                      ctypes.byref(self.energy._desc), nat.dptr(self.u32),
                                      nat.dptr(dmap), self.n, nat.dptr(diag), nat.dptr(self.b32),
                                      nat.dptr(self.x32), float(srv.inner_tol),
                                      int(self.req.max_cg or srv.maxiter_factor * self.n),
                                      nat.dptr(self.ws),
                                      self.ws.numel(), bs.cuda_stream))
        self.solving = True
        return True

    def on_linear_done(self, st) -> bool:
        """Newton update u_free += x after the batch reports the solve done, and
        the next residual queued (on the current stream).  Returns False when
        the solve failed (the job is done)."""
        self.result.cg_iterations += int(st.iterations)
        if st.status == nat.EBREAKDOWN:
            return self._finish(False, "CG breakdown")
        if st.status == nat.ENOTCONV:
            return self._finish(False, "inner PCG iteration limit")
        L = nat.lib()
        nat.check(L.tatva_axpy_free_f32(self.n, nat.dptr(self.lifter.free_device()), 1.0,
                                        nat.dptr(self.x32), nat.dptr(self.u), nat.stream()))
        self._start_residual()
        return True

    def _finish(self, ok: bool, failure: str | None = None) -> bool:
        self.done = True
        self.result.converged = ok
        self.result.failure = failure
        self.result.load_steps = self.step
        return False

    def release(self):
        if self.server.keep_solutions:
            self.result.u = self.u
        self.op.close()
        self.ws = self.u32 = self.b32 = self.x32 = self._diag = self._r = None


@dataclass
class ServeReport:
    results: list
    seconds: float
    batch_launches: int
    peak_admitted: int
    peak_hbm_bytes: int
    dof_iterations: int = 0
    extra: dict = field(default_factory=dict)

    def latencies(self):
        """Latencies of the admitted requests (seconds)."""
        return np.array([r.latency for r in self.results if math.isfinite(r.admitted)])

    def summary(self) -> dict:
        lat = self.latencies()
        admitted = [r for r in self.results if math.isfinite(r.admitted)]
        done = [r for r in self.results if r.converged]
        failures: dict = {}
        for r in self.results:
            if r.failure:
                failures[r.failure] = failures.get(r.failure, 0) + 1
        return {
            "requests": len(self.results),
            "converged": len(done),
            "failures": failures,
            "seconds": self.seconds,
            "throughput_req_per_s": len(done) / self.seconds if self.seconds > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else float("nan"),
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else float("nan"),
            "mean_queue_ms": float(np.mean([r.queue_wait for r in admitted]) * 1e3)
            if admitted else float("nan"),
            "p99_queue_ms": float(np.percentile([r.queue_wait for r in admitted], 99) * 1e3)
            if admitted else float("nan"),
            "dof_iters_per_s": self.dof_iterations / self.seconds if self.seconds > 0 else 0.0,
            "batch_launches": self.batch_launches,
            "peak_admitted": self.peak_admitted,
            "peak_hbm_gb": self.peak_hbm_bytes / 1e9,
        }


def _warm_host() -> None:
    """Pays one-time host costs of the process at server start-up instead of
    inside a request.  The first numpy expression that could reuse a large
    temporary (arrays of 256 KiB and more) checks its caller with
    backtrace(), whose first call walks the unwind tables of every loaded
    shared library; with torch's ROCm libraries mapped that takes ~0.1-0.3 s
    and holds the GIL, stalling the serving loop with it."""
    a = np.ones(1 << 16)
    _ = (a + a) + a


class Server:
    """Continuous-batching solve server on one GPU.

    ``run(requests)`` replays the requests at their arrival times (seconds
    from the start of the run) and returns a ServeReport.  ``max_slots``
    bounds the requests in flight, ``hbm_budget`` their device memory
    (SolveRequest.hbm_bytes); ``chunk`` PCG iterations run between polls.
    After each poll the solves still iterating are launched again at once;
    the FP64 Newton work of the finished ones (update, residual and its norm,
    Jacobi diagonal) runs on a second stream, and the loop takes a job's
    Newton decision only when its residual norm has reached the host (an
    event it polls, never waits on while the batch has work): its next solve
    joins a later launch.  ``prefetch`` worker threads prepare arrived requests (mesh, tiling,
    constraints, device buffers) while the GPU iterates the others, each on
    a torch stream of its own so that preparation does not wait for the
    batch work queued on the serving stream; a prepared request is admitted
    as soon as it is ready and a slot is free (preparation starts in arrival
    order).  A request's bytes count against the budget from the start of
    its preparation.
    """

    def __init__(self, *, max_slots: int = 64, hbm_budget: float = 200e9, chunk: int = 8,
                 tol: float = 1e-8, inner_tol: float = 1e-3, max_newton: int = 30,
                 maxiter_factor: int = 10, keep_solutions: bool = False, prefetch: int = 2):
        if max_slots < 1 or chunk < 1 or prefetch < 1:
            raise ValueError("max_slots, chunk and prefetch must be >= 1")
        self.prefetch = int(prefetch)
        self.max_slots = int(max_slots)
        self.hbm_budget = float(hbm_budget)
        self.chunk = int(chunk)
        self.tol, self.inner_tol = float(tol), float(inner_tol)
        self.max_newton = int(max_newton)
        self.maxiter_factor = int(maxiter_factor)
        self.keep_solutions = keep_solutions
        self._device = nat.device()
        h = ctypes.c_void_p()
        nat.check(nat.lib().tatva_batch_create(self._device.index or 0, self.max_slots,
                                               ctypes.byref(h)))
        self._batch = h.value
        self._stats = (nat.PcgStats * self.max_slots)()
        self._local = threading.local()
        _warm_host()

    def _prepare(self, req: SolveRequest):
        """Worker thread: the request's _Job, built on the thread's own stream,
        and an event the serving stream waits for before using it.  (A new
        thread starts on device 0, so the server's device is entered first.
        A job's tensors are freed only after the streams of their last uses
        have been synchronised (the batch poll on the serving stream, the last
        residual norm on the Newton stream), so their reuse on this stream
        cannot overtake work queued there.)"""
        t = nat.torch()
        with t.cuda.device(self._device):
            st = getattr(self._local, "stream", None)
            if st is None:
                st = self._local.stream = t.cuda.Stream(device=self._device)
            with t.cuda.stream(st):
                job = _Job(req, self)
                ready = t.cuda.Event()
                ready.record(st)
        return job, ready

    def close(self):
        if getattr(self, "_batch", None):
            nat.lib().tatva_batch_destroy(self._batch)
            self._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def run(self, requests, *, realtime: bool = True) -> ServeReport:
        t = nat.torch()
        L = nat.lib()
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        # a request larger than the whole budget can never be admitted: rejected alone
        rejected = [r for r in pending if r.hbm_bytes() > self.hbm_budget]
        pending = [r for r in pending if r.hbm_bytes() <= self.hbm_budget]
        waiting: list[SolveRequest] = []
        preparing: list = []  # (request, future of its _Job), FIFO
        jobs: dict[int, _Job] = {}
        free_slots = list(range(self.max_slots))[::-1]
        used_bytes = 0
        results = []
        launches = 0
        peak_adm = 0
        peak_bytes = 0
        dof_iters = 0
        pool = ThreadPoolExecutor(max_workers=self.prefetch, thread_name_prefix="tatva-prepare")
        serving = t.cuda.current_stream()
        # Newton steps of finished solves run here while the next batch launch
        # runs on the serving stream; their next solves start on the serving
        # stream, after that launch
        newton = t.cuda.Stream(device=self._device)
        ahead = False  # a batch launch is queued that no poll has seen yet
        t.cuda.synchronize()
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        try:
            while pending or waiting or preparing or jobs:
                now = clock() if realtime else 0.0
                while pending and (not realtime or pending[0].arrival <= now):
                    waiting.append(pending.pop(0))
                # preparation: in arrival order, at most `prefetch` at once, within
                # the HBM budget
                while waiting and len(preparing) < self.prefetch and \
                        used_bytes + waiting[0].hbm_bytes() <= self.hbm_budget:
                    req = waiting.pop(0)
                    used_bytes += req.hbm_bytes()
                    preparing.append((req, pool.submit(self._prepare, req)))
                # admission: every prepared request, in arrival order, while slots last
                for item in [p for p in preparing if p[1].done()]:
                    if not free_slots:
                        break
                    preparing.remove(item)
                    req, fut = item
                    job, ready = fut.result()
                    newton.wait_event(ready)
                    slot = free_slots.pop()
                    job.attach(slot, time.perf_counter() - t0)
                    jobs[slot] = job
                    with t.cuda.stream(newton):
                        job.begin()
                peak_adm = max(peak_adm, len(jobs))
                peak_bytes = max(peak_bytes, used_bytes)
                # Newton decisions of the jobs whose residual norms have arrived
                with t.cuda.stream(newton):
                    for slot, job in list(jobs.items()):
                        if job.norm_ready() and not job.on_norm(serving):
                            used_bytes -= job.req.hbm_bytes()
                            self._retire(job, jobs, free_slots, results, t0)
                if not jobs:
                    if preparing:
                        futures_wait([p[1] for p in preparing], timeout=0.01,
                                     return_when=FIRST_COMPLETED)
                    elif pending and realtime:
                        time.sleep(max(0.0, min(pending[0].arrival - clock(), 0.01)))
                    continue
                if not ahead and not any(job.solving for job in jobs.values()):
                    # nothing in the batch: wait for the oldest residual norm
                    next(job for job in jobs.values() if not job.solving)._ev.synchronize()
                    continue
                if not ahead:
                    nat.check(L.tatva_batch_iterate(self._batch, self.chunk, serving.cuda_stream))
                    launches += 1
                running = ctypes.c_int32(0)
                nat.check(L.tatva_batch_poll(self._batch, self._stats, ctypes.byref(running),
                                             serving.cuda_stream))
                # the solves still iterating go on at once; the finished ones
                # take their Newton steps meanwhile and join the launch after
                ahead = running.value > 0
                if ahead:
                    nat.check(L.tatva_batch_iterate(self._batch, self.chunk, serving.cuda_stream))
                    launches += 1
                finished = [slot for slot, job in jobs.items()
                            if job.solving and self._stats[slot].reserved]
                if finished:
                    # (the poll synchronised the serving stream: the finished
                    # solves' x are complete; the launch queued since does not
                    # touch their slots, which it skips as done)
                    with t.cuda.stream(newton):
                        for slot in finished:
                            st = self._stats[slot]
                            job = jobs[slot]
                            dof_iters += job.n * int(st.iterations)
                            if not job.on_linear_done(st):
                                used_bytes -= job.req.hbm_bytes()
                                self._retire(job, jobs, free_slots, results, t0)
        finally:
            pool.shutdown(wait=True)
        t.cuda.synchronize()
        seconds = time.perf_counter() - t0
        for r in rejected:
            results.append(RequestResult(rid=r.rid, n=r.n, material=r.material, n_free=0,
                                         arrival=r.arrival, admitted=float("nan"),
                                         finished=r.arrival, failure="exceeds the HBM budget"))
        results.sort(key=lambda r: r.rid)
        return ServeReport(results=results, seconds=seconds, batch_launches=launches,
                           peak_admitted=peak_adm, peak_hbm_bytes=peak_bytes,
                           dof_iterations=dof_iters)

    def _retire(self, job, jobs, free_slots, results, t0):
        nat.check(nat.lib().tatva_batch_release(self._batch, job.slot))
        job.result.finished = time.perf_counter() - t0
        job.release()
        results.append(job.result)
        del jobs[job.slot]
        free_slots.append(job.slot)


__all__ = ["SolveRequest", "RequestResult", "ServeReport", "Server", "CubeProblem", "make_stream",
           "traction_loads", "lame"]
