"""Serving R/K evaluation requests: continuous batching on one GPU.

A request asks for the residual R (a line-search evaluation, ``Ronly``) or R and K (a
Newton step) of one mesh at a given state (ub, vb, h) — the calls ``uvhAssembly`` receives
from ``CalcCostFunctionNRuvh.m:32`` and ``SSTREAM_TransientImplicit.m:370``. The server keeps
every registered mesh resident in HBM (connectivity, CSR pattern, element-row buffers and the
fields that do not change between iterations), so a request uploads only its state.

Scheduling: a request is admitted when
  * a slot (a HIP stream with output buffers of the largest mesh; nothing is allocated while
    serving) is free,
  * its outputs (R, and K for a Newton step) fit in what is left of the HBM budget for
    in-flight outputs, and
  * no other request on the same mesh is in flight (they share the mesh's work buffers);
requests that arrive while others run are admitted as soon as slots and memory free up.
Waiting requests are taken in arrival order (``policy="fifo"``, the default), later ones that
fit passing a blocked one, or shortest expected service first (mesh elements, x 1.6 for a
Newton step's R and K; ``policy="sjf"``); either way nothing passes a request that later
arrivals have passed ``max_bypass`` times. (Measured at 70 % load on a 2k-128k-element mix,
sjf did not lower the latency tail: the small requests' tail comes from sharing the device
with large kernels, not from queueing.) Admitted requests run concurrently on their streams,
so their kernels overlap on the device.
"""
from __future__ import annotations

import gc
import time
from dataclasses import dataclass, field

import numpy as np

from . import uvh


@dataclass
class Request:
    rid: int
    mesh: int
    newton: bool                     # R and K (True) or R only (False)
    arrival: float                   # seconds from the start of the run
    ub: np.ndarray = None
    vb: np.ndarray = None
    h: np.ndarray = None
    # filled by the server
    start: float = float("nan")
    done: float = float("nan")
    R: np.ndarray | None = None
    K: np.ndarray | None = None
    error: Exception | None = None   # the assembly's NaN error (uvhMatrixAssembly.m:575-584)
    keep: bool = True                # copy R (and K) to the host at completion (with keep_outputs)
    bypassed: int = 0

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _Mesh:
    """One registered mesh: its context, resident fields and state buffers."""

    def __init__(self, torch, dev, CtrlVar, MUA, F0, F1, device):
        self.ctx = uvh.UvhContext(device)
        self.ctx.set_mesh(MUA)
        self.nele, self.nn = MUA.Nele, MUA.Nnodes
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
        self.f1 = {k: t(getattr(F1, k)) for k in uvh.F1_KEYS}
        self.f0 = {k: t(getattr(F0, k)) for k in uvh.F0_KEYS}

        class _F:
            pass
        a, b = _F(), _F()
        for k in uvh.F1_KEYS:
            setattr(a, k, self.f1[k])
        for k in uvh.F0_KEYS:
            setattr(b, k, self.f0[k])
        self.fields = uvh.make_fields(b, a)
        self.ctrl = uvh.make_ctrl(CtrlVar, F1)
        self.busy = False

    def out_bytes(self, newton: bool) -> int:
        return 8 * (3 * self.nn + (self.ctx.nnz if newton else 0))


class UvhServer:
    kWindow = 32   # waiting requests considered per admission round, earliest arrivals first

    def __init__(self, CtrlVar, meshes, slots: int = 8, hbm_budget: float = 16e9, device: int = 0,
                 fp32: bool = False, max_bypass: int = 16, keep_outputs: bool = True, policy: str = "fifo"):
        """``meshes``: list of (MUA, F0, F1). ``hbm_budget``: bytes for in-flight outputs."""
        import torch
        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        self.meshes = [_Mesh(torch, self.dev, CtrlVar, *m, device) for m in meshes]
        self.streams = [torch.cuda.Stream(self.dev) for _ in range(slots)]   # default priority
        # each slot owns output buffers of the largest registered mesh, so that serving
        # allocates nothing; a request's outputs are views of its slot's buffers
        n3 = max(3 * m.nn for m in self.meshes)
        nnz = max(m.ctx.nnz for m in self.meshes)
        self.slot_R = [torch.empty(n3, dtype=torch.float64, device=self.dev) for _ in range(slots)]
        self.slot_K = [torch.empty(nnz, dtype=torch.float64, device=self.dev) for _ in range(slots)]
        # one completion event per slot, recorded again for each request the slot takes (its
        # previous request has completed by then)
        self.slot_ev = [torch.cuda.Event() for _ in range(slots)]
        # each request's NaN flags, copied stream-ordered behind its kernels into its slot's
        # pinned word and read when the slot's event has completed
        self.slot_status = [torch.zeros(1, dtype=torch.int32).pin_memory() for _ in range(slots)]
        self.budget = hbm_budget
        self.flags = uvh.UVH_FP32 if fp32 else 0
        self.max_bypass = max_bypass
        self.keep_outputs = keep_outputs
        if policy not in ("sjf", "fifo"):
            raise ValueError(f"unknown admission policy {policy!r}")
        self.policy = policy
        self.peak_used = 0.0
        self.peak_inflight = 0
        self.stall_threshold = 2e-3
        self.stalls: list[float] = []   # host stalls of the last run, seconds

    # one request on one stream: state upload, assembly, result download
    def _launch(self, req: Request, slot: int):
        torch = self.torch
        m = self.meshes[req.mesh]
        s = self.streams[slot]
        with torch.cuda.stream(s):
            for name, val in (("ub", req.ub), ("vb", req.vb), ("h", req.h)):
                if val is None:
                    continue
                if isinstance(val, np.ndarray):
                    m.f1[name].copy_(torch.from_numpy(val), non_blocking=False)
                else:   # a pinned host tensor: asynchronous on the request's stream
                    m.f1[name].copy_(val, non_blocking=True)
            R = self.slot_R[slot][:3 * m.nn]
            K = self.slot_K[slot][:m.ctx.nnz] if req.newton else None
            flags = self.flags | (0 if req.newton else uvh.UVH_RONLY)
            m.ctx.assemble_device(m.ctrl, m.fields, flags, R, K, None, None, stream=s.cuda_stream)
            m.ctx.status_async(self.slot_status[slot], stream=s.cuda_stream)
            ev = self.slot_ev[slot]
            ev.record(s)
        return (req, slot, R, K, ev)

    def run(self, requests: list[Request], clock=time.perf_counter) -> list[Request]:
        """Serves ``requests`` in real time (arrival times from the call); returns them completed.
        A request whose outputs exceed the whole budget is refused (ValueError). Python's cyclic
        garbage collector is paused while serving (a collection over the requests' tensors can
        hold the polling loop for milliseconds)."""
        was_enabled = gc.isenabled()
        gc.disable()
        try:
            return self._run(requests, clock)
        finally:
            if was_enabled:
                gc.enable()

    def _run(self, requests: list[Request], clock) -> list[Request]:
        for r in requests:
            if self.meshes[r.mesh].out_bytes(r.newton) > self.budget:
                raise ValueError(f"request {r.rid}: outputs exceed the HBM budget of {self.budget:.3g} bytes")
        torch = self.torch
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        inflight = []
        free_slots = list(range(len(self.streams)))
        used = 0.0
        t0 = clock()
        i = 0
        self.stalls = []
        prev = 0.0
        while i < len(pending) or waiting or inflight:
            now = clock() - t0
            # a pass of this loop takes microseconds; one that took longer than stall_threshold
            # was held up inside a call (a result download with keep_outputs, or, rarely, a
            # state upload or kernel launch that held the host for 6-13 ms), which delays every
            # request in the system
            if now - prev > self.stall_threshold:
                self.stalls.append(now - prev)
            prev = now
            while i < len(pending) and pending[i].arrival <= now:
                waiting.append(pending[i])
                i += 1
            # completions
            still = []
            for item in inflight:
                req, slot, R, K, ev = item
                if ev.query():
                    req.done = clock() - t0
                    req.error = uvh.UvhContext.status_error(int(self.slot_status[slot][0]))
                    if self.keep_outputs and req.keep:
                        req.R = R.cpu().numpy()
                        req.K = None if K is None else K.cpu().numpy()
                    m = self.meshes[req.mesh]
                    m.busy = False
                    used -= m.out_bytes(req.newton)
                    free_slots.append(slot)
                else:
                    still.append(item)
            inflight = still
            # admission. fifo (default): arrival order, skipping requests that do not fit yet; a
            # skipped request counts as overtaken when a later one is admitted, and nothing
            # passes a request overtaken max_bypass times. sjf: the kWindow earliest waiting
            # requests by expected service (requests overtaken max_bypass times first); the
            # window bounds the host work per admission and holds every request a later
            # arrival can overtake
            admitted, skipped = [], []
            if self.policy == "sjf" and free_slots and waiting:
                window = waiting[:self.kWindow]
                order = sorted(window, key=lambda r: (r.bypassed < self.max_bypass, self._cost(r), r.arrival))
            else:
                window, order = None, waiting
            for req in order:
                if not free_slots:
                    break
                m = self.meshes[req.mesh]
                need = m.out_bytes(req.newton)
                if m.busy or used + need > self.budget:
                    if req.bypassed >= self.max_bypass:
                        break   # no one overtakes a request that has been passed over enough
                    skipped.append(req)
                    continue
                if window is None:
                    for w in skipped:   # each of these is overtaken once more
                        w.bypassed += 1
                else:
                    taken = {id(a) for a in admitted}
                    for w in window:    # every earlier arrival still waiting is overtaken once more
                        if w.arrival < req.arrival and w is not req and id(w) not in taken:
                            w.bypassed += 1
                slot = free_slots.pop()
                m.busy = True
                used += need
                req.start = clock() - t0
                inflight.append(self._launch(req, slot))
                admitted.append(req)
                self.peak_used = max(self.peak_used, used)
                self.peak_inflight = max(self.peak_inflight, len(inflight))
            if admitted:
                ids = {id(r) for r in admitted}
                waiting = [r for r in waiting if id(r) not in ids]
            if not admitted and not inflight and i < len(pending) and not waiting:
                # idle until the next arrival
                dt = pending[i].arrival - (clock() - t0)
                if dt > 2e-4:
                    time.sleep(dt - 1e-4)
                    prev = clock() - t0
        torch.cuda.synchronize()
        return requests

    def _cost(self, req: Request) -> float:
        """Expected service time, in element units (R and K cost ~1.6x R alone)."""
        return self.meshes[req.mesh].nele * (1.6 if req.newton else 1.0)

    def serve_alone(self, req: Request) -> Request:
        """The same request with nothing else on the device (for the mix-vs-alone check)."""
        self.torch.cuda.synchronize()
        r = Request(req.rid, req.mesh, req.newton, 0.0, req.ub, req.vb, req.h)
        item = self._launch(r, 0)
        self.torch.cuda.synchronize()
        _, _, R, K, _ = item
        r.error = uvh.UvhContext.status_error(int(self.slot_status[0][0]))
        r.R = R.cpu().numpy()
        r.K = None if K is None else K.cpu().numpy()
        return r

    def close(self):
        for m in self.meshes:
            m.ctx.close()


def make_requests(meshes, n: int, rate: float, seed: int = 5, newton_fraction: float = 0.25,
                  perturb: float = 1.0, pinned: bool = False) -> list[Request]:
    """``n`` requests with Poisson arrivals at ``rate`` per second (rate = inf: all at once),
    meshes drawn uniformly, Newton steps : line-search residuals = 1 : 3, each with its own
    seeded state (the mesh's F1 with u, v perturbed by up to ``perturb`` m/a). ``pinned``: the
    states are page-locked host tensors, uploaded asynchronously."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        if np.isfinite(rate):
            t += rng.exponential(1.0 / rate)
        mi = int(rng.integers(len(meshes)))
        M, F0, F1 = meshes[mi]
        newton = bool(rng.random() < newton_fraction)
        ub = F1.ub + perturb * rng.standard_normal(M.Nnodes)
        vb = F1.vb + perturb * rng.standard_normal(M.Nnodes)
        h = F1.h.copy()
        if pinned:
            import torch
            ub, vb, h = (torch.from_numpy(a).pin_memory() for a in (ub, vb, h))
        out.append(Request(k, mi, newton, t if np.isfinite(rate) else 0.0, ub, vb, h))
    return out


def percentiles(reqs, q=(50, 99)):
    lat = np.array([r.latency for r in reqs])
    return {f"p{p}": float(np.percentile(lat, p)) for p in q}
