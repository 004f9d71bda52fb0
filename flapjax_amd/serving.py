"""Serving form: continuous batching of unsteady UVLM cases on one GPU.

Cases of differing panel counts, wake lengths and step counts arrive over
time.  The engine admits a waiting case as soon as a slot is free and its
device footprint fits the HBM budget, keeps every admitted case's state
(bound and wake grids, circulations, factorised AIC, motion history)
resident on the device, and advances all active cases together, one time
step per engine tick:

  * per case: grid from the prescribed beam poses, collocation stencils, wake
    circulation shift, AIC assembly, right-hand side, circulation solve, forces;
  * across cases: the three Biot-Savart pair sums of the step (free-wake
    rollup with the convection epilogue, wake upwash at the collocation
    points, velocities at the filament midpoints) are each ONE fused launch
    over all active cases (a persistent grid striding over a device task
    table).

Cases that arrive while others run join at the next tick; finished cases
leave and free their slot and memory.  Latency = completion - arrival.

Per-step semantics are those of ``uvlm.UVLM`` (aero/uvlm.py:824-1117) with
the circulation solve done with the explicit inverse of the dynamic AIC kept
for the case (rigid motion keeps the AIC fixed up to rounding) and refined in
FP64 against the AIC re-assembled every step, so each step solves the current
system to FP64 accuracy; the stop-test outcome of every case's solve is read
back at the end of the tick.  A tick is one native call (fj_step_batched):
each stage -- geometry, tables, the three fused pair sums, AIC, right-hand
side, refinement, forces, lift -- is one launch over all active cases.  With
precision="mixed" the rollup and force pair sums take the mixed form (FP32
far field, FP64 near field) as uvlm.UVLM does; the upwash, AIC and solve
stay FP64.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _kernels as K
from . import _lib as L
from .flowfields import Constant
from .utils import EPSILON, R_CUTOFF

F64 = torch.float64


@dataclass
class CaseRequest:
    """One rigid-wing pitch case (chord 1, span 10, U = 1)."""

    case_id: int
    m: int
    n: int
    steps: int
    m_star: int
    pitch_amp_deg: float
    k_red: float
    alpha0_deg: float = 5.0
    span: float = 10.0
    arrival: float = 0.0


@dataclass
class CaseResult:
    case_id: int
    arrival: float
    start: float
    finish: float
    lift: np.ndarray            # total vertical steady force per step
    gamma_b_final: np.ndarray
    zeta_w_final: np.ndarray

    @property
    def latency(self):
        return self.finish - self.arrival


def motion(req: CaseRequest):
    """hg (steps, n+1, 4, 4) and its exact time derivative hg_dot for the
    pitch theta = a0 + A sin(w t) about the beam axis."""
    dt = 1.0 / req.m
    omega = 2.0 * req.k_red
    t = np.arange(req.steps) * dt
    th = np.deg2rad(req.alpha0_deg) + np.deg2rad(req.pitch_amp_deg) * np.sin(omega * t)
    thd = np.deg2rad(req.pitch_amp_deg) * omega * np.cos(omega * t)
    hg = np.zeros((req.steps, req.n + 1, 4, 4))
    c, s = np.cos(th), np.sin(th)
    hg[:, :, 0, 0], hg[:, :, 0, 2] = c[:, None], s[:, None]
    hg[:, :, 2, 0], hg[:, :, 2, 2] = -s[:, None], c[:, None]
    hg[:, :, 1, 1] = 1.0
    hg[:, :, 1, 3] = np.linspace(0.0, req.span, req.n + 1)[None, :]
    hg[:, :, 3, 3] = 1.0
    hgd = np.zeros_like(hg)
    hgd[:, :, 0, 0], hgd[:, :, 0, 2] = (-s * thd)[:, None], (c * thd)[:, None]
    hgd[:, :, 2, 0], hgd[:, :, 2, 2] = (-c * thd)[:, None], (-s * thd)[:, None]
    return hg, hgd


def footprint_bytes(req: CaseRequest, precision: str = "fp64") -> int:
    """Device bytes a case holds while active."""
    m, n, ms = req.m, req.n, req.m_star
    nb = m * n
    s_b = m * (n + 1) + (m + 1) * n
    s_w = ms * (n + 1) + (ms + 1) * n
    words = 0
    words += 2 * req.steps * (n + 1) * 16            # motion
    words += 6 * (m + 1) * (n + 1) * 3               # grids, velocities, forces
    words += 2 * (ms + 1) * (n + 1) * 3 + 2 * ms * n  # wake (double buffered)
    words += 6 * nb * 3 + 6 * nb                     # stencils, circulations
    words += nb * nb                                 # FP64 AIC
    words += nb * nb                                 # kept inverse
    words += 16 * (s_b + s_w)                        # filament table
    words += 3 * (ms * (n + 1) + s_b + nb)           # pair-sum targets/outputs
    words += req.steps                               # lift history
    words += 16 * (s_b + s_w) + 10 * nb + 6 * s_b     # tick workspace share: second table, stencils, midpoints
    words += 3 * (m + 1) * (n + 1)                   # steady forces
    # pair-sum partials (target tiles of 64, filament chunks of 256, at most
    # about 8192 work items per pair sum)
    for n_tgt, n_seg in (((ms + 1) * (n + 1), s_b + s_w), (nb, s_w), (s_b, s_b + s_w)):
        tiles = -(-n_tgt // 64)
        words += 3 * 64 * tiles * max(1, min(-(-n_seg // 256), -(-8192 // tiles)))
    return 8 * words + 4096


class _InverseCache:
    """Explicit inverses of initial dynamic AICs keyed by (lattice, kernel,
    initial pose): cases that start from the same pose of the same lattice
    have the same matrix, so they share one factorisation.  Holds at most
    ``max_bytes`` of matrices (oldest entries dropped first; a dropped entry
    stays alive while an active case still holds it)."""

    def __init__(self, max_bytes: int):
        self.max_bytes = max_bytes
        self.entries: dict = {}
        self.bytes = 0

    def get(self, key):
        return self.entries.get(key)

    def put(self, key, ent):
        nbytes = sum(t.numel() * t.element_size() for t in ent)
        if nbytes > self.max_bytes:
            return
        while self.entries and self.bytes + nbytes > self.max_bytes:
            old = self.entries.pop(next(iter(self.entries)))
            self.bytes -= sum(t.numel() * t.element_size() for t in old)
        self.entries[key] = ent
        self.bytes += nbytes


def _static_solve(A, Ainv, U, te, b):
    """x = A_s^{-1} b for A_s = A + U E^T (E selects the columns `te`) from
    the kept inverse of A, by the Sherman-Morrison-Woodbury identity
    x = y - Z (I + E^T Z)^{-1} E^T y with y = A^{-1} b, Z = A^{-1} U, followed
    by one correction sweep with the FP64 residual b - A_s x.  The capacitance
    matrix I + E^T Z = E^T A^{-1} A_s E is as well conditioned as A_s relative
    to A (its size is the number of trailing-edge rings), and solving it with
    the small LU replaces a second full factorisation."""
    Z = Ainv @ U
    C = Z[te].clone()
    C.diagonal().add_(1.0)
    fac = K.lu_factor(C, overwrite=True)

    def apply(r):
        y = Ainv @ r
        return y - Z @ K.lu_solve(fac, y[te])

    x = apply(b)
    res = b - A @ x - U @ x[te]
    return x + apply(res)


class _Slot:
    """Resident state of one admitted case."""

    def __init__(self, req: CaseRequest, dev, precision: str, flow: Constant, kernel_kind: int,
                 cache: "_InverseCache | None" = None):
        self.req = req
        m, n, ms = req.m, req.n, req.m_star
        self.dt = 1.0 / m
        self.flow = flow
        hg, hgd = motion(req)
        # pinned staging, asynchronous upload (the staging block is held by
        # torch's host allocator until the copy has run)
        self.hg = torch.from_numpy(hg).pin_memory().to(dev, non_blocking=True)
        self.hgd = torch.from_numpy(hgd).pin_memory().to(dev, non_blocking=True)
        self.dof = torch.arange(n + 1, dtype=torch.int32, device=dev)
        z0 = np.zeros((m + 1, n + 1, 3))
        z0[..., 0] = (np.linspace(0.0, 1.0, m + 1) - 0.25)[:, None]
        self.zeta0 = torch.as_tensor(z0, device=dev)
        self.kind = kernel_kind
        self.precision = precision
        self.prec = L.PREC_MIXED if precision == "mixed" else L.PREC_FP64
        self.i = 0  # last completed step index
        self.lift = torch.zeros(req.steps, dtype=F64, device=dev)
        self.s_b = m * (n + 1) + (m + 1) * n
        self.s_w = ms * (n + 1) + (ms + 1) * n
        # state buffers
        self.zb, _ = K.grid_from_hg(self.hg[0], None, self.dof, self.zeta0)
        self.zb_prev = self.zb.clone()
        self.zbd = torch.zeros_like(self.zb)
        s = torch.as_tensor(np.concatenate(([0.0], np.cumsum(np.ones(ms) * self.dt * flow.u_inf_mag))), device=dev)
        self.zw = K.wake_initialise(self.zb, s, flow.u_inf_dir)
        self.zw_next = torch.empty_like(self.zw)
        self.gw = torch.empty((ms, n), dtype=F64, device=dev)
        self.gw_next = torch.empty_like(self.gw)
        self.gbd = torch.zeros((m, n), dtype=F64, device=dev)
        nb = m * n
        self.A = torch.empty((nb, nb), dtype=F64, device=dev)  # this step's AIC, written by the ticks
        c, nc, _ = K.colloc_normals(self.zb)
        rhs = K.rhs(c, nc, None, None, flow, 0.0)
        # the dynamic (bound-only) AIC at the initial pose and its inverse,
        # kept for the refinement of every step; shared with every case whose
        # initial pose, lattice and kernel are the same (the same matrix)
        key = (m, n, kernel_kind, hg[0].tobytes())
        ent = cache.get(key) if cache is not None else None
        if ent is None:
            A0 = torch.empty((nb, nb), dtype=F64, device=dev)
            K.aic_assemble(c, nc, self.zb, kernel_kind, EPSILON, R_CUTOFF, A0, 0)
            ent = (A0, K.lu_inverse(K.lu_factor(A0)))
            if cache is not None:
                cache.put(key, ent)
        A0, self.Ainv = ent
        # the static AIC (aero/uvlm.py:895-1012) is the dynamic AIC plus the
        # lumped wake in the n trailing-edge ring columns (add_wake_influence,
        # aero/aic.py:190-205): A_s = A_0 + U E^T, U = (TE ring + wake column
        # sums) - TE ring, from the trailing-edge ring row alone
        te = slice((m - 1) * n, m * n)
        U = torch.empty((nb, n), dtype=F64, device=dev)
        K.aic_assemble(c, nc, self.zb[m - 1:], kernel_kind, EPSILON, R_CUTOFF, U, 0, zeta_w=self.zw)
        U -= A0[:, te]
        self.gb = _static_solve(A0, self.Ainv, U, te, rhs).reshape(m, n).contiguous()
        K.static_wake_gamma(self.gb, ms, out=self.gw)
        self.gb_prev = self.gb.clone()
        self.ws = torch.empty(L.lib().fj_refine_workspace_bytes(m * n) // 8 + 1, dtype=F64, device=dev)
        self.fs = torch.empty((m + 1, n + 1, 3), dtype=F64, device=dev)
        # (A, B) buffer pairs of the step state: step i writes into B when i
        # is odd and reads the previous step from A, and the other way round
        # when i is even; the step-0 state is in A
        self.zb_pair = (self.zb, self.zb_prev)
        self.gb_pair = (self.gb, self.gb_prev)
        self.zw_pair = (self.zw, self.zw_next)
        self.gw_pair = (self.gw, self.gw_next)

    def final_state(self):
        """(circulation, wake grid) of the last step: the step-i state is in
        the second buffer of each pair when i is odd."""
        k = self.i & 1
        return self.gb_pair[k], self.zw_pair[k]

    @property
    def done(self):
        return self.i >= self.req.steps - 1


class ServingEngine:
    def __init__(self, precision: str = "fp64", max_slots: int = 32, hbm_budget_bytes: int = 64 << 30,
                 free_wake: bool = True, device=None, inverse_cache_bytes: int = 8 << 30,
                 graph_ticks: bool = False):
        """graph_ticks: replay each tick from a captured HIP graph (one per
        staging parity, captured again whenever the active set changes)
        instead of issuing its launches; the host then only restages the
        step's records (fj_step_batched_restage)."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.max_slots = max_slots
        self.budget = hbm_budget_bytes
        self.free_wake = free_wake
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.flow = Constant(np.array([1.0, 0.0, 0.0]), 1.225, True)
        self.kind = L.KERNEL_EPSILON
        self._staging = {}
        self._bufs = {}
        self._adm = None
        self._plan = None
        self._cache = _InverseCache(inverse_cache_bytes) if inverse_cache_bytes > 0 else None
        self.graph_ticks = bool(graph_ticks)
        self._graphs = {}
        self._graph_set = None
        self.graph_captures = 0

    # ---- fused pair sums ---------------------------------------------------
    def _batched(self, tasks, seg, targets, out, prec):
        """tasks: list of (tgt_off, n_tgt, seg_off, n_seg, out_off, mode, dt, time)."""
        nt = len(tasks)
        if nt == 0:
            return
        cols = list(zip(*tasks))
        ia = lambda v: (ctypes.c_int * nt)(*[int(x) for x in v])  # noqa: E731
        da = lambda v: (ctypes.c_double * nt)(*[float(x) for x in v])  # noqa: E731
        tgt_off, n_tgt, seg_off, n_seg, out_off, mode = (ia(cols[k]) for k in range(6))
        dt, tm = da(cols[6]), da(cols[7])
        flows = (L.FjFlowfield * nt)(*([self.flow.descriptor()] * nt))
        tb = L.lib().fj_batched_task_bytes(nt)
        table = torch.empty(tb // 8 + 1, dtype=F64, device=self.dev)
        staging = ctypes.create_string_buffer(tb)
        wsb = L.lib().fj_batched_workspace_bytes(nt, ctypes.cast(n_tgt, ctypes.c_void_p),
                                                  ctypes.cast(n_seg, ctypes.c_void_p))
        ws = torch.empty(wsb // 8 + 1, dtype=F64, device=self.dev)
        if prec != L.PREC_FP64:
            # the fused launch runs the FP64 evaluation; the mixed far field is
            # applied per case through fj_pair_sum
            raise NotImplementedError
        rc = L.lib().fj_pair_sum_batched(nt, ctypes.cast(tgt_off, ctypes.c_void_p), ctypes.cast(n_tgt, ctypes.c_void_p),
                                         ctypes.cast(seg_off, ctypes.c_void_p), ctypes.cast(n_seg, ctypes.c_void_p),
                                         ctypes.cast(out_off, ctypes.c_void_p), ctypes.cast(mode, ctypes.c_void_p),
                                         ctypes.cast(dt, ctypes.c_void_p), ctypes.cast(tm, ctypes.c_void_p),
                                         ctypes.cast(flows, ctypes.c_void_p), L.ptr(seg), L.ptr(targets),
                                         L.ptr(out), self.kind, EPSILON, R_CUTOFF, L.ptr(table),
                                         ctypes.cast(staging, ctypes.c_void_p), L.ptr(ws), L.stream())
        L.check(rc, "fj_pair_sum_batched")
        self._keep = (table, ws)

    # ---- one engine tick: every active case advances one step ----------------
    def _tick_plan(self, slots):
        """Case records of a set of active cases: the fields that stay fixed
        while the set does, the two buffers of each (previous, current) pair,
        the per-step strides, and the workspace and table sizes.  Rebuilt
        when the set changes."""
        nc = len(slots)
        rec = np.zeros(nc, dtype=L.STEP_CASE_DTYPE)
        col = lambda f: np.array([f(sl) for sl in slots], dtype=np.uint64)  # noqa: E731
        rec["m"] = [sl.req.m for sl in slots]
        rec["n"] = [sl.req.n for sl in slots]
        rec["m_star"] = [sl.req.m_star for sl in slots]
        rec["dt"] = [sl.dt for sl in slots]
        for field, f in (("dof", lambda sl: sl.dof.data_ptr()), ("zeta0", lambda sl: sl.zeta0.data_ptr()),
                         ("zeta_b_dot", lambda sl: sl.zbd.data_ptr()), ("aic", lambda sl: sl.A.data_ptr()),
                         ("aic_inv", lambda sl: sl.Ainv.data_ptr()), ("gamma_b_dot", lambda sl: sl.gbd.data_ptr()),
                         ("f_steady", lambda sl: sl.fs.data_ptr()), ("refine_ws", lambda sl: sl.ws.data_ptr())):
            rec[field] = col(f)
        iters = self._grow("iters", 4 * nc)  # one stop-test word per case of the tick
        rec["iters"] = iters.data_ptr() + 4 * np.arange(nc, dtype=np.uint64)
        plan = {
            "key": tuple(id(sl) for sl in slots), "rec": rec, "iters": iters,
            "dt": rec["dt"].copy(),
            "hg": col(lambda sl: sl.hg.data_ptr()), "hgd": col(lambda sl: sl.hgd.data_ptr()),
            "hg_stride": col(lambda sl: sl.hg.stride(0) * 8), "lift": col(lambda sl: sl.lift.data_ptr()),
        }
        for name in ("zb", "gb", "zw", "gw"):
            plan[name] = (col(lambda sl, nm=name: getattr(sl, nm + "_pair")[0].data_ptr()),
                          col(lambda sl, nm=name: getattr(sl, nm + "_pair")[1].data_ptr()))
        lib = L.lib()
        plan["ws"] = lib.fj_step_batched_workspace_bytes(nc, ctypes.c_void_p(rec.ctypes.data))
        plan["table"] = lib.fj_step_batched_table_bytes(nc)
        return plan

    def _tick(self, slots, parity):
        """Enqueue one time step of every active case in one native call
        (fj_step_batched): geometry, wake shift and rollup, upwash, AIC,
        right-hand side and kept-inverse refinement, forces and lift of all
        cases, each stage one launch over all cases.  Returns (event, host
        status, slots); nothing waits for the device here.  ``parity``
        selects one of two host staging / status buffers: the tick enqueued
        two ticks earlier has completed when this one is built."""
        nc = len(slots)
        plan = self._plan
        if plan is None or plan["key"] != tuple(id(sl) for sl in slots):
            plan = self._plan = self._tick_plan(slots)
        rec = self._fill_records(plan, slots)
        lib = L.lib()
        ws = self._grow("ws", plan["ws"])
        table = self._grow("table", plan["table"])
        staging = self._grow(f"staging{parity}", plan["table"], pinned=True)
        fd = self.flow.descriptor()
        prec = L.PREC_MIXED if self.precision == "mixed" else L.PREC_FP64
        args = (nc, ctypes.c_void_p(rec.ctypes.data), ctypes.byref(fd), self.flow.rho, 0.7, self.kind, EPSILON,
                R_CUTOFF, prec, int(self.free_wake), 3, 0.0, L.ptr(table), ctypes.c_void_p(staging.data_ptr()),
                L.ptr(ws))
        if self.graph_ticks:
            # the graph holds this set's buffers, so a new set drops the
            # graphs of the previous one (its buffers may have been regrown)
            if self._graph_set != plan["key"]:
                self._graphs.clear()
                self._graph_set = plan["key"]
            g = self._graphs.get(parity)
            if g is None:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    rc = lib.fj_step_batched(*args, L.stream())
                L.check(rc, "fj_step_batched")
                self._graphs[parity] = g
                self.graph_captures += 1
            else:
                L.check(lib.fj_step_batched_restage(*args, L.stream()), "fj_step_batched_restage")
            g.replay()
        else:
            rc = lib.fj_step_batched(*args, L.stream())
            L.check(rc, "fj_step_batched")
        for sl in slots:
            sl.i += 1
        status_host = self._grow(f"status{parity}", 4 * nc, pinned=True)[:4 * nc].view(torch.int32)
        status_host.copy_(plan["iters"][:4 * nc].view(torch.int32), non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        return ev, status_host, list(slots)

    @staticmethod
    def _fill_records(plan, slots):
        """The per-step fields of the case records for the next step of
        every case (time, pose, lift slot, the state buffer pairs' roles)."""
        rec = plan["rec"]
        i = np.fromiter((sl.i + 1 for sl in slots), dtype=np.uint64, count=len(slots))
        odd = (i & np.uint64(1)).astype(bool)
        rec["time"] = i.astype(np.float64) * plan["dt"]
        rec["hg"] = plan["hg"] + i * plan["hg_stride"]
        rec["hg_dot"] = plan["hgd"] + i * plan["hg_stride"]
        rec["lift"] = plan["lift"] + np.uint64(8) * i
        # buffer pairs (A, B): step i writes its state into B and reads the
        # previous step's from A when i is odd, the other way round when even
        for cur, prev, name in (("zeta_b", "zeta_b_nm1", "zb"), ("gamma_b", "gamma_b_nm1", "gb"),
                                ("zeta_w", "zeta_w_nm1", "zw"), ("gamma_w", "gamma_w_nm1", "gw")):
            a, b = plan[name]
            rec[cur] = np.where(odd, b, a)
            rec[prev] = np.where(odd, a, b)
        return rec

    @staticmethod
    def _check(tick):
        """Wait for an enqueued tick and check every case's stop test."""
        ev, status_host, slots = tick
        ev.synchronize()
        bad = [sl.req.case_id for sl, it in zip(slots, status_host.tolist()) if it < 0]
        if bad:
            raise L.FlapjaxLibError(f"circulation solve failed its stop test for cases {bad}")

    def _grow(self, name, nbytes, pinned=False):
        """Byte buffer kept across ticks, grown when a tick needs more."""
        buf = self._bufs.get(name)
        if buf is None or buf.numel() < nbytes:
            size = max(int(nbytes), 256)
            if pinned:
                buf = torch.empty(size, dtype=torch.uint8, pin_memory=True)
            else:
                buf = torch.empty(size, dtype=torch.uint8, device=self.dev)
            self._bufs[name] = buf
        return buf

    def run(self, requests, time_source=time.perf_counter, log=None):
        """Serve requests (sorted by arrival, seconds from start).  Returns
        (results, stats)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        active: list[_Slot] = []
        warming: list[_Slot] = []
        used = 0
        results = []
        t0 = time_source()
        ticks = 0
        pairs = 0
        peak = 0
        main = torch.cuda.current_stream(self.dev)
        if self._adm is None:
            self._adm = torch.cuda.Stream(device=self.dev)
        adm = self._adm
        inflight = None  # the enqueued tick not yet waited for, and the cases it finishes
        while pending or active or warming or inflight is not None:
            now = time_source() - t0
            # admission: arrived, slot free, footprint fits.  A new case's
            # set-up (grids, AIC, factorisation and inverse, static solve) is
            # enqueued on the admission stream, so it runs on the device
            # beside the ticks of the cases already active; the case joins
            # the ticks once its set-up has completed.  The admission stream
            # first waits for the work already on the main stream (memory
            # that finished cases released may still be read there); a
            # case's buffers are only released after the tick that last used
            # them has been synchronised.
            while pending and pending[0].arrival <= now and len(active) + len(warming) < self.max_slots:
                fp = footprint_bytes(pending[0], self.precision)
                if used + fp > self.budget and (active or warming or inflight is not None):
                    break
                req = pending.pop(0)
                adm.wait_stream(main)
                with torch.cuda.stream(adm):
                    sl = _Slot(req, self.dev, self.precision, self.flow, self.kind, self._cache)
                    sl.ready = torch.cuda.Event()
                    sl.ready.record(adm)
                sl.start = time_source() - t0
                sl.fp = fp
                used += fp
                peak = max(peak, used)
                warming.append(sl)
            if warming and not active and inflight is None:
                warming[0].ready.synchronize()
            for sl in [w for w in warming if w.ready.query()]:
                main.wait_event(sl.ready)
                warming.remove(sl)
                active.append(sl)
            # ticks are pipelined: tick k is enqueued before the host waits
            # for tick k-1, so the device always has the next tick queued
            # while the host checks the previous one and admits cases
            cur = None
            if active:
                tick = self._tick(active, ticks % 2)
                ticks += 1
                for sl in active:
                    r = sl.req
                    s_b, s_w = sl.s_b, sl.s_w
                    pairs += (r.m_star * (r.n + 1) * (s_b + s_w) if self.free_wake else 0) + \
                        r.m * r.n * s_w + s_b * (s_b + s_w)
                cur = (tick, [sl for sl in active if sl.done])
                active = [sl for sl in active if not sl.done]
            if inflight is not None:
                tick, finished = inflight
                self._check(tick)
                now = time_source() - t0
                for sl in finished:
                    r = sl.req
                    gb, zw = sl.final_state()
                    results.append(CaseResult(r.case_id, r.arrival, sl.start, now, sl.lift.cpu().numpy(),
                                              gb.cpu().numpy(), zw.cpu().numpy()))
                    used -= sl.fp
            inflight = cur
            if cur is None and not active and not warming and pending:
                wait = pending[0].arrival - (time_source() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            if log is not None and ticks % 50 == 0:
                log(f"tick {ticks}: active {len(active)}, pending {len(pending)}, done {len(results)}")
        makespan = time_source() - t0
        lat = np.array([r.latency for r in results])
        stats = {"cases": len(results), "makespan_s": makespan, "cases_per_s": len(results) / makespan,
                 "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)), "ticks": ticks,
                 "pairs": pairs, "gpairs_per_s": pairs / makespan / 1e9, "peak_bytes": peak}
        return results, stats


def poisson_stream(n_cases: int, rate: float, seed: int = 2, m_star: int | None = None, steps_range=(50, 400)):
    """The served mixed-size stream: M in {4,8,16,32}, N in {16,32,64,128},
    steps uniform in steps_range, pitch amplitude 0-5 deg, k 0.05-0.5,
    Poisson arrivals at `rate` cases/s."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_cases))
    reqs = []
    for i in range(n_cases):
        m = int(rng.choice([4, 8, 16, 32]))
        n = int(rng.choice([16, 32, 64, 128]))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        ms = min(steps, m_star) if m_star is not None else steps
        reqs.append(CaseRequest(i, m, n, steps, ms, float(rng.uniform(0, 5)), float(rng.uniform(0.05, 0.5)),
                                arrival=float(t[i])))
    return reqs
