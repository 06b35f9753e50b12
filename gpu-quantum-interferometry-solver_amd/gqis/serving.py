"""Continuous-batching sweep server on one GPU.

Requests (a compiled model, its sweep axes, tlist, solver, precision, output mode) arrive over
time.  Each is flattened to per-point parameter lists; pending requests with the same batch key
(model, solver, precision, output mode, time grid, constants) are merged into one launch
(``param_mode`` 1 of ``gq_sweep_args``), up to ``max_batch_points``.  A launch is admitted when a
stream slot is free and its device memory fits the HBM budget; launches on different streams
overlap on the device, and requests arriving while others run are admitted as slots and memory
free up.  Results are copied to pinned host memory on the launch's stream; completion is polled
with events, and each request's latency is arrival -> host result.

Per-point arithmetic does not depend on the batch (one lane per point, same kernel), so a
request's result is bitwise the same alone and in the mix.
"""
from collections import deque
from dataclasses import dataclass, field
import itertools
import time

import numpy as np

from . import codegen, runtime
from .solver import _emit_source, _runtime_const_values, _check_tlist
from .solvers import get_solver_spec
from .tableaux import rows_per_step


@dataclass
class SweepRequest:
    """One sweep: ``compiled`` from compile_lindblad_rhs / compile_ode_rhs, 1 or 2 axes."""
    compiled: object
    tlist: object
    axes: tuple
    solver: str = "rk4"
    fp64: bool = False
    output_mode: str = "mean"
    adaptive_options: dict = None
    runtime_consts: dict = None
    warmup_time: float = 0.0
    arrival: float = 0.0
    rid: int = field(default_factory=itertools.count().__next__)
    # filled by the server
    result: object = None
    done_at: float = None
    status: int = 0
    batch_size: int = 0

    @property
    def shape(self):
        nx = len(self.axes[0])
        ny = len(self.axes[1]) if len(self.axes) > 1 else 1
        return nx, ny


class _Launch:
    def __init__(self, requests, stream, event, result, host, status_dev, status_host, nbytes,
                 offsets, mode, size, start_event, stop_event, host_key):
        self.requests, self.stream, self.event = requests, stream, event
        self.result, self.host, self.status_dev, self.status_host = result, host, status_dev, status_host
        self.nbytes, self.offsets, self.mode, self.size = nbytes, offsets, mode, size
        self.start_event, self.stop_event, self.host_key = start_event, stop_event, host_key


class SweepServer:
    def __init__(self, device=0, streams=4, hbm_budget_bytes=16 << 30,
                 max_batch_points=1 << 22, block_size=256):
        import torch
        self.torch = torch
        self.ctx = runtime.Context.get(device)
        self.dev = torch.device("cuda", self.ctx.device)
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(streams)]
        self.free_streams = deque(self.streams)
        self.budget = int(hbm_budget_bytes)
        self.in_use = 0
        self.max_batch_points = int(max_batch_points)
        self.block = int(block_size)
        self.pending = {}        # batch key -> deque of requests
        self.running = []
        self.finished = []
        self.modules = {}
        self.peak_in_use = 0
        self.launches = 0
        self.gpu_kernel_s = 0.0       # sum of launch durations (they may overlap)
        self.host_launch_s = 0.0
        self.host_pool = {}           # pinned host buffers by (bytes class)
        self.gpu_by_kind = {}         # (N, solver, precision) -> [launches, points, GPU seconds]

    # -- preparation ---------------------------------------------------------------------------
    def _key(self, r):
        num_steps, dt, _, _, _ = _check_tlist(r.tlist)
        spec = get_solver_spec(r.solver, step_size=dt, adaptive_options=r.adaptive_options,
                               fp64=r.fp64)
        consts = tuple(_runtime_const_values(r.compiled, r.runtime_consts, None, True))
        warm = int(np.floor(float(r.warmup_time) * num_steps)) if r.warmup_time < 1 else num_steps
        opts = tuple(sorted((k, str(v)) for k, v in spec.options.items()))
        return (r.compiled.cache_key, spec.name, opts, r.output_mode, bool(r.fp64), num_steps,
                float(dt), warm, consts)

    def _module(self, r, key):
        m = self.modules.get(key)
        if m is None:
            num_steps, dt, _, _, _ = _check_tlist(r.tlist)
            spec = get_solver_spec(r.solver, step_size=dt, adaptive_options=r.adaptive_options,
                                   fp64=r.fp64)
            src = _emit_source(r.compiled, spec, r.output_mode, False, False, False, r.fp64, True,
                               self.block, False)
            mod, _, _ = self.ctx.compile(src, codegen.HIPRTC_BASE_OPTIONS)
            m = (mod, spec, num_steps, dt)
            self.modules[key] = m
        return m

    def warm(self, requests):
        """Compile every kernel variant the requests need (before the stream starts)."""
        for r in requests:
            self._module(r, self._key(r))

    # -- queueing ------------------------------------------------------------------------------
    def submit(self, r):
        self.pending.setdefault(self._key(r), deque()).append(r)

    def _bytes(self, r, npts):
        item = 8 if r.fp64 else 4
        size = r.compiled.state_size
        out = npts * (size if r.output_mode in ("final_rho", "mean_state") else 2) * item
        return out + 2 * npts * item + 4096

    def _admit(self, now):
        started = 0
        while self.free_streams:
            # oldest waiting request first
            best = None
            for key, q in self.pending.items():
                if q and (best is None or q[0].arrival < self.pending[best][0].arrival):
                    best = key
            if best is None:
                break
            q = self.pending[best]
            batch, npts = [], 0
            while q:
                nx, ny = q[0].shape
                if batch and npts + nx * ny > self.max_batch_points:
                    break
                r = q.popleft()
                batch.append(r)
                npts += nx * ny
            need = self._bytes(batch[0], npts)
            if self.in_use + need > self.budget and self.running:
                for r in reversed(batch):
                    q.appendleft(r)
                break
            self._launch(best, batch, npts, need)
            started += 1
        return started

    def _launch(self, key, batch, npts, nbytes):
        torch = self.torch
        mod, spec, num_steps, dt = self._module(batch[0], key)
        r0 = batch[0]
        rdt = torch.float64 if r0.fp64 else torch.float32
        cdt = torch.complex128 if r0.fp64 else torch.complex64
        h0 = time.perf_counter()
        offsets, xs, ys = [], [], []
        off = 0
        for r in batch:
            x = np.asarray(r.axes[0], np.float64)
            y = np.asarray(r.axes[1], np.float64) if len(r.axes) > 1 else np.zeros(1)
            xs.append(x)
            ys.append(y)
            offsets.append(off)
            off += len(x) * len(y)
        stream = self.free_streams.popleft()
        size = r0.compiled.state_size
        mode = r0.output_mode
        with torch.cuda.stream(stream):
            start_ev = torch.cuda.Event(enable_timing=True)
            stop_ev = torch.cuda.Event(enable_timing=True)
            # axes go up as they are (small); the per-point lists are expanded on the device
            ax = torch.as_tensor(np.concatenate(xs + ys), dtype=rdt).to(self.dev, non_blocking=True)
            parts_x, parts_y, pos = [], [], 0
            for x, y in zip(xs, ys):
                parts_x.append(ax[pos:pos + len(x)].repeat_interleave(len(y)))
                pos += len(x)
            for x, y in zip(xs, ys):
                parts_y.append(ax[pos:pos + len(y)].repeat(len(x)))
                pos += len(y)
            pxd = torch.cat(parts_x)
            pyd = torch.cat(parts_y)
            consts = torch.as_tensor(np.asarray(key[-1] or (0.0,), np.float64), dtype=rdt,
                                     device=self.dev)
            if mode in ("final_rho", "mean_state"):
                result = torch.empty((size, npts), dtype=rdt, device=self.dev)
            else:
                result = torch.empty(npts, dtype=cdt, device=self.dev)
            status = torch.zeros(1, dtype=torch.int32, device=self.dev)
            cursor = torch.zeros(1, dtype=torch.int32, device=self.dev)
            natoms = len(r0.compiled.plan.time_atoms)
            rows = rows_per_step(spec.name) * num_steps + 1
            table = (torch.empty(rows * natoms, dtype=rdt, device=self.dev)
                     if not spec.adaptive and natoms else None)
            a = runtime.SweepArgs()
            a.dt, a.t_in = float(dt), 0.0
            a.num_x, a.num_y = npts, 1
            a.point_begin, a.point_count = 0, npts
            a.num_steps, a.warmup_steps = num_steps, int(key[7])
            a.trace_stride, a.trace_offset, a.num_trace, a.trace_point_stride = 1, 0, 0, npts
            a.param_mode = 1
            a.par_x, a.par_y, a.consts = pxd.data_ptr(), pyd.data_ptr(), consts.data_ptr()
            a.time_table = table.data_ptr() if table is not None else None
            for i, v in enumerate(spec.parameters[:runtime.MAX_SOLVER_PARAMS]):
                a.solver_param[i] = float(v)
            a.result, a.status, a.work_cursor = result.data_ptr(), status.data_ptr(), cursor.data_ptr()
            sid = stream.cuda_stream
            start_ev.record(stream)
            if table is not None:
                self.ctx.launch(mod, codegen.kernel_name(
                    "gq_rk4_time_table" if spec.template == "fixed_rk4.hip" else "gq_fixed_time_table",
                    r0.fp64), (rows + 255) // 256, 256, a, sid)
            kname = codegen.kernel_name(spec.kernel, r0.fp64)
            grid = (npts + self.block - 1) // self.block
            if spec.adaptive and spec.options.get("persistent", True):
                grid = min(grid, self.ctx.resident_groups(mod, kname, self.block))
            self.ctx.launch(mod, kname, max(1, grid), self.block, a, sid)
            stop_ev.record(stream)
            nb = result.numel() * result.element_size()
            host_key = (1 << max(12, (nb - 1).bit_length()))
            pool = self.host_pool.setdefault(host_key, [])
            buf = pool.pop() if pool else torch.empty(host_key, dtype=torch.uint8, pin_memory=True)
            host = buf[:nb].view(result.dtype).view(result.shape)
            host.copy_(result, non_blocking=True)
            status_host = torch.empty(1, dtype=torch.int32, pin_memory=True)
            status_host.copy_(status, non_blocking=True)
            event = torch.cuda.Event()
            event.record(stream)
        # keep device tensors alive until the launch completes
        launch = _Launch(batch, stream, event, (result, pxd, pyd, consts, status, cursor, table, ax),
                         host, status, status_host, nbytes, offsets, mode, size, start_ev, stop_ev,
                         (host_key, buf))
        self.host_launch_s += time.perf_counter() - h0
        self.running.append(launch)
        self.in_use += nbytes
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        self.launches += 1
        for r in batch:
            r.batch_size = len(batch)

    def _complete(self, now_fn):
        done = [l for l in self.running if l.event.query()]
        for l in done:
            self.running.remove(l)
            self.in_use -= l.nbytes
            self.free_streams.append(l.stream)
            t = now_fn()
            flags = int(l.status_host[0])
            ks = l.start_event.elapsed_time(l.stop_event) / 1000.0
            self.gpu_kernel_s += ks
            r0 = l.requests[0]
            kind = (r0.compiled.state_size, r0.solver, "fp64" if r0.fp64 else "fp32")
            acc = self.gpu_by_kind.setdefault(kind, [0, 0, 0.0])
            acc[0] += 1
            acc[1] += sum(r.shape[0] * r.shape[1] for r in l.requests)
            acc[2] += ks
            host = l.host.numpy().copy()
            self.host_pool.setdefault(l.host_key[0], []).append(l.host_key[1])
            for r, off in zip(l.requests, l.offsets):
                nx, ny = r.shape
                if l.mode in ("final_rho", "mean_state"):
                    r.result = host[:, off:off + nx * ny].T.reshape(nx, ny, l.size)
                else:
                    r.result = host[off:off + nx * ny].reshape(nx, ny)
                # the launch's status word covers the whole batch; each request checks its own
                # slice so one bad request does not mark the others
                r.status = int(flags != 0 and not np.all(np.isfinite(r.result)))
                r.done_at = t
                self.finished.append(r)
        return len(done)

    # -- driving -------------------------------------------------------------------------------
    def serve(self, requests, realtime=True, poll_s=0.0005):
        """Feed requests at their arrival times (seconds from start; all at once when
        ``realtime`` is False) and run to completion.  Returns the finished requests with
        ``done_at`` relative to the start."""
        order = sorted(requests, key=lambda r: r.arrival)
        if not realtime:
            for r in order:
                r.arrival = 0.0
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0
        i = 0
        while i < len(order) or self.running or any(self.pending.values()):
            t = now()
            while i < len(order) and order[i].arrival <= t:
                self.submit(order[i])
                i += 1
            self._admit(t)
            if not self._complete(now):
                if realtime and i < len(order) and not self.running:
                    time.sleep(max(0.0, min(poll_s, order[i].arrival - now())))
                else:
                    time.sleep(poll_s)
        return self.finished
