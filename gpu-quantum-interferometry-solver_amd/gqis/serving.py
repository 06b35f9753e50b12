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

from . import autotune, codegen, device, runtime
from .solver import _emit_source, _runtime_const_values, _check_tlist, grouped_applies
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
                 max_batch_points=1 << 22, block_size="auto"):
        import torch
        self.torch = torch
        self.ctx = runtime.Context.get(device)
        self.dev = torch.device("cuda", self.ctx.device)
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(streams)]
        self.free_streams = deque(self.streams)
        self.budget = int(hbm_budget_bytes)
        self.in_use = 0
        self.max_batch_points = int(max_batch_points)
        self.block = block_size if block_size == "auto" else int(block_size)
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
        """Kernel, constants and time table of a batch key, resident on the device."""
        m = self.modules.get(key)
        if m is not None:
            return m
        torch = self.torch
        num_steps, dt, _, _, _ = _check_tlist(r.tlist)
        spec = get_solver_spec(r.solver, step_size=dt, adaptive_options=r.adaptive_options,
                               fp64=r.fp64)
        rdt = torch.float64 if r.fp64 else torch.float32
        consts_vals = list(key[-1]) or [0.0]
        hm = self._hilbert_model(r) if spec.name == "rk4" else None
        if hm is not None:
            from . import hilbert
            from .solver import SPLIT_F16_MAX_ENTRY
            # FP32 requests use the FP16 hi/lo products, as mesolve_2D does
            split = ("f16" if not r.fp64 and np.max(np.abs(hm.h_eff)) <= SPLIT_F16_MAX_ENTRY
                     else None)
            cvals, offs = hilbert.constants(hm, list(key[-1]), split)
            block = hilbert.block_for(hm, r.fp64, 256 if self.block == "auto" else self.block,
                                      split=split)
            src = hilbert.emit(hm, offs, r.output_mode, block, r.fp64, split).replace(
                "GQ_RHO0_FROM_ARGS", "0")
            kname = codegen.kernel_name(hilbert.kernel_base(split, r.fp64), r.fp64)
            natoms = int(src.split("#define GQ_NATOMS ")[1].split()[0])
            cands = tuple((1, b) for b in autotune.CANDIDATES
                          if hilbert.lds_bytes(hm.n, r.fp64, b, hilbert.jump_width(hm), split,
                                               mono=hilbert.n_mono(hm), nbasis=len(hm.basis))
                          <= hilbert.LDS_BYTES_PER_CU)

            def variant(c):
                return (autotune.with_block(src, c[1]), kname,
                        (c[1] // 64) * hilbert.points_per_wave(hm.n, r.fp64, split=split))
            consts_vals = cvals
            table_kernel = codegen.kernel_name("gq_rk4_time_table", r.fp64)
            rows = 2 * num_steps + 1
        else:
            block = 256 if self.block == "auto" else self.block
            src = _emit_source(r.compiled, spec, r.output_mode, False, False, False, r.fp64, True,
                               block, False)
            kname = codegen.kernel_name(spec.kernel, r.fp64)
            natoms = len(r.compiled.plan.time_atoms)
            size = r.compiled.state_size
            cands = tuple((1, b) for b in autotune.CANDIDATES)
            if size >= device.MIN_GROUPED_STATES:
                cands += tuple((g, 64 * g) for g in device.GROUP_CHOICES
                               if grouped_applies(spec, size, r.fp64, g, natoms))

            def variant(c):
                g, b = c
                text = src if g == 1 else _emit_source(r.compiled, spec, r.output_mode, False,
                                                       False, False, r.fp64, True, b, False, g)
                return (autotune.with_block(text, b), device._kernel_for(spec, g, r.fp64),
                        64 if g > 1 else b)
            table_kernel = codegen.kernel_name(
                "gq_rk4_time_table" if spec.template == "fixed_rk4.hip" else "gq_fixed_time_table",
                r.fp64)
            rows = rows_per_step(spec.name) * num_steps + 1
        mod, _, _ = self.ctx.compile(src, codegen.HIPRTC_BASE_OPTIONS)
        consts = torch.as_tensor(np.asarray(consts_vals, np.float64), dtype=rdt, device=self.dev)
        table = None
        if not spec.adaptive and natoms:
            table = torch.empty(rows * natoms, dtype=rdt, device=self.dev)
            a = runtime.SweepArgs()
            a.dt, a.t_in, a.num_steps = float(dt), 0.0, num_steps
            a.consts, a.time_table = consts.data_ptr(), table.data_ptr()
            self.ctx.launch(mod, table_kernel, (rows + 255) // 256, 256, a,
                            torch.cuda.current_stream(self.dev).cuda_stream)
            torch.cuda.synchronize(self.dev)
        m = {"module": mod, "spec": spec, "num_steps": num_steps, "dt": dt, "kernel": kname,
             "block": block, "per_block": variant((1, block))[2], "groups": 1,
             "consts": consts, "table": table, "hilbert": hm is not None, "autotune": {}}
        if self.block == "auto":
            choice, m["autotune"] = autotune.choose(
                self.ctx, src, codegen.HIPRTC_BASE_OPTIONS,
                lambda module, c: self._trial(m, r, module, c, *variant(c)[1:]), cands,
                source_of=lambda c: variant(c)[0])
            text, m["kernel"], m["per_block"] = variant(choice)
            m["module"], _, _ = self.ctx.compile(text, codegen.HIPRTC_BASE_OPTIONS)
            m["groups"], m["block"] = choice
        self.modules[key] = m
        return m

    def _trial(self, m, r, module, cand, kernel, per_block, side=256, steps=64):
        """One timed launch of a candidate (lanes per point, work-group size): side x side
        points spread over the request's parameter ranges, the first ``steps`` steps of its
        time grid."""
        torch = self.torch
        rdt = torch.float64 if r.fp64 else torch.float32
        x = np.linspace(float(np.min(r.axes[0])), float(np.max(r.axes[0])), side)
        y = (np.linspace(float(np.min(r.axes[1])), float(np.max(r.axes[1])), side)
             if len(r.axes) > 1 else np.zeros(1))
        npts = x.size * y.size
        pxd = torch.as_tensor(np.repeat(x, y.size), dtype=rdt, device=self.dev)
        pyd = torch.as_tensor(np.tile(y, x.size), dtype=rdt, device=self.dev)
        size = r.compiled.state_size
        if r.output_mode in ("final_rho", "mean_state"):
            result = torch.empty((size, npts), dtype=rdt, device=self.dev)
        else:
            result = torch.empty(npts, dtype=torch.complex128 if r.fp64 else torch.complex64,
                                 device=self.dev)
        flags = torch.zeros(2, dtype=torch.int32, device=self.dev)
        spec = m["spec"]
        a = runtime.SweepArgs()
        a.dt, a.t_in = float(m["dt"]), 0.0
        a.num_x, a.num_y, a.point_begin, a.point_count = npts, 1, 0, npts
        a.num_steps, a.warmup_steps = min(steps, m["num_steps"]), 0
        a.trace_stride, a.trace_point_stride, a.param_mode = 1, npts, 1
        a.par_x, a.par_y, a.consts = pxd.data_ptr(), pyd.data_ptr(), m["consts"].data_ptr()
        a.time_table = m["table"].data_ptr() if m["table"] is not None else None
        for i, v in enumerate(spec.parameters[:runtime.MAX_SOLVER_PARAMS]):
            a.solver_param[i] = float(v)
        a.result, a.status, a.work_cursor = result.data_ptr(), flags.data_ptr(), flags.data_ptr() + 4
        groups, block = cand
        grid = (npts + per_block - 1) // per_block
        if spec.adaptive and spec.options.get("persistent", True) and groups == 1:
            grid = min(grid, self.ctx.resident_groups(module, kernel, block))
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        return device._timed_launches(self.ctx, module, kernel, max(1, grid), block, a, stream,
                                      flags[1:])

    def _hilbert_model(self, r):
        """Matrix-core model of a Lindblad request whose operators fit the Hilbert-space kernel
        (N = 32; N = 16 with a dense H_eff; N = 8 with a dense H_eff in FP64), else None
        (smaller or sparser models run faster on the per-point kernels); the same choice
        mesolve_2D makes."""
        if r.compiled.kind != "lindblad" or r.output_mode not in ("mean", "final", "final_rho"):
            return None
        n = int(round(np.sqrt(r.compiled.state_size + 1)))
        if n < (8 if r.fp64 else 16):
            return None
        src = getattr(r.compiled, "source_model", None)
        if src is None:
            return None
        from . import hilbert
        from .solver import _drive_map
        hm = hilbert.extract(src, _drive_map(src["Drive"]))
        if hm is not None and n < 32 and not hilbert.dense_enough(hm):
            return None
        return hm

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

    def _pinned(self, nbytes):
        key = 1 << max(12, (int(nbytes) - 1).bit_length())
        pool = self.host_pool.setdefault(key, [])
        buf = pool.pop() if pool else self.torch.empty(key, dtype=self.torch.uint8, pin_memory=True)
        return key, buf

    def _launch(self, key, batch, npts, nbytes):
        torch = self.torch
        h0 = time.perf_counter()
        ent = self._module(batch[0], key)
        spec, num_steps, dt = ent["spec"], ent["num_steps"], ent["dt"]
        r0 = batch[0]
        rdt = torch.float64 if r0.fp64 else torch.float32
        cdt = torch.complex128 if r0.fp64 else torch.complex64
        npdt = np.float64 if r0.fp64 else np.float32
        offsets, xs, ys = [], [], []
        off = 0
        for r in batch:
            x = np.asarray(r.axes[0], np.float64)
            y = np.asarray(r.axes[1], np.float64) if len(r.axes) > 1 else np.zeros(1)
            xs.append(x)
            ys.append(y)
            offsets.append(off)
            off += len(x) * len(y)
        axes_host = np.concatenate(xs + ys).astype(npdt)
        stage_key, stage_buf = self._pinned(axes_host.nbytes)
        stage = stage_buf[:axes_host.nbytes].view(rdt)
        stage.numpy()[:] = axes_host
        stream = self.free_streams.popleft()
        size = r0.compiled.state_size
        mode = r0.output_mode
        with torch.cuda.stream(stream):
            start_ev = torch.cuda.Event(enable_timing=True)
            stop_ev = torch.cuda.Event(enable_timing=True)
            ax = stage.to(self.dev, non_blocking=True)
            parts_x, parts_y, pos = [], [], 0
            for x, y in zip(xs, ys):
                parts_x.append(ax[pos:pos + len(x)].repeat_interleave(len(y)))
                pos += len(x)
            for x, y in zip(xs, ys):
                parts_y.append(ax[pos:pos + len(y)].repeat(len(x)))
                pos += len(y)
            pxd = torch.cat(parts_x)
            pyd = torch.cat(parts_y)
            if mode in ("final_rho", "mean_state"):
                result = torch.empty((size, npts), dtype=rdt, device=self.dev)
            else:
                result = torch.empty(npts, dtype=cdt, device=self.dev)
            flags_dev = torch.zeros(2, dtype=torch.int32, device=self.dev)   # status, cursor
            a = runtime.SweepArgs()
            a.dt, a.t_in = float(dt), 0.0
            a.num_x, a.num_y = npts, 1
            a.point_begin, a.point_count = 0, npts
            a.num_steps, a.warmup_steps = num_steps, int(key[7])
            a.trace_stride, a.trace_offset, a.num_trace, a.trace_point_stride = 1, 0, 0, npts
            a.param_mode = 1
            a.par_x, a.par_y, a.consts = pxd.data_ptr(), pyd.data_ptr(), ent["consts"].data_ptr()
            a.time_table = ent["table"].data_ptr() if ent["table"] is not None else None
            for i, v in enumerate(spec.parameters[:runtime.MAX_SOLVER_PARAMS]):
                a.solver_param[i] = float(v)
            a.result = result.data_ptr()
            a.status = flags_dev.data_ptr()
            a.work_cursor = flags_dev.data_ptr() + 4
            sid = stream.cuda_stream
            start_ev.record(stream)
            grid = (npts + ent["per_block"] - 1) // ent["per_block"]
            if spec.adaptive and spec.options.get("persistent", True) and ent["groups"] == 1:
                if "resident" not in ent:
                    ent["resident"] = self.ctx.resident_groups(ent["module"], ent["kernel"],
                                                               ent["block"])
                grid = min(grid, ent["resident"])
            self.ctx.launch(ent["module"], ent["kernel"], max(1, grid), ent["block"], a, sid)
            stop_ev.record(stream)
            nb = result.numel() * result.element_size()
            host_key, buf = self._pinned(nb + 8)
            host = buf[:nb].view(result.dtype).view(result.shape)
            host.copy_(result, non_blocking=True)
            status_host = buf[nb:nb + 8].view(torch.int32)
            status_host.copy_(flags_dev, non_blocking=True)
            event = torch.cuda.Event()
            event.record(stream)
        launch = _Launch(batch, stream, event, (result, pxd, pyd, flags_dev, ax), host, flags_dev,
                         status_host, nbytes, offsets, mode, size, start_ev, stop_ev,
                         [(host_key, buf), (stage_key, stage_buf)])
        self.running.append(launch)
        self.in_use += nbytes
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        self.launches += 1
        for r in batch:
            r.batch_size = len(batch)
        self.host_launch_s += time.perf_counter() - h0

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
            for k, b in l.host_key:
                self.host_pool.setdefault(k, []).append(b)
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
    def serve(self, requests, realtime=True, poll_s=0.0005, progress=None, progress_every=100):
        """Feed requests at their arrival times (seconds from start; all at once when
        ``realtime`` is False) and run to completion.  Returns the finished requests with
        ``done_at`` relative to the start.  ``progress(finished, submitted, t)`` is called each
        time another ``progress_every`` requests have finished."""
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
            before = len(self.finished)
            completed = self._complete(now)
            if progress is not None and (len(self.finished) // progress_every
                                         > before // progress_every):
                progress(len(self.finished), i, now())
            if not completed:
                if realtime and i < len(order) and not self.running:
                    time.sleep(max(0.0, min(poll_s, order[i].arrival - now())))
                else:
                    time.sleep(poll_s)
        return self.finished
