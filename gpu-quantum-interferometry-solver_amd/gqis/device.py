"""Device side of one sweep call: buffers (torch), hiprtc module, launches, status, layouts.

Replaces the reference's CuPy upload / launch / copy-back block (``gqis/solver.py:2268-2362``).
Outputs are written by the kernels in coalesced point-minor layouts and returned in the
reference's shapes: ``(nx, ny)`` complex, ``(nx, ny, D)`` real, traces ``(nx, ny, n[, D])``.
"""
import time

import numpy as np

from . import autotune, codegen, runtime


def _axis(values, dtype, device):
    import torch
    if isinstance(values, torch.Tensor):
        return values.to(device=device, dtype=dtype).reshape(-1).contiguous()
    return torch.as_tensor(np.asarray(values, dtype=np.float64).reshape(-1), dtype=dtype,
                           device=device)


def _initial_states(rho0_values, nx, ny, size, dtype, device):
    import torch
    if isinstance(rho0_values, torch.Tensor):
        r = rho0_values.to(device=device, dtype=dtype)
    else:
        arr = np.asarray(rho0_values)
        if np.iscomplexobj(arr):
            raise ValueError("rho0_values must be real.")
        r = torch.as_tensor(arr.astype(np.float64), dtype=dtype, device=device)
    if tuple(r.shape) == (nx, size) and ny == 1:
        r = r.reshape(nx, 1, size)
    if tuple(r.shape) != (nx, ny, size):
        raise ValueError(f"rho0_values must have shape ({nx}, {ny}, {size}) "
                         f"or ({nx}, {size}) when num_Y=1; got {tuple(r.shape)}.")
    return r.reshape(nx * ny, size).contiguous()


GROUP_CHOICES = (2, 4, 8, 16)
MIN_GROUPED_STATES = 24     # below this one lane holds a point's state without spills


def _kernel_for(spec, groups, fp64):
    if groups > 1:
        name = ("gq_adaptive_grouped" if spec.adaptive else
                "gq_rk4_grouped" if spec.template == "fixed_rk4.hip" else "gq_fixed_grouped")
        return codegen.kernel_name(name, fp64)
    return codegen.kernel_name(spec.kernel, fp64)


def _grid(npts, groups, block):
    """Work-groups of a launch: one point per lane, or 64 points per group of waves."""
    per = 64 if groups > 1 else block
    return (npts + per - 1) // per


def run_sweep(*, source, variant, spec, compiled, mode, var_arrays, rho0_values, const_vals,
              fp64, dt, t_in, num_steps, warmup_steps, num_trace, trace_stride, trace_offset,
              state_trace, time_table, block_size, options, reuse, collect, emit=None,
              lanes_per_point=1, prepare=False):
    """Buffers, time table, launch and output layouts of one sweep.  ``source`` is the kernel
    for ``lanes_per_point`` lanes per point; ``emit(groups)`` returns the source of another
    grouping (used when ``block_size`` is "auto").  With ``prepare`` the buffers, module and
    launch chain are returned without launching (``collect_outputs`` turns them into results
    after a launch)."""
    import torch
    setup_start = time.time()
    ctx = runtime.Context.get()
    dev = torch.device("cuda", ctx.device)
    rdt = torch.float64 if fp64 else torch.float32
    cdt = torch.complex128 if fp64 else torch.complex64
    # the mixed-precision RK4 form keeps its sweep values and time table in FP64
    mixed = spec.template == "fixed_rk4_mixed.hip"
    adt = torch.float64 if mixed else rdt
    axes = list(var_arrays.values())
    px = _axis(axes[0], adt, dev)
    py = _axis(axes[1], adt, dev) if len(axes) > 1 else torch.zeros(1, dtype=adt, device=dev)
    nx, ny = int(px.numel()), int(py.numel())
    npts = nx * ny
    size = compiled.state_size

    options = codegen.HIPRTC_BASE_OPTIONS + tuple(options)
    groups = int(lanes_per_point)
    tune = block_size == "auto"
    if tune:
        block_size = 256       # the emitted source's size; measured below
    if groups > 1:
        block_size = 64 * groups
    module, cached, build_s = ctx.compile(source, options)
    rho0 = (_initial_states(rho0_values, nx, ny, size, rdt, dev)
            if rho0_values is not None else None)
    consts = torch.as_tensor(np.asarray(const_vals if const_vals else [0.0], np.float64),
                             dtype=rdt, device=dev)
    if mode in ("final_rho", "mean_state"):
        result = torch.empty((size, max(npts, 1)), dtype=rdt, device=dev)
    else:
        result = torch.empty(max(npts, 1), dtype=cdt, device=dev)
    trace = None
    if num_trace:
        # zero-filled: an adaptive point that fails stops writing its trace
        alloc = torch.zeros if spec.adaptive else torch.empty
        trace = (alloc((num_trace, size, npts), dtype=rdt, device=dev) if state_trace
                 else alloc((num_trace, npts), dtype=cdt, device=dev))
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    cursor = torch.zeros(1, dtype=torch.int32, device=dev)
    natoms = len(compiled.plan.time_atoms)
    table = None
    use_table = bool(time_table) and not spec.adaptive and natoms > 0
    from .tableaux import rows_per_step
    rows = rows_per_step(spec.name) * num_steps + 1
    if use_table:
        table = torch.empty(rows * natoms, dtype=adt, device=dev)
    steps_out = (torch.zeros((npts, 2), dtype=torch.int32, device=dev)
                 if spec.adaptive and spec.options.get("count_steps") else None)

    a = runtime.SweepArgs()
    a.dt = float(dt)
    a.t_in = float(t_in)
    a.num_x, a.num_y = nx, ny
    a.point_begin, a.point_count = 0, npts
    a.num_steps, a.warmup_steps = int(num_steps), int(warmup_steps)
    a.trace_stride, a.trace_offset, a.num_trace = int(trace_stride), int(trace_offset), int(num_trace)
    a.trace_point_stride = npts
    a.par_x, a.par_y, a.consts = px.data_ptr(), py.data_ptr(), consts.data_ptr()
    a.rho0 = rho0.data_ptr() if rho0 is not None else None
    a.time_table = table.data_ptr() if table is not None else None
    for i, v in enumerate(spec.parameters[:runtime.MAX_SOLVER_PARAMS]):
        a.solver_param[i] = float(v)
    a.result = result.data_ptr()
    a.trace = trace.data_ptr() if trace is not None else None
    a.adaptive_steps = steps_out.data_ptr() if steps_out is not None else None
    a.status = status.data_ptr()
    a.work_cursor = cursor.data_ptr()

    stream = torch.cuda.current_stream(dev).cuda_stream
    persistent = spec.adaptive and spec.options.get("persistent", True)
    tune_s, tune_times = 0.0, {}
    if tune and groups == 1 and npts > 0 and num_steps > 0:
        t0 = time.time()
        if use_table and rows > 0:
            ctx.launch(module, codegen.kernel_name(table_kernel(spec), fp64), (rows + 255) // 256,
                       256, a, stream)
        trial_steps = min(int(num_steps), 64)
        from .solver import grouped_applies
        cands = [(1, b) for b in autotune.CANDIDATES]
        if emit is not None and size >= MIN_GROUPED_STATES:
            cands += [(g, 64 * g) for g in GROUP_CHOICES
                      if grouped_applies(spec, size, fp64, g, natoms)]

        def source_of(cand):
            return autotune.with_block(source if cand[0] == 1 else emit(cand[0]), cand[1])

        def trial(mod, cand):
            g, block = cand
            t = runtime.SweepArgs.from_buffer_copy(a)
            t.num_steps, t.warmup_steps, t.num_trace, t.trace = trial_steps, 0, 0, None
            kn = _kernel_for(spec, g, fp64)
            grid = _grid(npts, g, block)
            if persistent and g == 1:
                grid = min(ctx.resident_groups(mod, kn, block), grid)
            return _timed_launches(ctx, mod, kn, max(grid, 1), block, t, stream, cursor)

        if npts * trial_steps >= autotune.MIN_TRIAL_POINT_STEPS:
            choice, tune_times = autotune.choose(ctx, source, options, trial, cands, source_of)
        else:
            key = autotune.source_key(source, options, ctx.arch)
            choice = autotune.lookup(ctx, key, cands) or (1, 256)
        groups, block_size = choice
        module, _, _ = ctx.compile(source_of(choice), options)
        ctx.read_status(status, stream, clear=True)
        cursor.zero_()
        tune_s = time.time() - t0
    kname = _kernel_for(spec, groups, fp64)
    launches = []
    if npts > 0 and num_steps >= 0:
        if use_table and rows > 0:
            launches.append((codegen.kernel_name(table_kernel(spec), fp64), (rows + 255) // 256,
                             256))
        if persistent and groups == 1:
            grid = min(ctx.resident_groups(module, kname, block_size),
                       (npts + block_size - 1) // block_size)
        else:
            grid = _grid(npts, groups, block_size)
        launches.append((kname, max(grid, 1), block_size))
    plan = {"ctx": ctx, "module": module, "args": a, "launches": launches, "fp64": fp64,
            "mode": mode, "nx": nx, "ny": ny, "size": size, "num_trace": num_trace,
            "state_trace": state_trace, "adaptive": spec.adaptive,
            "buffers": {"px": px, "py": py, "consts": consts, "rho0": rho0, "result": result,
                        "trace": trace, "status": status, "cursor": cursor, "table": table,
                        "steps": steps_out},
            "info": {"build_s": build_s, "cached_kernel": cached, "block_size": int(block_size),
                     "lanes_per_point": int(groups), "autotune_s": tune_s,
                     "autotune": tune_times, "kernel": kname}}
    if prepare:
        return plan
    timer = runtime.EventTimer(ctx, stream) if collect else None
    setup_s = time.time() - setup_start
    if timer:
        timer.record_start()
    for name, grid, block in launches:
        ctx.launch(module, name, grid, block, a, stream)
    if timer:
        timer.record_stop()
    flags = ctx.read_status(status, stream)
    gpu_s = timer.elapsed_ms() / 1000.0 if timer else 0.0
    if timer:
        timer.close()
    out = collect_outputs(plan, flags)
    out.update(gpu_s=gpu_s, setup_s=setup_s)
    return out


def collect_outputs(plan, flags):
    """Results of a launched plan in the reference's shapes (device tensors) and its info."""
    b = plan["buffers"]
    nx, ny, size, num_trace = plan["nx"], plan["ny"], plan["size"], plan["num_trace"]
    npts = nx * ny
    if plan["mode"] in ("final_rho", "mean_state"):
        out = b["result"][:, :npts].t().reshape(nx, ny, size).contiguous()
    else:
        out = b["result"][:npts].reshape(nx, ny).clone()
    tr = None
    if b["trace"] is not None:
        if plan["state_trace"]:
            tr = b["trace"].permute(2, 0, 1).reshape(nx, ny, num_trace, size).contiguous()
        else:
            tr = b["trace"].t().reshape(nx, ny, num_trace).contiguous()
    steps = b["steps"].reshape(nx, ny, 2).cpu().numpy() if b["steps"] is not None else None
    return {"result": out, "trace": tr, "status": flags, "adaptive_steps": steps,
            "gpu_s": 0.0, "setup_s": 0.0, **plan["info"]}


def _timed_launches(ctx, module, kname, grid, block, args, stream, cursor, repeats=2):
    """Shortest of ``repeats`` timed launches after one untimed (first use of a module)."""
    timer = runtime.EventTimer(ctx, stream)
    best = float("inf")
    try:
        for rep in range(repeats + 1):
            cursor.zero_()
            timer.record_start()
            ctx.launch(module, kname, grid, block, args, stream)
            timer.record_stop()
            ms = timer.elapsed_ms()
            if rep:
                best = min(best, ms)
    finally:
        timer.close()
    return best / 1000.0


def table_kernel(spec):
    return ("gq_rk4_time_table" if spec.template in ("fixed_rk4.hip", "fixed_rk4_mixed.hip")
            else "gq_fixed_time_table")
