"""A sweep kept resident on the device for repeated runs, replayed as one HIP graph (per-point
kernels and the matrix-core Hilbert-space kernels alike).

Fitting loops call the same sweep many times with new sweep values, runtime constants or
initial states (the reference re-enters ``mesolve_2D`` each time, ``gqis/solver.py:1712-2391``:
cache lookups, uploads, allocation, one launch, finite check).  ``prepare_sweep`` does all of
that once: buffers for the axes, constants, initial states, time table, result and status stay
allocated, and the launch chain (status clear, time table, sweep kernel) is recorded as a HIP
graph with its arguments fixed.  ``run`` copies the new inputs into those buffers, replays the
graph with one call and returns the result in ``mesolve_2D``'s shapes and dtypes.  The status
word is cleared by a kernel at the head of the chain (a replayed graph holds no memset).
"""
import time

import numpy as np

from . import codegen, runtime
from .device import _axis, _initial_states, collect_outputs


class PreparedSweep:
    """Returned by ``prepare_sweep`` (same arguments as ``mesolve_2D``)."""

    def __init__(self, plan, *, compiled, spec, const_values, runtime_consts, use_defaults,
                 sweep_symbols, uses_rho0_values, finish):
        import torch
        self.torch = torch
        self.plan = plan
        self.ctx = plan["ctx"]
        self.compiled, self.spec = compiled, spec
        self.const_values = dict(const_values or {})
        self.runtime_consts = dict(runtime_consts or {})
        self.use_defaults = use_defaults
        self.sweep_symbols = tuple(sweep_symbols)
        self.uses_rho0_values = uses_rho0_values
        self._finish = finish
        self.graph = None
        self.last_gpu_s = 0.0
        launches = list(plan["launches"])
        if launches:
            clear = codegen.kernel_name("gq_clear_flags", plan["fp64"])
            self.graph = self.ctx.graph_create(plan["module"], [(clear, 1, 64)] + launches,
                                               plan["args"])

    @property
    def shape(self):
        return self.plan["nx"], self.plan["ny"]

    def _set_axes(self, var_arrays):
        b = self.plan["buffers"]
        if tuple(var_arrays) != self.sweep_symbols[:len(var_arrays)]:
            raise ValueError(f"var_arrays must sweep {self.sweep_symbols}, in the order given "
                             "when the sweep was prepared.")
        for key, (sym, vals) in zip(("px", "py"), var_arrays.items()):
            new = _axis(vals, b[key].dtype, b[key].device)
            if new.numel() != b[key].numel():
                raise ValueError(f"Sweep axis {sym} has {new.numel()} values; the prepared sweep "
                                 f"has {b[key].numel()}.")
            b[key].copy_(new)

    def _set_consts(self, runtime_consts):
        merged = dict(self.runtime_consts)
        merged.update(runtime_consts)
        if "set_consts" in self.plan:
            # matrix-core plan: its runtime constants lead a buffer of packed operators
            syms = self.plan["runtime_syms"]
            if not syms:
                raise ValueError("This model has no runtime constants.")
            values = dict(self.plan["runtime_defaults"])
            values.update(merged)
            self.plan["set_consts"]([float(values[s]) for s in syms])
            self.runtime_consts = merged
            return
        from .solver import _runtime_const_values
        vals = _runtime_const_values(self.compiled, merged, self.const_values, self.use_defaults)
        if not vals:
            raise ValueError("This model has no runtime constants.")
        b = self.plan["buffers"]["consts"]
        b.copy_(self.torch.as_tensor(np.asarray(vals, np.float64), dtype=b.dtype,
                                     device=b.device))
        self.runtime_consts = merged

    def run(self, var_arrays=None, runtime_consts=None, rho0_values=None, *,
            return_device=False, ignore_non_finite_output=False, timed=False):
        """Replay the sweep with new sweep values (same lengths), runtime constants and / or
        initial states; inputs not given keep their last values.  Returns what ``mesolve_2D``
        returns for the prepared arguments (without the timing dictionary)."""
        torch = self.torch
        plan = self.plan
        if var_arrays:
            self._set_axes(dict(var_arrays))
        if runtime_consts:
            self._set_consts(runtime_consts)
        if rho0_values is not None:
            if not self.uses_rho0_values:
                raise ValueError("The sweep was prepared without rho0_values.")
            b = plan["buffers"]["rho0"]
            b.copy_(_initial_states(rho0_values, plan["nx"], plan["ny"], plan["size"], b.dtype,
                                    b.device))
        if plan["adaptive"] and plan["buffers"]["trace"] is not None:
            plan["buffers"]["trace"].zero_()     # a failing point stops writing its trace
        stream = torch.cuda.current_stream(plan["buffers"]["status"].device).cuda_stream
        timer = runtime.EventTimer(self.ctx, stream) if timed else None
        if timer:
            timer.record_start()
        if self.graph is not None:
            self.ctx.graph_launch(self.graph, stream)
        if timer:
            timer.record_stop()
        flags = self.ctx.read_status(plan["buffers"]["status"], stream)
        if timer:
            self.last_gpu_s = timer.elapsed_ms() / 1000.0
            timer.close()
        outputs = collect_outputs(plan, flags)
        return self._finish(outputs, return_device=return_device,
                            ignore_non_finite_output=ignore_non_finite_output,
                            total_start=time.time())

    def close(self):
        if self.graph is not None:
            self.ctx.graph_destroy(self.graph)
            self.graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
