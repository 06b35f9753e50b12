"""On-disk code-object cache: a new process loads the kernel compiled by an earlier one."""
import os
import subprocess
import sys
import uuid

import numpy as np
import pytest

from gqis import mesolve_2D, workloads
from gqis.runtime import Context

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys, numpy as np
sys.path[:0] = [{pkg!r}, {root!r}]
from gqis import mesolve_2D, workloads
w = workloads.lzs_two_level(n_eps=8, n_amp=8, n_t=201, t_end=20.0)
res, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"], w["tlist"],
                       var_arrays=w["var_arrays"], fp64=True, hiprtc_options=({opt!r},),
                       return_timing_info=True)
np.save({out!r}, res)
print("cached_kernel", info["cached_kernel"], flush=True)
"""


def test_new_process_loads_code_object_from_disk(gpu, tmp_path):
    opt = f"-DGQ_CACHE_PROBE_{uuid.uuid4().hex}=1"     # a key no earlier run has stored
    w = workloads.lzs_two_level(n_eps=8, n_amp=8, n_t=201, t_end=20.0)
    ctx = Context.get()
    ctx.set_cache_dir(tmp_path)
    try:
        first, info = mesolve_2D(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                                 w["tlist"], var_arrays=w["var_arrays"], fp64=True,
                                 hiprtc_options=(opt,), return_timing_info=True)
    finally:
        ctx.set_cache_dir(None)
    assert info["cached_kernel"] == "miss"
    stored = [f for f in os.listdir(tmp_path) if f.endswith(".hsaco")]
    assert len(stored) >= 1 and all(f.startswith(ctx.arch + "-") for f in stored)
    out = str(tmp_path / "child.npy")
    code = CHILD.format(pkg=os.path.join(ROOT, "gpu-quantum-interferometry-solver_amd"),
                        root=ROOT, opt=opt, out=out)
    env = dict(os.environ, GQIS_KERNEL_CACHE_DIR=str(tmp_path))
    print("child process: same sweep with the on-disk cache", flush=True)
    proc = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                          timeout=240)
    assert proc.returncode == 0, proc.stderr[-2000:]
    assert "cached_kernel disk" in proc.stdout
    np.testing.assert_array_equal(np.load(out), first)


def _probe_source(value):
    from gqis import codegen
    with open(codegen._header_path()) as fh:
        header = fh.read()
    return (header + '\nextern "C" __global__ void gq_cache_probe(const gq_sweep_args a) '
            f"{{ if (threadIdx.x == 0) *a.status = {value}u; }}\n")


def _probe(ctx, module, stream):
    import torch
    from gqis import runtime
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    a = runtime.SweepArgs()
    a.status = status.data_ptr()
    ctx.launch(module, "gq_cache_probe", 1, 64, a, stream)
    return ctx.read_status(status, stream)


def test_module_cache_keeps_held_modules_and_unloads_unheld_ones(gpu):
    """The context keeps at most 128 modules that no caller holds: compiling 140 throw-away
    modules unloads the oldest unheld ones (a new compile of one of them is not a cache hit),
    while a module the caller still holds stays loaded and launchable.  Two threads compiling
    the same new source get the same module."""
    import threading
    import torch
    ctx = Context.get()
    stream = torch.cuda.current_stream().cuda_stream
    tag = uuid.uuid4().hex[:8]
    opts = ["-O3", f"-DGQ_CACHE_TAG_{tag}=1"]
    held, cached, _ = ctx.compile(_probe_source(7), opts)
    assert cached == 0
    first = None
    for i in range(140):
        m, c, _ = ctx.compile(_probe_source(1000 + i), opts)
        assert c == 0
        if i == 0:
            first = m._as_parameter_.value
        del m
    assert _probe(ctx, held, stream) == 7
    again, c, _ = ctx.compile(_probe_source(7), opts)
    assert c == 1 and again._as_parameter_.value == held._as_parameter_.value
    m, c, _ = ctx.compile(_probe_source(1000), opts)
    assert c == 0                     # it was unloaded and is compiled anew
    assert _probe(ctx, m, stream) == 1000
    del first

    results = []

    def worker():
        mod, _, _ = ctx.compile(_probe_source(4242), opts)
        results.append(mod)
    threads = [threading.Thread(target=worker) for _ in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert len({r._as_parameter_.value for r in results}) == 1
    assert _probe(ctx, results[0], stream) == 4242
