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
