"""Work-group size search (gqis/autotune.py) on the CPU: the choice, its memory and disk
records, skipped candidates, and the Hilbert kernel's LDS formula against hiprtc's own count."""
import json

import pytest

from gqis import autotune, codegen, hilbert, runtime, solver, workloads

SRC = "#define GQ_BLOCK 256\nkernel body\n"


class FakeCtx:
    arch = "gfx950"

    def __init__(self, cache_dir=None, bad=()):
        self.cache_dir, self.bad, self.compiled = cache_dir, set(bad), []

    def compile(self, source, options):
        block = int(source.split("#define GQ_BLOCK ")[1].split()[0])
        if block in self.bad:
            raise runtime.GqError("local memory exceeds limit")
        self.compiled.append(block)
        return block, 0, 0.0


@pytest.fixture(autouse=True)
def _fresh():
    autotune.clear_memory()
    yield
    autotune.clear_memory()


def test_with_block_and_key():
    assert autotune.with_block(SRC, 64).startswith("#define GQ_BLOCK 64\n")
    assert autotune.source_key(SRC, ("-O3",), "gfx950") == autotune.source_key(
        autotune.with_block(SRC, 512), ("-O3",), "gfx950")
    assert autotune.source_key(SRC, ("-O3",), "gfx950") != autotune.source_key(
        SRC, ("-O2",), "gfx950")
    with pytest.raises(ValueError):
        autotune.with_block("no define", 64)


def test_choose_measures_once_and_remembers(tmp_path):
    times = {64: 3.0, 128: 1.0, 256: 2.0, 512: 5.0}
    ctx = FakeCtx(cache_dir=str(tmp_path))
    calls = []

    def trial(module, block):
        calls.append(block)
        return times[block]
    best, measured = autotune.choose(ctx, SRC, ("-O3",), trial)
    assert best == 128 and measured == times and calls == [64, 128, 256, 512]
    best2, measured2 = autotune.choose(ctx, SRC, ("-O3",), trial)
    assert best2 == 128 and measured2 == {} and len(calls) == 4
    rec = json.load(open(tmp_path / "autotune.json"))
    assert list(rec.values())[0]["choice"] == 128
    autotune.clear_memory()                       # a new process reads the disk record
    assert autotune.lookup(FakeCtx(cache_dir=str(tmp_path)),
                           autotune.source_key(SRC, ("-O3",), "gfx950")) == 128
    assert autotune.lookup(FakeCtx(), autotune.source_key(SRC, ("-O3",), "gfx950")) == 128


def test_variant_candidates_with_their_own_sources(tmp_path):
    ctx = FakeCtx(cache_dir=str(tmp_path))
    cands = [(1, 64), (1, 256), (4, 256)]
    t = {(1, 64): 2.0, (1, 256): 3.0, (4, 256): 1.0}
    best, measured = autotune.choose(ctx, SRC, (), lambda m, c: t[c], cands,
                                     source_of=lambda c: autotune.with_block(SRC, c[1]))
    assert best == (4, 256) and measured == t
    autotune.clear_memory()
    key = autotune.source_key(SRC, (), "gfx950")
    assert autotune.lookup(ctx, key, cands) == (4, 256)      # tuples survive the JSON record
    assert autotune.lookup(ctx, key, [(1, 64)]) is None        # not a candidate any more


def test_candidates_that_do_not_compile_are_skipped():
    ctx = FakeCtx(bad={256, 512})
    best, measured = autotune.choose(ctx, SRC, (), lambda m, b: {64: 2.0, 128: 1.5}[b])
    assert best == 128 and set(measured) == {64, 128}


def test_hilbert_lds_formula_matches_the_compiler():
    w = workloads.random_n_level(32, n_eps=4, n_amp=4)
    model = solver._prepare_model(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                                  sweep_symbols=tuple(w["var_arrays"]), const_values=None,
                                  runtime_consts=None, keep_symbolic_consts=None,
                                  auto_runtime_consts=False, rho0=None, ode_size=None)
    hm = hilbert.extract(model, solver._drive_map(model["Drive"]))
    _, offs = hilbert.constants(hm, [])
    # FP64 VALU form at 256 threads and FP64 matrix-core form at 512 threads: both over the
    # limit, and the compiler's figure is the formula's
    for block, valu in ((256, True), (512, False)):
        src = hilbert.emit(hm, offs, "mean", block, True, valu=valu).replace(
            "GQ_RHO0_FROM_ARGS", "0")
        est = hilbert.lds_bytes(32, True, block, hilbert.jump_width(hm), valu=valu,
                                nbasis=len(hm.basis))
        assert est > hilbert.LDS_BYTES_PER_CU
        with pytest.raises(runtime.GqError, match=rf"local memory \({est}\) exceeds limit"):
            runtime.compile_check(src, codegen.HIPRTC_BASE_OPTIONS)


def test_hilbert_lds_formula_with_general_jump_operators():
    # the 5-qubit chain (lowering operators with one nonzero per row) adds the per-wavefront
    # state copy and row tables: at 256 threads the FP64 kernel exceeds the limit by the
    # formula's count, and block_for halves the work-group
    w = workloads.multi_qubit(5, n_eps=2, n_amp=2)
    model = solver._prepare_model(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                                  sweep_symbols=tuple(w["var_arrays"]), const_values=None,
                                  runtime_consts=None, keep_symbolic_consts=None,
                                  auto_runtime_consts=False, rho0=None, ode_size=None)
    hm = hilbert.extract(model, solver._drive_map(model["Drive"]))
    assert hilbert.n_mono(hm) == 5
    _, offs = hilbert.constants(hm, [])
    src = hilbert.emit(hm, offs, "mean", 256, True).replace("GQ_RHO0_FROM_ARGS", "0")
    est = hilbert.lds_bytes(32, True, 256, hilbert.jump_width(hm), mono=5, nbasis=len(hm.basis))
    assert est > hilbert.LDS_BYTES_PER_CU
    with pytest.raises(runtime.GqError, match=rf"local memory \({est}\) exceeds limit"):
        runtime.compile_check(src, codegen.HIPRTC_BASE_OPTIONS)
    assert hilbert.block_for(hm, True, 256) == 128
