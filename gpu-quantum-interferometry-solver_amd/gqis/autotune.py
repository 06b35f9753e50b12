"""Work-group size of a sweep kernel chosen by measurement (``block_size="auto"``).

The per-point kernels keep one parameter point per lane, so the work-group size sets the
register budget the compiler gets (``__launch_bounds__``), how many work-groups stay resident,
and, for the persistent adaptive kernels, how many lanes refill together.  Measured on MI355X
(gfx950): the four-level RK4 kernels (15 states) change by less than 1% between 64 and 512
threads, while an 8-level chain (63 states) runs its Tsit5A FP64 kernel 26% faster at 64 threads
than at 256 and its RK4 FP64 kernel 9x slower at 512 (register spills).  So the size is measured
once per kernel source and device: each candidate is compiled (hiprtc, cached like any module)
and timed with events over a short prefix of the real launch (a few steps of the same points),
the fastest is kept in memory and, when an on-disk kernel cache directory is set, in
``autotune.json`` there.  The reference fixes its block at 16 x 8 threads
(``gqis/solver.py:2322-2340``); there is no counterpart to this search there.
"""
import hashlib
import json
import os
import re
import threading

CANDIDATES = (64, 128, 256, 512)
MIN_TRIAL_POINT_STEPS = 1 << 20     # below this the launch latency dominates the measurement
_BLOCK_LINE = re.compile(r"^#define GQ_BLOCK \d+$", re.M)
_memory = {}
_lock = threading.Lock()


def with_block(source, block):
    """The kernel source with its work-group size set to ``block``."""
    out, n = _BLOCK_LINE.subn(f"#define GQ_BLOCK {int(block)}", source)
    if n != 1:
        raise ValueError("kernel source has no single GQ_BLOCK definition")
    return out


def source_key(source, options, arch):
    text = _BLOCK_LINE.sub("#define GQ_BLOCK ?", source)
    h = hashlib.sha1()
    for part in (arch, "\0".join(options), text):
        h.update(part.encode())
        h.update(b"\0")
    return h.hexdigest()


def _disk_path(ctx):
    return os.path.join(ctx.cache_dir, "autotune.json") if getattr(ctx, "cache_dir", None) else None


def _disk_load(ctx):
    path = _disk_path(ctx)
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _disk_store(ctx, key, record):
    path = _disk_path(ctx)
    if not path:
        return
    data = _disk_load(ctx)
    data[key] = record
    tmp = f"{path}.{os.getpid()}.tmp"
    try:
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        if os.path.exists(tmp):
            os.remove(tmp)


def _as_choice(value):
    return tuple(value) if isinstance(value, list) else value


def lookup(ctx, key, candidates=CANDIDATES):
    """The recorded choice for ``key`` if it is one of ``candidates``, else None."""
    with _lock:
        if key in _memory and _memory[key] in candidates:
            return _memory[key]
    rec = _disk_load(ctx).get(key)
    if isinstance(rec, dict) and _as_choice(rec.get("choice")) in candidates:
        choice = _as_choice(rec["choice"])
        with _lock:
            _memory[key] = choice
        return choice
    return None


def choose(ctx, source, options, trial, candidates=CANDIDATES, source_of=None):
    """Fastest candidate for the kernel of ``source``.  A candidate is a work-group size (the
    source's GQ_BLOCK is replaced) or, with ``source_of(candidate)`` giving its source, any
    variant such as (lanes per point, work-group size).  ``trial(module, candidate)`` runs one
    measured launch and returns its duration in seconds.  Returns ``(choice, {candidate:
    seconds})``; the timings are empty when the choice came from memory or disk."""
    candidates = tuple(candidates)
    key = source_key(source, options, ctx.arch)
    hit = lookup(ctx, key, candidates)
    if hit is not None:
        return hit, {}
    times = {}
    for cand in candidates:
        src = source_of(cand) if source_of is not None else with_block(source, cand)
        try:
            module, _, _ = ctx.compile(src, options)
        except Exception:       # e.g. the LDS of a larger work-group does not fit: skip it
            if cand == candidates[-1] and not times:
                raise
            continue
        times[cand] = float(trial(module, cand))
    best = min(times, key=times.get)
    with _lock:
        _memory[key] = best
    _disk_store(ctx, key, {"choice": list(best) if isinstance(best, tuple) else best,
                           "seconds": {str(c): t for c, t in times.items()}})
    return best, times


def clear_memory():
    with _lock:
        _memory.clear()
