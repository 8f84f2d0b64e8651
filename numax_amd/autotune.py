"""Launch-shape autotuning for the device paths, cached per device.

The kernels take their shapes from two knobs: the resident workgroups per CU
of each batched size class (``nmx_batch_cfg.teams_per_cu``) and the panel
width of the mixed-precision factorization (``nmx_ir_cfg.block``).  The
sweeps themselves are native (``nmx_batch_autotune``, ``nmx_ir_autotune``:
synthetic inputs, event-timed runs); this module runs them and keeps the
winners in a JSON file keyed by device name and library version, which
:func:`numax_amd.linalg.batched_solve` and :func:`numax_amd.linalg.solve_mpir`
consult when the caller does not pass the knob.  The file is
``$NUMAX_AMD_TUNE_CACHE`` if set, else ``~/.cache/numax_amd/tune.json``.
"""
from __future__ import annotations

import ctypes
import json
import os
import threading

from numax_amd import _lib

_lock = threading.Lock()


def cache_path() -> str:
    return os.environ.get("NUMAX_AMD_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "numax_amd",
                                                                    "tune.json")


def _device_key(device: int) -> str:
    import torch
    return f"{torch.cuda.get_device_name(device)}|lib{_lib.load().nmx_version()}"


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, value) -> None:
    with _lock:
        data = _load()
        data[key] = value
        path = cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def cached(kind: str, device: int = 0, **shape):
    """The stored choice for ``kind`` ("batched" or "mpir") and shape, or None."""
    key = "|".join([_device_key(device), kind] + [f"{k}={v}" for k, v in sorted(shape.items())])
    return _load().get(key)


def _context(device: int):
    import torch
    torch.zeros(1, device=f"cuda:{device}")  # the framework's runtime up first, as in every other entry
    return _lib.context_for(device, torch.cuda.current_stream(device).cuda_stream)


def tune_batched(device: int = 0, samples: int = 2000, store: bool = True) -> list:
    """Sweeps resident workgroups per CU for each batched size class (n > 128,
    64 < n <= 128, n <= 64); returns ``teams_per_cu`` (three ints)."""
    lib = _lib.load()
    cfg = _lib.BatchCfg()
    lib.nmx_batch_cfg_default(ctypes.byref(cfg))
    _lib.check(lib.nmx_batch_autotune(_context(device).handle, int(samples), ctypes.byref(cfg)), "batch autotune")
    teams = [int(cfg.teams_per_cu[i]) for i in range(3)]
    if store:
        key = "|".join([_device_key(device), "batched"])
        _store(key, teams)
    return teams


def tune_batched_on(run, candidates=(1, 2, 3, 4, 6, 8), reps: int = 2) -> list:
    """Tunes ``teams_per_cu`` on the caller's own batch: ``run(cfg)`` solves
    it once with the given ``nmx_batch_cfg`` (and may be called ~50 times).

    The size classes' launches share the chip (the Cholesky classes run on
    a second stream beside the LU classes), so the best value for one class
    depends on the others and on the batch's mix; the synthetic per-class
    sweep of :func:`tune_batched` cannot see that.  This is one pass of
    coordinate search from the library's defaults: each class in turn takes
    the candidate with the shortest wall time (best of ``reps`` runs after a
    warm-up, the device synchronized around each) with the others held."""
    import time

    import torch
    lib = _lib.load()
    cfg = _lib.BatchCfg()
    lib.nmx_batch_cfg_default(ctypes.byref(cfg))

    def timed(c) -> float:
        run(c)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            run(c)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    cur = timed(cfg)
    for i in range(3):
        start = keep = int(cfg.teams_per_cu[i])
        for v in candidates:
            if v == start:  # already timed
                continue
            cfg.teams_per_cu[i] = v
            t = timed(cfg)
            if t < cur:
                cur, keep = t, v
        cfg.teams_per_cu[i] = keep
    return [int(cfg.teams_per_cu[i]) for i in range(3)]


def tune_mpir(n: int, device: int = 0, precision: str = "fp16s", store: bool = True) -> int:
    """Sweeps the mixed-precision panel width (128, 256, 512) at order ``n``;
    returns the fastest."""
    from numax_amd.linalg import PRECISIONS
    lib = _lib.load()
    cfg = _lib.IrCfg()
    lib.nmx_ir_cfg_default(ctypes.byref(cfg))
    cfg.precision = PRECISIONS[precision]
    _lib.check(lib.nmx_ir_autotune(_context(device).handle, int(n), ctypes.byref(cfg)), "mpir autotune")
    block = int(cfg.block)
    if store:
        key = "|".join([_device_key(device), "mpir", f"n={int(n)}", f"precision={precision}"])
        _store(key, block)
    return block


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="tune the device paths' launch shapes and cache them")
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--mpir-n", type=int, nargs="*", default=[4096, 16384])
    a = ap.parse_args()
    print("batched teams_per_cu:", tune_batched(samples=a.samples))
    for nn in a.mpir_n:
        print(f"mpir block at n={nn}:", tune_mpir(nn))
    print("cache:", cache_path())
