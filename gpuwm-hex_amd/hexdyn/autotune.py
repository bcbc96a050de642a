"""Launch-shape autotuner of the gather kernels.

Each gather kernel runs one wavefront per cell, edge or vertex row (level k
on lane k, the whole column of a row in one wavefront, so the level block is
the row).  What is left to tune is the tile: wavefronts (rows) per block,
1, 2, 4, 8 or 16.  A block's rows are consecutive along the space-filling curve and
share the block's XCD, so the tile sets how much of a neighbourhood one
scheduling unit covers (the rows of one block share the CU's L1, and
consecutive rows on the curve share most of their neighbour rows), and how
many blocks a CU holds.  The tuner times
every gather kernel at each tile with HIP events over profiled large steps
of the handle's own mesh, keeps each kernel's fastest tile, checks that the
whole step with the chosen tiles is not slower than with the default (4),
and stores the choice in a JSON file keyed by device, mesh size, levels,
scalars, mixing and numerics; ``apply_saved`` installs it after ``hexdyn_create``.
(The reference tunes the level chunking of its hoisted edge kernels by
longest-processing-time packing, cuda_edge_hoist.py:63-101.)
"""

from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Dict, Optional

from .lib import lib

CANDIDATES = (1, 2, 4, 8, 16)
COLUMN_KERNELS = {"k_vcoef", "k_ac_cell", "k_diag_edge_flux_t"}  # fixed block shapes


def _cache_dir() -> Path:
    root = os.environ.get("HEXDYN_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "hexdyn")
    p = Path(root)
    p.mkdir(parents=True, exist_ok=True)
    return p


def tuning_key(drv) -> str:
    """Mesh and configuration of a stored table (the cache directory belongs
    to one machine; the device name is left out because a profiler attached
    to the process can hide it)."""
    c = drv.case
    return (f"C{c.mesh.n_cells}-E{c.mesh.n_edges}-L{c.n_levels}-T{c.n_scalars}-"
            f"{c.config.config_horiz_mixing}-{c.config.numerics}")


def set_launch(drv, kernel: str, waves_per_block: int) -> None:
    rc = lib().hexdyn_set_launch(drv.handle, kernel.encode(), int(waves_per_block))
    if rc != 0:
        raise ValueError(f"hexdyn_set_launch({kernel}, {waves_per_block}) refused ({rc})")


def get_launch(drv, kernel: str) -> int:
    return int(lib().hexdyn_get_launch(drv.handle, kernel.encode()))


def apply(drv, table: Dict[str, int]) -> None:
    set_launch(drv, "*", 4)
    for k, w in table.items():
        set_launch(drv, k, w)


def _kernel_times(drv, steps: int) -> Dict[str, float]:
    import torch

    drv.profile(True)
    drv.step_device(steps)
    torch.cuda.synchronize()
    out = {}
    for name in drv.profile_kernels():
        avg, n = drv.profile_read(name)
        out[name] = avg * n / steps
    drv.profile(False)
    return out


def _step_ms(drv, steps: int) -> float:
    import torch

    drv.step_device(1)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    drv.step_device(steps)
    torch.cuda.synchronize()
    return 1000.0 * (time.perf_counter() - w0) / steps


def tune(drv, steps: int = 2, check_steps: int = 10) -> dict:
    """Pick each gather kernel's wavefronts per block by timing; install and
    return {"table", "per_kernel_ms", "default_ms_per_step", "tuned_ms_per_step"}."""
    times = {}
    for w in CANDIDATES:
        set_launch(drv, "*", w)
        _kernel_times(drv, 1)  # warm (graphs are off while profiling; first launches compile nothing)
        times[w] = _kernel_times(drv, steps)
    names = sorted(set().union(*[set(t) for t in times.values()]) - COLUMN_KERNELS)
    table = {}
    per_kernel = {}
    for n in names:
        best = min(CANDIDATES, key=lambda w: times[w].get(n, float("inf")))
        per_kernel[n] = {str(w): round(times[w].get(n, float("nan")), 5) for w in CANDIDATES}
        if best != 4:
            table[n] = best
    apply(drv, {})
    default_ms = _step_ms(drv, check_steps)
    apply(drv, table)
    tuned_ms = _step_ms(drv, check_steps)
    if tuned_ms > default_ms:
        # the whole step decides: keep the default when the per-kernel picks lose
        table = {}
        apply(drv, table)
    return {"table": table, "per_kernel_ms": per_kernel, "default_ms_per_step": round(default_ms, 4),
            "tuned_ms_per_step": round(tuned_ms, 4), "candidates": list(CANDIDATES), "key": tuning_key(drv)}


def save(record: dict, path: Optional[Path] = None) -> Path:
    path = path or _cache_dir() / f"tune-{record['key']}.json"
    path.write_text(json.dumps(record, indent=1) + "\n")
    return path


def apply_saved(drv) -> Optional[dict]:
    """Install the stored tuning of this device and mesh, if any."""
    p = _cache_dir() / f"tune-{tuning_key(drv)}.json"
    if not p.exists():
        return None
    rec = json.loads(p.read_text())
    apply(drv, rec.get("table", {}))
    return rec
