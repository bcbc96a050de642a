"""TEST INFRASTRUCTURE — surface pressure of the FP64 restatement after a
forecast, the judge of the plain and learned-corrected GPU runs' 24 h error.

Only bench.py's CPU leg and tests/ read this module or the files it writes;
the product never does.  The FP64 run of the x1.40962 headline case takes
several minutes of host time, so it is made once per case and stored as
``oracle/data/ps_<workload>_<hours>h_fp64.npz`` with a digest of the case's
initial state and configuration; a reader refuses a file whose digest does
not match the case it is asked about.

    python oracle/surface_pressure.py x1.40962-moist 24

Surface pressure is the hydrostatic weight of the column,
ps = g * sum_k rho_zz(k) * (1 + sum_s q_s(k)) / rdzw(k)
(the dry air plus the water species the mass loading carries).
"""

from __future__ import annotations

import hashlib
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
DATA = HERE / "data"
GRAVITY = 9.80616


def surface_pressure(case, state) -> np.ndarray:
    """Column weight (Pa) of a logical-layout state (float64)."""
    rdzw = np.asarray(case.grid.rdzw, dtype=np.float64)
    rho = np.asarray(state["rho_zz"], dtype=np.float64)
    load = np.ones_like(rho)
    if case.n_moist > 0:
        load = load + np.asarray(state["scalars"][: case.n_moist], dtype=np.float64).sum(axis=0)
    return GRAVITY * (rho * load / rdzw[:, None]).sum(axis=0)


def case_digest(case) -> str:
    from hexdyn.case import initial_state_arrays

    h = hashlib.sha256()
    for k, a in sorted(initial_state_arrays(case).items()):
        h.update(k.encode())
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(repr((case.config.config_dt, case.config.config_number_of_sub_steps,
                   case.config.config_horiz_mixing)).encode())
    h.update(np.ascontiguousarray(case.grid.zz).tobytes())
    return h.hexdigest()[:24]


def path_for(workload: str, hours: float) -> Path:
    return DATA / f"ps_{workload}_{hours:g}h_fp64.npz"


def load(case, workload: str, hours: float):
    """The stored FP64 surface pressure after `hours`, or None (missing or
    made from another case)."""
    p = path_for(workload, hours)
    if not p.exists():
        return None
    z = np.load(p, allow_pickle=False)
    if str(z["digest"]) != case_digest(case):
        return None
    return np.asarray(z["ps"], dtype=np.float64)


def make(workload: str, hours: float, mixing: str = "2d_smagorinsky") -> Path:
    # run as a script, this folder is first on the path and would shadow the package
    sys.path[:] = [p for p in sys.path if Path(p or ".").resolve() != HERE]
    sys.path.insert(0, str(ROOT / "gpuwm-hex_amd"))
    sys.path.insert(0, str(ROOT))
    from hexdyn.case import WORKLOADS, build_case, initial_state_arrays
    from hexdyn.config import DycoreConfig
    from oracle.oracle import Oracle

    dt = WORKLOADS[workload]["dt"]
    case = build_case(workload, config=DycoreConfig(config_dt=dt, config_horiz_mixing=mixing))
    n = int(round(hours * 3600.0 / dt))
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    t0 = time.perf_counter()
    for i in range(n):
        rc = o.step(1)
        if rc != 0:
            raise RuntimeError(f"FP64 restatement flagged step {i + 1}")
        if (i + 1) % 10 == 0:
            print(f"# {workload} FP64 step {i + 1}/{n} ({time.perf_counter() - t0:.0f} s)", flush=True)
    ps = surface_pressure(case, o.get_state())
    DATA.mkdir(exist_ok=True)
    out = path_for(workload, hours)
    np.savez_compressed(out, ps=ps, digest=np.array(case_digest(case)), steps=np.array(n), dt=np.array(dt),
                        mixing=np.array(mixing))
    return out


if __name__ == "__main__":
    wl = sys.argv[1] if len(sys.argv) > 1 else "x1.40962-moist"
    hrs = float(sys.argv[2]) if len(sys.argv) > 2 else 24.0
    print(make(wl, hrs))
