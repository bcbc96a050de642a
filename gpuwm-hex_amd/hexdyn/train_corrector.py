"""Train the learned corrector of one workload on the GPU and save it.

    python -m hexdyn.train_corrector --workload x1.40962-moist --pairs 2000 \
        --out gpuwm-hex_amd/hexdyn/data/corrector_x1.40962-moist.npz

Pairs (one coarse step, two fine steps) come from the dycore itself along
seeded trajectories: the workload's initial state with a seeded theta bubble,
a seeded spin-up of 0..spinup fine steps, then consecutive pairs along the
fine trajectory.  The file holds the fp32 master weights, the feature and
output scales, the check envelope of the training states and the training
record.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path


def main(argv=None) -> int:
    p = argparse.ArgumentParser()
    p.add_argument("--workload", default="x1.40962-moist")
    p.add_argument("--pairs", type=int, default=2000)
    p.add_argument("--trajectories", type=int, default=8)
    p.add_argument("--layers", type=int, default=4)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--lr-final", type=float, default=1e-4)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--spinup-min", type=int, default=2)
    p.add_argument("--spinup", type=int, default=8)
    p.add_argument("--epochs", type=int, default=1)
    p.add_argument("--coarse-substeps", type=int, default=3)
    p.add_argument("--coarse-subcycles", type=int, default=3, choices=[1, 3])
    p.add_argument("--mixing", default="2d_smagorinsky", choices=["2d_smagorinsky", "off"])
    p.add_argument("--out", required=True)
    a = p.parse_args(argv)

    import numpy as np
    import torch

    from .case import WORKLOADS, build_case
    from .config import DycoreConfig
    from . import learned

    assert torch.cuda.is_available(), "training runs on the GPU"
    spec = WORKLOADS[a.workload]
    cfg = DycoreConfig(config_dt=spec["dt"], config_horiz_mixing=a.mixing)
    t0 = time.perf_counter()
    case = build_case(a.workload, config=cfg)
    print(f"# case {a.workload} built in {time.perf_counter() - t0:.1f} s", flush=True)
    tc = learned.TrainConfig(pairs=a.pairs, trajectories=a.trajectories, layers=a.layers, lr=a.lr,
                             lr_final=a.lr_final, seed=a.seed, spinup_min=a.spinup_min,
                             spinup_max=a.spinup, epochs=a.epochs,
                             coarse=(a.coarse_substeps, a.coarse_subcycles))
    t0 = time.perf_counter()
    net, hist = learned.train(case, tc, log=lambda m: print(m, flush=True))
    wall = time.perf_counter() - t0
    record = {"workload": a.workload, "pairs": a.pairs, "trajectories": a.trajectories, "layers": a.layers,
              "lr": a.lr, "lr_final": a.lr_final, "epochs": a.epochs, "seed": a.seed, "spinup": [a.spinup_min, a.spinup], "mixing": a.mixing,
              "train_seconds": wall, "loss_first50": float(np.mean(hist[:50])),
              "loss_last50": float(np.mean(hist[-50:])), "envelope": [float(x) for x in net.envelope],
              "coarse": {"dt": cfg.config_dt * 2, "number_of_sub_steps": a.coarse_substeps,
                         "dynamics_split_steps": a.coarse_subcycles}}
    out = Path(a.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    learned.save(net, out, history=np.asarray(hist, np.float32), record=json.dumps(record))
    print(json.dumps(record), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
