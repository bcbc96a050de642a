"""RK schedule against the native-compiled MPAS schedule bits
(verification/step-anchors/native-schedule*.json, checked by the
reference's tests/test_rk2_v841.py:90)."""
import json
from pathlib import Path

import numpy as np
import pytest

from hexdyn.setup import rk_schedule
from oracle.oracle import schedule_f32

GOLDEN = json.loads((Path(__file__).parent / "golden" / "native_schedule.json").read_text())


@pytest.mark.parametrize("row", GOLDEN["rows"], ids=lambda r: f"dt{r['dt']}-n{r['substeps']}-o{r['order']}")
def test_python_schedule_bits(row):
    sched = rk_schedule(row["dt"], order=row["order"], acoustic_substeps=row["substeps"],
                        dynamics_splits=GOLDEN["dynamics_split_steps"])
    large = np.array([s.large_timestep for s in sched], dtype=np.float32).view(np.uint32).tolist()
    acoustic = np.array([s.acoustic_timestep for s in sched], dtype=np.float32).view(np.uint32).tolist()
    assert large == row["large_bits"]
    assert acoustic == row["acoustic_bits"]
    assert [s.acoustic_steps for s in sched] == row["counts"]


@pytest.mark.parametrize("row", GOLDEN["rows"], ids=lambda r: f"dt{r['dt']}-n{r['substeps']}-o{r['order']}")
def test_oracle_schedule_bits(row):
    large, acoustic, steps = schedule_f32(row["dt"], row["order"], row["substeps"], GOLDEN["dynamics_split_steps"])
    assert large.view(np.uint32).tolist() == row["large_bits"]
    assert acoustic.view(np.uint32).tolist() == row["acoustic_bits"]
    assert steps.tolist() == row["counts"]


def test_schedule_mutation_is_caught():
    # a plausible mistake (acoustic dt of stage 1 = dt_dyn/n instead of dt_dyn/3) must fail the bits
    row = [r for r in GOLDEN["rows"] if r["order"] == 3][0]
    dt_dyn = row["dt"] / 3
    wrong = np.float32(dt_dyn / row["substeps"]).view(np.uint32)
    assert int(wrong) != row["acoustic_bits"][0]
