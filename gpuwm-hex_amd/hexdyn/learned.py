"""Learned large-step corrector over the hexdyn C ABI.

A coarse large step at twice the workload's dt, with fewer acoustic substeps
than the two fine steps it replaces, followed by the graph network's
correction stands in for two fine large steps (the workload's dt with its 2
acoustic substeps and 3 dynamics subcycles: 24 acoustic substeps).  Two
coarse steps are trained and measured:

* 3 acoustic substeps, 3 dynamics subcycles (15 substeps per coarse step):
  the more accurate one, and this module's default;
* 8 acoustic substeps, 1 dynamics subcycle (13 substeps, 3 RK stages
  instead of 9): the faster one, used for the headline.

The literal "1 acoustic substep per stage" is unstable at doubled dt: two
coarse steps of the x1.10242 case (dt 2880 s) set the validation flag at
step 2 with 1 substep per stage and at step 4 with 2 (stronger divergence
damping, smdiv 0.2-0.4, makes it fail sooner); 3 substeps with 3 subcycles,
and 8 substeps with 1 subcycle, run.

The network lives in libhexdyn.so (``hexdyn_net_*``: bf16 MFMA forward,
backward and Adam kernels, fp32 master weights); this module holds the host
side: parameter layout and initialisation, feature and target scales,
training on pairs the dycore generates, and the checked step with the plain
dycore as its fallback.

The reference dycore has no learned component.  The corrector wraps the
reference's step boundary (``cuda_driver.py:5751`` step_device): a corrected
step publishes a new state or, when a check trips, the plain steps'
state, and the checks extend the reference's per-step validation flag
(``cuda_dynamics_v841.py:376-410``, read at ``cuda_driver.py:5355-5403``):
the flag itself (non-finite or non-positive state), a horizontal CFL limit,
a vertical CFL limit, a divergence limit and the FP64 dry-mass check.
"""

from __future__ import annotations

import ctypes as ct
import time
from dataclasses import dataclass, replace
from pathlib import Path
from typing import Callable, Dict, List, Optional

import numpy as np

from . import capi
from .config import DycoreConfig
from .lib import lib

HIDDEN = 128
N_IN = 384
N_OUT = 320
ROW = 64
DATA = Path(__file__).resolve().parent / "data"


def _torch():
    import torch

    return torch


def param_layout(layers: int) -> List[tuple]:
    """(name, shape) of every parameter block in the flat vector's order."""
    out = [("W_in", (N_IN, HIDDEN)), ("b_in", (HIDDEN,))]
    for l in range(layers):
        out += [(f"W_{l}", (2 * HIDDEN, HIDDEN)), (f"b_{l}", (HIDDEN,))]
    out += [("W_out", (HIDDEN, N_OUT)), ("b_out", (N_OUT,))]
    return out


def unpack(flat: np.ndarray, layers: int) -> Dict[str, np.ndarray]:
    out, off = {}, 0
    for name, shape in param_layout(layers):
        n = int(np.prod(shape))
        out[name] = flat[off:off + n].reshape(shape)
        off += n
    assert off == flat.size
    return out


def init_params(layers: int, seed: int = 0, out_gain: float = 0.01) -> np.ndarray:
    """He-normal matrices, zero biases; the decoder scaled by out_gain so the
    untrained correction is small."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in param_layout(layers):
        if len(shape) == 1:
            parts.append(np.zeros(shape, dtype=np.float32))
        else:
            std = np.sqrt(2.0 / shape[0]) * (out_gain if name == "W_out" else 1.0)
            parts.append((rng.standard_normal(shape) * std).astype(np.float32))
    return np.concatenate([p.ravel() for p in parts])


COARSE_DEFAULT = (3, 3)   # (acoustic substeps per stage, dynamics subcycles) of the doubled-dt step


def coarse_config(cfg: DycoreConfig, factor: int = 2, sub_steps: int = 3, splits: int = 3) -> DycoreConfig:
    """The coarse step: dt x factor with its own acoustic substep count and
    dynamics subcycles (1 or 3)."""
    return replace(cfg, config_dt=cfg.config_dt * factor, config_number_of_sub_steps=sub_steps,
                   config_dynamics_split_steps=splits)


class NetError(RuntimeError):
    pass


class CorrectorNet:
    """The device network (hexdyn_net_*) bound to one driver's mesh."""

    def __init__(self, drv, layers: int = 4) -> None:
        self.layers = layers
        self.drv = drv
        h = ct.c_void_p()
        rc = lib().hexdyn_net_create(drv.handle, layers, ct.byref(h))
        if rc != 0:
            raise NetError(f"hexdyn_net_create failed ({rc})")
        self._n = h
        self.n_params = int(lib().hexdyn_net_param_count(self._n))
        self.in_mean = np.zeros(N_IN, np.float32)
        self.in_scale = np.ones(N_IN, np.float32)
        self.out_scale = np.ones(N_OUT, np.float32)
        self.envelope = np.zeros(3)  # max check values of the training states at the coarse dt
        self.coarse = COARSE_DEFAULT  # (substeps, subcycles) of the coarse step it was trained after

    def close(self) -> None:
        if getattr(self, "_n", None):
            lib().hexdyn_net_destroy(self._n)
            self._n = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _s():
        return ct.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            buf = ct.create_string_buffer(256)
            lib().hexdyn_net_last_error(self._n, buf, len(buf))
            raise NetError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")

    # parameters
    def set_params(self, flat: np.ndarray) -> None:
        a = np.ascontiguousarray(flat, dtype=np.float32)
        assert a.size == self.n_params
        self._check(lib().hexdyn_net_set_params(self._n, a.ctypes.data_as(ct.POINTER(ct.c_float)), a.size, self._s()),
                    "hexdyn_net_set_params")

    def get_params(self) -> np.ndarray:
        a = np.empty(self.n_params, np.float32)
        self._check(lib().hexdyn_net_get_params(self._n, a.ctypes.data_as(ct.POINTER(ct.c_float)), a.size, self._s()),
                    "hexdyn_net_get_params")
        return a

    def get_grads(self) -> np.ndarray:
        a = np.empty(self.n_params, np.float32)
        self._check(lib().hexdyn_net_get_grads(self._n, a.ctypes.data_as(ct.POINTER(ct.c_float)), a.size, self._s()),
                    "hexdyn_net_get_grads")
        return a

    def set_scales(self, in_mean, in_scale, out_scale) -> None:
        self.in_mean = np.ascontiguousarray(in_mean, np.float32).ravel()
        self.in_scale = np.ascontiguousarray(in_scale, np.float32).ravel()
        self.out_scale = np.ascontiguousarray(out_scale, np.float32).ravel()
        p = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_float))  # noqa: E731
        self._check(lib().hexdyn_net_set_scales(self._n, p(self.in_mean), p(self.in_scale), p(self.out_scale),
                                                self._s()), "hexdyn_net_set_scales")

    # passes
    def forward(self, drv, keep: bool = False) -> None:
        self._check(lib().hexdyn_net_forward(self._n, drv.handle, 1 if keep else 0, self._s()), "hexdyn_net_forward")

    def set_target(self, coarse, fine) -> None:
        self._check(lib().hexdyn_net_set_target(self._n, coarse.handle, fine.handle, self._s()),
                    "hexdyn_net_set_target")

    def loss(self) -> Dict[str, float]:
        out = (ct.c_double * 5)()
        self._check(lib().hexdyn_net_loss(self._n, out, self._s()), "hexdyn_net_loss")
        return {"total": out[0], "u": out[1], "theta_m": out[2], "rho_zz": out[3], "w": out[4]}

    def backward(self) -> None:
        self._check(lib().hexdyn_net_backward(self._n, self._s()), "hexdyn_net_backward")

    def adam(self, lr: float, step: int, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self._check(lib().hexdyn_net_adam(self._n, lr, beta1, beta2, eps, step, self._s()), "hexdyn_net_adam")

    def apply(self, drv, gain: float = 1.0) -> None:
        self._check(lib().hexdyn_net_apply(self._n, drv.handle, gain, self._s()), "hexdyn_net_apply")

    def checks(self, drv) -> tuple:
        """(max horizontal CFL, max vertical CFL, max |divergence| dt, dry mass)"""
        out = (ct.c_float * 3)()
        mass = ct.c_double()
        self._check(lib().hexdyn_net_checks(self._n, drv.handle, out, ct.byref(mass), self._s()),
                    "hexdyn_net_checks")
        return float(out[0]), float(out[1]), float(out[2]), float(mass.value)

    def raw_features(self, drv):
        torch = _torch()
        x = torch.empty((drv.case.mesh.n_cells, N_IN), dtype=torch.float32, device="cuda")
        self._check(lib().hexdyn_net_features(self._n, drv.handle, ct.c_void_p(x.data_ptr()), self._s()),
                    "hexdyn_net_features")
        return x

    def export(self, what: str):
        torch = _torch()
        C, E = self.drv.case.mesh.n_cells, self.drv.case.mesh.n_edges
        code, shape = {"X": (0, (C, N_IN)), "Y": (1, (C, N_OUT)), "target_cells": (2, (C, 3, ROW)),
                       "target_edges": (3, (E, ROW))}[what]
        t = torch.empty(shape, dtype=torch.float32, device="cuda")
        self._check(lib().hexdyn_net_export(self._n, code, ct.c_void_p(t.data_ptr()), self._s()), "hexdyn_net_export")
        return t

    # persistence
    def state_dict(self) -> Dict[str, np.ndarray]:
        return {"params": self.get_params(), "in_mean": self.in_mean, "in_scale": self.in_scale,
                "out_scale": self.out_scale, "layers": np.array(self.layers),
                "envelope": np.asarray(self.envelope, np.float64), "coarse": np.asarray(self.coarse)}

    def load_state(self, d) -> None:
        assert int(d["layers"]) == self.layers
        self.set_params(np.asarray(d["params"], np.float32))
        self.set_scales(d["in_mean"], d["in_scale"], d["out_scale"])
        if "envelope" in d:
            self.envelope = np.asarray(d["envelope"], np.float64)
        if "coarse" in d:
            self.coarse = tuple(int(x) for x in np.asarray(d["coarse"]).ravel())


def copy_state(dst, src) -> None:
    rc = lib().hexdyn_copy_state(dst.handle, src.handle, CorrectorNet._s())
    if rc != 0:
        raise RuntimeError(f"hexdyn_copy_state failed ({rc})")


# --------------------------------------------------------------- scales
class ScaleStats:
    """Pooled statistics of raw features (mean and std per channel-row over
    cells) and raw fine-minus-coarse differences (root mean square per row
    over cells, over edges for u) of several pairs."""

    def __init__(self) -> None:
        self.n = 0
        self.fsum = self.fsq = None
        self.tc_sq = self.te_sq = None
        self.n_te = 0

    def add(self, net: "CorrectorNet", coarse, fine) -> None:
        x = net.raw_features(coarse).double()
        net.set_scales(net.in_mean, net.in_scale, np.ones(N_OUT, np.float32))
        net.set_target(coarse, fine)
        tc = net.export("target_cells").double()   # [C][3][64] raw differences
        te = net.export("target_edges").double()   # [E][64]
        f1, f2 = x.sum(dim=0), (x * x).sum(dim=0)
        c2, e2 = (tc * tc).sum(dim=0), (te * te).sum(dim=0)
        if self.fsum is None:
            self.fsum, self.fsq, self.tc_sq, self.te_sq = f1, f2, c2, e2
        else:
            self.fsum, self.fsq = self.fsum + f1, self.fsq + f2
            self.tc_sq, self.te_sq = self.tc_sq + c2, self.te_sq + e2
        self.n += x.shape[0]
        self.n_te += te.shape[0]

    def scales(self, L: int):
        torch = _torch()
        mean = self.fsum / self.n
        var = (self.fsq / self.n - mean * mean).clamp_min(0.0)
        std = var.sqrt()
        scale = torch.where(std > 0, 1.0 / std.clamp_min(1e-30), torch.zeros_like(std))
        valid = torch.zeros(6, ROW, dtype=torch.bool, device=mean.device)
        valid[:, :L] = True
        valid[2, L] = True          # w has L + 1 rows
        valid[0, L] = True          # Coriolis parameter
        v = valid.ravel()
        mean = torch.where(v, mean, torch.zeros_like(mean)).float().cpu().numpy()
        scale = torch.where(v, scale, torch.zeros_like(scale)).float().cpu().numpy()
        n_c = self.n
        out = np.ones((5, ROW), np.float64)

        def rms(sq, n):
            r = np.sqrt(sq.cpu().numpy() / n)
            return np.where(r > 0, r, 1.0)

        out[0] = out[1] = rms(self.te_sq, self.n_te)
        out[2] = rms(self.tc_sq[0], n_c)
        out[3] = rms(self.tc_sq[1], n_c)
        out[4] = rms(self.tc_sq[2], n_c)
        out[:, L + 1:] = 1.0
        return mean, scale, out.astype(np.float32).ravel()


# --------------------------------------------------------------- training
def perturbed_state(case, seed: int, amplitude: float = 1.0, radius: float = 1.2e6) -> Dict[str, np.ndarray]:
    """The case's initial state with a seeded Gaussian theta bubble (random
    centre, amplitude in K) added to rho*theta at constant density."""
    from .case import initial_state_arrays

    st = {k: np.array(v, copy=True) for k, v in initial_state_arrays(case).items()}
    rng = np.random.default_rng(seed)
    m = case.mesh
    xyz = np.stack([m.arrays["xCell"], m.arrays["yCell"], m.arrays["zCell"]], axis=1).astype(np.float64)
    r = np.linalg.norm(xyz, axis=1)
    p = rng.standard_normal(3)
    p /= np.linalg.norm(p)
    ang = np.arccos(np.clip((xyz @ p) / r, -1.0, 1.0))
    dist = ang * r
    bump = amplitude * rng.uniform(-1.0, 1.0) * np.exp(-(dist / radius) ** 2)
    theta = st["rtheta_m"].astype(np.float64) / st["rho_zz"].astype(np.float64)
    st["rtheta_m"] = (st["rho_zz"] * (theta + bump[None, :])).astype(np.float32)
    return st


@dataclass
class TrainConfig:
    pairs: int = 2000
    trajectories: int = 8
    layers: int = 4
    lr: float = 1e-3
    lr_final: float = 1e-4
    seed: int = 0
    spinup_min: int = 2          # fine steps before a trajectory's first pair: seeded in [min, max]
    spinup_max: int = 8
    stats_pairs: int = 8         # pairs of a separate trajectory that set the scales
    epochs: int = 1              # passes over the same seeded trajectories (the pairs are regenerated)
    coarse: tuple = COARSE_DEFAULT  # (acoustic substeps, dynamics subcycles) of the doubled-dt step


def make_pair_drivers(case, coarse=COARSE_DEFAULT):
    """Fine and coarse drivers of one case (the fine one at the case's dt;
    coarse = (substeps, subcycles) of the doubled-dt step)."""
    from .driver import HexDycoreDriver

    fine = HexDycoreDriver.from_host(case)
    c = HexDycoreDriver.from_host(replace(case, config=coarse_config(case.config, 2, int(coarse[0]),
                                                                     int(coarse[1]))))
    return fine, c


def make_pair(fine, coarse) -> None:
    """From the fine driver's state: one coarse step into the coarse driver
    and two fine steps in place (the fine driver then holds the target)."""
    copy_state(coarse, fine)
    coarse.step_device(1)
    fine.step_device(2)


def fit_scales(net: "CorrectorNet", case, fine, coarse, tc: "TrainConfig") -> None:
    """Scales from `stats_pairs` pairs spread along one seeded trajectory as
    long as a training trajectory."""
    per = max(1, tc.pairs // tc.trajectories)
    stride = max(1, per // max(1, tc.stats_pairs))
    stats = ScaleStats()
    fine.set_state(perturbed_state(case, seed=1000 * (tc.seed + 1) + 999))
    fine.step_device(tc.spinup_min)
    for _ in range(tc.stats_pairs):
        make_pair(fine, coarse)
        stats.add(net, coarse, fine)
        fine.step_device(2 * (stride - 1))
    net.set_scales(*stats.scales(case.n_levels))


def train(case, tc: TrainConfig = TrainConfig(), log: Callable[[str], None] = print,
          net: Optional[CorrectorNet] = None, drivers=None, fit: bool = True):
    """Train the corrector on (coarse step, two fine steps) pairs generated
    by the dycore along seeded trajectories (fit: set the scales first).
    Returns (net, history)."""
    torch = _torch()
    fine, coarse = drivers or make_pair_drivers(case, tc.coarse)
    if net is None:
        net = CorrectorNet(coarse, tc.layers)
        net.set_params(init_params(tc.layers, tc.seed))
    if fit:
        fit_scales(net, case, fine, coarse, tc)
    rng = np.random.default_rng(tc.seed + 17)
    spins = [int(rng.integers(tc.spinup_min, tc.spinup_max + 1)) for _ in range(tc.trajectories)]
    per = max(1, tc.pairs // tc.trajectories)
    total = per * tc.trajectories * tc.epochs
    history = []
    step = 0
    t0 = time.perf_counter()
    ratio = coarse.case.config.config_dt / fine.case.config.config_dt
    envelope = np.zeros(3)
    for epoch in range(tc.epochs):
        for traj in range(tc.trajectories):
            fine.set_state(perturbed_state(case, seed=1000 * (tc.seed + 1) + traj))
            if spins[traj]:
                fine.step_device(spins[traj])
            for _ in range(per):
                make_pair(fine, coarse)
                net.forward(coarse, keep=True)
                net.set_target(coarse, fine)
                lo = net.loss()
                if epoch == 0:
                    envelope = np.maximum(envelope, ratio * np.asarray(net.checks(fine)[:3]))
                net.backward()
                frac = step / max(1, total - 1)
                lr = tc.lr * (tc.lr_final / tc.lr) ** frac
                net.adam(lr, step + 1)
                step += 1
                history.append(lo["total"])
                if step % 50 == 0 or step == 1:
                    log(f"# pair {step}/{total} loss {lo['total']:.4f} (u {lo['u']:.3f} theta {lo['theta_m']:.3f} "
                        f"rho {lo['rho_zz']:.3f} w {lo['w']:.3f}) {time.perf_counter() - t0:.0f} s")
    torch.cuda.synchronize()
    net.envelope = envelope
    net.coarse = tuple(int(x) for x in tc.coarse)
    return net, history


def validation_loss(net: "CorrectorNet", case, fine, coarse, seeds, spins) -> List[float]:
    """Loss of the network on held-out pairs (seeded states, spin-ups)."""
    out = []
    for seed, spin in zip(seeds, spins):
        fine.set_state(perturbed_state(case, seed=seed))
        if spin:
            fine.step_device(spin)
        make_pair(fine, coarse)
        net.forward(coarse)
        net.set_target(coarse, fine)
        out.append(net.loss()["total"])
    return out


# --------------------------------------------------------------- stepping
@dataclass
class Limits:
    """Per-step checks of the corrected state; any trip falls back to the
    plain steps.  Values are at the coarse dt: max |u| dt / dcEdge,
    max |w| dt rdzw / zz, max |divergence| dt, and the relative dry-mass
    change over the step.  ``from_envelope`` sets the first three to a
    margin over what the fine trajectories of training reached."""
    cfl_h: float = 1.0
    cfl_v: float = float("inf")
    div: float = float("inf")
    mass_rel: float = 1e-9

    @classmethod
    def from_envelope(cls, envelope, margin: float = 1.5, mass_rel: float = 1e-9) -> "Limits":
        e = [float(x) for x in np.asarray(envelope).ravel()]
        return cls(cfl_h=min(1.0, margin * e[0]), cfl_v=margin * e[1], div=margin * e[2], mass_rel=mass_rel)


class LearnedStepper:
    """Corrected coarse steps with the plain fine steps as the fallback.

    The state lives in the coarse driver.  Each step snapshots it into the
    fine driver (a device copy), takes one coarse step, applies the
    network's correction and checks the result; a trip (validation flag,
    CFL, divergence or dry mass) discards it and advances the snapshot with
    two plain fine steps instead."""

    def __init__(self, case, net_state=None, *, layers: int = 4, limits: Optional[Limits] = None,
                 drivers=None) -> None:
        self.case = case
        coarse = tuple(int(x) for x in np.asarray(net_state["coarse"]).ravel()) \
            if net_state is not None and "coarse" in net_state else COARSE_DEFAULT
        self.fine, self.coarse = drivers or make_pair_drivers(case, coarse)
        self.net = CorrectorNet(self.coarse, layers)
        if net_state is not None:
            self.net.load_state(net_state)
        if limits is None:
            limits = Limits.from_envelope(self.net.envelope) if np.any(self.net.envelope > 0) else Limits()
        self.limits = limits
        self.gain = 1.0   # fraction of the predicted correction that is applied
        self.fallbacks = 0
        self.steps = 0
        self.trips: List[str] = []
        self._mass = None

    def set_state(self, state) -> None:
        self.coarse.set_state(state)
        self._mass = None

    def step(self, n: int = 1) -> None:
        lim = self.limits
        for _ in range(n):
            if self._mass is None:
                self._mass = self.net.checks(self.coarse)[3]
            copy_state(self.fine, self.coarse)
            why = None
            try:
                self.coarse.step_device(1)
            except FloatingPointError:
                why = "flag"
            if why is None:
                self.net.forward(self.coarse)
                self.net.apply(self.coarse, self.gain)
                ch, cv, dv, m1 = self.net.checks(self.coarse)
                if not (ch <= lim.cfl_h):
                    why = "cfl_h"
                elif not (cv <= lim.cfl_v):
                    why = "cfl_v"
                elif not (dv <= lim.div):
                    why = "div"
                elif not (abs(m1 - self._mass) <= lim.mass_rel * abs(self._mass)):
                    why = "mass"
            if why is None:
                self._mass = m1
            else:
                self.fallbacks += 1
                self.trips.append(why)
                self.fine.step_device(2)
                copy_state(self.coarse, self.fine)
                self._mass = None
            self.steps += 1

    def get_state(self):
        return self.coarse.get_state()

    def close(self) -> None:
        self.net.close()
        self.fine.close()
        self.coarse.close()


def save(net: CorrectorNet, path, **meta) -> None:
    d = net.state_dict()
    d.update({k: np.array(v) for k, v in meta.items()})
    np.savez_compressed(path, **d)


def load(path) -> Dict[str, np.ndarray]:
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}
