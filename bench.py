#!/usr/bin/env python3
"""Headline benchmark: simulated days per wall-clock day of the plain dycore
on x1.40962 L55 moist (T=6, FCT transport), dt 720 s, one MI355X, with the
v8.4.1 2-D Smagorinsky horizontal mixing (the MPAS default; --mixing off for
the unmixed step).

A step is one large step (3 dynamics subcycles x RK3 with acoustic substeps,
then split FCT transport) of the whole mesh with the state resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line.  `roofline` reports the dominant kernel (largest total
time) with its algorithmic bytes per launch over its average launch time,
timed with HIP events on the launch stream; `step_roofline` reports the whole
step's algorithmic bytes (phase-level model) over the step time;
`cpu_baseline` times the CPU restatement of the same step (binary32 strict,
OpenMP over all host cores) on a bounded sample of the same workload;
`fp64_rel_l2` states the measured numerics' error after one large step
against the FP64 restatement (relative L2 on u, w, theta_m, rho_zz).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
HEADLINE_METRIC = "sim-days/wall-day x1.40962 L55, 1 GPU: plain vs learned-corrected; %HBM peak"
sys.path.insert(0, str(ROOT / "gpuwm-hex_amd"))


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=120)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="x1.40962-moist")
    p.add_argument("--numerics", default="strict", choices=["strict", "fast"])
    p.add_argument("--mixing", default="2d_smagorinsky", choices=["2d_smagorinsky", "off"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-steps", type=int, default=0, help="CPU sample size in large steps (0: auto ~15 s)")
    p.add_argument("--profile-steps", type=int, default=2)
    p.add_argument("--calibration-copy-mib", type=int, default=0,
                   help="also run a device copy of this many MiB (counter calibration for rocprofv3 --pmc)")
    p.add_argument("--form", default="both", choices=["both", "plain"],
                   help="both: also time the learned-corrected form (needs the trained corrector file)")
    p.add_argument("--error-hours", type=float, default=24.0,
                   help="forecast length of the surface-pressure error against the stored FP64 run")
    return p.parse_args()


def cpu_baseline(case, cpu_steps: int):
    """Time the CPU restatement (binary32 strict) on a bounded sample."""
    sys.path.insert(0, str(ROOT))
    from oracle.oracle import Oracle
    from hexdyn.case import initial_state_arrays

    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    t0 = time.perf_counter()
    o.step(1)
    first = time.perf_counter() - t0
    n = cpu_steps if cpu_steps > 0 else max(1, min(8, int(15.0 / max(first, 1e-3))))
    t0 = time.perf_counter()
    o.step(n)
    wall = time.perf_counter() - t0
    sim = n * case.config.config_dt
    return {
        "value": sim / wall,
        "unit": "sim-days/wall-day",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} large steps of {case.name} (after 1 untimed), {wall:.1f} s wall, "
                  f"binary32 strict CPU restatement, OpenMP",
        "seconds_per_step": wall / n,
    }


def fp64_error(case, got):
    """Relative L2 error of one GPU large step (`got`, the measured numerics)
    against the FP64 restatement of the same step, on u, w, theta_m, rho_zz.
    Part of the CPU leg: the restatement is the checker, never the thing timed."""
    import numpy as np

    sys.path.insert(0, str(ROOT))
    from oracle.oracle import Oracle
    from hexdyn.case import initial_state_arrays

    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    if o.step(1) != 0:
        return {"error": "FP64 restatement refused the step"}
    ref = o.get_state()
    out = {}
    for k in ("u", "w", "theta_m", "rho_zz"):
        r = np.asarray(ref[k], dtype=np.float64)
        out[k] = float(np.linalg.norm(np.asarray(got[k], dtype=np.float64) - r) / np.linalg.norm(r))
    out["after"] = "1 large step from the initial state"
    return out


def surface_pressure_rms(case, state, ps64):
    sys.path.insert(0, str(ROOT))
    from oracle.surface_pressure import surface_pressure
    import numpy as np

    d = surface_pressure(case, state) - ps64
    return float(np.sqrt(np.mean(d * d)))


def learned_leg(case, args, plain_value, plain_state_24h, variant=""):
    """Time the learned-corrected form (coarse step + correction + checks,
    plain fallback) on the same workload and state its 24 h error."""
    import numpy as np
    import torch

    from hexdyn import learned
    from hexdyn.case import initial_state_arrays

    wfile = learned.DATA / f"corrector_{args.workload}{variant}.npz"
    if not wfile.exists():
        return {"error": f"no trained corrector {wfile.name}"}
    net_state = learned.load(wfile)
    st = learned.LearnedStepper(case, net_state, layers=int(net_state["layers"]))
    ratio = 2
    for _ in range(max(1, args.warmup // ratio)):
        st.step(1)
    st.set_state(initial_state_arrays(case))
    torch.cuda.synchronize()
    n = max(1, args.steps // ratio)
    fb0 = st.fallbacks
    w0 = time.perf_counter()
    st.step(n)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    value = n * ratio * case.config.config_dt / wall
    out = {"form": "learned", "weights": wfile.name, "value": round(value, 1), "unit": "sim-days/wall-day",
           "ms_per_coarse_step": round(1000.0 * wall / n, 4), "coarse_steps": n,
           "coarse_dt": case.config.config_dt * ratio,
           "coarse_number_of_sub_steps": int(st.net.coarse[0]), "coarse_dynamics_split_steps": int(st.net.coarse[1]),
           "speedup_vs_plain": round(value / plain_value, 4), "fallbacks": st.fallbacks - fb0,
           "trips": sorted(set(st.trips)), "limits": st.limits.__dict__,
           "network": {"layers": int(net_state["layers"]), "hidden": learned.HIDDEN,
                       "params": int(np.asarray(net_state["params"]).size)}}
    # MFMA roofline of the network's GEMMs (HIP events on the launch stream)
    st.coarse.profile(True)
    st.net.forward(st.coarse)
    st.net.forward(st.coarse)
    torch.cuda.synchronize()
    avg, launches = st.coarse.profile_read("k_net_gemm")
    st.coarse.profile(False)
    C, L = case.mesh.n_cells, case.n_levels
    layers = int(net_state["layers"])
    flops = 2.0 * C * (learned.N_IN * learned.HIDDEN + layers * 2 * learned.HIDDEN * learned.HIDDEN
                       + learned.HIDDEN * learned.N_OUT)
    gemm_ms = avg * launches / 2.0
    if gemm_ms > 0:
        tf = flops / (gemm_ms * 1e-3) / 1e12
        out["roofline_mfma"] = {"bound": "mfma", "achieved": round(tf, 2), "peak": 2500.0, "unit": "TFLOP/s",
                                "frac": round(tf / 2500.0, 5), "flops_per_forward": flops,
                                "gemm_ms_per_forward": round(gemm_ms, 5), "gemm_launches_per_forward": launches / 2}
    # 24 h run of the learned form: FP64 integrals every step, and the
    # surface-pressure error against the stored FP64 run beside plain's
    if plain_state_24h is not None:
        sys.path.insert(0, str(ROOT))
        from oracle import surface_pressure as sp

        st2 = learned.LearnedStepper(case, net_state, layers=layers, drivers=(st.fine, st.coarse))
        st2.set_state(initial_state_arrays(case))
        nsteps = int(round(args.error_hours * 3600.0 / (ratio * case.config.config_dt)))
        first = st2.coarse.integrals()
        last = first
        for _ in range(nsteps):  # FP64 integrals every (coarse) step
            st2.step(1)
            last = st2.coarse.integrals()
        out["learned_24h_fp64_integral_drift"] = {
            "dry_mass": (last["dry_mass"] - first["dry_mass"]) / first["dry_mass"],
            "total_energy": (last["total_energy"] - first["total_energy"]) / first["total_energy"],
            "fallbacks": st2.fallbacks, "steps": st2.steps}
        learned24 = st2.get_state()
        # the same coarse steps without the correction (what the network adds)
        st.coarse.set_state(initial_state_arrays(case))
        st.coarse.step_device(nsteps)
        coarse24 = st.coarse.get_state()
        ps_plain = sp.surface_pressure(case, plain_state_24h)
        out["ps_rms_vs_plain_pa"] = {"learned": surface_pressure_rms(case, learned24, ps_plain),
                                     "uncorrected_coarse": surface_pressure_rms(case, coarse24, ps_plain)}
        ps64 = sp.load(case, args.workload, args.error_hours)
        if ps64 is not None:
            lerr = surface_pressure_rms(case, learned24, ps64)
            perr = surface_pressure_rms(case, plain_state_24h, ps64)
            out["ps_rms_vs_fp64"] = {"hours": args.error_hours, "plain_pa": perr, "learned_pa": lerr,
                                     "uncorrected_coarse_pa": surface_pressure_rms(case, coarse24, ps64),
                                     "ratio": lerr / perr if perr > 0 else None}
        else:
            out["ps_rms_vs_fp64"] = {"error": "no stored FP64 surface pressure for this case"}
    st.close()
    return out


ENSEMBLE = "ensemble-8x40962-24x10242"


def ensemble_bench(args) -> None:
    """Batched ensemble: 8 members on x1.40962 (dt 720 s) and 24 on x1.10242
    (dt 1440 s), each mesh group one instance on the union of its members'
    meshes (one launch sequence per group), advanced 12 h after one warm-up
    interval.  Beside it, the same members as separate instances (timed on
    one member of each mesh over 2 h and scaled by the member counts)."""
    import torch

    from hexdyn import ensemble
    from hexdyn.driver import HexDycoreDriver

    t0 = time.perf_counter()
    ens = ensemble.config3()
    setup = time.perf_counter() - t0
    print(f"# ensemble set-up {setup:.1f} s", file=sys.stderr, flush=True)
    ens.advance(ens.interval)
    hours = 12.0
    wall = ens.advance(hours * 3600.0)
    members = sum(g.n for g in ens.groups)
    value = members * hours / 24.0 / (wall / 86400.0)
    sep = 0.0
    detail = []
    for g in ens.groups:
        d = HexDycoreDriver.from_host(g.case, state=ensemble.member_state(g.case, 0))
        d.step_device(1)
        n = int(round(2 * 3600.0 / g.dt))
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        d.step_device(n)
        torch.cuda.synchronize()
        one = (time.perf_counter() - w0) / n
        d.close()
        per_member_12h = one * hours * 3600.0 / g.dt
        sep += per_member_12h * g.n
        detail.append({"mesh": g.case.name, "members": g.n, "dt": g.dt, "cells_per_member": g.case.mesh.n_cells,
                       "ms_per_step_separate": round(1000 * one, 4)})
    ens.close()
    print(json.dumps({
        "metric": "member-sim-days/wall-day, batched ensemble 8 x x1.40962 + 24 x x1.10242 L55, 1 GPU",
        "value": round(value, 1), "unit": "member-sim-days/wall-day", "n_gpus": 1, "hours": hours,
        "wall_s": round(wall, 4), "members": members, "higher_is_better": True, "dtype": "f32",
        "data": "synthetic (seeded meshes, JW jet, seeded moisture, seeded smoothed theta noise sigma 0.1 K per member)",
        "config": {"workload": ENSEMBLE, "groups": detail, "form": "plain", "numerics": "strict",
                   "horiz_mixing": "2d_smagorinsky", "hipgraph": True},
        "separate_instances_wall_s_estimate": round(sep, 4),
        "batching_speedup": round(sep / wall, 4)}))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": f"--gpus {args.gpus}: this build runs on one GPU only; "
                                   "multi-GPU sharding is designed but not built"}))
        sys.exit(2)
    if args.workload == ENSEMBLE:
        ensemble_bench(args)
        return
    import torch

    from hexdyn.bytes_model import HBM_PEAK_GBS, kernel_bytes, step_bytes
    from hexdyn.case import build_case
    from hexdyn.config import DycoreConfig
    from hexdyn.driver import HexDycoreDriver
    from hexdyn.case import WORKLOADS

    spec = WORKLOADS[args.workload]
    cfg = DycoreConfig(config_dt=spec["dt"], numerics=args.numerics, config_horiz_mixing=args.mixing)
    t_setup = time.perf_counter()
    case = build_case(args.workload, config=cfg)
    drv = HexDycoreDriver.from_host(case)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    print(f"# setup {setup_s:.1f}s: C={case.mesh.n_cells} E={case.mesh.n_edges} L={case.n_levels} "
          f"T={case.n_scalars}", file=sys.stderr, flush=True)

    if args.calibration_copy_mib > 0:
        n = args.calibration_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    from hexdyn import autotune

    tuning = drv.tuning
    if tuning is None:
        tuning = autotune.tune(drv)
        try:
            autotune.save(tuning)
        except OSError:
            pass
    print(f"# launch tuning: {tuning['table']} (default {tuning['default_ms_per_step']} ms/step, "
          f"tuned {tuning['tuned_ms_per_step']} ms/step)", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        drv.step_device(1)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    ev0.record(stream)
    drv.step_device(args.steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = 1000.0 * wall / args.steps
    sim_seconds = args.steps * case.config.config_dt
    value = sim_seconds / wall
    print(f"# timed {args.steps} steps: {wall:.3f}s wall ({gpu_ms:.1f} ms events), {ms_per_step:.3f} ms/step",
          file=sys.stderr, flush=True)

    # per-kernel timing (events on the launch stream around every kernel)
    drv.profile(True)
    drv.step_device(args.profile_steps)
    torch.cuda.synchronize()
    C, E, V, L, T = case.mesh.n_cells, case.mesh.n_edges, case.mesh.n_vertices, case.n_levels, case.n_scalars
    kernels = []
    for name in drv.profile_kernels():
        avg_ms, launches = drv.profile_read(name)
        try:
            kb = kernel_bytes(name, C=C, E=E, V=V, L=L, T=T, max_edges=case.mesh.max_edges)
        except KeyError:
            kb = None
        kernels.append({"kernel": name, "avg_ms": avg_ms, "launches_per_step": launches / args.profile_steps,
                        "total_ms_per_step": avg_ms * launches / args.profile_steps,
                        "alg_bytes": kb, "gbs": (kb / (avg_ms * 1e-3) / 1e9) if kb and avg_ms > 0 else None})
    drv.profile(False)
    kernels.sort(key=lambda r: -r["total_ms_per_step"])
    dom = kernels[0]
    # rocprofv3 FETCH_SIZE/WRITE_SIZE bytes per launch of the same command (hexdyn.profile_summary)
    traffic = None
    tag = args.workload + ("" if args.mixing != "off" else "-nomix") + ("" if args.numerics == "strict" else "-fast")
    tfile = ROOT / "profiles" / f"traffic_{tag}.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text())["hbm_bytes_per_launch"].get(dom["kernel"])
        except (ValueError, KeyError):
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(dom["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom["kernel"],
                "avg_launch_ms": round(dom["avg_ms"], 5), "alg_bytes_per_launch": dom["alg_bytes"]}
    from hexdyn.setup import rk_schedule

    sched = rk_schedule(cfg.config_dt, order=cfg.config_time_integration_order,
                        acoustic_substeps=cfg.config_number_of_sub_steps,
                        dynamics_splits=cfg.config_dynamics_split_steps)
    n_ac = sum(s.acoustic_steps for s in sched)
    b_alg = step_bytes(C=C, L=L, T=T, max_edges=case.mesh.max_edges, n_ac=n_ac, mix=args.mixing != "off",
                       splits=cfg.config_dynamics_split_steps)
    step_gbs = b_alg / (ms_per_step * 1e-3) / 1e9
    step_roofline = {"alg_bytes_per_step": b_alg, "achieved": round(step_gbs, 1), "unit": "GB/s",
                     "frac": round(step_gbs / HBM_PEAK_GBS, 4)}

    # the CPU leg (rank 0, N=1): the CPU baseline and the FP64 error of one GPU step
    cpu = None
    err64 = None
    if not args.no_cpu_baseline:
        first = HexDycoreDriver.from_host(case)
        first.step_device(1)
        got = first.get_state()
        first.close()
        try:
            cpu = cpu_baseline(case, args.cpu_steps)
        except Exception as exc:  # reported, never silently substituted
            cpu = {"error": repr(exc)}
        try:
            err64 = fp64_error(case, got)
        except Exception as exc:
            err64 = {"error": repr(exc)}

    learned_out = None
    learned_alt = None
    if args.form == "both":
        plain24 = None
        try:
            hrs = int(round(args.error_hours * 3600.0 / case.config.config_dt))
            d24 = HexDycoreDriver.from_host(case)
            first = d24.integrals()
            last = first
            for _ in range(hrs):  # FP64 integrals every step (GPU, fixed-order sums)
                d24.step_device(1)
                last = d24.integrals()
            plain24 = d24.get_state()
            plain_drift = {"dry_mass": (last["dry_mass"] - first["dry_mass"]) / first["dry_mass"],
                           "total_energy": (last["total_energy"] - first["total_energy"]) / first["total_energy"]}
            d24.close()
        except Exception as exc:
            learned_out = {"error": f"plain 24 h run: {exc!r}"}
        if learned_out is None:
            try:
                learned_out = learned_leg(case, args, value, plain24)
                learned_out["plain_24h_fp64_integral_drift"] = plain_drift
            except Exception as exc:  # reported, never silently substituted
                learned_out = {"error": repr(exc)}
            # the more accurate, slower coarse step (3 substeps, 3 subcycles), when trained
            from hexdyn import learned as _l

            if (_l.DATA / f"corrector_{args.workload}.split3.npz").exists():
                try:
                    learned_alt = learned_leg(case, args, value, plain24, variant=".split3")
                except Exception as exc:
                    learned_alt = {"error": repr(exc)}

    out = {
        "metric": HEADLINE_METRIC if args.workload == "x1.40962-moist"
        else f"sim-days/wall-day {args.workload.split('-')[0]} L{L}, 1 GPU; %HBM peak",
        "value": round(value, 1),
        "unit": "sim-days/wall-day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded icosahedral SCVT mesh, JW baroclinic jet, random terrain, seeded moisture)",
        "config": {"workload": args.workload, "cells": C, "edges": E, "levels": L, "scalars": T,
                   "dt": cfg.config_dt, "form": "plain", "numerics": args.numerics,
                   "horiz_mixing": args.mixing,
                   "dynamics_split_steps": cfg.config_dynamics_split_steps,
                   "number_of_sub_steps": cfg.config_number_of_sub_steps, "parallelism": "single-gpu",
                   "launch_tuning": {"waves_per_block": tuning["table"], "default": 4,
                                     "default_ms_per_step": tuning["default_ms_per_step"],
                                     "tuned_ms_per_step": tuning["tuned_ms_per_step"]}},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "cpu_baseline": cpu,
        "fp64_rel_l2": err64,
        "learned": learned_out,
        "learned_split3": learned_alt,
        "kernels_top": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in r.items()} for r in kernels[:8]],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
