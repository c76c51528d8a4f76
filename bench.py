"""Benchmark of the veros mixing hot path on one MI355X.

One step = the kbot-masked TDMA solve (solve_implicit on 1440x720x80
vertical-diffusion systems) + isoneutral_diffusion_pre + isoneutral_diffusion
for temp and for salt on the 1440x720x80 quarter-degree grid, inputs resident
in HBM; the isoneutral part runs as one call (vx_iso_pre_diffusion_ts_f64),
with the separate pre and temperature + salinity calls timed beside it.
Timed with HIP events on the stream the kernels run on.

Prints ONE JSON line (rank 0). ``value`` is the step throughput in Mcells/s
(interior cells / step time) of the configured precision form; TDMA
Mcolumns/s and isoneutral Mcells/s are reported beside it for both forms.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

HBM_PEAK_GBS = 8000.0
# HBM bytes per call of vx_iso_pre_diffusion_ts_f64 (its seven kernels) on
# 1440x720x80, from the FETCH_SIZE and WRITE_SIZE rocprofv3 passes recorded in
# profiles/ (read bytes = 2.0 x FETCH_SIZE, write bytes = 1.0 x WRITE_SIZE,
# calibrated on a 1 GiB device copy)
PRE_TS_TRAFFIC_BYTES = {(1440, 720, 80): 27_204_216_832}
# algorithmic bytes per interior cell: the fused temperature + salinity step
# with the energy branch reads K_iso, Ai_* and K_11/22/33 once (8 + 128 + 24),
# P_diss read and written (16), and per tracer tr[tau], tr[taup1] read and
# written, dtr_iso read and written, int_drhodX (8 + 16 + 16 + 8)
TS_BYTES_PER_CELL = 8 + 128 + 24 + 16 + 2 * (8 + 16 + 16 + 8)
# pre + the temperature + salinity step in one call: K_iso read (8), Ai_* and
# K_11/22/33 written instead of read (128 + 24), the rest as above
PRE_TS_BYTES_PER_CELL = 8 + 128 + 24 + 16 + 2 * (8 + 16 + 16 + 8)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=1440)
    ap.add_argument("--ny", type=int, default=720)
    ap.add_argument("--nz", type=int, default=80)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--refine", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--batched", action="store_true",
                    help="instead of the step: 64 independent uneven grids, sequential launches vs one batched "
                         "plan (untuned, autotuned, hipGraph), fp64 and mixed")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy (a known-traffic dispatch for PMC scaling)")
    return ap.parse_args()


def build_workload(nx, ny, nz, seed=1002):
    """Config-2 inputs built on the device (same recipe as
    veros_amd.workloads.iso_config2 / tdma_config2, drawn with torch RNG)."""
    import math

    import numpy as np
    import torch

    from veros_amd import workloads as W
    from veros_amd.state import to_device_state

    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    X, Y = nx + 4, ny + 4
    rng = np.random.default_rng(seed)
    settings = W.config2_settings(nx, ny, nz)
    grid = W.make_grid(nx, ny, nz, dx_deg=0.25, dy_deg=0.25, y_origin=-89.75, dz_top=10.0, dz_bottom=250.0)
    kbot = W.make_kbot(rng, nx, ny, nz, kmin=min(5, nz), land_frac=0.35, cyclic=True)
    f64 = dict(dtype=torch.float64, device=dev)
    zt = torch.tensor(grid["zt"], **f64)
    T1 = 20.0 * torch.exp(zt / 800.0)[None, None, :] + 0.1 * torch.randn((X, Y, nz), generator=g, **f64)
    S1 = 35.0 + 0.3 * torch.randn((X, Y, nz), generator=g, **f64)
    for F in (T1, S1):
        F[-2:] = F[2:4].clone()
        F[:2] = F[-4:-2].clone()
    vs = dict(tau=0, taup1=1, kbot=torch.from_numpy(kbot).to(dev))
    vs["temp"] = T1[..., None].repeat(1, 1, 1, 3).contiguous()
    vs["salt"] = S1[..., None].repeat(1, 1, 1, 3).contiguous()
    del T1, S1
    absz = zt.abs()[None, None, :, None]
    vs["int_drhodT"] = (absz * settings.drhodT).expand(X, Y, nz, 3).contiguous()
    vs["int_drhodS"] = (absz * settings.drhodS).expand(X, Y, nz, 3).contiguous()
    vs["K_iso"] = torch.full((X, Y, nz), 1000.0, **f64)
    for k in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by"):
        vs[k] = torch.zeros((X, Y, nz, 2, 2), **f64)
    for k in ("K_11", "K_22", "K_33", "dtemp_iso", "dsalt_iso", "P_diss_iso"):
        vs[k] = torch.zeros((X, Y, nz), **f64)
    for k, v in grid.items():
        vs[k] = torch.tensor(v, **f64)
    # vertical mixing inputs (veros_amd.workloads.add_vertmix_fields): kappaH ~
    # LogU(1e-5, 1e-2), surface forcing ~ 1e-5 N(0, 1) (salt 1e-6)
    lo, hi = math.log(1e-5), math.log(1e-2)
    vs["kappaH"] = torch.exp(lo + (hi - lo) * torch.rand((X, Y, nz), generator=g, **f64))
    vs["forc_temp_surface"] = 1e-5 * torch.randn((X, Y), generator=g, **f64)
    vs["forc_salt_surface"] = 1e-6 * torch.randn((X, Y), generator=g, **f64)
    vs["dtemp_vmix"] = torch.zeros((X, Y, nz), **f64)
    vs["dsalt_vmix"] = torch.zeros((X, Y, nz), **f64)
    # vertical friction inputs (veros_amd.workloads.add_friction_fields):
    # u, v ~ 0.1 N(0, 1) m/s, kappaM ~ LogU(1e-5, 1e-2), the T and v-face areas
    for k in ("u", "v"):
        vs[k] = 0.1 * torch.randn((X, Y, nz, 3), generator=g, **f64)
    vs["kappaM"] = torch.exp(lo + (hi - lo) * torch.rand((X, Y, nz), generator=g, **f64))
    vs["area_t"] = vs["cost"][None, :] * vs["dyt"][None, :] * vs["dxt"][:, None]
    vs["area_v"] = vs["cosu"][None, :] * vs["dyu"][None, :] * vs["dxt"][:, None]
    for k in ("du_mix", "dv_mix", "K_diss_v"):
        vs[k] = torch.zeros((X, Y, nz), **f64)
    from types import SimpleNamespace
    state = SimpleNamespace(variables=SimpleNamespace(**vs), settings=settings)

    # TDMA systems: a = c = -dt*kappa/dz^2, kappa ~ LogU(1e-5, 1e-2), b = 1 - a - c, d ~ N(0, 1)
    dz = torch.tensor(W.geometric_dz(nz, 10.0, 250.0), **f64)
    lo, hi = math.log(1e-5), math.log(1e-2)
    kappa = torch.exp(lo + (hi - lo) * torch.rand((nx, ny, nz), generator=g, **f64))
    a = -3600.0 * kappa / dz[None, None, :] ** 2
    del kappa
    c = a.clone()
    b = 1.0 - a - c
    d = torch.randn((nx, ny, nz), generator=g, **f64)
    tk = torch.from_numpy(rng.integers(min(5, nz), nz + 1, size=(nx, ny)).astype(np.int32))
    tk[torch.from_numpy(rng.random((nx, ny)) < 0.35)] = 0
    tdma = dict(a=a, b=b, c=c, d=d, kbot=tk.to(dev), x=torch.empty_like(d))
    return state, tdma


def tdma_bytes(kbot, nz):
    import torch
    kb = kbot.to(torch.int64)
    wet = torch.where((kb >= 1) & (kb <= nz), nz - kb + 1, torch.zeros_like(kb))
    W = int(wet.sum().item())
    C = kb.numel()
    return 32 * W + 8 * C * nz + 4 * C, W


def cpu_baseline(nx, ny, nz, budget_s):
    """The oracle (NumPy float64; dgtsv for the TDMA) on a bounded sample of
    the same workload: a 1440 x ny_s x 80 TDMA block and a 360 x ny_s x 80
    isoneutral block, sized from a short calibration run."""
    import numpy as np

    from oracle import isoneutral as OI
    from oracle import tdma as OT
    from veros_amd import workloads as W

    def time_iso(nys):
        st = W.iso_config2(nx=360, ny=nys, nz=nz)
        st.settings = W.config2_settings(360, nys, nz)
        t0 = time.perf_counter()
        pre = OI.isoneutral_diffusion_pre(st)
        for k, v in pre._asdict().items():
            setattr(st.variables, k, v)
        OI.isoneutral_diffusion(st, "temp", True)
        OI.isoneutral_diffusion(st, "salt", False)
        return time.perf_counter() - t0, 360 * nys * nz

    def time_tdma(nys):
        p = W.tdma_config2(nx=nx, ny=nys, nz=nz)
        t0 = time.perf_counter()
        OT.solve_implicit_kbot(p.a, p.b, p.c, p.d, p.kbot)
        return time.perf_counter() - t0, nx * nys

    t_iso, n_iso = time_iso(4)
    t_td, n_td = time_tdma(8)
    scale_iso = max(1, int((budget_s * 0.7) / max(t_iso, 1e-3) * 4))
    scale_td = max(1, int((budget_s * 0.3) / max(t_td, 1e-3) * 8))
    t_iso, n_iso = time_iso(min(scale_iso, 720))
    t_td, n_td = time_tdma(min(scale_td, 720))
    iso_rate = n_iso / t_iso / 1e6
    td_rate = n_td / t_td / 1e6
    # step-equivalent: per interior cell, the TDMA costs (1/nz) columns
    N = nx * ny * nz
    t_step = N / (iso_rate * 1e6) + (nx * ny) / (td_rate * 1e6)
    return dict(value=N / t_step / 1e6, unit="Mcells/s", cores=1, kind="port",
                sample=f"oracle (NumPy float64, LAPACK dgtsv) on a 360x{min(scale_iso, 720)}x{nz} isoneutral "
                       f"block ({t_iso:.1f} s) and a {nx}x{min(scale_td, 720)}x{nz} TDMA block ({t_td:.1f} s)",
                tdma_mcolumns_per_s=td_rate, iso_mcells_per_s=iso_rate,
                host=dict(cpu_count=os.cpu_count(), affinity=len(os.sched_getaffinity(0)),
                          cpu=platform.processor() or platform.machine(), numpy=np.__version__))


def run_batched(args):
    """64 grids, nx, ny ~ U{48..720}, nz ~ U{15..100}, kbot ~ U{1..nz} with
    30% land: 64 sequential solve_implicit launches vs one batched plan."""
    import torch

    from veros_amd import utilities
    from veros_amd.batched import TdmaBatch, random_systems

    torch.cuda.set_device(0)
    print("[bench] building 64 uneven grids", file=sys.stderr, flush=True)
    systems = random_systems(64)
    torch.cuda.synchronize()
    cols = sum(s["a"].shape[0] * s["a"].shape[1] for s in systems)
    cells = sum(s["a"].numel() for s in systems)
    nbytes = sum(tdma_bytes(s["kbot"], s["a"].shape[2])[0] for s in systems)
    stream = torch.cuda.current_stream()

    def timed(fn, reps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for s0, e0 in evs:
            s0.record(stream)
            fn()
            e0.record(stream)
        torch.cuda.synchronize()
        ts = sorted(a.elapsed_time(b) for a, b in evs)
        return ts[len(ts) // 2]

    res = {}
    for prec in ("fp64", "mixed"):
        def sequential():
            for s0 in systems:
                utilities.solve_implicit_kbot(s0["a"], s0["b"], s0["c"], s0["d"], s0["kbot"], precision=prec,
                                              refine_steps=1, out=s0["x"])
        plan = TdmaBatch(systems, prec)
        r = {"sequential_64_launches_ms": timed(sequential, args.steps),
             "plan_untuned_ms": timed(plan.run, args.steps)}
        info = plan.tune()
        r["plan_tuned_ms"] = timed(plan.run, args.steps)
        r["plan_graph_ms"] = timed(plan.graph_launch, args.steps)
        r["tuned_choices"] = info["buckets"]
        best = min(r["plan_tuned_ms"], r["plan_graph_ms"])
        r["speedup_vs_sequential"] = round(r["sequential_64_launches_ms"] / best, 3)
        r["mcolumns_per_s"] = round(cols / (best * 1e-3) / 1e6, 1)
        r["algorithmic_GBps"] = round(nbytes / (best * 1e-3) / 1e9, 1)
        res[prec] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        plan.close()
    print(json.dumps({"workload": "64 independent grids, nx,ny ~ U{48..720}, nz ~ U{15..100}, 30% land",
                      "columns": cols, "cells": cells, "algorithmic_bytes": nbytes, "results": res}))


def main():
    args = parse()
    if args.batched:
        return run_batched(args)
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus above 1 is not supported (the TDMA columns would "
                                   "shard with no communication, the stencils by 2-D blocks with a 2-cell "
                                   "halo exchange of T and S per step; not built)"}))
        sys.exit(2)
    import torch

    from veros_amd import friction, isoneutral, thermodynamics, utilities

    torch.cuda.set_device(0)
    nx, ny, nz = args.nx, args.ny, args.nz
    N = nx * ny * nz
    print(f"[bench] building {nx}x{ny}x{nz} workload", file=sys.stderr, flush=True)
    state, td = build_workload(nx, ny, nz)
    torch.cuda.synchronize()
    vs = state.variables
    stream = torch.cuda.current_stream()
    # a plain 1 GiB device-to-device copy: the sustained HBM rate this box
    # reaches (read + write bytes / time), reported next to the roofline, and a
    # dispatch of known traffic for PMC scaling
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3 if args.calibrate_copy else 2):
        dst.copy_(src)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    for _ in range(5):
        dst.copy_(src)
    c1.record(stream)
    torch.cuda.synchronize()
    copy_gbps = 5 * 2 * src.numel() * 8 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del src, dst

    def run_tdma(prec):
        utilities.solve_implicit_kbot(td["a"], td["b"], td["c"], td["d"], td["kbot"], precision=prec,
                                      refine_steps=args.refine, out=td["x"])

    def run_pre():
        isoneutral.isoneutral_diffusion_pre(state)

    def run_diff():
        isoneutral.isoneutral_diffusion_temp_salt(state)

    def run_pre_diff():
        isoneutral.isoneutral_pre_and_diffusion_temp_salt(state)

    def run_vertmix():
        thermodynamics.vertmix_tempsalt(state)

    def run_friction():
        friction.implicit_vert_friction(state)

    def run_diff_two_calls():
        isoneutral.isoneutral_diffusion(state, vs.temp, True)
        isoneutral.isoneutral_diffusion(state, vs.salt, False)

    def timed(fn, reps):
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for s, e in evs:
            s.record(stream)
            fn()
            e.record(stream)
        torch.cuda.synchronize()
        ts = sorted(s.elapsed_time(e) for s, e in evs)
        return ts[len(ts) // 2], ts[0]

    # per-part timings (median, min) in ms
    parts = {}
    for name, fn in (("tdma_fp64", lambda: run_tdma("fp64")), ("tdma_mixed", lambda: run_tdma("mixed")),
                     ("iso_pre", run_pre), ("iso_diffusion_TS", run_diff),
                     ("iso_diffusion_T_then_S", run_diff_two_calls), ("iso_pre_and_diffusion_TS", run_pre_diff),
                     ("vertmix_TS", run_vertmix), ("friction_UV", run_friction)):
        for _ in range(args.warmup):
            fn()
        parts[name] = timed(fn, max(5, args.steps))
        print(f"[bench] {name}: median {parts[name][0]:.3f} ms, min {parts[name][1]:.3f} ms",
              file=sys.stderr, flush=True)

    prec = args.precision

    def step():
        run_tdma(prec)
        run_pre_diff()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(args.steps):
        step()
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps

    B_td, W_td = tdma_bytes(td["kbot"], nz)
    C = nx * ny
    B_pre = 176 * N + 4 * (nx + 4) * (ny + 4)
    B_diff_two = 2 * (224 * N + 4 * (nx + 4) * (ny + 4))
    B_diff = TS_BYTES_PER_CELL * N + 4 * (nx + 4) * (ny + 4)
    B_pre_diff = PRE_TS_BYTES_PER_CELL * N + 4 * (nx + 4) * (ny + 4)
    # vertmix_tempsalt: kappaH read and temp/salt[taup1] read and written on the
    # wet cells (8 + 2 * 16), the two tendencies written on every cell
    # (2 * 8 over X*Y*nz), kbot and the two forcings per column (4 + 16)
    _, W_v = tdma_bytes(vs.kbot[2:-2, 2:-2], nz)
    B_vmix = 40 * W_v + 16 * (nx + 4) * (ny + 4) * nz + 20 * (nx + 4) * (ny + 4)
    # implicit_vert_friction: per cell kappaM read (8), u and v at tau read and
    # at taup1 written or kept (4 * 8), du_mix and dv_mix written (16),
    # K_diss_v read and written (16), over the X*Y*nz array
    B_fric = 72 * (nx + 4) * (ny + 4) * nz
    rates = {
        "tdma_mcolumns_per_s": {p: C / (parts[f"tdma_{p}"][0] * 1e-3) / 1e6 for p in ("fp64", "mixed")},
        "iso_mcells_per_s": N / (parts["iso_pre_and_diffusion_TS"][0] * 1e-3) / 1e6,
        "iso_mcells_per_s_pre_then_TS": N / ((parts["iso_pre"][0] + parts["iso_diffusion_TS"][0]) * 1e-3) / 1e6,
    }
    kern = {
        "tdma_fp64": (B_td, parts["tdma_fp64"][0]),
        "tdma_mixed": (B_td, parts["tdma_mixed"][0]),
        "iso_pre": (B_pre, parts["iso_pre"][0]),
        "iso_diffusion_TS": (B_diff, parts["iso_diffusion_TS"][0]),
        "iso_diffusion_T_then_S": (B_diff_two, parts["iso_diffusion_T_then_S"][0]),
        "iso_pre_and_diffusion_TS": (B_pre_diff, parts["iso_pre_and_diffusion_TS"][0]),
        "vertmix_TS": (B_vmix, parts["vertmix_TS"][0]),
        "friction_UV": (B_fric, parts["friction_UV"][0]),
    }
    gbs = {k: b / (t * 1e-3) / 1e9 for k, (b, t) in kern.items()}
    # roofline: pre + the temperature + salinity diffusion in one C-ABI call,
    # the largest launch of the step (the wet-cell plan's three scan kernels,
    # the dry-column zero fill on a helper stream beside the wet-cell slopes +
    # face fluxes, the top-level K_33 zero, then the divergence + K33 solve +
    # updates over the wet cells)
    dom = "iso_pre_and_diffusion_TS"
    roof = {"bound": "hbm",
            "kernel": "vx_iso_pre_diffusion_ts_f64 = wet_scan_{count,chunks,columns}_kernel + "
                      "(iso_pre_dry_kernel || iso_pre_kernel<true> + k33_top_zero_kernel + iso_step_wet_kernel<2>)",
            "achieved": round(gbs[dom], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs[dom] / HBM_PEAK_GBS, 4),
            "traffic": PRE_TS_TRAFFIC_BYTES.get((nx, ny, nz)),
            "algorithmic_bytes_per_launch": kern[dom][0], "algorithmic_bytes_per_cell": PRE_TS_BYTES_PER_CELL,
            "avg_launch_ms": round(kern[dom][1], 4),
            "copy_GBps": round(copy_gbps, 1), "frac_of_copy": round(gbs[dom] / copy_gbps, 4)}
    out = {
        "metric": "TDMA Mcolumns/s + isoneutral Mcells/s; mixed vs FP64 GPU at <=1e-12 rel err",
        "value": round(N / (ms * 1e-3) / 1e6, 2),
        "unit": "Mcells/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"mixed-precision TDMA + isoneutral step, {nx}x{ny}x{nz} quarter-degree grid "
                               f"(35% land, kbot U{{5..80}})", "nx": nx, "ny": ny, "nz": nz,
                   "precision": prec, "refine_steps": args.refine, "tdma_wet_cells": W_td},
        "tdma_mcolumns_per_s": {k: round(v, 1) for k, v in rates["tdma_mcolumns_per_s"].items()},
        "iso_mcells_per_s": round(rates["iso_mcells_per_s"], 1),
        "iso_mcells_per_s_pre_then_TS": round(rates["iso_mcells_per_s_pre_then_TS"], 1),
        "mixed_over_fp64_tdma": round(parts["tdma_fp64"][0] / parts["tdma_mixed"][0], 3),
        "kernel_ms_median": {k: round(v[0], 4) for k, v in parts.items()},
        "kernel_algorithmic_GBps": {k: round(v, 1) for k, v in gbs.items()},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle on host cores)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(nx, ny, nz, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
