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
# profiles/mixing_step_1440x720x80.md (read bytes = 2.0 x FETCH_SIZE, write
# bytes = 1.0 x WRITE_SIZE, calibrated on a 1 GiB device copy), per layout of
# the time-level variables
# (the levels-innermost figure predates the segmented K33 solve of the step
# kernel and pre's unconditional neighbour loads)
PRE_TS_TRAFFIC_BYTES = {(1440, 720, 80, "outer"): 21_844_174_848, (1440, 720, 80, "inner"): 27_204_216_832}
# the minimum bytes per interior cell of pre + the temperature + salinity step
# in one call: K_iso read (8), Ai_* and K_11/22/33 written once and not read
# back (128 + 24), P_diss read and written (16), and per tracer tr[tau] (8),
# tr[taup1] read and written (16), dtr_iso read and written (16), int_drhodX (8)
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
    ap.add_argument("--levels", choices=["outer", "inner"], default="outer",
                    help="storage of the time-level variables (temp, salt, int_drhodT/S, u, v): levels outermost "
                         "(veros_amd.layout, the device state) or the reference's levels-innermost arrays")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--batched", action="store_true",
                    help="instead of the step: 64 independent uneven grids, sequential launches vs one batched "
                         "plan (untuned, autotuned, hipGraph), fp64 and mixed")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-grid batched leg")
    ap.add_argument("--only", default=None,
                    help="comma-separated parts to time (tdma_fp64, tdma_mixed, iso_pre, iso_diffusion_TS, "
                         "iso_diffusion_T_then_S, iso_pre_and_diffusion_TS, vertmix_TS, friction_UV) and "
                         "nothing else: one JSON line of their timings (for profiler runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy (a known-traffic dispatch for PMC scaling)")
    return ap.parse_args()


def build_workload(nx, ny, nz, seed=1002, levels="outer"):
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
    # TKE / EKE / IDEMIX inputs (veros_amd.workloads.add_turbulence_fields,
    # drawn on the device): energies (X, Y, nz, 3), scales, IDEMIX parameters
    # times maskW; the dissipation terms share four random fields
    kbx = vs["kbot"].long()
    kk = torch.arange(nz, device=dev)
    maskW = ((kbx[..., None] >= 1) & (kk[None, None, :] >= kbx[..., None] - 1)).to(torch.float64)
    del kbx
    u3 = lambda lo_, hi_: lo_ + (hi_ - lo_) * torch.rand((X, Y, nz), generator=g, **f64)  # noqa: E731
    vs["tke"] = 1e-5 + (1e-2 - 1e-5) * torch.rand((X, Y, nz, 3), generator=g, **f64)
    vs["sqrttke"] = torch.sqrt(vs["tke"][..., 0])
    vs["mxl"] = u3(1.0, 100.0)
    vs["forc_tke_surface"] = 2e-5 * torch.randn((X, Y), generator=g, **f64)
    vs["eke"] = 1e-4 + (1e-2 - 1e-4) * torch.rand((X, Y, nz, 3), generator=g, **f64)
    vs["sqrteke"] = torch.sqrt(vs["eke"][..., 0])
    vs["eke_len"] = u3(1e3, 1e5)
    vs["K_gm"] = u3(150.0, 450.0)
    vs["E_iw"] = -1e-5 + (1e-3 + 1e-5) * torch.rand((X, Y, nz, 3), generator=g, **f64)
    vs["dE_iw"] = torch.zeros((X, Y, nz, 3), **f64)
    vs["c0"] = u3(0.0, 1.0) * maskW
    vs["v0"] = u3(0.0, 0.5) * maskW
    vs["alpha_c"] = u3(0.0, 1e-3) * maskW
    vs["forc_iw_bottom"] = 1e-6 * torch.rand((X, Y), generator=g, **f64)
    vs["forc_iw_surface"] = 1e-6 * torch.rand((X, Y), generator=g, **f64)
    shared = [1e-9 * torch.randn((X, Y, nz), generator=g, **f64) * maskW for _ in range(4)]
    del maskW
    for q, k in enumerate(("K_diss_v", "P_diss_v", "P_diss_adv", "P_diss_nonlin", "eke_diss_tke", "K_diss_bot",
                           "K_diss_gm", "K_diss_h", "P_diss_skew", "P_diss_hmix")):
        vs[k] = shared[q % 4]
    for k in ("tke_diss", "eke_diss_iw", "iw_diss"):
        vs[k] = 1e-9 * torch.randn((X, Y, nz), generator=g, **f64)
    vs["eke_diss_tke_out"] = torch.zeros((X, Y, nz), **f64)
    vs["tke_surf_corr"] = torch.zeros((X, Y), **f64)
    from types import SimpleNamespace
    state = SimpleNamespace(variables=SimpleNamespace(**vs), settings=settings)
    if levels == "outer":
        from veros_amd.layout import LEVELS_OUTER, set_state_layout
        set_state_layout(state, LEVELS_OUTER)

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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


# windows of the 1440x720x80 state checked against the oracle: the west edge
# with its cyclic halo, the interior, the north-east corner (halo-inclusive
# index ranges; the oracle's cells [2:-2, 2:-2] of a window equal the whole
# grid's)
CHECK_WINDOWS = ((0, 100, 0, 84), (700, 800, 300, 384), (1344, 1444, 640, 724))
ISO_NAMES = ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by", "K_11", "K_22", "K_33", "temp", "salt", "dtemp_iso",
             "dsalt_iso", "P_diss_iso", "K_iso", "int_drhodT", "int_drhodS", "dxt", "dxu", "dyt", "dyu", "cost",
             "cosu", "dzt", "dzw", "zt")


def cpu_baseline(snap, budget_s):
    """The oracle (NumPy float64; LAPACK dgtsv for the TDMA) timed on the host
    cores on windows of the same 1440x720x80 state and on whole rows of the
    same TDMA columns, and the per-output error of the GPU results against
    it (max|x - x_ref| / max|x_ref| over the window cells [2:-2, 2:-2] / the
    sampled columns)."""
    import numpy as np

    from oracle import isoneutral as OI
    from oracle import tdma as OT

    def oracle_step(w):
        P0 = np.array(w.variables.P_diss_iso, copy=True)
        pre = OI.isoneutral_diffusion_pre(w)
        for k, v in pre._asdict().items():
            setattr(w.variables, k, v)
        OI.isoneutral_diffusion(w, "temp", True)
        Pt = np.array(w.variables.P_diss_iso, copy=True)
        OI.isoneutral_diffusion(w, "salt", False)
        return Pt - P0, w.variables.P_diss_iso - Pt

    def rel(x, r):
        s = np.abs(r).max()
        return float(np.abs(x - r).max() / s) if s > 0 else float(np.abs(x - r).max())

    errs = {}
    t_iso = 0.0
    n_iso = 0
    sl = np.s_[2:-2, 2:-2]
    for (before, after) in snap["iso_windows"]:
        nxw, nyw = before.settings.nx, before.settings.ny
        nz = before.settings.nz
        t0 = time.perf_counter()
        dPt, dPs = oracle_step(before)
        t_iso += time.perf_counter() - t0
        n_iso += nxw * nyw * nz
        ref = before.variables
        tau1 = ref.taup1
        for k in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by", "K_11", "K_22", "K_33", "temp", "salt", "dtemp_iso",
                  "dsalt_iso"):
            g, r = getattr(after, k), getattr(ref, k)
            if k in ("temp", "salt"):
                g, r = g[..., tau1], r[..., tau1]
            errs[k] = max(errs.get(k, 0.0), rel(g[sl], r[sl]))
        Pr = ref.P_diss_iso[sl]
        errs["P_diss_iso"] = max(errs.get("P_diss_iso", 0.0), rel(after.P_diss_iso[sl], Pr))
        scale = max(np.abs(Pr).max(), np.abs(dPt[sl]).max(), np.abs(dPs[sl]).max())
        errs["P_diss_iso_vs_largest_tracer_term"] = max(errs.get("P_diss_iso_vs_largest_tracer_term", 0.0),
                                                        float(np.abs(after.P_diss_iso[sl] - Pr).max() / scale))
    iso_rate = n_iso / t_iso / 1e6
    # TDMA: whole rows of columns, as many as about a third of the budget allows
    rows = snap["tdma_rows"]
    t0 = time.perf_counter()
    ref = OT.solve_implicit_kbot(rows["a"][:1], rows["b"][:1], rows["c"][:1], rows["d"][:1], rows["kbot"][:1])
    t1 = time.perf_counter() - t0
    nr = int(max(1, min(rows["a"].shape[0], (0.3 * budget_s) / max(t1, 1e-3))))
    t0 = time.perf_counter()
    ref = OT.solve_implicit_kbot(rows["a"][:nr], rows["b"][:nr], rows["c"][:nr], rows["d"][:nr],
                                 rows["kbot"][:nr])
    t_td = time.perf_counter() - t0
    n_td = nr * rows["a"].shape[1]
    td_rate = n_td / t_td / 1e6
    for prec, x in rows["x"].items():
        errs[f"tdma_x_{prec}"] = rel(x[:nr], ref)
    # more isoneutral cells for the timing, up to about two thirds of the budget
    extra = snap.get("iso_timing_window")
    if extra is not None and t_iso < 0.6 * budget_s:
        t0 = time.perf_counter()
        oracle_step(extra)
        t_iso += time.perf_counter() - t0
        n_iso += extra.settings.nx * extra.settings.ny * extra.settings.nz
        iso_rate = n_iso / t_iso / 1e6
    nx, ny, nz = snap["shape"]
    N = nx * ny * nz
    t_step = N / (iso_rate * 1e6) + (nx * ny) / (td_rate * 1e6)
    try:
        import scipy
        scipy_v = scipy.__version__
    except ImportError:
        scipy_v = None
    base = dict(value=round(N / t_step / 1e6, 4), unit="Mcells/s", cores=1, kind="port",
                sample=f"oracle (NumPy float64, LAPACK dgtsv) on {len(snap['iso_windows'])} windows of the "
                       f"{nx}x{ny}x{nz} state plus one timing window ({n_iso} isoneutral cells, {t_iso:.1f} s) "
                       f"and {nr} whole rows of the TDMA columns ({n_td} columns, {t_td:.1f} s)",
                tdma_mcolumns_per_s=round(td_rate, 4), iso_mcells_per_s=round(iso_rate, 4),
                host=dict(cpu_count=os.cpu_count(), affinity=len(os.sched_getaffinity(0)), cpu=cpu_model(),
                          numpy=np.__version__, scipy=scipy_v),
                jax_cpu="n/a (JAX is not installed on the host)")
    return base, errs


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


def timed_events(fn, reps, stream):
    """(median, min) ms of ``reps`` calls, each bracketed by HIP events on
    ``stream`` (the stream the entry points enqueue on)."""
    import torch
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for s0, e0 in evs:
        s0.record(stream)
        fn()
        e0.record(stream)
    torch.cuda.synchronize()
    ts = sorted(a.elapsed_time(b) for a, b in evs)
    return ts[len(ts) // 2], ts[0]


def batched_leg(args, stream):
    """The 64-grid batched workload: 64 sequential single-grid launches against
    one plan (one persistent launch per depth bucket), untuned, tuned and
    replayed as a hipGraph, for both precision forms."""
    import torch

    from veros_amd import utilities
    from veros_amd.batched import TdmaBatch, random_systems

    systems = random_systems(64)
    torch.cuda.synchronize()
    cols = sum(s["a"].shape[0] * s["a"].shape[1] for s in systems)
    cells = sum(s["a"].numel() for s in systems)
    nbytes = sum(tdma_bytes(s["kbot"], s["a"].shape[2])[0] for s in systems)
    reps = max(5, min(args.steps, 10))
    res = {}
    for prec in ("fp64", "mixed"):
        def sequential():
            for s0 in systems:
                utilities.solve_implicit_kbot(s0["a"], s0["b"], s0["c"], s0["d"], s0["kbot"], precision=prec,
                                              refine_steps=1, out=s0["x"])
        plan = TdmaBatch(systems, prec)
        for fn in (sequential, plan.run):
            for _ in range(args.warmup):
                fn()
        r = {"sequential_64_launches_ms": timed_events(sequential, reps, stream)[0],
             "plan_untuned_ms": timed_events(plan.run, reps, stream)[0]}
        info = plan.tune()
        r["plan_tuned_ms"] = timed_events(plan.run, reps, stream)[0]
        plan.graph_launch()
        r["plan_graph_ms"] = timed_events(plan.graph_launch, reps, stream)[0]
        r["tuned_choices"] = info["buckets"]
        best = min(r["plan_tuned_ms"], r["plan_graph_ms"])
        r["speedup_vs_sequential"] = r["sequential_64_launches_ms"] / best
        r["mcolumns_per_s"] = cols / (best * 1e-3) / 1e6
        r["algorithmic_GBps"] = nbytes / (best * 1e-3) / 1e9
        res[prec] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        plan.close()
        print(f"[bench] batched {prec}: {res[prec]['plan_graph_ms']} ms graph, "
              f"{res[prec]['sequential_64_launches_ms']} ms sequential", file=sys.stderr, flush=True)
    del systems
    torch.cuda.empty_cache()
    return {"workload": "64 independent grids, nx,ny ~ U{48..720}, nz ~ U{15..100}, kbot U{1..nz}, 30% land",
            "columns": cols, "cells": cells, "algorithmic_bytes": nbytes, "results": res}


def run_batched(args):
    import torch
    torch.cuda.set_device(0)
    print(json.dumps(batched_leg(args, torch.cuda.current_stream())))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus above 1 is not supported (the TDMA columns would "
                                   "shard with no communication, the stencils by 2-D blocks with a 2-cell "
                                   "halo exchange of T and S per step; not built)"}))
        sys.exit(2)
    if args.batched:
        return run_batched(args)
    import torch

    from veros_amd import friction, isoneutral, thermodynamics, utilities
    from veros_amd import workloads as W
    from veros_amd.state import to_device_state

    torch.cuda.set_device(0)
    nx, ny, nz = args.nx, args.ny, args.nz
    N = nx * ny * nz
    X, Y = nx + 4, ny + 4
    print(f"[bench] building {nx}x{ny}x{nz} workload", file=sys.stderr, flush=True)
    state, td = build_workload(nx, ny, nz, levels=args.levels)
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
    copy_ms = timed_events(lambda: dst.copy_(src), 5, stream)[0]
    copy_gbps = 2 * src.numel() * 8 / (copy_ms * 1e-3) / 1e9
    del src, dst

    def run_tdma(prec, out=None):
        utilities.solve_implicit_kbot(td["a"], td["b"], td["c"], td["d"], td["kbot"], precision=prec,
                                      refine_steps=args.refine, out=td["x"] if out is None else out)

    # ---- results checked against the oracle (cpu_baseline leg): inputs of
    # windows of the state before the one call, its outputs after it, and
    # whole rows of the TDMA columns solved by both forms
    snap = None
    if not args.no_cpu_baseline:
        ww = [w for w in CHECK_WINDOWS if w[1] <= X and w[3] <= Y]
        fm = W.whole_grid_masks(state)
        before = [W.window_state(state, *w, names=ISO_NAMES, full_masks=fm) for w in ww]
        # the timing window: 364 x 196 columns next to the first check window
        tw = (100, 464, 0, min(Y, 200))
        timing = W.window_state(state, *tw, names=ISO_NAMES, full_masks=fm) if tw[1] <= X else None
        del fm
        isoneutral.isoneutral_pre_and_diffusion_temp_salt(state)
        torch.cuda.synchronize()
        after = [W.window_state(state, *w, masks=False, names=ISO_NAMES).variables for w in ww]
        rows = torch.arange(0, min(nx, 64), device="cuda")
        xs = {}
        for prec in ("fp64", "mixed"):
            x = torch.empty_like(td["d"])
            run_tdma(prec, out=x)
            xs[prec] = x.index_select(0, rows).cpu().numpy()
            del x
        snap = dict(iso_windows=list(zip(before, after)), iso_timing_window=timing, shape=(nx, ny, nz),
                    tdma_rows=dict({k: td[k].index_select(0, rows).cpu().numpy()
                                    for k in ("a", "b", "c", "d", "kbot")}, x=xs))
        # back to the workload's initial state for the timings
        del state, td, vs
        torch.cuda.empty_cache()
        state, td = build_workload(nx, ny, nz, levels=args.levels)
        vs = state.variables
        torch.cuda.synchronize()

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

    from veros_amd.workloads import turbulence_settings
    turb_settings = {
        # TKE with IDEMIX and EKE on (forcing K_diss_v - P_diss_v - P_diss_adv - P_diss_nonlin + eke_diss_tke +
        # iw_diss) and lateral diffusion; EKE as veros runs it; IDEMIX with the EKE dissipation moved to the
        # bottom cell and lateral diffusion
        "tke": turbulence_settings(nz=nz, enable_eke=True, enable_idemix=True, enable_tke_hor_diffusion=True),
        "eke": turbulence_settings(nz=nz, enable_eke=True),
        "idemix": turbulence_settings(nz=nz, enable_eke=True, enable_eke_diss_bottom=True,
                                      enable_idemix_hor_diffusion=True),
    }

    def run_turb(model):
        from types import SimpleNamespace
        from veros_amd import turbulence
        st_ = SimpleNamespace(variables=vs, settings=turb_settings[model])
        fn = {"tke": turbulence.integrate_tke_kernel, "eke": turbulence.integrate_eke_kernel,
              "idemix": turbulence.integrate_idemix_kernel}[model]
        if model == "eke":  # its eke_diss_tke output is a separate buffer from TKE's input
            saved = vs.eke_diss_tke
            vs.eke_diss_tke = vs.eke_diss_tke_out
            fn(st_)
            vs.eke_diss_tke = saved
        else:
            fn(st_)

    def run_diff_two_calls():
        isoneutral.isoneutral_diffusion(state, vs.temp, True)
        isoneutral.isoneutral_diffusion(state, vs.salt, False)

    # per-part timings (median, min) in ms
    parts = {}
    only = None if args.only is None else set(args.only.split(","))
    for name, fn in (("tdma_fp64", lambda: run_tdma("fp64")), ("tdma_mixed", lambda: run_tdma("mixed")),
                     ("iso_pre", run_pre), ("iso_diffusion_TS", run_diff),
                     ("iso_diffusion_T_then_S", run_diff_two_calls), ("iso_pre_and_diffusion_TS", run_pre_diff),
                     ("vertmix_TS", run_vertmix), ("friction_UV", run_friction),
                     ("tke", lambda: run_turb("tke")), ("eke", lambda: run_turb("eke")),
                     ("idemix", lambda: run_turb("idemix"))):
        if only is not None and name not in only:
            continue
        for _ in range(args.warmup):
            fn()
        parts[name] = timed_events(fn, max(5, args.steps), stream)
        print(f"[bench] {name}: median {parts[name][0]:.3f} ms, min {parts[name][1]:.3f} ms",
              file=sys.stderr, flush=True)

    if only is not None:
        print(json.dumps({"parts_ms_median": {k: round(v[0], 4) for k, v in parts.items()},
                          "config": {"nx": nx, "ny": ny, "nz": nz}}))
        return

    def step_of(prec):
        def step():
            run_tdma(prec)
            run_pre_diff()
        return step

    # the timed region: the configured precision form's step, K steps
    prec = args.precision
    step = step_of(prec)
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
    step_ms = {prec: ms}
    other = "fp64" if prec == "mixed" else "mixed"
    step_ms[other] = timed_events(step_of(other), max(5, args.steps), stream)[0]
    print(f"[bench] step: {step_ms}", file=sys.stderr, flush=True)

    B_td, W_td = tdma_bytes(td["kbot"], nz)
    C = nx * ny
    cols_halo = X * Y
    # algorithmic bytes per interior cell with each call counted on its own: pre 176 (T, S, K_iso
    # read; Ai_* and K_11/22/33 written), each tracer's isoneutral_diffusion
    # 200 (tr[tau], tr[taup1], K_iso, Ai_*, K_*, dtr read; tr[taup1], dtr
    # written) + 24 with the energy branch; 4 B of kbot per column and kernel
    B8_pre = 176 * N + 4 * cols_halo
    B8_tracer = (200 + 24) * N + 4 * cols_halo
    B8_step = B8_pre + 2 * B8_tracer
    # the one call's own minimum (the tensor written once and never re-read)
    B_pre_diff = PRE_TS_BYTES_PER_CELL * N + 4 * cols_halo
    # vertmix_tempsalt: kappaH read and temp/salt[taup1] read and written on the
    # wet cells (8 + 2 * 16), the two tendencies written on every cell
    # (2 * 8 over X*Y*nz), kbot and the two forcings per column (4 + 16)
    _, W_v = tdma_bytes(vs.kbot[2:-2, 2:-2], nz)
    B_vmix = 40 * W_v + 16 * cols_halo * nz + 20 * cols_halo
    # implicit_vert_friction: per cell kappaM read (8), u and v at tau read and
    # at taup1 written or kept (4 * 8), du_mix and dv_mix written (16),
    # K_diss_v read and written (16), over the X*Y*nz array
    B_fric = 72 * cols_halo * nz
    # TKE / EKE / IDEMIX, over the X*Y*nz array, 8 B per array and cell: TKE
    # reads kappaM, sqrttke, mxl, tke[tau], tke[taup1] and 6 forcing terms and
    # writes tke[taup1], tke_diss (13); EKE reads kappaM, sqrteke, eke_len,
    # K_gm, eke[tau], eke[taup1] and 5 terms and writes eke[taup1],
    # eke_diss_iw, eke_diss_tke (14); IDEMIX reads c0, v0, alpha_c, E_iw[tau],
    # E_iw[taup1] and 2 terms and writes E_iw[taup1], iw_diss (9)
    B_turb = {"tke": 13 * 8 * cols_halo * nz, "eke": 14 * 8 * cols_halo * nz, "idemix": 9 * 8 * cols_halo * nz}
    kern = {
        "tdma_fp64": (B_td, parts["tdma_fp64"][0]),
        "tdma_mixed": (B_td, parts["tdma_mixed"][0]),
        "iso_pre": (B8_pre, parts["iso_pre"][0]),
        "iso_diffusion_TS": (2 * B8_tracer, parts["iso_diffusion_TS"][0]),
        "iso_diffusion_T_then_S": (2 * B8_tracer, parts["iso_diffusion_T_then_S"][0]),
        "iso_pre_and_diffusion_TS": (B8_step, parts["iso_pre_and_diffusion_TS"][0]),
        "vertmix_TS": (B_vmix, parts["vertmix_TS"][0]),
        "friction_UV": (B_fric, parts["friction_UV"][0]),
        **{m: (B_turb[m], parts[m][0]) for m in ("tke", "eke", "idemix")},
    }
    gbs = {k: b / (t * 1e-3) / 1e9 for k, (b, t) in kern.items()}
    # roofline: pre + the temperature + salinity diffusion in one C-ABI call,
    # the largest launch of the step
    dom = "iso_pre_and_diffusion_TS"
    t_dom = parts[dom][0]
    traffic = PRE_TS_TRAFFIC_BYTES.get((nx, ny, nz, args.levels))
    roof = {"bound": "hbm",
            "kernel": "vx_iso_pre_diffusion_ts_f64 = wet_scan_{count,chunks,columns}_kernel + "
                      "(iso_pre_dry_kernel || iso_pre_kernel<true> + k33_top_zero_kernel + iso_step_wet_kernel<2>)",
            "achieved": round(gbs[dom], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs[dom] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over this call on this workload "
                              f"with the time levels {args.levels}most (read = 2.0 x FETCH_SIZE, write = 1.0 x "
                              "WRITE_SIZE, factors from a 1 GiB copy); profiles/mixing_step_1440x720x80.md",
            "algorithmic_bytes_per_launch": B8_step,
            "algorithmic_bytes_per_cell": 176 + 2 * 224,
            "avg_launch_ms": round(t_dom, 4),
            "alternative_accountings": {
                "fused_minimum": {"bytes_per_cell": PRE_TS_BYTES_PER_CELL, "bytes": B_pre_diff,
                                  "achieved": round(B_pre_diff / (t_dom * 1e-3) / 1e9, 1),
                                  "frac": round(B_pre_diff / (t_dom * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
                "measured_traffic": None if traffic is None else {
                    "bytes": traffic, "achieved": round(traffic / (t_dom * 1e-3) / 1e9, 1),
                    "frac": round(traffic / (t_dom * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
            "copy_GBps": round(copy_gbps, 1), "frac_of_copy": round(gbs[dom] / copy_gbps, 4),
            "note": "achieved/frac count the per-cell bytes of pre + T + S as separate calls (the tensor written "
                    "once and read back for each tracer); the one call never reads the tensor back, so they can "
                    "exceed the peak: fused_minimum (the call's own minimum bytes) and measured_traffic (HBM "
                    "counters) give the bytes it moves"}

    # the 360x160x60 ACC-like channel (FP64), isoneutral one call
    st1 = to_device_state(W.iso_config1())
    for _ in range(args.warmup):
        isoneutral.isoneutral_pre_and_diffusion_temp_salt(st1)
    t1_ms = timed_events(lambda: isoneutral.isoneutral_pre_and_diffusion_temp_salt(st1), max(5, args.steps),
                         stream)[0]
    N1 = 360 * 160 * 60
    del st1

    batched = None
    if not args.no_batched:
        del state, td, vs
        torch.cuda.empty_cache()
        batched = batched_leg(args, stream)

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
        "step_ms": {k: round(v, 4) for k, v in step_ms.items()},
        "step_mcells_per_s": {k: round(N / (v * 1e-3) / 1e6, 1) for k, v in step_ms.items()},
        "tdma_mcolumns_per_s": {p: round(C / (parts[f"tdma_{p}"][0] * 1e-3) / 1e6, 1) for p in ("fp64", "mixed")},
        "tdma_roofline_frac": {p: round(gbs[f"tdma_{p}"] / HBM_PEAK_GBS, 4) for p in ("fp64", "mixed")},
        "mixed_over_fp64_tdma": round(parts["tdma_fp64"][0] / parts["tdma_mixed"][0], 3),
        "iso_mcells_per_s": round(N / (t_dom * 1e-3) / 1e6, 1),
        "iso_mcells_per_s_pre_then_TS": round(N / ((parts["iso_pre"][0] + parts["iso_diffusion_TS"][0]) * 1e-3)
                                              / 1e6, 1),
        "config1_iso_fp64": {"workload": "360x160x60 ACC-like channel, one call pre + T + S", "ms": round(t1_ms, 4),
                             "mcells_per_s": round(N1 / (t1_ms * 1e-3) / 1e6, 1)},
        "kernel_ms_median": {k: round(v[0], 4) for k, v in parts.items()},
        "kernel_algorithmic_GBps": {k: round(v, 1) for k, v in gbs.items()},
        "roofline": roof,
    }
    if batched is not None:
        out["batched"] = batched
    if snap is not None:
        print("[bench] CPU baseline and per-output errors (oracle on host cores)", file=sys.stderr, flush=True)
        base, errs = cpu_baseline(snap, args.cpu_budget)
        out["cpu_baseline"] = base
        out["max_rel_err_vs_fp64_oracle"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
