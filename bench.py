#!/usr/bin/env python3
"""Benchmark of the D3Q27 step behind sim_tcpc on one MI355X.

Workload (N=1): the synthetic TCPC junction 256x192x128 (IVC R=20 from x-min,
SVC R=16 from x-max, axes 6 cells apart in y, pulmonary R=14 along y to the
label-5/6 outlets), cumulant collision, FP32, Carreau-Yasuda set A
(mu0=0.056 Pa s, mu_inf=0.00345 Pa s, lambda=3.313 s, n=0.3568, a=2,
rho=1060 kg/m^3, dx=0.25 mm) with the local non-equilibrium viscosity,
parabolic inflows u_max 0.04 (IVC) / 0.03 (SVC), outlets at rho=1, lattice
viscosity nu_inf = 0.01.  A step is one time step of the whole case; the
timed region is K consecutive steps with the state resident in HBM.

value = fluid MLUPS = non-wall cells x K / time of the K steps.  The fused
kernel's roofline is its algorithmic bytes (2 x 27 x 4 B per non-wall cell)
over its average launch duration (HIP events around back-to-back launches on
the case's stream), against 8 TB/s; the same kernel on the 2x
scaled junction (512x384x256, 2.9M fluid cells, ~0.7 GB of populations,
well above the 256 MB Infinity Cache) is reported as the HBM-resident line.
Also reported: each collision model on the same case, a 15-junction batch
against the same cases one after another, the first --sweep-requests
requests of the 200-request serving stream (Poisson arrivals at 2/s, a
240 GB budget) against plain FP32 service of the same requests, and the
64x32x32 FP64 channel for its full 1,000 steps on the GPU and the CPU, and
the storage precision study: the junction scaled to 512x384x256 with
pulsatile inflow (A = 0.3, T = 4,000 steps), cumulant + CY-A, 40,000 steps
stored as FP64, FP32, FP16- and BF16-offset, each field's error against the
FP64 run (about 30 s of GPU time); and each collision model at FP32 against
its FP64 run after 20,000 steps of the 256x192x128 junction (cfg1_parity).
The
CPU baseline is the oracle (a dense A/B OpenMP restatement of the
reference's Host build) on --cpu-steps steps of the benchmark case, with the
thread count used and lscpu's physical core count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "fluid MLUPS & %HBM roofline (D3Q27 cumulant+CY); sweep p50/p99 case latency"
PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"


def profiled_traffic(name: str) -> int | None:
    """HBM bytes per k_step launch from the rocprofv3 PMC summary of this
    workload (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel factors),
    the line 'traffic_bytes_per_launch: N' of profiles/<name>.md."""
    path = PROFILES / f"{name}.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.strip().startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1].strip())
    return None


def make_case(scale: float, storage: int, collision: int, lbm_viscosity: float, cy_model: int | None = None):
    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.geometry import tcpc_mask
    from nonnewlbm_amd.materials import case_kwargs, material_by_name

    g = tcpc_mask(scale)
    mat = material_by_name("cy-a")
    kw = case_kwargs(mat, g.dl, lbm_viscosity, cy_model=cy_model)
    kw.update(storage=storage, collision=collision, profile=_abi.PROFILE_PARABOLIC, ivc_umax=0.04,
              svc_umax=0.03, ivc_center=g.meta["ivc"]["center"], svc_center=g.meta["svc"]["center"],
              ivc_radius=g.meta["ivc"]["radius"], svc_radius=g.meta["svc"]["radius"], stats_every=1)
    return g, kw


FULL_GRIDS = ((128, 128, 128), (192, 144, 96), (256, 192, 128), (384, 288, 192), (512, 384, 256))
STREAM_BUDGET = 240 * 10**9


def stream_requests(n: int = 200, rate: float = 2.0, seed: int = 0) -> list[dict]:
    """The serving stream: Poisson arrivals at `rate` per second; IVC radius
    R ~ U[10, 28] cells (SVC 0.8 R, pulmonary 0.7 R, the 20/16/14 ratios of the
    256x192x128 junction), caval offset ~ U[0, 12] cells, grid from FULL_GRIDS
    (128^3 .. 512x384x256, dl = 64 mm / X), collision in BGK/MRT/KBC/CUM,
    material in Newtonian/CY-A/CY-B, FP32 or FP16-offset storage,
    5,000-50,000 steps, parabolic u_max ~ U[0.01, 0.06] (numpy
    default_rng(seed)).  The first k requests of the stream are the same for
    every n >= k."""
    from nonnewlbm_amd import _abi
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate))
        out.append(dict(arrival=t, radius=float(rng.uniform(10.0, 28.0)), offset=float(rng.uniform(0.0, 12.0)),
                        grid=FULL_GRIDS[int(rng.integers(len(FULL_GRIDS)))],
                        collision=int(rng.choice([_abi.BGK, _abi.MRT, _abi.KBC, _abi.CUM])),
                        material=str(rng.choice(["blood-newtonian", "cy-a", "cy-b"])),
                        storage=int(rng.choice([_abi.F32, _abi.F16_OFFSET])),
                        steps=int(rng.integers(5000, 50001)),
                        umax=(float(rng.uniform(0.01, 0.06)), float(rng.uniform(0.01, 0.06)))))
    return out


def stream_case(req: dict, storage=None):
    """(int8 x-fastest wall mask, dims, case keywords) of a stream request."""
    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.geometry import tcpc_mask
    from nonnewlbm_amd.materials import case_kwargs, material_by_name
    X, Y, Z = req["grid"]
    R = req["radius"]
    g = tcpc_mask(1.0, r_ivc=R, r_svc=0.8 * R, r_pa=0.7 * R, offset=req["offset"], dims=(X, Y, Z), dl=0.064 / X)
    kw = case_kwargs(material_by_name(req["material"]), g.dl, 0.01)
    kw.update(storage=req["storage"] if storage is None else storage, collision=req["collision"],
              profile=_abi.PROFILE_PARABOLIC, ivc_umax=req["umax"][0], svc_umax=req["umax"][1],
              ivc_center=g.meta["ivc"]["center"], svc_center=g.meta["svc"]["center"],
              ivc_radius=g.meta["ivc"]["radius"], svc_radius=g.meta["svc"]["radius"], stats_every=1)
    return g.flat().astype(np.int8), g.dims, kw


def _pcts(a: np.ndarray) -> dict:
    return {k: round(float(np.percentile(a, q)), 3) for k, q in (("p50", 50), ("p95", 95), ("p99", 99))} | \
        {"max": round(float(a.max()), 3)}


def sweep(n: int, rate: float, tick: int, budget: int = STREAM_BUDGET, graph: bool = True) -> dict:
    """The first n requests of the serving stream arrive on their Poisson
    schedule into one Pool (continuous batching, `budget` bytes of device
    memory); latency = completion - submission on the pool clock.

    Baseline: plain FP32 service of the same n requests, one case at a time
    with the single-case kernel.  Each request's FP32 run is timed on the GPU;
    a FIFO queue over the same arrival times then gives the sequential
    service's latencies and makespan (request k starts at max(arrival_k,
    end_{k-1})).  Aggregate fluid MLUPS = fluid-cell updates of all n
    requests / makespan (first arrival to last completion); "busy" divides by
    the seconds the pool spent stepping instead."""
    from concurrent.futures import ThreadPoolExecutor

    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.lbm355 import Case, Pool
    reqs = stream_requests(n, rate)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=8) as ex:
        specs = list(ex.map(stream_case, reqs))
    print(f"# sweep: {n} geometries in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    pool = Pool(budget_bytes=budget, steps_per_tick=tick, graph=graph)
    t0 = time.perf_counter()
    ids = []
    for r, (flat, dims, kw) in zip(reqs, specs):
        wait = r["arrival"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ids.append(pool.submit(flat, dims, r["steps"], report_every=1000, **kw))
    while not pool.drain(timeout=20.0):
        info = pool.stats()
        print(f"# sweep: {info['done']} done, {info['running']} running, {info['queued']} queued, "
              f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    sts = [pool.poll(i) for i in ids]
    info = pool.stats()
    pool.close()
    done = [s for s in sts if s["state_name"] == "done"]
    lat = np.array([s["t_done"] - s["t_submit"] for s in done])
    first = np.array([s["t_first_report"] - s["t_submit"] for s in done])
    span = max(s["t_done"] for s in done) - min(s["t_submit"] for s in done)
    updates = info["cell_updates"]
    print(f"# sweep: {len(done)}/{n} done, p50 {np.percentile(lat, 50):.2f} s, p99 {np.percentile(lat, 99):.2f} s, "
          f"makespan {span:.1f} s, {updates / span / 1e6:.0f} fluid MLUPS over the makespan", file=sys.stderr,
          flush=True)
    # plain FP32 service, one request at a time
    seq_t = []
    seq_updates = 0
    for k, (r, (flat, dims, kw)) in enumerate(zip(reqs, specs)):
        c = Case(flat, dims, **dict(kw, storage=_abi.F32))
        c.sync()
        t1 = time.perf_counter()
        c.step(r["steps"])
        c.sync()
        seq_t.append(time.perf_counter() - t1)
        seq_updates += c.info()["active_cells"] * r["steps"]
        c.close()
        if (k + 1) % 16 == 0:
            print(f"# sweep: sequential FP32 {k + 1}/{n}", file=sys.stderr, flush=True)
    end = 0.0
    seq_lat = []
    for r, dt in zip(reqs, seq_t):
        end = max(end, r["arrival"]) + dt
        seq_lat.append(end - r["arrival"])
    seq_span = end - reqs[0]["arrival"]
    seq_busy = sum(seq_t)
    mk = updates / span / 1e6
    seq_mk = seq_updates / seq_span / 1e6
    seq_rate = seq_updates / seq_busy / 1e6
    return {
        "workload": f"the first {n} requests of the 200-request serving stream: TCPC junctions, IVC R ~ U[10,28], "
                    "caval offset ~ U[0,12], grids 128^3 .. 512x384x256, BGK/MRT/KBC/CUM, Newtonian/CY-A/CY-B, "
                    "FP32/FP16-offset, parabolic u_max U[0.01,0.06], 5000-50000 steps, Poisson arrivals",
        "requests": n, "done": len(done), "failed": n - len(done), "arrival_rate_per_s": rate, "seed": 0,
        "steps_per_tick": tick, "budget_bytes": info["budget_bytes"],
        "last_arrival_s": round(reqs[-1]["arrival"], 3),
        "latency_s": _pcts(lat),
        "first_report_s": _pcts(first),
        "makespan_s": round(span, 3),
        "fluid_mlups_makespan": round(mk, 1),
        "fluid_mlups_busy": round(updates / info["busy_seconds"] / 1e6, 1),
        "sequential_fp32": {"latency_s": _pcts(np.array(seq_lat)), "makespan_s": round(seq_span, 3),
                            "fluid_mlups_makespan": round(seq_mk, 1), "fluid_mlups_busy": round(seq_rate, 1),
                            "gpu_seconds": round(seq_busy, 3)},
        "speedup_makespan_mlups": round(mk / seq_mk, 3),
        "speedup_vs_sequential_busy_mlups": round(mk / seq_rate, 3),
    }


def batch15(steps: int, graph: bool = True, small: bool = True) -> dict:
    """Fifteen independent junctions stepped as one batch (one launch per
    step over their concatenated brick lists, graph replay) against the same
    cases stepped one after another, each with the plain single-case launch:
    5 parabolic inflow pairs u_max ~ U[0.01, 0.06] x {Newtonian 0.0035 Pa s,
    CY-A, CY-B}, grid scale per case from {0.5, 0.75, 1.0, 1.5}
    (numpy default_rng(0)), cumulant, FP32."""
    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.geometry import tcpc_mask
    from nonnewlbm_amd.lbm355 import Batch, Case
    from nonnewlbm_amd.materials import case_kwargs, material_by_name
    rng = np.random.default_rng(0)
    pairs = [(float(rng.uniform(0.01, 0.06)), float(rng.uniform(0.01, 0.06))) for _ in range(5)]
    geoms: dict = {}
    specs = []
    for umax in pairs:
        for mat in ("blood-newtonian", "cy-a", "cy-b"):
            sc = float(rng.choice([0.5, 0.75, 1.0, 1.5]))
            if sc not in geoms:
                geoms[sc] = tcpc_mask(sc)
            g = geoms[sc]
            kw = case_kwargs(material_by_name(mat), g.dl, 0.01)
            kw.update(storage=_abi.F32, collision=_abi.CUM, profile=_abi.PROFILE_PARABOLIC, ivc_umax=umax[0],
                      svc_umax=umax[1], ivc_center=g.meta["ivc"]["center"], svc_center=g.meta["svc"]["center"],
                      ivc_radius=g.meta["ivc"]["radius"], svc_radius=g.meta["svc"]["radius"], stats_every=1)
            specs.append((sc, g, kw))
    # sequential: one case at a time, plain launches
    seq_s, cells = 0.0, 0
    for sc, g, kw in specs:
        c = Case(g.flat(), g.dims, **kw)
        c.step(2)
        c.sync()
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        seq_s += time.perf_counter() - t0
        cells += c.info()["active_cells"]
        c.close()
    # batched
    cases = [Case(g.flat(), g.dims, **kw) for sc, g, kw in specs]
    b = Batch(cases, graph=graph)
    b.step(2)
    t0 = time.perf_counter()
    b.step(steps)
    bat_s = time.perf_counter() - t0
    # the same batch as one group: every step sweeps all fifteen cases
    b.set_group_bytes(0)
    b.step(2)
    t0 = time.perf_counter()
    b.step(steps)
    one_s = time.perf_counter() - t0
    finite = all(bool(np.isfinite(c.field("vx")).all()) for c in cases)
    b.close()
    for c in cases:
        c.close()
    seq = cells * steps / seq_s / 1e6
    bat = cells * steps / bat_s / 1e6
    print(f"# batch of 15: {bat:.0f} fluid MLUPS batched, {seq:.0f} sequential ({bat / seq:.2f}x)", file=sys.stderr,
          flush=True)
    out = {"workload": "15 TCPC junctions (5 parabolic u_max pairs ~ U[0.01,0.06] x Newtonian/CY-A/CY-B, scale "
                        "per case from 0.5/0.75/1.0/1.5, default_rng(0)), cumulant, FP32",
            "scales": [sc for sc, g, kw in specs], "fluid_cells": cells, "steps": steps,
            "batched_fluid_mlups": round(bat, 1), "sequential_fluid_mlups": round(seq, 1),
            "speedup": round(bat / seq, 3),
            "one_group_fluid_mlups": round(cells * steps / one_s / 1e6, 1),
            "grouping": "consecutive members up to 256 MiB of device memory per group, 64 steps per group in turn",
            "results_finite": finite, "graph": graph}
    if small:
        out["small_cases"] = small_batch(steps, graph)
    return out


def small_batch(steps: int, graph: bool = True) -> dict:
    """Sixteen 128x96x64 junctions (44,420 fluid cells each, CY-A, cumulant,
    FP32, u_max pairs ~ U[0.01, 0.06], default_rng(1)): one such case leaves
    most of the GPU idle, a batch fills it."""
    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.geometry import tcpc_mask
    from nonnewlbm_amd.lbm355 import Batch, Case
    from nonnewlbm_amd.materials import case_kwargs, material_by_name
    rng = np.random.default_rng(1)
    g = tcpc_mask(0.5)
    specs = []
    for _ in range(16):
        kw = case_kwargs(material_by_name("cy-a"), g.dl, 0.01)
        kw.update(storage=_abi.F32, collision=_abi.CUM, profile=_abi.PROFILE_PARABOLIC,
                  ivc_umax=float(rng.uniform(0.01, 0.06)), svc_umax=float(rng.uniform(0.01, 0.06)),
                  ivc_center=g.meta["ivc"]["center"], svc_center=g.meta["svc"]["center"],
                  ivc_radius=g.meta["ivc"]["radius"], svc_radius=g.meta["svc"]["radius"], stats_every=1)
        specs.append(kw)
    cases = [Case(g.flat(), g.dims, **kw) for kw in specs]
    cells = sum(c.info()["active_cells"] for c in cases)
    seq_s = 0.0
    for c in cases:
        c.step(2)
        c.sync()
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        seq_s += time.perf_counter() - t0
    b = Batch(cases, graph=graph)
    b.step(2)
    t0 = time.perf_counter()
    b.step(steps)
    bat_s = time.perf_counter() - t0
    b.close()
    for c in cases:
        c.close()
    seq = cells * steps / seq_s / 1e6
    bat = cells * steps / bat_s / 1e6
    print(f"# batch of 16 small: {bat:.0f} fluid MLUPS batched, {seq:.0f} sequential ({bat / seq:.2f}x)",
          file=sys.stderr, flush=True)
    return {"workload": "16 TCPC junctions 128x96x64, CY-A, cumulant, FP32", "fluid_cells": cells, "steps": steps,
            "batched_fluid_mlups": round(bat, 1), "sequential_fluid_mlups": round(seq, 1),
            "speedup": round(bat / seq, 3)}


def models(steps: int) -> dict:
    """The 256x192x128 junction workload with each collision model (CY-A neq,
    FP32, statistics every step), autotuned launch."""
    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.lbm355 import Case
    out = {}
    for name, coll in (("BGK", _abi.BGK), ("MRT_LES", _abi.MRT), ("KBC_N4", _abi.KBC), ("CUM", _abi.CUM)):
        g, kw = make_case(1.0, _abi.F32, coll, 0.01)
        c = Case(g.flat(), g.dims, **kw)
        c.step(50)
        bpg, nt, cap = c.autotune()
        c.sync()
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        dt = time.perf_counter() - t0
        active = c.info()["active_cells"]
        out[name] = {"fluid_mlups": round(active * steps / dt / 1e6, 1), "step_us": round(dt / steps * 1e6, 2),
                     "launch": [bpg, nt, cap]}
        c.close()
    print(f"# models: {out}", file=sys.stderr, flush=True)
    return out


def precision_study(scale: float, steps: int, collision: int | None = None, storages=None,
                    pulse_amp: float = 0.3) -> dict:
    """Storage precision against FP64 on the pulsatile junction: cumulant (or
    ``collision``) + CY-A (neq), parabolic inflows 0.04/0.03 with u(t) = u (1 +
    0.3 sin(2 pi t / 4000)) (steady with pulse_amp = 0), FP64 statistics.
    Relative L2 over the non-wall cells of u and of the output-path gamma_dot
    and wss_proxy (the VTK formulas, sim_tcpc.cu:697-714), and the LPA share
    of the mean outlet flow."""
    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.lbm355 import Case
    from nonnewlbm_amd.materials import material_by_name
    mat = material_by_name("cy-a")
    model = _abi.CUM if collision is None else collision
    name = {v: k for k, v in _abi.COLLISIONS.items() if k in ("cumulant", "bgk", "mrt_les", "kbc_n4")}[model]
    inflow = f"pulsatile (A={pulse_amp}, T=4000 steps)" if pulse_amp else "steady parabolic inflow"
    res: dict = {"workload": f"TCPC junction x{scale} {inflow}, {name} + CY-A neq, {steps} steps",
                 "bar": "u rel L2 <= 1e-3 and LPA split within 0.5 % of FP64", "storages": {}}
    ref = None
    for st in storages or (_abi.F64, _abi.F32, _abi.F16_OFFSET, _abi.BF16_OFFSET):
        g, kw = make_case(scale, st, model, 0.01)
        if pulse_amp:
            kw.update(pulse_amp=pulse_amp, pulse_period=4000.0)
        c = Case(g.flat(), g.dims, **kw)
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        secs = time.perf_counter() - t0
        lab = c.field("label")
        act = lab != 1
        u = np.stack([c.field(k)[act] for k in ("vx", "vy", "vz")])
        rho1 = c.field("rho")[act] - 1.0
        c.compute_strain()
        S = {k: c.field(k)[act] for k in ("s11", "s12", "s13", "s22", "s32", "s33")}
        gam = np.sqrt(2.0) * np.sqrt(S["s11"] ** 2 + S["s22"] ** 2 + S["s33"] ** 2
                                     + 2 * (S["s12"] ** 2 + S["s13"] ** 2 + S["s32"] ** 2))
        dt, dl = kw["dt"], kw["dl"]
        nu = kw["nu_ref"] + (kw["cy_nu0"] - kw["nu_ref"]) * (1 + (gam * kw["cy_lambda"]) ** kw["cy_a"]) ** (
            (kw["cy_n"] - 1) / kw["cy_a"])
        wallm = c.field("wallmask") != 0
        p = np.pad(wallm, 1, constant_values=False)
        near = (p[:-2, 1:-1, 1:-1] | p[2:, 1:-1, 1:-1] | p[1:-1, :-2, 1:-1] | p[1:-1, 2:, 1:-1]
                | p[1:-1, 1:-1, :-2] | p[1:-1, 1:-1, 2:])[act]
        gdot = gam / dt
        wss = np.where(near, nu * dl * dl / dt * mat.density * gdot, 0.0)
        mf = c.flux(mean=True)
        split = mf[3] / (mf[2] + mf[3])
        c.close()
        row = {"seconds": round(secs, 3), "fluid_mlups": round(int(act.sum()) * steps / secs / 1e6, 1),
               "lpa_split": round(float(split), 6)}
        if ref is None:
            ref = (u, gdot, wss, split, rho1)
        else:
            def rl2(a, b):
                return float(np.sqrt(np.sum((a - b) ** 2)) / np.sqrt(np.sum(b ** 2)))
            row.update(u_rel_l2=rl2(u, ref[0]), rho_minus_1_rel_l2=rl2(rho1, ref[4]), gamma_dot_rel_l2=rl2(gdot, ref[1]),
                       wss_rel_l2=rl2(wss, ref[2]), split_abs_diff_percent=round(100 * abs(split - ref[3]), 5))
            # the accuracy bar of reduced storage: velocity within 1e-3
            # relative L2 and the flow split within 0.5 % of the FP64 run
            row["meets_bar"] = bool(row["u_rel_l2"] <= 1e-3 and row["split_abs_diff_percent"] <= 0.5)
        res["storages"][_abi.STORAGE_NAMES[st]] = row
        print(f"# precision {name} {_abi.STORAGE_NAMES[st]}: {row}", file=sys.stderr, flush=True)
    return res


def cfg1_parity(steps: int) -> dict:
    """The 256x192x128 junction (steady parabolic inflows 0.04/0.03, CY-A
    neq) for ``steps`` steps under each collision model, FP32 storage against
    the FP64 run of the same model on the GPU.  The FP64 path is the one the
    parity tests hold to the CPU oracle (populations to 1e-12); the oracle
    itself would take about 40 minutes per model at this size."""
    from nonnewlbm_amd import _abi
    out = {}
    for name, m in (("cumulant", _abi.CUM), ("bgk", _abi.BGK), ("mrt_les", _abi.MRT), ("kbc_n4", _abi.KBC)):
        r = precision_study(1.0, steps, m, (_abi.F64, _abi.F32), pulse_amp=0.0)
        f32 = r["storages"]["f32"]
        out[name] = {k: f32[k] for k in ("u_rel_l2", "rho_minus_1_rel_l2", "gamma_dot_rel_l2", "wss_rel_l2",
                                         "split_abs_diff_percent", "meets_bar")}
    return {"workload": f"TCPC junction 256x192x128, steady parabolic inflow 0.04/0.03, CY-A neq, {steps} steps, "
                        "FP32 against FP64 per collision model (GPU)", "bar": "u rel L2 <= 1e-3", "models": out}


def configure(case, launch: str) -> tuple[int, int, int]:
    """Autotune ("auto") or set 'bricks_per_group,nontemporal,waves_per_cu'."""
    if launch == "auto":
        return case.autotune()
    bpg, nt, cap = (int(v) for v in launch.split(","))
    case.set_shape(bpg)
    case.set_nontemporal(nt)
    case.set_occupancy(cap)
    return bpg, nt, cap


def host_cpus() -> dict:
    """Host CPU of this process: lscpu's model and physical cores (sockets x
    cores per socket), and the CPUs this process may run on."""
    import subprocess
    out = {"model": None, "physical_cores": None, "threads_per_core": None,
           "cpus_available": len(os.sched_getaffinity(0))}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return out
    kv = {}
    for line in txt.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    out["model"] = kv.get("Model name")
    try:
        out["physical_cores"] = int(kv["Socket(s)"]) * int(kv["Core(s) per socket"])
        out["threads_per_core"] = int(kv["Thread(s) per core"])
    except (KeyError, ValueError):
        pass
    return out


def channel_case(storage: int):
    """The 64x32x32 channel: radius 12 centred in y-z, parabolic inflow
    u_max = 0.05 at x-min, pressure outlet (rho = 1) at x-max, cumulant,
    Newtonian tau = 0.6 (nu = 1/30)."""
    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.geometry import channel_mask
    g = channel_mask()
    kw = dict(storage=storage, collision=_abi.CUM, nu_ref=(0.6 - 0.5) / 3.0, profile=_abi.PROFILE_PARABOLIC,
              ivc_umax=0.05, ivc_radius=12.0, ivc_center=(15.5, 15.5), face_bc=_abi.CHANNEL_FACES,
              dl=1.0e-3, dt=1.0e-3)
    return g, kw


def _timed_oracle(g, kw, steps: int) -> tuple:
    from oracle.oracle import OracleCase
    o = OracleCase(g.flat(), g.dims, **kw)
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    return o, time.perf_counter() - t0


def _rel(a, b) -> float:
    return float(np.sqrt(np.sum((a - b) ** 2)) / max(np.sqrt(np.sum(b ** 2)), 1e-300))


def channel_line() -> dict:
    """The 64x32x32 channel (cumulant, FP64, tau = 0.6) for its full 1,000
    steps on the GPU and on the CPU oracle, with the relative L2 of u and of
    rho - 1 between the two runs."""
    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.lbm355 import Case
    from oracle import oracle
    g, kw = channel_case(_abi.F64)
    c = Case(g.flat(), g.dims, **kw)
    c.step(1)
    c.sync()
    t0 = time.perf_counter()
    c.step(1000)
    c.sync()
    gpu_s = time.perf_counter() - t0
    o, cpu_s = _timed_oracle(g, kw, 1000)
    act = o.field("label") != 1
    n = int(act.sum())
    u = np.stack([c.field(k)[act] for k in ("vx", "vy", "vz")])
    uo = np.stack([o.field(k)[act] for k in ("vx", "vy", "vz")])
    out = {"workload": "channel 64x32x32, radius 12, cumulant, FP64, tau=0.6, parabolic inflow u_max=0.05, "
                       "1000 steps (after one untimed step)",
           "fluid_cells": n,
           "gpu_fluid_mlups": round(n * 1000 / gpu_s / 1e6, 2), "gpu_seconds": round(gpu_s, 4),
           "cpu_fluid_mlups": round(n * 1000 / cpu_s / 1e6, 3), "cpu_seconds": round(cpu_s, 3),
           "cpu_threads": oracle.threads(),
           "u_rel_l2_gpu_vs_cpu": _rel(u, uo),
           "rho_rel_l2_gpu_vs_cpu": _rel(c.field("rho")[act] - 1.0, o.field("rho")[act] - 1.0)}
    c.close()
    print(f"# channel: GPU {out['gpu_fluid_mlups']} / CPU {out['cpu_fluid_mlups']} fluid MLUPS, "
          f"u rel L2 {out['u_rel_l2_gpu_vs_cpu']:.2e}", file=sys.stderr, flush=True)
    return out


def cpu_baseline(g, kw, steps: int, all_models: bool) -> dict:
    """The oracle (a dense A/B OpenMP restatement of the reference's Host
    build, populations held in the case's storage type) on `steps` steps of
    the benchmark case after one untimed step, on every thread OpenMP is
    given; with all_models also BGK, MRT_LES and KBC_N4 on the same steps."""
    from nonnewlbm_amd import _abi
    from oracle import oracle
    o, dt = _timed_oracle(g, kw, steps)
    active = o.info()["active_cells"]
    del o
    host = host_cpus()
    out = {"value": round(active * steps / dt / 1e6, 3), "unit": "fluid MLUPS", "cores": oracle.threads(),
           "kind": "port",
           "all_sites_mlups": round(int(np.prod(g.dims)) * steps / dt / 1e6, 3),
           "sample": f"{steps} steps of the benchmark case (after 1 untimed step), {dt:.1f} s, dense A/B over "
                     f"{int(np.prod(g.dims))} cells, {oracle.threads()} OpenMP threads",
           "host": host}
    if all_models:
        rates = {_abi.COLLISION_IDS[kw["collision"]]: out["value"]}
        for coll in (_abi.BGK, _abi.MRT, _abi.KBC):
            if coll == kw["collision"]:
                continue
            o, dt = _timed_oracle(g, dict(kw, collision=coll), steps)
            rates[_abi.COLLISION_IDS[coll]] = round(active * steps / dt / 1e6, 3)
            del o
        out["models"] = rates
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--storage", default="fp32")
    ap.add_argument("--collision", default="cumulant")
    ap.add_argument("--lbm-viscosity", type=float, default=0.01)
    ap.add_argument("--cy-model", default="neq", choices=["neq", "fd-force", "off"],
                    help="Carreau-Yasuda coupling of the workload (fd-force adds the reference's pre-kernels)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-hbm", action="store_true", help="skip the 512x384x256 HBM-resident line")
    ap.add_argument("--cpu-steps", type=int, default=500, help="timed steps of the CPU baseline")
    ap.add_argument("--cpu-models", action="store_true", help="CPU baseline for all four collision models")
    ap.add_argument("--no-channel", action="store_true", help="skip the 64x32x32 channel GPU/CPU line")
    ap.add_argument("--no-sweep", action="store_true", help="skip the serving stream")
    ap.add_argument("--precision-study", type=int, default=40000, metavar="STEPS",
                    help="steps of the FP64/FP32/FP16/BF16 storage study on the 512x384x256 pulsatile junction "
                         "(0: skip)")
    ap.add_argument("--precision-scale", type=float, default=2.0)
    ap.add_argument("--cfg1-steps", type=int, default=20000,
                    help="steps of the FP32-against-FP64 check of every collision model at 256x192x128 (0: skip)")
    ap.add_argument("--precision-models", type=int, default=0, metavar="STEPS",
                    help="the same study for BGK, MRT_LES and KBC_N4 (FP64 / FP32 / FP16-offset)")
    ap.add_argument("--launch", default="auto", metavar="BPG,NT,CAP",
                    help="launch configuration of the 256x192x128 line (bricks per workgroup, non-temporal even "
                         "step, wavefronts per CU); default: autotune")
    ap.add_argument("--launch-hbm", default="auto", metavar="BPG,NT,CAP",
                    help="the same for the 512x384x256 line")
    ap.add_argument("--batch-steps", type=int, default=2000,
                    help="steps of the 15-junction batch against sequential stepping (0: skip)")
    ap.add_argument("--model-steps", type=int, default=2000,
                    help="steps per collision model on the 256x192x128 workload (0: skip)")
    ap.add_argument("--sweep-requests", type=int, default=48,
                    help="serve the first N requests of the 200-request stream (200: all of it)")
    ap.add_argument("--sweep-rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--sweep-tick", type=int, default=64, help="steps per pool tick")
    ap.add_argument("--graph", type=int, default=1, choices=[0, 1],
                    help="batches and the pool replay captured graphs (1) or launch directly (0)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: only the one-GPU measurement exists; independent cases spread over GPUs "
              "as separate processes and one case's slab decomposition is not built, so --gpus must be 1",
              file=sys.stderr)
        return 2

    from nonnewlbm_amd import _abi
    from nonnewlbm_amd.lbm355 import Case, copy_bandwidth

    storage = _abi.STORAGES[args.storage]
    collision = _abi.COLLISIONS[args.collision]
    sbytes = _abi.STORAGE_BYTES[storage]

    cy_model = {"neq": _abi.CY_NEQ, "fd-force": _abi.CY_FD_FORCE, "off": _abi.CY_OFF}[args.cy_model]
    g, kw = make_case(1.0, storage, collision, args.lbm_viscosity, cy_model)
    case = Case(g.flat(), g.dims, **kw)
    info = case.info()
    active = info["active_cells"]
    print(f"# case {g.dims} active={active} bricks={info['bricks_active']}/{info['bricks_stored']} "
          f"device={info['device_bytes'] / 2**20:.1f} MiB", file=sys.stderr, flush=True)

    case.step(args.warmup)
    # launch shape, even-step cache policy and wavefront cap: the fastest of
    # real steps, or the given configuration
    bpg, nt, cap = configure(case, args.launch)
    case.sync()
    print(f"# launch: {bpg} bricks per workgroup, non-temporal even step {nt}, wavefront cap {cap} ({args.launch})",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    case.step(args.steps)
    case.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    mlups = active * args.steps / wall / 1e6
    print(f"# timed {args.steps} steps in {wall:.3f} s -> {mlups:.1f} fluid MLUPS", file=sys.stderr, flush=True)

    # avg_step_ms: HIP events on the case's stream around the first 512 of
    # 1024 steps, launched back to back (32 whole statistics groups), one
    # k_step launch each; kern_ms: an event pair around each of the other 512
    step_ms, kern_ms = case.time_kernel(1024)
    rho = case.field("rho")
    vx = case.field("vx")
    lab = case.field("label")
    act = lab != 1
    finite = bool(np.isfinite(rho[act]).all() and np.isfinite(vx[act]).all())
    alg_bytes = active * 2 * 27 * sbytes
    achieved = alg_bytes / (step_ms * 1e-3) / 1e9
    flux = case.flux(mean=False)
    # host handover (outputs only; a step moves no host data): one dense
    # double field read back, and the step rate when rho and u are read back
    # every 100 steps
    ncell = int(np.prod(g.dims))
    t0 = time.perf_counter()
    for _ in range(5):
        case.field("vx")
    readback_gbs = 5 * ncell * 8 / (time.perf_counter() - t0) / 1e9
    t0 = time.perf_counter()
    for _ in range(10):
        case.step(100)
        for name in ("rho", "vx", "vy", "vz"):
            case.field(name)
    mlups_readback = active * 1000 / (time.perf_counter() - t0) / 1e6
    copy_gbs = copy_bandwidth(1 << 30, 10)
    print(f"# k_step {step_ms * 1e3:.2f} us per launch back to back -> {achieved:.0f} GB/s algorithmic "
          f"({kern_ms * 1e3:.2f} us with an event pair per launch); copy {copy_gbs:.0f} GB/s; finite={finite}",
          file=sys.stderr, flush=True)
    default_workload = (storage == _abi.F32 and collision == _abi.CUM and args.lbm_viscosity == 0.01
                        and cy_model == _abi.CY_NEQ)
    traffic = profiled_traffic("k_step_tcpc256_cum_f32_cy_stats") if default_workload else None

    out = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "fluid MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if storage == _abi.F64 else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"TCPC junction 256x192x128 ({active} fluid cells), {_abi.COLLISION_IDS[collision]}, "
                        f"{_abi.STORAGE_NAMES[storage]} storage, Carreau-Yasuda set A (neq local viscosity), "
                        f"parabolic inflow 0.04/0.03, nu_lbm={args.lbm_viscosity}",
            "storage": _abi.STORAGE_NAMES[storage],
            "collision": _abi.COLLISION_IDS[collision],
            "fluid_cells": active,
            "grid": list(g.dims),
            "statistics": "mean/rms sums every step (FP64)",
            "launch": {"bricks_per_workgroup": bpg, "nontemporal_even_step": nt, "waves_per_cu_cap": cap},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_step (fused AA stream-collide)",
            "bytes_per_launch": alg_bytes,
            "avg_launch_us": round(step_ms * 1e3, 3),
            "event_pair_launch_us": round(kern_ms * 1e3, 3),
            "working_set_bytes": info["device_bytes"],
            "note": "working set below the 256 MB Infinity Cache: this fraction is cache-assisted",
        },
        "all_sites_mlups": round(int(np.prod(g.dims)) * args.steps / wall / 1e6, 1),
        "host_handover": {"field_readback_gbs": round(readback_gbs, 2),
                          "fluid_mlups_reading_rho_u_every_100_steps": round(mlups_readback, 1)},
        "step_us": round(step_ms * 1e3, 3),
        "copy_bandwidth_gbs": round(copy_gbs, 1),
        "results_finite": finite,
        "plane_flux_lattice": [round(float(v), 6) for v in flux],
    }

    if not args.no_hbm:
        g2, kw2 = make_case(2.0, storage, collision, args.lbm_viscosity, cy_model)
        c2 = Case(g2.flat(), g2.dims, **kw2)
        i2 = c2.info()
        bpg2, nt2, cap2 = configure(c2, args.launch_hbm)
        c2.step(50)
        s2, k2 = c2.time_kernel(512)  # 256 back-to-back steps: 16 whole statistics groups
        b2 = i2["active_cells"] * 2 * 27 * sbytes
        a2 = b2 / (s2 * 1e-3) / 1e9
        out["roofline_hbm_resident"] = {
            "bound": "hbm", "achieved": round(a2, 1), "peak": PEAK_GBS, "unit": "GB/s", "frac": round(a2 / PEAK_GBS, 4),
            "workload": f"TCPC junction 512x384x256 ({i2['active_cells']} fluid cells), same model and storage",
            "avg_launch_us": round(s2 * 1e3, 3), "event_pair_launch_us": round(k2 * 1e3, 3),
            "fluid_mlups": round(i2["active_cells"] / (s2 * 1e3), 1),
            "traffic": profiled_traffic("k_step_tcpc512_cum_f32_cy_stats") if default_workload else None,
            "working_set_bytes": i2["device_bytes"],
            "frac_of_copy": round(a2 / copy_gbs, 4) if copy_gbs > 0 else None,
            "launch": {"bricks_per_workgroup": bpg2, "nontemporal_even_step": nt2, "waves_per_cu_cap": cap2},
        }
        print(f"# HBM-resident: k_step {s2 * 1e3:.1f} us -> {a2:.0f} GB/s ({a2 / PEAK_GBS:.1%})", file=sys.stderr,
              flush=True)
        c2.close()

    if args.model_steps > 0:
        out["models"] = models(args.model_steps)

    if args.batch_steps > 0:
        out["batch15"] = batch15(args.batch_steps, bool(args.graph))

    if not args.no_sweep:
        out["sweep"] = sweep(args.sweep_requests, args.sweep_rate, args.sweep_tick, graph=bool(args.graph))

    if args.precision_study > 0:
        out["precision_study"] = precision_study(args.precision_scale, args.precision_study)
    if args.cfg1_steps > 0:
        out["cfg1_parity"] = cfg1_parity(args.cfg1_steps)
    if args.precision_models > 0:
        from nonnewlbm_amd import _abi
        out["precision_models"] = {
            name: precision_study(args.precision_scale, args.precision_models, m, (_abi.F64, _abi.F32, _abi.F16_OFFSET))
            for name, m in (("bgk", _abi.BGK), ("mrt_les", _abi.MRT), ("kbc_n4", _abi.KBC))}

    if not args.no_channel:
        out["channel"] = channel_line()

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g, kw, args.cpu_steps, args.cpu_models)
        print(f"# cpu baseline {out['cpu_baseline']}", file=sys.stderr, flush=True)

    print(json.dumps(out), flush=True)
    return 0 if finite else 1


if __name__ == "__main__":
    sys.exit(main())
