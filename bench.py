#!/usr/bin/env python3
"""Benchmark: plain RKDP54 LLG integration with demag on one MI355X.

Workload (the large 3D plain-method case): 512 x 512 x 32 cells of 4 nm cubes, Ms = 1e6 A/m,
A = 1.5e-11 J/m, uniaxial K1 = 5e5 J/m^3 along z, alpha = 0.1, Hz = 50 mT; demag + exchange
+ anisotropy + Zeeman; initial M i.i.d. uniform on the sphere from the reference's BorisRand
sequence (seed 1); FP32 per-step arithmetic, FP64 demag-kernel setup once (not timed).

A step is one accepted RKDP54 step (6 field evaluations, each a full FFT demag convolution,
plus the fused field/stage kernels); rejected attempts cost time but are not counted, like
the reference's iteration counter.  value = cells * accepted steps / wall seconds.

The JSON line also carries:
  roofline     — the kernel with the largest total time in the timed region: algorithmic
                 HBM bytes per launch / its average launch duration (HIP events recorded on
                 its own stream around every launch in the timed region), against 8 TB/s;
  step_roofline — the whole step: algorithmic bytes of every pass launched in the timed
                 region (rejected attempts included) / wall time, against 8 TB/s;
  cpu_baseline — the oracle (this repository's CPU restatement of the reference path, FP64,
                 OpenMP over all host threads) timed on a bounded sample of the same
                 workload on this machine's host cores;
  eval_speedup — the same workload with the demag evaluation speedup (saved demag fields
                 extrapolated in time, one convolution per step);
  batched      — 64 films of uneven sizes, batched RK4 vs the same films one by one, and
                 the same with RKDP54 (per-film adaptive steps controlled on the device);
  adjoint      — 128x128x4 RK4 rollout with its discrete adjoint: backward / forward cost;
  multilayer   — 16 stacks of 3 layers of 256x256 cells, batched RK4 vs one by one, and one
                 stack's adjoint (per-layer Ms and field-ramp slope gradients).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "LLG cell-updates/s (RKDP54+demag, FP32) batched vs plain; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0


def workload():
    from boris2_amd import api
    n, h = (512, 512, 32), (4e-9, 4e-9, 4e-9)
    p = api.FMParams(Ms=1e6, A=1.5e-11, alpha=0.1, K1=5e5, ea=(0.0, 0.0, 1.0), Ha=(0.0, 0.0, 50e-3 / api.MU0),
                     modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)
    return n, h, p


def algorithmic_bytes(kind: str, n, N) -> float:
    """HBM bytes one launch of `kind` must move at minimum (FP32, complex 8 B, 3 components)."""
    nx, ny, nz = n
    Nx, Ny, Nz = N
    NC, NYH, NZH = Nx // 2 + 1, Ny // 2 + 1, Nz // 2 + 1
    cells = nx * ny * nz
    half = 24.0 * NC * ny * nz          # spectra with only the ny data rows, 3 components
    full = 24.0 * NC * Ny * nz          # spectra after the y transform
    kern = 24.0 * NC * NYH * NZH        # 6-component FP32 kernel octant
    return {
        "fwd_x_r2c": 12.0 * cells + half,
        "y_fft": half + full,
        "z_fft_kernel_ifft": full + kern + full,
        "y_ifft": full + half,
        "inv_x_c2r": half + 12.0 * cells,
        "y2d_fft_kernel_ifft": half + 16.0 * NC * NYH + half,
        # RKDP54 stages 0..5 move 4,6,7,8,9,10 float3 arrays of the mesh (planar Hd included)
        "llg_stage": 44 / 6 * 12.0 * cells,
        "rkdp_advance": 4 * 12.0 * cells,
        # fused row kernel, average over the 6 launches of an accepted RKDP54 step: stage 0
        # reads the spectra rows + 3 float3 arrays (M and the partial embedded solution in,
        # E0 out); stages 1..5 read and write the spectra rows and move 5,6,7,8,9 arrays
        "row_c2r_stage_r2c": (11 * half + 38 * 12.0 * cells) / 6,
        # x-R2C fused with the RKDP54 advance: M, E0 in; sM1, M1 out; spectra rows out
        "fwd_x_advance": 4 * 12.0 * cells + half,
    }[kind]


def conv_N(n):
    """N of the zero-padded convolution of an open-boundary mesh whose nx and 2 ny have plans:
    N.x = 2 nx, N.y = 2 ny, N.z = 1 (2D) or the next power of two >= 2 nz."""
    nz = n[2]
    Nz = 1
    if nz > 1:
        while Nz < 2 * nz:
            Nz *= 2
    return (2 * n[0], 2 * n[1], Nz)


def unfused_eval_bytes(n) -> float:
    """Algorithmic HBM bytes of one unfused field evaluation of a mesh: the convolution passes
    (x R2C, the y / z / y passes or the 2D y pass, x C2R to planar H_demag) plus the stage
    kernel's read of H_demag (12 B/cell)."""
    N = conv_N(n)
    cells = n[0] * n[1] * n[2]
    mid = (algorithmic_bytes("y2d_fft_kernel_ifft", n, N) if n[2] == 1 else
           algorithmic_bytes("y_fft", n, N) + algorithmic_bytes("z_fft_kernel_ifft", n, N) + algorithmic_bytes("y_ifft", n, N))
    return algorithmic_bytes("fwd_x_r2c", n, N) + mid + algorithmic_bytes("inv_x_c2r", n, N) + 12.0 * cells


def batched_roofline(meshes, evals: int, stage_array_passes: int, ms_per_step: float) -> dict:
    """Whole-step roofline of a batched run: algorithmic bytes of one accepted step of every
    mesh (evals unfused evaluations + the stage kernels' array passes of 12 B/cell) over the
    measured time per step, against the 8 TB/s peak.  Rejected RKDP54 attempts are not counted."""
    b = sum(evals * unfused_eval_bytes(n) + stage_array_passes * 12.0 * n[0] * n[1] * n[2] for n, *_ in meshes)
    ach = b / (ms_per_step * 1e-3) / 1e9
    return {"bound": "hbm", "alg_GB_per_step": round(b / 1e9, 4), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4)}


def host_info() -> dict:
    """The host CPU model and thread count, and whether FFTW can be loaded (dlopen of
    libfftw3.so.3): the oracle uses its own FFT either way."""
    import ctypes
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ctypes.CDLL("libfftw3.so.3")
        fftw = "found (libfftw3.so.3 loads; the oracle still uses its own FFT)"
    except OSError as e:
        fftw = f"not found (dlopen libfftw3.so.3: {e})"
    return {"cpu_model": model, "host_threads": os.cpu_count(), "fftw3": fftw}


# rocprof kernel-name prefix of each launch kind (for the PMC traffic table in profiles/)
KIND_TO_KERNEL = {"fwd_x_r2c": "k_fwd_x<512, 8, false>", "y_fft": "k_y<1024, -1>", "z_fft_kernel_ifft": "k_z<",
                  "y_ifft": "k_y<1024, 1>", "inv_x_c2r": "k_inv_x<", "llg_stage": "k_stage<", "rkdp_advance": "k_rkdp_advance",
                  "row_c2r_stage_r2c": "k_row<", "fwd_x_advance": "k_fwd_x<512, 8, true>"}
TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_rkdp54_512x512x32.json"


def measured_traffic(kind: str):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs (copy-calibrated)."""
    if not TRAFFIC_FILE.exists():
        return None
    t = json.loads(TRAFFIC_FILE.read_text())["kernels"]
    pre = KIND_TO_KERNEL.get(kind)
    vals = [v["traffic_bytes"] for k, v in t.items() if pre and pre in k]
    return round(sum(vals) / len(vals)) if vals else None


def gpu_run(steps: int, warmup: int):
    import torch
    from boris2_amd import api, _lib
    n, h, p = workload()
    ncell = n[0] * n[1] * n[2]
    M0 = api.random_magnetisation(ncell, p.Ms, seed=1).astype(np.float32)
    t0 = time.perf_counter()
    s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    N, dev_bytes = s.demag.info()
    s.run_to_iteration(warmup)
    torch.cuda.synchronize()
    st0 = s.state()
    L = _lib.load()
    _lib.check(L.bm_prof_begin(64 * (steps + 8) * 2))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run_to_iteration(warmup + steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    L.bm_prof_end()
    st1 = s.state()
    kinds = {}
    for k, name in enumerate(_lib.PROF_KINDS):
        cnt, tot, mn, mx = C.c_longlong(), C.c_double(), C.c_double(), C.c_double()
        _lib.check(L.bm_prof_query(k, C.byref(cnt), C.byref(tot), C.byref(mn), C.byref(mx)))
        if cnt.value:
            avg = tot.value / cnt.value
            b = algorithmic_bytes(name, n, N)
            kinds[name] = {"launches": cnt.value, "total_ms": round(tot.value, 3), "avg_ms": round(avg, 5),
                           "alg_GB_per_launch": round(b / 1e9, 4), "GB_per_s": round(b / (avg * 1e-3) / 1e9, 1)}
    accepted = st1["iteration"] - st0["iteration"]
    m_avg = s.average_m()
    s.close()
    return {"n": n, "N": N, "ncell": ncell, "wall": wall, "accepted": accepted, "warmup": warmup,
            "rejected": st1["rejections"] - st0["rejections"], "evals": st1["evaluations"] - st0["evaluations"],
            "kinds": kinds, "setup_s": setup_s, "dev_bytes": dev_bytes, "sim_time": st1["time"], "m_avg": m_avg.tolist()}


def speedup_run(steps: int, warmup: int, mode: int, plain):
    """The same workload with the demag evaluation speedup (mode = number of saved demag fields,
    extrapolated in time between convolutions): one convolution per accepted step instead of
    six.  Reports the rate, the speedup over the plain run and <m> of both after the same
    iteration count (the trajectories differ by the extrapolation error)."""
    import torch
    from boris2_amd import api
    n, h, p = workload()
    ncell = n[0] * n[1] * n[2]
    M0 = api.random_magnetisation(ncell, p.Ms, seed=1).astype(np.float32)
    s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
    s.set_eval_speedup(mode)
    s.run_to_iteration(warmup)
    torch.cuda.synchronize()
    st0, c0 = s.state(), s.speedup_state()["convolutions"]
    t0 = time.perf_counter()
    s.run_to_iteration(warmup + steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st1, c1 = s.state(), s.speedup_state()["convolutions"]
    m_avg = s.average_m()
    s.close()
    acc = st1["iteration"] - st0["iteration"]
    rate = ncell * acc / wall
    plain_rate = plain["ncell"] * plain["accepted"] / plain["wall"]
    out = {"mode": mode, "method": "RKDP54 adaptive", "steps": acc, "rejected_attempts": st1["rejections"] - st0["rejections"],
           "convolutions": c1 - c0, "field_evaluations": st1["evaluations"] - st0["evaluations"],
           "cell_updates_per_s": rate, "ms_per_step": 1e3 * wall / max(1, acc), "over_plain": rate / plain_rate,
           "sim_time_s": st1["time"], "m_avg": m_avg.tolist()}
    if plain["accepted"] == acc and plain.get("warmup") == warmup:
        out["plain_m_avg"] = plain["m_avg"]
        out["plain_sim_time_s"] = plain["sim_time"]
        out["max_abs_dm_avg"] = float(np.abs(np.array(plain["m_avg"]) - m_avg).max())
    return out


def batched_meshes(seed: int = 64, count: int = 64):
    """64 independent films: nx, ny from {64, 96, 128, 192, 256}, nz from {1, 2, 4, 8}, 4 nm cells,
    Ms uniform in [4e5, 1.2e6] A/m, A 1.3e-11 J/m, alpha uniform in [0.01, 0.5], demag + exchange,
    M0 random on the sphere (seeded)."""
    from boris2_amd import api
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        n = (int(rng.choice([64, 96, 128, 192, 256])), int(rng.choice([64, 96, 128, 192, 256])), int(rng.choice([1, 2, 4, 8])))
        p = api.FMParams(Ms=float(rng.uniform(4e5, 1.2e6)), A=1.3e-11, alpha=float(rng.uniform(0.01, 0.5)),
                         modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE)
        M0 = api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 1000 + b).astype(np.float32)
        out.append((n, (4e-9, 4e-9, 4e-9), p, M0))
    return out


def batch_run(steps: int):
    """Batched RK4 (dt 50 fs) over the 64 films vs the same films run one by one on the plain
    single-mesh path; both timed over `steps` steps after one warm-up step."""
    import torch
    from boris2_amd import api
    meshes = batched_meshes()
    dt = 50e-15
    cells = sum(n[0] * n[1] * n[2] for n, _, _, _ in meshes)
    bs = api.BatchSolver(api.BM_EVAL_RK4, [(n, h, p) for n, h, p, _ in meshes], [M0 for *_, M0 in meshes], autotune="radix")
    bs.set_dt(dt)
    t0 = time.perf_counter()
    bs.run(1)                 # the autotuner runs before the first step
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    bs.run(steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    tuned = bs.autotune_report()
    bs.close()
    solvers = []
    for n, h, p, M0 in meshes:
        s = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
        s.set_dt(dt)
        s.run_to_iteration(1)
        solvers.append(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in solvers:
        s.run_to_iteration(1 + steps)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    for s in solvers:
        s.close()
    return {"meshes": len(meshes), "cells": cells, "steps": steps, "method": "RK4 fixed dt 50 fs",
            "batched_cell_updates_per_s": cells * steps / tb, "plain_cell_updates_per_s": cells * steps / tp,
            "batched_over_plain": tp / tb, "batched_ms_per_step": 1e3 * tb / steps, "plain_ms_per_step": 1e3 * tp / steps,
            "roofline": batched_roofline(meshes, 4, 22, 1e3 * tb / steps),
            "autotune_s": round(t_tune, 2), "autotuned_passes": tuned}


def batch_rkdp_run(steps: int, warm: int = 5):
    """The 64 films with RKDP54 (each its own adaptive step, the reference's default control):
    batched with the step control on the device vs the same films one by one on the plain
    path (host step control, one lte read-back per step); `steps` accepted steps of every
    film after `warm` warm-up steps (the first steps from random states are rejected and
    redone)."""
    import torch
    from boris2_amd import api
    meshes = batched_meshes()
    cells = sum(n[0] * n[1] * n[2] for n, _, _, _ in meshes)
    bs = api.BatchSolver(api.BM_EVAL_RKDP54, [(n, h, p) for n, h, p, _ in meshes], [M0 for *_, M0 in meshes], autotune="radix")
    bs.run(warm)
    torch.cuda.synchronize()
    rej0 = sum(bs.mesh_state(b)["rejections"] for b in range(len(meshes)))
    t0 = time.perf_counter()
    bs.run(steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    rej = sum(bs.mesh_state(b)["rejections"] for b in range(len(meshes))) - rej0
    bs.close()
    solvers = []
    for n, h, p, M0 in meshes:
        s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
        s.run_to_iteration(warm)
        solvers.append(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in solvers:
        s.run_to_iteration(warm + steps)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    for s in solvers:
        s.close()
    return {"meshes": len(meshes), "cells": cells, "steps": steps, "method": "RKDP54 adaptive, per-mesh dT",
            "rejected_attempts": rej, "batched_cell_updates_per_s": cells * steps / tb,
            "plain_cell_updates_per_s": cells * steps / tp, "batched_over_plain": tp / tb,
            "batched_ms_per_step": 1e3 * tb / steps, "plain_ms_per_step": 1e3 * tp / steps,
            "roofline": batched_roofline(meshes, 6, 52, 1e3 * tb / steps)}


def smooth_unit_field(n, rng, kmax: int = 3) -> np.ndarray:
    """A seeded smooth random unit vector field on an n = (nx, ny, nz) mesh, (ncell, 3):
    each component a sum of sine/cosine modes of wavenumber 0..kmax per in-plane axis (whole
    mesh lengths; uniform through the thickness) with normally distributed amplitudes,
    normalised per cell."""
    nx, ny, nz = n
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    comps = []
    for _ in range(3):
        f = np.zeros((ny, nx))
        for kx in range(kmax + 1):
            for ky in range(kmax + 1):
                a, b, c, d = rng.normal(size=4) / (1.0 + kx + ky)
                f += (a * np.cos(2 * np.pi * kx * X) + b * np.sin(2 * np.pi * kx * X)) * \
                     (c * np.cos(2 * np.pi * ky * Y) + d * np.sin(2 * np.pi * ky * Y))
        comps.append(f)
    m = np.stack(comps, axis=-1)
    m /= np.linalg.norm(m, axis=-1, keepdims=True)
    return np.tile(m.reshape(-1, 3), (nz, 1))


def adjoint_problem(seed: int = 4):
    """The differentiation workload: a 128x128x4 film of 4 nm cells; Ms map = sum of 8 seeded
    Gaussian blobs mapped to [5e5, 1.1e6] A/m; A 1.3e-11 J/m, alpha 0.05; a 1 ns sinc pulse of
    20 mT amplitude in a seeded direction; m0 and m_target seeded smooth random unit fields.
    Returns (n, h, params, Ms (ncell,), m0 (ncell, 3), m_target (ncell, 3))."""
    from boris2_amd import api
    rng = np.random.default_rng(seed)
    n = (128, 128, 4)
    x, y = np.meshgrid(np.arange(n[0]), np.arange(n[1]))
    blob = np.zeros((n[1], n[0]))
    for _ in range(8):
        cx, cy, r = rng.uniform(0, n[0]), rng.uniform(0, n[1]), rng.uniform(6, 20)
        blob += np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * r * r))
    blob = (blob - blob.min()) / (blob.max() - blob.min())
    Ms = np.tile((5e5 + 6e5 * blob).reshape(-1), n[2]).astype(np.float32)
    d = rng.normal(size=3); d /= np.linalg.norm(d)
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.05, modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN,
                     wave=1, amp=20e-3 / api.MU0, w=2 * np.pi / 1e-9, t0=0.5e-9, dir=tuple(d))
    m0 = smooth_unit_field(n, rng)
    mt = smooth_unit_field(n, rng)
    return n, (4e-9, 4e-9, 4e-9), p, Ms, m0, mt


def adjoint_run(steps: int, ckpt: int):
    """The differentiation workload (adjoint_problem): RK4 dt 250 fs, checkpoints every `ckpt`
    steps; loss 1 - <m_final . m_target>.  Forward and backward wall times (the backward
    includes the segment recomputation)."""
    import torch
    from boris2_amd import api
    n, h, p, Ms, m0, mt = adjoint_problem()
    ncell = n[0] * n[1] * n[2]
    r = api.RK4Adjoint(n, h, p, torch.tensor(Ms), 250e-15, steps, ckpt)
    m0t, mtt = torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, _ = r.forward(m0t, mtt)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    g, ga = r.backward()
    torch.cuda.synchronize()
    tbk = time.perf_counter() - t0
    r.close()
    return {"mesh": list(n), "steps": steps, "checkpoint_every": ckpt, "forward_s": round(tf, 3), "backward_s": round(tbk, 3),
            "adjoint_over_forward": round(tbk / tf, 3), "loss": loss, "dL_damp": ga,
            "forward_cell_updates_per_s": ncell * steps / tf}


def multilayer_instances(seed: int = 5, count: int = 16):
    """16 stacks of 3 layers of 256x256 cells of 4 nm: 2/1/4-cell layers (8/4/16 nm) with
    1-cell spacers, per-layer Ms uniform in [6e5, 1.2e6] A/m, A 1.3e-11, K1 3e5 J/m^3 along z
    in the thin layer, alpha 0.02, a seeded field ramp per stack, M0 random on the sphere."""
    from boris2_amd import api
    rng = np.random.default_rng(seed)
    hz, z0 = [8e-9, 4e-9, 16e-9], [0.0, 12e-9, 20e-9]
    out = []
    for b in range(count):
        lay = dict(hz=hz, z0=z0, Ms=[float(v) for v in rng.uniform(6e5, 1.2e6, 3)], A=[1.3e-11] * 3,
                   alpha=[0.02] * 3, K1=[0.0, 3e5, 0.0])
        p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, ea=(0.0, 0.0, 1.0),
                         Ha=tuple(float(v) for v in rng.normal(0.0, 8e3, 3)), wave=2,
                         slope=tuple(float(v) for v in rng.normal(0.0, 2e13, 3)),
                         modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)
        ms = np.repeat(lay["Ms"], 256 * 256)
        M0 = (api.random_magnetisation(ms.size, 1.0, 2000 + b) * ms[:, None]).astype(np.float32)
        out.append(((256, 256, 3), (4e-9, 4e-9, 16e-9), p, lay, M0))
    return out


def multilayer_run(steps: int, adj_steps: int, ckpt: int = 100):
    """The 16 layer stacks: batched RK4 (dt 100 fs) vs the same stacks one by one on the
    single-stack path over `steps` steps; then the adjoint of one stack's `adj_steps`-step
    rollout (per-layer Ms and ramp-slope gradients), forward and backward wall times, and the
    adjoints of all 16 stacks run in lockstep (one stream each) against 16 times one stack."""
    import torch
    from boris2_amd import api
    items = multilayer_instances()
    dt = 100e-15
    cells = sum(n[0] * n[1] * n[2] for n, *_ in items)
    bs = api.BatchSolver(api.BM_EVAL_RK4, [(n, h, p) for n, h, p, _, _ in items], [M0 for *_, M0 in items],
                         layers=[lay for *_, lay, _ in items], autotune="radix")
    bs.set_dt(dt)
    bs.run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bs.run(steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    bs.close()
    solvers = []
    for n, h, p, lay, M0 in items:
        s = api.LayerStackSolver(api.BM_EVAL_RK4, n[0], n[1], h[0], h[1], lay["hz"], lay["z0"], p, Ms=lay["Ms"],
                                 A=lay["A"], alpha=lay["alpha"], K1=lay["K1"], M0=M0)
        s.set_dt(dt)
        s.run_to_iteration(1)
        solvers.append(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in solvers:
        s.run_to_iteration(1 + steps)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    for s in solvers:
        s.close()
    out = {"stacks": len(items), "layers": 3, "cells": cells, "steps": steps, "method": "RK4 fixed dt 100 fs",
           "batched_cell_updates_per_s": cells * steps / tb, "plain_cell_updates_per_s": cells * steps / tp,
           "batched_over_plain": tp / tb, "batched_ms_per_step": 1e3 * tb / steps, "plain_ms_per_step": 1e3 * tp / steps}
    if adj_steps > 0:
        n, h, p, lay, M0 = items[0]
        rng = np.random.default_rng(8)
        mt = rng.normal(size=M0.shape); mt /= np.linalg.norm(mt, axis=1, keepdims=True)
        m0 = M0 / np.linalg.norm(M0, axis=1, keepdims=True)
        Ms = torch.tensor(np.repeat(lay["Ms"], n[0] * n[1]), dtype=torch.float32)
        r = api.RK4Adjoint(n, h, p, Ms, dt, adj_steps, ckpt, layers=lay)
        m0t, mtt = torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss, _ = r.forward(m0t, mtt)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        g, _ = r.backward()
        torch.cuda.synchronize()
        tbk = time.perf_counter() - t0
        gz = r.zeeman_grads()
        gl = g.double().cpu().reshape(3, -1).sum(1).tolist()
        r.close()
        out["adjoint"] = {"stack": 0, "steps": adj_steps, "checkpoint_every": ckpt, "forward_s": round(tf, 3),
                          "backward_s": round(tbk, 3), "adjoint_over_forward": round(tbk / tf, 3), "loss": loss,
                          "dL_dMs_layer": gl, "dL_dslope": gz[1:].tolist()}
        # all 16 stacks' adjoints in lockstep, one stream each
        rs, m0s, mts = [], [], []
        for k, (n, h, p, lay, M0) in enumerate(items):
            rng = np.random.default_rng(8 + k)
            mt = rng.normal(size=M0.shape); mt /= np.linalg.norm(mt, axis=1, keepdims=True)
            Ms = torch.tensor(np.repeat(lay["Ms"], n[0] * n[1]), dtype=torch.float32)
            rs.append(api.RK4Adjoint(n, h, p, Ms, dt, adj_steps, ckpt, layers=lay))
            m0s.append(torch.tensor(M0 / np.linalg.norm(M0, axis=1, keepdims=True), dtype=torch.float32))
            mts.append(torch.tensor(mt, dtype=torch.float32))
        bt = api.RK4AdjointBatch(rs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses, _ = bt.forward(m0s, mts)
        torch.cuda.synchronize()
        tbf = time.perf_counter() - t0
        t0 = time.perf_counter()
        grads = bt.backward()
        torch.cuda.synchronize()
        tbb = time.perf_counter() - t0
        for r in rs:
            r.close()
        out["batched_adjoint"] = {"stacks": len(items), "steps": adj_steps, "checkpoint_every": ckpt,
                                  "forward_s": round(tbf, 3), "backward_s": round(tbb, 3),
                                  "adjoint_over_forward": round(tbb / tbf, 3),
                                  "one_stack_alone_x16_s": round(16 * (tf + tbk), 3),
                                  "batched_over_one_by_one": round(16 * (tf + tbk) / (tbf + tbb), 3),
                                  "mean_loss": float(np.mean(losses))}
    return out


def cpu_baseline(cpu_steps: int):
    """Oracle FP64 on the host: kernel setup (reported separately) then `cpu_steps` steps."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    from boris2_amd import api
    n, h, p = workload()
    ncell = n[0] * n[1] * n[2]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    o = O.OracleSim(api.BM_EVAL_RKDP54, n, h, p, Nz_override=64)
    setup = time.perf_counter() - t0
    o.set_M(api.random_magnetisation(ncell, p.Ms, seed=1))
    t0 = time.perf_counter()
    o.run_to_iteration(cpu_steps)
    dt = time.perf_counter() - t0
    hi = host_info()
    return {"value": ncell * cpu_steps / dt, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"{cpu_steps} RKDP54 steps ({6 * cpu_steps} field evaluations) of the same 512x512x32 workload, "
                      f"oracle FP64 with its own mixed-radix FFT, OpenMP {cores} threads; "
                      f"{dt:.1f} s timed, FP64 kernel setup {setup:.1f} s not included",
            **hi}


def trilayer_run(steps: int):
    """The reference's own multilayered example of 3D meshes (MLConv Paper/3D/trilayer_switch.bsm:
    permalloy meshes 128x64x{4,2,4} of 5 nm with 1 nm gaps, per-mesh fields, RK4 at 0.5 ps) as a
    stack of 10 planes: wall time of `steps` RK4 steps after 20 warm-up steps."""
    import time
    import numpy as np
    import torch
    from boris2_amd import api
    nx, ny = 128, 64
    hz, z0, grp = api.stack_of_meshes([(4, 5e-9, -21e-9), (2, 5e-9, 0.0), (4, 5e-9, 11e-9)])
    b_out, b_mid = np.array([0.025037111148343075, 0.0021904566281595616, 0.0]), np.array([0.006259271503900462, 0.0005476160419954826, 0.0])
    Ha = np.array([(b_out if g != 1 else b_mid) / api.MU0 for g in grp])
    p = api.FMParams(Ms=8e5, A=13e-12, alpha=0.02, modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN)
    nl = nx * ny
    M0 = np.zeros((len(grp) * nl, 3), np.float32)
    M0[:, 0] = np.repeat(np.where(grp == 1, 8e5, -8e5), nl)
    s = api.LayerStackSolver(api.BM_EVAL_RK4, nx, ny, 5e-9, 5e-9, hz, z0, p, M0=M0)
    s.set_layer_meshes(grp, Ha)
    s.set_dt(5e-13)
    s.run_to_iteration(20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run_to_iteration(20 + steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s.close()
    ncell = len(grp) * nl
    return {"workload": "reference trilayer (3 meshes, 10 planes of 128x64, RK4 0.5 ps)", "cells": ncell, "steps": steps,
            "ms_per_step": round(1e3 * wall / steps, 4), "cell_updates_per_s": ncell * steps / wall,
            "wall_3ns_s": round(6000 * wall / steps, 3)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-steps", type=int, default=50, help="steps of the batched vs one-by-one comparison (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=4000, help="RK4 steps of the adjoint rollout (0: skip)")
    ap.add_argument("--ml-steps", type=int, default=50, help="steps of the batched layer-stack comparison (0: skip)")
    ap.add_argument("--ml-adjoint-steps", type=int, default=1000, help="RK4 steps of the layer-stack adjoint (0: skip)")
    ap.add_argument("--trilayer-steps", type=int, default=1000, help="RK4 steps of the reference trilayer stack (0: skip)")
    ap.add_argument("--speedup-steps", type=int, default=500, help="RKDP54 steps with the demag evaluation speedup (0: skip)")
    ap.add_argument("--speedup-mode", type=int, default=3, help="evaluation speedup mode (saved demag fields, 1..6)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 512 MiB device copy 5 times (scales the HBM byte counters)")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        x = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each": 2 * x.numel() * 4, "copies": 5}))
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; a mesh or batch sharded over several GPUs is not built.",
              file=sys.stderr)
        return 2

    r = gpu_run(a.steps, a.warmup)
    kinds = r["kinds"]
    dom = max(kinds, key=lambda k: kinds[k]["total_ms"])
    d = kinds[dom]
    achieved = d["GB_per_s"]
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
            "traffic_source": str(TRAFFIC_FILE.relative_to(ROOT)), "kernel": dom,
            "alg_bytes_per_launch": round(d["alg_GB_per_launch"] * 1e9), "avg_launch_ms": d["avg_ms"]}
    speedup = speedup_run(a.speedup_steps, a.warmup, a.speedup_mode, r) if a.speedup_steps > 0 else None
    batched = batch_run(a.batch_steps) if a.batch_steps > 0 else None
    if batched is not None:
        batched["rkdp54"] = batch_rkdp_run(a.batch_steps)
    adjoint = adjoint_run(a.adjoint_steps, 100) if a.adjoint_steps > 0 else None
    multilayer = multilayer_run(a.ml_steps, a.ml_adjoint_steps) if a.ml_steps > 0 else None
    if multilayer is not None and a.trilayer_steps > 0:
        multilayer["trilayer"] = trilayer_run(a.trilayer_steps)
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_steps)
    value = r["ncell"] * r["accepted"] / r["wall"]
    # the whole step against the same peak: algorithmic bytes of every pass launched in the
    # timed region (rejected attempts included) over the wall time
    step_bytes = sum(v["launches"] * v["alg_GB_per_launch"] for v in kinds.values()) * 1e9
    step_roof = {"bound": "hbm", "alg_GB_per_accepted_step": round(step_bytes / 1e9 / max(1, r["accepted"]), 3),
                 "achieved": round(step_bytes / r["wall"] / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(step_bytes / r["wall"] / 1e9 / HBM_PEAK_GBS, 4)}
    line = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": r["accepted"],
        "warmup": a.warmup, "ms_per_step": 1e3 * r["wall"] / max(1, r["accepted"]), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain RKDP54 + demag/exchange/uniaxial anisotropy/Zeeman, 512x512x32 cells of 4 nm, "
                               "FP32 steps, FP64 kernel setup", "cells": r["ncell"], "fft_N": list(r["N"]),
                   "accepted_steps": r["accepted"], "rejected_attempts": r["rejected"],
                   "field_evaluations": r["evals"], "sim_time_s": r["sim_time"], "setup_s": round(r["setup_s"], 2),
                   "parallelism": "single GPU"},
        "roofline": roof,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
        "eval_speedup": speedup,
        "batched": batched,
        "adjoint": adjoint,
        "multilayer": multilayer,
        "kernels": kinds,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
