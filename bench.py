#!/usr/bin/env python3
"""Benchmark of the MI355X hot path.

Headline (``value``): eikonal source-stages per second.  One step = one
forward pass of the multistage eikonal solver over all OBS of the workload:
for every OBS the twelve time fields of the five phases (Pg, PsedP, PmP,
Pg_ss, Pg_ss2), solved on the device with the inputs (model, geometry)
already resident.  A source-stage is one (OBS, phase) chain, so a step
completes 5 x nobs source-stages.  Default workload: the batched forward on
the 0.1 km grid (2001x301 nodes), 200 OBS every 1 km with the sediment base
and Moho perturbed by seeded Fourier modes: 1000 source-stages (2400 fields)
in one fmt2d_forward.  ``multistage`` (60 OBS every 3.3 km) is reported
beside it, and ``per_source``: the same 200-OBS forward solved one OBS at a
time (the per-source stage loop of the reference, on the GPU), timed on four
OBS across the model and scaled to all 200, with the batched step's speedup
over it.  ``level_launch``: the same forward with the tile solver's launches
run by one cooperative dispatch per stage level instead of replayed graphs
(its time, and a check that every field has the same bits).

Second figure (``iteration``): seconds per inversion iteration = forward +
rays + Frechet rows + system assembly + one LSQR solve (fixed damping, update
capped) on the plain case (401x61, 20 OBS, 399 shots, Pg; picks from the
checkerboard model with 50 ms seeded noise, starting model without it,
Laplacian smoothing), with the reference's own inversion timed beside it.
The line search that completes the iteration (forwards without Frechet rows,
model steps) is reported apart, for both.

Third figure (``programs``): wall seconds of the Fortran host programs
(fmt2d_amd/fortran/fmt2d_hip and invert2d_hip, process start to exit) on the
plain case, beside the reference's own fmt2d and invert2d on the same input
files: the forward program over the 20 OBS (all 12 fields, 7980 Pg rays,
arrival files), and a two-iteration inversion (automatic damping as in the
reference's example scripts, with the initial and final forwards and every
output file).

Fourth figure (``gradient``): seconds for the misfit 1/2 |t - t_obs|^2 and
its gradient with respect to the velocity grid and the Moho depths through
fmt2d_amd.autograd (forward, rays, adjoint-state backward) at 2001x301, 60
OBS, Pg and PmP (picks from a +-5 % checkerboard of 10 km x 3 km cells plus
30 ms seeded noise, starting model without it), with the ray-based
backward (Frechet rows + G^T r) timed beside it.

Sixth figure (``fp32_storage``): the multistage forward with the tile
solver's grid-node fields stored in FP32 beside FP64 storage, with the
largest |dT| per stage field against the FP64 fields and the fields that
have nodes beyond 1e-4 s named.

Fifth figure (``joint``, ``--joint N``, default 4, 0 skips): seconds per iteration of a
joint velocity + interface inversion on an inversion mesh coarser than the
propagation grid: 2001x301, 200 OBS every 1 km, shots every 0.2 km, five
phases (434k rays); velocity parameters on a 0.5 km mesh (401x61), sediment
base and Moho depths every 1 km (2x201); N iterations of the alternating
schedule with hybrid smoothing, damping 0.1 and the line search.  The
reference has no such parameterisation, so no CPU figure stands beside it.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
BYTES_PER_NODE = 32  # FP64: read velocity 8, write T 8, write grad T 16


def workload(name: str):
    from fmt2d_amd import synth
    if name == "batched":
        m, g = synth.batched_case()
        return m, g, ("batched forward, 2001x301 grid (0.1 km), 200 OBS every 1 km, sediment base and Moho "
                      "perturbed by seeded Fourier modes (seed 3), 5 phases / 12 fields per OBS (1000 "
                      "source-stages, 2400 fields in one fmt2d_forward)")
    if name == "multistage":
        m, g = synth.multistage_case(dx=0.1, nobs=60)
        return m, g, "multistage forward, 2001x301 grid (0.1 km), 60 OBS every 3.3 km, 5 phases / 12 fields per OBS"
    if name == "plain":
        m, g = synth.plain_pg_case()
        return m, g, "plain forward, 401x61 grid (0.5 km), 20 OBS every 10 km, 5 phases / 12 fields per OBS"
    raise SystemExit(f"unknown workload {name}")


def write_case(d, m, g):
    from fmt2d_amd import synth
    synth.write_fmt2d_case(d, m, g)


def cpu_baseline(case_dir: str, obs_x, nsample: int):
    """Reference FMT2D modules (oracle/_ref/ref_fields, built from the reference
    sources) timed on one host core for `nsample` OBS of the same model; falls
    back to the C restatement of the oracle when that build is absent."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_fields")
    sample = list(obs_x[:nsample])
    if os.path.exists(exe):
        with open(os.path.join(case_dir, "ref_fields.in"), "w") as f:
            f.write("model.smesh\nmoho.dat\nsed.dat\nnone\n")
            f.write(f"{len(sample)}\n" + "".join(f"{x}\n" for x in sample) + "12\n")
        cmd = [exe]
        if os.path.exists("/usr/bin/taskset"):
            cmd = ["/usr/bin/taskset", "-c", "0", exe]
        r = subprocess.run(cmd, cwd=case_dir, capture_output=True, text=True, check=True)
        sec = float([ln for ln in r.stdout.splitlines() if "stage loop seconds" in ln][-1].split()[-1])
        compiler = open(os.path.join(ROOT, "oracle", "_ref", "COMPILER")).read().strip()
        kind = "reference"
        desc = f"reference modules ({compiler}), 1 core, run_forward2d stage schedule"
    else:
        from oracle.pyoracle import OracleModel
        from fmt2d_amd import formats
        sm = formats.read_smesh(os.path.join(case_dir, "model.smesh"))
        sx, sz = formats.read_refl(os.path.join(case_dir, "sed.dat"))
        mx, mz = formats.read_refl(os.path.join(case_dir, "moho.dat"))
        om = OracleModel(sm, sx, sz, mx, mz)
        t0 = time.perf_counter()
        for x in sample:
            om.forward_source(x, 12, headwave=True)
        sec = time.perf_counter() - t0
        kind = "port"
        desc = "C restatement of the reference FMM (gcc -O2), 1 core"
    return {"value": 5 * len(sample) / sec, "unit": "source-stages/s", "cores": 1, "kind": kind,
            "sample": f"{len(sample)} OBS x 12 fields of the same model ({desc}), {sec:.2f} s"}


def measured_traffic(workload_name: str):
    """HBM bytes per tile-kernel dispatch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy
    kernel of known size), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"fim_tiles_kernel_traffic_{workload_name}.txt")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("hbm_bytes_per_dispatch:"):
            return float(ln.split(":")[1])
    return None


def inversion_case_dir(d):
    """Plain inversion case in directory d (true and starting models, geometry
    without picks); returns (true model, starting model, geometry)."""
    from fmt2d_amd import synth
    m_true, g = synth.plain_pg_case()
    m0 = synth.make_model(dx=0.5, checker_amp=0.0)
    return m_true, m0, g


def cpu_iteration_baseline(d, m0, gi):
    """The reference inversion (oracle/_ref/invert2d, built from the
    reference sources) on the same case, one core: seconds of its first
    iteration's forward (FMM + rays + Frechet) and LSQR solve, from its log."""
    from fmt2d_amd import synth
    exe = os.path.join(ROOT, "oracle", "_ref", "invert2d")
    if not os.path.exists(exe):
        return None
    synth.write_fmt2d_case(d, m0, gi)
    with open(os.path.join(d, "frechet.in"), "w") as f:
        f.write("1\n1\n1\n1\n0\n")
    with open(os.path.join(d, "invert2d.in"), "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
        for v in [1, 0.1, 0.3, 1.0, 400, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0]:
            f.write(f"{v}\n")
        f.write("1 0.0 0.5 0.3\n-1\n")
    cmd = [exe]
    if os.path.exists("/usr/bin/taskset"):
        cmd = ["/usr/bin/taskset", "-c", "0", exe]
    r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=600)
    fw = [ln for ln in r.stdout.splitlines() if "forward breakdown" in ln]
    ls = [ln for ln in r.stdout.splitlines() if "LSQR converged" in ln]
    lsf = [ln for ln in r.stdout.splitlines() if "no frechet" in ln]
    if len(fw) < 2 or not ls:
        return None
    nums = [float(x) for x in fw[1].replace(",", " ").split() if x.replace(".", "", 1).isdigit()]
    fwd = sum(nums[:3])
    lsqr = float(ls[0].split("time=")[1].split()[0])
    line = sum(float(ln.split(":")[1].split()[0]) for ln in lsf)
    return {"value": round(fwd + lsqr, 4), "unit": "s/iteration", "cores": 1, "kind": "reference",
            "line_search_s": round(line, 4),
            "sample": f"one iteration of the reference inversion on the same case: forward {fwd:.2f} s "
                      f"(FMM+rays+Frechet), LSQR {lsqr:.2f} s; its line search ({len(lsf)} forward(s) without "
                      f"Frechet rows) {line:.2f} s, reported apart"}


def inversion_iteration(steps: int, warmup: int, cpu: bool):
    """Seconds per inversion iteration on the GPU (see module docstring)."""
    from fmt2d_amd import formats, hip, synth
    m_true, m0, g = inversion_case_dir(None)
    with tempfile.TemporaryDirectory() as d:
        synth.write_fmt2d_case(d, m_true, g)
        sm = formats.read_smesh(os.path.join(d, "model.smesh"))
        sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
        mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
        geom = formats.read_geom(os.path.join(d, "geom.dat"))
        ctx = hip.Context(0)
        ctx.set_model(sm, sx, sz, mx, mz)
        ctx.forward(list(geom.obs_x))
        t, v = ctx.trace(geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
        ctx.close()
        gi = synth.with_picks(g, t, v)
        synth.write_fmt2d_case(d, m0, gi)
        sm = formats.read_smesh(os.path.join(d, "model.smesh"))
        geom = formats.read_geom(os.path.join(d, "geom.dat"))
        ctx = hip.Context(0)
        ctx.set_model(sm, sx, sz, mx, mz)
        ctx.set_inversion(True, [])
        parts = {"forward": [], "rays": [], "frechet_system": [], "lsqr": []}
        line = []
        iters = 0
        ls_forwards = 0
        for k in range(warmup + steps):
            ctx.set_model(sm, sx, sz, mx, mz)  # every step starts from the same model
            ctx.set_inversion(True, [])
            t0 = time.perf_counter()
            ctx.forward(list(geom.obs_x))
            t1 = time.perf_counter()
            ctx.trace(geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
            t2 = time.perf_counter()
            nd, nnz, eps = ctx.build_system(geom.t_obs, crit_chi=1.0, refl_weight=1.0)
            t3 = time.perf_counter()
            dm, dr = ctx.damped_update(lam_damp=0.1, lam_damp_iface=0.1, target_dv=0.05, max_rel_pert=0.3,
                                       w_smooth=0.3, w_smooth_iface=0.3, max_iter=400)
            t4 = time.perf_counter()
            dm2, ls = ctx.line_search(geom.t_obs, dm, crit_chi=1.0, iface_step_ratio=0.5, first_step=1)
            t5 = time.perf_counter()
            if k >= warmup:
                for key, dt in zip(parts, (t1 - t0, t2 - t1, t3 - t2, t4 - t3)):
                    parts[key].append(dt)
                line.append(t5 - t4)
            iters = dr.iterations
            ls_forwards = ls.forwards
        ctx.close()
        cpu_b = None
        if cpu:
            with tempfile.TemporaryDirectory() as d2:
                cpu_b = cpu_iteration_baseline(d2, m0, gi)
    per = {k: float(np.mean(v_)) for k, v_ in parts.items()}
    total = sum(per.values())
    return {"metric": "forward+Frechet+LSQR s/iteration", "value": round(total, 4), "unit": "s/iteration",
            "higher_is_better": False,
            "workload": "plain inversion case: 401x61 grid (0.5 km), 20 OBS, 399 shots, Pg; one iteration "
                        "(forward, rays, Frechet CSR, system, one LSQR solve, Laplacian smoothing, damping 0.1, "
                        "cap of the update)",
            "breakdown_ms": {k: round(1e3 * v_, 2) for k, v_ in per.items()},
            "line_search_ms": round(1e3 * float(np.mean(line)), 2), "line_search_forwards": ls_forwards,
            "ndata": nd, "nnz": nnz, "lsqr_iterations": iters, "cpu_baseline": cpu_b}


def run_program(exe, d, timeout=900):
    t0 = time.perf_counter()
    r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout)
    sec = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(f"{exe} failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
    return sec


def programs_figure(cpu: bool):
    """Wall seconds of the Fortran host programs on the plain case, and of the
    reference's programs on the same files (see module docstring)."""
    from fmt2d_amd import synth
    from fmt2d_amd.build import FMT2D, INVERT2D
    if not (os.path.exists(FMT2D) and os.path.exists(INVERT2D)):
        return None
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    out = {"workload": "plain case: 401x61 grid (0.5 km), 20 OBS, 399 shots per OBS at 10 m depth, Pg"}
    m_true, m0, g = inversion_case_dir(None)
    with tempfile.TemporaryDirectory() as d:
        synth.write_fmt2d_case(d, m_true, g)
        ours = run_program(FMT2D, d)
        arr = open(os.path.join(d, "arrivals.dat")).read()
        ref = None
        if cpu and os.path.exists(os.path.join(ref_dir, "fmt2d")):
            ref = run_program(os.path.join(ref_dir, "fmt2d"), d)
            ref_same = open(os.path.join(d, "arrivals.dat")).read() == arr
        out["fmt2d"] = {"seconds": round(ours, 3), "reference_seconds": None if ref is None else round(ref, 3),
                        "arrivals_identical": None if ref is None else ref_same}
        # picks from the forward program's arrivals (valid rays), as the
        # example scripts make them; then a two-iteration inversion
        rows = [ln.split() for ln in arr.splitlines()]
        t = np.array([float(r[3]) for r in rows])
        gi = synth.with_picks(g, t, t > 0)
        synth.write_fmt2d_case(d, m0, gi)
        with open(os.path.join(d, "frechet.in"), "w") as f:
            f.write("1\n1\n1\n1\n0\n")
        with open(os.path.join(d, "invert2d.in"), "w") as f:
            f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
            for v in [2, 0.0, 0.3, 1.0, 400, 0.03, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0]:
                f.write(f"{v}\n")
            f.write("1 0.0 0.5 0.3\n-1\n")
        ours = run_program(INVERT2D, d)
        log = open(os.path.join(d, "invert_log_latest.txt")).read()
        ref = None
        if cpu and os.path.exists(os.path.join(ref_dir, "invert2d")):
            ref = run_program(os.path.join(ref_dir, "invert2d"), d)
            ref_same = open(os.path.join(d, "invert_log_latest.txt")).read() == log
        out["invert2d"] = {"seconds": round(ours, 3), "reference_seconds": None if ref is None else round(ref, 3),
                           "iterations": 2, "log_identical": None if ref is None else ref_same}
    return out


def gradient_figure(steps: int):
    """Misfit and gradient at the adjoint-gradient workload (see module
    docstring) through fmt2d_amd.autograd: the forward with rays, the
    adjoint-state backward (fmt2d_adjoint_gradient) and, for comparison, the
    ray-based backward (Frechet rows + G^T r) timed apart."""
    import torch
    from fmt2d_amd import formats, hip, synth
    from fmt2d_amd.autograd import TravelTimes
    m_true, m0, g = synth.gradient_case()
    with tempfile.TemporaryDirectory() as d:
        synth.write_fmt2d_case(d, m_true, g)
        sm = formats.read_smesh(os.path.join(d, "model.smesh"))
        sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
        mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
        geom = formats.read_geom(os.path.join(d, "geom.dat"))
        ctx = hip.Context(0)
        ctx.set_model(sm, sx, sz, mx, mz)
        ctx.forward(list(geom.obs_x))
        t, v = ctx.trace(geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
        noise = np.random.default_rng(4).normal(0.0, 0.03, size=len(t))
        t_obs = torch.tensor(np.where(v, t + noise, np.nan), dtype=torch.float64, device="cuda:0")
        synth.write_fmt2d_case(d, m0, g)
        sm0 = formats.read_smesh(os.path.join(d, "model.smesh"))
        ctx.set_model(sm0, sx, sz, mx, mz)
        rays = (geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
        res = {}
        for mode in ("adjoint", "rays"):
            tt = TravelTimes(ctx, geom.obs_x, *rays, invert_velocity=True, ifaces=[4], backward=mode)
            vel = torch.tensor(ctx.model(0), dtype=torch.float64, device="cuda:0", requires_grad=True)
            moho = torch.tensor(ctx.model(4), dtype=torch.float64, device="cuda:0", requires_grad=True)
            fw, bw, stats = [], [], []
            for _ in range(steps):
                vel.grad = None
                moho.grad = None
                t0 = time.perf_counter()
                tc = tt(vel, moho)
                ok = torch.isfinite(tc) & torch.isfinite(t_obs)
                r = torch.where(ok, tc - torch.nan_to_num(t_obs), torch.zeros_like(tc))
                misfit = 0.5 * torch.sum(r * r)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                misfit.backward()
                torch.cuda.synchronize()
                t2 = time.perf_counter()
                fw.append(t1 - t0)
                bw.append(t2 - t1)
                stats.append(tt.last_stats)
            res[mode] = dict(fw=float(np.mean(fw)), bw=float(np.mean(bw)), stats=stats[-1], n_ok=int(ok.sum()),
                             misfit=float(misfit.detach()), gn=float(torch.linalg.vector_norm(vel.grad)),
                             gz=float(torch.linalg.vector_norm(moho.grad)))
        ctx.close()
    a, b = res["adjoint"], res["rays"]
    st = a["stats"]
    return {"metric": "misfit + gradient seconds", "value": round(a["fw"] + a["bw"], 4),
            "unit": "s", "higher_is_better": False, "backward": "adjoint state",
            "workload": "2001x301 grid (0.1 km), 60 OBS, 1000 shots, Pg at every offset and PmP at 3-30 km; "
                        "gradient with respect to every grid node's velocity and every Moho node's depth",
            "forward_rays_s": round(a["fw"], 4), "backward_s": round(a["bw"], 4),
            "adjoint": {"fields": st.fields, "rounds": st.rounds, "record_ms": round(st.ms_record, 1),
                        "sweep_ms": round(st.ms_reverse, 1), "total_ms": round(st.ms_total, 1),
                        "peak_gib": round(st.peak_bytes / 2**30, 3), "unmatched_nodes": st.leaves_unmatched},
            "ray_backward_s": round(b["bw"], 4),
            "ray_backward_note": "Frechet rows of compute_frechet_for_ray2d + G^T r (the reference's linearisation), "
                                 "inversion parameters only",
            "rays": len(geom.rec_x), "rays_with_picks_and_times": a["n_ok"],
            "misfit": a["misfit"], "grad_norm_velocity": a["gn"], "grad_norm_moho": a["gz"]}


def joint_cpu_sample(m0, g, t_obs, keep_obs=(50, 150)):
    """The reference inversion (oracle/_ref/invert2d, one core) on a bounded
    sample of the joint workload: the same starting model, grid and picks for
    two of its 200 OBS, its own parameterisation (the propagation grid: the
    reference cannot run the coarser mesh), velocity + sediment base + Moho.
    Timed: its first forward with Frechet rows (its own "forward breakdown"
    print), which scales with the OBS count.  The run is stopped there: on
    the 2001x301 grid the reference's LSQR set-up sorts the 542 k column
    norms with an O(n^2) exchange sort (sort_array, inversion2d.f90:6297-6312,
    called at :479), hours at this size, so no LSQR time stands beside it."""
    from fmt2d_amd import synth
    exe = os.path.join(ROOT, "oracle", "_ref", "invert2d")
    if not os.path.exists(exe):
        return None
    recs, k = [], 0
    for i, rl in enumerate(g.receivers):
        n = len(rl)
        if i in keep_obs:
            recs.append([(x, z, rt, float(t_obs[k + j]), u) for j, (x, z, rt, t, u) in enumerate(rl)])
        k += n
    gs = synth.Geometry(obs_x=g.obs_x[list(keep_obs)], obs_z=g.obs_z[list(keep_obs)], receivers=recs)
    line = None
    with tempfile.TemporaryDirectory() as d:
        synth.write_fmt2d_case(d, m0, gs)
        with open(os.path.join(d, "frechet.in"), "w") as f:
            f.write("1\n1\n1\n1\n2\n3 4\n")
        with open(os.path.join(d, "invert2d.in"), "w") as f:
            f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
            for v in [1, 0.1, 0.3, 1.0, 400, 0.03, 0.1, 0.3, 10.0, 1.0, 1.0, 1.0, 10.0, 3.0]:
                f.write(f"{v}\n")
            f.write("3 0.5 0.5 0.3\n0\n8.0\n")
        cmd = [exe]
        if os.path.exists("/usr/bin/taskset"):
            cmd = ["/usr/bin/taskset", "-c", "0", exe]
        t0 = time.perf_counter()
        proc = subprocess.Popen(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
        try:
            for ln in proc.stdout:
                if "forward breakdown" in ln:
                    line = ln
                    break
                if time.perf_counter() - t0 > 600:
                    break
        finally:
            proc.kill()
            proc.wait()
        wall = time.perf_counter() - t0
    if line is None:
        return {"error": "no forward breakdown from the reference within 600 s"}
    nums = [float(x) for x in line.replace(",", " ").replace("=", " ").split() if x.replace(".", "", 1).isdigit()]
    fmm, trace, fre = nums[:3]
    scale = len(g.obs_x) / len(keep_obs)
    compiler = open(os.path.join(ROOT, "oracle", "_ref", "COMPILER")).read().strip()
    return {"value": round((fmm + trace + fre) * scale, 2), "unit": "s per forward with Frechet rows (estimate)",
            "cores": 1, "kind": "reference",
            "sample": f"reference inversion ({compiler}, 1 core) on {len(keep_obs)} of the {len(g.obs_x)} OBS "
                      f"(x = {', '.join(str(float(g.obs_x[i])) for i in keep_obs)} km), same grid, model and "
                      f"picks, its own parameterisation (propagation grid): first forward with Frechet rows "
                      f"FMM {fmm:.2f} s + rays {trace:.2f} s + Frechet {fre:.2f} s, scaled by {scale:.0f} to all "
                      f"OBS; stopped after it ({wall:.1f} s): its LSQR set-up sorts 542 k column norms with an "
                      f"O(n^2) exchange sort (inversion2d.f90:6297-6312), hours at this size"}


def joint_figure(iters: int, cpu: bool = False):
    """Joint velocity + interface inversion on an inversion mesh coarser than
    the grid (see module docstring): the iterations of the alternating
    schedule, each phase timed."""
    from fmt2d_amd import formats, hip, synth
    obs_x = np.round(np.arange(0.5, 200.0, 1.0), 6)
    m_true, m0, g = synth.inversion_case(dx=0.1, obs_x=obs_x, shot_dx=0.2, phases=True)
    with tempfile.TemporaryDirectory() as d:
        def load(m):
            synth.write_fmt2d_case(d, m, g)
            return (formats.read_smesh(os.path.join(d, "model.smesh")), *formats.read_refl(os.path.join(d, "sed.dat")),
                    *formats.read_refl(os.path.join(d, "moho.dat")))
        true_in = load(m_true)
        geom = formats.read_geom(os.path.join(d, "geom.dat"))
        start_in = load(m0)
    rays = (geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    ctx = hip.Context(0)
    ctx.set_model(*true_in)
    ctx.forward(list(geom.obs_x))
    t_true, v_true = ctx.trace(*rays)
    noise = np.random.default_rng(5).normal(0.0, 0.05, size=len(t_true))
    t_obs = np.where(v_true, t_true + noise, -1.0)
    ctx.set_model(*start_in)
    ctx.set_inversion(True, [3, 4])
    ctx.set_inversion_mesh(5, 10)
    n_vel, n_par = ctx.inversion_info()
    vnx, vnz, n_fre = ctx.mesh_info()
    dm_sum = np.zeros(n_par)
    first = 1
    log, phases = [], {"forward": [], "rays": [], "system": [], "lsqr": [], "line_search": []}
    modes = {}
    for it in range(1, iters + 1):
        mode = 2 if it % 4 == 0 else 1        # three velocity iterations, then one interface iteration
        t0 = time.perf_counter()
        ctx.forward(list(geom.obs_x))
        t1 = time.perf_counter()
        reused = ctx.stats().launches == 0     # the previous line search left this model's fields
        t, valid = ctx.trace(*rays)
        t2 = time.perf_counter()
        nd, nnz, _ = ctx.build_system(t_obs, crit_chi=1.0, refl_weight=10.0, alt_mode=mode)
        t3 = time.perf_counter()
        dm, dres = ctx.damped_update(dm_sum, lam_damp=0.1, lam_damp_iface=0.1, smooth_type=3, w_smooth=0.3,
                                     w_smooth_iface=0.3, max_iter=400, corr_x=10.0, corr_z=3.0, corr_x_iface=8.0)
        t4 = time.perf_counter()
        if it == iters:
            # the last iteration's system solved once more in each LSQR
            # reduction mode (fmt2d_set_lsqr_mode): ordered (the reference's
            # rounding) and parallel (tree / wave reductions), same damping
            for lm in (0, 1):
                ctx.set_lsqr_mode(lm)
                modes[lm] = ctx.lsqr(dm_sum, wdv=dres.wd, wdv_iface=dres.wd_iface, w_smooth=0.3,
                                     w_smooth_iface=0.3, smooth_type=3, max_iter=400, corr_x=10.0, corr_z=3.0,
                                     corr_x_iface=8.0)
            ctx.set_lsqr_mode(0)
        t4b = time.perf_counter()
        alive = np.ascontiguousarray(valid, dtype=np.int32)
        dm, r = ctx.line_search(t_obs, dm, crit_chi=1.0, first_step=first, valid=alive)
        t5 = time.perf_counter()
        first = r.next_first_step
        dm_sum += dm
        for k, v in zip(phases, (t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4b)):
            phases[k].append(v)
        log.append({"iteration": it, "mode": "velocity" if mode == 1 else "interfaces", "ndata": nd, "nnz": nnz,
                    "rms": round(r.rms_orig, 6), "rms_new": round(r.rms_new, 6), "step": r.step,
                    "lsqr_iterations": dres.iterations, "line_search_forwards": r.forwards,
                    "forward_reused": bool(reused),
                    "seconds": round(t5 - t0 - (t4b - t4), 3)})
        print(f"joint iteration {it}: {log[-1]}", file=sys.stderr, flush=True)
    ctx.close()
    tot = [sum(v) for v in zip(*phases.values())]
    # LSQR: one solve per iteration (fixed damping); algorithmic bytes of an
    # iteration: G v and G^T u each read every non-zero (FP64 value + int32
    # index), plus the vectors of the hybrid system (3 m_aug + 8 n doubles)
    lsqr_iters = sum(e["lsqr_iterations"] for e in log)
    lsqr_s = float(np.sum(phases["lsqr"]))
    nnz_m = float(np.mean([e["nnz"] for e in log]))
    nd_m = float(np.mean([e["ndata"] for e in log]))
    m_aug = nd_m + 2 * n_par
    lsqr_bytes = 2 * nnz_m * 12 + (3 * m_aug + 8 * n_par) * 8
    lsqr_ms_it = 1e3 * lsqr_s / max(lsqr_iters, 1)
    lsqr_fig = {"iterations": lsqr_iters, "ms_per_iteration": round(lsqr_ms_it, 3),
                "algorithmic_bytes_per_iteration": int(lsqr_bytes),
                "achieved_GBps": round(lsqr_bytes / (lsqr_ms_it * 1e-3) / 1e9, 2),
                "frac_of_hbm_peak": round(lsqr_bytes / (lsqr_ms_it * 1e-3) / 1e9 / PEAK_HBM_GBS, 5)}
    last_bytes = 2 * log[-1]["nnz"] * 12 + (3 * (log[-1]["ndata"] + 2 * n_par) + 8 * n_par) * 8
    (dm0, s0), (dm1, s1) = modes[0], modes[1]
    for key, (dmx, sx) in (("last_system_ordered", modes[0]), ("last_system_parallel", modes[1])):
        ms_it = sx.ms / max(sx.iterations, 1)
        lsqr_fig[key] = {"iterations": sx.iterations, "ms_per_iteration": round(ms_it, 3),
                         "achieved_GBps": round(last_bytes / (ms_it * 1e-3) / 1e9, 2),
                         "frac_of_hbm_peak": round(last_bytes / (ms_it * 1e-3) / 1e9 / PEAK_HBM_GBS, 5)}
    lsqr_fig["last_system_parallel"]["max_rel_dm_vs_ordered"] = float(
        np.max(np.abs(dm1 - dm0)) / max(np.max(np.abs(dm0)), 1e-300))
    cpu_b = joint_cpu_sample(m0, g, t_obs) if cpu else None
    return {"metric": "joint inversion seconds per iteration", "value": round(float(np.mean(tot)), 3), "unit": "s",
            "higher_is_better": False,
            "workload": "2001x301 grid (0.1 km), 200 OBS every 1 km, 1000 shots every 0.2 km, five phases in the "
                        "reference's windows; velocity on a 0.5 km mesh (401x61) and sediment base + Moho every "
                        "1 km (2x201); alternating schedule (three velocity iterations, one interface iteration), "
                        "hybrid smoothing, damping 0.1, LSQR at most 400 iterations, line search",
            "rays": int(len(t_true)), "picks": int((t_obs >= 0).sum()), "mesh": [vnx, vnz], "n_par": n_par,
            "n_frechet_columns": n_fre,
            "phase_seconds": {k: round(float(np.mean(v)), 3) for k, v in phases.items()}, "lsqr": lsqr_fig,
            "cpu_baseline": cpu_b, "log": log}


def fp32_figure(steps: int = 2):
    """The multistage forward (2001x301, 60 OBS) with the tile solver's
    grid-node fields stored in FP32 (fmt2d_set_field_precision 32: T as an
    32-bit fixed point of 2^-24 s, gradients FP32, compute FP64)
    beside FP64 storage: time, bytes per node of the solver's field storage
    and, per stage field, the largest |dT| against the FP64 fields; the
    fields with nodes beyond 1e-4 s are named."""
    from fmt2d_amd import hip, formats
    m, g, desc = workload("multistage")
    with tempfile.TemporaryDirectory() as d:
        write_case(d, m, g)
        sm = formats.read_smesh(os.path.join(d, "model.smesh"))
        sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
        mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    obs = list(g.obs_x)
    out = {}
    for bits in (64, 32):
        ctx.set_field_precision(bits)
        ctx.forward(obs)
        walls, fims = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            ctx.forward(obs)
            walls.append(time.perf_counter() - t0)
            fims.append(ctx.stats().ms_fim)
        out[bits] = {"ms_per_step": round(1e3 * float(np.mean(walls)), 1),
                     "fim_kernel_ms_per_step": round(float(np.mean(fims)), 1),
                     "tile_visits": int(ctx.stats().tiles)}
        if bits == 64:
            ctx.field_snapshot()
    mdt, nd, nover = ctx.field_compare(1e-4)
    summ = ctx.summary()
    ctx.close()
    per = mdt.reshape(len(obs), 12)
    named = [{"obs": int(k // 12), "tf": int(k % 12 + 1), "nodes_over_1e-4": int(nover[k]),
              "max_abs_dt_s": float(mdt[k])} for k in np.flatnonzero(nover > 0)]
    return {"workload": desc, "fp64": out[64], "fp32": out[32],
            "solver_field_bytes_per_node": {"fp64": 25, "fp32": 13,
                                            "note": "T, dT/dx, dT/dz and the tie byte the tile solver reads and "
                                                    "writes per grid node and visit (model: velocity 8 B and "
                                                    "flags 2 B read in both)"},
            "max_abs_dt_per_stage_field_s": [float(x) for x in per.max(axis=0)],
            "fields": int(len(mdt)), "fields_with_nodes_over_1e-4": len(named),
            "nodes_over_1e-4": int(np.sum(nover)), "reached_slots": int(np.sum(summ["reached"])),
            "reach_mismatches": int(np.sum(nd)), "max_abs_dt_s": float(mdt.max()), "named_fields": named[:40]}


def forward_figure(workload_name: str, steps: int, warmup: int, cpu_sample: int, autotune: bool = False,
                   per_source: int = 0, level_steps: int = 0):
    """Source-stages per second of one workload's forward (the headline's
    measurement; see the module docstring), with the tile-kernel time (the
    dispatches' spans, stamped on the device clock inside the replayed
    graphs) and, when cpu_sample > 0, the reference timed on that many of
    the workload's OBS.  autotune: the tile schedule is first picked by
    fmt2d_autotune on this model and these OBS (before the warm-up, outside
    the timed steps); every schedule gives the same fields bit for bit.
    per_source > 0: the same forward solved one OBS at a time (the per-stage
    path of the reference's loop, each OBS's stage chain on its own) for
    that many OBS spread over the model, after the timed steps.
    level_steps > 0: that many more forwards in the level launch mode (one
    cooperative dispatch per stage level runs all its solver launches),
    timed the same way, with their fields checked against the graph
    mode's (hash of every field)."""
    from fmt2d_amd import hip, formats
    m, g, desc = workload(workload_name)
    with tempfile.TemporaryDirectory() as d:
        write_case(d, m, g)
        sm = formats.read_smesh(os.path.join(d, "model.smesh"))
        sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
        mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
        ctx = hip.Context(0)
        ctx.set_model(sm, sx, sz, mx, mz)
        obs = list(g.obs_x)
        sched = None
        if autotune:
            r = ctx.autotune(obs, level=0)
            sched = {"tile": [r.tile_x, r.tile_z], "max_sweeps": r.max_sweeps, "max_sweeps_heavy": r.max_sweeps_heavy,
                     "colors": r.colors, "window_tiles": r.window_tiles, "tuner_trials": r.trials,
                     "tuner_ms_default": round(r.ms_start, 1), "tuner_ms_chosen": round(r.ms_best, 1)}
            print(f"autotune: {sched}", file=sys.stderr, flush=True)
        for _ in range(warmup):
            ctx.forward(obs)
        walls, fims, nodes, launches, tiles, disp = [], [], 0, [], [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            ctx.forward(obs)
            walls.append(time.perf_counter() - t0)
            st = ctx.stats()
            fims.append(st.ms_fim / 1e3)
            nodes = st.region_nodes
            launches.append(st.launches)
            tiles.append(st.tiles)
            disp.append(st.dispatches)
        single = None
        if per_source > 0:
            pick = sorted({int(round(k * (len(obs) - 1) / max(per_source - 1, 1))) for k in range(per_source)})
            ctx.forward([obs[pick[0]]])  # field storage for one OBS, outside the timing
            sw, sl = [], []
            for k in pick:
                t0 = time.perf_counter()
                ctx.forward([obs[k]])
                sw.append(time.perf_counter() - t0)
                sl.append(ctx.stats().launches)
            per_obs = float(np.mean(sw))
            single = {"obs_timed": [obs[k] for k in pick], "seconds_per_obs": round(per_obs, 4),
                      "launches_per_obs": int(np.mean(sl)), "seconds_all_obs_estimate": round(per_obs * len(obs), 2)}
        level = None
        if level_steps > 0:
            ctx.forward(obs)
            ref_hash = ctx.summary()["hash"].copy()
            try:
                ctx.set_solver_launch("level")
                ctx.forward(obs)
                lw, ld = [], []
                for _ in range(level_steps):
                    t0 = time.perf_counter()
                    ctx.forward(obs)
                    lw.append(time.perf_counter() - t0)
                    ld.append(ctx.stats().dispatches)
                same = bool(np.array_equal(ctx.summary()["hash"], ref_hash))
            finally:
                ctx.set_solver_launch("graph")
            level = {"ms_per_step": round(1e3 * float(np.mean(lw)), 1), "solver_dispatches_per_step": int(np.mean(ld)),
                     "fields_same_bits_as_graph_mode": same}
        cpu = cpu_baseline(d, obs, cpu_sample) if cpu_sample > 0 else None
        ctx.close()
    return dict(desc=desc, obs=obs, walls=walls, fims=fims, nodes=nodes, launches=launches, tiles=tiles, disp=disp,
                cpu=cpu, schedule=sched, single=single, level=level)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batched", choices=["batched", "multistage", "plain"])
    ap.add_argument("--cpu-sample", type=int, default=12, help="OBS timed on the host CPU (0: skip)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile schedule (8x8 tiles, caps 8/8, window 3, two colours)")
    ap.add_argument("--no-multistage", action="store_true", help="skip the 60-OBS multistage figure")
    ap.add_argument("--no-iteration", action="store_true", help="skip the inversion-iteration figure")
    ap.add_argument("--no-programs", action="store_true", help="skip the host-program figure")
    ap.add_argument("--no-gradient", action="store_true", help="skip the misfit-gradient figure")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 field-storage figure")
    ap.add_argument("--no-per-source", action="store_true", help="skip the one-OBS-at-a-time forward figure")
    ap.add_argument("--no-level", action="store_true", help="skip the level-launch-mode forward figure")
    ap.add_argument("--joint", type=int, default=4,
                    help="iterations of the joint inversion on the 0.5 km / 1 km mesh (2001x301, 200 OBS, five "
                         "phases; about a minute for 4); 0 skips it")
    ap.add_argument("--joint-only", action="store_true", help="only the joint-inversion figure")
    ap.add_argument("--iteration-only", action="store_true",
                    help="only the inversion-iteration figure (for profiling its kernels)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 times): the reference point that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        import torch
        n = args.copy_probe_mib << 20
        src = torch.ones(n // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "dispatches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    if args.joint_only:
        print(json.dumps({"joint": joint_figure(max(args.joint, 1), args.cpu_sample > 0)}))
        return

    if args.iteration_only:
        print(json.dumps({"iteration": inversion_iteration(max(1, min(args.steps, 3)), args.warmup,
                                                           args.cpu_sample > 0)}))
        return

    head = forward_figure(args.workload, args.steps, args.warmup, args.cpu_sample, autotune=not args.no_autotune,
                          per_source=0 if args.no_per_source else 4, level_steps=0 if args.no_level else 2)
    print(f"headline forward done: {head['walls']}", file=sys.stderr, flush=True)
    ms = None
    if args.workload != "multistage" and not args.no_multistage:
        f = forward_figure("multistage", max(1, min(args.steps, 3)), 1, 0)
        w = float(np.mean(f["walls"]))
        ms = {"metric": "eikonal source-stages/s", "value": round(5 * len(f["obs"]) / w, 3),
              "unit": "source-stages/s", "workload": f["desc"], "ms_per_step": round(1e3 * w, 3),
              "solver_launches_per_step": int(np.mean(f["launches"])), "tiles_per_step": int(np.mean(f["tiles"])),
              "fim_kernel_ms_per_step": round(1e3 * float(np.mean(f["fims"])), 3)}
        print(f"multistage figure done: {ms['value']}", file=sys.stderr, flush=True)
    f32 = None if args.no_fp32 else fp32_figure()
    it = None if args.no_iteration else inversion_iteration(max(1, min(args.steps, 3)), 1, args.cpu_sample > 0)
    progs = None if args.no_programs else programs_figure(args.cpu_sample > 0)
    grad = None if args.no_gradient else gradient_figure(max(1, min(args.steps, 2)))
    joint = joint_figure(args.joint, args.cpu_sample > 0) if args.joint > 0 else None
    walls, fims, nodes, disp = head["walls"], head["fims"], head["nodes"], head["disp"]
    obs = head["obs"]
    wall = float(np.sum(walls))
    stages = 5 * len(obs)
    value = stages * args.steps / wall
    fim_s = float(np.mean(fims))
    alg_bytes = nodes * BYTES_PER_NODE
    # per dispatch: the step's algorithmic bytes spread over its dispatches,
    # over the mean dispatch duration (the same ratio as bytes / summed time)
    ndisp = float(np.mean(disp))
    achieved = (alg_bytes / ndisp) / (fim_s / ndisp) / 1e9
    traffic = measured_traffic(args.workload)
    out = {
        "metric": "eikonal source-stages/s",
        "value": round(value, 3),
        "unit": "source-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded model generator of the FMT2D model family)",
        "config": {"workload": head["desc"], "nobs": len(obs), "source_stages_per_step": stages,
                   "fields_per_step": 12 * len(obs), "region_nodes_per_step": int(nodes),
                   "solver_launches_per_step": int(np.mean(head["launches"])),
                   "tiles_per_step": int(np.mean(head["tiles"])), "fim_kernel_ms_per_step": round(1e3 * fim_s, 3),
                   "fim_kernel_dispatches_per_step": int(ndisp),
                   "fim_kernel_us_per_dispatch": round(1e6 * fim_s / ndisp, 2),
                   "solver_launch": "8-launch batches captured once per dependency level in a HIP graph and "
                                    "replayed; dispatch spans from the device clock",
                   "schedule": head["schedule"]},
        "roofline": {"bound": "hbm", "kernel": "fim_tiles_kernel", "achieved": round(achieved, 3),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 6),
                     "traffic": traffic,
                     "algorithmic_bytes": f"{BYTES_PER_NODE} B per region node per field (FP64 velocity "
                                          f"read, T and grad T written) = {alg_bytes} B per step, "
                                          f"{alg_bytes / ndisp:.0f} B per dispatch"},
        "cpu_baseline": head["cpu"],
        "level_launch": head["level"],
        "per_source": None if head["single"] is None else dict(
            head["single"], workload="the same forward solved one OBS at a time (each OBS's twelve stage fields "
                                     "on their own, as the reference's per-source loop runs them), timed on the "
                                     "OBS listed and scaled to all of them",
            batched_speedup=round(head["single"]["seconds_all_obs_estimate"] / (wall / args.steps), 1)),
        "multistage": ms,
        "fp32_storage": f32,
        "iteration": it,
        "programs": progs,
        "gradient": grad,
        "joint": joint,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
