#!/usr/bin/env python3
"""Benchmark of the MI355X forward hot path: eikonal source-stages per second.

One step = one forward pass of the multistage eikonal solver over all OBS of
the workload: for every OBS the twelve time fields of the five phases (Pg,
PsedP, PmP, Pg_ss, Pg_ss2), solved on the device with the inputs (model,
geometry) already resident.  A source-stage is one (OBS, phase) chain, so a
step completes 5 x nobs source-stages.

Prints ONE JSON line (rank 0).  Default workload: the full multistage forward
on the 0.1 km grid (2001x301 nodes), 60 OBS every 3.3 km.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="multistage", choices=["multistage", "plain"])
    ap.add_argument("--cpu-sample", type=int, default=4, help="OBS timed on the host CPU (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from fmt2d_amd import hip, formats
    m, g, desc = workload(args.workload)
    with tempfile.TemporaryDirectory() as d:
        write_case(d, m, g)
        sm = formats.read_smesh(os.path.join(d, "model.smesh"))
        sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
        mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
        ctx = hip.Context(0)
        ctx.set_model(sm, sx, sz, mx, mz)
        obs = list(g.obs_x)
        for _ in range(args.warmup):
            ctx.forward(obs)
        walls, fims, nodes, launches, tiles = [], [], 0, [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            ctx.forward(obs)
            walls.append(time.perf_counter() - t0)
            st = ctx.stats()
            fims.append(st.ms_fim / 1e3)
            nodes = st.region_nodes
            launches.append(st.launches)
            tiles.append(st.tiles)
        cpu = cpu_baseline(d, obs, args.cpu_sample) if args.cpu_sample > 0 else None
        ctx.close()

    wall = float(np.sum(walls))
    stages = 5 * len(obs)
    value = stages * args.steps / wall
    fim_s = float(np.mean(fims))
    alg_bytes = nodes * BYTES_PER_NODE
    achieved = alg_bytes / fim_s / 1e9
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
        "config": {"workload": desc, "nobs": len(obs), "source_stages_per_step": stages,
                   "fields_per_step": 12 * len(obs), "region_nodes_per_step": int(nodes),
                   "solver_launches_per_step": int(np.mean(launches)),
                   "tiles_per_step": int(np.mean(tiles)), "fim_kernel_ms_per_step": round(1e3 * fim_s, 3)},
        "roofline": {"bound": "hbm", "kernel": "fim_tiles_kernel", "achieved": round(achieved, 3),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 6),
                     "traffic": None,
                     "algorithmic_bytes": f"{BYTES_PER_NODE} B per region node per field (FP64 velocity "
                                          f"read, T and grad T written) = {alg_bytes} B per step"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
