#!/usr/bin/env python3
"""Explicit face-operator sweep on one MI355X (the 256^3 jittered, randomly
renumbered hex mesh; FP64 and FP32): achieved GB/s against the HBM roofline.

A step is one pass of every operator of the sweep over the device-resident
synthetic fields: interpolate(linear), gaussGreenGrad, div(Gauss linear),
div(Gauss upwind), laplacian(Gauss linear uncorrected), snGrad(uncorrected),
each in FP64 and in FP32 (benchmarks/bench_explicitOperators.cpp: the
"No allocation" sections).  Throughput counts algorithmic bytes: every
API-level input array once at its stored precision plus every output once
(per internal face / cell / boundary face coefficients in BYTES below).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
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

HBM_PEAK_GBS = 8000.0

# bytes per internal face, per cell, per boundary face (int32 indices)
BYTES = {
    ("interp_linear", "f64"): (24, 8, 16), ("interp_linear", "f32"): (16, 4, 8),
    ("grad", "f64"): (40, 40, 36), ("grad", "f32"): (28, 20, 20),
    ("div_linear", "f64"): (24, 24, 20), ("div_linear", "f32"): (16, 12, 12),
    ("div_upwind", "f64"): (16, 24, 20), ("div_upwind", "f32"): (12, 12, 12),
    ("laplacian", "f64"): (32, 24, 36), ("laplacian", "f32"): (20, 12, 20),
    ("sngrad", "f64"): (24, 8, 28), ("sngrad", "f32"): (16, 4, 16),
}
OPS = ["interp_linear", "grad", "div_linear", "div_upwind", "laplacian", "sngrad"]


def alg_bytes(op, prec, N, F, B):
    f, c, b = BYTES[(op, prec)]
    return f * F + c * N + b * B


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_case(n, seed):
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc

    t0 = time.time()
    poly = fa.PolyMesh.hex(n, jitter=0.2, jitter_seed=seed, perm_seed=seed)
    log(f"hex mesh {n}^3 generated in {time.time() - t0:.1f}s")
    ctx = fa.Context(0)
    t0 = time.time()
    mesh = fa.Mesh.from_poly(ctx, poly)
    log(f"device mesh (GPU geometry) built in {time.time() - t0:.1f}s")
    del poly
    N, F, B = mesh.n_cells, mesh.n_internal_faces, mesh.n_boundary_faces
    bc = fa.BoundaryConditions(mesh, ["zeroGradient"] * mesh.n_patches)
    # T ~ U[-1,1] per cell; zeroGradient walls give T_b = T_P
    T = ctx.uniform(N, -1.0, 1.0, seed, 0)
    Tb = ctx.zeros((B,))
    Tf = fvcc.VolumeField(mesh, T, Tb, bc, "T")
    Tf.correct_boundary_conditions()
    # phi = Sf . interpolate(U), U_i ~ U[-1,1], U_b = U_P
    U = ctx.uniform(3 * N, -1.0, 1.0, seed, 1)
    fc = mesh.get("faceCells")
    Uh = U.numpy().reshape(N, 3)
    Ub = ctx.upload(np.ascontiguousarray(Uh[fc]))
    del Uh
    Uf = fvcc.VolumeField(mesh, U, Ub)
    phi = fvcc.flux(ctx, mesh, Uf)
    U.close()
    Ub.close()
    gamma = fvcc.SurfaceField(mesh, ctx.uniform(F + B, 0.5, 2.0, seed, 2))
    fields = {"f64": (Tf, phi, gamma)}
    T32 = fvcc.VolumeField(mesh, Tf.internal.astype(np.float32), Tf.boundary.astype(np.float32), bc)
    fields["f32"] = (T32, fvcc.SurfaceField(mesh, phi.values.astype(np.float32)),
                     fvcc.SurfaceField(mesh, gamma.values.astype(np.float32)))
    ctx.sync()
    return ctx, mesh, bc, fields


def make_ops(ctx, mesh, fields):
    """Returns {(op, prec): launch()} over preallocated outputs."""
    from foamadapter_amd import fvcc

    nFB = mesh.n_internal_faces + mesh.n_boundary_faces
    ops = {}
    for prec, (T, phi, gamma) in fields.items():
        dt = np.float64 if prec == "f64" else np.float32
        out_f = fvcc.SurfaceField(mesh, ctx.empty((nFB,), dt))
        out_c = ctx.empty((mesh.n_cells,), dt)
        out_g = ctx.empty((mesh.n_cells, 3), dt)
        lin = fvcc.SurfaceInterpolation(ctx, mesh, ["linear"])
        grd = fvcc.GaussGreenGrad(ctx, mesh)
        dl = fvcc.GaussGreenDiv(ctx, mesh, ["Gauss", "linear"])
        du = fvcc.GaussGreenDiv(ctx, mesh, ["Gauss", "upwind"])
        lap = fvcc.GaussGreenLaplacian(ctx, mesh, ["Gauss", "linear", "uncorrected"])
        sng = fvcc.FaceNormalGradient(ctx, mesh, ["uncorrected"])
        ops[("interp_linear", prec)] = (lambda T=T, o=out_f, f=lin: f.interpolate(T, o))
        ops[("grad", prec)] = (lambda T=T, o=out_g, f=grd: f.grad(T, 1.0, o))
        ops[("div_linear", prec)] = (lambda T=T, p=phi, o=out_c, f=dl: f.div(o, p, T, 1.0))
        ops[("div_upwind", prec)] = (lambda T=T, p=phi, o=out_c, f=du: f.div(o, p, T, 1.0))
        ops[("laplacian", prec)] = (lambda T=T, g=gamma, o=out_c, f=lap: f.laplacian(o, g, T, 1.0))
        ops[("sngrad", prec)] = (lambda T=T, o=out_f, f=sng: f.faceNormalGrad(T, o))
    return ops


def cpu_baseline(mesh, fields, budget_s=20.0):
    """The oracle (serial C++ restatement, 1 thread) on the same mesh and fields,
    FP64 sweep, each op run until about budget_s/6 seconds have passed."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_cpu as orc

    g = mesh.geometry()
    T, phi, gamma = fields["f64"]
    Th, Tbh, ph, gh = T.internal.numpy(), T.boundary.numpy(), phi.numpy(), gamma.numpy()
    N, F, B = mesh.n_cells, mesh.n_internal_faces, mesh.n_boundary_faces
    kinds = [1] * mesh.n_patches
    calls = {
        "interp_linear": lambda: orc.interpolate(g, Th, Tbh, 0),
        "grad": lambda: orc.grad(g, Th, Tbh),
        "div_linear": lambda: orc.div(g, ph, Th, Tbh, 0),
        "div_upwind": lambda: orc.div(g, ph, Th, Tbh, 1),
        "laplacian": lambda: orc.laplacian(g, gh, Th, Tbh, kinds),
        "sngrad": lambda: orc.sngrad(g, Th, Tbh, kinds),
    }
    tot_b, tot_t, runs = 0.0, 0.0, {}
    for op, fn in calls.items():
        n, t = 0, 0.0
        while t < budget_s / len(calls) or n == 0:
            t0 = time.perf_counter()
            fn()
            t += time.perf_counter() - t0
            n += 1
        tot_b += n * alg_bytes(op, "f64", N, F, B)
        tot_t += t
        runs[op] = n
    return {"value": round(tot_b / tot_t / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle (serial C++ restatement, 1 thread) FP64 sweep of the 6 operators on "
                      f"the same {N}-cell mesh and fields, runs per op {runs}, {tot_t:.1f}s total"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the face-operator sweep has no multi-GPU path in "
              "this build (sharding is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        return 2

    ctx, mesh, bc, fields = build_case(args.n, args.seed)
    N, F, B = mesh.n_cells, mesh.n_internal_faces, mesh.n_boundary_faces
    ops = make_ops(ctx, mesh, fields)
    keys = [(op, p) for p in ("f64", "f32") for op in OPS]
    for _ in range(args.warmup):
        for k in keys:
            ops[k]()
    ctx.sync()
    per = {k: 0.0 for k in keys}
    ctx.record(62)
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        for i, k in enumerate(keys):
            ctx.record(2 * i)
            ops[k]()
            ctx.record(2 * i + 1)
        for i, k in enumerate(keys):
            per[k] += ctx.elapsed_ms(2 * i, 2 * i + 1)
    ctx.record(63)
    ctx.sync()
    t_wall = time.perf_counter() - t_wall
    total_ms = ctx.elapsed_ms(62, 63)
    ms_step = total_ms / args.steps
    bytes_step = sum(alg_bytes(op, p, N, F, B) for op, p in keys)
    # whole-job throughput: all algorithmic bytes of the K steps over the
    # event-bracketed time of the K steps
    value = bytes_step / (ms_step * 1e-3) / 1e9
    per_op = {}
    for (op, p) in keys:
        ms = per[(op, p)] / args.steps
        gbs = alg_bytes(op, p, N, F, B) / (ms * 1e-3) / 1e9
        per_op[f"{op}_{p}"] = {"ms": round(ms, 4), "GB/s": round(gbs, 1),
                               "frac_peak": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(keys, key=lambda k: per[k])
    dom_ms = per[dom] / args.steps
    dom_bytes = alg_bytes(dom[0], dom[1], N, F, B)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    rec = {
        "metric": "face-op GB/s vs HBM peak (explicit operator sweep, FP64+FP32)",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"explicit operator sweep, {args.n}^3 hex cells, interior "
                               f"vertices jittered +-0.2h, cell numbering randomly permuted "
                               f"(seed {args.seed}), numbering as given",
                   "cells": N, "internal_faces": F, "boundary_faces": B,
                   "ops": OPS, "precisions": ["f64", "f32"]},
        "per_op": per_op,
        "roofline": {"bound": "hbm", "kernel": f"{dom[0]}_{dom[1]}", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_launch": dom_bytes,
                     "avg_launch_ms": round(dom_ms, 4)},
        "wall_s_timed": round(t_wall, 3),
    }
    if not args.no_cpu:
        log("timing the CPU baseline (oracle, 1 thread)")
        rec["cpu_baseline"] = cpu_baseline(mesh, fields)
    print(json.dumps(rec), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
