#!/usr/bin/env python3
"""Face-operator sweep and implicit solves on one MI355X.

Headline (value): the explicit operator sweep on the 256^3 hex mesh with
jittered interior vertices and a randomly permuted cell numbering, FP64 and
FP32.  A step is one pass of every operator of the sweep over device-resident
synthetic fields: interpolate(linear), gaussGreenGrad, div(Gauss linear),
div(Gauss upwind), laplacian(Gauss linear uncorrected), snGrad(uncorrected),
in FP64 and in FP32 (benchmarks/bench_explicitOperators.cpp, "No allocation"
sections).  Throughput counts algorithmic bytes: every API-level input array
once at its stored precision plus every output once (per internal face / cell
/ boundary face coefficients in BYTES).

The device mesh is renumbered at import (Z-order of cell centres, the role of
OpenFOAM's renumberMesh); the caller's fields are mapped into that numbering once when
they are uploaded (cost reported as boundary_mapping).  The same sweep on the
numbering as given is reported beside it (as_given), and the results of the
two are compared.

Also reported: the implicit heat-transfer case (32^3, 10 Euler steps,
Jacobi-CG to 1e-10; the steps run twice and the second run is timed, the first
reported beside it) and the pressure-Poisson case (128^3 jittered, permuted,
FP64 CG vs FP32-inner/FP64-refined CG to 1e-8).

Also reported by default: 512 batched heat problems (batched_heat), the
served request stream (served: solves/s, p50/p99 latency, against plain FP64
per-request solves) and the neoIcoFoam PISO loop on a 64^3 lid-driven cavity
(icofoam_cavity: ms per time step); --no-batched / --no-served / --no-icofoam
skip them.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
                       [--no-implicit] [--no-as-given] [--no-batched] [--no-served]
                       [--no-icofoam]
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


# kernel that runs each (op, precision) of the sweep on a renumbered mesh, as rocprofv3 names it
KERNEL_OF = {
    ("interp_linear", "f64"): "k_interp4<double>", ("interp_linear", "f32"): "k_interp4<float>",
    ("sngrad", "f64"): "k_sngrad4<double>", ("sngrad", "f32"): "k_sngrad4<float>",
    ("grad", "f64"): "k_gather<double, GradOp<double>, 256, 256>",
    ("grad", "f32"): "k_gather<float, GradOp<float>, 256, 256>",
    ("div_linear", "f64"): "k_gather<double, DivOp<double, 0>, 512, 512>",
    ("div_linear", "f32"): "k_gather<float, DivOp<float, 0>, 1024, 512>",
    ("div_upwind", "f64"): "k_gather<double, DivOp<double, 1>, 512, 512>",
    ("div_upwind", "f32"): "k_gather<float, DivOp<float, 1>, 1024, 512>",
    ("laplacian", "f64"): "k_gather<double, LapOp<double>, 256, 256>",
    ("laplacian", "f32"): "k_gather<float, LapOp<float>, 1024, 512>",
}
TRAFFIC_PROFILE = ROOT / "profiles" / "faceops_sweep_256cube.md"
# the sources whose kernels the traffic profile measured; the profile records
# their hash, and a profile of other sources is not used
KERNEL_SOURCES = ("faceops.hip", "mesh.hip", "mesh.hpp", "common.hpp")


def kernel_stamp():
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update((ROOT / "foamadapter_amd" / "csrc" / f).read_bytes())
    return h.hexdigest()[:16]


def measured_traffic(kernel, n):
    """(HBM bytes per launch of the sweep's `kernel`, named by op and precision as
    "grad_f64", source note) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this sweep (256^3 only), provided the summary was measured on the
    current kernel sources (its recorded hash); otherwise (None, why)."""
    import re
    if n != 256 or not TRAFFIC_PROFILE.exists():
        return None, "no traffic profile for this size"
    text = TRAFFIC_PROFILE.read_text()
    m = re.search(r"kernel sources sha256\[:16\]: `([0-9a-f]{16})`", text)
    if not m or m.group(1) != kernel_stamp():
        return None, "the committed traffic profile measured other kernel sources"
    # rows: | `kernel symbol` | op prec | mean us | alg GB | alg TB/s | fetch GB | write GB | ...
    op = kernel.rsplit("_", 1)
    want = f"{op[0]} {op[1]}"
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 7 and cells[1] == want:
            try:
                hbm = float(cells[5]) + float(cells[6])
            except ValueError:
                return None, "unreadable profile row"
            return int(hbm * 1e9), f"{TRAFFIC_PROFILE.relative_to(ROOT)} ({m.group(1)})"
    return None, "kernel not in the traffic profile"


def roofline_record(dominant, n):
    """The bench line's roofline object for the sweep's dominant kernel:
    dominant = ((op, precision), mean launch ms, algorithmic bytes per launch)."""
    (op, prec), ms, nbytes = dominant
    achieved = nbytes / (ms * 1e-3) / 1e9
    traffic, note = measured_traffic(f"{op}_{prec}", n)
    return {"bound": "hbm", "kernel": f"{op}_{prec}", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "traffic_source": note, "bytes_per_launch": nbytes,
            "avg_launch_ms": round(ms, 4), "kernel_name": KERNEL_OF[(op, prec)]}


def stream_ceiling(ctx, n_doubles=1 << 28, runs=10):
    """The hand-written streaming kernels (fv_stream: copy a = b, triad a = b + s c,
    16-byte loads and stores) on 2 GiB arrays: the HBM rate a kernel that only
    streams reaches on this part, the practical ceiling of the face operators."""
    from foamadapter_amd import _lib
    a, b, c = (ctx.empty((n_doubles,)) for _ in range(3))
    b.fill(1.0)
    c.fill(2.0)
    out = {}
    # bytes moved per double of the arrays: copy 8 read + 8 written, triad 16 + 8,
    # read 8 (its block sums are 1/2048 of that)
    for mode, name, moved in ((0, "copy", 16), (1, "triad", 24), (2, "read", 8)):
        for _ in range(3):
            _lib.call("fv_stream", ctx.handle, mode, n_doubles, b.ptr, c.ptr, a.ptr, 0.5)
        ctx.sync()
        ctx.record(56)
        for _ in range(runs):
            _lib.call("fv_stream", ctx.handle, mode, n_doubles, b.ptr, c.ptr, a.ptr, 0.5)
        ctx.record(57)
        ms = ctx.elapsed_ms(56, 57) / runs
        gbs = moved * n_doubles / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1),
                     "frac_peak": round(gbs / HBM_PEAK_GBS, 4)}
    for x in (a, b, c):
        x.close()
    out["bytes"] = (f"{8 * n_doubles >> 30} GiB per array; copy moves 2 arrays, triad 3, "
                    f"read 1")
    return out


def mapback_cost(ctx, case, sweep, runs=10):
    """Per op of the sweep: the time to map its output back to the caller's
    numbering (the scatter of fv_permute through the renumbered mesh's map,
    oriented for snGrad) and the op's algorithmic GB/s with that mapping
    included."""
    from foamadapter_amd import _lib
    from foamadapter_amd.runtime import prec_of
    mesh = case.mesh
    N, F, B = mesh.n_cells, mesh.n_internal_faces, mesh.n_boundary_faces
    out, tot_ms = {}, 0.0
    for (op, prec), (fn, res) in sweep["ops"].items():
        face = op in ("interp_linear", "sngrad")
        mp = mesh.face_map if face else mesh.cell_map
        ncomp = 3 if op == "grad" else 1
        dst = ctx.empty(res.shape, res.dtype)
        args = (ctx.handle, prec_of(res.dtype), res.shape[0], ncomp, 1, 1 if op == "sngrad" else 0,
                mp.ptr, res.ptr, dst.ptr)
        _lib.call("fv_permute", *args)
        ctx.sync()
        ctx.record(54)
        for _ in range(runs):
            _lib.call("fv_permute", *args)
        ctx.record(55)
        ms = ctx.elapsed_ms(54, 55) / runs
        op_ms = sweep["per_op"][f"{op}_{prec}"]["ms"]
        tot_ms += op_ms + ms
        out[f"{op}_{prec}"] = {"map_ms": round(ms, 4),
                               "GB/s_with_map": round(alg_bytes(op, prec, N, F, B) /
                                                      ((op_ms + ms) * 1e-3) / 1e9, 1)}
        dst.close()
    step_b = sum(alg_bytes(op, p, N, F, B) for op, p in sweep["ops"])
    return {"per_op": out, "sweep_GB/s_with_map": round(step_b / (tot_ms * 1e-3) / 1e9, 1),
            "what": "each output scattered back to the caller's cell / face numbering after "
                    "the op (fv_permute), timed with HIP events"}


def alg_bytes(op, prec, N, F, B):
    f, c, b = BYTES[(op, prec)]
    return f * F + c * N + b * B


def cg_bytes(N, nnz, iters, reps, mixed, starts=1, finals=0):
    """Algorithmic bytes of Jacobi-PCG on an N-row CSR system with nnz entries
    (int32 columns and row offsets), run as CG on the symmetrically scaled
    system (unit diagonal, off-diagonal entries only):
    * each solve start scales the matrix: reads the diagonal (row offset,
      diagonal slot, value) and writes g and d = 0 (FP64 32 N, FP32 24 N), reads
      each off-diagonal entry's source index, column and value and writes its
      scaled value (FP64 24 (nnz - N), FP32 20 (nnz - N));
    * each FP64 residual r = b - A x (the start, every replacement of the mixed
      solver, and the `finals` explicit residuals a plain FP64 solve computes
      at exit) reads the matrix, x, b and writes r (12 nnz + 4(N+1) + 24 N);
    * each iteration reads the off-diagonal entries and the vectors it needs
      once: SpMV p, q; update d, p, r, q, g read, d, r written; direction r, p
      read, p written (FP64 12 (nnz - N) + 96 N, FP32 inner 8 (nnz - N) + 48 N)."""
    off = nnz - N
    scale = (20 * off + 24 * N) if mixed else (24 * off + 32 * N)
    res64 = 12 * nnz + 4 * (N + 1) + 24 * N
    it = (8 * off + 48 * N) if mixed else (12 * off + 96 * N)
    return starts * scale + (starts + reps + finals) * res64 + iters * it


def heat_assembly_bytes(N, F, B):
    """ddt + laplacian assembly (implicit): 56 / 40 / 36 bytes per internal face /
    cell / boundary face."""
    return 56 * F + 40 * N + 36 * B


def roof(bytes_, ms):
    gbs = bytes_ / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"algorithmic_GB": round(bytes_ / 1e9, 4), "effective_GB/s": round(gbs, 1),
            "frac_peak": round(gbs / HBM_PEAK_GBS, 4)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Case:
    """Device mesh + the sweep's fields in the mesh's numbering."""

    def __init__(self, ctx, poly, seed, renumber):
        import foamadapter_amd as fa
        from foamadapter_amd import fvcc

        t0 = time.time()
        self.mesh = mesh = fa.Mesh.from_poly(ctx, poly, renumber=renumber)
        self.setup_s = time.time() - t0
        N, F, B = mesh.n_cells, mesh.n_internal_faces, mesh.n_boundary_faces
        self.bc = fa.BoundaryConditions(mesh, ["zeroGradient"] * mesh.n_patches)
        # caller-numbered inputs: T ~ U[-1,1] per cell, U_i ~ U[-1,1] per cell,
        # gamma ~ U[0.5,2] per face (the generator indexes the caller's numbering)
        Tc = ctx.uniform(N, -1.0, 1.0, seed, 0)
        Uc = ctx.uniform(3 * N, -1.0, 1.0, seed, 1)
        Uc.shape = (N, 3)
        gc = ctx.uniform(F + B, 0.5, 2.0, seed, 2)
        ctx.sync()
        ctx.record(60)
        T = mesh.to_device(Tc, "cell")
        U = mesh.to_device(Uc, "cell")
        gamma = mesh.to_device(gc, "face")
        ctx.record(61)
        self.mapping_ms = ctx.elapsed_ms(60, 61) if renumber else 0.0
        Tf = fvcc.VolumeField(mesh, T, ctx.zeros((B,)), self.bc, "T")
        Tf.correct_boundary_conditions()          # zeroGradient: T_b = T_P
        fc = mesh.get("faceCells")
        Ub = ctx.upload(np.ascontiguousarray(U.numpy()[fc]))   # U_b = U_P
        phi = fvcc.flux(ctx, mesh, fvcc.VolumeField(mesh, U, Ub))
        self.fields = {"f64": (Tf, phi, fvcc.SurfaceField(mesh, gamma))}
        T32 = fvcc.VolumeField(mesh, T.astype(np.float32), Tf.boundary.astype(np.float32), self.bc)
        self.fields["f32"] = (T32, fvcc.SurfaceField(mesh, phi.values.astype(np.float32)),
                              fvcc.SurfaceField(mesh, gamma.astype(np.float32)))
        # free the caller-numbered copies (on a mesh that is not renumbered they
        # are the device fields themselves and stay)
        for dev, caller in ((T, Tc), (gamma, gc)):
            if dev is not caller:
                caller.close()
        for a in (U, Uc, Ub):
            a.close()
        ctx.sync()

    def ops(self, ctx):
        from foamadapter_amd import fvcc
        mesh = self.mesh
        nFB = mesh.n_internal_faces + mesh.n_boundary_faces
        ops = {}
        for prec, (T, phi, gamma) in self.fields.items():
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
            ops[("interp_linear", prec)] = (lambda T=T, o=out_f, f=lin: f.interpolate(T, o), out_f.values)
            ops[("grad", prec)] = (lambda T=T, o=out_g, f=grd: f.grad(T, 1.0, o), out_g)
            ops[("div_linear", prec)] = (lambda T=T, p=phi, o=out_c, f=dl: f.div(o, p, T, 1.0), out_c)
            ops[("div_upwind", prec)] = (lambda T=T, p=phi, o=out_c, f=du: f.div(o, p, T, 1.0), out_c)
            ops[("laplacian", prec)] = (lambda T=T, g=gamma, o=out_c, f=lap: f.laplacian(o, g, T, 1.0), out_c)
            ops[("sngrad", prec)] = (lambda T=T, o=out_f, f=sng: f.faceNormalGrad(T, o), out_f.values)
        return ops


def time_sweep(ctx, case, steps, warmup):
    N, F, B = case.mesh.n_cells, case.mesh.n_internal_faces, case.mesh.n_boundary_faces
    ops = case.ops(ctx)
    keys = [(op, p) for p in ("f64", "f32") for op in OPS]
    for _ in range(warmup):
        for k in keys:
            ops[k][0]()
    ctx.sync()
    per = {k: 0.0 for k in keys}
    ctx.record(62)
    for _ in range(steps):
        for i, k in enumerate(keys):
            ctx.record(2 * i)
            ops[k][0]()
            ctx.record(2 * i + 1)
        for i, k in enumerate(keys):
            per[k] += ctx.elapsed_ms(2 * i, 2 * i + 1)
    ctx.record(63)
    ctx.sync()
    total_ms = ctx.elapsed_ms(62, 63)
    ms_step = total_ms / steps
    bytes_step = sum(alg_bytes(op, p, N, F, B) for op, p in keys)
    per_op = {}
    for (op, p) in keys:
        ms = per[(op, p)] / steps
        gbs = alg_bytes(op, p, N, F, B) / (ms * 1e-3) / 1e9
        per_op[f"{op}_{p}"] = {"ms": round(ms, 4), "GB/s": round(gbs, 1),
                               "frac_peak": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(keys, key=lambda k: per[k])
    dom_ms = per[dom] / steps
    dom_bytes = alg_bytes(dom[0], dom[1], N, F, B)
    return {"ms_per_step": ms_step, "GB/s": bytes_step / (ms_step * 1e-3) / 1e9,
            "per_op": per_op, "dominant": (dom, dom_ms, dom_bytes), "ops": ops}


def bench_box(ctx, steps, warmup, dims=(400, 300, 140), seed=1007):
    """The sweep on a box whose sides are not powers of two (400 x 300 x 140 =
    16.8M cells, the 256^3 count; jittered, permuted, renumbered): its Z-order
    runs are ragged, which the staged gathers' lists must absorb.  Reports the
    sweep and per-op GB/s and, per workgroup size, the workgroups that take the
    per-cell path."""
    import foamadapter_amd as fa
    n = max(dims)
    poly = fa.PolyMesh.hex(dims, lengths=tuple(d / n for d in dims), jitter=0.2,
                           jitter_seed=seed, perm_seed=seed)
    case = Case(ctx, poly, seed, renumber=True)
    del poly
    sw = time_sweep(ctx, case, steps, warmup)
    m = case.mesh
    out = {"mesh": f"{dims[0]}x{dims[1]}x{dims[2]} jittered, permuted, renumbered",
           "cells": m.n_cells, "GB/s": round(sw["GB/s"], 1), "ms_per_step": round(sw["ms_per_step"], 3),
           "per_op_GB/s": {k: v["GB/s"] for k, v in sw["per_op"].items()},
           "gather_workgroups_per_cell_path": {str(c): list(m.gather_stats(c)) for c in (256, 512, 1024)}}
    for k in list(case.fields):
        del case.fields[k]
    m.close()
    return out


def cpu_baseline(case, sweep, budget_s=24.0):
    """The CPU restatement on the same mesh and fields, FP64 sweep of the 6
    operators: the multi-core form (OpenMP cell walks in ascending face order,
    the role of NeoN's OpenMP executor; OMP_NUM_THREADS threads) and the serial
    form (1 thread, NeoN's serial executor), each op repeated until about
    budget_s/12 seconds have passed.  Both return the serial bits.

    The same outputs check the GPU at full size: every operator of the sweep is
    run once more on the GPU and compared with the multi-core restatement (FP64)
    and the serial FP32 restatement (FP32), bit for bit."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_cpu as orc

    mesh = case.mesh
    g = mesh.geometry()
    T, phi, gamma = case.fields["f64"]
    Th, Tbh, ph, gh = T.internal.numpy(), T.boundary.numpy(), phi.numpy(), gamma.numpy()
    N, F, B = mesh.n_cells, mesh.n_internal_faces, mesh.n_boundary_faces
    kinds = [1] * mesh.n_patches
    omp = orc.OmpOps(g)
    forms = {
        "multicore": {
            "interp_linear": lambda: omp.interpolate(Th, Tbh),
            "grad": lambda: omp.grad(Th, Tbh),
            "div_linear": lambda: omp.div(ph, Th, Tbh, 0),
            "div_upwind": lambda: omp.div(ph, Th, Tbh, 1),
            "laplacian": lambda: omp.laplacian(gh, Th, Tbh, kinds),
            "sngrad": lambda: omp.sngrad(Th, Tbh, kinds),
        },
        "serial": {
            "interp_linear": lambda: orc.interpolate(g, Th, Tbh, 0),
            "grad": lambda: orc.grad(g, Th, Tbh),
            "div_linear": lambda: orc.div(g, ph, Th, Tbh, 0),
            "div_upwind": lambda: orc.div(g, ph, Th, Tbh, 1),
            "laplacian": lambda: orc.laplacian(g, gh, Th, Tbh, kinds),
            "sngrad": lambda: orc.sngrad(g, Th, Tbh, kinds),
        },
    }
    res, ref64 = {}, {}
    for form, calls in forms.items():
        tot_b, tot_t, runs = 0.0, 0.0, {}
        for op, fn in calls.items():
            out = fn()                             # warm (first touch, stencil in cache)
            if form == "multicore":
                ref64[op] = out
            n, t = 0, 0.0
            while t < budget_s / 2 / len(calls) or n == 0:
                t0 = time.perf_counter()
                fn()
                t += time.perf_counter() - t0
                n += 1
            tot_b += n * alg_bytes(op, "f64", N, F, B)
            tot_t += t
            runs[op] = n
        res[form] = (round(tot_b / tot_t / 1e9, 3), runs, tot_t)
    omp.close()
    # full-size parity of the GPU sweep (the checker: the oracle outputs above)
    log("comparing the 256^3 GPU sweep with the oracle (FP64 multi-core, FP32 serial)")
    ops = sweep["ops"]
    T32, phi32, gam32 = case.fields["f32"]
    t32, tb32, p32, g32 = (T32.internal.numpy(), T32.boundary.numpy(), phi32.numpy(),
                           gam32.numpy())
    ref32 = {
        "interp_linear": lambda: orc.interpolate(g, t32, tb32, 0, None, np.float32),
        "grad": lambda: orc.grad(g, t32, tb32, np.float32),
        "div_linear": lambda: orc.div(g, p32, t32, tb32, 0, np.float32),
        "div_upwind": lambda: orc.div(g, p32, t32, tb32, 1, np.float32),
        "laplacian": lambda: orc.laplacian(g, g32, t32, tb32, kinds, None, np.float32),
        "sngrad": lambda: orc.sngrad(g, t32, tb32, kinds, None, np.float32),
    }
    parity = {}
    for prec in ("f64", "f32"):
        same, worst = [], 0.0
        for op in OPS:
            fn, out = ops[(op, prec)]
            fn()
            got = out.numpy()
            want = ref64[op] if prec == "f64" else ref32[op]()
            want = np.asarray(want).reshape(got.shape)
            if np.array_equal(got, want):
                same.append(op)
            else:
                d = float(np.max(np.abs(got.astype(np.float64) - want)) /
                          max(float(np.max(np.abs(want))), 1e-300))
                worst = max(worst, d)
        parity[prec] = {"ops": len(OPS), "bitwise": len(same),
                        "max_rel_diff": worst,
                        "against": "oracle multi-core FP64" if prec == "f64" else "oracle serial FP32"}
    threads = orc.OmpOps.threads()
    v, runs, t = res["multicore"]
    sv, sruns, st = res["serial"]
    return {"value": v, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"CPU restatement (NeoN is not buildable offline): OpenMP cell walks in "
                      f"ascending face order, {threads} threads, FP64 sweep of the 6 operators "
                      f"on the same {N}-cell renumbered mesh and fields (results bitwise the "
                      f"serial oracle's), runs per op {runs}, {t:.1f}s",
            "serial": {"value": sv, "unit": "GB/s", "cores": 1,
                       "sample": f"serial oracle, runs per op {sruns}, {st:.1f}s"}}, parity


# repeated CG solves are replayed as captured hipGraphs unless --cg-eager (the form
# to trace with rocprofv3 --kernel-trace: every kernel is launched by itself)
CG_GRAPH = True


def bench_heat(ctx, n=32, steps=10, cpu=True):
    """Implicit heat transfer: ddt(T) - laplacian(1e-2, T), Euler dt=1e-3, 10 steps,
    Jacobi-CG to 1e-10; walls T=300 (x0), 400 (x1), 350 elsewhere.  Bytes: per step
    the assembly and the CG (iterations, FP64 residuals)."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc
    poly = fa.PolyMesh.hex(n)
    m = fa.Mesh.from_poly(ctx, poly)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    nnz = N + 2 * F
    po = m.get("patchOffsets")
    Tb = np.full(B, 350.0)
    Tb[po[0]:po[1]] = 300.0
    Tb[po[1]:po[2]] = 400.0
    bc = fa.BoundaryConditions(m, ["fixedValue"] * 6)
    res = {}
    for mode in ("double", "mixed"):
        kap = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, 1e-2)))
        # the 10 steps run twice from the same initial field: the first run pays
        # the one-time setup (sparsity pattern, linear system, solver workspace),
        # the second is the steady time-stepping cost reported as ms_per_step
        runs = []
        for run in range(2):
            T = fvcc.VolumeField(m, ctx.uniform(N, 300.0, 400.0, 1001, 0), ctx.upload(Tb), bc, "T")
            eqn = dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T)
            stats = []
            ctx.sync()
            t0 = time.perf_counter()
            for s in range(steps):
                st = dsl.solve(eqn, T, s * 1e-3, 1e-3, solver={"relTol": 1e-10, "maxIter": 1000,
                                                                 "precision": mode,
                                                                 "graph": CG_GRAPH})
                stats.append(st)
            ctx.sync()
            runs.append((time.perf_counter() - t0, stats, T.numpy()))
        wall, stats, Tf = runs[1]
        assert np.array_equal(Tf, runs[0][2]), "heat run is not reproducible"
        mixed = mode == "mixed"
        nbytes = sum(heat_assembly_bytes(N, F, B) +
                     cg_bytes(N, nnz, st.numIter, st.outerIter if mixed else 0, mixed,
                              finals=0 if mixed else 1)   # FP64: the explicit residual at exit
                     for st in stats)
        res[mode] = {"wall_ms": round(wall * 1e3, 2), "ms_per_step": round(wall * 1e3 / steps, 3),
                     "first_run_ms_per_step": round(runs[0][0] * 1e3 / steps, 3),
                     "cg_iters": [st.numIter for st in stats],
                     "fp64_residual_replacements": [st.outerIter for st in stats] if mixed else None,
                     "Tmin": float(Tf.min()), "Tmax": float(Tf.max()),
                     "roofline": roof(nbytes, wall * 1e3)}
        res[mode]["_T"] = Tf
    out = {"mesh": f"{n}^3", "cells": N}
    if cpu:
        out["cpu_baseline"], Tcpu = cpu_heat(poly, Tb, steps)
        for mode in ("double", "mixed"):
            Tg = res[mode]["_T"]
            res[mode]["rel_diff_vs_oracle"] = float(np.max(np.abs(Tg - Tcpu)) / np.max(np.abs(Tcpu)))
    for mode in ("double", "mixed"):
        del res[mode]["_T"]
    m.close()
    return {**out, **res}


def cpu_heat(poly, Tb, steps):
    """The heat case on the oracle: serial assembly per step and the multi-core
    Jacobi-PCG (orc_omp_cg, OMP_NUM_THREADS threads) to 1e-10; the run is the
    CPU baseline of the heat leg and its end state the parity check."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_cpu as orc
    g = orc.geometry(poly)
    sp = orc.sparsity(g)
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    gam = np.full(F + B, 1e-2)
    T = orc.uniform(N, 300.0, 400.0, 1001, 0)
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ls = orc.LinSys(g, sp)
        ls.ddt(T, 1e-3)
        ls.laplacian(gam, Tb, [0] * 6, coeff=-1.0)
        T, st = ls.cg(T, 1e-10, 1000, omp=True)
        iters.append(st["iters"])
    wall = time.perf_counter() - t0
    threads = orc.OmpOps.threads()
    return {"value": round(wall * 1e3 / steps, 3), "unit": "ms/step", "cores": threads,
            "kind": "port", "cg_iters": iters,
            "sample": f"oracle: serial assembly + multi-core Jacobi-PCG ({threads} threads) per "
                      f"step, all {steps} steps of the same case"}, T


def bench_poisson(ctx, n=128, seed=1003, repeats=3, cpu=True):
    """laplacian(1,p) = div(phi(U)), p=0 on x0, zeroGradient elsewhere, 128^3
    jittered + permuted (renumbered at import); FP64 CG vs FP32 CG in FP64
    refinement, both to 1e-8 relative residual.  Bytes: cg_bytes."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc, la
    t0 = time.time()
    poly = fa.PolyMesh.hex(n, jitter=0.2, jitter_seed=seed, perm_seed=seed)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    del poly
    log(f"poisson mesh {n}^3 ready in {time.time() - t0:.1f}s")
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    bc = fa.BoundaryConditions(m, ["fixedValue"] + ["zeroGradient"] * (m.n_patches - 1))
    p = fvcc.VolumeField(m, ctx.zeros((N,)), ctx.zeros((B,)), bc, "p")
    U = fvcc.VolumeField(m, ctx.uniform(3 * N, -1, 1, seed, 1), ctx.uniform(3 * B, -1, 1, seed, 2))
    phi = fvcc.flux(ctx, m, U)
    one = fvcc.SurfaceField(m, ctx.upload(np.ones(F + B)))
    ls = dsl.assemble(dsl.imp.laplacian(one, p) - dsl.exp.div(phi), p, 0.0, 1.0)
    out = {}
    xs = {}
    nnz = N + 2 * F
    for mode in ("double", "mixed"):
        best = None
        for _ in range(repeats):
            x = ctx.zeros((N,))
            ctx.sync()
            ctx.record(50)
            st = la.solve_cg(ls, x, {"relTol": 1e-8, "maxIter": 20000, "precision": mode,
                                     "checkEvery": 64, "graph": CG_GRAPH})
            ctx.record(51)
            ms = ctx.elapsed_ms(50, 51)
            if best is None or ms < best[0]:
                best = (ms, st)
            xs[mode] = x
        ms, st = best
        mixed = mode == "mixed"
        nbytes = cg_bytes(N, nnz, st.numIter, st.outerIter if mixed else 0, mixed,
                          finals=0 if mixed else 1)   # FP64: the explicit residual at exit
        out[mode] = {"ms": round(ms, 3), "solves_per_s": round(1e3 / ms, 2), "iters": st.numIter,
                     "outer": st.outerIter, "final_rel_res": st.finalResNorm,
                     "true_rel_res": st.trueResNorm, "roofline": roof(nbytes, ms)}
    a, b = xs["mixed"].numpy(), xs["double"].numpy()
    out["mixed_vs_fp64_rel_err"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["speedup_mixed_over_fp64"] = round(out["double"]["ms"] / out["mixed"]["ms"], 3)
    if cpu:
        out["cpu_baseline"] = cpu_poisson(ls, b)
    m.close()
    return {"mesh": f"{n}^3 jittered, permuted, renumbered", "cells": N, **out}


def cpu_poisson(ls, x_gpu, serial_iters=40):
    """The same assembled system (values and right-hand side downloaded from the
    GPU) solved by the oracle's multi-core Jacobi-PCG (orc_omp_cg) from x = 0 to
    1e-8: the CPU baseline of the Poisson leg, and its solution against the GPU's.
    The serial CG runs a bounded sample of serial_iters iterations beside it."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_cpu as orc
    A = ls.scipy()
    rhs = ls.rhs()
    N = A.shape[0]
    sp = {"rowOffs": np.ascontiguousarray(A.indptr, np.int32),
          "colIdxs": np.ascontiguousarray(A.indices, np.int32)}
    vals = np.ascontiguousarray(A.data)
    stats = np.zeros(3)
    x = np.zeros(N)
    t0 = time.perf_counter()
    orc.lib.orc_omp_cg(N, orc.P(sp["rowOffs"]), orc.P(sp["colIdxs"]), orc.P(vals), orc.P(rhs),
                       orc.P(x), 1e-8, 20000, orc.P(stats))
    wall = time.perf_counter() - t0
    xs = np.zeros(N)
    st2 = np.zeros(3)
    t1 = time.perf_counter()
    orc.lib.orc_cg(N, orc.P(sp["rowOffs"]), orc.P(sp["colIdxs"]), orc.P(vals), orc.P(rhs),
                   orc.P(xs), 0.0, serial_iters, orc.P(st2))
    ts = time.perf_counter() - t1
    threads = orc.OmpOps.threads()
    iters = int(stats[2])
    return {"value": round(1.0 / wall, 4), "unit": "solves/s", "ms": round(wall * 1e3, 1),
            "cores": threads, "kind": "port", "iters": iters, "final_rel_res": float(stats[1]),
            "gpu_fp64_rel_diff": float(np.linalg.norm(x - x_gpu) / np.linalg.norm(x_gpu)),
            "sample": f"oracle multi-core Jacobi-PCG ({threads} threads), the whole solve of the "
                      f"same {N}-row system from x = 0 to 1e-8 ({iters} iterations)",
            "serial": {"ms_per_iter": round(ts * 1e3 / serial_iters, 2), "cores": 1,
                       "sample": f"serial oracle CG, {serial_iters} iterations of the same solve",
                       "solve_ms_estimate": round(ts * 1e3 / serial_iters * iters, 1)}}


def _loguniform(u, lo, hi):
    return np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))


def bench_batched(ctx, P=512, seed=1004, steps=5, use_graph=True):
    """512 independent heat problems on cubic meshes n ~ logU[12, 96], kappa ~
    logU[1e-4, 1e-1], dt = 1e-3, T0 ~ U[300,400], walls ~ U[300,400], 5 Euler
    steps, to 1e-8.  Plain: one problem at a time, FP64 fv_solve_cg per step.
    New: per step, the 512 assemblies in one launch (fv_assemble_ddt_laplacian_many)
    and all solves in one batched launch sequence (FP32 CG with FP64 residual
    replacement, hipGraph chunks, autotuned tile shape).  use_graph=False launches
    the chunks' kernels one by one with the default shape (no autotuning, which
    replays graphs): the same kernels, for kernel-level profiling."""
    import foamadapter_amd as fa
    from foamadapter_amd import _lib, la
    from foamadapter_amd.la import solver_opts
    import ctypes as C

    sys.path.insert(0, str(ROOT / "tests"))
    u = lambda n, s: _uniform_host(n, 0.0, 1.0, seed, s)  # noqa: E731
    ns = np.rint(_loguniform(u(P, 0), 12, 96)).astype(int)
    kappas = _loguniform(u(P, 1), 1e-4, 1e-1)
    t0 = time.time()
    meshes, sps = {}, {}
    for n in sorted(set(ns)):
        m = fa.Mesh.from_poly(ctx, fa.PolyMesh.hex(int(n)))
        meshes[n] = m
        sps[n] = la.SparsityPattern(m)
    log(f"batched: {len(meshes)} distinct meshes built in {time.time() - t0:.1f}s")
    probs = []
    for i in range(P):
        m = meshes[ns[i]]
        N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
        walls = 300.0 + 100.0 * _uniform_host(6, 0.0, 1.0, seed + 1 + i, 9)
        po = m.get("patchOffsets") if i == 0 or ns[i] != ns[i - 1] else po
        Tb = np.concatenate([np.full(po[k + 1] - po[k], walls[k]) for k in range(6)])
        bc = getattr(m, "_bc_fv", None) or fa.BoundaryConditions(m, ["fixedValue"] * 6)
        m._bc_fv = bc
        T0 = ctx.uniform(N, 300.0, 400.0, seed + 1 + i, 0)
        probs.append({"m": m, "bc": bc, "Tb": ctx.upload(Tb), "T0": T0,
                      "kap": ctx.upload(np.full(F + B, kappas[i])),
                      "ls": la.LinearSystem(sps[ns[i]]), "N": N, "nnz": N + 2 * F,
                      "asm_b": heat_assembly_bytes(N, F, B)})
    ctx.sync()
    total_cells = int(sum(p["N"] for p in probs))
    log(f"batched: {P} problems, {total_cells} cells, set up in {time.time() - t0:.1f}s")
    solver = {"relTol": 1e-8, "maxIter": 5000}

    def run_plain():
        o = solver_opts(solver)
        st = _lib.FvSolverStats()
        iters, nbytes = 0, 0
        for p in probs:
            x = p["x"]
            _lib.call("fv_memcpy_d2d", ctx.handle, x.ptr, p["T0"].ptr, x.nbytes)
            for s in range(steps):
                _lib.call("fv_memcpy_d2d", ctx.handle, p["xo"].ptr, x.ptr, x.nbytes)
                _lib.call("fv_assemble_ddt_laplacian", p["ls"].handle, p["xo"].ptr, 1e-3,
                          p["kap"].ptr, p["bc"].handle, p["Tb"].ptr, None, -1.0)
                _lib.call("fv_solve_cg", p["ls"].handle, x.ptr, C.byref(o), C.byref(st))
                iters += st.iters
                nbytes += p["asm_b"] + cg_bytes(p["N"], p["nnz"], st.iters, 0, False, finals=1)
        ctx.sync()
        return iters, nbytes

    def run_batched(batch, heat):
        for p in probs:
            _lib.call("fv_memcpy_d2d", ctx.handle, p["x"].ptr, p["T0"].ptr, p["x"].nbytes)
            _lib.call("fv_memcpy_d2d", ctx.handle, p["xo"].ptr, p["T0"].ptr, p["x"].nbytes)
        iters, nbytes, cgb = 0, 0, 0
        for s in range(steps):
            heat.assemble(-1.0)                           # all 512 systems, one launch
            stats = batch.solve(solver, use_graph=use_graph)   # x_old <- x after the solve
            iters += sum(st.numIter for st in stats)
            c = sum(cg_bytes(p["N"], p["nnz"], st.numIter, st.outerIter, True)
                    for p, st in zip(probs, stats))
            cgb += c
            nbytes += c + sum(p["asm_b"] for p in probs)
        ctx.sync()
        run_batched.cg_bytes = cgb
        return iters, nbytes

    for p in probs:
        p["x"] = ctx.empty((p["N"],))
        p["xo"] = ctx.empty((p["N"],))
    out = {"problems": P, "steps": steps, "cells_total": total_cells,
           "mesh_edges": [int(ns.min()), int(ns.max())], "solves": P * steps}
    # plain FP64 per-problem drop-in
    t = time.perf_counter()
    it_plain, b_plain = run_plain()
    wall_plain = time.perf_counter() - t
    x_plain = [p["x"].numpy() for p in probs]
    out["plain_fp64"] = {"wall_s": round(wall_plain, 3), "solves_per_s": round(P * steps / wall_plain, 1),
                         "cg_iters": it_plain, "roofline": roof(b_plain, wall_plain * 1e3)}
    # batched, mixed precision, autotuned shape
    batch = la.Batch(ctx, "mixed")
    for p in probs:
        batch.add(p["ls"], p["x"], p["xo"])
    for p in probs:
        _lib.call("fv_memcpy_d2d", ctx.handle, p["x"].ptr, p["T0"].ptr, p["x"].nbytes)
        _lib.call("fv_memcpy_d2d", ctx.handle, p["xo"].ptr, p["T0"].ptr, p["x"].nbytes)
        _lib.call("fv_assemble_ddt_laplacian", p["ls"].handle, p["xo"].ptr, 1e-3, p["kap"].ptr,
                  p["bc"].handle, p["Tb"].ptr, None, -1.0)
    t = time.perf_counter()
    if use_graph:
        shape = batch.autotune(solver, str(ROOT / "foamadapter_amd" / "batch_tuning.json"))
    else:
        shape = (4096, 16)
        batch.set_shape(*shape)
    out["autotune"] = {"tile_rows": shape[0], "chunk": shape[1], "graphs": use_graph,
                       "seconds": round(time.perf_counter() - t, 2)}
    heat = la.HeatAssembly([(p["ls"], p["xo"], 1e-3, p["kap"], p["bc"], p["Tb"], None)
                            for p in probs])
    run_batched(batch, heat)                              # warm (graph capture)
    ev0 = batch.timing()
    t = time.perf_counter()
    it_b, b_b = run_batched(batch, heat)
    wall_b = time.perf_counter() - t
    ev1 = batch.timing()
    ev_ms = (ev1["chunk_ms"] - ev0["chunk_ms"]) + (ev1["start_ms"] - ev0["start_ms"])
    errs = np.array([float(np.max(np.abs(p["x"].numpy() - xp)) / np.max(np.abs(xp)))
                     for p, xp in zip(probs, x_plain)])
    out["batched_mixed"] = {"wall_s": round(wall_b, 3), "solves_per_s": round(P * steps / wall_b, 1),
                            "cg_iters": it_b, "roofline": roof(b_b, wall_b * 1e3),
                            "gpu_events": {"chunks": ev1["chunks"] - ev0["chunks"],
                                           "chunk_ms": round(ev1["chunk_ms"] - ev0["chunk_ms"], 3),
                                           "start_ms": round(ev1["start_ms"] - ev0["start_ms"], 3),
                                           "solve_roofline": roof(run_batched.cg_bytes, ev_ms),
                                           "what": "HIP events around every chunk graph launch "
                                                   "and batched start (assembly and host work "
                                                   "excluded)"},
                            "rel_diff_vs_plain_fp64": {"members": len(errs),
                                                       "max": float(errs.max()),
                                                       "median": float(np.median(errs))}}
    out["speedup"] = round(wall_plain / wall_b, 2)
    batch.close()
    for p in probs:
        for k in ("Tb", "T0", "kap", "x", "xo"):
            p[k].close()
        p["ls"].close()
    for m in meshes.values():
        m.close()
    return out


def _uniform_host(n, lo, hi, seed, stream):
    """Same counter-based generator as fv_fill_uniform."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)

    def mix(x):
        with np.errstate(over="ignore"):
            x = (x + np.uint64(0x9E3779B97F4A7C15)) & M
            x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M
            x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M
            return x ^ (x >> np.uint64(31))

    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = mix(np.uint64(seed) ^ mix(np.uint64(stream)))
        v = (mix(k + i) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * v


def bench_icofoam(ctx, n=64, seed=1006, warmup=2, steps=5):
    """neoIcoFoam (examples/neoIcoFoam/neoIcoFoam.cpp:80-180) on a lid-driven
    cavity: n^3 jittered, permuted mesh renumbered at import, lid (patch y1)
    U = (1, 0, 0), walls no-slip, nu = 0.01, dt = 0.5 h (Co about 0.5), upwind
    convection, momentum predictor and two pressure correctors, the solver
    settings of OpenFOAM's cavity tutorial (U: PBiCGStab to an absolute 1e-5;
    p: PCG to 1e-6 absolute or 0.05 relative).  Timed: `steps` whole time steps
    after `warmup`, host loop included (per step: three momentum solves, two
    pressure solves, the coupling kernels, the Courant number and the
    continuity error read back)."""
    import contextlib
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    from foamadapter_amd.piso import IcoFoam
    poly = fa.PolyMesh.hex(n, jitter=0.2, jitter_seed=seed, perm_seed=seed)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    del poly
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    po = m.get("patchOffsets")
    Ub = np.zeros((B, 3))
    Ub[po[3]:po[4], 0] = 1.0
    U = fvcc.VolumeField(m, ctx.zeros((N, 3)), ctx.upload(Ub),
                         fa.BoundaryConditions(m, ["fixedValue"] * 6), "U")
    p = fvcc.VolumeField(m, ctx.zeros((N,)), ctx.zeros((B,)),
                         fa.BoundaryConditions(m, ["zeroGradient"] * 6), "p")
    phi = fvcc.flux(ctx, m, U)
    nu = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, 0.01)), "nu")
    sol = IcoFoam(U, p, phi, nu, 0.5 / n, fvSchemes={"div": "Gauss upwind"},
                  fvSolution={"U": {"solver": "PBiCGStab", "tolerance": 1e-5, "relTol": 0.0},
                              "p": {"solver": "PCG", "tolerance": 1e-6, "relTol": 0.05,
                                    "graph": CG_GRAPH},
                              "PISO": {"nCorrectors": 2, "pRefCell": 0, "pRefValue": 0}})
    with contextlib.redirect_stdout(sys.stderr):      # the solvers' log lines
        for _ in range(warmup):
            sol.step()
        ctx.sync()
        t0 = time.perf_counter()
        reps = [sol.step() for _ in range(steps)]
        ctx.sync()
        wall = time.perf_counter() - t0
    last = reps[-1]
    out = {"mesh": f"{n}^3 jittered, permuted, renumbered", "cells": N, "steps": steps,
           "ms_per_step": round(wall * 1e3 / steps, 3),
           "cell_steps_per_s": round(N * steps / wall, 1),
           "coNum": round(last.coNum, 4), "sumLocalContErr": last.sumLocalContErr,
           "U_iters": [[s.numIter for s in r.U] for r in reps],
           "p_iters": [[s.numIter for s in r.p] for r in reps],
           "max_Ux": float(np.max(U.numpy()[:, 0]))}
    m.close()
    return out


def bench_served(n_req=96, seed=1005, rate=200.0, tile_rows=4096, chunk=16, rate_low=40.0,
                 solo_rows=0):
    """Served request stream: Poisson arrivals, 70% heat (5 Euler steps, kappa ~
    logU[1e-4,1e-1]) / 30% pressure Poisson, on jittered permuted hex cubes with
    edge n ~ logU[16,160] (snapped to a catalogue of resident meshes), 200 GB
    HBM budget, 1e-8 relative residual.  The same stream is run through the
    plain FP64 per-request path and through continuous batching (mixed), at the
    saturating rate and at a lower one; every mixed result is compared with the
    plain FP64 result of the same request."""
    import foamadapter_amd as fa
    from foamadapter_amd.serve import Server

    cat = [16, 20, 25, 32, 40, 50, 64, 80, 100, 128, 160]
    u = lambda k: _uniform_host(n_req, 0.0, 1.0, seed, k)  # noqa: E731
    n_draw = _loguniform(u(0), 16, 160)
    mesh_idx = [int(np.argmin(np.abs(np.log(cat) - np.log(v)))) for v in n_draw]
    kind = (u(1) < 0.3).astype(int)          # 1: Poisson
    kappa = _loguniform(u(2), 1e-4, 1e-1)
    gaps = -np.log(1.0 - u(3))               # unit-rate exponential inter-arrival gaps
    t0 = time.time()
    polys = [fa.PolyMesh.hex(n, jitter=0.2, jitter_seed=seed + n, perm_seed=seed + n) for n in cat]
    log(f"served: catalogue of {len(cat)} meshes generated in {time.time() - t0:.1f}s")
    budget = 200 << 30

    def run(label, rate_, keep, **opts):
        srv = Server(0, hbm_budget=budget, **opts)
        t1 = time.time()
        ids = [srv.add_mesh(p, renumber=True) for p in polys]
        log(f"served[{label}]: meshes resident in {time.time() - t1:.1f}s")
        # warm-up request per kind
        for r in (srv.submit_heat(ids[0], 1e-2, steps=1, seed=1), srv.submit_poisson(ids[0], seed=2)):
            srv.wait(r)
        srv_stats0 = srv.stats()
        rids = []
        start = time.perf_counter()
        tnext = start
        for i in range(n_req):
            if rate_:
                tnext += gaps[i] / rate_
                d = tnext - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            mid = ids[mesh_idx[i]]
            if kind[i]:
                rids.append(srv.submit_poisson(mid, seed=seed + 10 * i, rel_tol=1e-8))
            else:
                rids.append(srv.submit_heat(mid, kappa=float(kappa[i]), dt=1e-3, steps=5,
                                            seed=seed + 10 * i, rel_tol=1e-8))
        lat, bad, results, iters = [], 0, [], 0
        it_bytes = 0
        for i, r in enumerate(rids):
            x, st = srv.wait(r, ids[mesh_idx[i]] if keep else None)
            results.append(x)
            n = cat[mesh_idx[i]]
            N, F = n ** 3, 3 * n * n * (n - 1)
            if st.status == 0:
                lat.append(st.latency_ms)
                iters += st.iters
                it_bytes += cg_bytes(N, N + 2 * F, st.iters, 0, opts["precision"] == "mixed",
                                     starts=0)
            else:
                bad += 1
        wall = time.perf_counter() - start
        s = srv.stats()
        srv.close()
        lat = np.array(lat)
        return {"solves_per_s": round(len(lat) / wall, 3), "wall_s": round(wall, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "mean_ms": round(float(lat.mean()), 2), "failed": bad,
                "peak_active": s["peak_active"], "peak_gb": round(s["peak_bytes"] / 1e9, 2),
                "cg_iters": iters,
                "cg_iteration_roofline": roof(it_bytes, wall * 1e3),
                "warmup_completed": srv_stats0["completed"],
                "worker_ms": {k: round(s[k] - srv_stats0[k], 1)
                              for k in ("admit_ms", "gpu_ms", "advance_ms", "solo_ms")},
                "chunks": s["chunks"] - srv_stats0["chunks"],
                "table_builds": s["table_builds"] - srv_stats0["table_builds"],
                "graph_builds": s["graph_builds"] - srv_stats0["graph_builds"],
                "gpu_event_ms": {k: round(s[k] - srv_stats0[k], 1)
                                 for k in ("chunk_event_ms", "start_event_ms")},
                "solo_runs": s["solo_runs"] - srv_stats0["solo_runs"]}, results

    results = {}
    plain, xp = run("plain", rate, True, precision="double", batched=False)
    mixed, xm = run("batched", rate, True, precision="mixed", batched=True,
                    tile_rows=tile_rows, chunk=chunk, solo_rows=solo_rows)
    errs = np.array([float(np.linalg.norm(a - b) / np.linalg.norm(b))
                     for a, b in zip(xm, xp) if a is not None and b is not None])
    results["plain_fp64"] = plain
    results["batched_mixed"] = mixed
    results["batched_mixed"]["rel_err_vs_fp64_per_request"] = {
        "requests": int(len(errs)), "max": float(errs.max()), "median": float(np.median(errs))}
    results["speedup_solves_per_s"] = round(mixed["solves_per_s"] / plain["solves_per_s"], 2)
    if rate_low:
        lo_p, _ = run("plain-low", rate_low, False, precision="double", batched=False)
        lo_m, _ = run("batched-low", rate_low, False, precision="mixed", batched=True,
                      tile_rows=tile_rows, chunk=chunk, solo_rows=solo_rows)
        results["below_saturation"] = {
            "rate_per_s": rate_low,
            "plain_fp64": {k: lo_p[k] for k in ("solves_per_s", "p50_ms", "p99_ms", "mean_ms")},
            "batched_mixed": {k: lo_m[k] for k in ("solves_per_s", "p50_ms", "p99_ms", "mean_ms")}}
    results["stream"] = {"requests": n_req, "heat": int((kind == 0).sum()),
                         "poisson": int(kind.sum()), "rate_per_s": rate,
                         "mesh_edges": cat}
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-implicit", action="store_true")
    ap.add_argument("--no-as-given", action="store_true")
    ap.add_argument("--no-batched", dest="batched", action="store_false",
                    help="skip the 512-problem batched heat workload")
    ap.add_argument("--batched", dest="batched", action="store_true",
                    help="run the batched heat workload (the default)")
    ap.add_argument("--batch-problems", type=int, default=512)
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="batched leg without hipGraph replay and autotuning (for kernel "
                         "profiling)")
    ap.add_argument("--no-served", dest="served", action="store_false",
                    help="skip the served request stream")
    ap.add_argument("--served", dest="served", action="store_true",
                    help="run the served request stream (the default)")
    ap.add_argument("--requests", type=int, default=96)
    ap.add_argument("--served-chunk", type=int, default=16,
                    help="CG iterations per chunk of the served batch")
    ap.add_argument("--served-tile-rows", type=int, default=4096)
    ap.add_argument("--served-solo-rows", type=int, default=0,
                    help="cells from which a served request runs alone (0: the library's "
                         "default, negative: never)")
    ap.add_argument("--rate", type=float, default=200.0,
                    help="Poisson arrival rate (requests/s) of the served stream (0: all at once); "
                         "the default offers more than the plain path sustains, so both paths "
                         "are measured at saturation")
    ap.add_argument("--rate-low", type=float, default=40.0,
                    help="a second, lower arrival rate of the served stream (0: skip)")
    ap.add_argument("--no-icofoam", action="store_true")
    ap.add_argument("--no-box", action="store_true",
                    help="skip the sweep on the 400 x 300 x 140 box")
    ap.add_argument("--cg-eager", action="store_true",
                    help="launch the implicit legs' repeated CG solves kernel by kernel instead "
                         "of replaying them as hipGraphs (for kernel traces)")
    ap.add_argument("--ico-n", type=int, default=64, help="cells per edge of the icoFoam cavity")
    ap.add_argument("--gather-run", type=int, default=0,
                    help="bricks per XCD run of the staged gathers (0: the library default)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 2 GiB device-to-device 3 times (a known-size reference "
                         "for hardware traffic counters)")
    args = ap.parse_args()
    global CG_GRAPH
    CG_GRAPH = not args.cg_eager
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the finite-volume path runs on a single GPU "
              "(no cross-GPU partitioning is built); run with --gpus 1", file=sys.stderr)
        return 2
    import foamadapter_amd as fa

    ctx = fa.Context(0)
    if args.copy_probe:
        from foamadapter_amd import _lib
        nb = 2 << 30
        a, b = ctx.empty((nb // 8,)), ctx.empty((nb // 8,))
        a.fill(1.0)
        for _ in range(3):
            _lib.call("fv_memcpy_d2d", ctx.handle, b.ptr, a.ptr, nb)
        ctx.sync()
        a.close()
        b.close()
    t0 = time.time()
    poly = fa.PolyMesh.hex(args.n, jitter=0.2, jitter_seed=args.seed, perm_seed=args.seed)
    log(f"hex mesh {args.n}^3 generated in {time.time() - t0:.1f}s")
    case = Case(ctx, poly, args.seed, renumber=True)
    if args.gather_run:
        case.mesh.set_gather_run(args.gather_run)
    log(f"renumbered device mesh + fields ready (mesh setup {case.setup_s:.1f}s)")
    sweep = time_sweep(ctx, case, args.steps, args.warmup)
    N, F, B = case.mesh.n_cells, case.mesh.n_internal_faces, case.mesh.n_boundary_faces
    (dop, dprec), dom_ms, dom_bytes = sweep["dominant"]
    rec = {
        "metric": "face-op GB/s vs HBM peak (explicit operator sweep, FP64+FP32)",
        "value": round(sweep["GB/s"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sweep["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"explicit operator sweep: {args.n}^3 hex cells, interior vertices "
                               f"jittered +-0.2h, cell numbering randomly permuted (seed "
                               f"{args.seed}); mesh renumbered at import (Z-order of cell centres), "
                               f"fields mapped once at upload",
                   "cells": N, "internal_faces": F, "boundary_faces": B, "ops": OPS,
                   "precisions": ["f64", "f32"]},
        "per_op": sweep["per_op"],
        "roofline": roofline_record(sweep["dominant"], args.n),
        "boundary_mapping": {"ms": round(case.mapping_ms, 3),
                             "what": "permuting the caller's T, U (cells) and gamma (faces) into "
                                     "the renumbered order once at upload"},
    }
    rec["caller_numbering"] = mapback_cost(ctx, case, sweep)
    log("streaming ceiling (copy, triad)")
    rec["ceiling"] = stream_ceiling(ctx)
    if not args.no_as_given:
        g = Case(ctx, poly, args.seed, renumber=False)
        ag = time_sweep(ctx, g, max(3, args.steps // 4), 1)
        # same operator, both numberings: gradient mapped back to the caller's numbering
        ren_grad = case.mesh.to_caller(sweep["ops"][("grad", "f64")][1], "cell").numpy()
        ag_grad = ag["ops"][("grad", "f64")][1].numpy()
        diff = float(np.max(np.abs(ren_grad - ag_grad)) / np.max(np.abs(ag_grad)))
        rec["as_given"] = {"GB/s": round(ag["GB/s"], 1), "ms_per_step": round(ag["ms_per_step"], 3),
                           "per_op": ag["per_op"], "grad_f64_rel_diff_vs_renumbered": diff}
        for k in list(g.fields):
            del g.fields[k]
        g.mesh.close()
    del poly
    if not args.no_box:
        log("sweep on a 400 x 300 x 140 box")
        rec["ragged_box"] = bench_box(ctx, max(3, args.steps // 4), 1)
    if not args.no_cpu:
        log("timing the CPU baseline (oracle: OpenMP and serial)")
        rec["cpu_baseline"], rec["parity_full_size"] = cpu_baseline(case, sweep)
    case.mesh.close()
    if not args.no_implicit:
        log("implicit heat transfer 32^3")
        rec["implicit_heat"] = bench_heat(ctx, cpu=not args.no_cpu)
        log("pressure Poisson 128^3, FP64 vs mixed")
        rec["implicit_poisson"] = bench_poisson(ctx, cpu=not args.no_cpu)
    if not args.no_icofoam:
        log("icoFoam lid-driven cavity")
        rec["icofoam_cavity"] = bench_icofoam(ctx, n=args.ico_n)
    if args.batched:
        log("batched heat solves")
        rec["batched_heat"] = bench_batched(ctx, P=args.batch_problems,
                                            use_graph=not args.batch_no_graph)
    if args.served:
        log("served request stream")
        rec["served"] = bench_served(n_req=args.requests, rate=args.rate,
                                     tile_rows=args.served_tile_rows, chunk=args.served_chunk,
                                     rate_low=args.rate_low, solo_rows=args.served_solo_rows)
    print(json.dumps(rec), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
