"""ctypes wrapper of the CPU oracle (oracle/fv_oracle.cpp -> oracle/_build/liboracle.so).

Test infrastructure only: the checker the HIP library is compared against.
Nothing in the product package imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
# FV_ORACLE_SO selects another build of the same oracle (the sanitizer build)
ORACLE_SO = Path(os.environ.get("FV_ORACLE_SO", ROOT / "oracle" / "_build" / "liboracle.so"))


def _load():
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    return C.CDLL(str(ORACLE_SO))


lib = _load()


class _Arg:
    """ctypes argument that keeps its array alive for the duration of the call
    (a bare c_void_p of a temporary conversion would dangle)."""

    __slots__ = ("_as_parameter_", "arr")

    def __init__(self, a):
        self.arr = a
        self._as_parameter_ = C.c_void_p(a.ctypes.data)


def P(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return _Arg(a)


def _f(x, dt=np.float64):
    return np.ascontiguousarray(x, dtype=dt)


def _i(x):
    return np.ascontiguousarray(x, dtype=np.int32)


lib.orc_assemble_ddt.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_double, C.c_double, C.c_void_p, C.c_void_p]
lib.orc_assemble_source.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_double, C.c_void_p]
lib.orc_assemble_laplacian.argtypes = [C.c_int] * 4 + [C.c_void_p] * 16 + [C.c_double, C.c_void_p, C.c_void_p]
lib.orc_assemble_div.argtypes = [C.c_int] * 5 + [C.c_void_p] * 14 + [C.c_double, C.c_void_p, C.c_void_p]
lib.orc_laplacian_bcoeffs.argtypes = [C.c_int] * 3 + [C.c_void_p] * 7 + [C.c_double, C.c_void_p,
                                                                         C.c_void_p]
lib.orc_set_reference.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]
lib.orc_cg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_double, C.c_int, C.c_void_p]
lib.orc_omp_cg.argtypes = lib.orc_cg.argtypes


def omp_set_threads(n: int) -> int:
    """OpenMP threads of the multi-core restatement (n <= 0: unchanged); returns the count."""
    return lib.orc_omp_set_threads(int(n))


def geometry(poly) -> dict:
    """NeoN-layout geometry (meshAdapter.cpp:59-136) from polyMesh primitives."""
    nF, nI, N = poly.n_faces, poly.n_internal_faces, poly.n_cells
    pts = _f(poly.points)
    foff, fpts = _i(poly.face_offsets), _i(poly.face_points)
    own, nbr = _i(poly.owner), _i(poly.neighbour)
    Cf = np.zeros((nF, 3))
    Sf = np.zeros((nF, 3))
    lib.orc_face_geometry(nF, P(foff), P(fpts), P(pts), P(Cf), P(Sf))
    Cc = np.zeros((N, 3))
    V = np.zeros(N)
    lib.orc_cell_geometry(N, nF, nI, P(own), P(nbr), P(Cf), P(Sf), P(Cc), P(V))
    g = {"V": V, "C": Cc, "owner": own[:nI].copy(), "neighbour": nbr.copy(),
         "Sf": Sf[:nI].copy(), "Cf": Cf[:nI].copy()}
    magSf, w, dc, nodc = (np.zeros(nI) for _ in range(4))
    lib.orc_internal_coeffs(nI, P(g["owner"]), P(nbr), P(Cc), P(g["Cf"]), P(g["Sf"]), P(magSf),
                            P(w), P(dc), P(nodc))
    g.update(magSf=magSf, weights=w, deltaCoeffs=dc, nonOrthDeltaCoeffs=nodc)
    bfaces, poff = [], [0]
    for p in poly.patches:
        if not p.empty:
            bfaces.extend(range(p.start, p.start + p.size))
        poff.append(len(bfaces))
    bfaces = np.array(bfaces, np.int64)
    B = len(bfaces)
    fc = _i(own[bfaces]) if B else np.zeros(0, np.int32)
    Sfb = np.ascontiguousarray(Sf[bfaces]) if B else np.zeros((0, 3))
    Cfb = np.ascontiguousarray(Cf[bfaces]) if B else np.zeros((0, 3))
    magSfb, wb, dcb, nodcb = (np.zeros(B) for _ in range(4))
    nfb, deltab = np.zeros((B, 3)), np.zeros((B, 3))
    lib.orc_boundary_coeffs(B, P(fc), P(Cc), P(Cfb), P(Sfb), P(magSfb), P(nfb), P(deltab),
                            P(wb), P(dcb), P(nodcb))
    g.update(faceCells=fc, Sf_b=Sfb, Cf_b=Cfb, magSf_b=magSfb, nf_b=nfb, delta_b=deltab,
             weights_b=wb, deltaCoeffs_b=dcb, nonOrthDeltaCoeffs_b=nodcb,
             patchOffsets=np.array(poff, np.int32))
    return g


def _sizes(g):
    return len(g["V"]), len(g["owner"]), len(g["faceCells"])


def _cast(g, dt):
    return {k: (_f(v, dt) if v.dtype.kind == "f" else v) for k, v in g.items()}


def interpolate(g, T, Tb, scheme=0, phi=None, dt=np.float64):
    N, F, B = _sizes(g)
    gg = _cast(g, dt)
    out = np.zeros(F + B, dt)
    fn = lib.orc_interp_f64 if dt == np.float64 else lib.orc_interp_f32
    fn(scheme, F, B, P(gg["owner"]), P(gg["neighbour"]), P(gg["weights"]),
       P(_f(phi, dt)) if phi is not None else None, P(_f(T, dt)), P(_f(Tb, dt)), P(out))
    return out


def grad(g, T, Tb, dt=np.float64):
    N, F, B = _sizes(g)
    gg = _cast(g, dt)
    out = np.zeros((N, 3), dt)
    fn = lib.orc_grad_f64 if dt == np.float64 else lib.orc_grad_f32
    fn(N, F, B, P(gg["owner"]), P(gg["neighbour"]), P(gg["faceCells"]), P(gg["weights"]),
       P(gg["Sf"]), P(gg["Sf_b"]), P(gg["V"]), P(_f(T, dt)), P(_f(Tb, dt)), P(out))
    return out


def div(g, phi, T, Tb, scheme=0, dt=np.float64):
    N, F, B = _sizes(g)
    gg = _cast(g, dt)
    out = np.zeros(N, dt)
    fn = lib.orc_div_f64 if dt == np.float64 else lib.orc_div_f32
    fn(scheme, N, F, B, P(gg["owner"]), P(gg["neighbour"]), P(gg["faceCells"]), P(gg["weights"]),
       P(gg["V"]), P(_f(phi, dt)), P(_f(T, dt)), P(_f(Tb, dt)), P(out))
    return out


def _kinds(g, kinds):
    return _i(kinds), g["patchOffsets"]


def sngrad(g, T, Tb, kinds, Tg=None, dt=np.float64):
    N, F, B = _sizes(g)
    gg = _cast(g, dt)
    out = np.zeros(F + B, dt)
    k, poff = _kinds(g, kinds)
    fn = lib.orc_sngrad_f64 if dt == np.float64 else lib.orc_sngrad_f32
    fn(F, B, len(poff) - 1, P(poff), P(k), P(gg["owner"]), P(gg["neighbour"]),
       P(gg["faceCells"]), P(gg["nonOrthDeltaCoeffs"]), P(gg["deltaCoeffs_b"]), P(_f(T, dt)),
       P(_f(Tb, dt)), P(_f(Tg, dt)) if Tg is not None else None, P(out))
    return out


def laplacian(g, gamma, T, Tb, kinds, Tg=None, dt=np.float64):
    N, F, B = _sizes(g)
    gg = _cast(g, dt)
    out = np.zeros(N, dt)
    k, poff = _kinds(g, kinds)
    fn = lib.orc_laplacian_f64 if dt == np.float64 else lib.orc_laplacian_f32
    fn(N, F, B, len(poff) - 1, P(poff), P(k), P(gg["owner"]), P(gg["neighbour"]),
       P(gg["faceCells"]), P(gg["magSf"]), P(gg["nonOrthDeltaCoeffs"]), P(gg["magSf_b"]),
       P(gg["deltaCoeffs_b"]), P(gg["V"]), P(_f(gamma, dt)), P(_f(T, dt)), P(_f(Tb, dt)),
       P(_f(Tg, dt)) if Tg is not None else None, P(out))
    return out


def flux(g, U, Ub):
    """phi = Sf . (w (U_P - U_N) + U_N); boundary Sf_b . U_b (FoamAdapter::flux)."""
    own, nbr, w = g["owner"], g["neighbour"], g["weights"]
    u = w[:, None] * (U[own] - U[nbr]) + U[nbr]
    Sf = g["Sf"]
    phi = Sf[:, 0] * u[:, 0] + Sf[:, 1] * u[:, 1] + Sf[:, 2] * u[:, 2]
    Sfb = g["Sf_b"]
    phib = Sfb[:, 0] * Ub[:, 0] + Sfb[:, 1] * Ub[:, 1] + Sfb[:, 2] * Ub[:, 2]
    return np.concatenate([phi, phib])


def correct_boundary(g, T, Tb, kinds, Tg=None):
    Tb = np.array(Tb, dtype=np.float64, copy=True)
    poff = g["patchOffsets"]
    for p, k in enumerate(kinds):
        if k == 1:
            s, e = poff[p], poff[p + 1]
            grad_ = Tg[s:e] if Tg is not None else np.zeros(e - s)
            Tb[s:e] = T[g["faceCells"][s:e]] + grad_ / g["deltaCoeffs_b"][s:e]
    return Tb


# ---- implicit path -----------------------------------------------------------

def sparsity(g):
    N, F, B = _sizes(g)
    rowOffs = np.zeros(N + 1, np.int32)
    colIdxs = np.zeros(N + 2 * F, np.int32)
    diag = np.zeros(N, np.int32)
    oo = np.zeros(F, np.int32)
    no = np.zeros(F, np.int32)
    lib.orc_sparsity(N, F, P(g["owner"]), P(g["neighbour"]), P(rowOffs), P(colIdxs), P(diag),
                     P(oo), P(no))
    return {"rowOffs": rowOffs, "colIdxs": colIdxs, "diagOffset": diag, "ownerOffset": oo,
            "neighbourOffset": no}


class LinSys:
    def __init__(self, g, sp):
        self.g, self.sp = g, sp
        N = len(g["V"])
        self.values = np.zeros(len(sp["colIdxs"]))
        self.rhs = np.zeros(N)
        # boundary coefficients of the laplacian terms (face flux = bRhs - bMat T_P)
        self.bMat = np.zeros(len(g["faceCells"]))
        self.bRhs = np.zeros(len(g["faceCells"]))

    def ddt(self, Told, dt, coeff=1.0):
        sp = self.sp
        lib.orc_assemble_ddt(len(self.rhs), P(sp["rowOffs"]), P(sp["diagOffset"]), P(self.g["V"]),
                             P(_f(Told)), dt, coeff, P(self.values), P(self.rhs))

    def source(self, cfield, coeff=1.0):
        sp = self.sp
        lib.orc_assemble_source(len(self.rhs), P(sp["rowOffs"]), P(sp["diagOffset"]),
                                P(self.g["V"]), P(_f(cfield)), coeff, P(self.values))

    def laplacian(self, gamma, Tb, kinds, Tg=None, coeff=1.0):
        g, sp = self.g, self.sp
        N, F, B = _sizes(g)
        k, poff = _kinds(g, kinds)
        lib.orc_assemble_laplacian(N, F, B, len(poff) - 1, P(poff), P(k), P(g["owner"]),
                                   P(g["neighbour"]), P(g["faceCells"]), P(g["magSf"]),
                                   P(g["nonOrthDeltaCoeffs"]), P(g["magSf_b"]),
                                   P(g["deltaCoeffs_b"]), P(_f(gamma)), P(_f(Tb)),
                                   P(_f(Tg)) if Tg is not None else None, P(sp["rowOffs"]),
                                   P(sp["diagOffset"]), P(sp["ownerOffset"]),
                                   P(sp["neighbourOffset"]), coeff, P(self.values), P(self.rhs))
        lib.orc_laplacian_bcoeffs(F, B, len(poff) - 1, P(poff), P(k), P(g["magSf_b"]),
                                  P(g["deltaCoeffs_b"]), P(_f(gamma)), P(_f(Tb)),
                                  P(_f(Tg)) if Tg is not None else None, coeff, P(self.bMat),
                                  P(self.bRhs))

    def div(self, phi, Tb, kinds, Tg=None, scheme=0, coeff=1.0):
        g, sp = self.g, self.sp
        N, F, B = _sizes(g)
        k, poff = _kinds(g, kinds)
        lib.orc_assemble_div(scheme, N, F, B, len(poff) - 1, P(poff), P(k), P(g["owner"]),
                             P(g["neighbour"]), P(g["faceCells"]), P(g["weights"]),
                             P(g["deltaCoeffs_b"]), P(_f(phi)), P(_f(Tb)),
                             P(_f(Tg)) if Tg is not None else None, P(sp["rowOffs"]),
                             P(sp["diagOffset"]), P(sp["ownerOffset"]), P(sp["neighbourOffset"]),
                             coeff, P(self.values), P(self.rhs))

    def set_reference(self, cell, value):
        sp = self.sp
        lib.orc_set_reference(P(sp["rowOffs"]), P(sp["diagOffset"]), int(cell), float(value),
                              P(self.values), P(self.rhs))

    def residual(self, x):
        res = np.zeros(len(self.rhs))
        sp = self.sp
        lib.orc_residual(len(self.rhs), P(sp["rowOffs"]), P(sp["colIdxs"]), P(self.values),
                         P(self.rhs), P(_f(x)), P(res))
        return res

    def scipy(self):
        import scipy.sparse as sps
        N = len(self.rhs)
        return sps.csr_matrix((self.values, self.sp["colIdxs"], self.sp["rowOffs"]), shape=(N, N))

    def cg(self, x0, rel_tol, max_iter=1000, omp=False):
        """Jacobi-PCG from x0: the serial oracle, or (omp=True) its multi-core
        form (same recurrence, block-ordered dot products)."""
        x = np.array(x0, dtype=np.float64, copy=True)
        stats = np.zeros(3)
        sp = self.sp
        fn = lib.orc_omp_cg if omp else lib.orc_cg
        rc = fn(len(self.rhs), P(sp["rowOffs"]), P(sp["colIdxs"]), P(self.values),
                P(self.rhs), P(x), rel_tol, max_iter, P(stats))
        return x, {"init_res": stats[0], "final_res": stats[1], "iters": int(stats[2]), "rc": rc}


# ---- pressure-velocity coupling helpers (pressureVelocityCoupling.cpp:14-211) -------

def rau(g, sp, values):
    N, F, B = _sizes(g)
    out, outb = np.zeros(N), np.zeros(B)
    lib.orc_rau(N, B, P(sp["rowOffs"]), P(sp["diagOffset"]), P(_f(values)), P(g["V"]),
                P(g["faceCells"]), P(out), P(outb))
    return out, outb


def hbya(g, sp, values, b3, U, rAU):
    N, F, B = _sizes(g)
    out, outb = np.zeros((N, 3)), np.zeros((B, 3))
    lib.orc_hbya(N, F, B, P(g["owner"]), P(g["neighbour"]), P(sp["rowOffs"]),
                 P(sp["ownerOffset"]), P(sp["neighbourOffset"]), P(_f(values)), P(_f(b3)),
                 P(_f(U)), P(_f(rAU)), P(g["V"]), P(g["faceCells"]), P(out), P(outb))
    return out, outb


def update_face_velocity(g, sp, values, bMat, bRhs, pred, p):
    N, F, B = _sizes(g)
    out = np.zeros(F + B)
    lib.orc_update_face_velocity(F, B, P(g["owner"]), P(g["neighbour"]), P(sp["rowOffs"]),
                                 P(sp["ownerOffset"]), P(sp["neighbourOffset"]), P(_f(values)),
                                 P(g["faceCells"]), P(_f(bMat)), P(_f(bRhs)), P(_f(pred)),
                                 P(_f(p)), P(out))
    return out


def update_velocity(hbya_, rAU, gradp):
    N = len(rAU)
    out = np.zeros((N, 3))
    lib.orc_update_velocity(N, P(_f(hbya_)), P(_f(rAU)), P(_f(gradp)), P(out))
    return out


lib.orc_courant.restype = C.c_double
lib.orc_courant.argtypes = [C.c_int] * 3 + [C.c_void_p] * 5 + [C.c_double]


def courant(g, phi, dt):
    N, F, B = _sizes(g)
    return lib.orc_courant(N, F, B, P(g["owner"]), P(g["neighbour"]), P(g["faceCells"]),
                           P(_f(phi)), P(g["V"]), float(dt))


def constrain_hbya(g, ukinds, Ub, hbyab):
    k, poff = _kinds(g, ukinds)
    out = np.array(hbyab, dtype=np.float64, copy=True)
    lib.orc_constrain_hbya(len(g["faceCells"]), len(poff) - 1, P(poff), P(k), P(_f(Ub)), P(out))
    return out


# ---- counter-based RNG (same hash as the library's fv_fill_uniform) ----------------

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(x):
    with np.errstate(over="ignore"):
        x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M
        x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M
        x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M
        return x ^ (x >> np.uint64(31))


def uniform(n, lo, hi, seed, stream):
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = _mix64(np.uint64(seed) ^ _mix64(np.uint64(stream)))
        u = (_mix64(k + i) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


# ---- multi-core restatement (the benchmark's CPU baseline) -----------------

lib.orc_omp_stencil.restype = C.c_void_p
lib.orc_omp_stencil.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
lib.orc_omp_free.argtypes = [C.c_void_p]
lib.orc_omp_grad.argtypes = [C.c_void_p] * 10
lib.orc_omp_div.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 8
lib.orc_omp_laplacian.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 13
lib.orc_omp_interp.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6
lib.orc_omp_sngrad.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 10


class OmpOps:
    """FP64 interpolate/grad/div/laplacian/snGrad of the multi-core restatement
    (cell walks in ascending face order: the serial oracle's bits at any thread
    count).  Holds the cell->face stencil it builds for g."""

    def __init__(self, g):
        self.g = g
        N, F, B = _sizes(g)
        self.N, self.F, self.B = N, F, B
        self.h = lib.orc_omp_stencil(N, F, B, P(_i(g["owner"])), P(_i(g["neighbour"])),
                                     P(_i(g["faceCells"])))

    @staticmethod
    def threads():
        return lib.orc_omp_threads()

    def interpolate(self, T, Tb):
        g = self.g
        out = np.empty(self.F + self.B)
        lib.orc_omp_interp(self.F, self.B, P(g["owner"]), P(g["neighbour"]), P(g["weights"]),
                           P(_f(T)), P(_f(Tb)), P(out))
        return out

    def grad(self, T, Tb):
        g = self.g
        out = np.empty((self.N, 3))
        lib.orc_omp_grad(self.h, P(g["owner"]), P(g["neighbour"]), P(g["weights"]), P(g["Sf"]),
                         P(g["Sf_b"]), P(g["V"]), P(_f(T)), P(_f(Tb)), P(out))
        return out

    def div(self, phi, T, Tb, scheme=0):
        g = self.g
        out = np.empty(self.N)
        lib.orc_omp_div(self.h, scheme, P(g["owner"]), P(g["neighbour"]), P(g["weights"]),
                        P(g["V"]), P(_f(phi)), P(_f(T)), P(_f(Tb)), P(out))
        return out

    def laplacian(self, gamma, T, Tb, kinds):
        g = self.g
        k, poff = _kinds(g, kinds)
        out = np.empty(self.N)
        lib.orc_omp_laplacian(self.h, len(poff) - 1, P(poff), P(k), P(g["owner"]),
                              P(g["neighbour"]), P(g["magSf"]), P(g["nonOrthDeltaCoeffs"]),
                              P(g["magSf_b"]), P(g["deltaCoeffs_b"]), P(g["V"]), P(_f(gamma)),
                              P(_f(T)), P(_f(Tb)), P(out))
        return out

    def sngrad(self, T, Tb, kinds):
        g = self.g
        k, poff = _kinds(g, kinds)
        out = np.empty(self.F + self.B)
        lib.orc_omp_sngrad(self.F, self.B, len(poff) - 1, P(poff), P(k), P(g["owner"]),
                           P(g["neighbour"]), P(g["faceCells"]), P(g["nonOrthDeltaCoeffs"]),
                           P(g["deltaCoeffs_b"]), P(_f(T)), P(_f(Tb)), P(out))
        return out

    def close(self):
        if self.h:
            lib.orc_omp_free(self.h)
            self.h = None

    def __del__(self):
        self.close()
