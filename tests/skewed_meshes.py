"""Meshes with planar faces whose geometry has a closed form: an affine image of
the uniform unit-cube mesh (skewed, non-orthogonal), a tapered mesh (y scaled by
1 + alpha*x, whose z-faces are trapezoids), and a polyhedral block with
non-conforming refinement (coarse cells with m*m face neighbours and hanging
points on their side faces)."""
import dataclasses

import numpy as np

from foamadapter_amd.mesh import Patch, PolyMesh

AFFINE_A = np.array([[1.0, 0.3, 0.1], [0.0, 0.8, -0.25], [0.15, 0.0, 1.2]])  # det 0.93675
AFFINE_B = np.array([0.2, -0.1, 0.05])
TAPER_ALPHA = 0.5


# a strong shear in the x-y plane: the x- and y-faces meet their cell deltas at
# about 88 degrees (cos = 1/sqrt(1 + 30^2) < 0.05), so their non-orthogonal delta
# coefficients take the 0.05|d| clamp; the z-faces stay orthogonal
SHEAR_A = np.array([[1.0, 30.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def affine_hex(n, perm_seed=0, A=AFFINE_A, b=AFFINE_B):
    """(uniform mesh, its image under x -> A x + b): same topology and numbering."""
    poly0 = PolyMesh.hex(n, perm_seed=perm_seed)
    return poly0, dataclasses.replace(poly0, points=poly0.points @ A.T + b)


def tapered_hex(n, alpha=TAPER_ALPHA, perm_seed=0):
    """(uniform mesh, its image under (x, y, z) -> (x, y (1 + alpha x), z)).  Planes
    x = const, y = const and z = const stay planes, so every face stays planar."""
    poly0 = PolyMesh.hex(n, perm_seed=perm_seed)
    p = poly0.points.copy()
    p[:, 1] = p[:, 1] * (1.0 + alpha * p[:, 0])
    return poly0, dataclasses.replace(poly0, points=p)


def refined_block(m=3):
    """Coarse cube [0,1]^3 | an m x m layer of fine cells [1,2] x [j/m,(j+1)/m] x
    [k/m,(k+1)/m] | coarse cube [2,3] x [0,1]^2.  Each coarse cell has the m*m
    fine cells as neighbours (its matrix row m*m + 1 entries) and its four side faces carry
    the fine layer's hanging points ((m + 3)-gons).  Cells: 0 the left coarse cell,
    1 + j + m*k the fine cells, m*m + 1 the right one (every one of whose
    neighbours has a lower number).  Patches: inlet (x = 0), outlet (x = 3),
    walls."""
    pts, pid = [], {}

    def P(X, Y, Z):  # integer grid point, coordinates X/m, Y/m, Z/m
        key = (X, Y, Z)
        if key not in pid:
            pid[key] = len(pts)
            pts.append((X / m, Y / m, Z / m))
        return pid[key]

    nc = m * m + 2
    right = m * m + 1
    lo = {0: (0, 0, 0), right: (2 * m, 0, 0)}
    hi = {0: (m, m, m), right: (3 * m, m, m)}
    for k in range(m):
        for j in range(m):
            c = 1 + j + m * k
            lo[c], hi[c] = (m, j, k), (2 * m, j + 1, k + 1)
    centre = {c: np.add(lo[c], hi[c]) / 2.0 for c in range(nc)}
    faces = []  # (vertex ids, owner, neighbour or patch name)

    def square(ax, v, a0, a1, b0, b1):  # axis-normal rectangle at coordinate v
        corners = [(a0, b0), (a1, b0), (a1, b1), (a0, b1)]
        out = []
        for a, b in corners:
            q = [0, 0, 0]
            q[ax], q[(ax + 1) % 3], q[(ax + 2) % 3] = v, a, b
            out.append(tuple(q))
        return out

    # coarse <-> fine sub-faces and fine <-> fine faces
    for k in range(m):
        for j in range(m):
            c = 1 + j + m * k
            faces.append((square(0, m, j, j + 1, k, k + 1), 0, c))
            faces.append((square(0, 2 * m, j, j + 1, k, k + 1), c, right))
            if j + 1 < m:
                faces.append((square(1, j + 1, k, k + 1, m, 2 * m), c, c + 1))
            if k + 1 < m:
                faces.append((square(2, k + 1, m, 2 * m, j, j + 1), c, c + m))
            if j == 0:
                faces.append((square(1, 0, k, k + 1, m, 2 * m), c, "walls"))
            if j == m - 1:
                faces.append((square(1, m, k, k + 1, m, 2 * m), c, "walls"))
            if k == 0:
                faces.append((square(2, 0, m, 2 * m, j, j + 1), c, "walls"))
            if k == m - 1:
                faces.append((square(2, m, m, 2 * m, j, j + 1), c, "walls"))
    # coarse cells: end faces and side faces with the hanging points at the
    # fine layer's edge (x = m for the left cell, x = 2m for the right one)
    for c, x0, x1, xe, end, patch in ((0, 0, m, m, 0, "inlet"), (right, 2 * m, 3 * m, 2 * m, 3 * m, "outlet")):
        faces.append((square(0, end, 0, m, 0, m), c, patch))
        xo = x0 if xe == x1 else x1  # the far edge of the side face
        for ax in (1, 2):
            for v in (0, m):
                poly = []
                for t in range(m + 1):  # along the fine-layer edge
                    q = [xe, 0, 0]
                    q[ax], q[3 - ax] = v, t
                    poly.append(tuple(q))
                for t in (m, 0):  # back along the far edge
                    q = [xo, 0, 0]
                    q[ax], q[3 - ax] = v, t
                    poly.append(tuple(q))
                faces.append((poly, c, "walls"))
    # orient: area vector from owner to neighbour (outward on the boundary)
    recs = []
    for poly, own, nb in faces:
        xyz = np.array(poly, float)
        n = 0.5 * np.cross(xyz, np.roll(xyz, -1, axis=0)).sum(0)
        d = (centre[nb] if isinstance(nb, int) else xyz.mean(0)) - centre[own]
        if n @ d < 0:
            poly = poly[::-1]
        recs.append(([P(*q) for q in poly], own, nb))
    internal = sorted((r for r in recs if isinstance(r[2], int)), key=lambda r: (r[1], r[2]))
    order = list(internal)
    patches, start = [], len(internal)
    for name, typ in (("inlet", "patch"), ("outlet", "patch"), ("walls", "wall")):
        fs = [r for r in recs if r[2] == name]
        order += fs
        patches.append(Patch(name, typ, start, len(fs)))
        start += len(fs)
    off = np.cumsum([0] + [len(r[0]) for r in order]).astype(np.int32)
    return PolyMesh(points=np.array(pts, float), face_offsets=off,
                    face_points=np.array([v for r in order for v in r[0]], np.int32),
                    owner=np.array([r[1] for r in order], np.int32),
                    neighbour=np.array([r[2] for r in internal], np.int32), patches=patches)
