"""Seeded synthetic meshes for the mesh-to-SDF workloads.

Every generator returns ``(vertices float32[N,3], indices uint32[T,3])`` of a
watertight, outward-oriented triangle mesh in millimetres (the reference's
``TriangleMesh`` convention, native/src/engine/io/mesh.hpp:14-47).

Workloads
---------
* ``perturbed_icosphere`` -- icosphere subdivided 5x (20,480 tris), R=20 mm,
  radial perturbation 2 mm*sin(3*theta)*sin(4*phi) plus N(0, 0.1 mm) vertex
  noise; the small plain-method case (128^3 grid at 0.5 mm).
* ``vessel`` -- thick-walled cup (outer R 40, wall 3, height 90, floor 2 mm)
  plus a separate closed torus shell (R 22, r 5) that passes through the wall,
  ~200k tris, N(0, 0.02 mm) vertex jitter.  The torus overlaps the wall, so the
  mesh has winding-number-2 junctions.
* ``with_slivers`` -- splits a seeded 1% of triangles at an interior point at
  barycentric distance 1e-4 from an edge (aspect ratio ~1e4, still watertight).
* ``random_vessel`` -- perturbed cups/tori of a requested triangle budget for
  the learned proposer's training set and the batched bake.
* reference fixtures ``open_box``, ``two_cubes`` (reference
  native/src/verification/harness_vessel_tests.cpp:59-99) and ``uv_sphere``.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "icosphere",
    "perturbed_icosphere",
    "cup",
    "torus",
    "vessel",
    "with_slivers",
    "random_vessel",
    "open_box",
    "two_cubes",
    "bake_batch",
    "uv_sphere",
    "moebius",
    "merge",
    "signed_volume",
    "orient_outward",
    "frame_grid",
    "jitter",
]


def _as_mesh(v, f):
    return np.ascontiguousarray(v, dtype=np.float32), np.ascontiguousarray(f, dtype=np.uint32)


def merge(*meshes):
    vs, fs, off = [], [], 0
    for v, f in meshes:
        vs.append(np.asarray(v, dtype=np.float64))
        fs.append(np.asarray(f, dtype=np.int64) + off)
        off += len(v)
    return _as_mesh(np.concatenate(vs), np.concatenate(fs))


def signed_volume(v, f) -> float:
    v = np.asarray(v, dtype=np.float64)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def orient_outward(v, f):
    """Flip every face if the enclosed signed volume is negative."""
    f = np.asarray(f)
    if signed_volume(v, f) < 0:
        f = f[:, [0, 2, 1]]
    return _as_mesh(v, f)


def jitter(v, sigma, seed):
    rng = np.random.default_rng(seed)
    return (np.asarray(v, dtype=np.float64) + rng.normal(0.0, sigma, size=np.shape(v))).astype(np.float32)


# --------------------------------------------------------------------------
# icosphere
# --------------------------------------------------------------------------
def icosphere(subdiv: int):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
         (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
         (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
         (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, dtype=np.int64)
    for _ in range(subdiv):
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        e_sorted = np.sort(e, axis=1)
        uniq, inv = np.unique(e_sorted, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(verts)
        verts = np.concatenate([verts, mid])
        nf = len(faces)
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        faces = np.concatenate([
            np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1),
            np.stack([c, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return orient_outward(verts, faces)


def perturbed_icosphere(seed: int = 1, subdiv: int = 5, radius: float = 20.0,
                        amp: float = 2.0, noise: float = 0.1):
    v, f = icosphere(subdiv)
    v = v.astype(np.float64)
    theta = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
    phi = np.arctan2(v[:, 1], v[:, 0])
    r = radius + amp * np.sin(3.0 * theta) * np.sin(4.0 * phi)
    v = v * r[:, None]
    v = jitter(v, noise, seed)
    return orient_outward(v, f)


# --------------------------------------------------------------------------
# surfaces of revolution / torus
# --------------------------------------------------------------------------
def _revolve(profile, n_az, z_axis_center=(0.0, 0.0)):
    """Revolve an open (r, z) polyline about the z axis.  Points with r == 0
    become single pole vertices.  Returns a closed mesh if both ends of the
    profile lie on the axis."""
    cx, cy = z_axis_center
    ang = np.linspace(0.0, 2.0 * math.pi, n_az, endpoint=False)
    ca, sa = np.cos(ang), np.sin(ang)
    verts, ring_ids = [], []
    for r, z in profile:
        if r == 0.0:
            ring_ids.append([len(verts)] * n_az)
            verts.append((cx, cy, z))
        else:
            ids = list(range(len(verts), len(verts) + n_az))
            ring_ids.append(ids)
            verts.extend(zip(cx + r * ca, cy + r * sa, np.full(n_az, z)))
    faces = []
    for k in range(len(profile) - 1):
        A, B = ring_ids[k], ring_ids[k + 1]
        for j in range(n_az):
            j1 = (j + 1) % n_az
            a0, a1, b0, b1 = A[j], A[j1], B[j], B[j1]
            if a0 != a1:
                faces.append((a0, a1, b1) if b0 == b1 else (a0, a1, b0))
            if b0 != b1:
                faces.append((a1, b1, b0))
    return np.array(verts, dtype=np.float64), np.array(faces, dtype=np.int64)


def _subdivide_polyline(points, step):
    out = [points[0]]
    for p, q in zip(points[:-1], points[1:]):
        seg = math.hypot(q[0] - p[0], q[1] - p[1])
        n = max(1, int(math.ceil(seg / step)))
        for i in range(1, n + 1):
            t = i / n
            out.append((p[0] + (q[0] - p[0]) * t, p[1] + (q[1] - p[1]) * t))
    return out


def cup(outer_r=40.0, wall=3.0, height=90.0, floor=2.0, step=0.76, n_az=None):
    inner = outer_r - wall
    prof = [(0.0, 0.0), (outer_r, 0.0), (outer_r, height), (inner, height),
            (inner, floor), (0.0, floor)]
    prof = _subdivide_polyline(prof, step)
    if n_az is None:
        n_az = max(16, int(round(2.0 * math.pi * outer_r / step)))
    v, f = _revolve(prof, n_az)
    return orient_outward(v, f)


def torus(R=22.0, r=5.0, center=(20.0, 0.0, 45.0), axis="y", nu=None, nv=None, step=0.76):
    if nu is None:
        nu = max(12, int(round(2.0 * math.pi * (R + r) / step)))
    if nv is None:
        nv = max(8, int(round(2.0 * math.pi * r / step)))
    u = np.linspace(0.0, 2.0 * math.pi, nu, endpoint=False)
    w = np.linspace(0.0, 2.0 * math.pi, nv, endpoint=False)
    U, W = np.meshgrid(u, w, indexing="ij")
    rad = R + r * np.cos(W)
    a = rad * np.cos(U)
    b = rad * np.sin(U)
    c = r * np.sin(W)
    if axis == "z":
        pts = np.stack([a, b, c], -1)
    elif axis == "y":
        pts = np.stack([a, c, b], -1)
    else:
        pts = np.stack([c, a, b], -1)
    v = pts.reshape(-1, 3) + np.asarray(center, dtype=np.float64)
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    i1, j1 = (ii + 1) % nu, (jj + 1) % nv
    q00 = ii * nv + jj
    q10 = i1 * nv + jj
    q11 = i1 * nv + j1
    q01 = ii * nv + j1
    f = np.concatenate([np.stack([q00, q10, q11], -1).reshape(-1, 3),
                        np.stack([q00, q11, q01], -1).reshape(-1, 3)])
    return orient_outward(v, f)


def vessel(seed: int = 2, step: float = 0.84, jitter_mm: float = 0.02,
           torus_center=(19.0, 0.0, 45.0)):
    """Cup + overlapping torus handle (~200k tris), seeded vertex jitter."""
    c = cup(step=step)
    t = torus(center=torus_center, step=step)
    v, f = merge(c, t)
    return _as_mesh(jitter(v, jitter_mm, seed), f)


def random_vessel(seed: int, target_tris: int, tri_bounds=None):
    """Perturbed cup and/or torus of about ``target_tris`` triangles.  With
    ``tri_bounds = (lo, hi)`` the edge step is corrected (up to four times)
    until the count lies in [lo, hi]."""
    rng = np.random.default_rng(seed)
    kind = rng.integers(0, 3)  # 0 cup+torus, 1 cup, 2 torus pair
    outer = rng.uniform(15.0, 40.0)
    wall = rng.uniform(1.5, 4.0)
    height = rng.uniform(20.0, 80.0)
    floor = rng.uniform(1.0, 3.0)
    Rt = rng.uniform(0.3, 0.6) * outer
    rt = rng.uniform(0.15, 0.3) * Rt
    # Choose the edge step so the triangle count lands near the target.
    area = 0.0
    if kind in (0, 1):
        area += 2 * math.pi * outer * height * 2 + 2 * math.pi * outer * outer
    if kind in (0, 2):
        area += 4 * math.pi ** 2 * Rt * rt * (1 if kind == 0 else 2)
    step = math.sqrt(2.0 * area / max(target_tris, 100))
    handle_x = outer * rng.uniform(0.4, 1.0) if kind == 0 else 0.0

    def build(step):
        parts = []
        if kind in (0, 1):
            parts.append(cup(outer, wall, height, floor, step=step))
        if kind == 0:
            parts.append(torus(Rt, rt, center=(handle_x, 0.0, height * 0.5), axis="y", step=step))
        if kind == 2:
            parts.append(torus(Rt, rt, center=(0.0, 0.0, 0.0), axis="z", step=step))
            parts.append(torus(Rt, rt, center=(Rt, 0.0, 0.0), axis="y", step=step))
        return merge(*parts)

    v, f = build(step)
    if tri_bounds is not None:
        lo, hi = tri_bounds
        want = min(max(target_tris, lo), hi)
        for _ in range(4):
            if lo <= len(f) <= hi:
                break
            # triangle count ~ 1 / step^2
            step *= math.sqrt(len(f) / (0.97 * want if len(f) > hi else 1.03 * want))
            v, f = build(step)
        if not lo <= len(f) <= hi:
            raise RuntimeError(f"random_vessel({seed}): {len(f)} triangles outside {tri_bounds}")
    v = v.astype(np.float64)
    # low-frequency radial wobble + jitter
    k = rng.uniform(0.5, 2.0)
    ph = rng.uniform(0, 2 * math.pi, size=3)
    wob = 1.0 + 0.02 * np.sin(k * v[:, 2] / 10.0 + ph[0]) * np.cos(np.arctan2(v[:, 1], v[:, 0]) * 3 + ph[1])
    v[:, 0] *= wob
    v[:, 1] *= wob
    v = jitter(v, 0.02, seed + 7919)
    return _as_mesh(v, f)


def bake_batch(n: int = 64, seed: int = 5000, tri_range=(2_000, 500_000), side_range=(64, 384),
               h_range=(0.25, 0.5)):
    """The batched-bake workload: ``n`` seeded random vessels with log-uniform
    triangle counts, each scaled so that, framed by the reference rule at a
    voxel size drawn from ``h_range``, its longest lattice side is drawn
    uniformly from ``side_range``.  Returns [(vertices, faces, h)]."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed + i)
        tri = int(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]))))
        side = int(rng.integers(side_range[0], side_range[1] + 1))
        h = float(np.float32(rng.uniform(*h_range)))
        v, f = random_vessel(seed + i, tri, tri_bounds=tri_range)
        ext = float(np.max(v.max(0) - v.min(0)))
        scale = (side - 22) * h / ext  # frame_grid pads 10 h on each side
        c = (v.max(0) + v.min(0)) / 2
        v = ((v - c) * scale).astype(np.float32)
        out.append((v, f, h))
    return out


def with_slivers(v, f, frac: float = 0.01, eps: float = 1e-4, seed: int = 3):
    """Split ``frac`` of the triangles 1-to-3 at a point with barycentric
    weight ``eps`` on the vertex opposite one edge: the triangle on that edge
    is a sliver of aspect ~1/eps; the mesh stays watertight."""
    rng = np.random.default_rng(seed)
    f = np.asarray(f, dtype=np.int64)
    v = np.asarray(v, dtype=np.float64)
    n = int(round(frac * len(f)))
    pick = rng.choice(len(f), size=n, replace=False)
    keep = np.ones(len(f), dtype=bool)
    keep[pick] = False
    a, b, c = f[pick, 0], f[pick, 1], f[pick, 2]
    q = (0.5 - eps / 2) * v[a] + (0.5 - eps / 2) * v[b] + eps * v[c]
    qi = len(v) + np.arange(n)
    new = np.concatenate([np.stack([a, b, qi], 1), np.stack([b, c, qi], 1), np.stack([c, a, qi], 1)])
    return _as_mesh(np.concatenate([v, q]), np.concatenate([f[keep], new]))


# --------------------------------------------------------------------------
# reference fixtures (harness_vessel_tests.cpp:59-99)
# --------------------------------------------------------------------------
def _box_verts(x0, x1):
    return [(x0, -10, -10), (x1, -10, -10), (x1, 10, -10), (x0, 10, -10),
            (x0, -10, 10), (x1, -10, 10), (x1, 10, 10), (x0, 10, 10)]


def open_box():
    v = _box_verts(-10, 10)
    f = [(0, 2, 1), (0, 3, 2), (0, 1, 5), (0, 5, 4), (1, 2, 6), (1, 6, 5),
         (2, 3, 7), (2, 7, 6), (3, 0, 4), (3, 4, 7)]
    return _as_mesh(v, f)


def two_cubes():
    v = _box_verts(-10, 10) + _box_verts(-2, 18)
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
         (1, 2, 6), (1, 6, 5), (2, 3, 7), (2, 7, 6), (3, 0, 4), (3, 4, 7),
         (8, 10, 9), (8, 11, 10), (12, 13, 14), (12, 14, 15), (8, 9, 13), (8, 13, 12),
         (9, 10, 14), (9, 14, 13), (10, 11, 15), (10, 15, 14), (11, 8, 12), (11, 12, 15)]
    return _as_mesh(v, f)


def moebius(R=20.0, half_width=4.0, n=96):
    """A Moebius band: n quads around a circle of radius R whose cross
    direction turns by half a turn; the last quad joins the first with top
    and bottom exchanged, so no consistent winding exists."""
    v = []
    for i in range(n):
        th = 2 * math.pi * i / n
        r = np.array([math.cos(th), math.sin(th), 0.0])
        d = math.cos(th / 2) * r + math.sin(th / 2) * np.array([0.0, 0.0, 1.0])
        c = R * r
        v += [c - half_width * d, c + half_width * d]  # bottom 2i, top 2i + 1
    f = []
    for i in range(n):
        b0, t0 = 2 * i, 2 * i + 1
        if i + 1 < n:
            b1, t1 = 2 * i + 2, 2 * i + 3
        else:  # half a turn later the cross direction is reversed
            b1, t1 = 1, 0
        f += [(b0, b1, t1), (b0, t1, t0)]
    return _as_mesh(np.array(v), f)


def uv_sphere(radius=15.0, stacks=48, slices=48):
    prof = [(radius * math.sin(math.pi * i / stacks), -radius * math.cos(math.pi * i / stacks))
            for i in range(stacks + 1)]
    prof[0] = (0.0, prof[0][1])
    prof[-1] = (0.0, prof[-1][1])
    v, f = _revolve(prof, slices)
    return orient_outward(v, f)


# --------------------------------------------------------------------------
# grid framing (reference build_from_mesh, cuda_rasterizer.cu:1358-1372)
# --------------------------------------------------------------------------
def frame_grid(v, h: float):
    """bbox padded by 10 h, min = floor(lo / h), max = ceil(hi / h), in FP32."""
    v = np.asarray(v, dtype=np.float32)
    h32 = np.float32(h)
    lo = v.min(axis=0) - np.float32(h32 * np.float32(10.0))
    hi = v.max(axis=0) + np.float32(h32 * np.float32(10.0))
    mn = np.floor(lo / h32).astype(np.int64)
    mx = np.ceil(hi / h32).astype(np.int64)
    return tuple(int(x) for x in mn), tuple(int(x) for x in mx)
