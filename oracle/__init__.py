"""ctypes binding of the CPU oracle (``sdf_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg as the checker / CPU baseline.  The product
library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdf_oracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [
        ("fast_min_voxels", C.c_int64),
        ("band_voxels", C.c_int),
        ("parity_max_hits", C.c_int),
        ("verify_samples", C.c_int),
        ("verify_gens", C.c_int),
        ("max_sites_per_probe", C.c_int),
        ("patch_radius", C.c_int),
        ("max_rounds", C.c_int),
        ("max_extra_rounds", C.c_int),
        ("sup", C.c_int),
        ("csr_max_entries", C.c_int64),
        ("jitter_x", C.c_float),
        ("jitter_y", C.c_float),
        ("clamp_intended_sign", C.c_int),
        ("fsm_converge", C.c_int),
        ("fsm_round_cap", C.c_int),
        ("winding_mode", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("decline_reason", C.c_int),
        ("unknown_flag", C.c_int),
        ("fixup_voxels", C.c_int64),
        ("eikonal_rounds", C.c_int),
        ("extra_rounds", C.c_int),
        ("far_residual", C.c_int64),
        ("safe_clamped", C.c_int64),
        ("probes", C.c_int),
        ("sites", C.c_int),
        ("patched_voxels", C.c_int64),
        ("patch_rounds", C.c_int),
        ("last_probe_m", C.c_int),
        ("csr_entries", C.c_int64),
        ("ms_splat", C.c_double),
        ("ms_parity", C.c_double),
        ("ms_fixup", C.c_double),
        ("ms_eikonal", C.c_double),
        ("ms_verify", C.c_double),
        ("ms_patch", C.c_double),
        ("ms_apply", C.c_double),
        ("ms_total", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    src = os.path.join(HERE, "sdf_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f3 = C.POINTER(C.c_float)
        L.orc_params_default.argtypes = [C.POINTER(Params)]
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_num_threads.restype = C.c_int
        L.orc_point_triangle_distance_sq.argtypes = [_f32p] * 4
        L.orc_point_triangle_distance_sq.restype = C.c_float
        L.orc_point_triangle_distance_sq_f64.argtypes = [_f64p] * 4
        L.orc_point_triangle_distance_sq_f64.restype = C.c_double
        L.orc_triangle_solid_angle.argtypes = [_f32p] * 4
        L.orc_triangle_solid_angle.restype = C.c_float
        L.orc_triangle_solid_angle_f64.argtypes = [_f64p] * 4
        L.orc_triangle_solid_angle_f64.restype = C.c_double
        L.orc_godunov.argtypes = [C.c_float] * 4
        L.orc_godunov.restype = C.c_float
        L.orc_build_triangles.argtypes = [_f32p, C.c_int64, _u32p, C.c_int64, _f32p]
        L.orc_build_triangles.restype = C.c_int64
        L.orc_frame_grid.argtypes = [_f32p, C.c_int64, C.c_float, _i32p, _i32p]
        L.orc_fill.argtypes = [_f32p, C.c_int64, C.c_float]
        L.orc_splat_band.argtypes = [_f32p, C.c_int64, C.c_float, C.c_float, _i32p, _i32p, _f32p]
        L.orc_csr_entries.argtypes = [_f32p, C.c_int64, C.c_float, C.c_int, _i32p, _i32p]
        L.orc_csr_entries.restype = C.c_int64
        L.orc_parity_signs.argtypes = [_f32p, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_int,
                                       _i32p, _i32p, _u8p, C.c_void_p]
        L.orc_parity_signs.restype = C.c_int
        L.orc_winding_fixup.argtypes = [_f32p, C.c_int64, C.c_float, _i32p, _i32p, _u8p, C.c_int]
        L.orc_winding_fixup.restype = C.c_int64
        L.orc_winding_at_voxel.argtypes = [_f32p, C.c_int64, C.c_float, _i32p, _i32p, C.c_int64, C.c_int]
        L.orc_winding_at_voxel.restype = C.c_double
        L.orc_fsm_sweep.argtypes = [_f32p, _i32p, C.c_float, C.c_int, C.c_int]
        L.orc_fsm_sweep.restype = C.c_int
        L.orc_fsm_sweep_db.argtypes = [_f32p, _f32p, _i32p, C.c_float, C.c_int, C.c_int]
        L.orc_fsm_sweep_db.restype = C.c_int
        L.orc_fsm_round.argtypes = [_f32p, _i32p, C.c_float]
        L.orc_fsm_round.restype = C.c_int
        L.orc_winding_voxels.argtypes = [_f32p, C.c_int64, C.c_float, _i32p, _i32p, _i64p, C.c_int64, C.c_int,
                                         _f64p]
        L.orc_fsm_sweep_inplace.argtypes = [_f32p, _i32p, C.c_float, C.c_int, C.c_int]
        L.orc_fsm_sweep_inplace.restype = C.c_int
        L.orc_fsm_sweep_octant.argtypes = [_f32p, _i32p, C.c_float, C.c_int, C.c_int, C.c_int]
        L.orc_fsm_sweep_octant.restype = C.c_int
        L.orc_fsm_octant_round.argtypes = [_f32p, _i32p, C.c_float]
        L.orc_fsm_octant_round.restype = C.c_int
        L.orc_count_far.argtypes = [_f32p, C.c_int64]
        L.orc_count_far.restype = C.c_int64
        L.orc_clamp_far.argtypes = [_f32p, _u8p, C.c_int64, C.c_float, C.c_int]
        L.orc_verify_indices.argtypes = [C.c_int64, C.c_int, C.c_int, _i64p]
        L.orc_band_f64.argtypes = [_f32p, C.c_int64, C.c_float, C.c_double, _i32p, _i32p, _f64p]
        L.orc_mark_patch.argtypes = [_u8p, _i64p, C.c_int, C.c_int, _i32p]
        L.orc_apply_signs.argtypes = [_f32p, _u8p, C.c_int64]
        L.orc_fast_sdf.argtypes = [_f32p, C.c_int64, C.c_float, _i32p, _i32p, C.POINTER(Params),
                                   _f32p, C.c_void_p, C.POINTER(Stats)]
        L.orc_fast_sdf.restype = C.c_int
        L.orc_plain_exact_sdf.argtypes = [_f32p, C.c_int64, C.c_float, _i32p, _i32p, C.c_int,
                                          _f32p, C.c_void_p, C.POINTER(C.c_int)]
        L.orc_plain_exact_sdf.restype = C.c_int
        L.orc_legacy_sdf.argtypes = [_f32p, C.c_int64, C.c_float, _i32p, _i32p, _f32p]
        L.orc_winding_points.argtypes = [_f32p, _u32p, C.c_int64, _f32p, C.c_int64, _f32p]
        L.orc_query_distances.argtypes = [_f32p, _u32p, C.c_int64, _f32p, C.c_int64, _f32p]
        L.orc_mc_table.argtypes = [C.c_int, C.POINTER(C.c_int)]
        L.orc_mc_table.restype = C.c_int
        L.orc_mc_count.argtypes = [_f32p, _i32p, C.c_float, C.c_float, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_mc_extract.argtypes = [_f32p, _i32p, C.c_float, _f32p, C.c_float, _f32p, _u32p]
        L.orc_interior_count.argtypes = [_f32p, C.c_int64]
        L.orc_interior_count.restype = C.c_int64
        L.orc_eikonal_redblack.argtypes = [_f32p, _i32p, C.c_float, C.c_int]
        L.orc_orient_mesh.argtypes = [_f32p, C.c_int64, _u32p, C.c_int64, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_signed_volume.argtypes = [_f32p, _u32p, C.c_int64]
        L.orc_signed_volume.restype = C.c_double
        L.orc_hash_01.argtypes = [C.c_uint32]
        L.orc_hash_01.restype = C.c_float
        L.orc_gwn_test_points.argtypes = [C.c_int, C.c_uint32, _f32p]
        del f3
        _lib = L
    return _lib


# ----------------------------------------------------------------------------
# numpy-level helpers
# ----------------------------------------------------------------------------
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_params(**kw) -> Params:
    p = Params()
    lib().orc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def build_triangles(verts, idx):
    """Indexed mesh -> float32[T', 12] triangles (v0, v1, v2, unit normal)."""
    v = _f32(verts).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    nt = len(i) // 3
    out = np.zeros((max(nt, 1), 12), dtype=np.float32)
    k = lib().orc_build_triangles(v, len(v) // 3, i, nt, out.reshape(-1))
    return out[:k].copy()


def dims_of(mn, mx):
    return np.array([mx[0] - mn[0] + 1, mx[1] - mn[1] + 1, mx[2] - mn[2] + 1], dtype=np.int32)


def splat_band(tris, h, mn, dims, band_voxels=8):
    n = int(np.prod(dims))
    d = np.full(n, 1e30, dtype=np.float32)
    lib().orc_splat_band(_f32(tris).reshape(-1), len(tris), h, np.float32(band_voxels) * np.float32(h),
                         _i32(mn), _i32(dims), d)
    return d


def band_f64(tris, h, mn, dims, band_voxels=8):
    """Binary64 band distances from the binary32 lattice points (1e30 outside)."""
    out = np.empty(int(np.prod(dims)), dtype=np.float64)
    lib().orc_band_f64(_f32(tris).reshape(-1), len(tris), h, band_voxels * float(np.float32(h)), _i32(mn),
                       _i32(dims), out)
    return out


def parity_signs(tris, h, mn, dims, jx=0.037, jy=0.051, max_hits=128, with_hits=False):
    n = int(np.prod(dims))
    s = np.zeros(n, dtype=np.uint8)
    hits = np.zeros(int(dims[0]) * int(dims[1]), dtype=np.int32) if with_hits else None
    flag = lib().orc_parity_signs(_f32(tris).reshape(-1), len(tris), h, jx, jy, max_hits,
                                  _i32(mn), _i32(dims), s,
                                  hits.ctypes.data if with_hits else None)
    return (s, flag, hits) if with_hits else (s, flag)


def winding_at_voxels(tris, h, mn, dims, voxels, mode=0):
    """Winding numbers at lattice voxels (threads over the list)."""
    t = _f32(tris).reshape(-1)
    v = np.ascontiguousarray(np.asarray(voxels).reshape(-1), dtype=np.int64)
    out = np.empty(len(v), dtype=np.float64)
    if len(v):
        lib().orc_winding_voxels(t, len(tris), h, _i32(mn), _i32(dims), v, len(v), mode, out)
    return out


def fsm_round(d, dims, h):
    """One round x+, x-, y+, y-, z+, z- of the pinned schedule, in place on
    float32 ``d``; returns 1 if a voxel decreased."""
    return lib().orc_fsm_round(d, _i32(dims), h)


def octant_converge(d, dims, h, max_rounds=512):
    """Rounds of the 8 classical orderings until nothing changes; returns rounds."""
    dm = _i32(dims)
    for r in range(max_rounds):
        if not lib().orc_fsm_octant_round(d, dm, h):
            return r + 1
    return max_rounds


def fsm_converge(d, dims, h, max_rounds=512):
    """Run reference 6-sweep rounds until nothing changes; returns rounds."""
    dm = _i32(dims)
    for r in range(max_rounds):
        ch = 0
        for ax in range(3):
            ch |= lib().orc_fsm_sweep(d, dm, h, ax, +1)
            ch |= lib().orc_fsm_sweep(d, dm, h, ax, -1)
        if not ch:
            return r + 1
    return max_rounds


def fast_sdf(tris, h, mn, mx, params=None, want_sign=False):
    p = params or default_params()
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    out = np.zeros(n, dtype=np.float32)
    sg = np.zeros(n, dtype=np.uint8) if want_sign else None
    st = Stats()
    rc = lib().orc_fast_sdf(_f32(tris).reshape(-1), len(tris), h, _i32(mn), _i32(mx), C.byref(p), out,
                            sg.ctypes.data if want_sign else None, C.byref(st))
    return rc, out, sg, st


def plain_exact_sdf(tris, h, mn, mx, winding_mode=0, want_sign=False):
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    out = np.zeros(n, dtype=np.float32)
    sg = np.zeros(n, dtype=np.uint8)
    rounds = C.c_int(0)
    lib().orc_plain_exact_sdf(_f32(tris).reshape(-1), len(tris), h, _i32(mn), _i32(mx), winding_mode,
                              out, sg.ctypes.data, C.byref(rounds))
    return (out, sg, rounds.value) if want_sign else (out, rounds.value)


def legacy_sdf(tris, h, mn, mx):
    dims = dims_of(mn, mx)
    out = np.zeros(int(np.prod(dims)), dtype=np.float32)
    lib().orc_legacy_sdf(_f32(tris).reshape(-1), len(tris), h, _i32(mn), _i32(mx), out)
    return out


def winding_points(verts, idx, pts):
    v = _f32(verts).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    p = _f32(pts).reshape(-1)
    out = np.zeros(len(p) // 3, dtype=np.float32)
    lib().orc_winding_points(v, i, len(i) // 3, p, len(out), out)
    return out


def query_distances(verts, idx, pts):
    v = _f32(verts).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    p = _f32(pts).reshape(-1)
    out = np.zeros(len(p) // 3, dtype=np.float32)
    lib().orc_query_distances(v, i, len(i) // 3, p, len(out), out)
    return out


def mc_table(c):
    e = (C.c_int * 16)()
    n = lib().orc_mc_table(int(c), e)
    return [e[k] for k in range(n)]


def mc_extract(grid, dims, h, origin, iso):
    """Marching cubes of an x-fastest grid: (verts (n, 3) float32, tris (m, 3) uint32)."""
    g = _f32(grid).reshape(-1)
    d = _i32(dims)
    if min(int(x) for x in dims) < 2 or len(g) == 0:  # cuda_marching_cubes.cu:842-845
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
    nv, nt = C.c_int64(0), C.c_int64(0)
    lib().orc_mc_count(g, d, float(h), float(iso), C.byref(nv), C.byref(nt))
    verts = np.zeros((nv.value, 3), np.float32)
    tris = np.zeros((nt.value, 3), np.uint32)
    if nv.value and nt.value:
        lib().orc_mc_extract(g, d, float(h), _f32(origin), float(iso), verts.reshape(-1), tris.reshape(-1))
    return verts, tris


def interior_count(grid):
    g = _f32(grid).reshape(-1)
    return int(lib().orc_interior_count(g, len(g)))


def gwn_test_points(mode, seed):
    out = np.zeros(30000, dtype=np.float32)
    lib().orc_gwn_test_points(mode, seed, out)
    return out.reshape(10000, 3)


def orient_mesh(verts, idx):
    """(oriented indices (n, 3) uint32, orientable, had_inconsistent,
    was_inverted, signed volume)."""
    v = _f32(verts).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1).copy()
    o, h, w, vol = C.c_int(0), C.c_int(0), C.c_int(0), C.c_double(0)
    lib().orc_orient_mesh(v, len(v) // 3, i, len(i) // 3, C.byref(o), C.byref(h), C.byref(w), C.byref(vol))
    return i.reshape(-1, 3), bool(o.value), bool(h.value), bool(w.value), vol.value


def signed_volume(verts, idx):
    v = _f32(verts).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    return lib().orc_signed_volume(v, i, len(i) // 3)
