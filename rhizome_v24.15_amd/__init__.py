"""MI355X (gfx950) mesh-to-SDF library: ctypes binding of ``librhizome_sdf.so``.

The product is the C-ABI library (``include/rhizome_sdf.h``) and its C++
drop-in ``rhizome::CudaRasterizer`` (``include/rhizome/cuda_rasterizer.hpp``).
This module exposes the same entry points to Python callers.  There is no CPU
fallback: every call goes to the HIP library and raises if it is missing or
no GPU is visible.

Reference interfaces (golemid/rhizome_v24.15):
  compute_sdf          CudaRasterizer::compute_mesh_sdf   native/src/engine/sdf/cuda_rasterizer.hpp:21-28
  build_from_mesh      CudaRasterizer::build_from_mesh    cuda_rasterizer.hpp:31-35
  winding_numbers      compute_winding_numbers_gpu        cuda_rasterizer.hpp:280-285
  eikonal_redblack     fast_sweeping_eikonal_gpu          cuda_rasterizer.hpp:288-291
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librhizome_sdf.so")

RHZ_OK = 0
RHZ_DECLINED = 1

METHOD_AUTO = 0
METHOD_FAST = 1
METHOD_LEGACY = 2
METHOD_PLAIN_EXACT = 3
METHOD_FAST_LEARNED = 4

FSM_REFERENCE = 0
FSM_CONVERGE = 1
FSM_OCTANT = 2

WINDING_FP32 = 0
WINDING_FP32_KAHAN = 1


class RhzError(RuntimeError):
    pass


class BakeItem(C.Structure):
    _fields_ = [
        ("d_tris", C.c_void_p),
        ("n_tris", C.c_int64),
        ("h", C.c_float),
        ("min_c", C.c_int32 * 3),
        ("max_c", C.c_int32 * 3),
        ("d_out", C.c_void_p),
    ]


class Params(C.Structure):
    _fields_ = [
        ("struct_size", C.c_int32),
        ("method", C.c_int32),
        ("fast_min_voxels", C.c_int64),
        ("band_voxels", C.c_int32),
        ("parity_max_hits", C.c_int32),
        ("verify_samples", C.c_int32),
        ("verify_gens", C.c_int32),
        ("max_sites_per_probe", C.c_int32),
        ("patch_radius", C.c_int32),
        ("max_rounds", C.c_int32),
        ("max_extra_rounds", C.c_int32),
        ("sup", C.c_int32),
        ("fsm_policy", C.c_int32),
        ("csr_max_entries", C.c_int64),
        ("jitter_x", C.c_float),
        ("jitter_y", C.c_float),
        ("clamp_intended_sign", C.c_int32),
        ("fsm_round_cap", C.c_int32),
        ("winding_precision", C.c_int32),
        ("fp64_band", C.c_float),
        ("reprobe_fraction", C.c_float),
        ("learned_max_patch_gens", C.c_int32),
        ("candidate_radius", C.c_float),
        ("scorer", C.c_int32),
        ("learned_certify", C.c_int32),
        ("far_field_beta", C.c_float),
        ("mlp_weights", C.c_void_p),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("decline_reason", C.c_int32),
        ("method_used", C.c_int32),
        ("unknown_flag", C.c_int32),
        ("dims", C.c_int32 * 3),
        ("eikonal_rounds", C.c_int32),
        ("extra_rounds", C.c_int32),
        ("sweeps", C.c_int32),
        ("n_triangles", C.c_int64),
        ("fixup_voxels", C.c_int64),
        ("far_residual", C.c_int64),
        ("safe_clamped", C.c_int64),
        ("probes", C.c_int32),
        ("sites", C.c_int32),
        ("patched_voxels", C.c_int64),
        ("patch_rounds", C.c_int32),
        ("last_probe_m", C.c_int32),
        ("csr_entries", C.c_int64),
        ("reprobed_voxels", C.c_int64),
        ("fp64_voxels", C.c_int64),
        ("ms_total", C.c_double),
        ("ms_splat", C.c_double),
        ("ms_parity", C.c_double),
        ("ms_fixup", C.c_double),
        ("ms_eikonal", C.c_double),
        ("ms_verify", C.c_double),
        ("ms_apply", C.c_double),
        ("ms_host_total", C.c_double),
        ("ms_sweeps", C.c_double),
        ("ms_transposes", C.c_double),
        ("ms_winding", C.c_double),
        ("winding_pairs", C.c_int64),
        ("candidates", C.c_int64),
        ("mismatches_found", C.c_int64),
        ("ms_proposer", C.c_double),
        ("winding_calls", C.c_int64),
        ("sweep_calls", C.c_int64),
        ("ranked_probes", C.c_int64),
        ("ranked_mismatches", C.c_int64),
        ("risky_candidates", C.c_int64),
        ("fallback_probes", C.c_int64),
        ("fallback_mismatches", C.c_int64),
        ("far_decided", C.c_int64),
        ("far_undecided", C.c_int64),
        ("n_clusters", C.c_int64),
        ("ms_far_field", C.c_double),
        ("n_parent_clusters", C.c_int64),
        ("far_second_pass", C.c_int64),
    ]

    def as_dict(self):
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = list(v) if k == "dims" else v
        return d

    def log_line(self, band_voxels: int = 8) -> str:
        """The reference's one-line summary (cuda_rasterizer.cu:1145-1157),
        formatted by rhz_sdf_log_line."""
        buf = C.create_string_buffer(512)
        check(lib().rhz_sdf_log_line(C.byref(self), int(band_voxels), buf, 512), "rhz_sdf_log_line")
        return buf.value.decode()


class Accuracy(C.Structure):
    """rhz_accuracy: error of a finished field against binary64 ground truth."""
    _fields_ = [
        ("samples", C.c_int64),
        ("band_samples", C.c_int64),
        ("far_samples", C.c_int64),
        ("unreached", C.c_int64),
        ("on_surface", C.c_int64),
        ("sign_mismatches", C.c_int64),
        ("band_max_err", C.c_double),
        ("band_rms_err", C.c_double),
        ("band_mean_err", C.c_double),
        ("far_max_err", C.c_double),
        ("far_rms_err", C.c_double),
        ("far_mean_err", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def library_path() -> str:
    return LIB_PATH


def lib():
    """Loads the HIP library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RhzError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       "(hipcc --offload-arch=gfx950)")
    # torch bundles its own libamdhip64.so.7 (same SONAME as /opt/rocm's).
    # Whichever copy is loaded first serves the whole process; loading torch
    # first keeps one HIP runtime that both torch tensors and this library use.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i64, i32, f32 = C.c_void_p, C.c_int64, C.c_int32, C.c_float
    L.rhz_sdf_params_default.argtypes = [C.POINTER(Params)]
    L.rhz_last_error.restype = C.c_char_p
    L.rhz_device_count.restype = C.c_int
    L.rhz_sdf_compute.argtypes = [vp, i64, f32, vp, vp, C.POINTER(Params), vp, C.POINTER(Stats)]
    L.rhz_sdf_compute_device.argtypes = [vp, i64, f32, vp, vp, C.POINTER(Params), vp, vp, C.POINTER(Stats)]
    L.rhz_build_triangles.argtypes = [vp, i64, vp, i64, vp, C.POINTER(C.c_int64)]
    L.rhz_mesh_bbox.argtypes = [vp, i64, vp, vp]
    L.rhz_frame_grid.argtypes = [vp, i64, f32, vp, vp]
    L.rhz_winding_numbers.argtypes = [vp, i64, vp, i64, vp, i64, vp]
    L.rhz_query_distances.argtypes = [vp, i64, vp, i64, vp, i64, vp]
    L.rhz_iso_surface.argtypes = [vp, vp, f32, vp, f32, vp, i64, vp, i64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.rhz_iso_surface_device.argtypes = [vp, vp, f32, vp, f32, vp, i64, vp, i64, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), vp]
    L.rhz_interior_count.argtypes = [vp, i64, C.POINTER(C.c_int64)]
    L.rhz_interior_count_device.argtypes = [vp, i64, C.POINTER(C.c_int64), vp]
    L.rhz_eikonal_redblack.argtypes = [vp, vp, f32, i32]
    L.rhz_stage_fill.argtypes = [vp, i64, f32, vp]
    L.rhz_stage_splat.argtypes = [vp, i64, f32, f32, vp, vp, vp, vp]
    L.rhz_stage_parity.argtypes = [vp, i64, f32, f32, f32, i32, i32, i64, vp, vp, vp, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int64), vp]
    L.rhz_stage_fixup.argtypes = [vp, i64, f32, vp, vp, i32, vp, C.POINTER(C.c_int64), vp]
    L.rhz_stage_sweep.argtypes = [vp, vp, vp, f32, i32, i32, vp, vp]
    L.rhz_stage_sweep_round.argtypes = [vp, vp, vp, f32, vp, vp]
    L.rhz_stage_sweep_octant.argtypes = [vp, vp, f32, i32, i32, i32, vp, vp]
    L.rhz_stage_count_far.argtypes = [vp, i64, C.POINTER(C.c_int64), vp]
    L.rhz_stage_clamp_far.argtypes = [vp, vp, i64, f32, i32, vp]
    L.rhz_stage_verify.argtypes = [vp, i64, f32, vp, vp, i32, i32, i32, vp, vp, C.POINTER(C.c_int32), vp]
    L.rhz_stage_mark_patch.argtypes = [vp, vp, i32, i32, vp, vp]
    L.rhz_stage_apply.argtypes = [vp, vp, i64, vp]
    L.rhz_stage_winding_voxels.argtypes = [vp, i64, f32, vp, vp, vp, i64, i32, vp, vp]
    L.rhz_sdf_bake.argtypes = [C.POINTER(BakeItem), i32, C.POINTER(Params), i32, C.POINTER(Stats),
                               C.POINTER(C.c_double), vp]
    L.rhz_stage_winding_voxels_noalloc.argtypes = [vp, i64, f32, vp, vp, vp, i64, i32, vp, vp]
    L.rhz_sdf_log_line.argtypes = [C.POINTER(Stats), i32, C.c_char_p, i64]
    L.rhz_orient_mesh.argtypes = [vp, i64, vp, i64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    L.rhz_signed_volume.argtypes = [vp, i64, vp, i64, C.POINTER(C.c_double)]
    L.rhz_sdf_accuracy.argtypes = [vp, i64, f32, vp, vp, vp, i64, i32, C.c_uint64, vp, C.POINTER(Accuracy)]
    L.rhz_tuning_set.argtypes = [C.c_char_p, i64]
    L.rhz_tuning_get.argtypes = [C.c_char_p, C.POINTER(C.c_int64)]
    L.rhz_sweep_shape_class.argtypes = [C.POINTER(C.c_int32)]
    L.rhz_selftest_sweep_sqrt.argtypes = [C.POINTER(C.c_uint64)]
    L.rhz_mlp_param_count.restype = C.c_int
    L.rhz_mlp_blob_count.restype = C.c_int
    L.rhz_mlp_builtin.argtypes = [vp]
    L.rhz_mlp_forward.argtypes = [vp, vp, i64, vp, vp]
    L.rhz_mlp_grad.argtypes = [vp, vp, vp, i64, f32, vp, vp, vp]
    L.rhz_mlp_adam.argtypes = [vp, vp, vp, vp, i32, f32, f32, f32, f32, i32, f32, f32, vp]
    L.rhz_learned_candidates.argtypes = [vp, i64, f32, vp, vp, C.POINTER(Params), vp, vp, vp, vp, i64,
                                         C.POINTER(C.c_int64), vp]
    _lib = L
    return L


def check(rc: int, what: str = "call") -> int:
    if rc < 0:
        msg = lib().rhz_last_error().decode(errors="replace")
        raise RhzError(f"{what} failed ({rc}): {msg}")
    return rc


def default_params(**kw) -> Params:
    p = Params()
    lib().rhz_sdf_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _i32x3(v):
    """Three int32 values as a ctypes array: ctypes keeps it alive for the
    duration of the call it is passed to (an address taken from a temporary
    numpy array would not be)."""
    a = np.asarray(v, dtype=np.int64).reshape(3)
    return (C.c_int32 * 3)(int(a[0]), int(a[1]), int(a[2]))


def _np3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(3))


def _ptr(a) -> int:
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def tuning_set(key: str, value: int) -> None:
    check(lib().rhz_tuning_set(key.encode(), int(value)), "rhz_tuning_set")


def tuning_get(key: str) -> int:
    v = C.c_int64(0)
    check(lib().rhz_tuning_get(key.encode(), C.byref(v)), "rhz_tuning_get")
    return v.value


def sweep_shape_class(dims) -> int:
    """Lattice-size class of the per-class sweep tuning keys ("sweep_shape_x@C",
    ...): 0 up to 2^21 voxels, 1 up to 2^24, 2 up to 2^27, 3 above."""
    return check(lib().rhz_sweep_shape_class(_i32x3(dims)), "rhz_sweep_shape_class")


def selftest_sweep_sqrt() -> int:
    """Mismatches of the sweeps' square root against IEEE sqrt over every
    binary32 argument in [2^-96, inf), checked on the device (0 = exact)."""
    m = C.c_uint64(0)
    check(lib().rhz_selftest_sweep_sqrt(C.byref(m)), "rhz_selftest_sweep_sqrt")
    return int(m.value)


def device_count() -> int:
    return lib().rhz_device_count()


def release_workspace() -> None:
    """Frees the cached device workspaces of the current device
    (rhz_release_workspace): the single-lattice calls' and the bake lanes'."""
    lib().rhz_release_workspace()


# ----------------------------------------------------------------------------
# pipeline-level API (host buffers)
# ----------------------------------------------------------------------------
def compute_sdf(tris, h, min_c, max_c, params: Params | None = None, **kw):
    """Signed field of ``tris`` (float32[T, 12]) on the lattice [min_c, max_c].

    Returns ``(rc, field, stats)``; ``field`` is float32[nz, ny, nx] on
    RHZ_OK, None when the method declined."""
    p = params if params is not None else default_params(**kw)
    t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 12)
    mn, mx = _np3(min_c), _np3(max_c)  # locals: alive across the call
    dims = mx - mn + 1
    out = np.empty(int(np.prod(dims.astype(np.int64))), dtype=np.float32)
    st = Stats()
    rc = check(lib().rhz_sdf_compute(t.ctypes.data, len(t), float(h), mn.ctypes.data, mx.ctypes.data, C.byref(p),
                                     out.ctypes.data, C.byref(st)), "rhz_sdf_compute")
    field = out.reshape(int(dims[2]), int(dims[1]), int(dims[0])) if rc == RHZ_OK else None
    return rc, field, st


def compute_sdf_device(d_tris, n_tris, h, min_c, max_c, d_out, params: Params | None = None, stream=None, **kw):
    """Device-resident variant: ``d_tris``/``d_out`` are device pointers or
    torch tensors on the GPU."""
    p = params if params is not None else default_params(**kw)
    mn, mx = _np3(min_c), _np3(max_c)  # locals: alive across the call
    st = Stats()
    s = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
    rc = check(lib().rhz_sdf_compute_device(_ptr(d_tris), int(n_tris), float(h), mn.ctypes.data, mx.ctypes.data,
                                            C.byref(p), _ptr(d_out), s, C.byref(st)), "rhz_sdf_compute_device")
    return rc, st


def bake(items, params: Params | None = None, n_lanes: int = 4, stream=None, **kw):
    """Batched bake.  ``items``: sequence of (d_tris, n_tris, h, min_c, max_c,
    d_out) with device tensors / pointers.  The bake starts after the work
    queued on ``stream`` (a torch stream or hipStream_t; None = the legacy
    default stream, torch's default).  Returns (rc, [Stats], ms per item)."""
    p = params if params is not None else default_params(**kw)
    s = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
    n = len(items)
    arr = (BakeItem * max(n, 1))()
    for i, (t, nt, h, mn, mx, out) in enumerate(items):
        arr[i].d_tris = _ptr(t)
        arr[i].n_tris = int(nt)
        arr[i].h = float(h)
        for j in range(3):
            arr[i].min_c[j] = int(mn[j])
            arr[i].max_c[j] = int(mx[j])
        arr[i].d_out = _ptr(out)
    st = (Stats * max(n, 1))()
    ms = (C.c_double * max(n, 1))()
    rc = check(lib().rhz_sdf_bake(arr, n, C.byref(p), int(n_lanes), st, ms, s), "rhz_sdf_bake")
    return rc, [st[i] for i in range(n)], [ms[i] for i in range(n)]


def accuracy(d_tris, n_tris, h, min_c, max_c, d_field, n_samples=262144, band_voxels=8, seed=1, stream=None):
    """Error of the device field ``d_field`` against binary64 ground truth at
    ``n_samples`` jittered stratified voxels (rhz_sdf_accuracy)."""
    mn, mx = _np3(min_c), _np3(max_c)
    a = Accuracy()
    s = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
    check(lib().rhz_sdf_accuracy(_ptr(d_tris), int(n_tris), float(h), mn.ctypes.data, mx.ctypes.data,
                                 _ptr(d_field), int(n_samples), int(band_voxels), int(seed), s, C.byref(a)),
          "rhz_sdf_accuracy")
    return a


def build_triangles(verts, idx):
    """Indexed mesh -> float32[T', 12] triangles (degenerates dropped, index order)."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    nt = len(i) // 3
    out = np.empty((max(nt, 1), 12), dtype=np.float32)
    k = C.c_int64(0)
    check(lib().rhz_build_triangles(v.ctypes.data, len(v) // 3, i.ctypes.data, nt, out.ctypes.data, C.byref(k)),
          "rhz_build_triangles")
    return out[: k.value].copy()


def mesh_bbox(verts):
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    check(lib().rhz_mesh_bbox(v.ctypes.data, len(v) // 3, lo.ctypes.data, hi.ctypes.data), "rhz_mesh_bbox")
    return lo, hi


def frame_grid(verts, h):
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
    mn = np.zeros(3, np.int32)
    mx = np.zeros(3, np.int32)
    check(lib().rhz_frame_grid(v.ctypes.data, len(v) // 3, float(h), mn.ctypes.data, mx.ctypes.data),
          "rhz_frame_grid")
    return tuple(int(x) for x in mn), tuple(int(x) for x in mx)


def build_from_mesh(verts, idx, h, params: Params | None = None, **kw):
    """Reference build_from_mesh: bbox framing (pad 10 h), triangle build,
    compute_sdf.  Returns ``(min_c, max_c, field, stats)`` or raises."""
    mn, mx = frame_grid(verts, h)
    tris = build_triangles(verts, idx)
    if len(tris) == 0:
        raise RhzError("mesh has no non-degenerate triangles")
    rc, field, st = compute_sdf(tris, h, mn, mx, params, **kw)
    if rc != RHZ_OK:
        raise RhzError(f"compute_sdf declined (reason {st.decline_reason})")
    return mn, mx, field, st


def winding_numbers(verts, idx, pts):
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    p = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1)
    out = np.empty(len(p) // 3, dtype=np.float32)
    check(lib().rhz_winding_numbers(v.ctypes.data, len(v) // 3, i.ctypes.data, len(i) // 3, p.ctypes.data,
                                    len(out), out.ctypes.data), "rhz_winding_numbers")
    return out


def orient_mesh(verts, idx):
    """Winding unification and outward orientation (rhz_orient_mesh).
    Returns (indices (n, 3) uint32, orientable, had_inconsistent,
    was_inverted, signed volume)."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1).copy()
    o, h, w, vol = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_double(0)
    check(lib().rhz_orient_mesh(v.ctypes.data, len(v) // 3, i.ctypes.data, len(i) // 3, C.byref(o), C.byref(h),
                                C.byref(w), C.byref(vol)), "rhz_orient_mesh")
    return i.reshape(-1, 3), bool(o.value), bool(h.value), bool(w.value), vol.value


def signed_volume(verts, idx):
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    vol = C.c_double(0)
    check(lib().rhz_signed_volume(v.ctypes.data, len(v) // 3, i.ctypes.data, len(i) // 3, C.byref(vol)),
          "rhz_signed_volume")
    return vol.value


def query_distances(verts, idx, pts):
    """Unsigned point-to-mesh distances (rhz_query_distances)."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
    i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    p = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1)
    out = np.empty(len(p) // 3, dtype=np.float32)
    check(lib().rhz_query_distances(v.ctypes.data, len(v) // 3, i.ctypes.data, len(i) // 3, p.ctypes.data,
                                    len(out), out.ctypes.data), "rhz_query_distances")
    return out


def iso_surface(grid, dims, h, origin, iso=0.0):
    """Marching cubes of an x-fastest field (rhz_iso_surface): (verts (n, 3)
    float32, tris (m, 3) uint32)."""
    g = np.ascontiguousarray(grid, dtype=np.float32).reshape(-1)
    d = (C.c_int32 * 3)(*[int(x) for x in dims])
    o = (C.c_float * 3)(*[float(x) for x in origin])
    nv, nt = C.c_int64(0), C.c_int64(0)
    check(lib().rhz_iso_surface(g.ctypes.data, d, float(h), o, float(iso), None, 0, None, 0, C.byref(nv),
                                C.byref(nt)), "rhz_iso_surface")
    verts = np.zeros((nv.value, 3), np.float32)
    tris = np.zeros((nt.value, 3), np.uint32)
    if nv.value and nt.value:
        check(lib().rhz_iso_surface(g.ctypes.data, d, float(h), o, float(iso), verts.ctypes.data, nv.value,
                                    tris.ctypes.data, nt.value, C.byref(nv), C.byref(nt)), "rhz_iso_surface")
    return verts, tris


def iso_surface_device(d_grid, dims, h, origin, iso, d_verts, verts_cap, d_tris, tris_cap, stream=None):
    """Device variant: returns (n_verts, n_tris); writes the buffers when they hold the surface."""
    d = (C.c_int32 * 3)(*[int(x) for x in dims])
    o = (C.c_float * 3)(*[float(x) for x in origin])
    nv, nt = C.c_int64(0), C.c_int64(0)
    pv = _ptr(d_verts) if d_verts is not None else None
    pt = _ptr(d_tris) if d_tris is not None else None
    check(lib().rhz_iso_surface_device(_ptr(d_grid), d, float(h), o, float(iso), pv, int(verts_cap), pt,
                                       int(tris_cap), C.byref(nv), C.byref(nt), stream),
          "rhz_iso_surface_device")
    return nv.value, nt.value


def interior_count(grid):
    g = np.ascontiguousarray(grid, dtype=np.float32).reshape(-1)
    c = C.c_int64(0)
    check(lib().rhz_interior_count(g.ctypes.data, len(g), C.byref(c)), "rhz_interior_count")
    return c.value


def interior_count_device(d_grid, n, stream=None):
    c = C.c_int64(0)
    check(lib().rhz_interior_count_device(_ptr(d_grid), int(n), C.byref(c), stream), "rhz_interior_count_device")
    return c.value


def eikonal_redblack(grid, h, iters=8):
    g = np.ascontiguousarray(grid, dtype=np.float32).copy()
    dims = np.array([g.shape[2], g.shape[1], g.shape[0]], dtype=np.int32)
    check(lib().rhz_eikonal_redblack(g.ctypes.data, dims.ctypes.data, float(h), int(iters)), "rhz_eikonal_redblack")
    return g


# ----------------------------------------------------------------------------
# stage-level API (device pointers or torch GPU tensors)
# ----------------------------------------------------------------------------
class Stages:
    """Thin wrappers of the rhz_stage_* entry points."""

    @staticmethod
    def fill(d, n, value, stream=None):
        check(lib().rhz_stage_fill(_ptr(d), int(n), float(value), stream), "rhz_stage_fill")

    @staticmethod
    def splat(d_tris, n_tris, h, band_width, min_c, dims, d, stream=None):
        check(lib().rhz_stage_splat(_ptr(d_tris), int(n_tris), float(h), float(band_width),
                                    _i32x3(min_c), _i32x3(dims), _ptr(d), stream),
              "rhz_stage_splat")

    @staticmethod
    def parity(d_tris, n_tris, h, min_c, dims, d_sign, jitter=(0.037, 0.051), max_hits=128, sup=4,
               csr_max_entries=100000000, stream=None):
        flag = C.c_int32(0)
        ent = C.c_int64(0)
        rc = check(lib().rhz_stage_parity(_ptr(d_tris), int(n_tris), float(h), float(jitter[0]), float(jitter[1]),
                                          int(max_hits), int(sup), int(csr_max_entries), _i32x3(min_c),
                                          _i32x3(dims), _ptr(d_sign), C.byref(flag), C.byref(ent),
                                          stream), "rhz_stage_parity")
        return rc, flag.value, ent.value

    @staticmethod
    def fixup(d_tris, n_tris, h, min_c, dims, d_sign, precision=WINDING_FP32_KAHAN, stream=None):
        k = C.c_int64(0)
        check(lib().rhz_stage_fixup(_ptr(d_tris), int(n_tris), float(h), _i32x3(min_c),
                                    _i32x3(dims), int(precision), _ptr(d_sign), C.byref(k), stream),
              "rhz_stage_fixup")
        return k.value

    @staticmethod
    def sweep(src, dst, dims, h, axis, direction, d_changed, stream=None):
        check(lib().rhz_stage_sweep(_ptr(src), _ptr(dst), _i32x3(dims), float(h), int(axis),
                                    int(direction), _ptr(d_changed), stream), "rhz_stage_sweep")

    @staticmethod
    def sweep_octant(d, dims, h, sx, sy, sz, d_changed, stream=None):
        check(lib().rhz_stage_sweep_octant(_ptr(d), _i32x3(dims), float(h), int(sx), int(sy), int(sz),
                                           _ptr(d_changed), stream), "rhz_stage_sweep_octant")

    @staticmethod
    def sweep_round(a, b, dims, h, d_changed, stream=None):
        check(lib().rhz_stage_sweep_round(_ptr(a), _ptr(b), _i32x3(dims), float(h), _ptr(d_changed),
                                          stream), "rhz_stage_sweep_round")

    @staticmethod
    def count_far(d, n, stream=None):
        k = C.c_int64(0)
        check(lib().rhz_stage_count_far(_ptr(d), int(n), C.byref(k), stream), "rhz_stage_count_far")
        return k.value

    @staticmethod
    def clamp_far(d, sign, n, band_width, intended=1, stream=None):
        check(lib().rhz_stage_clamp_far(_ptr(d), _ptr(sign), int(n), float(band_width), int(intended), stream),
              "rhz_stage_clamp_far")

    @staticmethod
    def verify(d_tris, n_tris, h, min_c, dims, samples, gen, d_sign, d_sites, precision=WINDING_FP32_KAHAN,
               stream=None):
        m = C.c_int32(0)
        check(lib().rhz_stage_verify(_ptr(d_tris), int(n_tris), float(h), _i32x3(min_c),
                                     _i32x3(dims), int(samples), int(gen), int(precision), _ptr(d_sign),
                                     _ptr(d_sites), C.byref(m), stream), "rhz_stage_verify")
        return m.value

    @staticmethod
    def mark_patch(d_sign, d_sites, n_sites, radius, dims, stream=None):
        check(lib().rhz_stage_mark_patch(_ptr(d_sign), _ptr(d_sites), int(n_sites), int(radius),
                                         _i32x3(dims), stream), "rhz_stage_mark_patch")

    @staticmethod
    def apply(d, sign, n, stream=None):
        check(lib().rhz_stage_apply(_ptr(d), _ptr(sign), int(n), stream), "rhz_stage_apply")

    @staticmethod
    def winding_voxels(d_tris, n_tris, h, min_c, dims, d_voxels, n_voxels, d_w, precision=WINDING_FP32_KAHAN,
                       stream=None):
        check(lib().rhz_stage_winding_voxels(_ptr(d_tris), int(n_tris), float(h), _i32x3(min_c),
                                             _i32x3(dims), _ptr(d_voxels), int(n_voxels),
                                             int(precision), _ptr(d_w), stream), "rhz_stage_winding_voxels")

    @staticmethod
    def winding_voxels_noalloc(d_tris, n_tris, h, min_c, dims, d_voxels, n_voxels, d_w,
                               precision=WINDING_FP32_KAHAN, stream=None):
        """rhz_stage_winding_voxels_noalloc: no device allocation, at most
        2^20 voxels, enqueued on ``stream`` without waiting."""
        s = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
        check(lib().rhz_stage_winding_voxels_noalloc(_ptr(d_tris), int(n_tris), float(h), _i32x3(min_c),
                                                     _i32x3(dims), _ptr(d_voxels), int(n_voxels), int(precision),
                                                     _ptr(d_w), s), "rhz_stage_winding_voxels_noalloc")


# ----------------------------------------------------------------------------
# sign-risk proposer (12 -> 64 -> 64 -> 64 -> 1 MLP) and its training kernels
# ----------------------------------------------------------------------------
class Proposer:
    """Wrappers of the rhz_mlp_* / rhz_learned_candidates entry points."""

    N_FEATURES = 12

    @staticmethod
    def param_count() -> int:
        return lib().rhz_mlp_param_count()

    @staticmethod
    def blob_count() -> int:
        return lib().rhz_mlp_blob_count()

    @staticmethod
    def builtin():
        """The trained blob compiled into the library, or None."""
        out = np.zeros(Proposer.blob_count(), dtype=np.float32)
        rc = check(lib().rhz_mlp_builtin(out.ctypes.data), "rhz_mlp_builtin")
        return out if rc == 0 else None

    @staticmethod
    def forward(d_blob, d_feats, n, d_logits, stream=None):
        check(lib().rhz_mlp_forward(_ptr(d_blob), _ptr(d_feats), int(n), _ptr(d_logits), stream),
              "rhz_mlp_forward")

    @staticmethod
    def grad(d_blob, d_feats, d_labels, n, pos_weight, d_grads, d_loss, stream=None):
        check(lib().rhz_mlp_grad(_ptr(d_blob), _ptr(d_feats), _ptr(d_labels), int(n), float(pos_weight),
                                 _ptr(d_grads), _ptr(d_loss), stream), "rhz_mlp_grad")

    @staticmethod
    def adam(d_blob, d_grads, d_m, d_v, n_params, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, grad_scale=1.0,
             weight_decay=0.0, stream=None):
        check(lib().rhz_mlp_adam(_ptr(d_blob), _ptr(d_grads), _ptr(d_m), _ptr(d_v), int(n_params), float(lr),
                                 float(beta1), float(beta2), float(eps), int(step), float(grad_scale),
                                 float(weight_decay), stream), "rhz_mlp_adam")

    @staticmethod
    def candidates(d_tris, n_tris, h, min_c, max_c, d_voxels, d_feats, cap, d_risk=None, d_scan=None,
                   params: Params | None = None, stream=None, **kw):
        p = params if params is not None else default_params(**kw)
        n = C.c_int64(0)
        opt = lambda a: _ptr(a) if a is not None else None  # noqa: E731
        rc = check(lib().rhz_learned_candidates(_ptr(d_tris), int(n_tris), float(h), _i32x3(min_c),
                                                _i32x3(max_c), C.byref(p), opt(d_voxels), opt(d_feats),
                                                opt(d_risk), opt(d_scan), int(cap), C.byref(n), stream),
                   "rhz_learned_candidates")
        if rc != RHZ_OK:
            raise RhzError(f"rhz_learned_candidates declined ({rc})")
        return n.value
