// C++ drop-in for the mesh-to-SDF entry points of rhizome::CudaRasterizer
// (native/src/engine/sdf/cuda_rasterizer.hpp:15-35, :73-86, :255-261, :280-307,
// :661-665) and of MeshWindingPass (engine/io/parsers.hpp:9-25), implemented
// over the C-ABI of rhizome_sdf.h (link with -lrhizome_sdf).  Header-only:
// nothing of the C++ standard library crosses the shared-library boundary.
//
// Semantics kept from the reference (cuda_rasterizer.cu:1163-1265, :1340-1435):
//  * compute_mesh_sdf returns false on failure (message on stderr); on
//    success out_distances is resized to the lattice size (new elements
//    1e6, as the reference's resize) and every element is written: signed
//    distance, negative inside (|w| >= 0.5), x fastest, lattice point
//    (min + i) * h.  Lattices of >= 32Mi voxels take the fast path, which
//    falls back to the exact legacy kernel when it declines; smaller ones go
//    to the legacy kernel directly.
//  * out_time_ms is host wall time of the call, transfers included.
//  * build_from_mesh frames the mesh (bbox padded by 10 h, floor/ceil),
//    builds the triangles (degenerates dropped) and returns nullptr on
//    failure; every grid it returns carries a new version number.
//  * Calls are synchronous.  Concurrent calls from several threads are
//    serialised inside the library (one cached device workspace per device).
//
// glm is not a dependency: ivec3 / vec3 are plain structs with glm's layout
// (three 32-bit members x, y, z), so a caller's glm values convert member-wise.
#pragma once

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "rhizome_sdf.h"

namespace rhizome {

struct ivec3 {
    int32_t x = 0, y = 0, z = 0;
};
struct vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
};

// cuda_types.hpp:18-21: {v0, v1, v2, normal}, 48 bytes, the layout of rhz_triangle.
struct CudaTriangle {
    vec3 v0, v1, v2;
    vec3 normal;
};
static_assert(sizeof(CudaTriangle) == sizeof(rhz_triangle), "CudaTriangle must match rhz_triangle");

// engine/io/mesh.hpp:14-47: flat xyz vertices and uint32 triangle indices (mm).
struct TriangleMesh {
    std::vector<float> vertices;
    std::vector<uint32_t> indices;
    size_t vertex_count() const noexcept { return vertices.size() / 3; }
    size_t triangle_count() const noexcept { return indices.size() / 3; }
};

// cuda_types.hpp:155-175 (the fields build_from_mesh fills).
class CudaVolumeGrid {
public:
    std::vector<float> data;
    ivec3 min_coord{};
    ivec3 max_coord{};
    ivec3 dims{};
    float voxel_size = 0.5f;
    uint64_t version = 1;
    bool is_valid() const { return !data.empty() && dims.x > 0 && dims.y > 0 && dims.z > 0; }
};

class CudaRasterizer {
public:
    static bool is_cuda_available() { return rhz_device_count() > 0; }

    static bool query_gpu_memory(size_t& out_free_bytes, size_t& out_total_bytes) {
        uint64_t f = 0, t = 0;
        if (rhz_device_memory(&f, &t) != RHZ_OK) return false;
        out_free_bytes = (size_t)f;
        out_total_bytes = (size_t)t;
        return true;
    }

    static bool compute_mesh_sdf(const std::vector<CudaTriangle>& triangles, float voxel_size, ivec3 min_coord,
                                 ivec3 max_coord, std::vector<float>& out_distances, float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        const int64_t dx = (int64_t)max_coord.x - min_coord.x + 1, dy = (int64_t)max_coord.y - min_coord.y + 1,
                      dz = (int64_t)max_coord.z - min_coord.z + 1;
        if (dx <= 0 || dy <= 0 || dz <= 0) {
            std::fprintf(stderr, "[CudaRasterizer] compute_mesh_sdf: empty lattice\n");
            return false;
        }
        out_distances.resize((size_t)(dx * dy * dz), 1e6f);
        rhz_sdf_params p;
        rhz_sdf_params_default(&p);
        p.method = RHZ_METHOD_AUTO;
        rhz_sdf_stats st;
        const int32_t mn[3] = {min_coord.x, min_coord.y, min_coord.z};
        const int32_t mx[3] = {max_coord.x, max_coord.y, max_coord.z};
        const int rc = rhz_sdf_compute(reinterpret_cast<const rhz_triangle*>(triangles.data()),
                                       (int64_t)triangles.size(), voxel_size, mn, mx, &p, out_distances.data(), &st);
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc != RHZ_OK) {
            std::fprintf(stderr, "[CudaRasterizer] compute_mesh_sdf failed (%d): %s\n", rc, rhz_last_error());
            return false;
        }
        if (st.method_used == RHZ_METHOD_FAST) {
            // the reference logs its fast-path summary on channel "CUDA"
            // (cuda_rasterizer.cu:1145-1157); out_time_ms is the whole call's
            st.ms_host_total = out_time_ms;
            char msg[512];
            if (rhz_sdf_log_line(&st, p.band_voxels, msg, sizeof(msg)) == RHZ_OK) log_sink()("CUDA", msg);
        }
        return true;
    }

    // Where compute_mesh_sdf sends its fast-path summary line (the reference
    // calls its RhizomeLog(channel, msg)); default: "[channel] msg" on stderr.
    using LogSink = void (*)(const char* channel, const char* msg);
    static void set_log_sink(LogSink sink) { log_sink() = sink ? sink : &stderr_sink; }

    static std::shared_ptr<CudaVolumeGrid> build_from_mesh(const TriangleMesh& mesh, float voxel_size,
                                                           float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        out_time_ms = 0.f;
        const int64_t nv = (int64_t)mesh.vertex_count(), nt = (int64_t)mesh.triangle_count();
        if (nv == 0 || nt == 0 || !(voxel_size > 0.f)) return nullptr;
        int32_t mn[3], mx[3];
        if (rhz_frame_grid(mesh.vertices.data(), nv, voxel_size, mn, mx) != RHZ_OK) return nullptr;
        std::vector<CudaTriangle> tris((size_t)nt);
        int64_t kept = 0;
        if (rhz_build_triangles(mesh.vertices.data(), nv, mesh.indices.data(), nt,
                                reinterpret_cast<rhz_triangle*>(tris.data()), &kept) != RHZ_OK || kept == 0)
            return nullptr;
        tris.resize((size_t)kept);
        auto grid = std::make_shared<CudaVolumeGrid>();
        grid->min_coord = {mn[0], mn[1], mn[2]};
        grid->max_coord = {mx[0], mx[1], mx[2]};
        grid->dims = {mx[0] - mn[0] + 1, mx[1] - mn[1] + 1, mx[2] - mn[2] + 1};
        grid->voxel_size = voxel_size;
        float ms = 0.f;
        if (!compute_mesh_sdf(tris, voxel_size, grid->min_coord, grid->max_coord, grid->data, ms)) return nullptr;
        grid->version = next_version();
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return grid;
    }

    static bool compute_winding_numbers_gpu(const TriangleMesh& mesh, const std::vector<vec3>& query_points,
                                            std::vector<float>& out_winding_numbers, float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        out_winding_numbers.assign(query_points.size(), 0.f);
        const int rc = rhz_winding_numbers(mesh.vertices.data(), (int64_t)mesh.vertex_count(), mesh.indices.data(),
                                           (int64_t)mesh.triangle_count(),
                                           reinterpret_cast<const float*>(query_points.data()),
                                           (int64_t)query_points.size(), out_winding_numbers.data());
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc != RHZ_OK) {
            std::fprintf(stderr, "[CudaRasterizer] compute_winding_numbers_gpu failed (%d): %s\n", rc,
                         rhz_last_error());
            return false;
        }
        return true;
    }

    // cuda_mesh_ops.cu:707-771: out_distances is resized (new elements 0); an
    // empty query list or mesh returns true with time 0 and nothing written.
    static bool query_distance_batch_gpu(const TriangleMesh& mesh, const std::vector<vec3>& query_points,
                                         std::vector<float>& out_distances, float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        out_distances.resize(query_points.size(), 0.0f);
        if (query_points.empty() || mesh.indices.empty()) {
            out_time_ms = 0.0f;
            return true;
        }
        const int rc = rhz_query_distances(mesh.vertices.data(), (int64_t)mesh.vertex_count(), mesh.indices.data(),
                                           (int64_t)mesh.triangle_count(),
                                           reinterpret_cast<const float*>(query_points.data()),
                                           (int64_t)query_points.size(), out_distances.data());
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc != RHZ_OK) {
            std::fprintf(stderr, "[CudaRasterizer] query_distance_batch_gpu failed (%d): %s\n", rc,
                         rhz_last_error());
            return false;
        }
        return true;
    }

    // cuda_marching_cubes.cu:829-962: out_mesh is cleared; a lattice with an
    // axis below 2 cells (or an empty grid) gives an empty mesh and true.
    // Unlike the reference, a grid smaller than dims or a device failure
    // returns false (the reference returns true with an empty mesh).
    static bool extract_iso_surface(const std::vector<float>& distance_grid, ivec3 grid_dims, float voxel_size,
                                    vec3 origin, float iso_value, TriangleMesh& out_mesh, float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        out_mesh.vertices.clear();
        out_mesh.indices.clear();
        out_time_ms = 0.0f;
        if (grid_dims.x < 2 || grid_dims.y < 2 || grid_dims.z < 2 || distance_grid.empty()) return true;
        if (distance_grid.size() < (size_t)grid_dims.x * (size_t)grid_dims.y * (size_t)grid_dims.z) return false;
        const int32_t dims[3] = {grid_dims.x, grid_dims.y, grid_dims.z};
        const float org[3] = {origin.x, origin.y, origin.z};
        int64_t nv = 0, nt = 0;
        int rc = rhz_iso_surface(distance_grid.data(), dims, voxel_size, org, iso_value, nullptr, 0, nullptr, 0, &nv,
                                 &nt);
        if (rc == RHZ_OK && nv > 0 && nt > 0) {
            out_mesh.vertices.resize((size_t)nv * 3);
            out_mesh.indices.resize((size_t)nt * 3);
            rc = rhz_iso_surface(distance_grid.data(), dims, voxel_size, org, iso_value, out_mesh.vertices.data(), nv,
                                 out_mesh.indices.data(), nt, &nv, &nt);
        }
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc != RHZ_OK) {
            out_mesh.vertices.clear();
            out_mesh.indices.clear();
            std::fprintf(stderr, "[CudaRasterizer] extract_iso_surface failed (%d): %s\n", rc, rhz_last_error());
            return false;
        }
        return true;
    }

    // cuda_marching_cubes.cu:952-962: origin = min_coord * voxel_size.
    static TriangleMesh extract_iso_surface(const CudaVolumeGrid& volume, float iso_value = 0.0f) {
        TriangleMesh mesh;
        float ms = 0.0f;
        const vec3 origin{(float)volume.min_coord.x * volume.voxel_size, (float)volume.min_coord.y * volume.voxel_size,
                          (float)volume.min_coord.z * volume.voxel_size};
        extract_iso_surface(volume.data, volume.dims, volume.voxel_size, origin, iso_value, mesh, ms);
        return mesh;
    }

    // ui/viewport/cuda_viewport_gl.cu:702-740: voxels with value <= 0, times h^3.
    static bool compute_sdf_interior_volume_gpu(const CudaVolumeGrid& grid, float& out_volume_mm3,
                                                float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        out_volume_mm3 = 0.0f;
        out_time_ms = 0.0f;
        if (!grid.is_valid()) return false;
        int64_t count = 0;
        if (rhz_interior_count(grid.data.data(), (int64_t)grid.data.size(), &count) != RHZ_OK) {
            std::fprintf(stderr, "[CudaRasterizer] compute_sdf_interior_volume_gpu failed: %s\n", rhz_last_error());
            return false;
        }
        const float vox = grid.voxel_size * grid.voxel_size * grid.voxel_size;
        out_volume_mm3 = (float)count * vox;
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return true;
    }

    static bool fast_sweeping_eikonal_gpu(CudaVolumeGrid& grid, float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        if (!grid.is_valid()) return false;
        const int32_t dims[3] = {grid.dims.x, grid.dims.y, grid.dims.z};
        const int rc = rhz_eikonal_redblack(grid.data.data(), dims, grid.voxel_size, 8);
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc != RHZ_OK) {
            std::fprintf(stderr, "[CudaRasterizer] fast_sweeping_eikonal_gpu failed (%d): %s\n", rc,
                         rhz_last_error());
            return false;
        }
        grid.version = next_version();
        return true;
    }

    // cuda_rasterizer.hpp:294-301 / cuda_mesh_topology.cu:707-842: winding
    // unification, non-orientable detection and outward orientation, in place.
    static bool orient_mesh_gpu(TriangleMesh& inout_mesh, bool& out_is_orientable, bool& out_had_inconsistent_winding,
                                bool& out_was_inverted, float& out_signed_volume, float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        out_is_orientable = true;
        out_had_inconsistent_winding = false;
        out_was_inverted = false;
        out_signed_volume = 0.f;
        int32_t o = 1, h = 0, w = 0;
        double vol = 0.0;
        const int rc = rhz_orient_mesh(inout_mesh.vertices.data(), (int64_t)inout_mesh.vertex_count(),
                                       inout_mesh.indices.data(), (int64_t)inout_mesh.triangle_count(), &o, &h, &w,
                                       &vol);
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc != RHZ_OK) {
            std::fprintf(stderr, "[CudaRasterizer] orient_mesh_gpu failed (%d): %s\n", rc, rhz_last_error());
            return false;
        }
        out_is_orientable = o != 0;
        out_had_inconsistent_winding = h != 0;
        out_was_inverted = w != 0;
        out_signed_volume = (float)vol;
        return true;
    }

    // cuda_rasterizer.hpp:303-307 / cuda_mesh_topology.cu:653-705.
    static bool compute_signed_volume_gpu(const TriangleMesh& mesh, float& out_signed_volume_mm3, float& out_time_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        double vol = 0.0;
        const int rc = rhz_signed_volume(mesh.vertices.data(), (int64_t)mesh.vertex_count(), mesh.indices.data(),
                                         (int64_t)mesh.triangle_count(), &vol);
        out_time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        out_signed_volume_mm3 = (float)vol;
        if (rc != RHZ_OK) {
            std::fprintf(stderr, "[CudaRasterizer] compute_signed_volume_gpu failed (%d): %s\n", rc, rhz_last_error());
            return false;
        }
        return true;
    }

private:
    static uint64_t next_version() {
        static std::atomic<uint64_t> v{1};
        return ++v;
    }
    static void stderr_sink(const char* channel, const char* msg) { std::fprintf(stderr, "[%s] %s\n", channel, msg); }
    static LogSink& log_sink() {
        static LogSink sink = &stderr_sink;
        return sink;
    }
};

// engine/io/parsers.hpp:9-25 and obj_parser.cpp:8-34: the winding pre-pass
// run on every loaded mesh, with the reference's messages.
struct WindingPrepassResult {
    bool is_orientable = true;
    bool had_inconsistent_winding = false;
    bool was_inverted = false;
    float signed_volume = 0.0f;
    std::string message;
};

class MeshWindingPass {
public:
    static WindingPrepassResult orient_mesh(TriangleMesh& mesh) {
        WindingPrepassResult r;
        if (mesh.triangle_count() == 0) return r;
        float ms = 0.f;
        CudaRasterizer::orient_mesh_gpu(mesh, r.is_orientable, r.had_inconsistent_winding, r.was_inverted,
                                        r.signed_volume, ms);
        if (!r.is_orientable)
            r.message = "Mesh contains non-orientable topology (e.g. Mobius band).";
        else if (r.was_inverted)
            r.message = "Mesh normals were inverted; corrected to outward orientation.";
        else if (r.had_inconsistent_winding)
            r.message = "Inconsistent face winding detected and unified.";
        else
            r.message = "Mesh orientation consistent and outward-facing.";
        return r;
    }
};

}  // namespace rhizome
