// Drives the C++ drop-in (include/rhizome/cuda_rasterizer.hpp) the way the
// reference's callers do (build_from_mesh, compute_mesh_sdf,
// compute_winding_numbers_gpu, query_distance_batch_gpu, extract_iso_surface,
// compute_sdf_interior_volume_gpu), for the parity tests.
//
//   dropin_main info
//   dropin_main sdf <mesh.bin> <h> <out.bin>
//     mesh.bin: int64 nv, int64 nt, float xyz[3 nv], uint32 idx[3 nt]
//     out.bin:  int32 min[3], int32 max[3], float field[V] (build_from_mesh)
//   dropin_main orient <mesh.bin> <out.bin>   (MeshWindingPass::orient_mesh; out: uint32 idx[3 nt])
//   dropin_main wind <mesh.bin> <pts.bin> <out.bin>
//     pts.bin: int64 np, float xyz[3 np]; out.bin: float w[np]
// Exit status: 0 ok, 1 usage or I/O error, 2 the library call failed.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rhizome/cuda_rasterizer.hpp"

using namespace rhizome;

static bool read_mesh(const char* path, TriangleMesh& m) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    int64_t nv = 0, nt = 0;
    bool ok = std::fread(&nv, 8, 1, f) == 1 && std::fread(&nt, 8, 1, f) == 1 && nv >= 0 && nt >= 0;
    if (ok) {
        m.vertices.resize((size_t)nv * 3);
        m.indices.resize((size_t)nt * 3);
        ok = std::fread(m.vertices.data(), 4, m.vertices.size(), f) == m.vertices.size() &&
             std::fread(m.indices.data(), 4, m.indices.size(), f) == m.indices.size();
    }
    std::fclose(f);
    return ok;
}

int main(int argc, char** argv) {
    if (argc >= 2 && !std::strcmp(argv[1], "info")) {
        size_t fr = 0, tot = 0;
        const bool mem = CudaRasterizer::query_gpu_memory(fr, tot);
        std::printf("available %d memory %d total %zu\n", CudaRasterizer::is_cuda_available() ? 1 : 0, mem ? 1 : 0,
                    tot);
        return 0;
    }
    if (argc == 5 && !std::strcmp(argv[1], "sdf")) {
        TriangleMesh m;
        if (!read_mesh(argv[2], m)) return 1;
        float ms = 0.f;
        auto g = CudaRasterizer::build_from_mesh(m, (float)std::atof(argv[3]), ms);
        if (!g) return 2;
        FILE* f = std::fopen(argv[4], "wb");
        if (!f) return 1;
        std::fwrite(&g->min_coord, 4, 3, f);
        std::fwrite(&g->max_coord, 4, 3, f);
        std::fwrite(g->data.data(), 4, g->data.size(), f);
        std::fclose(f);
        std::printf("dims %d %d %d time_ms %.3f version %llu\n", g->dims.x, g->dims.y, g->dims.z, ms,
                    (unsigned long long)g->version);
        return 0;
    }
    if (argc == 5 && (!std::strcmp(argv[1], "wind") || !std::strcmp(argv[1], "dist"))) {
        TriangleMesh m;
        if (!read_mesh(argv[2], m)) return 1;
        FILE* f = std::fopen(argv[3], "rb");
        if (!f) return 1;
        int64_t np = 0;
        if (std::fread(&np, 8, 1, f) != 1 || np < 0) {
            std::fclose(f);
            return 1;
        }
        std::vector<vec3> pts((size_t)np);
        const bool ok = std::fread(pts.data(), 12, pts.size(), f) == pts.size();
        std::fclose(f);
        if (!ok) return 1;
        std::vector<float> w;
        float ms = 0.f;
        const bool dist = !std::strcmp(argv[1], "dist");
        if (dist ? !CudaRasterizer::query_distance_batch_gpu(m, pts, w, ms)
                 : !CudaRasterizer::compute_winding_numbers_gpu(m, pts, w, ms))
            return 2;
        FILE* o = std::fopen(argv[4], "wb");
        if (!o) return 1;
        std::fwrite(w.data(), 4, w.size(), o);
        std::fclose(o);
        return 0;
    }
    if (argc == 4 && (!std::strcmp(argv[1], "mc") || !std::strcmp(argv[1], "vol"))) {
        // grid file: int32 dims[3], float origin[3], float h, float iso, dims product floats
        FILE* f = std::fopen(argv[2], "rb");
        if (!f) return 1;
        int32_t d[3];
        float org[3], h = 0.f, iso = 0.f;
        bool ok = std::fread(d, 4, 3, f) == 3 && std::fread(org, 4, 3, f) == 3 && std::fread(&h, 4, 1, f) == 1 &&
                  std::fread(&iso, 4, 1, f) == 1;
        CudaVolumeGrid g;
        if (ok) {
            g.data.resize((size_t)d[0] * (size_t)d[1] * (size_t)d[2]);
            ok = std::fread(g.data.data(), 4, g.data.size(), f) == g.data.size();
        }
        std::fclose(f);
        if (!ok) return 1;
        g.dims = {d[0], d[1], d[2]};
        g.voxel_size = h;
        FILE* o = std::fopen(argv[3], "wb");
        if (!o) return 1;
        float ms = 0.f;
        if (!std::strcmp(argv[1], "vol")) {
            float vol = 0.f;
            if (!CudaRasterizer::compute_sdf_interior_volume_gpu(g, vol, ms)) return 2;
            std::fwrite(&vol, 4, 1, o);
        } else {
            TriangleMesh m;
            if (!CudaRasterizer::extract_iso_surface(g.data, g.dims, h, vec3{org[0], org[1], org[2]}, iso, m, ms))
                return 2;
            const int64_t nv = (int64_t)m.vertex_count(), nt = (int64_t)m.triangle_count();
            std::fwrite(&nv, 8, 1, o);
            std::fwrite(&nt, 8, 1, o);
            std::fwrite(m.vertices.data(), 4, m.vertices.size(), o);
            std::fwrite(m.indices.data(), 4, m.indices.size(), o);
        }
        std::fclose(o);
        return 0;
    }
    if (argc == 4 && !std::strcmp(argv[1], "orient")) {
        // the winding pre-pass of the OBJ loader: out.bin = the oriented indices
        TriangleMesh m;
        if (!read_mesh(argv[2], m)) return 1;
        const WindingPrepassResult r = MeshWindingPass::orient_mesh(m);
        FILE* o = std::fopen(argv[3], "wb");
        if (!o) return 1;
        std::fwrite(m.indices.data(), 4, m.indices.size(), o);
        std::fclose(o);
        std::printf("orientable %d inconsistent %d inverted %d volume %.6f\n%s\n", r.is_orientable ? 1 : 0,
                    r.had_inconsistent_winding ? 1 : 0, r.was_inverted ? 1 : 0, r.signed_volume, r.message.c_str());
        return 0;
    }
    std::fprintf(stderr, "usage: dropin_main info | sdf mesh.bin h out.bin | wind|dist mesh.bin pts.bin out.bin | mc|vol grid.bin out.bin | orient mesh.bin out.bin\n");
    return 1;
}
