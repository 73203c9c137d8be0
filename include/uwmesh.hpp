// uwmesh.hpp -- C++ host facade over the C ABI (uwmesh.h), mirroring the reference crate's
// chunk-builder interface so that a caller of underwater_world/src/chunk.rs finds the same names,
// argument meaning and error behaviour:
//
//   uw::Perlin(seed)                 noise::Perlin::new(seed)               state.rs:358-359
//   uw::Chunk(x, y, z)               Chunk::new((x, y, z))                  chunk.rs:89-103
//   Chunk::build_full(perlin)        Chunk::build_full(&perlin, &device)    chunk.rs:266-268
//   Chunk::build_partial(perlin)     Chunk::build_partial -> finished       chunk.rs:270-313
//   Chunk::not_blank()               Chunk::not_blank                       chunk.rs:344
//   Chunk::verts_buffer_slice()      Chunk::verts_buffer_slice              chunk.rs:346
//   Chunk::inds_buffer_slice()       Chunk::inds_buffer_slice               chunk.rs:347
//   Chunk::num_inds()                Chunk::num_inds                        chunk.rs:348
//   Chunk::tris_around(pct, range)   Chunk::tris_around -> Vec<Tri>         chunk.rs:315-342
//   uw::Tri                          util::Tri {verts, normal}              util.rs:7-20
//   uw::ChunkBatch::build(...)       World::build_full_step / build_step    world.rs:113-145
//                                    for a whole candidate set in one launch
//
// Differences, by design: the build finishes in one call (build_partial returns true); vertices
// are xyz f32 world positions unless the chunk is made with Format::VertColor (the reference's
// 6-float VertColor, draw.rs:4-9); indices are u32 unless made with u16 = true (the reference's
// `ind as u16`, chunk.rs:243). The reference panics when a blank chunk's buffers are requested
// (chunk.rs:345-347); this facade throws uw::Error.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "uwmesh.h"

namespace uw {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const uw_ctx* ctx, const char* what) {
    if (rc != UW_OK) {
        std::string m = std::string(what) + " failed (" + std::to_string(rc) + ")";
        if (ctx) m += ": " + std::string(uw_last_error(ctx));
        throw Error(rc, m);
    }
}

// Density field on one GPU: permutation table of `seed` plus the field parameters.
class Perlin {
public:
    // The reference crate's constants (chunk.rs:5-12, perlin_util.rs:6-29) with this seed.
    explicit Perlin(uint32_t seed, int device = 0) {
        uw_params p;
        uw_params_reference(&p, seed);
        create(p, device);
    }
    Perlin(const uw_params& p, int device = 0) { create(p, device); }
    ~Perlin() {
        if (ctx_) uw_ctx_destroy(ctx_);
    }
    Perlin(const Perlin&) = delete;
    Perlin& operator=(const Perlin&) = delete;
    Perlin(Perlin&& o) noexcept : ctx_(o.ctx_), params_(o.params_) { o.ctx_ = nullptr; }

    uw_ctx* ctx() const { return ctx_; }
    const uw_params& params() const { return params_; }

private:
    void create(const uw_params& p, int device) {
        params_ = p;
        check(uw_ctx_create(&p, device, &ctx_), nullptr, "uw_ctx_create");
    }
    uw_ctx* ctx_ = nullptr;
    uw_params params_{};
};

// Collision triangle: world-space vertices and unit normal (util.rs:7-20).
struct Tri {
    float verts[3][3];
    float normal[3];
};

class Chunk {
public:
    static constexpr int32_t INTERNAL_SIZE = 12;  // chunk.rs:6
    enum class Format { Xyz, VertColor };

    Chunk(int32_t x, int32_t y, int32_t z, int32_t resolution = INTERNAL_SIZE, Format format = Format::Xyz,
          bool u16 = false)
        : pos_{x, y, z}, n_(resolution), fmt_(format), u16_(u16) {}

    void build_full(const Perlin& perlin) {
        while (!build_partial(perlin)) {
        }
    }

    // One call builds the whole chunk; returns true ("finished") like the last slice of the
    // reference's state machine.
    bool build_partial(const Perlin& perlin) {
        if (done_) return true;
        const size_t vfl = floats_per_vertex();
        const size_t icap = size_t(uw_max_inds(n_));
        verts_.resize(size_t(uw_max_verts(n_)) * vfl);
        if (u16_) inds16_.resize(icap);
        else inds_.resize(icap);
        normals_.resize(icap);  // 3 floats per triangle
        tri_cells_.resize(icap / 3);
        uw_host_out out = {verts_.data(),
                           verts_.size() / vfl,
                           u16_ ? static_cast<void*>(inds16_.data()) : static_cast<void*>(inds_.data()),
                           icap,
                           fmt_ == Format::VertColor ? uint32_t(UW_VERTEX_XYZ_RGB) : uint32_t(UW_VERTEX_XYZ),
                           u16_ ? uint32_t(UW_INDEX_U16) : uint32_t(UW_INDEX_U32),
                           normals_.data(),
                           tri_cells_.data()};
        check(uw_chunk_build_host_out(perlin.ctx(), pos_[0], pos_[1], pos_[2], n_, &out, &info_), perlin.ctx(),
              "uw_chunk_build_host_out");
        verts_.resize(size_t(info_.n_verts) * vfl);
        if (u16_) inds16_.resize(size_t(info_.n_inds));
        else inds_.resize(size_t(info_.n_inds));
        normals_.resize(size_t(info_.n_inds));
        tri_cells_.resize(size_t(info_.n_inds) / 3);
        verts_.shrink_to_fit();
        inds_.shrink_to_fit();
        inds16_.shrink_to_fit();
        normals_.shrink_to_fit();
        tri_cells_.shrink_to_fit();
        done_ = true;
        return true;
    }

    // Collision triangles of the cells within `range` cells of the chunk-local position
    // local_pos_percent (position / chunk extent), in the reference's order (chunk.rs:315-342).
    std::vector<Tri> tris_around(const float local_pos_percent[3], int32_t range) const {
        std::vector<Tri> tris;
        if (!done_ || tri_cells_.empty()) return tris;
        std::vector<uint32_t> ids(tri_cells_.size());
        uint32_t count = 0;
        check(uw_tris_around(tri_cells_.data(), uint32_t(tri_cells_.size()), n_, local_pos_percent[0],
                             local_pos_percent[1], local_pos_percent[2], range, ids.data(), uint32_t(ids.size()),
                             &count),
              nullptr, "uw_tris_around");
        const size_t vfl = floats_per_vertex();
        tris.resize(count);
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t t = ids[k];
            for (int c = 0; c < 3; ++c) {
                const size_t v = u16_ ? inds16_[3 * t + c] : inds_[3 * t + c];
                for (int d = 0; d < 3; ++d) tris[k].verts[c][d] = verts_[v * vfl + d];
            }
            for (int d = 0; d < 3; ++d) tris[k].normal[d] = normals_[3 * t + d];
        }
        return tris;
    }

    bool not_blank() const { return done_ && info_.n_inds > 0; }
    bool early_skipped() const { return done_ && (info_.flags & UW_CHUNK_EARLY_SKIP); }
    const std::vector<float>& verts_buffer_slice() const {
        if (!not_blank()) throw Error(UW_ERR_INVALID_ARG, "verts_buffer_slice() on a blank chunk");
        return verts_;
    }
    const std::vector<uint32_t>& inds_buffer_slice() const {
        if (!not_blank()) throw Error(UW_ERR_INVALID_ARG, "inds_buffer_slice() on a blank chunk");
        if (u16_) throw Error(UW_ERR_INVALID_ARG, "inds_buffer_slice() on a u16 chunk: use inds16_buffer_slice()");
        return inds_;
    }
    const std::vector<uint16_t>& inds16_buffer_slice() const {
        if (!not_blank()) throw Error(UW_ERR_INVALID_ARG, "inds16_buffer_slice() on a blank chunk");
        if (!u16_) throw Error(UW_ERR_INVALID_ARG, "inds16_buffer_slice() on a u32 chunk");
        return inds16_;
    }
    const std::vector<float>& tri_normals() const { return normals_; }
    const std::vector<uint32_t>& tri_cells() const { return tri_cells_; }
    size_t num_inds() const { return info_.n_inds; }
    size_t num_verts() const { return info_.n_verts; }
    const uw_chunk_info& info() const { return info_; }

private:
    size_t floats_per_vertex() const { return fmt_ == Format::VertColor ? 6 : 3; }

    int32_t pos_[3];
    int32_t n_;
    Format fmt_;
    bool u16_;
    bool done_ = false;
    uw_chunk_info info_{};
    std::vector<float> verts_;
    std::vector<uint32_t> inds_;
    std::vector<uint16_t> inds16_;
    std::vector<float> normals_;
    std::vector<uint32_t> tri_cells_;
};

// A whole chunk set built by one persistent launch (uw_build_chunks_batched); results copied back
// to host memory per chunk. Device buffers are reused across builds of the same or smaller size.
class ChunkBatch {
public:
    struct Mesh {
        uw_chunk_info info;
        std::vector<float> verts;
        std::vector<uint32_t> inds;
    };

    explicit ChunkBatch(const Perlin& perlin) : perlin_(perlin) {}
    ~ChunkBatch() { release(); }
    ChunkBatch(const ChunkBatch&) = delete;
    ChunkBatch& operator=(const ChunkBatch&) = delete;

    // coords: 3 ints per chunk; res: cells per axis per chunk (1..UW_BATCH_MAX_RESOLUTION).
    std::vector<Mesh> build(const std::vector<int32_t>& coords, const std::vector<int32_t>& res) {
        const size_t n = res.size();
        if (coords.size() != 3 * n) throw Error(UW_ERR_INVALID_ARG, "coords must hold 3 ints per chunk");
        uint64_t vcap = 0, icap = 0;
        for (int32_t r : res) {
            vcap += uw_max_verts(r);
            icap += uw_max_inds(r);
        }
        reserve(n, vcap, icap);
        hip(hipMemcpy(d_coords_, coords.data(), coords.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        hip(hipMemcpy(d_res_, res.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
        uw_mesh_out out = {d_verts_, d_inds_, d_info_, d_totals_, vcap, icap, UW_VERTEX_XYZ, UW_INDEX_U32, nullptr,
                           nullptr};
        check(uw_build_chunks_batched(perlin_.ctx(), d_coords_, d_res_, int32_t(n), &out, nullptr), perlin_.ctx(),
              "uw_build_chunks_batched");
        hip(hipDeviceSynchronize());
        std::vector<uw_chunk_info> info(n);
        hip(hipMemcpy(info.data(), d_info_, n * sizeof(uw_chunk_info), hipMemcpyDeviceToHost));
        std::vector<Mesh> meshes(n);
        for (size_t i = 0; i < n; ++i) {
            Mesh& m = meshes[i];
            m.info = info[i];
            m.verts.resize(size_t(info[i].n_verts) * 3);
            m.inds.resize(info[i].n_inds);
            if (info[i].n_inds == 0 || (info[i].flags & UW_CHUNK_OVERFLOW)) continue;
            hip(hipMemcpy(m.verts.data(), d_verts_ + 3 * info[i].vert_off, m.verts.size() * sizeof(float),
                          hipMemcpyDeviceToHost));
            hip(hipMemcpy(m.inds.data(), d_inds_ + info[i].ind_off, m.inds.size() * sizeof(uint32_t),
                          hipMemcpyDeviceToHost));
        }
        return meshes;
    }

private:
    static void hip(hipError_t e) {
        if (e != hipSuccess) throw Error(UW_ERR_HIP, hipGetErrorString(e));
    }
    void reserve(size_t n, uint64_t vcap, uint64_t icap) {
        if (n <= cap_n_ && vcap <= cap_v_ && icap <= cap_i_) return;
        release();
        hip(hipMalloc(&d_coords_, 3 * n * sizeof(int32_t)));
        hip(hipMalloc(&d_res_, n * sizeof(int32_t)));
        hip(hipMalloc(&d_info_, n * sizeof(uw_chunk_info)));
        hip(hipMalloc(&d_totals_, sizeof(uw_totals)));
        hip(hipMalloc(&d_verts_, vcap * 3 * sizeof(float)));
        hip(hipMalloc(&d_inds_, icap * sizeof(uint32_t)));
        cap_n_ = n;
        cap_v_ = vcap;
        cap_i_ = icap;
    }
    void release() {
        (void)hipFree(d_coords_);
        (void)hipFree(d_res_);
        (void)hipFree(d_info_);
        (void)hipFree(d_totals_);
        (void)hipFree(d_verts_);
        (void)hipFree(d_inds_);
        d_coords_ = d_res_ = nullptr;
        d_info_ = nullptr;
        d_totals_ = nullptr;
        d_verts_ = nullptr;
        d_inds_ = nullptr;
        cap_n_ = cap_v_ = cap_i_ = 0;
    }

    const Perlin& perlin_;
    int32_t* d_coords_ = nullptr;
    int32_t* d_res_ = nullptr;
    uw_chunk_info* d_info_ = nullptr;
    uw_totals* d_totals_ = nullptr;
    float* d_verts_ = nullptr;
    uint32_t* d_inds_ = nullptr;
    size_t cap_n_ = 0;
    uint64_t cap_v_ = 0, cap_i_ = 0;
};

}  // namespace uw
