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
//   uw::ChunkBatch::build(...)       World::build_full_step / build_step    world.rs:113-145
//                                    for a whole candidate set in one launch
//
// Differences, by design: the build finishes in one call (build_partial returns true); vertices
// are xyz f32 world positions (the reference's VertColor carries a colour too, draw.rs:4-9);
// indices are u32 (the reference truncates to u16, chunk.rs:243). The reference panics when a
// blank chunk's buffers are requested (chunk.rs:345-347); this facade throws uw::Error.
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

class Chunk {
public:
    static constexpr int32_t INTERNAL_SIZE = 12;  // chunk.rs:6

    Chunk(int32_t x, int32_t y, int32_t z, int32_t resolution = INTERNAL_SIZE) : pos_{x, y, z}, n_(resolution) {}

    void build_full(const Perlin& perlin) {
        while (!build_partial(perlin)) {
        }
    }

    // One call builds the whole chunk; returns true ("finished") like the last slice of the
    // reference's state machine.
    bool build_partial(const Perlin& perlin) {
        if (done_) return true;
        verts_.resize(size_t(uw_max_verts(n_)) * 3);
        inds_.resize(size_t(uw_max_inds(n_)));
        check(uw_chunk_build_host(perlin.ctx(), pos_[0], pos_[1], pos_[2], n_, verts_.data(), verts_.size() / 3,
                                  inds_.data(), inds_.size(), &info_),
              perlin.ctx(), "uw_chunk_build_host");
        verts_.resize(size_t(info_.n_verts) * 3);
        inds_.resize(size_t(info_.n_inds));
        verts_.shrink_to_fit();
        inds_.shrink_to_fit();
        done_ = true;
        return true;
    }

    bool not_blank() const { return done_ && info_.n_inds > 0; }
    bool early_skipped() const { return done_ && (info_.flags & UW_CHUNK_EARLY_SKIP); }
    const std::vector<float>& verts_buffer_slice() const {
        if (!not_blank()) throw Error(UW_ERR_INVALID_ARG, "verts_buffer_slice() on a blank chunk");
        return verts_;
    }
    const std::vector<uint32_t>& inds_buffer_slice() const {
        if (!not_blank()) throw Error(UW_ERR_INVALID_ARG, "inds_buffer_slice() on a blank chunk");
        return inds_;
    }
    size_t num_inds() const { return info_.n_inds; }
    size_t num_verts() const { return info_.n_verts; }
    const uw_chunk_info& info() const { return info_; }

private:
    int32_t pos_[3];
    int32_t n_;
    bool done_ = false;
    uw_chunk_info info_{};
    std::vector<float> verts_;
    std::vector<uint32_t> inds_;
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
        uw_mesh_out out = {d_verts_, d_inds_, d_info_, d_totals_, vcap, icap, UW_VERTEX_XYZ, 0u};
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
