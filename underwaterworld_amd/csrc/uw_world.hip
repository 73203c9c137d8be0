// World streaming with a batched generation step (underwater_world/src/world.rs:64-247).
//
// The planner restates World::update_nearby and GenPrio (world.rs:29-60,149-228) on the host in
// the reference's f32 operation order (this file is compiled with -ffp-contract=off). The build
// step differs by design: where the reference builds one chunk per frame (build_full_step /
// build_step, world.rs:113-145), uw_world_update builds up to max_batch pending candidates in one
// persistent launch (uw_build_chunks_batched) and then applies the reference's per-chunk rules to
// the results in pop order. Each launch's output is compacted into an exact-size arena; a chunk
// refers to its slice of an arena and the arena is freed when its last chunk is evicted.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <numeric>
#include <vector>

#include "uw_internal.h"

namespace {

using Key = std::array<int32_t, 3>;

constexpr int32_t kChunkSize = 16;  // chunk.rs:5 (world units per chunk edge)

float magnitude(float x, float y, float z) { return std::sqrt((x * x + y * y) + z * z); }

// util::in_frustum (util.rs:74-83): m column-major; row . (p, 1) summed left to right.
bool in_frustum(const float p[3], const float* m) {
    const float v[4] = {p[0], p[1], p[2], 1.0f};
    float r[4];
    for (int row = 0; row < 4; ++row) {
        float acc = m[row] * v[0];
        for (int col = 1; col < 4; ++col) acc = acc + m[4 * col + row] * v[col];
        r[row] = acc;
    }
    const float x = r[0] / r[3], y = r[1] / r[3], z = r[2] / r[3];
    return std::fabs(x) <= 1.0f && std::fabs(y) <= 1.0f && z >= 0.0f && z <= 1.0f;
}

int32_t rust_f32_to_i32(float v) {
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return -2147483647 - 1;
    return int32_t(v);
}

// Sub::chunk (sub.rs:457-463).
Key sub_chunk(const float pos[3]) {
    Key k;
    for (int i = 0; i < 3; ++i) k[i] = rust_f32_to_i32(std::floor(pos[i] / float(kChunkSize)));
    return k;
}

int64_t dist_sq(const Key& a, const Key& b) {
    int64_t s = 0;
    for (int i = 0; i < 3; ++i) s += int64_t(a[i] - b[i]) * int64_t(a[i] - b[i]);
    return s;
}

// cgmath Vector3::angle: atan2(|a x b|, a . b).
float angle(const float a[3], const float b[3]) {
    const float cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    const float dot = (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
    return std::atan2(magnitude(cx, cy, cz), dot);
}

bool config_ok(const uw_world_config* c) {
    return c && c->generation_dist >= 1 && c->generation_dist <= 64 && c->view_dist >= 0 && c->keep_dist >= 0 &&
           c->min_z <= c->max_z && c->resolution >= 1 && c->resolution <= UW_BATCH_MAX_RESOLUTION &&
           c->max_batch >= 0 && c->vertex_format <= 1 && c->index_format <= 1;
}

// World::update_nearby. lookup(pos) returns -1 (absent), 0 (blank) or 1 (not blank).
template <class Lookup>
void plan(const uw_world_config& cfg, const float sp[3], const float* view_proj, const float* gen_proj, Lookup lookup,
          std::vector<Key>& render, std::vector<uw_world_candidate>& cands) {
    render.clear();
    cands.clear();
    const Key sc = sub_chunk(sp);
    const float max_view = float(cfg.view_dist * kChunkSize);
    const float max_gen = float(cfg.generation_dist * kChunkSize);
    const int32_t start_z = std::max(sc[2] - cfg.generation_dist, cfg.min_z);
    const int32_t end_z = std::min(sc[2] + cfg.generation_dist, cfg.max_z);
    static const float kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                        {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    const float cs = float(kChunkSize);
    for (int32_t x = -cfg.generation_dist; x < cfg.generation_dist; ++x) {
        const int32_t cx = sc[0] + x;
        for (int32_t y = -cfg.generation_dist; y < cfg.generation_dist; ++y) {
            const int32_t cy = sc[1] + y;
            for (int32_t cz = start_z; cz <= end_z; ++cz) {
                const float center[3] = {(float(cx) + 0.5f) * cs, (float(cy) + 0.5f) * cs, (float(cz) + 0.5f) * cs};
                const float dist = magnitude(sp[0] - center[0], sp[1] - center[1], sp[2] - center[2]);
                if (dist > max_gen) continue;
                bool in_view = false, in_gen = false;
                for (const auto& c : kCorner) {
                    const float corner[3] = {(float(cx) + c[0]) * cs, (float(cy) + c[1]) * cs, (float(cz) + c[2]) * cs};
                    if (in_frustum(corner, gen_proj)) {
                        in_gen = true;
                        if (in_frustum(corner, view_proj)) {
                            in_view = true;
                            break;
                        }
                    }
                }
                const Key pos = {cx, cy, cz};
                const int known = lookup(pos);
                if (known >= 0) {
                    if (dist < max_view && known == 1 && in_view) render.push_back(pos);
                } else {
                    uw_world_candidate w;
                    w.pos[0] = cx;
                    w.pos[1] = cy;
                    w.pos[2] = cz;
                    w.dist = dist;
                    w.z = float(cz) * cs;
                    w.in_view = in_view;
                    w.in_gen = in_gen;
                    cands.push_back(w);
                }
            }
        }
    }
    // GenPrio::compare ascending = the reference's pop order (it sorts descending and pops the
    // back); exact ties in reverse generation order.
    std::vector<int32_t> idx(cands.size());
    std::iota(idx.begin(), idx.end(), 0);
    auto rank = [&](const uw_world_candidate& c) { return (c.in_view ? 0 : 2) + (c.in_gen ? 0 : 1); };
    std::sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
        const uw_world_candidate &p = cands[size_t(a)], &q = cands[size_t(b)];
        if (rank(p) != rank(q)) return rank(p) < rank(q);
        const float sp_ = p.dist * p.dist + p.z, sq_ = q.dist * q.dist + q.z;
        if (sp_ != sq_) return sp_ < sq_;
        return a > b;
    });
    std::vector<uw_world_candidate> sorted(cands.size());
    for (size_t i = 0; i < idx.size(); ++i) sorted[i] = cands[size_t(idx[i])];
    cands.swap(sorted);
}

}  // namespace

struct uw_world_arena {
    void* verts = nullptr;
    void* inds = nullptr;
    float* normals = nullptr;
    uint32_t* cells = nullptr;
    int32_t live = 0;
};

struct uw_world_stored {
    uw_chunk_info info;
    int32_t arena;  // -1 for a blank chunk
    uint64_t seq;   // insertion order (the eviction walk's key order)
};

struct uw_world {
    uw_ctx* ctx = nullptr;
    uw_world_config cfg{};
    std::map<Key, uw_world_stored> chunks;
    std::vector<Key> render;
    std::vector<uw_world_candidate> to_generate;  // pop order
    size_t gen_head = 0;                          // candidates before this index are consumed
    std::vector<Key> keys_left;
    bool should_full_build = true;
    float last_pos[3];
    float last_bearing[3];
    uint64_t next_seq = 0;
    std::vector<uw_world_arena> arenas;
    std::vector<int32_t> free_arenas;
    // staging of one launch
    int32_t* d_coords = nullptr;
    int32_t* d_res = nullptr;
    uw_chunk_info* d_info = nullptr;
    uw_totals* d_totals = nullptr;
    float* d_verts = nullptr;
    void* d_inds = nullptr;
    float* d_norm = nullptr;
    uint32_t* d_cells = nullptr;
    int32_t cap_chunks = 0;
};

namespace {

size_t vert_bytes(const uw_world_config& c) { return c.vertex_format == UW_VERTEX_XYZ_RGB ? 24 : 12; }
size_t ind_bytes(const uw_world_config& c) { return c.index_format == UW_INDEX_U16 ? 2 : 4; }

void release_staging(uw_world* w) {
    (void)hipFree(w->d_coords);
    (void)hipFree(w->d_res);
    (void)hipFree(w->d_info);
    (void)hipFree(w->d_totals);
    (void)hipFree(w->d_verts);
    (void)hipFree(w->d_inds);
    (void)hipFree(w->d_norm);
    (void)hipFree(w->d_cells);
    w->d_coords = w->d_res = nullptr;
    w->d_info = nullptr;
    w->d_totals = nullptr;
    w->d_verts = nullptr;
    w->d_inds = nullptr;
    w->d_norm = nullptr;
    w->d_cells = nullptr;
    w->cap_chunks = 0;
}

void release_arena(uw_world* w, int32_t a) {
    uw_world_arena& r = w->arenas[size_t(a)];
    if (--r.live > 0) return;
    (void)hipFree(r.verts);
    (void)hipFree(r.inds);
    (void)hipFree(r.normals);
    (void)hipFree(r.cells);
    r = uw_world_arena{};
    w->free_arenas.push_back(a);
}

void evict(uw_world* w, std::map<Key, uw_world_stored>::iterator it) {
    if (it->second.arena >= 0) release_arena(w, it->second.arena);
    w->chunks.erase(it);
}

// World::remove_far_way (world.rs:230-241): one key per update.
void remove_far_way(uw_world* w, const Key& sc) {
    if (!w->keys_left.empty()) {
        const Key pos = w->keys_left.back();
        w->keys_left.pop_back();
        if (dist_sq(pos, sc) >= int64_t(w->cfg.keep_dist) * w->cfg.keep_dist) {
            auto it = w->chunks.find(pos);
            if (it != w->chunks.end()) evict(w, it);
        }
    } else {
        std::vector<std::pair<uint64_t, Key>> order;
        order.reserve(w->chunks.size());
        for (const auto& kv : w->chunks) order.emplace_back(kv.second.seq, kv.first);
        std::sort(order.begin(), order.end());
        w->keys_left.clear();
        for (const auto& o : order) w->keys_left.push_back(o.second);
    }
}

int ensure_staging(uw_world* w, int32_t k) {
    if (k <= w->cap_chunks) return UW_OK;
    release_staging(w);
    const uint64_t n = uint64_t(k);
    const uint64_t vcap = n * uw_max_verts(w->cfg.resolution), icap = n * uw_max_inds(w->cfg.resolution);
    uw_ctx* ctx = w->ctx;
    UW_HIP(ctx, hipMalloc(&w->d_coords, n * 3 * sizeof(int32_t)));
    UW_HIP(ctx, hipMalloc(&w->d_res, n * sizeof(int32_t)));
    UW_HIP(ctx, hipMalloc(&w->d_info, n * sizeof(uw_chunk_info)));
    UW_HIP(ctx, hipMalloc(&w->d_totals, sizeof(uw_totals)));
    UW_HIP(ctx, hipMalloc(&w->d_verts, vcap * vert_bytes(w->cfg)));
    UW_HIP(ctx, hipMalloc(&w->d_inds, icap * ind_bytes(w->cfg)));
    if (w->cfg.collision) {
        UW_HIP(ctx, hipMalloc(&w->d_norm, icap / 3 * 3 * sizeof(float)));
        UW_HIP(ctx, hipMalloc(&w->d_cells, icap / 3 * sizeof(uint32_t)));
    }
    std::vector<int32_t> res(size_t(k), w->cfg.resolution);
    UW_HIP(ctx, hipMemcpy(w->d_res, res.data(), res.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    w->cap_chunks = k;
    return UW_OK;
}

// Builds the next k candidates in one launch and stores them in pop order.
int build_batch(uw_world* w, int32_t k, const Key& sc) {
    uw_ctx* ctx = w->ctx;
    if (int rc = ensure_staging(w, k)) return rc;
    std::vector<int32_t> coords(size_t(k) * 3);
    for (int32_t i = 0; i < k; ++i)
        for (int d = 0; d < 3; ++d) coords[size_t(i) * 3 + d] = w->to_generate[w->gen_head + size_t(i)].pos[d];
    UW_HIP(ctx, hipMemcpy(w->d_coords, coords.data(), coords.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    const uint64_t vcap = uint64_t(k) * uw_max_verts(w->cfg.resolution);
    const uint64_t icap = uint64_t(k) * uw_max_inds(w->cfg.resolution);
    uw_mesh_out out = {w->d_verts, static_cast<uint32_t*>(w->d_inds), w->d_info, w->d_totals, vcap, icap,
                       w->cfg.vertex_format, w->cfg.index_format, w->d_norm, w->d_cells};
    if (int rc = uw_build_chunks_batched(ctx, w->d_coords, w->d_res, k, &out, nullptr)) return rc;
    UW_HIP(ctx, hipDeviceSynchronize());
    std::vector<uw_chunk_info> info(static_cast<size_t>(k));
    uw_totals tot;
    UW_HIP(ctx, hipMemcpy(info.data(), w->d_info, info.size() * sizeof(uw_chunk_info), hipMemcpyDeviceToHost));
    UW_HIP(ctx, hipMemcpy(&tot, w->d_totals, sizeof(tot), hipMemcpyDeviceToHost));
    if (tot.overflow) return uw_fail(ctx, UW_ERR_OVERFLOW, "world batch overflowed its worst-case staging");
    int32_t arena = -1;
    if (tot.n_inds > 0) {
        if (!w->free_arenas.empty()) {
            arena = w->free_arenas.back();
            w->free_arenas.pop_back();
        } else {
            arena = int32_t(w->arenas.size());
            w->arenas.emplace_back();
        }
        uw_world_arena& a = w->arenas[size_t(arena)];
        const uint64_t ntri = tot.n_inds / 3;
        UW_HIP(ctx, hipMalloc(&a.verts, tot.n_verts * vert_bytes(w->cfg)));
        UW_HIP(ctx, hipMalloc(&a.inds, tot.n_inds * ind_bytes(w->cfg)));
        UW_HIP(ctx, hipMemcpy(a.verts, w->d_verts, tot.n_verts * vert_bytes(w->cfg), hipMemcpyDeviceToDevice));
        UW_HIP(ctx, hipMemcpy(a.inds, w->d_inds, tot.n_inds * ind_bytes(w->cfg), hipMemcpyDeviceToDevice));
        if (w->cfg.collision) {
            UW_HIP(ctx, hipMalloc(&a.normals, ntri * 3 * sizeof(float)));
            UW_HIP(ctx, hipMalloc(&a.cells, ntri * sizeof(uint32_t)));
            UW_HIP(ctx, hipMemcpy(a.normals, w->d_norm, ntri * 3 * sizeof(float), hipMemcpyDeviceToDevice));
            UW_HIP(ctx, hipMemcpy(a.cells, w->d_cells, ntri * sizeof(uint32_t), hipMemcpyDeviceToDevice));
        }
    }
    const bool full = w->should_full_build;
    const int64_t near = int64_t(w->cfg.view_dist + 1) * (w->cfg.view_dist + 1);
    for (int32_t i = 0; i < k; ++i) {
        const uw_world_candidate& c = w->to_generate[w->gen_head + size_t(i)];
        const Key pos = {c.pos[0], c.pos[1], c.pos[2]};
        const bool not_blank = info[size_t(i)].n_inds > 0;
        // build_full_step pushes every non-blank chunk; build_step only those near the sub
        if (not_blank && (full || dist_sq(pos, sc) <= near)) w->render.push_back(pos);
        auto it = w->chunks.find(pos);
        if (it != w->chunks.end()) evict(w, it);
        uw_world_stored s;
        s.info = info[size_t(i)];
        s.arena = not_blank ? arena : -1;
        s.seq = w->next_seq++;
        if (not_blank) w->arenas[size_t(arena)].live++;
        w->chunks.emplace(pos, s);
    }
    w->gen_head += size_t(k);
    return UW_OK;
}

}  // namespace

extern "C" {

void uw_world_config_reference(uw_world_config* c) {
    if (!c) return;
    c->view_dist = 4;
    c->generation_dist = 5;
    c->keep_dist = 6;
    c->max_z = 2;
    c->min_z = -2;
    c->recheck_dist = 4.0f;
    c->recheck_angle = 0.33f;
    c->resolution = 12;
    c->max_batch = 0;
    c->vertex_format = UW_VERTEX_XYZ;
    c->index_format = UW_INDEX_U32;
    c->collision = 1;
}

int uw_world_plan(const uw_world_config* cfg, const float sub_pos[3], const float view_proj[16],
                  const float gen_proj[16], const int32_t* known, const uint8_t* known_not_blank, int32_t nknown,
                  int32_t* render, int32_t render_cap, int32_t* n_render, uw_world_candidate* cands, int32_t cand_cap,
                  int32_t* n_cand) {
    if (!config_ok(cfg) || !sub_pos || !view_proj || !gen_proj || nknown < 0 ||
        (nknown > 0 && (!known || !known_not_blank)) || !n_render || !n_cand || render_cap < 0 || cand_cap < 0 ||
        (render_cap > 0 && !render) || (cand_cap > 0 && !cands))
        return UW_ERR_INVALID_ARG;
    std::map<Key, int> map;
    for (int32_t i = 0; i < nknown; ++i)
        map[Key{known[3 * i], known[3 * i + 1], known[3 * i + 2]}] = known_not_blank[i] ? 1 : 0;
    std::vector<Key> r;
    std::vector<uw_world_candidate> c;
    plan(*cfg, sub_pos, view_proj, gen_proj,
         [&](const Key& k) {
             auto it = map.find(k);
             return it == map.end() ? -1 : it->second;
         },
         r, c);
    *n_render = int32_t(r.size());
    *n_cand = int32_t(c.size());
    for (int32_t i = 0; i < std::min<int32_t>(render_cap, *n_render); ++i)
        for (int d = 0; d < 3; ++d) render[3 * i + d] = r[size_t(i)][d];
    for (int32_t i = 0; i < std::min<int32_t>(cand_cap, *n_cand); ++i) cands[i] = c[size_t(i)];
    return (*n_render > render_cap || *n_cand > cand_cap) ? UW_ERR_OVERFLOW : UW_OK;
}

int uw_world_create(uw_ctx* ctx, const uw_world_config* cfg, uw_world** out) {
    if (!ctx || !out) return UW_ERR_INVALID_ARG;
    if (!config_ok(cfg)) return uw_fail(ctx, UW_ERR_INVALID_ARG, "invalid world configuration");
    uw_world* w = new uw_world();
    w->ctx = ctx;
    w->cfg = *cfg;
    for (int i = 0; i < 3; ++i) w->last_pos[i] = w->last_bearing[i] = INFINITY;  // world.rs:84-85
    *out = w;
    return UW_OK;
}

void uw_world_destroy(uw_world* w) {
    if (!w) return;
    (void)hipSetDevice(w->ctx->device);
    release_staging(w);
    for (auto& a : w->arenas) {
        (void)hipFree(a.verts);
        (void)hipFree(a.inds);
        (void)hipFree(a.normals);
        (void)hipFree(a.cells);
    }
    delete w;
}

int uw_world_update(uw_world* w, const float sub_pos[3], const float sub_bearing[3], const float view_proj[16],
                    const float gen_proj[16], int sub_reset, int32_t* built) {
    if (!w || !sub_pos || !sub_bearing || !view_proj || !gen_proj) return UW_ERR_INVALID_ARG;
    uw_ctx* ctx = w->ctx;
    UW_HIP(ctx, hipSetDevice(ctx->device));
    if (built) *built = 0;
    const Key sc = sub_chunk(sub_pos);
    remove_far_way(w, sc);
    const float dist = magnitude(sub_pos[0] - w->last_pos[0], sub_pos[1] - w->last_pos[1], sub_pos[2] - w->last_pos[2]);
    const float ang = angle(sub_bearing, w->last_bearing);
    if (sub_reset || dist > w->cfg.recheck_dist || ang > w->cfg.recheck_angle) {
        plan(w->cfg, sub_pos, view_proj, gen_proj,
             [&](const Key& k) {
                 auto it = w->chunks.find(k);
                 return it == w->chunks.end() ? -1 : (it->second.info.n_inds > 0 ? 1 : 0);
             },
             w->render, w->to_generate);
        w->gen_head = 0;
        for (int i = 0; i < 3; ++i) {
            w->last_pos[i] = sub_pos[i];
            w->last_bearing[i] = sub_bearing[i];
        }
    }
    const size_t pending = w->to_generate.size() - w->gen_head;
    const int32_t k = int32_t(w->cfg.max_batch > 0 ? std::min<size_t>(size_t(w->cfg.max_batch), pending) : pending);
    if (k > 0) {
        if (int rc = build_batch(w, k, sc)) return rc;
        if (built) *built = k;
    }
    if (w->should_full_build)
        w->should_full_build =
            !(w->gen_head == w->to_generate.size() ||
              w->chunks.size() >= size_t(w->cfg.generation_dist) * w->cfg.generation_dist * w->cfg.generation_dist);
    return UW_OK;
}

int32_t uw_world_generate_count(const uw_world* w) { return w ? int32_t(w->to_generate.size() - w->gen_head) : 0; }
int32_t uw_world_render_count(const uw_world* w) { return w ? int32_t(w->render.size()) : 0; }
int32_t uw_world_total_count(const uw_world* w) { return w ? int32_t(w->chunks.size()) : 0; }

int32_t uw_world_chunks_to_render(const uw_world* w, int32_t* out, int32_t cap) {
    if (!w) return 0;
    const int32_t n = int32_t(w->render.size());
    for (int32_t i = 0; i < std::min(n, cap); ++i)
        for (int d = 0; d < 3; ++d) out[3 * i + d] = w->render[size_t(i)][d];
    return n;
}

int uw_world_get_chunk(const uw_world* w, int32_t x, int32_t y, int32_t z, uw_world_chunk* out) {
    if (!w || !out) return UW_ERR_INVALID_ARG;
    auto it = w->chunks.find(Key{x, y, z});
    if (it == w->chunks.end()) return UW_ERR_INVALID_ARG;
    const uw_world_stored& s = it->second;
    *out = uw_world_chunk{};
    out->info = s.info;
    out->info.vert_off = 0;
    out->info.ind_off = 0;
    if (s.arena >= 0) {
        const uw_world_arena& a = w->arenas[size_t(s.arena)];
        out->verts = reinterpret_cast<const float*>(static_cast<const char*>(a.verts) +
                                                    s.info.vert_off * vert_bytes(w->cfg));
        out->inds = static_cast<const char*>(a.inds) + s.info.ind_off * ind_bytes(w->cfg);
        if (a.normals) out->tri_normals = a.normals + s.info.ind_off / 3 * 3;
        if (a.cells) out->tri_cells = a.cells + s.info.ind_off / 3;
    }
    return UW_OK;
}

int uw_world_read_chunk(const uw_world* w, int32_t x, int32_t y, int32_t z, const uw_host_out* ho, uw_chunk_info* info) {
    if (!w || !ho || !info) return UW_ERR_INVALID_ARG;
    uw_world_chunk c;
    if (int rc = uw_world_get_chunk(w, x, y, z, &c)) return rc;
    *info = c.info;
    if (ho->vertex_format != w->cfg.vertex_format || ho->index_format != w->cfg.index_format)
        return uw_fail(w->ctx, UW_ERR_INVALID_ARG, "host formats differ from the world's");
    if ((ho->verts && c.info.n_verts > ho->vert_cap) || (ho->inds && c.info.n_inds > ho->ind_cap) ||
        ((ho->tri_normals || ho->tri_cells) && c.info.n_inds > ho->ind_cap))
        return uw_fail(w->ctx, UW_ERR_OVERFLOW, "output capacity too small for this chunk");
    if (c.info.n_inds == 0) return UW_OK;
    uw_ctx* ctx = w->ctx;
    UW_HIP(ctx, hipSetDevice(ctx->device));
    const size_t ntri = c.info.n_inds / 3;
    if (ho->verts)
        UW_HIP(ctx, hipMemcpy(ho->verts, c.verts, c.info.n_verts * vert_bytes(w->cfg), hipMemcpyDeviceToHost));
    if (ho->inds) UW_HIP(ctx, hipMemcpy(ho->inds, c.inds, c.info.n_inds * ind_bytes(w->cfg), hipMemcpyDeviceToHost));
    if (ho->tri_normals && c.tri_normals)
        UW_HIP(ctx, hipMemcpy(ho->tri_normals, c.tri_normals, ntri * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (ho->tri_cells && c.tri_cells)
        UW_HIP(ctx, hipMemcpy(ho->tri_cells, c.tri_cells, ntri * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return UW_OK;
}

}  // extern "C"
