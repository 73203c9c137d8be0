// Context management, parameter validation, permutation table and lookup-table derivation.
#include <math.h>
#include <string.h>

#include <new>

#include "uw_internal.h"

int uw_fail(uw_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int uw_hip_fail(uw_ctx* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return uw_fail(ctx, e == hipErrorOutOfMemory ? UW_ERR_OOM : UW_ERR_HIP, m);
}

int uw_refuse_capture(uw_ctx* ctx, hipStream_t st) {
    if (!st) return UW_OK;  // the null stream cannot be captured
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    UW_HIP(ctx, hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return uw_fail(ctx, UW_ERR_INVALID_ARG,
                       "stream is capturing: this entry point uses context-owned scratch that a later call may "
                       "reallocate; capture the batched build with uw_batched_graph_create");
    return UW_OK;
}

int uw_prof_start(uw_ctx* ctx, hipStream_t st, int stage) {
    if (!ctx->profiling) return UW_OK;
    if (ctx->prof_used + 2 > ctx->prof_events.size()) {
        for (int i = 0; i < 2; ++i) {
            hipEvent_t e;
            UW_HIP(ctx, hipEventCreate(&e));
            ctx->prof_events.push_back(e);
        }
    }
    if (ctx->prof_stage.size() < ctx->prof_events.size() / 2) ctx->prof_stage.resize(ctx->prof_events.size() / 2);
    ctx->prof_stage[ctx->prof_used / 2] = stage;
    UW_HIP(ctx, hipEventRecord(ctx->prof_events[ctx->prof_used], st));
    return UW_OK;
}

int uw_prof_end(uw_ctx* ctx, hipStream_t st) {
    if (!ctx->profiling) return UW_OK;
    UW_HIP(ctx, hipEventRecord(ctx->prof_events[ctx->prof_used + 1], st));
    ctx->prof_used += 2;
    return UW_OK;
}

namespace {

// Marsaglia xor128 step as implemented by rand_xorshift 0.3 XorShiftRng::next_u32.
struct XorShift128 {
    uint32_t x, y, z, w;
    uint32_t next() {
        uint32_t t = x ^ (x << 11);
        x = y;
        y = z;
        z = w;
        w = w ^ (w >> 19) ^ (t ^ (t >> 8));
        return w;
    }
    // rand 0.8 gen_range(0..bound) for u32: Lemire widening multiply, rejecting low words above
    // the zone (bound << leading_zeros(bound)) - 1.
    uint32_t below(uint32_t bound) {
        const uint32_t zone = (bound << __builtin_clz(bound)) - 1u;
        for (;;) {
            const uint64_t m = uint64_t(next()) * bound;
            if (uint32_t(m) <= zone) return uint32_t(m >> 32);
        }
    }
};

// noise 0.8.2 PermutationTable::new(seed): the 16 seed bytes are [1,0,0,0] followed by three
// little-endian copies of the seed, i.e. state (1, seed, seed, seed); the identity permutation
// of 0..=255 is then Fisher-Yates shuffled from the top (rand 0.8 SliceRandom::shuffle).
void build_permutation(uint32_t seed, uint8_t out[256]) {
    XorShift128 rng{1u, seed, seed, seed};
    for (int i = 0; i < 256; ++i) out[i] = uint8_t(i);
    for (uint32_t i = 255; i > 0; --i) {
        const uint32_t j = rng.below(i + 1);
        const uint8_t t = out[i];
        out[i] = out[j];
        out[j] = t;
    }
}

// Gradient vectors of noise 0.8.2 perlin_3d for hash & 15; codes 12..15 repeat four of the
// twelve edge directions.

void build_tables(uw_tables& t, const uint8_t perm[256]) {
    memset(&t, 0, sizeof(t));
    for (int c = 0; c < 256; ++c) {
        const uint64_t row = UW_MC_TRI_PACKED[c];
        t.tri[c] = row;
        const int n = 3 * int(row & 15u);
        uint16_t seen = 0;
        uint16_t before[12];
        for (int e = 0; e < 12; ++e) before[e] = 0xffff;
        for (int j = 0; j < n; ++j) {
            const int e = int((row >> (4 + 4 * j)) & 15u);
            if (!(seen & (1u << e))) {
                before[e] = seen;
                seen = uint16_t(seen | (1u << e));
            }
        }
        for (int e = 0; e < 12; ++e) t.before[c][e] = before[e] == 0xffff ? seen : before[e];
        t.used[c] = seen;
    }
    memcpy(t.perm, perm, 256);
    for (int h = 0; h < 16; ++h)
        t.grad[h] = make_float4(float(kUwGrad[h][0]), float(kUwGrad[h][1]), float(kUwGrad[h][2]), 0.0f);
}

bool params_valid(const uw_params* p, std::string* why) {
    if (p->octaves < 1 || p->octaves > UW_MAX_OCTAVES) { *why = "octaves must be in [1, 16]"; return false; }
    if (!(p->frequency > 0.0) || !isfinite(p->frequency)) { *why = "frequency must be positive"; return false; }
    if (!(p->lacunarity > 0.0) || !isfinite(p->lacunarity)) { *why = "lacunarity must be positive"; return false; }
    double s = 0.0;
    for (int k = 0; k < p->octaves; ++k) {
        if (!isfinite(p->amplitudes[k])) { *why = "amplitudes must be finite"; return false; }
        s += p->amplitudes[k];
    }
    if (s == 0.0) { *why = "amplitudes must not sum to zero"; return false; }
    if (p->bias_kind != UW_BIAS_TERRACE && p->bias_kind != UW_BIAS_LINEAR) { *why = "unknown bias_kind"; return false; }
    if (p->bias_kind == UW_BIAS_TERRACE && (p->bias_b == 0.0f || p->bias_c == 0.0f)) {
        *why = "terrace bias needs nonzero divisor and modulus";
        return false;
    }
    if (p->chunk_extent < 1 || p->chunk_extent > (1 << 20)) { *why = "chunk_extent out of range"; return false; }
    if (!isfinite(p->iso)) { *why = "iso must be finite"; return false; }
    for (int i = 0; i < 3; ++i)
        if (!isfinite(p->offset[i])) { *why = "offset must be finite"; return false; }
    return true;
}

}  // namespace

uw_field uw_make_field(const uw_params& p) {
    uw_field f;
    memset(&f, 0, sizeof(f));
    f.frequency = p.frequency;
    for (int i = 0; i < 3; ++i) f.offset[i] = p.offset[i];
    double fk = 1.0, sum = 0.0;
    for (int k = 0; k < p.octaves; ++k) {
        f.octave_freq[k] = fk;
        fk *= p.lacunarity;
        f.amp[k] = float(p.amplitudes[k]);
        sum += p.amplitudes[k];
    }
    f.amp_sum = float(sum);
    // 1/(sqrt(3)/2) and the normalisation folded into one weight per octave (f64, rounded once)
    const double scale = 1.1547005383792515;
    for (int k = 0; k < p.octaves; ++k) f.weight[k] = float(p.amplitudes[k] * scale / sum);
    f.clamp_hi = float(1.0 / scale);
    f.clamp_lo = -f.clamp_hi;
    f.octaves = p.octaves;
    f.extent = p.chunk_extent;
    f.bias_kind = p.bias_kind;
    f.bias_a = p.bias_a;
    f.bias_b = p.bias_b;
    f.bias_c = p.bias_c;
    f.iso = p.iso;
    return f;
}

int uw_check_resolution(uw_ctx* ctx, int32_t n) {
    if (n < 1 || n > UW_MAX_RESOLUTION) return uw_fail(ctx, UW_ERR_INVALID_ARG, "resolution must be in [1, 64]");
    return UW_OK;
}

extern "C" {

void uw_params_reference(uw_params* p, uint32_t seed) {
    memset(p, 0, sizeof(*p));
    p->seed = seed;
    p->octaves = 3;                 // chunk.rs:9
    p->frequency = 1.0 / 16.0;      // chunk.rs:108 (/ CHUNK_SIZE)
    p->lacunarity = 2.0;            // perlin_util.rs:18
    p->amplitudes[0] = 1.0;         // perlin_util.rs:10,17
    p->amplitudes[1] = 0.5;
    p->amplitudes[2] = 0.25;
    p->iso = -0.1f;                 // chunk.rs:10
    p->bias_kind = UW_BIAS_TERRACE; // perlin_util.rs:27-28
    p->bias_a = 16.0f;              // CHUNK_SIZE as f32
    p->bias_b = 32.0f;              // MAX_HEIGHT
    p->bias_c = 0.25f;              // ADJ_Z_MOD
    p->chunk_extent = 16;           // chunk.rs:5
}

int uw_ctx_create(const uw_params* p, int device, uw_ctx** out) {
    if (!p || !out) return UW_ERR_INVALID_ARG;
    *out = nullptr;
    std::string why;
    if (!params_valid(p, &why)) return UW_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UW_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return UW_ERR_INVALID_ARG;
    uw_ctx* ctx = new (std::nothrow) uw_ctx();
    if (!ctx) return UW_ERR_OOM;
    ctx->params = *p;
    ctx->device = device;
    uw_default_batch_tuning(&ctx->batch_tuning);
    build_permutation(p->seed, ctx->perm);
    build_tables(ctx->host_tables, ctx->perm);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cus = prop.multiProcessorCount;
    if (hipSetDevice(device) != hipSuccess ||
        hipMalloc(&ctx->d_tables, sizeof(uw_tables)) != hipSuccess ||
        hipMemcpy(ctx->d_tables, &ctx->host_tables, sizeof(uw_tables), hipMemcpyHostToDevice) != hipSuccess ||
        uw_noalloc_setup(ctx) != UW_OK) {
        uw_ctx_destroy(ctx);
        return UW_ERR_HIP;
    }
    *out = ctx;
    return UW_OK;
}

void uw_ctx_destroy(uw_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipFree(ctx->d_tables);
    uw_plain_scratch& s = ctx->plain;
    (void)hipFree(s.dens);
    (void)hipFree(s.below);
    (void)hipFree(s.cases);
    (void)hipFree(s.voff);
    (void)hipFree(s.act);
    (void)hipFree(s.ioff);
    (void)hipFree(s.ztab);
    (void)hipFree(ctx->d_coord1);
    (void)hipFree(ctx->d_verts1);
    (void)hipFree(ctx->d_inds1);
    (void)hipFree(ctx->d_info1);
    (void)hipFree(ctx->d_totals1);
    (void)hipFree(ctx->d_norm1);
    (void)hipFree(ctx->d_cells1);
    (void)hipFree(ctx->d_batch_head);
    (void)hipFree(ctx->d_batch_order);
    (void)hipFree(ctx->d_batch_ztab);
    (void)hipFree(ctx->d_batch_dscr);
    (void)hipFree(ctx->d_batch_bulk);
    (void)hipFree(ctx->d_vol);
    (void)hipFree(ctx->d_adj_part);
    (void)hipFree(ctx->d_adj_plan);
    (void)hipFree(ctx->d_adj_tabs);
    for (hipEvent_t e : ctx->prof_events) (void)hipEventDestroy(e);
    delete ctx;
}

const char* uw_last_error(const uw_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int uw_ctx_permutation(const uw_ctx* ctx, uint8_t* out256) {
    if (!ctx || !out256) return UW_ERR_INVALID_ARG;
    memcpy(out256, ctx->perm, 256);
    return UW_OK;
}

void uw_permutation_table(uint32_t seed, uint8_t* out256) {
    if (out256) build_permutation(seed, out256);
}

int uw_profile_enable(uw_ctx* ctx, int on) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    ctx->profiling = on != 0;
    return UW_OK;
}

int uw_profile_read_stages(uw_ctx* ctx, int32_t nstages, double* total_ms, uint64_t* launches) {
    if (!ctx || nstages < 1 || !total_ms || !launches) return UW_ERR_INVALID_ARG;
    for (int32_t k = 0; k < nstages; ++k) {
        total_ms[k] = 0.0;
        launches[k] = 0;
    }
    for (size_t i = 0; i + 1 < ctx->prof_used; i += 2) {
        UW_HIP(ctx, hipEventSynchronize(ctx->prof_events[i + 1]));
        const int k = ctx->prof_stage[i / 2];
        if (k < 0 || k >= nstages) continue;
        float ms = 0.0f;
        UW_HIP(ctx, hipEventElapsedTime(&ms, ctx->prof_events[i], ctx->prof_events[i + 1]));
        total_ms[k] += ms;
        launches[k] += 1;
    }
    ctx->prof_used = 0;
    return UW_OK;
}

int uw_profile_read(uw_ctx* ctx, double* total_ms, uint64_t* launches) {
    return uw_profile_read_stages(ctx, 1, total_ms, launches);
}

uint64_t uw_max_verts(int32_t n) { return n < 1 ? 0 : 5ull * uint64_t(n) * uint64_t(n + 1) * uint64_t(n + 1); }
uint64_t uw_max_inds(int32_t n) { return n < 1 ? 0 : 15ull * uint64_t(n) * uint64_t(n) * uint64_t(n); }

}  // extern "C"
