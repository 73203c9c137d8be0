/*
 * uw_oracle.c -- CPU restatement of the reference chunk builder, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE. Nothing in the engine (underwaterworld_amd/, include/) links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do,
 * and only to check or to time the reference algorithm beside the GPU path.
 *
 * It restates, operation for operation, the reference's hot path:
 *   - noise::Perlin::new(seed) -> PermutationTable::new (crate `noise` 0.8.2, built on
 *     rand 0.8.5 SliceRandom::shuffle and rand_xorshift 0.3 XorShiftRng). Neither crate is
 *     vendored in the reference nor present on this machine; the steps below restate their
 *     published algorithms (call site: underwater_world/src/state.rs:358-359).
 *   - noise::Perlin::get([f64;3]) -> core::perlin::perlin_3d (same crate; call site
 *     underwater_world/src/perlin_util.rs:13).
 *   - perlin_util::perlin_3d_octaves / iso_at (underwater_world/src/perlin_util.rs:6-29).
 *   - Chunk::build_iso + corner_to_iso_idx (underwater_world/src/chunk.rs:105-129,351-353).
 *   - Chunk::early_blank_check (chunk.rs:131-133).
 *   - Chunk::build_mesh classification, triangle-table walk, edge lerp and ordered-pair
 *     first-seen vertex dedup (chunk.rs:135-264), with the triangulation table of
 *     marching_table.rs:1-272.
 *
 * Two numeric variants:
 *   (i)  reference-faithful: f64 noise and coordinates, f32 density, f32 edge lerp, evaluated
 *        in the reference's operation order. Compile with -ffp-contract=off (Rust never
 *        contracts a*b+c), no fast-math.
 *   (ii) FP64: the same sample positions, but density, lerp and positions carried in double.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   uwo_xorshift_next  -- Marsaglia xor128 known answer (first output 3701687786 for the
 *                         published seed 123456789/362436069/521288629/88675123).
 *   uwo_gen_index      -- exhaustive uniformity of the accept set on a reduced word size
 *                         (same widening-multiply rule), and range bounds.
 *   uwo_perm           -- bijection and determinism. The table values for a given seed are
 *                         parity unpinned: the crate cannot be built or fetched here.
 *   uwo_perlin3        -- lattice zeros, lattice gradients equal the hashed gradient vector
 *                         (brute-force hash), closed form on lattice axes, 256-periodicity,
 *                         |n| <= 1, and an independent numpy restatement. The gradient-case
 *                         order for hash & 15 in 12..15, the final clamp and the
 *                         interpolation form are recalled from the crate, not verified
 *                         against it: parity to the crate itself is unpinned.
 *   uwo_density_*      -- lattice known answers (fbm == 0 at integer noise coordinates, so
 *                         rho is exactly the bias term), and the numpy restatement.
 *   uwo_mesh_*         -- all 256 single-cell sign patterns (triangle count = table row
 *                         length / 3), dyadic edge-lerp closed forms, literal linear-search
 *                         dedup vs hashed dedup, vertex count == number of distinct ordered
 *                         corner pairs, and the numpy restatement on small grids.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../underwaterworld_amd/csrc/mc_table.h"

#define UWO_MAX_OCTAVES 16

/* Parameters of the density field and the chunk geometry (mirrors the engine's uw_params
 * field-for-field so that the same values can be handed to both). */
typedef struct uwo_params {
    uint32_t seed;
    int32_t octaves;
    double frequency;   /* noise units per world unit (reference: 1/16, chunk.rs:108) */
    double lacunarity;  /* per-octave frequency multiplier (perlin_util.rs:18)        */
    double amplitudes[UWO_MAX_OCTAVES]; /* per-octave amplitude (perlin_util.rs:17)  */
    float iso;          /* isovalue (chunk.rs:10)                                      */
    int32_t bias_kind;  /* 0 = terrace (perlin_util.rs:27-28), 1 = linear             */
    float bias_a;       /* terrace: multiplier (16.0); linear: slope                   */
    float bias_b;       /* terrace: divisor (32.0);    linear: offset                  */
    float bias_c;       /* terrace: modulus (0.25)                                     */
    int32_t chunk_extent; /* world units per chunk edge (chunk.rs:5)                   */
    double offset[3];     /* world-space shift of every sample (0 = reference)         */
} uwo_params;

/* ------------------------------------------------------------------------------------- */
/* rand_xorshift 0.3: XorShiftRng::from_seed / next_u32                                    */

typedef struct { uint32_t x, y, z, w; } uwo_xorshift;

uint32_t uwo_xorshift_next(uwo_xorshift* r) {
    uint32_t x = r->x;
    uint32_t t = x ^ (x << 11);
    r->x = r->y;
    r->y = r->z;
    r->z = r->w;
    uint32_t w = r->w;
    r->w = w ^ (w >> 19) ^ (t ^ (t >> 8));
    return r->w;
}

/* Exposed for the known-answer test: state -> n outputs. */
void uwo_xorshift_stream(uint32_t x, uint32_t y, uint32_t z, uint32_t w, int n, uint32_t* out) {
    uwo_xorshift r = {x, y, z, w};
    for (int i = 0; i < n; ++i) out[i] = uwo_xorshift_next(&r);
}

/* rand 0.8.5 UniformInt<u32>::sample_single_inclusive(0, ubound-1) as used by gen_index:
 * widening multiply with the conservative rejection zone (range << lz) - 1. */
static uint32_t uwo_gen_index(uwo_xorshift* r, uint32_t ubound) {
    uint32_t range = ubound;
    uint32_t lz = (uint32_t)__builtin_clz(range);
    uint32_t zone = (range << lz) - 1u;
    for (;;) {
        uint32_t v = uwo_xorshift_next(r);
        uint64_t m = (uint64_t)v * (uint64_t)range;
        uint32_t lo = (uint32_t)m;
        if (lo <= zone) return (uint32_t)(m >> 32);
    }
}

/* Same rule on 16-bit words, for the exhaustive uniformity check in the tests:
 * returns the number of accepted v in [0, 2^16) mapping to each output in [0, range). */
void uwo_gen_index_histogram16(uint32_t range, uint32_t* hist) {
    uint32_t lz = (uint32_t)__builtin_clz(range) - 16u;
    uint32_t zone = ((range << lz) - 1u) & 0xffffu;
    for (uint32_t i = 0; i < range; ++i) hist[i] = 0;
    for (uint32_t v = 0; v < 65536u; ++v) {
        uint32_t m = v * range;
        if ((m & 0xffffu) <= zone) hist[m >> 16]++;
    }
}

/* noise 0.8.2 PermutationTable::new(seed): seed bytes [1,0,0,0, s,s,s] little-endian words
 * x = 1, y = z = w = seed, then SliceRandom::shuffle of [0, 1, ..., 255]. */
void uwo_perm(uint32_t seed, uint8_t* out) {
    uwo_xorshift r = {1u, seed, seed, seed};
    for (int i = 0; i < 256; ++i) out[i] = (uint8_t)i;
    for (uint32_t i = 255; i >= 1; --i) {
        uint32_t j = uwo_gen_index(&r, i + 1u);
        uint8_t t = out[i];
        out[i] = out[j];
        out[j] = t;
    }
}

/* ------------------------------------------------------------------------------------- */
/* noise 0.8.2 core::perlin::perlin_3d                                                     */

static inline double uwo_grad_dot(uint32_t h, double x, double y, double z) {
    switch (h & 15u) {
        case 0: return x + y;
        case 1: return -x + y;
        case 2: return x - y;
        case 3: return -x - y;
        case 4: return x + z;
        case 5: return -x + z;
        case 6: return x - z;
        case 7: return -x - z;
        case 8: return y + z;
        case 9: return -y + z;
        case 10: return y - z;
        case 11: return -y - z;
        case 12: return x + y;
        case 13: return -x + y;
        case 14: return -y + z;
        default: return -y - z;
    }
}

static inline uint32_t uwo_hash3(const uint8_t* P, int64_t x, int64_t y, int64_t z) {
    uint32_t a = P[(uint32_t)(x & 255)];
    uint32_t b = P[a ^ (uint32_t)(y & 255)];
    return P[b ^ (uint32_t)(z & 255)];
}

/* s_curve5 (interpolate.rs): t^3 (t (6t - 15) + 10) */
static inline double uwo_quintic(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

double uwo_perlin3(const uint8_t* P, double px, double py, double pz) {
    const double SCALE_FACTOR = 1.1547005383792515; /* 1/(sqrt(3)/2) */
    double fx = floor(px), fy = floor(py), fz = floor(pz);
    int64_t cx = (int64_t)fx, cy = (int64_t)fy, cz = (int64_t)fz;
    double dx = px - fx, dy = py - fy, dz = pz - fz;
#define UWO_G(ox, oy, oz) \
    uwo_grad_dot(uwo_hash3(P, cx + ox, cy + oy, cz + oz), dx - (double)ox, dy - (double)oy, dz - (double)oz)
    double g000 = UWO_G(0, 0, 0);
    double g100 = UWO_G(1, 0, 0);
    double g010 = UWO_G(0, 1, 0);
    double g110 = UWO_G(1, 1, 0);
    double g001 = UWO_G(0, 0, 1);
    double g101 = UWO_G(1, 0, 1);
    double g011 = UWO_G(0, 1, 1);
    double g111 = UWO_G(1, 1, 1);
#undef UWO_G
    double a = uwo_quintic(dx), b = uwo_quintic(dy), c = uwo_quintic(dz);
    double k0 = g000;
    double k1 = g100 - g000;
    double k2 = g010 - g000;
    double k3 = g001 - g000;
    double k4 = g000 + g110 - g100 - g010;
    double k5 = g000 + g101 - g100 - g001;
    double k6 = g000 + g011 - g010 - g001;
    double k7 = g100 + g010 + g001 + g111 - g000 - g110 - g101 - g011;
    double r = k0 + k1 * a + k2 * b + k3 * c + k4 * a * b + k5 * a * c + k6 * b * c + k7 * a * b * c;
    r = r * SCALE_FACTOR;
    if (r < -1.0) r = -1.0;
    if (r > 1.0) r = 1.0;
    return r;
}

/* Batch form for the tests: n points (xyz interleaved) -> n values. */
void uwo_perlin3_many(const uint8_t* P, const double* xyz, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = uwo_perlin3(P, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

/* perlin_util::perlin_3d_octaves (perlin_util.rs:6-22): sum_k A_k n(p f_k) / sum_k A_k with
 * f_0 = 1, f_{k+1} = f_k * lacunarity. The reference's A_k = 0.5^k. */
double uwo_fbm(const uint8_t* P, const uwo_params* prm, double x, double y, double z) {
    double total = 0.0, freq = 1.0, max_value = 0.0;
    for (int k = 0; k < prm->octaves; ++k) {
        double amp = prm->amplitudes[k];
        total += uwo_perlin3(P, x * freq, y * freq, z * freq) * amp;
        max_value += amp;
        freq *= prm->lacunarity;
    }
    return total / max_value;
}

/* perlin_util::iso_at (perlin_util.rs:24-29), reference-faithful f32 evaluation order:
 *   p = (f32) fbm;  adj = ((f32) z * A) / B;  rho = (adj + p) - fmodf(adj, C)     (terrace)
 *   p = (f32) fbm;  adj = (f32) z * slope;    rho = (adj + p) + offset            (linear)   */
float uwo_iso_at(const uint8_t* P, const uwo_params* prm, double x, double y, double z) {
    float p = (float)uwo_fbm(P, prm, x, y, z);
    float z32 = (float)z;
    if (prm->bias_kind == 0) {
        float adj = (z32 * prm->bias_a) / prm->bias_b;
        return (adj + p) - fmodf(adj, prm->bias_c);
    }
    float adj = z32 * prm->bias_a;
    return (adj + p) + prm->bias_b;
}

/* FP64 variant of iso_at: same formula carried in double. */
double uwo_iso_at_fp64(const uint8_t* P, const uwo_params* prm, double x, double y, double z) {
    double p = uwo_fbm(P, prm, x, y, z);
    if (prm->bias_kind == 0) {
        double adj = (z * (double)prm->bias_a) / (double)prm->bias_b;
        return (adj + p) - fmod(adj, (double)prm->bias_c);
    }
    return (z * (double)prm->bias_a + p) + (double)prm->bias_b;
}

/* Chunk::build_iso coordinate (chunk.rs:107-116) generalised to n cells per axis:
 *   step = (f32) extent / (f32) n;    p = ((f64) i * (f64) step + (f64)(extent * c)) * freq
 * For the reference, extent = 16, n = 12, freq = 1/16; multiplying by a power of two is
 * exact, so this equals the reference's division by CHUNK_SIZE bit for bit. */
float uwo_step(const uwo_params* prm, int32_t n) { return (float)prm->chunk_extent / (float)n; }

double uwo_coord_axis(const uwo_params* prm, int axis, int32_t c, int32_t n, int32_t i) {
    float step = uwo_step(prm, n);
    double local = (double)i * (double)step;
    return ((local + (double)(prm->chunk_extent * c)) + prm->offset[axis]) * prm->frequency;
}

double uwo_coord(const uwo_params* prm, int32_t c, int32_t n, int32_t i) { return uwo_coord_axis(prm, 0, c, n, i); }

/* Density grid of one chunk, (n+1)^3 samples, index x (n+1)^2 + y (n+1) + z. */
void uwo_density_ref(const uwo_params* prm, const int32_t* coord, int32_t n, float* out) {
    uint8_t P[256];
    uwo_perm(prm->seed, P);
    int32_t L = n + 1;
    for (int32_t x = 0; x < L; ++x) {
        double px = uwo_coord_axis(prm, 0, coord[0], n, x);
        for (int32_t y = 0; y < L; ++y) {
            double py = uwo_coord_axis(prm, 1, coord[1], n, y);
            for (int32_t z = 0; z < L; ++z) {
                double pz = uwo_coord_axis(prm, 2, coord[2], n, z);
                out[((int64_t)x * L + y) * L + z] = uwo_iso_at(P, prm, px, py, pz);
            }
        }
    }
}

void uwo_density_fp64(const uwo_params* prm, const int32_t* coord, int32_t n, double* out) {
    uint8_t P[256];
    uwo_perm(prm->seed, P);
    int32_t L = n + 1;
    for (int32_t x = 0; x < L; ++x) {
        double px = uwo_coord_axis(prm, 0, coord[0], n, x);
        for (int32_t y = 0; y < L; ++y) {
            double py = uwo_coord_axis(prm, 1, coord[1], n, y);
            for (int32_t z = 0; z < L; ++z) {
                double pz = uwo_coord_axis(prm, 2, coord[2], n, z);
                out[((int64_t)x * L + y) * L + z] = uwo_iso_at_fp64(P, prm, px, py, pz);
            }
        }
    }
}

/* Chunk::early_blank_check (chunk.rs:131-133): blank iff every sample is above iso. */
int uwo_early_blank(const float* dens, int64_t count, float iso) {
    for (int64_t i = 0; i < count; ++i)
        if (!(dens[i] > iso)) return 0;
    return 1;
}

/* ------------------------------------------------------------------------------------- */
/* Chunk::build_mesh (chunk.rs:135-264)                                                    */

typedef struct {
    uint64_t* keys; /* (a_lin << 32) | b_lin, + 1 so that 0 means empty */
    uint32_t* vals;
    uint64_t mask;
} uwo_map;

static int uwo_map_init(uwo_map* m, uint64_t min_cap) {
    uint64_t cap = 64;
    while (cap < 2 * min_cap) cap <<= 1;
    m->keys = (uint64_t*)calloc(cap, sizeof(uint64_t));
    m->vals = (uint32_t*)malloc(cap * sizeof(uint32_t));
    m->mask = cap - 1;
    return m->keys && m->vals;
}
static void uwo_map_free(uwo_map* m) { free(m->keys); free(m->vals); }
static uint32_t* uwo_map_slot(uwo_map* m, uint64_t key, int* found) {
    uint64_t k = key + 1, h = (k * 0x9E3779B97F4A7C15ull) >> 17;
    for (;;) {
        h &= m->mask;
        if (m->keys[h] == k) { *found = 1; return &m->vals[h]; }
        if (m->keys[h] == 0) { m->keys[h] = k; *found = 0; return &m->vals[h]; }
        ++h;
    }
}

typedef struct {
    int32_t n_verts, n_inds, n_active, overflow;
} uwo_mesh_counts;

/*
 * Terrain colour of a vertex (chunk.rs:215-222), f32 in the reference's operation order:
 *   world_z_ratio = world_z / CHUNK_SIZE; mix = create_mix_ratio(MIN_Z, MAX_Z, ratio)
 *   (util.rs:93-95, world.rs:11-12: -2, 2); value = BASE_VALUE + (corner_b_idx % 3) / 9;
 *   hue = MIN_HUE + (MAX_HUE - MIN_HUE) * mix (chunk.rs:14-17: -150, 60, saturation 0.6,
 *   base value 0.4); hsv_to_rgb (util.rs:122-153: rem_euclid 360, sextant by h = hue / 60,
 *   times 255) then to_srgb (util.rs:106-111: ((c / 255 + 0.055) / 1.055)^2.4).
 */
void uwo_vertex_color(float world_z, int corner_b_idx, float chunk_size, float rgb[3]) {
    const float ratio = world_z / chunk_size;
    const float mix = (ratio - (-2.0f)) / (2.0f - (-2.0f));
    const float value = 0.4f + (float)(corner_b_idx % 3) / 9.0f;
    float hue = -150.0f + (60.0f - (-150.0f)) * mix;
    float hr = fmodf(hue, 360.0f);
    if (hr < 0.0f) hr += 360.0f;
    const float c = value * 0.6f;
    const float h = hr / 60.0f;
    const float x = c * (1.0f - fabsf(fmodf(h, 2.0f) - 1.0f));
    const float m = value - c;
    float r, g, b;
    if (0.0f <= h && h < 1.0f) { r = c; g = x; b = 0.0f; }
    else if (1.0f <= h && h < 2.0f) { r = x; g = c; b = 0.0f; }
    else if (2.0f <= h && h < 3.0f) { r = 0.0f; g = c; b = x; }
    else if (3.0f <= h && h < 4.0f) { r = 0.0f; g = x; b = c; }
    else if (4.0f <= h && h < 5.0f) { r = x; g = 0.0f; b = c; }
    else { r = c; g = 0.0f; b = x; }
    const float lin[3] = {(r + m) * 255.0f, (g + m) * 255.0f, (b + m) * 255.0f};
    for (int k = 0; k < 3; ++k) rgb[k] = powf((lin[k] / 255.0f + 0.055f) / 1.055f, 2.4f);
}

/*
 * Collision-triangle normal, util::Tri::new (util.rs:12-20) with cgmath 0.18's Vector3 (crate not
 * vendored; its cross, dot-as-sum-of-products and magnitude = sqrt(dot) are recalled):
 * n = (b - a) x (c - a), then n / |n| unless |n| == 0 (safe_normalize, util.rs:60-63).
 */
void uwo_tri_normal(const float* a, const float* b, const float* c, float* out) {
    const float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const float n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const float mag = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) out[k] = mag == 0.0f ? n[k] : n[k] / mag;
}

/*
 * Chunk::tris_around (chunk.rs:315-342), literally: for x, y, z over the clamped box, append the
 * triangles of the map entry (x, y, z). The map is the per-triangle cell list tri_cells (the
 * Build.tris key of each triangle, chunk.rs:173); entries are looked up by a scan. A box whose end
 * lies below its start is empty. Returns the number of ids (writes at most cap).
 */
int64_t uwo_tris_around(const uint32_t* tri_cells, int64_t ntris, int32_t n, float px, float py, float pz,
                        int32_t range, uint32_t* ids, int64_t cap) {
    const float pct[3] = {px, py, pz};
    int64_t lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        float f = floorf(pct[k] * (float)n);
        int64_t m = isnan(f) ? 0 : f >= 2147483647.0f ? 2147483647 : f <= -2147483648.0f ? -2147483648LL : (int64_t)f;
        lo[k] = m - range < 0 ? 0 : m - range;
        hi[k] = m + range > n ? n : m + range;
    }
    int64_t cnt = 0;
    for (int64_t x = lo[0]; x <= hi[0]; ++x)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
            for (int64_t z = lo[2]; z <= hi[2]; ++z) {
                if (x >= n || y >= n || z >= n) continue; /* no map entry for cell index n */
                const uint32_t key = (uint32_t)((x * n + y) * n + z);
                for (int64_t t = 0; t < ntris; ++t)
                    if (tri_cells[t] == key) {
                        if (cnt < cap) ids[cnt] = (uint32_t)t;
                        ++cnt;
                    }
            }
    return cnt;
}

static int mesh_ref_impl(const uwo_params* prm, const int32_t* coord, int32_t n, const float* dens,
                         int dedup_mode, float* verts, float* colors, int64_t vcap, uint32_t* inds,
                         int64_t icap, uint8_t* cases, uint32_t* tri_cells, float* tri_normals,
                         uwo_mesh_counts* counts);

/*
 * Reference-faithful mesher. dedup_mode 0 = the reference's literal linear search over the
 * ordered corner pairs seen so far (chunk.rs:233-243, O(V*I)); 1 = the same first-seen
 * numbering through a hash map (identical output, used for large grids).
 * verts: 3 floats per vertex (world position, chunk.rs:224-229); inds: u32 (the reference
 * stores `ind as u16`, which equals this value whenever V <= 65535).
 * cases (optional): n^3 u8 cube-case indices in scan order. tri_cells (optional): for each
 * triangle, the linear index of the cell that emitted it (collision-triangle key, chunk.rs:173).
 * Returns 0, or -2 if a capacity was too small (counts still hold the required sizes).
 */
int uwo_mesh_ref(const uwo_params* prm, const int32_t* coord, int32_t n, const float* dens,
                 int dedup_mode, float* verts, int64_t vcap, uint32_t* inds, int64_t icap,
                 uint8_t* cases, uint32_t* tri_cells, uwo_mesh_counts* counts) {
    return mesh_ref_impl(prm, coord, n, dens, dedup_mode, verts, NULL, vcap, inds, icap, cases, tri_cells,
                         NULL, counts);
}

/* uwo_mesh_ref plus colors: 3 floats per vertex, the VertColor colour of the vertex's first-seen
 * emission (chunk.rs:215-231; the vertex pushed is the first one built for its corner pair). */
int uwo_mesh_ref_rgb(const uwo_params* prm, const int32_t* coord, int32_t n, const float* dens,
                     float* verts, float* colors, int64_t vcap, uint32_t* inds, int64_t icap,
                     uwo_mesh_counts* counts) {
    return mesh_ref_impl(prm, coord, n, dens, 1, verts, colors, vcap, inds, icap, NULL, NULL, NULL, counts);
}

/* uwo_mesh_ref plus the collision triangles of Build.tris (chunk.rs:167-175,245-250): per
 * triangle its cell key and the normal of util::Tri::new over the three emitted positions. */
int uwo_mesh_ref_tris(const uwo_params* prm, const int32_t* coord, int32_t n, const float* dens,
                      float* verts, int64_t vcap, uint32_t* inds, int64_t icap, uint32_t* tri_cells,
                      float* tri_normals, uwo_mesh_counts* counts) {
    return mesh_ref_impl(prm, coord, n, dens, 1, verts, NULL, vcap, inds, icap, NULL, tri_cells, tri_normals,
                         counts);
}

static int mesh_ref_impl(const uwo_params* prm, const int32_t* coord, int32_t n, const float* dens,
                         int dedup_mode, float* verts, float* colors, int64_t vcap, uint32_t* inds,
                         int64_t icap, uint8_t* cases, uint32_t* tri_cells, float* tri_normals,
                         uwo_mesh_counts* counts) {
    const int32_t L = n + 1;
    const float scale = uwo_step(prm, n);
    const float iso = prm->iso;
    const float off[3] = {(float)(prm->chunk_extent * coord[0]), (float)(prm->chunk_extent * coord[1]),
                          (float)(prm->chunk_extent * coord[2])};
    int64_t nv = 0, ni = 0, nact = 0;
    /* vert_pairs as linear corner indices a*S + b (chunk.rs:37,233-238) */
    uint64_t* pairs = NULL;
    int64_t pairs_cap = 0;
    uwo_map map = {0};
    if (dedup_mode == 1 && !uwo_map_init(&map, (uint64_t)5 * n * L * L + 16)) return -1;
    const uint64_t S = (uint64_t)L * L * L;

    for (int32_t x = 0; x < n; ++x)
        for (int32_t y = 0; y < n; ++y)
            for (int32_t z = 0; z < n; ++z) {
                int32_t cc[8][3];
                for (int i = 0; i < 8; ++i) {
                    uint8_t o = UW_MC_CORNER_OFS[i];
                    cc[i][0] = x + (o & 1);
                    cc[i][1] = y + ((o >> 1) & 1);
                    cc[i][2] = z + ((o >> 2) & 1);
                }
                uint32_t cs = 0;
                for (int i = 0; i < 8; ++i) {
                    float r = dens[((int64_t)cc[i][0] * L + cc[i][1]) * L + cc[i][2]];
                    if (r < iso) cs |= 1u << i;
                }
                int64_t cell = ((int64_t)x * n + y) * n + z;
                if (cases) cases[cell] = (uint8_t)cs;
                if (cs != 0 && cs != 255) ++nact;
                uint64_t row = UW_MC_TRI_PACKED[cs];
                int ntri = (int)(row & 15u);
                float cur_tri[3][3];
                for (int j = 0; j < 3 * ntri; ++j) {
                    int e = (int)((row >> (4 + 4 * j)) & 15u);
                    int ia = UW_MC_EDGE_A_IDX[e], ib = UW_MC_EDGE_B_IDX[e];
                    const int32_t* ca = cc[ia];
                    const int32_t* cb = cc[ib];
                    float sa[3], sb[3];
                    for (int k = 0; k < 3; ++k) {
                        sa[k] = (float)ca[k] * scale;
                        sb[k] = (float)cb[k] * scale;
                    }
                    float ra = dens[((int64_t)ca[0] * L + ca[1]) * L + ca[2]];
                    float rb = dens[((int64_t)cb[0] * L + cb[1]) * L + cb[2]];
                    float t = (iso - ra) / (rb - ra);
                    float m[3];
                    for (int k = 0; k < 3; ++k) {
                        float d = sb[k] - sa[k];
                        m[k] = sa[k] + t * d;
                    }
                    uint64_t a_lin = ((uint64_t)ca[0] * L + ca[1]) * L + ca[2];
                    uint64_t b_lin = ((uint64_t)cb[0] * L + cb[1]) * L + cb[2];
                    uint64_t key = a_lin * S + b_lin;
                    int64_t id = -1;
                    if (dedup_mode == 0) {
                        for (int64_t q = 0; q < nv; ++q)
                            if (pairs[q] == key) { id = q; break; }
                    } else {
                        int found;
                        uint32_t* slot = uwo_map_slot(&map, key, &found);
                        if (found) id = *slot; else *slot = (uint32_t)nv;
                    }
                    if (id < 0) {
                        id = nv;
                        if (dedup_mode == 0) {
                            if (nv == pairs_cap) {
                                pairs_cap = pairs_cap ? 2 * pairs_cap : 1024;
                                pairs = (uint64_t*)realloc(pairs, (size_t)pairs_cap * sizeof(uint64_t));
                            }
                            pairs[nv] = key;
                        }
                        if (nv < vcap && verts) {
                            verts[3 * nv + 0] = m[0] + off[0];
                            verts[3 * nv + 1] = m[1] + off[1];
                            verts[3 * nv + 2] = m[2] + off[2];
                            if (colors)
                                uwo_vertex_color(verts[3 * nv + 2], ib, (float)prm->chunk_extent,
                                                 colors + 3 * nv);
                        }
                        ++nv;
                    }
                    if (ni < icap && inds) inds[ni] = (uint32_t)id;
                    if (tri_cells && j % 3 == 0 && ni / 3 < icap / 3) tri_cells[ni / 3] = (uint32_t)cell;
                    /* current_tri holds this emission's vert.pos (chunk.rs:224-231,245-249) */
                    for (int k = 0; k < 3; ++k) cur_tri[j % 3][k] = m[k] + off[k];
                    if (tri_normals && j % 3 == 2 && ni / 3 < icap / 3)
                        uwo_tri_normal(cur_tri[0], cur_tri[1], cur_tri[2], tri_normals + 3 * (ni / 3));
                    ++ni;
                }
            }
    free(pairs);
    if (dedup_mode == 1) uwo_map_free(&map);
    counts->n_verts = (int32_t)nv;
    counts->n_inds = (int32_t)ni;
    counts->n_active = (int32_t)nact;
    counts->overflow = (nv > vcap || ni > icap);
    return counts->overflow ? -2 : 0;
}

/* FP64 mesher: same topology rule (strict < iso on the double densities), lerp and
 * positions in double, with the same sample positions (corner * (f64)(f32)step + offset). */
int uwo_mesh_fp64(const uwo_params* prm, const int32_t* coord, int32_t n, const double* dens,
                  double* verts, int64_t vcap, uint32_t* inds, int64_t icap, uint8_t* cases,
                  uwo_mesh_counts* counts) {
    const int32_t L = n + 1;
    const double scale = (double)uwo_step(prm, n);
    const double iso = (double)prm->iso;
    const double off[3] = {(double)(prm->chunk_extent * coord[0]), (double)(prm->chunk_extent * coord[1]),
                           (double)(prm->chunk_extent * coord[2])};
    int64_t nv = 0, ni = 0, nact = 0;
    uwo_map map;
    if (!uwo_map_init(&map, (uint64_t)5 * n * L * L + 16)) return -1;
    const uint64_t S = (uint64_t)L * L * L;
    for (int32_t x = 0; x < n; ++x)
        for (int32_t y = 0; y < n; ++y)
            for (int32_t z = 0; z < n; ++z) {
                int32_t cc[8][3];
                for (int i = 0; i < 8; ++i) {
                    uint8_t o = UW_MC_CORNER_OFS[i];
                    cc[i][0] = x + (o & 1);
                    cc[i][1] = y + ((o >> 1) & 1);
                    cc[i][2] = z + ((o >> 2) & 1);
                }
                uint32_t cs = 0;
                for (int i = 0; i < 8; ++i)
                    if (dens[((int64_t)cc[i][0] * L + cc[i][1]) * L + cc[i][2]] < iso) cs |= 1u << i;
                int64_t cell = ((int64_t)x * n + y) * n + z;
                if (cases) cases[cell] = (uint8_t)cs;
                if (cs != 0 && cs != 255) ++nact;
                uint64_t row = UW_MC_TRI_PACKED[cs];
                int ntri = (int)(row & 15u);
                for (int j = 0; j < 3 * ntri; ++j) {
                    int e = (int)((row >> (4 + 4 * j)) & 15u);
                    const int32_t* ca = cc[UW_MC_EDGE_A_IDX[e]];
                    const int32_t* cb = cc[UW_MC_EDGE_B_IDX[e]];
                    double ra = dens[((int64_t)ca[0] * L + ca[1]) * L + ca[2]];
                    double rb = dens[((int64_t)cb[0] * L + cb[1]) * L + cb[2]];
                    double t = (iso - ra) / (rb - ra);
                    uint64_t key = ((((uint64_t)ca[0] * L + ca[1]) * L + ca[2]) * S) +
                                   (((uint64_t)cb[0] * L + cb[1]) * L + cb[2]);
                    int found;
                    uint32_t* slot = uwo_map_slot(&map, key, &found);
                    int64_t id;
                    if (found) id = *slot;
                    else {
                        id = nv;
                        *slot = (uint32_t)nv;
                        if (nv < vcap && verts)
                            for (int k = 0; k < 3; ++k) {
                                double sa = (double)ca[k] * scale, sb = (double)cb[k] * scale;
                                verts[3 * nv + k] = sa + t * (sb - sa) + off[k];
                            }
                        ++nv;
                    }
                    if (ni < icap && inds) inds[ni] = (uint32_t)id;
                    ++ni;
                }
            }
    uwo_map_free(&map);
    counts->n_verts = (int32_t)nv;
    counts->n_inds = (int32_t)ni;
    counts->n_active = (int32_t)nact;
    counts->overflow = (nv > vcap || ni > icap);
    return counts->overflow ? -2 : 0;
}

/* ------------------------------------------------------------------------------------- */
/* Whole-chunk builds, for the CPU baseline and for bulk parity: density -> blank check ->
 * mesh, per chunk, optionally across threads (chunk-parallel). The mesh itself is not kept
 * beyond per-chunk counts and an order-independent checksum of the positions and indices. */

typedef struct {
    int32_t n_verts, n_inds, n_active, early_blank;
    uint64_t checksum;
} uwo_chunk_result;

typedef struct {
    const uwo_params* prm;
    const int32_t* coords;
    int32_t n;
    int dedup_mode;
    int64_t begin, end;
    uwo_chunk_result* res;
} uwo_job;

static uint64_t uwo_mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h;
}

static void* uwo_worker(void* arg) {
    uwo_job* jb = (uwo_job*)arg;
    int32_t n = jb->n, L = n + 1;
    int64_t S = (int64_t)L * L * L;
    int64_t vcap = 5LL * n * L * L + 16, icap = 15LL * n * n * n + 16;
    float* dens = (float*)malloc((size_t)S * sizeof(float));
    float* verts = (float*)malloc((size_t)vcap * 3 * sizeof(float));
    uint32_t* inds = (uint32_t*)malloc((size_t)icap * sizeof(uint32_t));
    for (int64_t c = jb->begin; c < jb->end; ++c) {
        const int32_t* co = jb->coords + 3 * c;
        uwo_chunk_result* r = &jb->res[c];
        uwo_density_ref(jb->prm, co, n, dens);
        memset(r, 0, sizeof(*r));
        if (uwo_early_blank(dens, S, jb->prm->iso)) { r->early_blank = 1; continue; }
        uwo_mesh_counts mc;
        uwo_mesh_ref(jb->prm, co, n, dens, jb->dedup_mode, verts, vcap, inds, icap, NULL, NULL, &mc);
        r->n_verts = mc.n_verts;
        r->n_inds = mc.n_inds;
        r->n_active = mc.n_active;
        uint64_t h = 0;
        for (int64_t i = 0; i < mc.n_inds; ++i) {
            uint32_t bits[3];
            memcpy(bits, &verts[3 * inds[i]], sizeof(bits));
            h = uwo_mix(h, ((uint64_t)i << 32) ^ bits[0] ^ ((uint64_t)bits[1] << 11) ^ ((uint64_t)bits[2] << 23));
        }
        r->checksum = h;
    }
    free(dens);
    free(verts);
    free(inds);
    return NULL;
}

int uwo_build_chunks(const uwo_params* prm, const int32_t* coords, int64_t nchunks, int32_t n,
                     int dedup_mode, int nthreads, uwo_chunk_result* res) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    uwo_job jobs[256];
    int64_t per = (nchunks + nthreads - 1) / nthreads;
    int started = 0;
    for (int t = 0; t < nthreads; ++t) {
        int64_t b = t * per, e = b + per;
        if (e > nchunks) e = nchunks;
        if (b >= e) break;
        jobs[t] = (uwo_job){prm, coords, n, dedup_mode, b, e, res};
        if (nthreads == 1) { uwo_worker(&jobs[t]); started = 0; break; }
        pthread_create(&th[t], NULL, uwo_worker, &jobs[t]);
        ++started;
    }
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* ------------------------------------------------------------------------------------- */
/* Adjoint oracle (FP64). The reference has no differentiation; this is the analytic gradient
 * the engine's adjoint kernel is checked against, plus the loss itself for finite differences.
 *
 * Loss of one chunk: L = sum over the chunk's vertices v (first-seen records, as emitted) of
 * (v_z - origin_z - target_z)^2. With the cube cases held fixed, L is smooth in the densities:
 *   v = s_a + t (s_b - s_a),  t = (iso - rho_a) / (rho_b - rho_a)
 *   dt/drho_a = (iso - rho_b) / D^2,  dt/drho_b = (rho_a - iso) / D^2,  D = rho_b - rho_a.
 * Densities depend on the octave amplitudes A_j and a world-space offset delta of every sample:
 *   rho = bias(p_z) + sum_k A_k n_k(f_k p) / sum_k A_k,
 *   drho/dA_j = (n_j - fbm) / sum A,
 *   drho/ddelta = freq (sum_k A_k f_k grad n_k / sum A + bias'(p_z) e_z)   (delta in world units),
 * with bias' = slope for the linear bias and 0 (almost everywhere) for the terrace. A clamped
 * octave (|n| = 1) has zero gradient. */

/* perlin_3d value and its gradient with respect to the point. */
double uwo_perlin3_grad(const uint8_t* P, double px, double py, double pz, double* grad) {
    const double SCALE_FACTOR = 1.1547005383792515;
    static const int8_t G[16][3] = {{1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1},
                                    {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1},
                                    {1, 1, 0},  {-1, 1, 0}, {0, -1, 1}, {0, -1, -1}};
    double fx = floor(px), fy = floor(py), fz = floor(pz);
    int64_t cx = (int64_t)fx, cy = (int64_t)fy, cz = (int64_t)fz;
    double d[3] = {px - fx, py - fy, pz - fz};
    double g[8], gg[8][3];
    for (int c = 0; c < 8; ++c) {
        int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        uint32_t h = uwo_hash3(P, cx + ox, cy + oy, cz + oz) & 15u;
        double v[3] = {d[0] - ox, d[1] - oy, d[2] - oz};
        g[c] = G[h][0] * v[0] + G[h][1] * v[1] + G[h][2] * v[2];
        for (int k = 0; k < 3; ++k) gg[c][k] = G[h][k];
    }
    double s[3], ds[3];
    for (int k = 0; k < 3; ++k) {
        double t = d[k];
        s[k] = uwo_quintic(t);
        ds[k] = 30.0 * t * t * (t - 1.0) * (t - 1.0);
    }
    /* trilinear interpolation with weights w_c = prod_k (o_k ? s_k : 1 - s_k) */
    double r = 0.0, dr[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 8; ++c) {
        int o[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
        double w[3], dw[3];
        for (int k = 0; k < 3; ++k) {
            w[k] = o[k] ? s[k] : 1.0 - s[k];
            dw[k] = o[k] ? ds[k] : -ds[k];
        }
        double W = w[0] * w[1] * w[2];
        r += W * g[c];
        dr[0] += dw[0] * w[1] * w[2] * g[c] + W * gg[c][0];
        dr[1] += w[0] * dw[1] * w[2] * g[c] + W * gg[c][1];
        dr[2] += w[0] * w[1] * dw[2] * g[c] + W * gg[c][2];
    }
    r *= SCALE_FACTOR;
    if (r < -1.0 || r > 1.0) {
        grad[0] = grad[1] = grad[2] = 0.0;
        return r < -1.0 ? -1.0 : 1.0;
    }
    for (int k = 0; k < 3; ++k) grad[k] = dr[k] * SCALE_FACTOR;
    return r;
}

/* Per-sample density derivatives: drho/dA (octaves values) and drho/ddelta (3). */
static void uwo_density_derivs(const uint8_t* P, const uwo_params* prm, double x, double y, double z, double* dA,
                               double* dD) {
    double nk[UWO_MAX_OCTAVES], total = 0.0, m = 0.0, freq = 1.0, gsum[3] = {0, 0, 0};
    for (int k = 0; k < prm->octaves; ++k) {
        double g[3];
        nk[k] = uwo_perlin3_grad(P, x * freq, y * freq, z * freq, g);
        total += nk[k] * prm->amplitudes[k];
        m += prm->amplitudes[k];
        for (int i = 0; i < 3; ++i) gsum[i] += prm->amplitudes[k] * freq * g[i];
        freq *= prm->lacunarity;
    }
    double fbm = total / m;
    for (int k = 0; k < prm->octaves; ++k) dA[k] = (nk[k] - fbm) / m;
    for (int i = 0; i < 3; ++i) dD[i] = gsum[i] / m;
    if (prm->bias_kind == 1) dD[2] += (double)prm->bias_a;
    for (int i = 0; i < 3; ++i) dD[i] *= prm->frequency; /* the offset is in world units */
}

/* FP64 plane loss of one chunk and (optionally) its gradients.
 * cases_in: NULL to classify the FP64 densities, else the frozen n^3 cube cases.
 * rho_in: NULL to evaluate the FP64 densities, else given densities (S doubles) for the lerps;
 * the amplitude/offset chain rule always evaluates the FP64 field derivatives.
 * g_dens (S doubles), g_amp (octaves), g_off (3): optional outputs.
 * Returns the loss; *nverts receives the number of vertices. */
double uwo_adjoint_chunk(const uwo_params* prm, const int32_t* coord, int32_t n, const uint8_t* cases_in,
                         const double* rho_in, double target_z, double* g_dens, double* g_amp, double* g_off,
                         int32_t* nverts) {
    uint8_t P[256];
    uwo_perm(prm->seed, P);
    const int32_t L = n + 1;
    const int64_t S = (int64_t)L * L * L;
    double* rho = (double*)malloc((size_t)S * sizeof(double));
    double* g = (double*)calloc((size_t)S, sizeof(double));
    uint8_t* seen_key = NULL;
    if (rho_in) memcpy(rho, rho_in, (size_t)S * sizeof(double));
    else uwo_density_fp64(prm, coord, n, rho);
    const double iso = (double)prm->iso;
    const double step = (double)uwo_step(prm, n);
    uwo_map map;
    uwo_map_init(&map, (uint64_t)5 * n * L * L + 16);
    double loss = 0.0;
    int64_t nv = 0;
    for (int32_t x = 0; x < n; ++x)
        for (int32_t y = 0; y < n; ++y)
            for (int32_t z = 0; z < n; ++z) {
                int32_t cc[8][3];
                for (int i = 0; i < 8; ++i) {
                    uint8_t o = UW_MC_CORNER_OFS[i];
                    cc[i][0] = x + (o & 1);
                    cc[i][1] = y + ((o >> 1) & 1);
                    cc[i][2] = z + ((o >> 2) & 1);
                }
                uint32_t cs = 0;
                int64_t cell = ((int64_t)x * n + y) * n + z;
                if (cases_in) cs = cases_in[cell];
                else
                    for (int i = 0; i < 8; ++i)
                        if (rho[((int64_t)cc[i][0] * L + cc[i][1]) * L + cc[i][2]] < iso) cs |= 1u << i;
                uint64_t row = UW_MC_TRI_PACKED[cs];
                int ntri = (int)(row & 15u);
                for (int j = 0; j < 3 * ntri; ++j) {
                    int e = (int)((row >> (4 + 4 * j)) & 15u);
                    const int32_t* ca = cc[UW_MC_EDGE_A_IDX[e]];
                    const int32_t* cb = cc[UW_MC_EDGE_B_IDX[e]];
                    int64_t ia = ((int64_t)ca[0] * L + ca[1]) * L + ca[2];
                    int64_t ib = ((int64_t)cb[0] * L + cb[1]) * L + cb[2];
                    int found;
                    uint32_t* slot = uwo_map_slot(&map, (uint64_t)ia * (uint64_t)S + (uint64_t)ib, &found);
                    if (found) continue;
                    *slot = (uint32_t)nv++;
                    double ra = rho[ia], rb = rho[ib], D = rb - ra;
                    double t = (iso - ra) / D;
                    double za = ca[2] * step, zb = cb[2] * step;
                    double r = (za + t * (zb - za)) - target_z;
                    loss += r * r;
                    double dz = 2.0 * r * (zb - za);
                    g[ia] += dz * (iso - rb) / (D * D);
                    g[ib] += dz * (ra - iso) / (D * D);
                }
            }
    uwo_map_free(&map);
    free(seen_key);
    if (g_dens) memcpy(g_dens, g, (size_t)S * sizeof(double));
    if (g_amp || g_off) {
        double ga[UWO_MAX_OCTAVES] = {0}, go[3] = {0, 0, 0};
        double dA[UWO_MAX_OCTAVES], dD[3];
        for (int32_t x = 0; x < L; ++x)
            for (int32_t y = 0; y < L; ++y)
                for (int32_t z = 0; z < L; ++z) {
                    int64_t s = ((int64_t)x * L + y) * L + z;
                    if (g[s] == 0.0) continue;
                    uwo_density_derivs(P, prm, uwo_coord_axis(prm, 0, coord[0], n, x), uwo_coord_axis(prm, 1, coord[1], n, y),
                                       uwo_coord_axis(prm, 2, coord[2], n, z), dA, dD);
                    for (int k = 0; k < prm->octaves; ++k) ga[k] += g[s] * dA[k];
                    for (int i = 0; i < 3; ++i) go[i] += g[s] * dD[i];
                }
        if (g_amp) for (int k = 0; k < prm->octaves; ++k) g_amp[k] = ga[k];
        if (g_off) for (int i = 0; i < 3; ++i) g_off[i] = go[i];
    }
    free(rho);
    free(g);
    if (nverts) *nverts = (int32_t)nv;
    return loss;
}

/* Plane loss and its density gradient from given FP64 densities and frozen cube cases (the mesh
 * topology). Used for finite differences on single density samples. */
double uwo_plane_loss_dens(int32_t n, double step, double iso, const double* rho, const uint8_t* cases,
                           double target_z, double* g_dens) {
    const int32_t L = n + 1;
    const int64_t S = (int64_t)L * L * L;
    if (g_dens) memset(g_dens, 0, (size_t)S * sizeof(double));
    uwo_map map;
    uwo_map_init(&map, (uint64_t)5 * n * L * L + 16);
    double loss = 0.0;
    int64_t nv = 0;
    for (int32_t x = 0; x < n; ++x)
        for (int32_t y = 0; y < n; ++y)
            for (int32_t z = 0; z < n; ++z) {
                uint32_t cs = cases[((int64_t)x * n + y) * n + z];
                uint64_t row = UW_MC_TRI_PACKED[cs];
                int ntri = (int)(row & 15u);
                for (int j = 0; j < 3 * ntri; ++j) {
                    int e = (int)((row >> (4 + 4 * j)) & 15u);
                    uint8_t oa = UW_MC_CORNER_OFS[UW_MC_EDGE_A_IDX[e]], ob = UW_MC_CORNER_OFS[UW_MC_EDGE_B_IDX[e]];
                    int64_t ia = ((int64_t)(x + (oa & 1)) * L + (y + ((oa >> 1) & 1))) * L + (z + ((oa >> 2) & 1));
                    int64_t ib = ((int64_t)(x + (ob & 1)) * L + (y + ((ob >> 1) & 1))) * L + (z + ((ob >> 2) & 1));
                    int found;
                    uint32_t* slot = uwo_map_slot(&map, (uint64_t)ia * (uint64_t)S + (uint64_t)ib, &found);
                    if (found) continue;
                    *slot = (uint32_t)nv++;
                    double ra = rho[ia], rb = rho[ib], D = rb - ra;
                    double t = (iso - ra) / D;
                    double za = (z + ((oa >> 2) & 1)) * step, zb = (z + ((ob >> 2) & 1)) * step;
                    double r = (za + t * (zb - za)) - target_z;
                    loss += r * r;
                    if (g_dens) {
                        double dz = 2.0 * r * (zb - za);
                        g_dens[ia] += dz * (iso - rb) / (D * D);
                        g_dens[ib] += dz * (ra - iso) / (D * D);
                    }
                }
            }
    uwo_map_free(&map);
    return loss;
}
