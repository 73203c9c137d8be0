// Device-side density field: sample coordinates, Perlin octaves and the z-bias.
//
// Restates, for FP32 VALU execution, the reference's density:
//   Chunk::build_iso coordinates           underwater_world/src/chunk.rs:107-116
//   perlin_util::perlin_3d_octaves          underwater_world/src/perlin_util.rs:6-22
//   perlin_util::iso_at (bias)              underwater_world/src/perlin_util.rs:24-29
//   noise 0.8.2 core::perlin::perlin_3d     (external crate, call site perlin_util.rs:13)
//
// Numerics. Sample coordinates are formed in f64 exactly as the reference does, and per axis the
// lattice index, the fraction d, d - 1 and the quintic fade s(d) are derived in f64 and rounded
// once to f32. These per-axis quantities are shared by every sample of a grid line, so they are
// computed once per (axis index, octave) into LDS. Only the per-sample combination runs in f32:
// eight gradient dots (one rounding each: the gradients are {-1,0,1} vectors with two nonzero
// components) and seven FMA lerps, which equal the crate's 8-term multilinear polynomial
// k0 + k1 a + ... + k7 a b c in exact arithmetic.
#pragma once

#include "uw_internal.h"

// One axis of one octave: hashed/low lattice bytes and the f32-rounded fraction data.
struct __align__(16) uw_axis {
    uint32_t bits;  // x axis: P[L & 255] | P[(L+1) & 255] << 8; y, z axes: (L & 255) | ((L+1) & 255) << 8
    float d;        // p - floor(p)
    float dm1;      // d - 1
    float fade;     // d^3 (d (6 d - 15) + 10)
};

// 1/(sqrt(3)/2), noise 0.8.2 perlin_3d SCALE_FACTOR, rounded to f32.
#define UW_PERLIN_SCALE 1.1547005383792515f

// Noise coordinate of sample i of chunk c along `axis` (chunk.rs:107-108): all f64; the product
// by a power-of-two frequency equals the reference's division by CHUNK_SIZE, and the world-space
// offset (0 for the reference, where adding it is exact) shifts every sample.
__device__ __forceinline__ double uw_sample_coord(const uw_field& f, int axis, int32_t c, int32_t i, float step) {
    const double local = double(i) * double(step);
    return ((local + double(f.extent * c)) + f.offset[axis]) * f.frequency;
}

__device__ __forceinline__ double uw_quintic64(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

__device__ __forceinline__ uw_axis uw_axis_of(double p, double fk) {
    const double pk = p * fk;
    const double fl = floor(pk);
    const int64_t L = int64_t(fl);
    const double d = pk - fl;
    uw_axis a;
    a.bits = uint32_t(L & 255) | (uint32_t((L + 1) & 255) << 8);
    a.d = float(d);
    a.dm1 = float(d - 1.0);
    a.fade = float(uw_quintic64(d));
    return a;
}

// Fold the first hash level into an x-axis entry.
__device__ __forceinline__ uw_axis uw_hash_x(uw_axis a, const uint8_t* perm) {
    a.bits = uint32_t(perm[a.bits & 255]) | (uint32_t(perm[(a.bits >> 8) & 255]) << 8);
    return a;
}

__device__ __forceinline__ float uw_gdot(float4 g, float x, float y, float z) {
    return fmaf(g.z, z, fmaf(g.y, y, g.x * x));
}

__device__ __forceinline__ float uw_lerp(float t, float a, float b) { return fmaf(t, b - a, a); }

// Two sources of the gradient of permutation slot j (the vector of code P[j] & 15), giving the same
// float4 (components exactly -1, 0 or 1):
//   const float* G3: an LDS table of three floats per slot (12-byte stride, so the slots of a
//     gather spread over all LDS banks): three LDS reads per gather, no arithmetic;
//   uw_gcode: an LDS byte per slot, GC[j] = 2 (P[j] & 15), and the components of the 16 gradients
//     packed two bits each (two's complement) into constants: one LDS read and six VALU
//     operations per gather.
// Which is faster depends on how often a line rebuilds its lattice faces: the byte codes win when
// the octaves are fine against the sample step (faces rebuilt at most samples, so LDS traffic and
// bank conflicts bound the loop), the float table when the loop is bound by its VALU work.
__device__ __forceinline__ float4 uw_grad3(const float* G3, uint32_t j) {
    const float* g = G3 + 3u * j;
    return make_float4(g[0], g[1], g[2], 0.0f);
}

struct uw_gcode {
    const uint8_t* c;
};
__device__ __forceinline__ float4 uw_grad3(uw_gcode g, uint32_t j) {
    constexpr uint32_t kx = uw_grad_pack(0), ky = uw_grad_pack(1), kz = uw_grad_pack(2);
    const uint32_t sh = g.c[j];
    const int gx = __builtin_amdgcn_sbfe(int(kx), sh, 2u);
    const int gy = __builtin_amdgcn_sbfe(int(ky), sh, 2u);
    const int gz = __builtin_amdgcn_sbfe(int(kz), sh, 2u);
    return make_float4(float(gx), float(gy), float(gz), 0.0f);
}
// Gradient vector of hash code h (0..15) from the packed constants: no table load, so filling a
// per-slot LDS table costs one memory round trip (the permutation) instead of two.
__device__ __forceinline__ float4 uw_grad_vec(uint32_t h) {
    constexpr uint32_t kx = uw_grad_pack(0), ky = uw_grad_pack(1), kz = uw_grad_pack(2);
    const uint32_t sh = 2u * h;
    const int gx = __builtin_amdgcn_sbfe(int(kx), sh, 2u);
    const int gy = __builtin_amdgcn_sbfe(int(ky), sh, 2u);
    const int gz = __builtin_amdgcn_sbfe(int(kz), sh, 2u);
    return make_float4(float(gx), float(gy), float(gz), 0.0f);
}

// Code table entry of permutation slot j for uw_gcode (perm_j = P[j]).
__host__ __device__ inline uint8_t uw_gcode_of(uint8_t perm_j) { return uint8_t((perm_j & 15u) << 1); }

// One octave of improved Perlin noise at the lattice cell / fractions given by X (hashed), Y, Z.
// P: permutation table; G3: gradient source (uw_grad3).
template <class GT>
__device__ __forceinline__ float uw_perlin_octave(const uw_axis& X, const uw_axis& Y, const uw_axis& Z,
                                                  const uint8_t* P, GT G3) {
    const uint32_t a0 = X.bits & 255, a1 = X.bits >> 8;
    const uint32_t y0 = Y.bits & 255, y1 = Y.bits >> 8;
    const uint32_t z0 = Z.bits & 255, z1 = Z.bits >> 8;
    const uint32_t b00 = P[a0 ^ y0], b01 = P[a0 ^ y1], b10 = P[a1 ^ y0], b11 = P[a1 ^ y1];
    const float g000 = uw_gdot(uw_grad3(G3, b00 ^ z0), X.d, Y.d, Z.d);
    const float g100 = uw_gdot(uw_grad3(G3, b10 ^ z0), X.dm1, Y.d, Z.d);
    const float g010 = uw_gdot(uw_grad3(G3, b01 ^ z0), X.d, Y.dm1, Z.d);
    const float g110 = uw_gdot(uw_grad3(G3, b11 ^ z0), X.dm1, Y.dm1, Z.d);
    const float g001 = uw_gdot(uw_grad3(G3, b00 ^ z1), X.d, Y.d, Z.dm1);
    const float g101 = uw_gdot(uw_grad3(G3, b10 ^ z1), X.dm1, Y.d, Z.dm1);
    const float g011 = uw_gdot(uw_grad3(G3, b01 ^ z1), X.d, Y.dm1, Z.dm1);
    const float g111 = uw_gdot(uw_grad3(G3, b11 ^ z1), X.dm1, Y.dm1, Z.dm1);
    const float x00 = uw_lerp(X.fade, g000, g100);
    const float x10 = uw_lerp(X.fade, g010, g110);
    const float x01 = uw_lerp(X.fade, g001, g101);
    const float x11 = uw_lerp(X.fade, g011, g111);
    const float y0v = uw_lerp(Y.fade, x00, x10);
    const float y1v = uw_lerp(Y.fade, x01, x11);
    const float r = uw_lerp(Z.fade, y0v, y1v) * UW_PERLIN_SCALE;
    return fminf(fmaxf(r, -1.0f), 1.0f);
}

// Bias split: rho = (adj + p) + post.  Terrace (perlin_util.rs:27-28): adj = (z * a) / b,
// post = -fmod(adj, c) (u - m == u + (-m) exactly). Linear: adj = z * a, post = b.
__device__ __forceinline__ void uw_bias_terms(const uw_field& f, double pz, float* adj, float* post) {
    const float z32 = float(pz);
    if (f.bias_kind == UW_BIAS_TERRACE) {
        const float a = (z32 * f.bias_a) / f.bias_b;
        *adj = a;
        *post = -fmodf(a, f.bias_c);
    } else {
        *adj = z32 * f.bias_a;
        *post = f.bias_b;
    }
}

__device__ __forceinline__ float uw_density_from(float total, float amp_sum, float adj, float post) {
    const float p = total / amp_sum;
    return (adj + p) + post;
}

// ---------------------------------------------------------------------------------------------
// Line evaluation. Along a grid line of fixed (x, y) only z varies, and within one lattice cell of
// octave k the eight corner gradients are fixed, so each corner term is affine in dz:
//   g_ijl(dz) = Gx dx_i + Gy dy_j + Gz (dz - l) = P_ijl + Gz_ijl dz_l.
// The x and y lerps are linear, so the two z faces of the cell reduce to
//   N_l(dz) = A_l + B_l dz_l,  A_l = lerp_y(lerp_x(P_.. l)),  B_l = lerp_y(lerp_x(Gz_.. l)),
// and a sample is N0 + s(dz) (N1 - N0) with N1 - N0 = C + D dz: three FMAs, the clamp and the
// octave accumulation. The A/B setup (gradient fetches and ~20 FLOPs per face) is redone only when
// the line enters a new lattice cell of that octave, which is wave-uniform because every lane of a
// wave walks its own line at the same z index.

struct uw_line_oct {
    float A0, B0, A1, B1;
    float C, D;  // N1 - N0 = C + D dz with C = (A1 - B1) - A0, D = B1 - B0
};

__device__ __forceinline__ void uw_line_diff(uw_line_oct& s) {
    s.C = (s.A1 - s.B1) - s.A0;
    s.D = s.B1 - s.B0;
}

// One z face of a line's lattice cell: A = lerp_y(lerp_x(P_..)), B = lerp_y(lerp_x(Gz_..)) for the
// face whose z lattice byte is zb. bb = b00 | b01 << 8 | b10 << 16 | b11 << 24 (second-level
// hashes of the line's x/y corners). The face depends on the z lattice index only through zb.
template <class GT>
__device__ __forceinline__ void uw_line_face(const uw_axis& X, const uw_axis& Y, uint32_t bb, uint32_t zb, GT G3,
                                             float* A, float* B) {
    const uint32_t b00 = bb & 255u, b01 = (bb >> 8) & 255u, b10 = (bb >> 16) & 255u, b11 = bb >> 24;
    const float4 G00 = uw_grad3(G3, b00 ^ zb), G10 = uw_grad3(G3, b10 ^ zb), G01 = uw_grad3(G3, b01 ^ zb),
                 G11 = uw_grad3(G3, b11 ^ zb);
    const float P00 = fmaf(G00.y, Y.d, G00.x * X.d), P10 = fmaf(G10.y, Y.d, G10.x * X.dm1);
    const float P01 = fmaf(G01.y, Y.dm1, G01.x * X.d), P11 = fmaf(G11.y, Y.dm1, G11.x * X.dm1);
    const float a = X.fade, b = Y.fade;
    *A = uw_lerp(b, uw_lerp(a, P00, P10), uw_lerp(a, P01, P11));
    *B = uw_lerp(b, uw_lerp(a, G00.z, G10.z), uw_lerp(a, G01.z, G11.z));
}

// Per-line, per-octave state for the lattice cell whose low/high z bytes are packed in zbits.
template <class GT>
__device__ __forceinline__ uw_line_oct uw_line_setup(const uw_axis& X, const uw_axis& Y, uint32_t bb, uint32_t zbits,
                                                     GT G3) {
    uw_line_oct s;
    uw_line_face(X, Y, bb, zbits & 255u, G3, &s.A0, &s.B0);
    uw_line_face(X, Y, bb, (zbits >> 8) & 255u, G3, &s.A1, &s.B1);
    uw_line_diff(s);
    return s;
}

// Unscaled octave value at the line sample, clamped to the unscaled image of [-1, 1]; the caller
// multiplies by the octave weight A_k (2/sqrt 3) / sum A.
// N0 + s(dz) (N1 - N0), with the face difference N1 - N0 = C + D dz formed by one FMA.
__device__ __forceinline__ float uw_line_sample(const uw_line_oct& s, float dz, float fade, float lo, float hi) {
    const float n0 = fmaf(s.B0, dz, s.A0);
    const float dn = fmaf(s.D, dz, s.C);
    return __builtin_amdgcn_fmed3f(fmaf(fade, dn, n0), lo, hi);
}

// Per-chunk LDS tables: every (axis index, octave) entry of the chunk and the z-bias split.
struct uw_chunk_tables {
    uw_axis* AX;   // [O][L], hashed
    uw_axis* AY;   // [O][L]
    uw_axis* AZ;   // [O][L]
    float* ADJ;    // [L]
    float* POST;   // [L]
};

__device__ __forceinline__ void uw_build_chunk_tables(const uw_field& f, const uint8_t* perm, int32_t cx, int32_t cy,
                                                      int32_t cz, int32_t L, float step, const uw_chunk_tables& T,
                                                      int tid, int nthr) {
    const int32_t O = f.octaves;
    const int32_t nent = 3 * O * L;
    for (int e = tid; e < nent; e += nthr) {
        const int32_t axis = e / (O * L);
        const int32_t r = e - axis * O * L;
        const int32_t k = r / L, i = r - k * L;
        const double fk = f.octave_freq[k];
        if (axis == 0) T.AX[r] = uw_hash_x(uw_axis_of(uw_sample_coord(f, 0, cx, i, step), fk), perm);
        else if (axis == 1) T.AY[r] = uw_axis_of(uw_sample_coord(f, 1, cy, i, step), fk);
        else T.AZ[r] = uw_axis_of(uw_sample_coord(f, 2, cz, i, step), fk);
    }
    for (int i = tid; i < L; i += nthr) uw_bias_terms(f, uw_sample_coord(f, 2, cz, i, step), &T.ADJ[i], &T.POST[i]);
}

// The x and y tables only (the density loops read the z entries from the chunk's z table in HBM,
// see uw_ztab): AX[k L + i] hashed, AY[k L + i].
// One thread per (axis, octave, sample) entry, so the short f64 chains of a chunk's entries run
// side by side; freq holds the octave frequencies (an LDS copy, indexed per lane).
__device__ __forceinline__ void uw_build_xy_tables(const uw_field& f, const double* freq, const uint8_t* perm,
                                                   int32_t cx, int32_t cy, int32_t L, float step, uw_axis* AX,
                                                   uw_axis* AY, int tid, int nthr) {
    const int32_t OL = f.octaves * L;
    const uw_divider dL = uw_make_divider(uint32_t(L));
    for (int e = tid; e < 2 * OL; e += nthr) {
        const bool is_y = e >= OL;
        const int32_t r = is_y ? e - OL : e;
        const int32_t k = int32_t(uw_div(uint32_t(r), dL)), i = r - k * L;
        const uw_axis v = uw_axis_of(uw_sample_coord(f, is_y ? 1 : 0, is_y ? cy : cx, i, step), freq[k]);
        if (is_y) AY[r] = v;
        else AX[r] = uw_hash_x(v, perm);
    }
}

// ---------------------------------------------------------------------------------------------
// Per-chunk z tables in HBM. Every lane of a wave walks its line at the same z index, so the z
// entries and the bias split of a sample are wave-uniform: kept in HBM (written by a previous
// kernel, uw_launch_ztables) they are fetched with scalar loads into scalar registers, costing
// neither vector instructions nor vector registers. One record of uw_zrec_words(O) 32-bit words
// per z sample, so a sample's fetch is one wide scalar load from one address:
//   [3k] packed z lattice bytes of octave k (low face | high face << 8), [3k + 1] dz,
//   [3k + 2] fade(dz); [3O] adj, [3O + 1] post (the z-bias split); [3O + 2] face flags: bit 2k,
//   octave k enters a new lattice cell at this sample (every octave does at sample 0); bit 2k + 1,
//   that cell's low face is the previous cell's high face (z byte stepped by one), so only the new
//   high face is built.
// Records 0 .. stride: the one past the last sample is zero and only ever fetched ahead.
__host__ __device__ constexpr int32_t uw_zrec_words(int32_t octaves) { return (3 * octaves + 3 + 3) & ~3; }

struct uw_ztab {
    const uint32_t* rec;
    int32_t rw;  // words per record
};

__host__ __device__ inline size_t uw_ztab_bytes(int32_t octaves, int32_t stride) {
    return size_t(stride + 1) * size_t(uw_zrec_words(octaves)) * sizeof(uint32_t);
}

__device__ __forceinline__ uw_ztab uw_ztab_of(const unsigned char* base, int32_t chunk, int32_t octaves,
                                              int32_t stride) {
    uw_ztab t;
    t.rec = reinterpret_cast<const uint32_t*>(base + size_t(chunk) * uw_ztab_bytes(octaves, stride));
    t.rw = uw_zrec_words(octaves);
    return t;
}

// Record i of a chunk's z table (sample coordinate cz, L samples, i <= L; i == L gives the zero
// record).
__device__ __forceinline__ void uw_fill_zrec(const uw_field& f, int32_t cz, int32_t L, float step, int32_t i,
                                             uint32_t* rec) {
    const int32_t O = f.octaves;
    const int32_t rw = uw_zrec_words(O);
    if (i >= L) {
        for (int32_t w = 0; w < rw; ++w) rec[w] = 0u;
        return;
    }
    const double pz = uw_sample_coord(f, 2, cz, i, step);
    const double pp = i > 0 ? uw_sample_coord(f, 2, cz, i - 1, step) : 0.0;
    uint32_t flags = 0;
    for (int32_t k = 0; k < O; ++k) {
        const uw_axis z = uw_axis_of(pz, f.octave_freq[k]);
        rec[3 * k] = z.bits;
        rec[3 * k + 1] = __float_as_uint(z.d);
        rec[3 * k + 2] = __float_as_uint(z.fade);
        if (i == 0) {
            flags |= 1u << (2 * k);
        } else {
            const uint32_t prev = uw_axis_of(pp, f.octave_freq[k]).bits;
            if (z.bits != prev) {
                flags |= 1u << (2 * k);
                if ((z.bits & 255u) == ((prev >> 8) & 255u)) flags |= 2u << (2 * k);
            }
        }
    }
    float adj, post;
    uw_bias_terms(f, pz, &adj, &post);
    rec[3 * O] = __float_as_uint(adj);
    rec[3 * O + 1] = __float_as_uint(post);
    rec[3 * O + 2] = flags;
    for (int32_t w = 3 * O + 3; w < rw; ++w) rec[w] = 0u;
}

// Record g of a launch's z tables (chunk g / (stride + 1), record g % (stride + 1)); res == null:
// every chunk has n_uniform cells per axis. Records past a chunk's spare, and chunks whose
// resolution does not fit the stride (flagged invalid by the mesher, which never reads them), are
// left unwritten.
__device__ __forceinline__ void uw_ztab_record(const uw_field& f, const int32_t* __restrict__ coords,
                                               const int32_t* __restrict__ res, int32_t n_uniform, int32_t nchunks,
                                               int32_t stride, unsigned char* __restrict__ out, uint64_t g) {
    const uint32_t per = uint32_t(stride) + 1u;  // records per chunk
    if (g >= uint64_t(nchunks) * per) return;
    const int32_t c = int32_t(g / per), i = int32_t(g - uint64_t(c) * per);
    const int32_t n = res ? res[c] : n_uniform;
    if (n < 1 || n + 1 > stride) return;
    if (i > n + 1) return;
    uint32_t* rec = reinterpret_cast<uint32_t*>(out + size_t(c) * uw_ztab_bytes(f.octaves, stride)) +
                    size_t(i) * size_t(uw_zrec_words(f.octaves));
    uw_fill_zrec(f, coords[3 * c + 2], n + 1, float(f.extent) / float(n), i, rec);
}

// Scalar (constant address space) loads of wave-uniform HBM data written by an earlier kernel.
typedef const __attribute__((address_space(4))) uint32_t* uw_const_u32_ptr;

// Minimum / maximum of two finite values as single instructions (fminf / fmaxf also quiet
// signalling NaN operands first, two more instructions per call in a loop).
__device__ __forceinline__ float uw_fmin_raw(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float uw_fmax_raw(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// One z record in scalar registers (OMAX octave slots, the first O used).
template <int OMAX>
struct uw_zsample {
    uint32_t bits[OMAX];
    float d[OMAX], fade[OMAX];
    float adj, post;
    uint32_t flags;
    __device__ __forceinline__ void load(const uint32_t* rec, int32_t O) {
        const uw_const_u32_ptr q = (uw_const_u32_ptr)(rec);
#pragma unroll
        for (int k = 0; k < OMAX; ++k) {
            if (k < O) {
                bits[k] = q[3 * k];
                d[k] = __uint_as_float(q[3 * k + 1]);
                fade[k] = __uint_as_float(q[3 * k + 2]);
            }
        }
        adj = __uint_as_float(q[3 * O]);
        post = __uint_as_float(q[3 * O + 1]);
        flags = q[3 * O + 2];
    }
};

// Sign summary bits of a set of samples: some sample not above iso (so not early-blank), some
// below iso, some at or above iso (a cell can straddle iso only if both of the last two hold).
#define UW_BELOW_NOT_ABOVE 1
#define UW_BELOW_LT 2
#define UW_BELOW_GE 4

// Octave-count specialisation of the density kernels: OCT > 0 compiles for exactly OCT octaves
// (the common counts, so the octave loop unrolls with no per-octave predicates); OCT == 0 reads
// the count from the field (at most UW_MAX_OCTAVES).
template <int OCT>
struct uw_octaves {
    static constexpr int kMax = OCT > 0 ? OCT : UW_MAX_OCTAVES;
    __device__ __forceinline__ static int32_t count(const uw_field& f) { return OCT > 0 ? OCT : f.octaves; }
    __host__ __device__ static constexpr bool accepts(int32_t o) { return OCT > 0 ? o == OCT : o <= UW_MAX_OCTAVES; }
};

// Face-flag word of a z record that sets up every octave's lattice cell from scratch (bit 2k per
// octave, no "low face is the previous high face" bits): the first sample of a line that starts
// inside the z range.
__device__ __forceinline__ uint32_t uw_full_setup_flags(int32_t octaves) {
    return 0x55555555u & ((octaves >= 16) ? 0xFFFFFFFFu : ((1u << (2 * octaves)) - 1u));
}

// One z-line of samples: the line whose octave-k x / y axis entries are AX(k) / AY(k) (hashed x),
// z records from zr on (rw words each; the record after the last sample is fetched ahead and
// discarded), `count` samples. FORCE: the first sample builds every octave's lattice cell whatever
// its record's face flags say (a line that starts inside a longer z range). store(iz, rho)
// receives the densities; vmin / vmax keep the running extremes.
template <int OCT, bool FORCE, class AxisX, class AxisY, class GT, class Store>
__device__ __forceinline__ void uw_noise_zline(const uw_field& f, AxisX AX, AxisY AY, const uint8_t* P, GT G3,
                                               const uint32_t* zr, int32_t rw, int32_t count,
                                               float& vmin, float& vmax, Store store) {
    constexpr int OMAX = uw_octaves<OCT>::kMax;
    const int32_t O = uw_octaves<OCT>::count(f);
    uint32_t bb[OMAX];
    uw_line_oct st[OMAX];
#pragma unroll
    for (int k = 0; k < OMAX; ++k) {
        if (k < O) {
            const uint32_t xb2 = AX(k).bits, yb = AY(k).bits;
            const uint32_t a0 = xb2 & 255u, a1 = xb2 >> 8, y0 = yb & 255u, y1 = yb >> 8;
            bb[k] = uint32_t(P[a0 ^ y0]) | (uint32_t(P[a0 ^ y1]) << 8) | (uint32_t(P[a1 ^ y0]) << 16) |
                    (uint32_t(P[a1 ^ y1]) << 24);
        }
    }
    // Sample iz, then the fetch of the record of sample iz + 1 (the fetch after the last sample
    // reads the spare record and is discarded). The fetch is issued after the face updates,
    // whose LDS reads would otherwise wait for it too, and before the sample arithmetic, which
    // covers its latency. Two records alternate, so the fetched registers are used in place.
    auto sample = [&](const uw_zsample<OMAX>& Z, uw_zsample<OMAX>& Znext, int32_t iz) {
        if (Z.flags) {  // some octave enters a new lattice cell
#pragma unroll
            for (int k = 0; k < OMAX; ++k) {
                if (k < O && (Z.flags & (1u << (2 * k)))) {
                    const uw_axis X = AX(k), Y = AY(k);
                    const uint32_t zb = Z.bits[k];
                    if (Z.flags & (2u << (2 * k))) {
                        st[k].A0 = st[k].A1;
                        st[k].B0 = st[k].B1;
                        uw_line_face(X, Y, bb[k], zb >> 8, G3, &st[k].A1, &st[k].B1);
                        uw_line_diff(st[k]);
                    } else {
                        st[k] = uw_line_setup(X, Y, bb[k], zb, G3);
                    }
                }
            }
        }
        zr += rw;
        Znext.load(zr, O);
        __builtin_amdgcn_sched_barrier(0);  // keep the fetch ahead of the arithmetic
        float total = 0.0f;
#pragma unroll
        for (int k = 0; k < OMAX; ++k)
            if (k < O) total = fmaf(f.weight[k], uw_line_sample(st[k], Z.d[k], Z.fade[k], f.clamp_lo, f.clamp_hi),
                                    total);
        const float rho = (Z.adj + total) + Z.post;
        store(iz, rho);
        vmin = uw_fmin_raw(vmin, rho);
        vmax = uw_fmax_raw(vmax, rho);
    };
    uw_zsample<OMAX> Za, Zb;
    Za.load(zr, O);
    if (FORCE) Za.flags = uw_full_setup_flags(O);
    int32_t iz = 0;
    for (; iz + 1 < count; iz += 2) {  // no exit between a fetch and its use
        sample(Za, Zb, iz);
        sample(Zb, Za, iz + 1);
    }
    if (iz < count) sample(Za, Zb, iz);
}

__device__ __forceinline__ int uw_below_bits(float vmin, float vmax, float iso) {
    return (vmin <= iso ? UW_BELOW_NOT_ABOVE : 0) | (vmin < iso ? UW_BELOW_LT : 0) | (vmax >= iso ? UW_BELOW_GE : 0);
}

// Densities of x-planes [xa, xb) of a chunk whose tables are in T, one thread per (x, y) line.
// Plane x is written to dst + slot(x) * L * L with slot(x) = x % nslot, z fastest.
// Returns, per thread, the OR of the UW_BELOW_* bits of its samples.
template <int OCT, class GT>
__device__ __forceinline__ int uw_noise_planes(const uw_field& f, const uw_chunk_tables& T, const uw_ztab& Zt,
                                               const uint8_t* P, GT G3, int32_t L, int32_t xa, int32_t xb,
                                               float* dst, int32_t nslot, int tid, int nthr) {
    const int32_t lines = (xb - xa) * L;
    const uw_divider dL = uw_make_divider(uint32_t(L)), dslot = uw_make_divider(uint32_t(nslot));
    const int32_t rw = OCT > 0 ? uw_zrec_words(OCT) : Zt.rw;
    // running extremes of the thread's samples; the sign bits follow from them at the end
    float vmin = INFINITY, vmax = -INFINITY;
    for (int32_t t = tid; t < lines; t += nthr) {
        const int32_t tq = int32_t(uw_div(uint32_t(t), dL));
        const int32_t ix = xa + tq, iy = t - tq * L;
        float* out = dst + int32_t(uw_mod(uint32_t(ix), dslot)) * L * L + iy * L;
        uw_noise_zline<OCT, false>(
            f, [&](int k) { return T.AX[k * L + ix]; }, [&](int k) { return T.AY[k * L + iy]; }, P, G3, Zt.rec, rw, L,
            vmin, vmax, [&](int32_t iz, float rho) { out[iz] = rho; });
    }
    return uw_below_bits(vmin, vmax, f.iso);
}

// Octave value and gradient with respect to the octave coordinate (f32), given the per-axis
// lattice data and fade derivatives s'(d) = 30 d^2 (d - 1)^2.
template <class GT>
__device__ __forceinline__ float uw_octave_grad(const uw_axis& X, const uw_axis& Y, const uw_axis& Z, float dfx,
                                                float dfy, float dfz, const uint8_t* P, GT G3, float g[3]) {
    const uint32_t a0 = X.bits & 255, a1 = X.bits >> 8;
    const uint32_t y0 = Y.bits & 255, y1 = Y.bits >> 8;
    const uint32_t z0 = Z.bits & 255, z1 = Z.bits >> 8;
    const uint32_t b00 = P[a0 ^ y0], b01 = P[a0 ^ y1], b10 = P[a1 ^ y0], b11 = P[a1 ^ y1];
    const float4 G000 = uw_grad3(G3, b00 ^ z0), G100 = uw_grad3(G3, b10 ^ z0), G010 = uw_grad3(G3, b01 ^ z0),
                 G110 = uw_grad3(G3, b11 ^ z0);
    const float4 G001 = uw_grad3(G3, b00 ^ z1), G101 = uw_grad3(G3, b10 ^ z1), G011 = uw_grad3(G3, b01 ^ z1),
                 G111 = uw_grad3(G3, b11 ^ z1);
    const float g000 = uw_gdot(G000, X.d, Y.d, Z.d), g100 = uw_gdot(G100, X.dm1, Y.d, Z.d);
    const float g010 = uw_gdot(G010, X.d, Y.dm1, Z.d), g110 = uw_gdot(G110, X.dm1, Y.dm1, Z.d);
    const float g001 = uw_gdot(G001, X.d, Y.d, Z.dm1), g101 = uw_gdot(G101, X.dm1, Y.d, Z.dm1);
    const float g011 = uw_gdot(G011, X.d, Y.dm1, Z.dm1), g111 = uw_gdot(G111, X.dm1, Y.dm1, Z.dm1);
    const float a = X.fade, b = Y.fade, c = Z.fade;
    const float x00 = uw_lerp(a, g000, g100), x10 = uw_lerp(a, g010, g110);
    const float x01 = uw_lerp(a, g001, g101), x11 = uw_lerp(a, g011, g111);
    const float yv0 = uw_lerp(b, x00, x10), yv1 = uw_lerp(b, x01, x11);
    float r = uw_lerp(c, yv0, yv1) * UW_PERLIN_SCALE;
    // d/dx
    const float dx00 = fmaf(dfx, g100 - g000, uw_lerp(a, G000.x, G100.x));
    const float dx10 = fmaf(dfx, g110 - g010, uw_lerp(a, G010.x, G110.x));
    const float dx01 = fmaf(dfx, g101 - g001, uw_lerp(a, G001.x, G101.x));
    const float dx11 = fmaf(dfx, g111 - g011, uw_lerp(a, G011.x, G111.x));
    const float drx = uw_lerp(c, uw_lerp(b, dx00, dx10), uw_lerp(b, dx01, dx11));
    // d/dy
    const float dy0 = fmaf(dfy, x10 - x00, uw_lerp(b, uw_lerp(a, G000.y, G100.y), uw_lerp(a, G010.y, G110.y)));
    const float dy1 = fmaf(dfy, x11 - x01, uw_lerp(b, uw_lerp(a, G001.y, G101.y), uw_lerp(a, G011.y, G111.y)));
    const float dry = uw_lerp(c, dy0, dy1);
    // d/dz
    const float ez0 = uw_lerp(b, uw_lerp(a, G000.z, G100.z), uw_lerp(a, G010.z, G110.z));
    const float ez1 = uw_lerp(b, uw_lerp(a, G001.z, G101.z), uw_lerp(a, G011.z, G111.z));
    const float drz = fmaf(dfz, yv1 - yv0, uw_lerp(c, ez0, ez1));
    if (r > 1.0f || r < -1.0f) {
        g[0] = g[1] = g[2] = 0.0f;
        return r > 1.0f ? 1.0f : -1.0f;
    }
    g[0] = drx * UW_PERLIN_SCALE;
    g[1] = dry * UW_PERLIN_SCALE;
    g[2] = drz * UW_PERLIN_SCALE;
    return r;
}
