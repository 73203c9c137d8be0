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

// Noise coordinate of sample i of chunk c along one axis (chunk.rs:107-108): all f64, the
// product by a power-of-two frequency equals the reference's division by CHUNK_SIZE.
__device__ __forceinline__ double uw_sample_coord(const uw_field& f, int32_t c, int32_t i, float step) {
    const double local = double(i) * double(step);
    return (local + double(f.extent * c)) * f.frequency;
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

// One octave of improved Perlin noise at the lattice cell / fractions given by X (hashed), Y, Z.
// P: permutation table; GP[j]: gradient vector of code P[j] & 15.
__device__ __forceinline__ float uw_perlin_octave(const uw_axis& X, const uw_axis& Y, const uw_axis& Z,
                                                  const uint8_t* P, const float4* GP) {
    const uint32_t a0 = X.bits & 255, a1 = X.bits >> 8;
    const uint32_t y0 = Y.bits & 255, y1 = Y.bits >> 8;
    const uint32_t z0 = Z.bits & 255, z1 = Z.bits >> 8;
    const uint32_t b00 = P[a0 ^ y0], b01 = P[a0 ^ y1], b10 = P[a1 ^ y0], b11 = P[a1 ^ y1];
    const float g000 = uw_gdot(GP[b00 ^ z0], X.d, Y.d, Z.d);
    const float g100 = uw_gdot(GP[b10 ^ z0], X.dm1, Y.d, Z.d);
    const float g010 = uw_gdot(GP[b01 ^ z0], X.d, Y.dm1, Z.d);
    const float g110 = uw_gdot(GP[b11 ^ z0], X.dm1, Y.dm1, Z.d);
    const float g001 = uw_gdot(GP[b00 ^ z1], X.d, Y.d, Z.dm1);
    const float g101 = uw_gdot(GP[b10 ^ z1], X.dm1, Y.d, Z.dm1);
    const float g011 = uw_gdot(GP[b01 ^ z1], X.d, Y.dm1, Z.dm1);
    const float g111 = uw_gdot(GP[b11 ^ z1], X.dm1, Y.dm1, Z.dm1);
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
