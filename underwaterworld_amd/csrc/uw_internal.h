// Internal declarations shared by the engine's translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/uwmesh.h"
#include "mc_table.h"
#include "uw_div.h"

// Device-resident lookup tables, uploaded once per context.
// Gradient vectors of the 16 hash codes (noise 0.8.2 perlin_3d gradient_dot_v), components in
// {-1, 0, 1}.
constexpr int8_t kUwGrad[16][3] = {
    {1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
    {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1}, {1, 1, 0},  {-1, 1, 0}, {0, -1, 1}, {0, -1, -1},
};
// Component `comp` of every code, two bits per code (two's complement), code h at bits 2h.
__host__ __device__ constexpr uint32_t uw_grad_pack(int comp) {
    uint32_t k = 0;
    for (int h = 0; h < 16; ++h) k |= uint32_t(kUwGrad[h][comp] & 3) << (2 * h);
    return k;
}

struct uw_tables {
    // Packed triangulation rows (see mc_table.h).
    uint64_t tri[256];
    // before[c][e]: 12-bit mask of the distinct edges whose first appearance in row c precedes
    // the first appearance of edge e. A cell's k-th new vertex is its k-th first-seen edge, so the
    // rank of edge e among a cell's owned edges is popcount(before[c][e] & owned).
    uint16_t before[256][12];
    // Edges used by row c (exactly the edges whose endpoints straddle the isovalue).
    uint16_t used[256];
    // Perlin permutation (noise 0.8.2 PermutationTable).
    uint8_t perm[256];
    // Gradient vector for hash code h & 15 (kUwGrad[h]); w unused.
    float4 grad[16];
};

// Scratch arrays of the plain path, grown on demand.
struct uw_plain_scratch {
    size_t chunks = 0;   // capacity in chunks
    int32_t n = 0;       // resolution the arrays were sized for
    float* dens = nullptr;
    uint32_t* below = nullptr;
    uint8_t* cases = nullptr;
    uint32_t* voff = nullptr;
    uint32_t* act = nullptr;   // (cell index) | (case << 24), compacted per chunk
    uint32_t* ioff = nullptr;  // index offset of each compacted active cell
    unsigned char* ztab = nullptr;  // per-chunk z tables of the density kernel
    size_t ztab_bytes = 0;
};

struct uw_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    uint32_t* buffers = nullptr;  // queue head, task records and z tables owned by the graph
};

void uw_default_batch_tuning(uw_batch_tuning* t);

struct uw_ctx {
    uw_params params;
    int device = 0;
    int num_cus = 256;
    uw_batch_tuning batch_tuning;
    uint32_t* d_batch_head = nullptr;
    void* d_batch_order = nullptr;  // task records of the batched queue
    size_t batch_order_cap = 0;
    unsigned char* d_batch_ztab = nullptr;  // per-chunk z tables of the eager batched path
    size_t batch_ztab_bytes = 0;
    float* d_batch_dscr = nullptr;  // per-workgroup density scratch of x-slab chunks (eager batched path)
    size_t batch_dscr_bytes = 0;
    unsigned char* d_batch_bulk = nullptr;  // bulk path scratch of the eager batched path
    size_t batch_bulk_bytes = 0;
    double* d_adj_part = nullptr;  // adjoint backward: per-(workgroup, chunk) parameter sums
    size_t adj_part_bytes = 0;
    uint32_t* d_adj_plan = nullptr;  // adjoint backward: piece plan (piece starts, workgroup first chunks)
    size_t adj_plan_bytes = 0;
    void* d_adj_tabs = nullptr;    // adjoint backward: per-chunk lattice tables
    size_t adj_tabs_bytes = 0;
    unsigned char* d_vol = nullptr;  // single-volume path: tables, density field, row counts, partial sums
    size_t vol_bytes = 0;
    uint8_t perm[256];
    uw_tables host_tables;
    uw_tables* d_tables = nullptr;
    uw_plain_scratch plain;
    // kernel timing (uw_profile_*)
    bool profiling = false;
    std::vector<hipEvent_t> prof_events;  // pairs (start, end), reused
    std::vector<int> prof_stage;          // stage (enum uw_profile_stage) of each recorded pair
    size_t prof_used = 0;                 // events recorded since the last read
    // single-chunk host facade buffers
    int32_t* d_coord1 = nullptr;
    float* d_verts1 = nullptr;
    uint32_t* d_inds1 = nullptr;
    uw_chunk_info* d_info1 = nullptr;
    uw_totals* d_totals1 = nullptr;
    float* d_norm1 = nullptr;     // collision normals of the single-chunk build
    uint32_t* d_cells1 = nullptr; // collision cells of the single-chunk build
    uint64_t verts1_cap = 0, inds1_cap = 0;
    std::string last_error;
};

int uw_fail(uw_ctx* ctx, int code, const std::string& msg);
// Takes the device address of the no-allocation scratch array on the context's (current) device,
// once per device.
int uw_noalloc_setup(uw_ctx* ctx);
// UW_ERR_INVALID_ARG when `st` is capturing a graph: entry points whose launches bind
// context-owned scratch (reallocated by later calls that need more) must not be captured.
int uw_refuse_capture(uw_ctx* ctx, hipStream_t st);
// Record the start / end event of a timed kernel on `st` (no-ops unless profiling is enabled).
// stage: enum uw_profile_stage (0 = the path's dominant kernel).
int uw_prof_start(uw_ctx* ctx, hipStream_t st, int stage = 0);
int uw_prof_end(uw_ctx* ctx, hipStream_t st);
int uw_hip_fail(uw_ctx* ctx, hipError_t e, const char* what);

#define UW_HIP(ctx, expr)                                        \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return uw_hip_fail((ctx), _e, #expr); \
    } while (0)

// Noise-side constants derived from uw_params, passed to kernels by value.
struct uw_field {
    double frequency;
    double offset[3];                    // world-space shift of every sample coordinate
    double octave_freq[UW_MAX_OCTAVES];  // lacunarity^k by repeated multiplication
    float amp[UW_MAX_OCTAVES];           // amplitudes rounded to f32
    float amp_sum;                       // sum of amplitudes (f64 sum rounded to f32)
    float weight[UW_MAX_OCTAVES];        // A_k * (2/sqrt 3) / sum A, rounded once to f32
    float clamp_lo, clamp_hi;            // -+ sqrt(3)/2: the crate's [-1, 1] clamp before scaling
    int32_t octaves;
    int32_t extent;
    int32_t bias_kind;
    float bias_a, bias_b, bias_c;
    float iso;
};

uw_field uw_make_field(const uw_params& p);
int uw_check_resolution(uw_ctx* ctx, int32_t n);
// Writes the per-chunk z tables (uw_ztab layout, `stride` samples per row) of a launch.
int uw_launch_ztables(uw_ctx* ctx, const uw_field& f, const int32_t* coords, const int32_t* res, int32_t n_uniform,
                      int32_t nchunks, int32_t stride, unsigned char* out, hipStream_t st);
