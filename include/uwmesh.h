/*
 * uwmesh.h -- C ABI of the MI355X (gfx950) chunk-meshing engine.
 *
 * The engine replaces the reference world generator's chunk builder: 3D multi-octave Perlin
 * density sampled over each chunk's voxel grid, then marching-cubes isosurface extraction with
 * the reference's vertex dedup. A chunk coordinate goes in; compact xyz-f32 vertex and u32 index
 * buffers come out; blank chunks are flagged and skipped.
 *
 * Reference interfaces replaced (underwater_world/src/...):
 *   noise::Perlin::new(seed)                    state.rs:358-359    -> uw_ctx_create
 *   perlin_util::iso_at / perlin_3d_octaves     perlin_util.rs:6-29 -> uw_sample_density
 *   Chunk::new + build_full / build_partial     chunk.rs:89-103,266-313 -> uw_chunk_build_host
 *   Chunk::not_blank / num_inds / *_buffer_slice chunk.rs:344-348   -> uw_chunk_info fields
 *   World::build_full_step / build_step loop    world.rs:113-145    -> uw_build_chunks_plain
 *
 * Conventions
 *   - Every entry point returns an int status (UW_OK = 0, negative on error); the context keeps a
 *     message for the last error (uw_last_error).
 *   - Pointers documented "device" are HIP device pointers; "host" pointers are ordinary memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream). Device-pointer entry
 *     points are asynchronous on that stream; host-pointer entry points are synchronous.
 *   - One context per host thread (or per stream); a context is not thread-safe.
 *   - Vertex positions are world-space (chunk.rs:224-229); indices are chunk-local, three per
 *     triangle, in the reference's table order (same winding). Vertex numbering follows the
 *     reference's first-seen order over ordered corner pairs (chunk.rs:233-243).
 */
#ifndef UWMESH_H
#define UWMESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UW_MAX_OCTAVES 16
#define UW_MAX_RESOLUTION 64 /* cells per chunk axis accepted by the chunk paths */

enum uw_status {
    UW_OK = 0,
    UW_ERR_INVALID_ARG = -1,
    UW_ERR_OVERFLOW = -2, /* output capacity too small; totals report the required sizes */
    UW_ERR_HIP = -3,
    UW_ERR_NO_DEVICE = -4,
    UW_ERR_OOM = -5,
};

enum uw_bias_kind {
    UW_BIAS_TERRACE = 0, /* rho = (adj + p) - fmod(adj, c), adj = (z * a) / b  (perlin_util.rs:27-28) */
    UW_BIAS_LINEAR = 1,  /* rho = (z * a + p) + b */
};

/* Density-field and chunk-geometry parameters. uw_params_reference() fills the reference
 * crate's constants (chunk.rs:5-12, perlin_util.rs:6-29). */
typedef struct uw_params {
    uint32_t seed;                        /* Perlin permutation seed (noise::Perlin::new)   */
    int32_t octaves;                      /* 1..UW_MAX_OCTAVES  (chunk.rs:9)                */
    double frequency;                     /* noise units per world unit (chunk.rs:108: 1/16) */
    double lacunarity;                    /* octave frequency multiplier (perlin_util.rs:18) */
    double amplitudes[UW_MAX_OCTAVES];    /* octave amplitudes (perlin_util.rs:17)           */
    float iso;                            /* isovalue (chunk.rs:10)                          */
    int32_t bias_kind;                    /* enum uw_bias_kind                               */
    float bias_a, bias_b, bias_c;         /* bias coefficients, see enum uw_bias_kind        */
    int32_t chunk_extent;                 /* world units per chunk edge (chunk.rs:5)         */
    double offset[3];                     /* world-space shift of every sample (0 = reference) */
} uw_params;

/* Per-chunk result record (32 bytes). */
enum uw_chunk_flags {
    UW_CHUNK_BLANK = 1,      /* no triangles: the reference's !not_blank() (chunk.rs:291,344) */
    UW_CHUNK_EARLY_SKIP = 2, /* every sample above iso: early_blank_check (chunk.rs:131-133)  */
    UW_CHUNK_OVERFLOW = 4,   /* did not fit the output capacity; nothing written for it       */
    UW_CHUNK_INVALID = 8,    /* resolution outside the accepted range; nothing built          */
    UW_CHUNK_INDEX_WRAP = 16, /* u16 indices and more than 65535 vertices: the indices wrapped,
                                 exactly as the reference's `ind as u16` (chunk.rs:243)        */
};

typedef struct uw_chunk_info {
    uint32_t flags;    /* enum uw_chunk_flags                                   */
    uint32_t n_verts;  /* vertices of this chunk (num verts of chunk.rs:36)     */
    uint32_t n_inds;   /* indices of this chunk (num_inds, chunk.rs:348)        */
    uint32_t n_active; /* cells whose case is neither 0 nor 255                 */
    uint64_t vert_off; /* first vertex of this chunk in uw_mesh_out.verts       */
    uint64_t ind_off;  /* first index of this chunk in uw_mesh_out.inds         */
} uw_chunk_info;

/* Batch totals (device, written by the build): required sizes even on overflow. */
typedef struct uw_totals {
    uint64_t n_verts;
    uint64_t n_inds;
    uint32_t n_nonblank;
    uint32_t overflow; /* 1 if any chunk overflowed */
} uw_totals;

/* Vertex wire formats. UW_VERTEX_XYZ: 3 floats (world position). UW_VERTEX_XYZ_RGB: 6 floats,
 * the reference's VertColor (draw.rs:4-9): position, then the sRGB terrain colour of
 * chunk.rs:215-222 (taken, like the reference, from the edge's first-seen cell). */
enum uw_vertex_format { UW_VERTEX_XYZ = 0, UW_VERTEX_XYZ_RGB = 1 };

/* Index formats. UW_INDEX_U16 is the reference's wire format (chunk.rs:243, drawn as
 * IndexFormat::Uint16 at state.rs:507): inds then points to uint16_t values. */
enum uw_index_format { UW_INDEX_U32 = 0, UW_INDEX_U16 = 1 };

/* Output description for the multi-chunk device entry points. All pointers are device
 * pointers: verts holds 3 (or 6, see vertex_format) floats per vertex (vert_cap vertices), inds
 * ind_cap u32 (or u16, see index_format) values, info one record per chunk, totals one record.
 *
 * Collision triangles (optional, NULL = not written; Build.tris of chunk.rs:42,167-175,245-250):
 * triangle t of chunk c (indices ind_off + 3t .. + 2) is record ind_off / 3 + t of
 *   tri_normals  3 floats: normalize((v1 - v0) x (v2 - v0)), or the raw cross product when its
 *                length is 0 (util::Tri::new, util.rs:12-20);
 *   tri_cells    the chunk-local cell (x n + y) n + z that emitted it (the Build.tris key). Within
 *                a chunk the values are non-decreasing, so they are also the CSR of the
 *                reference's cell -> triangles map (see uw_tris_around). */
typedef struct uw_mesh_out {
    float* verts;
    uint32_t* inds;
    uw_chunk_info* info;
    uw_totals* totals;
    uint64_t vert_cap;
    uint64_t ind_cap;
    uint32_t vertex_format; /* enum uw_vertex_format */
    uint32_t index_format;  /* enum uw_index_format  */
    float* tri_normals;
    uint32_t* tri_cells;
} uw_mesh_out;

typedef struct uw_ctx uw_ctx;

/* The reference crate's constants, with the given seed. */
void uw_params_reference(uw_params* p, uint32_t seed);

/* Create a context on HIP device `device`: validates the parameters, builds the permutation
 * table (noise::Perlin::new, state.rs:358-359) and uploads it with the marching-cubes tables. */
int uw_ctx_create(const uw_params* p, int device, uw_ctx** out);
void uw_ctx_destroy(uw_ctx* ctx);
const char* uw_last_error(const uw_ctx* ctx);

/* The context's 256-entry permutation table (host copy). */
int uw_ctx_permutation(const uw_ctx* ctx, uint8_t* out256);

/* Host-only: the permutation table noise::Perlin::new(seed) builds (no device needed). */
void uw_permutation_table(uint32_t seed, uint8_t* out256);

/* Kernel timing: while enabled, every build records HIP events on its stream around its kernels,
 * each pair tagged with a stage. Stage 0 is the path's dominant kernel: the density kernel of the
 * plain path, the persistent kernel of the batched path, the density kernel of the volume path and
 * the backward kernel(s) of the adjoint. uw_profile_read synchronises on the recorded events,
 * returns the summed milliseconds and the number of timed launches of stage 0 since the last read,
 * and resets every stage; uw_profile_read_stages does the same for stages [0, nstages) into
 * arrays. Not for use under graph capture. */
enum uw_profile_stage {
    UW_STAGE_DOMINANT = 0,
    UW_STAGE_PLAIN_CLASSIFY = 1, /* k_classify: cases, counts, compaction */
    UW_STAGE_PLAIN_OFFSETS = 2,  /* k_offsets: chunk offsets */
    UW_STAGE_PLAIN_EMIT = 3,     /* k_emit: vertices and indices */
    UW_STAGE_BATCH_SETUP = 4,    /* k_prep (+ k_place with the bulk path): z tables, queue order */
    UW_STAGE_VOLUME_COUNT = 5,   /* k_vol_count */
    UW_STAGE_VOLUME_SCAN = 6,    /* k_vol_scan_* */
    UW_STAGE_VOLUME_EMIT = 7,    /* k_vol_emit */
    UW_STAGE_VOLUME_TABLES = 8,  /* k_vol_tables */
    UW_STAGE_BULK_FRONT = 9,     /* k_bulk_density + k_bulk_count + k_bulk_scan (batched bulk path) */
    UW_STAGE_BULK_EMIT = 10,     /* k_bulk_emit */
    UW_STAGE_COUNT = 11
};
int uw_profile_enable(uw_ctx* ctx, int on);
int uw_profile_read(uw_ctx* ctx, double* total_ms, uint64_t* launches);
int uw_profile_read_stages(uw_ctx* ctx, int32_t nstages, double* total_ms, uint64_t* launches);

/* Upper bounds for one chunk of n cells per axis: vertices 5 n (n+1)^2, indices 15 n^3. */
uint64_t uw_max_verts(int32_t n);
uint64_t uw_max_inds(int32_t n);

/* Density sampling only (iso_at over every chunk sample, chunk.rs:105-129):
 * coords  device int32[3 * nchunks] chunk coordinates,
 * dens    device float[nchunks * (n+1)^3], per chunk index x (n+1)^2 + y (n+1) + z,
 * below   device uint32[nchunks] or NULL: sign summary of each chunk's samples, bit 0 = some
 *         sample not above iso (so not early-blank), bit 1 = some below iso, bit 2 = some at or
 *         above iso (a triangle needs bits 1 and 2). */
int uw_sample_density(uw_ctx* ctx, const int32_t* coords, int32_t nchunks, int32_t n, float* dens,
                      uint32_t* below, void* stream);

/* Density at arbitrary noise-space points: perlin_util::iso_at(perlin, x, y, z)
 * (perlin_util.rs:24-29), as the boids query it point by point (boid.rs:131-143,324-331).
 * xyz: device double[3 npts]; out: device float[npts]. */
int uw_density_at(uw_ctx* ctx, const double* xyz, int64_t npts, float* out, void* stream);

/* Plain path: every stage of the chunk builder as its own kernel launch over all chunks
 * (density -> classify + compact -> offsets -> emit). Uniform resolution n. `cases` may be NULL
 * or a device uint8[nchunks * n^3] that receives every cell's cube case (chunk.rs:155-162). */
int uw_build_chunks_plain(uw_ctx* ctx, const int32_t* coords, int32_t nchunks, int32_t n,
                          const uw_mesh_out* out, uint8_t* cases, void* stream);

/* Drop-in form of Chunk::new(pos) + build_full (chunk.rs:89-103,266-268): one chunk, host
 * buffers, synchronous. verts receives 3 floats per vertex, inds u32 indices. On
 * UW_ERR_OVERFLOW nothing is copied and info->n_verts / n_inds give the sizes needed. */
int uw_chunk_build_host(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, float* verts,
                        uint64_t vert_cap, uint32_t* inds, uint64_t ind_cap, uw_chunk_info* info);

/* Host output of one chunk with every wire option (the fields mean what they mean in
 * uw_mesh_out, for host memory): verts vert_cap vertices of 3 or 6 floats, inds ind_cap u32 or
 * u16 values, tri_normals / tri_cells (optional) ind_cap / 3 records. */
typedef struct uw_host_out {
    float* verts;
    uint64_t vert_cap;
    void* inds;
    uint64_t ind_cap;
    uint32_t vertex_format; /* enum uw_vertex_format */
    uint32_t index_format;  /* enum uw_index_format  */
    float* tri_normals;
    uint32_t* tri_cells;
} uw_host_out;

/* uw_chunk_build_host with the reference's full output: VertColor vertices (draw.rs:4-9), u16
 * indices (chunk.rs:243) and the collision triangles of Build.tris (chunk.rs:167-175,245-250). */
int uw_chunk_build_host_out(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, const uw_host_out* out,
                            uw_chunk_info* info);

/* Chunk::tris_around (chunk.rs:315-342), host-only: the triangles of the cells within `range`
 * cells of the point local_pos_percent (chunk-local position / chunk extent, per axis) of a chunk
 * of n cells per axis, given that chunk's tri_cells (ntris records). Cells are visited x -> y -> z
 * over [max(m - range, 0), min(m + range, n)] with m = floor(percent * n), and each cell's
 * triangles in emission order, so ids (chunk-local triangle numbers) come out in the reference's
 * order. A box whose end lies below its start is empty (the reference's loop would wrap there).
 * Writes at most cap ids; *count receives the full number (UW_ERR_OVERFLOW if it exceeds cap). */
int uw_tris_around(const uint32_t* tri_cells, uint32_t ntris, int32_t n, float px, float py, float pz,
                   int32_t range, uint32_t* ids, uint32_t cap, uint32_t* count);

/* ---- Batched path: one persistent launch over many chunks of mixed resolution ------------- */

#define UW_BATCH_MAX_RESOLUTION 32

/* Launch shape of the persistent kernel (chosen by the autotuner, see autotune.py).
 * block: threads per workgroup (128, 192, 256 or 512); grid: persistent workgroups (0 = as many
 * as fit on the device); max_res: largest resolution a launch accepts (sizes the LDS; larger
 * chunks are flagged UW_CHUNK_INVALID); slab[n]: cell planes per x-slab for chunks of n cells per
 * axis (slab[n] = n builds the chunk in one slab with one density evaluation; smaller slabs bound
 * LDS use and evaluate the densities twice); bulk_min: chunks of at least bulk_min cells per axis
 * (up to max_res) leave the persistent queue for the bulk path, whose launches spread each chunk's
 * densities and planes over the device (0 or a value above max_res: no bulk path). */
typedef struct uw_batch_tuning {
    int32_t block;
    int32_t grid;
    int32_t max_res;
    int8_t slab[UW_BATCH_MAX_RESOLUTION + 1];
    int32_t bulk_min;
} uw_batch_tuning;

int uw_get_batch_tuning(const uw_ctx* ctx, uw_batch_tuning* t);
int uw_set_batch_tuning(uw_ctx* ctx, const uw_batch_tuning* t);

/* Replaces the per-frame one-chunk-at-a-time loop of World::update (world.rs:93-145) with one
 * build of a whole chunk set. coords: device int32[3 n], res: device int32[n] cells per axis
 * (1..32, per chunk). Output as for uw_build_chunks_plain; chunk ranges are placed in completion
 * order (each chunk's record gives its offsets; contents are deterministic). Asynchronous on
 * `stream`, with no host synchronisation. Its launches bind scratch owned by the context, which a
 * later call with more chunks reallocates, so this entry point must not be captured into a graph:
 * on a capturing stream it returns UW_ERR_INVALID_ARG (as do uw_build_chunks_plain,
 * uw_sample_density, uw_build_chunks_batched_tape and uw_adjoint_backward). Graph use goes through
 * uw_batched_graph_create, whose graph owns its buffers. */
int uw_build_chunks_batched(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                            const uw_mesh_out* out, void* stream);

/* No counterpart in the reference: uw_build_chunks_batched for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It calls
 * no hipMalloc / hipFree: its scratch (queue state, task records, z tables) lives in a file-scope
 * device array of UW_NOALLOC_SCRATCH_BYTES per device, held until the process ends, and x-slab
 * chunks evaluate their densities twice instead of keeping them in a density scratch. Launches
 * needing more scratch than that are refused with UW_ERR_INVALID_ARG (uw_last_error gives the size
 * needed); about 8,000 chunks of up to 32^3 cells at 6 octaves fit. Every call on one device shares
 * that device's array, so calls that overlap on a device (concurrent streams, concurrently replayed
 * graphs) race: the caller serialises them. May be captured into a graph (the launches bind no
 * context-owned buffer). Output as for uw_build_chunks_batched. */
#define UW_NOALLOC_SCRATCH_BYTES (32u << 20)
int uw_build_chunks_batched_noalloc(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                                    const uw_mesh_out* out, void* stream);

/* ---- Differentiation (no counterpart in the reference, which has no gradients) ------------ */

/* Replace the octave amplitudes (octaves values) and/or the world-space sample offset (3 values)
 * of the context's field; NULL leaves a part unchanged. */
int uw_ctx_set_field(uw_ctx* ctx, const double* amplitudes, const double* offset);

/* Per-vertex record of the forward build that the backward pass needs: the chunk-local sample
 * indices (x (n+1)^2 + y (n+1) + z) of the edge's endpoints a -> b and their densities. */
typedef struct uw_vertex_tape {
    uint32_t a, b;
    float rho_a, rho_b;
} uw_vertex_tape;

/* Forward: the batched build of uw_build_chunks_batched that also writes tape[v] for every
 * vertex v it writes (tape: device, out->vert_cap entries). */
int uw_build_chunks_batched_tape(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                                 const uw_mesh_out* out, uw_vertex_tape* tape, void* stream);

/* Backward (vector-Jacobian product) of the forward through the edge lerps to the densities and,
 * with the octave noise recomputed at each vertex's endpoints (nothing but the tape is stored),
 * to the octave amplitudes and the world-space sample offset. The cube cases are held fixed.
 *   info        device records written by the forward (n_verts, vert_off per chunk)
 *   tape        device tape of the forward
 *   grad_verts  device float[3 V]: dL/d(vertex position), laid out like the forward's verts
 *   grad_dens   device float[nchunks * dens_stride] or NULL: dL/d(density) per chunk sample,
 *               written (all nchunks * dens_stride entries are zeroed first, then the
 *               vertices' shares added); dens_stride >= (n+1)^3 of every chunk
 *   grad_amp    device double[nchunks * octaves] or NULL: per-chunk dL/dA_k (includes the
 *               normalisation by the amplitude sum)
 *   grad_off    device double[nchunks * 3] or NULL: per-chunk dL/d(offset)
 * grad_amp and grad_off are summed in a fixed order for a given launch shape; grad_dens is summed
 * by atomics, so it can differ in the last bits between runs. */
int uw_adjoint_backward(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                        const uw_chunk_info* info, const uw_vertex_tape* tape, const float* grad_verts,
                        float* grad_dens, int64_t dens_stride, double* grad_amp, double* grad_off, void* stream);

/* ---- Single large volume: one chunk meshed as one grid --------------------------------------- */

#define UW_VOLUME_MAX_RESOLUTION 640 /* cells per axis; keeps every vertex id and triangle count < 2^32 */

/* Chunk::new(pos) + build_full (chunk.rs:89-103,266-313) for ONE chunk of n cells per axis
 * (1 <= n <= UW_VOLUME_MAX_RESOLUTION) at chunk coordinate (cx, cy, cz), with the context's chunk
 * extent: samples at ((f64) i * (f64) step + extent * c) for i in [0, n], step = (f32) extent / (f32) n
 * (chunk.rs:105-129), cube cases, lerps and the first-seen vertex numbering over the whole grid
 * (chunk.rs:141-243): the result the reference's builder gives for that chunk, as one mesh in one
 * index space. The per-chunk paths stop at 32 cells per axis; this one is tiled internally (32^3
 * tiles with a one-cell halo for the numbering across tile faces).
 * out: as for uw_build_chunks_plain with one chunk (info[0] receives the record, vert_off =
 * ind_off = 0). tape (optional, device, out->vert_cap entries): per vertex the endpoints' sample
 * indices x (n+1)^2 + y (n+1) + z and their densities, for uw_volume_backward. The density field,
 * the tables and the row counts live in a context-owned scratch of uw_volume_scratch_bytes.
 * Asynchronous on `stream`; refuses a capturing stream (UW_ERR_INVALID_ARG). */
int uw_build_volume(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, const uw_mesh_out* out,
                    uw_vertex_tape* tape, void* stream);

/* Backward of uw_build_volume (no counterpart in the reference): as uw_adjoint_backward for the
 * volume's one chunk, with the octave noise recomputed at every vertex's endpoints (no stored
 * field). info: the forward's device record; tape / grad_verts: device, laid out like its vertices;
 * grad_dens: device float[(n+1)^3] (index as the tape's) or NULL, accumulated; grad_amp: device
 * double[octaves] or NULL; grad_off: device double[3] or NULL (both written, summed in a fixed
 * order). */
int uw_volume_backward(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, const uw_chunk_info* info,
                       const uw_vertex_tape* tape, const float* grad_verts, float* grad_dens, double* grad_amp,
                       double* grad_off, void* stream);

/* Bytes of the context-owned scratch that uw_build_volume / uw_volume_backward use for n cells
 * per axis (the density field is 4 (n+1)^3 of it); 0 for invalid arguments. */
uint64_t uw_volume_scratch_bytes(const uw_ctx* ctx, int32_t n);

/* The batched build captured as an instantiated HIP graph over fixed buffers; relaunch it after
 * changing the contents (not the addresses or count) of coords/res. */
typedef struct uw_graph uw_graph;
int uw_batched_graph_create(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                            const uw_mesh_out* out, uw_graph** g);
int uw_graph_launch(uw_graph* g, void* stream);
void uw_graph_destroy(uw_graph* g);

/* ---------------------------------------------------------------------------------------------
 * World streaming (world.rs:64-241) with a batched generation step: World::update re-plans the
 * neighbourhood like the reference, but instead of building one chunk per frame
 * (build_full_step / build_step, world.rs:113-145) it builds up to max_batch pending candidates,
 * in the reference's priority (pop) order, in ONE batched launch. Chunks, the render list and the
 * eviction walk keep the reference's rules; results enter them in pop order.
 *
 * Matrices are cgmath Matrix4<f32> as 16 floats, column-major (m[4 * col + row]), i.e. what
 * Camera::chunk_generation_frustum_matrix(55 deg / 90 deg) returns (camera.rs:58-62).
 * ------------------------------------------------------------------------------------------- */
typedef struct uw_world_config {
    int32_t view_dist;       /* world.rs:8   VIEW_DIST       4  */
    int32_t generation_dist; /* world.rs:9   GENERATION_DIST 5  */
    int32_t keep_dist;       /* world.rs:10  KEEP_DIST       6  */
    int32_t max_z, min_z;    /* world.rs:11-12               2, -2 */
    float recheck_dist;      /* world.rs:5   4.0  */
    float recheck_angle;     /* world.rs:6   0.33 rad */
    int32_t resolution;      /* cells per chunk axis (chunk.rs:6: 12) */
    int32_t max_batch;       /* chunks built per update; 0 = every pending candidate */
    uint32_t vertex_format;  /* enum uw_vertex_format of the stored meshes */
    uint32_t index_format;   /* enum uw_index_format */
    int32_t collision;       /* nonzero: also store collision triangles (for tris_around) */
} uw_world_config;

/* One planned candidate: chunk position and its GenPrio (world.rs:29-34). */
typedef struct uw_world_candidate {
    int32_t pos[3];
    float dist;
    float z;
    int32_t in_view;
    int32_t in_gen;
} uw_world_candidate;

/* The reference's constants; max_batch 0, xyz vertices, u32 indices, collision on. */
void uw_world_config_reference(uw_world_config* cfg);

/* World::update_nearby (world.rs:149-228), host-only. known: nknown chunk positions (int32[3])
 * with their not_blank flags (uint8, may be NULL when nknown == 0). Writes the render list
 * (int32[3] each, at most render_cap) and the generation candidates in pop order (most important
 * first, at most cand_cap); *n_render / *n_cand receive the full counts (UW_ERR_OVERFLOW if a
 * capacity was too small). Priority ties, which the reference orders arbitrarily
 * (sort_unstable_by), come out in reverse generation order. */
int uw_world_plan(const uw_world_config* cfg, const float sub_pos[3], const float view_proj[16],
                  const float gen_proj[16], const int32_t* known, const uint8_t* known_not_blank, int32_t nknown,
                  int32_t* render, int32_t render_cap, int32_t* n_render, uw_world_candidate* cands,
                  int32_t cand_cap, int32_t* n_cand);

typedef struct uw_world uw_world;
int uw_world_create(uw_ctx* ctx, const uw_world_config* cfg, uw_world** out);
void uw_world_destroy(uw_world* w);

/* World::update (world.rs:93-111): remove_far_way, re-plan if reset / moved / turned, then one
 * batched build of up to max_batch candidates. Synchronous: on return the new chunks are stored
 * and the render list is final. *built (may be NULL) receives the number of chunks built. */
int uw_world_update(uw_world* w, const float sub_pos[3], const float sub_bearing[3], const float view_proj[16],
                    const float gen_proj[16], int sub_reset, int32_t* built);

/* world.rs:243-247 */
int32_t uw_world_generate_count(const uw_world* w);
int32_t uw_world_render_count(const uw_world* w);
int32_t uw_world_total_count(const uw_world* w);
/* chunks_to_render (world.rs:243): up to cap positions (int32[3] each); returns the count. */
int32_t uw_world_chunks_to_render(const uw_world* w, int32_t* out, int32_t cap);

/* World::get_chunk (world.rs:89-91). Returns UW_OK and the chunk's record, or UW_ERR_INVALID_ARG
 * when the chunk is not stored. Device pointers (valid until the chunk is evicted) to its
 * vertices, indices and, when the world stores them, collision records; NULL for a blank chunk. */
typedef struct uw_world_chunk {
    uw_chunk_info info;   /* vert_off / ind_off are relative to this chunk's own buffers (0) */
    const float* verts;
    const void* inds;
    const float* tri_normals;
    const uint32_t* tri_cells;
} uw_world_chunk;
int uw_world_get_chunk(const uw_world* w, int32_t x, int32_t y, int32_t z, uw_world_chunk* out);

/* Copies a stored chunk's buffers to host memory (the upload source of chunk.rs:291-303);
 * buffers may be NULL to skip them. UW_ERR_OVERFLOW if a capacity is too small. */
int uw_world_read_chunk(const uw_world* w, int32_t x, int32_t y, int32_t z, const uw_host_out* out,
                        uw_chunk_info* info);

#ifdef __cplusplus
}
#endif
#endif /* UWMESH_H */
