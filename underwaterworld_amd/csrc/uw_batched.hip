// Batched path: thousands of chunks of mixed resolution meshed by ONE persistent launch.
//
// A fixed grid of workgroups (a few per CU) pulls chunk tasks from a global queue (atomic head
// over a resolution-sorted order, largest chunks first). Each workgroup builds its chunk entirely
// in LDS:
//   1. per-axis octave tables of the chunk (uw_build_chunk_tables);
//   2. densities, x-slab by x-slab, by the line evaluator (uw_noise_planes) into a ring of
//      density planes;
//   3. a counting pass over the cells (cube case -> owned vertex records, triangles);
//   4. one atomic reservation of the chunk's contiguous vertex and index ranges;
//   5. emission: rounds of cells are classified and block-scanned; active cells are compacted
//      into an LDS list, and the list (flushed when full or at the end of a slab) drives the edge
//      lerps and the index writes. Vertex ids follow the reference's first-seen order through the
//      edge-ownership rule (uw_mesh.h); each owner's vertex base and case sit in an LDS ring of the
//      last cells.
// Chunks that fit one slab compute their densities once; larger chunks recompute them in step 5
// (the densities are not stored anywhere in HBM). The launch sequence (reset, order, persistent
// kernel) has no host synchronisation, so it can be captured in a hipGraph.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>

#include "uw_mesh.h"
#include "uw_noise.h"
#include "uw_queue.h"
#include "uw_wave.h"

namespace {

constexpr int kMaxN = UW_BATCH_MAX_RESOLUTION;
// x-slab cell ring (vertex base << 8 | case), a power of two: the larger ring (fewer, fuller
// emission flushes) when it costs no resident workgroups, or when the launch's chunks still get a
// workgroup each; else the smaller
constexpr int kRingMax = 4096, kRingMin = 2048;
constexpr int kOnePassMax = 2048; // most cells of a chunk built in one pass (the compacted list's cap)
constexpr int kMinList = 512;     // smallest compacted active-cell list
constexpr int kDensBudget = 20 * 1024;  // default LDS bytes for the density slab ring
constexpr int kDefaultBulkMin = 17;     // default bulk_min: the first resolution built in x-slabs

struct BatchArgs {
    uw_field f;
    const uw_tables* tab;
    const uw_task* tasks;
    int32_t nchunks;
    uint32_t* head;
    unsigned long long* resv;  // range counters (see kQueueBytes)
    int32_t packed;            // 1: one packed atomic per reservation (worst-case totals fit 32 bits)
    uw_chunk_info* info;
    uw_totals* totals;
    float* verts;
    uint32_t* inds;
    uw_vertex_tape* tape;  // optional per-vertex record for the adjoint
    const unsigned char* ztab;  // per-chunk z tables (uw_ztab layout), written by k_prep
    float* dscr;                // per-workgroup density scratch for x-slab chunks (null: recompute)
    uint64_t dscr_stride;       // floats per workgroup: (max_res + 1)^3
    int32_t zstride;            // samples per z-table row: max_res + 1
    uw_emit_opts eo;       // vertex / index formats and optional collision records
    uint64_t vcap, icap;
    int32_t max_res;      // largest resolution accepted by this launch (sizes the LDS)
    int32_t list_cap;     // entries of the compacted active-cell list
    int32_t ring_cap;     // entries of the cell ring: kRingMax or kRingMin (a power of two), or list_cap when every chunk is one-pass
    int32_t has_lioff;    // 1: a separate index-offset list follows the active list (x-slab chunks)
    int32_t axis_stride;  // entries per octave row of the x / y tables: max_res + 1
    int8_t sx[kMaxN + 1]; // cell planes per slab for each resolution
};

__global__ void k_reset(uint32_t* head, uw_totals* totals) {
    if (threadIdx.x == 0) {
        *head = 0u;
        uw_totals z = {0, 0, 0, 0};
        *totals = z;
    }
}

constexpr int kSurfBins = 16;                         // surface-distance classes per resolution
constexpr int kOrderBuckets = (kMaxN + 2) * kSurfBins;  // queue-order buckets

// Queue-order bucket of a chunk: resolution first (largest first, invalid last), then how far an
// estimate of the density at the chunk's centre lies from the isovalue (nearest first): the z-bias
// plus the fBm's lowest octave (the largest amplitude), in single precision. A chunk whose centre
// density is near the isovalue is the one likely to hold surface, and to cost a mesh on top of its
// densities; chunks far from it are blank or solid and cost the densities alone. Placing the costly
// ones first (longest task first) leaves cheap tasks for the end of the launch, when the last
// round's balance sets the kernel time. Only the processing order depends on it, never a result.
__device__ __forceinline__ uw_axis order_axis(float q) {
    const float fl = floorf(q);
    const int32_t l = int32_t(fl);
    const float d = q - fl;
    uw_axis a;
    a.bits = uint32_t(l & 255) | (uint32_t((l + 1) & 255) << 8);
    a.d = d;
    a.dm1 = d - 1.0f;
    a.fade = d * d * d * (d * (d * 6.0f - 15.0f) + 10.0f);
    return a;
}
__device__ __forceinline__ int order_bucket(const uw_field& f, const uint8_t* P, const float* G3, int32_t n,
                                            int32_t cx, int32_t cy, int32_t cz) {
    if (n < 1 || n > kMaxN) return (kMaxN + 1) * kSurfBins;
    const float h = 0.5f * float(f.extent), fr = float(f.frequency), f0 = float(f.octave_freq[0]);
    const float px = (float(f.extent * cx) + h + float(f.offset[0])) * fr;
    const float py = (float(f.extent * cy) + h + float(f.offset[1])) * fr;
    const float pz = (float(f.extent * cz) + h + float(f.offset[2])) * fr;
    const uw_axis X = uw_hash_x(order_axis(px * f0), P), Y = order_axis(py * f0), Z = order_axis(pz * f0);
    const float total = f.amp[0] * uw_perlin_octave(X, Y, Z, P, G3);
    float adj, post;
    uw_bias_terms(f, double(pz), &adj, &post);
    const float d = fabsf(uw_density_from(total, f.amp_sum, adj, post) - f.iso);
    const int k = d < 1.0f ? int(d * float(kSurfBins)) : kSurfBins - 1;
    return (kMaxN - n) * kSurfBins + (k < kSurfBins ? k : kSurfBins - 1);
}

// Queue set-up, a counting sort by order_bucket (resolution, largest first and invalid last, then
// nearest to the surface first) in two launches that each spread over the device (one workgroup
// would spend ~5 us on the bucket estimates of 4096 chunks alone):
//  k_prep:  workgroups [0, nest) take kPrepThreads chunks each: bucket estimate, rank within the
//           bucket among the workgroup's chunks (LDS atomic), the workgroup's bucket histogram;
//           the other workgroups write the chunks' z tables, one thread per record (ztab_block).
//  k_place: workgroup j places the chunks of estimate workgroup j: bucket offsets from all
//           histograms (total over all workgroups, plus the part of workgroups before j), then
//           slot = offset + rank. Workgroup 0 also resets the totals and the range counters, and
//           publishes the number of queue entries the bulk path takes (the chunks of buckets below
//           bulk_bucket: resolution bulk_min and up, which come first; 0: no bulk path) and the
//           queue head (that number plus first_task: the persistent kernel's workgroup g starts
//           with the task after the bulk entries and g others, without claiming it). Thread c clears
//           slot c of the bulk path's per-entry counters (early-blank bits, active cells; indexed
//           by queue entry, so clearing every slot below nchunks covers the bulk entries; null
//           without the bulk path).
// A launch of at most kPrepThreads chunks has one estimate workgroup, whose histogram is the
// launch's: it places its chunks and does workgroup 0's resets itself, and k_place is not launched
// (one short launch, ~4 us, less for such launches).
// A launch without the bulk path (whose entries must be an exact prefix of the queue) needs no
// global order either: each estimate workgroup sorts its own chunks by bucket and the queue
// interleaves the workgroups' sorted lists (round q holds every workgroup's q-th chunk), so the
// costly chunks of every workgroup still come first, and k_place is not launched (4096 x 12^3:
// step 0.0927 -> 0.0883 ms, the persistent kernel unchanged). Queue order decides load balance
// only: every chunk's output is the same in any order.
// The histograms are written whole (zeros included) by every launch, so no state carries over.
// Per-chunk key: bucket | rank << 16 (rank < kPrepThreads).
constexpr int kPrepThreads = 256;
static_assert(kOrderBuckets < 65536 && kPrepThreads <= 65536, "bucket and rank share a 32-bit key");

struct OrderBufs {
    size_t keys_off, hist_off, bytes;  // byte offsets after the task records, total bytes
};
__host__ __device__ inline int32_t order_blocks(int32_t nchunks) {
    return int32_t((int64_t(nchunks) + kPrepThreads - 1) / kPrepThreads);
}
inline OrderBufs order_bufs(int32_t nchunks) {
    OrderBufs o;
    o.keys_off = size_t(nchunks) * sizeof(uw_task);
    o.hist_off = (o.keys_off + size_t(nchunks) * sizeof(uint32_t) + 255) & ~size_t(255);
    o.bytes = o.hist_off + size_t(order_blocks(nchunks)) * kOrderBuckets * sizeof(uint32_t);
    return o;
}

// Queue placement and resets (k_place's work), passed to k_prep so that a launch of one estimate
// workgroup (at most kPrepThreads chunks) places its chunks in the same launch.
struct PlaceArgs {
    uw_task* tasks;
    uint32_t* head;
    uw_totals* totals;
    uint32_t first_task, bulk_bucket;
    uint32_t* bulk_below;
    uint32_t* bulk_act;
};

// Workgroup 0's resets of k_place: range and exit counters, launch totals.
__device__ __forceinline__ void reset_queue(const PlaceArgs& pa) {
    unsigned long long* resv = uw_queue_resv(pa.head);
    resv[0] = 0ull;
    resv[kResvIndsOffset] = 0ull;
    resv[kExitOffset] = 0ull;
    uw_totals z = {0, 0, 0, 0};
    *pa.totals = z;
}

// Task record of chunk c at queue slot `slot`, and the clearing of its bulk counters.
__device__ __forceinline__ void place_chunk(const PlaceArgs& pa, int32_t c, int32_t n, int32_t cx, int32_t cy,
                                            int32_t cz, uint32_t slot) {
    if (pa.bulk_below) {
        pa.bulk_below[c] = 0u;
        pa.bulk_act[c] = 0u;
    }
    uw_task t;
    t.c = c;
    t.n = n;
    t.cx = cx;
    t.cy = cy;
    t.cz = cz;
    t.pad[0] = t.pad[1] = t.pad[2] = 0;
    pa.tasks[slot] = t;
}

// Bulk entry count and queue head from the global bucket offsets (one thread).
__device__ __forceinline__ void publish_head(const PlaceArgs& pa, const uint32_t* base) {
    const uint32_t nbig =
        (pa.bulk_bucket > 0u && pa.bulk_bucket < uint32_t(kOrderBuckets)) ? base[pa.bulk_bucket] : 0u;
    *uw_queue_bulk_count(pa.head) = nbig;
    *pa.head = nbig + pa.first_task;
}

__device__ __forceinline__ void estimate_block(const uw_field& f, const uw_tables* __restrict__ tab,
                                               const int32_t* __restrict__ coords, const int32_t* __restrict__ res,
                                               int32_t nchunks, uint32_t* __restrict__ keys,
                                               uint32_t* __restrict__ hist, const PlaceArgs* place, int32_t nest) {
    __shared__ uint32_t cnt[kOrderBuckets];
    __shared__ uint8_t P[256];
    __shared__ float G3[3 * 256];
    __shared__ float4 G16[16];
    const int tid = threadIdx.x;
    const int32_t c = int32_t(blockIdx.x) * kPrepThreads + tid;
    const bool valid = c < nchunks;
    // the noise tables and the chunk, loaded side by side
    const uint32_t perm = tab->perm[tid];
    float4 g16 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid < 16) g16 = tab->grad[tid];
    const int32_t n = valid ? res[c] : 0;
    const int32_t cx = valid ? coords[3 * c] : 0, cy = valid ? coords[3 * c + 1] : 0,
                  cz = valid ? coords[3 * c + 2] : 0;
    for (int k = tid; k < kOrderBuckets; k += kPrepThreads) cnt[k] = 0u;
    if (tid < 16) G16[tid] = g16;
    __syncthreads();
    P[tid] = uint8_t(perm);
    {  // slot j holds the gradient of code P[j] & 15, as the density loops' tables do
        const float4 g = G16[perm & 15];
        G3[3 * tid] = g.x;
        G3[3 * tid + 1] = g.y;
        G3[3 * tid + 2] = g.z;
    }
    __syncthreads();
    const int bk = order_bucket(f, P, G3, n, cx, cy, cz);
    const uint32_t rank = valid ? atomicAdd(&cnt[bk], 1u) : 0u;
    __syncthreads();
    if (place) {
        // the workgroup's own bucket order (exclusive prefix of its bucket counts, three buckets
        // per thread, wave scans): the launch's order when it is the only estimate workgroup,
        // else one list of the interleaved queue (launches without the bulk path)
        __shared__ uint32_t wsum[kPrepThreads / 64];
        uint32_t t3[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) t3[k] = 3 * tid + k < kOrderBuckets ? cnt[3 * tid + k] : 0u;
        const uint32_t mine = t3[0] + t3[1] + t3[2];
        const uint32_t inc = uw_wave_incl_scan(mine);
        if ((tid & 63) == 63) wsum[tid >> 6] = inc;
        __syncthreads();
        uint32_t off = inc - mine;
        for (int w = 0; w < (tid >> 6); ++w) off += wsum[w];
        __syncthreads();  // every count is read: cnt[] takes the offsets
        if (3 * tid < kOrderBuckets) cnt[3 * tid] = off;
        if (3 * tid + 1 < kOrderBuckets) cnt[3 * tid + 1] = off + t3[0];
        if (3 * tid + 2 < kOrderBuckets) cnt[3 * tid + 2] = off + t3[0] + t3[1];
        __syncthreads();
        if (tid == 0 && blockIdx.x == 0) {
            reset_queue(*place);
            publish_head(*place, cnt);
        }
        if (valid) {
            const uint32_t q = cnt[bk] + rank;  // position in this workgroup's bucket order
            uint32_t slot = q;
            if (nest > 1) {
                // round q of the interleaved queue holds every workgroup's q-th chunk; the last
                // workgroup has r chunks, the others kPrepThreads
                const uint32_t j = blockIdx.x, m = uint32_t(nest);
                const uint32_t r = uint32_t(nchunks) - uint32_t(kPrepThreads) * (m - 1u);
                slot = q < r ? q * m + j : r * m + (q - r) * (m - 1u) + j;
            }
            place_chunk(*place, c, n, cx, cy, cz, slot);
        }
        return;
    }
    uint32_t* h = hist + size_t(blockIdx.x) * kOrderBuckets;
    for (int k = tid; k < kOrderBuckets; k += kPrepThreads) h[k] = cnt[k];
    if (valid) keys[c] = uint32_t(bk) | (rank << 16);
}

static_assert(kPrepThreads == 256 && kPrepThreads * 3 >= kOrderBuckets, "three buckets per thread in k_place");

// The z records g0 .. g0 + kPrepThreads - 1 (one per thread; the layout of uw_ztab_record, so
// consecutive records are contiguous), staged in LDS at a pitch of rw + 1 words (no bank aliasing)
// and written as one block with 16-byte stores. A thread writing its own record with 4-byte stores
// at a stride of rw words touched a separate cache line with every lane of every store. Records
// uw_ztab_record leaves unwritten (invalid chunks, past a chunk's last sample) are written as zeros.
__device__ __forceinline__ void ztab_block(const uw_field& f, const int32_t* __restrict__ coords,
                                           const int32_t* __restrict__ res, int32_t nchunks, int32_t stride,
                                           unsigned char* __restrict__ out, uint64_t g0, uint32_t* stage) {
    const int tid = threadIdx.x;
    const int32_t rw = uw_zrec_words(f.octaves), pitch = rw + 1;
    const uint32_t per = uint32_t(stride) + 1u;
    const uint64_t total = uint64_t(nchunks) * per;
    const uint64_t g = g0 + uint64_t(tid);
    uint32_t* rec = stage + tid * pitch;
    bool filled = false;
    if (g < total) {
        const int32_t c = int32_t(g / per), i = int32_t(g - uint64_t(c) * per);
        const int32_t n = res[c];
        if (n >= 1 && n + 1 <= stride && i <= n + 1) {
            uw_fill_zrec(f, coords[3 * c + 2], n + 1, float(f.extent) / float(n), i, rec);
            filled = true;
        }
    }
    if (!filled)
        for (int32_t w = 0; w < rw; ++w) rec[w] = 0u;
    __syncthreads();
    if (g0 >= total) return;  // (block-uniform)
    const uint32_t nrec = uint32_t(min(uint64_t(kPrepThreads), total - g0));
    const uint32_t nq = nrec * uint32_t(rw) / 4u;  // 16-byte words of the block (rw is a multiple of 4)
    uint4* dst = reinterpret_cast<uint4*>(out + g0 * uint64_t(rw) * 4u);
    for (uint32_t q = uint32_t(tid); q < nq; q += kPrepThreads) {
        const uint32_t wd = 4u * q, r = wd / uint32_t(rw), w = wd - r * uint32_t(rw);
        const uint32_t* s = stage + r * uint32_t(pitch) + w;
        dst[q] = make_uint4(s[0], s[1], s[2], s[3]);
    }
}

__global__ void __launch_bounds__(kPrepThreads) k_prep(uw_field f, const uw_tables* __restrict__ tab,
                                                       const int32_t* __restrict__ coords,
                                                       const int32_t* __restrict__ res, int32_t nchunks,
                                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ hist,
                                                       int32_t nest, int32_t stride, unsigned char* __restrict__ ztab,
                                                       PlaceArgs pa) {
    if (int32_t(blockIdx.x) < nest) {
        const bool self_placed = nest == 1 || pa.bulk_bucket == 0u;
        estimate_block(f, tab, coords, res, nchunks, keys, hist, self_placed ? &pa : nullptr, nest);
        return;
    }
    extern __shared__ uint32_t zstage[];  // kPrepThreads records, pitch uw_zrec_words(O) + 1
    ztab_block(f, coords, res, nchunks, stride, ztab, uint64_t(blockIdx.x - uint32_t(nest)) * kPrepThreads, zstage);
}

__global__ void __launch_bounds__(kPrepThreads) k_place(const int32_t* __restrict__ coords,
                                                        const int32_t* __restrict__ res, int32_t nchunks,
                                                        const uint32_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ hist, int32_t nest, PlaceArgs pa) {
    __shared__ uint32_t base[3 * kPrepThreads];
    __shared__ uint32_t wsum[kPrepThreads / 64];
    const int tid = threadIdx.x;
    const int32_t j = int32_t(blockIdx.x);
    if (j == 0 && tid == 0) reset_queue(pa);
    const int32_t c = j * kPrepThreads + tid;
    const bool valid = c < nchunks;
    const uint32_t key = valid ? keys[c] : 0u;
    const int32_t n = valid ? res[c] : 0;
    const int32_t cx = valid ? coords[3 * c] : 0, cy = valid ? coords[3 * c + 1] : 0,
                  cz = valid ? coords[3 * c + 2] : 0;
    // buckets 3 tid .. 3 tid + 2: totals over all estimate workgroups, and over those before j
    // (kHistBatch workgroups' loads in flight together, one L2 round trip for launches of up to
    // 4096 chunks: a loop that waits per workgroup costs one round trip each)
    constexpr int kHistBatch = 16;
    uint32_t tot[3] = {0u, 0u, 0u}, before[3] = {0u, 0u, 0u};
    for (int32_t w0 = 0; w0 < nest; w0 += kHistBatch) {
        uint32_t v[kHistBatch][3];
#pragma unroll
        for (int u = 0; u < kHistBatch; ++u) {
            const int32_t w = w0 + u;
            const uint32_t* h = hist + size_t(w) * kOrderBuckets;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int b = 3 * tid + k;
                v[u][k] = (w < nest && b < kOrderBuckets) ? h[b] : 0u;
            }
        }
#pragma unroll
        for (int u = 0; u < kHistBatch; ++u)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                tot[k] += v[u][k];
                before[k] += w0 + u < j ? v[u][k] : 0u;
            }
    }
    // exclusive prefix of the bucket totals: three per thread, wave scans, then wave offsets
    const uint32_t mine = tot[0] + tot[1] + tot[2];
    const uint32_t inc = uw_wave_incl_scan(mine);
    if ((tid & 63) == 63) wsum[tid >> 6] = inc;
    __syncthreads();
    uint32_t off = inc - mine;
    for (int w = 0; w < (tid >> 6); ++w) off += wsum[w];
    base[3 * tid] = off + before[0];
    base[3 * tid + 1] = off + tot[0] + before[1];
    base[3 * tid + 2] = off + tot[0] + tot[1] + before[2];
    __syncthreads();
    if (j == 0 && tid == 0) publish_head(pa, base);  // workgroup 0 has none before it: base[] are global
    if (valid) place_chunk(pa, c, n, cx, cy, cz, base[key & 0xFFFFu] + (key >> 16));
}

// Reserves a chunk's contiguous vertex and index ranges (one thread). A packed launch (worst-case
// totals below 2^32) takes both with one 64-bit atomic on vertices | indices << 32, halving the
// traffic on the one contended counter; otherwise one atomic per range. The launch totals are
// written by the last workgroup to leave, so no other atomic goes to a shared counter per chunk.
__device__ __forceinline__ void reserve_ranges(const BatchArgs& a, uint32_t nv, uint32_t ni, unsigned long long* vo,
                                               unsigned long long* io, uint32_t* flags) {
    uw_reserve_ranges(a.resv, a.packed, a.vcap, a.icap, a.totals, nv, ni, vo, io, flags);
}

// Copies the record of queue task t (c = -1 past the end of the queue) and t to shared memory.
__device__ __forceinline__ void stage_task(const BatchArgs& a, uint32_t t, int32_t* rec) {
    int4 r0 = make_int4(-1, 0, 0, 0);
    int32_t cz = 0;
    if (t < uint32_t(a.nchunks)) {
        const int4* p = reinterpret_cast<const int4*>(a.tasks + t);
        r0 = p[0];
        cz = p[1].x;
    }
    rec[0] = r0.x;
    rec[1] = r0.y;
    rec[2] = r0.z;
    rec[3] = r0.w;
    rec[4] = cz;
    rec[5] = int32_t(t);
}

template <int BLOCK>
__device__ __forceinline__ unsigned long long block_scan_u64(unsigned long long v, unsigned long long* total,
                                                             unsigned long long* lds) {
    constexpr int W = BLOCK / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long inc = uw_wave_incl_scan(v);
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    unsigned long long base = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const unsigned long long t = lds[w];
        base += (w < wave) ? t : 0ull;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return base + inc - v;
}

struct CellCounts {
    uint32_t cs, nv, nt, act;
};

// Corner densities of cell (x, y, z) from the density-plane ring (plane x at slot x % ns).
__device__ __forceinline__ void cell_corners(const float* ring, uw_divider dns, int32_t L, int32_t x, int32_t y,
                                             int32_t z, float r[8]) {
    const int32_t ns = int32_t(dns.d);
    const int32_t s0 = int32_t(uw_mod(uint32_t(x), dns));
    const int32_t s1 = (s0 + 1 == ns) ? 0 : s0 + 1;
    const float* p0 = ring + s0 * L * L + y * L + z;
    const float* p1 = ring + s1 * L * L + y * L + z;
    r[0] = p0[0];
    r[1] = p1[0];
    r[2] = p1[1];
    r[3] = p0[1];
    r[4] = p0[L];
    r[5] = p1[L];
    r[6] = p1[L + 1];
    r[7] = p0[L + 1];
}

__device__ __forceinline__ CellCounts classify_cell(const float* ring, uw_divider dns, int32_t L, int32_t x,
                                                    int32_t y, int32_t z, float iso, const uint32_t* lut) {
    float r[8];
    cell_corners(ring, dns, L, x, y, z, r);
    CellCounts c;
    c.cs = uw_case_of(r, iso);
    const uint32_t w = lut[c.cs];
    c.nv = __popc(w & 0xFFFu & uw_owned_mask(x, y, z));
    c.nt = w >> 16;
    c.act = (c.cs != 0u && c.cs != 255u) ? 1u : 0u;
    return c;
}

// LDS carve-up (bytes): fixed part, then the per-octave axis tables, then the density ring.
__host__ __device__ constexpr size_t lds_fixed_bytes(size_t list_cap, size_t ring_cap, int has_lioff) {
    return 256 * 12 /*G3*/ + 256 * 8 /*TRI*/ + 256 * 4 /*LUT*/ + 256 * 12 * 2 /*BEF*/ + ring_cap * 4 +
           list_cap * 4 * (has_lioff ? 2 : 1) + 256 /*P*/;
}
__host__ __device__ inline size_t lds_table_bytes(int octaves, int axis_stride) {
    return 2 * size_t(octaves) * size_t(axis_stride) * sizeof(uw_axis);
}

// Emission of one active cell: owned-edge lerps into verts (and the adjoint tape), then the
// cell's indices. vbase: chunk-local vertex id of the cell's first owned record; ibase: global
// index of its first index; owner_voff(q) gives the chunk-local vertex base of cell q; RING holds
// (local base << 8 | case) per cell.
// Vertices of one active cell: one per owned edge of the case, at its rank in first-appearance
// order. vlim ends the chunk's reserved vertex range: writes stay inside it whatever the ids.
__device__ __forceinline__ void emit_vertices(const BatchArgs& a, int32_t L, int32_t x, int32_t y, int32_t z,
                                              uint32_t cs, uint32_t owned, const float r[8], uint64_t vb, float step,
                                              float ox, float oy, float oz, const uint32_t* LUT, const uint16_t* BEF,
                                              uint64_t vlim) {
    const uw_emit_opts& eo = a.eo;
    for (uint32_t vm = LUT[cs] & 0xFFFu & owned; vm; vm &= vm - 1u) {
        const uint32_t ed = uint32_t(__builtin_ctz(vm));
        const uint32_t rank = __popc(uint32_t(BEF[cs * 12 + ed]) & owned);
        if (vb + rank >= vlim) continue;
        float p[3];
        uw_edge_vertex(ed, x, y, z, r, a.f.iso, step, ox, oy, oz, p);
        float* dst = a.verts + eo.vstride * (vb + rank);
        dst[0] = p[0];
        dst[1] = p[1];
        dst[2] = p[2];
        if (eo.vstride == 6) uw_vertex_color(p[2], uw_edge_b(ed), float(a.f.extent), dst + 3);
        if (a.tape) {
            const uint32_t ia = uw_edge_a(ed), ib2 = uw_edge_b(ed);
            const uint32_t ca = uw_corner_ofs(ia), cb = uw_corner_ofs(ib2);
            uw_vertex_tape tv;
            tv.a = uint32_t(((x + int32_t(ca & 1u)) * L + y + int32_t((ca >> 1) & 1u)) * L + z + int32_t((ca >> 2) & 1u));
            tv.b = uint32_t(((x + int32_t(cb & 1u)) * L + y + int32_t((cb >> 1) & 1u)) * L + z + int32_t((cb >> 2) & 1u));
            tv.rho_a = r[ia];
            tv.rho_b = r[ib2];
            a.tape[vb + rank] = tv;
        }
    }
}

// Chunk-local vertex id of the record cell (x, y, z) sees as edge ed: its own rank when it owns
// the edge, else the owner's (owner_voff(q) gives the vertex base of cell q; RING holds
// (local base << 8 | case) per cell).
template <class OwnerVoff>
__device__ __forceinline__ uint32_t edge_id(int32_t n, int32_t x, int32_t y, int32_t z, uint32_t cs, uint32_t owned,
                                            uint32_t ed, uint32_t myv, const uint16_t* BEF, const uint32_t* RING,
                                            uint32_t ring_mask, OwnerVoff owner_voff) {
    if (owned & (1u << ed)) return myv + __popc(uint32_t(BEF[cs * 12 + ed]) & owned);
    int32_t qx, qy, qz;
    uint32_t qe;
    uw_edge_owner(x, y, z, ed, &qx, &qy, &qz, &qe);
    const uint32_t q = uint32_t((qx * n + qy) * n + qz);
    const uint32_t qw = RING[q & ring_mask];
    return owner_voff(q, qw) + __popc(uint32_t(BEF[(qw & 255u) * 12 + qe]) & uw_owned_mask(qx, qy, qz));
}

// Emission of one active cell: its vertices (and the adjoint tape), then its index entries in the
// row's order, and, when requested, its collision records. ib: global index of its first index;
// vlim / ilim end the chunk's reserved ranges.
template <class OwnerVoff>
__device__ __forceinline__ void emit_cell(const BatchArgs& a, int32_t n, int32_t L, int32_t x, int32_t y, int32_t z,
                                          uint32_t cs, const float r[8], uint32_t myv, uint64_t vo, uint64_t ib,
                                          float step, float ox, float oy, float oz, const uint64_t* TRI,
                                          const uint32_t* LUT, const uint16_t* BEF, const uint32_t* RING,
                                          uint32_t ring_mask, OwnerVoff owner_voff, uint64_t vlim, uint64_t ilim) {
    const uint32_t owned = uw_owned_mask(x, y, z);
    const uw_emit_opts& eo = a.eo;
    emit_vertices(a, L, x, y, z, cs, owned, r, vo + myv, step, ox, oy, oz, LUT, BEF, vlim);
    const uint64_t row = TRI[cs];
    const int nidx = 3 * int(row & 15u);
    const bool coll = eo.tri_normals || eo.tri_cells;
    float q0[3] = {0.f, 0.f, 0.f}, q1[3] = {0.f, 0.f, 0.f};
    for (int j = 0; j < nidx; ++j) {
        const uint32_t ed = uint32_t(row >> (4 + 4 * j)) & 15u;
        const uint32_t id = edge_id(n, x, y, z, cs, owned, ed, myv, BEF, RING, ring_mask, owner_voff);
        if (ib + j >= ilim) continue;
        uw_store_index(a.inds, eo.ind16, ib + j, id);
        if (coll) {
            float p[3];
            uw_edge_vertex(ed, x, y, z, r, a.f.iso, step, ox, oy, oz, p);
            const int k = j % 3;
            if (k == 0) {
#pragma unroll
                for (int d = 0; d < 3; ++d) q0[d] = p[d];
            } else if (k == 1) {
#pragma unroll
                for (int d = 0; d < 3; ++d) q1[d] = p[d];
            } else {
                const uint64_t t = (ib + j) / 3;
                if (eo.tri_normals) {
                    float nm[3];
                    uw_tri_normal(q0, q1, p, nm);
                    eo.tri_normals[3 * t] = nm[0];
                    eo.tri_normals[3 * t + 1] = nm[1];
                    eo.tri_normals[3 * t + 2] = nm[2];
                }
                if (eo.tri_cells) eo.tri_cells[t] = uint32_t((x * n + y) * n + z);
            }
        }
    }
}

template <int BLOCK, int OCT>
__global__ void __launch_bounds__(BLOCK) k_batched(BatchArgs a) {
    constexpr int W = BLOCK / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    const int32_t O = a.f.octaves;
    float* G3 = reinterpret_cast<float*>(smem);  // gradient of each permutation slot, 3 floats
    uint64_t* TRI = reinterpret_cast<uint64_t*>(G3 + 3 * 256);
    uint32_t* LUT = reinterpret_cast<uint32_t*>(TRI + 256);
    uint16_t* BEF = reinterpret_cast<uint16_t*>(LUT + 256);  // first-seen rank masks [256][12]
    uint32_t* RING = reinterpret_cast<uint32_t*>(BEF + 256 * 12);
    uint32_t* LIST = RING + a.ring_cap;
    uint32_t* LIOFF = LIST + a.list_cap;  // x-slab chunks only (has_lioff)
    // list_cap and ring_cap are multiples of 4, so the tables below stay 16-byte aligned
    uint8_t* P = reinterpret_cast<uint8_t*>(LIST + a.list_cap * (a.has_lioff ? 2 : 1));
    uw_axis* AX = reinterpret_cast<uw_axis*>(P + 256);
    uw_axis* AY = AX + O * a.axis_stride;
    float* DENS = reinterpret_cast<float*>(AY + O * a.axis_stride);
    __shared__ unsigned long long scan_lds[W];
    __shared__ uint32_t wtot[W][3];
    __shared__ int32_t s_next[6];  // c, n, cx, cy, cz and queue index of the next task
    __shared__ uint32_t s_bits;
    __shared__ unsigned long long s_vo, s_io;
    __shared__ int s_emit;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uw_tables* tab = a.tab;
    const bool pf = tid == BLOCK - 1;
    const uint32_t nbig = *uw_queue_bulk_count(a.head);
    if (pf) stage_task(a, nbig + blockIdx.x, s_next);
    __syncthreads();
    // a workgroup whose first task is past the end of the queue (a launch whose chunks all went to
    // the bulk path, or fewer tasks than workgroups) skips the table loads and goes to the exit
    const int32_t c_first = __builtin_amdgcn_readfirstlane(s_next[0]);
    if (c_first >= 0 && c_first < a.nchunks) {
        for (int i = tid; i < 256; i += BLOCK) {
            const uint8_t p = tab->perm[i];
            P[i] = p;
            const float4 g = uw_grad_vec(p & 15u);
            G3[3 * i] = g.x;
            G3[3 * i + 1] = g.y;
            G3[3 * i + 2] = g.z;
            TRI[i] = tab->tri[i];
            LUT[i] = uint32_t(tab->used[i]) | (uint32_t(tab->tri[i] & 15u) << 16);
        }
        uw_lds_load_before(BEF, tab, tid, BLOCK);
    }
    __shared__ double s_freq[UW_MAX_OCTAVES];
    if (tid < UW_MAX_OCTAVES) s_freq[tid] = a.f.octave_freq[tid];
    const float iso = a.f.iso;

    // Task queue. Workgroup g starts with task nbig + g (nbig: the entries taken by the bulk path;
    // the queue head starts at nbig plus the grid size). While
    // at least two grids' worth of tasks lie beyond the current one, a one-pass task is given one
    // task of lookahead: the block's last lane, the one with the least density work, claims the
    // next task and stages its record in shared memory once its own share of the density
    // evaluation is done, so the two dependent memory round trips (queue head, then record)
    // overlap the other lanes' work instead of opening the next task. Near the end of the queue,
    // and for multi-slab tasks (long enough that the round trips hardly count), the next task is
    // claimed when the block is free, which keeps the end of the launch balanced.
    uint32_t nonblank = 0;  // thread 0: chunks of this workgroup with indices
    bool staged = true;  // block-uniform: s_next holds the next task

    for (;;) {
        __syncthreads();  // every wave is done with the previous task's shared state
        if (!staged) {
            if (pf) stage_task(a, atomicAdd(a.head, 1u), s_next);
            __syncthreads();
        }
        // chunk-level values are wave-uniform: keep them in scalar registers
        const int32_t c = __builtin_amdgcn_readfirstlane(s_next[0]);
        if (c < 0 || c >= a.nchunks) break;  // end of the queue (the second test guards the info writes)
        const int32_t n = __builtin_amdgcn_readfirstlane(s_next[1]);
        const int32_t cx = __builtin_amdgcn_readfirstlane(s_next[2]);
        const int32_t cy = __builtin_amdgcn_readfirstlane(s_next[3]);
        const int32_t cz = __builtin_amdgcn_readfirstlane(s_next[4]);
        const bool lookahead = uint32_t(__builtin_amdgcn_readfirstlane(s_next[5])) + 2u * gridDim.x <
                               uint32_t(a.nchunks);
        staged = false;
        if (tid == 0) s_bits = 0u;  // first set after the table barrier below
        if (n < 1 || n > a.max_res || !uw_octaves<OCT>::accepts(O)) {
            if (tid == 0) {
                uw_chunk_info r = {UW_CHUNK_INVALID | UW_CHUNK_BLANK, 0, 0, 0, 0, 0};
                a.info[c] = r;
            }
            continue;
        }
        const int32_t L = n + 1, nn = n * n, C = nn * n;
        const uw_divider dn = uw_make_divider(uint32_t(n)), dnn = uw_make_divider(uint32_t(nn));
        const float step = float(a.f.extent) / float(n);
        const float ox = float(a.f.extent * cx), oy = float(a.f.extent * cy), oz = float(a.f.extent * cz);
        const uw_chunk_tables T = {AX, AY, nullptr, nullptr, nullptr};
        uw_build_xy_tables(a.f, s_freq, P, cx, cy, L, step, AX, AY, tid, BLOCK);
        const uw_ztab Zt = uw_ztab_of(a.ztab, c, O, a.zstride);
        __syncthreads();

        const int32_t sx = a.sx[n];
        const int32_t ns = sx + 1;
        const int32_t nslab = (n + sx - 1) / sx;
        const uw_divider dns = uw_make_divider(uint32_t(ns));

        if (nslab == 1 && C <= a.list_cap) {
            // ---------- one density evaluation, one classification pass ----------
            const int bits = uw_noise_planes<OCT>(a.f, T, Zt, P, G3, L, 0, L, DENS, ns, tid, BLOCK);
            // s_next was read before the table barrier; the next read follows the end-of-task barrier
            if (lookahead) {
                if (pf) stage_task(a, atomicAdd(a.head, 1u), s_next);
                staged = true;
            }
            if (bits) atomicOr(&s_bits, uint32_t(bits));
            __syncthreads();
            const uint32_t sb = s_bits;
            const bool straddle = (sb & UW_BELOW_LT) && (sb & UW_BELOW_GE);
            if (!straddle) {
                if (tid == 0) {
                    const uint32_t fl = (sb & UW_BELOW_NOT_ABOVE) ? uint32_t(UW_CHUNK_BLANK)
                                                                : uint32_t(UW_CHUNK_BLANK | UW_CHUNK_EARLY_SKIP);
                    uw_chunk_info r = {fl, 0, 0, 0, 0, 0};
                    a.info[c] = r;
                }
                continue;
            }
            // wave w owns cells [w cpw, (w+1) cpw): wave-level scans, no block barrier
            const uint32_t cpw = (uint32_t(C) + W - 1) / W;
            const uint32_t c0 = uint32_t(wave) * cpw, c1 = min(uint32_t(C), c0 + cpw);
            uint32_t cv = 0, ct = 0, ca = 0;
            for (uint32_t base = c0; base < c1; base += 64) {
                const uint32_t g = base + lane;
                CellCounts cc = {0, 0, 0, 0};
                if (g < c1) {
                    int32_t x, y, z;
                    uw_cell_xyz(g, dn, dnn, &x, &y, &z);
                    cc = classify_cell(DENS, dns, L, x, y, z, iso, LUT);
                }
                // per wave of 64 cells: vertices <= 768 (10 bits), triangles <= 320 (9 bits),
                // active cells <= 64 (7 bits), so one 32-bit scan carries all three
                const uint32_t v = cc.nv | (cc.nt << 10) | (cc.act << 19);
                const uint32_t inc = uw_wave_incl_scan(v);
                const uint32_t ex = inc - v;
                if (g < c1) {
                    RING[g] = ((cv + (ex & 0x3FFu)) << 8) | cc.cs;
                    if (cc.act) {
                        // cell (< 2048) | wave-local triangle prefix << 11
                        const uint32_t slot = c0 + ca + (ex >> 19);
                        LIST[slot] = g | ((ct + ((ex >> 10) & 0x1FFu)) << 11);
                    }
                }
                const uint32_t tot = uw_wave_last(inc);
                cv += tot & 0x3FFu;
                ct += (tot >> 10) & 0x1FFu;
                ca += tot >> 19;
            }
            if (lane == 0) {
                wtot[wave][0] = cv;
                wtot[wave][1] = ct;
                wtot[wave][2] = ca;
            }
            __syncthreads();
            uint32_t vbase[W], ibase[W], abase[W + 1];
            uint32_t nv = 0, nt = 0, na = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                vbase[w] = nv;
                ibase[w] = 3u * nt;
                abase[w] = na;
                nv += wtot[w][0];
                nt += wtot[w][1];
                na += wtot[w][2];
            }
            abase[W] = na;
            const uint32_t ni = 3u * nt;
            if (tid == 0) {
                uint32_t flags = ni == 0 ? uint32_t(UW_CHUNK_BLANK) : 0u;
                unsigned long long vo = 0, io = 0;
                if (ni > 0) reserve_ranges(a, nv, ni, &vo, &io, &flags);
                nonblank += ni > 0 ? 1u : 0u;
                if (a.eo.ind16 && nv > 65535u) flags |= UW_CHUNK_INDEX_WRAP;
                uw_chunk_info r = {flags, nv, ni, na, vo, io};
                a.info[c] = r;
                s_vo = vo;
                s_io = io;
                s_emit = (ni > 0 && !(flags & UW_CHUNK_OVERFLOW)) ? 1 : 0;
            }
            __syncthreads();
            if (!s_emit) continue;
            const unsigned long long vo = s_vo, io = s_io;
            const uw_divider dcpw = uw_make_divider(cpw);
            auto owner_voff = [&](uint32_t q, uint32_t qw) -> uint32_t {
                const uint32_t wq = uw_div(q, dcpw);
                uint32_t b = 0;
#pragma unroll
                for (int w = 0; w < W; ++w) b = (uint32_t(w) == wq) ? vbase[w] : b;
                return b + (qw >> 8);
            };
            // Triangle-parallel indices: when the chunk's triangle -> list-slot map fits the density
            // slab (which the index pass no longer reads), the per-cell loop writes only vertices and
            // every thread then takes whole triangles, so no lane waits for the cell with the most
            // index entries. Collision records need the densities and keep the per-cell loop.
            const bool flat = !(a.eo.tri_normals || a.eo.tri_cells) && nt <= uint32_t(ns * L * L * 2);
            for (uint32_t e = tid; e < na; e += BLOCK) {
                uint32_t w = 0;
#pragma unroll
                for (int k = 1; k < W; ++k) w += (e >= abase[k]) ? 1u : 0u;
                uint32_t wb = 0, wi = 0, wa = 0;
#pragma unroll
                for (int k = 0; k < W; ++k)
                    if (uint32_t(k) == w) {
                        wb = vbase[k];
                        wi = ibase[k];
                        wa = abase[k];
                    }
                const uint32_t slot = w * cpw + (e - wa);
                const uint32_t ent = LIST[slot];
                const uint32_t g = ent & 0x7FFu;
                const uint32_t rw = RING[g];
                int32_t x, y, z;
                uw_cell_xyz(g, dn, dnn, &x, &y, &z);
                float r[8];
                cell_corners(DENS, dns, L, x, y, z, r);
                if (flat)
                    emit_vertices(a, L, x, y, z, rw & 255u, uw_owned_mask(x, y, z), r, vo + wb + (rw >> 8), step, ox,
                                  oy, oz, LUT, BEF, vo + nv);
                else
                    emit_cell(a, n, L, x, y, z, rw & 255u, r, wb + (rw >> 8), vo, io + wi + 3u * (ent >> 11), step,
                              ox, oy, oz, TRI, LUT, BEF, RING, 0xFFFFFFFFu, owner_voff, vo + nv, io + ni);
            }
            if (flat) {
                __syncthreads();  // the density slab is read no more in this chunk
                uint16_t* T2C = reinterpret_cast<uint16_t*>(DENS);
                for (uint32_t e = tid; e < na; e += BLOCK) {
                    uint32_t w = 0;
#pragma unroll
                    for (int k = 1; k < W; ++k) w += (e >= abase[k]) ? 1u : 0u;
                    uint32_t wi = 0, wa = 0;
#pragma unroll
                    for (int k = 0; k < W; ++k)
                        if (uint32_t(k) == w) {
                            wi = ibase[k];
                            wa = abase[k];
                        }
                    const uint32_t slot = w * cpw + (e - wa);
                    const uint32_t ent = LIST[slot];
                    const uint32_t tp = wi / 3u + (ent >> 11);
                    const uint32_t ntri = uint32_t(TRI[RING[ent & 0x7FFu] & 255u] & 15u);
                    for (uint32_t i = 0; i < ntri; ++i) T2C[tp + i] = uint16_t(slot);
                }
                __syncthreads();
                for (uint32_t t = tid; t < nt; t += BLOCK) {
                    const uint32_t slot = T2C[t];
                    const uint32_t w = uw_div(slot, dcpw);
                    uint32_t wb = 0, wi = 0;
#pragma unroll
                    for (int k = 0; k < W; ++k)
                        if (uint32_t(k) == w) {
                            wb = vbase[k];
                            wi = ibase[k];
                        }
                    const uint32_t ent = LIST[slot];
                    const uint32_t g = ent & 0x7FFu;
                    const uint32_t rw = RING[g];
                    const uint32_t cs = rw & 255u, myv = wb + (rw >> 8);
                    const uint32_t j0 = 3u * (t - (wi / 3u + (ent >> 11)));
                    int32_t x, y, z;
                    uw_cell_xyz(g, dn, dnn, &x, &y, &z);
                    const uint32_t owned = uw_owned_mask(x, y, z);
                    const uint64_t row = TRI[cs];
#pragma unroll
                    for (uint32_t k = 0; k < 3; ++k) {
                        const uint32_t ed = uint32_t(row >> (4 + 4 * (j0 + k))) & 15u;
                        const uint32_t id = edge_id(n, x, y, z, cs, owned, ed, myv, BEF, RING, 0xFFFFFFFFu, owner_voff);
                        const uint64_t gi = io + 3ull * t + k;
                        if (gi < io + ni) uw_store_index(a.inds, a.eo.ind16, gi, id);
                    }
                }
            }
            continue;
        }

        // ---------- x-slabs: counting pass, reservation, emission pass ----------
        int below = 0;
        unsigned long long cnt = 0;  // nv | nt << 24 | act << 48 (per thread)
        for (int32_t s = 0; s < nslab; ++s) {
            const int32_t xa = s * sx, xe = min(n, xa + sx);
            below |= uw_noise_planes<OCT>(a.f, T, Zt, P, G3, L, s == 0 ? 0 : xa + 1, xe + 1, DENS, ns, tid, BLOCK) &
                     UW_BELOW_NOT_ABOVE;

            __syncthreads();
            if (nslab > 1 && a.dscr) {
                // the slab's new planes also go to the workgroup's scratch, for the emission pass
                float* gs = a.dscr + uint64_t(blockIdx.x) * a.dscr_stride;
                const int32_t LL = L * L;
                for (int32_t px = s == 0 ? 0 : xa + 1; px <= xe; ++px) {
                    const float* src = DENS + int32_t(uw_mod(uint32_t(px), dns)) * LL;
                    float* dst = gs + px * LL;
                    for (int32_t r = tid; r < LL; r += BLOCK) dst[r] = src[r];
                }
            }
            for (int32_t g = xa * nn + tid; g < xe * nn; g += BLOCK) {
                int32_t x, y, z;
                uw_cell_xyz(uint32_t(g), dn, dnn, &x, &y, &z);
                const CellCounts cc = classify_cell(DENS, dns, L, x, y, z, iso, LUT);
                cnt += (unsigned long long)cc.nv | ((unsigned long long)cc.nt << 24) |
                       ((unsigned long long)cc.act << 48);
            }
            if (nslab > 1) __syncthreads();
        }
        below = __syncthreads_or(below);
        unsigned long long tot;
        block_scan_u64<BLOCK>(cnt, &tot, scan_lds);
        const uint32_t nv = uint32_t(tot & 0xFFFFFFull), nt = uint32_t((tot >> 24) & 0xFFFFFFull),
                       na = uint32_t(tot >> 48);
        const uint32_t ni = 3u * nt;
        if (tid == 0) {
            uint32_t flags = 0;
            unsigned long long vo = 0, io = 0;
            if (!below) flags = UW_CHUNK_BLANK | UW_CHUNK_EARLY_SKIP;
            else if (ni == 0) flags = UW_CHUNK_BLANK;
            if (ni > 0) reserve_ranges(a, nv, ni, &vo, &io, &flags);
            nonblank += ni > 0 ? 1u : 0u;
            if (a.eo.ind16 && nv > 65535u) flags |= UW_CHUNK_INDEX_WRAP;
            uw_chunk_info r = {flags, nv, ni, na, vo, io};
            a.info[c] = r;
            s_vo = vo;
            s_io = io;
            s_emit = (ni > 0 && !(flags & UW_CHUNK_OVERFLOW)) ? 1 : 0;
        }
        __syncthreads();
        if (!s_emit) continue;

        const unsigned long long vo = s_vo, io = s_io;
        // rounds that may stay pending before the cell ring would overwrite an owner
        const int32_t ring = a.ring_cap;  // x-slab launches: a power of two
        const int32_t max_pending_rounds = max(1, (ring - nn - 1) / BLOCK - 1);
        const uint32_t list_limit = uint32_t(a.list_cap - BLOCK);
        auto owner_voff = [](uint32_t, uint32_t qw) -> uint32_t { return qw >> 8; };
        uint32_t carry_v = 0, carry_i = 0;
        for (int32_t s = 0; s < nslab; ++s) {
            const int32_t xa = s * sx, xe = min(n, xa + sx);
            if (nslab > 1) {
                if (a.dscr) {  // the counting pass's densities, back from the scratch
                    const float* gs = a.dscr + uint64_t(blockIdx.x) * a.dscr_stride;
                    const int32_t LL = L * L;
                    for (int32_t px = s == 0 ? 0 : xa + 1; px <= xe; ++px) {
                        const float* src = gs + px * LL;
                        float* dst = DENS + int32_t(uw_mod(uint32_t(px), dns)) * LL;
                        for (int32_t r = tid; r < LL; r += BLOCK) dst[r] = src[r];
                    }
                } else {
                    uw_noise_planes<OCT>(a.f, T, Zt, P, G3, L, s == 0 ? 0 : xa + 1, xe + 1, DENS, ns, tid, BLOCK);
                }
                __syncthreads();
            }
            uint32_t list_n = 0;
            int32_t pending_rounds = 0;
            const int32_t gend = xe * nn;
            for (int32_t g0 = xa * nn; g0 < gend; g0 += BLOCK) {
                const int32_t g = g0 + tid;
                CellCounts cc = {0, 0, 0, 0};
                if (g < gend) {
                    int32_t x, y, z;
                    uw_cell_xyz(uint32_t(g), dn, dnn, &x, &y, &z);
                    cc = classify_cell(DENS, dns, L, x, y, z, iso, LUT);
                }
                const unsigned long long packed = (unsigned long long)cc.nv | ((unsigned long long)cc.nt << 16) |
                                                  ((unsigned long long)cc.act << 32);
                unsigned long long rt;
                const unsigned long long pre = block_scan_u64<BLOCK>(packed, &rt, scan_lds);
                if (g < gend) {
                    RING[g & (ring - 1)] = ((carry_v + uint32_t(pre & 0xFFFFull)) << 8) | cc.cs;
                    if (cc.act) {
                        const uint32_t slot = list_n + uint32_t(pre >> 32);
                        LIST[slot] = uint32_t(g);
                        LIOFF[slot] = carry_i + 3u * uint32_t((pre >> 16) & 0xFFFFull);
                    }
                }
                carry_v += uint32_t(rt & 0xFFFFull);
                carry_i += 3u * uint32_t((rt >> 16) & 0xFFFFull);
                list_n += uint32_t(rt >> 32);
                ++pending_rounds;
                const bool last = g0 + BLOCK >= gend;
                if (last || list_n > list_limit || pending_rounds >= max_pending_rounds) {
                    __syncthreads();
                    for (uint32_t e = tid; e < list_n; e += BLOCK) {
                        const uint32_t gg = LIST[e];
                        const uint32_t w = RING[gg & (ring - 1)];
                        int32_t x, y, z;
                        uw_cell_xyz(gg, dn, dnn, &x, &y, &z);
                        float r[8];
                        cell_corners(DENS, dns, L, x, y, z, r);
                        emit_cell(a, n, L, x, y, z, w & 255u, r, w >> 8, vo, io + LIOFF[e], step, ox, oy, oz, TRI,
                                  LUT, BEF, RING, uint32_t(ring - 1), owner_voff, vo + nv, io + ni);
                    }
                    __syncthreads();
                    list_n = 0;
                    pending_rounds = 0;
                }
            }
        }
    }
    // Launch totals, written by the last workgroup to leave: one 64-bit exit atomic per workgroup
    // (done | non-blank chunks << 32) on its own line, spread over the end of the launch; the
    // reservation atomics of every workgroup returned before its exit atomic, so the range
    // counters read here (atomically, at the same coherence point) are final. The workgroups that
    // had no task (g >= nchunks - nbig: they all leave at once, and their atomics would queue on
    // the one line) are counted by one atomic of the first of them.
    const uint32_t first_idle = nbig < uint32_t(a.nchunks) ? uint32_t(a.nchunks) - nbig : 0u;
    const uint32_t add = blockIdx.x < first_idle ? 1u : (blockIdx.x == first_idle ? gridDim.x - first_idle : 0u);
    if (tid == 0 && add) {
        unsigned long long* exitc = a.resv + kExitOffset;
        const unsigned long long old = atomicAdd(exitc, add | (static_cast<unsigned long long>(nonblank) << 32));
        if (uint32_t(old) + add == gridDim.x) {
            const unsigned long long r0 = atomicAdd(a.resv, 0ull);
            a.totals->n_verts = a.packed ? (r0 & 0xFFFFFFFFull) : r0;
            a.totals->n_inds = a.packed ? (r0 >> 32) : atomicAdd(a.resv + kResvIndsOffset, 0ull);
            a.totals->n_nonblank = uint32_t(old >> 32) + nonblank;
        }
    }
}

using BatchedKernel = void (*)(BatchArgs);

template <int BLOCK>
BatchedKernel pick_octaves(int octaves) {
    switch (octaves) {
        case 3: return k_batched<BLOCK, 3>;
        case 4: return k_batched<BLOCK, 4>;
        case 6: return k_batched<BLOCK, 6>;
        case 8: return k_batched<BLOCK, 8>;
        default: return k_batched<BLOCK, 0>;
    }
}

BatchedKernel pick_kernel(int block, int octaves) {
    switch (block) {
        case 128: return pick_octaves<128>(octaves);
        case 192: return pick_octaves<192>(octaves);
        case 512: return pick_octaves<512>(octaves);
        default: return pick_octaves<256>(octaves);
    }
}

// True when every resolution up to max_res is built in one pass (one slab, n^3 <= list_cap):
// the launch then needs no x-slab buffers (a cell ring and the separate offset list).
bool one_pass_only(const uw_batch_tuning& t, int32_t list_cap) {
    for (int n = 1; n <= t.max_res; ++n)
        if (t.slab[n] != n || n * n * n > list_cap) return false;
    return true;
}

// Compacted-list capacity: every chunk built in one pass (one slab, n^3 <= kOnePassMax cells) must fit,
// and a round of `block` cells must fit on top of a flushable list.
int32_t list_cap_for(const uw_batch_tuning& t) {
    int32_t cap = kMinList > t.block ? kMinList : t.block;
    for (int n = 1; n <= t.max_res; ++n)
        if (t.slab[n] == n && n * n * n <= kOnePassMax) cap = n * n * n > cap ? n * n * n : cap;
    cap = (cap + 3) & ~3;  // keeps the LDS tables after the list 16-byte aligned
    return cap > kOnePassMax ? kOnePassMax : cap;
}

size_t dens_bytes_for(const uw_batch_tuning& t) {
    size_t m = 0;
    for (int n = 1; n <= t.max_res; ++n) {
        const size_t L = size_t(n) + 1;
        const size_t ns = size_t(t.slab[n]) + 1;
        m = ns * L * L * 4 > m ? ns * L * L * 4 : m;
    }
    return m;
}

}  // namespace

// Default tuning: whole-chunk slabs while the (n+1)^3 densities fit the slab budget, else the
// largest slab that does.
void uw_default_batch_tuning(uw_batch_tuning* t) {
    memset(t, 0, sizeof(*t));
    t->block = 256;
    t->grid = 0;
    t->max_res = kMaxN;
    for (int n = 1; n <= kMaxN; ++n) {
        const int L = n + 1;
        int sx = n;
        if (L * L * L * 4 > kDensBudget) sx = max(1, kDensBudget / (4 * L * L) - 1);
        t->slab[n] = int8_t(min(sx, n));
    }
    t->bulk_min = kDefaultBulkMin;
}

int uw_check_batch_tuning(const uw_batch_tuning& t) {
    if (t.block != 128 && t.block != 192 && t.block != 256 && t.block != 512) return UW_ERR_INVALID_ARG;
    if (t.grid < 0 || t.max_res < 1 || t.max_res > kMaxN || t.bulk_min < 0) return UW_ERR_INVALID_ARG;
    for (int n = 1; n <= kMaxN; ++n)
        if (t.slab[n] < 1 || t.slab[n] > n) return UW_ERR_INVALID_ARG;
    return UW_OK;
}

namespace {

// Everything a batched launch needs that must not happen inside stream capture: buffer
// allocation, the kernel's LDS attribute and the persistent grid size.
struct BatchLaunch {
    uw_batch_tuning pt;  // the persistent kernel's shape: max_res below bulk_min when the bulk path is on
    uw_bulk_shape bulk;
    BatchedKernel k;
    size_t lds;
    int grid;
    int32_t ring;       // cell ring entries (list_cap when every chunk is one-pass)
    size_t ztab_bytes;  // z tables of nchunks chunks
    uint64_t dscr_stride;  // density scratch floats per workgroup (0: no resolution of the shape has x-slabs)
    size_t dscr_bytes;     // density scratch of the whole grid
};

// Launch shape of a batched launch (kernel, LDS, grid, ring, scratch sizes); allocates nothing.
// allow_bulk = false keeps every chunk on the persistent kernel.
int shape_batched(uw_ctx* ctx, int32_t nchunks, BatchLaunch* bl, bool allow_bulk = true) {
    const int32_t octaves = ctx->params.octaves;
    bl->ztab_bytes = size_t(nchunks) * uw_ztab_bytes(octaves, ctx->batch_tuning.max_res + 1);
    memset(&bl->bulk, 0, sizeof(bl->bulk));
    if (allow_bulk)
        if (int rc = uw_bulk_shape_for(ctx, ctx->batch_tuning, nchunks, &bl->bulk)) return rc;
    bl->pt = ctx->batch_tuning;
    if (bl->bulk.on) bl->pt.max_res = max(1, ctx->batch_tuning.bulk_min - 1);
    const uw_batch_tuning& tune = bl->pt;
    bl->k = pick_kernel(tune.block, octaves);
    const int32_t list_cap = list_cap_for(tune);
    const bool one_pass = one_pass_only(tune, list_cap);
    const size_t rest = lds_table_bytes(octaves, tune.max_res + 1) + dens_bytes_for(tune);
    auto lds_for = [&](int32_t ring) { return lds_fixed_bytes(size_t(list_cap), size_t(ring), one_pass ? 0 : 1) + rest; };
    auto fits_per_cu = [&](size_t lds, int* per_cu) -> int {
        UW_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(bl->k),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        *per_cu = 0;
        UW_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, reinterpret_cast<const void*>(bl->k),
                                                                 tune.block, lds));
        return UW_OK;
    };
    int per_cu = 0;
    if (one_pass) {
        bl->ring = list_cap;
        bl->lds = lds_for(list_cap);
        if (int rc = fits_per_cu(bl->lds, &per_cu)) return rc;
    } else {  // the larger ring unless it costs resident workgroups that the launch's chunks would use
        int per_max = 0;
        if (int rc = fits_per_cu(lds_for(kRingMax), &per_max)) return rc;
        if (int rc = fits_per_cu(lds_for(kRingMin), &per_cu)) return rc;
        const bool no_loss = per_max >= per_cu || int64_t(nchunks) <= int64_t(per_max) * ctx->num_cus;
        bl->ring = (per_max > 0 && no_loss) ? kRingMax : kRingMin;
        if (bl->ring == kRingMax) per_cu = per_max;
        bl->lds = lds_for(bl->ring);
        if (int rc = fits_per_cu(bl->lds, &per_cu)) return rc;
    }
    if (per_cu < 1) return uw_fail(ctx, UW_ERR_INVALID_ARG, "batched kernel does not fit on a CU with this tuning");
    bl->grid = tune.grid;
    if (bl->grid == 0) bl->grid = per_cu * ctx->num_cus;
    if (bl->grid > nchunks) bl->grid = max(1, nchunks);  // workgroups past the last task would only exit
    // x-slab chunks keep their counting-pass densities in a per-workgroup scratch for the emission
    // pass (4 (n+1)^3 bytes of HBM traffic each way instead of a second density evaluation)
    bool slabs = false;
    for (int n = 1; n <= tune.max_res; ++n) slabs = slabs || tune.slab[n] < n;
    const uint64_t Lm = uint64_t(tune.max_res) + 1;
    bl->dscr_stride = slabs ? Lm * Lm * Lm : 0;
    bl->dscr_bytes = size_t(bl->dscr_stride) * size_t(bl->grid) * sizeof(float);
    return UW_OK;
}

// shape_batched plus the context-owned scratch of the eager path (grown as needed).
int prepare_batched(uw_ctx* ctx, int32_t nchunks, BatchLaunch* bl) {
    if (int rc = shape_batched(ctx, nchunks, bl)) return rc;
    if (size_t(nchunks) > ctx->batch_order_cap) {
        (void)hipFree(ctx->d_batch_order);
        ctx->d_batch_order = nullptr;
        ctx->batch_order_cap = 0;
        UW_HIP(ctx, hipMalloc(&ctx->d_batch_order, order_bufs(nchunks).bytes));
        ctx->batch_order_cap = size_t(nchunks);
    }
    if (!ctx->d_batch_head) UW_HIP(ctx, hipMalloc(&ctx->d_batch_head, kQueueBytes));
    if (bl->ztab_bytes > ctx->batch_ztab_bytes) {
        (void)hipFree(ctx->d_batch_ztab);
        ctx->d_batch_ztab = nullptr;
        ctx->batch_ztab_bytes = 0;
        UW_HIP(ctx, hipMalloc(&ctx->d_batch_ztab, bl->ztab_bytes));
        ctx->batch_ztab_bytes = bl->ztab_bytes;
    }
    if (bl->dscr_bytes > ctx->batch_dscr_bytes) {
        (void)hipFree(ctx->d_batch_dscr);
        ctx->d_batch_dscr = nullptr;
        ctx->batch_dscr_bytes = 0;
        UW_HIP(ctx, hipMalloc(&ctx->d_batch_dscr, bl->dscr_bytes));
        ctx->batch_dscr_bytes = bl->dscr_bytes;
    }
    if (bl->bulk.bytes > ctx->batch_bulk_bytes) {
        (void)hipFree(ctx->d_batch_bulk);
        ctx->d_batch_bulk = nullptr;
        ctx->batch_bulk_bytes = 0;
        UW_HIP(ctx, hipMalloc(&ctx->d_batch_bulk, bl->bulk.bytes));
        ctx->batch_bulk_bytes = bl->bulk.bytes;
    }
    return UW_OK;
}

int launch_batched(uw_ctx* ctx, const BatchLaunch& bl, const int32_t* coords, const int32_t* res, int32_t nchunks,
                   const uw_mesh_out* out, uint32_t* head, void* task_buf, unsigned char* ztab, float* dscr,
                   unsigned char* bulk_scratch, hipStream_t st, uw_vertex_tape* tape = nullptr) {
    const uw_batch_tuning& tune = bl.pt;
    const int32_t max_res = ctx->batch_tuning.max_res;  // z tables and reservations cover bulk chunks too
    BatchArgs a;
    memset(&a, 0, sizeof(a));
    a.f = uw_make_field(ctx->params);
    a.tab = ctx->d_tables;
    a.nchunks = nchunks;
    a.info = out->info;
    a.totals = out->totals;
    a.verts = out->verts;
    a.inds = out->inds;
    a.vcap = out->vert_cap;
    a.icap = out->ind_cap;
    a.tape = tape;
    a.eo = uw_make_emit_opts(out);
    a.max_res = tune.max_res;
    a.list_cap = list_cap_for(tune);
    const bool one_pass = one_pass_only(tune, a.list_cap);
    a.ring_cap = bl.ring;
    a.has_lioff = one_pass ? 0 : 1;
    a.axis_stride = tune.max_res + 1;
    for (int n = 0; n <= kMaxN; ++n) a.sx[n] = n == 0 ? 1 : tune.slab[n];
    uw_task* tasks = static_cast<uw_task*>(task_buf);
    a.tasks = tasks;
    a.resv = reinterpret_cast<unsigned long long*>(reinterpret_cast<unsigned char*>(head) + 128);
    // packed reservations need every launch total below 2^32 (chunks above max_res emit nothing)
    const uint64_t worst_v = uint64_t(nchunks) * uw_max_verts(max_res),
                   worst_i = uint64_t(nchunks) * uw_max_inds(max_res);
    a.packed = (worst_v < (1ull << 32) && worst_i < (1ull << 32)) ? 1 : 0;
    a.head = head;
    a.ztab = ztab;
    a.dscr = bl.dscr_stride ? dscr : nullptr;
    a.dscr_stride = bl.dscr_stride;
    a.zstride = max_res + 1;
    const uint64_t zblocks = (uint64_t(nchunks) * uint64_t(a.zstride + 1) + kPrepThreads - 1) / kPrepThreads;
    const int32_t nest = order_blocks(nchunks);
    if (zblocks + uint64_t(nest) > 0x7FFFFFFFull) return uw_fail(ctx, UW_ERR_INVALID_ARG, "too many chunks for one launch");
    const OrderBufs ob = order_bufs(nchunks);
    uint32_t* keys = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(task_buf) + ob.keys_off);
    uint32_t* hist = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(task_buf) + ob.hist_off);
    uw_bulk_io bio;
    memset(&bio, 0, sizeof(bio));
    PlaceArgs pa;
    pa.tasks = tasks;
    pa.head = head;
    pa.totals = a.totals;
    pa.first_task = uint32_t(bl.grid);
    pa.bulk_bucket = 0;
    pa.bulk_below = nullptr;
    pa.bulk_act = nullptr;
    if (bl.bulk.on) {
        pa.bulk_bucket = uint32_t(kMaxN - ctx->batch_tuning.bulk_min + 1) * kSurfBins;
        bio.tasks = tasks;
        bio.qhead = head;
        bio.ztab = ztab;
        bio.zstride = a.zstride;
        bio.packed = a.packed;
        bio.out = out;
        bio.tape = tape;
        bio.scratch = bulk_scratch;
        pa.bulk_below = reinterpret_cast<uint32_t*>(bulk_scratch + bl.bulk.below_off);
        pa.bulk_act = reinterpret_cast<uint32_t*>(bulk_scratch + bl.bulk.act_off);
    }
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_BATCH_SETUP)) return rc;
    // a launch of at most kPrepThreads chunks, or one without the bulk path, is placed by its
    // estimate workgroups themselves
    const size_t zstage_bytes = size_t(kPrepThreads) * size_t(uw_zrec_words(a.f.octaves) + 1) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_prep, dim3(uint32_t(zblocks) + uint32_t(nest)), dim3(kPrepThreads), zstage_bytes, st, a.f,
                       a.tab, coords,
                       res, nchunks, keys, hist, nest, a.zstride, ztab, pa);
    if (nest > 1 && pa.bulk_bucket != 0u)
        hipLaunchKernelGGL(k_place, dim3(uint32_t(nest)), dim3(kPrepThreads), 0, st, coords, res, nchunks, keys, hist,
                           nest, pa);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    // bulk chunks: densities, counts and range reservations before the persistent kernel (whose
    // last workgroup writes the launch totals from the shared range counter), emission after it
    if (int rc = uw_bulk_launch_front(ctx, bl.bulk, bio, st)) return rc;
    if (int rc = uw_prof_start(ctx, st)) return rc;
    hipLaunchKernelGGL(bl.k, dim3(uint32_t(bl.grid)), dim3(uint32_t(tune.block)), bl.lds, st, a);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    if (int rc = uw_bulk_launch_back(ctx, bl.bulk, bio, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}

bool args_ok(int32_t nchunks, const int32_t* coords, const int32_t* res, const uw_mesh_out* out) {
    return nchunks >= 0 && out && out->info && out->totals && out->vertex_format <= 1 && out->index_format <= 1 &&
           (nchunks == 0 || (coords && res)) &&
           (out->vert_cap == 0 || out->verts) && (out->ind_cap == 0 || out->inds);
}

}  // namespace

extern "C" {

int uw_build_chunks_batched(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                            const uw_mesh_out* out, void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (!args_ok(nchunks, coords, res, out)) return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    UW_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int rc = uw_refuse_capture(ctx, st)) return rc;
    BatchLaunch bl;
    if (int rc = prepare_batched(ctx, nchunks > 0 ? nchunks : 1, &bl)) return rc;
    if (nchunks == 0) {
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, ctx->d_batch_head, out->totals);
        UW_HIP(ctx, hipGetLastError());
        return UW_OK;
    }
    return launch_batched(ctx, bl, coords, res, nchunks, out, ctx->d_batch_head, ctx->d_batch_order, ctx->d_batch_ztab,
                          ctx->d_batch_dscr, ctx->d_batch_bulk, st);
}

int uw_build_chunks_batched_tape(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                                 const uw_mesh_out* out, uw_vertex_tape* tape, void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (!args_ok(nchunks, coords, res, out) || (out->vert_cap > 0 && !tape))
        return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    UW_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int rc = uw_refuse_capture(ctx, st)) return rc;
    BatchLaunch bl;
    if (int rc = prepare_batched(ctx, nchunks > 0 ? nchunks : 1, &bl)) return rc;
    if (nchunks == 0) {
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, ctx->d_batch_head, out->totals);
        UW_HIP(ctx, hipGetLastError());
        return UW_OK;
    }
    return launch_batched(ctx, bl, coords, res, nchunks, out, ctx->d_batch_head, ctx->d_batch_order, ctx->d_batch_ztab,
                          ctx->d_batch_dscr, ctx->d_batch_bulk, st, tape);
}

int uw_get_batch_tuning(const uw_ctx* ctx, uw_batch_tuning* t) {
    if (!ctx || !t) return UW_ERR_INVALID_ARG;
    *t = ctx->batch_tuning;
    return UW_OK;
}

int uw_set_batch_tuning(uw_ctx* ctx, const uw_batch_tuning* t) {
    if (!ctx || !t) return UW_ERR_INVALID_ARG;
    if (uw_check_batch_tuning(*t)) return uw_fail(ctx, UW_ERR_INVALID_ARG, "invalid batch tuning");
    ctx->batch_tuning = *t;
    return UW_OK;
}

int uw_batched_graph_create(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                            const uw_mesh_out* out, uw_graph** g) {
    if (!ctx || !g) return UW_ERR_INVALID_ARG;
    *g = nullptr;
    if (nchunks < 1 || !args_ok(nchunks, coords, res, out)) return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    UW_HIP(ctx, hipSetDevice(ctx->device));
    BatchLaunch bl;
    if (int rc = prepare_batched(ctx, nchunks, &bl)) return rc;
    // the graph owns its queue state, task records and z tables (the context's may be reallocated):
    // queue state at byte 0, task records (with the ordering keys and histograms) after it, z tables
    // after them, then the x-slab density scratch and the bulk path's scratch
    uint32_t* gbuf = nullptr;
    const size_t ord_bytes = (kQueueBytes + order_bufs(nchunks).bytes + 255) & ~size_t(255);
    const size_t ztab_bytes = (bl.ztab_bytes + 255) & ~size_t(255);
    const size_t dscr_bytes = (bl.dscr_bytes + 255) & ~size_t(255);
    UW_HIP(ctx, hipMalloc(&gbuf, ord_bytes + ztab_bytes + dscr_bytes + bl.bulk.bytes));
    unsigned char* gtasks = reinterpret_cast<unsigned char*>(gbuf) + kQueueBytes;
    unsigned char* gztab = reinterpret_cast<unsigned char*>(gbuf) + ord_bytes;
    float* gdscr = reinterpret_cast<float*>(gztab + ztab_bytes);
    unsigned char* gbulk = gztab + ztab_bytes + dscr_bytes;
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(gbuf);
        return uw_fail(ctx, UW_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = UW_OK;
    const bool was_profiling = ctx->profiling;
    ctx->profiling = false;  // no timing events inside a graph
    if (e == hipSuccess) rc = launch_batched(ctx, bl, coords, res, nchunks, out, gbuf, gtasks, gztab, gdscr, gbulk, cs);
    ctx->profiling = was_profiling;
    hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(cs, &graph) : hipSuccess;
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess || rc != UW_OK || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipFree(gbuf);
        if (e != hipSuccess) return uw_hip_fail(ctx, e, "hipStreamBeginCapture");
        if (rc != UW_OK) return rc;
        return uw_hip_fail(ctx, e2, "hipStreamEndCapture");
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        (void)hipFree(gbuf);
        return uw_hip_fail(ctx, e, "hipGraphInstantiate");
    }
    uw_graph* out_g = new uw_graph();
    out_g->graph = graph;
    out_g->exec = exec;
    out_g->device = ctx->device;
    out_g->buffers = gbuf;
    *g = out_g;
    return UW_OK;
}

}  // extern "C"

// ---- No-allocation form -------------------------------------------------------------------------
// The batched build with its scratch (queue state, task records with their ordering keys and
// histograms, z tables) in one file-scope device array instead of context-owned allocations, for
// callers that cannot allocate (code under graph capture, tight serving loops). x-slab chunks
// evaluate their densities again in the emission pass instead of keeping them in a density scratch.
// Each device holds its own copy of the array until the process ends; its device address is taken
// once per device when a context is created on it. Every call on a device shares that device's
// array: calls that overlap on one device (concurrent streams, or graphs replayed concurrently)
// race, and must be serialised by the caller. Each launch rewrites everything it reads (the queue
// set-up kernels reset the counters and write every record), so no clearing is needed between calls.
namespace {
constexpr size_t kNoAllocBytes = size_t(32) << 20;  // UW_NOALLOC_SCRATCH_BYTES
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
unsigned char* g_scratch_ptr[kNoAllocDevices] = {};
std::mutex g_scratch_mu;

size_t noalloc_bytes(const BatchLaunch& bl, int32_t nchunks) {
    return kQueueBytes + ((order_bufs(nchunks).bytes + 255) & ~size_t(255)) + ((bl.ztab_bytes + 255) & ~size_t(255));
}
}  // namespace

// Device address of this device's scratch array (once per device, with the device current).
int uw_noalloc_setup(uw_ctx* ctx) {
    static_assert(kNoAllocBytes == UW_NOALLOC_SCRATCH_BYTES, "scratch size and header disagree");
    if (ctx->device < 0 || ctx->device >= kNoAllocDevices) return UW_OK;  // the no-alloc form refuses such devices
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return UW_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return uw_hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[ctx->device] = static_cast<unsigned char*>(ptr);
    return UW_OK;
}

extern "C" {

int uw_build_chunks_batched_noalloc(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                                    const uw_mesh_out* out, void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (!args_ok(nchunks, coords, res, out)) return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    if (ctx->device < 0 || ctx->device >= kNoAllocDevices)
        return uw_fail(ctx, UW_ERR_INVALID_ARG, "no-allocation scratch: device index out of range");
    UW_HIP(ctx, hipSetDevice(ctx->device));
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[ctx->device];
    }
    if (!base) return uw_fail(ctx, UW_ERR_INVALID_ARG, "no-allocation scratch not set up on this device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    BatchLaunch bl;
    if (int rc = shape_batched(ctx, nchunks > 0 ? nchunks : 1, &bl, false)) return rc;
    const size_t need = noalloc_bytes(bl, nchunks > 0 ? nchunks : 1);
    if (need > kNoAllocBytes)
        return uw_fail(ctx, UW_ERR_INVALID_ARG,
                       "launch needs " + std::to_string(need) + " bytes of scratch; the no-allocation form has " +
                           std::to_string(kNoAllocBytes) + " (fewer chunks, a smaller max_res or fewer octaves fit)");
    uint32_t* head = reinterpret_cast<uint32_t*>(base);
    if (nchunks == 0) {
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, head, out->totals);
        UW_HIP(ctx, hipGetLastError());
        return UW_OK;
    }
    unsigned char* tasks = base + kQueueBytes;
    unsigned char* ztab = tasks + ((order_bufs(nchunks).bytes + 255) & ~size_t(255));
    // timing events only outside graph capture
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (st) UW_HIP(ctx, hipStreamIsCapturing(st, &cs));
    const bool was_profiling = ctx->profiling;
    if (cs != hipStreamCaptureStatusNone) ctx->profiling = false;
    const int rc = launch_batched(ctx, bl, coords, res, nchunks, out, head, tasks, ztab, nullptr, nullptr, st);
    ctx->profiling = was_profiling;
    return rc;
}

int uw_graph_launch(uw_graph* g, void* stream) {
    if (!g) return UW_ERR_INVALID_ARG;
    if (hipSetDevice(g->device) != hipSuccess) return UW_ERR_HIP;
    return hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)) == hipSuccess ? UW_OK : UW_ERR_HIP;
}

void uw_graph_destroy(uw_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    (void)hipFree(g->buffers);
    delete g;
}

}  // extern "C"
