// Queue state and task records of the batched build, shared by the persistent kernel
// (uw_batched.hip) and the bulk path for large chunks (uw_bulk.hip).
#pragma once

#include "uw_internal.h"

// Queue state (one 512-byte block, each counter on its own cache line):
//   byte   0  u32 task head (next queue index to claim)
//   byte 128  u64 vertex range counter (packed launches: vertices | indices << 32)
//   byte 256  u64 index range counter (unpacked launches)
//   byte 384  u64 exit counter: workgroups done | non-blank chunks << 32
//   byte 448  u32 number of queue entries at the front taken by the bulk path
constexpr size_t kQueueBytes = 512;
constexpr int kResvIndsOffset = 16;  // u64 words from the vertex counter to the index counter
constexpr int kExitOffset = 32;      // u64 words from the vertex counter to the exit counter
constexpr size_t kBulkCountByte = 448;

// One queued chunk, in queue order (32 bytes, two 16-byte loads): chunk index c (-1: queue
// empty), resolution n and chunk coordinates. Written by k_place, read once per task.
struct __align__(16) uw_task {
    int32_t c, n, cx, cy, cz, pad[3];
};

__device__ __forceinline__ unsigned long long* uw_queue_resv(uint32_t* head) {
    return reinterpret_cast<unsigned long long*>(reinterpret_cast<unsigned char*>(head) + 128);
}
__device__ __forceinline__ uint32_t* uw_queue_bulk_count(uint32_t* head) {
    return reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned char*>(head) + kBulkCountByte);
}

// Reserves a chunk's contiguous vertex and index ranges (one thread). A packed launch (worst-case
// totals below 2^32) takes both with one 64-bit atomic on vertices | indices << 32; otherwise one
// atomic per range. Flags overflow (and the launch's overflow word) when the ranges pass the caps.
__device__ __forceinline__ void uw_reserve_ranges(unsigned long long* resv, int packed, uint64_t vcap, uint64_t icap,
                                                  uw_totals* totals, uint32_t nv, uint32_t ni, unsigned long long* vo,
                                                  unsigned long long* io, uint32_t* flags) {
    if (packed) {
        const unsigned long long old = atomicAdd(resv, (unsigned long long)nv | ((unsigned long long)ni << 32));
        *vo = old & 0xFFFFFFFFull;
        *io = old >> 32;
    } else {
        *vo = atomicAdd(resv, (unsigned long long)nv);
        *io = atomicAdd(resv + kResvIndsOffset, (unsigned long long)ni);
    }
    if (*vo + nv > vcap || *io + ni > icap) {
        *flags |= UW_CHUNK_OVERFLOW;
        atomicOr(&totals->overflow, 1u);
    }
}

// ---- Bulk path (uw_bulk.hip): large chunks of a batched launch built by plane-parallel launches --
// Launch shape and scratch layout of the bulk path for a launch of `nchunks` chunks (scratch sized
// for every chunk being a bulk chunk of max_res cells: the resolutions are on the device).
struct uw_bulk_shape {
    bool on;         // tuning.bulk_min within [1, max_res]
    int32_t max_res; // largest resolution of the launch
    int32_t entries; // chunks of the launch (count and emit grids: entries x slabs, slab-major)
    int32_t wpc;     // density workgroup items per chunk
    int32_t nx_tab;  // x entries per octave in the density kernel's LDS
    size_t lds_density;
    int grid_density, grid_count, grid_scan, grid_emit;
    int count_planes;  // cell planes per count workgroup: 4 or 8
    int emit_planes;   // cell planes per emit workgroup: 1, 2 or 4
    size_t dens_off, rows_off, cases_off, below_off, act_off, rec_off, bytes;
};

// Buffers of one batched launch that the bulk kernels read or write.
struct uw_bulk_io {
    const uw_task* tasks;
    uint32_t* qhead;
    const unsigned char* ztab;
    int32_t zstride;
    int32_t packed;
    const uw_mesh_out* out;
    uw_vertex_tape* tape;
    unsigned char* scratch;  // uw_bulk_shape::bytes
};

int uw_bulk_shape_for(uw_ctx* ctx, const uw_batch_tuning& t, int32_t nchunks, uw_bulk_shape* s);
// density, count and scan (the range reservations): before the persistent kernel, whose last
// workgroup writes the launch totals from the shared range counter
int uw_bulk_launch_front(uw_ctx* ctx, const uw_bulk_shape& s, const uw_bulk_io& io, hipStream_t st);
// emission
int uw_bulk_launch_back(uw_ctx* ctx, const uw_bulk_shape& s, const uw_bulk_io& io, hipStream_t st);
