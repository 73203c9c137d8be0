// Exercises the C++ facade (include/uwmesh.hpp) the way the reference's world loop drives
// chunk.rs: Perlin::new(seed), Chunk::new(pos), build_full, not_blank, buffer slices, num_inds;
// then the same chunks through ChunkBatch. Prints one line per chunk:
//   <path> <x> <y> <z> <not_blank> <num_verts> <num_inds> <fnv1a of vertex and index bytes>
// and, for each non-blank chunk, the collision query of boid.rs:205 at the chunk's centre:
//   tris <x> <y> <z> <count> <fnv1a of the returned Tri records>
// tests/test_gpu_facade.py compares the lines with the Python binding's results.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/uwmesh.hpp"

static uint64_t fnv1a(const void* p, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

int main(int argc, char** argv) {
    const uint32_t seed = argc > 1 ? uint32_t(strtoul(argv[1], nullptr, 10)) : 7u;
    const int coords[][3] = {{0, 0, -1}, {0, 0, 0}, {1, 2, 0}, {-4, 3, -1}, {0, 0, 2}, {3, 1, -1}};
    const int nc = int(sizeof(coords) / sizeof(coords[0]));
    try {
        uw::Perlin perlin(seed);
        std::vector<int32_t> flat, res;
        for (int i = 0; i < nc; ++i) {
            uw::Chunk ch(coords[i][0], coords[i][1], coords[i][2]);
            ch.build_full(perlin);
            uint64_t h = 0;
            if (ch.not_blank()) {
                const auto& v = ch.verts_buffer_slice();
                const auto& ix = ch.inds_buffer_slice();
                h = fnv1a(ix.data(), ix.size() * 4, fnv1a(v.data(), v.size() * 4));
            } else {
                bool threw = false;
                try {
                    (void)ch.verts_buffer_slice();
                } catch (const uw::Error&) {
                    threw = true;
                }
                if (!threw) {
                    fprintf(stderr, "blank chunk did not refuse its buffers\n");
                    return 3;
                }
            }
            printf("chunk %d %d %d %d %zu %zu %016llx\n", coords[i][0], coords[i][1], coords[i][2], int(ch.not_blank()),
                   ch.num_verts(), ch.num_inds(), (unsigned long long)h);
            if (ch.not_blank()) {
                const float pct[3] = {0.5f, 0.5f, 0.5f};
                const std::vector<uw::Tri> tris = ch.tris_around(pct, 3);
                const uint64_t th = tris.empty() ? 0 : fnv1a(tris.data(), tris.size() * sizeof(uw::Tri));
                printf("tris %d %d %d %zu %016llx\n", coords[i][0], coords[i][1], coords[i][2], tris.size(),
                       (unsigned long long)th);
            }
            flat.insert(flat.end(), coords[i], coords[i] + 3);
            res.push_back(12);
        }
        uw::ChunkBatch batch(perlin);
        auto meshes = batch.build(flat, res);
        for (int i = 0; i < nc; ++i) {
            const auto& m = meshes[size_t(i)];
            uint64_t h = 0;
            if (m.info.n_inds) h = fnv1a(m.inds.data(), m.inds.size() * 4, fnv1a(m.verts.data(), m.verts.size() * 4));
            printf("batch %d %d %d %d %u %u %016llx\n", coords[i][0], coords[i][1], coords[i][2], int(m.info.n_inds > 0),
                   m.info.n_verts, m.info.n_inds, (unsigned long long)h);
        }
    } catch (const uw::Error& e) {
        fprintf(stderr, "uw::Error %d: %s\n", e.code(), e.what());
        return 2;
    }
    return 0;
}
