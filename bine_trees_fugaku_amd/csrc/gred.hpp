// gred.hpp -- work-item format shared by the host plan compiler and the HIP
// gather-reduce kernel.
#pragma once

#include <cstdint>

namespace bine {

// Location of one leaf or output: (buffer, rank, element offset).
struct PoolEntry {
    uint32_t rb;   // rank in bits 0..27, buffer id in bits 28..31
    uint32_t off;  // element offset inside the rank's partition (at tile start minus delta)
};

inline PoolEntry make_entry(uint32_t buf, uint32_t rank, uint32_t off) { return PoolEntry{rank | (buf << 28), off}; }

enum : uint16_t {
    ITEM_VEC = 1,  // the tile moves as 16-byte vectors: whole vectors, every leaf/output address dword-aligned
    ITEM_RUN = 2,  // a run of `pad` one-leaf, one-output copies (see GredItem)
};

// Longest run of copies merged into one item: two, one pass of the copy-run kernel.  Lowering also cuts a
// long single copy into runs of two consecutive tiles.  On the 8x8x8 alltoall
// (int64, 4 KiB per rank pair) runs of 2 moved 6.2 TB/s, of 4 5.9, of 8 5.9, of 16 5.6 and of 32 5.3 TB/s:
// longer runs serialise passes inside one workgroup and leave fewer, longer items; unmerged copies ran 5.75.
constexpr uint32_t kMaxRun = 2;

// One tile of one group: elements [delta, delta + n) of the group.
struct alignas(16) GredItem {
    uint32_t leaf_off;  // first leaf in the pool
    uint32_t out_off;   // first output in the pool
    uint16_t lg;        // log2(number of leaves)
    uint16_t flags;     // ITEM_VEC, ITEM_RUN
    uint32_t nouts;
    uint32_t delta;
    uint32_t n;
    uint32_t bucket;    // buffer table index (batches); 0 for single calls
    uint32_t pad;       // ITEM_RUN: number of copies k >= 2.  The pool holds {first leaf, leaf step}
                        // at leaf_off and {first output, output step} at out_off; copy j reads
                        // first leaf + j * step (rank and offset both advance) and writes likewise.
};
static_assert(sizeof(GredItem) == 32, "GredItem layout");

// Per-call buffer table: base address and per-rank stride of the four buffers.
struct BufTable {
    uint64_t base[4];
    uint64_t stride[4];
};

// Launch shape of the gather-reduce kernel.
struct LaunchShape {
    int vpt = 1;          // 16-byte vectors per thread per pass: 1, 2 or 4
    int threads = 256;    // workgroup size: 64..512 (a multiple of 64)
    int persistent = 0;   // 0: one workgroup per item; >0: grid of that many workgroups
    int nontemporal = 0;  // streaming loads/stores
    int two_phase = 0;    // reductions with several outputs: 0 one launch; 1 two launches (the reductions
                          // write their first output, then copies of it fill the others), copies in group
                          // order; 2 the same, copies tile by tile
    int double_tiles = 0; // two phases: the reduction launch's items span two tiles each
};

// The fourth launch-shape word of the C interface: bit 0 streaming, bits 1-2 the two-phase mode, bit 3 reduction
// items of two tiles in a two-phase list.
inline LaunchShape shape_from_words(int vpt, int threads, int persistent, int flags) {
    return LaunchShape{vpt, threads, persistent, flags & 1, (flags >> 1) & 3, (flags >> 3) & 1};
}
inline int shape_flags(const LaunchShape& s) {
    return (s.nontemporal ? 1 : 0) | (s.two_phase << 1) | (s.double_tiles ? 8 : 0);
}

// Items whose tree has at least 2^(log2(waves) + 3) leaves are reduced with
// their leaves split across the waves of the workgroup; their tile is one
// wave wide.  Shallower items spread their elements over all threads.
inline int split_lg(const LaunchShape& s) { return __builtin_ctz((unsigned)(s.threads / 64)) + 3; }
inline int64_t shallow_tile_elems(const LaunchShape& s, size_t ds) { return (int64_t)s.threads * s.vpt * (int64_t)(16 / ds); }
inline int64_t deep_tile_elems(const LaunchShape& s, size_t ds) { return (int64_t)64 * s.vpt * (int64_t)(16 / ds); }

}  // namespace bine
