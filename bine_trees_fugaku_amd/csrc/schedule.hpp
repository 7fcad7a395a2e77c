// schedule.hpp -- Bine schedule primitives for P virtual ranks on a (1-3)D torus.
//
// Host-side C++; everything here is integer index math that the plan compiler
// turns into device work lists.  Semantics follow the reference library
// (HLC-Lab/bine_trees_fugaku, lib/libbine_coll.cc and lib/libbine_common.cc);
// each function names the reference routine whose behaviour it reproduces.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace bine {

constexpr int kMaxDims = 3;
constexpr int kMaxPorts = 6;
constexpr int kMaxSteps = 40;

enum Family : int { FAM_DEFAULT = 0, FAM_BINE = 1, FAM_RECDOUB = 2, FAM_BRUCK = 3, FAM_RING = 4 };
enum Distance : int { DIST_INCREASING = 0, DIST_DECREASING = 1 };

int ceil_log2(uint64_t x);                       // lib/libbine_coll.cc:65-87
inline int pmod(int a, int b) { int r = a % b; return r < 0 ? r + b : r; }  // lib/libbine_coll.cc:59-62
inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
int jacobsthal(int k);                           // rho_k = (1 - (-2)^(k+1)) / 3, lib/libbine_common.h:27

// Negabinary helpers (lib/libbine_coll_bcast.cc:622-633, lib/libbine_coll_alltoall.cc:325-378).
uint32_t to_negabinary(int32_t v);
int32_t from_negabinary(uint32_t nb);
uint32_t bit_reverse32(uint32_t x);
uint32_t negabinary_rank(uint32_t rank, uint32_t size);  // get_rank_negabinary_representation
uint32_t remap_rank(uint32_t rank, uint32_t size);       // remap_rank: reverse(gray(nb)) >> (32 - L)
int remap_distance_doubling(uint32_t num);               // Mersenne decomposition

struct Torus {
    uint32_t dims[kMaxDims] = {1, 1, 1};
    uint32_t nd = 1;
    uint32_t size = 1;
    uint32_t steps = 0;
    uint32_t steps_per_dim[kMaxDims] = {0, 0, 0};

    static bool make(const uint32_t* d, uint32_t n, Torus* out);
    void coords(uint32_t id, int* c) const;   // row-major, last dimension fastest
    uint32_t id(const int* c) const;          // periodic wrap
};

// One schedule step: which dimension it moves in and its index within that dimension.
struct StepDim {
    uint32_t d;
    uint32_t k;
};

bool mirrored_port(int port, uint32_t nd);                            // lib/libbine_coll.cc:15-31
std::vector<StepDim> step_dims(const Torus& t, uint32_t port);         // compute_step_info
// Peer of the rank at `coord` for global step `step` (get_peer_c).
uint32_t peer_of(const Torus& t, const int* coord, uint32_t step, uint32_t port,
                 const std::vector<StepDim>& sd, int family);
// Peer sequence of `rank` (compute_peers, lib/libbine_coll.cc:443-504).
std::vector<uint32_t> peer_sequence(const Torus& t, uint32_t rank, uint32_t port, int family);

// Per-port reduce-scatter block schedule: send_step[r*P + b] is the step at
// which rank r forwards block b (steps() == never).  It reproduces
// BineBitmapCalculator's block_step (lib/libbine_common.cc:1491-1510,
// 2036-2045): the largest starting step i whose subtree below peer_i(r)
// contains b.  The receive set of r at step s is the send set of its peer.
struct BlockSchedule {
    uint32_t P = 0, L = 0, port = 0;
    int family = FAM_BINE;
    std::vector<uint32_t> peers;      // [r*L + s]
    std::vector<uint8_t> send_step;   // [r*P + b]

    void build(const Torus& t, uint32_t port, int family);
    bool sends(uint32_t r, uint32_t s, uint32_t b) const { return send_step[(size_t)r * P + b] == s; }
    uint32_t peer(uint32_t r, uint32_t s) const { return peers[(size_t)r * L + s]; }
};

// Bine binomial tree (get_tree, lib/libbine_coll.cc:122-151, 282-304, 334-389).
struct Tree {
    std::vector<uint32_t> parent, reached_at_step, remapped, remapped_max, subtree_root;
};
bool build_tree(const Torus& t, uint32_t root, uint32_t port, int family, int dist, Tree* out);

// Remapped rank of the B_CONT variant: post-order index in the tree of
// rank 0 (dfs, lib/libbine_common.cc:1981-1993).
uint32_t cont_remapped_rank(const Torus& t, uint32_t rank, uint32_t port, int family);

}  // namespace bine
