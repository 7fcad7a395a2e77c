// sim.hpp -- torus link-traffic simulation of collective schedules.
//
// A case (torus, collective, algorithm) is a list of step classes.  Every
// message of a class carries the same share num/den of the per-rank vector,
// and a class may repeat (the P-1 identical steps of a ring).  For each class
// the GPU kernel counts, per directed link, how many messages cross it under
// dimension-ordered minimal routing; link bytes at any message size are
// count * floor(bytes * num / den), so every figure is an exact integer.
#pragma once

#include <cstdint>
#include <vector>

#include "schedule.hpp"
#include "status.hpp"

namespace bine {

enum SimAlgo : int { SIM_BINE = 0, SIM_RECDOUB = 1, SIM_RING = 2, SIM_BRUCK = 3, SIM_SIMPY = 4 };

// How a class derives a rank's destination (k: step index, d: torus dimension).
enum SimKind : uint32_t {
    SK_TORUS_BINE = 0,     // torus peer, Bine distance rho_k on dimension d (compute_peers, port 0)
    SK_TORUS_RECDOUB = 1,  // torus peer, recursive-doubling distance on dimension d
    SK_RING = 2,           // r -> r + 1 (linear rank order)
    SK_LIN_BINE = 3,       // 1D Bine butterfly partner with mask 2^k (bine_alltoall_mpi)
    SK_LIN_XOR = 4,        // r -> r xor 2^k
    SK_BRUCK = 5,          // r -> r + 2^k (bruck_alltoall)
    SK_TREE_BINE_DOWN = 6, // bine_bcast_l_mpi / bine_scatter_mpi tree, root 0, mask 2^k
    SK_TREE_BINE_UP = 7,   // bine_reduce_mpi / bine_gather_mpi tree, root 0, mask 2^k
    SK_TREE_BINO_DOWN = 8, // binomial tree, root 0: r with r % 2^(k+1) == 0 sends to r + 2^k
    SK_TREE_BINO_UP = 9,   // binomial tree, root 0: r with r % 2^(k+1) == 2^k sends to r - 2^k
    SK_SIMPY = 10,         // simulation/simulator.py peer of step k (its id order and distances)
};

struct SimClass {
    uint32_t kind, d, k;
    uint32_t num, den, rep;  // bytes share num/den of the vector; repetitions
};

struct SimCase {
    Torus t;
    int coll, algo;
    std::vector<SimClass> classes;
};

// Build the class list of a case; false when the algorithm does not exist for
// that collective or torus.
bool sim_build_case(const Torus& t, int coll, int algo, SimCase* out);

// Destination of `rank` in a class (host reference of the device function;
// UINT32_MAX when the rank sends nothing).
uint32_t sim_dest(const Torus& t, const SimClass& c, uint32_t rank);

// Dimension-ordered minimal route: calls f(link) for every directed link crossed,
// link = (node * nd + d) * 2 + (negative ? 1 : 0).
template <typename F>
inline void sim_route(const Torus& t, uint32_t src, uint32_t dst, F f) {
    int cs[kMaxDims], cd[kMaxDims];
    t.coords(src, cs);
    t.coords(dst, cd);
    int cur[kMaxDims] = {cs[0], cs[1], cs[2]};
    for (uint32_t d = 0; d < t.nd; ++d) {
        const int dim = (int)t.dims[d];
        int delta = pmod(cd[d] - cs[d], dim);
        if (!delta) continue;
        const bool neg = 2 * delta > dim;
        const int hops = neg ? dim - delta : delta;
        for (int h = 0; h < hops; ++h) {
            f(((uint64_t)t.id(cur) * t.nd + d) * 2 + (neg ? 1 : 0));
            cur[d] = pmod(cur[d] + (neg ? -1 : 1), dim);
        }
    }
}

}  // namespace bine
