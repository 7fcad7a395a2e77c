// sim_dest.hpp -- destination of one rank's message in a simulation step
// class; shared by the host reference and the GPU kernel.
#pragma once

#include <cstdint>

#ifndef BINE_HD
#if defined(__HIPCC__)
#define BINE_HD __host__ __device__
#else
#define BINE_HD
#endif
#endif

namespace bine {
namespace simd {

struct TorusD {
    uint32_t dims[3];
    uint32_t nd, size;
};

// Device records of the simulation (one per case, one per step class).
struct SimCaseRec {
    TorusD t;
    uint32_t L, nlinks;
    uint64_t pad;
};
struct SimClassRec {
    uint32_t kind, d, k, case_idx;
    uint64_t mult_off;  // first link counter of the class
};

BINE_HD inline int pm(int a, int b) {
    int r = a % b;
    return r < 0 ? r + b : r;
}
BINE_HD inline uint32_t to_nb(int32_t v) { return (0xAAAAAAAAu + (uint32_t)v) ^ 0xAAAAAAAAu; }
BINE_HD inline int32_t from_nb(uint32_t n) { return (int32_t)((0xAAAAAAAAu ^ n) - 0xAAAAAAAAu); }
BINE_HD inline int rho(uint32_t k) {
    int64_t p = 1;
    for (uint32_t i = 0; i <= k; ++i) p *= -2;
    return (int)((1 - p) / 3);
}
BINE_HD inline void coords(const TorusD& t, uint32_t id, int* c) {
    for (int i = (int)t.nd - 1; i >= 0; --i) {
        c[i] = (int)(id % t.dims[i]);
        id /= t.dims[i];
    }
}
BINE_HD inline uint32_t ident(const TorusD& t, const int* c) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < t.nd; ++i) r = r * t.dims[i] + (uint32_t)pm(c[i], (int)t.dims[i]);
    return r;
}
// simulation/simulator.py:11-26 (for 3D: x varies with id / d1 mod d0, y fastest, z slowest)
BINE_HD inline void simpy_coords(const TorusD& t, uint32_t id, int* c) {
    if (t.nd == 1) {
        c[0] = (int)id;
    } else if (t.nd == 2) {
        c[0] = (int)(id / t.dims[1]);
        c[1] = (int)(id % t.dims[1]);
    } else {
        c[0] = (int)((id / t.dims[1]) % t.dims[0]);
        c[1] = (int)(id % t.dims[1]);
        c[2] = (int)(id / (t.dims[0] * t.dims[1]));
    }
}
BINE_HD inline uint32_t simpy_ident(const TorusD& t, const int* c) {
    if (t.nd == 1) return (uint32_t)c[0];
    if (t.nd == 2) return (uint32_t)(c[0] * (int)t.dims[1] + c[1]);
    return (uint32_t)(c[2] * (int)(t.dims[0] * t.dims[1]) + c[0] * (int)t.dims[1] + c[1]);
}

// Bine tree tests (lib/libbine_coll_bcast.cc:645-663, lib/libbine_coll_reduce.cc:587-611), root 0
BINE_HD inline bool equal_lsbs(uint32_t v, uint32_t mask_lsbs) {
    uint32_t l = to_nb((int32_t)v) & mask_lsbs;
    return l == 0 || l == mask_lsbs;
}
BINE_HD inline uint32_t tree_partner(uint32_t r, uint32_t mask, uint32_t P) {
    return (uint32_t)pm(from_nb(to_nb((int32_t)r) ^ ((mask << 1) - 1)), (int)P);
}

// Destination of rank r (library rank order) or 0xFFFFFFFF when r is silent.
// L = log2(P) for the 1D kinds.
BINE_HD inline uint32_t dest(const TorusD& t, uint32_t kind, uint32_t d, uint32_t k, uint32_t r, uint32_t L) {
    const uint32_t P = t.size;
    switch (kind) {
    case 0:
    case 1: {  // torus peer on dimension d, step k of that dimension
        int c[3];
        coords(t, r, c);
        int dist;
        if (kind == 0) {
            dist = rho(k);
            if (c[d] & 1) dist = -dist;
        } else {
            dist = (c[d] ^ (1 << k)) - c[d];
        }
        c[d] = pm(c[d] + dist, (int)t.dims[d]);
        return ident(t, c);
    }
    case 2: return (r + 1) % P;
    case 3: {
        int dist = from_nb((2u << k) - 1);
        return (uint32_t)((r % 2 == 0) ? pm((int)r + dist, (int)P) : pm((int)r - dist, (int)P));
    }
    case 4: return r ^ (1u << k);
    case 5: return (r + (1u << k)) % P;
    case 6: {  // bcast order: masks 2^(L-1) .. 1; r sends at mask m if it was reached at a larger mask
        const uint32_t m = 1u << k;
        bool recvd = (r == 0);
        for (uint32_t mm = 1u << (L - 1); mm > m && !recvd; mm >>= 1) recvd = equal_lsbs(r, (mm << 1) - 1);
        return recvd ? tree_partner(r, m, P) : 0xFFFFFFFFu;
    }
    case 7: {  // reduce order: masks 1 .. 2^(L-1); r sends at the first mask where it stops receiving
        const uint32_t m = 1u << k;
        for (uint32_t mm = 1; mm <= m; mm <<= 1) {
            bool send = !equal_lsbs(r, (mm << 2) - 1) || ((mm << 1) >= P && r != 0);
            if (send) return mm == m ? tree_partner(r, m, P) : 0xFFFFFFFFu;
        }
        return 0xFFFFFFFFu;
    }
    case 8: {
        const uint32_t m = 1u << k;
        return (r % (2 * m) == 0 && r + m < P) ? r + m : 0xFFFFFFFFu;
    }
    case 9: {
        const uint32_t m = 1u << k;
        return (r % (2 * m) == m) ? r - m : 0xFFFFFFFFu;
    }
    case 10: {  // simulator.py: r's peer at step k; ranks are read in simulator.py's id order
        const int dist_tab[10] = {1, 1, 3, 5, 11, 21, 43, 85, 171, 341};
        int c[3];
        simpy_coords(t, r, c);
        const uint32_t dim = k % t.nd, rel = k / t.nd;
        if (rel >= 10) return 0xFFFFFFFFu;  // beyond simulator.py's table; such cases are refused when built
        int sgn = (rel % 2 == 0) ? 1 : -1;
        if (c[dim] % 2) sgn = -sgn;
        c[dim] = pm(c[dim] + sgn * dist_tab[rel], (int)t.dims[dim]);
        return simpy_ident(t, c);
    }
    default: return 0xFFFFFFFFu;
    }
}

}  // namespace simd
}  // namespace bine
