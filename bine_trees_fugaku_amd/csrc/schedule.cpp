// schedule.cpp -- Bine schedule primitives (see schedule.hpp).
#include "schedule.hpp"

#include <cstring>

namespace bine {

int ceil_log2(uint64_t x) {
    if (x <= 1) return 0;
    return 64 - __builtin_clzll(x - 1);
}

int jacobsthal(int k) {
    // (1 - (-2)^(k+1)) / 3: 1, -1, 3, -5, 11, ...  Valid for k < 31.
    int64_t p = 1;
    for (int i = 0; i <= k; ++i) p *= -2;
    return (int)((1 - p) / 3);
}

uint32_t to_negabinary(int32_t v) {
    const uint32_t m = 0xAAAAAAAAu;
    return (m + (uint32_t)v) ^ m;
}

int32_t from_negabinary(uint32_t nb) {
    const uint32_t m = 0xAAAAAAAAu;
    return (int32_t)((m ^ nb) - m);
}

uint32_t bit_reverse32(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}

// Smallest / largest value representable with n negabinary digits.
static int64_t nb_min(uint32_t n) {
    int64_t v = 0;
    for (uint32_t i = 1; i < n; i += 2) v -= (int64_t)1 << i;
    return v;
}
static int64_t nb_max(uint32_t n) {
    int64_t v = 0;
    for (uint32_t i = 0; i < n; i += 2) v += (int64_t)1 << i;
    return v;
}
static bool nb_fits(int64_t x, uint32_t n) { return x >= nb_min(n) && x <= nb_max(n); }

uint32_t negabinary_rank(uint32_t rank, uint32_t size) {
    // Two candidate representatives of the rank's distance class modulo `size`;
    // when both fit in L digits the reference keeps the one passing its MSB
    // test, written with the decimal constant 80000000 (an intended hex MSB).
    const uint32_t nbits = (uint32_t)ceil_log2(size);
    const int64_t r = rank, s = size;
    int64_t a = (rank & 1) ? r : -r;
    int64_t b = (rank & 1) ? r - s : -r + s;
    bool fa = nb_fits(a, nbits), fb = nb_fits(b, nbits);
    uint32_t na = fa ? to_negabinary((int32_t)a) : UINT32_MAX;
    uint32_t nbv = fb ? to_negabinary((int32_t)b) : UINT32_MAX;
    if (!fa) return nbv;
    if (!fb) return na;
    const uint32_t msb_probe = 80000000u >> (32 - nbits);
    return (na & msb_probe) ? na : nbv;
}

uint32_t remap_rank(uint32_t rank, uint32_t size) {
    uint32_t nb = negabinary_rank(rank, size);
    uint32_t gray = nb ^ (nb >> 1);
    return bit_reverse32(gray) >> (32 - ceil_log2(size));
}

int remap_distance_doubling(uint32_t num) {
    int out = 0;
    while (num) {
        int k = 31 - __builtin_clz(num);
        out ^= 1 << k;
        num ^= (uint32_t)((1ull << (k + 1)) - 1);
    }
    return out;
}

bool Torus::make(const uint32_t* d, uint32_t n, Torus* out) {
    if (n < 1 || n > (uint32_t)kMaxDims) return false;
    Torus t;
    t.nd = n;
    t.size = 1;
    t.steps = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (d[i] < 1 || d[i] > 65536) return false;
        t.dims[i] = d[i];
        t.size *= d[i];
        t.steps_per_dim[i] = (uint32_t)ceil_log2(d[i]);
        t.steps += t.steps_per_dim[i];
    }
    if (t.steps > (uint32_t)kMaxSteps || t.size > (1u << 20)) return false;
    *out = t;
    return true;
}

void Torus::coords(uint32_t id, int* c) const {
    for (int i = (int)nd - 1; i >= 0; --i) {
        c[i] = (int)(id % dims[i]);
        id /= dims[i];
    }
}

uint32_t Torus::id(const int* c) const {
    uint32_t r = 0;
    for (uint32_t i = 0; i < nd; ++i) r = r * dims[i] + (uint32_t)pmod(c[i], (int)dims[i]);
    return r;
}

bool mirrored_port(int port, uint32_t nd) {
    if (nd == 3) return port >= 3;
    if (nd == 2) return port == 2 || port == 3;
    if (nd == 1) return port & 1;
    return false;
}

std::vector<StepDim> step_dims(const Torus& t, uint32_t port) {
    // Round-robin over dimensions starting at port % nd, skipping a dimension
    // once its ceil(log2(dim)) steps are used.
    std::vector<StepDim> out(t.steps);
    uint32_t used[kMaxDims] = {0, 0, 0};
    uint32_t d = port % t.nd;
    for (uint32_t i = 0; i < t.steps; ++i) {
        out[i] = {d, used[d]};
        if (used[d] < t.steps_per_dim[d]) ++used[d];
        if (i + 1 == t.steps) break;
        do {
            d = (d + 1) % t.nd;
        } while (used[d] >= t.steps_per_dim[d]);
    }
    return out;
}

static int step_distance(int coord, uint32_t k, uint32_t port, uint32_t nd, int family) {
    int dist;
    if (family == FAM_RECDOUB) {
        dist = (coord ^ (1 << k)) - coord;
    } else {
        dist = jacobsthal((int)k);
        if (coord & 1) dist = -dist;
    }
    if (mirrored_port((int)port, nd)) dist = -dist;
    return dist;
}

uint32_t peer_of(const Torus& t, const int* coord, uint32_t step, uint32_t port, const std::vector<StepDim>& sd,
                 int family) {
    int c[kMaxDims] = {0, 0, 0};
    for (uint32_t i = 0; i < t.nd; ++i) c[i] = coord[i];
    const StepDim s = sd[step];
    c[s.d] = pmod(c[s.d] + step_distance(coord[s.d], s.k, port, t.nd, family), (int)t.dims[s.d]);
    return t.id(c);
}

std::vector<uint32_t> peer_sequence(const Torus& t, uint32_t rank, uint32_t port, int family) {
    // Same walk as step_dims, but a dimension whose steps are exhausted is
    // only discovered (and skipped) when it comes up, as compute_peers does.
    std::vector<uint32_t> out(t.steps);
    int coord[kMaxDims];
    t.coords(rank, coord);
    if (t.nd == 1) {
        for (uint32_t i = 0; i < t.steps; ++i) {
            int c = pmod((int)rank + step_distance((int)rank, i, port, 1, family), (int)t.dims[0]);
            out[i] = (uint32_t)c;
        }
        return out;
    }
    bool done[kMaxDims] = {false, false, false};
    uint32_t next[kMaxDims] = {0, 0, 0};
    int last = (int)port - 1;
    for (uint32_t i = 0; i < t.steps;) {
        int o = 0, d;
        do {
            d = (last + 1 + o) % (int)t.nd;
            ++o;
        } while (done[d]);
        uint32_t k = next[d]++;
        last = d;
        if (k < t.steps_per_dim[d]) {
            int c[kMaxDims];
            for (uint32_t j = 0; j < t.nd; ++j) c[j] = coord[j];
            c[d] = pmod(c[d] + step_distance(coord[d], k, port, t.nd, family), (int)t.dims[d]);
            out[i++] = t.id(c);
        } else {
            done[d] = true;
        }
    }
    return out;
}

void BlockSchedule::build(const Torus& t, uint32_t prt, int fam) {
    P = t.size;
    L = t.steps;
    port = prt;
    family = fam;
    peers.assign((size_t)P * L, 0);
    send_step.assign((size_t)P * P, (uint8_t)L);
    const std::vector<StepDim> sd = step_dims(t, prt);
    // next_hop[node*L + s] = peer of `node` at step s
    std::vector<uint32_t> hop((size_t)P * L);
    for (uint32_t r = 0; r < P; ++r) {
        int c[kMaxDims];
        t.coords(r, c);
        for (uint32_t s = 0; s < L; ++s) hop[(size_t)r * L + s] = peer_of(t, c, s, prt, sd, fam);
        std::vector<uint32_t> seq = peer_sequence(t, r, prt, fam);
        for (uint32_t s = 0; s < L; ++s) peers[(size_t)r * L + s] = seq[s];
    }
    // For rank r and starting step i (largest first), every node reachable
    // from hop(r, i) through strictly later steps, taken in increasing order,
    // is forwarded at step i unless a later start already claimed it.
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    for (uint32_t r = 0; r < P; ++r) {
        uint8_t* row = &send_step[(size_t)r * P];
        for (int i = (int)L - 1; i >= 0; --i) {
            stack.clear();
            stack.push_back({hop[(size_t)r * L + i], (uint32_t)i});
            while (!stack.empty()) {
                auto [node, s] = stack.back();
                stack.pop_back();
                if (row[node] == (uint8_t)L) row[node] = (uint8_t)i;
                for (uint32_t j = s + 1; j < L; ++j) stack.push_back({hop[(size_t)node * L + j], j});
            }
        }
        row[r] = (uint8_t)L;
    }
}

// Tree construction follows get_tree: a node's parent is the first visitor
// (in depth-first order over steps) reaching it at the smallest step; ranks
// are then renumbered in post-order from P-1 down so every subtree covers a
// contiguous range and the root gets 0.
namespace {
struct TreeBuilder {
    const Torus& t;
    uint32_t port;
    int family, dist;
    std::vector<StepDim> sd;
    Tree* tr;
    uint32_t real_step(uint32_t i) const { return dist == DIST_DECREASING ? t.steps - 1 - i : i; }
    uint32_t step_peer(uint32_t node, uint32_t i) const {
        int c[kMaxDims];
        t.coords(node, c);
        return peer_of(t, c, real_step(i), port, sd, family);
    }
    void grow(uint32_t node, uint32_t step, uint32_t subroot) {
        for (uint32_t i = step; i < t.steps; ++i) {
            uint32_t p = step_peer(node, i);
            if (tr->parent[p] == UINT32_MAX || i < tr->reached_at_step[p]) {
                tr->parent[p] = node;
                tr->reached_at_step[p] = i;
                tr->subtree_root[p] = step == 0 ? p : subroot;
            }
            grow(p, i + 1, step == 0 ? p : subroot);
        }
    }
    void number(uint32_t node, uint32_t step, uint32_t* next) {
        tr->remapped_max[node] = *next;
        for (uint32_t i = step; i < t.steps; ++i) {
            uint32_t p = step_peer(node, i);
            if (tr->parent[p] == node) number(p, i + 1, next);
        }
        tr->remapped[node] = (*next)--;
    }
};
}  // namespace

bool build_tree(const Torus& t, uint32_t root, uint32_t port, int family, int dist, Tree* out) {
    const uint32_t P = t.size;
    out->parent.assign(P, UINT32_MAX);
    out->reached_at_step.assign(P, t.steps);
    out->remapped.assign(P, 0);
    out->remapped_max.assign(P, 0);
    out->subtree_root.assign(P, 0);
    TreeBuilder b{t, port, family, dist, step_dims(t, port), out};
    b.grow(root, 0, 0);
    out->parent[root] = UINT32_MAX;
    out->reached_at_step[root] = 0;
    out->subtree_root[root] = UINT32_MAX;
    uint32_t next = P - 1;
    b.number(root, 0, &next);
    return next == UINT32_MAX;
}

uint32_t cont_remapped_rank(const Torus& t, uint32_t rank, uint32_t port, int family) {
    // Post-order position of `rank` among all depth-first visits from rank 0.
    const std::vector<StepDim> sd = step_dims(t, port);
    uint32_t remap = t.size - 1;
    bool found = false;
    struct Frame {
        uint32_t node, step, next_i;
    };
    std::vector<Frame> st;
    st.push_back({0, 0, 0});
    while (!st.empty()) {
        Frame& f = st.back();
        if (f.next_i < f.step) f.next_i = f.step;
        if (f.next_i < t.steps) {
            int c[kMaxDims];
            t.coords(f.node, c);
            uint32_t i = f.next_i++;
            uint32_t p = peer_of(t, c, i, port, sd, family);
            st.push_back({p, i + 1, i + 1});
            continue;
        }
        if (found) --remap;
        if (f.node == rank) found = true;
        st.pop_back();
    }
    return remap;
}

}  // namespace bine
