/*
 * bine_oracle.c -- CPU oracle (test infrastructure only; see bine_oracle.h).
 *
 * Plain-C restatement of the reference's MPI-layer Bine collectives over P
 * virtual ranks in host memory.  Point-to-point messages are simulated
 * step-synchronously: within one step every posted send is snapshot first and
 * then delivered to the matching receive (same source, destination and port,
 * in posting order), which is what MPI's non-overtaking rule gives for these
 * schedules.  Reductions use the reference's operand order
 * (MPI_Reduce_local(in, inout): inout = in (+) inout).
 */
#include "bine_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* scalar helpers                                                            */
/* ------------------------------------------------------------------------ */

/* lib/libbine_common.h:27 (rhos), extended by the closed form (1-(-2)^(k+1))/3 */
static const int RHOS[20] = {1, -1, 3, -5, 11, -21, 43, -85, 171, -341, 683, -1365,
                             2731, -5461, 10923, -21845, 43691, -87381, 174763, -349525};
/* lib/libbine_common.h:28-29 */
static const int SMALLEST_NB[20] = {0, 0, -2, -2, -10, -10, -42, -42, -170, -170, -682, -682,
                                    -2730, -2730, -10922, -10922, -43690, -43690, -174762, -174762};
static const int LARGEST_NB[20] = {0, 1, 1, 5, 5, 21, 21, 85, 85, 341, 341, 1365,
                                   1365, 5461, 5461, 21845, 21845, 87381, 87381, 349525};

size_t or_dtsize(int dt) {
    switch (dt) {
    case OR_INT32: return 4;
    case OR_INT64: return 8;
    case OR_FLOAT: return 4;
    case OR_DOUBLE: return 8;
    case OR_INT8: return 1;
    default: return 0;
    }
}

#ifdef _OPENMP
#include <omp.h>
#endif
int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int or_rho(int k) { return (k >= 0 && k < 20) ? RHOS[k] : 0; }

/* lib/libbine_coll.cc:65-87 */
int or_ceil_log2(unsigned long long x) {
    static const unsigned long long t[6] = {0xFFFFFFFF00000000ull, 0x00000000FFFF0000ull,
                                            0x000000000000FF00ull, 0x00000000000000F0ull,
                                            0x000000000000000Cull, 0x0000000000000002ull};
    int y = (((x & (x - 1)) == 0) ? 0 : 1);
    int j = 32;
    for (int i = 0; i < 6; i++) {
        int k = (((x & t[i]) == 0) ? 0 : j);
        y += k;
        x >>= k;
        j >>= 1;
    }
    return y;
}

/* lib/libbine_coll.cc:59-62 */
int or_mod(int a, int b) {
    int r = a % b;
    return r < 0 ? r + b : r;
}

/* lib/libbine_coll.cc:55-57 */
int or_is_power_of_two(int x) { return (x != 0) && ((x & (x - 1)) == 0); }

/* lib/libbine_coll_bcast.cc:622-626 */
uint32_t or_btonb(int32_t bin) {
    const uint32_t mask = 0xAAAAAAAAu;
    return (mask + (uint32_t)bin) ^ mask;
}

/* lib/libbine_coll_bcast.cc:628-633 */
int32_t or_nbtob(uint32_t neg) {
    const uint32_t mask = 0xAAAAAAAAu;
    return (int32_t)((mask ^ neg) - mask);
}

/* lib/libbine_coll_alltoall.cc:325-332 */
uint32_t or_reverse(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    x = ((x >> 16) & 0xffffu) | ((x & 0xffffu) << 16);
    return x;
}

/* lib/libbine_coll_alltoall.cc:335-337 */
int or_in_range(int x, uint32_t nbits) {
    if (nbits >= 20) return 0;
    return x >= SMALLEST_NB[nbits] && x <= LARGEST_NB[nbits];
}

/* lib/libbine_coll_alltoall.cc:339-369, including the decimal 80000000 MSB test */
uint32_t or_rank_negabinary(uint32_t rank, uint32_t size) {
    uint32_t nba = UINT32_MAX, nbb = UINT32_MAX;
    uint32_t num_bits = (uint32_t)or_ceil_log2(size);
    int r = (int)rank, s = (int)size;
    if (rank % 2) {
        if (or_in_range(r, num_bits)) nba = or_btonb(r);
        if (or_in_range(r - s, num_bits)) nbb = or_btonb(r - s);
    } else {
        if (or_in_range(-r, num_bits)) nba = or_btonb(-r);
        if (or_in_range(-r + s, num_bits)) nbb = or_btonb(-r + s);
    }
    if (nba == UINT32_MAX && nbb != UINT32_MAX) return nbb;
    if (nba != UINT32_MAX && nbb == UINT32_MAX) return nba;
    if (nba & (80000000u >> (32 - num_bits))) return nba;
    return nbb;
}

/* lib/libbine_coll_alltoall.cc:372-378 */
uint32_t or_remap_rank(uint32_t rank, uint32_t size) {
    uint32_t v = or_rank_negabinary(rank, size);
    v = v ^ (v >> 1);
    uint32_t num_bits = (uint32_t)or_ceil_log2(size);
    return or_reverse(v) >> (32 - num_bits);
}

/* lib/libbine_coll_allgather.cc:1121-1123 */
static uint32_t nb_to_nu(uint32_t nb, uint32_t size) {
    return or_reverse(nb ^ (nb >> 1)) >> (32 - or_ceil_log2(size));
}

/* lib/libbine_coll_allgather.cc:1125-1158 (smaller of the two candidates) */
uint32_t or_get_nu(uint32_t rank, uint32_t size) {
    uint32_t nba = UINT32_MAX, nbb = UINT32_MAX;
    uint32_t num_bits = (uint32_t)or_ceil_log2(size);
    int r = (int)rank, s = (int)size;
    if (rank % 2) {
        if (or_in_range(r, num_bits)) nba = or_btonb(r);
        if (or_in_range(r - s, num_bits)) nbb = or_btonb(r - s);
    } else {
        if (or_in_range(-r, num_bits)) nba = or_btonb(-r);
        if (or_in_range(-r + s, num_bits)) nbb = or_btonb(-r + s);
    }
    if (nba == UINT32_MAX && nbb != UINT32_MAX) return nb_to_nu(nbb, size);
    if (nba != UINT32_MAX && nbb == UINT32_MAX) return nb_to_nu(nba, size);
    int nu_a = (int)nb_to_nu(nba, size), nu_b = (int)nb_to_nu(nbb, size);
    return (uint32_t)(nu_a < nu_b ? nu_a : nu_b);
}

/* lib/libbine_coll_alltoall.cc:381-383, 401-409 */
int or_remap_distance_doubling(uint32_t num) {
    int remapped = 0;
    while (num > 0) {
        int k = 31 - __builtin_clz(num);
        remapped ^= (0x1 << k);
        num ^= (uint32_t)((1ULL << (k + 1)) - 1);
    }
    return remapped;
}

/* ------------------------------------------------------------------------ */
/* torus and schedule                                                        */
/* ------------------------------------------------------------------------ */

/* BineCoordConverter ctor (lib/libbine_common.cc:248-259) */
int or_torus_init(or_torus_t* t, const uint32_t* dims, uint32_t ndims) {
    if (ndims < 1 || ndims > OR_MAX_DIMS) return OR_ERR_ARG;
    memset(t, 0, sizeof(*t));
    t->ndims = ndims;
    t->size = 1;
    for (uint32_t d = 0; d < ndims; d++) {
        if (dims[d] < 1) return OR_ERR_ARG;
        t->dims[d] = dims[d];
        t->size *= dims[d];
        t->steps_per_dim[d] = (uint32_t)or_ceil_log2(dims[d]);
        t->num_steps += t->steps_per_dim[d];
    }
    return OR_OK;
}

/* lib/libbine_common.cc:206-216 (row-major, last dimension fastest) */
void or_coord_from_id(const or_torus_t* t, int id, int* coord) {
    int nnodes = (int)t->size;
    for (uint32_t i = 0; i < t->ndims; i++) {
        nnodes = nnodes / (int)t->dims[i];
        coord[i] = id / nnodes;
        id = id % nnodes;
    }
}

/* lib/libbine_common.cc:219-238 (periodic wrap) */
int or_id_from_coord(const or_torus_t* t, const int* coords) {
    int rank = 0, multiplier = 1;
    for (int i = (int)t->ndims - 1; i >= 0; i--) {
        int coord = coords[i];
        int dim = (int)t->dims[i];
        if (coord >= dim) {
            coord = coord % dim;
        } else if (coord < 0) {
            coord = coord % dim;
            if (coord) coord = dim + coord;
        }
        rank += multiplier * coord;
        multiplier *= dim;
    }
    return rank;
}

/* lib/libbine_coll.cc:15-31 */
int or_is_mirroring_port(int port, uint32_t ndims) {
    if (ndims == 3) return port >= (int)ndims;
    if (ndims == 2) {
        if (port == 0 || port == 1) return 0;
        if (port == 2 || port == 3) return 1;
        return 0;
    }
    if (ndims == 1) return port % 2;
    return 0;
}

/* lib/libbine_coll.cc:33-42 */
static int distance_sign(int coord, int port, uint32_t ndims) {
    int m = 1;
    if (coord & 1) m = -m;
    if (or_is_mirroring_port(port, ndims)) m = -m;
    return m;
}

/* lib/libbine_coll.cc:306-331 */
void or_step_info(const or_torus_t* t, uint32_t port, uint32_t* sd, uint32_t* ssd) {
    size_t next_step_per_dim[OR_MAX_DIMS] = {0, 0, 0};
    size_t current_d = port % t->ndims;
    for (size_t i = 0; i < t->num_steps; i++) {
        sd[i] = (uint32_t)current_d;
        ssd[i] = (uint32_t)next_step_per_dim[current_d];
        size_t d = current_d;
        if (next_step_per_dim[d] < (size_t)or_ceil_log2(t->dims[d])) next_step_per_dim[d] += 1;
        if (i != t->num_steps - 1) {
            do {
                current_d = (current_d + 1) % t->ndims;
                d = current_d;
            } while (next_step_per_dim[d] >= (size_t)or_ceil_log2(t->dims[d]));
        }
    }
}

/* lib/libbine_coll.cc:89-107 */
void or_get_peer_c(const or_torus_t* t, const int* coord, uint32_t step, uint32_t port,
                   const uint32_t* sd, const uint32_t* ssd, int family, int* coord_peer) {
    memcpy(coord_peer, coord, sizeof(int) * t->ndims);
    uint32_t d = sd[step];
    uint32_t step_in_d = ssd[step];
    int distance;
    if (family == OR_FAMILY_RECDOUB) {
        distance = (coord_peer[d] ^ (1 << step_in_d)) - coord_peer[d];
        if (or_is_mirroring_port((int)port, t->ndims)) distance = -distance;
    } else {
        distance = or_rho((int)step_in_d) * distance_sign(coord[d], (int)port, t->ndims);
    }
    coord_peer[d] = or_mod(coord_peer[d] + distance, (int)t->dims[d]);
}

/* lib/libbine_coll.cc:443-504 */
void or_compute_peers(const or_torus_t* t, uint32_t rank, int port, int family, uint32_t* peers) {
    int terminated[OR_MAX_DIMS] = {0, 0, 0};
    uint8_t next_step_per_dim[OR_MAX_DIMS] = {0, 0, 0};
    int coord[OR_MAX_DIMS];
    or_coord_from_id(t, (int)rank, coord);
    int target_dim, relative_step, distance, last_dim = port - 1;
    for (uint32_t i = 0; i < t->num_steps;) {
        if (t->ndims > 1) {
            or_coord_from_id(t, (int)rank, coord);
            int o = 0;
            do {
                target_dim = (last_dim + 1 + o) % (int)t->ndims;
                o++;
            } while (terminated[target_dim]);
            relative_step = next_step_per_dim[target_dim];
            ++next_step_per_dim[target_dim];
            last_dim = target_dim;
        } else {
            target_dim = 0;
            relative_step = (int)i;
            coord[0] = (int)rank;
        }
        if (family == OR_FAMILY_RECDOUB) {
            distance = (coord[target_dim] ^ (1 << relative_step)) - coord[target_dim];
        } else {
            distance = or_rho(relative_step);
            if (coord[target_dim] & 1) distance = -distance;
        }
        if (or_is_mirroring_port(port, t->ndims)) distance = -distance;
        if ((uint32_t)relative_step < t->steps_per_dim[target_dim]) {
            coord[target_dim] = or_mod(coord[target_dim] + distance, (int)t->dims[target_dim]);
            peers[i] = (t->ndims > 1) ? (uint32_t)or_id_from_coord(t, coord) : (uint32_t)coord[0];
            i += 1;
        } else {
            terminated[target_dim] = 1;
        }
    }
}

typedef struct {
    const or_torus_t* t;
    uint32_t port;
    int family;
    const uint32_t* sd;
    const uint32_t* ssd;
} sched_ctx_t;

/* lib/libbine_common.cc:1491-1510 */
static void compute_block_step(const sched_ctx_t* c, const int* coord_rank, size_t starting_step, size_t step,
                               size_t num_steps, uint32_t* block_step) {
    if (step < num_steps) {
        for (size_t i = step + 1; i < num_steps; i++) {
            int peer[OR_MAX_DIMS];
            or_get_peer_c(c->t, coord_rank, (uint32_t)i, c->port, c->sd, c->ssd, c->family, peer);
            compute_block_step(c, peer, starting_step, i, num_steps, block_step);
        }
        int rank = or_id_from_coord(c->t, coord_rank);
        if (block_step[rank] == num_steps || starting_step > block_step[rank]) block_step[rank] = (uint32_t)starting_step;
    }
}

/* BineBitmapCalculator ctor, non-remapped branch (lib/libbine_common.cc:2036-2045) */
void or_block_step(const or_torus_t* t, uint32_t rank, uint32_t port, int family, uint32_t* block_step) {
    uint32_t L = t->num_steps;
    uint32_t sd[OR_MAX_STEPS], ssd[OR_MAX_STEPS];
    or_step_info(t, port, sd, ssd);
    sched_ctx_t c = {t, port, family, sd, ssd};
    int coord_mine[OR_MAX_DIMS];
    or_coord_from_id(t, (int)rank, coord_mine);
    for (uint32_t i = 0; i < t->size; i++) block_step[i] = L;
    for (uint32_t i = 0; i < L; i++) {
        int peer[OR_MAX_DIMS];
        or_get_peer_c(t, coord_mine, i, port, sd, ssd, family, peer);
        compute_block_step(&c, peer, i, i, L, block_step);
    }
    block_step[rank] = L;
}

/* compute_next_bitmaps, non-remapped branch (lib/libbine_common.cc:1552-1579) */
void or_step_bitmaps(const or_torus_t* t, uint32_t rank, uint32_t port, int family, uint32_t step,
                     uint8_t* send_bitmap, uint8_t* recv_bitmap) {
    uint32_t P = t->size, L = t->num_steps;
    uint32_t* bs = (uint32_t*)malloc(sizeof(uint32_t) * P);
    or_block_step(t, rank, port, family, bs);
    for (uint32_t i = 0; i < P; i++) send_bitmap[i] = (bs[i] == step);
    uint32_t peers[OR_MAX_STEPS];
    or_compute_peers(t, rank, (int)port, family, peers);
    uint32_t sd[OR_MAX_STEPS], ssd[OR_MAX_STEPS];
    or_step_info(t, port, sd, ssd);
    sched_ctx_t c = {t, port, family, sd, ssd};
    for (uint32_t i = 0; i < P; i++) bs[i] = L;
    uint32_t peer = peers[step];
    int peer_coord[OR_MAX_DIMS];
    or_coord_from_id(t, (int)peer, peer_coord);
    for (uint32_t j = step; j < L; j++) {
        int ppr[OR_MAX_DIMS];
        or_get_peer_c(t, peer_coord, j, port, sd, ssd, family, ppr);
        compute_block_step(&c, ppr, j, j, L, bs);
    }
    bs[peer] = L;
    for (uint32_t i = 0; i < P; i++) recv_bitmap[i] = (bs[i] == step);
    free(bs);
}

/* build_tree (lib/libbine_coll.cc:122-151) */
static void build_tree(const sched_ctx_t* c, const int* coord_root, size_t step, int dist, uint32_t* reached,
                       uint32_t subtree_root, uint32_t* parent, uint32_t* subtree_roots) {
    size_t L = c->t->num_steps;
    for (size_t i = step; i < L; i++) {
        int peer[OR_MAX_DIMS];
        size_t real_step = (dist == OR_DIST_DECREASING) ? L - 1 - i : i;
        or_get_peer_c(c->t, coord_root, (uint32_t)real_step, c->port, c->sd, c->ssd, c->family, peer);
        uint32_t rank = (uint32_t)or_id_from_coord(c->t, peer);
        if (parent[rank] == UINT32_MAX || i < reached[rank]) {
            parent[rank] = (uint32_t)or_id_from_coord(c->t, coord_root);
            reached[rank] = (uint32_t)i;
            subtree_roots[rank] = (step == 0) ? rank : subtree_root;
        }
        uint32_t actual_subtree_root = (step == 0) ? rank : subtree_root;
        build_tree(c, peer, i + 1, dist, reached, actual_subtree_root, parent, subtree_roots);
    }
}

/* remap_ranks (lib/libbine_coll.cc:282-304) */
static void remap_ranks(const sched_ctx_t* c, const int* coord_root, size_t step, int dist, uint32_t* next_rank,
                        const uint32_t* parent, uint32_t* remapped, uint32_t* remapped_max) {
    size_t L = c->t->num_steps;
    uint32_t me = (uint32_t)or_id_from_coord(c->t, coord_root);
    remapped_max[me] = *next_rank;
    for (size_t i = step; i < L; i++) {
        int peer[OR_MAX_DIMS];
        size_t real_step = (dist == OR_DIST_DECREASING) ? L - 1 - i : i;
        or_get_peer_c(c->t, coord_root, (uint32_t)real_step, c->port, c->sd, c->ssd, c->family, peer);
        if (parent[or_id_from_coord(c->t, peer)] == me)
            remap_ranks(c, peer, i + 1, dist, next_rank, parent, remapped, remapped_max);
    }
    remapped[me] = *next_rank;
    (*next_rank)--;
}

/* get_tree (lib/libbine_coll.cc:334-389) */
int or_get_tree(const or_torus_t* t, uint32_t root, uint32_t port, int family, int dist, uint32_t* parent,
                uint32_t* reached, uint32_t* remapped, uint32_t* remapped_max, uint32_t* subtree_roots) {
    uint32_t sd[OR_MAX_STEPS], ssd[OR_MAX_STEPS];
    or_step_info(t, port, sd, ssd);
    sched_ctx_t c = {t, port, family, sd, ssd};
    int coord_root[OR_MAX_DIMS];
    or_coord_from_id(t, (int)root, coord_root);
    for (uint32_t i = 0; i < t->size; i++) {
        parent[i] = UINT32_MAX;
        reached[i] = t->num_steps;
        subtree_roots[i] = 0;
        remapped[i] = 0;
        remapped_max[i] = 0;
    }
    build_tree(&c, coord_root, 0, dist, reached, 0, parent, subtree_roots);
    parent[root] = UINT32_MAX;
    reached[root] = 0;
    subtree_roots[root] = UINT32_MAX;
    uint32_t next_rank = t->size - 1;
    remap_ranks(&c, coord_root, 0, dist, &next_rank, parent, remapped, remapped_max);
    return next_rank == UINT32_MAX ? OR_OK : OR_ERR_MATCH;
}

/* dfs (lib/libbine_common.cc:1981-1993) */
static void cont_dfs(const sched_ctx_t* c, const int* coord_rank, size_t step, const int* target,
                     uint32_t* remap, int* found) {
    for (size_t i = step; i < c->t->num_steps; i++) {
        int peer[OR_MAX_DIMS];
        or_get_peer_c(c->t, coord_rank, (uint32_t)i, c->port, c->sd, c->ssd, c->family, peer);
        cont_dfs(c, peer, i + 1, target, remap, found);
    }
    if (*found) (*remap)--;
    if (memcmp(coord_rank, target, sizeof(int) * c->t->ndims) == 0) *found = 1;
}

/* lib/libbine_common.cc:2023-2034 */
uint32_t or_cont_remapped_rank(const or_torus_t* t, uint32_t rank, uint32_t port, int family) {
    uint32_t sd[OR_MAX_STEPS], ssd[OR_MAX_STEPS];
    or_step_info(t, port, sd, ssd);
    sched_ctx_t c = {t, port, family, sd, ssd};
    uint32_t remap = t->size - 1;
    int c0[OR_MAX_DIMS], me[OR_MAX_DIMS];
    or_coord_from_id(t, 0, c0);
    or_coord_from_id(t, (int)rank, me);
    int found = 0;
    cont_dfs(&c, c0, 0, me, &remap, &found);
    return remap;
}

/* ------------------------------------------------------------------------ */
/* local reductions                                                          */
/* ------------------------------------------------------------------------ */

/* MPI_Reduce_local(in, inout): inout = in + inout (SUM); VOID leaves inout.
 * lib/libbine_coll.cc:153-207 for the types it covers; int64/double follow MPI. */
static void reduce_local(int dt, int op, const void* in, void* inout, size_t n) {
    if (op != OR_SUM) return;
    switch (dt) {
    case OR_INT32: {
        const int32_t* a = (const int32_t*)in;
        int32_t* b = (int32_t*)inout;
        for (size_t i = 0; i < n; i++) b[i] = (int32_t)((uint32_t)a[i] + (uint32_t)b[i]);
        break;
    }
    case OR_INT64: {
        const int64_t* a = (const int64_t*)in;
        int64_t* b = (int64_t*)inout;
        for (size_t i = 0; i < n; i++) b[i] = (int64_t)((uint64_t)a[i] + (uint64_t)b[i]);
        break;
    }
    case OR_FLOAT: {
        const float* a = (const float*)in;
        float* b = (float*)inout;
        for (size_t i = 0; i < n; i++) b[i] = a[i] + b[i];
        break;
    }
    case OR_DOUBLE: {
        const double* a = (const double*)in;
        double* b = (double*)inout;
        for (size_t i = 0; i < n; i++) b[i] = a[i] + b[i];
        break;
    }
    case OR_INT8: {
        const int8_t* a = (const int8_t*)in;
        int8_t* b = (int8_t*)inout;
        for (size_t i = 0; i < n; i++) b[i] = (int8_t)((uint8_t)a[i] + (uint8_t)b[i]);
        break;
    }
    }
}

/* 3-argument reduce_local: out = a + b (lib/libbine_coll.cc:209-267); VOID leaves out untouched */
static void reduce_local3(int dt, int op, const void* a, const void* b, void* out, size_t n) {
    if (op != OR_SUM) return;
    size_t ds = or_dtsize(dt);
    uint8_t* tmp = (uint8_t*)malloc(n * ds + 1);
    memcpy(tmp, b, n * ds);
    reduce_local(dt, op, a, tmp, n);
    memcpy(out, tmp, n * ds);
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* step-synchronous message engine                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int src, dst, tag;
    const uint8_t* ptr;
    size_t bytes;
} msg_t;

typedef struct {
    msg_t* sends;
    size_t ns, cs;
    msg_t* recvs; /* src = peer we receive from, dst = receiver */
    size_t nr, cr;
} step_t;

static void step_init(step_t* s) { memset(s, 0, sizeof(*s)); }
static void step_free(step_t* s) {
    free(s->sends);
    free(s->recvs);
}
static void post_send(step_t* s, int src, int dst, int tag, const void* p, size_t bytes) {
    if (s->ns == s->cs) {
        s->cs = s->cs ? s->cs * 2 : 64;
        s->sends = (msg_t*)realloc(s->sends, sizeof(msg_t) * s->cs);
    }
    msg_t m = {src, dst, tag, (const uint8_t*)p, bytes};
    s->sends[s->ns++] = m;
}
static void post_recv(step_t* s, int me, int from, int tag, void* p, size_t bytes) {
    if (s->nr == s->cr) {
        s->cr = s->cr ? s->cr * 2 : 64;
        s->recvs = (msg_t*)realloc(s->recvs, sizeof(msg_t) * s->cr);
    }
    msg_t m = {from, me, tag, (const uint8_t*)p, bytes};
    s->recvs[s->nr++] = m;
}

/* Snapshot every send, then match the k-th send (src->dst, tag) with the k-th
 * receive posted by dst for (src, tag).  A shorter message into a larger
 * receive is legal MPI; a longer one is a truncation error. */
static const msg_t* g_sort_base;
static int msg_key_cmp(const void* a, const void* b) {
    const msg_t* x = &g_sort_base[*(const size_t*)a];
    const msg_t* y = &g_sort_base[*(const size_t*)b];
    if (x->src != y->src) return x->src < y->src ? -1 : 1;
    if (x->dst != y->dst) return x->dst < y->dst ? -1 : 1;
    if (x->tag != y->tag) return x->tag < y->tag ? -1 : 1;
    size_t i = *(const size_t*)a, j = *(const size_t*)b; /* posting order */
    return i < j ? -1 : (i > j);
}

/* message trace (or_trace_*) */
static int g_trace_on = 0;
static int32_t g_trace_step = 0;
static size_t g_trace_n = 0, g_trace_cap = 0;
static int32_t* g_trace_i32 = NULL; /* step, src, dst triples */
static uint64_t* g_trace_bytes = NULL;

void or_trace_start(void) {
    g_trace_on = 1;
    g_trace_step = 0;
    g_trace_n = 0;
}
size_t or_trace_stop(void) {
    g_trace_on = 0;
    return g_trace_n;
}
void or_trace_fetch(int32_t* step, int32_t* src, int32_t* dst, uint64_t* bytes) {
    for (size_t i = 0; i < g_trace_n; i++) {
        step[i] = g_trace_i32[3 * i];
        src[i] = g_trace_i32[3 * i + 1];
        dst[i] = g_trace_i32[3 * i + 2];
        bytes[i] = g_trace_bytes[i];
    }
}
static void trace_msg(int src, int dst, size_t bytes) {
    if (g_trace_n == g_trace_cap) {
        g_trace_cap = g_trace_cap ? 2 * g_trace_cap : 4096;
        g_trace_i32 = (int32_t*)realloc(g_trace_i32, sizeof(int32_t) * 3 * g_trace_cap);
        g_trace_bytes = (uint64_t*)realloc(g_trace_bytes, sizeof(uint64_t) * g_trace_cap);
    }
    g_trace_i32[3 * g_trace_n] = g_trace_step;
    g_trace_i32[3 * g_trace_n + 1] = src;
    g_trace_i32[3 * g_trace_n + 2] = dst;
    g_trace_bytes[g_trace_n++] = bytes;
}

uint32_t or_route_bytes(const uint32_t* dims, uint32_t ndims, size_t n, const int32_t* src, const int32_t* dst,
                        const uint64_t* bytes, uint64_t* link_bytes) {
    or_torus_t t;
    if (or_torus_init(&t, dims, ndims) != OR_OK) return 0;
    uint32_t longest = 0;
    for (size_t i = 0; i < n; i++) {
        int a[OR_MAX_DIMS], b[OR_MAX_DIMS];
        or_coord_from_id(&t, src[i], a);
        or_coord_from_id(&t, dst[i], b);
        uint32_t hops = 0;
        for (uint32_t d = 0; d < ndims; d++) {
            int dim = (int)dims[d];
            int fwd = or_mod(b[d] - a[d], dim); /* hops going the positive way */
            int neg = 2 * fwd > dim;
            int h = neg ? dim - fwd : fwd;
            for (int k = 0; k < h; k++) {
                size_t node = (size_t)or_id_from_coord(&t, a);
                link_bytes[(node * ndims + d) * 2 + (size_t)neg] += bytes[i];
                a[d] = or_mod(a[d] + (neg ? -1 : 1), dim);
            }
            hops += (uint32_t)h;
        }
        if (hops > longest) longest = hops;
    }
    return longest;
}

static int step_run(step_t* s) {
    size_t total = 0;
    for (size_t i = 0; i < s->ns; i++) total += s->sends[i].bytes;
    uint8_t* stage = (uint8_t*)malloc(total + 1);
    size_t* soff = (size_t*)malloc(sizeof(size_t) * (s->ns + 1));
    size_t o = 0;
    for (size_t i = 0; i < s->ns; i++) {
        soff[i] = o;
        o += s->sends[i].bytes;
    }
    /* snapshot every send (messages are independent: one thread each) */
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < (long)s->ns; i++) memcpy(stage + soff[i], s->sends[i].ptr, s->sends[i].bytes);
    int rc = (s->ns == s->nr) ? OR_OK : OR_ERR_MATCH;
    size_t* si = (size_t*)malloc(sizeof(size_t) * (s->ns + 1));
    size_t* ri = (size_t*)malloc(sizeof(size_t) * (s->nr + 1));
    for (size_t i = 0; i < s->ns; i++) si[i] = i;
    for (size_t i = 0; i < s->nr; i++) ri[i] = i;
    g_sort_base = s->sends;
    qsort(si, s->ns, sizeof(size_t), msg_key_cmp);
    g_sort_base = s->recvs;
    qsort(ri, s->nr, sizeof(size_t), msg_key_cmp);
    for (size_t k = 0; k < s->nr && rc == OR_OK; k++) {
        const msg_t* snd = &s->sends[si[k]];
        const msg_t* r = &s->recvs[ri[k]];
        if (snd->src != r->src || snd->dst != r->dst || snd->tag != r->tag || snd->bytes > r->bytes) {
            rc = OR_ERR_MATCH;
            break;
        }
        if (g_trace_on && snd->src != snd->dst) trace_msg(snd->src, snd->dst, snd->bytes);
    }
    /* deliver: every receive buffer is written by exactly one matched message */
    if (rc == OR_OK) {
#pragma omp parallel for schedule(dynamic, 16)
        for (long k = 0; k < (long)s->nr; k++)
            memcpy((void*)s->recvs[ri[k]].ptr, stage + soff[si[k]], s->sends[si[k]].bytes);
    }
    if (g_trace_on) g_trace_step++;
    free(si);
    free(ri);
    free(soff);
    free(stage);
    s->ns = s->nr = 0;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* allreduce                                                                 */
/* ------------------------------------------------------------------------ */

/* get_blocks_info: split by port, then by block (lib/libbine.cc:971-1010) */
static void blocks_info_allreduce(size_t count, uint32_t P, int ports, size_t ds, size_t* off, size_t* cnt) {
    size_t part = count / ports, rem = count % ports, so_far = 0;
    for (int p = 0; p < ports; p++) {
        size_t cp = part + ((size_t)p < rem ? 1 : 0);
        size_t op = so_far * ds;
        so_far += cp;
        size_t bpart = cp / P, brem = cp % P, bsf = 0;
        for (uint32_t i = 0; i < P; i++) {
            size_t cb = bpart + (i < brem ? 1 : 0);
            off[(size_t)p * P + i] = op + bsf * ds;
            cnt[(size_t)p * P + i] = cb;
            bsf += cb;
        }
    }
}

/* bine_coll_b for ALLREDUCE with the B / B_COALESCE step (lib/libbine_common.cc:1631-1735, 930-1159).
 * COALESCE merges consecutive blocks into one message; the bytes moved and the
 * arithmetic are the same, so both are restated by the per-block form. */
static int allreduce_b(const or_torus_t* t, int ports, int family, size_t count, int dt, int op,
                       const uint8_t* send, uint8_t* recv) {
    uint32_t P = t->size, L = t->num_steps;
    size_t ds = or_dtsize(dt), rb = count * ds;
    size_t* off = (size_t*)malloc(sizeof(size_t) * ports * P);
    size_t* cnt = (size_t*)malloc(sizeof(size_t) * ports * P);
    blocks_info_allreduce(count, P, ports, ds, off, cnt);
    uint32_t* peers = (uint32_t*)malloc(sizeof(uint32_t) * ports * P * L);
    for (int p = 0; p < ports; p++)
        for (uint32_t r = 0; r < P; r++) or_compute_peers(t, r, p, family, peers + ((size_t)p * P + r) * L);
    /* bitmaps[(p*P + r)*L + s][i] */
    uint8_t* sbm = (uint8_t*)malloc((size_t)ports * P * L * P);
    uint8_t* rbm = (uint8_t*)malloc((size_t)ports * P * L * P);
    for (int p = 0; p < ports; p++)
        for (uint32_t r = 0; r < P; r++)
            for (uint32_t s = 0; s < L; s++) {
                size_t idx = (((size_t)p * P + r) * L + s) * P;
                or_step_bitmaps(t, r, (uint32_t)p, family, s, sbm + idx, rbm + idx);
            }
    memcpy(recv, send, (size_t)P * rb);
    uint8_t* tmp = (uint8_t*)malloc((size_t)P * rb + 1);
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    for (int coll = 0; coll < 2 && rc == OR_OK; coll++) {
        for (uint32_t step = 0; step < L && rc == OR_OK; step++) {
            uint32_t bstep = (coll == 0) ? step : L - 1 - step;
            for (int p = 0; p < ports; p++) {
                for (uint32_t r = 0; r < P; r++) {
                    uint32_t peer = peers[((size_t)p * P + r) * L + bstep];
                    size_t idx = (((size_t)p * P + r) * L + bstep) * P;
                    const uint8_t* snd = (coll == 0) ? sbm + idx : rbm + idx;
                    const uint8_t* rcv = (coll == 0) ? rbm + idx : sbm + idx;
                    uint8_t* dst = (coll == 0) ? tmp : recv;
                    for (uint32_t i = 0; i < P; i++) {
                        size_t bo = off[(size_t)p * P + i], bc = cnt[(size_t)p * P + i];
                        if (snd[i]) post_send(&st, (int)r, (int)peer, p, recv + r * rb + bo, bc * ds);
                        if (rcv[i]) post_recv(&st, (int)r, (int)peer, p, dst + r * rb + bo, bc * ds);
                    }
                }
            }
            rc = step_run(&st);
            if (rc != OR_OK || coll == 1) continue;
            /* every (port, rank) reduces its own partition: independent */
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
            for (int p = 0; p < ports; p++)
                for (uint32_t r = 0; r < P; r++) {
                    size_t idx = (((size_t)p * P + r) * L + bstep) * P;
                    for (uint32_t i = 0; i < P; i++)
                        if (rbm[idx + i]) {
                            size_t bo = off[(size_t)p * P + i], bc = cnt[(size_t)p * P + i];
                            reduce_local(dt, op, tmp + r * rb + bo, recv + r * rb + bo, bc);
                        }
                }
        }
    }
    step_free(&st);
    free(tmp);
    free(sbm);
    free(rbm);
    free(peers);
    free(off);
    free(cnt);
    return rc;
}

/* bine_coll_b with the B_CONT step (lib/libbine_common.cc:1162-1237, chunk params :1524-1550, 1905-1915) */
static int allreduce_b_cont(const or_torus_t* t, int ports, int family, size_t count, int dt, int op,
                            const uint8_t* send, uint8_t* recv) {
    uint32_t P = t->size, L = t->num_steps;
    if (!or_is_power_of_two((int)P)) return OR_ERR_UNSUPPORTED;
    size_t ds = or_dtsize(dt), rb = count * ds;
    size_t* off = (size_t*)malloc(sizeof(size_t) * ports * P);
    size_t* cnt = (size_t*)malloc(sizeof(size_t) * ports * P);
    blocks_info_allreduce(count, P, ports, ds, off, cnt);
    /* chunk params per (port, rank, step): send_off, send_cnt, recv_off, recv_cnt */
    size_t* cp = (size_t*)malloc(sizeof(size_t) * 4 * ports * P * L);
    uint32_t* peers = (uint32_t*)malloc(sizeof(uint32_t) * ports * P * L);
    for (int p = 0; p < ports; p++)
        for (uint32_t r = 0; r < P; r++) {
            or_compute_peers(t, r, p, family, peers + ((size_t)p * P + r) * L);
            uint32_t rr = or_cont_remapped_rank(t, r, (uint32_t)p, family);
            size_t min_r = 0, max_r = P, min_s, max_s;
            for (uint32_t s = 0; s < L; s++) {
                min_s = min_r;
                max_s = max_r;
                size_t middle = (min_r + max_r + 1) / 2;
                if (rr < middle) {
                    min_s = middle;
                    max_r = middle;
                } else {
                    max_s = middle;
                    min_r = middle;
                }
                size_t* c = cp + 4 * (((size_t)p * P + r) * L + s);
                c[0] = off[(size_t)p * P + min_s];
                c[1] = 0;
                for (size_t i = min_s; i < max_s; i++) c[1] += cnt[(size_t)p * P + i];
                c[2] = off[(size_t)p * P + min_r];
                c[3] = 0;
                for (size_t i = min_r; i < max_r; i++) c[3] += cnt[(size_t)p * P + i];
            }
        }
    memcpy(recv, send, (size_t)P * rb);
    uint8_t* tmp = (uint8_t*)malloc((size_t)P * rb + 1);
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    for (int coll = 0; coll < 2 && rc == OR_OK; coll++) {
        for (uint32_t step = 0; step < L && rc == OR_OK; step++) {
            uint32_t bstep = (coll == 0) ? step : L - 1 - step;
            for (int p = 0; p < ports; p++)
                for (uint32_t r = 0; r < P; r++) {
                    uint32_t peer = peers[((size_t)p * P + r) * L + bstep];
                    const size_t* c = cp + 4 * (((size_t)p * P + r) * L + bstep);
                    size_t so = coll == 0 ? c[0] : c[2], sc = coll == 0 ? c[1] : c[3];
                    size_t ro = coll == 0 ? c[2] : c[0], rcn = coll == 0 ? c[3] : c[1];
                    post_send(&st, (int)r, (int)peer, p, recv + r * rb + so, sc * ds);
                    post_recv(&st, (int)r, (int)peer, p, (coll == 0 ? tmp : recv) + r * rb + ro, rcn * ds);
                }
            rc = step_run(&st);
            if (rc != OR_OK || coll == 1) continue;
            for (int p = 0; p < ports; p++)
                for (uint32_t r = 0; r < P; r++) {
                    const size_t* c = cp + 4 * (((size_t)p * P + r) * L + bstep);
                    reduce_local(dt, op, tmp + r * rb + c[2], recv + r * rb + c[2], c[3]);
                }
        }
    }
    step_free(&st);
    free(tmp);
    free(cp);
    free(peers);
    free(off);
    free(cnt);
    return rc;
}

/* bine_coll_l_mpi with shrink/enlarge (lib/libbine_common.cc:733-846, 269-353) */
static int allreduce_l(const or_torus_t* t, int ports, int family, size_t count, int dt, int op,
                       const uint8_t* send, uint8_t* recv) {
    uint32_t P = t->size, D = t->ndims;
    size_t ds = or_dtsize(dt), rb = count * ds;
    uint32_t vd[OR_MAX_DIMS];
    int all_p2 = 1;
    for (uint32_t d = 0; d < D; d++) {
        if (!or_is_power_of_two((int)t->dims[d])) {
            all_p2 = 0;
            vd[d] = 1u << (t->steps_per_dim[d] - 1);
        } else {
            vd[d] = t->dims[d];
        }
    }
    or_torus_t tv;
    or_torus_init(&tv, vd, D);
    uint8_t* tmp = (uint8_t*)malloc((size_t)P * rb + 1);
    int* idle = (int*)calloc(P, sizeof(int));
    int* first_copy_done = (int*)calloc(P, sizeof(int));
    int* rank_virtual = (int*)malloc(sizeof(int) * P);
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    for (uint32_t r = 0; r < P; r++) rank_virtual[r] = (int)r;
    if (!all_p2) {
        /* shrink, one dimension at a time; every rank walks dims in order and stops once idle */
        for (uint32_t i = 0; i < D && rc == OR_OK; i++) {
            if (or_is_power_of_two((int)t->dims[i])) continue;
            int extra = (int)t->dims[i] - (int)vd[i];
            int* role = (int*)calloc(P, sizeof(int)); /* 1 = sender, 2 = receiver */
            for (uint32_t r = 0; r < P; r++) {
                if (idle[r]) continue;
                int coord[OR_MAX_DIMS], cp[OR_MAX_DIMS];
                or_coord_from_id(t, (int)r, coord);
                memcpy(cp, coord, sizeof(coord));
                const uint8_t* rs = first_copy_done[r] ? recv + r * rb : send + r * rb;
                uint8_t* rr = first_copy_done[r] ? tmp + r * rb : recv + r * rb;
                if (coord[i] >= (int)vd[i]) {
                    cp[i] = coord[i] - extra;
                    post_send(&st, (int)r, or_id_from_coord(t, cp), 0, rs, rb);
                    role[r] = 1;
                } else if (coord[i] + extra >= (int)vd[i]) {
                    cp[i] = coord[i] + extra;
                    post_recv(&st, (int)r, or_id_from_coord(t, cp), 0, rr, rb);
                    role[r] = 2;
                }
            }
            rc = step_run(&st);
            for (uint32_t r = 0; r < P && rc == OR_OK; r++) {
                if (role[r] == 1) idle[r] = 1;
                if (role[r] == 2) {
                    const uint8_t* agg = first_copy_done[r] ? tmp + r * rb : send + r * rb;
                    reduce_local(dt, op, agg, recv + r * rb, count);
                    first_copy_done[r] = 1;
                }
            }
            free(role);
        }
        for (uint32_t r = 0; r < P; r++) {
            int coord[OR_MAX_DIMS];
            or_coord_from_id(t, (int)r, coord);
            rank_virtual[r] = or_id_from_coord(&tv, coord);
        }
    }
    uint32_t Lv = tv.num_steps;
    uint32_t* vpeers = (uint32_t*)malloc(sizeof(uint32_t) * ports * P * (Lv + 1));
    for (int p = 0; p < ports; p++)
        for (uint32_t r = 0; r < P; r++)
            if (!idle[r]) or_compute_peers(&tv, (uint32_t)rank_virtual[r], p, family, vpeers + ((size_t)p * P + r) * (Lv + 1));
    size_t part = count / ports, rem = count % ports;
    for (uint32_t step = 0; step < Lv && rc == OR_OK; step++) {
        for (uint32_t r = 0; r < P; r++) {
            if (idle[r]) continue;
            int s0 = (step == 0) && !first_copy_done[r];
            const uint8_t* bs = s0 ? send + r * rb : recv + r * rb;
            uint8_t* br = s0 ? recv + r * rb : tmp + r * rb;
            size_t so_far = 0;
            for (int p = 0; p < ports; p++) {
                int cpv[OR_MAX_DIMS];
                uint32_t vpeer = vpeers[((size_t)p * P + r) * (Lv + 1) + step];
                or_coord_from_id(&tv, (int)vpeer, cpv);
                int peer = or_id_from_coord(t, cpv);
                size_t cport = part + ((size_t)p < rem ? 1 : 0);
                size_t oport = so_far * ds;
                so_far += cport;
                post_send(&st, (int)r, peer, p, bs + oport, cport * ds);
                post_recv(&st, (int)r, peer, p, br + oport, cport * ds);
            }
        }
        rc = step_run(&st);
        for (uint32_t r = 0; r < P && rc == OR_OK; r++) {
            if (idle[r]) continue;
            int s0 = (step == 0) && !first_copy_done[r];
            const uint8_t* agg = s0 ? send + r * rb : tmp + r * rb;
            reduce_local(dt, op, agg, recv + r * rb, count);
        }
    }
    if (!all_p2 && rc == OR_OK) {
        /* enlarge, dimensions from last to first */
        for (int d = (int)D - 1; d >= 0 && rc == OR_OK; d--) {
            if (or_is_power_of_two((int)t->dims[d])) continue;
            int extra = (int)t->dims[d] - (int)vd[d];
            for (uint32_t r = 0; r < P; r++) {
                int coord[OR_MAX_DIMS], cp[OR_MAX_DIMS];
                or_coord_from_id(t, (int)r, coord);
                memcpy(cp, coord, sizeof(coord));
                if (coord[d] >= (int)vd[d]) {
                    cp[d] = coord[d] - extra;
                    post_recv(&st, (int)r, or_id_from_coord(t, cp), 0, recv + r * rb, rb);
                } else if (coord[d] + extra >= (int)vd[d]) {
                    cp[d] = coord[d] + extra;
                    post_send(&st, (int)r, or_id_from_coord(t, cp), 0, recv + r * rb, rb);
                }
            }
            rc = step_run(&st);
        }
    }
    step_free(&st);
    free(vpeers);
    free(rank_virtual);
    free(first_copy_done);
    free(idle);
    free(tmp);
    return rc;
}

/* MPI_Allreduce (lib/libbine.cc:1012-1062) */
int or_allreduce(const or_torus_t* t, int ports, int family, int algo, int count, int dt, int op,
                 const void* send, void* recv) {
    if (!or_dtsize(dt) || ports < 1 || ports > OR_MAX_PORTS || count < 0) return OR_ERR_ARG;
    if (family != OR_FAMILY_BINE && family != OR_FAMILY_RECDOUB) return OR_ERR_UNSUPPORTED;
    switch (algo) {
    case OR_AR_L:
        return allreduce_l(t, ports, family, (size_t)count, dt, op, (const uint8_t*)send, (uint8_t*)recv);
    case OR_AR_B:
    case OR_AR_B_COALESCE:
        if ((size_t)count < (size_t)ports * t->size) return OR_ERR_ARG; /* lib/libbine.cc:1036 */
        return allreduce_b(t, ports, family, (size_t)count, dt, op, (const uint8_t*)send, (uint8_t*)recv);
    case OR_AR_B_CONT:
        if ((size_t)count < (size_t)ports * t->size) return OR_ERR_ARG;
        return allreduce_b_cont(t, ports, family, (size_t)count, dt, op, (const uint8_t*)send, (uint8_t*)recv);
    default:
        return OR_ERR_UNSUPPORTED;
    }
}

/* ------------------------------------------------------------------------ */
/* reduce-scatter (contiguous, permute first)                                */
/* ------------------------------------------------------------------------ */

/* bine_reduce_scatter_mpi_contiguous (lib/libbine_coll_reduce_scatter.cc:660-758), single port
 * (the last-step write at recvbuf + offset_port only holds for port 0, :707,743). */
int or_reduce_scatter(const or_torus_t* t, int ports, int family, int dist, int recvcount, int dt, int op,
                      const void* sendv, void* recvv) {
    uint32_t P = t->size, L = t->num_steps;
    if (ports != 1) return OR_ERR_UNSUPPORTED;
    if (!or_is_power_of_two((int)P)) return OR_ERR_UNSUPPORTED; /* :764-770 */
    if ((size_t)recvcount * P < P) return OR_ERR_ARG;               /* lib/libbine.cc:1085 */
    size_t ds = or_dtsize(dt);
    if (!ds) return OR_ERR_ARG;
    const uint8_t* send = (const uint8_t*)sendv;
    uint8_t* recv = (uint8_t*)recvv;
    size_t sb = (size_t)recvcount * P * ds, bcnt = (size_t)recvcount, bb = bcnt * ds;
    uint8_t* tmp = (uint8_t*)malloc((size_t)P * sb + 1);
    uint32_t* tr = (uint32_t*)malloc(sizeof(uint32_t) * 5 * P);
    uint32_t* peers = (uint32_t*)malloc(sizeof(uint32_t) * P * L);
    int rc = OR_OK;
    for (uint32_t r = 0; r < P && rc == OR_OK; r++) {
        or_compute_peers(t, r, 0, family, peers + (size_t)r * L);
        rc = or_get_tree(t, r, 0, family, dist, tr, tr + P, tr + 2 * P, tr + 3 * P, tr + 4 * P);
        for (uint32_t i = 0; i < P; i++) memcpy(tmp + r * sb + bb * tr[2 * P + i], send + r * sb + i * bb, bb);
    }
    step_t st;
    step_init(&st);
    for (uint32_t step = 0; step < L && rc == OR_OK; step++) {
        size_t nblocks = P >> (step + 1);
        size_t cnt = nblocks * bcnt;
        for (uint32_t r = 0; r < P; r++) {
            uint32_t peer = (dist == OR_DIST_DECREASING) ? peers[(size_t)r * L + L - step - 1] : peers[(size_t)r * L + step];
            /* MPI_Sendrecv_replace on the upper half */
            post_send(&st, (int)r, (int)peer, 0, tmp + r * sb + cnt * ds, cnt * ds);
            post_recv(&st, (int)r, (int)peer, 0, tmp + r * sb + cnt * ds, cnt * ds);
        }
        rc = step_run(&st);
        for (uint32_t r = 0; r < P && rc == OR_OK; r++) {
            uint8_t* lo = tmp + r * sb;
            uint8_t* hi = lo + cnt * ds;
            if (step == L - 1) reduce_local3(dt, op, hi, lo, recv + r * bb, cnt);
            else reduce_local(dt, op, hi, lo, cnt);
        }
    }
    step_free(&st);
    free(peers);
    free(tr);
    free(tmp);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* allgather                                                                 */
/* ------------------------------------------------------------------------ */

/* variant 0: bine_allgather_mpi_contiguous (lib/libbine_coll_allgather.cc:289-374)
 * variant 1: bine_allgather_send_mpi (:492-574)
 * variant 2: bine_allgather_blocks_mpi (:1161-1228) */
int or_allgather(const or_torus_t* t, int ports, int family, int dist, int variant, int count, int dt,
                 const void* sendv, void* recvv) {
    uint32_t P = t->size, L = t->num_steps;
    size_t ds = or_dtsize(dt);
    if (!ds || count < 1) return OR_ERR_ARG;
    const uint8_t* send = (const uint8_t*)sendv;
    uint8_t* recv = (uint8_t*)recvv;
    size_t rcb = (size_t)count * P * ds; /* per-rank recv bytes */
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    if (variant == 0) {
        if (!or_is_power_of_two((int)P)) return OR_ERR_UNSUPPORTED;
        /* blocks: split by block, then by port (lib/libbine.cc:1185-1206) */
        size_t part = (size_t)count / ports, rem = (size_t)count % ports;
        size_t per_port_tmp = (size_t)ceil((float)count / ports) * ds * P; /* tmpbuf_size / ports */
        uint8_t* tmp = (uint8_t*)calloc((size_t)P * per_port_tmp * ports + 1, 1);
        uint32_t* peers = (uint32_t*)malloc(sizeof(uint32_t) * P * L);
        uint32_t* tr = (uint32_t*)malloc(sizeof(uint32_t) * 5 * P);
        for (int p = 0; p < ports && rc == OR_OK; p++) {
            size_t cport = part + ((size_t)p < rem ? 1 : 0);
            size_t pofs = 0;
            for (int q = 0; q < p; q++) pofs += (part + ((size_t)q < rem ? 1 : 0)) * ds;
            size_t c0 = cport; /* blocks_info[port][0].count */
            for (uint32_t r = 0; r < P; r++) {
                or_compute_peers(t, r, p, family, peers + (size_t)r * L);
                uint8_t* tb = tmp + ((size_t)r * ports + p) * per_port_tmp;
                memcpy(tb, send + (size_t)r * count * ds + pofs, c0 * ds);
            }
            for (uint32_t step = 0; step < L && rc == OR_OK; step++) {
                size_t n = ((size_t)1 << step) * c0;
                for (uint32_t r = 0; r < P; r++) {
                    uint32_t peer = (dist == OR_DIST_DECREASING) ? peers[(size_t)r * L + L - step - 1] : peers[(size_t)r * L + step];
                    uint8_t* tb = tmp + ((size_t)r * ports + p) * per_port_tmp;
                    post_send(&st, (int)r, (int)peer, 0, tb, n * ds);
                    post_recv(&st, (int)r, (int)peer, 0, tb + n * ds, n * ds);
                }
                rc = step_run(&st);
            }
            int odist = (dist == OR_DIST_DECREASING) ? OR_DIST_INCREASING : OR_DIST_DECREASING;
            for (uint32_t r = 0; r < P && rc == OR_OK; r++) {
                rc = or_get_tree(t, r, (uint32_t)p, family, odist, tr, tr + P, tr + 2 * P, tr + 3 * P, tr + 4 * P);
                uint8_t* tb = tmp + ((size_t)r * ports + p) * per_port_tmp;
                for (uint32_t i = 0; i < P; i++)
                    memcpy(recv + r * rcb + (size_t)i * count * ds + pofs, tb + (size_t)tr[2 * P + i] * c0 * ds, cport * ds);
            }
        }
        free(tr);
        free(peers);
        free(tmp);
    } else if (variant == 1) {
        if (ports != 1) return OR_ERR_UNSUPPORTED;
        int odist = (dist == OR_DIST_DECREASING) ? OR_DIST_INCREASING : OR_DIST_DECREASING;
        uint32_t* tr = (uint32_t*)malloc(sizeof(uint32_t) * 5 * P);
        rc = or_get_tree(t, 0, 0, family, odist, tr, tr + P, tr + 2 * P, tr + 3 * P, tr + 4 * P);
        uint32_t* remap = tr + 2 * P;
        uint32_t* inv = (uint32_t*)malloc(sizeof(uint32_t) * P);
        for (uint32_t i = 0; i < P; i++) inv[remap[i]] = i;
        uint32_t* peers = (uint32_t*)malloc(sizeof(uint32_t) * P * L);
        size_t* min_res = (size_t*)malloc(sizeof(size_t) * P);
        size_t cb = (size_t)count * ds;
        for (uint32_t r = 0; r < P; r++) {
            or_compute_peers(t, r, 0, family, peers + (size_t)r * L);
            post_send(&st, (int)r, (int)inv[r], 0, send + r * cb, cb);
            post_recv(&st, (int)r, (int)remap[r], 0, recv + r * rcb + remap[r] * cb, cb);
            min_res[r] = remap[r];
        }
        if (rc == OR_OK) rc = step_run(&st);
        size_t nb = 1;
        for (uint32_t step = 0; step < L && rc == OR_OK; step++) {
            size_t* mrr = (size_t*)malloc(sizeof(size_t) * P);
            for (uint32_t r = 0; r < P; r++) {
                uint32_t peer = (dist == OR_DIST_DECREASING) ? peers[(size_t)r * L + L - step - 1] : peers[(size_t)r * L + step];
                mrr[r] = remap[peer] & ~((1u << step) - 1);
                post_send(&st, (int)r, (int)peer, 0, recv + r * rcb + min_res[r] * cb, nb * cb);
                post_recv(&st, (int)r, (int)peer, 0, recv + r * rcb + mrr[r] * cb, nb * cb);
            }
            rc = step_run(&st);
            for (uint32_t r = 0; r < P; r++) min_res[r] = min_res[r] < mrr[r] ? min_res[r] : mrr[r];
            free(mrr);
            nb *= 2;
        }
        free(min_res);
        free(peers);
        free(inv);
        free(tr);
    } else if (variant == 2) {
        /* 1D, linear ranks */
        size_t cb = (size_t)count * ds;
        for (uint32_t r = 0; r < P; r++) memcpy(recv + r * rcb + r * cb, send + r * cb, cb);
        int inverse_mask = 0x1 << (int)(ceil(log2((double)P)) - 1);
        int step = 0;
        while (inverse_mask > 0 && rc == OR_OK) {
            for (uint32_t block = 1; block < P && rc == OR_OK; block++) {
                int k = 31 - __builtin_clz(or_get_nu(block, P));
                if (k != step) continue;
                for (uint32_t r = 0; r < P; r++) {
                    int partner = (r % 2 == 0) ? or_mod((int)r + or_nbtob((uint32_t)((inverse_mask << 1) - 1)), (int)P)
                                               : or_mod((int)r - or_nbtob((uint32_t)((inverse_mask << 1) - 1)), (int)P);
                    size_t bts, btr;
                    if (r % 2 == 0) {
                        btr = (size_t)or_mod((int)(block + r), (int)P);
                        bts = (size_t)or_mod(partner - (int)block, (int)P);
                    } else {
                        btr = (size_t)or_mod((int)r - (int)block, (int)P);
                        bts = (size_t)or_mod((int)block + partner, (int)P);
                    }
                    if (bts != (size_t)partner) post_send(&st, (int)r, partner, 0, recv + r * rcb + bts * cb, cb);
                    if (btr != r) post_recv(&st, (int)r, partner, 0, recv + r * rcb + btr * cb, cb);
                }
                rc = step_run(&st);
            }
            inverse_mask >>= 1;
            step++;
        }
    } else {
        rc = OR_ERR_UNSUPPORTED;
    }
    step_free(&st);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* alltoall                                                                  */
/* ------------------------------------------------------------------------ */

/* bine_alltoall_mpi (lib/libbine_coll_alltoall.cc:411-540); family BRUCK -> bruck_alltoall (:18-107) */
int or_alltoall(uint32_t P, int family, int count, int dt, const void* sendv, void* recvv) {
    size_t ds = or_dtsize(dt);
    if (!ds || count < 1 || P < 2) return OR_ERR_ARG;
    const uint8_t* send = (const uint8_t*)sendv;
    uint8_t* recv = (uint8_t*)recvv;
    size_t cb = (size_t)count * ds, rbytes = cb * P;
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    if (family == OR_FAMILY_BRUCK) {
        uint8_t* tsend = (uint8_t*)calloc((size_t)P * rbytes, 1);
        uint8_t* tbuf = (uint8_t*)malloc((size_t)P * cb * ((P + 1) / 2) + 1);
        uint8_t* trecv = (uint8_t*)malloc((size_t)P * cb * ((P + 1) / 2) + 1);
        for (uint32_t r = 0; r < P; r++)
            for (uint32_t i = 0; i < P; i++) {
                uint32_t index = (i - r + P) % P;
                memcpy(tsend + r * rbytes + index * cb, send + r * rbytes + i * cb, cb);
            }
        for (uint32_t k = 1; k < P && rc == OR_OK; k <<= 1) {
            uint32_t idx[4096];
            uint32_t nsb = 0;
            for (uint32_t i = k; i < P; i++)
                if (i & k) idx[nsb++] = i;
            for (uint32_t r = 0; r < P; r++) {
                for (uint32_t i = 0; i < nsb; i++)
                    memcpy(tbuf + (size_t)r * cb * ((P + 1) / 2) + i * cb, tsend + r * rbytes + idx[i] * cb, cb);
                uint32_t recv_proc = (r - k + P) % P, send_proc = (r + k) % P;
                post_send(&st, (int)r, (int)send_proc, 0, tbuf + (size_t)r * cb * ((P + 1) / 2), nsb * cb);
                post_recv(&st, (int)r, (int)recv_proc, 0, trecv + (size_t)r * cb * ((P + 1) / 2), nsb * cb);
            }
            rc = step_run(&st);
            for (uint32_t r = 0; r < P; r++)
                for (uint32_t i = 0; i < nsb; i++)
                    memcpy(tsend + r * rbytes + idx[i] * cb, trecv + (size_t)r * cb * ((P + 1) / 2) + i * cb, cb);
        }
        for (uint32_t r = 0; r < P; r++)
            for (uint32_t i = 0; i < P; i++) {
                uint32_t index = (r - i + P) % P;
                memcpy(recv + r * rbytes + index * cb, tsend + r * rbytes + i * cb, cb);
            }
        free(tsend);
        free(tbuf);
        free(trecv);
        step_free(&st);
        return rc;
    }
    if (!or_is_power_of_two((int)P)) return OR_ERR_UNSUPPORTED;
    uint8_t* tmp = (uint8_t*)malloc((size_t)P * rbytes + 1);
    uint32_t* res = (uint32_t*)malloc(sizeof(uint32_t) * P * P);
    uint32_t* resn = (uint32_t*)malloc(sizeof(uint32_t) * P * P);
    size_t* nres = (size_t*)malloc(sizeof(size_t) * P);
    memcpy(tmp, send, (size_t)P * rbytes);
    for (uint32_t r = 0; r < P; r++) {
        for (uint32_t i = 0; i < P; i++) res[(size_t)r * P + i] = i;
        nres[r] = P;
    }
    int mask = 0x1;
    int inverse_mask = 0x1 << (int)(ceil(log2((double)P)) - 1);
    int block_first_mask = ~(inverse_mask - 1);
    while (mask < (int)P && rc == OR_OK) {
        for (uint32_t r = 0; r < P && rc == OR_OK; r++) {
            int partner = (r % 2 == 0) ? or_mod((int)r + or_nbtob((uint32_t)((mask << 1) - 1)), (int)P)
                                       : or_mod((int)r - or_nbtob((uint32_t)((mask << 1) - 1)), (int)P);
            size_t min_s = or_remap_rank((uint32_t)partner, P) & (uint32_t)block_first_mask;
            size_t max_s = min_s + inverse_mask - 1;
            size_t scnt = 0, kcnt = 0, off_send = 0, off_keep = 0, nn = 0;
            uint8_t* tb = tmp + r * rbytes;
            uint8_t* rbf = recv + r * rbytes;
            for (uint32_t i = 0; i < P; i++) {
                uint32_t block = res[(size_t)r * P + (i % nres[r])];
                uint32_t rb = or_remap_rank(block, P);
                size_t off = (size_t)i * cb;
                if (rb >= min_s && rb <= max_s) {
                    memcpy(rbf + off_send, tb + off, cb);
                    off_send += cb;
                    scnt++;
                } else {
                    if (off != off_keep) memmove(tb + off_keep, tb + off, cb);
                    off_keep += cb;
                    kcnt++;
                    resn[(size_t)r * P + nn++] = block;
                }
            }
            if (kcnt != P / 2 || scnt != P / 2) rc = OR_ERR_MATCH;
            nres[r] = nres[r] / 2;
            post_send(&st, (int)r, partner, 0, rbf, scnt * cb);
            post_recv(&st, (int)r, partner, 0, tb + (P / 2) * cb, scnt * cb);
        }
        if (rc == OR_OK) rc = step_run(&st);
        uint32_t* sw = res;
        res = resn;
        resn = sw;
        mask <<= 1;
        inverse_mask >>= 1;
        block_first_mask >>= 1;
    }
    int nbits = (int)ceil(log2((double)P));
    for (uint32_t r = 0; r < P && rc == OR_OK; r++)
        for (uint32_t i = 0; i < P; i++) {
            int rot = (r % 2 == 0) ? or_mod((int)i - (int)r, (int)P) : or_mod((int)r - (int)i, (int)P);
            uint32_t repr = or_in_range(rot, (uint32_t)nbits) ? or_btonb(rot) : or_btonb(rot - (int)P);
            int index = or_remap_distance_doubling(repr);
            memcpy(recv + r * rbytes + (size_t)i * cb, tmp + r * rbytes + (size_t)index * cb, cb);
        }
    free(tmp);
    free(res);
    free(resn);
    free(nres);
    step_free(&st);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* rooted 1D collectives                                                     */
/* ------------------------------------------------------------------------ */

/* variant 0: bine_bcast_l_mpi (lib/libbine_coll_bcast.cc:635-666)
 * variant 2: bine_bcast_scatter_allgather_mpi (:985-1095), root 0 only */
int or_bcast(uint32_t P, int variant, int count, int dt, int root, void* bufv) {
    size_t ds = or_dtsize(dt);
    if (!ds || count < 0 || root < 0 || root >= (int)P || P < 2) return OR_ERR_ARG;
    uint8_t* buf = (uint8_t*)bufv;
    size_t cb = (size_t)count * ds;
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    int nbits = (int)ceil(log2((double)P));
    if (variant == 0) {
        int* recvd = (int*)calloc(P, sizeof(int));
        recvd[root] = 1;
        int mask = 0x1 << (nbits - 1);
        while (mask > 0 && rc == OR_OK) {
            int* got = (int*)calloc(P, sizeof(int));
            for (uint32_t r = 0; r < P; r++) {
                int vrank = or_mod((int)r - root, (int)P);
                int partner = (int)(or_btonb(vrank) ^ (uint32_t)((mask << 1) - 1));
                partner = or_mod(or_nbtob((uint32_t)partner) + root, (int)P);
                int mask_lsbs = (mask << 1) - 1;
                int lsbs = (int)(or_btonb(vrank) & (uint32_t)mask_lsbs);
                int equal = (lsbs == 0 || lsbs == mask_lsbs);
                if (recvd[r]) post_send(&st, (int)r, partner, 0, buf + r * cb, cb);
                else if (equal) {
                    post_recv(&st, (int)r, partner, 0, buf + r * cb, cb);
                    got[r] = 1;
                }
            }
            rc = step_run(&st);
            for (uint32_t r = 0; r < P; r++)
                if (got[r]) recvd[r] = 1;
            free(got);
            mask >>= 1;
        }
        free(recvd);
    } else if (variant == 2) {
        if (root != 0) return OR_ERR_UNSUPPORTED;
        int* displs = (int*)malloc(sizeof(int) * P);
        int* rcounts = (int*)malloc(sizeof(int) * P);
        int cpr = count / (int)P, rem = count % (int)P;
        for (int i = 0; i < (int)P; i++) {
            displs[i] = cpr * i + (i < rem ? i : rem);
            rcounts[i] = cpr + (i < rem ? 1 : 0);
        }
        int* remapped = (int*)malloc(sizeof(int) * P);
        int* recvmask = (int*)malloc(sizeof(int) * P);
        int* recvd = (int*)calloc(P, sizeof(int));
        int inverse_mask0 = 0x1 << (nbits - 1);
        for (uint32_t r = 0; r < P; r++) {
            remapped[r] = (int)or_remap_rank(r, P);
            recvmask[r] = (r != (uint32_t)root) ? (0x1 << (__builtin_ffs(remapped[r]) - 1)) : (inverse_mask0 << 1);
        }
        recvd[root] = 1;
        int mask = 0x1, inverse_mask = inverse_mask0, bfm = ~(inverse_mask - 1);
        while (mask < (int)P && rc == OR_OK) {
            int* got = (int*)calloc(P, sizeof(int));
            for (uint32_t r = 0; r < P; r++) {
                int partner = (r % 2 == 0) ? or_mod((int)r + or_nbtob((uint32_t)((mask << 1) - 1)), (int)P)
                                           : or_mod((int)r - or_nbtob((uint32_t)((mask << 1) - 1)), (int)P);
                int sbf = (int)or_remap_rank((uint32_t)partner, P) & bfm;
                int sbl = sbf + inverse_mask - 1;
                int scnt = displs[sbl] - displs[sbf] + rcounts[sbl];
                int rbf = remapped[r] & bfm;
                int rbl = rbf + inverse_mask - 1;
                int rcnt = displs[rbl] - displs[rbf] + rcounts[rbl];
                if (recvd[r]) post_send(&st, (int)r, partner, 0, buf + r * cb + (size_t)displs[sbf] * ds, (size_t)scnt * ds);
                else if (inverse_mask == recvmask[r] || partner == root) {
                    post_recv(&st, (int)r, partner, 0, buf + r * cb + (size_t)displs[rbf] * ds, (size_t)rcnt * ds);
                    got[r] = 1;
                }
            }
            rc = step_run(&st);
            for (uint32_t r = 0; r < P; r++)
                if (got[r]) recvd[r] = 1;
            free(got);
            mask <<= 1;
            inverse_mask >>= 1;
            bfm >>= 1;
        }
        mask >>= 1;
        inverse_mask = 0x1;
        bfm = ~0x0;
        while (mask > 0 && rc == OR_OK) {
            for (uint32_t r = 0; r < P; r++) {
                int partner = (r % 2 == 0) ? or_mod((int)r + or_nbtob((uint32_t)((mask << 1) - 1)), (int)P)
                                           : or_mod((int)r - or_nbtob((uint32_t)((mask << 1) - 1)), (int)P);
                int has_r = !(inverse_mask < recvmask[r]);
                int has_s = !(inverse_mask == recvmask[r]);
                if (has_s) {
                    int f = remapped[r] & bfm, l = f + inverse_mask - 1;
                    int c = displs[l] - displs[f] + rcounts[l];
                    post_send(&st, (int)r, partner, 0, buf + r * cb + (size_t)displs[f] * ds, (size_t)c * ds);
                }
                if (has_r) {
                    int f = (int)or_remap_rank((uint32_t)partner, P) & bfm, l = f + inverse_mask - 1;
                    int c = displs[l] - displs[f] + rcounts[l];
                    post_recv(&st, (int)r, partner, 0, buf + r * cb + (size_t)displs[f] * ds, (size_t)c * ds);
                }
            }
            rc = step_run(&st);
            mask >>= 1;
            inverse_mask <<= 1;
            bfm <<= 1;
        }
        free(displs);
        free(rcounts);
        free(remapped);
        free(recvmask);
        free(recvd);
    } else {
        rc = OR_ERR_UNSUPPORTED;
    }
    step_free(&st);
    return rc;
}

/* variant 0: bine_reduce_mpi (lib/libbine_coll_reduce.cc:575-615)
 * variant 1: bine_reduce_redscat_gather_mpi (:1042-1152), root 0 only */
int or_reduce(uint32_t P, int variant, int count, int dt, int op, int root, const void* sendv, void* recv_root) {
    size_t ds = or_dtsize(dt);
    if (!ds || count < 0 || root < 0 || root >= (int)P || P < 2) return OR_ERR_ARG;
    const uint8_t* send = (const uint8_t*)sendv;
    size_t cb = (size_t)count * ds;
    uint8_t* res = (uint8_t*)malloc((size_t)P * cb + 1); /* recvbuf at root, scratch elsewhere */
    uint8_t* tmp = (uint8_t*)malloc((size_t)P * cb + 1);
    memcpy(res, send, (size_t)P * cb);
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    int nbits = (int)ceil(log2((double)P));
    if (variant == 0) {
        int* active = (int*)malloc(sizeof(int) * P);
        for (uint32_t r = 0; r < P; r++) active[r] = 1;
        int mask = 0x1;
        while (mask < (int)P && rc == OR_OK) {
            int* rcv = (int*)calloc(P, sizeof(int));
            for (uint32_t r = 0; r < P; r++) {
                if (!active[r]) continue;
                int vrank = or_mod((int)r - root, (int)P);
                int partner = (int)(or_btonb(vrank) ^ (uint32_t)((mask << 1) - 1));
                partner = or_mod(or_nbtob((uint32_t)partner) + root, (int)P);
                int mask_lsbs = (mask << 2) - 1;
                int lsbs = (int)(or_btonb(vrank) & (uint32_t)mask_lsbs);
                int equal = (lsbs == 0 || lsbs == mask_lsbs);
                if (!equal || ((mask << 1) >= (int)P && ((int)r != root))) {
                    post_send(&st, (int)r, partner, 0, res + r * cb, cb);
                    active[r] = 0;
                } else {
                    post_recv(&st, (int)r, partner, 0, tmp + r * cb, cb);
                    rcv[r] = 1;
                }
            }
            rc = step_run(&st);
            for (uint32_t r = 0; r < P && rc == OR_OK; r++)
                if (rcv[r]) reduce_local(dt, op, tmp + r * cb, res + r * cb, (size_t)count);
            free(rcv);
            mask <<= 1;
        }
        free(active);
    } else if (variant == 1) {
        if (root != 0) {
            rc = OR_ERR_UNSUPPORTED;
        } else {
            int* displs = (int*)malloc(sizeof(int) * P);
            int* rcounts = (int*)malloc(sizeof(int) * P);
            int cpr = count / (int)P, rem = count % (int)P;
            for (int i = 0; i < (int)P; i++) {
                displs[i] = cpr * i + (i < rem ? i : rem);
                rcounts[i] = cpr + (i < rem ? 1 : 0);
            }
            int mask = 0x1, inverse_mask = 0x1 << (nbits - 1), bfm = ~(inverse_mask - 1);
            while (mask < (int)P && rc == OR_OK) {
                int* rf = (int*)malloc(sizeof(int) * P);
                int* rcn = (int*)malloc(sizeof(int) * P);
                for (uint32_t r = 0; r < P; r++) {
                    int partner = (r % 2 == 0) ? or_mod((int)r + or_nbtob((uint32_t)((mask << 1) - 1)), (int)P)
                                               : or_mod((int)r - or_nbtob((uint32_t)((mask << 1) - 1)), (int)P);
                    int sbf = (int)or_remap_rank((uint32_t)partner, P) & bfm, sbl = sbf + inverse_mask - 1;
                    int scnt = displs[sbl] - displs[sbf] + rcounts[sbl];
                    int rbf = (int)or_remap_rank(r, P) & bfm, rbl = rbf + inverse_mask - 1;
                    rf[r] = rbf;
                    rcn[r] = displs[rbl] - displs[rbf] + rcounts[rbl];
                    post_send(&st, (int)r, partner, 0, res + r * cb + (size_t)displs[sbf] * ds, (size_t)scnt * ds);
                    post_recv(&st, (int)r, partner, 0, tmp + r * cb + (size_t)displs[rbf] * ds, (size_t)rcn[r] * ds);
                }
                rc = step_run(&st);
                for (uint32_t r = 0; r < P && rc == OR_OK; r++)
                    reduce_local(dt, op, tmp + r * cb + (size_t)displs[rf[r]] * ds, res + r * cb + (size_t)displs[rf[r]] * ds,
                                 (size_t)rcn[r]);
                free(rf);
                free(rcn);
                mask <<= 1;
                inverse_mask >>= 1;
                bfm >>= 1;
            }
            mask >>= 1;
            inverse_mask = 0x1;
            bfm = ~0x0;
            int* active = (int*)malloc(sizeof(int) * P);
            int* recvmask = (int*)malloc(sizeof(int) * P);
            for (uint32_t r = 0; r < P; r++) {
                active[r] = 1;
                recvmask[r] = 0x1 << (__builtin_ffs((int)or_remap_rank(r, P)) - 1);
            }
            while (mask > 0 && rc == OR_OK) {
                for (uint32_t r = 0; r < P; r++) {
                    if (!active[r]) continue;
                    int partner = (r % 2 == 0) ? or_mod((int)r + or_nbtob((uint32_t)((mask << 1) - 1)), (int)P)
                                               : or_mod((int)r - or_nbtob((uint32_t)((mask << 1) - 1)), (int)P);
                    if (inverse_mask & recvmask[r]) {
                        int f = (int)or_remap_rank(r, P) & bfm, l = f + inverse_mask - 1;
                        int c = displs[l] - displs[f] + rcounts[l];
                        post_send(&st, (int)r, partner, 0, res + r * cb + (size_t)displs[f] * ds, (size_t)c * ds);
                        active[r] = 0;
                    } else {
                        int f = (int)or_remap_rank((uint32_t)partner, P) & bfm, l = f + inverse_mask - 1;
                        int c = displs[l] - displs[f] + rcounts[l];
                        post_recv(&st, (int)r, partner, 0, res + r * cb + (size_t)displs[f] * ds, (size_t)c * ds);
                    }
                }
                rc = step_run(&st);
                mask >>= 1;
                inverse_mask <<= 1;
                bfm <<= 1;
            }
            free(active);
            free(recvmask);
            free(displs);
            free(rcounts);
        }
    } else {
        rc = OR_ERR_UNSUPPORTED;
    }
    if (rc == OR_OK) memcpy(recv_root, res + (size_t)root * cb, cb);
    step_free(&st);
    free(res);
    free(tmp);
    return rc;
}

/* bine_scatter_mpi (lib/libbine_coll_scatter.cc:187-324); send_root holds P*count elements. */
int or_scatter(uint32_t P, int count, int dt, int root, const void* send_root, void* recvv) {
    size_t ds = or_dtsize(dt);
    if (!ds || count < 1 || root < 0 || root >= (int)P || P < 2) return OR_ERR_ARG;
    uint8_t* recv = (uint8_t*)recvv;
    size_t cb = (size_t)count * ds;
    int nbits = (int)ceil(log2((double)P));
    /* per-rank state */
    int* hd = (int*)malloc(sizeof(int) * P);
    size_t* minr = (size_t*)malloc(sizeof(size_t) * P);
    size_t* maxr = (size_t*)malloc(sizeof(size_t) * P);
    int* recvd = (int*)calloc(P, sizeof(int));
    int* leaf = (int*)calloc(P, sizeof(int));
    int* sboff = (int*)malloc(sizeof(int) * P);
    uint8_t** sbuf = (uint8_t**)calloc(P, sizeof(uint8_t*));
    uint8_t** tmpb = (uint8_t**)calloc(P, sizeof(uint8_t*));
    for (uint32_t r = 0; r < P; r++) {
        hd[r] = (r % 2) ? -1 : 1;
        if (nbits % 2 == 0) hd[r] *= -1;
        unsigned msk = (0x1u << nbits) - 1;
        if (r % 2 == 0) {
            maxr[r] = (size_t)or_mod((int)(((unsigned)r + 0x55555555u) & msk), (int)P);
            minr[r] = (size_t)or_mod((int)(((unsigned)r - 0xAAAAAAAAu) & msk), (int)P);
        } else {
            minr[r] = (size_t)or_mod((int)(((unsigned)r - 0x55555555u) & msk), (int)P);
            maxr[r] = (size_t)or_mod((int)(((unsigned)r + 0xAAAAAAAAu) & msk), (int)P);
        }
        sboff[r] = (int)r;
        if ((int)r == root) {
            recvd[r] = 1;
            sbuf[r] = (uint8_t*)send_root;
        }
    }
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    int mask = 0x1 << (nbits - 1);
    while (mask > 0 && rc == OR_OK) {
        int* got = (int*)calloc(P, sizeof(int));
        for (uint32_t r = 0; r < P; r++) {
            int vrank = or_mod((int)r - root, (int)P);
            int vnb = (int)or_btonb(vrank);
            int partner = vnb ^ ((mask << 1) - 1);
            partner = or_mod(or_nbtob((uint32_t)partner) + root, (int)P);
            int mask_lsbs = (mask << 1) - 1;
            int lsbs = vnb & mask_lsbs;
            int equal = (lsbs == 0 || lsbs == mask_lsbs);
            size_t ts = minr[r], te = (size_t)or_mod((int)minr[r] + mask - 1, (int)P);
            size_t bs = (size_t)or_mod((int)te + 1, (int)P), be = maxr[r];
            size_t ss, se, rs, re;
            if (hd[r] == 1) {
                ss = bs; se = be; rs = ts; re = te;
                maxr[r] = (size_t)or_mod((int)maxr[r] - mask, (int)P);
            } else {
                ss = ts; se = te; rs = bs; re = be;
                minr[r] = (size_t)or_mod((int)minr[r] + mask, (int)P);
            }
            if (recvd[r]) {
                if (se >= ss) {
                    post_send(&st, (int)r, partner, 0, sbuf[r] + ss * cb, cb * (se - ss + 1));
                } else {
                    post_send(&st, (int)r, partner, 0, sbuf[r] + ss * cb, cb * ((P - 1) - ss + 1));
                    post_send(&st, (int)r, partner, 0, sbuf[r], cb * (se + 1));
                }
            } else if (equal) {
                size_t nb = (size_t)or_mod((int)(re - rs + 1), (int)P);
                uint8_t* rbuf;
                if (rs == re) {
                    rbuf = recv + r * cb;
                    leaf[r] = 1;
                } else {
                    tmpb[r] = (uint8_t*)malloc(cb * nb + 1);
                    sbuf[r] = tmpb[r];
                    rbuf = tmpb[r];
                    minr[r] = 0;
                    maxr[r] = nb - 1;
                    sboff[r] = or_mod((int)r - (int)rs, (int)P);
                }
                if (re >= rs) {
                    post_recv(&st, (int)r, partner, 0, rbuf, cb * nb);
                } else {
                    post_recv(&st, (int)r, partner, 0, rbuf, cb * ((P - 1) - rs + 1));
                    post_recv(&st, (int)r, partner, 0, rbuf + cb * ((P - 1) - rs + 1), cb * (re + 1));
                }
                got[r] = 1;
            }
        }
        rc = step_run(&st);
        for (uint32_t r = 0; r < P; r++)
            if (got[r]) recvd[r] = 1;
        free(got);
        mask >>= 1;
        for (uint32_t r = 0; r < P; r++) hd[r] *= -1;
    }
    for (uint32_t r = 0; r < P && rc == OR_OK; r++)
        if (!leaf[r]) memcpy(recv + r * cb, sbuf[r] + (size_t)sboff[r] * cb, cb);
    for (uint32_t r = 0; r < P; r++) free(tmpb[r]);
    step_free(&st);
    free(hd);
    free(minr);
    free(maxr);
    free(recvd);
    free(leaf);
    free(sboff);
    free(sbuf);
    free(tmpb);
    return rc;
}

/* bine_gather_mpi (lib/libbine_coll_gather.cc:202-272); recv_root receives P*count elements. */
int or_gather(uint32_t P, int count, int dt, int root, const void* sendv, void* recv_root) {
    size_t ds = or_dtsize(dt);
    if (!ds || count < 1 || root < 0 || root >= (int)P || P < 2) return OR_ERR_ARG;
    const uint8_t* send = (const uint8_t*)sendv;
    size_t cb = (size_t)count * ds, rb = cb * P;
    uint8_t* rbuf = (uint8_t*)malloc((size_t)P * rb + 1);
    size_t* minr = (size_t*)malloc(sizeof(size_t) * P);
    size_t* maxr = (size_t*)malloc(sizeof(size_t) * P);
    int* ext = (int*)malloc(sizeof(int) * P);
    int* active = (int*)malloc(sizeof(int) * P);
    for (uint32_t r = 0; r < P; r++) {
        memcpy(rbuf + r * rb + r * cb, send + r * cb, cb);
        minr[r] = maxr[r] = r;
        ext[r] = (r % 2) ? -1 : 1;
        active[r] = 1;
    }
    int rc = OR_OK;
    step_t st;
    step_init(&st);
    int mask = 0x1;
    while (mask < (int)P && rc == OR_OK) {
        for (uint32_t r = 0; r < P; r++) {
            if (!active[r]) continue;
            int vrank = or_mod((int)r - root, (int)P);
            int partner = (int)(or_btonb(vrank) ^ (uint32_t)((mask << 1) - 1));
            partner = or_mod(or_nbtob((uint32_t)partner) + root, (int)P);
            int mask_lsbs = (mask << 2) - 1;
            int lsbs = (int)(or_btonb(vrank) & (uint32_t)mask_lsbs);
            int equal = (lsbs == 0 || lsbs == mask_lsbs);
            uint8_t* b = rbuf + r * rb;
            if (!equal || ((mask << 1) >= (int)P && ((int)r != root))) {
                if (maxr[r] >= minr[r]) {
                    post_send(&st, (int)r, partner, 0, b + minr[r] * cb, cb * (maxr[r] - minr[r] + 1));
                } else {
                    post_send(&st, (int)r, partner, 0, b + minr[r] * cb, cb * ((P - 1) - minr[r] + 1));
                    post_send(&st, (int)r, partner, 0, b, cb * (maxr[r] + 1));
                }
                active[r] = 0;
            } else {
                size_t rs, re;
                if (ext[r] == 1) {
                    rs = (size_t)or_mod((int)maxr[r] + 1, (int)P);
                    re = (size_t)or_mod((int)maxr[r] + mask, (int)P);
                    maxr[r] = re;
                } else {
                    re = (size_t)or_mod((int)minr[r] - 1, (int)P);
                    rs = (size_t)or_mod((int)minr[r] - mask, (int)P);
                    minr[r] = rs;
                }
                if (re >= rs) {
                    post_recv(&st, (int)r, partner, 0, b + rs * cb, cb * (re - rs + 1));
                } else {
                    post_recv(&st, (int)r, partner, 0, b + rs * cb, cb * ((P - 1) - rs + 1));
                    post_recv(&st, (int)r, partner, 0, b, cb * (re + 1));
                }
                ext[r] *= -1;
            }
        }
        rc = step_run(&st);
        mask <<= 1;
    }
    if (rc == OR_OK) memcpy(recv_root, rbuf + (size_t)root * rb, rb);
    step_free(&st);
    free(rbuf);
    free(minr);
    free(maxr);
    free(ext);
    free(active);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* simulation/simulator.py                                                   */
/* ------------------------------------------------------------------------ */

/* simulator.py:11-17 (its own 3D id convention) */
int or_sim_coord_from_id(int id, const uint32_t* dims, uint32_t nd, int* coord) {
    if (nd == 1) {
        coord[0] = id;
    } else if (nd == 2) {
        coord[0] = id / (int)dims[1];
        coord[1] = id % (int)dims[1];
    } else if (nd == 3) {
        coord[0] = (id / (int)dims[1]) % (int)dims[0];
        coord[1] = id % (int)dims[1];
        coord[2] = id / (int)(dims[0] * dims[1]);
    } else {
        return OR_ERR_ARG;
    }
    return OR_OK;
}

/* simulator.py:20-26 */
int or_sim_id_from_coord(const int* coord, const uint32_t* dims, uint32_t nd) {
    if (nd == 1) return coord[0];
    if (nd == 2) return coord[0] * (int)dims[1] + coord[1];
    return coord[2] * (int)(dims[0] * dims[1]) + coord[0] * (int)dims[1] + coord[1];
}

/* Python's floor-division and modulo */
static int py_mod(int a, int b) {
    int r = a % b;
    return (r != 0 && ((r < 0) != (b < 0))) ? r + b : r;
}

/* simulator.py:52-67 */
static int sim_peer_distance(int sender, int step, const uint32_t* dims, uint32_t nd, int allgather,
                             const int* distances) {
    int dim = step % (int)nd;
    int rel = step / (int)nd;
    int coord[3];
    or_sim_coord_from_id(sender, dims, nd, coord);
    int sgn = (rel % 2 == 0) ? 1 : -1;
    int next_direction = (coord[dim] % 2 == 0) ? sgn : -sgn;
    if (allgather) return next_direction * distances[(int)log2((double)dims[0]) - rel - 1];
    return next_direction * distances[rel];
}

/* simulator.py:70-76 */
int or_sim_peer(int sender, int step, const uint32_t* dims, uint32_t nd, int allgather, const int* distances) {
    int distance = sim_peer_distance(sender, step, dims, nd, allgather, distances);
    int coord[3];
    or_sim_coord_from_id(sender, dims, nd, coord);
    int td = step % (int)nd;
    coord[td] = py_mod(coord[td] + distance, (int)dims[td]);
    return or_sim_id_from_coord(coord, dims, nd);
}

/* simulator.py:80-92 (appends into out, returns new length) */
static int sim_rs_indexes(int sender, int step, const uint32_t* dims, uint32_t nd, const int* distances, int p,
                          int* out, int n) {
    if (step >= (int)log2((double)p)) return n;
    for (int s = step; s < (int)log2((double)p); s++) {
        int peer = or_sim_peer(sender, s, dims, nd, 0, distances);
        out[n++] = peer;
        n = sim_rs_indexes(peer, s + 1, dims, nd, distances, p, out, n);
    }
    return n;
}

/* run_collectives (simulator.py:125-177); collective 0 = ALLREDUCE, 1 = REDUCESCATTER.
 * ALLGATHER (2) is not restated: simulator.py:101,104 call find_allgather_indexes
 * with a missing argument and raise TypeError before moving any data. */
int or_sim_run(const uint32_t* dims, uint32_t nd, int collective, const int* distances, double* data,
               double* out_stats) {
    if (collective == 2) return OR_ERR_UNSUPPORTED;
    int p = 1;
    for (uint32_t d = 0; d < nd; d++) p *= (int)dims[d];
    double sum_d = 0, sum_dr = 0, bw_new = 0, bw_ref = 0, bw_ideal = 0;
    double* nc = (double*)malloc(sizeof(double) * p * p);
    int* idx = (int*)malloc(sizeof(int) * (size_t)p * 4 + 16);
    char* rf = (char*)malloc(p);
    int rc = OR_OK;
    int steps = (int)log2((double)p);
    for (int step = 0; step < steps && rc == OR_OK; step++) {
        memcpy(nc, data, sizeof(double) * p * p);
        memset(rf, 0, p);
        for (int rank = 0; rank < p; rank++) {
            int peer = or_sim_peer(rank, step, dims, nd, 0, distances);
            if (rf[peer]) {
                rc = OR_ERR_MATCH;
                break;
            }
            rf[peer] = 1;
            /* send(peer -> rank) */
            if (collective == 0) {
                for (int k = 0; k < p; k++) nc[(size_t)rank * p + k] += data[(size_t)peer * p + k];
            } else {
                int n = sim_rs_indexes(peer, step, dims, nd, distances, p, idx, 0);
                for (int i = 0; i < n; i++) nc[(size_t)rank * p + idx[i]] += data[(size_t)peer * p + idx[i]];
            }
        }
        for (int k = 0; k < p && rc == OR_OK; k++)
            if (!rf[k]) rc = OR_ERR_MATCH;
        if (rc != OR_OK) break;
        double distance = fabs((double)sim_peer_distance(0, step, dims, nd, 0, distances));
        sum_d += distance;
        double distance_ref = pow(2.0, (double)step / (double)nd);
        bw_ideal += 1.0 / pow(2.0, (double)(step + 1));
        bw_new += (1.0 / pow(2.0, (double)(step + 1))) * distance;
        bw_ref += (1.0 / pow(2.0, (double)(step + 1))) * distance_ref;
        sum_dr += distance_ref;
        memcpy(data, nc, sizeof(double) * p * p);
    }
    out_stats[0] = sum_d;
    out_stats[1] = sum_dr;
    out_stats[2] = bw_new;
    out_stats[3] = bw_ref;
    out_stats[4] = bw_ideal;
    free(nc);
    free(idx);
    free(rf);
    return rc;
}
