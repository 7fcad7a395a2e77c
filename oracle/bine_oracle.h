/*
 * bine_oracle.h -- CPU oracle for the Bine-tree collective data path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * library's MPI-layer algorithms (HLC-Lab/bine_trees_fugaku, lib/ and
 * simulation/), executed over P virtual ranks held in host memory.  It is the
 * checker for the HIP product and the timed CPU baseline; nothing in the
 * product links, loads or calls it.  Every function cites the reference
 * file:line it follows.
 *
 * Buffers: a "virtual-rank buffer" is P consecutive per-rank partitions of
 * `per_rank_bytes` each (rank r at base + r*per_rank_bytes), the same layout
 * the product's C-ABI uses.
 *
 * Pinning.  Every function is checked by the tests against closed forms,
 * invariants, brute force, MPI semantics, the reference's own
 * simulation/remap_rank.c (built under _ref/ and as a committed fixture) or
 * the oracle's own message traces.  On power-of-two sizes the floating-point
 * association of every reducing collective (or_allreduce B / B_COALESCE /
 * B_CONT / L with 1-6 ports, or_reduce_scatter in both distance orders,
 * or_reduce as binomial tree and as reduce-scatter + gather) is checked bit
 * for bit against replays of the reduction trees built from the peer
 * function and remap_rank alone (tests/test_oracle_association.py).
 * Parity unpinned: the fp association and block ownership that or_allreduce
 * / or_get_tree / or_block_step produce on NON-power-of-two sizes (where
 * compute_block_step's tie-break and the "disconnect myself" rule act,
 * lib/libbine_common.cc:1499-1500, 1574, 2045, and the shrink/enlarge folds of
 * allreduce L, :269-353) follow the code as read but are confirmed by no
 * reference output; and MPI_Reduce_local for fp32/fp64 is taken as IEEE
 * a + b (the system MPI is not part of the reference).
 */
#ifndef BINE_ORACLE_H
#define BINE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_DIMS 3
#define OR_MAX_PORTS 6
#define OR_MAX_STEPS 40

/* lib/libbine_common.h:75-86 */
enum { OR_FAMILY_DEFAULT = 0, OR_FAMILY_BINE = 1, OR_FAMILY_RECDOUB = 2, OR_FAMILY_BRUCK = 3, OR_FAMILY_RING = 4 };
/* lib/libbine_common.h:65-68 */
enum { OR_DIST_INCREASING = 0, OR_DIST_DECREASING = 1 };
/* lib/libbine_common.h:93-99 */
enum { OR_AR_L = 0, OR_AR_B = 1, OR_AR_REDUCE_BCAST = 2, OR_AR_B_CONT = 3, OR_AR_B_COALESCE = 4 };
/* element types (MPI_INT32_T, MPI_INT64_T, MPI_FLOAT, MPI_DOUBLE, MPI_CHAR) */
enum { OR_INT32 = 0, OR_INT64 = 1, OR_FLOAT = 2, OR_DOUBLE = 3, OR_INT8 = 4 };
/* MPI_SUM and the bench's no-op "VOID" operator (bench/bench.cc:240-256) */
enum { OR_SUM = 0, OR_VOID = 1 };

enum { OR_OK = 0, OR_ERR_ARG = 1, OR_ERR_UNSUPPORTED = 2, OR_ERR_MATCH = 3 };

typedef struct {
    uint32_t dims[OR_MAX_DIMS];
    uint32_t ndims;
    uint32_t size;
    uint32_t num_steps;
    uint32_t steps_per_dim[OR_MAX_DIMS];
} or_torus_t;

size_t or_dtsize(int dt);

/* ---- scalar helpers -------------------------------------------------- */
int or_ceil_log2(unsigned long long x);            /* lib/libbine_coll.cc:65-87 */
int or_mod(int a, int b);                          /* lib/libbine_coll.cc:59-62 */
int or_is_power_of_two(int x);                     /* lib/libbine_coll.cc:55-57 */
int or_rho(int k);                                 /* lib/libbine_common.h:27 */
uint32_t or_btonb(int32_t bin);                    /* lib/libbine_coll_bcast.cc:622-626 */
int32_t or_nbtob(uint32_t neg);                    /* lib/libbine_coll_bcast.cc:628-633 */
uint32_t or_reverse(uint32_t x);                   /* lib/libbine_coll_alltoall.cc:325-332 */
int or_in_range(int x, uint32_t nbits);            /* lib/libbine_coll_alltoall.cc:335-337 */
uint32_t or_rank_negabinary(uint32_t rank, uint32_t size); /* lib/libbine_coll_alltoall.cc:339-369 */
uint32_t or_remap_rank(uint32_t rank, uint32_t size);      /* lib/libbine_coll_alltoall.cc:372-378 */
uint32_t or_get_nu(uint32_t rank, uint32_t size);          /* lib/libbine_coll_allgather.cc:1125-1158 */
int or_remap_distance_doubling(uint32_t num);              /* lib/libbine_coll_alltoall.cc:401-409 */

/* OpenMP threads the oracle's data movement and reductions use (1 without OpenMP). */
int or_threads(void);

/* ---- torus / schedule ------------------------------------------------ */
int or_torus_init(or_torus_t* t, const uint32_t* dims, uint32_t ndims);
void or_coord_from_id(const or_torus_t* t, int id, int* coord);   /* lib/libbine_common.cc:206-216 */
int or_id_from_coord(const or_torus_t* t, const int* coord);      /* lib/libbine_common.cc:219-238 */
int or_is_mirroring_port(int port, uint32_t ndims);               /* lib/libbine_coll.cc:15-31 */
void or_step_info(const or_torus_t* t, uint32_t port, uint32_t* d, uint32_t* step_in_d); /* :306-331 */
void or_get_peer_c(const or_torus_t* t, const int* coord, uint32_t step, uint32_t port,
                   const uint32_t* sd, const uint32_t* ssd, int family, int* coord_peer); /* :89-107 */
void or_compute_peers(const or_torus_t* t, uint32_t rank, int port, int family, uint32_t* peers); /* :443-504 */
/* BineBitmapCalculator ctor, block_step table (lib/libbine_common.cc:2036-2045, 1491-1510) */
void or_block_step(const or_torus_t* t, uint32_t rank, uint32_t port, int family, uint32_t* block_step);
/* compute_next_bitmaps for reduce-scatter step `step` (lib/libbine_common.cc:1512-1579) */
void or_step_bitmaps(const or_torus_t* t, uint32_t rank, uint32_t port, int family, uint32_t step,
                     uint8_t* send_bitmap, uint8_t* recv_bitmap);
/* get_tree (lib/libbine_coll.cc:334-389): five arrays of t->size entries each */
int or_get_tree(const or_torus_t* t, uint32_t root, uint32_t port, int family, int dist,
                uint32_t* parent, uint32_t* reached_at_step, uint32_t* remapped_ranks,
                uint32_t* remapped_ranks_max, uint32_t* subtree_roots);
/* BineBitmapCalculator remap_blocks=true: remapped rank via dfs (lib/libbine_common.cc:1981-1993, 2023-2034) */
uint32_t or_cont_remapped_rank(const or_torus_t* t, uint32_t rank, uint32_t port, int family);

/* ---- collectives over P virtual ranks -------------------------------- */
/* MPI_Allreduce -> bine_coll_b / bine_coll_l (lib/libbine.cc:1012-1062) */
int or_allreduce(const or_torus_t* t, int ports, int family, int algo, int count, int dt, int op,
                 const void* send, void* recv);
/* MPI_Reduce_scatter (block) -> bine_reduce_scatter_mpi_contiguous (lib/libbine.cc:1064-1149) */
int or_reduce_scatter(const or_torus_t* t, int ports, int family, int dist, int recvcount, int dt, int op,
                      const void* send, void* recv);
/* MPI_Allgather -> bine_allgather_mpi_contiguous / _blocks_mpi / _send_mpi (lib/libbine.cc:1168-1250) */
int or_allgather(const or_torus_t* t, int ports, int family, int dist, int variant, int count, int dt,
                 const void* send, void* recv);
/* MPI_Alltoall -> bine_alltoall_mpi / bruck_alltoall (lib/libbine.cc:1291-1309) */
int or_alltoall(uint32_t P, int family, int count, int dt, const void* send, void* recv);
/* MPI_Bcast_f -> bine_bcast_l_mpi / bine_bcast_scatter_allgather_mpi (lib/libbine.cc:1252-1289) */
int or_bcast(uint32_t P, int variant, int count, int dt, int root, void* buffer);
/* MPI_Reduce -> bine_reduce_mpi / bine_reduce_redscat_gather_mpi (lib/libbine.cc:1440-1475) */
int or_reduce(uint32_t P, int variant, int count, int dt, int op, int root, const void* send, void* recv_root);
/* MPI_Scatter -> bine_scatter_mpi (lib/libbine.cc:1312-1374) */
int or_scatter(uint32_t P, int count, int dt, int root, const void* send_root, void* recv);
/* MPI_Gather -> bine_gather_mpi (lib/libbine.cc:1376-1438) */
int or_gather(uint32_t P, int count, int dt, int root, const void* send, void* recv_root);

/* ---- torus traffic simulator (simulation/simulator.py) ----------------- */
int or_sim_coord_from_id(int id, const uint32_t* dims, uint32_t nd, int* coord); /* simulator.py:11-17 */
int or_sim_id_from_coord(const int* coord, const uint32_t* dims, uint32_t nd);   /* simulator.py:20-26 */
int or_sim_peer(int sender, int step, const uint32_t* dims, uint32_t nd, int allgather,
                const int* distances);                                          /* simulator.py:52-76 */
/* run_collectives stats (simulator.py:125-177): out = {sum_dist, sum_dist_ref, bw_new, bw_ref, bw_ideal};
 * data: P*P doubles, updated in place; returns OR_OK or OR_ERR_MATCH on a double receive / no send. */
int or_sim_run(const uint32_t* dims, uint32_t nd, int collective, const int* distances,
               double* data, double* out_stats);

/* ---- message trace and link routing (extension: the reference has no
 * per-link accounting) -------------------------------------------------- */
/* While tracing is on, every message the step engine delivers is recorded as
 * (engine step, src, dst, bytes); the step counter restarts at or_trace_start. */
void or_trace_start(void);
size_t or_trace_stop(void); /* returns the number of recorded messages */
void or_trace_fetch(int32_t* step, int32_t* src, int32_t* dst, uint64_t* bytes);
/* Dimension-ordered minimal routing on the torus (row-major ids, last dimension
 * fastest): dimension 0 first; on a dimension the shorter way round, the
 * positive way on a tie.  Adds bytes[i] to every directed link message i
 * crosses; link id = (node*ndims + d)*2 + (negative ? 1 : 0).  Returns the hops
 * of the longest route. */
uint32_t or_route_bytes(const uint32_t* dims, uint32_t ndims, size_t n, const int32_t* src, const int32_t* dst,
                        const uint64_t* bytes, uint64_t* link_bytes);

#ifdef __cplusplus
}
#endif
#endif
