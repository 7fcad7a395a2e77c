/*
 * bine_hip.h -- C ABI of libbine_hip.so: Bine-tree collectives over P virtual
 * ranks that live as HBM partitions of one AMD Instinct MI355X (gfx950).
 *
 * Drop-in boundary for the MPI-layer data path of HLC-Lab/bine_trees_fugaku.
 * Each collective keeps the argument order of the PMPI-interposed entry point
 * it replaces (lib/libbine.cc), with three substitutions:
 *   - MPI_Comm      -> bine_vcomm_t (a virtual communicator: torus dims, ports,
 *                      per-collective algorithm family/variant/distance);
 *   - MPI_Datatype  -> bine_datatype_t, MPI_Op -> bine_op_t (SUM, and the
 *                      bench's no-op VOID operator, bench/bench.cc:240-256);
 *   - a trailing `stream` (hipStream_t passed as void*; NULL = default stream).
 *
 * Buffer convention: a buffer argument is a device pointer to P consecutive
 * per-rank partitions, rank r's partition at buf + r * (per-rank bytes), where
 * the per-rank bytes are the MPI per-rank buffer size of that argument
 * (e.g. count * sizeof(type) for MPI_Allreduce).  Arguments that MPI only reads
 * or writes at the root (MPI_Scatter sendbuf, MPI_Gather recvbuf, MPI_Reduce
 * recvbuf) are a single partition: the root's.
 *
 * All calls are stream-ordered and return a bine_status_t.  Preconditions the
 * reference asserts (lib/libbine.cc:1035-1036, 1085, 148, lib/libbine_coll.cc:204)
 * are reported as BINE_ERR_ARG / BINE_ERR_UNSUPPORTED instead of aborting.
 * Schedules are compiled on the first call for a given (collective, count,
 * type, op, root, in place or not) and cached in the vcomm; a call with a
 * cached schedule does not allocate (a per-step run also needs its
 * temporaries, allocated on its first run) and may be captured into a HIP
 * graph.  Such a graph stays valid while that schedule stays cached: the
 * cache keeps the 4096 most recently used schedules (see
 * bine_vcomm_set_cache_limit), and bine_vcomm_trim empties it; batches hold
 * their own schedules.
 *
 * MPI_IN_PLACE: bine_allreduce with sendbuf == recvbuf reads the input from
 * recvbuf.  It runs the per-step schedule compiled for that aliasing (every
 * algorithm: B and B_CONT drop their initial copy, L keeps its running value
 * off RECV at the first step); other collectives reject overlapping buffers.
 *
 * Per-step runs of one schedule share its temporaries.  Outside graph
 * capture, each such run waits on the GPU for the previous one, whatever
 * streams they were issued on; calls captured into graphs must be ordered by
 * the caller.
 */
#ifndef BINE_HIP_H
#define BINE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BINE_OK = 0,
    BINE_ERR_ARG = 1,         /* invalid argument / violated precondition */
    BINE_ERR_UNSUPPORTED = 2, /* configuration the reference schedule does not realise */
    BINE_ERR_HIP = 3,         /* HIP runtime error */
    BINE_ERR_NOMEM = 4,
    BINE_ERR_NO_DEVICE = 5
} bine_status_t;

/* MPI_INT32_T, MPI_INT64_T, MPI_FLOAT, MPI_DOUBLE, MPI_CHAR */
typedef enum { BINE_INT32 = 0, BINE_INT64 = 1, BINE_FLOAT = 2, BINE_DOUBLE = 3, BINE_INT8 = 4 } bine_datatype_t;
/* MPI_SUM, and the bench's VOID op (MPI_Reduce_local leaves inout unchanged) */
typedef enum { BINE_OP_SUM = 0, BINE_OP_VOID = 1 } bine_op_t;
/* bine_algo_family_t, lib/libbine_common.h:75-86 */
typedef enum {
    BINE_FAMILY_DEFAULT = 0, BINE_FAMILY_BINE = 1, BINE_FAMILY_RECDOUB = 2, BINE_FAMILY_BRUCK = 3, BINE_FAMILY_RING = 4
} bine_family_t;
/* bine_distance_type_t, lib/libbine_common.h:65-68 */
typedef enum { BINE_DISTANCE_INCREASING = 0, BINE_DISTANCE_DECREASING = 1 } bine_distance_t;
/* bine_allreduce_algo_t, lib/libbine_common.h:93-99 */
typedef enum {
    BINE_ALLREDUCE_L = 0, BINE_ALLREDUCE_B = 1, BINE_ALLREDUCE_REDUCE_BCAST = 2, BINE_ALLREDUCE_B_CONT = 3,
    BINE_ALLREDUCE_B_COALESCE = 4
} bine_allreduce_algo_t;
/* bine_allgather_algo_t, lib/libbine_common.h:101-106 */
typedef enum {
    BINE_ALLGATHER_CONT_PERMUTE = 0, BINE_ALLGATHER_CONT_SEND = 1, BINE_ALLGATHER_BLOCKS = 2,
    BINE_ALLGATHER_GATHER_BCAST = 3
} bine_allgather_algo_t;
/* bine_bcast_algo_t, lib/libbine_common.h:115-119 */
typedef enum { BINE_BCAST_BINOMIAL_TREE = 0, BINE_BCAST_BINOMIAL_TREE_TMPBUF = 1, BINE_BCAST_SCATTER_ALLGATHER = 2 } bine_bcast_algo_t;
/* bine_reduce_algo_t, lib/libbine_common.h:137-140 */
typedef enum { BINE_REDUCE_BINOMIAL_TREE = 0, BINE_REDUCE_REDUCE_SCATTER_GATHER = 1 } bine_reduce_algo_t;

/* How a collective executes on the GPU.  Both produce bit-identical results:
 * STEP launches one gather-reduce kernel per schedule step (the reference's
 * message structure); FUSED compiles the whole schedule into one launch that
 * reads every input once and writes every output once, replaying each output
 * element's reduction tree in the reference's association order. */
typedef enum { BINE_EXEC_AUTO = 0, BINE_EXEC_STEP = 1, BINE_EXEC_FUSED = 2 } bine_exec_t;

/* One collective's configuration: bine_*_config_t (lib/libbine_common.h:142-196).
 * `algo` takes the collective's own enum above. */
typedef struct {
    int32_t family;   /* bine_family_t */
    int32_t algo;     /* per-collective algorithm enum */
    int32_t distance; /* bine_distance_t */
} bine_coll_config_t;

/* bine_env_t (lib/libbine_common.h:246-263) minus the uTofu/prealloc fields. */
typedef struct {
    uint32_t dimensions[3];
    uint32_t dimensions_num;
    uint32_t num_ports;
    bine_coll_config_t allreduce, allgather, reduce_scatter, bcast, alltoall, scatter, gather, reduce;
    int32_t exec; /* bine_exec_t */
    int32_t device; /* HIP device ordinal the vcomm runs on */
} bine_vcomm_config_t;

typedef struct bine_vcomm_s* bine_vcomm_t;

/* Defaults of init_env (lib/libbine.cc:28-73): 1D torus of `size` ranks, one
 * port, family BINE, allreduce B, allgather CONT_PERMUTE + DECREASING, bcast
 * and gather DECREASING, everything else INCREASING; exec AUTO, device 0. */
int bine_vcomm_config_init(bine_vcomm_config_t* cfg, uint32_t size);
/* Overlay the LIBBINE_* environment variables read by read_env (lib/libbine.cc:75-575):
 * LIBBINE_DIMENSIONS, LIBBINE_NUM_PORTS, LIBBINE_<COLL>_ALGO_FAMILY / _ALGO / _DISTANCE,
 * plus LIBBINE_EXEC = AUTO|STEP|FUSED. */
int bine_vcomm_config_from_env(bine_vcomm_config_t* cfg);
/* Replaces `new BineCommon(comm, env)` (lib/libbine_common.cc:116-179). */
int bine_vcomm_create(const bine_vcomm_config_t* cfg, bine_vcomm_t* out);
int bine_vcomm_destroy(bine_vcomm_t comm);
/* MPI_Comm_size on the virtual communicator. */
int bine_vcomm_size(bine_vcomm_t comm, uint32_t* size);
/* Replace the execution mode of an existing vcomm (cached schedules are kept). */
int bine_vcomm_set_exec(bine_vcomm_t comm, int32_t exec);
/* Plan cache (no counterpart in the reference, whose bitmap and tree caches are never evicted,
 * lib/libbine_common.cc:1671-1675, lib/libbine_coll.cc:334-389).  At most max_plans schedules stay
 * cached, least recently used evicted first (0 = no bound; the default is 4096).  When a per-step
 * run cannot allocate its temporaries, the temporaries of schedules that only the cache holds are
 * released, least recently used first, and the allocation is retried. */
int bine_vcomm_set_cache_limit(bine_vcomm_t comm, uint32_t max_plans);
/* Drop every cached schedule (batches keep theirs). */
int bine_vcomm_trim(bine_vcomm_t comm);
/* Schedules cached and bytes of per-step temporaries they hold. */
int bine_vcomm_cache_info(bine_vcomm_t comm, uint32_t* plans, uint64_t* tmp_bytes);

/* MPI_Allreduce, lib/libbine.cc:1012 */
int bine_allreduce(const void* sendbuf, void* recvbuf, int count, bine_datatype_t datatype, bine_op_t op,
                   bine_vcomm_t comm, void* stream);
/* bine_allreduce for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments and results as bine_allreduce, in place included, but it never compiles, allocates
 * or frees: the schedule must already be cached (compile it at set-up with one bine_allreduce
 * call of the same count, type, op and in-place-ness; otherwise BINE_ERR_UNSUPPORTED), and a
 * per-step run takes its temporaries from a 64 MiB device array the library holds per GPU until
 * the process ends (a larger need is BINE_ERR_ARG).  Every call on one GPU shares that array and
 * leaves its contents behind (the schedules write it before reading it); calls that overlap in
 * time on one GPU race, so issue them on one stream or order them. */
int bine_allreduce_noalloc(const void* sendbuf, void* recvbuf, int count, bine_datatype_t datatype, bine_op_t op,
                           bine_vcomm_t comm, void* stream);
/* MPI_Reduce_scatter, lib/libbine.cc:1064 (equal recvcounts, as the reference assumes :1065) */
int bine_reduce_scatter(const void* sendbuf, void* recvbuf, const int recvcounts[], bine_datatype_t datatype,
                        bine_op_t op, bine_vcomm_t comm, void* stream);
/* MPI_Allgather, lib/libbine.cc:1168 */
int bine_allgather(const void* sendbuf, int sendcount, bine_datatype_t sendtype, void* recvbuf, int recvcount,
                   bine_datatype_t recvtype, bine_vcomm_t comm, void* stream);
/* MPI_Bcast_f, lib/libbine.cc:1252 */
int bine_bcast(void* buffer, int count, bine_datatype_t datatype, int root, bine_vcomm_t comm, void* stream);
/* MPI_Alltoall, lib/libbine.cc:1291 */
int bine_alltoall(const void* sendbuf, int sendcount, bine_datatype_t sendtype, void* recvbuf, int recvcount,
                  bine_datatype_t recvtype, bine_vcomm_t comm, void* stream);
/* MPI_Scatter, lib/libbine.cc:1312 */
int bine_scatter(const void* sendbuf, int sendcount, bine_datatype_t sendtype, void* recvbuf, int recvcount,
                 bine_datatype_t recvtype, int root, bine_vcomm_t comm, void* stream);
/* MPI_Gather, lib/libbine.cc:1376 */
int bine_gather(const void* sendbuf, int sendcount, bine_datatype_t sendtype, void* recvbuf, int recvcount,
                bine_datatype_t recvtype, int root, bine_vcomm_t comm, void* stream);
/* MPI_Reduce, lib/libbine.cc:1440 */
int bine_reduce(const void* sendbuf, void* recvbuf, int count, bine_datatype_t datatype, bine_op_t op, int root,
                bine_vcomm_t comm, void* stream);

/* ---- schedule helpers (lib/libbine_coll.h:8-33), plain C for golden-vector tests ---- */
/* compute_peers, lib/libbine_coll.cc:443: peers[num_steps] */
int bine_compute_peers(const uint32_t* dims, uint32_t ndims, uint32_t rank, int port, int family, uint32_t* peers);
/* compute_step_info, lib/libbine_coll.cc:306: d[num_steps], step_in_d[num_steps] */
int bine_compute_step_info(const uint32_t* dims, uint32_t ndims, uint32_t port, uint32_t* d, uint32_t* step_in_d);
/* get_peer_c, lib/libbine_coll.cc:89-107: the coordinates of coord's peer at schedule step `step` of `port`
 * (coord and coord_peer hold ndims entries) */
int bine_get_peer(const uint32_t* dims, uint32_t ndims, const int32_t* coord, uint32_t step, uint32_t port,
                  int family, int32_t* coord_peer);
/* get_mirroring_port, lib/libbine_coll.cc:44-52: the first of num_ports ports whose schedule is mirrored
 * (distances of opposite sign), or -1 (also for ndims outside 1..3) */
int bine_get_mirroring_port(int num_ports, uint32_t ndims);
/* BineBitmapCalculator::block_step, lib/libbine_common.cc:1491-1510, 2036-2045: block_step[P] */
int bine_block_step(const uint32_t* dims, uint32_t ndims, uint32_t rank, uint32_t port, int family,
                    uint32_t* block_step);
/* get_tree, lib/libbine_coll.cc:334: five arrays of P entries */
int bine_get_tree(const uint32_t* dims, uint32_t ndims, uint32_t root, uint32_t port, int family, int distance,
                  uint32_t* parent, uint32_t* reached_at_step, uint32_t* remapped_ranks,
                  uint32_t* remapped_ranks_max, uint32_t* subtree_roots);
/* remap_rank, lib/libbine_coll_alltoall.cc:372 */
uint32_t bine_remap_rank(uint32_t rank, uint32_t size);
/* reduce_local, lib/libbine_coll.cc:153-207 (2-argument form: inout = in + inout), on device.
 * SUM, or VOID (no effect, as the reference's INT32 branch treats a non-SUM op); any other op is
 * BINE_ERR_ARG where the reference exits.  All five element types (the reference has no
 * int64/double).  inbuf may be inoutbuf itself; a partial overlap is BINE_ERR_ARG. */
int bine_reduce_local(const void* inbuf, void* inoutbuf, int count, bine_datatype_t datatype, bine_op_t op,
                      void* stream);
/* reduce_local, lib/libbine_coll.cc:209-267 (3-argument form: out = a + b), on device.  VOID writes
 * nothing.  outbuf may be either input itself; a partial overlap is BINE_ERR_ARG. */
int bine_reduce_local3(const void* inbuf_a, const void* inbuf_b, void* outbuf, int count,
                       bine_datatype_t datatype, bine_op_t op, void* stream);

/* ---- introspection: bytes a call moves (for roofline accounting) ---- */
typedef struct {
    uint64_t launches;        /* kernel launches per call */
    uint64_t items;           /* gather-reduce work items over all launches */
    uint64_t bytes_read;      /* HBM bytes read by the kernels (schedule bytes) */
    uint64_t bytes_written;   /* HBM bytes written by the kernels */
    uint64_t max_tree_leaves; /* widest reduction tree replayed by one item */
} bine_plan_stats_t;

/* Collective codes for plan queries and batches. */
typedef enum {
    BINE_COLL_ALLREDUCE = 0, BINE_COLL_REDUCE_SCATTER = 1, BINE_COLL_ALLGATHER = 2, BINE_COLL_BCAST = 3,
    BINE_COLL_ALLTOALL = 4, BINE_COLL_SCATTER = 5, BINE_COLL_GATHER = 6, BINE_COLL_REDUCE = 7
} bine_coll_t;

/* Compile (or fetch) the schedule of a collective without running it. */
int bine_plan_stats(bine_vcomm_t comm, int32_t coll, int count, bine_datatype_t datatype, bine_op_t op, int root,
                    int32_t exec, bine_plan_stats_t* out);

/* Time candidate launch shapes (vectors per thread, workgroup size, persistent
 * grid, streaming) of one collective's cached schedule on the caller's buffers
 * and keep the fastest for later calls.  Runs the collective (iters + 1) times
 * per candidate (bine_tuning_candidates lists them; 28 on an MI355X), then,
 * where the schedule holds reductions with several outputs, at the three
 * fastest shapes in both two-phase modes, with reduction items of one and two
 * tiles (twelve more; see bine_batch_set_tuning).  The
 * work list in place before the call stays allocated until the vcomm is
 * destroyed, so graphs captured earlier keep running the old launch shape.
 * Must not run concurrently with other calls on the same vcomm. */
int bine_plan_autotune(bine_vcomm_t comm, int32_t coll, const void* sendbuf, void* recvbuf, int count,
                       bine_datatype_t datatype, bine_op_t op, int root, int32_t exec, void* stream, int iters);

/* Host-only introspection: compile a schedule without a GPU and return its
 * lowered work (the arrays one call would upload).  Call once with NULL
 * arrays to size them.  items: info->n_items records of 32 bytes (bine
 * gather-reduce items); pool: info->n_pool records of 8 bytes; launches:
 * info->n_launches pairs (first item, item count).  in_place: the allreduce
 * schedule for sendbuf == recvbuf (SEND has no extent).  two_phase: 0, or the
 * two-phase mode 1 or 2 of bine_batch_set_tuning, plus 4 for reduction items of
 * two tiles (the flags word shifted right by one).  No HIP call is made. */
typedef struct {
    uint64_t n_items, n_pool, n_launches;
    int64_t extent[4];  /* per-rank elements of SEND, RECV, TMP, TMP2 */
    int32_t single[4];  /* buffer is one partition (the root's) */
    uint32_t P;
    int32_t fused_ok;
} bine_plan_dump_info_t;
int bine_plan_dump(const bine_vcomm_config_t* cfg, int32_t coll, int count, bine_datatype_t datatype, bine_op_t op,
                   int root, int32_t in_place, int32_t exec, int64_t tile_elems, int32_t two_phase,
                   bine_plan_dump_info_t* info, void* items, void* pool, uint32_t* launches);

/* ---- batching: many collectives in one launch (no counterpart in the reference:
 * it replaces the per-call read_env/BlockInfo/tmpbuf set-up, lib/libbine.cc:1013-1050,
 * for callers that issue many buckets at once) ---- */
typedef struct {
    int32_t coll;        /* bine_coll_t */
    const void* sendbuf; /* layout as in the single-call entry point */
    void* recvbuf;
    int32_t count;       /* count (allreduce, bcast, reduce) or per-rank block count (others) */
    int32_t datatype;    /* bine_datatype_t */
    int32_t op;          /* bine_op_t */
    int32_t root;
} bine_batch_item_t;

typedef struct bine_batch_s* bine_batch_t;

/* Compile `n` collectives on one vcomm into one fused work list.  The buffers
 * are bound here and must stay allocated, and the vcomm alive, until
 * bine_batch_destroy.  An allreduce item with sendbuf == recvbuf runs in place,
 * per step, after the fused lists.  Retuning a batch (set_tuning, autotune)
 * replaces its work lists: graphs the caller captured from bine_batch_launch
 * before that must be captured again (bine_batch_launch_graph does so itself). */
int bine_batch_create(bine_vcomm_t comm, const bine_batch_item_t* items, int n, bine_batch_t* out);
/* Run the whole batch: one launch per element type present (two for a list tuned to two phases). */
int bine_batch_launch(bine_batch_t batch, void* stream);
/* Same work captured once into a HIP graph and replayed. */
int bine_batch_launch_graph(bine_batch_t batch, void* stream);
/* Time candidate (vector width, workgroup, persistence, streaming) shapes and keep the fastest, chosen
 * separately for each element type's work list; a list holding reductions with several outputs is also
 * timed at its three fastest shapes in both two-phase modes, with reduction items of one and two tiles. */
int bine_batch_autotune(bine_batch_t batch, void* stream, int iters);
int bine_batch_stats(bine_batch_t batch, bine_plan_stats_t* out);
/* Launch shape: 16-byte vectors per thread (1, 2, 4), threads per workgroup (64..512, power of two),
 * persistent grid (0 = one workgroup per item), flags: bit 0 (BINE_TUNE_NONTEMPORAL) streaming loads and
 * stores; bits 1-2 the two-phase mode for reductions with several outputs (an allreduce's blocks):
 * 0 one launch, the reductions writing every output; BINE_TUNE_TWO_PHASE (2) the reductions write their
 * first output, then a second launch copies it to the others, so the memory serves a read phase and a
 * write phase instead of both at once (same bits); BINE_TUNE_TWO_PHASE_BY_TILE (4) the same with the
 * copies issued tile by tile across the blocks instead of block by block; BINE_TUNE_DOUBLE_TILES (8), with a
 * two-phase mode, gives the reduction launch items of two tiles (half as many).  get returns the last shape
 * set, or after autotuning the shape of the list that takes longest, or before either the first list's
 * starting shape (4, 64, 0, streaming where reductions of 64 leaves or more over blocks of two whole
 * tiles or more hold 90 % of the list's bytes in at least 1024 tiles; 1, 256, 0, 0 otherwise); set
 * applies one shape to every list. */
enum { BINE_TUNE_NONTEMPORAL = 1, BINE_TUNE_TWO_PHASE = 2, BINE_TUNE_TWO_PHASE_BY_TILE = 4, BINE_TUNE_DOUBLE_TILES = 8 };
int bine_batch_get_tuning(bine_batch_t batch, int32_t* vpt, int32_t* threads, int32_t* persistent_wgs,
                          int32_t* flags);
int bine_batch_set_tuning(bine_batch_t batch, int32_t vpt, int32_t threads, int32_t persistent_wgs,
                          int32_t flags);
int bine_batch_destroy(bine_batch_t batch);
/* The autotuner's candidate launch shapes on `device`: up to cap rows of
 * {vectors per thread, threads, persistent workgroups, flags}; *n = total.  The
 * two-phase trials an autotune adds are not listed. */
int bine_tuning_candidates(int device, int32_t* shapes, int cap, int* n);
/* Average time (ms) of one batch launch at each of n shapes (rows as above):
 * one warm-up launch, then iters timed ones.  The batch keeps its shape. */
int bine_batch_time_shapes(bine_batch_t batch, void* stream, int iters, const int32_t* shapes, int n, float* ms);

/* ---- torus link-traffic simulation --------------------------------------
 * Replaces simulation/simulator.py (run_collectives :125-177, get_peer :70-76)
 * and extends it with per-link accounting, which the reference lacks: every
 * message of a schedule step is routed over the torus (dimension-ordered,
 * shorter way round, positive way on a tie) and the GPU counts, per directed
 * link, how many messages cross it.  Ranks sit on the torus in the library's
 * row-major order (lib/libbine_common.cc:206-246), except for the
 * simulator.py algorithm, whose own id order (simulator.py:11-26) is kept.
 *
 * A case is (torus, collective, algorithm).  Its schedule is a list of step
 * classes; every message of a class carries the share num/den of the per-rank
 * vector and the class repeats rep times, so the bytes a link carries in a
 * class at S bytes per rank are count * floor(S * num / den), an exact integer. */
typedef enum {
    BINE_SIM_BINE = 0,    /* Bine schedules (torus peers for allreduce/RS/AG, 1D trees and butterflies otherwise) */
    BINE_SIM_RECDOUB = 1, /* recursive doubling / binomial trees */
    BINE_SIM_RING = 2,    /* ring reduce-scatter / allgather / allreduce */
    BINE_SIM_BRUCK = 3,   /* Bruck alltoall (lib/libbine_coll_alltoall.cc:18-107) */
    BINE_SIM_SIMPY = 4    /* simulation/simulator.py's own peers and id order (equal dimensions) */
} bine_sim_algo_t;

typedef struct {
    uint32_t dimensions[3];
    uint32_t dimensions_num;
    int32_t coll; /* bine_coll_t */
    int32_t algo; /* bine_sim_algo_t */
} bine_sim_case_t;

typedef struct bine_sim_s* bine_sim_t;

/* Step classes of one case (host only).  Arrays may be NULL; otherwise they
 * hold at least *nclasses entries (call once with NULL arrays to size them;
 * a case has at most 2 * 40 classes). */
int bine_sim_classes(const bine_sim_case_t* c, uint32_t* nclasses, uint32_t* kind, uint32_t* d, uint32_t* k,
                     uint32_t* num, uint32_t* den, uint32_t* rep);
/* Destination of every rank's message in class cls (host only; P entries,
 * UINT32_MAX for a rank that sends nothing).  For BINE_SIM_SIMPY, ranks and
 * destinations are simulator.py ids. */
int bine_sim_peers(const bine_sim_case_t* c, uint32_t cls, uint32_t* dest);
/* Upload n cases and allocate their link counters on `device`. */
int bine_sim_create(const bine_sim_case_t* cases, int n, int device, bine_sim_t* out);
/* Count every class of every case in one GPU pass: integer difference updates
 * per dimension run of each message, then prefix sums along every ring. */
int bine_sim_run(bine_sim_t sim, void* stream);
int bine_sim_case_info(bine_sim_t sim, int i, uint32_t* nclasses, uint32_t* nlinks);
/* Per class of case i (synchronous copies after the last run's stream): share, repetitions, the most
 * loaded link's message count, the sum of link crossings, the longest route. */
int bine_sim_class_stats(bine_sim_t sim, int i, uint32_t* num, uint32_t* den, uint32_t* rep, uint32_t* max_mult,
                         uint64_t* sum_mult, uint32_t* max_hops);
/* Message count of every directed link (nlinks entries) in class cls of case i. */
int bine_sim_link_counts(bine_sim_t sim, int i, uint32_t cls, uint32_t* counts);
int bine_sim_destroy(bine_sim_t sim);

/* ---- cost-model network of the algorithm selector -------------------------
 * No counterpart in the reference, whose selection is the fixed defaults of
 * lib/libbine.cc:28-73 (and Fujitsu MPI's hand-written fjmpi-dectree.conf).
 * An fp32 multilayer perceptron: dims = {inputs, hidden..., outputs}; ReLU
 * after every hidden layer, linear output.  Inputs x (n x inputs), targets t
 * and masks (n x outputs) are row-major fp32 device buffers owned by the
 * caller; the loss is sum(mask * (y - t)^2) / count(mask > 0).  Training uses
 * Adam (beta 0.9 / 0.999, eps 1e-8).  Calls are stream-ordered. */
typedef struct bine_mlp_s* bine_mlp_t;

/* Weights uniform in +-1/sqrt(fan_in) from splitmix64(seed), zero biases; buffers for batches up to max_batch rows. */
int bine_mlp_create(const int32_t* dims, int32_t ndims, int32_t max_batch, uint64_t seed, int device, bine_mlp_t* out);
int bine_mlp_num_params(bine_mlp_t mlp, uint64_t* n);
/* Flat parameters, layer by layer: W (outputs x inputs, row-major) then b. Synchronous. */
int bine_mlp_get_params(bine_mlp_t mlp, float* host);
int bine_mlp_set_params(bine_mlp_t mlp, const float* host); /* also resets the Adam state */
int bine_mlp_forward(bine_mlp_t mlp, const float* x, int32_t n, float* y, void* stream);
/* Loss (one float, device) and its gradient w.r.t. every parameter (device, num_params floats); no update. */
int bine_mlp_gradients(bine_mlp_t mlp, const float* x, const float* t, const float* mask, int32_t n, float* grads,
                       float* loss, void* stream);
/* One Adam step on the batch; loss (device float, may be NULL) is the loss before the step. */
int bine_mlp_train_step(bine_mlp_t mlp, const float* x, const float* t, const float* mask, int32_t n, float lr,
                        float* loss, void* stream);
int bine_mlp_destroy(bine_mlp_t mlp);

/* Last error message of the calling thread (static storage). */
const char* bine_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* BINE_HIP_H */
