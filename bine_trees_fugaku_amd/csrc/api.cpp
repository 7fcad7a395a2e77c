// api.cpp -- C ABI of libbine_hip.so (declared in include/bine_hip.h).
//
// A vcomm owns the torus, the per-collective configuration and a cache of
// compiled schedules.  A schedule is compiled once per (collective, count,
// type, op, root): the step program is built, checked symbolically, lowered
// to work items and uploaded with its temporary buffers.  A call then only
// validates its buffers and enqueues kernel launches on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "bine_hip.h"
#include "gred.hpp"
#include "plan.hpp"
#include "status.hpp"

namespace bine {
hipError_t launch_gred(int dtype, int max_lg, const GredItem* items, uint32_t n, const PoolEntry* pool,
                       const BufTable& tab, const BufTable* tabs, bool batch, uint32_t allow_vec,
                       const LaunchShape& sh, hipStream_t st);
hipError_t launch_grun(uint32_t esize, const GredItem* items, uint32_t n, const PoolEntry* pool, const BufTable& tab,
                       const BufTable* tabs, bool batch, uint32_t allow_vec, const LaunchShape& sh, hipStream_t st);
hipError_t launch_reduce_local(int dtype, const void* a, const void* b, void* out, uint64_t n, hipStream_t st);
size_t scratch_bytes();
hipError_t scratch_init(int device);
void* scratch_ptr(int device);
}  // namespace bine

using namespace bine;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return fail(BINE_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t dtsize(int dt) {
    switch (dt) {
    case BINE_INT32: return 4;
    case BINE_INT64: return 8;
    case BINE_FLOAT: return 4;
    case BINE_DOUBLE: return 8;
    case BINE_INT8: return 1;
    default: return 0;
    }
}

struct DevArrays {
    Lowered low;
    GredItem* items = nullptr;
    PoolEntry* pool = nullptr;
    int max_lg = 0;
    int64_t tile_elems = 0;
    LaunchShape shape;
    void release() {
        if (items) (void)hipFree(items);
        if (pool) (void)hipFree(pool);
        items = nullptr;
        pool = nullptr;
        low = Lowered();
    }
};

// Device arrays that a HIP graph captured earlier may still name: kept until their owner goes away.
struct Retired {
    std::vector<void*> ptrs;
    void keep(DevArrays* d) {
        if (d->items) ptrs.push_back(d->items);
        if (d->pool) ptrs.push_back(d->pool);
        d->items = nullptr;
        d->pool = nullptr;
    }
    ~Retired() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// Copy a lowered work list to the device.  The new arrays replace d's only once every allocation and copy
// has succeeded; on failure d is left as it was.  The old arrays are freed, or kept in `retire` when a
// captured graph may name them.
int install(Lowered&& low, int max_lg, int64_t tile_elems, const LaunchShape& sh, DevArrays* d, Retired* retire) {
    DevArrays n;
    n.low = std::move(low);
    n.max_lg = max_lg;
    n.tile_elems = tile_elems;
    n.shape = sh;
    hipError_t e = hipSuccess;
    if (!n.low.items.empty()) {
        e = hipMalloc(&n.items, n.low.items.size() * sizeof(GredItem));
        if (e == hipSuccess)
            e = hipMemcpy(n.items, n.low.items.data(), n.low.items.size() * sizeof(GredItem), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !n.low.pool.empty()) {
        e = hipMalloc(&n.pool, n.low.pool.size() * sizeof(PoolEntry));
        if (e == hipSuccess)
            e = hipMemcpy(n.pool, n.low.pool.data(), n.low.pool.size() * sizeof(PoolEntry), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n.release();
        return fail(BINE_ERR_HIP, std::string("work-list upload: ") + hipGetErrorString(e));
    }
    if (retire) retire->keep(d);
    d->release();
    *d = std::move(n);
    n.items = nullptr;
    n.pool = nullptr;
    return BINE_OK;
}

int max_lg_of(const std::vector<std::vector<Group>>& gs) {
    int m = 0;
    for (const auto& v : gs)
        for (const auto& g : v) m = std::max<int>(m, g.lg);
    return m;
}

int upload(const std::vector<std::vector<Group>>& gs, size_t ds, const LaunchShape& sh, uint32_t bucket, DevArrays* d,
           Retired* retire = nullptr) {
    Lowered low;
    lower_groups(gs, shallow_tile_elems(sh, ds), (int)(16 / ds), bucket, ds, &low, deep_tile_elems(sh, ds), split_lg(sh),
                 nullptr, nullptr, sh.two_phase, sh.double_tiles ? 2 : 1);
    return install(std::move(low), max_lg_of(gs), shallow_tile_elems(sh, ds), sh, d, retire);
}

struct DevPlan {
    CompiledPlan cp;
    int dtype = 0;
    size_t ds = 0;
    DevArrays step, fused;
    Retired retired;                      // work lists replaced by an autotune (graphs may name them)
    void* tmp[2] = {nullptr, nullptr};  // BUF_TMP, BUF_TMP2 workspaces, allocated on the first per-step run
    size_t tmp_bytes[2] = {0, 0};         // 0: the per-step schedule never touches that buffer
    // Per-step runs that use the temporaries are ordered across streams: each run waits for the event the
    // previous one recorded (outside graph capture).
    hipEvent_t tmp_done = nullptr;
    bool tmp_recorded = false;
    uint64_t last_use = 0;  // LRU stamp
    int pins = 0;           // batches that hold this plan (never evicted while > 0)
    ~DevPlan() {
        step.release();
        fused.release();
        for (void* p : tmp)
            if (p) (void)hipFree(p);
        if (tmp_done) (void)hipEventDestroy(tmp_done);
    }
    size_t tmp_total() const { return tmp_bytes[0] + tmp_bytes[1]; }
};

struct PlanKey {
    int coll, dtype, op, root;
    int64_t count;
    bool in_place;
    bool operator==(const PlanKey& o) const {
        return coll == o.coll && dtype == o.dtype && op == o.op && root == o.root && count == o.count &&
               in_place == o.in_place;
    }
};
struct PlanKeyHash {
    size_t operator()(const PlanKey& k) const {
        uint64_t h = (uint64_t)k.count * 0x9e3779b97f4a7c15ull;
        h ^= ((uint64_t)k.coll << 48) ^ ((uint64_t)k.dtype << 40) ^ ((uint64_t)k.op << 32) ^ (uint32_t)k.root ^
             ((uint64_t)k.in_place << 56);
        return (size_t)(h ^ (h >> 29));
    }
};

// Default launch shape: 256 threads, one 16-byte vector each (4 KiB tiles).
LaunchShape default_shape() { return LaunchShape{1, 256, 0, 0}; }
int64_t default_tile_elems(size_t ds) { return shallow_tile_elems(default_shape(), ds); }

// The shape a fused work list starts with before any tuning.  Four vectors per thread, one-wave workgroups and
// streaming loads and stores (the one-launch shape the autotuner keeps on long blocks) when the reductions of 64
// leaves or more hold at least 90 % of the list's bytes in blocks of two whole tiles or more at that shape, and
// those blocks give at least 1024 tiles, so every CU gets several.  Untuned, that shape moved the 512-rank fp64
// allreduce of 4 MiB per rank at 6.1 TB/s against 4.5 with the general shape; on shorter blocks or fewer tiles
// (the same torus at 256 KiB per rank, or 64 ranks of 2 MiB) it ran 10-67 % slower than the general shape, which
// every other list starts at.
LaunchShape initial_shape(const std::vector<std::vector<Group>>& gs, size_t ds) {
    const LaunchShape wide{4, 64, 0, 1};
    const int64_t tile = deep_tile_elems(wide, ds);
    uint64_t bytes = 0, long_bytes = 0, tiles = 0;
    for (const auto& v : gs)
        for (const Group& g : v) {
            if (g.lg == 0) continue;
            const uint64_t b = (uint64_t)g.n * (g.leaves.size() + g.outs.size()) * ds;
            bytes += b;
            if (g.lg >= 6 && g.n >= 2 * tile) {
                long_bytes += b;
                tiles += (uint64_t)(g.n / tile);
            }
        }
    if (tiles >= 1024 && long_bytes * 10 >= bytes * 9) return wide;
    return default_shape();
}

}  // namespace

// Error text shared with the other C-ABI translation units (sim, selector).
int bine::record_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Default bound of a vcomm's plan cache (bine_vcomm_set_cache_limit changes it).
constexpr uint32_t kDefaultMaxPlans = 4096;

struct bine_vcomm_s {
    bine_vcomm_config_t cfg;
    Torus t;
    std::mutex mu;
    // Compiled schedules, least recently used evicted beyond max_plans.  A batch or a call in flight holds
    // its plan through the shared pointer, so eviction never frees a plan still in use by the host.
    std::unordered_map<PlanKey, std::shared_ptr<DevPlan>, PlanKeyHash> plans;
    uint32_t max_plans = kDefaultMaxPlans;
    uint64_t clock = 0;
    hipError_t scratch_err = hipSuccess;  // resolving the allocation-free entry point's workspace at creation
};

namespace {

// Drop least recently used plans until at most `keep` remain (caller holds comm->mu).  A plan that a batch or
// a call in progress still holds lives on until they drop it.  The last holder's release frees the device
// arrays with hipFree, which synchronizes the device first, so a launch of an earlier call that is still
// running finishes before its work list goes away.
void evict_lru(bine_vcomm_t comm, size_t keep) {
    while (comm->plans.size() > keep) {
        auto victim = comm->plans.begin();
        for (auto it = comm->plans.begin(); it != comm->plans.end(); ++it)
            if (it->second->last_use < victim->second->last_use) victim = it;
        comm->plans.erase(victim);
    }
}

// Compile (or fetch) the schedule for a request.
int get_plan(bine_vcomm_t comm, const CollRequest& req, int dtype, std::shared_ptr<DevPlan>* out) {
    const size_t ds = dtsize(dtype);
    if (!ds) return fail(BINE_ERR_ARG, "unknown datatype");
    PlanKey key{req.coll, dtype, req.op, req.root, req.count, req.in_place};
    std::lock_guard<std::mutex> lock(comm->mu);
    auto it = comm->plans.find(key);
    if (it != comm->plans.end()) {
        it->second->last_use = ++comm->clock;
        *out = it->second;
        return BINE_OK;
    }
    HIPCHK(hipSetDevice(comm->cfg.device));
    Program prog;
    std::string err;
    if (!build_program(comm->t, comm->cfg, req, &prog, &err)) return fail(BINE_ERR_UNSUPPORTED, err);
    for (int b = 0; b < kNumBufs; ++b)
        if ((uint64_t)prog.extent[b] >= (1ull << 32)) return fail(BINE_ERR_UNSUPPORTED, "per-rank buffer exceeds 2^32 elements");
    auto dp = std::make_shared<DevPlan>();
    if (!compile_program(prog, true, &dp->cp, &err)) {
        if (req.in_place) err = "schedule cannot run in place (sendbuf == recvbuf): " + err;
        return fail(BINE_ERR_UNSUPPORTED, err);
    }
    dp->dtype = dtype;
    dp->ds = ds;
    int rc = upload(dp->cp.step_groups, ds, default_shape(), 0, &dp->step);
    if (rc) return rc;
    if (dp->cp.fused_ok && max_lg_of(std::vector<std::vector<Group>>(1, dp->cp.fused)) > 14) {
        dp->cp.fused_ok = false;  // the kernel's register stack holds trees of up to 2^14 leaves
        dp->cp.fused.clear();
        dp->cp.fused_why = "reduction tree wider than 2^14 leaves";
    }
    if (dp->cp.fused_ok) {
        std::vector<std::vector<Group>> one(1, dp->cp.fused);
        rc = upload(one, ds, initial_shape(one, ds), 0, &dp->fused);
        if (rc) return rc;
    }
    // temporaries are allocated only when some step touches them
    bool used[kNumBufs] = {false, false, false, false};
    for (const auto& launch : dp->cp.step_groups)
        for (const Group& g : launch) {
            for (const PoolEntry& e : g.leaves) used[e.rb >> 28] = true;
            for (const PoolEntry& e : g.outs) used[e.rb >> 28] = true;
        }
    for (int b = BUF_TMP; b <= BUF_TMP2; ++b)
        if (prog.extent[b] > 0 && used[b]) dp->tmp_bytes[b - BUF_TMP] = (size_t)prog.extent[b] * ds * (prog.single[b] ? 1 : prog.P);
    dp->last_use = ++comm->clock;
    *out = dp;
    if (comm->max_plans > 0) evict_lru(comm, comm->max_plans - 1);
    comm->plans.emplace(key, std::move(dp));
    return BINE_OK;
}

// Kernel launches issued for a lowered work list (a launch range with both
// kinds of items issues the reduction kernel and the copy-run kernel).
uint64_t kernel_launches(const Lowered& low) {
    uint64_t n = 0;
    for (size_t i = 0; i < low.launches.size(); ++i)
        n += (low.launches[i].second > low.runs[i] ? 1 : 0) + (low.runs[i] ? 1 : 0);
    return n;
}

// Release the temporaries of cached plans that nothing but the cache holds, least recently used first,
// until `need` bytes were freed (caller holds comm->mu).  Returns the bytes freed.
size_t release_idle_tmp(bine_vcomm_t comm, size_t need, const DevPlan* except) {
    std::vector<DevPlan*> idle;
    for (auto& kv : comm->plans)
        if (kv.second.get() != except && kv.second.use_count() == 1 && (kv.second->tmp[0] || kv.second->tmp[1]))
            idle.push_back(kv.second.get());
    std::sort(idle.begin(), idle.end(), [](const DevPlan* a, const DevPlan* b) { return a->last_use < b->last_use; });
    size_t freed = 0;
    for (DevPlan* p : idle) {
        if (freed >= need) break;
        for (int i = 0; i < 2; ++i)
            if (p->tmp[i]) {
                (void)hipFree(p->tmp[i]);
                p->tmp[i] = nullptr;
                freed += p->tmp_bytes[i];
            }
        p->tmp_recorded = false;
    }
    return freed;
}

// Allocate the per-step schedule's temporaries (the fused schedule needs none).
// Every rank keeps the reference's full-size temporary, so per-step execution
// of large rooted collectives can need P times the reference's per-process memory.
// When the device is out of memory, idle plans' temporaries are released and the allocation retried.
int ensure_tmp(bine_vcomm_t comm, DevPlan* dp) {
    std::lock_guard<std::mutex> lock(comm->mu);
    for (int i = 0; i < 2; ++i) {
        if (!dp->tmp_bytes[i] || dp->tmp[i]) continue;
        HIPCHK(hipSetDevice(comm->cfg.device));
        hipError_t e = hipMalloc(&dp->tmp[i], dp->tmp_bytes[i]);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            dp->tmp[i] = nullptr;
            if (release_idle_tmp(comm, dp->tmp_bytes[i], dp) > 0) e = hipMalloc(&dp->tmp[i], dp->tmp_bytes[i]);
        }
        if (e != hipSuccess) {
            dp->tmp[i] = nullptr;
            (void)hipGetLastError();
            return fail(BINE_ERR_HIP, "per-step execution needs " + std::to_string(dp->tmp_bytes[i] >> 20) +
                                          " MiB of temporaries: " + hipGetErrorString(e));
        }
    }
    if (dp->tmp_total() && !dp->tmp_done) HIPCHK(hipEventCreateWithFlags(&dp->tmp_done, hipEventDisableTiming));
    return BINE_OK;
}

bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs != hipStreamCaptureStatusNone;
}

// The device allocation holding [p, p + bytes) must contain the whole range.
int check_range(const void* p, size_t bytes, const char* what) {
    if (bytes == 0) return BINE_OK;
    if (!p) return fail(BINE_ERR_ARG, std::string(what) + " is NULL");
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    hipError_t e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(BINE_ERR_ARG, std::string(what) + " is not device memory allocated with HIP");
    }
    if ((const char*)p + bytes > (const char*)base + size)
        return fail(BINE_ERR_ARG, std::string(what) + " is smaller than P partitions of the per-rank size");
    return BINE_OK;
}

bool overlap(const void* a, size_t na, const void* b, size_t nb) {
    if (!a || !b || !na || !nb) return false;
    const char* x = (const char*)a;
    const char* y = (const char*)b;
    return x < y + nb && y < x + na;
}

size_t partition_bytes(const DevPlan* dp, int buf) {
    return (size_t)dp->cp.extent[buf] * dp->ds * (dp->cp.single[buf] ? 1 : dp->cp.P);
}

BufTable make_table(const DevPlan* dp, const void* send, void* recv) {
    BufTable t{};
    const void* bases[4] = {send, recv, dp->tmp[0], dp->tmp[1]};
    for (int b = 0; b < 4; ++b) {
        t.base[b] = (uint64_t)(uintptr_t)bases[b];
        t.stride[b] = dp->cp.single[b] ? 0 : (uint64_t)dp->cp.extent[b] * dp->ds;
    }
    return t;
}

// 16-byte vector accesses need 4-byte aligned addresses only (global memory runs in the
// unaligned access mode); partitions aligned below that (odd int8 counts) take the scalar path.
uint32_t vec_ok(const BufTable& t) {
    for (int b = 0; b < 4; ++b)
        if ((t.base[b] & 3) || (t.stride[b] & 3)) return 0;
    return 1;
}

// some buffer base or partition stride is not a multiple of 16 bytes
bool off16(const BufTable& t) {
    for (int b = 0; b < 4; ++b)
        if ((t.base[b] | t.stride[b]) & 15) return true;
    return false;
}

// The realigning kernels (aligned chunk stores built with lane shuffles) pay where stores dominate: an
// allgather-like list that writes many outputs per leaf it reads ran 16 % faster with them, while 1:1 copies
// ran 16 % slower and reductions 1.5 % slower than with plain unaligned 16-byte stores.
bool realign_pays(const Lowered& low, bool tables_off16) {
    return (low.unaligned_outs || tables_off16) && low.bytes_written >= 2 * low.bytes_read;
}

// allow_vec bit 0: vector items may run; the launch adds bit 1 (realigning kernels) where realign_pays.
int launch_arrays(const DevArrays& d, int dtype, const BufTable& tab, uint32_t allow_vec, hipStream_t st) {
    if ((allow_vec & 1u) && realign_pays(d.low, off16(tab))) allow_vec |= 2u;
    for (size_t li = 0; li < d.low.launches.size(); ++li) {
        const auto& l = d.low.launches[li];
        const uint32_t nrun = d.low.runs[li], nred = l.second - nrun;
        if (nred) {
            hipError_t e = launch_gred(dtype, d.max_lg, d.items + l.first, nred, d.pool, tab, nullptr, false,
                                       allow_vec, d.shape, st);
            if (e != hipSuccess) return fail(BINE_ERR_HIP, std::string("gather-reduce launch: ") + hipGetErrorString(e));
        }
        if (nrun) {
            hipError_t e = launch_grun((uint32_t)dtsize(dtype), d.items + l.first + nred, nrun, d.pool, tab, nullptr,
                                       false, allow_vec, d.shape, st);
            if (e != hipSuccess) return fail(BINE_ERR_HIP, std::string("copy-run launch: ") + hipGetErrorString(e));
        }
    }
    return BINE_OK;
}

// Launch a per-step work list whose temporaries the plan owns.  Outside graph capture the launches wait for
// the previous per-step run of the same plan (on whatever stream it ran) and record their own end, so calls on
// different streams never share the temporaries at once.  Under capture the caller orders the graph.
int launch_step_ordered(bine_vcomm_t comm, DevPlan* dp, const BufTable& tab, hipStream_t st) {
    const uint32_t av = vec_ok(tab);
    if (!dp->tmp_total() || capturing(st)) return launch_arrays(dp->step, dp->dtype, tab, av, st);
    std::lock_guard<std::mutex> lock(comm->mu);
    if (dp->tmp_recorded) HIPCHK(hipStreamWaitEvent(st, dp->tmp_done, 0));
    int rc = launch_arrays(dp->step, dp->dtype, tab, av, st);
    if (rc) return rc;
    HIPCHK(hipEventRecord(dp->tmp_done, st));
    dp->tmp_recorded = true;
    return BINE_OK;
}

// Run a collective call end to end.  sendbuf == recvbuf (allreduce only) is MPI_IN_PLACE: the input is read
// from recvbuf by a schedule compiled and race-checked for that aliasing.
int run_coll(bine_vcomm_t comm, const CollRequest& req0, int dtype, const void* send, void* recv, void* stream) {
    if (!comm) return fail(BINE_ERR_ARG, "NULL vcomm");
    CollRequest req = req0;
    std::shared_ptr<DevPlan> dpp;
    int rc = get_plan(comm, req, dtype, &dpp);
    if (rc) return rc;
    size_t sbytes = partition_bytes(dpp.get(), BUF_SEND), rbytes = partition_bytes(dpp.get(), BUF_RECV);
    if (overlap(send, sbytes, recv, rbytes)) {
        if (!(req.coll == BINE_COLL_ALLREDUCE && send == recv && sbytes == rbytes))
            return fail(BINE_ERR_ARG, "sendbuf and recvbuf overlap");
        req.in_place = true;
        if ((rc = get_plan(comm, req, dtype, &dpp))) return rc;
        sbytes = 0;
    }
    DevPlan* dp = dpp.get();
    if ((rc = check_range(send, sbytes, "sendbuf"))) return rc;
    if ((rc = check_range(recv, rbytes, "recvbuf"))) return rc;
    int exec = comm->cfg.exec;
    if (req.in_place && exec == BINE_EXEC_FUSED)
        return fail(BINE_ERR_UNSUPPORTED, "fused execution cannot run in place (sendbuf == recvbuf)");
    if (exec == BINE_EXEC_FUSED && !dp->cp.fused_ok)
        return fail(BINE_ERR_UNSUPPORTED, "schedule cannot be fused: " + dp->cp.fused_why);
    if (exec == BINE_EXEC_AUTO) exec = dp->cp.fused_ok ? BINE_EXEC_FUSED : BINE_EXEC_STEP;
    if (exec == BINE_EXEC_STEP && (rc = ensure_tmp(comm, dp))) return rc;
    HIPCHK(hipSetDevice(comm->cfg.device));
    BufTable tab = make_table(dp, req.in_place ? recv : send, recv);
    if (exec == BINE_EXEC_STEP) return launch_step_ordered(comm, dp, tab, (hipStream_t)stream);
    return launch_arrays(dp->fused, dtype, tab, vec_ok(tab), (hipStream_t)stream);
}

// A cached schedule, or nullptr; never compiles (the allocation-free entry point uses it).
std::shared_ptr<DevPlan> find_plan(bine_vcomm_t comm, const CollRequest& req, int dtype) {
    PlanKey key{req.coll, dtype, req.op, req.root, req.count, req.in_place};
    std::lock_guard<std::mutex> lock(comm->mu);
    auto it = comm->plans.find(key);
    if (it == comm->plans.end()) return nullptr;
    it->second->last_use = ++comm->clock;
    return it->second;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int parse_family(const char* s, int* out) {
    if (!strcmp(s, "DEFAULT")) *out = BINE_FAMILY_DEFAULT;
    else if (!strcmp(s, "BINE")) *out = BINE_FAMILY_BINE;
    else if (!strcmp(s, "RECDOUB")) *out = BINE_FAMILY_RECDOUB;
    else if (!strcmp(s, "BRUCK")) *out = BINE_FAMILY_BRUCK;
    else if (!strcmp(s, "RING")) *out = BINE_FAMILY_RING;
    else return BINE_ERR_ARG;
    return BINE_OK;
}

struct AlgoName {
    const char* name;
    int value;
};

}  // namespace

// ============================================================================
// vcomm
// ============================================================================

extern "C" int bine_vcomm_config_init(bine_vcomm_config_t* cfg, uint32_t size) {
    if (!cfg || size < 1) return fail(BINE_ERR_ARG, "bad config or size");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->dimensions[0] = size;
    cfg->dimensions[1] = cfg->dimensions[2] = 1;
    cfg->dimensions_num = 1;
    cfg->num_ports = 1;
    bine_coll_config_t* all[8] = {&cfg->allreduce, &cfg->allgather, &cfg->reduce_scatter, &cfg->bcast,
                                  &cfg->alltoall, &cfg->scatter, &cfg->gather, &cfg->reduce};
    for (auto* c : all) {
        c->family = BINE_FAMILY_BINE;
        c->algo = 0;
        c->distance = BINE_DISTANCE_INCREASING;
    }
    cfg->allreduce.algo = BINE_ALLREDUCE_B;
    cfg->allgather.algo = BINE_ALLGATHER_CONT_PERMUTE;
    cfg->allgather.distance = BINE_DISTANCE_DECREASING;
    cfg->bcast.distance = BINE_DISTANCE_DECREASING;
    cfg->gather.distance = BINE_DISTANCE_DECREASING;
    cfg->exec = BINE_EXEC_AUTO;
    cfg->device = 0;
    return BINE_OK;
}

extern "C" int bine_vcomm_config_from_env(bine_vcomm_config_t* cfg) {
    if (!cfg) return fail(BINE_ERR_ARG, "NULL config");
    const char* s;
    if ((s = getenv("LIBBINE_DIMENSIONS"))) {
        uint32_t d[3] = {1, 1, 1}, n = 0;
        std::string str(s);
        size_t pos = 0;
        while (pos <= str.size() && n < 4) {
            size_t x = str.find('x', pos);
            std::string tok = str.substr(pos, x == std::string::npos ? std::string::npos : x - pos);
            if (tok.empty() || n == 3) return fail(BINE_ERR_ARG, "bad LIBBINE_DIMENSIONS");
            d[n++] = (uint32_t)atoi(tok.c_str());
            if (x == std::string::npos) break;
            pos = x + 1;
        }
        for (uint32_t i = 0; i < 3; ++i) cfg->dimensions[i] = d[i];
        cfg->dimensions_num = n;
    }
    if ((s = getenv("LIBBINE_NUM_PORTS"))) cfg->num_ports = (uint32_t)atoi(s);
    if ((s = getenv("LIBBINE_EXEC"))) {
        if (!strcmp(s, "AUTO")) cfg->exec = BINE_EXEC_AUTO;
        else if (!strcmp(s, "STEP")) cfg->exec = BINE_EXEC_STEP;
        else if (!strcmp(s, "FUSED")) cfg->exec = BINE_EXEC_FUSED;
        else return fail(BINE_ERR_ARG, "bad LIBBINE_EXEC");
    }
    struct CollEnv {
        const char* name;
        bine_coll_config_t* c;
        std::vector<AlgoName> algos;
    } colls[] = {
        {"ALLREDUCE", &cfg->allreduce, {{"L", 0}, {"B", 1}, {"REDUCE_BCAST", 2}, {"B_CONT", 3}, {"B_COALESCE", 4}}},
        {"ALLGATHER", &cfg->allgather,
         {{"VEC_DOUBLING_CONT_PERMUTE", 0}, {"VEC_DOUBLING_CONT_SEND", 1}, {"VEC_DOUBLING_BLOCKS", 2}, {"GATHER_BCAST", 3}}},
        {"REDUCE_SCATTER", &cfg->reduce_scatter,
         {{"VEC_HALVING_CONT_PERMUTE", 0}, {"VEC_HALVING_CONT_SEND", 1}, {"VEC_HALVING_BLOCKS", 2}, {"REDUCE_SCATTER", 3}}},
        {"BCAST", &cfg->bcast, {{"BINOMIAL_TREE", 0}, {"BINOMIAL_TREE_TMPBUF", 1}, {"SCATTER_ALLGATHER", 2}}},
        {"ALLTOALL", &cfg->alltoall, {{"LOG", 0}}},
        {"SCATTER", &cfg->scatter, {{"BINOMIAL_TREE_CONT_PERMUTE", 0}, {"BINOMIAL_TREE_CONT_SEND", 1}}},
        {"GATHER", &cfg->gather, {{"BINOMIAL_TREE_CONT_PERMUTE", 0}, {"BINOMIAL_TREE_CONT_SEND", 1}}},
        {"REDUCE", &cfg->reduce, {{"BINOMIAL_TREE", 0}, {"REDUCE_SCATTER_GATHER", 1}}},
    };
    for (auto& ce : colls) {
        std::string base = std::string("LIBBINE_") + ce.name;
        if ((s = getenv((base + "_ALGO_FAMILY").c_str())) && parse_family(s, &ce.c->family))
            return fail(BINE_ERR_ARG, "bad " + base + "_ALGO_FAMILY");
        if ((s = getenv((base + "_ALGO").c_str()))) {
            bool found = false;
            for (auto& a : ce.algos)
                if (!strcmp(a.name, s)) {
                    ce.c->algo = a.value;
                    found = true;
                }
            if (!found) return fail(BINE_ERR_ARG, "bad " + base + "_ALGO");
        }
        if ((s = getenv((base + "_DISTANCE").c_str())))
            ce.c->distance = !strcmp(s, "INCREASING") ? BINE_DISTANCE_INCREASING : BINE_DISTANCE_DECREASING;
    }
    return BINE_OK;
}

extern "C" int bine_vcomm_create(const bine_vcomm_config_t* cfg, bine_vcomm_t* out) {
    if (!cfg || !out) return fail(BINE_ERR_ARG, "NULL argument");
    if (cfg->num_ports < 1 || cfg->num_ports > (uint32_t)kMaxPorts) return fail(BINE_ERR_ARG, "num_ports must be 1..6");
    Torus t;
    if (!Torus::make(cfg->dimensions, cfg->dimensions_num, &t)) return fail(BINE_ERR_ARG, "bad torus dimensions");
    if (t.size < 2) return fail(BINE_ERR_ARG, "a virtual communicator needs at least two ranks");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(BINE_ERR_NO_DEVICE, "no HIP device");
    }
    if (cfg->device < 0 || cfg->device >= ndev) return fail(BINE_ERR_ARG, "bad device ordinal");
    auto* c = new bine_vcomm_s();
    c->cfg = *cfg;
    c->t = t;
    // set-up of bine_allreduce_noalloc: the device address of its workspace, once per device
    c->scratch_err = scratch_init(cfg->device);
    *out = c;
    return BINE_OK;
}

extern "C" int bine_vcomm_destroy(bine_vcomm_t comm) {
    if (!comm) return BINE_OK;
    (void)hipSetDevice(comm->cfg.device);
    delete comm;
    return BINE_OK;
}

extern "C" int bine_vcomm_size(bine_vcomm_t comm, uint32_t* size) {
    if (!comm || !size) return fail(BINE_ERR_ARG, "NULL argument");
    *size = comm->t.size;
    return BINE_OK;
}

extern "C" int bine_vcomm_set_exec(bine_vcomm_t comm, int32_t exec) {
    if (!comm || exec < 0 || exec > 2) return fail(BINE_ERR_ARG, "bad exec");
    comm->cfg.exec = exec;
    return BINE_OK;
}

extern "C" int bine_vcomm_set_cache_limit(bine_vcomm_t comm, uint32_t max_plans) {
    if (!comm) return fail(BINE_ERR_ARG, "NULL vcomm");
    std::lock_guard<std::mutex> lock(comm->mu);
    (void)hipSetDevice(comm->cfg.device);
    comm->max_plans = max_plans;
    if (max_plans > 0) evict_lru(comm, max_plans);
    return BINE_OK;
}

extern "C" int bine_vcomm_trim(bine_vcomm_t comm) {
    if (!comm) return fail(BINE_ERR_ARG, "NULL vcomm");
    std::lock_guard<std::mutex> lock(comm->mu);
    (void)hipSetDevice(comm->cfg.device);
    comm->plans.clear();
    return BINE_OK;
}

extern "C" int bine_vcomm_cache_info(bine_vcomm_t comm, uint32_t* plans, uint64_t* tmp_bytes) {
    if (!comm) return fail(BINE_ERR_ARG, "NULL vcomm");
    std::lock_guard<std::mutex> lock(comm->mu);
    uint64_t t = 0;
    for (auto& kv : comm->plans)
        for (int i = 0; i < 2; ++i)
            if (kv.second->tmp[i]) t += kv.second->tmp_bytes[i];
    if (plans) *plans = (uint32_t)comm->plans.size();
    if (tmp_bytes) *tmp_bytes = t;
    return BINE_OK;
}

// ============================================================================
// collectives
// ============================================================================

extern "C" int bine_allreduce(const void* sendbuf, void* recvbuf, int count, bine_datatype_t datatype, bine_op_t op,
                              bine_vcomm_t comm, void* stream) {
    if (count < 0) return fail(BINE_ERR_ARG, "negative count");
    CollRequest r{BINE_COLL_ALLREDUCE, count, (int)op, 0};
    return run_coll(comm, r, datatype, sendbuf, recvbuf, stream);
}

extern "C" int bine_allreduce_noalloc(const void* sendbuf, void* recvbuf, int count, bine_datatype_t datatype,
                                      bine_op_t op, bine_vcomm_t comm, void* stream) {
    if (!comm) return fail(BINE_ERR_ARG, "NULL vcomm");
    if (count < 0) return fail(BINE_ERR_ARG, "negative count");
    const size_t ds = dtsize(datatype);
    if (!ds) return fail(BINE_ERR_ARG, "unknown datatype");
    const size_t bytes = (size_t)count * ds * comm->t.size;
    CollRequest req{BINE_COLL_ALLREDUCE, count, (int)op, 0};
    if (overlap(sendbuf, bytes, recvbuf, bytes)) {
        if (sendbuf != recvbuf) return fail(BINE_ERR_ARG, "sendbuf and recvbuf overlap");
        req.in_place = true;
    }
    std::shared_ptr<DevPlan> dpp = find_plan(comm, req, datatype);
    if (!dpp)
        return fail(BINE_ERR_UNSUPPORTED, "no cached schedule for this allreduce: compile it at set-up with a "
                                          "bine_allreduce call of the same count, type, op and in-place-ness");
    DevPlan* dp = dpp.get();
    int rc;
    if ((rc = check_range(req.in_place ? nullptr : sendbuf, req.in_place ? 0 : bytes, "sendbuf"))) return rc;
    if ((rc = check_range(recvbuf, bytes, "recvbuf"))) return rc;
    int exec = comm->cfg.exec;
    if (req.in_place && exec == BINE_EXEC_FUSED)
        return fail(BINE_ERR_UNSUPPORTED, "fused execution cannot run in place (sendbuf == recvbuf)");
    if (exec == BINE_EXEC_FUSED && !dp->cp.fused_ok)
        return fail(BINE_ERR_UNSUPPORTED, "schedule cannot be fused: " + dp->cp.fused_why);
    if (exec == BINE_EXEC_AUTO) exec = dp->cp.fused_ok ? BINE_EXEC_FUSED : BINE_EXEC_STEP;
    HIPCHK(hipSetDevice(comm->cfg.device));
    BufTable tab = make_table(dp, req.in_place ? recvbuf : sendbuf, recvbuf);
    if (exec == BINE_EXEC_STEP && dp->tmp_total()) {
        // the per-step temporaries come from the device's scratch array instead of the plan's allocations
        const size_t need = align256(dp->tmp_bytes[0]) + dp->tmp_bytes[1];
        if (need > scratch_bytes())
            return fail(BINE_ERR_ARG, "this schedule needs " + std::to_string(need) + " bytes of per-step temporaries; "
                                      "the allocation-free entry point holds at most " +
                                          std::to_string(scratch_bytes()) + " bytes");
        char* ptr = (char*)scratch_ptr(comm->cfg.device);
        if (!ptr)
            return fail(BINE_ERR_HIP, std::string("scratch workspace unavailable: hipGetSymbolAddress: ") +
                                          hipGetErrorString(comm->scratch_err));
        tab.base[BUF_TMP] = dp->tmp_bytes[0] ? (uint64_t)(uintptr_t)ptr : 0;
        tab.base[BUF_TMP2] = dp->tmp_bytes[1] ? (uint64_t)(uintptr_t)(ptr + align256(dp->tmp_bytes[0])) : 0;
    }
    return launch_arrays(exec == BINE_EXEC_FUSED ? dp->fused : dp->step, datatype, tab, vec_ok(tab),
                         (hipStream_t)stream);
}

extern "C" int bine_reduce_scatter(const void* sendbuf, void* recvbuf, const int recvcounts[], bine_datatype_t datatype,
                                   bine_op_t op, bine_vcomm_t comm, void* stream) {
    if (!comm || !recvcounts) return fail(BINE_ERR_ARG, "NULL argument");
    for (uint32_t i = 1; i < comm->t.size; ++i)
        if (recvcounts[i] != recvcounts[0])
            return fail(BINE_ERR_UNSUPPORTED, "unequal recvcounts (the reference assumes a reduce_scatter_block)");
    if (recvcounts[0] < 0) return fail(BINE_ERR_ARG, "negative count");
    CollRequest r{BINE_COLL_REDUCE_SCATTER, recvcounts[0], (int)op, 0};
    return run_coll(comm, r, datatype, sendbuf, recvbuf, stream);
}

extern "C" int bine_allgather(const void* sendbuf, int sendcount, bine_datatype_t sendtype, void* recvbuf,
                              int recvcount, bine_datatype_t recvtype, bine_vcomm_t comm, void* stream) {
    if (sendtype != recvtype || sendcount != recvcount)
        return fail(BINE_ERR_UNSUPPORTED, "send and receive type/count must match (lib/libbine.cc:1170-1171)");
    CollRequest r{BINE_COLL_ALLGATHER, sendcount, BINE_OP_SUM, 0};
    return run_coll(comm, r, sendtype, sendbuf, recvbuf, stream);
}

extern "C" int bine_bcast(void* buffer, int count, bine_datatype_t datatype, int root, bine_vcomm_t comm, void* stream) {
    if (count < 0) return fail(BINE_ERR_ARG, "negative count");
    CollRequest r{BINE_COLL_BCAST, count, BINE_OP_SUM, root};
    return run_coll(comm, r, datatype, nullptr, buffer, stream);
}

extern "C" int bine_alltoall(const void* sendbuf, int sendcount, bine_datatype_t sendtype, void* recvbuf, int recvcount,
                             bine_datatype_t recvtype, bine_vcomm_t comm, void* stream) {
    if (sendtype != recvtype || sendcount != recvcount)
        return fail(BINE_ERR_UNSUPPORTED, "send and receive type/count must match");
    CollRequest r{BINE_COLL_ALLTOALL, sendcount, BINE_OP_SUM, 0};
    return run_coll(comm, r, sendtype, sendbuf, recvbuf, stream);
}

extern "C" int bine_scatter(const void* sendbuf, int sendcount, bine_datatype_t sendtype, void* recvbuf, int recvcount,
                            bine_datatype_t recvtype, int root, bine_vcomm_t comm, void* stream) {
    if (sendtype != recvtype || sendcount != recvcount)
        return fail(BINE_ERR_UNSUPPORTED, "send and receive type/count must match (lib/libbine_coll_scatter.cc:195-196)");
    CollRequest r{BINE_COLL_SCATTER, recvcount, BINE_OP_SUM, root};
    return run_coll(comm, r, sendtype, sendbuf, recvbuf, stream);
}

extern "C" int bine_gather(const void* sendbuf, int sendcount, bine_datatype_t sendtype, void* recvbuf, int recvcount,
                           bine_datatype_t recvtype, int root, bine_vcomm_t comm, void* stream) {
    if (sendtype != recvtype || sendcount != recvcount)
        return fail(BINE_ERR_UNSUPPORTED, "send and receive type/count must match (lib/libbine_coll_gather.cc:209)");
    CollRequest r{BINE_COLL_GATHER, sendcount, BINE_OP_SUM, root};
    return run_coll(comm, r, sendtype, sendbuf, recvbuf, stream);
}

extern "C" int bine_reduce(const void* sendbuf, void* recvbuf, int count, bine_datatype_t datatype, bine_op_t op, int root,
                           bine_vcomm_t comm, void* stream) {
    if (count < 0) return fail(BINE_ERR_ARG, "negative count");
    CollRequest r{BINE_COLL_REDUCE, count, (int)op, root};
    return run_coll(comm, r, datatype, sendbuf, recvbuf, stream);
}

extern "C" int bine_plan_stats(bine_vcomm_t comm, int32_t coll, int count, bine_datatype_t datatype, bine_op_t op,
                               int root, int32_t exec, bine_plan_stats_t* out) {
    if (!comm || !out) return fail(BINE_ERR_ARG, "NULL argument");
    CollRequest r{coll, count, (int)op, root};
    std::shared_ptr<DevPlan> dpp;
    int rc = get_plan(comm, r, datatype, &dpp);
    if (rc) return rc;
    const DevPlan* dp = dpp.get();
    if (exec == BINE_EXEC_AUTO) exec = dp->cp.fused_ok ? BINE_EXEC_FUSED : BINE_EXEC_STEP;
    if (exec == BINE_EXEC_FUSED && !dp->cp.fused_ok) return fail(BINE_ERR_UNSUPPORTED, "schedule cannot be fused: " + dp->cp.fused_why);
    const DevArrays& d = exec == BINE_EXEC_FUSED ? dp->fused : dp->step;
    out->launches = kernel_launches(d.low);
    out->items = d.low.items.size();
    out->bytes_read = d.low.bytes_read;
    out->bytes_written = d.low.bytes_written;
    out->max_tree_leaves = d.low.max_leaves;
    return BINE_OK;
}

namespace {

// Launch shapes the autotuners time: vectors per thread, workgroup size,
// persistent grid, streaming loads/stores.
std::vector<LaunchShape> tuning_candidates(int cus) {
    return {{1, 64, 0, 0},  {1, 128, 0, 0}, {1, 256, 0, 0}, {1, 512, 0, 0},       {2, 64, 0, 0},
            {2, 128, 0, 0}, {2, 256, 0, 0}, {2, 512, 0, 0}, {4, 64, 0, 0},        {4, 128, 0, 0},
            {1, 256, 0, 1}, {2, 256, 0, 1}, {1, 512, 0, 1}, {2, 128, 0, 1},       {2, 64, 0, 1},
            {4, 64, 0, 1},  {4, 128, 0, 1}, {1, 64, 0, 1},  {4, 64, 8 * cus, 0},  {4, 64, 16 * cus, 1},
            {1, 256, 4 * cus, 0}, {2, 128, 8 * cus, 0}, {4, 256, 0, 1}, {4, 256, 0, 0}, {2, 256, 4 * cus, 1},
            {4, 128, 8 * cus, 1}, {4, 512, 0, 1}, {4, 512, 4 * cus, 1}};
}

// The two-phase trials of an autotune: the three fastest one-launch shapes, each with its copies in group order
// and tile by tile, and with reduction items of one and of two tiles.  Which order wins depends on the list
// (measured: group order for 512-leaf trees of two tiles, tile by tile for 64-leaf trees of many tiles), and so
// does the item length (two tiles: 1-2.5 % faster on 64 ranks of 64 tiles per block, 10 % slower on 512 ranks of
// two), so none of them wins everywhere.
std::vector<LaunchShape> two_phase_trials(std::vector<std::pair<float, LaunchShape>> timed) {
    std::stable_sort(timed.begin(), timed.end(),
                     [](const std::pair<float, LaunchShape>& a, const std::pair<float, LaunchShape>& b) { return a.first < b.first; });
    std::vector<LaunchShape> out;
    for (size_t i = 0; i < timed.size() && i < 3; ++i)
        for (int m = 1; m <= 2; ++m)
            for (int d = 0; d <= 1; ++d) {
                LaunchShape c = timed[i].second;
                c.two_phase = m;
                c.double_tiles = d;
                out.push_back(c);
            }
    return out;
}

// Issue one call of a collective (argument layout of its entry point).
int issue(bine_vcomm_t comm, int32_t coll, const void* send, void* recv, int count, int dtype, int op, int root,
          void* stream) {
    switch (coll) {
    case BINE_COLL_ALLREDUCE: return bine_allreduce(send, recv, count, (bine_datatype_t)dtype, (bine_op_t)op, comm, stream);
    case BINE_COLL_REDUCE_SCATTER: {
        std::vector<int> rc(comm->t.size, count);
        return bine_reduce_scatter(send, recv, rc.data(), (bine_datatype_t)dtype, (bine_op_t)op, comm, stream);
    }
    case BINE_COLL_ALLGATHER:
        return bine_allgather(send, count, (bine_datatype_t)dtype, recv, count, (bine_datatype_t)dtype, comm, stream);
    case BINE_COLL_BCAST: return bine_bcast(recv, count, (bine_datatype_t)dtype, root, comm, stream);
    case BINE_COLL_ALLTOALL:
        return bine_alltoall(send, count, (bine_datatype_t)dtype, recv, count, (bine_datatype_t)dtype, comm, stream);
    case BINE_COLL_SCATTER:
        return bine_scatter(send, count, (bine_datatype_t)dtype, recv, count, (bine_datatype_t)dtype, root, comm, stream);
    case BINE_COLL_GATHER:
        return bine_gather(send, count, (bine_datatype_t)dtype, recv, count, (bine_datatype_t)dtype, root, comm, stream);
    case BINE_COLL_REDUCE: return bine_reduce(send, recv, count, (bine_datatype_t)dtype, (bine_op_t)op, root, comm, stream);
    default: return fail(BINE_ERR_ARG, "unknown collective");
    }
}

}  // namespace

extern "C" int bine_plan_autotune(bine_vcomm_t comm, int32_t coll, const void* sendbuf, void* recvbuf, int count,
                                  bine_datatype_t datatype, bine_op_t op, int root, int32_t exec, void* stream,
                                  int iters) {
    if (!comm || iters < 1) return fail(BINE_ERR_ARG, "bad arguments");
    CollRequest r{coll, count, (int)op, root};
    r.in_place = coll == BINE_COLL_ALLREDUCE && sendbuf && sendbuf == recvbuf;
    std::shared_ptr<DevPlan> dpp;
    int rc = get_plan(comm, r, datatype, &dpp);
    if (rc) return rc;
    DevPlan* dp = dpp.get();
    if (exec == BINE_EXEC_AUTO) exec = dp->cp.fused_ok ? BINE_EXEC_FUSED : BINE_EXEC_STEP;
    if (exec == BINE_EXEC_FUSED && !dp->cp.fused_ok) return fail(BINE_ERR_UNSUPPORTED, "schedule cannot be fused");
    DevArrays& d = exec == BINE_EXEC_FUSED ? dp->fused : dp->step;
    const std::vector<std::vector<Group>> one(1, dp->cp.fused);
    const std::vector<std::vector<Group>>& groups = exec == BINE_EXEC_FUSED ? one : dp->cp.step_groups;
    const int32_t saved_exec = comm->cfg.exec;
    comm->cfg.exec = exec;
    HIPCHK(hipSetDevice(comm->cfg.device));
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, comm->cfg.device) == hipSuccess) cus = prop.multiProcessorCount;
    const std::vector<LaunchShape> cands = tuning_candidates(cus);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipStream_t st = (hipStream_t)stream;
    float best = 1e30f;
    LaunchShape bestc = d.shape;
    rc = BINE_OK;
    bool first = true;
    std::vector<std::pair<float, LaunchShape>> timed;
    auto trial = [&](const LaunchShape& c) {
        // the arrays in place before the call may be named by captured graphs: they are kept, not freed;
        // the candidates' own arrays never leave this loop
        if ((rc = upload(groups, dp->ds, c, 0, &d, first ? &dp->retired : nullptr))) return;
        first = false;
        if ((rc = issue(comm, coll, sendbuf, recvbuf, count, datatype, op, root, stream))) return;
        if (hipEventRecord(e0, st) != hipSuccess) {
            rc = fail(BINE_ERR_HIP, "event");
            return;
        }
        for (int i = 0; i < iters && !rc; ++i) rc = issue(comm, coll, sendbuf, recvbuf, count, datatype, op, root, stream);
        if (rc) return;
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        timed.emplace_back(ms, c);
        if (ms < best) {
            best = ms;
            bestc = c;
        }
    };
    for (const LaunchShape& c : cands) {
        trial(c);
        if (rc) break;
    }
    if (!rc && has_multi_output_reduction(groups))
        for (const LaunchShape& c : two_phase_trials(timed)) {
            trial(c);
            if (rc) break;
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    comm->cfg.exec = saved_exec;
    if (rc) return rc;
    return upload(groups, dp->ds, bestc, 0, &d);
}

extern "C" int bine_plan_dump(const bine_vcomm_config_t* cfg, int32_t coll, int count, bine_datatype_t datatype,
                              bine_op_t op, int root, int32_t in_place, int32_t exec, int64_t tile_elems,
                              int32_t two_phase, bine_plan_dump_info_t* info, void* items, void* pool,
                              uint32_t* launches) {
    if (!cfg || !info) return fail(BINE_ERR_ARG, "NULL argument");
    if (two_phase < 0 || two_phase > 7 || (two_phase & 3) == 3)
        return fail(BINE_ERR_ARG, "two_phase must be 0, 1 or 2, plus 4 for reduction items of two tiles");
    const size_t ds = dtsize(datatype);
    if (!ds || count < 0) return fail(BINE_ERR_ARG, "bad datatype or count");
    if (cfg->num_ports < 1 || cfg->num_ports > (uint32_t)kMaxPorts) return fail(BINE_ERR_ARG, "num_ports must be 1..6");
    Torus t;
    if (!Torus::make(cfg->dimensions, cfg->dimensions_num, &t) || t.size < 2) return fail(BINE_ERR_ARG, "bad torus");
    CollRequest r{coll, count, (int)op, root};
    r.in_place = in_place != 0;
    Program prog;
    CompiledPlan cp;
    std::string err;
    if (!build_program(t, *cfg, r, &prog, &err)) return fail(BINE_ERR_UNSUPPORTED, err);
    if (!compile_program(prog, true, &cp, &err)) return fail(BINE_ERR_UNSUPPORTED, err);
    if (exec == BINE_EXEC_AUTO) exec = cp.fused_ok ? BINE_EXEC_FUSED : BINE_EXEC_STEP;
    if (exec == BINE_EXEC_FUSED && !cp.fused_ok) return fail(BINE_ERR_UNSUPPORTED, "schedule cannot be fused: " + cp.fused_why);
    LaunchShape sh = default_shape();
    int64_t deep = deep_tile_elems(sh, ds);
    if (tile_elems <= 0) tile_elems = default_tile_elems(ds);
    else deep = tile_elems;
    Lowered low;
    if (exec == BINE_EXEC_FUSED)
        lower_groups(std::vector<std::vector<Group>>(1, cp.fused), tile_elems, (int)(16 / ds), 0, ds, &low, deep, split_lg(sh),
                     nullptr, nullptr, two_phase & 3, (two_phase & 4) ? 2 : 1);
    else
        lower_groups(cp.step_groups, tile_elems, (int)(16 / ds), 0, ds, &low, deep, split_lg(sh), nullptr, nullptr,
                     two_phase & 3, (two_phase & 4) ? 2 : 1);
    info->n_items = low.items.size();
    info->n_pool = low.pool.size();
    info->n_launches = low.launches.size();
    for (int b = 0; b < 4; ++b) {
        info->extent[b] = cp.extent[b];
        info->single[b] = cp.single[b] ? 1 : 0;
    }
    info->P = cp.P;
    info->fused_ok = cp.fused_ok ? 1 : 0;
    if (items) std::memcpy(items, low.items.data(), low.items.size() * sizeof(GredItem));
    if (pool) std::memcpy(pool, low.pool.data(), low.pool.size() * sizeof(PoolEntry));
    if (launches)
        for (size_t i = 0; i < low.launches.size(); ++i) {
            launches[2 * i] = low.launches[i].first;
            launches[2 * i + 1] = low.launches[i].second;
        }
    return BINE_OK;
}

// ============================================================================
// schedule helpers
// ============================================================================

extern "C" int bine_compute_peers(const uint32_t* dims, uint32_t ndims, uint32_t rank, int port, int family,
                                  uint32_t* peers) {
    Torus t;
    if (!dims || !peers || !Torus::make(dims, ndims, &t) || rank >= t.size || port < 0) return fail(BINE_ERR_ARG, "bad arguments");
    if (family != FAM_BINE && family != FAM_RECDOUB) return fail(BINE_ERR_ARG, "family must be BINE or RECDOUB");
    std::vector<uint32_t> v = peer_sequence(t, rank, (uint32_t)port, family);
    std::copy(v.begin(), v.end(), peers);
    return BINE_OK;
}

extern "C" int bine_compute_step_info(const uint32_t* dims, uint32_t ndims, uint32_t port, uint32_t* d, uint32_t* step_in_d) {
    Torus t;
    if (!dims || !d || !step_in_d || !Torus::make(dims, ndims, &t)) return fail(BINE_ERR_ARG, "bad arguments");
    std::vector<StepDim> v = step_dims(t, port);
    for (size_t i = 0; i < v.size(); ++i) {
        d[i] = v[i].d;
        step_in_d[i] = v[i].k;
    }
    return BINE_OK;
}

extern "C" int bine_get_peer(const uint32_t* dims, uint32_t ndims, const int32_t* coord, uint32_t step, uint32_t port,
                             int family, int32_t* coord_peer) {
    Torus t;
    if (!dims || !coord || !coord_peer || !Torus::make(dims, ndims, &t) || step >= t.steps)
        return fail(BINE_ERR_ARG, "bad arguments");
    if (family != FAM_BINE && family != FAM_RECDOUB) return fail(BINE_ERR_ARG, "family must be BINE or RECDOUB");
    int c[kMaxDims] = {0, 0, 0};
    for (uint32_t i = 0; i < t.nd; ++i) {
        if (coord[i] < 0 || coord[i] >= (int32_t)t.dims[i]) return fail(BINE_ERR_ARG, "coordinate out of range");
        c[i] = coord[i];
    }
    const uint32_t peer = peer_of(t, c, step, port, step_dims(t, port), family);
    t.coords(peer, c);
    for (uint32_t i = 0; i < t.nd; ++i) coord_peer[i] = c[i];
    return BINE_OK;
}

extern "C" int bine_get_mirroring_port(int num_ports, uint32_t ndims) {
    if (ndims < 1 || ndims > kMaxDims) return -1;
    for (int p = 0; p < num_ports; ++p)
        if (mirrored_port(p, ndims)) return p;
    return -1;
}

extern "C" int bine_block_step(const uint32_t* dims, uint32_t ndims, uint32_t rank, uint32_t port, int family,
                               uint32_t* block_step) {
    Torus t;
    if (!dims || !block_step || !Torus::make(dims, ndims, &t) || rank >= t.size) return fail(BINE_ERR_ARG, "bad arguments");
    if (family != FAM_BINE && family != FAM_RECDOUB) return fail(BINE_ERR_ARG, "family must be BINE or RECDOUB");
    BlockSchedule bs;
    bs.build(t, port, family);
    for (uint32_t b = 0; b < t.size; ++b) block_step[b] = bs.send_step[(size_t)rank * t.size + b];
    return BINE_OK;
}

extern "C" int bine_get_tree(const uint32_t* dims, uint32_t ndims, uint32_t root, uint32_t port, int family, int distance,
                             uint32_t* parent, uint32_t* reached_at_step, uint32_t* remapped_ranks,
                             uint32_t* remapped_ranks_max, uint32_t* subtree_roots) {
    Torus t;
    if (!dims || !Torus::make(dims, ndims, &t) || root >= t.size) return fail(BINE_ERR_ARG, "bad arguments");
    Tree tr;
    if (!build_tree(t, root, port, family, distance, &tr)) return fail(BINE_ERR_UNSUPPORTED, "tree numbering failed");
    std::copy(tr.parent.begin(), tr.parent.end(), parent);
    std::copy(tr.reached_at_step.begin(), tr.reached_at_step.end(), reached_at_step);
    std::copy(tr.remapped.begin(), tr.remapped.end(), remapped_ranks);
    std::copy(tr.remapped_max.begin(), tr.remapped_max.end(), remapped_ranks_max);
    std::copy(tr.subtree_root.begin(), tr.subtree_root.end(), subtree_roots);
    return BINE_OK;
}

extern "C" uint32_t bine_remap_rank(uint32_t rank, uint32_t size) { return remap_rank(rank, size); }

namespace {

// Shared checks of the two reduce_local forms: out = a + b.  out may be a or b exactly; any other overlap of
// out with an input would race in the grid-stride kernel.  VOID leaves out as it is.
int reduce_local_common(const void* a, const void* b, void* out, int count, bine_datatype_t datatype, bine_op_t op,
                        void* stream) {
    const size_t ds = dtsize(datatype);
    if (!ds) return fail(BINE_ERR_ARG, "unknown datatype");
    if (count < 0) return fail(BINE_ERR_ARG, "negative count");
    if (op != BINE_OP_SUM && op != BINE_OP_VOID) return fail(BINE_ERR_ARG, "reduce_local supports SUM and VOID only");
    if (op == BINE_OP_VOID || count == 0) return BINE_OK;
    const size_t bytes = (size_t)count * ds;
    int rc;
    if ((rc = check_range(a, bytes, "inbuf"))) return rc;
    if ((rc = check_range(b, bytes, "inbuf"))) return rc;
    if ((rc = check_range(out, bytes, "outbuf"))) return rc;
    if ((out != a && overlap(out, bytes, a, bytes)) || (out != b && overlap(out, bytes, b, bytes)))
        return fail(BINE_ERR_ARG, "output partially overlaps an input (only the identical buffer is allowed)");
    hipError_t e = launch_reduce_local(datatype, a, b, out, (uint64_t)count, (hipStream_t)stream);
    if (e != hipSuccess) return fail(BINE_ERR_HIP, std::string("reduce_local launch: ") + hipGetErrorString(e));
    return BINE_OK;
}

}  // namespace

extern "C" int bine_reduce_local(const void* inbuf, void* inoutbuf, int count, bine_datatype_t datatype, bine_op_t op,
                                 void* stream) {
    return reduce_local_common(inbuf, inoutbuf, inoutbuf, count, datatype, op, stream);
}

extern "C" int bine_reduce_local3(const void* inbuf_a, const void* inbuf_b, void* outbuf, int count,
                                  bine_datatype_t datatype, bine_op_t op, void* stream) {
    return reduce_local_common(inbuf_a, inbuf_b, outbuf, count, datatype, op, stream);
}

extern "C" const char* bine_last_error(void) { return g_err.c_str(); }

// ============================================================================
// batches
// ============================================================================

struct bine_batch_s {
    bine_vcomm_t comm = nullptr;
    std::vector<bine_batch_item_t> items;
    std::vector<std::shared_ptr<DevPlan>> plans;  // held for the batch's life (the cache may evict them)
    std::vector<int> use_fused;  // per item
    // one fused work list per element type
    struct TypeList {
        int dtype;
        std::vector<std::vector<Group>> groups;  // one launch
        std::vector<uint32_t> group_bucket;      // bucket of each group
        DevArrays arr;
        LaunchShape shape = default_shape();     // each element type is tuned on its own
        uint32_t allow_vec = 1;                  // bit 1 set when some output can start off a 16-byte boundary
    };
    std::vector<TypeList> lists;
    BufTable* d_tabs = nullptr;
    std::vector<BufTable> tabs;
    LaunchShape shape = default_shape();
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipStream_t graph_stream = nullptr;
    ~bine_batch_s() {
        for (auto& l : lists) l.arr.release();
        if (d_tabs) (void)hipFree(d_tabs);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

// Lower and upload one element type's work list.  On failure the list keeps its previous arrays.
int lower_list(bine_batch_t b, bine_batch_s::TypeList& l) {
    {
        const size_t ds = dtsize(l.dtype);
        // one lowering for the whole list: every work item carries its group's bucket index, runs of copies
        // form within a bucket, and items are ordered tile-major across all buckets
        Lowered all;
        std::vector<uint8_t> bucket_vec(b->tabs.size());
        for (size_t i = 0; i < b->tabs.size(); ++i) bucket_vec[i] = (uint8_t)vec_ok(b->tabs[i]);
        lower_groups(l.groups, shallow_tile_elems(l.shape, ds), (int)(16 / ds), 0, ds, &all, deep_tile_elems(l.shape, ds),
                     split_lg(l.shape), &l.group_bucket, &bucket_vec, l.shape.two_phase, l.shape.double_tiles ? 2 : 1);
        int max_lg = max_lg_of(l.groups);
        bool tables_off16 = false;
        for (uint32_t gb : l.group_bucket) tables_off16 = tables_off16 || off16(b->tabs[gb]);
        const uint32_t allow = realign_pays(all, tables_off16) ? 3u : 1u;
        int rc = install(std::move(all), max_lg, shallow_tile_elems(l.shape, ds), l.shape, &l.arr, nullptr);
        if (rc) return rc;
        l.allow_vec = allow;
    }
    return BINE_OK;
}

int batch_lower(bine_batch_t b) {
    for (auto& l : b->lists) {
        int rc = lower_list(b, l);
        if (rc) return rc;
    }
    return BINE_OK;
}

// Launch one element type's work list: per launch of the list, the reduction kernel, then the copy runs.
int run_list(bine_batch_t b, const bine_batch_s::TypeList& l, hipStream_t st) {
    const Lowered& low = l.arr.low;
    for (size_t li = 0; li < low.launches.size(); ++li) {
        const uint32_t first = low.launches[li].first, nrun = low.runs[li], nred = low.launches[li].second - nrun;
        if (nred) {
            hipError_t e = launch_gred(l.dtype, l.arr.max_lg, l.arr.items + first, nred, l.arr.pool, BufTable{},
                                       b->d_tabs, true, l.allow_vec, l.shape, st);
            if (e != hipSuccess) return fail(BINE_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(e));
        }
        if (nrun) {
            hipError_t e = launch_grun((uint32_t)dtsize(l.dtype), l.arr.items + first + nred, nrun, l.arr.pool,
                                       BufTable{}, b->d_tabs, true, l.allow_vec, l.shape, st);
            if (e != hipSuccess) return fail(BINE_ERR_HIP, std::string("batch copy-run launch: ") + hipGetErrorString(e));
        }
    }
    return BINE_OK;
}

int batch_run(bine_batch_t b, hipStream_t st) {
    for (auto& l : b->lists) {
        int rc = run_list(b, l, st);
        if (rc) return rc;
    }
    // collectives whose schedule cannot be fused run their step launches after
    for (size_t i = 0; i < b->items.size(); ++i) {
        if (b->use_fused[i]) continue;
        int rc = launch_step_ordered(b->comm, b->plans[i].get(), b->tabs[i], st);
        if (rc) return rc;
    }
    return BINE_OK;
}

}  // namespace

extern "C" int bine_batch_create(bine_vcomm_t comm, const bine_batch_item_t* items, int n, bine_batch_t* out) {
    if (!comm || !items || n <= 0 || !out) return fail(BINE_ERR_ARG, "bad arguments");
    auto b = std::make_unique<bine_batch_s>();
    b->comm = comm;
    b->items.assign(items, items + n);
    for (int i = 0; i < n; ++i) {
        const bine_batch_item_t& it = items[i];
        if (it.count < 0) return fail(BINE_ERR_ARG, "negative count");
        CollRequest r{it.coll, it.count, it.op, it.root};
        // an allreduce with sendbuf == recvbuf runs its in-place schedule per step
        r.in_place = it.coll == BINE_COLL_ALLREDUCE && it.sendbuf && it.sendbuf == it.recvbuf;
        std::shared_ptr<DevPlan> dpp;
        int rc = get_plan(comm, r, it.datatype, &dpp);
        if (rc) return rc;
        DevPlan* dp = dpp.get();
        const void* send = it.coll == BINE_COLL_BCAST ? nullptr : it.sendbuf;
        if (r.in_place) send = nullptr;
        if ((rc = check_range(send, partition_bytes(dp, BUF_SEND), "batch sendbuf"))) return rc;
        if ((rc = check_range(it.recvbuf, partition_bytes(dp, BUF_RECV), "batch recvbuf"))) return rc;
        if (overlap(send, partition_bytes(dp, BUF_SEND), it.recvbuf, partition_bytes(dp, BUF_RECV)))
            return fail(BINE_ERR_ARG, "batch item buffers overlap");
        if (r.in_place && comm->cfg.exec == BINE_EXEC_FUSED)
            return fail(BINE_ERR_UNSUPPORTED, "fused execution cannot run in place (sendbuf == recvbuf)");
        const bool fuse = dp->cp.fused_ok && comm->cfg.exec != BINE_EXEC_STEP;
        if (!fuse && (rc = ensure_tmp(comm, dp))) return rc;
        b->plans.push_back(dpp);
        b->tabs.push_back(make_table(dp, r.in_place ? it.recvbuf : send, it.recvbuf));
        b->use_fused.push_back(fuse ? 1 : 0);
        if (!fuse) continue;
        auto lit = std::find_if(b->lists.begin(), b->lists.end(), [&](const bine_batch_s::TypeList& l) { return l.dtype == it.datatype; });
        if (lit == b->lists.end()) {
            b->lists.push_back(bine_batch_s::TypeList{it.datatype, std::vector<std::vector<Group>>(1), {}, DevArrays()});
            lit = b->lists.end() - 1;
        }
        for (const Group& g : dp->cp.fused) {
            lit->groups[0].push_back(g);
            lit->group_bucket.push_back((uint32_t)i);
        }
    }
    // buffers of different batch items must not overlap each other's outputs
    for (size_t i = 0; i < b->items.size(); ++i)
        for (size_t j = 0; j < b->items.size(); ++j) {
            if (i == j) continue;
            const DevPlan* pi = b->plans[i].get();
            const DevPlan* pj = b->plans[j].get();
            const void* si = b->items[i].coll == BINE_COLL_BCAST ? nullptr : b->items[i].sendbuf;
            if (overlap(b->items[j].recvbuf, partition_bytes(pj, BUF_RECV), b->items[i].recvbuf, partition_bytes(pi, BUF_RECV)) ||
                overlap(b->items[j].recvbuf, partition_bytes(pj, BUF_RECV), si, partition_bytes(pi, BUF_SEND)))
                return fail(BINE_ERR_ARG, "a batch item's recvbuf overlaps another item's buffers");
        }
    HIPCHK(hipSetDevice(comm->cfg.device));
    HIPCHK(hipMalloc(&b->d_tabs, sizeof(BufTable) * b->tabs.size()));
    HIPCHK(hipMemcpy(b->d_tabs, b->tabs.data(), sizeof(BufTable) * b->tabs.size(), hipMemcpyHostToDevice));
    for (auto& l : b->lists) l.shape = initial_shape(l.groups, dtsize(l.dtype));
    if (!b->lists.empty()) b->shape = b->lists[0].shape;
    int rc = batch_lower(b.get());
    if (rc) return rc;
    *out = b.release();
    return BINE_OK;
}

extern "C" int bine_batch_launch(bine_batch_t b, void* stream) {
    if (!b) return fail(BINE_ERR_ARG, "NULL batch");
    HIPCHK(hipSetDevice(b->comm->cfg.device));
    return batch_run(b, (hipStream_t)stream);
}

extern "C" int bine_batch_launch_graph(bine_batch_t b, void* stream) {
    if (!b) return fail(BINE_ERR_ARG, "NULL batch");
    HIPCHK(hipSetDevice(b->comm->cfg.device));
    hipStream_t st = (hipStream_t)stream;
    if (!b->gexec) {
        hipStream_t cap;
        HIPCHK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        int rc = batch_run(b, cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (rc) return rc;
        if (e != hipSuccess) return fail(BINE_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->gexec, st));
    return BINE_OK;
}

extern "C" int bine_batch_stats(bine_batch_t b, bine_plan_stats_t* out) {
    if (!b || !out) return fail(BINE_ERR_ARG, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    for (auto& l : b->lists) {
        out->launches += kernel_launches(l.arr.low);
        out->items += l.arr.low.items.size();
        out->bytes_read += l.arr.low.bytes_read;
        out->bytes_written += l.arr.low.bytes_written;
        out->max_tree_leaves = std::max(out->max_tree_leaves, l.arr.low.max_leaves);
    }
    for (size_t i = 0; i < b->items.size(); ++i) {
        if (b->use_fused[i]) continue;
        const DevArrays& d = b->plans[i]->step;
        out->launches += kernel_launches(d.low);
        out->items += d.low.items.size();
        out->bytes_read += d.low.bytes_read;
        out->bytes_written += d.low.bytes_written;
    }
    return BINE_OK;
}

extern "C" int bine_batch_get_tuning(bine_batch_t b, int32_t* vpt, int32_t* threads, int32_t* persistent_wgs,
                                     int32_t* flags) {
    if (!b) return fail(BINE_ERR_ARG, "NULL batch");
    if (vpt) *vpt = b->shape.vpt;
    if (threads) *threads = b->shape.threads;
    if (persistent_wgs) *persistent_wgs = b->shape.persistent;
    if (flags) *flags = shape_flags(b->shape);
    return BINE_OK;
}

namespace {
void drop_graph(bine_batch_t b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}
}  // namespace

extern "C" int bine_batch_set_tuning(bine_batch_t b, int32_t vpt, int32_t threads, int32_t persistent_wgs,
                                     int32_t flags) {
    if (!b) return fail(BINE_ERR_ARG, "NULL batch");
    if ((vpt != 1 && vpt != 2 && vpt != 4) || threads < 64 || threads > 512 || (threads & (threads - 1)) ||
        persistent_wgs < 0 || flags < 0 || flags > 15 || ((flags >> 1) & 3) == 3)
        return fail(BINE_ERR_ARG, "bad launch shape");
    const LaunchShape sh = shape_from_words(vpt, threads, persistent_wgs, flags);
    b->shape = sh;
    drop_graph(b);
    for (auto& l : b->lists) {
        const bool relower = sh.vpt != l.shape.vpt || sh.threads != l.shape.threads || sh.two_phase != l.shape.two_phase ||
                             sh.double_tiles != l.shape.double_tiles;
        l.shape = sh;
        if (relower) {
            int rc = lower_list(b, l);
            if (rc) return rc;
        }
    }
    return BINE_OK;
}

namespace {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

int device_cus(int device) {
    hipDeviceProp_t prop;
    return hipGetDeviceProperties(&prop, device) == hipSuccess ? prop.multiProcessorCount : 256;
}

// ms[i]: average time of one batch launch at shapes[i] (one warm-up launch, then iters timed).
int time_shapes(bine_batch_t b, hipStream_t st, int iters, const std::vector<LaunchShape>& shapes, float* ms) {
    EventPair ev;
    HIPCHK(hipEventCreate(&ev.a));
    HIPCHK(hipEventCreate(&ev.b));
    for (size_t i = 0; i < shapes.size(); ++i) {
        const LaunchShape& c = shapes[i];
        int rc = bine_batch_set_tuning(b, c.vpt, c.threads, c.persistent, shape_flags(c));
        if (rc) return rc;
        if ((rc = batch_run(b, st))) return rc;  // warm-up
        HIPCHK(hipEventRecord(ev.a, st));
        for (int k = 0; k < iters; ++k)
            if ((rc = batch_run(b, st))) return rc;
        HIPCHK(hipEventRecord(ev.b, st));
        HIPCHK(hipEventSynchronize(ev.b));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, ev.a, ev.b));
        ms[i] = t / (float)iters;
    }
    return BINE_OK;
}

}  // namespace

extern "C" int bine_tuning_candidates(int device, int32_t* shapes, int cap, int* n) {
    if (!n || cap < 0 || (cap > 0 && !shapes)) return fail(BINE_ERR_ARG, "bad arguments");
    const std::vector<LaunchShape> c = tuning_candidates(device_cus(device));
    *n = (int)c.size();
    for (int i = 0; i < (int)c.size() && i < cap; ++i) {
        shapes[4 * i] = c[i].vpt;
        shapes[4 * i + 1] = c[i].threads;
        shapes[4 * i + 2] = c[i].persistent;
        shapes[4 * i + 3] = shape_flags(c[i]);
    }
    return BINE_OK;
}

extern "C" int bine_batch_time_shapes(bine_batch_t b, void* stream, int iters, const int32_t* shapes, int n,
                                      float* ms) {
    if (!b || iters < 1 || n < 1 || !shapes || !ms) return fail(BINE_ERR_ARG, "bad arguments");
    std::vector<LaunchShape> v;
    for (int i = 0; i < n; ++i) v.push_back(shape_from_words(shapes[4 * i], shapes[4 * i + 1], shapes[4 * i + 2], shapes[4 * i + 3]));
    HIPCHK(hipSetDevice(b->comm->cfg.device));
    const LaunchShape keep = b->shape;
    int rc = time_shapes(b, (hipStream_t)stream, iters, v, ms);
    int rc2 = bine_batch_set_tuning(b, keep.vpt, keep.threads, keep.persistent, shape_flags(keep));
    return rc ? rc : rc2;
}

extern "C" int bine_batch_autotune(bine_batch_t b, void* stream, int iters) {
    if (!b || iters < 1) return fail(BINE_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(b->comm->cfg.device));
    hipStream_t st = (hipStream_t)stream;
    const std::vector<LaunchShape> cands = tuning_candidates(device_cus(b->comm->cfg.device));
    // every element type's list is timed alone at every candidate and keeps its fastest shape
    const size_t nl = b->lists.size();
    std::vector<float> best(nl, 1e30f);
    std::vector<LaunchShape> pick(nl, default_shape());
    EventPair ev;
    HIPCHK(hipEventCreate(&ev.a));
    HIPCHK(hipEventCreate(&ev.b));
    drop_graph(b);
    std::vector<std::vector<std::pair<float, LaunchShape>>> timed(nl);
    auto trial = [&](size_t i, const LaunchShape& c) -> int {
        auto& l = b->lists[i];
        const bool relower = c.vpt != l.shape.vpt || c.threads != l.shape.threads || c.two_phase != l.shape.two_phase ||
                             c.double_tiles != l.shape.double_tiles;
        l.shape = c;
        int rc = relower ? lower_list(b, l) : BINE_OK;
        if (rc) return rc;
        if ((rc = run_list(b, l, st))) return rc;  // warm-up
        HIPCHK(hipEventRecord(ev.a, st));
        for (int k = 0; k < iters; ++k)
            if ((rc = run_list(b, l, st))) return rc;
        HIPCHK(hipEventRecord(ev.b, st));
        HIPCHK(hipEventSynchronize(ev.b));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, ev.a, ev.b));
        timed[i].emplace_back(t, c);
        if (t < best[i]) {
            best[i] = t;
            pick[i] = c;
        }
        return BINE_OK;
    };
    for (const LaunchShape& c : cands)
        for (size_t i = 0; i < nl; ++i)
            if (int rc = trial(i, c)) return rc;
    for (size_t i = 0; i < nl; ++i)
        if (has_multi_output_reduction(b->lists[i].groups))
            for (const LaunchShape& c : two_phase_trials(timed[i]))
                if (int rc = trial(i, c)) return rc;
    size_t dominant = 0;
    for (size_t i = 0; i < nl; ++i) {
        auto& l = b->lists[i];
        const bool relower = pick[i].vpt != l.shape.vpt || pick[i].threads != l.shape.threads ||
                             pick[i].two_phase != l.shape.two_phase || pick[i].double_tiles != l.shape.double_tiles;
        l.shape = pick[i];
        if (relower) {
            int rc = lower_list(b, l);
            if (rc) return rc;
        }
        if (best[i] > best[dominant]) dominant = i;
    }
    if (nl) b->shape = pick[dominant];
    return BINE_OK;
}

extern "C" int bine_batch_destroy(bine_batch_t b) {
    if (!b) return BINE_OK;
    (void)hipSetDevice(b->comm->cfg.device);
    delete b;
    return BINE_OK;
}
