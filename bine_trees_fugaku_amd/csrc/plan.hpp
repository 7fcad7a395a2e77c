// plan.hpp -- collective schedules as step programs, their symbolic check, and
// lowering to gather-reduce work items for the HIP kernel.
//
// A Program is the reference's message schedule rewritten as data movement
// between per-rank partitions of up to four device buffers (SEND, RECV and two
// temporaries).  Every step is a set of COPY (dst = a) and ADD (dst = a + b)
// operations that read the state before the step and never touch what another
// operation of the same step writes, so one kernel launch can run a step.
//
// The symbolic executor runs a Program on expression trees instead of data.
// It proves the step-level race freedom, checks the final state against the
// collective's MPI semantics, and groups the output ranges by their reduction
// tree; a group whose tree is perfect (every addition joins two equal-sized
// subtrees) becomes one fused work item that reads each leaf once, replays the
// tree pairwise in registers and writes every output copy once.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "bine_hip.h"
#include "gred.hpp"
#include "schedule.hpp"

namespace bine {

enum BufId : uint8_t { BUF_SEND = 0, BUF_RECV = 1, BUF_TMP = 2, BUF_TMP2 = 3 };
constexpr int kNumBufs = 4;

enum OpKind : uint8_t { OP_COPY = 0, OP_ADD = 1 };

struct Loc {
    uint8_t buf;
    uint32_t rank;
    int64_t off;  // elements from the start of that rank's partition
};

struct Op {
    OpKind kind;
    Loc dst, a, b;  // dst = a (COPY) or dst = a + b (ADD)
    int64_t n;
};

// Expected result of one output partition, for verification.
struct Expect {
    enum Kind { NONE = 0, SUM_ALL = 1, COPY_FROM = 2 } kind = NONE;
    uint8_t src_buf = BUF_SEND;
    // element e of output (rank r) depends on input offset src_off(r, e)
    std::function<int64_t(uint32_t, int64_t)> src_off;
    // for COPY_FROM: the input rank
    std::function<uint32_t(uint32_t, int64_t)> src_rank;
};

struct Program {
    uint32_t P = 0;
    int64_t extent[kNumBufs] = {0, 0, 0, 0};  // elements per rank partition
    bool single[kNumBufs] = {false, false, false, false};  // one partition (the root's)
    uint32_t root = 0;
    std::vector<std::vector<Op>> steps;
    struct Output {
        uint8_t buf;
        uint32_t rank;
        int64_t n;
    };
    std::vector<Output> outputs;
    Expect expect;
};

// One gather-reduce group: value(e) = tree over leaves (in order) at e,
// written to every out, for e in [0, n).
struct Group {
    std::vector<PoolEntry> leaves;  // 1 or 2^lg entries, offsets at e = 0
    std::vector<PoolEntry> outs;
    int64_t n = 0;
    uint16_t lg = 0;
};

struct CompiledPlan {
    uint32_t P = 0;
    int64_t extent[kNumBufs] = {0, 0, 0, 0};
    bool single[kNumBufs] = {false, false, false, false};
    std::vector<std::vector<Group>> step_groups;  // STEP execution: one launch per entry
    std::vector<Group> fused;                     // FUSED execution (empty when not realisable)
    bool fused_ok = false;
    std::string fused_why;
};

struct CollRequest {
    int coll = BINE_COLL_ALLREDUCE;
    int64_t count = 0;  // per-rank count as in the MPI call (block count for RS/AG/A2A/scatter/gather)
    int op = BINE_OP_SUM;
    int root = 0;
    bool in_place = false;  // allreduce with sendbuf == recvbuf: the input starts in RECV
};

// Build the step program of one collective for the given vcomm configuration.
// An in-place request gets a program over RECV alone (SEND has no extent), so
// the symbolic race check sees the aliasing.
bool build_program(const Torus& t, const bine_vcomm_config_t& cfg, const CollRequest& req, Program* out,
                   std::string* err);

// Run the program symbolically: race check, semantic check, fused grouping.
bool compile_program(const Program& prog, bool want_fused, CompiledPlan* out, std::string* err);

// Work items for the kernel.
struct Lowered {
    std::vector<GredItem> items;
    std::vector<PoolEntry> pool;
    std::vector<std::pair<uint32_t, uint32_t>> launches;  // (first item, item count)
    std::vector<uint32_t> runs;  // per launch: how many of its items are ITEM_RUN items (they come last)
    uint64_t bytes_read = 0, bytes_written = 0, max_leaves = 0;
    bool unaligned_outs = false;  // some vector item writes an output whose offset is not a whole number of vectors
};
// tile_elems: elements per work item; vec_elems: elements per 16-byte vector.
// Groups whose tree has at least 2^deep_lg leaves use deep_tile_elems instead:
// the kernel splits their leaves across the waves of a workgroup, so their
// tile is one wave wide.
// group_bucket (one launch only): the buffer-table index of every group, instead of `bucket`;
// bucket_vec: per buffer table, whether its addresses allow 16-byte vector access (no ITEM_VEC otherwise).
// two_phase (1 or 2): a launch holding reductions with several outputs becomes two launches, the reductions
// writing their first output, then copies of it to the other outputs (1: copies in group order, 2: tile by
// tile across the groups, as the reductions).  reduction_tiles: tiles per item in the reduction launch of a
// two-phase lowering (1 or 2).
void lower_groups(const std::vector<std::vector<Group>>& launches, int64_t tile_elems, int vec_elems,
                  uint32_t bucket, size_t dtsize, Lowered* out, int64_t deep_tile_elems = 0, int deep_lg = 99,
                  const std::vector<uint32_t>* group_bucket = nullptr, const std::vector<uint8_t>* bucket_vec = nullptr,
                  int two_phase = 0, int reduction_tiles = 1);
// Whether some launch holds a reduction with several outputs (the lists a two-phase lowering changes).
bool has_multi_output_reduction(const std::vector<std::vector<Group>>& launches);

}  // namespace bine
