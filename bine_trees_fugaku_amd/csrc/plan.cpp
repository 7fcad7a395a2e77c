// plan.cpp -- step programs for the eight collectives, symbolic execution,
// fused grouping and lowering to work items (see plan.hpp).
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>

namespace bine {

namespace {

inline Loc at(uint8_t buf, uint32_t rank, int64_t off) { return Loc{buf, rank, off}; }

// Program builder: appends operations to the current step.
struct PB {
    Program* p;
    int op;
    void step() { p->steps.emplace_back(); }
    std::vector<Op>& cur() { return p->steps.back(); }
    // Append an operation, extending the previous one when it continues it
    // exactly (consecutive blocks of the same ranks become one operation).
    static bool follows(const Loc& prev, int64_t n, const Loc& next) {
        return prev.buf == next.buf && prev.rank == next.rank && prev.off + n == next.off;
    }
    void push(const Op& o) {
        auto& v = cur();
        if (!v.empty()) {
            Op& l = v.back();
            if (l.kind == o.kind && follows(l.dst, l.n, o.dst) && follows(l.a, l.n, o.a) &&
                (o.kind == OP_COPY || follows(l.b, l.n, o.b))) {
                l.n += o.n;
                return;
            }
        }
        v.push_back(o);
    }
    void copy(Loc d, Loc a, int64_t n) {
        if (n <= 0) return;
        if (d.buf == a.buf && d.rank == a.rank && d.off == a.off) return;
        push(Op{OP_COPY, d, a, a, n});
    }
    // MPI_Reduce_local(a, b) into d: d = a (+) b.  VOID leaves b as it is.
    void reduce(Loc d, Loc a, Loc b, int64_t n) {
        if (n <= 0) return;
        if (op == BINE_OP_VOID) {
            copy(d, b, n);
            return;
        }
        push(Op{OP_ADD, d, a, b, n});
    }
    // 3-argument reduce_local (lib/libbine_coll.cc:209-267): VOID writes nothing.
    void reduce3(Loc d, Loc a, Loc b, int64_t n) {
        if (n <= 0 || op == BINE_OP_VOID) return;
        push(Op{OP_ADD, d, a, b, n});
    }
};

// Point-to-point messages of one step, matched per (source, destination) in
// posting order, as MPI's non-overtaking rule does.  A matched pair becomes a
// COPY of the sent elements (a shorter message into a larger receive is legal).
struct Matcher {
    struct S {
        uint32_t src, dst;
        Loc from;
        int64_t n;
        size_t seq;
    };
    std::vector<S> sends, recvs;  // recvs: src = sender, dst = receiver, from = receive location, n = capacity
    void send(uint32_t src, uint32_t dst, Loc from, int64_t n) { sends.push_back({src, dst, from, n, sends.size()}); }
    void recv(uint32_t me, uint32_t from_rank, Loc into, int64_t cap) {
        recvs.push_back({from_rank, me, into, cap, recvs.size()});
    }
    bool emit(PB& b, std::string* err) {
        auto key = [](const S& x, const S& y) {
            if (x.src != y.src) return x.src < y.src;
            if (x.dst != y.dst) return x.dst < y.dst;
            return x.seq < y.seq;
        };
        std::sort(sends.begin(), sends.end(), key);
        std::sort(recvs.begin(), recvs.end(), key);
        if (sends.size() != recvs.size()) {
            *err = "schedule posts unmatched messages";
            return false;
        }
        for (size_t i = 0; i < sends.size(); ++i) {
            const S& s = sends[i];
            const S& r = recvs[i];
            if (s.src != r.src || s.dst != r.dst || s.n > r.n) {
                *err = "schedule messages do not match";
                return false;
            }
            b.copy(r.from, s.from, s.n);
        }
        sends.clear();
        recvs.clear();
        return true;
    }
};

// Contiguous runs of blocks i (in block order) for which pred(i) holds.
template <typename F>
void for_runs(uint32_t P, const std::vector<int64_t>& off, const std::vector<int64_t>& cnt, size_t base, F pred,
              const std::function<void(int64_t, int64_t)>& emit) {
    uint32_t i = 0;
    while (i < P) {
        if (!pred(i)) {
            ++i;
            continue;
        }
        int64_t o = off[base + i], n = 0;
        while (i < P && pred(i)) n += cnt[base + i++];
        emit(o, n);
    }
}

// get_blocks_info: ports first, then blocks (lib/libbine.cc:971-1010); element units.
void allreduce_blocks(int64_t count, uint32_t P, int ports, std::vector<int64_t>* off, std::vector<int64_t>* cnt) {
    off->assign((size_t)ports * P, 0);
    cnt->assign((size_t)ports * P, 0);
    int64_t part = count / ports, rem = count % ports, sofar = 0;
    for (int p = 0; p < ports; ++p) {
        int64_t cp = part + (p < rem ? 1 : 0), op = sofar;
        sofar += cp;
        int64_t bp = cp / P, br = cp % P, bs = 0;
        for (uint32_t i = 0; i < P; ++i) {
            int64_t cb = bp + ((int64_t)i < br ? 1 : 0);
            (*off)[(size_t)p * P + i] = op + bs;
            (*cnt)[(size_t)p * P + i] = cb;
            bs += cb;
        }
    }
}

void expect_sum(Program* p, std::function<int64_t(uint32_t, int64_t)> src_off, int op) {
    if (op == BINE_OP_VOID) return;
    p->expect.kind = Expect::SUM_ALL;
    p->expect.src_buf = BUF_SEND;
    p->expect.src_off = std::move(src_off);
}

void expect_copy(Program* p, uint8_t src_buf, std::function<uint32_t(uint32_t, int64_t)> src_rank,
                 std::function<int64_t(uint32_t, int64_t)> src_off) {
    p->expect.kind = Expect::COPY_FROM;
    p->expect.src_buf = src_buf;
    p->expect.src_rank = std::move(src_rank);
    p->expect.src_off = std::move(src_off);
}

// ---- allreduce -------------------------------------------------------------

// bine_coll_b with the B / B_COALESCE step (lib/libbine_common.cc:1631-1735, 930-1159).
bool allreduce_b(const Torus& t, int ports, int family, int64_t count, int op, Program* p, std::string* err) {
    const uint32_t P = t.size, L = t.steps;
    p->extent[BUF_SEND] = p->extent[BUF_RECV] = p->extent[BUF_TMP2] = count;
    std::vector<int64_t> off, cnt;
    allreduce_blocks(count, P, ports, &off, &cnt);
    std::vector<BlockSchedule> bs(ports);
    for (int q = 0; q < ports; ++q) bs[q].build(t, (uint32_t)q, family);
    PB b{p, op};
    b.step();
    for (uint32_t r = 0; r < P; ++r) b.copy(at(BUF_RECV, r, 0), at(BUF_SEND, r, 0), count);
    for (uint32_t s = 0; s < L; ++s) {  // reduce-scatter: receive the peer's send set, reduce into it
        b.step();
        for (int q = 0; q < ports; ++q)
            for (uint32_t r = 0; r < P; ++r) {
                uint32_t peer = bs[q].peer(r, s);
                if (bs[q].peer(peer, s) != r) {
                    *err = "allreduce step is not a perfect matching on this torus";
                    return false;
                }
                for_runs(P, off, cnt, (size_t)q * P, [&](uint32_t i) { return bs[q].sends(peer, s, i); },
                         [&](int64_t o, int64_t n) { b.reduce(at(BUF_RECV, r, o), at(BUF_RECV, peer, o), at(BUF_RECV, r, o), n); });
            }
    }
    for (uint32_t s = 0; s < L; ++s) {  // allgather: steps in reverse, send and receive sets swapped
        const uint32_t bst = L - 1 - s;
        b.step();
        for (int q = 0; q < ports; ++q)
            for (uint32_t r = 0; r < P; ++r) {
                uint32_t peer = bs[q].peer(r, bst);
                for_runs(P, off, cnt, (size_t)q * P, [&](uint32_t i) { return bs[q].sends(r, bst, i); },
                         [&](int64_t o, int64_t n) { b.copy(at(BUF_RECV, r, o), at(BUF_RECV, peer, o), n); });
            }
    }
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, count});
    expect_sum(p, [](uint32_t, int64_t e) { return e; }, op);
    return true;
}

// bine_coll_b with the B_CONT step (lib/libbine_common.cc:1162-1237, chunk params :1524-1550).
bool allreduce_b_cont(const Torus& t, int ports, int family, int64_t count, int op, Program* p, std::string* err) {
    const uint32_t P = t.size, L = t.steps;
    if (!is_pow2(P)) {
        *err = "B_CONT needs a power-of-two number of ranks (lib/libbine_common.cc:1247-1251)";
        return false;
    }
    p->extent[BUF_SEND] = p->extent[BUF_RECV] = count;
    std::vector<int64_t> off, cnt;
    allreduce_blocks(count, P, ports, &off, &cnt);
    struct CP {
        int64_t so, sc, ro, rc;
    };
    std::vector<CP> cp((size_t)ports * P * L);
    std::vector<uint32_t> peers((size_t)ports * P * L);
    for (int q = 0; q < ports; ++q)
        for (uint32_t r = 0; r < P; ++r) {
            std::vector<uint32_t> ps = peer_sequence(t, r, (uint32_t)q, family);
            uint32_t rr = cont_remapped_rank(t, r, (uint32_t)q, family);
            uint32_t min_r = 0, max_r = P;
            for (uint32_t s = 0; s < L; ++s) {
                peers[((size_t)q * P + r) * L + s] = ps[s];
                uint32_t min_s = min_r, max_s = max_r, mid = (min_r + max_r + 1) / 2;
                if (rr < mid) {
                    min_s = mid;
                    max_r = mid;
                } else {
                    max_s = mid;
                    min_r = mid;
                }
                CP c{off[(size_t)q * P + min_s], 0, off[(size_t)q * P + min_r], 0};
                for (uint32_t i = min_s; i < max_s; ++i) c.sc += cnt[(size_t)q * P + i];
                for (uint32_t i = min_r; i < max_r; ++i) c.rc += cnt[(size_t)q * P + i];
                cp[((size_t)q * P + r) * L + s] = c;
            }
        }
    PB b{p, op};
    b.step();
    for (uint32_t r = 0; r < P; ++r) b.copy(at(BUF_RECV, r, 0), at(BUF_SEND, r, 0), count);
    for (int coll = 0; coll < 2; ++coll)
        for (uint32_t s = 0; s < L; ++s) {
            const uint32_t bst = coll == 0 ? s : L - 1 - s;
            b.step();
            for (int q = 0; q < ports; ++q)
                for (uint32_t r = 0; r < P; ++r) {
                    size_t ir = ((size_t)q * P + r) * L + bst;
                    uint32_t peer = peers[ir];
                    size_t ip = ((size_t)q * P + peer) * L + bst;
                    const CP& me = cp[ir];
                    const CP& pe = cp[ip];
                    if (peers[ip] != r) {
                        *err = "B_CONT step is not a matching";
                        return false;
                    }
                    if (coll == 0) {  // receive the peer's send chunk into my receive chunk
                        if (pe.sc != me.rc) { *err = "B_CONT chunk sizes differ"; return false; }
                        b.reduce(at(BUF_RECV, r, me.ro), at(BUF_RECV, peer, pe.so), at(BUF_RECV, r, me.ro), me.rc);
                    } else {          // allgather: roles of the chunks swap
                        if (pe.rc != me.sc) { *err = "B_CONT chunk sizes differ"; return false; }
                        b.copy(at(BUF_RECV, r, me.so), at(BUF_RECV, peer, pe.ro), me.sc);
                    }
                }
        }
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, count});
    expect_sum(p, [](uint32_t, int64_t e) { return e; }, op);
    return true;
}

// bine_coll_l_mpi with shrink / enlarge for non power-of-two dimensions
// (lib/libbine_common.cc:733-846, 269-353).  Each rank's running value moves
// between RECV and TMP so that a step never overwrites what a peer reads.
// In place, the running value starts in RECV instead of SEND, so the first
// step writes TMP (a peer still reads RECV) and an odd step count ends with a
// copy back into RECV.
bool allreduce_l(const Torus& t, int ports, int family, int64_t count, int op, bool in_place, Program* p,
                 std::string* err) {
    const uint32_t P = t.size, D = t.nd;
    p->extent[BUF_SEND] = p->extent[BUF_RECV] = p->extent[BUF_TMP] = count;
    uint32_t vd[kMaxDims] = {1, 1, 1};
    bool all_p2 = true;
    for (uint32_t d = 0; d < D; ++d) {
        if (!is_pow2(t.dims[d])) {
            all_p2 = false;
            vd[d] = 1u << (t.steps_per_dim[d] - 1);
        } else {
            vd[d] = t.dims[d];
        }
    }
    Torus tv;
    if (!Torus::make(vd, D, &tv)) {
        *err = "bad virtual torus";
        return false;
    }
    PB b{p, op};
    std::vector<uint8_t> cur(P, in_place ? BUF_RECV : BUF_SEND);
    std::vector<char> idle(P, 0);
    if (!all_p2) {
        for (uint32_t i = 0; i < D; ++i) {
            if (is_pow2(t.dims[i])) continue;
            const int extra = (int)t.dims[i] - (int)vd[i];
            b.step();
            std::vector<uint32_t> recv_from(P, UINT32_MAX);
            std::vector<char> sends(P, 0);
            for (uint32_t r = 0; r < P; ++r) {
                if (idle[r]) continue;
                int c[kMaxDims];
                t.coords(r, c);
                if (c[i] >= (int)vd[i]) {
                    sends[r] = 1;
                } else if (c[i] + extra >= (int)vd[i]) {
                    int cp[kMaxDims] = {c[0], c[1], c[2]};
                    cp[i] = c[i] + extra;
                    recv_from[r] = t.id(cp);
                }
            }
            for (uint32_t r = 0; r < P; ++r) {
                if (recv_from[r] == UINT32_MAX) continue;
                uint32_t q = recv_from[r];
                if (!sends[q] || idle[q]) {
                    *err = "shrink step has no matching sender";
                    return false;
                }
                // first receive: recvbuf = sendbuf + message; later: recvbuf = tmp + recvbuf
                b.reduce(at(BUF_RECV, r, 0), at(cur[q], q, 0), at(cur[r], r, 0), count);
                cur[r] = BUF_RECV;
            }
            for (uint32_t r = 0; r < P; ++r)
                if (sends[r]) idle[r] = 1;
        }
    }
    const uint32_t Lv = tv.steps;
    std::vector<std::vector<uint32_t>> vpeer((size_t)ports * P);
    std::vector<int> vrank(P);
    for (uint32_t r = 0; r < P; ++r) {
        int c[kMaxDims];
        t.coords(r, c);
        vrank[r] = (int)tv.id(c);
    }
    for (int q = 0; q < ports; ++q)
        for (uint32_t r = 0; r < P; ++r)
            if (!idle[r]) vpeer[(size_t)q * P + r] = peer_sequence(tv, (uint32_t)vrank[r], (uint32_t)q, family);
    const int64_t part = count / ports, rem = count % ports;
    for (uint32_t s = 0; s < Lv; ++s) {
        b.step();
        std::vector<uint8_t> nxt(cur);
        for (uint32_t r = 0; r < P; ++r) {
            if (idle[r]) continue;
            uint8_t target;
            if (cur[r] == BUF_SEND) target = ((Lv - s) & 1) ? BUF_RECV : BUF_TMP;
            else target = cur[r] == BUF_RECV ? BUF_TMP : BUF_RECV;
            int64_t sofar = 0;
            for (int q = 0; q < ports; ++q) {
                int64_t cpn = part + (q < rem ? 1 : 0);
                uint32_t vp = vpeer[(size_t)q * P + r][s];
                int c[kMaxDims];
                tv.coords(vp, c);
                uint32_t peer = t.id(c);
                if (idle[peer]) {
                    *err = "latency-optimal step reaches an idle rank";
                    return false;
                }
                b.reduce(at(target, r, sofar), at(cur[peer], peer, sofar), at(cur[r], r, sofar), cpn);
                sofar += cpn;
            }
            nxt[r] = target;
        }
        cur = nxt;
    }
    bool fix = false;
    for (uint32_t r = 0; r < P; ++r) fix |= (!idle[r] && cur[r] != BUF_RECV);
    if (fix) {
        b.step();
        for (uint32_t r = 0; r < P; ++r)
            if (!idle[r] && cur[r] != BUF_RECV) b.copy(at(BUF_RECV, r, 0), at(cur[r], r, 0), count);
    }
    if (!all_p2) {
        for (int d = (int)D - 1; d >= 0; --d) {
            if (is_pow2(t.dims[d])) continue;
            const int extra = (int)t.dims[d] - (int)vd[d];
            b.step();
            for (uint32_t r = 0; r < P; ++r) {
                int c[kMaxDims];
                t.coords(r, c);
                if (c[d] >= (int)vd[d]) {
                    int cp[kMaxDims] = {c[0], c[1], c[2]};
                    cp[d] = c[d] - extra;
                    b.copy(at(BUF_RECV, r, 0), at(BUF_RECV, t.id(cp), 0), count);
                }
            }
        }
    }
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, count});
    expect_sum(p, [](uint32_t, int64_t e) { return e; }, op);
    return true;
}

// ---- reduce-scatter (contiguous, permute first) -------------------------------

// bine_reduce_scatter_mpi_contiguous (lib/libbine_coll_reduce_scatter.cc:660-758): one port.
bool reduce_scatter_cont(const Torus& t, int ports, int family, int dist, int64_t c, int op, Program* p,
                         std::string* err) {
    const uint32_t P = t.size, L = t.steps;
    if (ports != 1) {
        *err = "reduce_scatter writes recvbuf + offset_port, valid for one port only "
               "(lib/libbine_coll_reduce_scatter.cc:707,743)";
        return false;
    }
    if (!is_pow2(P)) {
        *err = "reduce_scatter needs a power-of-two number of ranks (lib/libbine_coll_reduce_scatter.cc:764-770)";
        return false;
    }
    p->extent[BUF_SEND] = (int64_t)P * c;
    p->extent[BUF_RECV] = c;
    p->extent[BUF_TMP] = (int64_t)P * c;
    PB b{p, op};
    b.step();
    std::vector<std::vector<uint32_t>> peers(P);
    for (uint32_t r = 0; r < P; ++r) {
        Tree tr;
        if (!build_tree(t, r, 0, family, dist, &tr)) {
            *err = "tree numbering failed";
            return false;
        }
        peers[r] = peer_sequence(t, r, 0, family);
        for (uint32_t i = 0; i < P; ++i) b.copy(at(BUF_TMP, r, (int64_t)tr.remapped[i] * c), at(BUF_SEND, r, (int64_t)i * c), c);
    }
    for (uint32_t s = 0; s < L; ++s) {
        b.step();
        const int64_t cnt = (int64_t)(P >> (s + 1)) * c;
        for (uint32_t r = 0; r < P; ++r) {
            uint32_t peer = dist == DIST_DECREASING ? peers[r][L - 1 - s] : peers[r][s];
            uint32_t back = dist == DIST_DECREASING ? peers[peer][L - 1 - s] : peers[peer][s];
            if (back != r) {
                *err = "reduce_scatter step is not a matching";
                return false;
            }
            // Sendrecv_replace of the upper half, then lower = received + lower
            if (s + 1 == L) b.reduce3(at(BUF_RECV, r, 0), at(BUF_TMP, peer, cnt), at(BUF_TMP, r, 0), cnt);
            else b.reduce(at(BUF_TMP, r, 0), at(BUF_TMP, peer, cnt), at(BUF_TMP, r, 0), cnt);
        }
    }
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, c});
    expect_sum(p, [c](uint32_t r, int64_t e) { return (int64_t)r * c + e; }, op);
    return true;
}

// ---- allgather ----------------------------------------------------------------

// bine_allgather_mpi_contiguous (lib/libbine_coll_allgather.cc:289-374)
bool allgather_cont(const Torus& t, int ports, int family, int dist, int64_t c, Program* p, std::string* err) {
    const uint32_t P = t.size, L = t.steps;
    if (!is_pow2(P)) {
        *err = "contiguous allgather needs a power-of-two number of ranks";
        return false;
    }
    const int64_t part = c / ports, rem = c % ports;
    const int64_t c0max = part + (rem ? 1 : 0);
    p->extent[BUF_SEND] = c;
    p->extent[BUF_RECV] = (int64_t)P * c;
    p->extent[BUF_TMP] = (int64_t)ports * P * c0max;
    PB b{p, BINE_OP_SUM};
    const int odist = dist == DIST_DECREASING ? DIST_INCREASING : DIST_DECREASING;
    std::vector<std::vector<std::vector<uint32_t>>> peers(ports, std::vector<std::vector<uint32_t>>(P));
    b.step();
    int64_t pofs = 0;
    std::vector<int64_t> port_off(ports), port_cnt(ports);
    for (int q = 0; q < ports; ++q) {
        port_cnt[q] = part + (q < rem ? 1 : 0);
        port_off[q] = pofs;
        pofs += port_cnt[q];
        for (uint32_t r = 0; r < P; ++r) {
            peers[q][r] = peer_sequence(t, r, (uint32_t)q, family);
            b.copy(at(BUF_TMP, r, (int64_t)q * P * c0max), at(BUF_SEND, r, port_off[q]), port_cnt[q]);
        }
    }
    for (uint32_t s = 0; s < L; ++s) {
        b.step();
        for (int q = 0; q < ports; ++q) {
            const int64_t n = ((int64_t)1 << s) * port_cnt[q], reg = (int64_t)q * P * c0max;
            for (uint32_t r = 0; r < P; ++r) {
                uint32_t peer = dist == DIST_DECREASING ? peers[q][r][L - 1 - s] : peers[q][r][s];
                uint32_t back = dist == DIST_DECREASING ? peers[q][peer][L - 1 - s] : peers[q][peer][s];
                if (back != r) {
                    *err = "allgather step is not a matching";
                    return false;
                }
                b.copy(at(BUF_TMP, r, reg + n), at(BUF_TMP, peer, reg), n);
            }
        }
    }
    b.step();
    for (int q = 0; q < ports; ++q)
        for (uint32_t r = 0; r < P; ++r) {
            Tree tr;
            if (!build_tree(t, r, (uint32_t)q, family, odist, &tr)) {
                *err = "tree numbering failed";
                return false;
            }
            for (uint32_t i = 0; i < P; ++i)
                b.copy(at(BUF_RECV, r, (int64_t)i * c + port_off[q]),
                       at(BUF_TMP, r, (int64_t)q * P * c0max + (int64_t)tr.remapped[i] * port_cnt[q]), port_cnt[q]);
        }
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, (int64_t)P * c});
    expect_copy(p, BUF_SEND, [c](uint32_t, int64_t e) { return (uint32_t)(e / c); },
                [c](uint32_t, int64_t e) { return e % c; });
    return true;
}

// bine_allgather_send_mpi (lib/libbine_coll_allgather.cc:492-574): one port.
bool allgather_send(const Torus& t, int ports, int family, int dist, int64_t c, Program* p, std::string* err) {
    const uint32_t P = t.size, L = t.steps;
    if (ports != 1 || !is_pow2(P)) {
        *err = "CONT_SEND allgather needs one port and a power-of-two number of ranks";
        return false;
    }
    p->extent[BUF_SEND] = c;
    p->extent[BUF_RECV] = (int64_t)P * c;
    PB b{p, BINE_OP_SUM};
    const int odist = dist == DIST_DECREASING ? DIST_INCREASING : DIST_DECREASING;
    Tree tr;
    if (!build_tree(t, 0, 0, family, odist, &tr)) {
        *err = "tree numbering failed";
        return false;
    }
    std::vector<uint32_t> inv(P);
    for (uint32_t i = 0; i < P; ++i) inv[tr.remapped[i]] = i;
    std::vector<std::vector<uint32_t>> peers(P);
    std::vector<int64_t> minres(P);
    Matcher m;
    b.step();
    for (uint32_t r = 0; r < P; ++r) {
        peers[r] = peer_sequence(t, r, 0, family);
        m.send(r, inv[r], at(BUF_SEND, r, 0), c);
        m.recv(r, tr.remapped[r], at(BUF_RECV, r, (int64_t)tr.remapped[r] * c), c);
        minres[r] = tr.remapped[r];
    }
    if (!m.emit(b, err)) return false;
    int64_t nb = 1;
    for (uint32_t s = 0; s < L; ++s) {
        b.step();
        std::vector<int64_t> mrr(P);
        for (uint32_t r = 0; r < P; ++r) {
            uint32_t peer = dist == DIST_DECREASING ? peers[r][L - 1 - s] : peers[r][s];
            mrr[r] = tr.remapped[peer] & ~((1u << s) - 1);
            m.send(r, peer, at(BUF_RECV, r, minres[r] * c), nb * c);
            m.recv(r, peer, at(BUF_RECV, r, mrr[r] * c), nb * c);
        }
        if (!m.emit(b, err)) return false;
        for (uint32_t r = 0; r < P; ++r) minres[r] = std::min(minres[r], mrr[r]);
        nb *= 2;
    }
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, (int64_t)P * c});
    expect_copy(p, BUF_SEND, [c](uint32_t, int64_t e) { return (uint32_t)(e / c); },
                [c](uint32_t, int64_t e) { return e % c; });
    return true;
}

// bine_allgather_blocks_mpi (lib/libbine_coll_allgather.cc:1161-1228): 1D ranks.
uint32_t get_nu(uint32_t rank, uint32_t size) {
    const uint32_t nbits = (uint32_t)ceil_log2(size);
    auto nu = [&](uint32_t nb) { return bit_reverse32(nb ^ (nb >> 1)) >> (32 - nbits); };
    const int64_t r = rank, s = size;
    int64_t a = (rank & 1) ? r : -r, bb = (rank & 1) ? r - s : -r + s;
    auto fits = [&](int64_t x) {
        int64_t lo = 0, hi = 0;
        for (uint32_t i = 1; i < nbits; i += 2) lo -= (int64_t)1 << i;
        for (uint32_t i = 0; i < nbits; i += 2) hi += (int64_t)1 << i;
        return x >= lo && x <= hi;
    };
    bool fa = fits(a), fb = fits(bb);
    if (!fa) return nu(to_negabinary((int32_t)bb));
    if (!fb) return nu(to_negabinary((int32_t)a));
    return std::min(nu(to_negabinary((int32_t)a)), nu(to_negabinary((int32_t)bb)));
}

bool allgather_blocks(uint32_t P, int64_t c, Program* p, std::string* err) {
    if (!is_pow2(P)) {
        *err = "BLOCKS allgather needs a power-of-two number of ranks";
        return false;
    }
    p->extent[BUF_SEND] = c;
    p->extent[BUF_RECV] = (int64_t)P * c;
    PB b{p, BINE_OP_SUM};
    b.step();
    for (uint32_t r = 0; r < P; ++r) b.copy(at(BUF_RECV, r, (int64_t)r * c), at(BUF_SEND, r, 0), c);
    const int nbits = ceil_log2(P);
    Matcher m;
    for (int step = 0, inv = 1 << (nbits - 1); inv > 0; inv >>= 1, ++step) {
        for (uint32_t block = 1; block < P; ++block) {
            int k = 31 - __builtin_clz(get_nu(block, P));
            if (k != step) continue;
            b.step();
            for (uint32_t r = 0; r < P; ++r) {
                int dist = from_negabinary((uint32_t)((inv << 1) - 1));
                int partner = (r % 2 == 0) ? pmod((int)r + dist, (int)P) : pmod((int)r - dist, (int)P);
                int64_t bts, btr;
                if (r % 2 == 0) {
                    btr = pmod((int)(block + r), (int)P);
                    bts = pmod(partner - (int)block, (int)P);
                } else {
                    btr = pmod((int)r - (int)block, (int)P);
                    bts = pmod((int)block + partner, (int)P);
                }
                if (bts != partner) m.send(r, (uint32_t)partner, at(BUF_RECV, r, bts * c), c);
                if (btr != (int64_t)r) m.recv(r, (uint32_t)partner, at(BUF_RECV, r, btr * c), c);
            }
            if (!m.emit(b, err)) return false;
        }
    }
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, (int64_t)P * c});
    expect_copy(p, BUF_SEND, [c](uint32_t, int64_t e) { return (uint32_t)(e / c); },
                [c](uint32_t, int64_t e) { return e % c; });
    return true;
}

// ---- alltoall -----------------------------------------------------------------

// bine_alltoall_mpi (lib/libbine_coll_alltoall.cc:411-540).  The reference packs
// the blocks to send into recvbuf and compacts the kept ones inside tmpbuf;
// here the kept blocks move into the other temporary so a launch never reads
// what it writes.
bool alltoall_bine(uint32_t P, int64_t c, Program* p, std::string* err) {
    if (!is_pow2(P) || P < 2) {
        *err = "Bine alltoall needs a power-of-two number of ranks (lib/libbine_coll_alltoall.cc:484-485)";
        return false;
    }
    const int64_t E = (int64_t)P * c;
    p->extent[BUF_SEND] = p->extent[BUF_RECV] = p->extent[BUF_TMP] = p->extent[BUF_TMP2] = E;
    PB b{p, BINE_OP_SUM};
    std::vector<std::vector<uint32_t>> res(P, std::vector<uint32_t>(P));
    std::vector<size_t> nres(P, P);
    for (uint32_t r = 0; r < P; ++r)
        for (uint32_t i = 0; i < P; ++i) res[r][i] = i;
    std::vector<uint32_t> rmap(P);
    for (uint32_t i = 0; i < P; ++i) rmap[i] = remap_rank(i, P);
    uint8_t cur = BUF_SEND, nxt = BUF_TMP;
    const int nbits = ceil_log2(P);
    int mask = 1, inv = 1 << (nbits - 1), bfm = ~(inv - 1);
    while (mask < (int)P) {
        b.step();  // pack the blocks to send into recvbuf, keep the rest in order
        std::vector<uint32_t> partner(P);
        std::vector<std::vector<uint32_t>> resn(P);
        for (uint32_t r = 0; r < P; ++r) {
            int dist = from_negabinary((uint32_t)((mask << 1) - 1));
            partner[r] = (uint32_t)((r % 2 == 0) ? pmod((int)r + dist, (int)P) : pmod((int)r - dist, (int)P));
            uint32_t min_s = rmap[partner[r]] & (uint32_t)bfm, max_s = min_s + inv - 1;
            int64_t ns = 0, nk = 0;
            for (uint32_t i = 0; i < P; ++i) {
                uint32_t block = res[r][i % nres[r]];
                uint32_t rb = rmap[block];
                if (rb >= min_s && rb <= max_s) {
                    b.copy(at(BUF_RECV, r, ns++ * c), at(cur, r, (int64_t)i * c), c);
                } else {
                    b.copy(at(nxt, r, nk++ * c), at(cur, r, (int64_t)i * c), c);
                    resn[r].push_back(block);
                }
            }
            if (ns != (int64_t)P / 2 || nk != (int64_t)P / 2) {
                *err = "alltoall split is not balanced";
                return false;
            }
        }
        b.step();  // exchange the packed halves
        for (uint32_t r = 0; r < P; ++r) {
            if (partner[partner[r]] != r) {
                *err = "alltoall step is not a matching";
                return false;
            }
            b.copy(at(nxt, r, (int64_t)(P / 2) * c), at(BUF_RECV, partner[r], 0), (int64_t)(P / 2) * c);
            res[r] = resn[r];
            nres[r] /= 2;
        }
        cur = nxt;
        nxt = (nxt == BUF_TMP) ? BUF_TMP2 : BUF_TMP;
        mask <<= 1;
        inv >>= 1;
        bfm >>= 1;
    }
    b.step();  // final permutation by the Mersenne decomposition of the distance
    for (uint32_t r = 0; r < P; ++r)
        for (uint32_t i = 0; i < P; ++i) {
            int rot = (r % 2 == 0) ? pmod((int)i - (int)r, (int)P) : pmod((int)r - (int)i, (int)P);
            int64_t lo = 0, hi = 0;
            for (int k = 1; k < nbits; k += 2) lo -= (int64_t)1 << k;
            for (int k = 0; k < nbits; k += 2) hi += (int64_t)1 << k;
            uint32_t repr = (rot >= lo && rot <= hi) ? to_negabinary(rot) : to_negabinary(rot - (int)P);
            int index = remap_distance_doubling(repr);
            b.copy(at(BUF_RECV, r, (int64_t)i * c), at(cur, r, (int64_t)index * c), c);
        }
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, E});
    expect_copy(p, BUF_SEND, [c](uint32_t, int64_t e) { return (uint32_t)(e / c); },
                [c](uint32_t r, int64_t e) { return (int64_t)r * c + e % c; });
    return true;
}

// bruck_alltoall (lib/libbine_coll_alltoall.cc:18-107)
bool alltoall_bruck(uint32_t P, int64_t c, Program* p, std::string* err) {
    const int64_t E = (int64_t)P * c, half = (int64_t)((P + 1) / 2) * c;
    p->extent[BUF_SEND] = p->extent[BUF_RECV] = p->extent[BUF_TMP] = E;
    p->extent[BUF_TMP2] = 2 * half;
    PB b{p, BINE_OP_SUM};
    b.step();
    for (uint32_t r = 0; r < P; ++r)
        for (uint32_t i = 0; i < P; ++i) b.copy(at(BUF_TMP, r, (int64_t)((i - r + P) % P) * c), at(BUF_SEND, r, (int64_t)i * c), c);
    for (uint32_t k = 1; k < P; k <<= 1) {
        std::vector<uint32_t> idx;
        for (uint32_t i = k; i < P; ++i)
            if (i & k) idx.push_back(i);
        b.step();
        for (uint32_t r = 0; r < P; ++r)
            for (size_t j = 0; j < idx.size(); ++j) b.copy(at(BUF_TMP2, r, (int64_t)j * c), at(BUF_TMP, r, (int64_t)idx[j] * c), c);
        b.step();
        for (uint32_t r = 0; r < P; ++r) {
            uint32_t from = (r - k + P) % P;
            for (size_t j = 0; j < idx.size(); ++j) b.copy(at(BUF_TMP, r, (int64_t)idx[j] * c), at(BUF_TMP2, from, (int64_t)j * c), c);
        }
    }
    b.step();
    for (uint32_t r = 0; r < P; ++r)
        for (uint32_t i = 0; i < P; ++i) b.copy(at(BUF_RECV, r, (int64_t)((r - i + P) % P) * c), at(BUF_TMP, r, (int64_t)i * c), c);
    (void)err;
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, E});
    expect_copy(p, BUF_SEND, [c](uint32_t, int64_t e) { return (uint32_t)(e / c); },
                [c](uint32_t r, int64_t e) { return (int64_t)r * c + e % c; });
    return true;
}

// ---- rooted 1D collectives -------------------------------------------------------

inline int bine_partner(uint32_t r, int mask, uint32_t P) {
    int dist = from_negabinary((uint32_t)((mask << 1) - 1));
    return (r % 2 == 0) ? pmod((int)r + dist, (int)P) : pmod((int)r - dist, (int)P);
}

void displs_of(int64_t count, uint32_t P, std::vector<int64_t>* d, std::vector<int64_t>* rc) {
    d->resize(P);
    rc->resize(P);
    int64_t cpr = count / P, rem = count % P;
    for (uint32_t i = 0; i < P; ++i) {
        (*d)[i] = cpr * i + ((int64_t)i < rem ? i : rem);
        (*rc)[i] = cpr + ((int64_t)i < rem ? 1 : 0);
    }
}

// bine_bcast_l_mpi (lib/libbine_coll_bcast.cc:635-666); the buffer is RECV.
bool bcast_tree(uint32_t P, int64_t count, int root, Program* p, std::string* err) {
    p->extent[BUF_RECV] = count;
    PB b{p, BINE_OP_SUM};
    const int nbits = ceil_log2(P);
    std::vector<char> recvd(P, 0);
    recvd[root] = 1;
    Matcher m;
    for (int mask = 1 << (nbits - 1); mask > 0; mask >>= 1) {
        b.step();
        std::vector<char> got(P, 0);
        for (uint32_t r = 0; r < P; ++r) {
            int vr = pmod((int)r - root, (int)P);
            uint32_t vnb = to_negabinary(vr);
            int partner = pmod(from_negabinary(vnb ^ (uint32_t)((mask << 1) - 1)) + root, (int)P);
            uint32_t ml = (uint32_t)((mask << 1) - 1), lsbs = vnb & ml;
            bool eq = lsbs == 0 || lsbs == ml;
            if (recvd[r]) m.send(r, (uint32_t)partner, at(BUF_RECV, r, 0), count);
            else if (eq) {
                m.recv(r, (uint32_t)partner, at(BUF_RECV, r, 0), count);
                got[r] = 1;
            }
        }
        if (!m.emit(b, err)) return false;
        for (uint32_t r = 0; r < P; ++r) recvd[r] |= got[r];
    }
    p->root = (uint32_t)root;
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, count});
    expect_copy(p, BUF_RECV, [root](uint32_t, int64_t) { return (uint32_t)root; }, [](uint32_t, int64_t e) { return e; });
    return true;
}

// bine_bcast_scatter_allgather_mpi (lib/libbine_coll_bcast.cc:985-1095), root 0.
bool bcast_sa(uint32_t P, int64_t count, int root, Program* p, std::string* err) {
    if (root != 0 || !is_pow2(P)) {
        *err = "scatter-allgather bcast supports root 0 and power-of-two ranks (lib/libbine_coll_bcast.cc:993)";
        return false;
    }
    p->extent[BUF_RECV] = count;
    PB b{p, BINE_OP_SUM};
    std::vector<int64_t> dis, rcs;
    displs_of(count, P, &dis, &rcs);
    const int nbits = ceil_log2(P);
    const int inv0 = 1 << (nbits - 1);
    std::vector<int> rm(P), rmask(P);
    for (uint32_t r = 0; r < P; ++r) {
        rm[r] = (int)remap_rank(r, P);
        rmask[r] = r != (uint32_t)root ? (1 << (__builtin_ffs(rm[r]) - 1)) : (inv0 << 1);
    }
    auto rng = [&](int first, int inv, int64_t* o, int64_t* n) {
        int last = first + inv - 1;
        *o = dis[first];
        *n = dis[last] - dis[first] + rcs[last];
    };
    std::vector<char> recvd(P, 0);
    recvd[root] = 1;
    Matcher m;
    int mask = 1, inv = inv0, bfm = ~(inv - 1);
    while (mask < (int)P) {
        b.step();
        std::vector<char> got(P, 0);
        for (uint32_t r = 0; r < P; ++r) {
            int partner = bine_partner(r, mask, P);
            int64_t so, sn, ro, rn;
            rng((int)remap_rank((uint32_t)partner, P) & bfm, inv, &so, &sn);
            rng(rm[r] & bfm, inv, &ro, &rn);
            if (recvd[r]) m.send(r, (uint32_t)partner, at(BUF_RECV, r, so), sn);
            else if (inv == rmask[r] || partner == root) {
                m.recv(r, (uint32_t)partner, at(BUF_RECV, r, ro), rn);
                got[r] = 1;
            }
        }
        if (!m.emit(b, err)) return false;
        for (uint32_t r = 0; r < P; ++r) recvd[r] |= got[r];
        mask <<= 1;
        inv >>= 1;
        bfm >>= 1;
    }
    mask >>= 1;
    inv = 1;
    bfm = ~0;
    while (mask > 0) {
        b.step();
        for (uint32_t r = 0; r < P; ++r) {
            int partner = bine_partner(r, mask, P);
            if (!(inv == rmask[r])) {
                int64_t o, n;
                rng(rm[r] & bfm, inv, &o, &n);
                m.send(r, (uint32_t)partner, at(BUF_RECV, r, o), n);
            }
            if (!(inv < rmask[r])) {
                int64_t o, n;
                rng((int)remap_rank((uint32_t)partner, P) & bfm, inv, &o, &n);
                m.recv(r, (uint32_t)partner, at(BUF_RECV, r, o), n);
            }
        }
        if (!m.emit(b, err)) return false;
        mask >>= 1;
        inv <<= 1;
        bfm <<= 1;
    }
    p->root = (uint32_t)root;
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, count});
    expect_copy(p, BUF_RECV, [root](uint32_t, int64_t) { return (uint32_t)root; }, [](uint32_t, int64_t e) { return e; });
    return true;
}

// bine_reduce_mpi (lib/libbine_coll_reduce.cc:575-615).  Every rank accumulates
// in TMP (the reference's recvbuf at the root, a scratch copy elsewhere); the
// root's result is copied to the single-partition RECV at the end.
bool reduce_tree(uint32_t P, int64_t count, int root, int op, Program* p, std::string* err) {
    p->extent[BUF_SEND] = p->extent[BUF_TMP] = count;
    p->extent[BUF_RECV] = count;
    p->single[BUF_RECV] = true;
    p->root = (uint32_t)root;
    PB b{p, op};
    b.step();
    for (uint32_t r = 0; r < P; ++r) b.copy(at(BUF_TMP, r, 0), at(BUF_SEND, r, 0), count);
    std::vector<char> active(P, 1);
    for (int mask = 1; mask < (int)P; mask <<= 1) {
        b.step();
        std::vector<int> sender_to(P, -1), recv_from(P, -1);
        for (uint32_t r = 0; r < P; ++r) {
            if (!active[r]) continue;
            int vr = pmod((int)r - root, (int)P);
            uint32_t vnb = to_negabinary(vr);
            int partner = pmod(from_negabinary(vnb ^ (uint32_t)((mask << 1) - 1)) + root, (int)P);
            uint32_t ml = (uint32_t)((mask << 2) - 1), lsbs = vnb & ml;
            bool eq = lsbs == 0 || lsbs == ml;
            if (!eq || ((mask << 1) >= (int)P && (int)r != root)) sender_to[r] = partner;
            else recv_from[r] = partner;
        }
        for (uint32_t r = 0; r < P; ++r) {
            if (recv_from[r] < 0) continue;
            int q = recv_from[r];
            if (sender_to[q] != (int)r) {
                *err = "reduce step has no matching sender";
                return false;
            }
            b.reduce(at(BUF_TMP, r, 0), at(BUF_TMP, (uint32_t)q, 0), at(BUF_TMP, r, 0), count);
        }
        for (uint32_t r = 0; r < P; ++r) {
            if (sender_to[r] >= 0) {
                if (recv_from[sender_to[r]] != (int)r) {
                    *err = "reduce step sends to a rank that does not receive";
                    return false;
                }
                active[r] = 0;
            }
        }
    }
    b.step();
    b.copy(at(BUF_RECV, (uint32_t)root, 0), at(BUF_TMP, (uint32_t)root, 0), count);
    p->outputs.push_back({BUF_RECV, (uint32_t)root, count});
    expect_sum(p, [](uint32_t, int64_t e) { return e; }, op);
    return true;
}

// bine_reduce_redscat_gather_mpi (lib/libbine_coll_reduce.cc:1042-1152), root 0.
bool reduce_rsg(uint32_t P, int64_t count, int root, int op, Program* p, std::string* err) {
    if (root != 0 || !is_pow2(P)) {
        *err = "reduce-scatter-gather reduce supports root 0 and power-of-two ranks (lib/libbine_coll_reduce.cc:1049)";
        return false;
    }
    p->extent[BUF_SEND] = p->extent[BUF_TMP] = count;
    p->extent[BUF_RECV] = count;
    p->single[BUF_RECV] = true;
    p->root = (uint32_t)root;
    PB b{p, op};
    std::vector<int64_t> dis, rcs;
    displs_of(count, P, &dis, &rcs);
    auto rng = [&](int first, int inv, int64_t* o, int64_t* n) {
        int last = first + inv - 1;
        *o = dis[first];
        *n = dis[last] - dis[first] + rcs[last];
    };
    b.step();
    for (uint32_t r = 0; r < P; ++r) b.copy(at(BUF_TMP, r, 0), at(BUF_SEND, r, 0), count);
    const int nbits = ceil_log2(P);
    int mask = 1, inv = 1 << (nbits - 1), bfm = ~(inv - 1);
    while (mask < (int)P) {
        b.step();
        for (uint32_t r = 0; r < P; ++r) {
            int partner = bine_partner(r, mask, P);
            // I receive (and reduce) the range named by remap(me); my partner sends
            // exactly that range of its own buffer, since its partner is me.
            int64_t ro, rn;
            rng((int)remap_rank(r, P) & bfm, inv, &ro, &rn);
            if (bine_partner((uint32_t)partner, mask, P) != (int)r) {
                *err = "reduce-scatter step is not a matching";
                return false;
            }
            b.reduce(at(BUF_TMP, r, ro), at(BUF_TMP, (uint32_t)partner, ro), at(BUF_TMP, r, ro), rn);
        }
        mask <<= 1;
        inv >>= 1;
        bfm >>= 1;
    }
    mask >>= 1;
    inv = 1;
    bfm = ~0;
    std::vector<char> active(P, 1);
    std::vector<int> rmask(P);
    for (uint32_t r = 0; r < P; ++r) rmask[r] = 1 << (__builtin_ffs((int)remap_rank(r, P)) - 1);
    Matcher m;
    while (mask > 0) {
        b.step();
        for (uint32_t r = 0; r < P; ++r) {
            if (!active[r]) continue;
            int partner = bine_partner(r, mask, P);
            int64_t o, n;
            if (inv & rmask[r]) {
                rng((int)remap_rank(r, P) & bfm, inv, &o, &n);
                m.send(r, (uint32_t)partner, at(BUF_TMP, r, o), n);
                active[r] = 0;
            } else {
                rng((int)remap_rank((uint32_t)partner, P) & bfm, inv, &o, &n);
                m.recv(r, (uint32_t)partner, at(BUF_TMP, r, o), n);
            }
        }
        if (!m.emit(b, err)) return false;
        mask >>= 1;
        inv <<= 1;
        bfm <<= 1;
    }
    b.step();
    b.copy(at(BUF_RECV, 0, 0), at(BUF_TMP, 0, 0), count);
    p->outputs.push_back({BUF_RECV, 0, count});
    expect_sum(p, [](uint32_t, int64_t e) { return e; }, op);
    return true;
}

// bine_scatter_mpi (lib/libbine_coll_scatter.cc:187-324).  SEND is the root's
// single partition; receive windows of inner nodes live in TMP.
bool scatter_tree(uint32_t P, int64_t c, int root, Program* p, std::string* err) {
    if (root != 0) {
        *err = "scatter supports root 0: the reference halves by rank parity while rooting the tree at "
               "rank - root (lib/libbine_coll_scatter.cc:206-209)";
        return false;
    }
    if (!is_pow2(P)) {
        *err = "scatter needs a power-of-two number of ranks";
        return false;
    }
    p->extent[BUF_SEND] = (int64_t)P * c;
    p->single[BUF_SEND] = true;
    p->extent[BUF_RECV] = c;
    p->extent[BUF_TMP] = (int64_t)P * c;
    p->root = (uint32_t)root;
    PB b{p, BINE_OP_SUM};
    const int nbits = ceil_log2(P);
    std::vector<int> hd(P), sboff(P);
    std::vector<int64_t> minr(P), maxr(P);
    std::vector<char> recvd(P, 0), leaf(P, 0);
    std::vector<uint8_t> sbuf(P, BUF_TMP);
    const uint32_t msk = (1u << nbits) - 1;
    for (uint32_t r = 0; r < P; ++r) {
        hd[r] = (r % 2) ? -1 : 1;
        if (nbits % 2 == 0) hd[r] = -hd[r];
        if (r % 2 == 0) {
            maxr[r] = pmod((int)((r + 0x55555555u) & msk), (int)P);
            minr[r] = pmod((int)((r - 0xAAAAAAAAu) & msk), (int)P);
        } else {
            minr[r] = pmod((int)((r - 0x55555555u) & msk), (int)P);
            maxr[r] = pmod((int)((r + 0xAAAAAAAAu) & msk), (int)P);
        }
        sboff[r] = (int)r;
    }
    recvd[root] = 1;
    sbuf[root] = BUF_SEND;
    Matcher m;
    for (int mask = 1 << (nbits - 1); mask > 0; mask >>= 1) {
        b.step();
        std::vector<char> got(P, 0);
        for (uint32_t r = 0; r < P; ++r) {
            int vr = pmod((int)r - root, (int)P);
            uint32_t vnb = to_negabinary(vr);
            int partner = pmod(from_negabinary(vnb ^ (uint32_t)((mask << 1) - 1)) + root, (int)P);
            uint32_t ml = (uint32_t)((mask << 1) - 1), lsbs = vnb & ml;
            bool eq = lsbs == 0 || lsbs == ml;
            int64_t ts = minr[r], te = pmod((int)minr[r] + mask - 1, (int)P);
            int64_t bs = pmod((int)te + 1, (int)P), be = maxr[r];
            int64_t ss, se, rs, re;
            if (hd[r] == 1) {
                ss = bs; se = be; rs = ts; re = te;
                maxr[r] = pmod((int)maxr[r] - mask, (int)P);
            } else {
                ss = ts; se = te; rs = bs; re = be;
                minr[r] = pmod((int)minr[r] + mask, (int)P);
            }
            uint32_t sb_rank = sbuf[r] == BUF_SEND ? (uint32_t)root : r;
            if (recvd[r]) {
                if (se >= ss) {
                    m.send(r, (uint32_t)partner, at(sbuf[r], sb_rank, ss * c), c * (se - ss + 1));
                } else {
                    m.send(r, (uint32_t)partner, at(sbuf[r], sb_rank, ss * c), c * ((int64_t)P - ss));
                    m.send(r, (uint32_t)partner, at(sbuf[r], sb_rank, 0), c * (se + 1));
                }
            } else if (eq) {
                int64_t nb = pmod((int)(re - rs + 1), (int)P);
                uint8_t rbuf;
                if (rs == re) {
                    rbuf = BUF_RECV;
                    leaf[r] = 1;
                } else {
                    rbuf = BUF_TMP;
                    sbuf[r] = BUF_TMP;
                    minr[r] = 0;
                    maxr[r] = nb - 1;
                    sboff[r] = pmod((int)r - (int)rs, (int)P);
                }
                if (re >= rs) {
                    m.recv(r, (uint32_t)partner, at(rbuf, r, 0), c * nb);
                } else {
                    m.recv(r, (uint32_t)partner, at(rbuf, r, 0), c * ((int64_t)P - rs));
                    m.recv(r, (uint32_t)partner, at(rbuf, r, c * ((int64_t)P - rs)), c * (re + 1));
                }
                got[r] = 1;
            }
        }
        if (!m.emit(b, err)) return false;
        for (uint32_t r = 0; r < P; ++r) recvd[r] |= got[r];
        for (uint32_t r = 0; r < P; ++r) hd[r] = -hd[r];
    }
    b.step();
    for (uint32_t r = 0; r < P; ++r)
        if (!leaf[r]) b.copy(at(BUF_RECV, r, 0), at(sbuf[r], sbuf[r] == BUF_SEND ? (uint32_t)root : r, (int64_t)sboff[r] * c), c);
    for (uint32_t r = 0; r < P; ++r) p->outputs.push_back({BUF_RECV, r, c});
    expect_copy(p, BUF_SEND, [root](uint32_t, int64_t) { return (uint32_t)root; },
                [c](uint32_t r, int64_t e) { return (int64_t)r * c + e; });
    return true;
}

// bine_gather_mpi (lib/libbine_coll_gather.cc:202-272).  Every rank gathers in
// TMP (the reference's recvbuf at the root, a scratch copy elsewhere).
bool gather_tree(uint32_t P, int64_t c, int root, Program* p, std::string* err) {
    if (root % 2 != 0) {
        *err = "gather supports even roots: the reference grows its window by rank parity while rooting the "
               "tree at rank - root (lib/libbine_coll_gather.cc:221-224)";
        return false;
    }
    if (!is_pow2(P)) {
        *err = "gather needs a power-of-two number of ranks (the reference schedule deadlocks otherwise)";
        return false;
    }
    const int64_t E = (int64_t)P * c;
    p->extent[BUF_SEND] = c;
    p->extent[BUF_TMP] = E;
    p->extent[BUF_RECV] = E;
    p->single[BUF_RECV] = true;
    p->root = (uint32_t)root;
    PB b{p, BINE_OP_SUM};
    b.step();
    std::vector<int64_t> minr(P), maxr(P);
    std::vector<int> ext(P);
    std::vector<char> active(P, 1);
    for (uint32_t r = 0; r < P; ++r) {
        b.copy(at(BUF_TMP, r, (int64_t)r * c), at(BUF_SEND, r, 0), c);
        minr[r] = maxr[r] = r;
        ext[r] = (r % 2) ? -1 : 1;
    }
    Matcher m;
    for (int mask = 1; mask < (int)P; mask <<= 1) {
        b.step();
        for (uint32_t r = 0; r < P; ++r) {
            if (!active[r]) continue;
            int vr = pmod((int)r - root, (int)P);
            uint32_t vnb = to_negabinary(vr);
            int partner = pmod(from_negabinary(vnb ^ (uint32_t)((mask << 1) - 1)) + root, (int)P);
            uint32_t ml = (uint32_t)((mask << 2) - 1), lsbs = vnb & ml;
            bool eq = lsbs == 0 || lsbs == ml;
            if (!eq || ((mask << 1) >= (int)P && (int)r != root)) {
                if (maxr[r] >= minr[r]) {
                    m.send(r, (uint32_t)partner, at(BUF_TMP, r, minr[r] * c), c * (maxr[r] - minr[r] + 1));
                } else {
                    m.send(r, (uint32_t)partner, at(BUF_TMP, r, minr[r] * c), c * ((int64_t)P - minr[r]));
                    m.send(r, (uint32_t)partner, at(BUF_TMP, r, 0), c * (maxr[r] + 1));
                }
                active[r] = 0;
            } else {
                int64_t rs, re;
                if (ext[r] == 1) {
                    rs = pmod((int)maxr[r] + 1, (int)P);
                    re = pmod((int)maxr[r] + mask, (int)P);
                    maxr[r] = re;
                } else {
                    re = pmod((int)minr[r] - 1, (int)P);
                    rs = pmod((int)minr[r] - mask, (int)P);
                    minr[r] = rs;
                }
                if (re >= rs) {
                    m.recv(r, (uint32_t)partner, at(BUF_TMP, r, rs * c), c * (re - rs + 1));
                } else {
                    m.recv(r, (uint32_t)partner, at(BUF_TMP, r, rs * c), c * ((int64_t)P - rs));
                    m.recv(r, (uint32_t)partner, at(BUF_TMP, r, 0), c * (re + 1));
                }
                ext[r] = -ext[r];
            }
        }
        if (!m.emit(b, err)) return false;
    }
    b.step();
    b.copy(at(BUF_RECV, (uint32_t)root, 0), at(BUF_TMP, (uint32_t)root, 0), E);
    p->outputs.push_back({BUF_RECV, (uint32_t)root, E});
    expect_copy(p, BUF_SEND, [c](uint32_t, int64_t e) { return (uint32_t)(e / c); },
                [c](uint32_t, int64_t e) { return e % c; });
    return true;
}

// MPI_IN_PLACE: SEND and RECV are one buffer.  Every SEND location becomes the
// same RECV location and copies that turn into identities are dropped (the
// initial SEND -> RECV copy of B).  The symbolic race check then runs on the
// aliased program, so a schedule that reads a SEND partition after another
// operation of the step wrote that RECV partition is rejected.
void alias_send_onto_recv(Program* p) {
    auto fix = [](Loc& l) {
        if (l.buf == BUF_SEND) l.buf = BUF_RECV;
    };
    for (auto& st : p->steps) {
        std::vector<Op> kept;
        kept.reserve(st.size());
        for (Op o : st) {
            fix(o.dst);
            fix(o.a);
            fix(o.b);
            if (o.kind == OP_COPY && o.dst.buf == o.a.buf && o.dst.rank == o.a.rank && o.dst.off == o.a.off) continue;
            kept.push_back(o);
        }
        st.swap(kept);
    }
    p->extent[BUF_SEND] = 0;
    if (p->expect.src_buf == BUF_SEND) p->expect.src_buf = BUF_RECV;
}

}  // namespace

bool build_program(const Torus& t, const bine_vcomm_config_t& cfg, const CollRequest& req, Program* p,
                   std::string* err) {
    *p = Program();
    p->P = t.size;
    const int ports = (int)cfg.num_ports;
    const uint32_t P = t.size;
    if (req.count < 0) {
        *err = "negative count";
        return false;
    }
    if (req.in_place && req.coll != BINE_COLL_ALLREDUCE) {
        *err = "only allreduce runs in place (sendbuf == recvbuf)";
        return false;
    }
    if (req.op != BINE_OP_SUM && req.op != BINE_OP_VOID) {
        *err = "unknown op";
        return false;
    }
    auto fam_ok = [&](int f) { return f == FAM_BINE || f == FAM_RECDOUB; };
    // Family DEFAULT: the reference hands the call to the vendor MPI (PMPI_*, lib/libbine.cc:1016, 1069,
    // 1174, 1257, 1295, 1317, 1381, 1445).  The stand-in here is the recursive-doubling family for the
    // torus collectives on power-of-two tori (recursive halving then doubling: the classic vendor
    // allreduce), Bruck for the alltoall, and the Bine trees for the rooted collectives, whose MPI-layer
    // code the reference runs for BINE and RECDOUB alike.
    bool all_p2 = true;
    for (uint32_t d = 0; d < t.nd; ++d) all_p2 = all_p2 && is_pow2(t.dims[d]);
    auto resolve = [&](bine_coll_config_t c, int coll) {
        if (c.family != FAM_DEFAULT) return c;
        if (coll == BINE_COLL_ALLTOALL) c.family = FAM_BRUCK;
        else if (coll == BINE_COLL_ALLREDUCE || coll == BINE_COLL_REDUCE_SCATTER || coll == BINE_COLL_ALLGATHER)
            c.family = all_p2 ? FAM_RECDOUB : FAM_BINE;
        else c.family = FAM_BINE;
        return c;
    };
    switch (req.coll) {
    case BINE_COLL_ALLREDUCE: {
        const bine_coll_config_t c = resolve(cfg.allreduce, req.coll);
        if (!fam_ok(c.family)) { *err = "allreduce family must be BINE or RECDOUB"; return false; }
        bool ok;
        switch (c.algo) {
        case BINE_ALLREDUCE_L:
            ok = allreduce_l(t, ports, c.family, req.count, req.op, req.in_place, p, err);
            break;
        case BINE_ALLREDUCE_B:
        case BINE_ALLREDUCE_B_COALESCE:
        case BINE_ALLREDUCE_B_CONT:
            if (c.distance != BINE_DISTANCE_INCREASING) { *err = "allreduce B needs INCREASING distance (lib/libbine.cc:1035)"; return false; }
            if (req.count < (int64_t)ports * P) { *err = "count < num_ports * P (lib/libbine.cc:1036)"; return false; }
            ok = c.algo == BINE_ALLREDUCE_B_CONT ? allreduce_b_cont(t, ports, c.family, req.count, req.op, p, err)
                                                 : allreduce_b(t, ports, c.family, req.count, req.op, p, err);
            break;
        default:
            *err = "allreduce algorithm not available in the reference MPI layer";
            return false;
        }
        if (ok && req.in_place) alias_send_onto_recv(p);
        return ok;
    }
    case BINE_COLL_REDUCE_SCATTER: {
        const bine_coll_config_t c = resolve(cfg.reduce_scatter, req.coll);
        if (!fam_ok(c.family)) { *err = "reduce_scatter family must be BINE or RECDOUB"; return false; }
        if (c.algo != 0) { *err = "reduce_scatter algorithm must be VEC_HALVING_CONT_PERMUTE"; return false; }
        if (req.count * (int64_t)P < (int64_t)ports * P) { *err = "count < num_ports * P"; return false; }
        return reduce_scatter_cont(t, ports, c.family, c.distance, req.count, req.op, p, err);
    }
    case BINE_COLL_ALLGATHER: {
        const bine_coll_config_t c = resolve(cfg.allgather, req.coll);
        if (!fam_ok(c.family)) { *err = "allgather family must be BINE or RECDOUB"; return false; }
        if (req.count < 1) { *err = "allgather needs count >= 1"; return false; }
        if (c.algo == BINE_ALLGATHER_CONT_PERMUTE) {
            if (req.count < ports) { *err = "recvcount < num_ports"; return false; }
            return allgather_cont(t, ports, c.family, c.distance, req.count, p, err);
        }
        if (c.algo == BINE_ALLGATHER_CONT_SEND) return allgather_send(t, ports, c.family, c.distance, req.count, p, err);
        if (c.algo == BINE_ALLGATHER_BLOCKS) {
            if (t.nd != 1 || c.family != FAM_BINE) { *err = "BLOCKS allgather is 1D Bine only"; return false; }
            return allgather_blocks(P, req.count, p, err);
        }
        *err = "allgather algorithm not available in the reference MPI layer";
        return false;
    }
    case BINE_COLL_ALLTOALL: {
        const bine_coll_config_t c = resolve(cfg.alltoall, req.coll);
        if (req.count < 1) { *err = "alltoall needs count >= 1"; return false; }
        if (c.family == FAM_BRUCK) return alltoall_bruck(P, req.count, p, err);
        if (c.family == FAM_BINE) return alltoall_bine(P, req.count, p, err);
        *err = "alltoall family must be BINE or BRUCK";
        return false;
    }
    case BINE_COLL_BCAST: {
        const bine_coll_config_t c = resolve(cfg.bcast, req.coll);
        if (!fam_ok(c.family)) { *err = "bcast family must be BINE or RECDOUB"; return false; }
        if (req.root < 0 || req.root >= (int)P) { *err = "root out of range"; return false; }
        if (c.algo == BINE_BCAST_BINOMIAL_TREE) return bcast_tree(P, req.count, req.root, p, err);
        if (c.algo == BINE_BCAST_SCATTER_ALLGATHER) {
            if (c.distance != BINE_DISTANCE_INCREASING) { *err = "scatter-allgather bcast needs INCREASING distance"; return false; }
            return bcast_sa(P, req.count, req.root, p, err);
        }
        *err = "bcast algorithm not available in the reference MPI layer";
        return false;
    }
    case BINE_COLL_REDUCE: {
        const bine_coll_config_t c = resolve(cfg.reduce, req.coll);
        if (!fam_ok(c.family)) { *err = "reduce family must be BINE or RECDOUB"; return false; }
        if (req.root < 0 || req.root >= (int)P) { *err = "root out of range"; return false; }
        if (c.algo == BINE_REDUCE_BINOMIAL_TREE) return reduce_tree(P, req.count, req.root, req.op, p, err);
        if (c.algo == BINE_REDUCE_REDUCE_SCATTER_GATHER) return reduce_rsg(P, req.count, req.root, req.op, p, err);
        *err = "reduce algorithm not available";
        return false;
    }
    case BINE_COLL_SCATTER: {
        const bine_coll_config_t c = resolve(cfg.scatter, req.coll);
        if (!fam_ok(c.family)) { *err = "scatter family must be BINE or RECDOUB"; return false; }
        if (req.root < 0 || req.root >= (int)P) { *err = "root out of range"; return false; }
        if (req.count < ports) { *err = "recvcount < num_ports (lib/libbine.cc:1329)"; return false; }
        return scatter_tree(P, req.count, req.root, p, err);
    }
    case BINE_COLL_GATHER: {
        const bine_coll_config_t c = resolve(cfg.gather, req.coll);
        if (!fam_ok(c.family)) { *err = "gather family must be BINE or RECDOUB"; return false; }
        if (req.root < 0 || req.root >= (int)P) { *err = "root out of range"; return false; }
        if (req.count < ports) { *err = "recvcount < num_ports (lib/libbine.cc:1393)"; return false; }
        return gather_tree(P, req.count, req.root, p, err);
    }
    default:
        *err = "unknown collective";
        return false;
    }
}

// ===========================================================================
// Symbolic execution
// ===========================================================================

namespace {

struct SymVal {
    int32_t node;
    int64_t delta;  // value at position x is node(x + delta)
};

struct Node {
    uint8_t kind;  // 0 leaf, 1 add
    uint8_t buf;
    uint32_t rank;
    SymVal a, b;
    uint32_t nleaves;
    uint16_t height;
    bool perfect;
};

struct Piece {
    int64_t end;
    SymVal v;
};

struct Seg {
    int64_t lo, hi;  // relative to the op start
    SymVal v;        // relative: value at relative x is node(x + delta)
};

inline uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

class Sym {
public:
    explicit Sym(const Program& p) : prog(p), P(p.P) {
        maps.resize((size_t)kNumBufs * P);
        for (int b = 0; b < kNumBufs; ++b) {
            if (p.extent[b] <= 0) continue;
            for (uint32_t r = 0; r < P; ++r) {
                if (p.single[b] && r != p.root) continue;
                Node n{0, (uint8_t)b, r, {-1, 0}, {-1, 0}, 1, 0, true};
                nodes.push_back(n);
                maps[(size_t)b * P + r][0] = Piece{p.extent[b], SymVal{(int32_t)nodes.size() - 1, 0}};
            }
        }
    }

    bool valid(const Loc& l, int64_t n) const {
        if (l.buf >= kNumBufs || l.rank >= P || n < 0) return false;
        if (prog.extent[l.buf] <= 0) return false;
        if (prog.single[l.buf] && l.rank != prog.root) return false;
        return l.off >= 0 && l.off + n <= prog.extent[l.buf];
    }

    void read(const Loc& l, int64_t n, std::vector<Seg>* out) const {
        out->clear();
        const auto& m = maps[(size_t)l.buf * P + l.rank];
        auto it = m.upper_bound(l.off);
        --it;
        const int64_t end = l.off + n;
        for (; it != m.end() && it->first < end; ++it) {
            int64_t lo = std::max(it->first, l.off), hi = std::min(it->second.end, end);
            if (hi <= lo) continue;
            out->push_back(Seg{lo - l.off, hi - l.off, SymVal{it->second.v.node, it->second.v.delta + l.off}});
        }
    }

    int32_t add_node(SymVal a, SymVal b) {
        const Node& na = nodes[a.node];
        const Node& nb = nodes[b.node];
        Node n;
        n.kind = 1;
        n.buf = 0;
        n.rank = 0;
        n.a = a;
        n.b = b;
        n.nleaves = na.nleaves + nb.nleaves;
        n.height = (uint16_t)(std::max(na.height, nb.height) + 1);
        n.perfect = na.perfect && nb.perfect && na.height == nb.height;
        nodes.push_back(n);
        return (int32_t)nodes.size() - 1;
    }

    // Replace [dst.off, dst.off + n) of dst with the relative segments.
    void write(const Loc& d, int64_t n, const std::vector<Seg>& segs) {
        auto& m = maps[(size_t)d.buf * P + d.rank];
        const int64_t lo = d.off, hi = d.off + n;
        split(m, lo);
        split(m, hi);
        m.erase(m.lower_bound(lo), m.lower_bound(hi));
        for (const Seg& s : segs) m[lo + s.lo] = Piece{lo + s.hi, SymVal{s.v.node, s.v.delta - lo}};
    }

    const Program& prog;
    uint32_t P;
    std::vector<Node> nodes;
    std::vector<std::map<int64_t, Piece>> maps;

private:
    static void split(std::map<int64_t, Piece>& m, int64_t x) {
        auto it = m.upper_bound(x);
        if (it == m.begin()) return;
        --it;
        if (it->first < x && it->second.end > x) {
            Piece tail{it->second.end, it->second.v};
            it->second.end = x;
            m[x] = tail;
        }
    }
};

struct Region {
    int64_t lo, hi;
    size_t op;
    bool write;
    int64_t off_key;  // op-relative origin (dst.off for writes, src.off for reads)
};

// Race check of one step: no two writes overlap, and a read overlapping a write
// must be the same op's in-place operand (same location, same offset).
bool step_is_race_free(const std::vector<Op>& ops, uint32_t P) {
    std::unordered_map<uint64_t, std::vector<Region>> by;
    auto key = [P](uint8_t buf, uint32_t rank) { return (uint64_t)buf * P + rank; };
    for (size_t i = 0; i < ops.size(); ++i) {
        const Op& o = ops[i];
        by[key(o.dst.buf, o.dst.rank)].push_back({o.dst.off, o.dst.off + o.n, i, true, o.dst.off});
        by[key(o.a.buf, o.a.rank)].push_back({o.a.off, o.a.off + o.n, i, false, o.a.off});
        if (o.kind == OP_ADD) by[key(o.b.buf, o.b.rank)].push_back({o.b.off, o.b.off + o.n, i, false, o.b.off});
    }
    for (auto& kv : by) {
        auto& v = kv.second;
        std::vector<Region> w;
        for (auto& r : v)
            if (r.write) w.push_back(r);
        if (w.empty()) continue;
        std::sort(w.begin(), w.end(), [](const Region& x, const Region& y) { return x.lo < y.lo; });
        for (size_t i = 1; i < w.size(); ++i)
            if (w[i].lo < w[i - 1].hi) return false;
        for (auto& r : v) {
            if (r.write) continue;
            auto it = std::upper_bound(w.begin(), w.end(), r.lo, [](int64_t x, const Region& y) { return x < y.lo; });
            if (it != w.begin()) --it;
            for (; it != w.end() && it->lo < r.hi; ++it) {
                if (it->hi <= r.lo) continue;
                if (it->op != r.op || it->off_key != r.off_key) return false;
            }
        }
    }
    return true;
}

// Tree flattening, hashing and leaf order.
struct Flat {
    uint8_t buf;
    uint32_t rank;
    int64_t pos;
};

void flatten(const Sym& s, int32_t node, int64_t pos, std::vector<Flat>* out) {
    const Node& n = s.nodes[node];
    if (n.kind == 0) {
        out->push_back({n.buf, n.rank, pos});
        return;
    }
    flatten(s, n.a.node, pos + n.a.delta, out);
    flatten(s, n.b.node, pos + n.b.delta, out);
}

struct PairHash {
    size_t operator()(const std::pair<int32_t, int64_t>& p) const { return mix64((uint64_t)p.first * 1000003ull ^ (uint64_t)p.second); }
};

uint64_t canon_hash(const Sym& s, int32_t node, int64_t pos, std::unordered_map<std::pair<int32_t, int64_t>, uint64_t, PairHash>* memo) {
    auto k = std::make_pair(node, pos);
    auto it = memo->find(k);
    if (it != memo->end()) return it->second;
    const Node& n = s.nodes[node];
    uint64_t h;
    if (n.kind == 0) {
        h = mix64(((uint64_t)n.buf << 56) ^ ((uint64_t)n.rank << 36) ^ (uint64_t)pos);
    } else {
        uint64_t ha = canon_hash(s, n.a.node, pos + n.a.delta, memo);
        uint64_t hb = canon_hash(s, n.b.node, pos + n.b.delta, memo);
        if (ha > hb) std::swap(ha, hb);
        h = mix64(ha * 31 + mix64(hb ^ 0x5bd1e995ull));
    }
    (*memo)[k] = h;
    return h;
}

}  // namespace

bool compile_program(const Program& prog, bool want_fused, CompiledPlan* out, std::string* err) {
    *out = CompiledPlan();
    out->P = prog.P;
    for (int b = 0; b < kNumBufs; ++b) {
        out->extent[b] = prog.extent[b];
        out->single[b] = prog.single[b];
    }
    Sym sym(prog);
    std::vector<Seg> sa, sb, merged;
    std::vector<std::vector<Seg>> pending;
    std::vector<std::vector<Op>> steps;
    // Steps that touch each other's data may share a launch when the union is
    // race free and the later step reads nothing the earlier one writes.
    for (const auto& st : prog.steps) {
        if (st.empty()) continue;
        for (const Op& o : st) {
            if (!sym.valid(o.dst, o.n) || !sym.valid(o.a, o.n) || (o.kind == OP_ADD && !sym.valid(o.b, o.n))) {
                *err = "schedule operation out of bounds";
                return false;
            }
        }
        if (!step_is_race_free(st, prog.P)) {
            *err = "schedule step reads data another operation of the same step writes";
            return false;
        }
        steps.push_back(st);
    }
    // merge consecutive independent steps
    std::vector<std::vector<Op>> merged_steps;
    for (auto& st : steps) {
        if (!merged_steps.empty()) {
            std::vector<Op> cand = merged_steps.back();
            cand.insert(cand.end(), st.begin(), st.end());
            // RAW across the boundary is also caught: a read of the later step that
            // overlaps a write of the earlier one is a race in the merged launch.
            if (merged_steps.back().size() + st.size() < 200000 && step_is_race_free(cand, prog.P)) {
                merged_steps.back() = std::move(cand);
                continue;
            }
        }
        merged_steps.push_back(st);
    }
    for (const auto& st : merged_steps) {
        pending.assign(st.size(), {});
        for (size_t i = 0; i < st.size(); ++i) {
            const Op& o = st[i];
            sym.read(o.a, o.n, &sa);
            if (o.kind == OP_COPY) {
                pending[i] = sa;
                continue;
            }
            sym.read(o.b, o.n, &sb);
            merged.clear();
            size_t ia = 0, ib = 0;
            int64_t x = 0;
            while (x < o.n) {
                while (sa[ia].hi <= x) ++ia;
                while (sb[ib].hi <= x) ++ib;
                int64_t hi = std::min(sa[ia].hi, sb[ib].hi);
                int32_t nn = sym.add_node(sa[ia].v, sb[ib].v);
                merged.push_back(Seg{x, hi, SymVal{nn, 0}});
                x = hi;
            }
            pending[i] = merged;
        }
        for (size_t i = 0; i < st.size(); ++i) sym.write(st[i].dst, st[i].n, pending[i]);
        // per-step groups: one group per op
        std::vector<Group> gs;
        gs.reserve(st.size());
        for (const Op& o : st) {
            Group g;
            g.n = o.n;
            g.leaves.push_back(make_entry(o.a.buf, o.a.rank, (uint32_t)o.a.off));
            if (o.kind == OP_ADD) {
                g.leaves.push_back(make_entry(o.b.buf, o.b.rank, (uint32_t)o.b.off));
                g.lg = 1;
            }
            g.outs.push_back(make_entry(o.dst.buf, o.dst.rank, (uint32_t)o.dst.off));
            gs.push_back(std::move(g));
        }
        out->step_groups.push_back(std::move(gs));
    }

    // Final state: verify MPI semantics and group for the fused execution.
    std::unordered_map<std::pair<int32_t, int64_t>, uint64_t, PairHash> memo;
    std::unordered_map<uint64_t, size_t> group_of;  // (hash, len) -> group index
    std::unordered_map<uint64_t, char> verified;  // (node, pos, expected source) already checked
    bool fused_ok = want_fused;
    std::string why;
    std::vector<Flat> fl;
    std::vector<char> seen(prog.P);
    for (const auto& o : prog.outputs) {
        const auto& m = sym.maps[(size_t)o.buf * prog.P + o.rank];
        for (auto it = m.begin(); it != m.end() && it->first < o.n; ++it) {
            const int64_t lo = it->first, hi = std::min(it->second.end, o.n);
            const SymVal v = it->second.v;
            const Node& nd = sym.nodes[v.node];
            const int64_t pos = lo + v.delta;
            uint64_t vkey = 0;
            if (prog.expect.kind != Expect::NONE) {
                const Expect& ex = prog.expect;
                vkey = mix64(((uint64_t)v.node << 20) ^ mix64((uint64_t)pos) ^
                             mix64((uint64_t)ex.src_off(o.rank, lo) * 7919u) ^
                             (ex.kind == Expect::COPY_FROM ? mix64(ex.src_rank(o.rank, lo) + 17u) : 0));
            }
            if (prog.expect.kind != Expect::NONE && !verified.count(vkey)) {
                const Expect& ex = prog.expect;
                verified[vkey] = 1;
                // the expected input offset must advance one-for-one over the piece
                if (ex.src_off(o.rank, hi - 1) - ex.src_off(o.rank, lo) != hi - 1 - lo ||
                    (ex.kind == Expect::COPY_FROM && ex.src_rank(o.rank, lo) != ex.src_rank(o.rank, hi - 1))) {
                    // piece spans two expected source blocks: it cannot be right
                    *err = "schedule result differs from the collective's MPI semantics";
                    return false;
                }
                fl.clear();
                flatten(sym, v.node, pos, &fl);
                if (ex.kind == Expect::SUM_ALL) {
                    bool ok = fl.size() == prog.P;
                    std::fill(seen.begin(), seen.end(), 0);
                    for (const Flat& f : fl) {
                        if (!ok) break;
                        ok = f.buf == ex.src_buf && f.rank < prog.P && !seen[f.rank] && f.pos == ex.src_off(o.rank, lo);
                        if (ok) seen[f.rank] = 1;
                    }
                    if (!ok) {
                        *err = "schedule result differs from the collective's MPI semantics";
                        return false;
                    }
                } else {
                    if (fl.size() != 1 || fl[0].buf != ex.src_buf || fl[0].rank != ex.src_rank(o.rank, lo) ||
                        fl[0].pos != ex.src_off(o.rank, lo)) {
                        *err = "schedule result differs from the collective's MPI semantics";
                        return false;
                    }
                }
            }
            if (!fused_ok) continue;
            // identity output (an input left in place): nothing to write
            if (nd.kind == 0 && nd.buf == o.buf && nd.rank == o.rank && pos == lo) continue;
            if (!nd.perfect) {
                fused_ok = false;
                why = "a reduction tree is not perfect (non power-of-two schedule)";
                continue;
            }
            uint64_t h = canon_hash(sym, v.node, pos, &memo);
            uint64_t gk = mix64(h ^ mix64((uint64_t)(hi - lo)));
            auto git = group_of.find(gk);
            if (git == group_of.end()) {
                Group g;
                g.n = hi - lo;
                fl.clear();
                flatten(sym, v.node, pos, &fl);
                g.lg = (uint16_t)nd.height;
                for (const Flat& f : fl) g.leaves.push_back(make_entry(f.buf, f.rank, (uint32_t)f.pos));
                g.outs.push_back(make_entry(o.buf, o.rank, (uint32_t)lo));
                group_of[gk] = out->fused.size();
                out->fused.push_back(std::move(g));
            } else {
                out->fused[git->second].outs.push_back(make_entry(o.buf, o.rank, (uint32_t)lo));
            }
        }
    }
    // The fused launch reads the initial inputs while it writes the outputs:
    // no output may overlap a leaf's input range.
    if (fused_ok) {
        std::unordered_map<uint64_t, std::vector<std::pair<int64_t, int64_t>>> in_ranges;
        for (const Group& g : out->fused)
            for (const PoolEntry& l : g.leaves) in_ranges[l.rb].push_back({l.off, l.off + g.n});
        for (auto& kv : in_ranges) {
            auto& v = kv.second;
            std::sort(v.begin(), v.end());
            std::vector<std::pair<int64_t, int64_t>> u;
            for (auto& r : v) {
                if (!u.empty() && r.first <= u.back().second) u.back().second = std::max(u.back().second, r.second);
                else u.push_back(r);
            }
            v = u;
        }
        for (const Group& g : out->fused) {
            for (const PoolEntry& o : g.outs) {
                auto it = in_ranges.find(o.rb);
                if (it == in_ranges.end()) continue;
                for (auto& r : it->second)
                    if ((int64_t)o.off < r.second && (int64_t)o.off + g.n > r.first) {
                        fused_ok = false;
                        why = "an output overlaps an input of the fused launch";
                    }
            }
        }
    }
    if (want_fused && prog.outputs.empty()) fused_ok = false;
    out->fused_ok = fused_ok;
    out->fused_why = why;
    if (!fused_ok) out->fused.clear();
    return true;
}

// ===========================================================================
// Lowering
// ===========================================================================

void lower_groups(const std::vector<std::vector<Group>>& launches, int64_t shallow_tile, int vec, uint32_t bucket,
                  size_t dtsize, Lowered* out, int64_t deep_tile_elems, int deep_lg,
                  const std::vector<uint32_t>* group_bucket, const std::vector<uint8_t>* bucket_vec, int two_phase,
                  int reduction_tiles) {
    auto lower_launch = [&](const std::vector<Group>& gs, const std::vector<uint32_t>* gbk, bool copies_by_tile,
                            int64_t tiles_per_item) {
        const uint32_t first = (uint32_t)out->items.size();
        // Open run of aligned single copies: consecutive groups whose leaf and
        // output advance by one fixed (rank, offset) step become one ITEM_RUN item.
        int64_t run_item = -1;
        uint32_t run_len = 0;
        std::vector<uint32_t> tile_no;  // per emitted item: its tile's index within its group
        PoolEntry run_leaf{}, run_out{}, step_leaf{}, step_out{};
        auto vec_bucket = [&](uint32_t b) { return !bucket_vec || (*bucket_vec)[b]; };
        // 16-byte vector accesses need dword-aligned addresses only: loads run at full rate unaligned and
        // the kernels realign the stores, so leaves and outputs need not share a residue modulo the vector.
        auto word_ok = [&](const PoolEntry& e) { return ((uint64_t)e.off * dtsize) % 4 == 0; };
        auto single_copy = [&](const Group& g, uint32_t b) {
            return g.lg == 0 && g.leaves.size() == 1 && g.outs.size() == 1 && g.n <= shallow_tile && g.n % vec == 0 &&
                   word_ok(g.leaves[0]) && word_ok(g.outs[0]) && vec_bucket(b);
        };
        // step b - a with non-negative rank and offset parts in the same buffer
        auto step_of = [](const PoolEntry& a, const PoolEntry& b, PoolEntry* st) {
            if ((a.rb >> 28) != (b.rb >> 28) || (b.rb & 0x0FFFFFFFu) < (a.rb & 0x0FFFFFFFu) || b.off < a.off) return false;
            *st = PoolEntry{(b.rb & 0x0FFFFFFFu) - (a.rb & 0x0FFFFFFFu), b.off - a.off};
            return true;
        };
        auto advance = [](const PoolEntry& a, const PoolEntry& st, uint32_t k) {
            return PoolEntry{a.rb + st.rb * k, a.off + st.off * k};
        };
        for (size_t gi = 0; gi < gs.size(); ++gi) {
            const Group& g = gs[gi];
            const uint32_t gb = gbk ? (*gbk)[gi] : bucket;
            if (g.n <= 0) continue;
            if (single_copy(g, gb) && run_item >= 0 && run_len < kMaxRun && out->items[run_item].n == (uint32_t)g.n &&
                out->items[run_item].bucket == gb) {
                bool ok;
                if (run_len == 1) {
                    PoolEntry sl, so;
                    ok = step_of(run_leaf, g.leaves[0], &sl) && step_of(run_out, g.outs[0], &so) &&
                         (sl.rb | sl.off | so.rb | so.off) != 0;
                    if (ok) {
                        step_leaf = sl;
                        step_out = so;
                    }
                } else {
                    const PoolEntry l = advance(run_leaf, step_leaf, run_len), o = advance(run_out, step_out, run_len);
                    ok = l.rb == g.leaves[0].rb && l.off == g.leaves[0].off && o.rb == g.outs[0].rb && o.off == g.outs[0].off;
                }
                if (ok) {
                    ++run_len;
                    GredItem& it = out->items[run_item];
                    if (run_len == 2) {  // the head's two pool entries are the last ones: rewrite them as a run
                        out->pool.resize(it.leaf_off);
                        out->pool.push_back(run_leaf);
                        out->pool.push_back(step_leaf);
                        out->pool.push_back(run_out);
                        out->pool.push_back(step_out);
                        it.out_off = it.leaf_off + 2;
                        it.flags |= ITEM_RUN;
                    }
                    it.pad = run_len;
                    out->bytes_read += (uint64_t)g.n * dtsize;
                    out->bytes_written += (uint64_t)g.n * dtsize;
                    continue;
                }
            }
            run_item = -1;
            run_len = 0;
            const int64_t tile_elems =
                ((g.lg >= deep_lg && deep_tile_elems > 0) ? deep_tile_elems : shallow_tile) * tiles_per_item;
            const uint32_t leaf_off = (uint32_t)out->pool.size();
            out->pool.insert(out->pool.end(), g.leaves.begin(), g.leaves.end());
            const uint32_t out_off = (uint32_t)out->pool.size();
            out->pool.insert(out->pool.end(), g.outs.begin(), g.outs.end());
            // vectors need every address dword-aligned; the body starts where the first output's offset is
            // a whole number of vectors (outputs sharing its residue then store without realignment)
            const int64_t res = (g.outs.empty() ? g.leaves[0].off : g.outs[0].off) % vec;
            bool same = vec_bucket(gb);
            for (const PoolEntry& e : g.leaves) same &= word_ok(e);
            for (const PoolEntry& e : g.outs) same &= word_ok(e);
            uint32_t tiles = 0;
            auto emit = [&](int64_t d, int64_t n, bool v) {
                tile_no.push_back(tiles++);
                GredItem it{};
                it.leaf_off = leaf_off;
                it.out_off = out_off;
                it.lg = g.lg;
                it.flags = v ? ITEM_VEC : 0;
                it.nouts = (uint32_t)g.outs.size();
                it.delta = (uint32_t)d;
                it.n = (uint32_t)n;
                it.bucket = gb;
                out->items.push_back(it);
            };
            // two consecutive whole tiles of a long single copy as one copy run: one pass of the copy-run
            // kernel with two copies in flight, like the runs of short copies
            auto emit_tile_pair = [&](int64_t d0, int64_t n) {
                tile_no.push_back(tiles++);
                GredItem it{};
                it.leaf_off = (uint32_t)out->pool.size();
                out->pool.push_back(PoolEntry{g.leaves[0].rb, (uint32_t)(g.leaves[0].off + d0)});
                out->pool.push_back(PoolEntry{0, (uint32_t)n});  // step: same rank, next tile
                it.out_off = (uint32_t)out->pool.size();
                out->pool.push_back(PoolEntry{g.outs[0].rb, (uint32_t)(g.outs[0].off + d0)});
                out->pool.push_back(PoolEntry{0, (uint32_t)n});
                it.lg = 0;
                it.flags = ITEM_VEC | ITEM_RUN;
                it.nouts = 1;
                it.n = (uint32_t)n;
                it.bucket = gb;
                it.pad = 2;
                out->items.push_back(it);
            };
            const bool long_copy = g.lg == 0 && g.leaves.size() == 1 && g.outs.size() == 1 && kMaxRun >= 2;
            int64_t d = 0;
            if (same) {
                int64_t head = (vec - res) % vec;
                if (head > g.n) head = g.n;
                if (head) emit(0, head, false);
                d = head;
                int64_t body = ((g.n - d) / vec) * vec;
                int64_t x = 0;
                if (long_copy)
                    for (; x + 2 * tile_elems <= body; x += 2 * tile_elems) emit_tile_pair(d + x, tile_elems);
                for (; x < body; x += tile_elems) emit(d + x, std::min(tile_elems, body - x), true);
                d += body;
            }
            for (; d < g.n; d += tile_elems) emit(d, std::min(tile_elems, g.n - d), false);
            if (single_copy(g, gb)) {  // exactly one aligned item: a run may start here
                run_item = (int64_t)out->items.size() - 1;
                run_len = 1;
                run_leaf = g.leaves[0];
                run_out = g.outs[0];
            }
            const uint64_t nl = g.leaves.size();
            out->bytes_read += (uint64_t)g.n * nl * dtsize;
            out->bytes_written += (uint64_t)g.n * g.outs.size() * dtsize;
            out->max_leaves = std::max<uint64_t>(out->max_leaves, nl);
        }
        // Order: reductions first, tile-major within each bucket (tile 0 of every reducing group of the
        // bucket, then tile 1, ...; buckets in order), so that workgroups running at the same time spread
        // over many blocks of one collective; then plain copies in group order, so that neighbouring
        // workgroups stream neighbouring tiles of one block (tile-major order cost the copy collectives 7 %
        // at P = 512; group order cost the reductions 1-2 %).  Tile-major across all the buckets of a batch
        // spread concurrent workgroups over every bucket's pages: on 256 uneven fp32 buckets at P = 64 the
        // bucket-major order ran 3 % faster.  Run items go last: they are executed by their own kernel.
        const size_t cnt = out->items.size() - first;
        std::vector<uint32_t> order(cnt);
        for (size_t i = 0; i < cnt; ++i) order[i] = (uint32_t)i;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            const GredItem& ia = out->items[first + a];
            const GredItem& ib = out->items[first + b];
            const bool ra = ia.flags & ITEM_RUN, rb = ib.flags & ITEM_RUN;
            if (ra != rb) return rb;
            const bool ca = ia.lg == 0, cb = ib.lg == 0;
            if (ca != cb) return cb;
            if (ca && !copies_by_tile) return false;
            if (ia.bucket != ib.bucket) return ia.bucket < ib.bucket;
            return tile_no[a] < tile_no[b];
        });
        std::vector<GredItem> sorted(cnt);
        uint32_t nrun = 0;
        for (size_t i = 0; i < cnt; ++i) {
            sorted[i] = out->items[first + order[i]];
            nrun += (sorted[i].flags & ITEM_RUN) ? 1 : 0;
        }
        std::copy(sorted.begin(), sorted.end(), out->items.begin() + first);
        for (const GredItem& it : sorted) {
            if (!(it.flags & ITEM_VEC)) continue;
            if (it.flags & ITEM_RUN) {  // first output and the per-copy step
                out->unaligned_outs |= out->pool[it.out_off].off % vec != 0 || out->pool[it.out_off + 1].off % vec != 0;
            } else {
                for (uint32_t o = 0; o < it.nouts; ++o)
                    out->unaligned_outs |= ((int64_t)out->pool[it.out_off + o].off + it.delta) % vec != 0;
            }
        }
        out->launches.push_back({first, (uint32_t)cnt});
        out->runs.push_back(nrun);
    };
    for (size_t li = 0; li < launches.size(); ++li) {
        const std::vector<Group>& gs = launches[li];
        const std::vector<uint32_t>* gb = li == 0 ? group_bucket : nullptr;
        bool split = false;
        for (const Group& g : gs) split = split || (two_phase != 0 && g.lg > 0 && g.outs.size() > 1);
        if (!split) {
            lower_launch(gs, gb, false, 1);
            continue;
        }
        // Two phases: each reduction writes its first output only, and a second launch copies that output to
        // the others.  The first launch is all reads (the leaves), the second all writes (the copies), so the
        // memory does not alternate between the two directions.  Same bits: the copies move the sums unchanged.
        std::vector<Group> red, bc;
        std::vector<uint32_t> red_b, bc_b;
        for (size_t gi = 0; gi < gs.size(); ++gi) {
            const Group& g = gs[gi];
            const uint32_t b = gb ? (*gb)[gi] : bucket;
            red.push_back(g);
            red_b.push_back(b);
            if (g.lg > 0 && g.outs.size() > 1) {
                red.back().outs.resize(1);
                Group c;
                c.n = g.n;
                c.lg = 0;
                c.leaves.assign(1, g.outs[0]);
                c.outs.assign(g.outs.begin() + 1, g.outs.end());
                bc.push_back(std::move(c));
                bc_b.push_back(b);
            }
        }
        // reduction items of two tiles: half as many items, so a list of many tiles per block runs in fewer
        // rounds of resident workgroups (its reads then meet fewer of the reductions' writes)
        lower_launch(red, &red_b, false, reduction_tiles == 2 ? 2 : 1);
        lower_launch(bc, &bc_b, two_phase == 2, 1);
    }
}

bool has_multi_output_reduction(const std::vector<std::vector<Group>>& launches) {
    for (const auto& gs : launches)
        for (const Group& g : gs)
            if (g.lg > 0 && g.outs.size() > 1) return true;
    return false;
}

}  // namespace bine
