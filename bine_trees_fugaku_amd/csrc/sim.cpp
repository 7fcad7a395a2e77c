// sim.cpp -- link-traffic simulation: case construction and C ABI (see sim.hpp).
#include "sim.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bine_hip.h"
#include "sim_dest.hpp"

namespace bine {

hipError_t launch_sim(const void* cases, const void* classes, uint32_t ncls, const void* count_blocks,
                      uint32_t n_count_blocks, const void* scan_blocks, uint32_t n_scan_blocks, uint32_t* mult,
                      uint64_t mult_len, uint32_t* max_hops, uint32_t* max_mult, uint64_t* sum_mult, hipStream_t st);

static simd::TorusD dev_torus(const Torus& t) {
    simd::TorusD d;
    for (int i = 0; i < 3; ++i) d.dims[i] = t.dims[i];
    d.nd = t.nd;
    d.size = t.size;
    return d;
}

uint32_t sim_dest(const Torus& t, const SimClass& c, uint32_t rank) {
    return simd::dest(dev_torus(t), c.kind, c.d, c.k, rank, (uint32_t)ceil_log2(t.size));
}

bool sim_build_case(const Torus& t, int coll, int algo, SimCase* out) {
    out->t = t;
    out->coll = coll;
    out->algo = algo;
    out->classes.clear();
    const uint32_t P = t.size, L = t.steps;
    if (!is_pow2(P) || P < 2) return false;
    auto add = [&](uint32_t kind, uint32_t d, uint32_t k, uint32_t num, uint32_t den, uint32_t rep) {
        out->classes.push_back(SimClass{kind, d, k, num, den, rep});
    };
    const std::vector<StepDim> sd = step_dims(t, 0);
    const bool rs = coll == BINE_COLL_ALLREDUCE || coll == BINE_COLL_REDUCE_SCATTER;
    const bool ag = coll == BINE_COLL_ALLREDUCE || coll == BINE_COLL_ALLGATHER;
    switch (coll) {
    case BINE_COLL_ALLREDUCE:
    case BINE_COLL_REDUCE_SCATTER:
    case BINE_COLL_ALLGATHER:
        if (algo == SIM_BINE || algo == SIM_RECDOUB) {
            const uint32_t kind = algo == SIM_BINE ? SK_TORUS_BINE : SK_TORUS_RECDOUB;
            if (rs)  // reduce-scatter step s forwards P / 2^(s+1) blocks
                for (uint32_t s = 0; s < L; ++s) add(kind, sd[s].d, sd[s].k, 1, 2u << s, 1);
            if (ag)  // allgather step s forwards 2^s blocks along the step L-1-s peer
                for (uint32_t s = 0; s < L; ++s) add(kind, sd[L - 1 - s].d, sd[L - 1 - s].k, 1u << s, P, 1);
            return true;
        }
        if (algo == SIM_RING) {
            add(SK_RING, 0, 0, 1, P, (rs ? P - 1 : 0) + (ag ? P - 1 : 0));
            return true;
        }
        if (algo == SIM_SIMPY && coll != BINE_COLL_ALLGATHER) {
            for (uint32_t i = 1; i < t.nd; ++i)
                if (t.dims[i] != t.dims[0]) return false;  // simulator.py cycles dimensions step % D
            // step s takes distances[s // D] from a table of 10 (simulator.py:195), so at most 10 steps per
            // dimension; simulator.py raises IndexError beyond, and so does the case here
            if (t.steps_per_dim[0] > 10) return false;
            for (uint32_t s = 0; s < L; ++s) {
                if (coll == BINE_COLL_ALLREDUCE) add(SK_SIMPY, 0, s, 1, 1, 1);  // full vector (simulator.py:109-112)
                else add(SK_SIMPY, 0, s, 1, 2u << s, 1);                        // find_reducescatter_indexes
            }
            return true;
        }
        return false;
    case BINE_COLL_ALLTOALL:
        if (algo == SIM_BINE) {
            for (uint32_t k = 0; k < L; ++k) add(SK_LIN_BINE, 0, k, 1, 2, 1);
            return true;
        }
        if (algo == SIM_RECDOUB) {
            for (uint32_t k = 0; k < L; ++k) add(SK_LIN_XOR, 0, k, 1, 2, 1);
            return true;
        }
        if (algo == SIM_BRUCK) {
            for (uint32_t k = 0; k < L; ++k) add(SK_BRUCK, 0, k, 1, 2, 1);
            return true;
        }
        return false;
    case BINE_COLL_BCAST:
    case BINE_COLL_SCATTER:
        if (algo != SIM_BINE && algo != SIM_RECDOUB) return false;
        for (int k = (int)L - 1; k >= 0; --k)
            add(algo == SIM_BINE ? SK_TREE_BINE_DOWN : SK_TREE_BINO_DOWN, 0, (uint32_t)k,
                coll == BINE_COLL_BCAST ? 1 : (1u << k), coll == BINE_COLL_BCAST ? 1 : P, 1);
        return true;
    case BINE_COLL_REDUCE:
    case BINE_COLL_GATHER:
        if (algo != SIM_BINE && algo != SIM_RECDOUB) return false;
        for (uint32_t k = 0; k < L; ++k)
            add(algo == SIM_BINE ? SK_TREE_BINE_UP : SK_TREE_BINO_UP, 0, k, coll == BINE_COLL_REDUCE ? 1 : (1u << k),
                coll == BINE_COLL_REDUCE ? 1 : P, 1);
        return true;
    default:
        return false;
    }
}

}  // namespace bine

using namespace bine;
using simd::SimCaseRec;
using simd::SimClassRec;

namespace {
int sfail(int code, const std::string& m) { return record_error(code, m); }
}  // namespace


struct bine_sim_s {
    int device = 0;
    std::vector<SimCase> cases;
    std::vector<uint32_t> class_first;  // per case, index of its first class
    std::vector<uint64_t> mult_off;     // per flattened class
    std::vector<uint32_t> nlinks;       // per case
    SimCaseRec* d_cases = nullptr;
    SimClassRec* d_classes = nullptr;
    uint32_t* d_mult = nullptr;
    uint32_t* d_count_blocks = nullptr;  // per workgroup of the count kernel: (class, first rank)
    uint32_t* d_scan_blocks = nullptr;   // per workgroup of the ring scan: (class, first ring-direction)
    uint32_t n_count_blocks = 0, n_scan_blocks = 0;
    uint32_t* d_hops = nullptr;
    uint32_t* d_max = nullptr;
    uint64_t* d_sum = nullptr;
    uint64_t mult_len = 0;
    uint32_t ncls = 0;
    hipStream_t last = nullptr;  // stream of the last run; queries wait for it
    ~bine_sim_s() {
        for (void* p : {(void*)d_cases, (void*)d_classes, (void*)d_mult, (void*)d_hops, (void*)d_max, (void*)d_sum,
                        (void*)d_count_blocks, (void*)d_scan_blocks})
            if (p) (void)hipFree(p);
    }
};

#define SHIP(x)                                                                                     \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return sfail(BINE_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int bine_sim_classes(const bine_sim_case_t* c, uint32_t* nclasses, uint32_t* kind, uint32_t* d, uint32_t* k,
                                uint32_t* num, uint32_t* den, uint32_t* rep) {
    if (!c || !nclasses) return sfail(BINE_ERR_ARG, "NULL argument");
    Torus t;
    if (!Torus::make(c->dimensions, c->dimensions_num, &t)) return sfail(BINE_ERR_ARG, "bad torus");
    SimCase sc;
    if (!sim_build_case(t, c->coll, c->algo, &sc)) return sfail(BINE_ERR_UNSUPPORTED, "algorithm not defined for this case");
    *nclasses = (uint32_t)sc.classes.size();
    for (size_t i = 0; i < sc.classes.size(); ++i) {
        const SimClass& s = sc.classes[i];
        if (kind) kind[i] = s.kind;
        if (d) d[i] = s.d;
        if (k) k[i] = s.k;
        if (num) num[i] = s.num;
        if (den) den[i] = s.den;
        if (rep) rep[i] = s.rep;
    }
    return BINE_OK;
}

extern "C" int bine_sim_peers(const bine_sim_case_t* c, uint32_t cls, uint32_t* dest) {
    if (!c || !dest) return sfail(BINE_ERR_ARG, "NULL argument");
    Torus t;
    if (!Torus::make(c->dimensions, c->dimensions_num, &t)) return sfail(BINE_ERR_ARG, "bad torus");
    SimCase sc;
    if (!sim_build_case(t, c->coll, c->algo, &sc)) return sfail(BINE_ERR_UNSUPPORTED, "algorithm not defined for this case");
    if (cls >= sc.classes.size()) return sfail(BINE_ERR_ARG, "bad class index");
    for (uint32_t r = 0; r < t.size; ++r) dest[r] = sim_dest(t, sc.classes[cls], r);
    return BINE_OK;
}

extern "C" int bine_sim_create(const bine_sim_case_t* cases, int n, int device, bine_sim_t* out) {
    if (!cases || n <= 0 || !out) return sfail(BINE_ERR_ARG, "bad arguments");
    auto s = std::make_unique<bine_sim_s>();
    s->device = device;
    std::vector<SimCaseRec> crec;
    std::vector<SimClassRec> krec;
    std::vector<uint32_t> cblk, sblk;  // (class, first index) pairs
    uint64_t off = 0;
    for (int i = 0; i < n; ++i) {
        Torus t;
        if (!Torus::make(cases[i].dimensions, cases[i].dimensions_num, &t)) return sfail(BINE_ERR_ARG, "bad torus");
        SimCase sc;
        if (!sim_build_case(t, cases[i].coll, cases[i].algo, &sc))
            return sfail(BINE_ERR_UNSUPPORTED, "algorithm not defined for case " + std::to_string(i));
        const uint32_t nl = t.size * t.nd * 2;
        s->class_first.push_back((uint32_t)krec.size());
        s->nlinks.push_back(nl);
        crec.push_back(SimCaseRec{{{t.dims[0], t.dims[1], t.dims[2]}, t.nd, t.size}, (uint32_t)ceil_log2(t.size), nl, 0});
        uint32_t rings = 0;  // ring-directions of the torus: 2 * size / dims[d] per dimension
        for (uint32_t d = 0; d < t.nd; ++d) rings += 2 * (t.size / t.dims[d]);
        for (const SimClass& c : sc.classes) {
            const uint32_t cls = (uint32_t)krec.size();
            krec.push_back(SimClassRec{c.kind, c.d, c.k, (uint32_t)i, off});
            s->mult_off.push_back(off);
            off += nl;
            for (uint32_t r = 0; r < t.size; r += 256) {
                cblk.push_back(cls);
                cblk.push_back(r);
            }
            for (uint32_t q = 0; q < rings; q += 256) {
                sblk.push_back(cls);
                sblk.push_back(q);
            }
        }
        s->cases.push_back(std::move(sc));
    }
    s->ncls = (uint32_t)krec.size();
    s->mult_len = off;
    SHIP(hipSetDevice(device));
    SHIP(hipMalloc(&s->d_cases, sizeof(SimCaseRec) * crec.size()));
    SHIP(hipMemcpy(s->d_cases, crec.data(), sizeof(SimCaseRec) * crec.size(), hipMemcpyHostToDevice));
    SHIP(hipMalloc(&s->d_classes, sizeof(SimClassRec) * krec.size()));
    SHIP(hipMemcpy(s->d_classes, krec.data(), sizeof(SimClassRec) * krec.size(), hipMemcpyHostToDevice));
    SHIP(hipMalloc(&s->d_mult, sizeof(uint32_t) * off));
    s->n_count_blocks = (uint32_t)(cblk.size() / 2);
    s->n_scan_blocks = (uint32_t)(sblk.size() / 2);
    SHIP(hipMalloc(&s->d_count_blocks, sizeof(uint32_t) * cblk.size()));
    SHIP(hipMemcpy(s->d_count_blocks, cblk.data(), sizeof(uint32_t) * cblk.size(), hipMemcpyHostToDevice));
    SHIP(hipMalloc(&s->d_scan_blocks, sizeof(uint32_t) * sblk.size()));
    SHIP(hipMemcpy(s->d_scan_blocks, sblk.data(), sizeof(uint32_t) * sblk.size(), hipMemcpyHostToDevice));
    SHIP(hipMalloc(&s->d_hops, sizeof(uint32_t) * s->ncls));
    SHIP(hipMalloc(&s->d_max, sizeof(uint32_t) * s->ncls));
    SHIP(hipMalloc(&s->d_sum, sizeof(uint64_t) * s->ncls));
    *out = s.release();
    return BINE_OK;
}

extern "C" int bine_sim_run(bine_sim_t s, void* stream) {
    if (!s) return sfail(BINE_ERR_ARG, "NULL simulation");
    SHIP(hipSetDevice(s->device));
    SHIP(launch_sim(s->d_cases, s->d_classes, s->ncls, s->d_count_blocks, s->n_count_blocks, s->d_scan_blocks,
                    s->n_scan_blocks, s->d_mult, s->mult_len, s->d_hops, s->d_max, s->d_sum, (hipStream_t)stream));
    s->last = (hipStream_t)stream;
    return BINE_OK;
}

extern "C" int bine_sim_case_info(bine_sim_t s, int i, uint32_t* nclasses, uint32_t* nlinks) {
    if (!s || i < 0 || i >= (int)s->cases.size()) return sfail(BINE_ERR_ARG, "bad case index");
    if (nclasses) *nclasses = (uint32_t)s->cases[i].classes.size();
    if (nlinks) *nlinks = s->nlinks[i];
    return BINE_OK;
}

extern "C" int bine_sim_class_stats(bine_sim_t s, int i, uint32_t* num, uint32_t* den, uint32_t* rep,
                                    uint32_t* max_mult, uint64_t* sum_mult, uint32_t* max_hops) {
    if (!s || i < 0 || i >= (int)s->cases.size()) return sfail(BINE_ERR_ARG, "bad case index");
    SHIP(hipSetDevice(s->device));
    SHIP(hipStreamSynchronize(s->last));
    const uint32_t first = s->class_first[i];
    const size_t nc = s->cases[i].classes.size();
    for (size_t c = 0; c < nc; ++c) {
        const SimClass& k = s->cases[i].classes[c];
        if (num) num[c] = k.num;
        if (den) den[c] = k.den;
        if (rep) rep[c] = k.rep;
    }
    if (max_mult) SHIP(hipMemcpy(max_mult, s->d_max + first, nc * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (sum_mult) SHIP(hipMemcpy(sum_mult, s->d_sum + first, nc * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (max_hops) SHIP(hipMemcpy(max_hops, s->d_hops + first, nc * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return BINE_OK;
}

extern "C" int bine_sim_link_counts(bine_sim_t s, int i, uint32_t cls, uint32_t* counts) {
    if (!s || i < 0 || i >= (int)s->cases.size() || !counts) return sfail(BINE_ERR_ARG, "bad arguments");
    if (cls >= s->cases[i].classes.size()) return sfail(BINE_ERR_ARG, "bad class index");
    SHIP(hipSetDevice(s->device));
    SHIP(hipStreamSynchronize(s->last));
    const uint64_t off = s->mult_off[s->class_first[i] + cls];
    SHIP(hipMemcpy(counts, s->d_mult + off, s->nlinks[i] * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return BINE_OK;
}

extern "C" int bine_sim_destroy(bine_sim_t s) {
    if (!s) return BINE_OK;
    (void)hipSetDevice(s->device);
    delete s;
    return BINE_OK;
}
