// sim.hip -- link counting for the torus traffic simulation.
//
// Under dimension-ordered minimal routing a message crosses, per dimension,
// one contiguous run of links on one ring (the nodes that differ from its
// position only in that dimension), all in one direction.  So instead of one
// atomic add per hop, a run adds +1 at its first link and -1 just past its
// last (one more +1 at the ring's start when the run wraps round), in a
// difference array; a prefix sum along every ring and direction then turns the
// differences into the number of messages crossing each directed link.  The
// counts are the same integers a per-hop count gives (wrapping arithmetic
// modulo 2^32 cancels in the sums).
//
// Grids are sized per class: a table maps every workgroup to its class and to
// its block within the class, so small tori do not launch the largest torus's
// grid.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sim_dest.hpp"

namespace bine {

using simd::SimCaseRec;
using simd::SimClassRec;

namespace {

// Node stride of dimension d (row-major, last dimension fastest).
__device__ __forceinline__ uint32_t dim_stride(const simd::TorusD& t, uint32_t d) {
    uint32_t s = 1;
    for (uint32_t i = d + 1; i < t.nd; ++i) s *= t.dims[i];
    return s;
}

}  // namespace

__global__ void sim_zero(uint32_t* __restrict__ p, uint64_t n, uint32_t* __restrict__ hops, uint32_t* __restrict__ mx,
                         unsigned long long* __restrict__ sum, uint32_t ncls) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncls; i += (uint64_t)gridDim.x * blockDim.x) {
        hops[i] = 0;
        mx[i] = 0;
        sum[i] = 0;
    }
}

// One thread per (class, rank): the rank's message, as difference updates on the rings it crosses.
// blocks[b] = (class, first rank of the block).
__global__ void sim_count(const SimCaseRec* __restrict__ cases, const SimClassRec* __restrict__ classes,
                          const uint2* __restrict__ blocks, uint32_t* __restrict__ diff, uint32_t* __restrict__ max_hops) {
    const uint2 bk = blocks[blockIdx.x];
    const SimClassRec c = classes[bk.x];
    const SimCaseRec cs = cases[c.case_idx];
    const uint32_t r = bk.y + threadIdx.x;
    if (r >= cs.t.size) return;
    uint32_t dst = simd::dest(cs.t, c.kind, c.d, c.k, r, cs.L);
    if (dst == 0xFFFFFFFFu || dst >= cs.t.size) return;
    uint32_t src = r;
    if (c.kind == 10) {  // simulator.py ids -> torus coordinates
        int a[3], b[3];
        simd::simpy_coords(cs.t, src, a);
        simd::simpy_coords(cs.t, dst, b);
        src = simd::ident(cs.t, a);
        dst = simd::ident(cs.t, b);
    }
    if (src == dst) return;
    int s[3], d[3];
    simd::coords(cs.t, src, s);
    simd::coords(cs.t, dst, d);
    uint32_t* m = diff + c.mult_off;
    const uint32_t nd = cs.t.nd;
    uint32_t hops = 0;
    for (uint32_t dd = 0; dd < nd; ++dd) {
        const int n = (int)cs.t.dims[dd];
        const int delta = simd::pm(d[dd] - s[dd], n);
        if (!delta) continue;
        const bool neg = 2 * delta > n;
        const int h = neg ? n - delta : delta;
        // link slot of ring position j on this message's ring
        const int p = s[dd];
        s[dd] = 0;
        const uint32_t base = simd::ident(cs.t, s);
        const uint32_t stride = dim_stride(cs.t, dd);
        auto at = [&](int j) { return m + ((uint64_t)(base + (uint32_t)j * stride) * nd + dd) * 2 + (neg ? 1 : 0); };
        // the run covers ring positions [lo, lo + h) (mod n): from p upwards, or the h positions ending at p
        const int lo = neg ? p - h + 1 : p;
        if (lo >= 0 && lo + h <= n) {
            atomicAdd(at(lo), 1u);
            if (lo + h < n) atomicAdd(at(lo + h), 0xFFFFFFFFu);
        } else {  // wraps: [lo mod n, n) and [0, (lo + h) mod n)
            const int a = lo < 0 ? lo + n : lo, e = lo + h - (lo < 0 ? 0 : n);
            atomicAdd(at(a), 1u);
            atomicAdd(at(0), 1u);
            atomicAdd(at(e), 0xFFFFFFFFu);
        }
        s[dd] = d[dd];
        hops += (uint32_t)h;
    }
    atomicMax(max_hops + bk.x, hops);
}

// One thread per (class, ring, direction): prefix sum along the ring, then the class's most loaded link and
// its total link crossings.  blocks[b] = (class, first ring-direction of the block).
__global__ void sim_scan(const SimCaseRec* __restrict__ cases, const SimClassRec* __restrict__ classes,
                         const uint2* __restrict__ blocks, uint32_t* __restrict__ mult, uint32_t* __restrict__ max_mult,
                         unsigned long long* __restrict__ sum_mult) {
    const uint2 bk = blocks[blockIdx.x];
    const SimClassRec c = classes[bk.x];
    const SimCaseRec cs = cases[c.case_idx];
    const simd::TorusD& t = cs.t;
    uint32_t q = bk.y + threadIdx.x;  // ring-direction index within the class
    uint32_t mx = 0;
    unsigned long long sm = 0;
    uint32_t dd = 0;
    for (; dd < t.nd; ++dd) {  // rings of dimension dd: 2 * size / dims[dd] of them (two directions)
        const uint32_t nr = 2 * (t.size / t.dims[dd]);
        if (q < nr) break;
        q -= nr;
    }
    if (dd < t.nd) {
        const uint32_t dir = q & 1u, ring = q >> 1;
        const uint32_t stride = dim_stride(t, dd), n = t.dims[dd];
        // ring -> its position-0 node: the other coordinates enumerate the rings, dimension dd is held at 0
        const uint32_t hi = ring / stride, lo = ring % stride;
        const uint32_t base = hi * n * stride + lo;
        uint32_t* m = mult + c.mult_off;
        uint32_t acc = 0;
        for (uint32_t j = 0; j < n; ++j) {
            uint32_t* x = m + ((uint64_t)(base + j * stride) * t.nd + dd) * 2 + dir;
            acc += *x;
            *x = acc;
            mx = acc > mx ? acc : mx;
            sm += acc;
        }
    }
    // workgroup reduction, then one atomic per workgroup and class
    __shared__ uint32_t smax[256];
    __shared__ unsigned long long ssum[256];
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            smax[threadIdx.x] = smax[threadIdx.x] > smax[threadIdx.x + st] ? smax[threadIdx.x] : smax[threadIdx.x + st];
            ssum[threadIdx.x] += ssum[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax(max_mult + bk.x, smax[0]);
        atomicAdd(sum_mult + bk.x, ssum[0]);
    }
}

hipError_t launch_sim(const void* cases, const void* classes, uint32_t ncls, const void* count_blocks,
                      uint32_t n_count_blocks, const void* scan_blocks, uint32_t n_scan_blocks, uint32_t* mult,
                      uint64_t mult_len, uint32_t* max_hops, uint32_t* max_mult, uint64_t* sum_mult, hipStream_t st) {
    hipLaunchKernelGGL(sim_zero, dim3(4096), dim3(256), 0, st, mult, mult_len, max_hops, max_mult,
                       (unsigned long long*)sum_mult, ncls);
    if (n_count_blocks)
        hipLaunchKernelGGL(sim_count, dim3(n_count_blocks), dim3(256), 0, st, (const SimCaseRec*)cases,
                           (const SimClassRec*)classes, (const uint2*)count_blocks, mult, max_hops);
    if (n_scan_blocks)
        hipLaunchKernelGGL(sim_scan, dim3(n_scan_blocks), dim3(256), 0, st, (const SimCaseRec*)cases,
                           (const SimClassRec*)classes, (const uint2*)scan_blocks, mult, max_mult,
                           (unsigned long long*)sum_mult);
    return hipGetLastError();
}

}  // namespace bine
