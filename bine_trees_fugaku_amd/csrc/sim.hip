// sim.hip -- integer-atomic link counting for the torus traffic simulation.
//
// One thread per (step class, rank): the rank's destination under the class's
// schedule, then a dimension-ordered minimal route with one 32-bit atomic add
// per crossed directed link.  A second kernel reduces each class to its most
// loaded link and its total link crossings.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sim_dest.hpp"

namespace bine {

using simd::SimCaseRec;
using simd::SimClassRec;

__global__ void sim_zero(uint32_t* __restrict__ p, uint64_t n, uint32_t* __restrict__ hops, uint32_t ncls) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncls; i += (uint64_t)gridDim.x * blockDim.x) hops[i] = 0;
}

__global__ void sim_count(const SimCaseRec* __restrict__ cases, const SimClassRec* __restrict__ classes,
                          uint32_t* __restrict__ mult, uint32_t* __restrict__ max_hops) {
    const SimClassRec c = classes[blockIdx.y];
    const SimCaseRec cs = cases[c.case_idx];
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
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
    uint32_t* m = mult + c.mult_off;
    uint32_t hops = 0;
    for (uint32_t dd = 0; dd < cs.t.nd; ++dd) {
        const int dim = (int)cs.t.dims[dd];
        const int delta = simd::pm(d[dd] - s[dd], dim);
        if (!delta) continue;
        const bool neg = 2 * delta > dim;
        const int nh = neg ? dim - delta : delta;
        for (int h = 0; h < nh; ++h) {
            const uint32_t node = simd::ident(cs.t, s);
            atomicAdd(m + ((uint64_t)node * cs.t.nd + dd) * 2 + (neg ? 1 : 0), 1u);
            s[dd] = simd::pm(s[dd] + (neg ? -1 : 1), dim);
        }
        hops += (uint32_t)nh;
    }
    atomicMax(max_hops + blockIdx.y, hops);
}

__global__ void sim_reduce(const SimCaseRec* __restrict__ cases, const SimClassRec* __restrict__ classes,
                           const uint32_t* __restrict__ mult, uint32_t* __restrict__ max_mult, uint64_t* __restrict__ sum_mult) {
    __shared__ uint32_t smax[256];
    __shared__ uint64_t ssum[256];
    const SimClassRec c = classes[blockIdx.x];
    const uint32_t nl = cases[c.case_idx].nlinks;
    const uint32_t* m = mult + c.mult_off;
    uint32_t mx = 0;
    uint64_t sm = 0;
    for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) {
        const uint32_t v = m[i];
        mx = v > mx ? v : mx;
        sm += v;
    }
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
        max_mult[blockIdx.x] = smax[0];
        sum_mult[blockIdx.x] = ssum[0];
    }
}

hipError_t launch_sim(const void* cases, const void* classes, uint32_t ncls, uint32_t max_p, uint32_t* mult,
                      uint64_t mult_len, uint32_t* max_hops, uint32_t* max_mult, uint64_t* sum_mult, hipStream_t st) {
    hipLaunchKernelGGL(sim_zero, dim3(4096), dim3(256), 0, st, mult, mult_len, max_hops, ncls);
    for (uint32_t first = 0; first < ncls; first += 65535) {
        const uint32_t n = ncls - first < 65535 ? ncls - first : 65535;
        hipLaunchKernelGGL(sim_count, dim3((max_p + 255) / 256, n), dim3(256), 0, st,
                           (const SimCaseRec*)cases, (const SimClassRec*)classes + first, mult, max_hops + first);
        hipLaunchKernelGGL(sim_reduce, dim3(n), dim3(256), 0, st, (const SimCaseRec*)cases,
                           (const SimClassRec*)classes + first, mult, max_mult + first, sum_mult + first);
    }
    return hipGetLastError();
}

}  // namespace bine
