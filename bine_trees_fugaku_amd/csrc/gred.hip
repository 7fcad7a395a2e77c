// gred.hip -- the gather-reduce kernel that runs every Bine schedule on gfx950.
//
// A work item names a group of leaves (buffer, rank, offset) whose element-wise
// reduction tree is perfect, a list of outputs, and a tile [delta, delta + n)
// of elements.  A workgroup streams the tile: for each 16-byte vector it loads
// the leaves in groups of four or eight, reduces each group pairwise in
// registers and folds the group sums with a binary-counter stack, so the
// association is the balanced tree ((l0+l1)+(l2+l3))+... over the leaf order
// the host chose.  The
// sum is then stored to every output.  One-leaf items are copies (permutes),
// two-leaf items are the per-step "received + local" reductions.
//
// Leaf and output entries are the same for the whole workgroup, so they are
// fetched with scalar loads; data moves with 128-bit global loads/stores.
// Compile with -Xclang -target-feature -Xclang -packed-fp32-ops.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gred.hpp"

#ifndef BINE_GRED_REALIGN
#define BINE_GRED_REALIGN 0  // gred_realign.hip compiles this file again with 1
#endif

namespace bine {

typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// Address-space views.  Data goes through global (address space 1) pointers, so loads and stores issue
// as global_* instructions, which count against the vector-memory counter only; generic pointers would
// give flat_* instructions, which also hold the scalar/LDS counter, so that every wait for a scalar load
// would drain the data loads in flight as well.  Work items, pool entries and buffer tables are read
// through constant (address space 4) pointers with a wave-uniform address, so they issue as scalar loads
// (one s_load_dwordx8 for four consecutive pool entries) that never wait on the data loads.
#define BINE_GLOBAL __attribute__((address_space(1)))
#define BINE_CONST __attribute__((address_space(4)))
typedef BINE_GLOBAL u4 gu4;

template <typename X>
__device__ __forceinline__ BINE_GLOBAL X* gptr(const void* p) {
    return (BINE_GLOBAL X*)(uintptr_t)p;
}

// p must be the same for every lane of the wave
template <typename X>
__device__ __forceinline__ const BINE_CONST X* cptr(const X* p) {
    const uint64_t a = (uint64_t)(uintptr_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return (const BINE_CONST X*)(((uint64_t)hi << 32) | lo);
}

// A struct read word by word through a constant pointer (a plain copy would need a copy constructor that
// takes the generic address space).
template <typename X>
__device__ __forceinline__ X cload(const BINE_CONST X* p) {
    static_assert(sizeof(X) % 4 == 0, "struct of whole words");
    const BINE_CONST uint32_t* q = (const BINE_CONST uint32_t*)p;
    uint32_t w[sizeof(X) / 4];
#pragma unroll
    for (unsigned i = 0; i < sizeof(X) / 4; ++i) w[i] = q[i];
    X r;
    __builtin_memcpy(&r, w, sizeof(X));
    return r;
}

// 16 bytes of elements of type T, held as one 128-bit register quad.
template <typename T>
struct Vec {
    static constexpr int N = 16 / sizeof(T);
    u4 raw;
};

template <typename T>
__device__ __forceinline__ T add1(T a, T b) {
    return a + b;
}
template <>
__device__ __forceinline__ int32_t add1<int32_t>(int32_t a, int32_t b) {
    return (int32_t)((uint32_t)a + (uint32_t)b);
}
template <>
__device__ __forceinline__ int64_t add1<int64_t>(int64_t a, int64_t b) {
    return (int64_t)((uint64_t)a + (uint64_t)b);
}
template <>
__device__ __forceinline__ int8_t add1<int8_t>(int8_t a, int8_t b) {
    return (int8_t)((uint8_t)a + (uint8_t)b);
}

template <typename T>
__device__ __forceinline__ Vec<T> vadd(const Vec<T>& a, const Vec<T>& b) {
    T x[Vec<T>::N], y[Vec<T>::N];
    __builtin_memcpy(x, &a.raw, 16);
    __builtin_memcpy(y, &b.raw, 16);
#pragma unroll
    for (int i = 0; i < Vec<T>::N; ++i) x[i] = add1<T>(x[i], y[i]);
    Vec<T> r;
    __builtin_memcpy(&r.raw, x, 16);
    return r;
}

// 16 bytes added lane by lane (mod 256) on four 32-bit words: the low seven
// bits of each byte cannot carry out of the byte, the top bit is the xor of
// the carry into it with both operands' top bits.
template <>
__device__ __forceinline__ Vec<int8_t> vadd<int8_t>(const Vec<int8_t>& a, const Vec<int8_t>& b) {
    const u4 x = a.raw, y = b.raw;
    Vec<int8_t> r;
    r.raw = ((x & 0x7F7F7F7Fu) + (y & 0x7F7F7F7Fu)) ^ ((x ^ y) & 0x80808080u);
    return r;
}

template <typename T, bool NT>
__device__ __forceinline__ Vec<T> vload(const char* p) {
    Vec<T> r;
    if (NT) r.raw = __builtin_nontemporal_load(gptr<const u4>(p));
    else r.raw = *gptr<const u4>(p);
    return r;
}

template <typename T, bool NT>
__device__ __forceinline__ void vstore(char* p, const Vec<T>& v) {
    if (NT) __builtin_nontemporal_store(v.raw, gptr<u4>(p));
    else *gptr<u4>(p) = v.raw;
}

// ---- realigned stores ----------------------------------------------------------
// A vector tile's output can start 4, 8 or 12 bytes past a 16-byte boundary
// (partition strides that are not multiples of 16 bytes, e.g. odd fp32
// counts, or an output block whose offset has another residue than its
// source's).  gfx950 serves unaligned 16-byte loads at full rate but unaligned
// 16-byte stores about 15 % slower than aligned ones, so the realigning
// gather-reduce kernels (REALIGN = true; the host launches them for
// store-heavy work lists, where stores dominate, that can have unaligned
// outputs) write such outputs as aligned chunks: with the output k dwords past a boundary,
// chunk c holds the last k dwords of vector c - 1 and the first 4 - k of
// vector c.  Each wave holds 64 * VPT consecutive vectors (lane l of round u
// holds vector 64u + l); a lane takes its predecessor's last three words by a
// lane rotation (lane 0 of round u from lane 63 of round u - 1), the wave's
// first vector writes only its own part of its chunk and the wave's last
// vector also writes its last k dwords into the following chunk.  Every byte
// of the output range is written exactly once.
typedef unsigned int u2 __attribute__((ext_vector_type(2)));

struct Prev {
    unsigned y, z, w;
};

// lane l receives the words of lane (l - 1) mod 64
__device__ __forceinline__ Prev rotate_prev(const u4& v, uint32_t lane) {
    const int src = (int)((lane + 63u) & 63u);
    return Prev{(unsigned)__shfl((int)v.y, src, 64), (unsigned)__shfl((int)v.z, src, 64),
                (unsigned)__shfl((int)v.w, src, 64)};
}

template <bool NT>
__device__ __forceinline__ void st_word(char* p, unsigned v) {
    if (NT) __builtin_nontemporal_store(v, gptr<unsigned>(p));
    else *gptr<unsigned>(p) = v;
}

template <bool NT>
__device__ __forceinline__ void st_pair(char* p, const u2& v) {  // p 8-byte aligned
    if (NT) __builtin_nontemporal_store(v, gptr<u2>(p));
    else *gptr<u2>(p) = v;
}

template <bool NT>
__device__ __forceinline__ void st_quad(char* p, const u4& v) {
    if (NT) __builtin_nontemporal_store(v, gptr<u4>(p));
    else *gptr<u4>(p) = v;
}

// p: this lane's vector address, k = (p mod 16) / 4 in 1..3 (the same for the
// whole wave); pv: the previous vector's last three words; first: the previous
// vector is not this wave's; last: the next vector is not this wave's.
template <bool NT>
__device__ __forceinline__ void store_shifted(char* p, uint32_t k, const u4& o, const Prev& pv, bool first, bool last) {
    char* c = p - 4 * k;  // the aligned chunk that holds the vector's first 4 - k dwords
    if (!first) {
        const u4 v = k == 1 ? u4{pv.w, o.x, o.y, o.z} : k == 2 ? u4{pv.z, pv.w, o.x, o.y} : u4{pv.y, pv.z, pv.w, o.x};
        st_quad<NT>(c, v);
    } else if (k == 1) {  // own dwords 0..2 at chunk dwords 1..3
        st_word<NT>(p, o.x);
        st_pair<NT>(p + 4, u2{o.y, o.z});
    } else if (k == 2) {
        st_pair<NT>(p, u2{o.x, o.y});
    } else {
        st_word<NT>(p, o.x);
    }
    if (last) {  // own dwords 4-k..3 at the next chunk's dwords 0..k-1
        char* n = c + 16;
        if (k == 1) {
            st_word<NT>(n, o.w);
        } else if (k == 2) {
            st_pair<NT>(n, u2{o.z, o.w});
        } else {
            st_pair<NT>(n, u2{o.y, o.z});
            st_word<NT>(n + 8, o.w);
        }
    }
}

// One wave's 64 * VPT consecutive vectors (round u, lane l: vector index
// vi[u]) of one output p, realigned; rot[u] = rotate_prev of round u.
template <int VPT, bool NT>
__device__ __forceinline__ void store_rounds_shifted(char* p, uint32_t k, const u4 (&v)[VPT], const Prev (&rot)[VPT],
                                                     const uint32_t (&vi)[VPT], const bool (&ok)[VPT], uint32_t lane,
                                                     uint32_t nvec) {
#pragma unroll
    for (int u = 0; u < VPT; ++u) {
        if (ok[u]) {
            const Prev pv = (u > 0 && lane == 0) ? rot[u > 0 ? u - 1 : 0] : rot[u];
            store_shifted<NT>(p + (uint64_t)vi[u] * 16, k, v[u], pv, u == 0 && lane == 0,
                              vi[u] + 1 == nvec || (u == VPT - 1 && lane == 63u));
        }
    }
}

// Buffer address of a (buffer, rank, element offset) entry.  The buffer is
// picked with masks rather than an indexed or selected load so the table stays
// in scalar registers.
struct Table {
    uint64_t b0, b1, b2, b3, s0, s1, s2, s3;
    __device__ __forceinline__ char* at(uint32_t rb, uint64_t elem_off, uint32_t tsize) const {
        const uint32_t buf = rb >> 28, rank = rb & 0x0FFFFFFFu;
        const uint64_t m0 = 0ull - (uint64_t)(buf == 0), m1 = 0ull - (uint64_t)(buf == 1);
        const uint64_t m2 = 0ull - (uint64_t)(buf == 2), m3 = 0ull - (uint64_t)(buf == 3);
        const uint64_t base = (b0 & m0) | (b1 & m1) | (b2 & m2) | (b3 & m3);
        const uint64_t stride = (s0 & m0) | (s1 & m1) | (s2 & m2) | (s3 & m3);
        return reinterpret_cast<char*>(base + (uint64_t)rank * stride + elem_off * tsize);
    }
};

// Evaluate the balanced tree over 2^lg leaves; `load(k)` fetches leaf k.
// Leaves are taken in groups of G (4 or 8), each reduced pairwise, and the
// group sums are folded with a binary counter: the association is the same
// balanced tree for either G; G = 4 halves the loads in flight per thread.
template <typename W, int MAXS, int G, typename AddF, typename LoadF>
__device__ __forceinline__ W tree_eval(uint32_t lg, AddF add, LoadF load) {
    if (lg == 0) return load(0);
    if (lg == 1) return add(load(0), load(1));
    if (lg == 2) return add(add(load(0), load(1)), add(load(2), load(3)));
    W stk[MAXS];
    const uint32_t S = lg - (G == 8 ? 3 : 2);
    const uint32_t ngroups = 1u << S;
    W res;
    for (uint32_t j = 0; j < ngroups; ++j) {
        const uint32_t k = j * G;
        W v;
        if (G == 8) {
            W x0 = load(k + 0), x1 = load(k + 1), x2 = load(k + 2), x3 = load(k + 3);
            W x4 = load(k + 4), x5 = load(k + 5), x6 = load(k + 6), x7 = load(k + 7);
            v = add(add(add(x0, x1), add(x2, x3)), add(add(x4, x5), add(x6, x7)));
        } else {
            W x0 = load(k + 0), x1 = load(k + 1), x2 = load(k + 2), x3 = load(k + 3);
            v = add(add(x0, x1), add(x2, x3));
        }
        // binary-counter carry: level l holds the sum of the previous 2^l groups
        bool carry = true;
#pragma unroll
        for (int l = 0; l < MAXS; ++l) {
            if (carry && l < (int)S) {
                if ((j >> l) & 1u) {
                    v = add(stk[l], v);
                } else {
                    stk[l] = v;
                    carry = false;
                }
            }
        }
        res = v;
    }
    return res;
}

// ITEM_RUN items: k copies; copy j reads first leaf + j * step and writes first
// output + j * step.  Element-type agnostic: raw 16-byte vectors (esize only
// scales offsets).  A kernel of its own so the reduction kernel's register
// allocation is untouched.
template <int VPT, bool BATCH, bool NT>
__global__ void __launch_bounds__(512) grun_kernel(const GredItem* __restrict__ items, uint32_t nitems,
                                                   const PoolEntry* __restrict__ pool, BufTable tab,
                                                   const BufTable* __restrict__ tabs, uint32_t allow_vec,
                                                   uint32_t esize) {
    for (uint32_t ii = blockIdx.x; ii < nitems; ii += gridDim.x) {
        const GredItem it = cload(cptr(items + ii));
        Table t;
        if (BATCH) {
            const BINE_CONST BufTable* bt = cptr(tabs + it.bucket);
            t = Table{bt->base[0], bt->base[1], bt->base[2], bt->base[3], bt->stride[0], bt->stride[1], bt->stride[2], bt->stride[3]};
        } else {
            t = Table{tab.base[0], tab.base[1], tab.base[2], tab.base[3], tab.stride[0], tab.stride[1], tab.stride[2], tab.stride[3]};
        }
        const BINE_CONST PoolEntry* lp = cptr(pool + it.leaf_off);
        const BINE_CONST PoolEntry* op = cptr(pool + it.out_off);
        const PoolEntry l0 = cload(lp), ls = cload(lp + 1);
        const PoolEntry o0 = cload(op), os = cload(op + 1);
        const uint64_t delta = it.delta;
        const bool vec = (it.flags & ITEM_VEC) && (allow_vec & 1u);
        // two copies per pass, so every thread keeps 2 * VPT 16-byte loads in flight
        for (uint32_t j = 0; j < it.pad; j += 2) {
            const uint32_t j1 = j + 1 < it.pad ? j + 1 : j;  // an odd tail copies its last block twice
            const char* s0 = t.at(l0.rb + ls.rb * j, (uint64_t)l0.off + (uint64_t)ls.off * j + delta, esize);
            const char* s1 = t.at(l0.rb + ls.rb * j1, (uint64_t)l0.off + (uint64_t)ls.off * j1 + delta, esize);
            char* d0 = t.at(o0.rb + os.rb * j, (uint64_t)o0.off + (uint64_t)os.off * j + delta, esize);
            char* d1 = t.at(o0.rb + os.rb * j1, (uint64_t)o0.off + (uint64_t)os.off * j1 + delta, esize);
            if (vec) {
                const uint32_t nvec = (uint32_t)(((uint64_t)it.n * esize) / 16);
                for (uint32_t base = 0; base < nvec; base += blockDim.x * VPT) {
                    u4 a[VPT], c[VPT];
#pragma unroll
                    for (int u = 0; u < VPT; ++u) {
                        const uint32_t vi = base + threadIdx.x + u * blockDim.x;
                        if (vi < nvec) {
                            const BINE_GLOBAL u4* q0 = gptr<const u4>(s0 + (uint64_t)vi * 16);
                            const BINE_GLOBAL u4* q1 = gptr<const u4>(s1 + (uint64_t)vi * 16);
                            a[u] = NT ? __builtin_nontemporal_load(q0) : *q0;
                            c[u] = NT ? __builtin_nontemporal_load(q1) : *q1;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < VPT; ++u) {
                        const uint32_t vi = base + threadIdx.x + u * blockDim.x;
                        if (vi < nvec) {
                            BINE_GLOBAL u4* q0 = gptr<u4>(d0 + (uint64_t)vi * 16);
                            BINE_GLOBAL u4* q1 = gptr<u4>(d1 + (uint64_t)vi * 16);
                            if (NT) {
                                __builtin_nontemporal_store(a[u], q0);
                                __builtin_nontemporal_store(c[u], q1);
                            } else {
                                *q0 = a[u];
                                *q1 = c[u];
                            }
                        }
                    }
                }
            } else {
                const uint64_t nb = (uint64_t)it.n * esize;
                for (uint64_t b = threadIdx.x; b < nb; b += blockDim.x) {
                    const char x0 = *gptr<const char>(s0 + b), x1 = *gptr<const char>(s1 + b);
                    *gptr<char>(d0 + b) = x0;
                    *gptr<char>(d1 + b) = x1;
                }
            }
        }
    }
}

template <typename T, int VPT>
struct Pack {
    Vec<T> x[VPT];
};

// Leaves per group of tree_eval: eight where the registers allow (two vectors per thread, or four vectors with
// a stack of six levels, trees of up to 512 leaves), four otherwise.
template <int VPT, int MAXS>
constexpr int leaf_group() {
    return VPT == 4 && MAXS != 6 ? 4 : 8;
}

template <typename T, int VPT, int MAXS, bool BATCH, bool NT, bool REALIGN>
__global__ void __launch_bounds__(512) gred_kernel(const GredItem* __restrict__ items, uint32_t nitems,
                                                   const PoolEntry* __restrict__ pool, BufTable tab,
                                                   const BufTable* __restrict__ tabs, uint32_t allow_vec) {
    constexpr int VE = Vec<T>::N;
    constexpr int G = leaf_group<VPT, MAXS>();
    // partial subtree sums of the waves, [wave][u][lane] 16-byte vectors
    extern __shared__ __align__(16) unsigned char lds_raw[];
    Vec<T>* lds = reinterpret_cast<Vec<T>*>(lds_raw);
    const uint32_t nw = blockDim.x >> 6;
    const uint32_t lgw = __builtin_ctz(nw);
    // the wave index is the same on every lane of a wave: said so, the output loop `o = wave; o += nw` reads
    // its pool entries with scalar loads
    const uint32_t lane = threadIdx.x & 63u, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t ii = blockIdx.x; ii < nitems; ii += gridDim.x) {
        const GredItem it = cload(cptr(items + ii));
        Table t;
        if (BATCH) {
            const BINE_CONST BufTable* bt = cptr(tabs + it.bucket);
            t = Table{bt->base[0], bt->base[1], bt->base[2], bt->base[3], bt->stride[0], bt->stride[1], bt->stride[2], bt->stride[3]};
        } else {
            t = Table{tab.base[0], tab.base[1], tab.base[2], tab.base[3], tab.stride[0], tab.stride[1], tab.stride[2], tab.stride[3]};
        }
        const BINE_CONST PoolEntry* leaves = cptr(pool + it.leaf_off);
        const BINE_CONST PoolEntry* outs = cptr(pool + it.out_off);
        const uint32_t lg = it.lg;
        const uint64_t delta = it.delta;
        auto addp = [](const Pack<T, VPT>& a, const Pack<T, VPT>& b) __attribute__((always_inline)) {
            Pack<T, VPT> r;
#pragma unroll
            for (int u = 0; u < VPT; ++u) r.x[u] = vadd<T>(a.x[u], b.x[u]);
            return r;
        };
        const bool vec = (it.flags & ITEM_VEC) && (allow_vec & 1u);
        if (vec && nw > 1 && lg >= lgw + 3) {
            // Deep tree: wave w reduces the contiguous leaf range [w << lgs, (w+1) << lgs),
            // a whole subtree of the balanced tree; the top lgw levels are then
            // joined pairwise from LDS, so the association is unchanged.
            const uint32_t lgs = lg - lgw;
            const BINE_CONST PoolEntry* my = cptr(pool + it.leaf_off + ((size_t)wave << lgs));
            const uint32_t nvec = it.n / VE;
            for (uint32_t base = 0; base < nvec; base += 64 * VPT) {
                uint32_t vi[VPT];
                bool ok[VPT];
#pragma unroll
                for (int u = 0; u < VPT; ++u) {
                    vi[u] = base + lane + u * 64;
                    ok[u] = vi[u] < nvec;
                }
                auto loadp = [&](uint32_t k) __attribute__((always_inline)) {
                    const PoolEntry e = cload(my + k);
                    const char* p = t.at(e.rb, (uint64_t)e.off + delta, sizeof(T));
                    Pack<T, VPT> r;
#pragma unroll
                    for (int u = 0; u < VPT; ++u) {
                        if (ok[u]) r.x[u] = vload<T, NT>(p + (uint64_t)vi[u] * 16);
                        else r.x[u] = Vec<T>{};
                    }
                    return r;
                };
                Pack<T, VPT> part = tree_eval<Pack<T, VPT>, MAXS, G>(lgs, addp, loadp);
#pragma unroll
                for (int u = 0; u < VPT; ++u) lds[((size_t)wave * VPT + u) * 64 + lane] = part.x[u];
                __syncthreads();
                // join the wave subtrees pairwise in place: level st adds wave w+st into w
                for (uint32_t st = 1; st < nw; st <<= 1) {
                    if ((wave & (2 * st - 1)) == 0) {
#pragma unroll
                        for (int u = 0; u < VPT; ++u) {
                            Vec<T>& a = lds[((size_t)wave * VPT + u) * 64 + lane];
                            a = vadd<T>(a, lds[((size_t)(wave + st) * VPT + u) * 64 + lane]);
                        }
                    }
                    __syncthreads();
                }
                Pack<T, VPT> sum;
#pragma unroll
                for (int u = 0; u < VPT; ++u) sum.x[u] = lds[(size_t)u * 64 + lane];
                __syncthreads();  // the next pass reuses the LDS slots
                if (REALIGN) {
                    u4 v[VPT];
                    Prev rot[VPT];
#pragma unroll
                    for (int u = 0; u < VPT; ++u) {
                        v[u] = sum.x[u].raw;
                        rot[u] = rotate_prev(v[u], lane);
                    }
                    for (uint32_t o = wave; o < it.nouts; o += nw) {
                        const PoolEntry e = cload(outs + o);
                        char* p = t.at(e.rb, (uint64_t)e.off + delta, sizeof(T));
                        const uint32_t m = (uint32_t)(uintptr_t)p & 15u;
                        if (m != 0 && (m & 3u) == 0) {
                            store_rounds_shifted<VPT, NT>(p, m >> 2, v, rot, vi, ok, lane, nvec);
                        } else {
#pragma unroll
                            for (int u = 0; u < VPT; ++u)
                                if (ok[u]) vstore<T, NT>(p + (uint64_t)vi[u] * 16, sum.x[u]);
                        }
                    }
                } else {
                    for (uint32_t o = wave; o < it.nouts; o += nw) {
                        const PoolEntry e = cload(outs + o);
                        char* p = t.at(e.rb, (uint64_t)e.off + delta, sizeof(T));
#pragma unroll
                        for (int u = 0; u < VPT; ++u)
                            if (ok[u]) vstore<T, NT>(p + (uint64_t)vi[u] * 16, sum.x[u]);
                    }
                }
            }
        } else if (vec) {
            const uint32_t nvec = it.n / VE;
            for (uint32_t base = 0; base < nvec; base += blockDim.x * VPT) {
                uint32_t vi[VPT];
                bool ok[VPT];
#pragma unroll
                for (int u = 0; u < VPT; ++u) {
                    // realigning: each wave holds 64 * VPT consecutive vectors
                    vi[u] = REALIGN ? base + wave * (64u * VPT) + u * 64u + lane : base + threadIdx.x + u * blockDim.x;
                    ok[u] = vi[u] < nvec;
                }
                auto loadp = [&](uint32_t k) __attribute__((always_inline)) {
                    const PoolEntry e = cload(leaves + k);
                    const char* p = t.at(e.rb, (uint64_t)e.off + delta, sizeof(T));
                    Pack<T, VPT> r;
#pragma unroll
                    for (int u = 0; u < VPT; ++u) {
                        if (ok[u]) r.x[u] = vload<T, NT>(p + (uint64_t)vi[u] * 16);
                        else r.x[u] = Vec<T>{};
                    }
                    return r;
                };
                Pack<T, VPT> s = tree_eval<Pack<T, VPT>, MAXS, G>(lg, addp, loadp);
                if (REALIGN) {
                    u4 v[VPT];
                    Prev rot[VPT];
#pragma unroll
                    for (int u = 0; u < VPT; ++u) {
                        v[u] = s.x[u].raw;
                        rot[u] = rotate_prev(v[u], lane);
                    }
                    for (uint32_t o = 0; o < it.nouts; ++o) {
                        const PoolEntry e = cload(outs + o);
                        char* p = t.at(e.rb, (uint64_t)e.off + delta, sizeof(T));
                        const uint32_t m = (uint32_t)(uintptr_t)p & 15u;
                        if (m != 0 && (m & 3u) == 0) {
                            store_rounds_shifted<VPT, NT>(p, m >> 2, v, rot, vi, ok, lane, nvec);
                        } else {
#pragma unroll
                            for (int u = 0; u < VPT; ++u)
                                if (ok[u]) vstore<T, NT>(p + (uint64_t)vi[u] * 16, s.x[u]);
                        }
                    }
                } else {
                    for (uint32_t o = 0; o < it.nouts; ++o) {
                        const PoolEntry e = cload(outs + o);
                        char* p = t.at(e.rb, (uint64_t)e.off + delta, sizeof(T));
#pragma unroll
                        for (int u = 0; u < VPT; ++u)
                            if (ok[u]) vstore<T, NT>(p + (uint64_t)vi[u] * 16, s.x[u]);
                    }
                }
            }
        } else {
            for (uint32_t e0 = threadIdx.x; e0 < it.n; e0 += blockDim.x) {
                auto adds = [](T a, T b) __attribute__((always_inline)) { return add1<T>(a, b); };
                auto loads = [&](uint32_t k) __attribute__((always_inline)) {
                    const PoolEntry e = cload(leaves + k);
                    return *gptr<const T>(t.at(e.rb, (uint64_t)e.off + delta + e0, sizeof(T)));
                };
                T s = tree_eval<T, MAXS, 8>(lg, adds, loads);
                for (uint32_t o = 0; o < it.nouts; ++o) {
                    const PoolEntry e = cload(outs + o);
                    *gptr<T>(t.at(e.rb, (uint64_t)e.off + delta + e0, sizeof(T))) = s;
                }
            }
        }
    }
}

// out = a + b element-wise (reduce_local, lib/libbine_coll.cc:153-267).  out may be a or b itself (each
// element is read before it is written, by the same lane).  VEC: the first nvec * (16 / sizeof(T))
// elements move as 16-byte vectors (every base dword-aligned; the global memory serves unaligned 16-byte
// accesses), VPT vectors per lane per pass; the rest, and everything when !VEC, element by element.
template <typename T, bool VEC, int VPT>
__global__ void __launch_bounds__(256) reduce_local_kernel(const T* a, const T* b, T* out, uint64_t nvec,
                                                           uint64_t n) {
    const uint64_t lanes = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t done = 0;
    if (VEC) {
        const char* pa = reinterpret_cast<const char*>(a);
        const char* pb = reinterpret_cast<const char*>(b);
        char* po = reinterpret_cast<char*>(out);
        for (uint64_t v0 = tid; v0 < nvec; v0 += lanes * VPT) {
            Vec<T> x[VPT], y[VPT];
#pragma unroll
            for (int u = 0; u < VPT; ++u) {
                const uint64_t v = v0 + (uint64_t)u * lanes;
                if (v < nvec) {
                    x[u] = vload<T, true>(pa + v * 16);
                    y[u] = vload<T, true>(pb + v * 16);
                }
            }
#pragma unroll
            for (int u = 0; u < VPT; ++u) {
                const uint64_t v = v0 + (uint64_t)u * lanes;
                if (v < nvec) vstore<T, true>(po + v * 16, vadd<T>(x[u], y[u]));
            }
        }
        done = nvec * (uint64_t)Vec<T>::N;
    }
    for (uint64_t i = done + tid; i < n; i += lanes) *gptr<T>(out + i) = add1<T>(*gptr<const T>(a + i), *gptr<const T>(b + i));
}

// ---- host-side dispatch ------------------------------------------------------

template <typename T, int VPT, int MAXS, bool BATCH, bool RA>
static hipError_t launch_t(const GredItem* items, uint32_t n, const PoolEntry* pool, const BufTable& tab,
                           const BufTable* tabs, uint32_t allow_vec, const LaunchShape& sh, hipStream_t st) {
    uint32_t grid = sh.persistent > 0 ? (uint32_t)sh.persistent : n;
    if (grid > n) grid = n;
    if (grid == 0) return hipSuccess;
    if (sh.threads < 64 || sh.threads > 512 || (sh.threads & (sh.threads - 1))) return hipErrorInvalidValue;
    const size_t lds = (size_t)(sh.threads / 64) * VPT * 64 * 16;
    if (sh.nontemporal)
        hipLaunchKernelGGL((gred_kernel<T, VPT, MAXS, BATCH, true, RA>), dim3(grid), dim3(sh.threads), lds, st, items,
                           n, pool, tab, tabs, allow_vec);
    else
        hipLaunchKernelGGL((gred_kernel<T, VPT, MAXS, BATCH, false, RA>), dim3(grid), dim3(sh.threads), lds, st, items,
                           n, pool, tab, tabs, allow_vec);
    return hipGetLastError();
}

template <typename T, bool BATCH, bool RA>
static hipError_t launch_v(int max_lg, const GredItem* items, uint32_t n, const PoolEntry* pool, const BufTable& tab,
                           const BufTable* tabs, uint32_t allow_vec, const LaunchShape& sh, hipStream_t st) {
    const bool small = max_lg <= 9;
    switch (sh.vpt) {
    case 1:
        return small ? launch_t<T, 1, 6, BATCH, RA>(items, n, pool, tab, tabs, allow_vec, sh, st)
                     : launch_t<T, 1, 11, BATCH, RA>(items, n, pool, tab, tabs, allow_vec, sh, st);
    case 2:
        return small ? launch_t<T, 2, 6, BATCH, RA>(items, n, pool, tab, tabs, allow_vec, sh, st)
                     : launch_t<T, 2, 11, BATCH, RA>(items, n, pool, tab, tabs, allow_vec, sh, st);
    default:  // four vectors per thread; deep trees run with two
        // trees of up to 64 leaves: groups of 4 leaves, four stack levels, 163 VGPRs, three waves per SIMD
        if (max_lg <= 6) return launch_t<T, 4, 4, BATCH, RA>(items, n, pool, tab, tabs, allow_vec, sh, st);
        if (!small) return launch_t<T, 2, 11, BATCH, RA>(items, n, pool, tab, tabs, allow_vec, sh, st);
        // up to 512 leaves: groups of 8 leaves and six stack levels (250 VGPRs, no scratch; 0.3-1.4 % faster on
        // the 512-rank allreduce than groups of 4 with seven levels), except where those registers do not fit:
        // int8's byte arithmetic and the realigning stores
        if constexpr (sizeof(T) > 1 && !RA)
            return launch_t<T, 4, 6, BATCH, RA>(items, n, pool, tab, tabs, allow_vec, sh, st);
        else
            return launch_t<T, 4, 7, BATCH, RA>(items, n, pool, tab, tabs, allow_vec, sh, st);
    }
}

// dtype codes follow bine_datatype_t
template <bool RA>
hipError_t launch_gred_t(int dtype, int max_lg, const GredItem* items, uint32_t n, const PoolEntry* pool,
                         const BufTable& tab, const BufTable* tabs, bool batch, uint32_t allow_vec,
                         const LaunchShape& sh, hipStream_t st) {
#define BINE_DISPATCH(T)                                                                               \
    return batch ? launch_v<T, true, RA>(max_lg, items, n, pool, tab, tabs, allow_vec, sh, st)         \
                 : launch_v<T, false, RA>(max_lg, items, n, pool, tab, tabs, allow_vec, sh, st)
    switch (dtype) {
    case 0: BINE_DISPATCH(int32_t);
    case 1: BINE_DISPATCH(int64_t);
    case 2: BINE_DISPATCH(float);
    case 3: BINE_DISPATCH(double);
    case 4: BINE_DISPATCH(int8_t);
    default: return hipErrorInvalidValue;
    }
#undef BINE_DISPATCH
}

#if BINE_GRED_REALIGN
// this unit holds the realigning instantiations only
template hipError_t launch_gred_t<true>(int, int, const GredItem*, uint32_t, const PoolEntry*, const BufTable&,
                                        const BufTable*, bool, uint32_t, const LaunchShape&, hipStream_t);
#else
extern template hipError_t launch_gred_t<true>(int, int, const GredItem*, uint32_t, const PoolEntry*, const BufTable&,
                                               const BufTable*, bool, uint32_t, const LaunchShape&, hipStream_t);

// allow_vec: bit 0 permits ITEM_VEC items; bit 1 selects the realigning kernels
hipError_t launch_gred(int dtype, int max_lg, const GredItem* items, uint32_t n, const PoolEntry* pool,
                       const BufTable& tab, const BufTable* tabs, bool batch, uint32_t allow_vec,
                       const LaunchShape& sh, hipStream_t st) {
    if (n == 0) return hipSuccess;
    return (allow_vec & 2u) ? launch_gred_t<true>(dtype, max_lg, items, n, pool, tab, tabs, batch, allow_vec, sh, st)
                            : launch_gred_t<false>(dtype, max_lg, items, n, pool, tab, tabs, batch, allow_vec, sh, st);
}

hipError_t launch_grun(uint32_t esize, const GredItem* items, uint32_t n, const PoolEntry* pool, const BufTable& tab,
                       const BufTable* tabs, bool batch, uint32_t allow_vec, const LaunchShape& sh, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (sh.threads < 64 || sh.threads > 512 || (sh.threads & (sh.threads - 1))) return hipErrorInvalidValue;
    uint32_t grid = sh.persistent > 0 ? (uint32_t)sh.persistent : n;
    if (grid > n) grid = n;
#define BINE_RUN(V, B, N)                                                                                     \
    hipLaunchKernelGGL((grun_kernel<V, B, N>), dim3(grid), dim3(sh.threads), 0, st, items, n, pool, tab, tabs, \
                       allow_vec, esize)
#define BINE_RUN_V(B, N)           \
    switch (sh.vpt) {              \
    case 1: BINE_RUN(1, B, N); break; \
    case 2: BINE_RUN(2, B, N); break; \
    default: BINE_RUN(4, B, N); break; \
    }
    if (batch) {
        if (sh.nontemporal) { BINE_RUN_V(true, true) } else { BINE_RUN_V(true, false) }
    } else {
        if (sh.nontemporal) { BINE_RUN_V(false, true) } else { BINE_RUN_V(false, false) }
    }
#undef BINE_RUN_V
#undef BINE_RUN
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_rl(const void* a, const void* b, void* out, uint64_t n, hipStream_t st) {
    const bool vec = (((uintptr_t)a | (uintptr_t)b | (uintptr_t)out) & 3u) == 0;
    const uint64_t nvec = vec ? n / (uint64_t)Vec<T>::N : 0;
    constexpr int kVpt = 4;
    // enough lanes for every vector to be one lane's within one pass of VPT, capped at 8 workgroups per CU
    uint64_t work = vec ? (nvec + kVpt - 1) / kVpt : n;
    if (work == 0) work = n;  // only a scalar tail
    uint64_t blocks = (work + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks == 0) blocks = 1;
    if (vec)
        hipLaunchKernelGGL((reduce_local_kernel<T, true, kVpt>), dim3(blocks), dim3(256), 0, st, (const T*)a,
                           (const T*)b, (T*)out, nvec, n);
    else
        hipLaunchKernelGGL((reduce_local_kernel<T, false, kVpt>), dim3(blocks), dim3(256), 0, st, (const T*)a,
                           (const T*)b, (T*)out, nvec, n);
    return hipGetLastError();
}

// out = a + b over n elements of bine_datatype_t dtype
hipError_t launch_reduce_local(int dtype, const void* a, const void* b, void* out, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    switch (dtype) {
    case 0: return launch_rl<int32_t>(a, b, out, n, st);
    case 1: return launch_rl<int64_t>(a, b, out, n, st);
    case 2: return launch_rl<float>(a, b, out, n, st);
    case 3: return launch_rl<double>(a, b, out, n, st);
    case 4: return launch_rl<int8_t>(a, b, out, n, st);
    default: return hipErrorInvalidValue;
    }
}

#endif  // BINE_GRED_REALIGN

}  // namespace bine
