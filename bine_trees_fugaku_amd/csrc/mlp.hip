// mlp.hip -- fp32 multilayer perceptron (ReLU hidden layers) for the
// algorithm selector: forward, backward and Adam on the GPU.
//
// Layer l maps A_{l-1} (n x in_l) to Z_l = A_{l-1} W_l^T + b_l (n x out_l);
// hidden layers apply ReLU, the last layer is linear.  Parameters live in one
// flat fp32 buffer, layer by layer: W_l (out_l x in_l, row-major), then b_l.
//
// Every product is one tiled GEMM on the matrix cores: a 64x64 output tile
// per 256-thread workgroup, each wave owning a 32x32 quarter built from 2x2
// v_mfma_f32_16x16x4f32 blocks, operands staged through LDS in K chunks of 64
// (16-byte loads and LDS stores), the next chunk loaded into registers while
// the current one is multiplied.
// The loss is a masked mean squared error; every reduction runs in a fixed
// order, so gradients are deterministic run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bine_hip.h"
#include "status.hpp"

namespace bine {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 64;

enum Epi : int {
    EPI_BIAS = 0,       // C = acc + bias[col]
    EPI_BIAS_RELU = 1,  // C = max(acc + bias[col], 0)
    EPI_RELU_MASK = 2,  // C = acc * (mask[row, col] > 0)   (dZ of a ReLU layer)
    EPI_PLAIN = 3,      // C = acc
};

constexpr int SK = TK + 4;   // LDS row stride of a panel held [row][k]
constexpr int SR = TM + 16;  // LDS row stride of a panel held [k][row]

// One operand's 64 x 64 panel of op(X): rows [r0, r0 + 64), k in [k0, k0 + 64).
// Element (r, k) is X[r * ld + k] when KC (k contiguous in memory), else
// X[k * ld + r].  Thread t holds four quads: quad q has the outer index
// t / 16 + 16 q and the inner indices 4 (t % 16) + 0..3, the inner index
// running along X's contiguous dimension, so a whole panel moves as 16-byte
// loads and 16-byte LDS stores.  LDS holds the panel in X's own orientation:
// [r][k] with row stride SK when KC, [k][r] with row stride SR otherwise; both
// strides spread a fragment read's 64 lanes over 64 consecutive banks.
template <bool KC>
struct Panel {
    const float* X;
    int ld, r0, R;
    bool vec;  // X and ld allow 16-byte quads

    __device__ __forceinline__ size_t at(int r, int k) const { return KC ? (size_t)r * ld + k : (size_t)k * ld + r; }
    __device__ __forceinline__ void rk(int k0, int o, int c, int& r, int& k) const {
        if (KC) {
            r = r0 + o;
            k = k0 + c;
        } else {
            k = k0 + o;
            r = r0 + c;
        }
    }
    // the whole panel lies inside the operand and moves as quads (the same for the whole workgroup)
    __device__ __forceinline__ bool whole(int k0, int kend) const { return vec && r0 + TM <= R && k0 + TK <= kend; }

    __device__ __forceinline__ void fetch(floatx4 (&v)[4], int t, int k0, int kend, int kbeg) const {
        if (whole(k0, kend)) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int r, k;
                rk(k0, t / 16 + 16 * q, 4 * (t % 16), r, k);
                v[q] = *reinterpret_cast<const floatx4*>(X + at(r, k));
            }
        } else {  // element by element; an element out of range loads the panel's first one, replaced when staged
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    int r, k;
                    rk(k0, t / 16 + 16 * q, 4 * (t % 16) + c, r, k);
                    const bool ok = r < R && k < kend;
                    v[q][c] = X[ok ? at(r, k) : at(r0, kbeg)];
                }
        }
    }
    __device__ __forceinline__ void stage(float* S, const floatx4 (&v)[4], int t, int k0, int kend) const {
        const bool w = whole(k0, kend);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            floatx4 x = v[q];
            if (!w) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    int r, k;
                    rk(k0, t / 16 + 16 * q, 4 * (t % 16) + c, r, k);
                    if (!(r < R && k < kend)) x[c] = 0.f;
                }
            }
            *reinterpret_cast<floatx4*>(S + (t / 16 + 16 * q) * (KC ? SK : SR) + 4 * (t % 16)) = x;
        }
    }
    // panel element (row rr, k kk), from LDS
    __device__ __forceinline__ float frag(const float* S, int rr, int kk) const {
        return KC ? S[rr * SK + kk] : S[kk * SR + rr];
    }
};

// C[M x N] = op(A)[M x K] * op(B)[K x N] with the epilogue above.
// op(A)[m][k] = TA ? A[k * lda + m] : A[m * lda + k]
// op(B)[k][n] = TB ? B[n * ldb + k] : B[k * ldb + n]
// Split K: workgroup z covers k in [z * kc, min(K, (z + 1) * kc)) and, when
// gridDim.z > 1, writes its partial product to C + z * M * ldc (EPI_PLAIN only;
// sum_partials adds the slices in order).
//
// The K loop is software-pipelined: the next chunk's operands are loaded into
// registers while the matrix cores work on the current chunk from LDS.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f32(int M, int N, int K, int kc, const float* __restrict__ A, int lda,
                                                const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                                const float* __restrict__ bias, const float* __restrict__ mask, int epi) {
    __shared__ __align__(16) float As[TK * SR];  // room for either orientation (SR >= SK)
    __shared__ __align__(16) float Bs[TK * SR];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kbeg = blockIdx.z * kc;
    const int kend = kbeg + kc < K ? kbeg + kc : K;
    C += (size_t)blockIdx.z * M * ldc;
    const Panel<!TA> pa{A, lda, m0, M, (lda & 3) == 0 && ((uintptr_t)A & 15) == 0};
    const Panel<TB> pb{B, ldb, n0, N, (ldb & 3) == 0 && ((uintptr_t)B & 15) == 0};
    floatx4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    floatx4 va[4], vb[4];
    pa.fetch(va, tid, kbeg, kend, kbeg);
    pb.fetch(vb, tid, kbeg, kend, kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        pa.stage(As, va, tid, k0, kend);
        pb.stage(Bs, vb, tid, k0, kend);
        __syncthreads();
        if (k0 + TK < kend) {  // in flight during the products below
            pa.fetch(va, tid, k0 + TK, kend, kbeg);
            pb.fetch(vb, tid, k0 + TK, kend, kbeg);
        }
        // 16x16x4: lane supplies A[row = lane % 16][k = lane / 16] and B[k = lane / 16][col = lane % 16].
        // The chunk's fragments go to registers ahead of the products that use them, so LDS latency overlaps
        // the matrix cores.
        float af[TK / 4][2], bf[TK / 4][2];
#pragma unroll
        for (int q = 0; q < TK / 4; ++q) {
            const int kk = 4 * q + (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                af[q][i] = pa.frag(As, wm + i * 16 + (lane & 15), kk);
                bf[q][i] = pb.frag(Bs, wn + i * 16 + (lane & 15), kk);
            }
        }
#pragma unroll
        for (int q = 0; q < TK / 4; ++q)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[q][i], bf[q][j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // Output layout of a 16x16 block: lane holds rows 4 * (lane / 16) + r, column lane % 16.
    // The epilogue's operands (bias, ReLU mask) are loaded unconditionally from clamped, in-range positions
    // and all values are formed before the first store, so the lane waits for memory once.
    const bool has_bias = epi == EPI_BIAS || epi == EPI_BIAS_RELU;
    float bj[2] = {0.f, 0.f};
    if (has_bias) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + j * 16 + (lane & 15);
            bj[j] = bias[col < N ? col : n0];
        }
    }
    float mk[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) mk[i][j][r] = 1.f;
    if (epi == EPI_RELU_MASK) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                    const int col = n0 + wn + j * 16 + (lane & 15);
                    mk[i][j][r] = mask[(size_t)(row < M ? row : m0) * ldc + (col < N ? col : n0)];
                }
    }
    float out[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = acc[i][j][r] + bj[j];  // bj = 0 without a bias
                if (epi == EPI_BIAS_RELU) v = v > 0.f ? v : 0.f;
                if (epi == EPI_RELU_MASK) v = mk[i][j][r] > 0.f ? v : 0.f;
                out[i][j][r] = v;
            }
    // The values are pinned in registers here (an empty asm that takes them), so the arithmetic above, and
    // its wait for the loaded operands, is not sunk into the per-element store branches below: a wait there
    // would also wait for the stores issued before it.
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) asm volatile("" : "+v"(out[i][j][r]));
    if (m0 + TM <= M && n0 + TN <= N) {  // a whole tile: the same for every lane of the workgroup
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                    const int col = n0 + wn + j * 16 + (lane & 15);
                    C[(size_t)row * ldc + col] = out[i][j][r];
                }
    } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                    const int col = n0 + wn + j * 16 + (lane & 15);
                    if (row < M && col < N) C[(size_t)row * ldc + col] = out[i][j][r];
                }
    }
}

// dY = 2 * mask * (Y - T) / max(count, 1); per-block partial sums of the loss
// terms mask * (Y - T)^2 and of the mask, reduced in a fixed order.
__global__ __launch_bounds__(256) void mse_grad(int n_elems, const float* __restrict__ Y, const float* __restrict__ T,
                                                const float* __restrict__ Mk, float* __restrict__ dY,
                                                float* __restrict__ partial, int pass, const float* __restrict__ totals) {
    __shared__ float s_l[256], s_c[256];
    float l = 0.f, c = 0.f;
    const float inv = pass ? 2.f / fmaxf(totals[1], 1.f) : 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n_elems; i += gridDim.x * 256) {
        const float m = Mk[i];
        const float d = m > 0.f ? Y[i] - T[i] : 0.f;
        if (pass) dY[i] = inv * d;
        l += d * d;
        c += m > 0.f ? 1.f : 0.f;
    }
    s_l[threadIdx.x] = l;
    s_c[threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_l[threadIdx.x] += s_l[threadIdx.x + s];
            s_c[threadIdx.x] += s_c[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && !pass) {
        partial[2 * blockIdx.x] = s_l[0];
        partial[2 * blockIdx.x + 1] = s_c[0];
    }
}

// totals = {sum of loss terms, mask count}; loss = totals[0] / max(totals[1], 1).
__global__ void reduce_partials(int nb, const float* __restrict__ partial, float* __restrict__ totals,
                                float* __restrict__ loss_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float l = 0.f, c = 0.f;
    int i = 0;
    for (; i + 8 <= nb; i += 8) {  // eight pairs loaded, then added in order
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = partial[2 * i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            l += v[2 * u];
            c += v[2 * u + 1];
        }
    }
    for (; i < nb; ++i) {
        l += partial[2 * i];
        c += partial[2 * i + 1];
    }
    totals[0] = l;
    totals[1] = c;
    if (loss_out) loss_out[0] = l / fmaxf(c, 1.f);
}

// Partial column sums: workgroup (x, y) sums rows [y * rc, min(n, (y + 1) * rc))
// of 64 columns, rows split over 4 waves, into part[y * cols + col].
__global__ __launch_bounds__(256) void col_sum(int n, int cols, int rc, const float* __restrict__ dZ,
                                               float* __restrict__ part) {
    __shared__ float s[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + lane;
    const int r0 = blockIdx.y * rc;
    const int r1 = r0 + rc < n ? r0 + rc : n;
    float acc = 0.f;
    if (col < cols) {
        int i = r0 + wave;
        for (; i + 28 < r1; i += 32) {  // eight rows loaded, then added in order
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = dZ[(size_t)(i + 4 * u) * cols + col];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        for (; i < r1; i += 4) acc += dZ[(size_t)i * cols + col];
    }
    s[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && col < cols) part[(size_t)blockIdx.y * cols + col] = (s[0][lane] + s[1][lane]) + (s[2][lane] + s[3][lane]);
}

// out[i] = sum over z = 0 .. nz-1 of part[z * n + i], in that order.  The
// slices are loaded eight at a time before they are added, so one thread waits
// for memory once per eight slices rather than once per slice.
__global__ void sum_partials(int n, int nz, const float* __restrict__ part, float* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float acc = part[i];
        int z = 1;
        for (; z + 8 <= nz; z += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[(size_t)(z + u) * n + i];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        for (; z < nz; ++z) acc += part[(size_t)z * n + i];
        out[i] = acc;
    }
}

__global__ void adam_update(uint64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

template <bool TA, bool TB>
void gemm(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc, const float* bias,
          const float* mask, int epi, hipStream_t st) {
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, 1);
    hipLaunchKernelGGL((gemm_f32<TA, TB>), grid, dim3(256), 0, st, M, N, K, K, A, lda, B, ldb, C, ldc, bias, mask, epi);
}

// Number of K slices for a product whose output has `tiles` 64x64 tiles:
// enough workgroups to cover the chip, slices of at least 128 rows, at most kMaxSplit.
constexpr int kMaxSplit = 32;
int k_slices(int K, int tiles) {
    int nz = (512 + tiles - 1) / tiles;
    nz = std::min(nz, std::max(1, K / 128));
    return std::max(1, std::min(nz, kMaxSplit));
}

// C[M x N] = op(A) op(B) with K split into slices summed in a fixed order
// (work: nz * M * N floats).
template <bool TA, bool TB>
void gemm_split(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, float* work,
                hipStream_t st) {
    const int tiles = ((N + TN - 1) / TN) * ((M + TM - 1) / TM);
    const int nz = k_slices(K, tiles);
    if (nz == 1) {
        gemm<TA, TB>(M, N, K, A, lda, B, ldb, C, N, nullptr, nullptr, EPI_PLAIN, st);
        return;
    }
    const int kc = ((K + nz - 1) / nz + TK - 1) / TK * TK;
    const int nzr = (K + kc - 1) / kc;
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, nzr);
    hipLaunchKernelGGL((gemm_f32<TA, TB>), grid, dim3(256), 0, st, M, N, K, kc, A, lda, B, ldb, work, N, nullptr,
                       nullptr, (int)EPI_PLAIN);
    const int n = M * N;
    hipLaunchKernelGGL(sum_partials, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, st, n, nzr, work, C);
}

// db = column sums of dZ (n x cols), rows split into slices summed in a fixed order.
void col_sums(int n, int cols, const float* dz, float* db, float* work, hipStream_t st) {
    const int groups = (cols + 63) / 64;
    const int nz = k_slices(n, groups);
    const int rc = (n + nz - 1) / nz;
    const int nzr = (n + rc - 1) / rc;
    hipLaunchKernelGGL(col_sum, dim3(groups, nzr), dim3(256), 0, st, n, cols, rc, dz, work);
    hipLaunchKernelGGL(sum_partials, dim3((cols + 255) / 256), dim3(256), 0, st, cols, nzr, work, db);
}

constexpr int kLossBlocks = 120;

}  // namespace
}  // namespace bine

using namespace bine;

struct bine_mlp_s {
    int device = 0;
    std::vector<int> dims;  // in, hidden..., out
    int max_batch = 0;
    uint64_t nparams = 0;
    std::vector<uint64_t> w_off, b_off;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    std::vector<float*> act;  // act[l]: n x dims[l+1] outputs of layer l (post-ReLU for hidden layers)
    float *dz0 = nullptr, *dz1 = nullptr;  // ping-pong n x max width
    float *partial = nullptr, *totals = nullptr;
    float* work = nullptr;  // split-K partial products
    uint64_t step = 0;
    hipStream_t last = nullptr;  // stream of the last launch; host copies wait for it
    ~bine_mlp_s() {
        for (float* p : {params, grads, adam_m, adam_v, dz0, dz1, partial, totals, work})
            if (p) (void)hipFree(p);
        for (float* p : act)
            if (p) (void)hipFree(p);
    }
    int layers() const { return (int)dims.size() - 1; }
};

namespace {

int mfail(int code, const std::string& m) { return record_error(code, m); }

#define MHIP(x)                                                                                           \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess) return mfail(BINE_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Forward pass over n rows of x; leaves every layer's output in m->act.
void forward(bine_mlp_t m, const float* x, int n, hipStream_t st) {
    const float* in = x;
    for (int l = 0; l < m->layers(); ++l) {
        const int fi = m->dims[l], fo = m->dims[l + 1];
        const bool last = l == m->layers() - 1;
        gemm<false, true>(n, fo, fi, in, fi, m->params + m->w_off[l], fi, m->act[l], fo, m->params + m->b_off[l],
                          nullptr, last ? EPI_BIAS : EPI_BIAS_RELU, st);
        in = m->act[l];
    }
}

// Loss and gradients of the masked MSE into m->grads (and *loss_out when given).
void backward(bine_mlp_t m, const float* x, const float* t, const float* mask, int n, float* loss_out, hipStream_t st) {
    const int L = m->layers(), fo = m->dims[L];
    const int ne = n * fo;
    const int nb = std::min(kLossBlocks, (ne + 255) / 256);
    float* y = m->act[L - 1];
    hipLaunchKernelGGL(mse_grad, dim3(nb), dim3(256), 0, st, ne, y, t, mask, m->dz0, m->partial, 0, m->totals);
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(1), 0, st, nb, m->partial, m->totals, loss_out);
    hipLaunchKernelGGL(mse_grad, dim3(nb), dim3(256), 0, st, ne, y, t, mask, m->dz0, m->partial, 1, m->totals);
    float* dz = m->dz0;
    float* other = m->dz1;
    for (int l = L - 1; l >= 0; --l) {
        const int fi = m->dims[l], fout = m->dims[l + 1];
        const float* a_prev = l ? m->act[l - 1] : x;
        // dW_l = dZ_l^T A_{l-1}  (fout x fi), K = n
        gemm_split<true, false>(fout, fi, n, dz, fout, a_prev, fi, m->grads + m->w_off[l], m->work, st);
        col_sums(n, fout, dz, m->grads + m->b_off[l], m->work, st);
        if (l) {  // dZ_{l-1} = (dZ_l W_l) * (A_{l-1} > 0)
            gemm<false, false>(n, fi, fout, dz, fout, m->params + m->w_off[l], fi, other, fi, nullptr, m->act[l - 1],
                               EPI_RELU_MASK, st);
            std::swap(dz, other);
        }
    }
}

int check_batch(bine_mlp_t m, const void* x, int n) {
    if (!m) return mfail(BINE_ERR_ARG, "NULL network");
    if (!x) return mfail(BINE_ERR_ARG, "NULL input");
    if (n <= 0 || n > m->max_batch) return mfail(BINE_ERR_ARG, "batch size must be in [1, max_batch]");
    return BINE_OK;
}

}  // namespace

extern "C" int bine_mlp_create(const int32_t* dims, int32_t ndims, int32_t max_batch, uint64_t seed, int device,
                               bine_mlp_t* out) {
    if (!dims || ndims < 2 || ndims > 16 || max_batch <= 0 || !out) return mfail(BINE_ERR_ARG, "bad arguments");
    for (int i = 0; i < ndims; ++i)
        if (dims[i] <= 0 || dims[i] > 4096) return mfail(BINE_ERR_ARG, "layer widths must be in [1, 4096]");
    auto m = std::make_unique<bine_mlp_s>();
    m->device = device;
    m->dims.assign(dims, dims + ndims);
    m->max_batch = max_batch;
    uint64_t off = 0;
    int widest = 0;
    for (int l = 0; l + 1 < ndims; ++l) {
        m->w_off.push_back(off);
        off += (uint64_t)dims[l] * dims[l + 1];
        m->b_off.push_back(off);
        off += (uint64_t)dims[l + 1];
    }
    for (int i = 0; i < ndims; ++i) widest = std::max(widest, (int)dims[i]);
    m->nparams = off;
    // Weights uniform in +-1/sqrt(fan_in) from splitmix64, zero biases.
    std::vector<float> h(off, 0.f);
    uint64_t s = seed;
    for (int l = 0; l + 1 < ndims; ++l) {
        const float bound = 1.f / std::sqrt((float)dims[l]);
        for (uint64_t i = 0; i < (uint64_t)dims[l] * dims[l + 1]; ++i) {
            const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
            h[m->w_off[l] + i] = (float)((2.0 * u - 1.0) * bound);
        }
    }
    MHIP(hipSetDevice(device));
    const size_t pb = sizeof(float) * off;
    MHIP(hipMalloc(&m->params, pb));
    MHIP(hipMalloc(&m->grads, pb));
    MHIP(hipMalloc(&m->adam_m, pb));
    MHIP(hipMalloc(&m->adam_v, pb));
    MHIP(hipMemcpy(m->params, h.data(), pb, hipMemcpyHostToDevice));
    MHIP(hipMemset(m->adam_m, 0, pb));
    MHIP(hipMemset(m->adam_v, 0, pb));
    MHIP(hipMemset(m->grads, 0, pb));
    for (int l = 0; l + 1 < ndims; ++l) {
        float* a = nullptr;
        MHIP(hipMalloc(&a, sizeof(float) * (size_t)max_batch * dims[l + 1]));
        m->act.push_back(a);
    }
    MHIP(hipMalloc(&m->dz0, sizeof(float) * (size_t)max_batch * widest));
    MHIP(hipMalloc(&m->dz1, sizeof(float) * (size_t)max_batch * widest));
    size_t widest_w = 0;
    for (int l = 0; l + 1 < ndims; ++l) widest_w = std::max(widest_w, (size_t)dims[l] * dims[l + 1]);
    MHIP(hipMalloc(&m->work, sizeof(float) * kMaxSplit * std::max(widest_w, (size_t)widest)));
    MHIP(hipMalloc(&m->partial, sizeof(float) * 2 * kLossBlocks));
    MHIP(hipMalloc(&m->totals, sizeof(float) * 2));
    *out = m.release();
    return BINE_OK;
}

extern "C" int bine_mlp_num_params(bine_mlp_t m, uint64_t* n) {
    if (!m || !n) return mfail(BINE_ERR_ARG, "NULL argument");
    *n = m->nparams;
    return BINE_OK;
}

extern "C" int bine_mlp_get_params(bine_mlp_t m, float* host) {
    if (!m || !host) return mfail(BINE_ERR_ARG, "NULL argument");
    MHIP(hipSetDevice(m->device));
    MHIP(hipStreamSynchronize(m->last));
    MHIP(hipMemcpy(host, m->params, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
    return BINE_OK;
}

extern "C" int bine_mlp_set_params(bine_mlp_t m, const float* host) {
    if (!m || !host) return mfail(BINE_ERR_ARG, "NULL argument");
    MHIP(hipSetDevice(m->device));
    MHIP(hipStreamSynchronize(m->last));
    MHIP(hipMemcpy(m->params, host, sizeof(float) * m->nparams, hipMemcpyHostToDevice));
    const size_t pb = sizeof(float) * m->nparams;
    MHIP(hipMemset(m->adam_m, 0, pb));
    MHIP(hipMemset(m->adam_v, 0, pb));
    m->step = 0;
    return BINE_OK;
}

extern "C" int bine_mlp_forward(bine_mlp_t m, const float* x, int32_t n, float* y, void* stream) {
    int rc = check_batch(m, x, n);
    if (rc) return rc;
    if (!y) return mfail(BINE_ERR_ARG, "NULL output");
    MHIP(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    m->last = st;
    forward(m, x, n, st);
    MHIP(hipMemcpyAsync(y, m->act.back(), sizeof(float) * (size_t)n * m->dims.back(), hipMemcpyDeviceToDevice, st));
    MHIP(hipGetLastError());
    return BINE_OK;
}

extern "C" int bine_mlp_gradients(bine_mlp_t m, const float* x, const float* t, const float* mask, int32_t n,
                                  float* grads, float* loss, void* stream) {
    int rc = check_batch(m, x, n);
    if (rc) return rc;
    if (!t || !mask || !grads) return mfail(BINE_ERR_ARG, "NULL target, mask or gradient buffer");
    MHIP(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    m->last = st;
    forward(m, x, n, st);
    backward(m, x, t, mask, n, loss, st);
    MHIP(hipMemcpyAsync(grads, m->grads, sizeof(float) * m->nparams, hipMemcpyDeviceToDevice, st));
    MHIP(hipGetLastError());
    return BINE_OK;
}

extern "C" int bine_mlp_train_step(bine_mlp_t m, const float* x, const float* t, const float* mask, int32_t n,
                                   float lr, float* loss, void* stream) {
    int rc = check_batch(m, x, n);
    if (rc) return rc;
    if (!t || !mask) return mfail(BINE_ERR_ARG, "NULL target or mask");
    if (!(lr > 0.f)) return mfail(BINE_ERR_ARG, "learning rate must be positive");
    MHIP(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    m->last = st;
    forward(m, x, n, st);
    backward(m, x, t, mask, n, loss, st);
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    const uint64_t np = m->nparams;
    const int blocks = (int)std::min<uint64_t>(1024, (np + 255) / 256);
    hipLaunchKernelGGL(adam_update, dim3(blocks), dim3(256), 0, st, np, m->params, m->grads, m->adam_m, m->adam_v, lr,
                       b1, b2, 1e-8f, c1, c2);
    MHIP(hipGetLastError());
    return BINE_OK;
}

extern "C" int bine_mlp_destroy(bine_mlp_t m) {
    if (!m) return BINE_OK;
    (void)hipSetDevice(m->device);
    delete m;
    return BINE_OK;
}
