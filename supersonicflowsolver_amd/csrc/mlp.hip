// mlp.hip -- the warm-start surrogate: a coordinate MLP trained and run on
// gfx950 with hand-written kernels (C-ABI in include/psst_mlp.h).
//
// Hidden layers are BF16 GEMMs on the MFMA units (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation) with fused epilogues:
//   forward   Z = H W + b, H' = bf16(SiLU(Z))          (Z kept in FP32)
//   backward  dZ = bf16((dZ' W^T) * SiLU'(Z))
//             dW = H^T dZ' (split over the batch, fixed-order reduction)
// The input layer (n_in = 4 coordinates) and the output layer (n_out = 4
// fields) are too narrow for MFMA tiles and keep their inputs in FP32; they
// run as plain FP32 kernels on the BF16-rounded weights.  Every reduction
// is a fixed-order sum (no floating-point atomics), so training is
// deterministic.  The optimiser is Adam on FP32 master weights; after each
// step the BF16 copies the kernels read are refreshed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "psst_mlp.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32, LDS_K = BK + 8;  // +8: row stride of 80 bytes
constexpr int kThreads = 256;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

enum { EPI_FWD = 0, EPI_DH = 1, EPI_PART = 2 };

// C = A B over a 64x64 tile per workgroup (4 waves, each 32x32 = 2x2 MFMA
// 16x16 tiles), K in steps of 32 staged through LDS.
//   A(m,k) = TA ? A[k*lda + m] : A[m*lda + k]
//   B(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n]
// LDS holds the A tile as [m][k] and the B tile as [n][k], so that each MFMA
// operand fragment (8 consecutive k of one row / one column) is one 16-byte
// read.  Fragment layout of v_mfma_f32_16x16x32_bf16: lane l supplies row
// (l & 15) of A and column (l & 15) of B at k = 8*(l >> 4) .. +7, and
// receives C rows 4*(l >> 4) .. +3 of column (l & 15).
// blockIdx.z selects a K range [z*kps, (z+1)*kps) (split-K for EPI_PART).
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(kThreads) k_gemm(int M, int N, int K, const bf16 *__restrict__ A, int lda,
                                                   const bf16 *__restrict__ B, int ldb, const float *__restrict__ bias,
                                                   const float *__restrict__ zprev, float *__restrict__ zout,
                                                   bf16 *__restrict__ hout, int kps) {
    __shared__ __align__(16) bf16 As[BM][LDS_K];
    __shared__ __align__(16) bf16 Bs[BN][LDS_K];
    const int tm = blockIdx.x * BM, tn = blockIdx.y * BN;
    const int k0 = blockIdx.z * kps, k1 = min(K, k0 + kps);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kk = k0; kk < k1; kk += BK) {
        if (!TA) {
            const int r = tid >> 2, c = (tid & 3) * 8;
            *reinterpret_cast<bf16x8 *>(&As[r][c]) =
                *reinterpret_cast<const bf16x8 *>(A + (size_t)(tm + r) * lda + kk + c);
        } else {
            const int kr = tid >> 3, m = (tid & 7) * 8;
            const bf16x8 v = *reinterpret_cast<const bf16x8 *>(A + (size_t)(kk + kr) * lda + tm + m);
#pragma unroll
            for (int e = 0; e < 8; ++e) As[m + e][kr] = v[e];
        }
        if (TB) {
            const int n = tid >> 2, c = (tid & 3) * 8;
            *reinterpret_cast<bf16x8 *>(&Bs[n][c]) =
                *reinterpret_cast<const bf16x8 *>(B + (size_t)(tn + n) * ldb + kk + c);
        } else {
            const int kr = tid >> 3, n = (tid & 7) * 8;
            const bf16x8 v = *reinterpret_cast<const bf16x8 *>(B + (size_t)(kk + kr) * ldb + tn + n);
#pragma unroll
            for (int e = 0; e < 8; ++e) Bs[n + e][kr] = v[e];
        }
        __syncthreads();
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + i * 16 + fr][fk]);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + j * 16 + fr][fk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = tm + wm + i * 16 + (lane >> 4) * 4 + r, n = tn + wn + j * 16 + (lane & 15);
                const size_t o = (size_t)m * N + n;
                const float c = acc[i][j][r];
                if (EPI == EPI_FWD) {
                    const float z = c + bias[n];
                    if (zout) zout[o] = z;
                    hout[o] = (bf16)silu(z);
                } else if (EPI == EPI_DH) {
                    hout[o] = (bf16)(c * dsilu(zprev[o]));
                } else {
                    zout[(size_t)blockIdx.z * M * N + o] = c;
                }
            }
}

// Input layer: Z[m][j] = sum_i x[m][i] w[i][j] (i ascending) + b[j] in FP32,
// H = bf16(SiLU(Z)); rows m >= m_real are zero inputs.
__global__ void k_in_fwd(int M, int m_real, int n_in, int W, const float *__restrict__ x, const float *__restrict__ w,
                         const float *__restrict__ b, float *__restrict__ Z, bf16 *__restrict__ H) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)M * W) return;
    const int m = (int)(idx / W), j = (int)(idx % W);
    float z = 0.f;
    if (m < m_real)
        for (int i = 0; i < n_in; ++i) z += x[(size_t)m * n_in + i] * w[(size_t)i * W + j];
    z += b[j];
    if (Z) Z[idx] = z;
    H[idx] = (bf16)silu(z);
}

// Output layer, one wave per row: out[m][o] = sum_k H[m][k] w[k][o] + b[o]
// (each lane sums every 64th k, then a fixed butterfly over the lanes).  With
// targets: d = out - y, dOut = 2 d / count, per-block partial sums of d^2.
template <int NOUT>
__global__ void __launch_bounds__(kThreads) k_out_fwd(int M, int m_real, int W, const bf16 *__restrict__ H,
                                                      const float *__restrict__ w, const float *__restrict__ b,
                                                      float *__restrict__ out, const float *__restrict__ y,
                                                      float *__restrict__ dout, float scale, double *__restrict__ loss_part) {
    __shared__ double sl[kThreads / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = blockIdx.x * (kThreads / 64) + wv;
    double l2 = 0.0;
    if (m < M) {
        float acc[NOUT];
#pragma unroll
        for (int o = 0; o < NOUT; ++o) acc[o] = 0.f;
        for (int k = lane; k < W; k += 64) {
            const float h = (float)H[(size_t)m * W + k];
#pragma unroll
            for (int o = 0; o < NOUT; ++o) acc[o] += h * w[(size_t)k * NOUT + o];
        }
#pragma unroll
        for (int o = 0; o < NOUT; ++o)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[o] += __shfl_xor(acc[o], off, 64);
        if (lane == 0 && m < m_real) {
#pragma unroll
            for (int o = 0; o < NOUT; ++o) {
                const float v = acc[o] + b[o];
                if (out) out[(size_t)m * NOUT + o] = v;
                if (y) {
                    const float d = v - y[(size_t)m * NOUT + o];
                    dout[(size_t)m * NOUT + o] = 2.0f * d * scale;
                    l2 += (double)d * d;
                }
            }
        } else if (lane == 0 && y) {
#pragma unroll
            for (int o = 0; o < NOUT; ++o) dout[(size_t)m * NOUT + o] = 0.f;
        }
    }
    if (lane == 0) sl[wv] = l2;
    __syncthreads();
    if (threadIdx.x == 0 && loss_part) {
        double s = 0.0;
        for (int k = 0; k < kThreads / 64; ++k) s += sl[k];
        loss_part[blockIdx.x] = s;
    }
}

// dZ[m][k] = bf16((sum_o dOut[m][o] w[k][o]) * SiLU'(Z[m][k])).
template <int NOUT>
__global__ void k_out_bwd(int M, int W, const float *__restrict__ dout, const float *__restrict__ w,
                          const float *__restrict__ Z, bf16 *__restrict__ dZ) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)M * W) return;
    const int m = (int)(idx / W), k = (int)(idx % W);
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) s += dout[(size_t)m * NOUT + o] * w[(size_t)k * NOUT + o];
    dZ[idx] = (bf16)(s * dsilu(Z[idx]));
}

// part[s][i][j] = sum over rows m of chunk s of a(m,i) b(m,j); a == nullptr: 1.
template <typename TA_, typename TB_>
__global__ void k_outer_part(int M, int chunk, int mi, int nj, const TA_ *__restrict__ a, int lda,
                             const TB_ *__restrict__ b, int ldb, float *__restrict__ part) {
    const int e = (int)(blockIdx.y * blockDim.x + threadIdx.x);
    if (e >= mi * nj) return;
    const int i = e / nj, j = e % nj;
    const int m0 = blockIdx.x * chunk, m1 = min(M, m0 + chunk);
    float acc = 0.f;
#pragma unroll 8
    for (int m = m0; m < m1; ++m) acc += (a ? (float)a[(size_t)m * lda + i] : 1.0f) * (float)b[(size_t)m * ldb + j];
    part[(size_t)blockIdx.x * mi * nj + e] = acc;
}

// out[e] = sum_s part[s][e], s ascending.
__global__ void k_reduce(int S, int count, const float *__restrict__ part, float *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    float acc = 0.f;
#pragma unroll 8
    for (int s = 0; s < S; ++s) acc += part[(size_t)s * count + e];
    out[e] = acc;
}

__global__ void k_adam(int64_t P, float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// BF16 copies of the weights: as bf16 (hidden layers) or as the FP32 value
// of the bf16 rounding (input/output layers and biases).
__global__ void k_round_bf16(int64_t n, const float *__restrict__ src, bf16 *__restrict__ dst_b,
                             float *__restrict__ dst_f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bf16 h = (bf16)src[i];
    if (dst_b) dst_b[i] = h;
    if (dst_f) dst_f[i] = (float)h;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Minibatch of row `b`: a seeded hash of (seed, step, b) picks a training row.
__global__ void k_gather(int64_t n, int B, int n_in, int n_out, uint64_t seed, int step, const float *__restrict__ X,
                         const float *__restrict__ Y, float *__restrict__ xb, float *__restrict__ yb) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= B) return;
    const uint64_t h = mix64(seed ^ mix64(((uint64_t)step << 32) | (uint32_t)r));
    const int64_t src = (int64_t)(h % (uint64_t)n);
    for (int i = 0; i < n_in; ++i) xb[(size_t)r * n_in + i] = X[(size_t)src * n_in + i];
    for (int o = 0; o < n_out; ++o) yb[(size_t)r * n_out + o] = Y[(size_t)src * n_out + o];
}

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline unsigned nblk(size_t n, int b = kThreads) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct psst_mlp {
    int n_in = 4, W = 256, depth = 5, n_out = 4;
    int64_t P = 0;
    std::vector<int64_t> w_off, b_off;  // offsets of each layer's W and b in the parameter vector
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // masters, Adam moments, gradient
    std::vector<bf16 *> wb;     // hidden-layer BF16 weights (layers 1..depth-1)
    float *wr = nullptr;        // BF16-rounded FP32 copy of the whole parameter vector
    int64_t adam_t = 0;
    // activation buffers for up to cap rows
    int cap = 0;
    float *xb = nullptr, *yb = nullptr, *out = nullptr, *dout = nullptr, *part = nullptr;
    double *loss_part = nullptr;
    std::vector<float *> Z;  // [depth] pre-activations
    std::vector<bf16 *> H;   // [depth] activations
    bf16 *dz[2] = {nullptr, nullptr};
    size_t part_cap = 0;
    // training data
    float *X = nullptr, *Y = nullptr;
    int64_t n_data = 0;
    double ms_per_step = 0.0;
    std::string err;

    template <typename T>
    T *alloc(size_t n) {
        void *q = nullptr;
        hip_check(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
        allocs.push_back(q);
        return static_cast<T *>(q);
    }
    void release(void *q) {
        if (!q) return;
        auto it = std::find(allocs.begin(), allocs.end(), q);
        if (it != allocs.end()) allocs.erase(it);
        (void)hipFree(q);
    }
    int fan_in(int l) const { return l == 0 ? n_in : W; }
    int fan_out(int l) const { return l == depth ? n_out : W; }

    void refresh_bf16() {
        hipLaunchKernelGGL(k_round_bf16, dim3(nblk((size_t)P)), dim3(kThreads), 0, stream, P, (const float *)p,
                           (bf16 *)nullptr, wr);
        for (int l = 1; l < depth; ++l)
            hipLaunchKernelGGL(k_round_bf16, dim3(nblk((size_t)W * W)), dim3(kThreads), 0, stream, (int64_t)W * W,
                               (const float *)(p + w_off[l]), wb[l], (float *)nullptr);
        hip_check(hipGetLastError(), "refresh");
    }

    void ensure_cap(int rows) {
        if (rows <= cap) return;
        for (void *q : {(void *)xb, (void *)yb, (void *)out, (void *)dout, (void *)loss_part, (void *)dz[0],
                        (void *)dz[1]})
            release(q);
        for (auto q : Z) release(q);
        for (auto q : H) release(q);
        cap = rows;
        xb = alloc<float>((size_t)rows * n_in);
        yb = alloc<float>((size_t)rows * n_out);
        out = alloc<float>((size_t)rows * n_out);
        dout = alloc<float>((size_t)rows * n_out);
        loss_part = alloc<double>(nblk((size_t)rows, kThreads / 64));
        Z.assign((size_t)depth, nullptr);
        H.assign((size_t)depth, nullptr);
        for (int l = 0; l < depth; ++l) {
            Z[(size_t)l] = alloc<float>((size_t)rows * W);
            H[(size_t)l] = alloc<bf16>((size_t)rows * W);
        }
        dz[0] = alloc<bf16>((size_t)rows * W);
        dz[1] = alloc<bf16>((size_t)rows * W);
    }

    // forward over M rows (multiple of 64) of xb; m_real rows carry data;
    // with targets (yb) also the loss partials and dOut
    void forward(int M, int m_real, bool train, bool with_y) {
        const int L = depth;
        hipLaunchKernelGGL(k_in_fwd, dim3(nblk((size_t)M * W)), dim3(kThreads), 0, stream, M, m_real, n_in, W,
                           (const float *)xb, (const float *)(wr + w_off[0]), (const float *)(wr + b_off[0]),
                           train ? Z[0] : (float *)nullptr, H[0]);
        for (int l = 1; l < L; ++l) {
            hipLaunchKernelGGL((k_gemm<false, false, EPI_FWD>), dim3(M / BM, W / BN, 1), dim3(kThreads), 0, stream, M,
                               W, W, (const bf16 *)H[(size_t)l - 1], W, (const bf16 *)wb[(size_t)l], W,
                               (const float *)(wr + b_off[(size_t)l]), (const float *)nullptr,
                               train ? Z[(size_t)l] : (float *)nullptr, H[(size_t)l], W);
        }
        const float scale = 1.0f / (float)((double)m_real * n_out);
        const dim3 g(nblk((size_t)M, kThreads / 64));
        auto launch_out = [&](auto kern) {
            hipLaunchKernelGGL(kern, g, dim3(kThreads), 0, stream, M, m_real, W, (const bf16 *)H[(size_t)L - 1],
                               (const float *)(wr + w_off[(size_t)L]), (const float *)(wr + b_off[(size_t)L]), out,
                               with_y ? (const float *)yb : (const float *)nullptr, dout, scale,
                               with_y ? loss_part : (double *)nullptr);
        };
        switch (n_out) {
        case 1: launch_out(k_out_fwd<1>); break;
        case 2: launch_out(k_out_fwd<2>); break;
        case 3: launch_out(k_out_fwd<3>); break;
        case 4: launch_out(k_out_fwd<4>); break;
        case 5: launch_out(k_out_fwd<5>); break;
        case 6: launch_out(k_out_fwd<6>); break;
        case 7: launch_out(k_out_fwd<7>); break;
        default: launch_out(k_out_fwd<8>); break;
        }
        hip_check(hipGetLastError(), "forward");
    }

    // split reduction of a(m,i) b(m,j) over M rows into g[dst .. dst + mi*nj)
    // Rows per partial sum of outer(): 32 rows give M/32 workgroups, enough
    // to fill the GPU at the training batch sizes (256 rows left 64
    // workgroups, each walking 256 dependent loads: 58 us per call at
    // M = 16384, 4 us with 32 rows).
    static constexpr int kOuterChunk = 32;
    template <typename TA_, typename TB_>
    void outer(int M, int mi, int nj, const TA_ *a, int lda, const TB_ *b, int ldb, int64_t dst) {
        const int chunk = kOuterChunk, S = (M + chunk - 1) / chunk;
        const int count = mi * nj;
        hipLaunchKernelGGL((k_outer_part<TA_, TB_>), dim3(S, nblk((size_t)count)), dim3(kThreads), 0, stream, M,
                           chunk, mi, nj, a, lda, b, ldb, part);
        hipLaunchKernelGGL(k_reduce, dim3(nblk((size_t)count)), dim3(kThreads), 0, stream, S, count,
                           (const float *)part, g + dst);
    }

    void backward(int M) {
        const int L = depth;
        // output layer
        outer<bf16, float>(M, W, n_out, H[(size_t)L - 1], W, dout, n_out, w_off[(size_t)L]);
        outer<float, float>(M, 1, n_out, (const float *)nullptr, 0, dout, n_out, b_off[(size_t)L]);
        auto launch_ob = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(nblk((size_t)M * W)), dim3(kThreads), 0, stream, M, W, (const float *)dout,
                               (const float *)(wr + w_off[(size_t)L]), (const float *)Z[(size_t)L - 1], dz[0]);
        };
        switch (n_out) {
        case 1: launch_ob(k_out_bwd<1>); break;
        case 2: launch_ob(k_out_bwd<2>); break;
        case 3: launch_ob(k_out_bwd<3>); break;
        case 4: launch_ob(k_out_bwd<4>); break;
        case 5: launch_ob(k_out_bwd<5>); break;
        case 6: launch_ob(k_out_bwd<6>); break;
        case 7: launch_ob(k_out_bwd<7>); break;
        default: launch_ob(k_out_bwd<8>); break;
        }
        int cur = 0;
        for (int l = L - 1; l >= 1; --l) {
            // dW_l = H_{l-1}^T dZ  (split over the rows), db_l = column sums of dZ
            const int S = std::max(1, std::min(M / 512, 64));
            const int kps = ((M / S + BK - 1) / BK) * BK;
            const int splits = (M + kps - 1) / kps;
            hipLaunchKernelGGL((k_gemm<true, false, EPI_PART>), dim3(W / BM, W / BN, splits), dim3(kThreads), 0, stream,
                               W, W, M, (const bf16 *)H[(size_t)l - 1], W, (const bf16 *)dz[cur], W,
                               (const float *)nullptr, (const float *)nullptr, part, (bf16 *)nullptr, kps);
            hipLaunchKernelGGL(k_reduce, dim3(nblk((size_t)W * W)), dim3(kThreads), 0, stream, splits, W * W,
                               (const float *)part, g + w_off[(size_t)l]);
            outer<float, bf16>(M, 1, W, (const float *)nullptr, 0, dz[cur], W, b_off[(size_t)l]);
            // dZ_{l-1} = (dZ W_l^T) * SiLU'(Z_{l-1})
            hipLaunchKernelGGL((k_gemm<false, true, EPI_DH>), dim3(M / BM, W / BN, 1), dim3(kThreads), 0, stream, M, W,
                               W, (const bf16 *)dz[cur], W, (const bf16 *)wb[(size_t)l], W, (const float *)nullptr,
                               (const float *)Z[(size_t)l - 1], (float *)nullptr, dz[cur ^ 1], W);
            cur ^= 1;
        }
        // input layer
        outer<float, bf16>(M, n_in, W, xb, n_in, dz[cur], W, w_off[0]);
        outer<float, bf16>(M, 1, W, (const float *)nullptr, 0, dz[cur], W, b_off[0]);
        hip_check(hipGetLastError(), "backward");
    }

    void ensure_part(int M) {
        // split-K partials of the hidden dW, and outer()'s partials (at most
        // W * max(n_in, n_out) entries per row chunk)
        const size_t need = std::max<size_t>((size_t)64 * W * W, (size_t)((M + kOuterChunk - 1) / kOuterChunk) * W *
                                                                     std::max(std::max(n_in, n_out), 1));
        if (need <= part_cap) return;
        release(part);
        part = alloc<float>(need);
        part_cap = need;
    }

    double read_loss(int M) {
        const size_t nb = nblk((size_t)M, kThreads / 64);
        std::vector<double> lp(nb);
        hip_check(hipMemcpyAsync(lp.data(), loss_part, nb * sizeof(double), hipMemcpyDeviceToHost, stream), "D2H");
        hip_check(hipStreamSynchronize(stream), "loss");
        double s = 0.0;
        for (double x : lp) s += x;
        return s;
    }
};

namespace {
thread_local std::string g_mlp_error;

psst_status fail(psst_mlp *m, const std::exception &e, psst_status st = PSST_HIP_ERROR) {
    if (m) m->err = e.what();
    else g_mlp_error = e.what();
    return st;
}

// rows rounded up to the GEMM tile
int padded(int64_t n) { return (int)((n + BM - 1) / BM * BM); }
}  // namespace

extern "C" {

const char *psst_mlp_last_error(const psst_mlp *m) { return m ? m->err.c_str() : g_mlp_error.c_str(); }

psst_status psst_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, uint64_t seed, psst_mlp **out) {
    if (!out || n_in < 1 || n_in > 8 || n_out < 1 || n_out > 8 || width < 64 || width > 1024 || width % 64 ||
        depth < 2) {
        g_mlp_error = "psst_mlp_create: n_in/n_out 1..8, width a multiple of 64 up to 1024, depth >= 2";
        return PSST_BAD_ARG;
    }
    *out = nullptr;
    auto m = std::make_unique<psst_mlp>();
    try {
        m->n_in = n_in;
        m->W = width;
        m->depth = depth;
        m->n_out = n_out;
        int64_t off = 0;
        for (int l = 0; l <= depth; ++l) {
            m->w_off.push_back(off);
            off += (int64_t)m->fan_in(l) * m->fan_out(l);
            m->b_off.push_back(off);
            off += m->fan_out(l);
        }
        m->P = off;
        // Glorot-uniform weights from splitmix64-seeded xoshiro256**, zero biases
        std::vector<float> h((size_t)m->P, 0.f);
        uint64_t st[4], x = seed;
        for (auto &s : st) {
            uint64_t z = (x += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s = z ^ (z >> 31);
        }
        auto next = [&]() {
            auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
            uint64_t r = rotl(st[1] * 5, 7) * 9, t = st[1] << 17;
            st[2] ^= st[0];
            st[3] ^= st[1];
            st[1] ^= st[2];
            st[0] ^= st[3];
            st[2] ^= t;
            st[3] = rotl(st[3], 45);
            return (double)(r >> 11) * (1.0 / 9007199254740992.0);
        };
        for (int l = 0; l <= depth; ++l) {
            const double a = std::sqrt(6.0 / (m->fan_in(l) + m->fan_out(l)));
            for (int64_t k = 0; k < (int64_t)m->fan_in(l) * m->fan_out(l); ++k)
                h[(size_t)(m->w_off[(size_t)l] + k)] = (float)((2.0 * next() - 1.0) * a);
        }
        hip_check(hipStreamCreate(&m->stream), "hipStreamCreate");
        m->p = m->alloc<float>((size_t)m->P);
        m->m = m->alloc<float>((size_t)m->P);
        m->v = m->alloc<float>((size_t)m->P);
        m->g = m->alloc<float>((size_t)m->P);
        m->wr = m->alloc<float>((size_t)m->P);
        m->wb.assign((size_t)depth, nullptr);
        for (int l = 1; l < depth; ++l) m->wb[(size_t)l] = m->alloc<bf16>((size_t)width * width);
        hip_check(hipMemcpy(m->p, h.data(), sizeof(float) * m->P, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemset(m->m, 0, sizeof(float) * m->P), "hipMemset");
        hip_check(hipMemset(m->v, 0, sizeof(float) * m->P), "hipMemset");
        m->refresh_bf16();
        hip_check(hipStreamSynchronize(m->stream), "create");
    } catch (const std::exception &e) {
        g_mlp_error = e.what();
        for (void *q : m->allocs) (void)hipFree(q);
        if (m->stream) (void)hipStreamDestroy(m->stream);
        return PSST_HIP_ERROR;
    }
    *out = m.release();
    return PSST_OK;
}

int64_t psst_mlp_param_count(const psst_mlp *m) { return m ? m->P : 0; }

psst_status psst_mlp_get_params(psst_mlp *m, float *params) {
    if (!m || !params) return PSST_BAD_ARG;
    try {
        hip_check(hipMemcpy(params, m->p, sizeof(float) * m->P, hipMemcpyDeviceToHost), "hipMemcpy");
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(m, e);
    }
}

psst_status psst_mlp_set_params(psst_mlp *m, const float *params) {
    if (!m || !params) return PSST_BAD_ARG;
    try {
        hip_check(hipMemcpy(m->p, params, sizeof(float) * m->P, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemset(m->m, 0, sizeof(float) * m->P), "hipMemset");
        hip_check(hipMemset(m->v, 0, sizeof(float) * m->P), "hipMemset");
        m->adam_t = 0;
        m->refresh_bf16();
        hip_check(hipStreamSynchronize(m->stream), "set_params");
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(m, e);
    }
}

psst_status psst_mlp_forward(psst_mlp *m, const float *x, int64_t n, float *y) {
    if (!m || !x || !y || n <= 0) return PSST_BAD_ARG;
    try {
        const int chunk = 65536;
        for (int64_t b = 0; b < n; b += chunk) {
            const int rows = (int)std::min<int64_t>(chunk, n - b), M = padded(rows);
            m->ensure_cap(M);
            hip_check(hipMemcpyAsync(m->xb, x + b * m->n_in, sizeof(float) * rows * m->n_in, hipMemcpyHostToDevice,
                                     m->stream),
                      "H2D");
            m->forward(M, rows, false, false);
            hip_check(hipMemcpyAsync(y + b * m->n_out, m->out, sizeof(float) * rows * m->n_out, hipMemcpyDeviceToHost,
                                     m->stream),
                      "D2H");
            hip_check(hipStreamSynchronize(m->stream), "forward");
        }
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(m, e);
    }
}

psst_status psst_mlp_loss_grad(psst_mlp *m, const float *x, const float *y, int64_t n, double *loss, float *grad) {
    if (!m || !x || !y || n <= 0 || n > (1 << 24)) return PSST_BAD_ARG;
    try {
        const int M = padded(n);
        m->ensure_cap(M);
        m->ensure_part(M);
        hip_check(hipMemsetAsync(m->xb, 0, sizeof(float) * M * m->n_in, m->stream), "hipMemset");
        hip_check(hipMemcpyAsync(m->xb, x, sizeof(float) * n * m->n_in, hipMemcpyHostToDevice, m->stream), "H2D");
        hip_check(hipMemcpyAsync(m->yb, y, sizeof(float) * n * m->n_out, hipMemcpyHostToDevice, m->stream), "H2D");
        m->forward(M, (int)n, true, true);
        m->backward(M);
        const double l2 = m->read_loss(M);
        if (loss) *loss = l2 / ((double)n * m->n_out);
        if (grad) hip_check(hipMemcpy(grad, m->g, sizeof(float) * m->P, hipMemcpyDeviceToHost), "D2H");
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(m, e);
    }
}

psst_status psst_mlp_set_data(psst_mlp *m, const float *x, const float *y, int64_t n) {
    if (!m || !x || !y || n <= 0) return PSST_BAD_ARG;
    try {
        m->release(m->X);
        m->release(m->Y);
        m->X = m->alloc<float>((size_t)n * m->n_in);
        m->Y = m->alloc<float>((size_t)n * m->n_out);
        hip_check(hipMemcpy(m->X, x, sizeof(float) * n * m->n_in, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(m->Y, y, sizeof(float) * n * m->n_out, hipMemcpyHostToDevice), "hipMemcpy");
        m->n_data = n;
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(m, e);
    }
}

psst_status psst_mlp_train(psst_mlp *m, int32_t steps, int32_t batch, float lr, float lr_end, uint64_t seed,
                           float *loss_hist) {
    if (!m || steps <= 0 || batch <= 0 || batch % 1024) {
        if (m) m->err = "psst_mlp_train: steps > 0 and batch a positive multiple of 1024";
        return PSST_BAD_ARG;
    }
    if (!m->X) {
        m->err = "psst_mlp_train: no training data (psst_mlp_set_data)";
        return PSST_BAD_ARG;
    }
    try {
        m->ensure_cap(batch);
        m->ensure_part(batch);
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, m->stream), "hipEventRecord");
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const size_t nb = nblk((size_t)batch, kThreads / 64);
        std::vector<double> lp(nb);
        for (int s = 0; s < steps; ++s) {
            hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)batch)), dim3(kThreads), 0, m->stream, m->n_data, batch,
                               m->n_in, m->n_out, seed, s, (const float *)m->X, (const float *)m->Y, m->xb, m->yb);
            m->forward(batch, batch, true, true);
            m->backward(batch);
            ++m->adam_t;
            const double prog = steps > 1 ? (double)s / (steps - 1) : 1.0;
            const float lr_s = (float)(lr_end + 0.5 * (lr - lr_end) * (1.0 + std::cos(3.141592653589793 * prog)));
            const float bc1 = 1.0f - std::pow(b1, (float)m->adam_t), bc2 = 1.0f - std::pow(b2, (float)m->adam_t);
            hipLaunchKernelGGL(k_adam, dim3(nblk((size_t)m->P)), dim3(kThreads), 0, m->stream, m->P, m->p, m->m, m->v,
                               (const float *)m->g, lr_s, b1, b2, eps, bc1, bc2);
            m->refresh_bf16();
            if (loss_hist) {
                // partial sums of this step's loss, summed on the host when the step is read back
                hip_check(hipMemcpyAsync(lp.data(), m->loss_part, nb * sizeof(double), hipMemcpyDeviceToHost, m->stream),
                          "D2H");
                hip_check(hipStreamSynchronize(m->stream), "train");
                double l2 = 0.0;
                for (double x : lp) l2 += x;
                loss_hist[s] = (float)(l2 / ((double)batch * m->n_out));
            }
        }
        hip_check(hipEventRecord(e1, m->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "train");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        m->ms_per_step = ms / steps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(m, e);
    }
}

psst_status psst_mlp_train_stats(const psst_mlp *m, double *ms_per_step, double *gemm_flop_per_step) {
    if (!m) return PSST_BAD_ARG;
    if (ms_per_step) *ms_per_step = m->ms_per_step;
    // hidden GEMMs: forward 2 M W^2 per layer; backward dW and dH, 2 M W^2 each
    if (gemm_flop_per_step) *gemm_flop_per_step = 6.0 * m->cap * (double)m->W * m->W * (m->depth - 1);
    return PSST_OK;
}

void psst_mlp_destroy(psst_mlp *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *q : m->allocs) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

}  // extern "C"
