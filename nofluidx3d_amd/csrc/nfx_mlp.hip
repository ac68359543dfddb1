// libnfx: MLP surrogate kernels (BF16 operands on MFMA, FP32 accumulation) -- forward, backward
// and Adam for a fully connected GELU network.  No reference counterpart: the surrogate maps cell
// and indentation parameters to the force-indentation curve for parameter fitting.
//
// Every GEMM is C[M x N] = A[M x K] * B[N x K]^T with A and B row-major BF16 (K contiguous), the
// operand layout v_mfma_f32_16x16x32_bf16 reads directly: lane l holds 8 consecutive k of row
// (l % 16) of A and of row (l % 16) of B at k offset 8 * (l / 16); the 16 x 16 FP32 result tile
// leaves row 4 * (l / 16) + r, column l % 16 in accumulator element r.  A wavefront computes a
// 32 x 32 tile (2 x 2 MFMA tiles); a workgroup of four wavefronts stacked along M a 128 x 32 tile.
// Split-K (split_k > 1, for the weight-gradient GEMMs, whose M x N output is small and K the
// batch): workgroup z of the third grid dimension accumulates K-slice z into an FP32 workspace
// plane; a finish kernel sums the planes in slice order (deterministic) and applies the epilogue.
//
// Epilogues fuse the elementwise work into the GEMM that produces it:
//   STORE     C = acc (+ bias)
//   FWD_GELU  Z = acc + bias (FP32, kept for the backward pass), H = gelu(Z) as BF16 in [M x N] and
//             transposed [N x M] (the transposed copy feeds the weight-gradient GEMM)
//   BWD_GELU  dZ = acc * gelu'(Z) as BF16 in [M x N] and [N x M]
//   ADAM      acc is the weight gradient: Adam update of the FP32 master weights C (eps added to the
//             bias-corrected sqrt(v)), BF16 copies of the weights in [M x N] and [N x M]
#include <hip/hip_runtime.h>

#include <cmath>

#include "nfx_common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace nfx {

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// the epilogue of output element (row, col) with accumulated value v
template <int EPI>
__device__ __forceinline__ void epilogue(const nfx_mlp_gemm_args& g, int row, int col, float v) {
    const int M = g.M, N = g.N;
    const int64_t o = (int64_t)row * N + col, ot = (int64_t)col * M + row;
    if constexpr (EPI == NFX_MLP_STORE) {
        g.C[o] = g.bias ? v + g.bias[col] : v;
    } else if constexpr (EPI == NFX_MLP_FWD_GELU) {
        const float z = v + g.bias[col];
        if (g.C) g.C[o] = z;
        const __bf16 h = to_bf16(gelu(z));
        static_cast<__bf16*>(g.H)[o] = h;
        if (g.HT) static_cast<__bf16*>(g.HT)[ot] = h;
    } else if constexpr (EPI == NFX_MLP_BWD_GELU) {
        const __bf16 d = to_bf16(v * gelu_grad(g.Z[o]));
        static_cast<__bf16*>(g.H)[o] = d;
        static_cast<__bf16*>(g.HT)[ot] = d;
    } else {  // ADAM
        const float gr = v * g.grad_scale;
        const float m = g.beta1 * g.adam_m[o] + (1.0f - g.beta1) * gr;
        const float s = g.beta2 * g.adam_v[o] + (1.0f - g.beta2) * gr * gr;
        g.adam_m[o] = m;
        g.adam_v[o] = s;
        const float w = g.C[o] - g.lr_t * m / (sqrtf(s) * g.bc2 + g.eps);
        g.C[o] = w;
        static_cast<__bf16*>(g.H)[o] = to_bf16(w);
        static_cast<__bf16*>(g.HT)[ot] = to_bf16(w);
    }
}

// PARTIAL: store the raw accumulator of K-slice blockIdx.z into workspace plane blockIdx.z
// (the host checks K % (32 * gridDim.z) == 0)
template <int EPI, bool PARTIAL>
__global__ __launch_bounds__(256) void k_gemm_nt(nfx_mlp_gemm_args g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tm = blockIdx.y * 128 + wave * 32;
    const int tn = blockIdx.x * 32;
    const __bf16* __restrict__ A = static_cast<const __bf16*>(g.A);
    const __bf16* __restrict__ B = static_cast<const __bf16*>(g.B);
    const int r16 = lane & 15, kq = 8 * (lane >> 4);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16* a0 = A + (int64_t)(tm + r16) * g.lda + kq;
    const __bf16* a1 = a0 + (int64_t)16 * g.lda;
    const __bf16* b0 = B + (int64_t)(tn + r16) * g.ldb + kq;
    const __bf16* b1 = b0 + (int64_t)16 * g.ldb;
    const int kspan = g.K / (int)gridDim.z;
    const int kb = (int)blockIdx.z * kspan, ke = kb + kspan;
    // operands of the next k-step are loaded before the current step's MFMAs (the last step
    // re-reads its own, in-bounds, address), so one step of loads is always in flight
    bf16x8 va0 = *reinterpret_cast<const bf16x8*>(a0 + kb), va1 = *reinterpret_cast<const bf16x8*>(a1 + kb);
    bf16x8 vb0 = *reinterpret_cast<const bf16x8*>(b0 + kb), vb1 = *reinterpret_cast<const bf16x8*>(b1 + kb);
    for (int k0 = kb; k0 < ke; k0 += 32) {
        const int kn = k0 + 32 < ke ? k0 + 32 : k0;
        const bf16x8 na0 = *reinterpret_cast<const bf16x8*>(a0 + kn);
        const bf16x8 na1 = *reinterpret_cast<const bf16x8*>(a1 + kn);
        const bf16x8 nb0 = *reinterpret_cast<const bf16x8*>(b0 + kn);
        const bf16x8 nb1 = *reinterpret_cast<const bf16x8*>(b1 + kn);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb1, acc[1][1], 0, 0, 0);
        va0 = na0;
        va1 = na1;
        vb0 = nb0;
        vb1 = nb1;
    }
    float* ws = PARTIAL ? g.workspace + (int64_t)blockIdx.z * g.M * g.N : nullptr;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = tm + mi * 16 + 4 * (lane >> 4) + r;
                const int col = tn + ni * 16 + r16;
                if constexpr (PARTIAL) ws[(int64_t)row * g.N + col] = acc[mi][ni][r];
                else epilogue<EPI>(g, row, col, acc[mi][ni][r]);
            }
}

// The same product for grids that fill the chip (M, N multiples of 128; one pass over K): a
// workgroup tile of 128 x 128 staged through LDS, each wave a 64 x 64 sub-tile (4 x 4 MFMA tiles,
// 16 MFMAs per k-step from 8 LDS fragment reads instead of 4 MFMAs per 4 global reads).  Double
// buffered: the next k-slice goes global -> registers during the MFMAs and registers -> the other
// LDS buffer after them, one barrier per k-step.  Rows are padded by 8 BF16 against bank
// conflicts.  Accumulator layout and epilogue as k_gemm_nt.  Four waves per SIMD (the accumulators
// then live in VGPRs, 124 of them; the 40 KB of LDS per workgroup fit four workgroups per CU): 13-30 %
// faster than the three the unbounded allocation gives.
constexpr int kLdsTile = 128, kLdsK = 32, kLdsPitch = kLdsK + 8;
template <int EPI>
__global__ __launch_bounds__(256, 4) void k_gemm_lds(nfx_mlp_gemm_args g) {
    __shared__ __align__(16) __bf16 As[2][kLdsTile * kLdsPitch];
    __shared__ __align__(16) __bf16 Bs[2][kLdsTile * kLdsPitch];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int m0 = blockIdx.y * kLdsTile, n0 = blockIdx.x * kLdsTile;
    const __bf16* __restrict__ A = static_cast<const __bf16*>(g.A);
    const __bf16* __restrict__ B = static_cast<const __bf16*>(g.B);
    const int r16 = lane & 15, kq = 8 * (lane >> 4);
    // copy assignment: 128 rows x 4 chunks of 8 BF16 per operand, two chunks per thread
    int crow[2], ck[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int c = tid + 256 * i;
        crow[i] = c >> 2;
        ck[i] = (c & 3) * 8;
    }
    bf16x8 ra[2], rb[2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            ra[i] = *reinterpret_cast<const bf16x8*>(A + (int64_t)(m0 + crow[i]) * g.lda + k0 + ck[i]);
            rb[i] = *reinterpret_cast<const bf16x8*>(B + (int64_t)(n0 + crow[i]) * g.ldb + k0 + ck[i]);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            *reinterpret_cast<bf16x8*>(&As[buf][crow[i] * kLdsPitch + ck[i]]) = ra[i];
            *reinterpret_cast<bf16x8*>(&Bs[buf][crow[i] * kLdsPitch + ck[i]]) = rb[i];
        }
    };
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    gload(0);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < g.K; k0 += kLdsK) {
        const bool more = k0 + kLdsK < g.K;
        if (more) gload(k0 + kLdsK);
        bf16x8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = *reinterpret_cast<const bf16x8*>(&As[buf][(wm + 16 * i + r16) * kLdsPitch + kq]);
            b[i] = *reinterpret_cast<const bf16x8*>(&Bs[buf][(wn + 16 * i + r16) * kLdsPitch + kq]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (more) sstore(buf ^ 1);  // nobody reads buf ^ 1 now: the last barrier ended its reads
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                epilogue<EPI>(g, m0 + wm + 16 * i + 4 * (lane >> 4) + r, n0 + wn + 16 * j + r16, acc[i][j][r]);
}

// split-K finish: one thread per output element sums the split_k workspace planes in order
template <int EPI>
__global__ __launch_bounds__(256) void k_splitk_finish(nfx_mlp_gemm_args g) {
    const int64_t MN = (int64_t)g.M * g.N;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= MN) return;
    float v = 0.f;
    for (int s = 0; s < g.split_k; ++s) v += g.workspace[(int64_t)s * MN + i];
    epilogue<EPI>(g, (int)(i / g.N), (int)(i % g.N), v);
}

// loss = sum_ij w_i (Y - T)^2 / (M N) (w_i = 1 without row weights); dZ = 2 w_i (Y - T) / (M N) as
// BF16 in [M x N] and [N x M].  One workgroup per 32 x 32 tile (M, N multiples of 32, checked on
// the host): the tile is staged in LDS so that both the row-major and the transposed copy are
// written along contiguous rows.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ Tg,
                                             const float* __restrict__ w, int32_t M, int32_t N, __bf16* dZ,
                                             __bf16* dZT, double* loss_sum) {
    __shared__ __bf16 tile[32][34];
    __shared__ double ws[4];
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    const float sc = 2.0f / ((float)M * (float)N);
    double s = 0.0;
#pragma unroll
    for (int j = ty; j < 32; j += 8) {
        const int row = r0 + j;
        const int64_t i = (int64_t)row * N + c0 + tx;
        const float wr = w ? w[row] : 1.0f;
        const float d = Y[i] - Tg[i];
        s += (double)wr * d * d;
        const __bf16 b = to_bf16(sc * wr * d);
        dZ[i] = b;
        tile[j][tx] = b;
    }
    __syncthreads();
#pragma unroll
    for (int j = ty; j < 32; j += 8) dZT[(int64_t)(c0 + j) * M + r0 + tx] = tile[tx][j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss_sum, (ws[0] + ws[1] + ws[2] + ws[3]) / ((double)M * (double)N));
}

// bias gradient (row sums of dZ^T) and its Adam update; one workgroup per output feature
__global__ __launch_bounds__(256) void k_bias_adam(const __bf16* __restrict__ dZT, int32_t M, float* b, float* m,
                                                   float* v, float lr_t, float bc2, float beta1, float beta2,
                                                   float eps, float grad_scale) {
    const int j = blockIdx.x;
    float s = 0.f;
    for (int i = threadIdx.x; i < M; i += 256) s += (float)dZT[(int64_t)j * M + i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ float ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float gr = (ws[0] + ws[1] + ws[2] + ws[3]) * grad_scale;
        const float mm = beta1 * m[j] + (1.0f - beta1) * gr;
        const float vv = beta2 * v[j] + (1.0f - beta2) * gr * gr;
        m[j] = mm;
        v[j] = vv;
        b[j] -= lr_t * mm / (sqrtf(vv) * bc2 + eps);
    }
}

}  // namespace nfx

using namespace nfx;

extern "C" {

int nfx_mlp_gemm(nfx_ctx* c, const nfx_mlp_gemm_args* g) {
    NFX_REQUIRE(c && g, "NULL argument");
    NFX_REQUIRE(g->M > 0 && g->N > 0 && g->K > 0 && g->M % 128 == 0 && g->N % 32 == 0 && g->K % 32 == 0,
                "mlp gemm: M must be a multiple of 128, N and K of 32 (got %d %d %d)", g->M, g->N, g->K);
    NFX_REQUIRE(g->A && g->B && g->lda >= g->K && g->ldb >= g->K && g->lda % 8 == 0 && g->ldb % 8 == 0,
                "mlp gemm: bad operands");
    switch (g->epilogue) {
        case NFX_MLP_STORE: NFX_REQUIRE(g->C, "STORE needs C"); break;
        case NFX_MLP_FWD_GELU: NFX_REQUIRE(g->bias && g->H, "FWD_GELU needs bias and H"); break;
        case NFX_MLP_BWD_GELU: NFX_REQUIRE(g->Z && g->H && g->HT, "BWD_GELU needs Z, H and HT"); break;
        case NFX_MLP_ADAM: NFX_REQUIRE(g->C && g->H && g->HT && g->adam_m && g->adam_v, "ADAM needs C, H, HT, m, v"); break;
        default: NFX_REQUIRE(false, "unknown epilogue %d", g->epilogue);
    }
    const int S = g->split_k > 1 ? g->split_k : 1;
    NFX_REQUIRE(g->split_k >= 0 && S <= 64 && g->K % (32 * S) == 0,
                "mlp gemm: split_k %d must be at most 64 and divide K / 32 (K %d)", g->split_k, g->K);
    NFX_REQUIRE(S == 1 || g->workspace, "mlp gemm: split_k > 1 needs a workspace of split_k * M * N floats");
    // the LDS-tiled kernel where its 128 x 128 tiles still give a full grid (>= 256 workgroups)
    const bool lds = S == 1 && g->M % kLdsTile == 0 && g->N % kLdsTile == 0 &&
                     (int64_t)(g->M / kLdsTile) * (g->N / kLdsTile) >= 256;
    if (lds) {
        const dim3 grid(g->N / kLdsTile, g->M / kLdsTile);
        switch (g->epilogue) {
            case NFX_MLP_STORE: k_gemm_lds<NFX_MLP_STORE><<<grid, 256, 0, c->stream>>>(*g); break;
            case NFX_MLP_FWD_GELU: k_gemm_lds<NFX_MLP_FWD_GELU><<<grid, 256, 0, c->stream>>>(*g); break;
            case NFX_MLP_BWD_GELU: k_gemm_lds<NFX_MLP_BWD_GELU><<<grid, 256, 0, c->stream>>>(*g); break;
            default: k_gemm_lds<NFX_MLP_ADAM><<<grid, 256, 0, c->stream>>>(*g); break;
        }
    } else if (S == 1) {
        const dim3 grid(g->N / 32, g->M / 128);
        switch (g->epilogue) {
            case NFX_MLP_STORE: k_gemm_nt<NFX_MLP_STORE, false><<<grid, 256, 0, c->stream>>>(*g); break;
            case NFX_MLP_FWD_GELU: k_gemm_nt<NFX_MLP_FWD_GELU, false><<<grid, 256, 0, c->stream>>>(*g); break;
            case NFX_MLP_BWD_GELU: k_gemm_nt<NFX_MLP_BWD_GELU, false><<<grid, 256, 0, c->stream>>>(*g); break;
            default: k_gemm_nt<NFX_MLP_ADAM, false><<<grid, 256, 0, c->stream>>>(*g); break;
        }
    } else {
        k_gemm_nt<NFX_MLP_STORE, true><<<dim3(g->N / 32, g->M / 128, S), 256, 0, c->stream>>>(*g);
        NFX_CHECK_LAUNCH();
        const unsigned int nb = (unsigned int)(((int64_t)g->M * g->N + 255) / 256);
        switch (g->epilogue) {
            case NFX_MLP_STORE: k_splitk_finish<NFX_MLP_STORE><<<nb, 256, 0, c->stream>>>(*g); break;
            case NFX_MLP_FWD_GELU: k_splitk_finish<NFX_MLP_FWD_GELU><<<nb, 256, 0, c->stream>>>(*g); break;
            case NFX_MLP_BWD_GELU: k_splitk_finish<NFX_MLP_BWD_GELU><<<nb, 256, 0, c->stream>>>(*g); break;
            default: k_splitk_finish<NFX_MLP_ADAM><<<nb, 256, 0, c->stream>>>(*g); break;
        }
    }
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_mlp_mse(nfx_ctx* c, const float* Y, const float* target, const float* row_weight, int32_t M, int32_t N,
                void* dZ, void* dZT, double* loss_sum) {
    NFX_REQUIRE(c && Y && target && dZ && dZT && loss_sum, "NULL argument");
    NFX_REQUIRE(M > 0 && N > 0 && M % 32 == 0 && N % 32 == 0, "mlp mse: M and N must be multiples of 32 (got %d %d)",
                M, N);
    k_mse<<<dim3(N / 32, M / 32), 256, 0, c->stream>>>(Y, target, row_weight, M, N, (__bf16*)dZ, (__bf16*)dZT,
                                                       loss_sum);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_mlp_bias_adam(nfx_ctx* c, const void* dZT, int32_t M, int32_t N, float* b, float* m, float* v, float lr_t,
                      float bc2, float beta1, float beta2, float eps, float grad_scale) {
    NFX_REQUIRE(c && dZT && b && m && v && M > 0 && N > 0, "bad argument");
    k_bias_adam<<<N, 256, 0, c->stream>>>((const __bf16*)dZT, M, b, m, v, lr_t, bc2, beta1, beta2, eps, grad_scale);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

}  // extern "C"
