// libnfx: MLP surrogate kernels (BF16 operands on MFMA, FP32 accumulation) -- forward, backward
// and Adam for a fully connected GELU network.  No reference counterpart: the surrogate maps cell
// and indentation parameters to the force-indentation curve for parameter fitting.
//
// Every GEMM is C[M x N] = A[M x K] * B[N x K]^T with A and B row-major BF16 (K contiguous), the
// operand layout v_mfma_f32_16x16x32_bf16 reads directly: lane l holds 8 consecutive k of row
// (l % 16) of A and of row (l % 16) of B at k offset 8 * (l / 16); the 16 x 16 FP32 result tile
// leaves row 4 * (l / 16) + r, column l % 16 in accumulator element r.  A wavefront computes a
// 32 x 32 tile (2 x 2 MFMA tiles); a workgroup of four wavefronts stacked along M a 128 x 32 tile.
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

template <int EPI>
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
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        const bf16x8 va0 = *reinterpret_cast<const bf16x8*>(a0 + k0);
        const bf16x8 va1 = *reinterpret_cast<const bf16x8*>(a1 + k0);
        const bf16x8 vb0 = *reinterpret_cast<const bf16x8*>(b0 + k0);
        const bf16x8 vb1 = *reinterpret_cast<const bf16x8*>(b1 + k0);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb1, acc[1][1], 0, 0, 0);
    }
    const int M = g.M, N = g.N;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = tm + mi * 16 + 4 * (lane >> 4) + r;
                const int col = tn + ni * 16 + r16;
                const float v = acc[mi][ni][r];
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
}

// loss = sum_ij w_i (Y - T)^2 / (M N) (w_i = 1 without row weights); dZ = 2 w_i (Y - T) / (M N) as
// BF16 in [M x N] and [N x M]
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ Tg,
                                             const float* __restrict__ w, int32_t M, int32_t N, __bf16* dZ,
                                             __bf16* dZT, double* loss_sum) {
    const int64_t n = (int64_t)M * N;
    const float sc = 2.0f / (float)n;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int row = (int)(i / N), col = (int)(i % N);
        const float wr = w ? w[row] : 1.0f;
        const float d = Y[i] - Tg[i];
        s += (double)wr * d * d;
        const __bf16 b = to_bf16(sc * wr * d);
        dZ[i] = b;
        dZT[(int64_t)col * M + row] = b;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss_sum, (ws[0] + ws[1] + ws[2] + ws[3]) / (double)n);
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
    const dim3 grid(g->N / 32, g->M / 128);
    switch (g->epilogue) {
        case NFX_MLP_STORE: k_gemm_nt<NFX_MLP_STORE><<<grid, 256, 0, c->stream>>>(*g); break;
        case NFX_MLP_FWD_GELU: k_gemm_nt<NFX_MLP_FWD_GELU><<<grid, 256, 0, c->stream>>>(*g); break;
        case NFX_MLP_BWD_GELU: k_gemm_nt<NFX_MLP_BWD_GELU><<<grid, 256, 0, c->stream>>>(*g); break;
        default: k_gemm_nt<NFX_MLP_ADAM><<<grid, 256, 0, c->stream>>>(*g); break;
    }
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_mlp_mse(nfx_ctx* c, const float* Y, const float* target, const float* row_weight, int32_t M, int32_t N,
                void* dZ, void* dZT, double* loss_sum) {
    NFX_REQUIRE(c && Y && target && dZ && dZT && loss_sum && M > 0 && N > 0, "bad argument");
    const int64_t n = (int64_t)M * N;
    const unsigned int nb = (unsigned int)std::min<int64_t>((n + 255) / 256, 4096);
    k_mse<<<nb, 256, 0, c->stream>>>(Y, target, row_weight, M, N, (__bf16*)dZ, (__bf16*)dZT, loss_sum);
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
