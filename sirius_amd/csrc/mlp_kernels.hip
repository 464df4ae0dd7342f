// mlp_kernels.hip — BF16 MFMA kernels of the far-field surrogate network (gfx950).
//
// One GEMM kernel serves every layer, forward and backward: a 256-thread block computes a
// 64 x 64 tile of C as four 32 x 32 wave tiles with v_mfma_f32_32x32x8_bf16 (wave64; A and B
// fragments are 4 BF16 values per lane, C is 16 FP32 values per lane). Tiles of A (64 x 32)
// and B (32 x 64) are staged through LDS with 16-byte global loads along each operand's
// contiguous dimension, so transposed operands (the backward pass) cost no extra pass over
// memory. Epilogues fuse bias + GELU + BF16 rounding (forward), the GELU derivative
// (backward), or write split-K partials (weight gradients, reduced deterministically).
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace sirius {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (NaN/Inf pass through)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu_f(float z) {
    float t = tanhf(0.7978845608028654f * (z + 0.044715f * z * z * z));
    return 0.5f * z * (1.f + t);
}
__device__ __forceinline__ float gelu_df(float z) {
    float t = tanhf(0.7978845608028654f * (z + 0.044715f * z * z * z));
    return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * 0.7978845608028654f * (1.f + 3.f * 0.044715f * z * z);
}

constexpr int LDS_PITCH = 40;  // halves per LDS row: 32 + 8 pad keeps 16-byte alignment (80 B)

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[64 * LDS_PITCH];  // As[i][k]
    __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * LDS_PITCH];  // Bs[j][k]
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int kbeg = blockIdx.z * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    v16f acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        if (g.sak == 1) {  // A contiguous along k: row t/4, k 8*(t%4) .. +7
            const int r = t >> 2, kk = (t & 3) * 8;
            *(uint4*)&As[r * LDS_PITCH + kk] = *(const uint4*)(g.A + (int64_t)(row0 + r) * g.sai + k0 + kk);
        } else {  // A contiguous along i: k t/8, rows 8*(t%8) .. +7
            const int kk = t >> 3, r = (t & 7) * 8;
            uint4 v = *(const uint4*)(g.A + (int64_t)(k0 + kk) * g.sak + row0 + r);
            const uint16_t* e = (const uint16_t*)&v;
#pragma unroll
            for (int q = 0; q < 8; ++q) As[(r + q) * LDS_PITCH + kk] = e[q];
        }
        if (g.sbk == 1) {  // B contiguous along k: column t/4, k 8*(t%4) .. +7
            const int j = t >> 2, kk = (t & 3) * 8;
            *(uint4*)&Bs[j * LDS_PITCH + kk] = *(const uint4*)(g.B + (int64_t)(col0 + j) * g.sbj + k0 + kk);
        } else {  // B contiguous along j: k t/8, columns 8*(t%8) .. +7
            const int kk = t >> 3, j = (t & 7) * 8;
            uint4 v = *(const uint4*)(g.B + (int64_t)(k0 + kk) * g.sbk + col0 + j);
            const uint16_t* e = (const uint16_t*)&v;
#pragma unroll
            for (int q = 0; q < 8; ++q) Bs[(j + q) * LDS_PITCH + kk] = e[q];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 32; kk += 8) {
            const int ko = kk + 4 * (lane >> 5);
            v4s a = *(const v4s*)&As[(wr + (lane & 31)) * LDS_PITCH + ko];
            v4s b = *(const v4s*)&Bs[(wc + (lane & 31)) * LDS_PITCH + ko];
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // C fragment: lane holds column wc + lane % 32, rows 8 (r / 4) + 4 (lane / 32) + r % 4
    const int j = col0 + wc + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = row0 + wr + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t o = (int64_t)i * g.N + j;
        const float v = acc[r];
        if constexpr (EPI == EPI_HIDDEN) {
            const float z = v + g.bias[j];
            if (g.zout) g.zout[o] = z;
            g.hout[o] = f2bf(gelu_f(z));
        } else if constexpr (EPI == EPI_OUT) {
            g.out[o] = v + g.bias[j];
        } else if constexpr (EPI == EPI_BWD) {
            const float gz = v * gelu_df(g.zin[o]);
            g.gzf[o] = gz;
            g.gzb[o] = f2bf(gz);
        } else {
            g.out[(int64_t)blockIdx.z * g.M * g.N + o] = v;
        }
    }
}

hipError_t launch_gemm(int epi, const GemmArgs& g, int nsplit, hipStream_t st) {
    dim3 grid((unsigned)(g.N / 64), (unsigned)(g.M / 64), (unsigned)nsplit), block(256);
    switch (epi) {
        case EPI_HIDDEN: hipLaunchKernelGGL(k_gemm<EPI_HIDDEN>, grid, block, 0, st, g); break;
        case EPI_OUT: hipLaunchKernelGGL(k_gemm<EPI_OUT>, grid, block, 0, st, g); break;
        case EPI_BWD: hipLaunchKernelGGL(k_gemm<EPI_BWD>, grid, block, 0, st, g); break;
        default: hipLaunchKernelGGL(k_gemm<EPI_PART>, grid, block, 0, st, g); break;
    }
    return hipGetLastError();
}

// Input rows: h0[i][0..4] = bf16(x[i][0..4]), the rest of the padded 64-wide row zero; rows
// n .. n_pad - 1 are zero.
__global__ void k_mlp_input(const float* __restrict__ x, int64_t n, int64_t n_pad, uint16_t* __restrict__ h0) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_pad * MLP_X0) return;
    int64_t i = e / MLP_X0;
    int k = (int)(e % MLP_X0);
    h0[e] = (i < n && k < SIRIUS_MLP_IN) ? f2bf(x[i * SIRIUS_MLP_IN + k]) : (uint16_t)0;
}

__global__ void k_mlp_extract(const float* __restrict__ y64, int64_t n, float* __restrict__ y) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * SIRIUS_MLP_OUT) return;
    int64_t i = e / SIRIUS_MLP_OUT;
    y[e] = y64[i * MLP_N4 + (e % SIRIUS_MLP_OUT)];
}

// Per-sample loss and the output gradient (1/n of the batch mean included), padded to 64
// columns; rows past n are zero.
__global__ void k_mlp_loss(const float* __restrict__ y64, const int32_t* __restrict__ cls,
                           const float* __restrict__ dir, int64_t n, int64_t n_pad, float* __restrict__ gyf,
                           uint16_t* __restrict__ gyb, float* __restrict__ loss_i) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    float g[MLP_N4];
#pragma unroll
    for (int o = 0; o < MLP_N4; ++o) g[o] = 0.f;
    float loss = 0.f;
    if (i < n) {
        const float* y = y64 + i * MLP_N4;
        int c = cls[i];
        c = c < 0 ? 0 : (c > 2 ? 2 : c);
        float m = fmaxf(y[0], fmaxf(y[1], y[2]));
        float e0 = expf(y[0] - m), e1 = expf(y[1] - m), e2 = expf(y[2] - m);
        float s = e0 + e1 + e2;
        float yc = c == 0 ? y[0] : (c == 1 ? y[1] : y[2]);
        loss = -(yc - m - logf(s));
        const float inv_n = 1.f / (float)n;
        g[0] = (e0 / s - (c == 0)) * inv_n;
        g[1] = (e1 / s - (c == 1)) * inv_n;
        g[2] = (e2 / s - (c == 2)) * inv_n;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float d = c == 0 ? y[3 + k] - dir[i * 3 + k] : 0.f;
            loss += (float)SIRIUS_MLP_DIR_WEIGHT * d * d;
            g[3 + k] = 2.f * (float)SIRIUS_MLP_DIR_WEIGHT * d * inv_n;
        }
    }
#pragma unroll
    for (int o = 0; o < MLP_N4; ++o) {
        gyf[i * MLP_N4 + o] = g[o];
        gyb[i * MLP_N4 + o] = f2bf(g[o]);
    }
    loss_i[i] = loss;
}

// Column sums of src[rows][ld] over the first `cols` columns: nsplit row chunks into
// part[nsplit][cols], then dst[c] = sum of the chunks (fixed order: deterministic).
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ src, int64_t rows, int cols, int ld,
                                                     int64_t chunk, float* __restrict__ part) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
    float s = 0.f;
    if (c < cols)
        for (int64_t r = r0 + q; r < r1; r += 4) s += src[r * ld + c];
    red[q][threadIdx.x & 63] = s;
    __syncthreads();
    if (q == 0 && c < cols)
        part[(int64_t)blockIdx.y * cols + c] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                               red[3][threadIdx.x];
}
__global__ void k_colsum_final(const float* __restrict__ part, int nsplit, int cols, float* __restrict__ dst) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int k = 0; k < nsplit; ++k) s += part[(int64_t)k * cols + c];
    dst[c] = s;
}

hipError_t launch_colsum(const float* src, int64_t rows, int cols, int ld, float* part, int nsplit, float* dst,
                         hipStream_t st) {
    int64_t chunk = (rows + nsplit - 1) / nsplit;
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)((cols + 63) / 64), (unsigned)nsplit), dim3(256), 0, st, src,
                       rows, cols, ld, chunk, part);
    hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, st, part, nsplit, cols,
                       dst);
    return hipGetLastError();
}

// dst[r][c] (rows x cols, row-major) = sum over splits of part[s][r][c] (part is M x N).
__global__ void k_reduce_partials(const float* __restrict__ part, int nsplit, int M, int N, int rows, int cols,
                                  float* __restrict__ dst) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    int r = e / cols, c = e % cols;
    float s = 0.f;
    for (int k = 0; k < nsplit; ++k) s += part[(int64_t)k * M * N + (int64_t)r * N + c];
    dst[e] = s;
}
hipError_t launch_reduce_partials(const float* part, int nsplit, int M, int N, int rows, int cols, float* dst,
                                  hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)((rows * cols + 255) / 256)), dim3(256), 0, st, part, nsplit,
                       M, N, rows, cols, dst);
    return hipGetLastError();
}

// Sum of n floats: 256 block partials, then one block (deterministic order).
__global__ __launch_bounds__(256) void k_sum_part(const float* __restrict__ src, int64_t n, float* __restrict__ part) {
    __shared__ float red[256];
    float s = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += src[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
hipError_t launch_sum(const float* src, int64_t n, float* part, float* dst, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_part, dim3(256), dim3(256), 0, st, src, n, part);
    hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(256), 0, st, part, (int64_t)256, dst);
    return hipGetLastError();
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8); c1 = 1 - beta1^t, c2 = 1 - beta2^t.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t P, float lr, float c1, float c2) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P) return;
    float gi = g[e];
    float mi = 0.9f * m[e] + 0.1f * gi;
    float vi = 0.999f * v[e] + 0.001f * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}
hipError_t launch_adam(float* params, const float* grad, float* m, float* v, int64_t P, float lr, float c1, float c2,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, params, grad, m, v, P, lr, c1, c2);
    return hipGetLastError();
}

// Padded BF16 weights and padded biases of one layer from the FP32 masters.
__global__ void k_pack_layer(const float* __restrict__ W, const float* __restrict__ b, int out, int in, int NP, int KP,
                             uint16_t* __restrict__ w, float* __restrict__ bp) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NP * KP) return;
    int j = e / KP, k = e % KP;
    w[e] = (j < out && k < in) ? f2bf(W[j * in + k]) : (uint16_t)0;
    if (k == 0) bp[j] = j < out ? b[j] : 0.f;
}
hipError_t launch_pack(const float* params, uint16_t* const* w, float* const* b, hipStream_t st) {
    for (int l = 0; l < MLP_NL; ++l) {
        int n = MLP_NP[l] * MLP_KP[l];
        hipLaunchKernelGGL(k_pack_layer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           params + mlp_param_offset_w(l), params + mlp_param_offset_b(l), MLP_OUTD[l], MLP_IN[l],
                           MLP_NP[l], MLP_KP[l], w[l], b[l]);
    }
    return hipGetLastError();
}

hipError_t launch_mlp_input(const float* x, int64_t n, int64_t n_pad, uint16_t* h0, hipStream_t st) {
    int64_t e = n_pad * MLP_X0;
    hipLaunchKernelGGL(k_mlp_input, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, st, x, n, n_pad, h0);
    return hipGetLastError();
}
hipError_t launch_mlp_extract(const float* y64, int64_t n, float* y, hipStream_t st) {
    int64_t e = n * SIRIUS_MLP_OUT;
    hipLaunchKernelGGL(k_mlp_extract, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, st, y64, n, y);
    return hipGetLastError();
}
hipError_t launch_mlp_loss(const float* y64, const int32_t* cls, const float* dir, int64_t n, int64_t n_pad,
                           float* gyf, uint16_t* gyb, float* loss_i, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_loss, dim3((unsigned)((n_pad + 255) / 256)), dim3(256), 0, st, y64, cls, dir, n, n_pad,
                       gyf, gyb, loss_i);
    return hipGetLastError();
}

}  // namespace sirius
