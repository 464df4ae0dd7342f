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

// GELU, tanh form: 0.5 z (1 + tanh(u)) = z / (1 + e^{-2u}), u = 0.7978845608 (z + 0.044715 z^3),
// with the hardware base-2 exponential and reciprocal (the same function in training and
// inference)
__device__ __forceinline__ float gelu_f(float z) {
    const float u2 = z * (2.3022081983f + 0.1029432617f * z * z);  // 2u log2(e)
    return z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-u2));
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

// ------------------------------------------------------------------------------------------
// Fused inference: the whole network for a tile of 128 rows in one kernel, with gfx950's
// v_mfma_f32_32x32x16_bf16 (A 32 x 16 and B 16 x 32 BF16 operands, 8 values per lane: lane
// l holds row / column l % 32 and k = 8 (l / 32) .. +7; C 32 x 32 FP32, 16 values per lane:
// column l % 32, rows 8 (r / 4) + 4 (l / 32) + r % 4).
//  - Each layer is computed transposed, H'^T = W H^T: the weights are the A operand (16 bytes per
//    lane, stored in fragment order so that a wave's load for one k-step is one contiguous
//    kilobyte, streamed from L2 one k-step ahead of the MFMAs) and the tile's
//    activations the B operand, read from LDS ([128][264] BF16, row = sample; rows padded by 8
//    values so that the 16-byte reads of eight consecutive lanes cover every bank once).
//  - A lane's accumulators then hold, for one sample, runs of four consecutive neurons: the bias
//    + GELU epilogue converts each run to BF16 with the hardware's paired conversion and writes
//    it back to the sample's LDS row with one 8-byte store, between two barriers.
//  - Wave w computes neurons 64 w .. 64 w + 63 for all 128 samples (2 x 4 tiles of 32 x 32, 128
//    accumulator registers).
//  - The output head (256 -> 6) keeps FP32 precision with three BF16 products: the last hidden
//    activation h = hi + lo and the head's weights W = W_hi + W_lo, each split into BF16 parts,
//    and y = W_hi hi + W_lo hi + W_hi lo (the W_lo lo term is below FP32's resolution of y).
//    hi, then lo, pass through the same LDS buffer.
// ------------------------------------------------------------------------------------------
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

constexpr int MF_ROWS = 128;
constexpr int MF_PITCH = 264;

__device__ __forceinline__ v8bf ld8(const uint16_t* p) { return *(const v8bf*)p; }

__global__ __launch_bounds__(256, 2) void k_mlp_fused(MlpFusedArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t act[MF_ROWS * MF_PITCH];
    __shared__ __attribute__((aligned(16))) float bias_s[4 * MLP_H];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 31, lh = lane >> 5;
    for (int e = t; e < 4 * MLP_H; e += 256) bias_s[e] = a.b[e / MLP_H][e % MLP_H];
    const int64_t n = a.n_dev ? (int64_t)*a.n_dev : a.n;
    const int j0 = wave * 64;  // this wave's neurons
    // tiles of 128 rows, grid-stride (the row count may live on the device)
    for (int64_t row0 = (int64_t)blockIdx.x * MF_ROWS; row0 < n; row0 += (int64_t)gridDim.x * MF_ROWS) {
        __syncthreads();  // bias_s written; the previous tile's last reads of act are done
        // layer-0 input: x padded to 16 columns (row r of the tile is x[idx[r]] when gathered)
        for (int e = t; e < MF_ROWS * 16; e += 256) {
            const int r = e >> 4, k = e & 15;
            const int64_t gr = row0 + r;
            float v = 0.f;
            if (k < SIRIUS_MLP_IN && gr < n) v = a.x[(a.idx ? (int64_t)a.idx[gr] : gr) * a.x_stride + k];
            act[r * MF_PITCH + k] = f2bf(v);
        }
        __syncthreads();
        v16f acc[2][4];
        uint32_t lo_pk[2][4][8];  // low BF16 parts of the last hidden activation, four to two registers
        // the accumulators start at the layer's bias (row 8 q + 4 lh + e of tile mt is neuron
        // j0 + 32 mt + 8 q + 4 lh + e)
        auto init_bias = [&](int l) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 bb = *(const float4*)&bias_s[l * MLP_H + j0 + mt * 32 + 8 * q + 4 * lh];
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt) {
                        acc[mt][nt][4 * q + 0] = bb.x;
                        acc[mt][nt][4 * q + 1] = bb.y;
                        acc[mt][nt][4 * q + 2] = bb.z;
                        acc[mt][nt][4 * q + 3] = bb.w;
                    }
                }
        };
        init_bias(0);
        {
            v8bf w[2];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) w[mt] = ld8(a.w0 + (j0 + mt * 32 + li) * 16 + 8 * lh);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const v8bf hv = ld8(&act[(nt * 32 + li) * MF_PITCH + 8 * lh]);
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[mt], hv, acc[mt][nt], 0, 0, 0);
            }
        }
        for (int l = 0; l < 4; ++l) {
            if (l > 0) {
                const uint16_t* Wl = a.w[l - 1];
                init_bias(l);
                // this lane's fragment of k-step ks, row tile mt: Wp + 1024 ks + 512 mt (one running
                // pointer; the row-tile offset is an immediate)
                const uint16_t* Wp = Wl + wave * 16384 + lane * 8;
                v8bf w[2], wn[2];
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) w[mt] = ld8(Wp + 512 * mt);
#pragma unroll
                for (int ks = 0; ks < MLP_H / 16; ++ks) {
                    const int k0 = ks * 16;
                    if (ks + 1 < MLP_H / 16) {
                        Wp += 1024;
#pragma unroll
                        for (int mt = 0; mt < 2; ++mt) wn[mt] = ld8(Wp + 512 * mt);
                    }
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt) {
                        const v8bf hv = ld8(&act[(nt * 32 + li) * MF_PITCH + k0 + 8 * lh]);
#pragma unroll
                        for (int mt = 0; mt < 2; ++mt)
                            acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[mt], hv, acc[mt][nt], 0, 0, 0);
                    }
                    if (ks + 1 < MLP_H / 16) {
#pragma unroll
                        for (int mt = 0; mt < 2; ++mt) w[mt] = wn[mt];
                    }
                }
            }
            __syncthreads();  // every wave has read this layer's input
            // GELU on runs of four neurons; BF16 (the last layer: its high parts) back into the
            // sample's LDS row, the last layer's low parts kept in registers
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int jb = j0 + mt * 32 + 8 * q + 4 * lh;
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt) {
                        const float h0 = gelu_f(acc[mt][nt][4 * q + 0]), h1 = gelu_f(acc[mt][nt][4 * q + 1]);
                        const float h2 = gelu_f(acc[mt][nt][4 * q + 2]), h3 = gelu_f(acc[mt][nt][4 * q + 3]);
                        const v4bf hv = {(__bf16)h0, (__bf16)h1, (__bf16)h2, (__bf16)h3};
                        *(v4bf*)&act[(nt * 32 + li) * MF_PITCH + jb] = hv;
                        if (l == 3) {
                            const v4bf lv = {(__bf16)(h0 - (float)hv[0]), (__bf16)(h1 - (float)hv[1]),
                                             (__bf16)(h2 - (float)hv[2]), (__bf16)(h3 - (float)hv[3])};
                            uint32_t u[2];
                            __builtin_memcpy(u, &lv, 8);
                            lo_pk[mt][nt][2 * q] = u[0];
                            lo_pk[mt][nt][2 * q + 1] = u[1];
                        }
                    }
                }
            __syncthreads();
        }
        // output head, samples 32 w .. 32 w + 31 of the tile: W_hi hi + W_lo hi, then W_hi lo. Rows
        // 6..31 of the head's weights are zero: the lanes holding them all read row 6, so a head
        // load fetches the six live rows and one more line.
        const uint16_t* Hh = a.w4h + (li < SIRIUS_MLP_OUT ? li : SIRIUS_MLP_OUT) * MLP_H + 8 * lh;
        const uint16_t* Hl = a.w4l + (li < SIRIUS_MLP_OUT ? li : SIRIUS_MLP_OUT) * MLP_H + 8 * lh;
        v16f y;
#pragma unroll
        for (int r = 0; r < 16; ++r) y[r] = 0.f;
#pragma unroll
        for (int ks = 0; ks < MLP_H / 16; ++ks) {
            const int k0 = ks * 16;
            const v8bf hv = ld8(&act[(wave * 32 + li) * MF_PITCH + k0 + 8 * lh]);
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(Hh + k0), hv, y, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(Hl + k0), hv, y, 0, 0, 0);
        }
        __syncthreads();  // every wave has read the high parts
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int jb = j0 + mt * 32 + 8 * q + 4 * lh;
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    const uint32_t u[2] = {lo_pk[mt][nt][2 * q], lo_pk[mt][nt][2 * q + 1]};
                    __builtin_memcpy(&act[(nt * 32 + li) * MF_PITCH + jb], u, 8);
                }
            }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < MLP_H / 16; ++ks) {
            const int k0 = ks * 16;
            const v8bf hv = ld8(&act[(wave * 32 + li) * MF_PITCH + k0 + 8 * lh]);
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(Hh + k0), hv, y, 0, 0, 0);
        }
        // y^T: this lane's sample is 32 w + li, its outputs are rows 4 lh + r (r < 4) of the first
        // 8-row group
        const int64_t gr = row0 + wave * 32 + li;
        if (gr < n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 4 * lh + r;
                if (o < SIRIUS_MLP_OUT) a.y[gr * SIRIUS_MLP_OUT + o] = y[r] + a.b4[o];
            }
        }
    }
}

// Grid: one block per tile up to `max_blocks` (two blocks fit a CU), each looping over tiles.
hipError_t launch_mlp_fused(const MlpFusedArgs& a, int64_t cap, int max_blocks, hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    int64_t blocks = (cap + MF_ROWS - 1) / MF_ROWS;
    if (blocks > max_blocks) blocks = max_blocks;
    hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// The output head in FP32 for training (the fused inference keeps FP32 precision with split
// BF16 products): h4 = gelu(z3) from the stored pre-activations, W4 the FP32 masters.
// ------------------------------------------------------------------------------------------
// y64[i][o] = b4[o] + sum_k h4[i][k] W4[o][k] (o < 6; columns 6..63 zero). One wave per row
// group: each lane takes four k, the six partial sums are reduced across the wave.
__global__ __launch_bounds__(256) void k_head_fwd(const float* __restrict__ z3, const float* __restrict__ W4,
                                                  const float* __restrict__ b4, int64_t n_pad, float* __restrict__ y64) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n_pad) return;
    float part[SIRIUS_MLP_OUT] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = lane * 4 + q;
        const float h = gelu_f(z3[i * MLP_H + k]);
#pragma unroll
        for (int o = 0; o < SIRIUS_MLP_OUT; ++o) part[o] = fmaf(h, W4[o * MLP_H + k], part[o]);
    }
#pragma unroll
    for (int o = 0; o < SIRIUS_MLP_OUT; ++o)
        for (int off = 32; off > 0; off >>= 1) part[o] += __shfl_xor(part[o], off);
    if (lane < MLP_N4) {
        float v = 0.f;
#pragma unroll
        for (int o = 0; o < SIRIUS_MLP_OUT; ++o) v = lane == o ? part[o] + b4[o] : v;
        y64[i * MLP_N4 + lane] = v;
    }
}

// gz3[i][k] = (sum_o gy[i][o] W4[o][k]) gelu'(z3[i][k]), in FP32 and BF16.
__global__ void k_head_bwd(const float* __restrict__ gy64, const float* __restrict__ W4, const float* __restrict__ z3,
                           int64_t n_pad, float* __restrict__ gzf, uint16_t* __restrict__ gzb) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_pad * MLP_H) return;
    const int64_t i = e / MLP_H;
    const int k = (int)(e % MLP_H);
    float dh = 0.f;
#pragma unroll
    for (int o = 0; o < SIRIUS_MLP_OUT; ++o) dh = fmaf(gy64[i * MLP_N4 + o], W4[o * MLP_H + k], dh);
    const float g = dh * gelu_df(z3[e]);
    gzf[e] = g;
    gzb[e] = f2bf(g);
}

// dW4[o][k] = sum_i gy[i][o] gelu(z3[i][k]): row chunks into part[split][6][256], then a fixed-
// order sum over the chunks (deterministic).
__global__ __launch_bounds__(256) void k_head_wgrad_part(const float* __restrict__ gy64, const float* __restrict__ z3,
                                                         int64_t n_pad, int64_t chunk, float* __restrict__ part) {
    const int k = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * chunk, r1 = min(n_pad, r0 + chunk);
    float acc[SIRIUS_MLP_OUT] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int64_t i = r0; i < r1; ++i) {
        const float h = gelu_f(z3[i * MLP_H + k]);
#pragma unroll
        for (int o = 0; o < SIRIUS_MLP_OUT; ++o) acc[o] = fmaf(gy64[i * MLP_N4 + o], h, acc[o]);
    }
#pragma unroll
    for (int o = 0; o < SIRIUS_MLP_OUT; ++o) part[((int64_t)blockIdx.x * SIRIUS_MLP_OUT + o) * MLP_H + k] = acc[o];
}
__global__ void k_head_wgrad_final(const float* __restrict__ part, int nsplit, float* __restrict__ dW4) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= SIRIUS_MLP_OUT * MLP_H) return;
    float s = 0.f;
    for (int q = 0; q < nsplit; ++q) s += part[(int64_t)q * SIRIUS_MLP_OUT * MLP_H + e];
    dW4[e] = s;
}

hipError_t launch_head_fwd(const float* z3, const float* W4, const float* b4, int64_t n_pad, float* y64,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_head_fwd, dim3((unsigned)((n_pad + 3) / 4)), dim3(256), 0, st, z3, W4, b4, n_pad, y64);
    return hipGetLastError();
}
hipError_t launch_head_bwd(const float* gy64, const float* W4, const float* z3, int64_t n_pad, float* gzf,
                           uint16_t* gzb, hipStream_t st) {
    const int64_t e = n_pad * MLP_H;
    hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, st, gy64, W4, z3, n_pad, gzf, gzb);
    return hipGetLastError();
}
hipError_t launch_head_wgrad(const float* gy64, const float* z3, int64_t n_pad, float* part, int nsplit, float* dW4,
                             hipStream_t st) {
    const int64_t chunk = (n_pad + nsplit - 1) / nsplit;
    hipLaunchKernelGGL(k_head_wgrad_part, dim3((unsigned)nsplit), dim3(MLP_H), 0, st, gy64, z3, n_pad, chunk, part);
    hipLaunchKernelGGL(k_head_wgrad_final, dim3((SIRIUS_MLP_OUT * MLP_H + 255) / 256), dim3(256), 0, st, part, nsplit,
                       dW4);
    return hipGetLastError();
}

// Operands of the fused kernel from the FP32 masters: W0 padded to 16 input columns, W1..W3 as
// they are, the head's weights split into BF16 high and low parts (rows 6..31 zero).
// Fused-inference operands from the FP32 masters: W0 [256][16] (input columns padded), and the
// hidden layers' and the head's weights in MFMA fragment order: the 16-byte A operand of lane l
// for k-step ks, row tile mt of wave w sits at (((w 16 + ks) 2 + mt) 64 + l) 8 (hidden layer:
// row 64 w + 32 mt + l % 32, columns 16 ks + 8 (l / 32) .. +7), so each wave's weight load is one
// contiguous kilobyte; the head's high and low BF16 parts row-major [32][256] (rows 6..31 zero).
__global__ void k_pack_fused(const float* __restrict__ P, uint16_t* __restrict__ w0, uint16_t* __restrict__ wh,
                             uint16_t* __restrict__ w4h, uint16_t* __restrict__ w4l, size_t off_w0, size_t off_w1,
                             size_t off_w2, size_t off_w3, size_t off_w4) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < MLP_H * 16) {
        const int j = e / 16, k = e % 16;
        w0[e] = k < SIRIUS_MLP_IN ? f2bf(P[off_w0 + (size_t)j * SIRIUS_MLP_IN + k]) : (uint16_t)0;
    }
    if (e < 3 * MLP_H * MLP_H) {
        const int l = e / (MLP_H * MLP_H), r = e % (MLP_H * MLP_H);
        const int i = r & 7, lane = (r >> 3) & 63, mt = (r >> 9) & 1, ks = (r >> 10) & 15, w = r >> 14;
        const int j = 64 * w + 32 * mt + (lane & 31), k = 16 * ks + 8 * (lane >> 5) + i;
        const size_t off = l == 0 ? off_w1 : (l == 1 ? off_w2 : off_w3);
        wh[e] = f2bf(P[off + (size_t)j * MLP_H + k]);
    }
    if (e < 32 * MLP_H) {
        const int j = e / MLP_H, k = e % MLP_H;
        const float w = j < SIRIUS_MLP_OUT ? P[off_w4 + (size_t)j * MLP_H + k] : 0.f;
        const uint16_t hi = f2bf(w);
        w4h[e] = hi;
        w4l[e] = f2bf(w - __uint_as_float((uint32_t)hi << 16));
    }
}
hipError_t launch_pack_fused(const float* params, uint16_t* w0, uint16_t* wh, uint16_t* w4h, uint16_t* w4l,
                             const size_t off_w[5], hipStream_t st) {
    hipLaunchKernelGGL(k_pack_fused, dim3((3 * MLP_H * MLP_H + 255) / 256), dim3(256), 0, st, params, w0, wh, w4h, w4l,
                       off_w[0], off_w[1], off_w[2], off_w[3], off_w[4]);
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
