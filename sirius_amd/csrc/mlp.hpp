// mlp.hpp — the far-field surrogate network on the device: BF16 MFMA GEMMs with FP32
// accumulation and fused epilogues, the loss, Adam, and the surrogate's per-ray kernels.
//
// Shapes are padded so every GEMM tile is whole: rows (samples) to a multiple of 64, output
// columns to a multiple of 64, the reduction dimension to a multiple of 32. Padded weight
// rows/columns and padded activations are zero, so padding never changes a result.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sirius_hip.h"

namespace sirius {

constexpr int MLP_NL = 5;                      // weight layers
constexpr int MLP_H = SIRIUS_MLP_HIDDEN;       // 256
constexpr int MLP_K0 = 32;                     // padded input width of layer 0
constexpr int MLP_X0 = 64;                     // padded width of the stored input rows
constexpr int MLP_N4 = 64;                     // padded output width of the last layer

// Layer dims (unpadded) and padded GEMM shapes
constexpr int MLP_IN[MLP_NL] = {SIRIUS_MLP_IN, MLP_H, MLP_H, MLP_H, MLP_H};
constexpr int MLP_OUTD[MLP_NL] = {MLP_H, MLP_H, MLP_H, MLP_H, SIRIUS_MLP_OUT};
constexpr int MLP_KP[MLP_NL] = {MLP_K0, MLP_H, MLP_H, MLP_H, MLP_H};
constexpr int MLP_NP[MLP_NL] = {MLP_H, MLP_H, MLP_H, MLP_H, MLP_N4};

size_t mlp_param_count();
size_t mlp_param_offset_w(int l);  // offset of W_l in the public layout
size_t mlp_param_offset_b(int l);

// Generic GEMM C(M x N) = A(M x K) * B(K x N) over BF16 operands, FP32 accumulation.
// A(i, k) = A[i * sai + k * sak], B(k, j) = B[k * sbk + j * sbj]; exactly one stride of each
// operand is 1 (the kernel stages tiles through LDS with 16-byte loads along it).
enum { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_BWD = 2, EPI_PART = 3 };
struct GemmArgs {
    const uint16_t* A;
    int64_t sai, sak;
    const uint16_t* B;
    int64_t sbk, sbj;
    int M, N, K, kchunk;
    const float* bias;   // EPI_HIDDEN / EPI_OUT
    float* zout;         // EPI_HIDDEN: pre-activation (optional)
    uint16_t* hout;      // EPI_HIDDEN: bf16(gelu(z))
    const float* zin;    // EPI_BWD: pre-activation of the layer below
    float* gzf;          // EPI_BWD: gradient w.r.t. that pre-activation (FP32)
    uint16_t* gzb;       // EPI_BWD: the same in BF16
    float* out;          // EPI_OUT: outputs; EPI_PART: split-K partials [split][M][N]
};
hipError_t launch_gemm(int epi, const GemmArgs& g, int nsplit, hipStream_t st);

hipError_t launch_mlp_input(const float* x, int64_t n, int64_t n_pad, uint16_t* h0, hipStream_t st);
hipError_t launch_mlp_extract(const float* y64, int64_t n, float* y, hipStream_t st);
hipError_t launch_mlp_loss(const float* y64, const int32_t* cls, const float* dir, int64_t n, int64_t n_pad,
                           float* gyf, uint16_t* gyb, float* loss_i, hipStream_t st);
hipError_t launch_colsum(const float* src, int64_t rows, int cols, int ld, float* part, int nsplit, float* dst,
                         hipStream_t st);
hipError_t launch_reduce_partials(const float* part, int nsplit, int M, int N, int rows, int cols, float* dst,
                                  hipStream_t st);
hipError_t launch_sum(const float* src, int64_t n, float* part, float* dst, hipStream_t st);
hipError_t launch_adam(float* params, const float* grad, float* m, float* v, int64_t P, float lr, float c1, float c2,
                       hipStream_t st);
hipError_t launch_pack(const float* params, uint16_t* const* w, float* const* b, hipStream_t st);

// Fused inference (k_mlp_fused): the whole network over rows of x in one kernel.
struct MlpFusedArgs {
    const float* x;        // inputs: row r at x[r * x_stride] (x[idx[r] * x_stride] when idx is set)
    int64_t n;             // rows (when n_dev is null)
    const uint32_t* n_dev; // rows, read on the device (null: n)
    const int32_t* idx;    // optional gather list
    int64_t x_stride;      // floats per input record (5 for packed inputs)
    const uint16_t* w0;    // [256][16] BF16, input columns padded to 16
    const uint16_t* w[3];  // layers 1..3, [256][256] BF16 in MFMA fragment order (k_pack_fused)
    const float* b[4];     // biases of layers 0..3, [256]
    const uint16_t* w4h;   // head weights [32][256], BF16 high parts (rows 6..31 zero)
    const uint16_t* w4l;   // and low parts
    const float* b4;       // head bias [6]
    float* y;              // n x 6 outputs
};
hipError_t launch_mlp_fused(const MlpFusedArgs& a, int64_t cap, int max_blocks, hipStream_t st);

// Surrogate records (one per candidate ray, by ray id): network inputs (5), signed distance from
// the critical curve, the reference orbit's impact parameter, padding, the reference exit
// direction (3), padding. Calibration bins of the error gate, and the held-out rays a bin needs.
constexpr int SUR_REC = 12;
// The reference orbit of a ray starts at x0 + SUR_SPIN_SHIFT a e_phi (e_phi: the observer's
// azimuthal direction): the lens displaced as frame dragging displaces it to first order in a.
constexpr double SUR_SPIN_SHIFT = 1.1;
constexpr int SUR_BINS = 270;
constexpr uint32_t SUR_CAL_MIN = 20;
// FP32 output head of the training passes (z3: the last hidden pre-activations [n_pad][256])
hipError_t launch_head_fwd(const float* z3, const float* W4, const float* b4, int64_t n_pad, float* y64,
                           hipStream_t st);
hipError_t launch_head_bwd(const float* gy64, const float* W4, const float* z3, int64_t n_pad, float* gzf,
                           uint16_t* gzb, hipStream_t st);
hipError_t launch_head_wgrad(const float* gy64, const float* z3, int64_t n_pad, float* part, int nsplit, float* dW4,
                             hipStream_t st);
// off_w: parameter offsets of W_0 .. W_4 (mlp_param_offset_w)
hipError_t launch_pack_fused(const float* params, uint16_t* w0, uint16_t* wh, uint16_t* w4h, uint16_t* w4l,
                             const size_t off_w[5], hipStream_t st);

// Device state of the network of one handle.
struct MlpState {
    float* params = nullptr;  // FP32 masters, public layout
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;
    int64_t step = 0;
    uint16_t* w[MLP_NL] = {};  // padded BF16 weights, W_l[NP][KP]
    float* b[MLP_NL] = {};     // padded biases [NP]
    uint16_t* fw0 = nullptr;   // fused inference: W0 [256][16], hidden layers 1..3 (3 x 256 x 256,
    uint16_t* fwh = nullptr;   // fragment order), head W_hi / W_lo [32][256]
    uint16_t* fw4h = nullptr;
    uint16_t* fw4l = nullptr;
    // calibration of the error gate (SUR_BINS each): held-out rays the network would answer,
    // misclassified ones, and the largest direction error (rad); all zero until calibrated, so
    // an uncalibrated network answers nothing
    float* cal_err = nullptr;
    uint32_t* cal_n = nullptr;
    uint32_t* cal_bad = nullptr;
    // the largest distance from the critical curve among a bin's held-out answers: the bin
    // answers no ray beyond it (the calibration holds no evidence there)
    float* cal_dmax = nullptr;
    // the training cameras' field of view: the largest |x| and |y| of a pixel centre's direction
    // on the image plane at unit distance. Rays outside this box lie outside the training set's
    // support in (alpha, beta) for their observer and are never answered.
    float sup_x = 0.f, sup_y = 0.f;
    bool ready = false;
};

}  // namespace sirius
