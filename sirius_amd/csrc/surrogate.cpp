// surrogate.cpp — host orchestration of the MLP far-field surrogate: parameters, forward,
// loss and gradient, Adam training steps (the network contract is in sirius_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sirius_hip.h"
#include "mlp.hpp"
#include "sirius_internal.hpp"

namespace sirius {

size_t mlp_param_offset_w(int l) {
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += (size_t)MLP_OUTD[i] * MLP_IN[i] + MLP_OUTD[i];
    return o;
}
size_t mlp_param_offset_b(int l) { return mlp_param_offset_w(l) + (size_t)MLP_OUTD[l] * MLP_IN[l]; }
size_t mlp_param_count() { return mlp_param_offset_w(MLP_NL); }

void mlp_release(Handle& H) {
    if (!H.mlp) return;
    MlpState& S = *H.mlp;
    for (void* p : {(void*)S.params, (void*)S.adam_m, (void*)S.adam_v, (void*)S.grad})
        if (p) (void)hipFree(p);
    for (int l = 0; l < MLP_NL; ++l) {
        if (S.w[l]) (void)hipFree(S.w[l]);
        if (S.b[l]) (void)hipFree(S.b[l]);
    }
    for (void* p : {(void*)S.fw0, (void*)S.fwh, (void*)S.fw4h, (void*)S.fw4l, (void*)S.cal_err, (void*)S.cal_n,
                    (void*)S.cal_bad, (void*)S.cal_dmax})
        if (p) (void)hipFree(p);
    delete H.mlp;
    H.mlp = nullptr;
}

static int mlp_alloc(Handle& H) {
    if (H.mlp) return SIRIUS_HIP_OK;
    int rc = H.ensure_device();
    if (rc) return rc;
    H.mlp = new MlpState();
    MlpState& S = *H.mlp;
    size_t P = mlp_param_count();
    hipError_t e;
    for (float** p : {&S.params, &S.adam_m, &S.adam_v, &S.grad})
        if ((e = hipMalloc((void**)p, P * sizeof(float))) != hipSuccess) return H.hip_fail(e, "mlp alloc");
    for (int l = 0; l < MLP_NL; ++l) {
        if ((e = hipMalloc((void**)&S.w[l], (size_t)MLP_NP[l] * MLP_KP[l] * 2)) != hipSuccess ||
            (e = hipMalloc((void**)&S.b[l], (size_t)MLP_NP[l] * 4)) != hipSuccess)
            return H.hip_fail(e, "mlp alloc");
    }
    if ((e = hipMalloc((void**)&S.fw0, (size_t)MLP_H * 16 * 2)) != hipSuccess ||
        (e = hipMalloc((void**)&S.fwh, (size_t)3 * MLP_H * MLP_H * 2)) != hipSuccess ||
        (e = hipMalloc((void**)&S.fw4h, (size_t)32 * MLP_H * 2)) != hipSuccess ||
        (e = hipMalloc((void**)&S.fw4l, (size_t)32 * MLP_H * 2)) != hipSuccess ||
        (e = hipMalloc((void**)&S.cal_err, SUR_BINS * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&S.cal_n, SUR_BINS * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&S.cal_bad, SUR_BINS * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&S.cal_dmax, SUR_BINS * 4)) != hipSuccess)
        return H.hip_fail(e, "mlp alloc");
    return SIRIUS_HIP_OK;
}

hipError_t launch_fill(uint32_t* p, int n, uint32_t v, hipStream_t st);

// New parameters invalidate the calibration: the error gate answers nothing until the network
// is calibrated again (sirius_hip_surrogate_train calibrates on held-out rays).
static hipError_t mlp_clear_calibration(MlpState& S, hipStream_t st) {
    S.sup_x = S.sup_y = 0.f;
    hipError_t e;
    if ((e = launch_fill((uint32_t*)S.cal_err, SUR_BINS, 0u, st)) != hipSuccess ||
        (e = launch_fill(S.cal_n, SUR_BINS, 0u, st)) != hipSuccess)
        return e;
    if ((e = launch_fill(S.cal_bad, SUR_BINS, 0u, st)) != hipSuccess) return e;
    return launch_fill((uint32_t*)S.cal_dmax, SUR_BINS, 0u, st);
}

// BF16 operands of every kernel from the FP32 masters (after each update).
static hipError_t mlp_pack(MlpState& S, hipStream_t st) {
    hipError_t e = launch_pack(S.params, S.w, S.b, st);
    if (e != hipSuccess) return e;
    const size_t off_w[5] = {mlp_param_offset_w(0), mlp_param_offset_w(1), mlp_param_offset_w(2),
                             mlp_param_offset_w(3), mlp_param_offset_w(4)};
    return launch_pack_fused(S.params, S.fw0, S.fwh, S.fw4h, S.fw4l, off_w, st);
}

static int mlp_reset_adam(Handle& H) {
    MlpState& S = *H.mlp;
    size_t P = mlp_param_count();
    hipError_t e;
    if ((e = hipMemsetAsync(S.adam_m, 0, P * 4, H.stream)) != hipSuccess ||
        (e = hipMemsetAsync(S.adam_v, 0, P * 4, H.stream)) != hipSuccess)
        return H.hip_fail(e, "adam reset");
    S.step = 0;
    if ((e = mlp_pack(S, H.stream)) != hipSuccess || (e = mlp_clear_calibration(S, H.stream)) != hipSuccess)
        return H.hip_fail(e, "pack");
    if ((e = hipStreamSynchronize(H.stream)) != hipSuccess) return H.hip_fail(e, "synchronize");
    S.ready = true;
    return SIRIUS_HIP_OK;
}

static inline int64_t pad64(int64_t n) { return (n + 63) / 64 * 64; }

// Workspace of a training batch of n_pad rows
struct MlpWork {
    uint16_t* h[MLP_NL];  // h[0]: [n_pad][64], h[1..4]: [n_pad][256]
    float* z[MLP_NL - 1];
    float* y64;
    float* gyf;
    uint16_t* gyb;
    float* gzf[2];
    uint16_t* gzb[2];
    float* part;
    float* colpart;
    float* loss_i;
    float* loss_part;
    float* loss;
};

constexpr int COLSPLIT = 64;

static int mlp_workspace(Handle& H, int64_t n_pad, bool train, MlpWork& W) {
    int rc = SIRIUS_HIP_OK;
    const size_t r = (size_t)n_pad;
    W.h[0] = (uint16_t*)H.scratch("mlp_h0", r * MLP_X0 * 2, &rc);
    const char* hn[4] = {"mlp_h1", "mlp_h2", "mlp_h3", "mlp_h4"};
    const char* zn[4] = {"mlp_z0", "mlp_z1", "mlp_z2", "mlp_z3"};
    for (int l = 1; l < MLP_NL; ++l)
        W.h[l] = (uint16_t*)H.scratch(train ? hn[l - 1] : hn[(l - 1) & 1], r * MLP_H * 2, &rc);
    for (int l = 0; l < MLP_NL - 1; ++l) W.z[l] = train ? (float*)H.scratch(zn[l], r * MLP_H * 4, &rc) : nullptr;
    W.y64 = (float*)H.scratch("mlp_y64", r * MLP_N4 * 4, &rc);
    if (train) {
        W.gyf = (float*)H.scratch("mlp_gyf", r * MLP_N4 * 4, &rc);
        W.gyb = (uint16_t*)H.scratch("mlp_gyb", r * MLP_N4 * 2, &rc);
        W.gzf[0] = (float*)H.scratch("mlp_gzf0", r * MLP_H * 4, &rc);
        W.gzf[1] = (float*)H.scratch("mlp_gzf1", r * MLP_H * 4, &rc);
        W.gzb[0] = (uint16_t*)H.scratch("mlp_gzb0", r * MLP_H * 2, &rc);
        W.gzb[1] = (uint16_t*)H.scratch("mlp_gzb1", r * MLP_H * 2, &rc);
        W.part = (float*)H.scratch("mlp_part", (size_t)64 * MLP_H * MLP_H * 4, &rc);
        W.colpart = (float*)H.scratch("mlp_colpart", (size_t)COLSPLIT * MLP_H * 4, &rc);
        W.loss_i = (float*)H.scratch("mlp_loss_i", r * 4, &rc);
        W.loss_part = (float*)H.scratch("mlp_loss_part", 256 * 4, &rc);
        W.loss = (float*)H.scratch("mlp_loss", 256, &rc);
    }
    return rc;
}

static GemmArgs gemm(const uint16_t* A, int64_t sai, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbj, int M,
                     int N, int K) {
    GemmArgs g;
    std::memset(&g, 0, sizeof(g));
    g.A = A; g.sai = sai; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbj = sbj;
    g.M = M; g.N = N; g.K = K; g.kchunk = K;
    return g;
}

// Training forward of n_pad rows already in W.h[0]: the hidden layers on BF16 MFMA (keeping
// their pre-activations z), the output head in FP32 from z3.
static hipError_t mlp_forward_rows(MlpState& S, MlpWork& W, int64_t n_pad, hipStream_t st) {
    hipError_t e;
    for (int l = 0; l < MLP_NL - 1; ++l) {
        const int64_t lda = l == 0 ? MLP_X0 : MLP_H;
        GemmArgs g = gemm(W.h[l], lda, 1, S.w[l], 1, MLP_KP[l], (int)n_pad, MLP_H, MLP_KP[l]);
        g.bias = S.b[l];
        g.zout = W.z[l];
        g.hout = W.h[l + 1];
        if ((e = launch_gemm(EPI_HIDDEN, g, 1, st)) != hipSuccess) return e;
    }
    return launch_head_fwd(W.z[3], S.params + mlp_param_offset_w(4), S.params + mlp_param_offset_b(4), n_pad, W.y64,
                           st);
}

// Split of the sample dimension for the weight-gradient GEMMs
static void split_k(int64_t n_pad, int* nsplit, int* kchunk) {
    int s = (int)std::min<int64_t>(64, std::max<int64_t>(1, n_pad / 1024));
    int64_t c = (n_pad + s - 1) / s;
    c = (c + 31) / 32 * 32;
    *kchunk = (int)c;
    *nsplit = (int)((n_pad + c - 1) / c);
}

// Loss and gradient (into grad, public layout) of a batch; *loss_host gets the mean loss.
static int mlp_loss_grad_dev(Handle& H, const float* x, const int32_t* cls, const float* dir, int n, float* grad,
                             float* loss_host) {
    MlpState& S = *H.mlp;
    const int64_t n_pad = pad64(n);
    MlpWork W;
    int rc = mlp_workspace(H, n_pad, true, W);
    if (rc) return rc;
    hipStream_t st = H.stream;
    hipError_t e;
    if ((e = launch_mlp_input(x, n, n_pad, W.h[0], st)) != hipSuccess) return H.hip_fail(e, "mlp input");
    if ((e = mlp_forward_rows(S, W, n_pad, st)) != hipSuccess) return H.hip_fail(e, "mlp forward");
    if ((e = launch_mlp_loss(W.y64, cls, dir, n, n_pad, W.gyf, W.gyb, W.loss_i, st)) != hipSuccess)
        return H.hip_fail(e, "mlp loss");
    int nsplit, kchunk;
    split_k(n_pad, &nsplit, &kchunk);
    // layer 4 (the FP32 head): dW4 = gy^T gelu(z3), db4 = colsum(gy), gz3 = (gy W4) * gelu'(z3)
    {
        const int hsplit = (int)std::min<int64_t>(128, std::max<int64_t>(1, n_pad / 64));
        if ((e = launch_head_wgrad(W.gyf, W.z[3], n_pad, W.part, hsplit, grad + mlp_param_offset_w(4), st)) !=
            hipSuccess)
            return H.hip_fail(e, "dW4");
        if ((e = launch_colsum(W.gyf, n_pad, SIRIUS_MLP_OUT, MLP_N4, W.colpart, COLSPLIT, grad + mlp_param_offset_b(4),
                               st)) != hipSuccess)
            return H.hip_fail(e, "db4");
        if ((e = launch_head_bwd(W.gyf, S.params + mlp_param_offset_w(4), W.z[3], n_pad, W.gzf[0], W.gzb[0], st)) !=
            hipSuccess)
            return H.hip_fail(e, "dX4");
    }
    int cur = 0;
    for (int l = MLP_NL - 2; l >= 0; --l) {
        const int ncols = l == 0 ? MLP_X0 : MLP_H;  // width of the stored layer input rows
        GemmArgs g = gemm(W.gzb[cur], 1, MLP_H, W.h[l], ncols, 1, MLP_H, ncols, (int)n_pad);
        g.kchunk = kchunk;
        g.out = W.part;
        if ((e = launch_gemm(EPI_PART, g, nsplit, st)) != hipSuccess) return H.hip_fail(e, "dW");
        if ((e = launch_reduce_partials(W.part, nsplit, MLP_H, ncols, MLP_H, MLP_IN[l], grad + mlp_param_offset_w(l),
                                        st)) != hipSuccess)
            return H.hip_fail(e, "dW reduce");
        if ((e = launch_colsum(W.gzf[cur], n_pad, MLP_H, MLP_H, W.colpart, COLSPLIT, grad + mlp_param_offset_b(l),
                               st)) != hipSuccess)
            return H.hip_fail(e, "db");
        if (l == 0) break;
        GemmArgs d = gemm(W.gzb[cur], MLP_H, 1, S.w[l], MLP_H, 1, (int)n_pad, MLP_H, MLP_H);
        d.zin = W.z[l - 1];
        d.gzf = W.gzf[cur ^ 1];
        d.gzb = W.gzb[cur ^ 1];
        if ((e = launch_gemm(EPI_BWD, d, 1, st)) != hipSuccess) return H.hip_fail(e, "dX");
        cur ^= 1;
    }
    if ((e = launch_sum(W.loss_i, n_pad, W.loss_part, W.loss, st)) != hipSuccess) return H.hip_fail(e, "loss sum");
    float total = 0.f;
    if ((e = hipMemcpyAsync(&total, W.loss, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return H.hip_fail(e, "loss copy");
    if (loss_host) *loss_host = total / (float)n;
    return SIRIUS_HIP_OK;
}

MlpFusedArgs mlp_fused_args(const MlpState& S, const float* x, float* y, int64_t n) {
    MlpFusedArgs a;
    a.x = x;
    a.n = n;
    a.n_dev = nullptr;
    a.idx = nullptr;
    a.x_stride = SIRIUS_MLP_IN;
    a.w0 = S.fw0;
    for (int l = 1; l < 4; ++l) a.w[l - 1] = S.fwh + (size_t)(l - 1) * MLP_H * MLP_H;
    for (int l = 0; l < 4; ++l) a.b[l] = S.b[l];
    a.w4h = S.fw4h;
    a.w4l = S.fw4l;
    a.b4 = S.params + mlp_param_offset_b(4);
    a.y = y;
    return a;
}

// Forward pass of n samples x[n*5] -> y[n*6] (device): the fused inference kernel.
int mlp_forward_dev(Handle& H, const float* x, float* y, int64_t n) {
    hipError_t e = launch_mlp_fused(mlp_fused_args(*H.mlp, x, y, n), n, 1 << 30, H.stream);
    return e == hipSuccess ? SIRIUS_HIP_OK : H.hip_fail(e, "mlp forward");
}

hipError_t launch_sur_features(const CamConst* cams, int n_cam, const float* tabs, int64_t n, double sur_band,
                               double margin, double r_exit, float* x, float* ref, uint8_t* cand, float* dist,
                               float* rec, hipStream_t st);
hipError_t launch_sur_targets(const DevOut& o, const float* ref, int64_t n, int32_t* cls, float* dir, hipStream_t st);
hipError_t launch_sur_gate(const CamConst* cams, int n_cam, const float* tabs, int64_t n, double band, double crit_band,
                           double long_band, double sur_band, double margin, double r_exit, int32_t* list_a,
                           int32_t* list_b, int32_t* list_l, int32_t* list_s, uint32_t* counts, uint8_t* flags,
                           float* rec, const float* cal_err, const uint32_t* cal_n, const uint32_t* cal_bad,
                           const float* cal_dmax, double err_max, float sup_x, float sup_y, hipStream_t st);
hipError_t launch_sur_decide(const int32_t* list_s, const uint32_t* count_s, int64_t cap, const float* y,
                             const float* rec, const float* cal_err, const uint32_t* cal_n, const uint32_t* cal_bad,
                             const float* cal_dmax,
                             double pmin, double err_max, double audit_fraction, uint64_t audit_seed, const DevOut& o,
                             int n_bins, int32_t* list_a, uint32_t* count_a, int32_t* list_audit,
                             uint32_t* count_audit, uint32_t* n_gated, hipStream_t st);
hipError_t launch_sur_calibrate(const int32_t* list, int64_t m, const float* y, const float* rec, const DevOut& o,
                                double pmin, float* cal_err, uint32_t* cal_n, uint32_t* cal_bad, float* cal_dmax,
                                hipStream_t st);
hipError_t launch_sur_gather(const int32_t* perm, int64_t start, int64_t m, const float* x, const int32_t* cls,
                             const float* dir, float* xb, int32_t* cb, float* db, hipStream_t st);

// Surrogate stage of the FP32 path: the gate sorts the batch into lists A (FP32 integrator),
// B (FP64 band), L (taken first: near the critical curve) and S (the network's candidates, by the
// analytic test); the fused network runs on S, and the decision answers candidates or appends
// them to list A. work[16..19] = |A|, |B|, |L|, |S|, work[10] = |audit|, work[20] = candidates
// the error gate refused. Every count stays on the device.
int surrogate_stage(Handle& H, Batch& B, const SiriusHipTraceOpts& o, DevOut& dout, int32_t* list_a,
                    int32_t* list_b, int32_t* list_l, int32_t* list_audit, uint32_t* work) {
    int rc = SIRIUS_HIP_OK;
    float* tabs = upload_critical_tables(H, B, &rc);
    const int64_t n = B.n;
    int32_t* list_s = (int32_t*)H.scratch("sur_s", (size_t)n * 4, &rc);
    float* rec = (float*)H.scratch("sur_rec", (size_t)n * SUR_REC * 4, &rc);
    float* y = (float*)H.scratch("sur_y", (size_t)n * SIRIUS_MLP_OUT * 4, &rc);
    if (rc) return rc;
    MlpState& S = *H.mlp;
    hipError_t e;
    if ((e = launch_sur_gate(B.cams_dev, (int)B.cc.size(), tabs, n, std::max(o.escalate_band, 0.0),
                             std::max(o.critical_band, 0.0), SIRIUS_LONG_BAND, o.surrogate_band, o.surrogate_margin, o.r_escape,
                             list_a, list_b, list_l, list_s, work + 16, dout.flags, rec, S.cal_err, S.cal_n, S.cal_bad,
                             S.cal_dmax, o.surrogate_err_max, S.sup_x, S.sup_y, H.stream)) != hipSuccess)
        return H.hip_fail(e, "surrogate gate");
    MlpFusedArgs a = mlp_fused_args(S, rec, y, 0);
    a.n_dev = work + 19;
    a.idx = list_s;
    a.x_stride = SUR_REC;
    if ((e = launch_mlp_fused(a, n, 2 * std::max(H.cu_count, 1), H.stream)) != hipSuccess)
        return H.hip_fail(e, "surrogate network");
    if ((e = launch_sur_decide(list_s, work + 19, n, y, rec, S.cal_err, S.cal_n, S.cal_bad, S.cal_dmax, o.surrogate_pmin,
                               o.surrogate_err_max, o.audit_fraction, o.audit_seed, dout, o.n_bins, list_a, work + 16,
                               list_audit, work + 10, work + 20, H.stream)) != hipSuccess)
        return H.hip_fail(e, "surrogate decision");
    return SIRIUS_HIP_OK;
}

// splitmix64 stream for the initial weights
static uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace sirius

using namespace sirius;

#define MLP_GUARD_BEGIN try {
#define MLP_GUARD_END(h)                                                                   \
    }                                                                                      \
    catch (const std::bad_alloc&) { return (h)->fail(SIRIUS_HIP_E_NOMEM, "host allocation"); } \
    catch (...) { return (h)->fail(SIRIUS_HIP_E_DEVICE, "unexpected exception"); }

extern "C" {

size_t sirius_hip_mlp_param_count(void) { return mlp_param_count(); }

int sirius_hip_mlp_init(SiriusHipHandle hh, uint64_t seed) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    MLP_GUARD_BEGIN
    int rc = mlp_alloc(*H);
    if (rc) return rc;
    std::vector<float> p(mlp_param_count(), 0.f);
    uint64_t s = seed;
    for (int l = 0; l < MLP_NL; ++l) {
        const double lim = std::sqrt(6.0 / MLP_IN[l]);
        float* w = p.data() + mlp_param_offset_w(l);
        for (int i = 0; i < MLP_OUTD[l] * MLP_IN[l]; ++i) {
            double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            w[i] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    hipError_t e;
    if ((e = hipStreamSynchronize(H->stream)) != hipSuccess ||
        (e = hipMemcpy(H->mlp->params, p.data(), p.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return H->hip_fail(e, "params upload");
    return mlp_reset_adam(*H);
    MLP_GUARD_END(H)
}

int sirius_hip_mlp_get_params(SiriusHipHandle hh, float* host_params, size_t n) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    if (!H->mlp || !H->mlp->ready) return H->fail(SIRIUS_HIP_E_STATE, "network not initialised");
    if (!host_params || n != mlp_param_count()) return H->fail(SIRIUS_HIP_E_INVALID, "parameter count");
    hipError_t e = hipStreamSynchronize(H->stream);  // training steps run asynchronously on the stream
    if (e == hipSuccess) e = hipMemcpy(host_params, H->mlp->params, n * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? SIRIUS_HIP_OK : H->hip_fail(e, "params download");
}

int sirius_hip_mlp_set_params(SiriusHipHandle hh, const float* host_params, size_t n) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    MLP_GUARD_BEGIN
    if (!host_params || n != mlp_param_count()) return H->fail(SIRIUS_HIP_E_INVALID, "parameter count");
    int rc = mlp_alloc(*H);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(H->stream);
    if (e == hipSuccess) e = hipMemcpy(H->mlp->params, host_params, n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return H->hip_fail(e, "params upload");
    return mlp_reset_adam(*H);
    MLP_GUARD_END(H)
}

int sirius_hip_mlp_forward(SiriusHipHandle hh, const float* x, float* y, int n) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    MLP_GUARD_BEGIN
    if (!H->mlp || !H->mlp->ready) return H->fail(SIRIUS_HIP_E_STATE, "network not initialised");
    if (!x || !y || n <= 0) return H->fail(SIRIUS_HIP_E_INVALID, "x, y and n > 0 are required");
    int rc = mlp_forward_dev(*H, x, y, n);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(H->stream);
    return e == hipSuccess ? SIRIUS_HIP_OK : H->hip_fail(e, "synchronize");
    MLP_GUARD_END(H)
}

int sirius_hip_mlp_loss_grad(SiriusHipHandle hh, const float* x, const int32_t* cls, const float* dir, int n,
                             float* grad, float* loss) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    MLP_GUARD_BEGIN
    if (!H->mlp || !H->mlp->ready) return H->fail(SIRIUS_HIP_E_STATE, "network not initialised");
    if (!x || !cls || !dir || !grad || n <= 0) return H->fail(SIRIUS_HIP_E_INVALID, "x, cls, dir, grad, n > 0");
    return mlp_loss_grad_dev(*H, x, cls, dir, n, grad, loss);
    MLP_GUARD_END(H)
}

static int mlp_adam_step(Handle& H, const float* x, const int32_t* cls, const float* dir, int n, float lr,
                         float* loss) {
    MlpState& S = *H.mlp;
    int rc = mlp_loss_grad_dev(H, x, cls, dir, n, S.grad, loss);
    if (rc) return rc;
    S.step++;
    const float c1 = 1.f - std::pow(0.9f, (float)S.step), c2 = 1.f - std::pow(0.999f, (float)S.step);
    hipError_t e;
    if ((e = launch_adam(S.params, S.grad, S.adam_m, S.adam_v, (int64_t)mlp_param_count(), lr, c1, c2, H.stream)) !=
            hipSuccess ||
        (e = mlp_pack(S, H.stream)) != hipSuccess)
        return H.hip_fail(e, "adam");
    return SIRIUS_HIP_OK;
}

int sirius_hip_surrogate_features(SiriusHipHandle hh, const SiriusHipCamera* cams, int n_cam,
                                  const SiriusHipMetric* metrics, int n_metric, float* x, float* ref) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    MLP_GUARD_BEGIN
    if (!x || !ref) return H->fail(SIRIUS_HIP_E_INVALID, "x and ref are required");
    int rc = H->ensure_device();
    if (rc) return rc;
    SiriusHipTraceOpts o;
    sirius_hip_default_opts(&o, SIRIUS_PRECISION_FP32_ESCALATE);
    Batch B;
    if ((rc = prepare_batch(*H, cams, n_cam, metrics, n_metric, o, B))) return rc;
    if (B.family != SIRIUS_METRIC_KERR_SCHILD) return H->fail(SIRIUS_HIP_E_UNSUPPORTED, "Kerr-Schild cameras only");
    float* tabs = upload_critical_tables(*H, B, &rc);
    if (rc) return rc;
    hipError_t e = launch_sur_features(B.cams_dev, (int)B.cc.size(), tabs, B.n, o.surrogate_band, o.surrogate_margin,
                                       o.r_escape, x, ref, nullptr, nullptr, nullptr, H->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(H->stream);
    return e == hipSuccess ? SIRIUS_HIP_OK : H->hip_fail(e, "surrogate inputs");
    MLP_GUARD_END(H)
}

int sirius_hip_surrogate_train(SiriusHipHandle hh, const SiriusHipCamera* cams, int n_cam,
                               const SiriusHipMetric* metrics, int n_metric, const SiriusHipTraceOpts* label_opts,
                               int epochs, int batch, float lr, uint64_t seed, float* loss) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    MLP_GUARD_BEGIN
    if (!H->mlp || !H->mlp->ready) return H->fail(SIRIUS_HIP_E_STATE, "network not initialised");
    if (!label_opts || epochs < 0 || batch <= 0 || !(lr > 0))
        return H->fail(SIRIUS_HIP_E_INVALID, "label_opts, epochs >= 0, batch > 0, lr > 0");
    if (label_opts->use_surrogate) return H->fail(SIRIUS_HIP_E_INVALID, "labels come from the integrator");
    int rc;
    Batch B;
    if ((rc = prepare_batch(*H, cams, n_cam, metrics, n_metric, *label_opts, B))) return rc;
    if (B.family != SIRIUS_METRIC_KERR_SCHILD) return H->fail(SIRIUS_HIP_E_UNSUPPORTED, "Kerr-Schild cameras only");
    DevOut dout;
    if ((rc = bind_outputs(*H, nullptr, B.n, label_opts->n_bins, dout, label_opts->emission == SIRIUS_EMISSION_VOLUMETRIC))) return rc;
    if ((rc = run_trace(*H, B, *label_opts, dout, nullptr, nullptr, 0, nullptr, false))) return rc;
    const int64_t n = B.n;
    float* tabs = upload_critical_tables(*H, B, &rc);
    float* x = (float*)H->scratch("tr_x", (size_t)n * 5 * 4, &rc);
    float* ref = (float*)H->scratch("tr_ref", (size_t)n * 3 * 4, &rc);
    uint8_t* cand = (uint8_t*)H->scratch("tr_cand", (size_t)n, &rc);
    float* rec = (float*)H->scratch("tr_rec", (size_t)n * SUR_REC * 4, &rc);
    int32_t* cls = (int32_t*)H->scratch("tr_cls", (size_t)n * 4, &rc);
    float* dir = (float*)H->scratch("tr_dir", (size_t)n * 3 * 4, &rc);
    if (rc) return rc;
    // the candidate test and the reference are the trace path's, with label_opts' band, margin and
    // escape radius (the decision uses the tracing options' own)
    hipError_t e;
    if ((e = launch_sur_features(B.cams_dev, (int)B.cc.size(), tabs, n, label_opts->surrogate_band,
                                 label_opts->surrogate_margin, label_opts->r_escape, x, ref, cand, nullptr, rec,
                                 H->stream)) != hipSuccess ||
        (e = launch_sur_targets(dout, ref, n, cls, dir, H->stream)) != hipSuccess)
        return H->hip_fail(e, "training set");
    // training and held-out candidates: a seeded 1 in 4 of them calibrates the error gate
    std::vector<uint8_t> c((size_t)n);
    if ((e = hipMemcpyAsync(c.data(), cand, (size_t)n, hipMemcpyDeviceToHost, H->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(H->stream)) != hipSuccess)
        return H->hip_fail(e, "candidates");
    std::vector<int32_t> tr, ho;
    for (int64_t i = 0; i < n; ++i) {
        if (!c[(size_t)i]) continue;
        uint64_t z = seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(i + 1));
        (((splitmix(z) >> 32) & 3u) == 0 ? ho : tr).push_back((int32_t)i);
    }
    if (tr.empty()) return H->fail(SIRIUS_HIP_E_INVALID, "no ray of these cameras is a surrogate candidate");
    const int64_t nt = (int64_t)tr.size(), nh = (int64_t)ho.size();
    int32_t* perm = (int32_t*)H->scratch("tr_perm", (size_t)nt * 4, &rc);
    int32_t* hol = (int32_t*)H->scratch("tr_hold", (size_t)std::max<int64_t>(nh, 1) * 4, &rc);
    float* yh = (float*)H->scratch("tr_yhold", (size_t)std::max<int64_t>(nh, 1) * SIRIUS_MLP_OUT * 4, &rc);
    const int64_t m = std::min<int64_t>(batch, nt);
    float* xb = (float*)H->scratch("tr_xb", (size_t)m * 5 * 4, &rc);
    int32_t* cb = (int32_t*)H->scratch("tr_cb", (size_t)m * 4, &rc);
    float* db = (float*)H->scratch("tr_db", (size_t)m * 3 * 4, &rc);
    if (rc) return rc;
    if (nh && (e = hipMemcpyAsync(hol, ho.data(), (size_t)nh * 4, hipMemcpyHostToDevice, H->stream)) != hipSuccess)
        return H->hip_fail(e, "held-out upload");
    uint64_t s = seed;
    const int64_t steps_per_epoch = std::max<int64_t>(1, nt / m);
    const int64_t total = steps_per_epoch * epochs;
    int64_t step = 0;
    double last = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t i = nt - 1; i > 0; --i) std::swap(tr[(size_t)i], tr[(size_t)(splitmix(s) % (uint64_t)(i + 1))]);
        if ((e = hipMemcpyAsync(perm, tr.data(), (size_t)nt * 4, hipMemcpyHostToDevice, H->stream)) != hipSuccess)
            return H->hip_fail(e, "permutation upload");
        double sum = 0;
        for (int64_t b = 0; b < steps_per_epoch; ++b, ++step) {
            if ((e = launch_sur_gather(perm, b * m, m, x, cls, dir, xb, cb, db, H->stream)) != hipSuccess)
                return H->hip_fail(e, "gather");
            const double f = 0.5 * (1.0 + std::cos(3.141592653589793 * (double)step / (double)std::max<int64_t>(total, 1)));
            const float lr_t = (float)(lr * (0.05 + 0.95 * f));
            float l = 0.f;
            if ((rc = mlp_adam_step(*H, xb, cb, db, (int)m, lr_t, &l))) return rc;
            sum += l;
        }
        last = sum / (double)steps_per_epoch;
        // the permutation buffer is rewritten next epoch: its uploads read tr, which the next
        // shuffle changes, so the stream finishes this epoch first
        if ((e = hipStreamSynchronize(H->stream)) != hipSuccess) return H->hip_fail(e, "synchronize");
    }
    // calibration of the error gate on the held-out candidates
    MlpState& S = *H->mlp;
    if ((e = mlp_clear_calibration(S, H->stream)) != hipSuccess) return H->hip_fail(e, "calibration");
    if (nh) {
        MlpFusedArgs a = mlp_fused_args(S, rec, yh, nh);
        a.idx = hol;
        a.x_stride = SUR_REC;
        if ((e = launch_mlp_fused(a, nh, 1 << 30, H->stream)) != hipSuccess ||
            (e = launch_sur_calibrate(hol, nh, yh, rec, dout, label_opts->surrogate_pmin, S.cal_err, S.cal_n,
                                      S.cal_bad, S.cal_dmax, H->stream)) != hipSuccess)
            return H->hip_fail(e, "calibration");
    }
    // the support of the training set on the image plane: the largest pixel-centre coordinates
    // of the training cameras (as the device's ray construction computes them, in FP32)
    float sx = 0.f, sy = 0.f;
    for (const CamConst& c : B.cc) {
        sx = std::max(sx, (1.f - 1.f / (float)c.width) * (float)c.aspect * (float)c.tan_half);
        sy = std::max(sy, (1.f - 1.f / (float)c.height) * (float)c.tan_half);
    }
    S.sup_x = sx;
    S.sup_y = sy;
    if ((e = hipStreamSynchronize(H->stream)) != hipSuccess) return H->hip_fail(e, "synchronize");
    if (loss) *loss = (float)last;
    return SIRIUS_HIP_OK;
    MLP_GUARD_END(H)
}

int sirius_hip_surrogate_calibration(SiriusHipHandle hh, float* max_err, uint32_t* answered, uint32_t* misclassified,
                                     int capacity, int* n_bins) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    MLP_GUARD_BEGIN
    if (n_bins) *n_bins = SUR_BINS;
    if (!H->mlp || !H->mlp->ready) return H->fail(SIRIUS_HIP_E_STATE, "network not initialised");
    if (!max_err || !answered || !misclassified || capacity < SUR_BINS)
        return H->fail(SIRIUS_HIP_E_INVALID, "three arrays of at least n_bins entries are required");
    MlpState& S = *H->mlp;
    hipError_t e;
    if ((e = hipMemcpyAsync(max_err, S.cal_err, SUR_BINS * 4, hipMemcpyDeviceToHost, H->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(answered, S.cal_n, SUR_BINS * 4, hipMemcpyDeviceToHost, H->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(misclassified, S.cal_bad, SUR_BINS * 4, hipMemcpyDeviceToHost, H->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(H->stream)) != hipSuccess)
        return H->hip_fail(e, "calibration copy");
    return SIRIUS_HIP_OK;
    MLP_GUARD_END(H)
}

int sirius_hip_mlp_train_step(SiriusHipHandle hh, const float* x, const int32_t* cls, const float* dir, int n,
                              float lr, float* loss) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    MLP_GUARD_BEGIN
    if (!H->mlp || !H->mlp->ready) return H->fail(SIRIUS_HIP_E_STATE, "network not initialised");
    if (!x || !cls || !dir || n <= 0 || !(lr > 0)) return H->fail(SIRIUS_HIP_E_INVALID, "x, cls, dir, n > 0, lr > 0");
    return mlp_adam_step(*H, x, cls, dir, n, lr, loss);
    MLP_GUARD_END(H)
}

}  // extern "C"
