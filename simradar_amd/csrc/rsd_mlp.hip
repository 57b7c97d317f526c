// rsd_mlp.hip -- training and inference of the debris surrogate network with
// hand-written BF16 MFMA kernels (FP32 accumulate) on gfx950.
//
// One training step over a batch of B samples (B a multiple of 256):
//   k_mlp_fwd_bwd  per 16-sample tile: forward (activations kept in registers),
//                  output error dY = 2 (y - t) / (12 B), backward through the
//                  transposed weights and the ReLU masks; stores every layer's input X_L and
//                  pre-activation gradient dZ_L feature-major in bf16
//   k_mlp_wgrad    dW_L = dZ_L X_L^T as MFMA tiles over K-slices of the batch
//   k_mlp_bgrad    db_L = sum over the batch of dZ_L
//   k_mlp_adam     sums the K-slices in a fixed order, Adam update of the FP32
//                  master parameters, refresh of the bf16 images (W and W^T)
// No counterpart in the reference (the table path is rs.cl:195-359).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rsd.h"
#include "rsd_mlp.h"
#include "rsd_mlp_host.h"

#define KSPLIT 8
#define MLP_PN (MLP_WN + MLP_BN)      // padded parameter image (weights + biases + scaling)

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Copy the weight image (and its transposes) from the FP32 master copy.
__global__ void k_mlp_pack(const float *mw, uint16_t *w, uint16_t *wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_WN) return;
    const uint16_t h = mlp_bf16(mw[i]);
    w[i] = h;
    if (i >= MLP_W1 && i < MLP_W4) {
        const int L = (i - MLP_W1) / (MLP_H * MLP_H), o = ((i - MLP_W1) / MLP_H) % MLP_H, n = (i - MLP_W1) % MLP_H;
        wt[MLP_T1 + L * MLP_H * MLP_H + n * MLP_H + o] = h;
    } else if (i >= MLP_W4) {
        const int o = (i - MLP_W4) / MLP_H, n = (i - MLP_W4) % MLP_H;
        wt[MLP_T4 + n * MLP_N4 + o] = h;
    }
}

struct TrainK {
    MlpDev net;
    const float *x;         // [n][8] inputs (feature 7 = 0)
    const float *t;         // [n][12] normalised targets
    size_t start;           // first sample of the batch
    int B;                  // batch size
    uint16_t *act;          // [5][64][B] layer inputs X_L, feature-major
    uint16_t *dz;           // [5][64][B] dZ_L (L = 4: dY, 16 rows)
    float *loss;            // [B / 16] per-tile sum of squared errors
};

__device__ __forceinline__ void store_fm(uint16_t *base, int B, int feat0, int s, mlp_s4 v) {
    base[(size_t)(feat0 + 0) * B + s] = (uint16_t)v.x;
    base[(size_t)(feat0 + 1) * B + s] = (uint16_t)v.y;
    base[(size_t)(feat0 + 2) * B + s] = (uint16_t)v.z;
    base[(size_t)(feat0 + 3) * B + s] = (uint16_t)v.w;
}

static_assert(MLP_WN % 8 == 0 && MLP_TN % 8 == 0 && MLP_BN % 4 == 0, "images staged in 16-byte pieces");

__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(TrainK a) {
    __shared__ __attribute__((aligned(16))) uint16_t w[MLP_WN];
    __shared__ __attribute__((aligned(16))) uint16_t wt[MLP_TN];
    __shared__ __attribute__((aligned(16))) float b[MLP_BN];
    // the images in 16-byte pieces (their device copies are 16-byte aligned)
    for (int i = threadIdx.x; i < MLP_WN / 8; i += blockDim.x) ((uint4 *)w)[i] = ((const uint4 *)a.net.w)[i];
    for (int i = threadIdx.x; i < MLP_TN / 8; i += blockDim.x) ((uint4 *)wt)[i] = ((const uint4 *)a.net.wt)[i];
    for (int i = threadIdx.x; i < MLP_BN / 4; i += blockDim.x) ((float4 *)b)[i] = ((const float4 *)a.net.b)[i];
    __syncthreads();
    const int l = threadIdx.x & 63, row = l & 15, grp = l >> 4;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile * 16 >= a.B) return;
    const int s = tile * 16 + row;                   // sample of this lane's column, within the batch
    const size_t gs = a.start + (size_t)s;
    // layer-0 input: features 4 grp .. 4 grp + 3 of sample s
    mlp_f4 xin = {0.0f, 0.0f, 0.0f, 0.0f};
    if (grp < 2) xin = *(const mlp_f4 *)(a.x + gs * 8 + 4 * grp);
    const mlp_s4 x = mlp_pack(xin);
    mlp_s4 h[4][4];
    const mlp_f4 y = mlp_tile_forward(w, b, x, h);
    store_fm(a.act, a.B, 4 * grp, s, x);
#pragma unroll
    for (int L = 0; L < 4; L++)
#pragma unroll
        for (int k = 0; k < 4; k++) store_fm(a.act + (size_t)(L + 1) * 64 * a.B, a.B, 16 * k + 4 * grp, s, h[L][k]);
    // output error
    mlp_f4 dy;
    float se = 0.0f;
    const float g = 2.0f / (12.0f * (float)a.B);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int o = 4 * grp + r;
        float e = 0.0f;
        if (o < MLP_OUT) e = y[r] - a.t[gs * MLP_OUT + o];
        se += e * e;
        dy[r] = g * e;
    }
    for (int m = 32; m >= 1; m >>= 1) se += __shfl_xor(se, m, 64);
    if (l == 0) a.loss[tile] = se;
    const mlp_s4 dyb = mlp_pack(dy);
    store_fm(a.dz + (size_t)4 * 64 * a.B, a.B, 4 * grp, s, dyb);
    // backward: dZ3 = (W4^T dY) * [h3 > 0]
    mlp_s4 dcur[4], dnxt[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mlp_mfma(mlp_ld4(wt + MLP_T4 + (16 * mt + row) * MLP_N4 + 4 * grp), dyb, acc);
        const mlp_s4 hv = h[3][mt];
        acc[0] = mlp_f32((uint16_t)hv.x) > 0.0f ? acc[0] : 0.0f;
        acc[1] = mlp_f32((uint16_t)hv.y) > 0.0f ? acc[1] : 0.0f;
        acc[2] = mlp_f32((uint16_t)hv.z) > 0.0f ? acc[2] : 0.0f;
        acc[3] = mlp_f32((uint16_t)hv.w) > 0.0f ? acc[3] : 0.0f;
        dcur[mt] = mlp_pack(acc);
        store_fm(a.dz + (size_t)3 * 64 * a.B, a.B, 16 * mt + 4 * grp, s, dcur[mt]);
    }
    // dZ(L-1) = (W_L^T dZ_L) * [h(L-1) > 0], L = 3, 2, 1
#pragma unroll
    for (int L = 3; L >= 1; L--) {
        const uint16_t *WT = wt + MLP_T1 + (L - 1) * MLP_H * MLP_H;
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            mlp_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++) acc = mlp_mfma(mlp_ld4(WT + (16 * mt + row) * MLP_H + 16 * kt + 4 * grp), dcur[kt], acc);
            const mlp_s4 hv = h[L - 1][mt];
            acc[0] = mlp_f32((uint16_t)hv.x) > 0.0f ? acc[0] : 0.0f;
            acc[1] = mlp_f32((uint16_t)hv.y) > 0.0f ? acc[1] : 0.0f;
            acc[2] = mlp_f32((uint16_t)hv.z) > 0.0f ? acc[2] : 0.0f;
            acc[3] = mlp_f32((uint16_t)hv.w) > 0.0f ? acc[3] : 0.0f;
            dnxt[mt] = mlp_pack(acc);
            store_fm(a.dz + (size_t)(L - 1) * 64 * a.B, a.B, 16 * mt + 4 * grp, s, dnxt[mt]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) dcur[k] = dnxt[k];
    }
}

// One 16x16 tile of dW_L = dZ_L X_L^T over K-slice ks of the batch; writes
// part[ks][image index].  Tiles: L0 4 (out) x 1 (in 16), L1..L3 4 x 4, L4 1 x 4.
__global__ __launch_bounds__(64) void k_mlp_wgrad(const uint16_t *act, const uint16_t *dz, int B, float *part) {
    int t = blockIdx.x;
    const int ks = blockIdx.y;
    int L, mt, it;
    if (t < 4) { L = 0; mt = t; it = 0; }
    else if (t < 52) { L = 1 + (t - 4) / 16; mt = ((t - 4) % 16) / 4; it = (t - 4) % 4; }
    else { L = 4; mt = 0; it = t - 52; }
    const int l = threadIdx.x, row = l & 15, grp = l >> 4;
    const uint16_t *Z = dz + (size_t)L * 64 * B;
    const uint16_t *X = act + (size_t)L * 64 * B;
    const int slice = B / KSPLIT, s0 = ks * slice;
    mlp_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int s = s0; s < s0 + slice; s += 16) {
        const mlp_s4 A = mlp_ld4(Z + (size_t)(16 * mt + row) * B + s + 4 * grp);
        const mlp_s4 Bv = mlp_ld4(X + (size_t)(16 * it + row) * B + s + 4 * grp);
        acc = mlp_mfma(A, Bv, acc);
    }
    // D: lane l holds dW[16 mt + 4 grp + r][16 it + row]
    const int in_dim = L == 0 ? MLP_K0 : MLP_H;
    const int base = L == 0 ? MLP_W0 : (L == 4 ? MLP_W4 : MLP_W1 + (L - 1) * MLP_H * MLP_H);
    float *P = part + (size_t)ks * MLP_PN;
#pragma unroll
    for (int r = 0; r < 4; r++) P[base + (16 * mt + 4 * grp + r) * in_dim + 16 * it + row] = acc[r];
}

// db_L[o] = sum over the batch of dZ_L[o]: one block per bias entry (272).
__global__ __launch_bounds__(256) void k_mlp_bgrad(const uint16_t *dz, int B, float *part) {
    __shared__ float red[256];
    const int j = blockIdx.x;                 // 0..271: layers 0..3 x 64, layer 4 x 16
    const int L = j < 256 ? j / 64 : 4, o = j < 256 ? j % 64 : j - 256;
    const uint16_t *Z = dz + ((size_t)L * 64 + o) * B;
    float s = 0.0f;
    for (int i = threadIdx.x; i < B; i += blockDim.x) s += mlp_f32(Z[i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[MLP_WN + j] = red[0];     // slice 0 holds the bias gradients
}

__device__ __forceinline__ bool mlp_trainable(int i) {
    if (i < MLP_W1) return (i % MLP_K0) < MLP_IN;                 // W0 columns >= 7 are padding
    if (i >= MLP_W4 && i < MLP_WN) return (i - MLP_W4) / MLP_H < MLP_OUT;
    if (i >= MLP_WN + MLP_B4 && i < MLP_WN + MLP_SCALE) return (i - MLP_WN - MLP_B4) < MLP_OUT;
    return i < MLP_WN + MLP_SCALE;                                // scale and shift are fixed
}

struct AdamK {
    const float *part;      // [KSPLIT][MLP_PN]
    float *master;          // [MLP_PN]
    float *m1, *m2;         // Adam moments
    float *grad;            // optional: summed gradient out
    float lr, beta1, beta2, eps, c1, c2;   // c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t)
    int update;
};

__global__ void k_mlp_adam(AdamK a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_PN) return;
    float g = 0.0f;
    if (i < MLP_WN) {
        for (int k = 0; k < KSPLIT; k++) g += a.part[(size_t)k * MLP_PN + i];
    } else {
        g = a.part[i];
    }
    if (!mlp_trainable(i)) g = 0.0f;
    if (a.grad) a.grad[i] = g;
    if (!a.update || !mlp_trainable(i)) return;
    const float m1 = a.beta1 * a.m1[i] + (1.0f - a.beta1) * g;
    const float m2 = a.beta2 * a.m2[i] + (1.0f - a.beta2) * g * g;
    a.m1[i] = m1;
    a.m2[i] = m2;
    a.master[i] -= a.lr * (m1 * a.c1) / (sqrtf(m2 * a.c2) + a.eps);
}

__global__ __launch_bounds__(64) void k_mlp_infer(MlpDev net, const float *x, float *y, size_t n) {
    __shared__ __attribute__((aligned(16))) float scratch[64 * MLP_SCR];
    const size_t i = (size_t)blockIdx.x * 64 + threadIdx.x;
    float f[8];
#pragma unroll
    for (int k = 0; k < 8; k++) f[k] = i < n ? x[i * 8 + k] : 0.0f;
    float out[MLP_OUT];
    mlp_wave_eval(net.w, net.b, f, scratch, out);
    if (i < n)
#pragma unroll
        for (int k = 0; k < MLP_OUT; k++) y[i * MLP_OUT + k] = out[k];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct rsd_mlp {
    int device;
    hipStream_t stream;
    float *master, *m1, *m2, *part, *grad;   // padded images (MLP_PN)
    uint16_t *w, *wt;
    float *b;                                 // = master + MLP_WN (biases, scale, shift)
    float *x, *t;                             // training set
    size_t n;
    uint16_t *act, *dz;
    float *loss;
    int batch_cap;
    int step;
    char err[256];
};

#define MCHK(m, call)                                                                                      \
    do {                                                                                                   \
        hipError_t _r = (call);                                                                            \
        if (_r != hipSuccess) {                                                                            \
            snprintf((m)->err, sizeof((m)->err), "%s: %s", #call, hipGetErrorString(_r));                  \
            return RSD_ERR_HIP;                                                                            \
        }                                                                                                  \
    } while (0)

static void mfree(void *p) {
    if (p) (void)hipFree(p);
}

// Every copy and fill of a network's memory goes on the network's own stream
// (non-blocking, so work there is not ordered after the legacy default
// stream's): a host-side copy waits for the stream, a fill is ordered before
// the kernels that follow it.
static hipError_t mcopy(rsd_mlp *m, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    hipError_t r = hipMemcpyAsync(dst, src, bytes, kind, m->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(m->stream);
    return r;
}

static hipError_t mzero(rsd_mlp *m, void *dst, size_t bytes) { return hipMemsetAsync(dst, 0, bytes, m->stream); }

// Position of public parameter k (flat, unpadded) in the padded image, or -1.
static int mlp_image_index(int k) {
    if (k < 0) return -1;
    if (k < MLP_H * MLP_IN) return MLP_W0 + (k / MLP_IN) * MLP_K0 + k % MLP_IN;
    k -= MLP_H * MLP_IN;
    if (k < MLP_H) return MLP_WN + MLP_B0 + k;
    k -= MLP_H;
    for (int L = 1; L < 4; L++) {
        if (k < MLP_H * MLP_H) return MLP_W1 + (L - 1) * MLP_H * MLP_H + k;
        k -= MLP_H * MLP_H;
        if (k < MLP_H) return MLP_WN + MLP_H * L + k;
        k -= MLP_H;
    }
    if (k < MLP_OUT * MLP_H) return MLP_W4 + k;
    k -= MLP_OUT * MLP_H;
    if (k < MLP_OUT) return MLP_WN + MLP_B4 + k;
    return -1;
}

static int mlp_repack(rsd_mlp *m) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_WN + 255) / 256), dim3(256), 0, m->stream, (const float *)m->master, m->w, m->wt);
    MCHK(m, hipGetLastError());
    return RSD_OK;
}

extern "C" const char *rsd_mlp_last_error(const rsd_mlp *m) { return m ? m->err : "no network"; }

extern "C" int rsd_mlp_create(int device, uint64_t seed, rsd_mlp **out) {
    if (!out) return RSD_ERR_ARG;
    *out = NULL;
    if (device < 0 || device >= rsd_device_count()) return RSD_ERR_NO_DEVICE;
    rsd_mlp *m = (rsd_mlp *)calloc(1, sizeof(rsd_mlp));
    if (!m) return RSD_ERR_ALLOC;
    m->device = device;
    const size_t pb = MLP_PN * sizeof(float);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&m->master, pb) != hipSuccess || hipMalloc((void **)&m->m1, pb) != hipSuccess ||
        hipMalloc((void **)&m->m2, pb) != hipSuccess || hipMalloc((void **)&m->grad, pb) != hipSuccess ||
        hipMalloc((void **)&m->part, (size_t)KSPLIT * pb) != hipSuccess ||
        hipMalloc((void **)&m->w, MLP_WN * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc((void **)&m->wt, MLP_TN * sizeof(uint16_t)) != hipSuccess) {
        rsd_mlp_destroy(m);
        return RSD_ERR_ALLOC;
    }
    m->b = m->master + MLP_WN;
    // He-uniform weights (ReLU) from a splitmix64 stream, zero biases, unit scaling
    float *h = (float *)calloc(MLP_PN, sizeof(float));
    if (!h) {
        rsd_mlp_destroy(m);
        return RSD_ERR_ALLOC;
    }
    uint64_t s = seed ^ 0x5851F42D4C957F2Dull;
    for (int k = 0; k < RSD_MLP_PARAMS; k++) {
        const int i = mlp_image_index(k);
        if (i >= MLP_WN) continue;   // bias
        int fan_in, fan_out;
        if (i < MLP_W1) { fan_in = MLP_IN; fan_out = MLP_H; }
        else if (i < MLP_W4) { fan_in = MLP_H; fan_out = MLP_H; }
        else { fan_in = MLP_H; fan_out = MLP_OUT; }
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = ((double)(z >> 11) + 0.5) * 0x1.0p-53;
        const double lim = i >= MLP_W4 ? sqrt(6.0 / (fan_in + fan_out)) : sqrt(6.0 / fan_in);
        h[i] = (float)((2.0 * u - 1.0) * lim);
    }
    for (int k = 0; k < MLP_OUT; k++) h[MLP_WN + MLP_SCALE + k] = 1.0f;
    if (mcopy(m, m->master, h, pb, hipMemcpyHostToDevice) != hipSuccess || mzero(m, m->m1, pb) != hipSuccess ||
        mzero(m, m->m2, pb) != hipSuccess || mzero(m, m->w, MLP_WN * sizeof(uint16_t)) != hipSuccess ||
        mzero(m, m->wt, MLP_TN * sizeof(uint16_t)) != hipSuccess) {
        free(h);
        rsd_mlp_destroy(m);
        return RSD_ERR_HIP;
    }
    free(h);
    if (mlp_repack(m) != RSD_OK || hipStreamSynchronize(m->stream) != hipSuccess) {
        rsd_mlp_destroy(m);
        return RSD_ERR_HIP;
    }
    *out = m;
    return RSD_OK;
}

extern "C" void rsd_mlp_destroy(rsd_mlp *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    mfree(m->master); mfree(m->m1); mfree(m->m2); mfree(m->part); mfree(m->grad);
    mfree(m->w); mfree(m->wt); mfree(m->x); mfree(m->t); mfree(m->act); mfree(m->dz); mfree(m->loss);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    free(m);
}

extern "C" int rsd_mlp_set_params(rsd_mlp *m, const float *p) {
    if (!m || !p) return RSD_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    float *h = (float *)malloc(MLP_PN * sizeof(float));
    if (!h) return RSD_ERR_ALLOC;
    hipError_t r = mcopy(m, h, m->master, MLP_PN * sizeof(float), hipMemcpyDeviceToHost);
    if (r == hipSuccess) {
        for (int k = 0; k < RSD_MLP_PARAMS; k++) h[mlp_image_index(k)] = p[k];
        r = mcopy(m, m->master, h, MLP_PN * sizeof(float), hipMemcpyHostToDevice);
    }
    free(h);
    MCHK(m, r);
    m->step = 0;
    MCHK(m, mzero(m, m->m1, MLP_PN * sizeof(float)));
    MCHK(m, mzero(m, m->m2, MLP_PN * sizeof(float)));
    int st = mlp_repack(m);
    if (st) return st;
    MCHK(m, hipStreamSynchronize(m->stream));
    return RSD_OK;
}

static int get_image(rsd_mlp *m, const float *src, float *p) {
    float *h = (float *)malloc(MLP_PN * sizeof(float));
    if (!h) return RSD_ERR_ALLOC;
    hipError_t r = mcopy(m, h, src, MLP_PN * sizeof(float), hipMemcpyDeviceToHost);
    if (r == hipSuccess)
        for (int k = 0; k < RSD_MLP_PARAMS; k++) p[k] = h[mlp_image_index(k)];
    free(h);
    MCHK(m, r);
    return RSD_OK;
}

extern "C" int rsd_mlp_get_params(rsd_mlp *m, float *p) {
    if (!m || !p) return RSD_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipStreamSynchronize(m->stream));
    return get_image(m, m->master, p);
}

extern "C" int rsd_mlp_set_output_scaling(rsd_mlp *m, const float *scale, const float *shift) {
    if (!m || !scale || !shift) return RSD_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    float v[2 * MLP_N4] = {0};
    for (int k = 0; k < MLP_OUT; k++) {
        v[k] = scale[k];
        v[MLP_N4 + k] = shift[k];
    }
    MCHK(m, mcopy(m, m->master + MLP_WN + MLP_SCALE, v, sizeof(v), hipMemcpyHostToDevice));
    return RSD_OK;
}

static int ensure_batch(rsd_mlp *m, int B) {
    if (B <= m->batch_cap) return RSD_OK;
    MCHK(m, hipStreamSynchronize(m->stream));
    mfree(m->act); mfree(m->dz); mfree(m->loss);
    m->act = m->dz = NULL;
    m->loss = NULL;
    MCHK(m, hipMalloc((void **)&m->act, (size_t)5 * 64 * B * sizeof(uint16_t)));
    MCHK(m, hipMalloc((void **)&m->dz, (size_t)5 * 64 * B * sizeof(uint16_t)));
    MCHK(m, hipMalloc((void **)&m->loss, (size_t)(B / 16) * sizeof(float)));
    // rows of the padded features are never written by the kernels: zero them once
    MCHK(m, mzero(m, m->act, (size_t)5 * 64 * B * sizeof(uint16_t)));
    MCHK(m, mzero(m, m->dz, (size_t)5 * 64 * B * sizeof(uint16_t)));
    m->batch_cap = B;
    return RSD_OK;
}

// One forward + backward over samples [start, start + B) of x / t (device);
// the gradient lands in m->part (K-slices) ready for k_mlp_adam.
static int grad_step(rsd_mlp *m, const float *x, const float *t, size_t start, int B) {
    int st = ensure_batch(m, B);
    if (st) return st;
    TrainK a;
    a.net.w = m->w;
    a.net.wt = m->wt;
    a.net.b = m->b;
    a.x = x;
    a.t = t;
    a.start = start;
    a.B = B;
    a.act = m->act;
    a.dz = m->dz;
    a.loss = m->loss;
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(B / 64), dim3(256), 0, m->stream, a);
    MCHK(m, hipGetLastError());
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(56, KSPLIT), dim3(64), 0, m->stream, (const uint16_t *)m->act,
                       (const uint16_t *)m->dz, B, m->part);
    MCHK(m, hipGetLastError());
    hipLaunchKernelGGL(k_mlp_bgrad, dim3(4 * MLP_H + MLP_N4), dim3(256), 0, m->stream, (const uint16_t *)m->dz, B, m->part);
    MCHK(m, hipGetLastError());
    return RSD_OK;
}

static int apply(rsd_mlp *m, int update, float lr, float *grad_out) {
    AdamK k;
    k.part = m->part;
    k.master = m->master;
    k.m1 = m->m1;
    k.m2 = m->m2;
    k.grad = grad_out;
    k.beta1 = 0.9f;
    k.beta2 = 0.999f;
    k.eps = 1e-8f;
    k.lr = lr;
    k.update = update;
    if (update) m->step++;
    k.c1 = update ? (float)(1.0 / (1.0 - pow(0.9, m->step))) : 1.0f;
    k.c2 = update ? (float)(1.0 / (1.0 - pow(0.999, m->step))) : 1.0f;
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_PN + 255) / 256), dim3(256), 0, m->stream, k);
    MCHK(m, hipGetLastError());
    if (update) return mlp_repack(m);
    return RSD_OK;
}

static int upload_set(rsd_mlp *m, const float *x7, const float *t, size_t n, float **xd, float **td) {
    float *h = (float *)malloc(n * 8 * sizeof(float));
    if (!h) return RSD_ERR_ALLOC;
    for (size_t i = 0; i < n; i++) {
        memcpy(h + i * 8, x7 + i * MLP_IN, MLP_IN * sizeof(float));
        h[i * 8 + 7] = 0.0f;
    }
    hipError_t r = hipMalloc((void **)xd, n * 8 * sizeof(float));
    if (r == hipSuccess) r = mcopy(m, *xd, h, n * 8 * sizeof(float), hipMemcpyHostToDevice);
    free(h);
    if (r == hipSuccess && t) {
        r = hipMalloc((void **)td, n * MLP_OUT * sizeof(float));
        if (r == hipSuccess) r = mcopy(m, *td, t, n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice);
    }
    MCHK(m, r);
    return RSD_OK;
}

extern "C" int rsd_mlp_gradient(rsd_mlp *m, const float *x, const float *t, size_t n, float *grad, double *loss) {
    if (!m || !x || !t || !grad || n == 0 || n % 256 != 0 || n > (1u << 24)) return RSD_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    float *xd = NULL, *td = NULL;
    int st = upload_set(m, x, t, n, &xd, &td);
    if (!st) st = grad_step(m, xd, td, 0, (int)n);
    if (!st) st = apply(m, 0, 0.0f, m->grad);
    if (!st && hipStreamSynchronize(m->stream) != hipSuccess) st = RSD_ERR_HIP;
    if (!st) st = get_image(m, m->grad, grad);
    if (!st && loss) {
        float *l = (float *)malloc((n / 16) * sizeof(float));
        if (l && mcopy(m, l, m->loss, (n / 16) * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess) {
            double s = 0.0;
            for (size_t i = 0; i < n / 16; i++) s += l[i];
            *loss = s / (12.0 * (double)n);
        }
        free(l);
    }
    mfree(xd);
    mfree(td);
    return st;
}

extern "C" int rsd_mlp_set_data(rsd_mlp *m, const float *x, const float *t, size_t n) {
    if (!m || !x || !t || n < 256) return RSD_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipStreamSynchronize(m->stream));
    mfree(m->x);
    mfree(m->t);
    m->x = m->t = NULL;
    m->n = 0;
    int st = upload_set(m, x, t, n, &m->x, &m->t);
    if (st) return st;
    m->n = n;
    return RSD_OK;
}

extern "C" int rsd_mlp_train(rsd_mlp *m, int epochs, int batch, float lr, double *last_loss) {
    if (!m || epochs < 1 || batch < 256 || batch % 256 != 0 || lr <= 0.0f) return RSD_ERR_ARG;
    if (!m->x || m->n < (size_t)batch) {
        snprintf(m->err, sizeof(m->err), "no training set of at least one batch");
        return RSD_ERR_STATE;
    }
    MCHK(m, hipSetDevice(m->device));
    const size_t steps = m->n / (size_t)batch;
    const size_t total = steps * (size_t)epochs;
    size_t k = 0;
    for (int e = 0; e < epochs; e++) {
        for (size_t s = 0; s < steps; s++, k++) {
            // cosine decay from lr to lr / 20 over the run
            const float f = (float)(0.05 + 0.95 * 0.5 * (1.0 + cos(M_PI * (double)k / (double)total)));
            int st = grad_step(m, m->x, m->t, s * (size_t)batch, batch);
            if (!st) st = apply(m, 1, lr * f, NULL);
            if (st) return st;
        }
    }
    if (last_loss) {
        const int nb = batch / 16;
        float *l = (float *)malloc(nb * sizeof(float));
        if (!l) return RSD_ERR_ALLOC;
        hipError_t r = mcopy(m, l, m->loss, nb * sizeof(float), hipMemcpyDeviceToHost);
        double s = 0.0;
        for (int i = 0; i < nb; i++) s += l[i];
        free(l);
        MCHK(m, r);
        *last_loss = s / (12.0 * batch);
    }
    MCHK(m, hipStreamSynchronize(m->stream));
    return RSD_OK;
}

extern "C" int rsd_mlp_forward(rsd_mlp *m, const float *x, float *y, size_t n) {
    if (!m || !x || !y || n == 0) return RSD_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    float *xd = NULL, *yd = NULL;
    int st = upload_set(m, x, NULL, n, &xd, NULL);
    if (st) return st;
    hipError_t r = hipMalloc((void **)&yd, n * MLP_OUT * sizeof(float));
    if (r == hipSuccess) {
        MlpDev net;
        net.w = m->w;
        net.wt = m->wt;
        net.b = m->b;
        hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, m->stream, net, (const float *)xd, yd, n);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipMemcpyAsync(y, yd, n * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(m->stream);
    mfree(xd);
    mfree(yd);
    MCHK(m, r);
    return RSD_OK;
}

extern "C" int rsd_mlp_device_view(const rsd_mlp *m, MlpDev *out, int *device) {
    if (!m || !out) return RSD_ERR_ARG;
    out->w = m->w;
    out->wt = m->wt;
    out->b = m->b;
    if (device) *device = m->device;
    return RSD_OK;
}

extern "C" int rsd_mlp_sync(rsd_mlp *m) {
    if (!m) return RSD_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipStreamSynchronize(m->stream));
    return RSD_OK;
}
