// rsd_mlp.h -- debris surrogate network: shapes, parameter images and the
// wave-level BF16 MFMA forward tile shared by the engine's debris kernel and the
// training kernels.
//
// Network: 7 inputs -> 4 hidden layers of 64 (ReLU) -> 12 outputs (linear),
// i.e. 7 -> 64 -> 64 -> 64 -> 64 -> 12.  Inputs: the radar-relative orientation
// quaternion R = ori (x) o_conj(el, az) (4) and the body-frame unit air-flow
// direction u = rot(normalize(v_bg - v), conj(ori)) (3).  Outputs (before the
// per-output scale and shift): drag coefficients cd (3) and moment
// coefficients cm (3) of the ADM table at u, and the gamma-rotated scattering
// amplitudes hh, vv, hv (re, im) of the RCS table at R -- the quantities the
// table path reads (rs.cl:207-243, 287-356) after its lookups.
//
// Device layout (the "transposed" GEMM form): a layer computes Z = W X for 16
// samples at a time with v_mfma_f32_16x16x16_bf16, A = a 16x16 tile of W
// (lane l holds W[16 m + l % 16][16 k + 4 (l / 16) + 0..3]), B = the layer input
// (lane l holds features 16 k + 4 (l / 16) + 0..3 of sample l % 16) and D in
// the same lane layout as B (features 16 m + 4 (l / 16) + 0..3 of sample
// l % 16).  The output of one layer is therefore, tile for tile, the B operand
// of the next: activations never leave the registers.  Inference
// (mlp_forward4) runs the 64-wide layers on v_mfma_f32_16x16x32_bf16 instead,
// two such B tiles forming one K = 32 operand.
#ifndef RSD_MLP_H
#define RSD_MLP_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MLP_IN 7
#define MLP_OUT 12
#define MLP_H 64
#define MLP_K0 16                   // padded input width
#define MLP_N4 16                   // padded output width

// bf16 weight image (row-major [out][in]): W0 64x16, W1..W3 64x64, W4 16x64
#define MLP_W0 0
#define MLP_W1 (MLP_W0 + MLP_H * MLP_K0)
#define MLP_W2 (MLP_W1 + MLP_H * MLP_H)
#define MLP_W3 (MLP_W2 + MLP_H * MLP_H)
#define MLP_W4 (MLP_W3 + MLP_H * MLP_H)
#define MLP_WN (MLP_W4 + MLP_N4 * MLP_H)            // 14336 bf16
// transposed image ([in][out]) of W1..W4 for the backward pass
#define MLP_T1 0
#define MLP_T2 (MLP_T1 + MLP_H * MLP_H)
#define MLP_T3 (MLP_T2 + MLP_H * MLP_H)
#define MLP_T4 (MLP_T3 + MLP_H * MLP_H)
#define MLP_TN (MLP_T4 + MLP_H * MLP_N4)            // 13312 bf16
// float biases: b0..b3 (64 each), b4 (16), then output scale (16) and shift (16)
#define MLP_B0 0
#define MLP_B4 (4 * MLP_H)
#define MLP_SCALE (MLP_B4 + MLP_N4)
#define MLP_SHIFT (MLP_SCALE + MLP_N4)
#define MLP_BN (MLP_SHIFT + MLP_N4)                 // 304 floats

// Padded image of the weights for inference out of LDS.  W0: rows at a stride
// of 20 bf16 (40 bytes).  W1..W4: rows at 72 bf16 (144 bytes), each row's 64
// inputs reordered for v_mfma_f32_16x16x32_bf16 (mlp_mfma32): position
// 32 c + 8 g + j holds input 32 c + 4 g + j for j < 4 and 32 c + 16 + 4 g +
// (j - 4) for j >= 4, so one A operand is one aligned 16-byte read, and the 16
// rows one such read touches in a quarter wave cover the 64 LDS banks once (at
// the packed strides of 32 and 128 bytes they share 2 and 8 ways).
#define MLP_PK0 20
#define MLP_PH 72
#define MLP_PW1 (MLP_H * MLP_PK0)
#define MLP_PW4 (MLP_PW1 + 3 * MLP_H * MLP_PH)
#define MLP_PWN (MLP_PW4 + MLP_N4 * MLP_PH)         // 16256 bf16
#define MLP_SCR 12                                  // floats per sample in the output exchange

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef short mlp_s8 __attribute__((ext_vector_type(8)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 mlp_bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ mlp_s4 mlp_ld4(const uint16_t *p) {  // 4 consecutive bf16 (8-byte aligned)
    return *(const mlp_s4 *)p;
}

__device__ __forceinline__ mlp_s8 mlp_cat(mlp_s4 lo, mlp_s4 hi) {
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

// Row strides and offsets of a weight image, and the K = 32 A operand of lane
// group g for input chunk c of a W1..W4 row (inputs 32 c + 4 g + 0..3, then
// 32 c + 16 + 4 g + 0..3).
struct MlpPacked {
    static constexpr int S0 = MLP_K0, SH = MLP_H, O1 = MLP_W1, O4 = MLP_W4;
    static __device__ __forceinline__ mlp_s8 a8(const uint16_t *row, int c, int g) {
        return mlp_cat(mlp_ld4(row + 32 * c + 4 * g), mlp_ld4(row + 32 * c + 16 + 4 * g));
    }
};
struct MlpPadded {
    static constexpr int S0 = MLP_PK0, SH = MLP_PH, O1 = MLP_PW1, O4 = MLP_PW4;
    static __device__ __forceinline__ mlp_s8 a8(const uint16_t *row, int c, int g) {
        return *(const mlp_s8 *)(row + 32 * c + 8 * g);
    }
};

// Device-side view of a network.
struct MlpDev {
    const uint16_t *w;      // MLP_WN bf16
    const uint16_t *wt;     // MLP_TN bf16
    const float *b;         // MLP_BN floats
};

__device__ __forceinline__ uint16_t mlp_bf16(float f) {   // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float mlp_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef __bf16 mlp_bf2 __attribute__((ext_vector_type(2)));
typedef short mlp_s2 __attribute__((ext_vector_type(2)));

// Two values to bf16 in one v_cvt_pk_bf16_f32 (round to nearest even).  The
// empty asm pins the packed pair in a register: without it the compiler merges
// two pairs into one four-wide conversion and lowers that as four single
// conversions and two permutes.
__device__ __forceinline__ mlp_s2 mlp_pack2(float a, float b) {
    mlp_bf2 h;
    h.x = (__bf16)a;
    h.y = (__bf16)b;
    uint32_t u = __builtin_bit_cast(uint32_t, h);
    asm("" : "+v"(u));
    return __builtin_bit_cast(mlp_s2, u);
}

// Four values to bf16 (round to nearest even).
__device__ __forceinline__ mlp_s4 mlp_pack(mlp_f4 v) {
    return __builtin_shufflevector(mlp_pack2(v.x, v.y), mlp_pack2(v.z, v.w), 0, 1, 2, 3);
}

__device__ __forceinline__ mlp_f4 mlp_mfma(mlp_s4 a, mlp_s4 b, mlp_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// D = A B + C on 16 x 32 tiles: lane l holds A[l % 16][8 (l / 16) + 0..7] and
// B[8 (l / 16) + 0..7][l % 16]; D as for the 16x16x16 form.  Only the pairing
// of A's and B's K entries within a lane matters, so the K order is free:
// chunk c of a 64-wide layer input pairs the 16x16x16 B tiles 2 c and 2 c + 1
// (features 32 c + 4 g + 0..3, then 32 c + 16 + 4 g + 0..3).
__device__ __forceinline__ mlp_f4 mlp_mfma32(mlp_s8 a, mlp_s8 b, mlp_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_bf8, a), __builtin_bit_cast(mlp_bf8, b), c,
                                                    0, 0, 0);
}

__device__ __forceinline__ float mlp_act(float x) { return __builtin_fmaxf(x, 0.0f); }   // ReLU

// Forward of one 16-sample tile.  x: layer-0 B operand (lanes with l / 16 >= 2
// hold zeros); w, b: weight and bias images (LDS or global).  Returns the
// output tile in D layout (unscaled); h[4][4] receives the hidden activations
// (bf16, B layout, per layer and K step) when non-null.
__device__ __forceinline__ mlp_f4 mlp_tile_forward(const uint16_t *w, const float *b, mlp_s4 x, mlp_s4 (*h)[4]) {
    const int l = threadIdx.x & 63, row = l & 15, grp = l >> 4;
    mlp_s4 cur[4], nxt[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mlp_mfma(mlp_ld4(w + MLP_W0 + (16 * mt + row) * MLP_K0 + 4 * grp), x, acc);
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] = mlp_act(acc[r] + b[MLP_B0 + 16 * mt + 4 * grp + r]);
        cur[mt] = mlp_pack(acc);
    }
    if (h) {
#pragma unroll
        for (int k = 0; k < 4; k++) h[0][k] = cur[k];
    }
#pragma unroll 1
    for (int L = 1; L < 4; L++) {
        const uint16_t *W = w + MLP_W1 + (L - 1) * MLP_H * MLP_H;
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            mlp_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++) acc = mlp_mfma(mlp_ld4(W + (16 * mt + row) * MLP_H + 16 * kt + 4 * grp), cur[kt], acc);
#pragma unroll
            for (int r = 0; r < 4; r++) acc[r] = mlp_act(acc[r] + b[MLP_H * L + 16 * mt + 4 * grp + r]);
            nxt[mt] = mlp_pack(acc);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) cur[k] = nxt[k];
        if (h) {
#pragma unroll
            for (int k = 0; k < 4; k++) h[L][k] = cur[k];
        }
    }
    mlp_f4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < 4; kt++) y = mlp_mfma(mlp_ld4(w + MLP_W4 + row * MLP_H + 16 * kt + 4 * grp), cur[kt], y);
#pragma unroll
    for (int r = 0; r < 4; r++) y[r] += b[MLP_B4 + 4 * grp + r];
    return y;
}

// Copy the packed weight image w into the padded image dst (LDS) with the
// threads of a block: W0 as 64 rows of four 8-byte chunks; W1..W4 (contiguous
// in both images) as 208 rows of eight 16-byte A operands, each joining the
// two 8-byte runs of its row that MlpPadded::a8 pairs.
__device__ __forceinline__ void mlp_stage_padded(uint16_t *dst, const uint16_t *w) {
    const uint2 *src = (const uint2 *)w;
    for (int c = threadIdx.x; c < MLP_H * 4; c += blockDim.x)
        *(uint2 *)(dst + (c >> 2) * MLP_PK0 + (c & 3) * 4) = src[((c >> 2) * MLP_K0 + (c & 3) * 4) >> 2];
    for (int c = threadIdx.x; c < (3 * MLP_H + MLP_N4) * 8; c += blockDim.x) {
        const int r = c >> 3, ch = (c >> 2) & 1, g = c & 3;
        const int s = MLP_W1 + r * MLP_H + 32 * ch + 4 * g;
        const uint2 lo = src[s >> 2], hi = src[(s + 16) >> 2];
        *(uint4 *)(dst + MLP_PW1 + r * MLP_PH + 32 * ch + 8 * g) = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
}

// ReLU of four values and their bf16 image: rounded first, then max with 0 as
// signed 16-bit integers (a negative bf16 is a negative int16, a positive one a
// positive int16), two packed integer maxima in place of four float ones; the
// same bf16 values as rounding max(v, 0).
__device__ __forceinline__ mlp_s4 mlp_relu_pack(mlp_f4 v) {
    const mlp_s4 h = mlp_pack(v);
    return __builtin_elementwise_max(h, (mlp_s4){0, 0, 0, 0});
}

// Forward of the wave's four 16-sample tiles together: every weight tile is
// read once and applied to the four tiles (four independent MFMA chains).  The
// bias enters as the accumulator the first MFMA of each output tile starts
// from.  Layer 0 (K = 16 for the 7 inputs) runs on the 16x16x16 form, the
// 64-wide layers on the 16x16x32 one (two MFMA per output tile and sample
// tile instead of four).  LAY: the weight image's layout (MlpPacked or
// MlpPadded).
template <class LAY, int P = 4>
__device__ __forceinline__ void mlp_forward4(const uint16_t *w, const float *b, const mlp_s4 x[4], mlp_f4 y[4]) {
    const int l = threadIdx.x & 63, row = l & 15, grp = l >> 4;
#pragma unroll
    for (int t0 = 0; t0 < 4; t0 += P) {
        mlp_s4 cur[P][4], nxt[P][4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const mlp_s4 A = mlp_ld4(w + (16 * mt + row) * LAY::S0 + 4 * grp);
            const mlp_f4 bias = *(const mlp_f4 *)(b + MLP_B0 + 16 * mt + 4 * grp);
#pragma unroll
            for (int t = 0; t < P; t++) cur[t][mt] = mlp_relu_pack(mlp_mfma(A, x[t0 + t], bias));
        }
#pragma unroll 1
        for (int L = 1; L < 4; L++) {
            const uint16_t *W = w + LAY::O1 + (L - 1) * MLP_H * LAY::SH;
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                const mlp_f4 bias = *(const mlp_f4 *)(b + MLP_H * L + 16 * mt + 4 * grp);
                mlp_f4 acc[P];
#pragma unroll
                for (int t = 0; t < P; t++) acc[t] = bias;
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const mlp_s8 A = LAY::a8(W + (16 * mt + row) * LAY::SH, c, grp);
#pragma unroll
                    for (int t = 0; t < P; t++) acc[t] = mlp_mfma32(A, mlp_cat(cur[t][2 * c], cur[t][2 * c + 1]), acc[t]);
                }
#pragma unroll
                for (int t = 0; t < P; t++) nxt[t][mt] = mlp_relu_pack(acc[t]);
            }
#pragma unroll
            for (int t = 0; t < P; t++)
#pragma unroll
                for (int k = 0; k < 4; k++) cur[t][k] = nxt[t][k];
        }
        const mlp_f4 bias = *(const mlp_f4 *)(b + MLP_B4 + 4 * grp);
#pragma unroll
        for (int t = 0; t < P; t++) y[t0 + t] = bias;
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const mlp_s8 A = LAY::a8(w + LAY::O4 + row * LAY::SH, c, grp);
#pragma unroll
            for (int t = 0; t < P; t++) y[t0 + t] = mlp_mfma32(A, mlp_cat(cur[t][2 * c], cur[t][2 * c + 1]), y[t0 + t]);
        }
    }
}

// Network outputs of every lane's own sample.  The inputs reach the layer-0 B
// operands through scratch (each lane's 8 features as two float4, stored
// [half][lane] so that both the stores and the tile reads cover the banks
// once); the four tiles of the wave are evaluated together and their D tiles
// exchanged through scratch, XR floats per sample per round: XR = 12
// (MLP_SCR) in one round (at 48 bytes a row the 16 rows one store touches fall
// in distinct banks), XR = 4 in three rounds of one lane group's four outputs.
// scratch: 64 x max(8, XR) floats per wave.  out[0..11] = scale * y + shift.
template <class LAY = MlpPacked, int P = 4, int XR = MLP_SCR>
__device__ __forceinline__ void mlp_wave_eval(const uint16_t *w, const float *b, const float f[8], float *scratch,
                                              float out[MLP_OUT]) {
    static_assert(XR == MLP_SCR || XR == 4, "exchange of 12 or 4 floats per sample");
    const int l = threadIdx.x & 63, row = l & 15, grp = l >> 4;
    mlp_s4 x[4];
    mlp_f4 y[4];
    mlp_f4 *sx = (mlp_f4 *)scratch;
    sx[l] = (mlp_f4){f[0], f[1], f[2], f[3]};
    sx[64 + l] = (mlp_f4){f[4], f[5], f[6], f[7]};
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
    for (int t = 0; t < 4; t++) {   // lanes with l / 16 >= 2 hold zeros (read in bounds, then dropped)
        const mlp_f4 v = sx[(grp & 1) * 64 + 16 * t + row];
        x[t] = grp < 2 ? mlp_pack(v) : (mlp_s4){0, 0, 0, 0};
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    mlp_forward4<LAY, P>(w, b, x, y);
#pragma unroll
    for (int q0 = 0; q0 < MLP_OUT / 4; q0 += XR / 4) {
        if (XR == MLP_SCR ? grp < MLP_OUT / 4 : grp == q0) {
#pragma unroll
            for (int t = 0; t < 4; t++) *(mlp_f4 *)(scratch + (16 * t + row) * XR + (XR == MLP_SCR ? 4 * grp : 0)) = y[t];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
        for (int q = q0; q < q0 + XR / 4; q++) {
            const mlp_f4 v = *(const mlp_f4 *)(scratch + l * XR + 4 * (q - q0));
#pragma unroll
            for (int r = 0; r < 4; r++) out[4 * q + r] = v[r] * b[MLP_SCALE + 4 * q + r] + b[MLP_SHIFT + 4 * q + r];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
}

#endif
