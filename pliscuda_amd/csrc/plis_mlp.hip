// plis_mlp.hip -- the learned tau proposal on gfx950 matrix cores.
//
// MLP 12 -> 64 -> 64 -> 2 (GELU or ReLU), bfloat16 operands with fp32
// accumulation on v_mfma_f32_32x32x16_bf16.  One wavefront owns a tile of 64
// rows (subvolumes or training samples); activations are staged in LDS in
// both row- and column-major form so every MFMA fragment is a contiguous
// 16-byte LDS read.
//
// Parameter blob (fp32 master, shared with the host):
//   W1[16][64] | b1[64] | W2[64][64] | b2[64] | W3[64][2] | b3[2]
// (W[i][o] multiplies input i into output o; rows 12..15 of W1 stay zero).
// Packed bf16 copies for the MFMA B operand: Wt[o][i] (forward) and W[i][o]
// (backward), W3 padded to 32 outputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "plis_numerics.h"

namespace plis {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kNin = 16, kNh = 64, kNo = 2, kNoPad = 32;
constexpr int kPW1 = 0, kPB1 = kPW1 + kNin * kNh, kPW2 = kPB1 + kNh, kPB2 = kPW2 + kNh * kNh;
constexpr int kPW3 = kPB2 + kNh, kPB3 = kPW3 + kNh * kNo, kPTotal = kPB3 + kNo;  // 5378
// packed bf16 layout (element offsets)
constexpr int kQW1t = 0;                         // [64][16]
constexpr int kQW2t = kQW1t + kNh * kNin;        // [64][64]
constexpr int kQW3t = kQW2t + kNh * kNh;         // [32][64]
constexpr int kQW2 = kQW3t + kNoPad * kNh;       // [64][64]  W2[i][o]
constexpr int kQW3 = kQW2 + kNh * kNh;           // [64][32]  W3[i][o]
constexpr int kQTotal = kQW3 + kNh * kNoPad;
// LDS row strides (bf16 elements), padded against bank conflicts
constexpr int kLdX = 24, kLdH = 72;

struct MlpDev {
    const float* w32;        // fp32 master blob
    const __bf16* q;         // packed bf16
    const float* norm;       // fmu[12] fsig[12] y_mu y_sig
};

// Activations: ACT 0 = ReLU, 1 = GELU.  GELU(z) = z Phi(z) is evaluated as
//     max(z, 0) - |z| max(0, p(|z|)),  p(a) = 0.5 + a (C1 + a (C2 + a C3)),
// a cubic fitted minimax over the real line to the correction
// |z| Phi(-|z|) = relu(z) - GELU(z), with p(0) = 1/2 so that the slope at 0
// is GELU's 1/2: largest error 0.0058 against the erf form, 7 VALU
// instructions, no transcendental.
constexpr float kGeluC1 = -0.47735822f, kGeluC2 = 0.15120307f, kGeluC3 = -0.01585033f;

template <int ACT>
__device__ __forceinline__ float act_f(float z) {
    if constexpr (ACT == 0) {
        return fmaxf(z, 0.0f);
    } else {
        const float a = __builtin_fabsf(z);
        const float p = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(kGeluC3, a, kGeluC2), a, kGeluC1), a, 0.5f);
        return fmaxf(z, 0.0f) - a * fmaxf(p, 0.0f);
    }
}

// d act / dz (the backward pass keeps it from the forward pass)
template <int ACT>
__device__ __forceinline__ float act_d(float z) {
    if constexpr (ACT == 0) {
        return z > 0.0f ? 1.0f : 0.0f;
    } else {
        const float a = __builtin_fabsf(z);
        const float p = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(kGeluC3, a, kGeluC2), a, kGeluC1), a, 0.5f);
        const float dp = __builtin_fmaf(__builtin_fmaf(3.0f * kGeluC3, a, 2.0f * kGeluC2), a, kGeluC1);
        const float q = p > 0.0f ? __builtin_fmaf(a, dp, p) : 0.0f;
        return z > 0.0f ? 1.0f - q : q;
    }
}

// lane-layout of v_mfma_f32_32x32x16_bf16: A fragment = 8 consecutive k of
// row (l % 32) starting at k = 8*(l/32); B fragment = 8 consecutive k of
// column (l % 32); accumulator element i of lane l sits at
// row 8*(i/4) + 4*(l/32) + i%4, column l % 32.
__device__ __forceinline__ bf16x8 lds_frag(const __bf16* M, int ld, int row0, int k0) {
    const int l = threadIdx.x & 63;
    return *reinterpret_cast<const bf16x8*>(M + (row0 + (l & 31)) * ld + k0 + 8 * (l >> 5));
}
__device__ __forceinline__ bf16x8 glb_frag(const __bf16* M, int ld, int row0, int k0) {
    const int l = threadIdx.x & 63;
    return *reinterpret_cast<const bf16x8*>(M + (row0 + (l & 31)) * ld + k0 + 8 * (l >> 5));
}
__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int acc_row(int i) { return 8 * (i >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (i & 3); }
__device__ __forceinline__ int acc_col() { return threadIdx.x & 31; }

struct MlpLds {
    __bf16 x[64 * kLdX];      // standardised inputs [row][16]
    __bf16 xt[16 * kLdH];     // transposed [feature][row]
    __bf16 h1[64 * kLdH], h1t[64 * kLdH];
    __bf16 h2[64 * kLdH], h2t[64 * kLdH];
    __bf16 dz[64 * kLdH], dzt[64 * kLdH];
    __bf16 g1[64 * kLdH], g2[64 * kLdH];  // act'(z) of the hidden layers
    float out[64 * 2];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of the inference-only path (one wavefront): one buffer of 64 rows
// holds the standardised inputs (columns 0..15) and then each hidden layer's
// activations, every layer reading a tile half before overwriting it.
struct MlpLdsInfer {
    __bf16 h[64 * kLdH];
};

// Standardise lane's feature row into LDS row r (stride ld; lane r owns row r).
__device__ __forceinline__ void stage_inputs(__bf16* x, __bf16* xt, const MlpDev& m, const float (&f)[12],
                                             bool valid, int ld = kLdX) {
    const int r = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < kNin; ++k) {
        float z = 0.0f;
        if (k < 12 && valid) z = (f[k] - m.norm[k]) * (1.0f / m.norm[12 + k]);
        const __bf16 b = (__bf16)z;
        x[r * ld + k] = b;
        if (xt) xt[k * kLdH + r] = b;
    }
}

// One dense layer on the 64-row tile: out[64][N] = act(in[64][K] . W + b).
// Writes bf16 activations (row-major and transposed) and, for the hidden
// layers, act'(z) in bf16 (g) for the backward pass.
template <int K, int NB, bool ACTON, int ACT>
__device__ __forceinline__ void dense(const __bf16* in, int ldin, const __bf16* wt, const float* bias,
                                      __bf16* h, __bf16* ht, __bf16* g, float* out2) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            f32x16 c = {};
#pragma unroll
            for (int kb = 0; kb < K / 16; ++kb)
                c = mfma(lds_frag(in, ldin, mb * 32, kb * 16), glb_frag(wt, K, nb * 32, kb * 16), c);
            const int col = nb * 32 + acc_col();
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int row = mb * 32 + acc_row(i);
                if (ACTON) {
                    const float z = c[i] + bias[col];
                    const __bf16 a = (__bf16)act_f<ACT>(z);
                    h[row * kLdH + col] = a;
                    if (ht) ht[col * kLdH + row] = a;
                    g[row * kLdH + col] = (__bf16)act_d<ACT>(z);
                } else if (col < kNo) {
                    out2[row * 2 + col] = c[i] + bias[col];
                }
            }
        }
}

// Forward pass of the tile; returns the lane's two outputs.
template <int ACT>
__device__ __forceinline__ void mlp_forward_tile(MlpLds& s, const MlpDev& m, float& o0, float& o1) {
    wave_sync();
    dense<16, 2, true, ACT>(s.x, kLdX, m.q + kQW1t, m.w32 + kPB1, s.h1, s.h1t, s.g1, nullptr);
    wave_sync();
    dense<64, 2, true, ACT>(s.h1, kLdH, m.q + kQW2t, m.w32 + kPB2, s.h2, s.h2t, s.g2, nullptr);
    wave_sync();
    dense<64, 1, false, ACT>(s.h2, kLdH, m.q + kQW3t, m.w32 + kPB3, nullptr, nullptr, nullptr, s.out);
    wave_sync();
    const int r = threadIdx.x & 63;
    o0 = s.out[r * 2];
    o1 = s.out[r * 2 + 1];
}

// ---------------------------------------------------------------------------
// training: forward + backward over tiles of 64 samples, per-wave gradient
// partials (fixed summation order afterwards: deterministic)
// ---------------------------------------------------------------------------
struct TrainArgs {
    MlpDev m;
    const float* feats;      // [n][12]
    const float* ylog;       // [n] standardised log tau target
    const float* ycrit;      // [n] 0/1
    uint64_t n;              // records
    uint32_t batch;          // samples per step (multiple of 64)
    uint32_t step;           // selects the minibatch
    uint32_t seed_lo, seed_hi;
    int32_t full_pass;       // 1: use records 0..batch-1 in order (gradient checks)
    float* partial;          // [gridDim.x][kPTotal + 1] (last = loss sum)
};

template <int ACT>
__global__ __launch_bounds__(64) void mlp_train_kernel(TrainArgs a) {
    __shared__ MlpLds s;
    const int l = threadIdx.x & 63;
    f32x16 dw1[2], dw2[4], dw3[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) dw1[i] = f32x16{};
#pragma unroll
    for (int i = 0; i < 4; ++i) dw2[i] = f32x16{};
#pragma unroll
    for (int i = 0; i < 2; ++i) dw3[i] = f32x16{};
    float db1 = 0.f, db2 = 0.f, db3 = 0.f, loss = 0.f;
    const uint32_t ntile = a.batch / 64u;
    for (uint32_t t = blockIdx.x; t < ntile; t += gridDim.x) {
        const uint32_t sidx = t * 64u + (uint32_t)l;
        uint64_t rec;
        if (a.full_pass) {
            rec = sidx;
        } else {
            const U4 o = philox10(sidx, a.step, 0x4D4C50u, 0u, a.seed_lo, a.seed_hi);
            rec = (((uint64_t)o.y << 32) | o.x) % a.n;
        }
        float f[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = a.feats[rec * 12 + k];
        stage_inputs(s.x, s.xt, a.m, f, true);
        float o0, o1;
        mlp_forward_tile<ACT>(s, a.m, o0, o1);
        const float y = a.ylog[rec], c = a.ycrit[rec];
        const float sp = fmaxf(o1, 0.f) + log1pf(expf(-fabsf(o1)));
        loss += (o0 - y) * (o0 - y) + sp - c * o1;
        const float inv_b = 1.0f / (float)a.batch;
        const float d0 = 2.0f * (o0 - y) * inv_b;
        const float d1 = (1.0f / (1.0f + expf(-o1)) - c) * inv_b;
        // dO staged as bf16 [row][32] (cols 0,1) and transposed
        for (int q = 0; q < 32; ++q) {
            s.dz[l * kLdH + q] = (__bf16)(q == 0 ? d0 : (q == 1 ? d1 : 0.0f));
            s.dzt[q * kLdH + l] = (__bf16)(q == 0 ? d0 : (q == 1 ? d1 : 0.0f));
        }
        wave_sync();
        // dW3[hid][out] += H2^T . dO ; db3
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int kb = 0; kb < 4; ++kb)
                dw3[mb] = mfma(lds_frag(s.h2t, kLdH, mb * 32, kb * 16), lds_frag(s.dzt, kLdH, 0, kb * 16), dw3[mb]);
        if (l < kNo)
            for (int r = 0; r < 64; ++r) db3 += (float)s.dz[r * kLdH + l];
        // dH2 = dO . W3^T -> dZ2 = dH2 * relu'(Z2)
        f32x16 dh[4];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                f32x16 cc = {};
#pragma unroll
                for (int kb = 0; kb < 2; ++kb)
                    cc = mfma(lds_frag(s.dz, kLdH, mb * 32, kb * 16), glb_frag(a.m.q + kQW3, kNoPad, nb * 32, kb * 16), cc);
                dh[mb * 2 + nb] = cc;
            }
        wave_sync();
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int row = mb * 32 + acc_row(i), col = nb * 32 + acc_col();
                    const __bf16 v = (__bf16)(dh[mb * 2 + nb][i] * (float)s.g2[row * kLdH + col]);
                    s.dz[row * kLdH + col] = v;
                    s.dzt[col * kLdH + row] = v;
                }
        wave_sync();
        // dW2[in][out] += H1^T . dZ2 ; db2
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                for (int kb = 0; kb < 4; ++kb)
                    dw2[mb * 2 + nb] = mfma(lds_frag(s.h1t, kLdH, mb * 32, kb * 16),
                                            lds_frag(s.dzt, kLdH, nb * 32, kb * 16), dw2[mb * 2 + nb]);
        for (int r = 0; r < 64; ++r) db2 += (float)s.dz[r * kLdH + l];
        // dH1 = dZ2 . W2^T -> dZ1
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                f32x16 cc = {};
#pragma unroll
                for (int kb = 0; kb < 4; ++kb)
                    cc = mfma(lds_frag(s.dz, kLdH, mb * 32, kb * 16), glb_frag(a.m.q + kQW2, kNh, nb * 32, kb * 16), cc);
                dh[mb * 2 + nb] = cc;
            }
        wave_sync();
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int row = mb * 32 + acc_row(i), col = nb * 32 + acc_col();
                    const __bf16 v = (__bf16)(dh[mb * 2 + nb][i] * (float)s.g1[row * kLdH + col]);
                    s.dz[row * kLdH + col] = v;
                    s.dzt[col * kLdH + row] = v;
                }
        wave_sync();
        // dW1[in][hid] += X^T . dZ1 (M = 16 inputs padded to 32 rows) ; db1
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                bf16x8 af;
                const int mrow = l & 31, k0 = kb * 16 + 8 * (l >> 5);
#pragma unroll
                for (int j = 0; j < 8; ++j) af[j] = mrow < kNin ? s.xt[mrow * kLdH + k0 + j] : (__bf16)0.0f;
                dw1[nb] = mfma(af, lds_frag(s.dzt, kLdH, nb * 32, kb * 16), dw1[nb]);
            }
        for (int r = 0; r < 64; ++r) db1 += (float)s.dz[r * kLdH + l];
        wave_sync();
    }
    // write this block's partial gradient in blob order
    float* P = a.partial + (size_t)blockIdx.x * (kPTotal + 1);
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = acc_row(i), col = nb * 32 + acc_col();
            if (row < kNin) P[kPW1 + row * kNh + col] = dw1[nb][i];
        }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int row = mb * 32 + acc_row(i), col = nb * 32 + acc_col();
                P[kPW2 + row * kNh + col] = dw2[mb * 2 + nb][i];
            }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = mb * 32 + acc_row(i), col = acc_col();
            if (col < kNo) P[kPW3 + row * kNo + col] = dw3[mb][i];
        }
    P[kPB1 + l] = db1;
    P[kPB2 + l] = db2;
    if (l < kNo) P[kPB3 + l] = db3;
    // loss: wave sum
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (l == 0) P[kPTotal] = loss;
}

// Sum the per-block partials in a fixed order; optional Adam update.
struct AdamArgs {
    float* w32;
    float* m;
    float* v;
    float* grad;             // out: summed gradient (+ loss at [kPTotal])
    const float* partial;
    uint32_t nparts;
    int32_t t;               // Adam step (1-based); 0 = only reduce
    float lr, b1, b2, eps;
};

// Block b reduces parameters 64b..64b+63: wave w sums partials w, w+4, ...
// (coalesced rows), then the four sums are combined in a fixed order, so the
// gradient is deterministic.
constexpr int kAdamParams = 64;
__global__ __launch_bounds__(256) void mlp_adam_kernel(AdamArgs a) {
    __shared__ float s_g[4][kAdamParams];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p = blockIdx.x * kAdamParams + lane;
    float g = 0.f;
    if (p <= kPTotal) {
#pragma unroll 4
        for (uint32_t q = (uint32_t)w; q < a.nparts; q += 4u) g += a.partial[(size_t)q * (kPTotal + 1) + p];
    }
    s_g[w][lane] = g;
    __syncthreads();
    if (w != 0 || p > kPTotal) return;
    g = (s_g[0][lane] + s_g[1][lane]) + (s_g[2][lane] + s_g[3][lane]);
    a.grad[p] = g;
    if (p == kPTotal || a.t <= 0) return;
    if (p >= 12 * kNh && p < kNin * kNh) return;  // padding rows of W1 stay zero
    const float m = a.b1 * a.m[p] + (1.0f - a.b1) * g;
    const float v = a.b2 * a.v[p] + (1.0f - a.b2) * g * g;
    a.m[p] = m;
    a.v[p] = v;
    const float mh = m / (1.0f - powf(a.b1, (float)a.t));
    const float vh = v / (1.0f - powf(a.b2, (float)a.t));
    a.w32[p] -= a.lr * mh / (sqrtf(vh) + a.eps);
}

// fp32 master -> packed bf16 copies
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* w32, __bf16* q) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kQTotal) return;
    float v = 0.f;
    if (e < kQW2t) {  // W1t[o][i]
        const int o = (e - kQW1t) / kNin, i = (e - kQW1t) % kNin;
        v = w32[kPW1 + i * kNh + o];
    } else if (e < kQW3t) {  // W2t[o][i]
        const int o = (e - kQW2t) / kNh, i = (e - kQW2t) % kNh;
        v = w32[kPW2 + i * kNh + o];
    } else if (e < kQW2) {  // W3t[o][i], o < 32
        const int o = (e - kQW3t) / kNh, i = (e - kQW3t) % kNh;
        v = o < kNo ? w32[kPW3 + i * kNo + o] : 0.f;
    } else if (e < kQW3) {  // W2[i][o]
        v = w32[kPW2 + (e - kQW2)];
    } else {  // W3[i][o], o < 32
        const int i = (e - kQW3) / kNoPad, o = (e - kQW3) % kNoPad;
        v = o < kNo ? w32[kPW3 + i * kNo + o] : 0.f;
    }
    q[e] = (__bf16)v;
}

// One hidden layer of the inference tile in transposed form,
// out^T = relu(W^T in^T + b): the A operand is the packed W^T (rows = units),
// the B operand the tile rows of `in`, so lane l's accumulator holds four
// consecutive units of one tile row and stores them with one 8-byte write
// into the row-major activations the next layer reads.  The two halves of
// the tile (rows 0..31, 32..63) go one after the other; a half's B
// fragments are read before its rows are written, so `out` may be `in`.
template <int K, int ACT>
__device__ __forceinline__ void infer_layer(const __bf16* in, int ldin, const __bf16* wt, const float* bias,
                                            __bf16* out) {
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    const int l = threadIdx.x & 63;
#pragma unroll 1
    for (int nb = 0; nb < 2; ++nb) {
        bf16x8 bfr[K / 16];
#pragma unroll
        for (int kb = 0; kb < K / 16; ++kb) bfr[kb] = lds_frag(in, ldin, nb * 32, kb * 16);
        wave_sync();
        const int row = nb * 32 + (l & 31);
#pragma unroll 1
        for (int mb = 0; mb < 2; ++mb) {
            f32x16 c = {};
#pragma unroll
            for (int kb = 0; kb < K / 16; ++kb) c = mfma(glb_frag(wt, K, mb * 32, kb * 16), bfr[kb], c);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int u0 = mb * 32 + 8 * j + 4 * (l >> 5);
                const float4 bb = *reinterpret_cast<const float4*>(bias + u0);
                bf16x4 v;
                v[0] = (__bf16)act_f<ACT>(c[4 * j + 0] + bb.x);
                v[1] = (__bf16)act_f<ACT>(c[4 * j + 1] + bb.y);
                v[2] = (__bf16)act_f<ACT>(c[4 * j + 2] + bb.z);
                v[3] = (__bf16)act_f<ACT>(c[4 * j + 3] + bb.w);
                *reinterpret_cast<bf16x4*>(out + row * kLdH + u0) = v;
            }
        }
        wave_sync();
    }
}

// Forward pass of one 64-row tile for inference (inputs staged in s.h);
// returns lane r's two outputs (row r).  The output layer is W3^T (padded
// to 32 rows) times the tile: lanes 0..31 receive rows l and 32 + l.
template <int ACT>
__device__ __forceinline__ void mlp_infer_tile(MlpLdsInfer& s, const MlpDev& m, float& o0, float& o1) {
    wave_sync();
    infer_layer<kNin, ACT>(s.h, kLdH, m.q + kQW1t, m.w32 + kPB1, s.h);
    wave_sync();
    infer_layer<kNh, ACT>(s.h, kLdH, m.q + kQW2t, m.w32 + kPB2, s.h);
    wave_sync();
    const int l = threadIdx.x & 63;
    float a0 = 0.0f, a1 = 0.0f, b0 = 0.0f, b1 = 0.0f;
#pragma unroll 1
    for (int nb = 0; nb < 2; ++nb) {
        f32x16 c = {};
#pragma unroll
        for (int kb = 0; kb < kNh / 16; ++kb)
            c = mfma(glb_frag(m.q + kQW3t, kNh, 0, kb * 16), lds_frag(s.h, kLdH, nb * 32, kb * 16), c);
        if (nb == 0) {
            a0 = c[0];
            a1 = c[1];
        } else {
            b0 = c[0];
            b1 = c[1];
        }
    }
    const float t0 = __shfl(b0, (l + 32) & 63), t1 = __shfl(b1, (l + 32) & 63);
    o0 = (l < 32 ? a0 : t0) + m.w32[kPB3 + 0];
    o1 = (l < 32 ? a1 : t1) + m.w32[kPB3 + 1];
}

// Inference over given feature rows (parity checks): out[n][2].
template <int ACT>
__global__ __launch_bounds__(64) void mlp_infer_kernel(MlpDev m, const float* feats, uint32_t n, float* out) {
    __shared__ MlpLdsInfer s;
    const int l = threadIdx.x & 63;
    for (uint32_t t = blockIdx.x; t * 64u < n; t += gridDim.x) {
        const uint32_t r = t * 64u + (uint32_t)l;
        float f[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = r < n ? feats[(size_t)r * 12 + k] : 0.f;
        stage_inputs(s.h, nullptr, m, f, r < n, kLdH);
        float o0, o1;
        mlp_infer_tile<ACT>(s, m, o0, o1);
        if (r < n) {
            out[(size_t)r * 2] = o0;
            out[(size_t)r * 2 + 1] = o1;
        }
        wave_sync();
    }
}

}  // namespace plis
