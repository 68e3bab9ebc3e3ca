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

// ---------------------------------------------------------------------------
// inference: the three layers chained in registers
//
// Each layer runs in transposed form, out^T = W^T in^T: the A operand is a
// fragment of W^T (staged in LDS once per block), the B operand the tile's
// activations.  A lane's accumulator holds units 8(i/4) + 4(l/32) + i%4 of
// tile row l%32 (rows 32..63 in the second half-tile nb), and the next
// layer's B fragment wants 8 consecutive units of that row: lane l < 32
// already holds half of them, its partner l + 32 the other half, so one
// v_permlane32_swap per packed bf16 pair builds the fragment without any
// LDS round trip.  Input rows enter the same way (lane r holds row r).
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

constexpr int kLdW1 = 24, kLdW2 = 72;  // LDS row strides (bf16), padded against bank conflicts

// Weights of the inference path in LDS: W1^T [64][16], W2^T [64][64],
// W3^T [32][64] (rows 2..31 zero), biases.
struct MlpLdsW {
    __bf16 w1t[kNh * kLdW1];
    __bf16 w2t[kNh * kLdW2];
    __bf16 w3t[kNoPad * kLdW2];
    float b1[kNh], b2[kNh], b3[4];
};

// Whole block, 16-byte copies; the caller synchronises the block after it.
__device__ __forceinline__ void stage_weights(MlpLdsW& W, const MlpDev& m) {
    const uint32_t nt = blockDim.x;
    const uint4* q = reinterpret_cast<const uint4*>(m.q);
    for (uint32_t e = threadIdx.x; e < (uint32_t)kNh * 2u; e += nt)
        *reinterpret_cast<uint4*>(&W.w1t[(e >> 1) * kLdW1 + (e & 1u) * 8u]) = q[(kQW1t >> 3) + e];
    for (uint32_t e = threadIdx.x; e < (uint32_t)kNh * 8u; e += nt)
        *reinterpret_cast<uint4*>(&W.w2t[(e >> 3) * kLdW2 + (e & 7u) * 8u]) = q[(kQW2t >> 3) + e];
    for (uint32_t e = threadIdx.x; e < (uint32_t)kNoPad * 8u; e += nt)
        *reinterpret_cast<uint4*>(&W.w3t[(e >> 3) * kLdW2 + (e & 7u) * 8u]) = q[(kQW3t >> 3) + e];
    for (uint32_t e = threadIdx.x; e < (uint32_t)kNh; e += nt) {
        W.b1[e] = m.w32[kPB1 + e];
        W.b2[e] = m.w32[kPB2 + e];
    }
    if (threadIdx.x < (uint32_t)kNo) W.b3[threadIdx.x] = m.w32[kPB3 + threadIdx.x];
}

// The dword of lane l ^ 32 (v_permlane32_swap exchanges lanes 32..63 of its
// first operand with lanes 0..31 of its second).
__device__ __forceinline__ uint32_t xhalf(uint32_t v) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32u) ? r[0] : r[1];
}

__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    const bf16x2 v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}

// act(acc + bias) of one 32-unit accumulator block as 8 packed bf16 pairs
// (pair j = elements 2j, 2j + 1); bias points at the block's first unit.
template <int ACT>
__device__ __forceinline__ void act_pack(const f32x16& c, const float* bias, uint32_t (&P)[8]) {
    const uint32_t hi = (threadIdx.x >> 5) & 1u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const f32x4 b = *reinterpret_cast<const f32x4*>(bias + 8 * j + 4 * (int)hi);
        P[2 * j] = pack_bf16(act_f<ACT>(c[4 * j] + b.x), act_f<ACT>(c[4 * j + 1] + b.y));
        P[2 * j + 1] = pack_bf16(act_f<ACT>(c[4 * j + 2] + b.z), act_f<ACT>(c[4 * j + 3] + b.w));
    }
}

// B fragment of the next layer for k block 2 mb + H from the packed
// activations P of accumulator block mb: lane l < 32 keeps units 16H + 0..3
// and receives 16H + 4..7 from its partner; lane l >= 32 receives
// 16H + 8..11 and keeps 16H + 12..15.
template <int H>
__device__ __forceinline__ bf16x8 next_frag(const uint32_t (&P)[8]) {
    const bool up = (threadIdx.x & 32u) != 0u;
    const uint32_t r0 = xhalf(up ? P[4 * H] : P[4 * H + 2]);
    const uint32_t r1 = xhalf(up ? P[4 * H + 1] : P[4 * H + 3]);
    const u32x4 f = up ? u32x4{r0, r1, P[4 * H + 2], P[4 * H + 3]} : u32x4{P[4 * H], P[4 * H + 1], r0, r1};
    return __builtin_bit_cast(bf16x8, f);
}

// Forward pass of one 64-row tile: lane r holds feature row r (valid: the
// row exists); returns row r's two outputs.  W holds the staged weights.
template <int ACT>
__device__ __forceinline__ void mlp_infer_regs(const MlpLdsW& W, const MlpDev& m, const float (&f)[12], bool valid,
                                               float& o0, float& o1) {
    const bool up = (threadIdx.x & 32u) != 0u;
    uint32_t q[8];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const float z0 = valid ? (f[2 * j] - m.norm[2 * j]) * (1.0f / m.norm[12 + 2 * j]) : 0.0f;
        const float z1 = valid ? (f[2 * j + 1] - m.norm[2 * j + 1]) * (1.0f / m.norm[13 + 2 * j]) : 0.0f;
        q[j] = pack_bf16(z0, z1);
    }
    q[6] = 0u;
    q[7] = 0u;
    // layer-1 B fragments: features 0..7 / 8..15 of rows 0..31 (nb 0) and 32..63 (nb 1)
    uint32_t r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = xhalf(up ? q[i] : q[4 + i]);
    bf16x8 B1[2];
    B1[0] = __builtin_bit_cast(bf16x8, up ? u32x4{r[0], r[1], r[2], r[3]} : u32x4{q[0], q[1], q[2], q[3]});
    B1[1] = __builtin_bit_cast(bf16x8, up ? u32x4{q[4], q[5], q[6], q[7]} : u32x4{r[0], r[1], r[2], r[3]});
    float out[2][2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        bf16x8 B2[4];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            const f32x16 c = mfma(lds_frag(W.w1t, kLdW1, mb * 32, 0), B1[nb], f32x16{});
            uint32_t P[8];
            act_pack<ACT>(c, W.b1 + mb * 32, P);
            B2[2 * mb] = next_frag<0>(P);
            B2[2 * mb + 1] = next_frag<1>(P);
        }
        bf16x8 B3[4];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            f32x16 c = {};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) c = mfma(lds_frag(W.w2t, kLdW2, mb * 32, kb * 16), B2[kb], c);
            uint32_t P[8];
            act_pack<ACT>(c, W.b2 + mb * 32, P);
            B3[2 * mb] = next_frag<0>(P);
            B3[2 * mb + 1] = next_frag<1>(P);
        }
        f32x16 c = {};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) c = mfma(lds_frag(W.w3t, kLdW2, 0, kb * 16), B3[kb], c);
        // units 0 and 1 of row nb 32 + l % 32 sit in lanes l < 32
        out[nb][0] = c[0];
        out[nb][1] = c[1];
    }
    const float t0 = __builtin_bit_cast(float, xhalf(__builtin_bit_cast(uint32_t, out[1][0])));
    const float t1 = __builtin_bit_cast(float, xhalf(__builtin_bit_cast(uint32_t, out[1][1])));
    o0 = (up ? t0 : out[0][0]) + W.b3[0];
    o1 = (up ? t1 : out[0][1]) + W.b3[1];
}

// Inference over given feature rows (parity checks and the host API):
// out[n][2].  One wavefront per block, grid-stride over tiles.
template <int ACT>
__global__ __launch_bounds__(64) void mlp_infer_kernel(MlpDev m, const float* feats, uint32_t n, float* out) {
    __shared__ MlpLdsW W;
    stage_weights(W, m);
    __syncthreads();
    const int l = threadIdx.x & 63;
    for (uint32_t t = blockIdx.x; t * 64u < n; t += gridDim.x) {
        const uint32_t r = t * 64u + (uint32_t)l;
        float f[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = r < n ? feats[(size_t)r * 12 + k] : 0.f;
        float o0, o1;
        mlp_infer_regs<ACT>(W, m, f, r < n, o0, o1);
        if (r < n) {
            out[(size_t)r * 2] = o0;
            out[(size_t)r * 2 + 1] = o1;
        }
    }
}

}  // namespace plis
