// mlp.hpp — the learned surrogate of the plastic projection of the
// Wolper/Drucker-Prager return map (case 3 of Wolper_Drucker_Prager,
// simulation/kernels.cuh:431-479): a 4 -> 64 -> 64 -> 3 MLP with SiLU,
// BF16 weights and activations on MFMA (v_mfma_f32_32x32x8_bf16) with FP32
// accumulation, its training kernels and its evaluation kernels.
//
// Features of a trial state (elastic predictor Fe with singular values
// s0 >= s1, plastic Jp_inv, cohesion c): the log singular values in the
// sum/difference basis, (log J, D) = (log s0 + log s1, log s0 - log s1), then
// Jp_inv and log(c / 1e5).  Targets: the projected state's (log J', D') and
// log(Jp_inv' / Jp_inv).  Features and targets are standardised with fixed
// per-component means and scales (MlpNorm).
//
// Targets are log-magnitudes, log(-log J') and log D' (log J' < 0 and D' > 0
// on the ridge-capped plastic branch), so that the network's absolute error
// is a relative error of p' and q'.
//
// Layout: every layer is computed transposed, Z^T = W H^T, one wave per 64
// points (two 32-point column blocks).  The D operand of a 32x32 block (lane
// l holds column l % 32 and rows 8 (j / 4) + 4 (l / 32) + j % 4 in element j)
// is exactly the B operand of the next layer's K chunks, so the forward and
// the input-gradient backward pass stay in registers.  Training uses
// v_mfma_f32_32x32x8_bf16 (chunk q = j / 4, element i = j % 4); weight
// gradients go through LDS transposes.  Inference uses
// v_mfma_f32_32x32x16_bf16: K chunk c of 16 takes elements 8 (c % 2) .. + 7
// of output block c / 2, and the weights are packed in that K order.  Its
// layer 1 is split-precision: x = x_hi + x_lo and W1 = W1_hi + W1_lo in BF16,
// one MFMA computes W1_hi x_hi + W1_hi x_lo + W1_lo x_hi + b1_hi + b1_lo (the
// inputs keep about 16 bits); the layer-2 bias is one more MFMA against a
// constant operand.
#pragma once
#include "point_ops.hpp"

namespace ice {

constexpr int kMlpIn = 4, kMlpHid = 64, kMlpOut = 3;
// FP32 master parameters: W1[64][4], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]
constexpr int kOffW1 = 0, kOffB1 = 256, kOffW2 = 320, kOffB2 = 4416, kOffW3 = 4480, kOffB3 = 4672;
constexpr int kMlpParams = 4675;
// Packed BF16 operands, in units of 64 lanes x 4 bf16 (one MFMA operand)
constexpr int kPkW1A = 0;        // [ob 2]                 A = W1 (rows out, k in padded to 8)
constexpr int kPkW2A = 2;        // [ob 2][kk 8]           A = W2
constexpr int kPkW3A = 18;       // [kk 8]                 A = W3 (rows out padded to 32)
constexpr int kPkW2T = 26;       // [ib 2][kk 8]           A = W2^T (rows in, k out)
constexpr int kPkW3T = 42;       // [ib 2]                 A = W3^T (rows in, k out padded to 8)
constexpr int kPkOps = 44;

// Packed BF16 operands of inference, in units of 64 lanes x 8 bf16
constexpr int kIpW1 = 0;    // [ob 2]       A = (W1_hi | W1_hi | W1_lo | b1_hi, b1_lo) against (x_hi | x_lo | x_hi | 1, 1)
constexpr int kIpW2 = 2;    // [ob 2][c 4]  A = W2 in inference K order
constexpr int kIpB2 = 10;   // [ob 2]       A = (b2_hi, b2_lo) against (1, 1)
constexpr int kIpW3 = 12;   // [c 4]        A = W3 (rows out padded to 32) in inference K order
constexpr int kIpOps = 16;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ __forceinline__ float bf16_float(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

struct MlpNorm {
    float xm[kMlpIn], xs[kMlpIn];     // feature = (raw - xm) / xs
    float ym[kMlpOut], ys[kMlpOut];   // raw target = ym + ys * output
};

struct MlpDev {
    const bf16x4 *pk;    // kPkOps * 64 packed training operands
    const float *p;      // kMlpParams FP32 master parameters (training biases, b3)
    const bf16x8 *ipk;   // kIpOps * 64 packed inference operands
};

// Hidden unit behind element i of inference K chunk c in lane half hk.
__host__ __device__ __forceinline__ int ihid(int c, int hk, int i) {
    return 32 * (c >> 1) + 8 * (2 * (c & 1) + (i >> 2)) + 4 * hk + (i & 3);
}

__device__ __forceinline__ __bf16 to_bf16(float f) { return __builtin_bit_cast(__bf16, bf16_bits(f)); }
// BF16 head and tail of an FP32 value: hi + lo carries about 16 significant bits.
__device__ __forceinline__ float bf16_hi(float f) { return bf16_float(bf16_bits(f)); }
__device__ __forceinline__ float bf16_lo(float f) { return bf16_float(bf16_bits(f - bf16_hi(f))); }

__device__ __forceinline__ int mrow(int j, int lane) { return 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3); }

// SiLU and its derivative from the hardware exp2 and reciprocal (each within
// a few ulp of FP32, far below the BF16 rounding the activations take next).
__device__ __forceinline__ float sigmoid_hw(float z) {
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_hw(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigmoid_hw(z);
    return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)bf16_bits(a);
    r[1] = (short)bf16_bits(b);
    r[2] = (short)bf16_bits(c);
    r[3] = (short)bf16_bits(d);
    return r;
}

// Raw features of a trial state from the invariants of Fe (s0 = Q + R, s1 = Q - R).
__device__ __forceinline__ void mlp_features(const FInv &v, float jp, float coh, float x[kMlpIn]) {
    x[0] = log1pf(v.Jm1);                                       // log J
    x[1] = (v.Q > 0.f) ? 2.f * atanhf(fminf(v.R / v.Q, 0.99999994f)) : 0.f;   // log(s0 / s1)
    x[2] = jp;
    x[3] = logf(fmaxf(coh, 1.f) * 1e-5f);
}

struct MlpFwd {
    f32x16 z1[2][2], z2[2][2];   // pre-activations [out block][point block]
    bf16x4 x[2];                  // layer-1 B operands [point block]
    bf16x4 h1[8][2], h2[8][2];    // layer-2/3 B operands [k chunk][point block]
    f32x16 y[2];                  // outputs [point block]: lane < 32, element j < 3
};

// Forward pass of a wave's 64 points; xs holds this lane's standardised
// features.  Must be called by all 64 lanes.
__device__ __forceinline__ void mlp_forward(const MlpDev &M, const float xs[kMlpIn], MlpFwd &F) {
    const int lane = threadIdx.x & 63;
    const f32x16 zero = {};
#pragma unroll
    for (int bb = 0; bb < 2; bb++) {
        const int src = bb * 32 + (lane & 31);
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = __shfl(xs[i], src, 64);
        F.x[bb] = lane < 32 ? pack4(v[0], v[1], v[2], v[3]) : pack4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int ob = 0; ob < 2; ob++)
#pragma unroll
        for (int bb = 0; bb < 2; bb++) {
            f32x16 z = mfma(M.pk[(kPkW1A + ob) * 64 + lane], F.x[bb], zero);
#pragma unroll
            for (int j = 0; j < 16; j++) z[j] += M.p[kOffB1 + ob * 32 + mrow(j, lane)];
            F.z1[ob][bb] = z;
        }
#pragma unroll
    for (int kk = 0; kk < 8; kk++)
#pragma unroll
        for (int bb = 0; bb < 2; bb++) {
            const f32x16 &z = F.z1[kk >> 2][bb];
            const int j0 = (kk & 3) * 4;
            F.h1[kk][bb] = pack4(silu(z[j0]), silu(z[j0 + 1]), silu(z[j0 + 2]), silu(z[j0 + 3]));
        }
#pragma unroll
    for (int ob = 0; ob < 2; ob++)
#pragma unroll
        for (int bb = 0; bb < 2; bb++) {
            f32x16 z = zero;
#pragma unroll
            for (int kk = 0; kk < 8; kk++) z = mfma(M.pk[(kPkW2A + ob * 8 + kk) * 64 + lane], F.h1[kk][bb], z);
#pragma unroll
            for (int j = 0; j < 16; j++) z[j] += M.p[kOffB2 + ob * 32 + mrow(j, lane)];
            F.z2[ob][bb] = z;
        }
#pragma unroll
    for (int kk = 0; kk < 8; kk++)
#pragma unroll
        for (int bb = 0; bb < 2; bb++) {
            const f32x16 &z = F.z2[kk >> 2][bb];
            const int j0 = (kk & 3) * 4;
            F.h2[kk][bb] = pack4(silu(z[j0]), silu(z[j0 + 1]), silu(z[j0 + 2]), silu(z[j0 + 3]));
        }
#pragma unroll
    for (int bb = 0; bb < 2; bb++) {
        f32x16 y = zero;
#pragma unroll
        for (int kk = 0; kk < 8; kk++) y = mfma(M.pk[(kPkW3A + kk) * 64 + lane], F.h2[kk][bb], y);
#pragma unroll
        for (int j = 0; j < 3; j++) y[j] += M.p[kOffB3 + j];
        F.y[bb] = y;
    }
}

// Inference of a wave's 64 points, one 32-point block at a time, on
// v_mfma_f32_32x32x16_bf16 with the inference operands; y receives this
// lane's point's standardised outputs.  Must be called by all 64 lanes.
// Register-lean for use inside the tile pass: the blocks run one after the
// other (the operand loads are issued per block, not hoisted out of the
// block loop), and layer 3 consumes each pair of layer-2 chunks as soon as
// it exists.
__device__ __forceinline__ void mlp_infer(const MlpDev &M, const float xs[kMlpIn], float y[kMlpOut]) {
    const int lane = threadIdx.x & 63, hk = lane >> 5;
    const f32x16 zero = {};
    const __bf16 one = to_bf16(1.f), nil = to_bf16(0.f);
    bf16x8 ones;
#pragma unroll
    for (int i = 0; i < 8; i++) ones[i] = (hk == 0 && i < 2) ? one : nil;
#pragma unroll 1
    for (int bb = 0; bb < 2; bb++) {
        const bf16x8 *ops = M.ipk;
        __asm__ volatile("" : "+s"(ops));   // keeps this block's operand loads inside the loop
        const int src = bb * 32 + (lane & 31);
        bf16x8 xop;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float v = __shfl(xs[i], src, 64);
            const float hi = bf16_hi(v);
            xop[i] = to_bf16(hi);
            xop[4 + i] = hk == 0 ? to_bf16(v - hi) : (i < 2 ? one : nil);
        }
        bf16x8 h1[4];
#pragma unroll
        for (int ob = 0; ob < 2; ob++) {
            const f32x16 z = mfma16(ops[(kIpW1 + ob) * 64 + lane], xop, zero);
#pragma unroll
            for (int hf = 0; hf < 2; hf++)
#pragma unroll
                for (int i = 0; i < 8; i++) h1[2 * ob + hf][i] = (__bf16)silu(z[8 * hf + i]);
        }
        f32x16 o = zero;
#pragma unroll
        for (int ob = 0; ob < 2; ob++) {
            f32x16 z = mfma16(ops[(kIpB2 + ob) * 64 + lane], ones, zero);
#pragma unroll
            for (int c = 0; c < 4; c++) z = mfma16(ops[(kIpW2 + ob * 4 + c) * 64 + lane], h1[c], z);
#pragma unroll
            for (int hf = 0; hf < 2; hf++) {
                bf16x8 h2;
#pragma unroll
                for (int i = 0; i < 8; i++) h2[i] = (__bf16)silu(z[8 * hf + i]);
                o = mfma16(ops[(kIpW3 + 2 * ob + hf) * 64 + lane], h2, o);
            }
        }
#pragma unroll
        for (int j = 0; j < kMlpOut; j++) {
            const float v = __shfl(o[j] + M.p[kOffB3 + j], lane & 31, 64);
            if (hk == bb) y[j] = v;
        }
    }
}

// ---- packing, training, evaluation kernels ----------------------------------

// FP32 master parameters -> BF16 operands (one thread per operand lane).
__global__ __launch_bounds__(64) void k_mlp_pack(const float *__restrict__ p, bf16x4 *__restrict__ pk) {
    const int op = blockIdx.x, lane = threadIdx.x;
    const int r = lane & 31, kq = 4 * (lane >> 5);
    float v[4];
    for (int i = 0; i < 4; i++) {
        const int k = kq + i;
        float w = 0.f;
        if (op < kPkW2A) {                       // W1: rows out, k in
            const int out = op * 32 + r;
            if (k < kMlpIn) w = p[kOffW1 + out * kMlpIn + k];
        } else if (op < kPkW3A) {                // W2: rows out, k in
            const int ob = (op - kPkW2A) >> 3, kk = (op - kPkW2A) & 7;
            w = p[kOffW2 + (ob * 32 + r) * kMlpHid + kk * 8 + k];
        } else if (op < kPkW2T) {                // W3: rows out (< 3), k in
            const int kk = op - kPkW3A;
            if (r < kMlpOut) w = p[kOffW3 + r * kMlpHid + kk * 8 + k];
        } else if (op < kPkW3T) {                // W2^T: rows in, k out
            const int ib = (op - kPkW2T) >> 3, kk = (op - kPkW2T) & 7;
            w = p[kOffW2 + (kk * 8 + k) * kMlpHid + ib * 32 + r];
        } else {                                  // W3^T: rows in, k out (< 3)
            const int ib = op - kPkW3T;
            if (k < kMlpOut) w = p[kOffW3 + k * kMlpHid + ib * 32 + r];
        }
        v[i] = w;
    }
    pk[op * 64 + lane] = pack4(v[0], v[1], v[2], v[3]);
}

// FP32 master parameters -> BF16 inference operands (one thread per operand
// lane; element i of lane l is K index 8 (l / 32) + i).
__global__ __launch_bounds__(64) void k_mlp_pack_infer(const float *__restrict__ p, bf16x8 *__restrict__ ipk) {
    const int op = blockIdx.x, lane = threadIdx.x;
    const int r = lane & 31, hk = lane >> 5;
    bf16x8 v;
    for (int i = 0; i < 8; i++) {
        const int k = 8 * hk + i;
        float w = 0.f;
        if (op < kIpW2) {                        // W1 split, b1 split: rows out
            const int out = op * 32 + r;
            if (k < 12) {
                const float a = p[kOffW1 + out * kMlpIn + (k & 3)];
                w = k < 8 ? bf16_hi(a) : bf16_lo(a);
            } else if (k < 14) {
                const float b = p[kOffB1 + out];
                w = k == 12 ? bf16_hi(b) : bf16_lo(b);
            }
        } else if (op < kIpB2) {                 // W2: rows out, K chunk c
            const int ob = (op - kIpW2) >> 2, c = (op - kIpW2) & 3;
            w = p[kOffW2 + (ob * 32 + r) * kMlpHid + ihid(c, hk, i)];
        } else if (op < kIpW3) {                 // b2 split at K 0, 1
            const int ob = op - kIpB2;
            if (k < 2) {
                const float b = p[kOffB2 + ob * 32 + r];
                w = k == 0 ? bf16_hi(b) : bf16_lo(b);
            }
        } else {                                  // W3: rows out (< 3), K chunk c
            const int c = op - kIpW3;
            if (r < kMlpOut) w = p[kOffW3 + r * kMlpHid + ihid(c, hk, i)];
        }
        v[i] = to_bf16(w);
    }
    ipk[op * 64 + lane] = v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t h) { return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// Material constants of the plastic projection in FP64.
struct MlpPhys {
    double kappa, mu, ridge, pmax, pmin, ellipse_c, tan_phi, thickness;
};

__device__ __forceinline__ double q_yield_d(double p, double dpt, const MlpPhys &M) {
    if (!(p < M.pmax)) return 0.0;
    const double q_fs = M.ellipse_c * sqrt((M.pmax - p) * (p - M.pmin));
    const double q_dp = fmax(0.0, (p - dpt) * M.tan_phi);
    return fmin(q_fs, q_dp);
}

// The exact plastic projection in FP64 for a trial state (log J, D >= 0) with
// Jp_inv and cohesion c (kernels.cuh:431-479 in (p, q) form).  Returns false
// when the exact map would not take its plastic branch, or, with
// ridge_capped, when the ridge cap is not active (p_tr below it).
__device__ __forceinline__ bool exact_projection_d(double logJ, double D, double jp, double coh, const MlpPhys &M,
                                                   bool ridge_capped, double y[kMlpOut]) {
    const double J = exp(logJ);
    const double p = -0.5 * M.kappa * expm1(2.0 * logJ);
    const double q = 2.0 * M.mu * sinh(D);        // mu |s0^2 - s1^2| / J
    const double dpt = -coh / M.tan_phi;
    if (p < dpt) return false;                    // tension branch
    if (!(q > q_yield_d(p, dpt, M))) return false;   // elastic
    if (ridge_capped && p < M.ridge * M.thickness * jp) return false;
    const double p1 = fmin(p, M.ridge * M.thickness * jp);
    const double q1 = q_yield_d(p1, dpt, M);
    const double Je2 = 1.0 - 2.0 * p1 / M.kappa;  // Je_new^2
    const double Je = sqrt(Je2);
    const double t = (D != 0.0) ? q1 * J / (2.0 * M.mu) : 0.0;
    // s0' = sqrt(Je + t), s1' = sqrt(Je - t)
    y[0] = 0.5 * log((Je + t) * (Je - t));        // log(s0' s1')
    y[1] = atanh(t / Je);                          // log(s0' / s1')
    y[2] = 0.5 * log(Je2) - logJ;                 // log(Jp' / Jp) = log(Je_new / J)
    return true;
}

// Training pairs: trial states that the exact map projects plastically with
// the ridge cap active (p_tr >= RidgeFormationCoeff g rho h Jp_inv: the
// pressure is capped, kernels.cuh:439), drawn by rejection from log J ~ U(lj0, lj1), D ~ logU(d0, d1),
// Jp_inv ~ logU(jp0, jp1), c ~ lognormal(log 1e5, 0.5) capped at 1e6.
struct MlpSampling {
    double lj0, lj1, d0, d1, jp0, jp1;
};

__global__ __launch_bounds__(256) void k_mlp_pairs(MlpPhys M, MlpSampling S, uint64_t seed, uint32_t n,
                                                   float *__restrict__ X, float *__restrict__ T,
                                                   uint32_t *__restrict__ n_failed) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        bool ok = false;
        double x[kMlpIn] = {0, 0, 1, 0}, y[kMlpOut] = {0, 0, 0};
        for (uint32_t a = 0; a < 64 && !ok; a++) {
            const uint64_t h = mix64(seed ^ mix64(((uint64_t)i << 6) | a));
            const double r0 = u01(h), r1 = u01(mix64(h + 1)), r2 = u01(mix64(h + 2));
            const double g1 = u01(mix64(h + 3)), g2 = u01(mix64(h + 4));
            const double logJ = S.lj0 + (S.lj1 - S.lj0) * r0;
            const double D = S.d0 * exp(log(S.d1 / S.d0) * r1);
            const double jp = S.jp0 * exp(log(S.jp1 / S.jp0) * r2);
            const double gauss = sqrt(-2.0 * log(g1)) * cos(6.283185307179586 * g2);
            const double coh = fmin(1e5 * exp(0.5 * gauss), 1e6);
            ok = exact_projection_d(logJ, D, jp, coh, M, true, y) && y[0] < 0.0 && y[1] > 0.0;
            x[0] = logJ;
            x[1] = D;
            x[2] = jp;
            x[3] = log(coh * 1e-5);
        }
        if (!ok) atomicAdd(n_failed, 1u);
        for (int k = 0; k < kMlpIn; k++) X[(size_t)i * kMlpIn + k] = (float)x[k];
        // targets: log(-log J'), log D', log(Jp' / Jp)
        const double t[kMlpOut] = {ok ? log(-y[0]) : 0.0, ok ? log(y[1]) : 0.0, ok ? y[2] : 0.0};
        for (int k = 0; k < kMlpOut; k++) T[(size_t)i * kMlpOut + k] = (float)t[k];
    }
}

// Training pairs from recorded trial states: pair i takes record
// hash(seed, i) mod n_rec (uniform, with replacement); its target is the
// exact FP64 projection, as in k_mlp_pairs.  Records whose state the exact
// map would not project with the ridge cap active count in n_failed.
__global__ __launch_bounds__(256) void k_mlp_pairs_recorded(MlpPhys M, const float *__restrict__ R, uint32_t n_rec,
                                                            uint64_t seed, uint32_t n, float *__restrict__ X,
                                                            float *__restrict__ T, uint32_t *__restrict__ n_failed) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(mix64(seed ^ mix64((uint64_t)i)) % n_rec);
        double x[kMlpIn], y[kMlpOut] = {0, 0, 0};
        for (int k = 0; k < kMlpIn; k++) x[k] = (double)R[(size_t)r * kMlpIn + k];
        const bool ok = exact_projection_d(x[0], x[1], x[2], 1e5 * exp(x[3]), M, true, y) && y[0] < 0.0 && y[1] > 0.0;
        if (!ok) atomicAdd(n_failed, 1u);
        for (int k = 0; k < kMlpIn; k++) X[(size_t)i * kMlpIn + k] = (float)x[k];
        const double t[kMlpOut] = {ok ? log(-y[0]) : 0.0, ok ? log(y[1]) : 0.0, ok ? y[2] : 0.0};
        for (int k = 0; k < kMlpOut; k++) T[(size_t)i * kMlpOut + k] = (float)t[k];
    }
}

// One forward + backward pass over `batch` samples; adds the gradient of the
// mean squared error (standardised targets, mean over samples and outputs)
// into grad and the loss into *loss.  Sample s is row s (seq != 0) or a hashed
// row of the n pairs.  One wave per 64 samples.
__global__ __launch_bounds__(64) void k_mlp_train(MlpDev M, MlpNorm N, const float *__restrict__ X,
                                                  const float *__restrict__ T, uint32_t n, uint32_t batch,
                                                  uint64_t seed, uint32_t step, int seq, float *__restrict__ grad,
                                                  float *__restrict__ loss) {
    __shared__ uint16_t sh_x[8][64];      // [feature][point]
    __shared__ uint16_t sh_h1[64][64];    // [unit][point]
    __shared__ uint16_t sh_h2[64][64];
    __shared__ uint16_t sh_dz1[64][64];
    __shared__ uint16_t sh_dz2[64][64];
    __shared__ uint16_t sh_dy[32][64];    // [output][point], rows >= 3 zero
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x * 64 + lane;
    const bool valid = s < batch;
    const uint32_t row = seq ? s : (uint32_t)(mix64(seed ^ mix64(((uint64_t)step << 32) | s)) % n);
    float xs[kMlpIn] = {0.f, 0.f, 0.f, 0.f}, ts[kMlpOut] = {0.f, 0.f, 0.f};
    if (valid) {
        for (int k = 0; k < kMlpIn; k++) xs[k] = (X[(size_t)row * kMlpIn + k] - N.xm[k]) / N.xs[k];
        for (int k = 0; k < kMlpOut; k++) ts[k] = (T[(size_t)row * kMlpOut + k] - N.ym[k]) / N.ys[k];
    }
    MlpFwd F;
    mlp_forward(M, xs, F);

    // dY in the output layout: lane < 32, element j < 3 is output j of point bb*32 + lane
    const float inv = 2.f / (3.f * (float)batch);
    float lsum = 0.f;
    bf16x4 dyop[2];
#pragma unroll
    for (int bb = 0; bb < 2; bb++) {
        const int src = bb * 32 + (lane & 31);
        const bool v = __shfl((int)valid, src, 64) != 0;
        float d[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float t = __shfl(ts[j], src, 64);
            const float e = (lane < 32 && v) ? F.y[bb][j] - t : 0.f;
            lsum += e * e;
            d[j] = e * inv;
        }
        dyop[bb] = lane < 32 ? pack4(d[0], d[1], d[2], 0.f) : pack4(0.f, 0.f, 0.f, 0.f);
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (lane == 0) atomicAdd(loss, lsum / (3.f * (float)batch));

    const f32x16 zero = {};
    // dZ2 = (W3^T dY) * silu'(Z2)
    bf16x4 dz2op[8][2];
#pragma unroll
    for (int ib = 0; ib < 2; ib++)
#pragma unroll
        for (int bb = 0; bb < 2; bb++) {
            const f32x16 dh = mfma(M.pk[(kPkW3T + ib) * 64 + lane], dyop[bb], zero);
            float d[16];
#pragma unroll
            for (int j = 0; j < 16; j++) d[j] = dh[j] * dsilu(F.z2[ib][bb][j]);
#pragma unroll
            for (int q = 0; q < 4; q++) dz2op[ib * 4 + q][bb] = pack4(d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
        }
    // dZ1 = (W2^T dZ2) * silu'(Z1)
    bf16x4 dz1op[8][2];
#pragma unroll
    for (int ib = 0; ib < 2; ib++)
#pragma unroll
        for (int bb = 0; bb < 2; bb++) {
            f32x16 dh = zero;
#pragma unroll
            for (int kk = 0; kk < 8; kk++) dh = mfma(M.pk[(kPkW2T + ib * 8 + kk) * 64 + lane], dz2op[kk][bb], dh);
            float d[16];
#pragma unroll
            for (int j = 0; j < 16; j++) d[j] = dh[j] * dsilu(F.z1[ib][bb][j]);
#pragma unroll
            for (int q = 0; q < 4; q++) dz1op[ib * 4 + q][bb] = pack4(d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
        }
    // transposes: operand element (k = kk*8 + 4*(lane/32) + i, point = bb*32 + lane%32) -> [k][point]
    const int kq = 4 * (lane >> 5), pt = lane & 31;
#pragma unroll
    for (int bb = 0; bb < 2; bb++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            sh_x[kq + i][bb * 32 + pt] = (uint16_t)F.x[bb][i];
            sh_dy[kq + i][bb * 32 + pt] = (uint16_t)dyop[bb][i];
        }
#pragma unroll
        for (int kk = 0; kk < 8; kk++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int k = kk * 8 + kq + i;
                sh_h1[k][bb * 32 + pt] = (uint16_t)F.h1[kk][bb][i];
                sh_h2[k][bb * 32 + pt] = (uint16_t)F.h2[kk][bb][i];
                sh_dz1[k][bb * 32 + pt] = (uint16_t)dz1op[kk][bb][i];
                sh_dz2[k][bb * 32 + pt] = (uint16_t)dz2op[kk][bb][i];
            }
    }
    // rows 3..7 of sh_dy hold zeros (padding of the operand); clear rows 8..31
    for (int r = 8; r < 32; r++) sh_dy[r][lane] = 0;
    __syncthreads();

    auto opA = [&](const uint16_t(*m)[64], int rowbase, int kk) -> bf16x4 {
        // lane holds m[rowbase + lane%32][kk*8 + 4*(lane/32) .. +3]
        return *reinterpret_cast<const bf16x4 *>(&m[rowbase + pt][kk * 8 + kq]);
    };
    // dW2 [out][in] = sum_points dZ2[out] H1[in]
#pragma unroll
    for (int ob = 0; ob < 2; ob++)
#pragma unroll
        for (int ib = 0; ib < 2; ib++) {
            f32x16 g = zero;
#pragma unroll
            for (int kk = 0; kk < 8; kk++) g = mfma(opA(sh_dz2, ob * 32, kk), opA(sh_h1, ib * 32, kk), g);
#pragma unroll
            for (int j = 0; j < 16; j++)
                unsafeAtomicAdd(&grad[kOffW2 + (ob * 32 + mrow(j, lane)) * kMlpHid + ib * 32 + pt], g[j]);
        }
    // dW1 [out][in < 4] = sum_points dZ1[out] X[in]
#pragma unroll
    for (int ob = 0; ob < 2; ob++) {
        f32x16 g = zero;
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            const bf16x4 b = pt < 8 ? opA(sh_x, 0, kk) : pack4(0.f, 0.f, 0.f, 0.f);
            g = mfma(opA(sh_dz1, ob * 32, kk), b, g);
        }
        if (pt < kMlpIn)
#pragma unroll
            for (int j = 0; j < 16; j++) unsafeAtomicAdd(&grad[kOffW1 + (ob * 32 + mrow(j, lane)) * kMlpIn + pt], g[j]);
    }
    // dW3 [out < 3][in] = sum_points dY[out] H2[in]
#pragma unroll
    for (int ib = 0; ib < 2; ib++) {
        f32x16 g = zero;
#pragma unroll
        for (int kk = 0; kk < 8; kk++) g = mfma(opA(sh_dy, 0, kk), opA(sh_h2, ib * 32, kk), g);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int out = mrow(j, lane);
            if (out < kMlpOut) unsafeAtomicAdd(&grad[kOffW3 + out * kMlpHid + ib * 32 + pt], g[j]);
        }
    }
    // biases: sums of the BF16 deltas the weight gradients use
    {
        float s1 = 0.f, s2 = 0.f;
        for (int k = 0; k < 64; k++) {
            s1 += bf16_float(sh_dz1[lane][k]);
            s2 += bf16_float(sh_dz2[lane][k]);
        }
        unsafeAtomicAdd(&grad[kOffB1 + lane], s1);
        unsafeAtomicAdd(&grad[kOffB2 + lane], s2);
        if (lane < kMlpOut) {
            float s3 = 0.f;
            for (int k = 0; k < 64; k++) s3 += bf16_float(sh_dy[lane][k]);
            unsafeAtomicAdd(&grad[kOffB3 + lane], s3);
        }
    }
}

// Adam (bias-corrected) on the FP32 master parameters; clears the gradient
// it consumed for the next step's accumulation.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ p, float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                                  float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    const float gi = g[i];
    g[i] = 0.f;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// Inference over n rows of raw features: standardised outputs into Y.
__global__ __launch_bounds__(64) void k_mlp_infer(MlpDev M, MlpNorm N, const float *__restrict__ X, uint32_t n,
                                                  float *__restrict__ Y) {
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x * 64 + lane;
    float xs[kMlpIn] = {0.f, 0.f, 0.f, 0.f};
    if (s < n)
        for (int k = 0; k < kMlpIn; k++) xs[k] = (X[(size_t)s * kMlpIn + k] - N.xm[k]) / N.xs[k];
    float y[kMlpOut];
    mlp_infer(M, xs, y);
    if (s < n)
        for (int k = 0; k < kMlpOut; k++) Y[(size_t)s * kMlpOut + k] = y[k];
}

// The surrogate's answer for one point: raw outputs from standardised ones,
// the projected state's (p', q'), the yield violation of the projection, and
// whether it is accepted at tolerance tol (else the exact map runs).  m0, m1
// are the new singular values minus one, jpf the factor on Jp_inv.
//
// jpf is Je_new / J_tr, as the exact map applies it (kernels.cuh:465), with
// Je_new = J' cosh D' from the two outputs the violation test checks, not the
// network's third output (log Jp'/Jp, trained with the others): that output
// is unchecked, and a relative error of a few percent in it, applied on
// every plastic substep, moves Jp_inv (and with it the ridge cap) away from
// the exact run by ~0.5 % over 1000 substeps, while the relation keeps
// Jp_inv as close as the accepted (p', q').
//
// The exact map (kernels.cuh:441-466) picks p1 = min(p_tr, p_ridge) and
// q1 = q_yield(p1), then builds the new singular values from
// Je_new = sqrt(1 - 2 p1 / kappa) and t = q1 J_tr / (2 mu) as
// s0'^2, s1'^2 = Je_new +- t; the new state itself lies off the yield
// surface by O(t^2) and by J_tr / J'.  From the network's log J' and
// D' = log(s0' / s1'): Je_new = J' cosh D' and t = J' sinh D', hence
// p1 = -kappa/2 (J'^2 cosh^2 D' - 1) and q1 = 2 mu J' sinh D' / J_tr, and the
// violation |q1 - q_yield(p1)| / q_yield(p1) is zero for the exact map.
__device__ __forceinline__ bool mlp_decode(const MlpNorm &N, const float y[kMlpOut], float logJ_tr, float jp,
                                           float dpt, float th, const DevParams &P, float tol, float &m0, float &m1,
                                           float &jpf, float &p_out, float &q_out, float &viol) {
    const float lj = -expf(N.ym[0] + N.ys[0] * y[0]);
    const float D = expf(N.ym[1] + N.ys[1] * y[1]);
    p_out = -P.half_kappa * expm1f(2.f * lj);
    q_out = 2.f * P.mu * sinhf(D);
    // log cosh D' (D'^2 / 2 - D'^4 / 12 below 0.1, where coshf would round to 1)
    const float lc = D < 0.1f ? 0.5f * D * D * (1.f - D * D * (1.f / 6.f)) : logf(coshf(D));
    const float p1 = -P.half_kappa * expm1f(2.f * (lj + lc));
    const float q1 = 2.f * P.mu * expf(lj - logJ_tr) * sinhf(D);
    const float qy = q_yield_at(p1, dpt, P);
    viol = fabsf(q1 - qy) / fmaxf(qy, 1.f);
    const float p_ridge = P.ridge * th * jp;
    m0 = expm1f(0.5f * (lj + D));   // new singular values minus one
    m1 = expm1f(0.5f * (lj - D));
    jpf = expf((lj + lc) - logJ_tr);   // Je_new / J_tr
    return viol <= tol && p1 <= p_ridge * (1.f + tol) + 1.f && D >= 0.f && __builtin_isfinite(m0) &&
           __builtin_isfinite(m1) && __builtin_isfinite(jpf);
}

// Surrogate vs exact FP64 for n raw trial states (logJ, D, Jp, log(c/1e5)):
// out[i] = {p'_s, q'_s, p'_e, q'_e, violation, accepted}.
__global__ __launch_bounds__(64) void k_mlp_eval(MlpDev M, MlpNorm N, DevParams P, MlpPhys Mp,
                                                 const float *__restrict__ X, uint32_t n, float tol,
                                                 float *__restrict__ out) {
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x * 64 + lane;
    float xr[kMlpIn] = {0.f, 0.f, 1.f, 0.f}, xs[kMlpIn];
    if (s < n)
        for (int k = 0; k < kMlpIn; k++) xr[k] = X[(size_t)s * kMlpIn + k];
    for (int k = 0; k < kMlpIn; k++) xs[k] = (xr[k] - N.xm[k]) / N.xs[k];
    float y[kMlpOut];
    mlp_infer(M, xs, y);
    if (s >= n) return;
    const float coh = 1e5f * expf(xr[3]);
    const float dpt = -coh * P.inv_tan_phi;
    float m0, m1, jpf, ps, qs, viol;
    const bool ok = mlp_decode(N, y, xr[0], xr[2], dpt, 1.f, P, tol, m0, m1, jpf, ps, qs, viol);
    double ye[kMlpOut];
    double pe = 0.0, qe = 0.0;
    if (exact_projection_d(xr[0], xr[1], xr[2], (double)coh, Mp, false, ye)) {
        pe = -0.5 * Mp.kappa * expm1(2.0 * ye[0]);
        qe = 2.0 * Mp.mu * sinh(ye[1]);
    }
    float *o = out + (size_t)s * 6;
    o[0] = ps;
    o[1] = qs;
    o[2] = (float)pe;
    o[3] = (float)qe;
    o[4] = viol;
    o[5] = ok ? 1.f : 0.f;
}

}  // namespace ice
