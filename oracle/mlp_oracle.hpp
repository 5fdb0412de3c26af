// mlp_oracle.hpp — TEST INFRASTRUCTURE ONLY.  Plain C++ FP64 restatement of
// the surrogate network of the learned plastic projection (a build addition
// with no reference counterpart): 4 -> 64 -> 64 -> 3 MLP, SiLU, forward pass,
// mean-squared-error loss and its gradient by the chain rule.
//
// Mode 1 (training) rounds, like the device's training kernels, every MFMA
// operand to BF16 (round to nearest even): weights, the standardised
// features, the hidden activations, and the backward deltas dY, dZ2, dZ1;
// sums are FP64 here (FP32 on the device).  Bias gradients are sums of the
// rounded deltas, as on the device.  Mode 2 (inference) is the device's
// deployed forward pass: layer 1 unrounded (the device splits its operands
// into BF16 head and tail, about 16 significant bits), layers 2 and 3 with
// BF16 operands.  Mode 0 is the plain FP64 network, which finite differences
// pin.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc_mlp {

constexpr int IN = 4, HID = 64, OUT = 3;
constexpr int W1 = 0, B1 = 256, W2 = 320, B2 = 4416, W3 = 4480, B3 = 4672, NPARAM = 4675;

inline double bf16(double x) {
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

inline double silu(double z) { return z / (1.0 + std::exp(-z)); }
inline double dsilu(double z) {
    const double s = 1.0 / (1.0 + std::exp(-z));
    return s * (1.0 + z * (1.0 - s));
}

struct Norm {
    double xm[IN], xs[IN], ym[OUT], ys[OUT];
};

// Forward of one standardised feature row; keeps what backward needs.
struct Trace {
    double x[IN], z1[HID], h1[HID], z2[HID], h2[HID], y[OUT];
};

inline void forward(const double *p, const double xs[IN], int mode, Trace &T) {
    auto r = [&](double v) { return mode != 0 ? bf16(v) : v; };
    auto r1 = [&](double v) { return mode == 1 ? bf16(v) : v; };   // layer 1
    for (int k = 0; k < IN; k++) T.x[k] = r1(xs[k]);
    for (int o = 0; o < HID; o++) {
        double z = 0;
        for (int k = 0; k < IN; k++) z += r1(p[W1 + o * IN + k]) * T.x[k];
        T.z1[o] = z + p[B1 + o];
        T.h1[o] = r(silu(T.z1[o]));
    }
    for (int o = 0; o < HID; o++) {
        double z = 0;
        for (int k = 0; k < HID; k++) z += r(p[W2 + o * HID + k]) * T.h1[k];
        T.z2[o] = z + p[B2 + o];
        T.h2[o] = r(silu(T.z2[o]));
    }
    for (int o = 0; o < OUT; o++) {
        double z = 0;
        for (int k = 0; k < HID; k++) z += r(p[W3 + o * HID + k]) * T.h2[k];
        T.y[o] = z + p[B3 + o];
    }
}

// Loss = sum over rows and outputs of (y - t)^2 / (3 n); gradient added into g.
inline double loss_grad(const double *p, const Norm &N, const double *x_raw, const double *t_raw, int n, bool b16,
                        double *g) {
    auto r = [&](double v) { return b16 ? bf16(v) : v; };
    double loss = 0;
    Trace T;
    for (int s = 0; s < n; s++) {
        double xs[IN], ts[OUT];
        for (int k = 0; k < IN; k++) xs[k] = (x_raw[s * IN + k] - N.xm[k]) / N.xs[k];
        for (int k = 0; k < OUT; k++) ts[k] = (t_raw[s * OUT + k] - N.ym[k]) / N.ys[k];
        forward(p, xs, b16 ? 1 : 0, T);
        double dy[OUT];
        for (int o = 0; o < OUT; o++) {
            const double e = T.y[o] - ts[o];
            loss += e * e;
            dy[o] = r(2.0 * e / (3.0 * n));
        }
        double dz2[HID], dz1[HID];
        for (int k = 0; k < HID; k++) {
            double dh = 0;
            for (int o = 0; o < OUT; o++) dh += r(p[W3 + o * HID + k]) * dy[o];
            dz2[k] = r(dh * dsilu(T.z2[k]));
        }
        for (int k = 0; k < HID; k++) {
            double dh = 0;
            for (int o = 0; o < HID; o++) dh += r(p[W2 + o * HID + k]) * dz2[o];
            dz1[k] = r(dh * dsilu(T.z1[k]));
        }
        for (int o = 0; o < OUT; o++) {
            for (int k = 0; k < HID; k++) g[W3 + o * HID + k] += dy[o] * T.h2[k];
            g[B3 + o] += dy[o];
        }
        for (int o = 0; o < HID; o++) {
            for (int k = 0; k < HID; k++) g[W2 + o * HID + k] += dz2[o] * T.h1[k];
            g[B2 + o] += dz2[o];
            for (int k = 0; k < IN; k++) g[W1 + o * IN + k] += dz1[o] * T.x[k];
            g[B1 + o] += dz1[o];
        }
    }
    return loss / (3.0 * n);
}

}  // namespace orc_mlp
