// device_math.hpp — FP32 constitutive math of the substep for gfx950.
//
// The reference evaluates Kirchhoff stress, (p, q) and a Givens 2x2 SVD in FP64
// (simulation/kernels.cuh:337-343, 486-535; helper_math.cuh:24-143).  Here
// the state is FP32, so the same quantities are computed from forms that do
// not cancel near F = I:
//
//   F = [[a, b], [c, d]],  E = (a+d)/2, H = (c-b)/2, D = (a-d)/2, G = (b+c)/2
//   F = Q * Rot + R * Refl,  Q = |(E, H)|, R = |(D, G)|
//   Rot  = [[E, -H], [H, E]] / Q      (the polar rotation U V^T)
//   Refl = [[D,  G], [G, -D]] / R
//   sigma0 = Q + R >= sigma1 = Q - R  (sigma1 < 0 when det F < 0)
//
// which is the sorted convention of the reference's SVD (largest first, the
// sign on the smaller magnitude).  Any F' = U diag(s0, s1) V^T with the same
// singular vectors is ((s0+s1)/2) Rot + ((s0-s1)/2) Refl, so the return map
// never forms U and V.  J - 1 is formed from e = F - I, and dev(F F^T) from
// (D, G, H), so stresses keep the relative accuracy of the stored F.
//
// The device stores e = Fe - I in the four Fe lines (upload subtracts I,
// download adds it back), not Fe: near Fe = I the spacing of FP32 is 6e-8,
// while one substep of a slow floe changes the strain by 1e-10 or less, which
// an FP32 Fe would round away.  invariants() reads e, the F update adds
// dt C (I + e), and rebuild() writes e' from the new singular values minus
// one, so no step of the substep forms 1 + (small) and subtracts 1 again.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ice {

constexpr uint32_t kCrushed = 0x10000u;
constexpr uint32_t kDisabled = 0x20000u;
// Bits above the reference's (which uses 0..18) that only this library uses;
// they are cleared before points leave the device.
constexpr uint32_t kMoved = 0x80000u;     // in a cell other than the one it was sorted into
constexpr uint32_t kEscaped = 0x100000u;  // left the 3 x 3 around that cell: handled off-tile
constexpr uint32_t kFresh = 0x200000u;    // escaped in this substep's tile pass: P2G still due
// The drift from the sorted cell, dx and dy in {-1, 0, 1} as two-bit two's
// complement fields (bits 22-23, 24-25): kMoved is set exactly when it is
// nonzero; kFar marks a cell change that took it past one cell.
constexpr uint32_t kDriftBits = 0x3C00000u;
constexpr uint32_t kFar = 0x4000000u;
constexpr uint32_t kInternalBits = kMoved | kEscaped | kFresh | kDriftBits | kFar;

// Adds a cell change (dx, dy) to the drift kept in u.
__device__ __forceinline__ void drift_cell(uint32_t &u, int dx, int dy) {
    const int nx = (((int)(u << 8)) >> 30) + dx;   // bits 22-23, sign-extended
    const int ny = (((int)(u << 6)) >> 30) + dy;   // bits 24-25
    if ((uint32_t)(nx + 1) > 2u || (uint32_t)(ny + 1) > 2u) {   // outside -1..1
        u |= kFar;
        return;
    }
    u = (u & ~(kDriftBits | kMoved)) | ((uint32_t)(nx & 3) << 22) | ((uint32_t)(ny & 3) << 24) |
        ((nx | ny) ? kMoved : 0u);
}

// Constants the kernels read, precomputed on the host in FP64 and rounded
// once to FP32.  Passed by value as a kernel argument.
struct DevParams {
    int gx, gy;                 // grid nodes
    float h, h_inv, dt;
    float dt_h_inv;             // dt / h
    float four_h_inv;           // 4 / h
    float mu, kappa, half_kappa;
    float inv_2mu;              // 1 / (2 mu)
    float inv_half_kappa;       // 2 / kappa
    float particle_mass;        // ParticleVolume * IceDensity
    float dt_vol_Dpinv;
    float vmax2_01;             // 0.1 * vmax^2
    float kL, kQ;               // water drag * dt
    float wind_dt;              // dt * windStressToAccel
    // strength (GetParametersForGrain, kernels.cuh:538-566)
    float pmax, pmin, qmax, pmin2, beta, mSq;
    float fs_k;                 // -q0 / pmin2 (tension line slope)
    float ellipse_c;            // 2 qmax / (pmax - pmin)
    float tan_phi, DP_threshold_p, inv_tan_phi;
    float ridge;                // RidgeFormationCoeff * g * IceDensity
    int has_cohesion;
    int has_current;
    int update_every;           // UpdateEveryNthStep
    double dt_d;                // dt in FP64 for the simulation time
};

struct M2 {
    float a00, a01, a10, a11;
};

// 1/b from the hardware reciprocal and one Newton step (within 1 ulp), in
// place of the 10-instruction IEEE division sequence.
__device__ __forceinline__ float rcp_nr(float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, r, 1.f);
    return __builtin_fmaf(r, e, r);
}

// Hardware square root (within 1 ulp).
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ void weights(float x, float w[3]) {
    // quadratic B-spline weights of nodes -1, 0, +1 (kernels.cuh:323-332)
    float v0 = 0.5f - x, v2 = x + 0.5f;
    w[0] = 0.5f * v0 * v0;
    w[1] = 0.75f - x * x;
    w[2] = 0.5f * v2 * v2;
}

// Invariants of F used by stress, (p, q) and the return map.
struct FInv {
    float E, H, D, G;   // conformal / anti-conformal parts
    float Q, R;         // sigma0 = Q + R, sigma1 = Q - R
    float J, Jm1;       // det F and det F - 1
    float iJ;           // 1 / J
};

// The invariants of F = I + e from e.
__device__ __forceinline__ FInv invariants(const M2 &e) {
    FInv v;
    v.E = 1.f + 0.5f * (e.a00 + e.a11);
    v.H = 0.5f * (e.a10 - e.a01);
    v.D = 0.5f * (e.a00 - e.a11);
    v.G = 0.5f * (e.a01 + e.a10);
    v.Q = sqrt_hw(v.E * v.E + v.H * v.H);
    v.R = sqrt_hw(v.D * v.D + v.G * v.G);
    v.Jm1 = e.a00 + e.a11 + (e.a00 * e.a11 - e.a01 * e.a10);
    v.J = 1.f + v.Jm1;
    v.iJ = rcp_nr(v.J);
    return v;
}

// Kirchhoff stress mu/J dev(F F^T) + kappa/2 (J^2 - 1) I (kernels.cuh:525-535),
// symmetric: returns (t00, t01 = t10, t11).
__device__ __forceinline__ void kirchhoff(const FInv &v, const DevParams &P, float &t00, float &t01, float &t11) {
    const float d = 2.f * (v.E * v.D - v.H * v.G);       // (b00 - b11) / 2
    const float b01 = 2.f * (v.E * v.G + v.D * v.H);
    const float muJ = P.mu * v.iJ;
    const float pi = P.half_kappa * v.Jm1 * (v.J + 1.f);
    t00 = muJ * d + pi;
    t11 = -muJ * d + pi;
    t01 = muJ * b01;
}

// ComputePQ (kernels.cuh:337-343): p = -kappa/2 (J^2-1), q = sqrt2 mu/J |dev(F F^T)|
// with |dev(F F^T)| = 2 sqrt2 Q R.
__device__ __forceinline__ void compute_pq(const FInv &v, const DevParams &P, float &p, float &q) {
    p = -P.half_kappa * v.Jm1 * (v.J + 1.f);
    q = 4.f * P.mu * v.Q * v.R * v.iJ;
}

// CheckIfPointIsInsideFailureSurface (kernels.cuh:349-369)
__device__ __forceinline__ bool outside_failure_surface(float p, float q, const DevParams &P) {
    if (p < 0.f) {
        if (p < P.pmin2) return true;
        return q > P.fs_k * (p - P.pmin2);
    }
    const float y = (1.f + 2.f * P.beta) * q * q + P.mSq * (p + P.beta * P.pmax) * (p - P.pmax);
    return y > 0.f;
}

__device__ __forceinline__ float q_yield_at(float p, float dpt, const DevParams &P) {
    // min(elliptic cap, linear DP), 0 above pmax (kernels.cuh:419-429)
    if (!(p < P.pmax)) return 0.f;
    const float q_fs = P.ellipse_c * sqrt_hw((P.pmax - p) * (p - P.pmin));
    const float q_dp = fmaxf(0.f, (p - dpt) * P.tan_phi);
    return fminf(q_fs, q_dp);
}

// e' = F' - I for F' = ((s0+s1)/2) Rot + ((s0-s1)/2) Refl, from the new
// singular values minus one, m0 = s0 - 1 and m1 = s1 - 1.  The diagonal of
// ((s0+s1)/2) Rot - I is ((m0+m1)/2) cos + (cos - 1), with cos - 1 = (E - Q)/Q
// = -H^2 / ((E + Q) Q) for E > 0, so nothing cancels.
__device__ __forceinline__ M2 rebuild(const FInv &v, float m0, float m1) {
    const float sbm1 = 0.5f * (m0 + m1), sd = 0.5f * (m0 - m1);
    float rc = 1.f, rcm1 = 0.f, rs = 0.f;      // Rot = I when Q = 0 (helper_math.cuh:31-39)
    if (v.Q != 0.f) {
        const float iq = rcp_nr(v.Q);
        rc = v.E * iq;
        rs = v.H * iq;
        rcm1 = (v.E > 0.f ? -v.H * v.H * rcp_nr(v.E + v.Q) : v.E - v.Q) * iq;
    }
    float fc = 0.f, fs = 0.f;
    if (v.R != 0.f) { const float ir = rcp_nr(v.R); fc = v.D * ir; fs = v.G * ir; }
    M2 e;
    e.a00 = sbm1 * rc + rcm1 + sd * fc;
    e.a01 = -(sbm1 * rs + rs) + sd * fs;
    e.a10 = (sbm1 * rs + rs) + sd * fs;
    e.a11 = sbm1 * rc + rcm1 - sd * fc;
    return e;
}

// Wolper_Drucker_Prager (kernels.cuh:374-483) on the invariants; e holds
// Fe - I.  Returns the case: 0 tension with Jp_inv < 0.1, 1 tension, 2
// elastic, 3 plastic.  Each new singular value s enters rebuild() as
// s - 1 = (s^2 - 1) / (s + 1).
__device__ __forceinline__ int return_map(const FInv &v, float p, float q, float thickness, float dpt,
                                          const DevParams &P, M2 &e, float &Jp_inv) {
    if (p < dpt) {
        if (Jp_inv < 0.1f) {
            const float m = v.Jm1 * rcp_nr(sqrt_hw(v.J) + 1.f);   // sqrt(J) - 1
            e = rebuild(v, m, m);
            return 0;
        }
        e = rebuild(v, 0.f, 0.f);
        Jp_inv *= v.iJ;
        return 1;
    }
    if (!(q > q_yield_at(p, dpt, P))) return 2;
    const float p_ridge_max = P.ridge * thickness * Jp_inv;
    const float p1 = fminf(p, p_ridge_max);
    const float q1 = q_yield_at(p1, dpt, P);
    const float Je_new = sqrt_hw(1.f - p1 * P.inv_half_kappa);
    const float jm = -p1 * P.inv_half_kappa * rcp_nr(Je_new + 1.f);   // Je_new - 1
    // normalized(s_hat) = sign(J) (1, -1)/sqrt2 when sigma0^2 != sigma1^2, else 0
    float t = 0.f;
    if (v.Q * v.R != 0.f) t = q1 * fabsf(v.J) * P.inv_2mu;
    const float m0 = (jm + t) * rcp_nr(sqrt_hw(Je_new + t) + 1.f);
    const float m1 = (jm - t) * rcp_nr(sqrt_hw(Je_new - t) + 1.f);
    e = rebuild(v, m0, m1);
    Jp_inv *= Je_new * v.iJ;
    return 3;
}

__device__ __forceinline__ bool isnan4(const M2 &F) {
    return __builtin_isnan(F.a00) || __builtin_isnan(F.a01) || __builtin_isnan(F.a10) || __builtin_isnan(F.a11);
}

// The grid-node update of partition_kernel_update_nodes (kernels.cuh:109-177)
// for one node with mass > 0, plus the wind-stress term of this library
// (after the water drag, before the clamp).  Returns the node velocity;
// *wall is set for sticky nodes, whose momentum the caller adds to fx, fy.
__device__ __forceinline__ void node_velocity(float m, float mx, float my, uint8_t status, float cx, float cy,
                                              float wx, float wy, float ramp, const DevParams &P, float &vx,
                                              float &vy, bool &wall) {
    if (m == 0.f) { vx = mx; vy = my; wall = false; return; }   // untouched node keeps (px, py)
    const float im = rcp_nr(m);
    vx = mx * im;
    vy = my * im;
    wall = status != 100;
    if (wall) {
        vx = 0.f; vy = 0.f;
    } else {
        const float ux = cx * ramp - vx, uy = cy * ramp - vy;
        const float k = fminf(P.kL + P.kQ * sqrt_hw(ux * ux + uy * uy), 0.1f);
        vx += k * ux;
        vy += k * uy;
        vx += P.wind_dt * wx;
        vy += P.wind_dt * wy;
    }
    if (vx * vx + vy * vy > P.vmax2_01) { vx = 0.f; vy = 0.f; }
}

// Time ramp of the water current (kernels.cuh:147-154).
__device__ __forceinline__ float current_ramp(double t) {
    double r;
    if (t < 300000.0) r = 1.0 + fmin(t / (3600.0 * 2), 2.0);
    else r = 3.0 + fmin((t - 300000.0) / (3600.0 * 10), 2.0);
    return (float)r;
}

}  // namespace ice
