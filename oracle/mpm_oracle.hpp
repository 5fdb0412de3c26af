// mpm_oracle.hpp — CPU restatement of the reference substep.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the tests, the smoke check and
// the CPU baseline of bench.py compare against; the product (the HIP library)
// never includes, links or calls it.
//
// It restates igorg520b/iceMPM_2d_sp's substep math, Eigen-free, templated on
// the arithmetic type T (double = the reference precision, float = a plain
// FP32 restatement used to show the FP32 floor).  Each function cites the
// reference file:line it follows.  Two additions beyond the reference are restated
// here too, since the reference does not have them: the wind-stress term in
// the node update and the optional per-particle cohesion (DP threshold).
// One case the reference leaves undefined is given the HIP library's meaning:
// a point whose 3x3 stencil leaves the grid (cell x or y at 0 or at the last
// node) raises error_code_point_left_area, as the reference does
// (kernels.cuh:102-103), but is skipped by P2G and G2P, where the reference
// reads and writes nodes outside the grid arrays.
//
// P2G is deterministic for any thread count: particles are binned into
// x-stripes of 4 columns; even stripes run in parallel, then odd stripes, and
// inside a stripe particles are added in index order.  A particle in stripe k
// touches columns [4k-1, 4k+4], which never overlap the columns of stripe k+2.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <numeric>
#include "../include/icempm.h"
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

constexpr uint32_t kCrushed = 0x10000u, kDisabled = 0x20000u;   // kernels.cuh:11-12
constexpr double kPi = 3.14159265358979323846;                    // parameters_sim.h:40
constexpr double kG = 9.8;                                        // parameters_sim.h:43

template <class T> struct Mat2 {
    T a00, a01, a10, a11;   // row-major (i,j)
};
template <class T> inline Mat2<T> mul(const Mat2<T> &A, const Mat2<T> &B) {
    return {A.a00 * B.a00 + A.a01 * B.a10, A.a00 * B.a01 + A.a01 * B.a11,
            A.a10 * B.a00 + A.a11 * B.a10, A.a10 * B.a01 + A.a11 * B.a11};
}
template <class T> inline T det(const Mat2<T> &A) { return A.a00 * A.a11 - A.a10 * A.a01; }

// Derived constants: SimParams::ComputeLame / ComputeHelperVariables
// (parameters_sim.cpp:122-140).  Restated here rather than read from the
// struct so the oracle does not trust the product's derivation.
struct Derived {
    double lambda, mu, kappa, ParticleMass, cellsize_inv, Dp_inv, dt_vol_Dpinv, vmax;
    int UpdateEveryNthStep;
};
inline Derived derive(const IceParams &p) {
    Derived d;
    d.lambda = p.YoungsModulus * p.PoissonsRatio / ((1 + p.PoissonsRatio) * (1 - 2 * p.PoissonsRatio));
    d.mu = p.YoungsModulus / (2 * (1 + p.PoissonsRatio));
    d.kappa = d.mu * 2. / 3. + d.lambda;
    d.ParticleMass = p.ParticleVolume * p.IceDensity;
    d.UpdateEveryNthStep = (int)(p.AnimationFramePeriod / p.InitialTimeStep);
    d.cellsize_inv = 1. / p.cellsize;
    d.Dp_inv = 4. / (p.cellsize * p.cellsize);
    d.dt_vol_Dpinv = p.InitialTimeStep * p.ParticleVolume * d.Dp_inv;
    d.vmax = 0.25 * p.cellsize / p.InitialTimeStep;
    return d;
}

// CalculateWeightCoeffs (kernels.cuh:323-332): quadratic B-spline weights of
// the nodes at offsets -1, 0, +1 for cell-local coordinate x.
template <class T> inline void weights(T x, T w[3]) {
    T v0 = T(0.5) - x, v2 = x + T(0.5);
    w[0] = T(0.5) * v0 * v0;
    w[1] = T(0.75) - x * x;
    w[2] = T(0.5) * v2 * v2;
}

// dev (kernels.cuh:518-521)
template <class T> inline Mat2<T> dev(const Mat2<T> &A) {
    T h = (A.a00 + A.a11) / T(2);
    return {A.a00 - h, A.a01, A.a10, A.a11 - h};
}

// KirchhoffStress_Wolper (kernels.cuh:525-535):
// tau = mu/J dev(F F^T) + kappa/2 (J^2-1) I
template <class T> inline Mat2<T> kirchhoff(const Mat2<T> &F, T mu, T kappa) {
    T Je = det(F);
    Mat2<T> b{F.a00 * F.a00 + F.a01 * F.a01, F.a00 * F.a10 + F.a01 * F.a11,
              F.a10 * F.a00 + F.a11 * F.a01, F.a10 * F.a10 + F.a11 * F.a11};
    Mat2<T> db = dev(b);
    T s = mu * (T(1) / Je);
    T v = kappa * T(0.5) * (Je * Je - T(1));
    return {s * db.a00 + v, s * db.a01, s * db.a10, s * db.a11 + v};
}

// ComputePQ (kernels.cuh:337-343); coeff1 = sqrt(2) (kernels.cuh:7)
template <class T> inline void compute_pq(const Mat2<T> &F, T mu, T kappa, T &Je, T &p, T &q) {
    const T coeff1 = T(1.4142135623730950);
    Je = det(F);
    p = -(kappa / T(2)) * (Je * Je - T(1));
    Mat2<T> b{F.a00 * F.a00 + F.a01 * F.a01, F.a00 * F.a10 + F.a01 * F.a11,
              F.a10 * F.a00 + F.a11 * F.a01, F.a10 * F.a10 + F.a11 * F.a11};
    Mat2<T> d = dev(b);
    T nrm = std::sqrt(d.a00 * d.a00 + d.a10 * d.a10 + d.a01 * d.a01 + d.a11 * d.a11);
    q = coeff1 * mu * (T(1) / Je) * nrm;
}

// Givens rotation as the reference stores it (givens.cuh:28-209): (c, s).
template <class T> struct Givens { T c = 1, s = 0; };

// svd2x2 (kernels.cuh:486-499) over singular_value_decomposition /
// polar_decomposition (helper_math.cuh:24-45, 78-143) and GivensRotation
// fill / mat_rotation / operator*= (givens.cuh:117-154, 190-197).
// mA.data() is column-major; the << comma initialiser is row-major, so the
// returned U, V are the transposes of the column-major Givens fills.
template <class T> inline void svd2x2(const Mat2<T> &F, Mat2<T> &U, T S[2], Mat2<T> &V) {
    const T a[4] = {F.a00, F.a10, F.a01, F.a11};   // column-major
    // polar_decomposition
    Givens<T> gu;
    T x0 = a[0] + a[3], x1 = a[1] - a[2];
    T den = std::sqrt(x0 * x0 + x1 * x1);
    gu.c = 1; gu.s = 0;
    if (den != 0) { gu.c = x0 / den; gu.s = -x1 / den; }
    T s[4] = {a[0], a[1], a[2], a[3]};
    for (int d = 0; d < 2; d++) {               // mat_rotation<2>, rowi=0, rowk=1
        T t1 = s[0 + d * 2], t2 = s[1 + d * 2];
        s[0 + d * 2] = gu.c * t1 - gu.s * t2;
        s[1 + d * 2] = gu.s * t1 + gu.c * t2;
    }
    // symmetric 2x2 diagonalisation
    T x = s[0], y = s[2], z = s[3];
    T y2 = y * y, cosine, sine;
    if (y2 == 0) {
        cosine = 1; sine = 0; S[0] = x; S[1] = z;
    } else {
        T tau = T(0.5) * (x - z);
        T w = std::sqrt(tau * tau + y2);
        T t = (tau > 0) ? y / (tau + w) : y / (tau - w);
        cosine = T(1) / std::sqrt(t * t + T(1));
        sine = -t * cosine;
        T c2 = cosine * cosine, csy = 2 * cosine * sine * y, s2 = sine * sine;
        S[0] = c2 * x - csy + s2 * z;
        S[1] = s2 * x + csy + c2 * z;
    }
    Givens<T> gv;
    if (S[0] < S[1]) { std::swap(S[0], S[1]); gv.c = -sine; gv.s = cosine; }
    else { gv.c = cosine; gv.s = sine; }
    // u *= v
    T nc = gu.c * gv.c - gu.s * gv.s, ns = gu.s * gv.c + gu.c * gv.s;
    gu.c = nc; gu.s = ns;
    // fill<2> -> {c, s, -s, c}; row-major comma init
    U = {gu.c, gu.s, -gu.s, gu.c};
    V = {gv.c, gv.s, -gv.s, gv.c};
}

// U diag(s) V^T
template <class T> inline Mat2<T> usvt(const Mat2<T> &U, T s0, T s1, const Mat2<T> &V) {
    Mat2<T> US{U.a00 * s0, U.a01 * s1, U.a10 * s0, U.a11 * s1};
    Mat2<T> Vt{V.a00, V.a10, V.a01, V.a11};
    return mul(US, Vt);
}

// Material strength constants in T, with the grain variability of the
// reference fixed at 1 (GetParametersForGrain, kernels.cuh:538-566).
template <class T> struct Strength {
    T pmin, pmax, qmax, beta, mSq, pmin2;
};
template <class T> inline Strength<T> strength(const IceParams &p) {
    Strength<T> s;
    s.pmax = T(p.IceCompressiveStrength);
    s.pmin = T(-p.IceTensileStrength);
    s.qmax = T(p.IceShearStrength);
    s.pmin2 = T(-p.IceTensileStrength2);
    s.beta = T(p.IceTensileStrength / p.IceCompressiveStrength);
    s.mSq = (T(4) * s.qmax * s.qmax * (T(1) + T(2) * s.beta)) / ((s.pmax - s.pmin) * (s.pmax - s.pmin));
    return s;
}

// CheckIfPointIsInsideFailureSurface (kernels.cuh:349-369): latch "crushed"
// when (p, q) leaves the initial failure surface.
template <class T> inline void failure_check(uint32_t &util, T p, T q, const Strength<T> &g) {
    if (p < 0) {
        if (p < g.pmin2) { util |= kCrushed; return; }
        T q0 = T(2) * std::sqrt(-g.pmax * g.pmin) * g.qmax / (g.pmax - g.pmin);
        T k = -q0 / g.pmin2;
        T q_limit = k * (p - g.pmin2);
        if (q > q_limit) { util |= kCrushed; return; }
    } else {
        T y = (T(1) + T(2) * g.beta) * q * q + g.mSq * (p + g.beta * g.pmax) * (p - g.pmax);
        if (y > 0) util |= kCrushed;
    }
}

// Wolper_Drucker_Prager (kernels.cuh:374-483).  DP_threshold_p is passed in:
// the reference's global value, or -cohesion/tan(phi) with the per-particle
// cohesion line.  Returns the case (0..3; 2 = elastic) and sets *nan_flag
// when Fe came out NaN (kernels.cuh:470-479).
template <class T>
inline int return_map(const IceParams &prm, T mu, T kappa, T DP_threshold_p, T thickness,
                      T p_tr, T q_tr, T Je_tr, const Mat2<T> &U, const Mat2<T> &V,
                      const T s_hat[2], Mat2<T> &Fe, T &Jp_inv, bool *nan_flag) {
    const T coeff1_inv = T(0.7071067811865475);   // kernels.cuh:8
    const T pmin = T(-prm.IceTensileStrength);
    const T pmax = T(prm.IceCompressiveStrength);
    const T qmax = T(prm.IceShearStrength);
    const T tan_phi = std::tan(T(prm.DP_phi) * T(kPi) / T(180));
    int case1;
    if (p_tr < DP_threshold_p) {
        if (Jp_inv < T(0.1)) {
            case1 = 0;
            T sq = std::sqrt(Je_tr);
            Fe = usvt(U, sq, sq, V);
        } else {
            case1 = 1;
            Fe = usvt(U, T(1), T(1), V);
            Jp_inv /= Je_tr;
        }
    } else {
        case1 = 2;
        T q_yield;
        if (p_tr < pmax) {
            T q_fs = T(2) * std::sqrt((pmax - p_tr) * (p_tr - pmin)) * qmax / (pmax - pmin);
            T q_dp = std::max(T(0), (p_tr - DP_threshold_p) * tan_phi);
            q_yield = std::min(q_fs, q_dp);
        } else {
            q_yield = 0;
        }
        if (q_tr > q_yield) {
            case1 = 3;
            const T p_ridge_max = T(prm.RidgeFormationCoeff) * T(kG) * T(prm.IceDensity) * thickness * Jp_inv;
            T p_n_1 = std::min(p_tr, p_ridge_max);
            T q_n_1;
            if (p_n_1 < pmax) {
                T q_dp = std::max(T(0), (p_n_1 - DP_threshold_p) * tan_phi);
                T q_fs = T(2) * std::sqrt((pmax - p_n_1) * (p_n_1 - pmin)) * qmax / (pmax - pmin);
                q_n_1 = std::min(q_fs, q_dp);
            } else {
                q_n_1 = 0;
            }
            const T Je_new = std::sqrt(T(-2) * p_n_1 / kappa + T(1));
            const T s_norm = q_n_1 * coeff1_inv;
            // Eigen normalized(): unchanged when the squared norm is 0
            T n0 = s_hat[0], n1 = s_hat[1];
            T z = n0 * n0 + n1 * n1;
            if (z > 0) { T r = std::sqrt(z); n0 /= r; n1 /= r; }
            T b0 = s_norm * (Je_tr / mu) * n0 + Je_new;
            T b1 = s_norm * (Je_tr / mu) * n1 + Je_new;
            Fe = usvt(U, std::sqrt(b0), std::sqrt(b1), V);
            Jp_inv *= Je_new / Je_tr;
        }
    }
    *nan_flag = std::isnan(Fe.a00) || std::isnan(Fe.a10) || std::isnan(Fe.a01) || std::isnan(Fe.a11);
    return case1;
}

// One point of the G2P constitutive part: ComputePQ, ComputeSVD, failure
// check, return map (kernels.cuh:267-282).  Returns the return-map case or -1
// when the point stayed uncrushed.
template <class T>
inline int constitutive(const IceParams &prm, const Derived &d, const Strength<T> &g, T DP_threshold_p,
                        T thickness, Mat2<T> &Fe, T &Jp_inv, uint32_t &util, T &p_tr, T &q_tr,
                        uint32_t &err) {
    const T mu = T(d.mu), kappa = T(d.kappa);
    T Je_tr;
    compute_pq(Fe, mu, kappa, Je_tr, p_tr, q_tr);
    Mat2<T> U, V;
    T S[2];
    svd2x2(Fe, U, S, V);
    T S2[2] = {S[0] * S[0], S[1] * S[1]};
    // ComputeSVD / dev_d (kernels.cuh:501-516)
    T half = (S2[0] + S2[1]) / T(2);
    T s_hat[2] = {mu / Je_tr * (S2[0] - half), mu / Je_tr * (S2[1] - half)};
    if (!(util & kCrushed)) failure_check(util, p_tr, q_tr, g);
    int c = -1;
    if (util & kCrushed) {
        bool nanf = false;
        c = return_map(prm, mu, kappa, DP_threshold_p, thickness, p_tr, q_tr, Je_tr, U, V, s_hat, Fe, Jp_inv, &nanf);
        if (nanf) err |= ICE_ERR_POINT_FE_NAN;
    }
    return c;
}

// Whole-domain state in the oracle's own layout (separate integer arrays).
template <class T> struct State {
    IceParams prm;
    Derived d;
    Strength<T> g;
    bool has_cohesion = false;
    int gx = 0, gy = 0;
    size_t n = 0;
    // points
    std::vector<uint32_t> util, cell, pidx;
    std::vector<T> P, Q, Jp, px, py, vx, vy, F00, F01, F10, F11, B00, B01, B10, B11, thick, coh;
    // grid
    std::vector<T> gm, gpx, gpy, gvx, gvy, gfx, gfy, cur_x, cur_y, wind_x, wind_y;
    // time-varying wind: wind_x/y hold the frame w0, wind_dx/dy = w1 - w0
    std::vector<T> wind_dx, wind_dy;
    bool wind_frames = false;
    double wind_t0 = 0, wind_t1 = 1;
    std::vector<uint8_t> status;
    uint32_t error_code = 0;
    int disabled = 0;
    // stripe binning scratch
    std::vector<uint32_t> bin_start, bin_list;

    void init(const IceParams &p) {
        prm = p;
        d = derive(p);
        g = strength<T>(p);
        gx = p.GridXTotal; gy = p.GridYTotal;
        has_cohesion = (p.flags & ICE_FLAG_COHESION_LINE) != 0;
        size_t N = (size_t)gx * gy;
        for (auto *v : {&gm, &gpx, &gpy, &gvx, &gvy, &gfx, &gfy, &cur_x, &cur_y, &wind_x, &wind_y, &wind_dx, &wind_dy})
            v->assign(N, T(0));
        status.assign(N, (uint8_t)ICE_GRID_STATUS_WATER);
    }
    void resize(size_t np) {
        n = np;
        for (auto *v : {&util, &cell, &pidx}) v->assign(np, 0u);
        for (auto *v : {&P, &Q, &Jp, &px, &py, &vx, &vy, &F00, &F01, &F10, &F11, &B00, &B01, &B10, &B11, &thick, &coh})
            v->assign(np, T(0));
    }
    T dp_threshold(size_t i) const {
        if (!has_cohesion) return T(prm.DP_threshold_p);
        const T tan_phi = std::tan(T(prm.DP_phi) * T(kPi) / T(180));
        return -coh[i] / tan_phi;
    }

    // reset_grid (gpu_partition.cpp:300-308)
    void reset_grid() {
        std::fill(gm.begin(), gm.end(), T(0));
        std::fill(gpx.begin(), gpx.end(), T(0));
        std::fill(gpy.begin(), gpy.end(), T(0));
    }
    // clear_force_accumulator (gpu_partition.cpp:310-319)
    void clear_forces() {
        std::fill(gfx.begin(), gfx.end(), T(0));
        std::fill(gfy.begin(), gfy.end(), T(0));
    }

    // P2G scatter of one point (kernels.cuh:39-104)
    void p2g_point(size_t i, uint32_t &err) {
        if (util[i] & kDisabled) return;
        const T h = T(prm.cellsize);
        const T mu = T(d.mu), kappa = T(d.kappa);
        Mat2<T> Fe{F00[i], F01[i], F10[i], F11[i]};
        Mat2<T> Bp{B00[i], B01[i], B10[i], B11[i]};
        const T particle_mass = T(d.ParticleMass) * thick[i];
        const int cx = (int)(cell[i] & 0xffff), cy = (int)(cell[i] >> 16);
        if (cx < 1 || cy < 1 || cx > gx - 2 || cy > gy - 2) {   // kernels.cuh:102-103; no scatter (see above)
            err |= ICE_ERR_POINT_LEFT_AREA;
            return;
        }
        Mat2<T> PFt = kirchhoff(Fe, mu, kappa);
        const T k = T(d.dt_vol_Dpinv) * Jp[i] * thick[i];
        Mat2<T> st{particle_mass * Bp.a00 - k * PFt.a00, particle_mass * Bp.a01 - k * PFt.a01,
                   particle_mass * Bp.a10 - k * PFt.a10, particle_mass * Bp.a11 - k * PFt.a11};
        T wx[3], wy[3];
        weights(px[i], wx);
        weights(py[i], wy);
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++) {
                const T Wip = wx[a + 1] * wy[b + 1];
                const T incM = Wip * particle_mass;
                const T dx = (T(a) - px[i]) * h, dy = (T(b) - py[i]) * h;
                const T incVx = Wip * (vx[i] * particle_mass + (st.a00 * dx + st.a01 * dy));
                const T incVy = Wip * (vy[i] * particle_mass + (st.a10 * dx + st.a11 * dy));
                const size_t node = (size_t)(b + cy) + (size_t)(a + cx) * gy;
                gm[node] += incM;
                gpx[node] += incVx;
                gpy[node] += incVy;
                if (std::isnan(incVx) || std::isnan(incVy)) err |= ICE_ERR_GRID_P2G_NAN_VEL;
                if (std::isnan(incM)) err |= ICE_ERR_GRID_P2G_NAN_MASS;
            }
    }

    void p2g() {
        const int SW = 4;
        const int nstripes = (gx + SW - 1) / SW + 1;
        bin_start.assign(nstripes + 1, 0);
        for (size_t i = 0; i < n; i++) {
            int cx = (int)(cell[i] & 0xffff);
            int s = std::min(std::max(cx, 0) / SW, nstripes - 1);
            bin_start[s + 1]++;
        }
        for (int s = 0; s < nstripes; s++) bin_start[s + 1] += bin_start[s];
        bin_list.resize(n);
        {
            std::vector<uint32_t> fill(bin_start.begin(), bin_start.end() - 1);
            for (size_t i = 0; i < n; i++) {
                int cx = (int)(cell[i] & 0xffff);
                int s = std::min(std::max(cx, 0) / SW, nstripes - 1);
                bin_list[fill[s]++] = (uint32_t)i;
            }
        }
        uint32_t err = 0;
        for (int colour = 0; colour < 2; colour++) {
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
            for (int s = colour; s < nstripes; s += 2)
                for (uint32_t k = bin_start[s]; k < bin_start[s + 1]; k++) p2g_point(bin_list[k], err);
        }
        error_code |= err;
    }

    // partition_kernel_update_nodes (kernels.cuh:109-177) plus the wind-stress
    // term (not in the reference): v += dt * windStressToAccel * tau_wind on water nodes,
    // after the water drag and before the velocity clamp.  With time-varying
    // wind, tau_wind = w0 + s (w1 - w0), s = clamp((t - t0) / (t1 - t0), 0, 1):
    // the time interpolation (1 - tb) w0 + tb w1 of the reference's unfinished
    // wind interpolation (kernels.cuh:580-622, commented out;
    // windandcurrentinterpolator.cpp:66-70).
    void update_nodes(double simulation_time) {
        const T dt = T(prm.InitialTimeStep);
        const T vmax = T(d.vmax);
        const T kL = T(prm.waterDragEffectiveLinear * prm.InitialTimeStep);
        const T kQp = T(prm.waterDragEffectiveQuadratic * prm.InitialTimeStep);
        const T wacc = T(prm.windStressToAccel);
        const T t = T(simulation_time);
        const T ws = wind_frames ? T(std::min(std::max((simulation_time - wind_t0) / (wind_t1 - wind_t0), 0.0), 1.0))
                                 : T(0);
        T ramp;
        if (t < T(300000)) ramp = T(1) + std::min(t / T(3600 * 2), T(2));
        else ramp = T(3) + std::min((t - T(300000)) / T(3600 * 10), T(2));
        const size_t N = (size_t)gx * gy;
        uint32_t err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
        for (size_t idx = 0; idx < N; idx++) {
            const T mass = gm[idx];
            if (mass == 0) { gvx[idx] = gpx[idx]; gvy[idx] = gpy[idx]; continue; }
            const T mx = gpx[idx], my = gpy[idx];
            T vx_ = mx / mass, vy_ = my / mass;
            if (status[idx] != ICE_GRID_STATUS_WATER) {
                vx_ = 0; vy_ = 0;
                gfx[idx] += mx;
                gfy[idx] += my;
            } else {
                const T wx = cur_x[idx] * ramp, wy = cur_y[idx] * ramp;
                const T ux = wx - vx_, uy = wy - vy_;
                const T umag = std::sqrt(ux * ux + uy * uy);
                T k = kL + kQp * umag;
                k = std::min(k, T(0.1));
                vx_ += k * ux;
                vy_ += k * uy;
                const T tx = wind_frames ? wind_x[idx] + ws * wind_dx[idx] : wind_x[idx];
                const T ty = wind_frames ? wind_y[idx] + ws * wind_dy[idx] : wind_y[idx];
                vx_ += dt * wacc * tx;
                vy_ += dt * wacc * ty;
            }
            if (vx_ * vx_ + vy_ * vy_ > vmax * vmax * T(0.1)) { vx_ = 0; vy_ = 0; }
            gvx[idx] = vx_;
            gvy[idx] = vy_;
            if (std::isnan(vx_) || std::isnan(vy_)) err |= ICE_ERR_GRID_NAN;
        }
        error_code |= err;
    }

    // partition_kernel_g2p (kernels.cuh:181-315)
    void g2p(bool recordPQ) {
        const T h_inv = T(d.cellsize_inv);
        const T dt = T(prm.InitialTimeStep);
        uint32_t err = 0;
        int ndis = 0;
#pragma omp parallel for schedule(static) reduction(| : err) reduction(+ : ndis)
        for (size_t i = 0; i < n; i++) {
            uint32_t u = util[i];
            if (u & kDisabled) continue;
            T x = px[i], y = py[i];
            Mat2<T> Fe{F00[i], F01[i], F10[i], F11[i]};
            T jp = Jp[i];
            int cx = (int)(cell[i] & 0xffff), cy = (int)(cell[i] >> 16);
            if (cx < 1 || cy < 1 || cx > gx - 2 || cy > gy - 2) {   // stencil outside the grid: skipped (see above)
                err |= ICE_ERR_POINT_LEFT_AREA;
                continue;
            }
            T wx[3], wy[3];
            weights(x, wx);
            weights(y, wy);
            T pvx = 0, pvy = 0;
            Mat2<T> C{0, 0, 0, 0};
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++) {
                    const T dpx = T(a) - x, dpy = T(b) - y;
                    const T w = wx[a + 1] * wy[b + 1];
                    const size_t node = (size_t)(b + cy) + (size_t)(a + cx) * gy;
                    const T nvx = gvx[node], nvy = gvy[node];
                    pvx += w * nvx;
                    pvy += w * nvy;
                    const T f = T(4) * h_inv * w;
                    C.a00 += f * (nvx * dpx);
                    C.a01 += f * (nvx * dpy);
                    C.a10 += f * (nvy * dpx);
                    C.a11 += f * (nvy * dpy);
                }
            x += pvx * (dt * h_inv);
            y += pvy * (dt * h_inv);
            if (std::isnan(pvx) || std::isnan(pvy)) err |= ICE_ERR_POINT_VEL_NAN;
            if (std::isnan(x) || std::isnan(y)) err |= ICE_ERR_POINT_POS_NAN;
            if (std::isnan(C.a00) || std::isnan(C.a10) || std::isnan(C.a01) || std::isnan(C.a11))
                err |= ICE_ERR_POINT_BP_NAN;
            bool cell_updated = false;
            if (x > T(0.5)) { x -= T(1); cx++; cell_updated = true; }
            else if (x < T(-0.5)) { x += T(1); cx--; cell_updated = true; }
            if (y > T(0.5)) { y -= T(1); cy++; cell_updated = true; }
            else if (y < T(-0.5)) { y += T(1); cy--; cell_updated = true; }
            if (cell_updated) {
                if (cx <= 1 || cx >= gx - 2 || cy <= 1 || cy >= gy - 2) {
                    u |= kDisabled;
                    ndis++;
                }
            }
            if (x > T(0.5) || x < T(-0.5) || y > T(0.5) || y < T(-0.5)) err |= ICE_ERR_POINT_JUMP_CELLS;
            Mat2<T> A{T(1) + dt * C.a00, dt * C.a01, dt * C.a10, T(1) + dt * C.a11};
            Fe = mul(A, Fe);
            if (std::isnan(Fe.a00) || std::isnan(Fe.a10) || std::isnan(Fe.a01) || std::isnan(Fe.a11))
                err |= ICE_ERR_POINT_FE_NAN;
            T p_tr, q_tr;
            constitutive(prm, d, g, dp_threshold(i), thick[i], Fe, jp, u, p_tr, q_tr, err);
            px[i] = x; py[i] = y;
            vx[i] = pvx; vy[i] = pvy;
            F00[i] = Fe.a00; F01[i] = Fe.a01; F10[i] = Fe.a10; F11[i] = Fe.a11;
            B00[i] = C.a00; B01[i] = C.a01; B10[i] = C.a10; B11[i] = C.a11;
            Jp[i] = jp;
            util[i] = u;
            if (cell_updated) cell[i] = ((uint32_t)cy << 16) | (uint32_t)cx;
            if (recordPQ) { P[i] = p_tr; Q[i] = q_tr; }
        }
        error_code |= err;
        disabled += ndis;
    }

    // One substep as Model::Step sequences it (model.cpp:18-39).
    void substep(double t, bool recordPQ) {
        reset_grid();
        p2g();
        update_nodes(t);
        g2p(recordPQ);
    }

    // RemoveDisabledAndSort (host_side_soa.cpp:36-47) with a stable order:
    // remove_if keeps relative order; stable_sort by x*GridY+y
    // (ProxyPoint::getCellIndex, proxypoint2d.cpp:117-123).
    std::vector<uint32_t> sort_permutation() const {
        std::vector<uint32_t> keep;
        keep.reserve(n);
        for (size_t i = 0; i < n; i++)
            if (!(util[i] & kDisabled)) keep.push_back((uint32_t)i);
        auto key = [&](uint32_t i) {
            uint32_t c = cell[i];
            return (uint64_t)(c & 0xffff) * (uint64_t)gy + (uint64_t)(c >> 16);
        };
        std::stable_sort(keep.begin(), keep.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
        return keep;
    }
    void apply_permutation(const std::vector<uint32_t> &perm) {
        auto pu = [&](std::vector<uint32_t> &v) { std::vector<uint32_t> o(perm.size()); for (size_t k = 0; k < perm.size(); k++) o[k] = v[perm[k]]; v.swap(o); };
        auto pt = [&](std::vector<T> &v) { std::vector<T> o(perm.size()); for (size_t k = 0; k < perm.size(); k++) o[k] = v[perm[k]]; v.swap(o); };
        pu(util); pu(cell); pu(pidx);
        for (auto *v : {&P, &Q, &Jp, &px, &py, &vx, &vy, &F00, &F01, &F10, &F11, &B00, &B01, &B10, &B11, &thick, &coh}) pt(*v);
        n = perm.size();
        disabled = 0;
    }
};

}  // namespace orc
