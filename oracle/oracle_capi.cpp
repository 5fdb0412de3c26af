// oracle_capi.cpp — extern "C" face of the oracle for the Python tests and the
// CPU baseline.  TEST INFRASTRUCTURE ONLY: never linked into the product.
#include <type_traits>
#include "mpm_oracle.hpp"
#include "mlp_oracle.hpp"
#include "frame_oracle.hpp"
#include <chrono>

using namespace orc;

namespace {

struct Any {
    int precision;
    State<double> *s64 = nullptr;
    State<float> *s32 = nullptr;
};

template <class T> State<T> *get(Any *a);
template <> State<double> *get<double>(Any *a) { return a->s64; }
template <> State<float> *get<float>(Any *a) { return a->s32; }

inline uint32_t as_u32(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float as_f32(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

template <class T> void set_points(State<T> &s, const float *soa, size_t pitch, uint32_t n, int nlines) {
    s.resize(n);
    auto L = [&](int line, size_t i) { return soa[(size_t)line * pitch + i]; };
    for (size_t i = 0; i < n; i++) {
        s.util[i] = as_u32(L(ICE_IDX_UTILITY, i));
        s.cell[i] = as_u32(L(ICE_IDX_CELL, i));
        s.pidx[i] = as_u32(L(ICE_IDX_POINT, i));
        s.P[i] = L(ICE_IDX_P, i); s.Q[i] = L(ICE_IDX_Q, i);
        s.Jp[i] = L(ICE_IDX_JP_INV, i);
        s.px[i] = L(ICE_IDX_POSX, i); s.py[i] = L(ICE_IDX_POSX + 1, i);
        s.vx[i] = L(ICE_IDX_VELX, i); s.vy[i] = L(ICE_IDX_VELX + 1, i);
        s.F00[i] = L(ICE_IDX_FE00, i); s.F01[i] = L(ICE_IDX_FE00 + 1, i);
        s.F10[i] = L(ICE_IDX_FE00 + 2, i); s.F11[i] = L(ICE_IDX_FE00 + 3, i);
        s.B00[i] = L(ICE_IDX_BP00, i); s.B01[i] = L(ICE_IDX_BP00 + 1, i);
        s.B10[i] = L(ICE_IDX_BP00 + 2, i); s.B11[i] = L(ICE_IDX_BP00 + 3, i);
        s.thick[i] = L(ICE_IDX_THICKNESS, i);
        s.coh[i] = nlines > ICE_IDX_COHESION ? L(ICE_IDX_COHESION, i) : 0.f;
    }
    s.disabled = 0;
    for (size_t i = 0; i < n; i++) s.disabled += (s.util[i] & kDisabled) ? 1 : 0;
}

// Float lines as T; integer lines as their exact value in the out type.
template <class T, class O> void get_points(const State<T> &s, O *out, size_t pitch, int nlines, bool bits) {
    auto W = [&](int line, size_t i, O v) { out[(size_t)line * pitch + i] = v; };
    auto I = [&](uint32_t u) -> O {
        if (bits) { float f = as_f32(u); return (O)f; }
        return (O)u;
    };
    for (size_t i = 0; i < s.n; i++) {
        W(ICE_IDX_UTILITY, i, I(s.util[i]));
        W(ICE_IDX_CELL, i, I(s.cell[i]));
        W(ICE_IDX_POINT, i, I(s.pidx[i]));
        W(ICE_IDX_P, i, (O)s.P[i]); W(ICE_IDX_Q, i, (O)s.Q[i]);
        W(ICE_IDX_JP_INV, i, (O)s.Jp[i]);
        W(ICE_IDX_POSX, i, (O)s.px[i]); W(ICE_IDX_POSX + 1, i, (O)s.py[i]);
        W(ICE_IDX_VELX, i, (O)s.vx[i]); W(ICE_IDX_VELX + 1, i, (O)s.vy[i]);
        W(ICE_IDX_FE00, i, (O)s.F00[i]); W(ICE_IDX_FE00 + 1, i, (O)s.F01[i]);
        W(ICE_IDX_FE00 + 2, i, (O)s.F10[i]); W(ICE_IDX_FE00 + 3, i, (O)s.F11[i]);
        W(ICE_IDX_BP00, i, (O)s.B00[i]); W(ICE_IDX_BP00 + 1, i, (O)s.B01[i]);
        W(ICE_IDX_BP00 + 2, i, (O)s.B10[i]); W(ICE_IDX_BP00 + 3, i, (O)s.B11[i]);
        W(ICE_IDX_THICKNESS, i, (O)s.thick[i]);
        if (nlines > ICE_IDX_COHESION) W(ICE_IDX_COHESION, i, (O)s.coh[i]);
    }
}

template <class T> void mat_in(const T F[4], Mat2<T> &M) { M = {F[0], F[1], F[2], F[3]}; }
template <class T> void mat_out(const Mat2<T> &M, T F[4]) { F[0] = M.a00; F[1] = M.a01; F[2] = M.a10; F[3] = M.a11; }

}  // namespace

extern "C" {

int orc_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void *orc_create(const IceParams *p, int precision) {
    Any *a = new Any;
    a->precision = precision;
    if (precision == 32) { a->s32 = new State<float>; a->s32->init(*p); }
    else { a->s64 = new State<double>; a->s64->init(*p); }
    return a;
}

void orc_destroy(void *h) {
    Any *a = (Any *)h;
    delete a->s32;
    delete a->s64;
    delete a;
}

#define DISPATCH(h, expr)                                   \
    do {                                                    \
        Any *a_ = (Any *)(h);                               \
        if (a_->precision == 32) { auto &s = *a_->s32; expr; } \
        else { auto &s = *a_->s64; expr; }                  \
    } while (0)

int orc_set_points(void *h, const float *soa, size_t pitch, uint32_t n, int nlines) {
    DISPATCH(h, set_points(s, soa, pitch, n, nlines));
    return 0;
}

// status [gx*gy]; current, wind: [2][gx*gy] floats or NULL
int orc_set_grid(void *h, const uint8_t *status, const float *current, const float *wind) {
    DISPATCH(h, {
        size_t N = (size_t)s.gx * s.gy;
        if (status) std::copy(status, status + N, s.status.begin());
        for (size_t i = 0; i < N; i++) {
            s.cur_x[i] = current ? current[i] : 0; s.cur_y[i] = current ? current[N + i] : 0;
            s.wind_x[i] = wind ? wind[i] : 0; s.wind_y[i] = wind ? wind[N + i] : 0;
        }
        s.wind_frames = false;
    });
    return 0;
}

// time-varying wind: frames w0, w1 [2][gx*gy] floats valid at t0 < t1
int orc_set_wind_frames(void *h, const float *w0, const float *w1, double t0, double t1) {
    if (!(t1 > t0)) return 1;
    DISPATCH(h, {
        size_t N = (size_t)s.gx * s.gy;
        using VT = typename std::decay<decltype(s.wind_dx)>::type::value_type;   // the state's real type
        for (size_t i = 0; i < N; i++) {
            s.wind_x[i] = w0[i]; s.wind_y[i] = w0[N + i];
            s.wind_dx[i] = VT(w1[i]) - s.wind_x[i];
            s.wind_dy[i] = VT(w1[N + i]) - s.wind_y[i];
        }
        s.wind_frames = true;
        s.wind_t0 = t0;
        s.wind_t1 = t1;
    });
    return 0;
}

int orc_substep(void *h, double t, int recordPQ) {
    DISPATCH(h, s.substep(t, recordPQ != 0));
    return 0;
}

// Model::Step's substep loop (model.cpp:18-39) for nsteps substeps.
int orc_run(void *h, int first_step, int nsteps) {
    DISPATCH(h, {
        for (int k = 0; k < nsteps; k++) {
            int step = first_step + k;
            double t = s.prm.InitialTimeStep * step;
            bool rec = s.d.UpdateEveryNthStep > 0 && step % s.d.UpdateEveryNthStep == 0;
            s.substep(t, rec);
        }
    });
    return 0;
}

double orc_bench(void *h, int first_step, int nsteps) {
    auto t0 = std::chrono::steady_clock::now();
    orc_run(h, first_step, nsteps);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// Points in the float hssoa layout (integer lines as bit patterns).
int orc_get_points(void *h, float *soa, size_t pitch, int nlines) {
    DISPATCH(h, {
        get_points(s, soa, pitch, nlines, false);
        for (size_t i = 0; i < s.n; i++) {
            soa[(size_t)ICE_IDX_UTILITY * pitch + i] = as_f32(s.util[i]);
            soa[(size_t)ICE_IDX_CELL * pitch + i] = as_f32(s.cell[i]);
            soa[(size_t)ICE_IDX_POINT * pitch + i] = as_f32(s.pidx[i]);
        }
    });
    return 0;
}

// Points as doubles; integer lines as exact integer values.
int orc_get_points_f64(void *h, double *out, size_t pitch, int nlines) {
    DISPATCH(h, get_points(s, out, pitch, nlines, false));
    return 0;
}

uint32_t orc_count(void *h) {
    uint32_t n = 0;
    DISPATCH(h, n = (uint32_t)s.n);
    return n;
}

// mpp [3][N] m, px, py after P2G; vel [2][N] after the update; forces [2][N]
int orc_get_grid(void *h, double *mpp, double *vel, double *forces) {
    DISPATCH(h, {
        size_t N = (size_t)s.gx * s.gy;
        for (size_t i = 0; i < N; i++) {
            if (mpp) { mpp[i] = s.gm[i]; mpp[N + i] = s.gpx[i]; mpp[2 * N + i] = s.gpy[i]; }
            if (vel) { vel[i] = s.gvx[i]; vel[N + i] = s.gvy[i]; }
            if (forces) { forces[i] = s.gfx[i]; forces[N + i] = s.gfy[i]; }
        }
    });
    return 0;
}

// Overwrite the node masses/momenta (after P2G) or node velocities (after the
// update), for phase-level tests.  Each [N] doubles or NULL.
int orc_set_grid_state(void *h, const double *m, const double *px, const double *py, const double *vx,
                       const double *vy) {
    DISPATCH(h, {
        size_t N = (size_t)s.gx * s.gy;
        for (size_t i = 0; i < N; i++) {
            if (m) s.gm[i] = m[i];
            if (px) s.gpx[i] = px[i];
            if (py) s.gpy[i] = py[i];
            if (vx) s.gvx[i] = vx[i];
            if (vy) s.gvy[i] = vy[i];
        }
    });
    return 0;
}

int orc_clear_forces(void *h) {
    DISPATCH(h, s.clear_forces());
    return 0;
}

uint32_t orc_error(void *h) {
    uint32_t e = 0;
    DISPATCH(h, e = s.error_code);
    return e;
}
int orc_disabled(void *h) {
    int d = 0;
    DISPATCH(h, d = s.disabled);
    return d;
}

// Separate phases, for phase-level parity.
int orc_reset_grid(void *h) { DISPATCH(h, s.reset_grid()); return 0; }
int orc_p2g(void *h) { DISPATCH(h, s.p2g()); return 0; }
int orc_update_nodes(void *h, double t) { DISPATCH(h, s.update_nodes(t)); return 0; }
int orc_g2p(void *h, int recordPQ) { DISPATCH(h, s.g2p(recordPQ != 0)); return 0; }

// RemoveDisabledAndSort with std::stable_sort order; perm[k] = input slot.
int orc_sort(void *h, uint32_t *n_out, uint32_t *perm) {
    DISPATCH(h, {
        std::vector<uint32_t> p = s.sort_permutation();
        if (perm) std::copy(p.begin(), p.end(), perm);
        s.apply_permutation(p);
        *n_out = (uint32_t)p.size();
    });
    return 0;
}

// ---- device-function level entry points ------------------------------------

void orc_params_derive(IceParams *p) {
    Derived d = derive(*p);
    p->lambda = d.lambda; p->mu = d.mu; p->kappa = d.kappa;
    p->ParticleMass = d.ParticleMass; p->cellsize_inv = d.cellsize_inv; p->Dp_inv = d.Dp_inv;
    p->dt_vol_Dpinv = d.dt_vol_Dpinv; p->vmax = d.vmax; p->UpdateEveryNthStep = d.UpdateEveryNthStep;
}

#define FUNCS(T, SUF)                                                                              \
    void orc_weights_##SUF(T x, T w[3]) { weights<T>(x, w); }                                      \
    void orc_kirchhoff_##SUF(const IceParams *p, const T F[4], T tau[4]) {                         \
        Derived d = derive(*p); Mat2<T> M; mat_in(F, M);                                           \
        mat_out(kirchhoff<T>(M, T(d.mu), T(d.kappa)), tau);                                        \
    }                                                                                              \
    void orc_pq_##SUF(const IceParams *p, const T F[4], T out[3]) {                                \
        Derived d = derive(*p); Mat2<T> M; mat_in(F, M);                                           \
        compute_pq<T>(M, T(d.mu), T(d.kappa), out[0], out[1], out[2]);                             \
    }                                                                                              \
    void orc_svd_##SUF(const T F[4], T U[4], T S[2], T V[4]) {                                     \
        Mat2<T> M, Um, Vm; mat_in(F, M); svd2x2<T>(M, Um, S, Vm); mat_out(Um, U); mat_out(Vm, V);  \
    }                                                                                              \
    uint32_t orc_failure_check_##SUF(const IceParams *p, uint32_t util, T pp, T q) {               \
        Strength<T> g = strength<T>(*p); failure_check<T>(util, pp, q, g); return util;            \
    }                                                                                              \
    /* constitutive update of one point: returns case (-1 uncrushed), writes Fe, Jp, util, p, q */ \
    int orc_constitutive_##SUF(const IceParams *p, T dp_threshold, T thickness, T F[4], T *Jp,     \
                               uint32_t *util, T *p_tr, T *q_tr, uint32_t *err) {                  \
        Derived d = derive(*p); Strength<T> g = strength<T>(*p); Mat2<T> M; mat_in(F, M);          \
        int c = constitutive<T>(*p, d, g, dp_threshold, thickness, M, *Jp, *util, *p_tr, *q_tr, *err); \
        mat_out(M, F); return c;                                                                   \
    }

FUNCS(double, f64)
FUNCS(float, f32)

// ---- surrogate network (mlp_oracle.hpp) ----
// x raw [n][4]; norm = {xm[4], xs[4], ym[3], ys[3]}; y standardised outputs [n][3];
// mode 0 FP64, 1 BF16 operands (training), 2 BF16 operands past layer 1 (inference).
int orc_mlp_forward(const double *params, const double *norm, const double *x, int n, int mode, double *y) {
    orc_mlp::Norm N;
    std::memcpy(&N, norm, sizeof(N));
    orc_mlp::Trace T;
    for (int s = 0; s < n; s++) {
        double xs[orc_mlp::IN];
        for (int k = 0; k < orc_mlp::IN; k++) xs[k] = (x[s * orc_mlp::IN + k] - N.xm[k]) / N.xs[k];
        orc_mlp::forward(params, xs, mode, T);
        for (int k = 0; k < orc_mlp::OUT; k++) y[s * orc_mlp::OUT + k] = T.y[k];
    }
    return 0;
}

double orc_mlp_grad(const double *params, const double *norm, const double *x, const double *t, int n, int bf16,
                    double *grad) {
    orc_mlp::Norm N;
    std::memcpy(&N, norm, sizeof(N));
    std::memset(grad, 0, sizeof(double) * orc_mlp::NPARAM);
    return orc_mlp::loss_grad(params, N, x, t, n, bf16 != 0, grad);
}

// ---- frame arrays (frame_oracle.hpp) ----
// soa [>= 19][pitch] float lines; forces [2][gx*gy]; fields [10][gx*gy]; any output may be NULL.
int orc_frame_arrays(const float *soa, size_t pitch, int n, int gx, int gy, double particle_mass,
                     const uint8_t *status, const uint32_t *colors, int ncolors, const float *forces, double fscale,
                     uint8_t *count, float *fields, uint8_t *rgb, uint8_t *mask, float *fpr) {
    orc_frame::prepare(soa, pitch, n, gx, gy, particle_mass, status, colors, ncolors, forces, fscale, count, fields,
                       rgb, mask, fpr);
    return 0;
}

}  // extern "C"
