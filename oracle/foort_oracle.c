/*
 * foort_oracle.c -- TEST INFRASTRUCTURE ONLY.  A CPU restatement, in plain
 * C, of the reference's per-pixel null-geodesic integration (the OpenMP loop
 * of FOORT/Main.cpp:313-419 and everything it calls per pixel).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the results to compare with; the product never links or calls it.
 *
 * It follows the reference's arithmetic operation by operation: dense 4x4
 * tensors, the full 256-term Christoffel contraction (FOORT/Metric.cpp:14-55),
 * the tensor operators of FOORT/Geometry.h:141-184 (scalar products as
 * element*scalar, division as multiplication by the reciprocal) and the
 * truncated pi of FOORT/Geometry.h:44.
 *
 * Two builds of this file:
 *   liboracle_libm.so  - sin/cos/exp/pow from the C library, as the reference.
 *                        Pinned bit-for-bit against the reference's own sources
 *                        compiled into oracle/_ref (tests/test_oracle_ref.py) and
 *                        against the reference's output files Output/Test_*.dat.
 *   liboracle_fm.so    - the same code with the portable functions of
 *                        foort_amd/csrc/fm_math.h (ORACLE_FM defined); it is the
 *                        bit-exact reference of the GPU path, whose kernels use the
 *                        same functions.
 * The camera's log(r) (RLogScale, FOORT/ViewScreen.cpp:14) always uses the C
 * library (the portable set has no logarithm); it is evaluated once per run on
 * the host by both oracle and product.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "foort_gpu.h"

#ifdef ORACLE_FM
#include "fm_math.h"
#define O_SIN(x) fm_sin(x)
#define O_COS(x) fm_cos(x)
#define O_EXP(x) fm_exp(x)
#define O_POW2(x) fm_pow2(x)
#define O_POW3(x) fm_pow3(x)
#define O_POW4(x) fm_pow4(x)
#define O_POW8(x) fm_pow8(x)
#else
#define O_SIN(x) sin(x)
#define O_COS(x) cos(x)
#define O_EXP(x) exp(x)
#define O_POW2(x) pow((x), 2.)
#define O_POW3(x) pow((x), 3.)
#define O_POW4(x) pow((x), 4.)
#define O_POW8(x) pow((x), 8.)
#endif

#define DIM 4
static const double o_pi = 3.1415926535; /* FOORT/Geometry.h:44 */

typedef double T2[DIM][DIM];
typedef double T3[DIM][DIM][DIM];

/* sin and cos of the theta at which a metric is evaluated.  The C-library build reads them
 * in the metric itself, as sin(p[2]) and cos(p[2]) written where the reference writes them
 * (the compiler merges a sin and a cos of one argument into one sincos call, whose last bit can
 * differ from sin's alone: the same source shape as the reference gives the same calls); the
 * portable build takes them from the caller, which shifts the theta stencil's values. */
typedef struct {
    double s, c;
} oangle;
#ifdef ORACLE_FM
#define ANG_SIN(p, A) ((A)->s)
#define ANG_COS(p, A) ((A)->c)
#else
#define ANG_SIN(p, A) O_SIN((p)[2])
#define ANG_COS(p, A) O_COS((p)[2])
#endif

static oangle o_angle(double theta)
{
    oangle A;
    A.s = O_SIN(theta);
    A.c = O_COS(theta);
    return A;
}

/* sin and cos at theta + h and theta - h, the points of the finite differences in theta
 * (FOORT/Metric.cpp:24-30).  The C library evaluates each point by itself, as the reference;
 * the portable set shifts them from theta's reduced sine and cosine (fm_sincos3, the GPU
 * path's function). */
static void o_angle_stencil(double theta, double tp, double tm, oangle *Ap, oangle *Am)
{
#ifdef ORACLE_FM
    double s0, c0;
    fm_sincos3(theta, tp, tm, &s0, &c0, &Ap->s, &Ap->c, &Am->s, &Am->c);
#else
    *Ap = o_angle(tp);
    *Am = o_angle(tm);
#endif
}

/* ------------------------------------------------------------------------- */
/* Metrics (FOORT/Metric.cpp)                                                */
/* ------------------------------------------------------------------------- */

typedef struct {
    int id;
    int rlog;
    int horizon_metric;  /* SphericalHorizonMetric descendant */
    double horizon;      /* m_HorizonRadius */
    double a, m, p, q;   /* Kerr a; RL (rescaled) m a p q */
    double a13, a22, a52, eps3; /* Johannsen */
    double alpha3, mn_alpha, mn_k; /* Manko-Novikov */
} ometric;

static void o_metric_init(ometric *M, const foort_run_desc *d)
{
    const double *P = d->metric_params;
    memset(M, 0, sizeof(*M));
    M->id = d->metric;
    M->rlog = d->rlog_scale;
    M->horizon_metric = 1;
    switch (d->metric) {
    case FOORT_METRIC_KERR: /* FOORT/Metric.cpp:111-128 */
        M->a = P[0];
        M->horizon = 1 + sqrt(1 - P[0] * P[0]);
        break;
    case FOORT_METRIC_FLATSPACE: /* FOORT/Metric.cpp:205-215: not a SphericalHorizonMetric */
        M->horizon_metric = 0;
        M->rlog = 0;
        break;
    case FOORT_METRIC_RASHEEDLARSEN: { /* FOORT/Metric.cpp:244-271, P = {m, a, p, q} */
        double mP = P[0], aP = P[1], pP = P[2], qP = P[3];
        M->a = aP / ((pP + qP) / 4.0);
        M->m = mP / ((pP + qP) / 4.0);
        M->p = pP / ((pP + qP) / 4.0);
        M->q = qP / ((pP + qP) / 4.0);
        M->horizon = (mP + sqrt(mP * mP - aP * aP)) / ((pP + qP) / 4.0);
        break;
    }
    case FOORT_METRIC_JOHANNSEN: /* FOORT/Metric.cpp:378-407 */
        M->a = P[0];
        M->a13 = P[1];
        M->a22 = P[2];
        M->a52 = P[3];
        M->eps3 = P[4];
        M->horizon = 1 + sqrt(1 - P[0] * P[0]);
        break;
    case FOORT_METRIC_MANKONOVIKOV: /* FOORT/Metric.cpp:490-509 */
        M->a = P[0];
        M->alpha3 = P[1];
        M->mn_alpha = (P[0] == 0.0) ? 0.0 : (-1. + sqrt(1. - P[0] * P[0])) / P[0];
        M->mn_k = sqrt(1. - P[0] * P[0]);
        M->horizon = 1 + sqrt(1. - P[0] * P[0]);
        break;
    }
}

static void set5(T2 g, double g00, double g03, double g11, double g22, double g33)
{
    memset(g, 0, sizeof(T2));
    g[0][0] = g00;
    g[0][3] = g03;
    g[1][1] = g11;
    g[2][2] = g22;
    g[3][0] = g03;
    g[3][3] = g33;
}

/* FOORT/Metric.cpp:131-159 */
static void kerr_dd(const ometric *M, const double *p, const oangle *A, T2 g)
{
    double a = M->a;
    double r = M->rlog ? O_EXP(p[1]) : p[1];
    double sint = ANG_SIN(p, A);
    double cost = ANG_COS(p, A);
    double sigma = r * r + a * a * cost * cost;
    double delta = r * r + a * a - 2. * r;
    double A_ = (r * r + a * a) * (r * r + a * a) - delta * a * a * sint * sint;
    double g00 = -(1. - 2. * r / sigma);
    double g11 = sigma / delta;
    double g22 = sigma;
    double g33 = A_ / sigma * sint * sint;
    double g03 = -2. * a * r * sint * sint / sigma;
    if (M->rlog)
        g11 *= (r * r);
    set5(g, g00, g03, g11, g22, g33);
}

/* FOORT/Metric.cpp:162-191 */
static void kerr_uu(const ometric *M, const double *p, const oangle *A, T2 g)
{
    double a = M->a;
    double r = M->rlog ? O_EXP(p[1]) : p[1];
    double sint = ANG_SIN(p, A);
    double cost = ANG_COS(p, A);
    double sigma = r * r + a * a * cost * cost;
    double delta = r * r + a * a - 2. * r;
    double A_ = (r * r + a * a) * (r * r + a * a) - delta * a * a * sint * sint;
    double g00 = -A_ / (sigma * delta);
    double g11 = delta / sigma;
    double g22 = 1. / sigma;
    double g33 = (delta - a * a * sint * sint) / (sigma * delta * sint * sint);
    double g03 = -2. * a * r / (sigma * delta);
    if (M->rlog)
        g11 *= 1.0 / (r * r);
    set5(g, g00, g03, g11, g22, g33);
}

/* FOORT/Metric.cpp:218-229 */
static void flat_dd(const ometric *M, const double *p, const oangle *A, T2 g)
{
    (void)M;
    set5(g, -1, 0, 1, p[1] * p[1], p[1] * p[1] * ANG_SIN(p, A) * ANG_SIN(p, A));
}

static void flat_uu(const ometric *M, const double *p, const oangle *A, T2 g)
{
    (void)M;
    set5(g, -1, 0, 1, 1 / (p[1] * p[1]), 1 / (p[1] * p[1] * ANG_SIN(p, A) * ANG_SIN(p, A)));
}

/* shared part of FOORT/Metric.cpp:276-300 and :321-345 */
#define RL_COMMON                                                                                   \
    double m = M->m, a = M->a, pp = M->p, qq = M->q;                                              \
    double r = M->rlog ? O_EXP(p[1]) : p[1];                                                      \
    double r2 = r * r;                                                                             \
    double sint = ANG_SIN(p, A);                                                                   \
    double cost = ANG_COS(p, A);                                                                   \
    double sint2 = sint * sint;                                                                    \
    double cost2 = cost * cost;                                                                    \
    double delta = r2 + a * a - 2. * r * m;                                                        \
    double H3 = r2 - 2. * r * m + a * a * cost2;                                                   \
    double Bp = sqrt(pp * qq) * a * sint2 *                                                        \
                (((pp * qq + 4. * m * m) * r - m * (pp - 2. * m) * (qq - 2. * m)) /                  \
                 (2. * m * (pp + qq) * H3));                                                       \
    double H1 = r2 + a * a * cost2 + r * (pp - 2. * m) +                                           \
                (pp / (pp + qq)) * ((pp - 2. * m) * (qq - 2. * m) / 2.) -                          \
                (pp / (2. * m * (pp + qq))) *                                                      \
                    sqrt((pp * pp - 4. * m * m) * (qq * qq - 4. * m * m)) * a * cost;               \
    double H2 = r2 + a * a * cost2 + r * (qq - 2. * m) +                                           \
                (qq / (pp + qq)) * ((pp - 2. * m) * (qq - 2. * m) / 2.) +                          \
                (qq / (2. * m * (pp + qq))) *                                                      \
                    sqrt((pp * pp - 4. * m * m) * (qq * qq - 4. * m * m)) * a * cost;               \
    double sqH1H2 = sqrt(H1 * H2);

static void rl_dd(const ometric *M, const double *p, const oangle *A, T2 g)
{
    RL_COMMON
    double g00 = -H3 / sqH1H2;
    double g11 = sqH1H2 / delta;
    double g22 = sqH1H2;
    double g33 = -(H3 * Bp * Bp) / sqH1H2 + (sqH1H2 * delta * sint2) / H3;
    double g03 = -(H3 * Bp) / sqH1H2;
    if (M->rlog)
        g11 *= (r * r);
    set5(g, g00, g03, g11, g22, g33);
}

static void rl_uu(const ometric *M, const double *p, const oangle *A, T2 g)
{
    RL_COMMON
    double g00 = ((H3 * H3 * Bp * Bp) / sint2 - H1 * H2 * delta) / (sqH1H2 * H3 * delta);
    double g11 = delta / sqH1H2;
    double g22 = 1. / sqH1H2;
    double g33 = H3 / (sqH1H2 * delta * sint2);
    double g03 = -(H3 * Bp) / (sqH1H2 * delta * sint2);
    if (M->rlog)
        g11 *= 1.0 / (r * r);
    set5(g, g00, g03, g11, g22, g33);
}

/* shared part of FOORT/Metric.cpp:412-424 and :445-457 */
#define JOH_COMMON                                                                                 \
    double a = M->a;                                                                              \
    double r = M->rlog ? O_EXP(p[1]) : p[1];                                                      \
    double sint = ANG_SIN(p, A);                                                                   \
    double cost = ANG_COS(p, A);                                                                   \
    double A1 = 1. + M->a13 / (r * r * r);                                                         \
    double A2 = 1. + M->a22 / (r * r);                                                             \
    double A5 = 1. + M->a52 / (r * r);                                                             \
    double eps_f = M->eps3 / r;                                                                    \
    double rho2 = r * r + a * a * cost * cost + eps_f;                                             \
    double delta = r * r + a * a - 2. * r;

static void joh_dd(const ometric *M, const double *p, const oangle *A, T2 g)
{
    JOH_COMMON
    double g00 = -rho2 * (delta - a * a * A2 * A2 * sint * sint) /
                 O_POW2((r * r + a * a) * A1 - a * a * A2 * sint * sint);
    double g11 = rho2 / (delta * A5);
    double g22 = rho2;
    double g33 = rho2 * sint * sint * (O_POW2((r * r + a * a) * A1) - a * a * delta * sint * sint) /
                 O_POW2((r * r + a * a) * A1 - a * a * A2 * sint * sint);
    double g03 = -a * ((r * r + a * a) * A1 * A2 - delta) * rho2 * sint * sint /
                 O_POW2((r * r + a * a) * A1 - a * a * A2 * sint * sint);
    if (M->rlog)
        g11 *= (r * r);
    set5(g, g00, g03, g11, g22, g33);
}

static void joh_uu(const ometric *M, const double *p, const oangle *A, T2 g)
{
    JOH_COMMON
    double g00 = (-1. * O_POW2((r * r + a * a) * A1) + a * a * delta * sint * sint) / (delta * rho2);
    double g11 = delta * A5 / rho2;
    double g22 = 1. / rho2;
    double g33 = (-a * a * A2 * A2 * sint * sint + delta) / (delta * rho2 * sint * sint);
    double g03 = -a * (A2 * A1 * (r * r + a * a) - delta) / (delta * rho2);
    if (M->rlog)
        g11 *= 1.0 / (r * r);
    set5(g, g00, g03, g11, g22, g33);
}

/* shared part of FOORT/Metric.cpp:516-548 and :572-604 */
#define MN_COMMON                                                                                  \
    double al = M->mn_alpha, kk = M->mn_k, a3 = M->alpha3;                                        \
    double r = M->rlog ? O_EXP(p[1]) : p[1];                                                      \
    double sint = ANG_SIN(p, A);                                                                   \
    double cost = ANG_COS(p, A);                                                                   \
    double xx = (r - 1.) / kk;                                                                     \
    double yy = cost;                                                                              \
    double R = sqrt(xx * xx + yy * yy - 1.);                                                       \
    double P1 = xx * yy / R;                                                                       \
    double P2 = 0.5 * (3. * xx * xx * yy * yy / (R * R) - 1.);                                     \
    double P3 = 0.5 * (5. * O_POW3(xx * yy / R) - 3. * xx * yy / R);                               \
    double P4 = 0.125 * (35. * O_POW4(xx * yy / R) - 30. * xx * xx * yy * yy / (R * R) + 3.);      \
    double aa = -al * O_EXP(2. * a3 * (1. - (xx - yy) * (1. / R + P1 / (R * R) + P2 / (R * R * R) + \
                                                        P3 / (R * R * R * R))));                   \
    double bb = al * O_EXP(2. * a3 * (-1. + (xx + yy) * (1. / R - P1 / (R * R) + P2 / (R * R * R) - \
                                                        P3 / (R * R * R * R))));                   \
    double AA = (xx * xx - 1.) * (1. + aa * bb) * (1. + aa * bb) - (1. - yy * yy) * (bb - aa) * (bb - aa); \
    double BB = O_POW2((xx + 1. + (xx - 1.) * aa * bb)) + O_POW2(((1. + yy) * aa + (1. - yy) * bb)); \
    double CC = (xx * xx - 1.) * (1. + aa * bb) * (bb - aa - yy * (aa + bb)) +                     \
                (1. - yy * yy) * (bb - aa) * (1. + aa * bb + xx * (1. - aa * bb));                  \
    double psi = a3 * P3 / (R * R * R * R);                                                        \
    double gamma_prime = 2. * a3 * a3 * (P4 * P4 - P3 * P3) / (O_POW8(R)) +                        \
                         2. * a3 * (-yy + xx * P1 / R - yy * P2 / (R * R) + xx * P3 / (R * R * R)) / R; \
    double f = O_EXP(2. * psi) * AA / BB;                                                          \
    double exp_2gamma = O_EXP(2. * gamma_prime) * AA / ((xx * xx - yy * yy) * (1. - al * al) * (1. - al * al)); \
    double omega = 2. * kk * O_EXP(-2. * psi) * CC / AA - 4. * kk * al / (1. - al * al);           \
    double rho_sq = (r - 1.) * (r - 1.) - kk * kk * cost * cost;                                   \
    double delta = (r - 1.) * (r - 1.) - kk * kk;

static void mn_dd(const ometric *M, const double *p, const oangle *A, T2 g)
{
    MN_COMMON
    double g00 = -f;
    double g22 = exp_2gamma * rho_sq / f;
    double g11 = g22 / delta;
    double g33 = -f * omega * omega + delta * sint * sint / f;
    double g03 = omega * f;
    if (M->rlog)
        g11 *= (r * r);
    set5(g, g00, g03, g11, g22, g33);
}

static void mn_uu(const ometric *M, const double *p, const oangle *A, T2 g)
{
    MN_COMMON
    double beta_d = omega * f;
    double gamma_dd = -f * omega * omega + delta * sint * sint / f;
    double beta_u = beta_d / gamma_dd;
    double alpha_sq = beta_u * beta_d + f;
    double g00 = -1. / alpha_sq;
    double g22 = f / (exp_2gamma * rho_sq);
    double g11 = delta * g22;
    double g33 = 1. / gamma_dd - beta_u * beta_u / alpha_sq;
    double g03 = beta_u / alpha_sq;
    if (M->rlog)
        g11 *= 1.0 / (r * r);
    set5(g, g00, g03, g11, g22, g33);
}

static void o_gdd(const ometric *M, const double *p, const oangle *A, T2 g)
{
    switch (M->id) {
    case FOORT_METRIC_KERR: kerr_dd(M, p, A, g); break;
    case FOORT_METRIC_FLATSPACE: flat_dd(M, p, A, g); break;
    case FOORT_METRIC_RASHEEDLARSEN: rl_dd(M, p, A, g); break;
    case FOORT_METRIC_JOHANNSEN: joh_dd(M, p, A, g); break;
    default: mn_dd(M, p, A, g); break;
    }
}

static void o_guu(const ometric *M, const double *p, const oangle *A, T2 g)
{
    switch (M->id) {
    case FOORT_METRIC_KERR: kerr_uu(M, p, A, g); break;
    case FOORT_METRIC_FLATSPACE: flat_uu(M, p, A, g); break;
    case FOORT_METRIC_RASHEEDLARSEN: rl_uu(M, p, A, g); break;
    case FOORT_METRIC_JOHANNSEN: joh_uu(M, p, A, g); break;
    default: mn_uu(M, p, A, g); break;
    }
}

/* Metric::getChristoffel_udd, FOORT/Metric.cpp:14-55.  All five metrics have
 * the symmetries {0, 3} (FOORT/Metric.cpp:127, :214, :270, :406, :508). */
static void o_christoffel(const ometric *M, double h, const double *p, T3 G)
{
    T3 dd;
    memset(dd, 0, sizeof(T3));
    const oangle A0 = o_angle(p[2]);
    for (int coord = 0; coord < DIM; ++coord) {
        if (coord == 0 || coord == 3)
            continue;
        double shift[DIM] = {0, 0, 0, 0};
        shift[coord] = h;
        double pp[DIM], pm[DIM];
        for (int i = 0; i < DIM; ++i) {
            pp[i] = p[i] + shift[i];
            pm[i] = p[i] - shift[i];
        }
        oangle Ap = A0, Am = A0;
        if (coord == 2)
            o_angle_stencil(p[2], pp[2], pm[2], &Ap, &Am);
        T2 gp, gm;
        o_gdd(M, pp, &Ap, gp);
        o_gdd(M, pm, &Am, gm);
        double inv = 1 / (2 * h);
        for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j)
                dd[coord][i][j] = (gp[i][j] - gm[i][j]) * inv;
    }
    T2 uu;
    o_guu(M, p, &A0, uu);
    memset(G, 0, sizeof(T3));
    for (int mu = 0; mu < DIM; ++mu)
        for (int nu = 0; nu < DIM; ++nu)
            for (int rho = 0; rho < DIM; ++rho)
                for (int sigma = 0; sigma < DIM; ++sigma)
                    G[mu][nu][rho] += 1.0 / 2 * uu[mu][sigma] *
                                      (dd[nu][rho][sigma] + dd[rho][nu][sigma] - dd[sigma][nu][rho]);
}

/* ------------------------------------------------------------------------- */
/* Integrators (FOORT/Integrators.cpp)                                       */
/* ------------------------------------------------------------------------- */

typedef struct {
    ometric M;
    int integrator;
    double eps, minstep, h_fd, verlet_tol;
    /* terminations */
    uint32_t terms;
    double h_eps, sphere_R, theta_eps;
    uint64_t h_n, s_n, t_n, th_n, max_steps;
    /* diagnostics */
    uint32_t diags;
    double ep_thr;
    foort_update_frequency ep_f, cr_f;
    uint64_t step_cap;
    /* camera */
    double cam[4], size[2], center[2];
    T2 cam_gdd;
    T2 vielbein;
    int cam_rlog;
} orun;

/* std::min / std::max semantics */
static double o_min(double a, double b) { return (b < a) ? b : a; }
static double o_max(double a, double b) { return (a < b) ? b : a; }

/* Integrators::GetAdaptiveStep, FOORT/Integrators.cpp:34-47 */
static double o_adaptive_step(const orun *R, const double *pos, const double *vel)
{
    double d1 = R->eps / (fabs(vel[1]) + 1e-20);
    double d2 = R->eps * o_min(pos[2], o_pi - pos[2]) / (fabs(vel[2]) + 1e-20);
    double d3 = R->eps / (fabs(vel[3] + 1e-20));
    double h = 1 / (1 / fabs(d1) + 1 / fabs(d2) + 1 / fabs(d3));
    h = o_max(h, R->minstep);
    return h;
}

/* geoRHS lambda, FOORT/Integrators.cpp:60-69 (Source = NoSource, zero) */
static void o_rhs(const orun *R, const double *p, const double *v, double *ret)
{
    T3 G;
    o_christoffel(&R->M, R->h_fd, p, G);
    for (int i = 0; i < DIM; ++i)
        ret[i] = 0;
    for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j)
            for (int k = 0; k < DIM; ++k)
                ret[i] -= G[i][j][k] * v[j] * v[k];
}

/* a + s*b elementwise, written as the reference's tensor expression a + (b * s) */
static void axpy(double *out, const double *a, double s, const double *b)
{
    for (int i = 0; i < DIM; ++i)
        out[i] = a[i] + b[i] * s;
}

/* IntegrateGeodesicStep_RK4, FOORT/Integrators.cpp:51-94 */
static void o_rk4(const orun *R, const double *x, const double *v, double *nx, double *nv, double *step)
{
    double h = o_adaptive_step(R, x, v);
    double k1[DIM], k2[DIM], k3[DIM], k4[DIM];
    double l1[DIM], l2[DIM], l3[DIM], l4[DIM];
    double tp[DIM], tv[DIM];
    o_rhs(R, x, v, k1);
    memcpy(l1, v, sizeof(l1));
    axpy(tp, x, 0.5 * h, l1);
    axpy(tv, v, 0.5 * h, k1);
    o_rhs(R, tp, tv, k2);
    axpy(l2, v, 0.5 * h, k1);
    axpy(tp, x, 0.5 * h, l2);
    axpy(tv, v, 0.5 * h, k2);
    o_rhs(R, tp, tv, k3);
    axpy(l3, v, 0.5 * h, k2);
    axpy(tp, x, h, l3);
    axpy(tv, v, h, k3);
    o_rhs(R, tp, tv, k4);
    axpy(l4, v, h, k3);
    for (int i = 0; i < DIM; ++i) {
        nv[i] = v[i] + (((k1[i] + k2[i] * 2) + k3[i] * 2) + k4[i]) * (h / 6.0);
        nx[i] = x[i] + (((l1[i] + l2[i] * 2) + l3[i] * 2) + l4[i]) * (h / 6.0);
    }
    *step = h;
}

static double cartvecsq(const double *v)
{
    double ret = 0;
    for (int i = 0; i < DIM; ++i)
        ret += v[i] * v[i];
    return ret;
}

/* Verlet iterations are capped (the reference loops without bound); the cap
 * is identical in the GPU kernels */
#define VERLET_MAX_ITER 1000

/* IntegrateGeodesicStep_Verlet, FOORT/Integrators.cpp:99-152 */
static void o_verlet(const orun *R, const double *x, const double *v, double *nx, double *nv, double *step)
{
    double h = o_adaptive_step(R, x, v);
    double ac[DIM], as[DIM], vi[DIM], d[DIM];
    o_rhs(R, x, v, ac);
    for (int i = 0; i < DIM; ++i)
        nx[i] = (x[i] + v[i] * h) + ac[i] * (h * h / 2.0);
    axpy(vi, v, h, ac);
    o_rhs(R, nx, vi, as);
    for (int i = 0; i < DIM; ++i)
        nv[i] = v[i] + (ac[i] + as[i]) * (h / 2.0);
    int it = 0;
    while (R->verlet_tol > 0.0 && it < VERLET_MAX_ITER) {
        for (int i = 0; i < DIM; ++i)
            d[i] = nv[i] - vi[i];
        if (!(cartvecsq(d) / cartvecsq(nv) > R->verlet_tol * R->verlet_tol))
            break;
        memcpy(vi, nv, sizeof(vi));
        o_rhs(R, nx, vi, as);
        for (int i = 0; i < DIM; ++i)
            nv[i] = v[i] + (ac[i] + as[i]) * (h / 2.0);
        ++it;
    }
    *step = h;
}

/* ------------------------------------------------------------------------- */
/* ViewScreen (FOORT/ViewScreen.cpp)                                         */
/* ------------------------------------------------------------------------- */

static double innerprod(T2 gdd, const double *a, const double *b)
{
    double ret = 0;
    for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j)
            ret += a[i] * gdd[i][j] * b[j];
    return ret;
}

/* ViewScreen::ConstructVielbein, FOORT/ViewScreen.cpp:10-73 (including the
 * theta row that omits the e_r projection and uses the zero-initialised row
 * e_theta in the e_phi projection, :56-58) */
static void o_vielbein(orun *R)
{
    double pos[DIM];
    memcpy(pos, R->cam, sizeof(pos));
    pos[1] = R->cam_rlog ? log(pos[1]) : pos[1];
    T2 gdd, guu;
    const oangle A = o_angle(pos[2]);
    o_gdd(&R->M, pos, &A, gdd);
    o_guu(&R->M, pos, &A, guu);
    memcpy(R->cam_gdd, gdd, sizeof(T2));
    double signs[DIM] = {0, 0, 0, 0};
    T2 V;
    memset(V, 0, sizeof(V));
    double n, t;
    /* t */
    memcpy(V[0], guu[0], sizeof(V[0]));
    n = innerprod(gdd, V[0], V[0]);
    signs[0] = (n < 0) ? -1 : 1;
    t = 1 / sqrt(signs[0] * n);
    for (int i = 0; i < DIM; ++i)
        V[0][i] = (V[0][i] * signs[0]) * t;
    /* phi */
    {
        double c0 = signs[0] * innerprod(gdd, guu[3], V[0]);
        for (int i = 0; i < DIM; ++i)
            V[3][i] = guu[3][i] - V[0][i] * c0;
        n = innerprod(gdd, V[3], V[3]);
        signs[3] = (n < 0) ? -1 : 1;
        double s = signs[3] * R->cam[1] * O_SIN(R->cam[2]);
        t = 1 / sqrt(signs[3] * n);
        for (int i = 0; i < DIM; ++i)
            V[3][i] = (V[3][i] * s) * t;
    }
    /* r */
    {
        double c0 = signs[0] * innerprod(gdd, guu[1], V[0]);
        double c3 = signs[3] * innerprod(gdd, guu[1], V[3]);
        for (int i = 0; i < DIM; ++i)
            V[1][i] = (guu[1][i] - V[0][i] * c0) - V[3][i] * c3;
        n = innerprod(gdd, V[1], V[1]);
        signs[1] = (n < 0) ? -1 : 1;
        t = 1 / sqrt(signs[1] * n);
        for (int i = 0; i < DIM; ++i)
            V[1][i] = (V[1][i] * signs[1]) * t;
    }
    /* theta */
    {
        double c0 = signs[0] * innerprod(gdd, guu[2], V[0]);
        double c3 = signs[3] * innerprod(gdd, guu[2], V[3]);
        double row[DIM];
        for (int i = 0; i < DIM; ++i)
            row[i] = (guu[2][i] - V[0][i] * c0) - V[2][i] * c3; /* V[2] is still zero here */
        memcpy(V[2], row, sizeof(row));
        n = innerprod(gdd, V[2], V[2]);
        signs[2] = (n < 0) ? -1 : 1;
        double s = signs[2] * R->cam[1];
        t = 1 / sqrt(signs[2] * n);
        for (int i = 0; i < DIM; ++i)
            V[2][i] = (V[2][i] * s) * t;
    }
    memcpy(R->vielbein, V, sizeof(T2));
}

/* ViewScreen::SetNewInitialConditions, FOORT/ViewScreen.cpp:75-119, given the
 * mesh's unit screen point */
static void o_initial(const orun *R, double ux, double uy, double *pos, double *vel)
{
    memcpy(pos, R->cam, sizeof(double) * DIM);
    pos[1] = R->cam_rlog ? log(pos[1]) : pos[1];
    double alpha = R->center[0] + R->size[0] * (ux - 0.5);
    double beta = R->center[1] + R->size[1] * (uy - 0.5);
    double sintheta0 = O_SIN(pos[2]);
    double rc = R->cam[1];
    double densqrt = sqrt(rc * rc + alpha * alpha + beta * beta);
    double pf[DIM] = {-1.0, -rc / densqrt, -beta / rc / densqrt, alpha / rc / sintheta0 / densqrt};
    for (int i = 0; i < DIM; ++i) {
        vel[i] = 0;
        for (int j = 0; j < DIM; ++j)
            vel[i] += R->vielbein[j][i] * pf[j];
    }
    double cur = 0;
    for (int i = 0; i < DIM; ++i)
        cur += R->cam_gdd[0][i] * vel[i];
    double inv = 1 / cur;
    for (int i = 0; i < DIM; ++i)
        vel[i] = (vel[i] * 1.0) * inv;
}

/* ------------------------------------------------------------------------- */
/* Geodesic with terminations and diagnostics (FOORT/Geodesic.cpp,           */
/* FOORT/Terminations.cpp, FOORT/Diagnostics.cpp)                            */
/* ------------------------------------------------------------------------- */

typedef struct {
    double pos[DIM], vel[DIM], lambda;
    int term;
    uint64_t steps;
    uint64_t hc, bc, tc, thc, tsteps; /* termination counters */
    int quadrant;
    int passes;
    double prevtheta;
    uint64_t epc;
    double closest;
    uint64_t crc;
} ogeo;

/* Termination::DecideUpdate, FOORT/Terminations.cpp:71-90 */
static int term_decide(uint64_t n, uint64_t *cnt)
{
    if (n == 0)
        return 1;
    ++*cnt;
    if (*cnt >= n) {
        *cnt = 0;
        return 1;
    }
    return 0;
}

/* Diagnostic::DecideUpdate, FOORT/Diagnostics.cpp:113-140 */
static int diag_decide(const foort_update_frequency *f, uint64_t *cnt, const ogeo *g)
{
    if (f->update_n_steps > 0) {
        ++*cnt;
        if (*cnt >= f->update_n_steps) {
            *cnt = 0;
            return 1;
        }
        return 0;
    } else if (f->update_start && g->lambda == 0.0) {
        return 1;
    } else if (f->update_finish && g->term != FOORT_TERM_CONTINUE) {
        return 1;
    }
    return 0;
}

static double o_true_r(const orun *R, double rc)
{
    return R->M.rlog ? O_EXP(rc) : rc;
}

/* UpdateData of every diagnostic that is on (order does not matter: they are independent) */
static void o_diag_update(const orun *R, ogeo *g)
{
    if (R->diags & FOORT_DIAGFLAG_FOURCOLORSCREEN) { /* FOORT/Diagnostics.cpp:159-200 */
        if (g->term == FOORT_TERM_BOUNDARYSPHERE) {
            double p3 = g->pos[3];
            if (fabs(p3) < 2 * o_pi * 1e5) {
                while (p3 > 2 * o_pi)
                    p3 -= 2 * o_pi;
                while (p3 < 0)
                    p3 += 2 * o_pi;
            }
            int quadrant;
            if (g->pos[2] < o_pi / 2)
                quadrant = (p3 < o_pi) ? 1 : 2;
            else
                quadrant = (p3 < o_pi) ? 3 : 4;
            g->quadrant = quadrant;
        }
    }
    if (R->diags & FOORT_DIAGFLAG_EQUATORIALPASSES) { /* FOORT/Diagnostics.cpp:351-380 */
        if (diag_decide(&R->ep_f, &g->epc, g)) {
            double cur = g->pos[2];
            if (fabs(cur - o_pi / 2.0) > o_pi / 2.0 * R->ep_thr) {
                if (g->prevtheta > 0 && ((g->prevtheta - o_pi / 2.0) * (cur - o_pi / 2.0)) < 0.0)
                    ++g->passes;
                g->prevtheta = cur;
            }
        }
        if (g->term == FOORT_TERM_HORIZON)
            g->passes = -g->passes;
    }
    if (R->diags & FOORT_DIAGFLAG_CLOSESTRADIUS) { /* FOORT/Diagnostics.cpp:427-446 */
        if (diag_decide(&R->cr_f, &g->crc, g)) {
            double cur = o_true_r(R, g->pos[1]);
            if (g->closest < 0 || cur < g->closest)
                g->closest = cur;
        }
        if (g->term == FOORT_TERM_HORIZON)
            g->closest = 0.0;
    }
}

/* Geodesic::Reset, FOORT/Geodesic.cpp:33-58 */
static void o_reset(const orun *R, ogeo *g, const double *pos, const double *vel)
{
    memset(g, 0, sizeof(*g));
    memcpy(g->pos, pos, sizeof(g->pos));
    memcpy(g->vel, vel, sizeof(g->vel));
    g->lambda = 0.0;
    g->term = FOORT_TERM_CONTINUE;
    g->quadrant = 0;
    g->passes = 0;
    g->prevtheta = -1;
    g->closest = -1;
    o_diag_update(R, g);
}

/* HorizonTermination etc., FOORT/Terminations.cpp:96-214, checked in the
 * vector order of CreateTerminationVector (:22-40) */
static int o_check_terms(const orun *R, ogeo *g)
{
    if (R->terms & FOORT_TERMFLAG_HORIZON) {
        if (term_decide(R->h_n, &g->hc)) {
            double r = o_true_r(R, g->pos[1]);
            if ((r < R->M.horizon * (1 + R->h_eps)) || (R->M.horizon == 0.0 && r < R->h_eps))
                return FOORT_TERM_HORIZON;
        }
    }
    if (R->terms & FOORT_TERMFLAG_BOUNDARYSPHERE) {
        if (term_decide(R->s_n, &g->bc)) {
            double r = o_true_r(R, g->pos[1]);
            if (r > R->sphere_R)
                return FOORT_TERM_BOUNDARYSPHERE;
        }
    }
    if (R->terms & FOORT_TERMFLAG_TIMEOUT) {
        if (term_decide(R->t_n, &g->tc)) {
            if (g->tsteps >= R->max_steps)
                return FOORT_TERM_TIMEOUT;
            ++g->tsteps;
        }
    }
    if (R->terms & FOORT_TERMFLAG_THETASINGULARITY) {
        if (term_decide(R->th_n, &g->thc)) {
            double th = g->pos[2];
            if (fabs(th) < R->theta_eps || fabs(o_pi - th) < R->theta_eps)
                return FOORT_TERM_THETASINGULARITY;
        }
    }
    return FOORT_TERM_CONTINUE;
}

/* Geodesic::Update, FOORT/Geodesic.cpp:61-88 */
static void o_update(const orun *R, ogeo *g)
{
    double np[DIM], nv[DIM], h;
    if (R->integrator == FOORT_INTEGRATOR_VERLET)
        o_verlet(R, g->pos, g->vel, np, nv, &h);
    else
        o_rk4(R, g->pos, g->vel, np, nv, &h);
    g->lambda += h;
    memcpy(g->pos, np, sizeof(np));
    memcpy(g->vel, nv, sizeof(nv));
    ++g->steps;
    g->term = o_check_terms(R, g);
    if (g->term == FOORT_TERM_CONTINUE && g->steps >= R->step_cap)
        g->term = FOORT_TERM_STEPCAP;
    o_diag_update(R, g);
}

static int o_setup(orun *R, const foort_run_desc *d)
{
    memset(R, 0, sizeof(*R));
    o_metric_init(&R->M, d);
    R->integrator = d->integrator;
    R->eps = d->step_size;
    R->minstep = d->smallest_step;
    R->h_fd = d->derivative_h;
    R->verlet_tol = d->verlet_tolerance;
    R->terms = d->term_flags;
    if (!R->M.horizon_metric)
        R->terms &= ~FOORT_TERMFLAG_HORIZON;
    R->h_eps = d->horizon_epsilon;
    R->h_n = d->horizon_update_n;
    R->sphere_R = d->sphere_radius;
    R->s_n = d->sphere_update_n;
    R->max_steps = d->timeout_max_steps;
    R->t_n = d->timeout_update_n;
    R->theta_eps = d->theta_epsilon;
    R->th_n = d->theta_update_n;
    R->diags = d->diag_flags;
    R->ep_thr = d->ep_threshold;
    R->ep_f = d->ep_update;
    R->cr_f = d->cr_update;
    R->step_cap = d->step_cap ? d->step_cap : 2147483647ull;
    memcpy(R->cam, d->camera_position, sizeof(R->cam));
    memcpy(R->size, d->screen_size, sizeof(R->size));
    memcpy(R->center, d->screen_center, sizeof(R->center));
    R->cam_rlog = R->M.rlog;
    o_vielbein(R);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Exported entry points (ctypes)                                            */
/* ------------------------------------------------------------------------- */

/* Integrate `count` pixels: screen_index (row, col) pairs, or, if NULL, grid
 * indices first..first+count-1 of a size x size SimpleSquareMesh
 * (FOORT/Mesh.cpp:31-50).  Results as in foort_geodesic_results. */
int oracle_trace(const foort_run_desc *d, uint64_t size, uint64_t first, uint64_t count,
                 const uint64_t *screen_index, int32_t *term, int32_t *quadrant, int32_t *passes,
                 double *closest, double *fpos, double *fvel, double *lambda, uint64_t *steps,
                 int nthreads)
{
    orun R;
    o_setup(&R, d);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t ii = 0; ii < (int64_t)count; ++ii) {
        uint64_t row, col;
        if (screen_index) {
            row = screen_index[2 * ii];
            col = screen_index[2 * ii + 1];
        } else {
            uint64_t index = first + (uint64_t)ii;
            row = index / size;
            col = index - row * size;
        }
        double ux = row * 1.0 / (double)(size - 1);
        double uy = col * 1.0 / (double)(size - 1);
        double pos[DIM], vel[DIM];
        o_initial(&R, ux, uy, pos, vel);
        ogeo g;
        o_reset(&R, &g, pos, vel);
        while (g.term == FOORT_TERM_CONTINUE)
            o_update(&R, &g);
        term[ii] = g.term;
        quadrant[ii] = g.quadrant;
        passes[ii] = g.passes;
        closest[ii] = g.closest;
        lambda[ii] = g.lambda;
        steps[ii] = g.steps;
        memcpy(&fpos[4 * ii], g.pos, sizeof(g.pos));
        memcpy(&fvel[4 * ii], g.vel, sizeof(g.vel));
    }
    return 0;
}

/* GeodesicPositionDiagnostic, FOORT/Diagnostics.cpp:239-284: every saved position of one
 * geodesic in a growing array; at termination the array is cut down to about OutputSteps
 * entries: keep index i when i % (size / OutputSteps) == 0, then put the last saved point in
 * place of the last kept one unless it is kept already. */
typedef struct {
    double *p;
    uint64_t n, cap;
} optlist;

static void optlist_push(optlist *l, const double *pos)
{
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 64;
        l->p = (double *)realloc(l->p, l->cap * 4 * sizeof(double));
    }
    memcpy(l->p + 4 * l->n, pos, 4 * sizeof(double));
    ++l->n;
}

static void o_gp_update(const foort_update_frequency *f, uint64_t keep, const ogeo *g, uint64_t *cnt, optlist *l)
{
    if (diag_decide(f, cnt, g))
        optlist_push(l, g->pos);
    if (g->term != FOORT_TERM_CONTINUE && keep > 0 && keep < l->n) {
        const uint64_t jettison = l->n / keep;
        optlist tmp = {0, 0, 0};
        for (uint64_t i = 0; i < l->n; ++i)
            if (i % jettison == 0)
                optlist_push(&tmp, l->p + 4 * i);
        if ((l->n - 1) % jettison != 0) {
            --tmp.n;
            optlist_push(&tmp, l->p + 4 * (l->n - 1));
        }
        free(l->p);
        *l = tmp;
    }
}

/* oracle_trace with GeodesicPosition: kept[i] = the points pixel i keeps; the points of all
 * pixels, in pixel order, go to `points` (4 doubles each) when they fit in `capacity` points.
 * Returns the total number of kept points. */
uint64_t oracle_trace_positions(const foort_run_desc *d, uint64_t size, uint64_t first, uint64_t count,
                                const uint64_t *screen_index, int32_t *term, int32_t *quadrant, int32_t *passes,
                                double *closest, double *fpos, double *fvel, double *lambda, uint64_t *steps,
                                int nthreads, uint64_t *kept, double *points, uint64_t capacity)
{
    orun R;
    o_setup(&R, d);
    optlist *lists = (optlist *)calloc(count ? count : 1, sizeof(optlist));
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t ii = 0; ii < (int64_t)count; ++ii) {
        uint64_t row, col;
        if (screen_index) {
            row = screen_index[2 * ii];
            col = screen_index[2 * ii + 1];
        } else {
            uint64_t index = first + (uint64_t)ii;
            row = index / size;
            col = index - row * size;
        }
        double ux = row * 1.0 / (double)(size - 1);
        double uy = col * 1.0 / (double)(size - 1);
        double pos[DIM], vel[DIM];
        o_initial(&R, ux, uy, pos, vel);
        ogeo g;
        uint64_t gpc = 0;
        o_reset(&R, &g, pos, vel);
        o_gp_update(&d->gp_update, d->gp_output_steps, &g, &gpc, &lists[ii]);
        while (g.term == FOORT_TERM_CONTINUE) {
            o_update(&R, &g);
            o_gp_update(&d->gp_update, d->gp_output_steps, &g, &gpc, &lists[ii]);
        }
        term[ii] = g.term;
        quadrant[ii] = g.quadrant;
        passes[ii] = g.passes;
        closest[ii] = g.closest;
        lambda[ii] = g.lambda;
        steps[ii] = g.steps;
        memcpy(&fpos[4 * ii], g.pos, sizeof(g.pos));
        memcpy(&fvel[4 * ii], g.vel, sizeof(g.vel));
    }
    uint64_t total = 0;
    for (uint64_t i = 0; i < count; ++i) {
        kept[i] = lists[i].n;
        total += lists[i].n;
    }
    if (points && total <= capacity) {
        uint64_t at = 0;
        for (uint64_t i = 0; i < count; ++i) {
            if (lists[i].n)
                memcpy(points + 4 * at, lists[i].p, lists[i].n * 4 * sizeof(double));
            at += lists[i].n;
        }
    }
    for (uint64_t i = 0; i < count; ++i)
        free(lists[i].p);
    free(lists);
    return total;
}

/* Unit-level entry points for tests */
void oracle_metric(const foort_run_desc *d, const double *p, double *gdd16, double *guu16)
{
    ometric M;
    o_metric_init(&M, d);
    const oangle A = o_angle(p[2]);
    o_gdd(&M, p, &A, (double(*)[DIM])gdd16);
    o_guu(&M, p, &A, (double(*)[DIM])guu16);
}

void oracle_christoffel(const foort_run_desc *d, const double *p, double *G64)
{
    ometric M;
    o_metric_init(&M, d);
    o_christoffel(&M, d->derivative_h, p, (double(*)[DIM][DIM])G64);
}

double oracle_horizon_radius(const foort_run_desc *d)
{
    ometric M;
    o_metric_init(&M, d);
    return M.horizon;
}

void oracle_step(const foort_run_desc *d, const double *x, const double *v, double *nx, double *nv, double *h)
{
    orun R;
    o_setup(&R, d);
    if (R.integrator == FOORT_INTEGRATOR_VERLET)
        o_verlet(&R, x, v, nx, nv, h);
    else
        o_rk4(&R, x, v, nx, nv, h);
}

void oracle_camera(const foort_run_desc *d, double *vielbein16, double *gdd_row0)
{
    orun R;
    o_setup(&R, d);
    memcpy(vielbein16, R.vielbein, sizeof(T2));
    memcpy(gdd_row0, R.cam_gdd[0], sizeof(double) * DIM);
}

void oracle_initial(const foort_run_desc *d, double ux, double uy, double *pos, double *vel)
{
    orun R;
    o_setup(&R, d);
    o_initial(&R, ux, uy, pos, vel);
}

#ifdef ORACLE_FM
/* The portable elementary functions, exported for their accuracy tests */
void oracle_fm_eval(const double *x, uint64_t n, int which, double *out1, double *out2)
{
    for (uint64_t i = 0; i < n; ++i) {
        switch (which) {
        case 0: fm_sincos(x[i], &out1[i], &out2[i]); break;
        case 1: out1[i] = fm_exp(x[i]); break;
        case 2: out1[i] = fm_pow3(x[i]); break;
        case 3: out1[i] = fm_pow4(x[i]); break;
        default: out1[i] = fm_pow8(x[i]); break;
        }
    }
}

/* sin and cos at x[i], xp[i], xm[i] by fm_sincos3 (out: 6 per i), for its tests */
void oracle_fm_stencil(const double *x, const double *xp, const double *xm, uint64_t n, double *out)
{
    for (uint64_t i = 0; i < n; ++i)
        fm_sincos3(x[i], xp[i], xm[i], &out[6 * i], &out[6 * i + 1], &out[6 * i + 2], &out[6 * i + 3],
                   &out[6 * i + 4], &out[6 * i + 5]);
}

/* sin and cos at y[i] shifted from x[i] (fm_sincos_near), for the accuracy tests */
void oracle_fm_near(const double *x, const double *y, uint64_t n, double *s, double *c)
{
    for (uint64_t i = 0; i < n; ++i) {
        double sh, sl, ch, cl;
        int q = fm_sincos_red(x[i], &sh, &sl, &ch, &cl);
        fm_sincos_near(q, sh, sl, ch, cl, x[i], y[i], &s[i], &c[i]);
    }
}
#endif
