// path_host.cpp -- the right-hand sides of foort_path.hpp compiled for the
// host, exported for the CPU tests of the analytic formulations (the GPU
// kernels instantiate the same templates).
#include <cmath>
#include <cstring>

#include "foort_path.hpp"

namespace {
// metric constructors as in foort_gpu.hip (FOORT/Metric.cpp:111-509)
void params(const foort_run_desc *d, foort::RunConst &R)
{
    std::memset(&R, 0, sizeof(R));
    const double *P = d->metric_params;
    auto &M = R.M;
    M.rlog = d->rlog_scale;
    switch (d->metric) {
    case FOORT_METRIC_KERR: M.a = P[0]; break;
    case FOORT_METRIC_FLATSPACE: M.rlog = 0; break;
    case FOORT_METRIC_RASHEEDLARSEN:
        M.a = P[1] / ((P[2] + P[3]) / 4.0);
        M.m = P[0] / ((P[2] + P[3]) / 4.0);
        M.p = P[2] / ((P[2] + P[3]) / 4.0);
        M.q = P[3] / ((P[2] + P[3]) / 4.0);
        M.rl_root_pq = std::sqrt(M.p * M.q);
        M.rl_root_pm = std::sqrt((M.p * M.p - 4. * M.m * M.m) * (M.q * M.q - 4. * M.m * M.m));
        break;
    case FOORT_METRIC_JOHANNSEN: M.a = P[0]; M.a13 = P[1]; M.a22 = P[2]; M.a52 = P[3]; M.eps3 = P[4]; break;
    default:
        M.a = P[0];
        M.alpha3 = P[1];
        M.mn_alpha = (P[0] == 0.0) ? 0.0 : (-1. + std::sqrt(1. - P[0] * P[0])) / P[0];
        M.mn_k = std::sqrt(1. - P[0] * P[0]);
    }
    R.h_fd = d->derivative_h;
    R.inv2h = 1 / (2 * d->derivative_h);
}

template <int M>
void dispatch(int kind, const foort::RunConst &R, const double *p, const double *v, double *out)
{
    switch (kind) {
    case 0: foort::rhs_fd<M>(R, p, v, out); break;
    case 1: foort::rhs_an<M, double>(R, p, v, out); break;                 // dual numbers, binary64
    case 2: foort::rhs_analytic<M, double>(R, p, v, out); break;           // closed form where built, binary64
    case 3: foort::rhs_analytic<M, float>(R, p, v, out); break;            // same in binary32
    // the same two with RLogScale fixed at compile time, as the integration kernels run them
    case 4:
        if (R.M.rlog)
            foort::rhs_analytic<M, double, double, 1>(R, p, v, out);
        else
            foort::rhs_analytic<M, double, double, 0>(R, p, v, out);
        break;
    default:
        if (R.M.rlog)
            foort::rhs_analytic<M, float, double, 1>(R, p, v, out);
        else
            foort::rhs_analytic<M, float, double, 0>(R, p, v, out);
    }
}
} // namespace

extern "C" void path_rhs(const foort_run_desc *d, int kind, const double *p, const double *v, double *out)
{
    foort::RunConst R;
    params(d, R);
    switch (d->metric) {
    case FOORT_METRIC_KERR: dispatch<FOORT_METRIC_KERR>(kind, R, p, v, out); break;
    case FOORT_METRIC_FLATSPACE: dispatch<FOORT_METRIC_FLATSPACE>(kind, R, p, v, out); break;
    case FOORT_METRIC_RASHEEDLARSEN: dispatch<FOORT_METRIC_RASHEEDLARSEN>(kind, R, p, v, out); break;
    case FOORT_METRIC_JOHANNSEN: dispatch<FOORT_METRIC_JOHANNSEN>(kind, R, p, v, out); break;
    default: dispatch<FOORT_METRIC_MANKONOVIKOV>(kind, R, p, v, out); break;
    }
}

// sin and cos of th0 + d from those of th0 (the analytic formulations' RK4 stages), for its
// accuracy test: s, c of th0[i] + d[i]
extern "C" void path_sincos_shift(const double *th0, const double *d, double *s, double *c, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) {
        double s0, c0;
        foort::ssincos(th0[i], &s0, &c0);
        foort::stage_sincos(s0, c0, th0[i], th0[i] + d[i], &s[i], &c[i]);
    }
}

// the binary64 exponential of the analytic formulations, for its accuracy test
extern "C" void path_exp(const double *x, double *out, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        out[i] = foort::sexp(x[i]);
}
