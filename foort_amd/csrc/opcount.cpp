// opcount.cpp -- counts the arithmetic of one integration step of each
// formulation by evaluating the path's own templates (foort_path.hpp) with a
// scalar type that counts its operations.  The counts define the algorithmic
// FLOPs per step used for the roofline figures (add, sub, mul, div, sqrt and
// comparisons-free: each counts 1 FLOP; an FMA formed by the compiler is not
// distinguished; sin, cos and exp are reported separately).
//
// Build and run: g++ -std=c++17 -O1 -I include -I foort_amd/csrc foort_amd/csrc/opcount.cpp -o opcount && ./opcount
// Output: JSON, one object per (metric, formulation); the committed copy is
// foort_amd/flops_per_step.json (written by foort_amd.build).
#include <cstdio>

#include "foort_path.hpp"

namespace cnt {
long long add, mul, div, sqrt_, trans;
}

struct Cnt {
    double v;
    Cnt() = default;
    Cnt(double x) : v(x) {}
    Cnt &operator*=(const Cnt &o) { return *this = *this * o; }
    friend Cnt operator+(Cnt a, Cnt b) { ++cnt::add; return Cnt(a.v + b.v); }
    friend Cnt operator-(Cnt a, Cnt b) { ++cnt::add; return Cnt(a.v - b.v); }
    friend Cnt operator-(Cnt a) { return Cnt(-a.v); }
    friend Cnt operator*(Cnt a, Cnt b) { ++cnt::mul; return Cnt(a.v * b.v); }
    friend Cnt operator/(Cnt a, Cnt b) { ++cnt::div; return Cnt(a.v / b.v); }
    friend bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
    explicit operator double() const { return v; }
};

// found by argument-dependent lookup from the path templates
inline Cnt sexp(Cnt x) { ++cnt::trans; return Cnt(fm_exp(x.v)); }
inline Cnt ssqrt(Cnt x) { ++cnt::sqrt_; return Cnt(__builtin_sqrt(x.v)); }
inline void ssincos(double x, Cnt *s, Cnt *c) { cnt::trans += 2; s->v = fm_sin(x); c->v = fm_cos(x); }
// the reference-exact functions of the finite-difference path (fm_math.h):
// exp is transcendental; the integer powers cost what their double-double
// evaluation costs (an fma counted as 2)
inline Cnt srcp(Cnt x) { ++cnt::div; return Cnt(1.0 / x.v); }
inline Cnt arcp(Cnt x) { ++cnt::div; return Cnt(1.0 / x.v); }
inline Cnt fexp(Cnt x) { ++cnt::trans; return Cnt(fm_exp(x.v)); }
inline Cnt fsqrt(Cnt x) { ++cnt::sqrt_; return Cnt(__builtin_sqrt(x.v)); }
inline Cnt fpow2(Cnt x) { ++cnt::mul; return Cnt(x.v * x.v); }
inline Cnt fpow3(Cnt x) { cnt::mul += 5; cnt::add += 4; return Cnt(fm_pow3(x.v)); }
inline Cnt fpow4(Cnt x) { cnt::mul += 5; cnt::add += 5; return Cnt(fm_pow4(x.v)); }
inline Cnt fpow8(Cnt x) { cnt::mul += 9; cnt::add += 10; return Cnt(fm_pow8(x.v)); }

static void reset() { cnt::add = cnt::mul = cnt::div = cnt::sqrt_ = cnt::trans = 0; }

// Step bookkeeping of RK4 outside the right-hand sides: stage arguments
// (3 stages x 4 components x 2 vectors x 2 ops = 48), the final update
// (2 vectors x 4 components x 6 ops = 48) and lambda += h (1); plus the
// adaptive step: the reference form (FOORT/Integrators.cpp:34-47: 3 divisions,
// 3 reciprocals, 3 adds inside the fabs terms, 1 product, 1 sub, 2 adds and a
// final reciprocal = 14) for the finite-difference path, the equivalent
// one-division form (1 sub, 3 adds inside the fabs terms and the sum, 2 products, 1 add,
// 1 division = 8) for
// the analytic ones.
constexpr long long kRK4Book = 48 + 48 + 1;
constexpr long long kStepRef = 14, kStepFast = 8;

static long long flops() { return cnt::add + cnt::mul + cnt::div + cnt::sqrt_; }

template <int M>
static void count_fd(const char *name, foort::RunConst &R, double u)
{
    reset();
    // rhs_fd with Cnt scalars: 4 shifted g_dd, g^uu, differences, contraction
    Cnt uu = Cnt(u), h = Cnt(R.h_fd), th = Cnt(1.4);
    Cnt up = uu + h, um = uu - h, tp = th + h, tm = th - h;
    Cnt s0, c0, sp, cp, sm, cm;
    ssincos(th.v, &s0, &c0);
    ssincos(tp.v, &sp, &cp);
    ssincos(tm.v, &sm, &cm);
    Cnt gp[5], gm[5], d1[5], d2[5], g[5], v[4] = {Cnt(-1), Cnt(-0.9), Cnt(1e-4), Cnt(2e-5)}, out[4];
    foort::metric_dd<M, Cnt>(R.M, up, s0, c0, gp);
    foort::metric_dd<M, Cnt>(R.M, um, s0, c0, gm);
    for (int k = 0; k < 5; ++k)
        d1[k] = (gp[k] - gm[k]) * Cnt(R.inv2h);
    foort::metric_dd<M, Cnt>(R.M, uu, sp, cp, gp);
    foort::metric_dd<M, Cnt>(R.M, uu, sm, cm, gm);
    for (int k = 0; k < 5; ++k)
        d2[k] = (gp[k] - gm[k]) * Cnt(R.inv2h);
    foort::metric_uu<M, Cnt>(R.M, uu, s0, c0, g);
    foort::geo_contract<Cnt>(d1, d2, g, v, out);
    const long long rhs = flops();
    std::printf("    {\"metric\": \"%s\", \"formulation\": \"fd\", \"rhs_flops\": %lld, \"rhs_div\": %lld, "
                "\"rhs_sqrt\": %lld, \"rhs_transcendental\": %lld, \"rk4_step_flops\": %lld, "
                "\"rk4_step_transcendental\": %lld}",
                name, rhs, cnt::div, cnt::sqrt_, cnt::trans, 4 * rhs + kRK4Book + kStepRef, 4 * cnt::trans);
}

template <int M>
static void count_an(const char *name, foort::RunConst &R, double u)
{
    double p[4] = {0, u, 1.4, 0}, v[4] = {-1, -0.9, 1e-4, 2e-5}, out[4];
    reset();
    foort::rhs_analytic<M, Cnt>(R, p, v, out);
    const long long rhs = flops();
    std::printf("    {\"metric\": \"%s\", \"formulation\": \"analytic\", \"rhs_flops\": %lld, \"rhs_div\": %lld, "
                "\"rhs_sqrt\": %lld, \"rhs_transcendental\": %lld, \"rk4_step_flops\": %lld, "
                "\"rk4_step_transcendental\": %lld}",
                name, rhs, cnt::div, cnt::sqrt_, cnt::trans, 4 * rhs + kRK4Book + kStepFast, 4 * cnt::trans);
}

template <int M>
static void both(const char *name, double a, bool rlog, bool last)
{
    foort::RunConst R{};
    R.M.a = a;
    R.M.m = 1.0; R.M.p = 2.5; R.M.q = 3.0; R.M.a13 = 2.0; R.M.alpha3 = 1.0;
    R.M.mn_alpha = -0.27; R.M.mn_k = 0.87;
    R.M.rlog = rlog;
    R.h_fd = 1e-7;
    R.inv2h = 1 / (2 * R.h_fd);
    const double u = rlog ? 3.9 : 50.0;
    count_fd<M>(name, R, u);
    std::printf(",\n");
    count_an<M>(name, R, u);
    std::printf(last ? "\n" : ",\n");
}

int main()
{
    std::printf("{\"unit\": \"one integration step of one geodesic (RK4, radial coordinate r)\",\n");
    std::printf(" \"counting\": \"add, sub, mul, div, sqrt = 1 FLOP each; sin, cos, exp counted as transcendental\",\n");
    std::printf(" \"counts\": [\n");
    both<FOORT_METRIC_KERR>("kerr", 0.99, false, false);
    both<FOORT_METRIC_FLATSPACE>("flatspace", 0.0, false, false);
    both<FOORT_METRIC_RASHEEDLARSEN>("rasheed_larsen", 0.6, false, false);
    both<FOORT_METRIC_JOHANNSEN>("johannsen", 0.7, false, false);
    both<FOORT_METRIC_MANKONOVIKOV>("manko_novikov", 0.5, false, true);
    std::printf(" ]}\n");
    return 0;
}
