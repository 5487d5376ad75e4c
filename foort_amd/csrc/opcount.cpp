// opcount.cpp -- counts the arithmetic of one integration step of each
// formulation by evaluating the path's own templates (foort_path.hpp) with a
// scalar type that counts its operations.  The counts define the algorithmic
// FLOPs per step used for the roofline figures (add, sub, mul, div, sqrt and
// comparisons-free: each counts 1 FLOP; an FMA formed by the compiler is not
// distinguished; sin, cos and exp are reported separately).
//
// Build and run: g++ -std=c++17 -O1 -I include -I foort_amd/csrc foort_amd/csrc/opcount.cpp -o opcount && ./opcount
// Output: JSON, one object per (metric, formulation).
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

static void reset() { cnt::add = cnt::mul = cnt::div = cnt::sqrt_ = cnt::trans = 0; }

template <int M>
static void analytic(const char *name, foort::RunConst &R, double u)
{
    double p[4] = {0, u, 1.4, 0}, v[4] = {-1, -0.9, 1e-4, 2e-5}, out[4];
    reset();
    foort::rhs_an<M, Cnt>(R, p, v, out);
    // one RK4 step = 4 right-hand sides + the step bookkeeping below
    const long long rhs = cnt::add + cnt::mul + cnt::div + cnt::sqrt_;
    // adaptive step (FOORT/Integrators.cpp:34-47): 2 add + 1 sub + 3 abs-free add, 3 div + 3 reciprocal
    // + 1 mul + 1 final reciprocal = 14; RK4 stage arguments 3 x 2 x 4 x 2 = 48; final update
    // 2 x 4 x 6 = 48; lambda += h: 1
    const long long book = 14 + 48 + 48 + 1;
    std::printf("  {\"metric\": \"%s\", \"formulation\": \"analytic\", \"rhs_flops\": %lld, \"rhs_div\": %lld, "
                "\"rhs_sqrt\": %lld, \"rhs_transcendental\": %lld, \"rk4_step_flops\": %lld}",
                name, rhs, cnt::div, cnt::sqrt_, cnt::trans, 4 * rhs + book);
}

int main()
{
    foort::RunConst R{};
    R.M.a = 0.99;
    R.M.rlog = 0;
    std::printf("[\n");
    analytic<FOORT_METRIC_KERR>("kerr", R, 50.0);
    std::printf("\n]\n");
    return 0;
}
