"""Generate contraction.inc: the Christoffel symbols and the geodesic
right-hand side for a stationary, axisymmetric metric, unrolled and with the
structurally zero terms removed, in the reference's summation order.

The reference computes (FOORT/Metric.cpp:39-52, FOORT/Integrators.cpp:60-69)

    G[mu][nu][rho] = sum_{sigma=0..3} (0.5 * guu[mu][sigma])
                       * ((d[nu][rho][sigma] + d[rho][nu][sigma]) - d[sigma][nu][rho])
    ret[i]         = 0 - sum_{j,k} (G[i][j][k] * v[j]) * v[k]     (j outer, k inner)

with d[c][a][b] = d g_ab / d x^c, summed from +0 left to right.  For the five
metrics every g_dd and g^uu has the same non-zero pattern {00, 03 = 30, 11,
22, 33} and derivatives only along r and theta (c = 1, 2).  A term that is
structurally zero adds +-0 to the running sum and leaves every non-zero
value unchanged, so dropping it keeps the result bit-identical (only the sign
of an exact zero can differ).  Inside a bracket, zero operands are dropped
the same way: (A + 0) - C == A - C, (0 + 0) - C == -C.

A Christoffel symbol's leading +0 (0 + t1 + t2 ...) and the 0 of 0 - C are
dropped as well: they change only the sign of a zero symbol, and a symbol
enters the result only as a product subtracted from a running sum that
starts at +0, which no subtraction of a zero (of either sign) changes and
which never becomes -0 (x - x is +0 in round-to-nearest).  The running sums
keep their leading +0, so every output, zeros included, is the reference's.

Inputs of the generated function: d1[5], d2[5] (d/dr, d/dtheta of g00 g03
g11 g22 g33), u[5] (g^uu in the same order), v[4].  Output: out[4].
Run: python gen_contraction.py > contraction.inc
"""

NZ = {(0, 0): 0, (0, 3): 1, (3, 0): 1, (1, 1): 2, (2, 2): 3, (3, 3): 4}


def dname(c, a, b):
    if c not in (1, 2) or (a, b) not in NZ:
        return None
    return "d%d[%d]" % (c, NZ[(a, b)])


def uname(a, b):
    return "hu%d" % NZ[(a, b)] if (a, b) in NZ else None


def bracket(nu, rho, sigma):
    A = dname(nu, rho, sigma)
    B = dname(rho, nu, sigma)
    Cc = dname(sigma, nu, rho)
    if A is None and B is None and Cc is None:
        return None
    if A and B:
        s = "(%s + %s)" % (A, B)
    else:
        s = A or B
    if Cc:
        s = "(%s - %s)" % (s, Cc) if s else "(-%s)" % Cc
    return s


def main():
    lines = ["/* Generated by gen_contraction.py -- do not edit. */"]
    w = lines.append
    w("template <class T>")
    w("FM_HD_INLINE void geo_contract(const T *d1, const T *d2, const T *u, const T *v, T *out)")
    w("{")
    for k in range(5):
        w("    const T hu%d = T(0.5) * u[%d];" % (k, k))
    gam = {}
    for mu in range(4):
        for nu in range(4):
            for rho in range(nu, 4):
                terms = []
                for sigma in range(4):
                    un = uname(mu, sigma)
                    br = bracket(nu, rho, sigma)
                    if un and br:
                        terms.append("%s * %s" % (un, br))
                if terms:
                    name = "G%d%d%d" % (mu, nu, rho)
                    expr = terms[0]
                    for t in terms[1:]:
                        expr = "(%s) + %s" % (expr, t)
                    w("    const T %s = %s;" % (name, expr))
                    gam[(mu, nu, rho)] = name
                    gam[(mu, rho, nu)] = name
    for i in range(4):
        acc = "T(0)"
        for j in range(4):
            for k in range(4):
                g = gam.get((i, j, k))
                if g:
                    acc = "(%s - (%s * v[%d]) * v[%d])" % (acc, g, j, k)
        w("    out[%d] = %s;" % (i, acc))
    w("}")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
