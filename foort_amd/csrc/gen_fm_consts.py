"""Generate fm_consts.inc: the constants of the portable binary64 elementary
functions in fm_math.h, computed from their mathematical definitions with
mpmath at 300 bits and rounded to nearest binary64.

  * pi/2 split in three binary64 parts (Cody-Waite argument reduction);
  * ln2/64 split in three parts and 64/ln2 (Tang's table-driven exp);
  * 2^(j/64), j = 0..63, each as an unevaluated sum hi + lo;
  * minimax polynomials for sin and cos on [-pi/4, pi/4] (Remez exchange on
    the relative error of the assembled sin x = x + x^3 S(x^2) and
    cos x = 1 - x^2/2 + x^4 C(x^2), six coefficients each);
  * Taylor coefficients 1/n! for exp.

Run: python gen_fm_consts.py > fm_consts.inc
"""
import mpmath as mp

mp.mp.prec = 300


def d(x):
    return float(x).hex()


def split3(c):
    p1 = mp.mpf(float(c))
    p2 = mp.mpf(float(c - p1))
    p3 = mp.mpf(float(c - p1 - p2))
    return p1, p2, p3


def remez(f, w, a, b, n, iters=12, grid=3000):
    """Coefficients c[0..n] of the degree-n polynomial p minimising max |w (p - f)| on [a, b]
    (Remez exchange: solve for equal alternating error on n + 2 points, move the points to
    the extrema of the error, repeat), and that maximum."""
    m = n + 2
    xs = [(a + b) / 2 - (b - a) / 2 * mp.cos(mp.pi * i / (m - 1)) for i in range(m)]
    grid = [a + (b - a) * (mp.mpf(i) / grid) ** 2 for i in range(grid + 1)]
    best = None
    for _ in range(iters):
        A = mp.matrix(m, m)
        rhs = mp.matrix(m, 1)
        for i, x in enumerate(xs):
            for j in range(n + 1):
                A[i, j] = x ** j
            A[i, n + 1] = (-1) ** i / w(x)
            rhs[i] = f(x)
        sol = mp.lu_solve(A, rhs)
        c = [sol[j] for j in range(n + 1)]
        ev = [w(x) * (mp.polyval(c[::-1], x) - f(x)) for x in grid]
        ext = []  # local extrema of |error| (the interval's ends always count)
        for i, v in enumerate(ev):
            if i in (0, len(ev) - 1) or (abs(v) >= abs(ev[i - 1]) and abs(v) >= abs(ev[i + 1])):
                if ext and mp.sign(ext[-1][1]) == mp.sign(v):
                    if abs(v) > abs(ext[-1][1]):
                        ext[-1] = (grid[i], v)
                else:
                    ext.append((grid[i], v))
        while len(ext) > m:
            ext.pop(0 if abs(ext[0][1]) < abs(ext[-1][1]) else -1)
        best = (c, max(abs(v) for v in ev))
        if len(ext) < m:
            break
        xs = [x for x, _ in ext]
    return best


def main():
    out = []
    w = out.append
    w("/* Generated by gen_fm_consts.py -- do not edit. */")
    p1, p2, p3 = split3(mp.pi / 2)
    w("#define FM_PIO2_1 %s" % d(p1))
    w("#define FM_PIO2_2 %s" % d(p2))
    w("#define FM_PIO2_3 %s" % d(p3))
    w("#define FM_TWO_OVER_PI %s" % d(2 / mp.pi))
    n = 64
    c1, c2, c3 = split3(mp.log(2) / n)
    w("#define FM_LN2_64_1 %s" % d(c1))
    w("#define FM_LN2_64_2 %s" % d(c2))
    w("#define FM_INV_LN2_64 %s" % d(n / mp.log(2)))
    # sin x = x + x^3 S(z), cos x = 1 - z/2 + z^2 C(z), z = x^2, S and C of degree 5, minimax
    # in the relative error of sin and cos for |x| <= pi/4 (a hair wider: the reduced argument
    # can exceed pi/4 by rounding); the lower end keeps the weight z of the relative error away
    # from zero, below it the polynomials' contribution is negligible
    prec = mp.mp.prec
    mp.mp.prec = 200
    Z = (mp.pi / 4 * (1 + mp.mpf(2) ** -40)) ** 2
    lo = Z * mp.mpf("1e-5")
    rt = mp.sqrt
    S, es = remez(lambda z: (mp.sin(rt(z)) - rt(z)) / rt(z) ** 3, lambda z: rt(z) ** 3 / mp.sin(rt(z)), lo, Z, 5)
    C, ec = remez(lambda z: (mp.cos(rt(z)) - 1 + z / 2) / z ** 2, lambda z: z ** 2 / mp.cos(rt(z)), lo, Z, 5)
    mp.mp.prec = prec
    w("/* minimax relative error: sin 2^%.1f, cos 2^%.1f */" % (float(mp.log(es, 2)), float(mp.log(ec, 2))))
    for k, c in enumerate(S):
        w("#define FM_S%d %s" % (2 * k + 3, d(c)))
    for k, c in enumerate(C):
        w("#define FM_C%d %s" % (2 * k + 4, d(c)))
    for k in range(2, 7):
        w("#define FM_E%d %s" % (k, d(mp.mpf(1) / mp.factorial(k))))
    w("#define FM_EXP2_TABLE_HI { \\")
    rows = []
    los = []
    for j in range(n):
        t = mp.power(2, mp.mpf(j) / n)
        hi = mp.mpf(float(t))
        lo = mp.mpf(float(t - hi))
        rows.append(d(hi))
        los.append(d(lo))
    for j in range(0, n, 4):
        w("  " + ", ".join(rows[j:j + 4]) + (", \\" if j + 4 < n else " }"))
    w("#define FM_EXP2_TABLE_LO { \\")
    for j in range(0, n, 4):
        w("  " + ", ".join(los[j:j + 4]) + (", \\" if j + 4 < n else " }"))
    print("\n".join(out))


if __name__ == "__main__":
    main()
