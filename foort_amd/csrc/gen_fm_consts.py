"""Generate fm_consts.inc: the constants of the portable binary64 elementary
functions in fm_math.h, computed from their mathematical definitions with
mpmath at 300 bits and rounded to nearest binary64.

  * pi/2 split in three binary64 parts (Cody-Waite argument reduction);
  * ln2/64 split in three parts and 64/ln2 (Tang's table-driven exp);
  * 2^(j/64), j = 0..63, each as an unevaluated sum hi + lo;
  * Taylor coefficients 1/n! (with sign) for sin, cos and exp.

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
    # Taylor coefficients
    for k in range(3, 18, 2):
        w("#define FM_S%d %s" % (k, d(mp.mpf((-1) ** ((k - 1) // 2)) / mp.factorial(k))))
    for k in range(4, 19, 2):
        w("#define FM_C%d %s" % (k, d(mp.mpf((-1) ** (k // 2)) / mp.factorial(k))))
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
