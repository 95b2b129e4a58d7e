"""Print the constants of kmath.h (200-bit mpmath): the three-part pi/2
split, pi and pi/2 as hi+lo, ln2 split 42+53 bits, the atan(j/8) table, the
series coefficients, and the tables of the double-double log and sin/cos.
Run: python kelp_amd/csrc/gen_kmath_tables.py"""
import mpmath as mp

mp.mp.prec = 200


def rnd(v, bits):
    m, e = mp.frexp(v)
    s = mp.mpf(2) ** bits
    return mp.ldexp(mp.nint(m * s) / s, e)


def h(v):
    return float(v).hex()


def main():
    pio2 = mp.pi / 2
    p1 = rnd(pio2, 33)
    p2 = rnd(pio2 - p1, 33)
    p3 = pio2 - p1 - p2
    print("KM_PIO2_1", h(p1), "KM_PIO2_2", h(p2), "KM_PIO2_3", h(p3))
    print("KM_2_OVER_PI", h(2 / mp.pi))
    for name, v in (("PI", mp.pi), ("PIO2", pio2)):
        hi = float(v)
        print(f"KM_{name}_HI", h(hi), f"KM_{name}_LO", h(v - hi))
    ln2 = mp.log(2)
    hi = rnd(ln2, 42)
    print("KM_LN2_HI", h(hi), "KM_LN2_LO", h(ln2 - hi), "KM_INV_LN2", h(1 / ln2))
    print("KM_SQRT2", h(mp.sqrt(2)))
    for j in range(2, 17):
        a = mp.atan(mp.mpf(j) / 8)
        print(f"atan({j}/8)", h(a), h(a - float(a)))
    print("sin", [h((-1) ** k / mp.factorial(2 * k + 1)) for k in range(1, 9)])
    print("cos", [h((-1) ** k / mp.factorial(2 * k)) for k in range(2, 10)])
    print("atan", [h(mp.mpf((-1) ** k) / (2 * k + 1)) for k in range(1, 14)])
    print("log", [h(mp.mpf(2) / (2 * k + 1)) for k in range(1, 12)])
    print("exp", [h(1 / mp.factorial(k)) for k in range(2, 16)])




def cr_tables():
    """The tables of the double-double log and sin/cos of kmath.h:
    c_i = RN(1 / (1 + i/128)) and -log(c_i) as hi + lo for i = -37..53;
    sin(j/64) and cos(j/64) as hi + lo for j = 0..51."""
    cs, th, tl = [], [], []
    for i in range(-37, 54):
        c = float(1 / (1 + mp.mpf(i) / 128))
        t = -mp.log(mp.mpf(c))
        hi = float(t)
        cs.append(h(c))
        th.append(h(hi))
        tl.append(h(t - hi))
    print("static const double KM_LOGT_C[91] = {" + ", ".join(cs) + "};")
    print("static const double KM_LOGT_H[91] = {" + ", ".join(th) + "};")
    print("static const double KM_LOGT_L[91] = {" + ", ".join(tl) + "};")
    sh, sl, ch, cl = [], [], [], []
    for j in range(52):
        a = mp.mpf(j) / 64
        s, c = mp.sin(a), mp.cos(a)
        sh.append(h(s)), sl.append(h(s - float(s))), ch.append(h(c)), cl.append(h(c - float(c)))
    for nm, arr in (("SH", sh), ("SL", sl), ("CH", ch), ("CL", cl)):
        print(f"static const double KM_SINT_{nm}[52] = {{" + ", ".join(arr) + "};")
    third = mp.mpf(1) / 3
    sixth = mp.mpf(1) / 6
    print("KM_THIRD", h(third), h(third - float(third)), "KM_SIXTH", h(sixth), h(sixth - float(sixth)))
    print("log1p tail", [h(mp.mpf((-1) ** (k + 1)) / k) for k in range(4, 13)])
    print("sin tail", [h(mp.mpf((-1) ** k) / mp.factorial(2 * k + 1)) for k in range(2, 6)])
    print("cos tail", [h(mp.mpf((-1) ** k) / mp.factorial(2 * k)) for k in range(2, 6)])


def cr_tables2():
    """The tables of the double-double atan and exp of kmath.h: atan(j/64)
    as hi + lo for j = 0..64, 2^(j/64) as hi + lo for j = 0..63, ln2/64 in
    three parts (36 + 36 bits + the rest) and 64/ln2."""
    ah, al = [], []
    for j in range(65):
        a = mp.atan(mp.mpf(j) / 64)
        ah.append(h(a)), al.append(h(a - float(a)))
    print("static const double KM_ATANT_H[65] = {" + ", ".join(ah) + "};")
    print("static const double KM_ATANT_L[65] = {" + ", ".join(al) + "};")
    eh, el = [], []
    for j in range(64):
        e = mp.mpf(2) ** (mp.mpf(j) / 64)
        eh.append(h(e)), el.append(h(e - float(e)))
    print("static const double KM_EXPT_H[64] = {" + ", ".join(eh) + "};")
    print("static const double KM_EXPT_L[64] = {" + ", ".join(el) + "};")
    l = mp.log(2) / 64
    l1 = rnd(l, 36)
    l2 = rnd(l - l1, 36)
    l3 = l - l1 - l2
    print("KM_LN2_64_1", h(l1), "KM_LN2_64_2", h(l2), "KM_LN2_64_3", h(l3), "KM_64_OVER_LN2", h(64 / mp.log(2)))
    print("atan dd tail", [h(mp.mpf((-1) ** k) / (2 * k + 1)) for k in range(2, 7)])
    print("exp tail", [h(1 / mp.factorial(k)) for k in range(4, 10)])


if __name__ == "__main__":
    main()
    cr_tables()
    cr_tables2()
