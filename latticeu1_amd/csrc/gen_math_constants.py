#!/usr/bin/env python3
"""Generate u1_math_constants.h: the constants of the shared elementary functions.

Every constant is derived here from first principles with Python's ``decimal``
module at 100 significant digits and rounded once to binary64 / binary32:

* pi from Machin's formula  pi = 16 atan(1/5) - 4 atan(1/239);
* ln 2 from the series      ln 2 = 2 atanh(1/3);
* Taylor coefficients 1/n! and 1/(2k+1) exactly as rationals (binary32
  sin/cos/atan of the measurement kernels);
* minimax tables from Remez exchanges, kept as literals: binary32 atan on
  [0, 1] for the FP32 update kernels (remez_atan01), and the binary64 sin,
  cos, exp and atan polynomials shared by the oracle and the FP64 kernels
  (remez_mp, in 60-digit mpmath arithmetic).

Cody-Waite splits (pi/2 = P1 + P2 + P3, ln2 = HI + LO) take the remainder of
the exact value after subtracting the exactly-represented leading parts, so the
sum P1 + P2 + P3 carries about 150 bits of pi/2.

Run:  python3 gen_math_constants.py > u1_math_constants.h
The test suite re-derives every value with these same functions and compares
against the committed header, so the header cannot drift from this recipe
(the minimax table: its error is re-measured and the exchange re-run).
"""
from decimal import Decimal, getcontext
from fractions import Fraction
import struct
import sys

getcontext().prec = 100


def _atan_inv(n: int) -> Decimal:
    """atan(1/n) by its Taylor series (n >= 2)."""
    x = Decimal(1) / n
    x2 = x * x
    term, k, s = x, 0, Decimal(0)
    eps = Decimal(10) ** -(getcontext().prec + 2)
    while abs(term) > eps:
        s += term / (2 * k + 1) * (-1 if k % 2 else 1)
        term *= x2
        k += 1
    return s


def pi_dec() -> Decimal:
    return 16 * _atan_inv(5) - 4 * _atan_inv(239)


def ln2_dec() -> Decimal:
    # ln 2 = 2 * atanh(1/3) = 2 * sum_k (1/3)^(2k+1)/(2k+1)
    x = Decimal(1) / 3
    x2 = x * x
    term, k, s = x, 0, Decimal(0)
    eps = Decimal(10) ** -(getcontext().prec + 2)
    while term > eps:
        s += term / (2 * k + 1)
        term *= x2
        k += 1
    return 2 * s


def to_f64(d) -> float:
    """Correctly rounded conversion of a Decimal/Fraction to binary64."""
    if isinstance(d, Fraction):
        return d.numerator / d.denominator  # Python int/int division is correctly rounded
    return float(d)  # float(Decimal) goes through the correctly rounded string path


def to_f32(d) -> float:
    """Correctly rounded conversion to binary32 (via exact rational rounding)."""
    fr = Fraction(d) if not isinstance(d, Fraction) else d
    if fr == 0:
        return 0.0
    sign = -1 if fr < 0 else 1
    fr = abs(fr)
    e = fr.numerator.bit_length() - fr.denominator.bit_length()
    # normalise so that 2^23 <= m < 2^24
    while fr >= Fraction(2) ** (e + 1):
        e += 1
    while fr < Fraction(2) ** e:
        e -= 1
    scale = Fraction(2) ** (23 - e)
    m = fr * scale
    q, r = divmod(m.numerator, m.denominator)
    rem = Fraction(r, m.denominator)
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and q % 2 == 1):
        q += 1
    val = Fraction(q) / scale
    out = float(val) * sign
    # sanity: representable in binary32
    assert struct.unpack("f", struct.pack("f", out))[0] == out
    return out


def hexf(x: float) -> str:
    return float.hex(x)


def split3(exact: Decimal, conv):
    p1 = conv(exact)
    r1 = exact - Decimal(p1)
    p2 = conv(r1)
    r2 = r1 - Decimal(p2)
    p3 = conv(r2)
    return p1, p2, p3


def split2(exact: Decimal, conv):
    hi = conv(exact)
    lo = conv(exact - Decimal(hi))
    return hi, lo


def constants():
    pi = pi_dec()
    ln2 = ln2_dec()
    c = {}
    c["PI"] = to_f64(pi)
    c["TWO_PI"] = to_f64(2 * pi)
    c["PI_LO"] = to_f64(pi - Decimal(c["PI"]))
    c["TWO_PI_LO"] = to_f64(2 * pi - Decimal(c["TWO_PI"]))
    c["PIO2_HI"], c["PIO2_LO"] = split2(pi / 2, to_f64)
    c["PIO4"] = to_f64(pi / 4)
    c["PIO4_LO"] = to_f64(pi / 4 - Decimal(c["PIO4"]))
    c["TWO_OVER_PI"] = to_f64(2 / pi)
    c["PIO2_1"], c["PIO2_2"], c["PIO2_3"] = split3(pi / 2, to_f64)
    c["LN2_HI"], c["LN2_LO"] = split2(ln2, to_f64)
    c["LOG2E"] = to_f64(1 / ln2)
    # tan(pi/8) = sqrt(2) - 1
    c["TAN_PIO8"] = to_f64(Decimal(2).sqrt() - 1)
    # float versions
    f = {}
    f["PI"] = to_f32(pi)
    f["TWO_PI"] = to_f32(2 * pi)
    f["PIO2_HI"], f["PIO2_LO"] = split2(pi / 2, to_f32)
    f["PIO4"] = to_f32(pi / 4)
    f["TWO_OVER_PI"] = to_f32(2 / pi)
    f["PIO2_1"], f["PIO2_2"], f["PIO2_3"] = split3(pi / 2, to_f32)
    f["TAN_PIO8"] = to_f32(Decimal(2).sqrt() - 1)
    return c, f


def inv_fact(n: int) -> Fraction:
    fct = 1
    for i in range(2, n + 1):
        fct *= i
    return Fraction(1, fct)


def coeff_tables():
    """Coefficient tables, highest order first (Horner order)."""
    t = {}
    # binary64 minimax (remez_mp): sin r = r + r^3 P(r^2) through r^13,
    # cos r = 1 + r^2 Q(r^2) through r^14, e^r = 1 + r + r^2 E(r) through r^12,
    # atan t = t + t^3 A(t^2) through t^23
    for name in MM64:
        t[name] = [float.fromhex(h) for h in MM64_HEX[name]]
    # float tables
    t["SIN32"] = [to_f32(inv_fact(2 * k + 3) * (-1) ** (k + 1)) for k in range(3, -1, -1)]  # through r^9
    t["COS32"] = [to_f32(inv_fact(2 * k + 2) * (-1) ** (k + 1)) for k in range(4, -1, -1)]  # through r^10
    t["ATAN32"] = [to_f32(Fraction((-1) ** (k + 1), 2 * k + 3)) for k in range(8, -1, -1)]  # through t^19
    # minimax atan z = z + z^3 A(z^2) on all of [0, 1] (binary32 update kernels)
    t["ATAN01_MM32"] = [float.fromhex(h) for h in ATAN01_MM32_HEX]
    return t


# Minimax coefficients of atan z = z + z^3 A(z^2) on 0 <= z <= 1, relative
# error weight, A of degree 7 in w = z^2, highest order first, rounded to
# binary32.  They are the output of remez_atan01(8) below, rounded once; the
# literal values are kept so the header does not depend on the LAPACK build.
# Max relative error of the polynomial: 1.67e-8 before rounding, 1.81e-8 after
# (the test suite re-measures it).
ATAN01_MM32_HEX = ["0x1.7ed232p-9", "-0x1.0c2c14p-6", "0x1.61fddap-5", "-0x1.3556b6p-4",
                   "0x1.b4e12ap-4", "-0x1.230adcp-3", "0x1.9978f4p-3", "-0x1.5554dcp-2"]


# Binary64 minimax tables (highest order first): the output of remez_mp(name,
# n) rounded once to binary64, kept as literals.  Weighted error (relative
# error of the function; sin, cos on |r| <= pi/4, e^r on |r| <= ln2/2, atan on
# |t| <= tan(pi/8)) after rounding, in units of 2^-53: sin 0.10, cos 0.039,
# exp 0.0064, atan 0.039 (the test suite re-measures them and re-runs the
# exchange).
MM64 = {"SIN64": 6, "COS64": 7, "EXP64": 11, "ATAN64": 11}
MM64_HEX = {
    "SIN64": ["0x1.5d8ff90144cdcp-33", "-0x1.ae5e5b1533e8ep-26", "0x1.71de356876082p-19", "-0x1.a01a019bfea3dp-13",
              "0x1.111111110f7d6p-7", "-0x1.5555555555548p-3"],
    "COS64": ["-0x1.8f7571988e6d6p-37", "0x1.1ee96d22197a9p-29", "-0x1.27e4f72c0286ep-22", "0x1.a01a019b2e602p-16",
              "-0x1.6c16c16c13bb1p-10", "0x1.5555555555539p-5", "-0x1.0000000000000p-1"],
    "EXP64": ["0x1.1e3d557e25f52p-29", "0x1.af6bad920703bp-26", "0x1.27e5b3f37a4fbp-22", "0x1.71ddf0419bf66p-19",
              "0x1.a01a0136d60b9p-16", "0x1.a01a01b3715f4p-13", "0x1.6c16c16c30321p-10", "0x1.111111110ec7ep-7",
              "0x1.5555555555503p-5", "0x1.555555555555bp-3", "0x1.0000000000000p-1"],
    "ATAN64": ["-0x1.2565a607b5743p-6", "0x1.37d2c7d80af0ap-5", "-0x1.9cd14a1039218p-5", "0x1.df0ee3b60d900p-5",
               "-0x1.10ebaf736a466p-4", "0x1.3b11182728061p-4", "-0x1.745cf8cb68c93p-4", "0x1.c71c70e552a3ap-4",
               "-0x1.24924922aaf63p-3", "0x1.99999999957e7p-3", "-0x1.555555555553dp-2"],
}


def _mm64_problem(name):
    """(g, W, a, b) of a binary64 minimax table: its polynomial p minimises
    max |W(x) (g(x) - p(x))| over [a, b]; W (g - p) is the relative error of
    the function the table builds."""
    import mpmath as mp

    def nz(x):  # the quotients below at x = 0 are limits; evaluate just beside it
        return x if x != 0 else mp.mpf(10) ** -40

    if name == "SIN64":  # sin r = r + r^3 P(r^2), w = r^2 <= (pi/4)^2
        return (lambda w: (mp.sin(mp.sqrt(nz(w))) - mp.sqrt(nz(w))) / nz(w) ** mp.mpf(1.5),
                lambda w: nz(w) ** mp.mpf(1.5) / mp.sin(mp.sqrt(nz(w))), mp.mpf(0), (mp.pi / 4) ** 2)
    if name == "COS64":  # cos r = 1 + r^2 Q(r^2)
        return (lambda w: (mp.cos(mp.sqrt(nz(w))) - 1) / nz(w), lambda w: nz(w) / mp.cos(mp.sqrt(nz(w))),
                mp.mpf(0), (mp.pi / 4) ** 2)
    if name == "EXP64":  # e^r = 1 + r + r^2 E(r), |r| <= ln2 / 2
        return (lambda r: (mp.exp(nz(r)) - 1 - nz(r)) / nz(r) ** 2, lambda r: nz(r) ** 2 / mp.exp(nz(r)),
                -mp.log(2) / 2, mp.log(2) / 2)
    if name == "ATAN64":  # atan t = t + t^3 A(t^2), w = t^2 <= tan(pi/8)^2
        return (lambda w: (mp.atan(mp.sqrt(nz(w))) - mp.sqrt(nz(w))) / nz(w) ** mp.mpf(1.5),
                lambda w: nz(w) ** mp.mpf(1.5) / mp.atan(mp.sqrt(nz(w))), mp.mpf(0), (mp.sqrt(2) - 1) ** 2)
    raise KeyError(name)


def remez_mp(name, n, iters=10, grid=3000):
    """Remez exchange (Remez 1934) in 60-digit arithmetic (mpmath) for the n
    coefficients of table `name` (see _mm64_problem): the reference set starts
    at the Chebyshev extrema and moves to the n + 1 alternating extrema of the
    weighted error on a grid.  Returns (coefficients lowest order first as
    binary64, levelled error)."""
    import mpmath as mp

    g, W, a, b = _mm64_problem(name)
    with mp.workdps(60):
        ref = [(a + b) / 2 - (b - a) / 2 * mp.cos(mp.pi * k / n) for k in range(n + 1)]
        ref = [x + (b - a) / 1000 if abs(x) < (b - a) / 10**6 else x for x in ref]  # W vanishes at 0
        xs = [a + (b - a) * (i + mp.mpf(0.5)) / (grid + 1) for i in range(grid + 1)]
        xs = [x for x in xs if abs(x) > (b - a) / 10**4]
        gx = [g(x) for x in xs]
        wx = [W(x) for x in xs]
        c, lev = None, None
        for _ in range(iters):
            A = mp.matrix(n + 1, n + 1)
            rhs = mp.matrix(n + 1, 1)
            for i, x in enumerate(ref):
                for j in range(n):
                    A[i, j] = x**j
                A[i, n] = (-1) ** i / W(x)
                rhs[i] = g(x)
            sol = mp.lu_solve(A, rhs)
            c = [sol[j] for j in range(n)]
            lev = abs(sol[n])
            e = []
            for x, gv, wv in zip(xs, gx, wx):
                p = mp.mpf(0)
                for cc in reversed(c):
                    p = p * x + cc
                e.append(wv * (gv - p))
            idx, start = [], 0  # the largest |error| of each run of equal sign
            for i in range(1, len(e) + 1):
                if i == len(e) or mp.sign(e[i]) != mp.sign(e[start]):
                    idx.append(max(range(start, i), key=lambda t: abs(e[t])))
                    start = i
            while len(idx) > n + 1:
                idx = idx[1:] if abs(e[idx[0]]) < abs(e[idx[-1]]) else idx[:-1]
            if len(idx) < n + 1:
                break
            ref = [xs[k] for k in idx]
        return [float(v) for v in c], float(lev)


def mm64_table_error(name, npts=2000):
    """Largest weighted error of the committed binary64 table, evaluated
    exactly (60 digits) at npts points of its interval."""
    import mpmath as mp

    g, W, a, b = _mm64_problem(name)
    coeffs = [float.fromhex(h) for h in reversed(MM64_HEX[name])]
    with mp.workdps(60):
        worst = mp.mpf(0)
        for i in range(npts):
            x = a + (b - a) * (i + mp.mpf(0.5)) / npts
            p = mp.mpf(0)
            for cc in reversed(coeffs):
                p = p * x + mp.mpf(cc)
            worst = max(worst, abs(W(x) * (g(x) - p)))
        return float(worst)


def _atan01_q(w):
    """(atan z - z) / z^3 at w = z^2, with its Taylor series near 0."""
    import numpy as np
    w = np.asarray(w, dtype=np.float64)
    z = np.sqrt(w)
    out = np.empty_like(w)
    small = w < 1e-3
    ws = w[small]
    out[small] = -1 / 3 + ws / 5 - ws**2 / 7 + ws**3 / 9 - ws**4 / 11 + ws**5 / 13
    zb = z[~small]
    out[~small] = (np.arctan(zb) - zb) / zb**3
    return out


def remez_atan01(n: int, iters: int = 30, grid: int = 200001):
    """Remez exchange (Remez 1934) for the n coefficients of A in
    atan z = z + z^3 A(z^2), minimising max |atan z - approx| / atan z over
    [0, 1].  A is carried in the Chebyshev basis of 2w - 1 for conditioning;
    the reference set moves to the n + 1 alternating extrema of the weighted
    error on a grid of w.  Returns (coefficients in w, lowest order first,
    as binary64; the levelled error)."""
    import numpy as np
    from numpy.polynomial import chebyshev as cheb
    from numpy.polynomial import polynomial as poly

    def weight(w):  # error of A times this = relative error of atan
        z = np.sqrt(w)
        return np.where(w > 0, w * z / np.arctan(np.where(w > 0, z, 1.0)), 0.0)

    xs = np.linspace(0.0, 1.0, grid)
    ref = np.sort(0.5 - 0.5 * np.cos(np.pi * np.arange(n + 1) / n))
    ref[0] = 0.02  # the weight vanishes at w = 0
    for _ in range(iters):
        a = np.zeros((n + 1, n + 1))
        a[:, :n] = cheb.chebvander(2 * ref - 1, n - 1)
        a[:, n] = (-1.0) ** np.arange(n + 1) / weight(ref)
        sol = np.linalg.solve(a, _atan01_q(ref))
        c, lev = sol[:n], abs(sol[n])
        err = weight(xs) * (_atan01_q(xs) - cheb.chebval(2 * xs - 1, c))
        # the largest |error| of each run of equal sign
        starts = np.concatenate([[0], np.nonzero(np.diff(np.sign(err)))[0] + 1])
        ends = np.concatenate([starts[1:], [grid]])
        idx = [a + int(np.argmax(np.abs(err[a:b]))) for a, b in zip(starts, ends)]
        while len(idx) > n + 1:
            idx = idx[1:] if abs(err[idx[0]]) < abs(err[idx[-1]]) else idx[:-1]
        ref = xs[idx]
    mono = cheb.cheb2poly(c)  # in u = 2w - 1
    cw = np.zeros(n)
    for k, m in enumerate(mono):
        cw[:k + 1] += m * poly.polypow([-1.0, 2.0], k)[:k + 1]
    return cw, lev


def render() -> str:
    c, f = constants()
    t = coeff_tables()
    out = []
    out.append("// Generated by gen_math_constants.py -- do not edit by hand.")
    out.append("// Each value is the correctly rounded binary64/binary32 image of an exact")
    out.append("// mathematical constant computed with 100-digit decimal arithmetic, except")
    out.append("// the minimax tables SIN64, COS64, EXP64, ATAN64 and ATAN01_MM32 (Remez")
    out.append("// exchanges).")
    out.append("#pragma once")
    out.append("")
    out.append("namespace u1m {")
    for k, v in c.items():
        out.append(f"static constexpr double K_{k} = {hexf(v)};  // {v!r}")
    for k, v in f.items():
        out.append(f"static constexpr float KF_{k} = {hexf(v)}f;  // {v!r}")
    for name, arr in t.items():
        ty = "double" if name.endswith("64") else "float"
        suf = "" if ty == "double" else "f"
        body = ", ".join(hexf(x) + suf for x in arr)
        out.append(f"static constexpr int N_{name} = {len(arr)};")
        out.append(f"#define U1M_{name}_COEFFS {body}")
    out.append("}  // namespace u1m")
    out.append("")
    return "\n".join(out)


if __name__ == "__main__":
    sys.stdout.write(render())
