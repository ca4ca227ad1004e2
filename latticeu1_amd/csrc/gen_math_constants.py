#!/usr/bin/env python3
"""Generate u1_math_constants.h: the constants of the shared elementary functions.

Every constant is derived here from first principles with Python's ``decimal``
module at 100 significant digits and rounded once to binary64 / binary32:

* pi from Machin's formula  pi = 16 atan(1/5) - 4 atan(1/239);
* ln 2 from the series      ln 2 = 2 atanh(1/3);
* Taylor coefficients 1/n! and 1/(2k+1) exactly as rationals;
* one minimax table, binary32 atan on [0, 1] for the FP32 update kernels,
  from a Remez exchange (remez_atan01), kept as literals.

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
    # sin r = r + r^3 * P(r^2), P has terms (-1)^k / (2k+3)!, k = 0..7  (through r^17)
    t["SIN64"] = [to_f64(inv_fact(2 * k + 3) * (-1) ** (k + 1)) for k in range(7, -1, -1)]
    # cos r = 1 + r^2 * Q(r^2), Q has (-1)^(k+1)/(2k+2)!, k = 0..7 (through r^16)
    t["COS64"] = [to_f64(inv_fact(2 * k + 2) * (-1) ** (k + 1)) for k in range(7, -1, -1)]
    # exp r = 1 + r + r^2 * E(r), E has 1/(k+2)!, k = 0..11 (through r^13)
    t["EXP64"] = [to_f64(inv_fact(k + 2)) for k in range(11, -1, -1)]
    # atan t = t + t^3 * A(t^2), A has (-1)^(k+1)/(2k+3), k = 0..19 (through t^41)
    t["ATAN64"] = [to_f64(Fraction((-1) ** (k + 1), 2 * k + 3)) for k in range(19, -1, -1)]
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
    out.append("// the ATAN01_MM32 table (minimax coefficients from a Remez exchange).")
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
