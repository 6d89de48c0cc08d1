"""Writes special_functions.json: erf, erfc, erf^-1 and erfc^-1 at sampled doubles,
each correctly rounded to double from a 60+ significant-digit evaluation in
Python's decimal module (no third-party code).

Inputs are the values the reference's test/NeboTest.cpp:472-477 fills its fields
with (spatialops/structured/FieldHelper.h:82-85: sin(start + x + y*X + z*X*Y),
start 0 for input1 and 3*total for input4, input4 made non-negative), taken over
a small window, plus edge points (tiny, near 1, tails).

* erf(x)  = 2/sqrt(pi) * sum_n (-1)^n x^(2n+1) / (n! (2n+1)), summed with
  x^2/ln 10 + 10 extra digits (the largest term is about e^(x^2)), pi by Machin's
  formula.
* erfc(x) = 1 - erf(x) with the same series at a precision that also covers the
  cancellation (x up to 27: e^(x^2) ~ 1e316, so up to ~700 digits); x < 0 via
  2 - erfc(-x).
* erf^-1(y), erfc^-1(q): Newton iteration on the series at high precision from a
  double start (bisection on math.erf / math.erfc), until the step is below
  1e-45 of the root; for |y| > 0.5 the residual is taken through erfc.

Run: python tests/golden/make_special_vectors.py   (takes a few seconds)
"""
import json
import math
import os
from decimal import Decimal, getcontext, localcontext

HERE = os.path.dirname(os.path.abspath(__file__))


def _pi(prec):
    # Machin: pi = 16 atan(1/5) - 4 atan(1/239)
    with localcontext() as c:
        c.prec = prec + 10

        def atan_inv(n):
            x = Decimal(1) / n
            x2 = x * x
            term, s, k = x, x, 1
            while True:
                term *= -x2
                k += 2
                t = term / k
                if abs(t) < Decimal(10) ** -(prec + 8):
                    break
                s += t
            return s
        return +(16 * atan_inv(5) - 4 * atan_inv(239))


def _erf_series(x, prec):
    """erf(x) for a Decimal x at `prec` digits (caller sets enough extra digits)."""
    with localcontext() as c:
        c.prec = prec
        x2 = x * x
        term = x          # x^(2n+1) / n!
        s = x
        n = 0
        tiny = Decimal(10) ** -(prec + 5)
        while True:
            n += 1
            term = term * (-x2) / n
            t = term / (2 * n + 1)
            s += t
            if abs(t) < tiny * abs(s) or t == 0:
                break
        return 2 * s / _pi(prec).sqrt()


def erf_dec(x, digits=60):
    x = Decimal(x)
    extra = int(float(x * x) / math.log(10)) + 10
    v = _erf_series(x, digits + extra)
    with localcontext() as c:
        c.prec = digits
        return +v


def erfc_dec(x, digits=60):
    x = Decimal(x)
    if x < 0:
        with localcontext() as c:
            c.prec = digits
            return +(2 - erfc_dec(-x, digits + 5))
    # 1 - erf(x) loses about x^2/ln 10 digits to cancellation
    extra = 2 * int(float(x * x) / math.log(10)) + 20
    e = _erf_series(x, digits + extra)
    with localcontext() as c:
        c.prec = digits + extra
        v = 1 - e
        c.prec = digits
        return +v


def _deriv(x, prec):
    with localcontext() as c:
        c.prec = prec
        return 2 / _pi(prec).sqrt() * (-(x * x)).exp()


def inv_erf_dec(y, start, digits=50):
    y = Decimal(y)
    x = Decimal(start)
    for _ in range(60):
        extra = int(float(x * x) / math.log(10)) * 2 + 20
        p = digits + extra
        if abs(y) > Decimal("0.5"):
            r = erfc_dec(x if y > 0 else -x, p) - (1 - abs(y))
            r = -r if y > 0 else r
        else:
            r = erf_dec(x, p) - y
        with localcontext() as c:
            c.prec = p
            step = r / _deriv(x, p)
            x = x - step
        if step == 0 or abs(step) <= abs(x) * Decimal(10) ** -(digits - 5):
            break
    return x


def inv_erfc_dec(q, start, digits=50):
    q = Decimal(q)
    if q == 1:
        return Decimal(0)   # exact; Newton would only approach 0 relative to itself
    x = Decimal(start)
    for _ in range(80):
        extra = int(float(x * x) / math.log(10)) * 2 + 20
        p = digits + extra
        r = erfc_dec(x, p) - q
        with localcontext() as c:
            c.prec = p
            step = -r / _deriv(x, p)
            x = x - step
        if step == 0 or abs(step) <= abs(x) * Decimal(10) ** -(digits - 5):
            break
    return x


def to_double(d):
    return float(str(d))   # correctly rounded decimal -> binary64


def reference_inputs():
    """sin(start + x + y*X + z*X*Y) over an 8 x 6 x 5 window (FieldHelper.h:82-85)."""
    X, Y, Z = 8, 6, 5
    total = X * Y * Z
    in1, in4 = [], []
    for z in range(1, Z + 1):
        for y in range(1, Y + 1):
            for x in range(1, X + 1):
                idx = x + y * X + z * X * Y
                in1.append(math.sin(0.0 + idx))
                in4.append(abs(math.sin(3.0 * total + idx)))
    return in1, in4


def main():
    getcontext().prec = 80
    in1, in4 = reference_inputs()
    edge_x = [0.0, 5e-324, 1e-300, 1e-20, 1e-8, 0.25, 0.4769362762044699, 0.5, 1.0, 2.0, 3.5, 5.0, 6.0,
              10.0, 15.0, 26.0, -0.75, -3.0, -6.0]
    edge_y = [5e-324, 1e-300, 1e-20, 1e-8, 0.3, 0.5, 0.5000000000000001, 0.9, 0.99999, 1 - 2.0 ** -40,
              1 - 2.0 ** -53, -0.5, -0.999999, -(1 - 2.0 ** -50)]
    edge_q = [5e-324, 1e-300, 1e-100, 1e-30, 1e-10, 0.01, 0.4999999999999999, 0.5, 0.75, 1.0,
              1.0 + 2.0 ** -52, 1.25, 1.5, 1.99999, 2.0 - 2.0 ** -51]
    xs = in1[:80] + edge_x
    ys = in1[:80] + edge_y
    qs = [q for q in in4[:80] if q > 0.0] + edge_q
    out = {"source": "tests/golden/make_special_vectors.py (decimal, >= 50 digits, rounded to double)",
           "erf": [], "erfc": [], "erf_inv": [], "erfc_inv": []}
    for x in xs:
        out["erf"].append([x.hex(), to_double(erf_dec(x)).hex()])
        out["erfc"].append([x.hex(), to_double(erfc_dec(x)).hex()])
    for y in ys:
        start = _start_inv_erf(y)
        out["erf_inv"].append([y.hex(), to_double(inv_erf_dec(y, start)).hex()])
    for q in qs:
        start = _start_inv_erfc(q)
        out["erfc_inv"].append([q.hex(), to_double(inv_erfc_dec(q, start)).hex()])
    with open(os.path.join(HERE, "special_functions.json"), "w") as f:
        f.write("{\n")
        items = list(out.items())
        for i, (k, v) in enumerate(items):
            end = ",\n" if i + 1 < len(items) else "\n"
            if isinstance(v, list):
                f.write(f' "{k}": [\n' + ",\n".join("  " + json.dumps(p) for p in v) + "\n ]" + end)
            else:
                f.write(f' "{k}": {json.dumps(v)}' + end)
        f.write("}\n")
    print({k: len(v) for k, v in out.items() if isinstance(v, list)})


def _start_inv_erf(y):
    """A double start for Newton (bisection on math.erf / math.erfc)."""
    if y == 0.0:
        return 0.0
    s = 1.0 if y > 0 else -1.0
    a = abs(y)
    lo, hi = 0.0, 30.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if (math.erf(mid) < a) if a <= 0.5 else (math.erfc(mid) > 1.0 - a):
            lo = mid
        else:
            hi = mid
    return s * 0.5 * (lo + hi)


def _start_inv_erfc(q):
    lo, hi = -30.0, 30.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if math.erfc(mid) > q:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


if __name__ == "__main__":
    main()
