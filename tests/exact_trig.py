"""Correctly rounded sin and cos from exact decimal arithmetic: the exact
binary value of x reduced modulo 2 pi (pi from Machin's formula, 200
digits), the Taylor series in 160-digit decimals, then Python's correctly
rounded decimal-to-double conversion.  Independent of every C library and of
both implementations under test; slow, for a few thousand arguments."""
from decimal import Decimal, localcontext


def _atan_inv(n, prec):
    with localcontext() as ctx:
        ctx.prec = prec
        x = Decimal(1) / n
        x2 = x * x
        term, total, k = x, x, 1
        while True:
            term = -term * x2
            t = term / (2 * k + 1)
            if t == 0 or abs(t) < Decimal(10) ** -(prec + 5):
                return total
            total += t
            k += 1


with localcontext() as _c:
    _c.prec = 200
    PI = 16 * _atan_inv(5, 200) - 4 * _atan_inv(239, 200)


def _series(x, odd):
    with localcontext() as ctx:
        ctx.prec = 160
        X = Decimal(x)  # exact value of the double
        two_pi = 2 * PI
        X = X - (X / two_pi).to_integral_value() * two_pi
        x2 = X * X
        term = X if odd else Decimal(1)
        total = term
        n = 1 if odd else 0
        tiny = Decimal(10) ** -150
        while True:
            term = -term * x2 / ((n + 1) * (n + 2))
            n += 2
            total += term
            if abs(term) < tiny and n > 4:
                return float(total)


def cr_sin(x):
    x = float(x)
    return x if x == 0.0 else _series(x, True)


def cr_cos(x):
    return _series(float(x), False)
