"""Generates pp_qng_tables.h: the abscissae and weights of the non-adaptive
Gauss-Kronrod-Patterson sequence (10-, 21-, 43- and 87-point rules on
[-1, 1]) that QUADPACK's QNG routine uses (Piessens, de Doncker-Kapenga,
Ueberhuber and Kahaner, QUADPACK, Springer 1983; T. N. L. Patterson, "The
optimum addition of points to quadrature formulae", Math. Comp. 22, 1968).
GSL's gsl_integration_qng, which pplacer calls through Integration.qng
(integration.ml:28), carries the same tables.

Every rule is derived here from its definition rather than copied:

* x1, w10: the 10-point Gauss-Legendre rule (zeros of P_10);
* x2: the zeros of the Stieltjes polynomial E_11, orthogonal to every
  polynomial of degree < 11 with weight P_10 (the 21-point Kronrod rule);
* x3: the zeros of the degree-22 polynomial orthogonal to degree < 22 with
  weight P_10 * E_11 (Patterson's 43-point extension);
* x4: the zeros of the degree-44 polynomial orthogonal to degree < 44 with
  weight P_10 * E_11 * F_22 (the 87-point extension).

The extension polynomials are solved for in exact rational arithmetic (their
moment conditions have rational coefficients); their roots and the weights of
each rule (exact on every even monomial the rule has free weights for) are
computed with mpmath at 250 digits.  Tables list the positive abscissae in
descending order; weights for a +/- pair multiply f(c + h x) + f(c - h x); the
last entry of w21b, w43b and w87b is the weight of the centre.

    python pplacer_amd/csrc/gen_qng_tables.py   # rewrites pp_qng_tables.h
"""
import os
from fractions import Fraction

import mpmath as mp

mp.mp.dps = 250
HERE = os.path.dirname(os.path.abspath(__file__))


# polynomials as coefficient lists, index = power, Fraction entries
def pmul(a, b):
    out = [Fraction(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                out[i + j] += x * y
    return out


def moment(n):
    """integral of x^n over [-1, 1]"""
    return Fraction(0) if n % 2 else Fraction(2, n + 1)


def integral(p):
    return sum(c * moment(i) for i, c in enumerate(p))


def legendre(n):
    p0, p1 = [Fraction(1)], [Fraction(0), Fraction(1)]
    for k in range(1, n):
        # (k+1) P_{k+1} = (2k+1) x P_k - k P_{k-1}
        xp = [Fraction(0)] + p1
        nxt = [Fraction(0)] * (k + 2)
        for i, c in enumerate(xp):
            nxt[i] += Fraction(2 * k + 1, k + 1) * c
        for i, c in enumerate(p0):
            nxt[i] -= Fraction(k, k + 1) * c
        p0, p1 = p1, nxt
    return p1


def solve_fraction(A, b):
    n = len(A)
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for col in range(n):
        piv = next(r for r in range(col, n) if M[r][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        pv = M[col][col]
        M[col] = [v / pv for v in M[col]]
        for r in range(n):
            if r != col and M[r][col] != 0:
                f = M[r][col]
                M[r] = [a - f * c for a, c in zip(M[r], M[col])]
    return [M[i][n] for i in range(n)]


def extension(weight, degree):
    """Monic polynomial of the given degree (same parity as degree) orthogonal
    to x^k, k < degree, under the weight polynomial on [-1, 1]."""
    powers = list(range(degree % 2, degree, 2))  # free lower coefficients
    # conditions: integral(weight * p * x^k) = 0 for k with a non-vanishing
    # integrand parity (weight * p * x^k even)
    ks = [k for k in range(degree) if (len(weight) - 1 + degree + k) % 2 == 0]
    assert len(ks) == len(powers)
    A, b = [], []
    for k in ks:
        wk = pmul(weight, [Fraction(0)] * k + [Fraction(1)])
        A.append([integral(pmul(wk, [Fraction(0)] * j + [Fraction(1)])) for j in powers])
        b.append(-integral(pmul(wk, [Fraction(0)] * degree + [Fraction(1)])))
    coef = solve_fraction(A, b)
    p = [Fraction(0)] * (degree + 1)
    p[degree] = Fraction(1)
    for j, c in zip(powers, coef):
        p[j] = c
    return p


def positive_roots(p):
    """Positive roots, descending, of an even or odd polynomial whose roots are
    real and simple: roots of q(y) = p(sqrt y) (p / x for odd p) in y = x^2."""
    deg = len(p) - 1
    shift = deg % 2
    q = [p[i] for i in range(shift, deg + 1, 2)]  # q[j] = coefficient of y^j
    qm = [mp.mpf(c.numerator) / c.denominator for c in reversed(q)]
    ys = mp.polyroots(qm, maxsteps=2000, extraprec=3000)
    xs = []
    for y in ys:
        assert abs(mp.im(y)) < mp.mpf(10) ** -100, y
        y = mp.re(y)
        if y > mp.mpf(10) ** -100:
            xs.append(mp.sqrt(y))
    return sorted(xs, reverse=True)


def weights(nodes, center):
    """Weights of the symmetric rule on +/-nodes (and 0 if center), exact on
    x^(2k), k < number of free weights."""
    n = len(nodes) + (1 if center else 0)
    A, b = [], []
    for k in range(n):
        row = [2 * x ** (2 * k) for x in nodes]
        if center:
            row.append(mp.mpf(1) if k == 0 else mp.mpf(0))
        A.append(row)
        b.append(mp.mpf(2) / (2 * k + 1))
    return list(mp.lu_solve(mp.matrix(A), mp.matrix(b)))


def tables():
    p10 = legendre(10)
    e11 = extension(p10, 11)
    f22 = extension(pmul(p10, e11), 22)
    g44 = extension(pmul(pmul(p10, e11), f22), 44)
    x1 = positive_roots(p10)
    x2 = positive_roots(e11)
    x3 = positive_roots(f22)
    x4 = positive_roots(g44)
    assert (len(x1), len(x2), len(x3), len(x4)) == (5, 5, 11, 22)
    w10 = weights(x1, False)
    w21 = weights(x1 + x2, True)
    w43 = weights(x1 + x2 + x3, True)
    w87 = weights(x1 + x2 + x3 + x4, True)
    return {
        "x1": x1, "w10": w10,
        "x2": x2, "w21a": w21[:5], "w21b": w21[5:],
        "x3": x3, "w43a": w43[:10], "w43b": w43[10:],
        "x4": x4, "w87a": w87[:21], "w87b": w87[21:],
    }


def write(path):
    t = tables()
    lines = [
        "// Generated by gen_qng_tables.py: abscissae and weights of the 10-, 21-,",
        "// 43- and 87-point Gauss-Kronrod-Patterson rules of QUADPACK's QNG on",
        "// [-1, 1], derived from their defining orthogonality conditions.  Positive",
        "// abscissae in descending order; a pair's weight multiplies f(c+hx)+f(c-hx);",
        "// the last entry of w21b, w43b and w87b weights the centre.  The includer",
        "// defines PP_QNG_STORAGE (storage class of the arrays).",
        "#pragma once",
        "",
    ]
    for name, vals in t.items():
        lines.append("PP_QNG_STORAGE double pp_qng_%s[%d] = {" % (name, len(vals)))
        for v in vals:
            lines.append("    %s," % mp.nstr(v, 25, min_fixed=-1, max_fixed=1))
        lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    write(os.path.join(HERE, "pp_qng_tables.h"))
