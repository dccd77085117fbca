"""ReactorArkode's explicit Runge-Kutta tables (pelephysics_amd/erk_tables.py,
the ode.rk_method codes of Source/Reactions/ReactorArkode.cpp:26-83) against
the Butcher order conditions.

For every rooted tree t with |t| <= q the weights b satisfy
sum_i b_i Phi_i(t) = 1 / gamma(t) (Phi the elementary weights of the stages,
gamma the tree's density), and the embedding d for |t| <= p; each table fails
some condition of order q + 1 (so q is its order), and the rows of A sum to
c.  The Kennedy-Carpenter tables are rational approximations of irrational
coefficients, held to 1e-24.  A planted mistake (one coefficient changed in
its last digit) fails.  The generated header holds the correctly rounded
doubles of the same rationals.
"""
import os
import re
from fractions import Fraction as F

import pytest

from pelephysics_amd import erk_tables as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "pelephysics_amd", "csrc", "erk_tables.h")
APPROX = {31, 41, 53}  # rational approximations (Kennedy & Carpenter)


def _trees(n, memo={1: [()]}):
    """Rooted trees with n vertices, each a sorted tuple of child trees."""
    if n in memo:
        return memo[n]
    out = set()

    def parts(rem, largest, acc):
        if rem == 0:
            out.add(tuple(sorted(acc)))
            return
        for m in range(1, rem + 1):
            for t in _trees(m):
                key = (m, t)
                if largest is not None and key > largest:
                    continue
                parts(rem - m, key, acc + [t])

    parts(n - 1, None, [])
    memo[n] = sorted(out)
    return memo[n]


def _order(t):
    return 1 + sum(_order(c) for c in t)


def _gamma(t):
    g = _order(t)
    for c in t:
        g *= _gamma(c)
    return g


def _phi(t, A, s):
    v = [F(1)] * s
    for ch in t:
        w = _phi(ch, A, s)
        v = [v[i] * sum(A[i][j] * w[j] for j in range(s)) for i in range(s)]
    return v


def _worst(A, w, order, s):
    worst = F(0)
    for n in range(1, order + 1):
        for t in _trees(n):
            ph = _phi(t, A, s)
            worst = max(worst, abs(sum(w[i] * ph[i] for i in range(s)) - F(1, _gamma(t))))
    return worst


def test_tree_counts():
    assert [len(_trees(n)) for n in range(1, 9)] == [1, 1, 2, 4, 9, 20, 48, 115]


@pytest.mark.parametrize("code", sorted(E.TABLES))
def test_order_conditions(code):
    name, s, q, p, c, A, b, d = E.dense(code)
    tol = F(1, 10 ** 24) if code in APPROX else F(0)
    assert all(abs(sum(A[i]) - c[i]) <= tol for i in range(s)), name
    assert all(A[i][j] == 0 for i in range(s) for j in range(i, s)), name  # explicit
    assert _worst(A, b, q, s) <= tol, name
    assert _worst(A, d, p, s) <= tol, name
    assert _worst(A, b, q + 1, s) > F(1, 10 ** 6), name  # order exactly q
    assert name.endswith(f"_{s}_{p}_{q}")


def test_planted_mistake_fails():
    name, s, q, p, c, A, b, d = E.dense(52)
    A = [row[:] for row in A]
    A[4][2] += F(1, 10 ** 6)  # Dormand-Prince a_53 off in its sixth digit
    assert _worst(A, b, q, s) > F(1, 10 ** 9)


def test_defaults_missing_and_fsal():
    assert E.DEFAULT_METHOD == 40 and E.method(999)[0] == "ARKODE_ZONNEVELD_5_3_4"
    for code in (43, 54):
        with pytest.raises(KeyError):
            E.method(code)
    assert {k for k in E.TABLES if E.fsal(k)} == {30, 52}


def test_header_holds_the_rounded_tables():
    with open(HEADER) as f:
        text = f.read()
    entries = re.findall(r"\{(\d+), \"(\w+)\", (\d+), (\d+), (\d+), (\d),\s*\{([^}]*)\},\s*\{([^}]*)\},\s*"
                         r"\{([^}]*)\},\s*\{([^}]*)\}\}", text)
    assert [int(e[0]) for e in entries] == sorted(E.TABLES)
    for code, name, s, q, p, fsal, c, A, b, d in entries:
        n2, s2, q2, p2, cc, AA, bb, dd = E.dense(int(code))
        assert (name, int(s), int(q), int(p), bool(int(fsal))) == (n2, s2, q2, p2, E.fsal(int(code)))
        flat = [AA[i][j] for i in range(s2) for j in range(s2)]
        for txt, vals in ((c, cc), (A, flat), (b, bb), (d, dd)):
            got = [float(x) for x in txt.split(",")]
            assert got == [float(v) for v in vals], (name, txt[:40])
