"""The sphere potentials' closed forms in ibem3d_math.h (pot_coefs), derived here independently.

For a ball B of radius a centred at 0 and w in {1, y_p, y_p y_q}, the kernels use
  Phi^w(x) = int_B w(y) / |x - y| dy,   Psi^w(x) = int_B w(y) |x - y| dy
in the form  Phi^1 = F0, Phi^p = d_p F1, Phi^pq = d_pq F2b + delta_pq F2a  (Psi likewise with G),
with radial functions of rho = |x|^2 inside (polynomials, exact for first and higher derivatives of
Phi and third and higher of Psi) and of r = |x| outside (exact values). The derivation here expands
both kernels in Legendre polynomials of the angle between x and y,
  1 / |x - y| = sum_l r<^l / r>^(l+1) P_l,
  |x - y|     = sum_l (r<^(l+2) / ((2l+3) r>^(l+1)) - r<^l / ((2l-1) r>^(l-1))) P_l,
(each checked numerically against the kernel itself), integrates the angles in closed form with
sympy (only l <= 2 survives for these w) and the radius shell by shell, split at |y| = |x| inside
the ball; the result is compared with the radial functions built from the host-compiled pot_coefs,
at points inside and outside the ball, to 1e-12."""
import ctypes

import numpy as np
import pytest

sp = pytest.importorskip("sympy")

PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
RADIUS = 0.6


def kernel_l(kind, l, rl, rg):
    if kind == "phi":
        return rl ** l / rg ** (l + 1)
    return rl ** (l + 2) / ((2 * l + 3) * rg ** (l + 1)) - rl ** l / ((2 * l - 1) * rg ** (l - 1))


def test_legendre_expansions():
    rng = np.random.default_rng(3)
    from numpy.polynomial import legendre
    for _ in range(20):
        s, r = rng.uniform(0.1, 1.0, 2)
        t = rng.uniform(-1, 1)
        rl, rg = min(s, r), max(s, r)
        d = np.sqrt(r * r + s * s - 2 * r * s * t)
        for kind, want in (("phi", 1 / d), ("psi", d)):
            L = 400 if rg / rl < 1.2 else 120
            coef = np.array([kernel_l(kind, l, rl, rg) for l in range(L)])
            got = legendre.legval(t, coef)
            assert abs(got - want) < 1e-9 * abs(want), (kind, s, r, t)


def derived(a):
    """{(kind, w): sympy expression in x, y, z} for inside and outside, w in 1, p, pq."""
    x, y, z = sp.symbols("x y z", real=True)
    X = (x, y, z)
    s, r = sp.symbols("s r", positive=True)
    th, ph = sp.symbols("theta phi", real=True)
    n = (sp.sin(th) * sp.cos(ph), sp.sin(th) * sp.sin(ph), sp.cos(th))

    def ang(f):  # integral over the unit sphere, the axis of P_l along z
        return sp.integrate(sp.integrate(f * sp.sin(th), (ph, 0, 2 * sp.pi)), (th, 0, sp.pi))
    P = [sp.legendre(l, sp.cos(th)) for l in range(3)]
    A0 = ang(P[0])                       # w = 1:       l = 0
    A1 = ang(n[2] * P[1])                # w = y_p:     l = 1, times xhat_p
    B0 = ang(n[0] * n[0] * P[0])         # w = y_p y_q: l = 0, times delta_pq
    al2 = ang(n[0] * n[0] * P[2])        #              l = 2, al2 delta_pq + be2 xhat_p xhat_q
    be2 = ang(n[2] * n[2] * P[2]) - al2
    assert ang(n[0] * n[1] * P[2]) == 0 and ang(n[0] * n[2] * P[2]) == 0

    def rad(kind, l, deg, inside):
        f = s ** (2 + deg)
        if inside:
            return sp.integrate(f * kernel_l(kind, l, s, r), (s, 0, r)) + \
                sp.integrate(f * kernel_l(kind, l, r, s), (s, r, a))
        return sp.integrate(f * kernel_l(kind, l, s, r), (s, 0, a))
    out = {}
    R = sp.sqrt(x * x + y * y + z * z)
    for inside in (True, False):
        for kind in ("phi", "psi"):
            R0 = rad(kind, 0, 0, inside)
            R1 = rad(kind, 1, 1, inside)
            R20 = rad(kind, 0, 2, inside)
            R22 = rad(kind, 2, 2, inside)
            out[(inside, kind, "1")] = (A0 * R0).subs(r, R)
            for p in range(3):
                out[(inside, kind, p)] = (A1 * R1).subs(r, R) * X[p] / R
            for p, q in PAIRS:
                dl = 1 if p == q else 0
                out[(inside, kind, (p, q))] = (B0 * R20 * dl + R22 * al2 * dl).subs(r, R) + \
                    (R22 * be2).subs(r, R) * X[p] * X[q] / R ** 2
    return X, out


@pytest.fixture(scope="module")
def forms(hostmath):
    hostmath.hc_pot_coefs.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    X, der = derived(sp.Rational(3, 5))
    x, y, z = X
    rho = x * x + y * y + z * z
    R = sp.sqrt(rho)
    lib = {}
    for inside in (True, False):
        C = np.zeros(32)
        hostmath.hc_pot_coefs(RADIUS, 1 if inside else 0, C.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        C = C.reshape(8, 4)
        basis = [rho, rho ** 2, rho ** 3, rho ** 4] if inside else [R, 1 / R, 0, 0]
        f = [sum(sp.Float(C[k, j], 17) * basis[j] for j in range(4)) for k in range(8)]
        F0, F1, F2a, F2b, G0, G1, G2a, G2b = f
        for kind, (H0, H1, H2a, H2b) in (("phi", (F0, F1, F2a, F2b)), ("psi", (G0, G1, G2a, G2b))):
            lib[(inside, kind, "1")] = H0
            for p in range(3):
                lib[(inside, kind, p)] = sp.diff(H1, X[p])
            for p, q in PAIRS:
                lib[(inside, kind, (p, q))] = sp.diff(H2b, X[p], X[q]) + (H2a if p == q else 0)
    return X, der, lib


def points(inside, k=4):
    rng = np.random.default_rng(11 if inside else 12)
    out = []
    while len(out) < k:
        v = rng.uniform(-1.5, 1.5, 3)
        rr = np.linalg.norm(v)
        if (rr < 0.9 * RADIUS and rr > 0.1) if inside else (rr > 1.1 * RADIUS):
            out.append(v)
    return out


@pytest.mark.parametrize("inside", [True, False])
@pytest.mark.parametrize("kind", ["phi", "psi"])
def test_potentials_match_derivation(forms, inside, kind):
    X, der, lib = forms
    # inside: compare the derivatives of the order the closed form is exact for; outside: values
    order = (1 if kind == "phi" else 3) if inside else 0
    combos = [()] if order == 0 else (
        [(i,) for i in range(3)] if order == 1 else
        [(i, j, k) for i in range(3) for j in range(i, 3) for k in range(j, 3)])
    ws = ["1"] + list(range(3)) + PAIRS
    for w in ws:
        for c in combos:
            e_der = sp.diff(der[(inside, kind, w)], *[X[i] for i in c]) if c else der[(inside, kind, w)]
            e_lib = sp.diff(lib[(inside, kind, w)], *[X[i] for i in c]) if c else lib[(inside, kind, w)]
            fd = sp.lambdify(X, e_der, "numpy")
            fl = sp.lambdify(X, e_lib, "numpy")
            for v in points(inside):
                a, b = float(fd(*v)), float(fl(*v))
                scale = max(abs(a), 1e-3 * RADIUS ** 3)
                assert abs(a - b) <= 1e-12 * scale, (inside, kind, w, c, v, a, b)
