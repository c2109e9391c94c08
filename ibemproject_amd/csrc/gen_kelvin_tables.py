"""Generate kelvin_tables.inc: straight-line code for the Kelvin derivative tables at one point.

The boundary->inclusion strain kernels need, at x = Gauss point - sphere centre,
  S(beta) = d^beta r      for 3 <= |beta| <= 6   (74 entries)
  T(beta) = d^beta r^-1   for 1 <= |beta| <= 4   (34 entries)
and the stress kernel, at x = Gauss point - stress point, S for |beta| = 3, 4 and T for |beta| = 1, 2.
For f(rho), rho = |x|^2, the Hermite-type expansion gives
  d^beta f(rho) = sum_{j <= beta/2} prod_c beta_c! / (j_c! (beta_c - 2 j_c)!) (2 x_c)^(beta_c - 2 j_c)
                  * f^(|beta| - |j|)(rho),
and with f = rho^(1/2) or rho^(-1/2) the radial derivatives are f^(m) = c_m r rho^-m or
d_m r^-1 rho^-m, c_m = prod_{q<m} (1/2 - q), d_m = prod_{q<m} (-1/2 - q). Every index below is known
here, so the generated code has no loops and no indexed private arrays; the entries are split into
four groups of about equal work, one per wavefront of the kernel that builds the tables.

Run: python3 ibemproject_amd/csrc/gen_kelvin_tables.py   (rewrites kelvin_tables.inc beside it)
"""
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
GROUPS = 4


def tidx(b0, b1, b2):
    d = b0 + b1 + b2
    e = d - b0
    return d * (d + 1) * (d + 2) // 6 + e * (e + 1) // 2 + (e - b1)


def multi(d):
    for b0 in range(d, -1, -1):
        for b1 in range(d - b0, -1, -1):
            yield (b0, b1, d - b0 - b1)


def hcoef(b, j):
    return math.factorial(b) // (math.factorial(j) * math.factorial(b - 2 * j))


def terms(beta):
    """{m: [(coef, (p0, p1, p2))]} for d^beta f(rho)."""
    k = sum(beta)
    out = {}
    for j0 in range(beta[0] // 2 + 1):
        for j1 in range(beta[1] // 2 + 1):
            for j2 in range(beta[2] // 2 + 1):
                c = hcoef(beta[0], j0) * hcoef(beta[1], j1) * hcoef(beta[2], j2)
                m = k - (j0 + j1 + j2)
                out.setdefault(m, []).append((c, (beta[0] - 2 * j0, beta[1] - 2 * j1, beta[2] - 2 * j2)))
    return out


def mono(p):
    f = [f"y{c}[{e}]" for c, e in enumerate(p) if e > 0]
    return "*".join(f) if f else "1.0"


def entry_code(table, beta):
    """One statement writing the entry; table 'S' (r) at slot tidx, 'T' (1/r) at 84 + tidx."""
    P = "PR" if table == "S" else "PT"
    slot = tidx(*beta) + (0 if table == "S" else 84)
    parts = []
    for m, ts in sorted(terms(beta).items()):
        inner = " + ".join((f"{c}.0*{mono(p)}" if c != 1 else mono(p)) for c, p in ts)
        parts.append(f"{P}[{m}]*({inner})")
    return f"  tab[{slot}] = " + " + ".join(parts) + ";", sum(len(v) for v in terms(beta).values())


def main():
    entries = [("S", b) for d in range(3, 7) for b in multi(d)] + \
              [("T", b) for d in range(1, 5) for b in multi(d)]
    coded = [entry_code(t, b) for t, b in entries]
    # greedy balance by term count (+1 per entry for the store)
    order = sorted(range(len(coded)), key=lambda i: -coded[i][1])
    load = [0] * GROUPS
    groups = [[] for _ in range(GROUPS)]
    for i in order:
        g = load.index(min(load))
        groups[g].append(i)
        load[g] += coded[i][1] + 1
    lines = [
        "// Generated by gen_kelvin_tables.py from the expansion documented there; do not edit.",
        "// kelvin_group<G>(y0, y1, y2, PR, PT, tab): y_c[p] = (2 x_c)^p (p <= 6),",
        "// PR[m] = d^m/drho^m rho^(1/2) (m <= 6), PT[m] = d^m/drho^m rho^(-1/2) (m <= 4);",
        "// writes tab[tidx(beta)] = d^beta r (3 <= |beta| <= 6) and tab[84 + tidx(beta)] = d^beta r^-1",
        "// (1 <= |beta| <= 4) for the entries of group G.",
        f"// entries: {len(entries)}, terms per group: {load}",
        "template <int G>",
        "IBEM_HD void kelvin_group(const double* y0, const double* y1, const double* y2,",
        "                          const double* PR, const double* PT, double* tab);",
    ]
    for g in range(GROUPS):
        lines.append("template <>")
        lines.append(f"IBEM_HD void kelvin_group<{g}>(const double* y0, const double* y1, const double* y2,")
        lines.append("                             const double* PR, const double* PT, double* tab) {")
        for i in sorted(groups[g]):
            lines.append(coded[i][0])
        lines.append("}")
    # the stress kernel's subset in one function: d^beta r for |beta| = 3, 4 at tidx(beta) and
    # d^beta r^-1 for |beta| = 1, 2 at 35 + tidx(beta)
    sub = [("S", b) for d in (3, 4) for b in multi(d)] + [("T", b) for d in (1, 2) for b in multi(d)]
    lines += [
        "// kelvin_stress(y0, y1, y2, PR, PT, tab): the same expansion for the stress kernel's entries,",
        "// tab[tidx(beta)] = d^beta r (|beta| = 3, 4) and tab[35 + tidx(beta)] = d^beta r^-1 (|beta| = 1, 2).",
        "IBEM_HD void kelvin_stress(const double* y0, const double* y1, const double* y2,",
        "                           const double* PR, const double* PT, double* tab) {",
    ]
    for t, b in sub:
        code = entry_code(t, b)[0]
        if t == "T":
            code = code.replace(f"tab[{tidx(*b) + 84}]", f"tab[{tidx(*b) + 35}]")
        lines.append(code)
    lines.append("}")
    # the stress kernel's inclusion term at points outside a sphere: every d^beta r (4 <= |beta| <= 6)
    # and d^beta r^-1 (2 <= |beta| <= 6) at x = point - centre, each contracted at once with the
    # sphere's six rows of moments (V[ij * 144 + s], s = tidx - 20 for r, 64 + tidx - 4 for r^-1)
    mom = [("S", b, tidx(*b) - 20) for d in range(4, 7) for b in multi(d)] + \
          [("T", b, 64 + tidx(*b) - 4) for d in range(2, 7) for b in multi(d)]
    assert sorted(s for _, _, s in mom) == list(range(144))
    lines += [
        "// incl_stress_moments(y0, y1, y2, PR, PT, V, e): e[ij] += sum_beta d^beta r V[144 ij + tidx(beta) - 20]",
        "// (4 <= |beta| <= 6) + sum_beta d^beta r^-1 V[144 ij + 60 + tidx(beta)] (2 <= |beta| <= 6);",
        "// PR[m], PT[m] as above with m <= 6 for both.",
        "IBEM_HD void incl_stress_moments(const double* y0, const double* y1, const double* y2,",
        "                                 const double* PR, const double* PT,",
        "                                 const double* __restrict__ V, double* e) {",
        "  double d;",
    ]
    for t, b, s in mom:
        code = entry_code(t, b)[0]
        lines.append("  d = " + code.split(" = ", 1)[1])
        lines.append("  " + " ".join(f"e[{ij}] += d * V[{144 * ij + s}];" for ij in range(6)))
    lines.append("}")
    # the field kernel's outside basis: d^beta r and d^beta r^-1 for 1 <= |beta| <= 5, each pair passed
    # to out(b0, b1, b2, d r, d r^-1) as soon as it is formed (out resolves the table slots)
    lines += [
        "// field_basis(y0, y1, y2, PR, PT, out): out(b0, b1, b2, d^beta r, d^beta r^-1) for every",
        "// 1 <= |beta| <= 5; PR[m], PT[m] as above with m <= 5.",
        "template <class Out>",
        "IBEM_HD void field_basis(const double* y0, const double* y1, const double* y2, const double* PR,",
        "                         const double* PT, Out& out) {",
    ]
    for d in range(1, 6):
        for b in multi(d):
            er = entry_code("S", b)[0].split(" = ", 1)[1].rstrip(";")
            et = entry_code("T", b)[0].split(" = ", 1)[1].rstrip(";")
            lines.append(f"  out({b[0]}, {b[1]}, {b[2]}, {er},")
            lines.append(f"      {et});")
    lines.append("}")
    with open(os.path.join(HERE, "kelvin_tables.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("groups:", load)


if __name__ == "__main__":
    main()
