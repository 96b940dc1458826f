#!/usr/bin/env python3
"""Exact derivation of the vertex-pyramid coefficient tables for the six piecewise-polynomial SPH
kernels, written to kernel_tables.h (C header shared by the HIP kernels and the CPU oracle).

The reference ships the same numbers as generated C++ (src/kernel_spec_data.cpp:12-163, produced by
tools/kernel_tables/gen_cpp.py from the derivation in tools/kernel_tables/derive.py:38-63).  This is
our own derivation from the mathematics, in exact rational arithmetic:

  kernel      W(r) = A3 / H^3 * w(q),  q = r / H,  w = 0 for q >= 1, A3 = A3_pi / pi,
              4 pi A3 int_0^1 w(q) q^2 dq = 1  (3D),  2 pi A2 int_0^1 w q dq = 1  (2D).
  pyramid     apex at the particle, base = the triangle (foot of the perpendicular on the wall plane,
              foot on the edge line, edge vertex at azimuth phi), wall distance r0, q0 = r0 / H.
              Its kernel mass is I_P(phi) = int_0^phi dI_P/dphi' dphi' with
              dI_P/dphi = A3 q0 int_{q0}^{q_b(phi)} m(q) / q^2 dq,   m(q) = int_0^q w(s) s^2 ds,
              q_b = vertex radius / H.  Writing F_k = antiderivative of m_k / q^2 in region k
              (a Laurent polynomial) and mu = q0 / q_b:
                 dI_P/dphi = A3 [ sum_i a_i^(k) mu^i q0^(1-i)  +  B^(j,k)(q0) ],
              a_i^(k) = coefficient of q^(-i) in F_k, and B^(j,k) collects -q0 F_j(q0) and the
              region-crossing constants q0 [F_j(c_j) + sum_{j<l<k} (F_l(c_l) - F_l(c_{l-1})) - F_k(c_{k-1})].

Regions are 1..n for the kernel's n pieces and n+1 for q > 1.  Table layout (C, 0-based):
a[k-1][1-i] for i = 1, 0, -1, .., -13 and B[j-1][k-1][p] = coefficient of q0^p, p = 0..15.

Run:  python3 kiwi_amd/csrc/gen_kernel_tables.py            (rewrites kernel_tables.h)
      python3 kiwi_amd/csrc/gen_kernel_tables.py --check    (exit 1 if the header is stale)
"""
from fractions import Fraction as Q
import os
import sys

NI, NB, MAXREG = 15, 16, 4


class Laurent:
    """Finite Laurent polynomial sum c_p x^p with exact rational coefficients."""

    def __init__(self, terms=None):
        self.t = {p: Q(c) for p, c in (terms or {}).items() if c != 0}

    def __add__(self, o):
        r = dict(self.t)
        for p, c in o.t.items():
            r[p] = r.get(p, Q(0)) + c
        return Laurent(r)

    def __mul__(self, o):
        if not isinstance(o, Laurent):
            return Laurent({p: c * o for p, c in self.t.items()})
        r = {}
        for p, c in self.t.items():
            for s, d in o.t.items():
                r[p + s] = r.get(p + s, Q(0)) + c * d
        return Laurent(r)

    __rmul__ = __mul__

    def __neg__(self):
        return self * Q(-1)

    def power(self, n):
        r = Laurent({0: 1})
        for _ in range(n):
            r = r * self
        return r

    def integral(self):
        if -1 in self.t:
            raise ValueError("1/x term: antiderivative would be a logarithm")
        return Laurent({p + 1: c / (p + 1) for p, c in self.t.items()})

    def __call__(self, x):
        return sum((c * Q(x) ** p for p, c in self.t.items()), Q(0))


X = Laurent({1: 1})
ONE = Laurent({0: 1})


def spline(terms):
    """sum coef * (shift - q)_+^n  ->  (breakpoints, [piece polynomials])."""
    shifts = sorted({Q(s) for _, s, _ in terms})
    pieces = []
    for hi in shifts:
        poly = Laurent()
        for c, s, n in terms:
            if Q(s) >= hi:
                poly = poly + Q(c) * (Laurent({0: Q(s)}) + -X).power(n)
        pieces.append(poly)
    return shifts, pieces


def wendland(n, coefs):
    poly = (ONE + -X).power(n) * Laurent({i: Q(c) for i, c in enumerate(coefs)})
    return [Q(1)], [poly]


# key, name, (breaks, pieces), radkern (support in units of h), hfact (Phantom's default h/(m/rho)^(1/3)).
# radkern and hfact are conventions, not derived: the values follow the reference's table
# (src/kernel_spec_data.cpp:16,41,66,91,116,141).
KERNELS = [
    ("M4", "Cubic spline", spline([(1, 1, 3), (-4, Q(1, 2), 3)]), 2.0, 1.2),
    ("M5", "Quartic spline", spline([(1, 1, 4), (-5, Q(3, 5), 4), (10, Q(1, 5), 4)]), 2.5, 1.1),
    ("M6", "Quintic spline", spline([(1, 1, 5), (-6, Q(2, 3), 5), (15, Q(1, 3), 5)]), 3.0, 1.0),
    ("WC2", "Wendland C2", wendland(4, [1, 4]), 2.0, 1.3),
    ("WC4", "Wendland C4", wendland(6, [1, 6, Q(35, 3)]), 2.0, 1.5),
    ("WC6", "Wendland C6", wendland(8, [1, 8, 25, 32]), 2.0, 1.6),
]


def derive(breaks, pieces):
    n = len(pieces)
    nreg = n + 1
    cs = [Q(0)] + list(breaks)
    # enclosed-mass shape m_k(q) = int_0^q w s^2 ds, continuous across the breakpoints
    ms, acc = [], Q(0)
    for k, poly in enumerate(pieces):
        P = (poly * Laurent({2: 1})).integral()
        mk = P + Laurent({0: acc - P(cs[k])})
        ms.append(mk)
        acc = mk(cs[k + 1])
    ms.append(Laurent({0: acc}))  # outside the support: the total mass shape
    norm3 = 1 / (4 * acc)          # A3_pi
    tot2 = Q(0)
    for k, poly in enumerate(pieces):
        P = (poly * X).integral()
        tot2 += P(cs[k + 1]) - P(cs[k])
    norm2 = 1 / (2 * tot2)         # A2_pi
    F = [(m * Laurent({-2: 1})).integral() for m in ms]
    a = [[Q(0)] * NI for _ in range(MAXREG)]
    for k in range(nreg):
        for p, c in F[k].t.items():
            i = -p
            assert i != 0 and 1 - i < NI, (p, c)
            a[k][1 - i] = c
    B = [[[Q(0)] * NB for _ in range(MAXREG)] for _ in range(MAXREG)]
    for j in range(1, nreg + 1):
        for k in range(j, nreg + 1):
            b = -(F[j - 1] * X)
            if k > j:
                const = F[j - 1](cs[j])
                for l in range(j + 1, k):
                    const += F[l - 1](cs[l]) - F[l - 1](cs[l - 1])
                const -= F[k - 1](cs[k - 1])
                b = b + Laurent({1: const})
            for p, c in b.t.items():
                assert 0 <= p < NB, (p, c)
                B[j - 1][k - 1][p] = c
    return nreg, norm2, norm3, a, B


def fmt(x):
    return repr(float(x))


def render():
    out = []
    out.append("// GENERATED by kiwi_amd/csrc/gen_kernel_tables.py -- exact re-derivation; do not edit.")
    out.append("// Vertex-pyramid coefficient tables of the six piecewise-polynomial SPH kernels (M4, M5, M6,")
    out.append("// Wendland C2/C4/C6).  Layout: a[k-1][1-i] multiplies mu^i q0^(1-i) (i = 1, 0, .., -13);")
    out.append("// B[j-1][k-1][p] multiplies q0^p; c[] are the region breakpoints in units of the support H.")
    out.append("#pragma once")
    out.append("")
    out.append(f"#define KIWI_TAB_NI {NI}")
    out.append(f"#define KIWI_TAB_NB {NB}")
    out.append(f"#define KIWI_TAB_MAXREG {MAXREG}")
    out.append(f"#define KIWI_TAB_NKERNELS {len(KERNELS)}")
    out.append("")
    out.append("typedef struct KiwiKernelTable {")
    out.append("    const char* key;")
    out.append("    int nreg;                 /* pieces + 1 */")
    out.append("    double c[KIWI_TAB_MAXREG]; /* breakpoints c_1..c_n (c_n = 1), zero padded */")
    out.append("    double A2_pi, A3_pi;      /* normalisation: A_d = A_d_pi / pi */")
    out.append("    double radkern;           /* support H in units of h */")
    out.append("    double hfact;             /* default h = hfact (m/rho)^(1/3) */")
    out.append("    double a[KIWI_TAB_MAXREG][KIWI_TAB_NI];")
    out.append("    double B[KIWI_TAB_MAXREG][KIWI_TAB_MAXREG][KIWI_TAB_NB];")
    out.append("} KiwiKernelTable;")
    out.append("")
    out.append("/* KIWI_KERNEL_TABLE_DATA expands to the six initialisers, in the order M4 M5 M6 WC2 WC4 WC6. */")
    out.append("#define KIWI_KERNEL_TABLE_DATA \\")
    rows = []
    for key, name, (breaks, pieces), radkern, hfact in KERNELS:
        nreg, n2, n3, a, B = derive(breaks, pieces)
        c = list(breaks) + [Q(0)] * (MAXREG - len(breaks))
        s = [f'{{ /* {name} */ "{key}", {nreg}, ', "{" + ", ".join(fmt(v) for v in c) + "}, ",
             f"{fmt(n2)}, {fmt(n3)}, {fmt(radkern)}, {fmt(hfact)}, ",
             "{" + ", ".join("{" + ", ".join(fmt(v) for v in row) + "}" for row in a) + "}, ",
             "{" + ", ".join("{" + ", ".join("{" + ", ".join(fmt(v) for v in r) + "}" for r in blk) + "}"
                             for blk in B) + "} }"]
        rows.append("".join(s))
    out.append(", \\\n".join("    " + r for r in rows))
    out.append("")
    return "\n".join(out)


def main():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernel_tables.h")
    text = render()
    if "--check" in sys.argv:
        with open(path) as f:
            sys.exit(0 if f.read() == text else 1)
    with open(path, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main()
