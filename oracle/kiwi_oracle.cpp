// KIWI oracle: a plain C++ (CPU, FP64) restatement of the reference's kernel-integral mapping.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path in kiwi_amd/csrc is compared
// against; it is imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
// and never by the product path (which fails loudly when its HIP library is missing).
//
// Each function restates one reference function and cites it.  Arithmetic order follows the
// reference closely on purpose (the point of an oracle is to reproduce the reference's numbers to
// the ulp level), but the code, data layout and names are this project's own.
//
// Parity pins (how each part is checked; see tests/test_oracle.py):
//   m4 tiers / face walk  : oracle/_ref (the reference built from source) on random inputs, closed forms
//                           (enclosing box = 1, partition additivity), brute-force quadrature
//   generic tiers         : oracle/_ref, quadrature, normalisation, generic == cubic for M4
//   sph-to-grid loop      : golden fixtures written by the reference CLI (tests/golden), mass conservation
//   Voronoi BFS           : the lattice cross-check (Voronoi cells of a cubic lattice are the Cartesian
//                           cells) and the reference CLI's own `map` output on a small fixture
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include <quadmath.h>

#include "../kiwi_amd/csrc/kernel_tables.h"

namespace {

const KiwiKernelTable kKiwiKernelTables[KIWI_TAB_NKERNELS] = {KIWI_KERNEL_TABLE_DATA};

// ------------------------------------------------------------------------------------------------
// Cubic spline (M4), analytic: the three tiers of src/kernel_integral.cpp:222-400.
// ------------------------------------------------------------------------------------------------

// The cubic path is written once for a scalar type T: T = double is the oracle proper (the
// reference's arithmetic, bit for bit); T = __float128 evaluates the same algorithm in binary128 so
// that cells where the double result carries visible rounding error can be arbitrated.  Decimal
// constants are written as ratios C(n, d) = T(n) / T(d): in double that is exactly the literal the
// reference uses (the nearest double to n/d), in binary128 it is the exact rational to 113 bits.
template <typename T>
inline T C(int n, int d) { return T(n) / T(d); }

inline double q_sqrt(double x) { return std::sqrt(x); }
inline double q_acos(double x) { return std::acos(x); }
inline double q_log(double x) { return std::log(x); }
inline double q_atan(double x) { return std::atan(x); }
inline double q_sin(double x) { return std::sin(x); }
inline double q_fabs(double x) { return std::fabs(x); }
inline double q_cos(double x) { return std::cos(x); }
inline double q_atanh(double x) { return std::atanh(x); }
inline double q_pow(double x, int n) { return std::pow(x, n); }
inline double q_fmin(double a, double b) { return std::fmin(a, b); }
inline __float128 q_sqrt(__float128 x) { return sqrtq(x); }
inline __float128 q_acos(__float128 x) { return acosq(x); }
inline __float128 q_log(__float128 x) { return logq(x); }
inline __float128 q_atan(__float128 x) { return atanq(x); }
inline __float128 q_sin(__float128 x) { return sinq(x); }
inline __float128 q_fabs(__float128 x) { return fabsq(x); }
inline __float128 q_cos(__float128 x) { return cosq(x); }
inline __float128 q_atanh(__float128 x) { return atanhq(x); }
inline __float128 q_pow(__float128 x, int n) { return powq(x, static_cast<__float128>(n)); }
inline __float128 q_fmin(__float128 a, __float128 b) { return fminq(a, b); }
template <typename T> inline T q_pi();
template <> inline double q_pi<double>() { return M_PI; }
template <> inline __float128 q_pi<__float128>() { return M_PIq; }

inline void sincos_pair(double x, double& s, double& c) {
#if defined(__GLIBC__)
    ::sincos(x, &s, &c);  // what the reference uses on glibc (src/kernel_integral.cpp:200-211)
#else
    s = std::sin(x);
    c = std::cos(x);
#endif
}
inline void sincos_pair(__float128 x, __float128& s, __float128& c) { sincosq(x, &s, &c); }

template <typename T>
struct M4Wall {  // include/kiwi/kernel_integral.hpp:209-214
    bool ok = false;
    T h = 0, r0 = 0, h2 = 0, r03 = 0, q2 = 0, q3 = 0, qi2 = 0, qi3 = 0;
    T B1 = 0, B2 = 0, B3 = 0;
};

template <typename T>
struct M4Edge {  // include/kiwi/kernel_integral.hpp:216-221
    bool ok = false;
    T R0 = 0, a = 0, a2 = 0, D2 = 0, D3 = 0;
};

// The I-series at angle phi (src/kernel_integral.cpp:278-294, 367-386; identical in all three places).
template <typename T>
struct M4Series {
    T I0, Im2, Im4, I1, Im1, Im3, Im5;
};

template <typename T>
M4Series<T> m4_series(T phi, T a, T a2) {
    T sn, cs;
    sincos_pair(phi, sn, cs);
    const T c2 = cs * cs;
    const T mu = cs / a / q_sqrt(T(1) + c2 / a2);
    const T tn = sn / cs;
    M4Series<T> s;
    s.I0 = phi;
    s.Im2 = phi + a2 * tn;
    s.Im4 = phi + 2 * a2 * tn + C<T>(1, 3) * a2 * a2 * tn * (T(2) + T(1) / c2);
    const T u = sn * q_sqrt(1 - mu * mu);
    const T lg = q_log(1 + u) - q_log(1 - u);
    s.I1 = q_atan(u / a);
    s.Im1 = a / T(2) * lg + s.I1;
    s.Im3 = s.Im1 + a * (T(1) + a2) / T(4) * (2 * u / (1 - u * u) + lg);
    s.Im5 = s.Im3 + a * (T(1) + a2) * (T(1) + a2) / T(16) * ((10 * u - 6 * u * u * u) / (1 - u * u) / (1 - u * u) + T(3) * lg);
    return s;
}

// compute_wall_constants, src/kernel_integral.cpp:222-247
template <typename T>
M4Wall<T> m4_wall(T h, T r0) {
    M4Wall<T> w;
    w.h = h;
    w.r0 = r0;
    if (r0 == T(0)) return w;
    w.ok = true;
    w.h2 = h * h;
    w.r03 = r0 * r0 * r0;
    w.q2 = r0 / h * r0 / h;
    w.q3 = w.q2 * r0 / h;
    w.qi2 = h / r0 * h / r0;
    w.qi3 = w.qi2 * h / r0;
    if (r0 >= T(2) * h) {
        w.B3 = w.h2 * h / T(4);
    } else if (r0 > h) {
        const T common = C<T>(-4, 3) + (r0 / h) - C<T>(3, 10) * w.q2 + C<T>(1, 30) * w.q3 - C<T>(1, 15) * w.qi3;
        w.B3 = w.r03 / T(4) * (common + C<T>(8, 5) * w.qi2);
        w.B2 = w.r03 / T(4) * common;
    } else {
        const T common = C<T>(-2, 3) + C<T>(3, 10) * w.q2 - C<T>(1, 10) * w.q3;
        w.B3 = w.r03 / T(4) * (common + C<T>(7, 5) * w.qi2);
        w.B2 = w.r03 / T(4) * (common - C<T>(1, 5) * w.qi2);
        w.B1 = w.r03 / T(4) * common;
    }
    return w;
}

// The region-1 vertex form inside the r0h3/pi bracket (src/kernel_integral.cpp:389-390).
template <typename T>
inline T m4_inner(const M4Wall<T>& w, const M4Series<T>& s) {
    return C<T>(1, 6) * s.Im2 - C<T>(3, 40) * w.q2 * s.Im4 + C<T>(1, 40) * w.q3 * s.Im5 + w.B1 / w.r03 * s.I0;
}

// compute_edge_constants, src/kernel_integral.cpp:253-350
template <typename T>
M4Edge<T> m4_edge(const M4Wall<T>& w, T R0) {
    M4Edge<T> e;
    e.R0 = R0;
    if (!w.ok || R0 == T(0)) return e;
    e.ok = true;
    const T h = w.h, r0 = w.r0;
    e.a = R0 / r0;
    e.a2 = e.a * e.a;
    const T line2 = r0 * r0 + R0 * R0;
    T D2 = 0, D3 = 0;
    if (line2 <= w.h2) {
        const T p1 = q_acos(R0 / q_sqrt(h * h - r0 * r0));
        const M4Series<T> s = m4_series(p1, e.a, e.a2);
        D2 = C<T>(-1, 6) * s.Im2 + C<T>(1, 4) * (r0 / h) * s.Im3 - C<T>(3, 20) * w.q2 * s.Im4 +
             C<T>(1, 30) * w.q3 * s.Im5 - C<T>(1, 60) * w.qi3 * s.I1 + (w.B1 - w.B2) / w.r03 * s.I0;
    }
    if (line2 <= T(4) * w.h2) {
        const T p2 = q_acos(R0 / q_sqrt(T(4) * h * h - r0 * r0));
        const M4Series<T> s = m4_series(p2, e.a, e.a2);
        D3 = C<T>(1, 3) * s.Im2 - C<T>(1, 4) * (r0 / h) * s.Im3 + C<T>(3, 40) * w.q2 * s.Im4 -
             C<T>(1, 120) * w.q3 * s.Im5 + C<T>(4, 15) * w.qi3 * s.I1 + (w.B2 - w.B3) / w.r03 * s.I0 + D2;
    }
    e.D2 = D2;
    e.D3 = D3;
    return e;
}

// vertex_integral, src/kernel_integral.cpp:356-400
template <typename T>
T m4_vertex(const M4Wall<T>& w, const M4Edge<T>& e, T phi) {
    if (!w.ok || !e.ok) return T(0);
    if (phi == T(0)) return T(0);
    T sn, cs;
    sincos_pair(phi, sn, cs);
    const T R = e.R0 / cs;
    const T r2 = w.r0 * w.r0 + R * R;
    const M4Series<T> s = m4_series(phi, e.a, e.a2);
    const T r0 = w.r0, h = w.h;
    if (r2 < w.h2) return w.q3 / q_pi<T>() * m4_inner(w, s);
    if (r2 < T(4) * w.h2)
        return w.q3 / q_pi<T>() *
               (C<T>(1, 4) * (C<T>(4, 3) * s.Im2 - (r0 / h) * s.Im3 + C<T>(3, 10) * w.q2 * s.Im4 -
                              C<T>(1, 30) * w.q3 * s.Im5 + C<T>(1, 15) * w.qi3 * s.I1) +
                w.B2 / w.r03 * s.I0 + e.D2);
    return w.q3 / q_pi<T>() * (C<T>(-1, 4) * w.qi3 * s.I1 + w.B3 / w.r03 * s.I0 + e.D3);
}

// ------------------------------------------------------------------------------------------------
// Tabulated cubic path: generate_density_table + gridded_integral, src/kernel_integral.cpp:57-190.
// ------------------------------------------------------------------------------------------------

struct M4Table {
    int nr01 = 50, nr02 = 199, nR01 = 150, nR02 = 150, nph1 = 150, nph2 = 150;
    double rl = 0.1, mul = 0.98, cphil = 0.98;
    int nr0 = 249, nR0 = 300, nph = 300;
    std::vector<double> v;  // (nr0+2) x (nR0+1) x (nph+1)
    double& at(int i, int j, int k) { return v[(static_cast<size_t>(i) * (nR0 + 1) + j) * (nph + 1) + k]; }
    double at(int i, int j, int k) const { return v[(static_cast<size_t>(i) * (nR0 + 1) + j) * (nph + 1) + k]; }
};

double m4_full(double h, double phi, double r0, double R0);

const M4Table& m4_table() {
    static const M4Table t = [] {
        M4Table t;
        t.v.assign(static_cast<size_t>(t.nr0 + 2) * (t.nR0 + 1) * (t.nph + 1), 0.0);
        const int n = t.nR01;
        for (int i = 1; i <= t.nr0; ++i) {
            const double r0 = (i <= t.nr01) ? (t.rl / t.nr01) * i : t.rl + ((2.0 - t.rl) / t.nr02) * (i - t.nr01);
            for (int j = 0; j < t.nR0; ++j) {
                const double mu0 = (j < n) ? (t.mul / (n - 1)) * j : t.mul + ((1.0 - t.mul) / n) * (j - n + 1);
                const double R0 = r0 * std::sqrt(1.0 - mu0 * mu0) / mu0;  // j = 0 gives +inf on purpose
                for (int k = 0; k < t.nph; ++k) {
                    const double phi = (k < n) ? std::acos((t.cphil / (n - 1)) * k)
                                               : std::acos(t.cphil + ((1.0 - t.cphil) / n) * (k - n + 1));
                    t.at(i, j, k) = m4_full(1.0, phi, r0, R0);
                }
            }
        }
        for (int j = 0; j < t.nR0; ++j)
            for (int k = 0; k < t.nph; ++k) t.at(t.nr0 + 1, j, k) = t.at(t.nr0, j, k);
        return t;
    }();
    return t;
}

int clamp_idx(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// gridded_integral, src/kernel_integral.cpp:136-190 (the clamp warning of :125-133 is not printed)
double m4_gridded(const M4Table& t, double h, double phi, double r0_in, double R0_in) {
    if (r0_in == 0.0 || R0_in == 0.0 || phi == 0.0) return 0.0;
    double r0 = r0_in / h, R0 = R0_in / h;
    const double cphi = std::cos(phi);
    if (r0 < 0.0) r0 = 0.0;
    if (R0 < 0.0) R0 = 0.0;
    const double mu0 = r0 / std::sqrt(r0 * r0 + R0 * R0);
    if (r0 > 2.0) r0 = 2.0;
    int i, j, k;
    if (r0 < t.rl) i = static_cast<int>(r0 / (t.rl / t.nr01));
    else i = t.nr01 + static_cast<int>((r0 - t.rl) / ((2.0 - t.rl) / t.nr02));
    if (mu0 < t.mul) j = static_cast<int>(mu0 / (t.mul / (t.nR01 - 1)));
    else j = t.nR01 - 1 + static_cast<int>((mu0 - t.mul) / ((1.0 - t.mul) / t.nR02));
    if (cphi < t.cphil) k = static_cast<int>(cphi / (t.cphil / (t.nph1 - 1)));
    else k = t.nph1 - 1 + static_cast<int>((cphi - t.cphil) / ((1.0 - t.cphil) / t.nph2));
    i = clamp_idx(i, 0, t.nr0);
    j = clamp_idx(j, 0, t.nR0 - 1);
    k = clamp_idx(k, 0, t.nph - 1);
    double fr;
    if (r0 < t.rl) fr = (r0 - (t.rl / t.nr01) * i) / (t.rl / t.nr01);
    else fr = (r0 - t.rl - ((2.0 - t.rl) / t.nr02) * (i - t.nr01)) / ((2.0 - t.rl) / t.nr02);
    const double fx1 = fr * t.at(i + 1, j, k) + (1 - fr) * t.at(i, j, k);
    const double fx2 = fr * t.at(i + 1, j + 1, k) + (1 - fr) * t.at(i, j + 1, k);
    const double fx3 = fr * t.at(i + 1, j, k + 1) + (1 - fr) * t.at(i, j, k + 1);
    const double fx4 = fr * t.at(i + 1, j + 1, k + 1) + (1 - fr) * t.at(i, j + 1, k + 1);
    if (mu0 < t.mul) fr = (mu0 - t.mul / (t.nR01 - 1) * j) / (t.mul / (t.nR01 - 1));
    else fr = (mu0 - t.mul - (1.0 - t.mul) / t.nR02 * (j - t.nR01 + 1)) / ((1.0 - t.mul) / t.nR02);
    const double fy1 = fr * fx2 + (1 - fr) * fx1;
    const double fy2 = fr * fx4 + (1 - fr) * fx3;
    if (cphi < t.cphil) fr = (cphi - t.cphil / (t.nph1 - 1) * k) / (t.cphil / (t.nph1 - 1));
    else fr = (cphi - t.cphil - (1.0 - t.cphil) / t.nph2 * (k - t.nph1 + 1)) / ((1.0 - t.cphil) / t.nph2);
    double fz = fr * fy2 + (1 - fr) * fy1;
    if (j == t.nR01 + t.nR02 - 1 || k == t.nph1 + t.nph2 - 1) fz = 0.0;
    return fz;
}

// full_integral_new, src/kernel_integral.cpp:410-418
double m4_full(double h, double phi, double r0, double R0) {
    if (r0 == 0.0 || R0 == 0.0 || phi == 0.0) return 0.0;
    const M4Wall<double> w = m4_wall(h, r0);
    const M4Edge<double> e = m4_edge(w, R0);
    return m4_vertex(w, e, phi);
}

// ------------------------------------------------------------------------------------------------
// Generic piecewise-polynomial kernels: src/kernel_integral_generic.cpp:19-163, driven by the
// coefficient tables of kiwi_amd/csrc/kernel_tables.h (our own exact re-derivation).
// ------------------------------------------------------------------------------------------------

constexpr int NI = KIWI_TAB_NI;
constexpr int NREG = KIWI_TAB_MAXREG;

template <typename T>
struct GenWall {  // include/kiwi/kernel_integral_generic.hpp:27-35
    bool ok = false;
    const KiwiKernelTable* K = nullptr;
    T H = 0, r0 = 0, q0 = 0;
    int j = 0;
    T coef[NREG][NI];
    T bphi[NREG];
};

template <typename T>
struct GenEdge {  // include/kiwi/kernel_integral_generic.hpp:37-44
    bool ok = false;
    T R0 = 0, a = 0, a2 = 0;
    int nseg = 0;
    int region[NREG];
    T phi_end[NREG];
    T offset[NREG];
};

// basis_integrals, src/kernel_integral_generic.cpp:19-66: I_i(phi) at out[1-i] for i = 1 .. imin.
template <typename T>
void gen_basis(T a, T a2, T phi, int imin, T* out) {
    const T sn = q_sin(phi), cs = q_cos(phi);
    const T den = a2 + cs * cs;
    const T u = a * sn / q_sqrt(den);
    const T omu2 = cs * cs * (T(1) + a2) / den;
    out[0] = q_atan(u / a);
    out[1] = phi;
    const int mmax = (-imin - 1) / 2;
    T Iodd = out[0];
    T K = q_atanh(u);
    T p1a2 = 1;
    for (int m = 0; m <= mmax && -2 * m - 1 >= imin; ++m) {
        if (m > 0) {
            K = u / (T(2 * m) * q_pow(omu2, m)) + T(2 * m - 1) / T(2 * m) * K;
            p1a2 *= (T(1) + a2);
        }
        Iodd += a * p1a2 * K;
        out[2 + 2 * m] = Iodd;
    }
    const int kmax = -imin / 2;
    if (kmax >= 1) {
        T J[NI];
        const T t = sn / cs, sec = T(1) / cs;
        J[0] = phi;
        J[2] = t;
        T sp = 1;
        for (int l = 2; l <= kmax; ++l) {
            const int m = 2 * l;
            sp *= sec * sec;
            J[m] = sp * t / T(m - 1) + T(m - 2) / T(m - 1) * J[m - 2];
        }
        for (int k = 1; k <= kmax; ++k) {
            T sum = J[0], ap = 1, b = 1;
            for (int l = 1; l <= k; ++l) {
                b = b * T(k - l + 1) / T(l);
                ap *= a2;
                sum += b * ap * J[2 * l];
            }
            out[1 + 2 * k] = sum;
        }
    }
}

// region_of, src/kernel_integral_generic.cpp:68-72 (1-based)
template <typename T>
int gen_region(const KiwiKernelTable& K, T q) {
    for (int k = 0; k < K.nreg - 1; ++k)
        if (q <= T(K.c[k])) return k + 1;
    return K.nreg;
}

// imin_of, src/kernel_integral_generic.cpp:75-79
template <typename T>
int gen_imin(const GenWall<T>& w, int k) {
    for (int idx = NI - 1; idx >= 0; --idx)
        if (w.coef[k - 1][idx] != T(0)) return 1 - idx;
    return 1;
}

// G, src/kernel_integral_generic.cpp:82-92
template <typename T>
T gen_G(const GenWall<T>& w, const GenEdge<T>& e, int k, T phi) {
    T I[NI];
    gen_basis(e.a, e.a2, phi, gen_imin(w, k), I);
    T sum = w.bphi[k - 1] * phi;
    for (int idx = 0; idx < NI; ++idx) {
        const T c = w.coef[k - 1][idx];
        if (c != T(0)) sum += c * I[idx];
    }
    return sum;
}

// generic_wall_constants, src/kernel_integral_generic.cpp:96-117
template <typename T>
GenWall<T> gen_wall(const KiwiKernelTable& K, T H, T r0) {
    GenWall<T> w;
    w.K = &K;
    w.H = H;
    w.r0 = r0;
    if (r0 == T(0)) return w;
    w.ok = true;
    w.q0 = r0 / H;
    w.j = gen_region(K, w.q0);
    const T A3 = T(K.A3_pi) / q_pi<T>();
    for (int k = 1; k <= K.nreg; ++k) {
        for (int idx = 0; idx < NI; ++idx) {
            const double av = K.a[k - 1][idx];
            w.coef[k - 1][idx] = (av == 0.0) ? T(0) : A3 * T(av) * q_pow(w.q0, idx);
        }
        T b = 0;
        if (k >= w.j)
            for (int p = KIWI_TAB_NB - 1; p >= 0; --p) b = b * w.q0 + T(K.B[w.j - 1][k - 1][p]);
        w.bphi[k - 1] = A3 * b;
    }
    return w;
}

// generic_edge_constants, src/kernel_integral_generic.cpp:119-148
template <typename T>
GenEdge<T> gen_edge(const GenWall<T>& w, T R0) {
    GenEdge<T> e;
    e.R0 = R0;
    if (!w.ok || R0 == T(0)) return e;
    e.ok = true;
    e.a = R0 / w.r0;
    e.a2 = e.a * e.a;
    const KiwiKernelTable& K = *w.K;
    const T qR = R0 / w.H;
    const T rs = q_sqrt(w.q0 * w.q0 + qR * qR);
    int k = gen_region(K, rs);
    e.nseg = 0;
    for (; k <= K.nreg; ++k) {
        e.region[e.nseg] = k;
        if (k == K.nreg) {
            e.phi_end[e.nseg] = q_pi<T>() / T(2);
        } else {
            const T c = T(K.c[k - 1]);
            e.phi_end[e.nseg] = q_acos(q_fmin(T(1), qR / q_sqrt(c * c - w.q0 * w.q0)));
        }
        ++e.nseg;
    }
    e.offset[0] = 0;
    for (int s = 0; s + 1 < e.nseg; ++s) {
        const T pb = e.phi_end[s];
        e.offset[s + 1] = e.offset[s] + gen_G(w, e, e.region[s], pb) - gen_G(w, e, e.region[s + 1], pb);
    }
    return e;
}

// generic_vertex_integral, src/kernel_integral_generic.cpp:150-156
template <typename T>
T gen_vertex(const GenWall<T>& w, const GenEdge<T>& e, T phi) {
    if (!w.ok || !e.ok) return T(0);
    if (phi == T(0)) return T(0);
    int s = 0;
    while (s + 1 < e.nseg && phi > e.phi_end[s]) ++s;
    return e.offset[s] + gen_G(w, e, e.region[s], phi);
}

// ------------------------------------------------------------------------------------------------
// Face walk: face_integral_impl, src/kernel_integral.cpp:471-647, over an evaluator that supplies
// the wall / edge / vertex tiers (src/kernel_integral.cpp:433-467).
// ------------------------------------------------------------------------------------------------

template <typename T>
struct EvalM4 {
    using scalar = T;
    T h;
    const M4Table* table;  // non-null: tabulated path (double only)
    M4Wall<T> w;
    M4Edge<T> e;
    T ar0 = 0, R0 = 0;
    void wall(T r) {
        ar0 = r;
        if (!table) w = m4_wall(h, r);
    }
    void edge(T R) {
        R0 = R;
        if (!table) e = m4_edge(w, R);
    }
    T at(T phi) const;
};
template <>
double EvalM4<double>::at(double phi) const {
    return table ? m4_gridded(*table, h, phi, ar0, R0) : m4_vertex(w, e, phi);
}
template <>
__float128 EvalM4<__float128>::at(__float128 phi) const {
    return m4_vertex(w, e, phi);
}

template <typename T>
struct EvalGen {
    using scalar = T;
    const KiwiKernelTable* K;
    T H;
    GenWall<T> w;
    GenEdge<T> e;
    void wall(T r) { w = gen_wall(*K, H, r); }
    void edge(T R) { e = gen_edge(w, R); }
    T at(T phi) const { return gen_vertex(w, e, phi); }
};

template <class Ev>
typename Ev::scalar face_walk(typename Ev::scalar x0, typename Ev::scalar y0, typename Ev::scalar z0,
                              typename Ev::scalar support, const double* vd, int n, int* out_vert, Ev& ev) {
    using T = typename Ev::scalar;
    if (n < 3) {
        if (out_vert) *out_vert = 0;
        return T(0);
    }
    auto v = [&](int i) { return T(vd[i]); };
    const T x1 = v(0), y1 = v(1), z1 = v(2);
    const T x2s = v(3), y2s = v(4), z2s = v(5);
    const T x3s = v(6), y3s = v(7), z3s = v(8);
    // plane through the first three vertices (:495-506)
    const T A = (y2s - y1) * (z3s - z1) - (y3s - y1) * (z2s - z1);
    const T B = (z2s - z1) * (x3s - x1) - (z3s - z1) * (x2s - x1);
    const T C = (x2s - x1) * (y3s - y1) - (x3s - x1) * (y2s - y1);
    const T D = -A * x1 - B * y1 - C * z1;
    const T nrm = q_sqrt(A * A + B * B + C * C);
    const T r0 = (A * x0 + B * y0 + C * z0 + D) / nrm;
    const T ar0 = q_fabs(r0);
    const T xp = x0 - r0 * A / nrm, yp = y0 - r0 * B / nrm, zp = z0 - r0 * C / nrm;
    // winding sign reference, triple product about the particle (:525-529)
    const T ax = x1 - x0, ay = y1 - y0, az = z1 - z0;
    const T bx = x2s - x0, by = y2s - y0, bz = z2s - z0;
    const T cx = x3s - x0, cy = y3s - y0, cz = z3s - z0;
    const T s2 = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    ev.wall(ar0);

    bool inside = true, have_sign = false;
    T sign_ref = 0;
    T min_edge = T(std::numeric_limits<double>::infinity());
    T total = 0;
    for (int j = 0; j < n; ++j) {
        const T x2 = v(3 * j), y2 = v(3 * j + 1), z2 = v(3 * j + 2);
        const int jn = (j == n - 1) ? 0 : j + 1;
        const T x3 = v(3 * jn), y3 = v(3 * jn + 1), z3 = v(3 * jn + 2);
        // edge geometry (:573-583)
        const T r23 = q_sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2) + (z3 - z2) * (z3 - z2));
        const T r12 = q_sqrt((xp - x2) * (xp - x2) + (yp - y2) * (yp - y2) + (zp - z2) * (zp - z2));
        const T r13 = q_sqrt((xp - x3) * (xp - x3) + (yp - y3) * (yp - y3) + (zp - z3) * (zp - z3));
        const T cosa = ((x3 - x2) * (xp - x2) + (y3 - y2) * (yp - y2) + (z3 - z2) * (zp - z2)) / r12 / r23;
        const T R0 = (q_fabs(cosa) < T(1)) ? r12 * q_sqrt(1 - cosa * cosa) : T(0);
        // s1 about the particle (:588-591)
        const T px = xp - x0, py = yp - y0, pz = zp - z0;
        const T ux = x2 - x0, uy = y2 - y0, uz = z2 - z0;
        const T wx = x3 - x0, wy = y3 - y0, wz = z3 - z0;
        const T s1 = px * (uy * wz - uz * wy) + py * (uz * wx - ux * wz) + pz * (ux * wy - uy * wx);
        const T phi1 = (R0 < r12) ? q_acos(R0 / r12) : T(0);
        const T phi2 = (R0 < r13) ? q_acos(R0 / r13) : T(0);
        // inside-polygon side test (:600-611)
        const T ex = x3 - x2, ey = y3 - y2, ez = z3 - z2;
        const T tx = xp - x2, ty = yp - y2, tz = zp - z2;
        const T side = (ey * tz - ez * ty) * A + (ez * tx - ex * tz) * B + (ex * ty - ey * tx) * C;
        const T side_eps = T(1e-9) * nrm * r23 * std::max({q_fabs(tx), q_fabs(ty), q_fabs(tz), T(1)});
        if (q_fabs(side) > side_eps) {
            if (!have_sign) {
                sign_ref = side;
                have_sign = true;
            } else if ((side > 0) != (sign_ref > 0)) {
                inside = false;
            }
        }
        const bool beyond = (r12 * q_sin(phi1) >= r23) || (r13 * q_sin(phi2) >= r23);
        min_edge = std::min(min_edge, beyond ? std::min(r12, r13) : R0);
        T M = (s1 * s2 * r0 <= 0) ? T(-1) : T(1);
        ev.edge(R0);
        if (beyond) {
            M = (phi1 >= phi2) ? M * (ev.at(phi1) - ev.at(phi2)) : M * (ev.at(phi2) - ev.at(phi1));
        } else {
            M = M * (ev.at(phi1) + ev.at(phi2));
        }
        total += M;
    }
    // reachability of the face (:642-645)
    const T fd = inside ? ar0 : q_sqrt(r0 * r0 + min_edge * min_edge);
    if (out_vert) *out_vert = (fd >= support) ? n : 0;
    return total;
}

// ------------------------------------------------------------------------------------------------
// Operation count of the cubic path (measurement only): the same templates over a scalar that
// tallies what it computes.  Convention: every +, -, *, /, sqrt and transcendental call counts one
// operation (sincos counts two); comparisons, fabs and selects count none; constants such as 3/40
// are literals in the reference and cost nothing.
// ------------------------------------------------------------------------------------------------
struct OpTally {
    long long add = 0, mul = 0, div = 0, sqrt = 0, trans = 0;
};
thread_local OpTally g_tally;

struct OpCount {
    double v;
    OpCount() : v(0) {}
    OpCount(double x) : v(x) {}
};
inline OpCount operator+(OpCount a, OpCount b) { ++g_tally.add; return OpCount(a.v + b.v); }
inline OpCount operator-(OpCount a, OpCount b) { ++g_tally.add; return OpCount(a.v - b.v); }
inline OpCount operator*(OpCount a, OpCount b) { ++g_tally.mul; return OpCount(a.v * b.v); }
inline OpCount operator/(OpCount a, OpCount b) { ++g_tally.div; return OpCount(a.v / b.v); }
inline OpCount operator-(OpCount a) { return OpCount(-a.v); }
inline OpCount& operator+=(OpCount& a, OpCount b) { return a = a + b; }
inline bool operator<(OpCount a, OpCount b) { return a.v < b.v; }
inline bool operator<=(OpCount a, OpCount b) { return a.v <= b.v; }
inline bool operator>(OpCount a, OpCount b) { return a.v > b.v; }
inline bool operator>=(OpCount a, OpCount b) { return a.v >= b.v; }
inline bool operator==(OpCount a, OpCount b) { return a.v == b.v; }
inline bool operator!=(OpCount a, OpCount b) { return a.v != b.v; }
inline OpCount q_sqrt(OpCount x) { ++g_tally.sqrt; return OpCount(std::sqrt(x.v)); }
inline OpCount q_acos(OpCount x) { ++g_tally.trans; return OpCount(std::acos(x.v)); }
inline OpCount q_log(OpCount x) { ++g_tally.trans; return OpCount(std::log(x.v)); }
inline OpCount q_atan(OpCount x) { ++g_tally.trans; return OpCount(std::atan(x.v)); }
inline OpCount q_sin(OpCount x) { ++g_tally.trans; return OpCount(std::sin(x.v)); }
inline OpCount q_fabs(OpCount x) { return OpCount(std::fabs(x.v)); }
template <> inline OpCount q_pi<OpCount>() { return OpCount(M_PI); }
template <> inline OpCount C<OpCount>(int n, int d) { return OpCount(static_cast<double>(n) / d); }
inline void sincos_pair(OpCount x, OpCount& s, OpCount& c) {
    g_tally.trans += 2;
    double a, b;
    sincos_pair(x.v, a, b);
    s = OpCount(a);
    c = OpCount(b);
}
template <>
OpCount EvalM4<OpCount>::at(OpCount phi) const {
    return m4_vertex(w, e, phi);
}

// box_cell_faces, src/cartesian_grid.cpp:54-72: the 6 faces of a box, in the winding the face walk needs.
void box_faces(double cx, double cy, double cz, double hx, double hy, double hz, double f[6][12]) {
    const double xm = cx - hx, xp = cx + hx, ym = cy - hy, yp = cy + hy, zm = cz - hz, zp = cz + hz;
    const double F[6][12] = {{xp, ym, zp, xp, yp, zp, xp, yp, zm, xp, ym, zm},
                             {xm, yp, zm, xm, yp, zp, xm, ym, zp, xm, ym, zm},
                             {xp, yp, zm, xp, yp, zp, xm, yp, zp, xm, yp, zm},
                             {xm, ym, zp, xp, ym, zp, xp, ym, zm, xm, ym, zm},
                             {xm, yp, zp, xp, yp, zp, xp, ym, zp, xm, ym, zp},
                             {xp, ym, zm, xp, yp, zm, xm, yp, zm, xm, ym, zm}};
    std::memcpy(f, F, sizeof(F));
}

// KernelIntegrator::face (include/kiwi/kernel_integrator.hpp:126-130)
struct Integrator {
    const KiwiKernelTable* gen = nullptr;  // null: cubic fast path
    double support_in_h = 2.0;
    bool tabulated = false;
    double face(double x0, double y0, double z0, double h, const double* v, int n, int* out_vert) const {
        if (gen) {
            EvalGen<double> ev{gen, support_in_h * h, {}, {}};
            return face_walk(x0, y0, z0, support_in_h * h, v, n, out_vert, ev);
        }
        EvalM4<double> ev{h, tabulated ? &m4_table() : nullptr, {}, {}};
        return face_walk(x0, y0, z0, 2 * h, v, n, out_vert, ev);
    }
};

}  // namespace

// ================================================================================================
// C ABI (ctypes)
// ================================================================================================

extern "C" {

struct OracleKernel {  // mirrors kiwi_kernel in include/kiwi_hip.h
    int kernel_id;        // KIWI_KERNEL_M4 .. KIWI_KERNEL_WC6
    double support_in_h;  // 0: the kernel's default radkern
    int generic_path;     // force the generic path
    int tabulated;        // cubic tabulated path
};

struct OracleStats {  // mirrors kiwi_stats
    long long skipped_outside, n_negative, n_negative_warn, pairs_bbox, pairs_nonzero;
    double v_mean[3];
};

int oracle_abi_version(void) { return 1; }

double oracle_m4_vertex(double h, double phi, double r0, double R0) { return m4_full(h, phi, r0, R0); }

double oracle_m4_gridded(double h, double phi, double r0, double R0) {
    return m4_gridded(m4_table(), h, phi, r0, R0);
}

double oracle_generic_vertex(int kernel_id, double H, double phi, double r0, double R0) {
    if (kernel_id < 0 || kernel_id >= KIWI_TAB_NKERNELS) return std::nan("");
    if (r0 == 0.0 || R0 == 0.0 || phi == 0.0) return 0.0;
    const GenWall<double> w = gen_wall(kKiwiKernelTables[kernel_id], H, r0);
    const GenEdge<double> e = gen_edge(w, R0);
    return gen_vertex(w, e, phi);
}

double oracle_face(const OracleKernel* k, double x0, double y0, double z0, double h, const double* verts,
                   int nverts, int* out_vert) {
    Integrator ki;
    const KiwiKernelTable& K = kKiwiKernelTables[k->kernel_id];
    ki.support_in_h = k->support_in_h > 0 ? k->support_in_h : K.radkern;
    // Config::use_generic_kernel_path, src/config.cpp:275-277
    if (k->generic_path || k->kernel_id != 0 || ki.support_in_h != 2.0) ki.gen = &K;
    else ki.tabulated = k->tabulated != 0;
    return ki.face(x0, y0, z0, h, verts, nverts, out_vert);
}

// oracle_face with the kernel mathematics in binary128 (the vertices as given); the analytic cubic or
// generic path only
double oracle_face_q(const OracleKernel* k, double x0, double y0, double z0, double h, const double* verts,
                     int nverts) {
    if (!k || k->kernel_id < 0 || k->kernel_id >= KIWI_TAB_NKERNELS || k->tabulated) return std::nan("");
    const KiwiKernelTable& K = kKiwiKernelTables[k->kernel_id];
    const double sup = k->support_in_h > 0 ? k->support_in_h : K.radkern;
    const __float128 hh = h;
    if (k->generic_path || k->kernel_id != 0 || sup != 2.0) {
        const __float128 H = static_cast<__float128>(sup) * hh;
        EvalGen<__float128> ev{&K, H, {}, {}};
        return static_cast<double>(face_walk<EvalGen<__float128>>(x0, y0, z0, H, verts, nverts, nullptr, ev));
    }
    EvalM4<__float128> ev{hh, nullptr, {}, {}};
    return static_cast<double>(face_walk<EvalM4<__float128>>(x0, y0, z0, 2 * hh, verts, nverts, nullptr, ev));
}

// The sph-to-grid particle loop (src/cartesian_grid.cpp:184-301): outputs are the UN-normalised sums
// mass = sum Msum*m and p = sum Msum*m*v per cell, x-major idx = (ix*ny+iy)*nz+iz.  With threads > 1
// the per-thread accumulators are reduced in thread order, as the reference does; threads == 1
// reproduces the reference's OMP_NUM_THREADS=1 summation order exactly.
int oracle_sph_to_grid(long long n, const double* x, const double* y, const double* z, const double* vx,
                       const double* vy, const double* vz, const double* m, const double* h,
                       const double* domain, int nx, int ny, int nz, const OracleKernel* k, int threads,
                       double* mass, double* px, double* py, double* pz, OracleStats* st) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !k || k->kernel_id < 0 || k->kernel_id >= KIWI_TAB_NKERNELS) return -1;
    const KiwiKernelTable& K = kKiwiKernelTables[k->kernel_id];
    Integrator ki;
    ki.support_in_h = k->support_in_h > 0 ? k->support_in_h : K.radkern;
    if (k->generic_path || k->kernel_id != 0 || ki.support_in_h != 2.0) ki.gen = &K;
    else ki.tabulated = k->tabulated != 0;
    if (ki.tabulated) (void)m4_table();

    const double xmin = domain[0], ymin = domain[2], zmin = domain[4];
    const double dpx = (domain[1] - domain[0]) / nx, dpy = (domain[3] - domain[2]) / ny,
                 dpz = (domain[5] - domain[4]) / nz;
    const double hpx = 0.5 * dpx, hpy = 0.5 * dpy, hpz = 0.5 * dpz;
    const size_t nc = static_cast<size_t>(nx) * ny * nz;
    int nt = threads > 0 ? threads : 1;
#ifndef _OPENMP
    nt = 1;
#endif
    std::vector<std::vector<double>> acc(static_cast<size_t>(nt), std::vector<double>(4 * nc, 0.0));
    double vmx = 0, vmy = 0, vmz = 0;
    long long skipped = 0, nneg = 0, nwarn = 0, pbox = 0, pnz = 0;

#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 256) \
    reduction(+ : vmx, vmy, vmz, skipped, nneg, nwarn, pbox, pnz)
#endif
    for (long long i = 0; i < n; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* A = acc[static_cast<size_t>(tid)].data();
        const double x0 = x[i], y0 = y[i], z0 = z[i], hh = h[i], mm = m[i];
        vmx += vx[i];
        vmy += vy[i];
        vmz += vz[i];
        if (hh <= 0.0) continue;
        const double rad = ki.support_in_h * hh;
        int ix0 = static_cast<int>(std::floor((x0 - rad - xmin) / dpx));
        int ix1 = static_cast<int>(std::floor((x0 + rad - xmin) / dpx));
        int iy0 = static_cast<int>(std::floor((y0 - rad - ymin) / dpy));
        int iy1 = static_cast<int>(std::floor((y0 + rad - ymin) / dpy));
        int iz0 = static_cast<int>(std::floor((z0 - rad - zmin) / dpz));
        int iz1 = static_cast<int>(std::floor((z0 + rad - zmin) / dpz));
        ix0 = std::max(ix0, 0);
        ix1 = std::min(ix1, nx - 1);
        iy0 = std::max(iy0, 0);
        iy1 = std::min(iy1, ny - 1);
        iz0 = std::max(iz0, 0);
        iz1 = std::min(iz1, nz - 1);
        if (ix0 > ix1 || iy0 > iy1 || iz0 > iz1) {
            ++skipped;
            continue;
        }
        pbox += static_cast<long long>(ix1 - ix0 + 1) * (iy1 - iy0 + 1) * (iz1 - iz0 + 1);
        for (int ix = ix0; ix <= ix1; ++ix) {
            const double cx = xmin + (ix + 0.5) * dpx;
            for (int iy = iy0; iy <= iy1; ++iy) {
                const double cy = ymin + (iy + 0.5) * dpy;
                for (int iz = iz0; iz <= iz1; ++iz) {
                    const double cz = zmin + (iz + 0.5) * dpz;
                    double f[6][12];
                    box_faces(cx, cy, cz, hpx, hpy, hpz, f);
                    double Ms = 0.0;
                    for (int q = 0; q < 6; ++q) Ms += ki.face(x0, y0, z0, hh, f[q], 4, nullptr);
                    if (Ms < 0.0) {
                        ++nneg;
                        if (Ms <= -0.01) ++nwarn;
                        Ms = 0.0;
                    }
                    if (Ms == 0.0) continue;
                    ++pnz;
                    const size_t c = (static_cast<size_t>(ix) * ny + iy) * nz + iz;
                    A[c] += Ms * mm;
                    A[nc + c] += Ms * mm * vx[i];
                    A[2 * nc + c] += Ms * mm * vy[i];
                    A[3 * nc + c] += Ms * mm * vz[i];
                }
            }
        }
    }
    for (size_t c = 0; c < nc; ++c) {
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int t = 0; t < nt; ++t) {
            const double* A = acc[static_cast<size_t>(t)].data();
            s0 += A[c];
            s1 += A[nc + c];
            s2 += A[2 * nc + c];
            s3 += A[3 * nc + c];
        }
        mass[c] = s0;
        px[c] = s1;
        py[c] = s2;
        pz[c] = s3;
    }
    if (st) {
        st->skipped_outside = skipped;
        st->n_negative = nneg;
        st->n_negative_warn = nwarn;
        st->pairs_bbox = pbox;
        st->pairs_nonzero = pnz;
        st->v_mean[0] = n > 0 ? vmx / static_cast<double>(n) : 0.0;
        st->v_mean[1] = n > 0 ? vmy / static_cast<double>(n) : 0.0;
        st->v_mean[2] = n > 0 ? vmz / static_cast<double>(n) : 0.0;
    }
    return 0;
}

// The reference's cubic sph-to-grid sums for a list of cells, with the kernel mathematics of
// src/kernel_integral.cpp evaluated in binary128 (the geometry, i.e. the bounding boxes and the cell
// face coordinates, is formed in double exactly as src/cartesian_grid.cpp:220-247 does).  Used to
// arbitrate cells where two double-precision evaluations of the same integral disagree.
int oracle_sph_to_grid_cells_q(long long n, const double* x, const double* y, const double* z, const double* vx,
                               const double* vy, const double* vz, const double* m, const double* h,
                               const double* domain, int nx, int ny, int nz, const OracleKernel* k,
                               long long ncells, const long long* cells, double* out4) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !k || k->kernel_id < 0 || k->kernel_id >= KIWI_TAB_NKERNELS) return -1;
    if (k->tabulated) return -1;  // the tabulated path is double precision by construction
    const KiwiKernelTable& K = kKiwiKernelTables[k->kernel_id];
    const double sup = k->support_in_h > 0 ? k->support_in_h : K.radkern;
    const bool generic = k->generic_path || k->kernel_id != 0 || sup != 2.0;
    const double xmin = domain[0], ymin = domain[2], zmin = domain[4];
    const double dpx = (domain[1] - domain[0]) / nx, dpy = (domain[3] - domain[2]) / ny,
                 dpz = (domain[5] - domain[4]) / nz;
    const double hpx = 0.5 * dpx, hpy = 0.5 * dpy, hpz = 0.5 * dpz;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (long long q = 0; q < ncells; ++q) {
        const long long c = cells[q];
        const int ix = static_cast<int>(c / (static_cast<long long>(ny) * nz));
        const int iy = static_cast<int>((c / nz) % ny);
        const int iz = static_cast<int>(c % nz);
        const double cx = xmin + (ix + 0.5) * dpx, cy = ymin + (iy + 0.5) * dpy, cz = zmin + (iz + 0.5) * dpz;
        double f[6][12];
        box_faces(cx, cy, cz, hpx, hpy, hpz, f);
        __float128 acc[4] = {0, 0, 0, 0};
        for (long long i = 0; i < n; ++i) {
            const double hh = h[i];
            if (hh <= 0.0) continue;
            const double rad = sup * hh;
            const int ix0 = std::max(static_cast<int>(std::floor((x[i] - rad - xmin) / dpx)), 0);
            const int ix1 = std::min(static_cast<int>(std::floor((x[i] + rad - xmin) / dpx)), nx - 1);
            const int iy0 = std::max(static_cast<int>(std::floor((y[i] - rad - ymin) / dpy)), 0);
            const int iy1 = std::min(static_cast<int>(std::floor((y[i] + rad - ymin) / dpy)), ny - 1);
            const int iz0 = std::max(static_cast<int>(std::floor((z[i] - rad - zmin) / dpz)), 0);
            const int iz1 = std::min(static_cast<int>(std::floor((z[i] + rad - zmin) / dpz)), nz - 1);
            if (ix < ix0 || ix > ix1 || iy < iy0 || iy > iy1 || iz < iz0 || iz > iz1) continue;
            __float128 Ms = 0;
            for (int q = 0; q < 6; ++q) {
                const __float128 H = static_cast<__float128>(sup) * hh;
                if (generic) {
                    EvalGen<__float128> ev{&K, H, {}, {}};
                    Ms += face_walk<EvalGen<__float128>>(x[i], y[i], z[i], H, f[q], 4, nullptr, ev);
                } else {
                    EvalM4<__float128> ev{hh, nullptr, {}, {}};
                    Ms += face_walk<EvalM4<__float128>>(x[i], y[i], z[i], H, f[q], 4, nullptr, ev);
                }
            }
            if (Ms < 0) Ms = 0;
            const __float128 t = Ms * m[i];
            acc[0] += t;
            acc[1] += t * vx[i];
            acc[2] += t * vy[i];
            acc[3] += t * vz[i];
        }
        for (int k = 0; k < 4; ++k) out4[4 * q + k] = static_cast<double>(acc[k]);
    }
    return 0;
}

// Voronoi mapping BFS (src/mapping.cpp:105-206, 400-460) over flattened cell geometry:
//   cell_face_ptr[c] .. cell_face_ptr[c+1]   faces of cell c (in voro++ face order)
//   face_vert_ptr[f] .. face_vert_ptr[f+1]   absolute vertex coordinates (xyz triples) of face f
//   face_nbr[f]                              neighbouring cell index, or -1 (domain wall / unresolved)
//   home[i]                                  home cell of particle i, or -1 (unresolved -> skipped)
//   pid[i]                                   the particle id used for the BFS visit stamps
// Accumulates the same un-normalised sums as oracle_sph_to_grid, per cell.
}  // extern "C"

namespace {
// The BFS of run_mapping_stage over flattened geometry; visit(i, cell, Ms) is called for every visited
// (particle, cell) pair with the clamped double-precision cell integral, in the reference's order.
struct VoronoiCounts {
    long long unresolved = 0, unmapped = 0, nneg = 0, nwarn = 0, visited = 0, face_evals = 0;
};
template <class Visit>
VoronoiCounts voronoi_bfs(const Integrator& ki, long long n, const double* x, const double* y, const double* z,
                          const double* h, const int* pid, const int* home, long long ncells,
                          const long long* cell_face_ptr, const long long* face_vert_ptr, const double* face_xyz,
                          const int* face_nbr, Visit visit) {
    std::vector<int> stamp(static_cast<size_t>(ncells), 0), queue(static_cast<size_t>(ncells), 0);
    std::unordered_map<int64_t, std::pair<int, double>> cache;
    VoronoiCounts k;
    for (long long i = 0; i < n; ++i) {
        const double x0 = x[i], y0 = y[i], z0 = z[i], hh = h[i];
        const int ind = pid[i];
        if (hh <= 0.0) continue;
        const int hc = home[i];
        if (hc < 0) {
            ++k.unresolved;
            continue;
        }
        int top = 1, bottom = 0;
        queue[0] = hc;
        stamp[static_cast<size_t>(hc)] = ind;
        cache.clear();
        while (bottom < top) {
            const int cell = queue[static_cast<size_t>(bottom)];
            double Ms = 0.0;
            for (long long f = cell_face_ptr[cell]; f < cell_face_ptr[cell + 1]; ++f) {
                const int nb = face_nbr[f];
                int64_t key = 0;
                bool cacheable = false;
                if (nb >= 0) {
                    const int lo = std::min(cell, nb), hi = std::max(cell, nb);
                    key = (static_cast<int64_t>(lo) << 32) | static_cast<uint32_t>(hi);
                    cacheable = true;
                    auto it = cache.find(key);
                    if (it != cache.end()) {
                        Ms += (it->second.first == cell) ? it->second.second : -it->second.second;
                        continue;
                    }
                }
                const long long v0 = face_vert_ptr[f], nv = face_vert_ptr[f + 1] - v0;
                int ov = 0;
                const double fm = ki.face(x0, y0, z0, hh, face_xyz + 3 * v0, static_cast<int>(nv), &ov);
                ++k.face_evals;
                Ms += fm;
                if (cacheable) cache.emplace(key, std::make_pair(cell, fm));
                if (ov != nv && nb >= 0 && stamp[static_cast<size_t>(nb)] != ind) {
                    queue[static_cast<size_t>(top++)] = nb;
                    stamp[static_cast<size_t>(nb)] = ind;
                }
            }
            if (Ms < 0.0) {
                ++k.nneg;
                if (Ms <= -0.01) ++k.nwarn;
                Ms = 0.0;
            }
            visit(i, cell, Ms);
            ++bottom;
        }
        k.visited += top;
        if (top == 1) ++k.unmapped;
    }
    return k;
}

bool make_integrator(const OracleKernel* k, Integrator* ki) {
    if (!k || k->kernel_id < 0 || k->kernel_id >= KIWI_TAB_NKERNELS) return false;
    const KiwiKernelTable& K = kKiwiKernelTables[k->kernel_id];
    ki->support_in_h = k->support_in_h > 0 ? k->support_in_h : K.radkern;
    if (k->generic_path || k->kernel_id != 0 || ki->support_in_h != 2.0) ki->gen = &K;
    else ki->tabulated = k->tabulated != 0;
    return true;
}
}  // namespace

extern "C" {

int oracle_voronoi_map(long long n, const double* x, const double* y, const double* z, const double* vx,
                       const double* vy, const double* vz, const double* m, const double* h, const int* pid,
                       const int* home, long long ncells, const long long* cell_face_ptr,
                       const long long* face_vert_ptr, const double* face_xyz, const int* face_nbr,
                       const OracleKernel* k, double* mass, double* px, double* py, double* pz, long long* counts) {
    Integrator ki;
    if (!make_integrator(k, &ki)) return -1;
    for (long long c = 0; c < ncells; ++c) mass[c] = px[c] = py[c] = pz[c] = 0.0;
    const VoronoiCounts vc =
        voronoi_bfs(ki, n, x, y, z, h, pid, home, ncells, cell_face_ptr, face_vert_ptr, face_xyz, face_nbr,
                    [&](long long i, int cell, double Ms) {
                        const double mm = m[i];
                        mass[cell] += Ms * mm;
                        px[cell] += Ms * mm * vx[i];
                        py[cell] += Ms * mm * vy[i];
                        pz[cell] += Ms * mm * vz[i];
                    });
    const long long unresolved = vc.unresolved, unmapped = vc.unmapped, nneg = vc.nneg, nwarn = vc.nwarn,
                    visited = vc.visited, face_evals = vc.face_evals;
    if (counts) {
        counts[0] = unresolved;
        counts[1] = unmapped;
        counts[2] = nneg;
        counts[3] = nwarn;
        counts[4] = visited;
        counts[5] = face_evals;
    }
    return 0;
}

// The inverse-mapping test (run_inverse_mapping_test_stage, src/mapping.cpp:540-701): the map's BFS per
// particle, with the particle's clamped cell integrals weighted back by the forward map's cell masses
// M_c (density x volume): m_recovered = sum over visited cells with M_c > 0 of (Ms m / M_c) M_c
// (:652-660), coverage = m_recovered / m for m > 0 (:666-667).  Particles with h <= 0 or no home cell
// keep 0 (:626-644).
int oracle_inverse_mapping(long long n, const double* x, const double* y, const double* z, const double* m,
                           const double* h, const int* pid, const int* home, long long ncells,
                           const long long* cell_face_ptr, const long long* face_vert_ptr, const double* face_xyz,
                           const int* face_nbr, const OracleKernel* k, const double* cell_mass,
                           double* m_recovered, double* coverage, long long* counts) {
    Integrator ki;
    if (!make_integrator(k, &ki)) return -1;
    for (long long i = 0; i < n; ++i) m_recovered[i] = coverage[i] = 0.0;
    const VoronoiCounts vc =
        voronoi_bfs(ki, n, x, y, z, h, pid, home, ncells, cell_face_ptr, face_vert_ptr, face_xyz, face_nbr,
                    [&](long long i, int cell, double Ms) {
                        const double mi = cell_mass[cell];
                        if (mi > 0.0) m_recovered[i] += (Ms * m[i] / mi) * mi;
                    });
    for (long long i = 0; i < n; ++i) coverage[i] = (m[i] > 0.0) ? m_recovered[i] / m[i] : 0.0;
    if (counts) {
        counts[0] = vc.unresolved;
        counts[1] = vc.unmapped;
        counts[2] = vc.nneg;
        counts[3] = vc.nwarn;
        counts[4] = vc.visited;
        counts[5] = vc.face_evals;
    }
    return 0;
}

// The Voronoi map's sums for a list of cells with the face integrals in binary128: the double-precision
// BFS above decides which (particle, cell) pairs are visited; each visited pair's cell integral is then
// redone in binary128 over the cell's faces (vertices as given), clamped at 0 and accumulated in
// binary128.  Used to arbitrate cells where two double-precision evaluations disagree.
//
// share_lower = 0: every cell is integrated over its own copy of each face.  share_lower = 1: a face
// whose neighbour has the lower index, is visited by the same particle and holds exactly one face back
// (the cell holding exactly one face towards it) takes the negated integral of the neighbour's copy --
// the rule by which the GPU map integrates a shared face once.  share_lower = 2: the reference's face
// cache (src/mapping.cpp:87-159) -- a face towards a neighbour the particle's BFS processed earlier takes
// the negated integral of the neighbour's first face towards this cell, and a second face towards the
// same neighbour repeats the first one's value.  voro++ computes each cell's vertices on its own, within
// its plane tolerance (which scales with the domain's extent), so the two copies of a face can differ by
// more than rounding; the map's result depends on which copy a cell is closed with.
int oracle_voronoi_map_cells_q(long long n, const double* x, const double* y, const double* z, const double* vx,
                               const double* vy, const double* vz, const double* m, const double* h, const int* pid,
                               const int* home, long long ncells, const long long* cell_face_ptr,
                               const long long* face_vert_ptr, const double* face_xyz, const int* face_nbr,
                               const OracleKernel* k, long long nlist, const long long* cells, int share_lower,
                               double* out4) {
    Integrator ki;
    if (!make_integrator(k, &ki) || ki.tabulated) return -1;
    std::vector<long long> slot(static_cast<size_t>(ncells), -1);
    for (long long q = 0; q < nlist; ++q) {
        if (cells[q] < 0 || cells[q] >= ncells) return -1;
        slot[static_cast<size_t>(cells[q])] = q;
    }
    std::vector<std::vector<long long>> visitors(static_cast<size_t>(nlist));  // particles per listed cell
    // (particle, cell) pairs the BFS visits, with their processing rank within the particle's BFS
    std::unordered_map<unsigned long long, long long> visited;
    long long last_i = -1, rank = 0;
    auto pair_key = [&](long long i, long long c) {
        return static_cast<unsigned long long>(i) * static_cast<unsigned long long>(ncells) +
               static_cast<unsigned long long>(c);
    };
    voronoi_bfs(ki, n, x, y, z, h, pid, home, ncells, cell_face_ptr, face_vert_ptr, face_xyz, face_nbr,
                [&](long long i, int cell, double) {
                    const long long sl = slot[static_cast<size_t>(cell)];
                    if (sl >= 0) visitors[static_cast<size_t>(sl)].push_back(i);
                    if (i != last_i) {
                        last_i = i;
                        rank = 0;
                    }
                    if (share_lower) visited.emplace(pair_key(i, cell), rank++);
                });
    // the face of the lower-index neighbour each listed cell's face is closed with (-1: its own copy)
    auto twin_of = [&](long long c, long long f) -> long long {
        const int nb = face_nbr[f];
        if (share_lower != 1 || nb < 0 || nb >= c) return -1;
        int own = 0, back = 0;
        long long t = -1;
        for (long long g = cell_face_ptr[c]; g < cell_face_ptr[c + 1]; ++g) own += face_nbr[g] == nb;
        for (long long g = cell_face_ptr[nb]; g < cell_face_ptr[nb + 1]; ++g)
            if (face_nbr[g] == c) {
                ++back;
                t = g;
            }
        return (own == 1 && back == 1) ? t : -1;
    };
    std::vector<__float128> acc(static_cast<size_t>(4 * nlist), 0);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (long long sl = 0; sl < nlist; ++sl) {
        const long long c = cells[sl];
        __float128* a = &acc[static_cast<size_t>(4 * sl)];
        for (const long long i : visitors[static_cast<size_t>(sl)]) {  // particle order, as the BFS
            const __float128 hh = h[i];
            __float128 Ms = 0;
            std::vector<std::pair<int, __float128>> seen;  // share_lower 2: neighbours met in this cell
            const auto rc = visited.find(pair_key(i, c));
            for (long long f = cell_face_ptr[c]; f < cell_face_ptr[c + 1]; ++f) {
                long long g = twin_of(c, f);
                if (g >= 0 && !visited.count(pair_key(i, face_nbr[f]))) g = -1;
                const int nb = face_nbr[f];
                if (share_lower == 2 && nb >= 0) {
                    bool repeat = false;
                    for (const auto& s : seen)
                        if (s.first == nb) {
                            Ms += s.second;
                            repeat = true;
                            break;
                        }
                    if (repeat) continue;
                    const auto rn = visited.find(pair_key(i, nb));
                    if (rn != visited.end() && rc != visited.end() && rn->second < rc->second)
                        for (long long q = cell_face_ptr[nb]; q < cell_face_ptr[nb + 1]; ++q)
                            if (face_nbr[q] == c) {
                                g = q;
                                break;
                            }
                }
                const long long fe = g >= 0 ? g : f;
                const long long v0 = face_vert_ptr[fe];
                const int nv = static_cast<int>(face_vert_ptr[fe + 1] - v0);
                __float128 val;
                if (ki.gen) {
                    const __float128 H = static_cast<__float128>(ki.support_in_h) * hh;
                    EvalGen<__float128> ev{ki.gen, H, {}, {}};
                    val = face_walk<EvalGen<__float128>>(x[i], y[i], z[i], H, face_xyz + 3 * v0, nv, nullptr, ev);
                } else {
                    EvalM4<__float128> ev{hh, nullptr, {}, {}};
                    val = face_walk<EvalM4<__float128>>(x[i], y[i], z[i], 2 * hh, face_xyz + 3 * v0, nv, nullptr, ev);
                }
                Ms += g >= 0 ? -val : val;
                if (share_lower == 2 && nb >= 0) seen.emplace_back(nb, g >= 0 ? -val : val);
            }
            if (Ms < 0) Ms = 0;
            const __float128 t = Ms * m[i];
            a[0] += t;
            a[1] += t * vx[i];
            a[2] += t * vy[i];
            a[3] += t * vz[i];
        }
    }
    for (long long q = 0; q < 4 * nlist; ++q) out4[q] = static_cast<double>(acc[static_cast<size_t>(q)]);
    return 0;
}

// One cubic face integral evaluated over the counting scalar: its value (which must equal
// oracle_face's, so the counted path is the computed path) and its operations (add, mul, div, sqrt,
// transcendental).
double oracle_face_counted(double x0, double y0, double z0, double h, const double* verts, int nverts,
                           long long* counts) {
    g_tally = OpTally();
    EvalM4<OpCount> ev{OpCount(h), nullptr, {}, {}};
    const OpCount v = face_walk(OpCount(x0), OpCount(y0), OpCount(z0), OpCount(2.0 * h), verts, nverts, nullptr, ev);
    counts[0] = g_tally.add;
    counts[1] = g_tally.mul;
    counts[2] = g_tally.div;
    counts[3] = g_tally.sqrt;
    counts[4] = g_tally.trans;
    return v.v;
}

// Operations of the reference's cubic cell loop (src/cartesian_grid.cpp:239-279 with the face
// integrals of src/kernel_integral.cpp:471-647) over every bounding-box pair of the given particles,
// FP64 path.  out: pairs, add, mul, div, sqrt, transcendental.  Per pair it also counts the 6 box-face
// coordinates (6 adds), the cell-centre z coordinate (2 adds, 1 mul) and the four accumulations of
// Msum m and Msum m v (4 adds, 4 muls).
int oracle_count_m4(long long n, const double* x, const double* y, const double* z, const double* h,
                    const double* domain, int nx, int ny, int nz, int threads, long long* out) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !out) return -1;
    const double xmin = domain[0], ymin = domain[2], zmin = domain[4];
    const double dpx = (domain[1] - domain[0]) / nx, dpy = (domain[3] - domain[2]) / ny,
                 dpz = (domain[5] - domain[4]) / nz;
    const double hpx = 0.5 * dpx, hpy = 0.5 * dpy, hpz = 0.5 * dpz;
    long long pairs = 0, add = 0, mul = 0, dv = 0, sq = 0, tr = 0;
    int nt = threads > 0 ? threads : 1;
#ifndef _OPENMP
    nt = 1;
#endif
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1) reduction(+ : pairs, add, mul, dv, sq, tr)
#endif
    for (long long i = 0; i < n; ++i) {
        const double x0 = x[i], y0 = y[i], z0 = z[i], hh = h[i];
        if (hh <= 0.0) continue;
        const double rad = 2.0 * hh;
        const int ix0 = std::max(static_cast<int>(std::floor((x0 - rad - xmin) / dpx)), 0);
        const int ix1 = std::min(static_cast<int>(std::floor((x0 + rad - xmin) / dpx)), nx - 1);
        const int iy0 = std::max(static_cast<int>(std::floor((y0 - rad - ymin) / dpy)), 0);
        const int iy1 = std::min(static_cast<int>(std::floor((y0 + rad - ymin) / dpy)), ny - 1);
        const int iz0 = std::max(static_cast<int>(std::floor((z0 - rad - zmin) / dpz)), 0);
        const int iz1 = std::min(static_cast<int>(std::floor((z0 + rad - zmin) / dpz)), nz - 1);
        if (ix0 > ix1 || iy0 > iy1 || iz0 > iz1) continue;
        g_tally = OpTally();
        long long np = 0;
        for (int ix = ix0; ix <= ix1; ++ix)
            for (int iy = iy0; iy <= iy1; ++iy)
                for (int iz = iz0; iz <= iz1; ++iz) {
                    double f[6][12];
                    box_faces(xmin + (ix + 0.5) * dpx, ymin + (iy + 0.5) * dpy, zmin + (iz + 0.5) * dpz, hpx, hpy,
                              hpz, f);
                    OpCount Ms = 0.0;
                    for (int q = 0; q < 6; ++q) {
                        EvalM4<OpCount> ev{OpCount(hh), nullptr, {}, {}};
                        Ms += face_walk(OpCount(x0), OpCount(y0), OpCount(z0), OpCount(2.0 * hh), f[q], 4, nullptr, ev);
                    }
                    ++np;
                }
        pairs += np;
        add += g_tally.add + np * (6 + 2 + 4);
        mul += g_tally.mul + np * (1 + 4);
        dv += g_tally.div;
        sq += g_tally.sqrt;
        tr += g_tally.trans;
    }
    out[0] = pairs;
    out[1] = add;
    out[2] = mul;
    out[3] = dv;
    out[4] = sq;
    out[5] = tr;
    return 0;
}

}  // extern "C"
