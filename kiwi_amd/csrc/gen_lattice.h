// Generic piecewise-polynomial kernels (M4, M5, M6, Wendland C2/C4/C6) on the grid lattice.
//
// The reference's generic vertex integral (src/kernel_integral_generic.cpp:82-156) is, for a vertex
// pyramid with wall distance r0, edge distance R0 and angle phi, in radial region k of the vertex,
//
//     V = offset_k(edge) + G_k(phi),   G_k = bphi_k phi + sum_i coef_k,i I_i(phi; a = R0/r0),
//     coef_k,i = A3 a_k,i q0^(1-i),    q0 = r0 / H,
//
// with the basis integrals I_1 = atan(u/a), I_0 = phi, I_-(2m+1) (an atanh recurrence) and I_-2k
// (a binomial sum over a secant recurrence J).  At a lattice point with absolute offsets
// (da, db, dc) in units of the support H (wall axis a, edge normal b, along-edge c), every quantity
// of the pyramid is rational in (da, db, dc, r) except three shared transcendentals:
//
//     u = dc / r,  1 - u^2 = (da^2 + db^2) / r^2,  a = db / da,  tan(phi) = dc / db,
//     I_1 = T_b = atan(da dc / (db r)),   atanh(u) = L_c / 2 = log((r + dc)^2 / (da^2 + db^2)) / 2,
//
// the same T_b and L_c as the cubic lattice (m4_lattice.h).  Every I_-2k holds phi exactly once, so
// the phi terms of the two pyramids sharing a wall axis add up to (pi/2) Phi_a(k) with
// Phi = bphi + sum over even-order I of coef, a constant of the wall plane.  The lattice value is
//
//     W = sum over the 6 pyramids of Gnon + (pi/2) sum_a Phi_a(k) + sum over the 3 axis pairs E_ab(k),
//
// Gnon being G without its phi terms and E_ab(k) = offset_k(a, b) + offset_k(b, a), the offsets of
// the two edge lines through the pair of lattice coordinates (evaluated at the virtual vertices where
// the vertex radius crosses a breakpoint, src/kernel_integral_generic.cpp:119-148).
//
// Written once for host (g++, tests) and device (hipcc); double precision throughout.  The evaluation
// reads a GenTab, which gen_tab_init fills from a KiwiKernelTable.
#pragma once

#include <cmath>

#include "kernel_tables.h"
#include "m4_lattice.h"

#ifdef __HIPCC__
#define KIWI_GHD __host__ __device__ __forceinline__
#else
#define KIWI_GHD inline
#endif

namespace kiwi_hip {

constexpr int GEN_NI = KIWI_TAB_NI;
constexpr int GEN_NREG = KIWI_TAB_MAXREG;
constexpr double GEN_PI = 3.141592653589793238462643383279502884;
constexpr int GEN_NODD = (GEN_NI - 1) / 2;   // odd orders I_-(2m+1), m = 0..6 (coefficient index 2m + 2)
constexpr int GEN_NEVEN = (GEN_NI - 1) / 2;  // even orders I_-2l, l = 1..7 (coefficient index 2l + 1)

// The coefficients of one radial region in the form the lattice evaluation uses, compiled once per
// kernel from the table (gen_tab_init): the loop bounds that the largest nonzero coefficient fixes,
// A3 folded into every coefficient, and the binomial sums of the even orders as polynomials in q0^2.
// Compiled in double; the single-precision copy is the double one rounded.
template <typename R>
struct GenRegionT {
    int n_odd;                        // odd orders m = 0 .. n_odd - 1
    int n_even;                       // even orders l = 1 .. n_even
    R c1;                             // A3 a_0: the I_1 term
    R odd[GEN_NODD];                  // A3 a_(2m+2)
    R even[GEN_NEVEN + 1][GEN_NEVEN]; // [l][j]: A3 a_(2(l+j)+1) C(l+j, l); row 0 unused
};
using GenRegion = GenRegionT<double>;

// A kernel's table as the device reads it: the KiwiKernelTable fields the evaluation needs plus the
// compiled regions.
struct GenTab {
    int nreg;
    double c[KIWI_TAB_MAXREG];
    double A3_pi;
    double a[KIWI_TAB_MAXREG][KIWI_TAB_NI];
    double B[KIWI_TAB_MAXREG][KIWI_TAB_MAXREG][KIWI_TAB_NB];
    GenRegionT<double> reg[KIWI_TAB_MAXREG];
    GenRegionT<float> regf[KIWI_TAB_MAXREG];
};

// The coefficient type of a value type: the value type itself, and for dual numbers (dual.h) the type of
// their components, so that coefficients stay plain constants.
template <typename R>
struct GenCoef {
    using type = R;
};
template <typename R>
using gen_coef_t = typename GenCoef<R>::type;

template <typename R>
KIWI_GHD const GenRegionT<R>* gen_regions(const GenTab& K);
template <>
KIWI_GHD const GenRegionT<double>* gen_regions<double>(const GenTab& K) {
    return K.reg;
}
template <>
KIWI_GHD const GenRegionT<float>* gen_regions<float>(const GenTab& K) {
    return K.regf;
}

constexpr void gen_tab_init(GenTab& G, const KiwiKernelTable& K) {
    G = GenTab{};
    G.nreg = K.nreg;
    for (int k = 0; k < KIWI_TAB_MAXREG; ++k) G.c[k] = K.c[k];
    G.A3_pi = K.A3_pi;
    for (int k = 0; k < KIWI_TAB_MAXREG; ++k)
        for (int i = 0; i < KIWI_TAB_NI; ++i) G.a[k][i] = K.a[k][i];
    for (int j = 0; j < KIWI_TAB_MAXREG; ++j)
        for (int k = 0; k < KIWI_TAB_MAXREG; ++k)
            for (int p = 0; p < KIWI_TAB_NB; ++p) G.B[j][k][p] = K.B[j][k][p];
    const double A3 = K.A3_pi / GEN_PI;
    for (int k = 0; k < KIWI_TAB_MAXREG; ++k) {
        const double* av = K.a[k];
        int imax = 0;  // largest nonzero coefficient index
        for (int idx = GEN_NI - 1; idx >= 0; --idx)
            if (av[idx] != 0.0) {
                imax = idx;
                break;
            }
        GenRegion& R = G.reg[k];
        R.c1 = A3 * av[0];
        R.n_odd = imax >= 2 ? (imax - 2) / 2 + 1 : 0;
        R.n_even = imax >= 2 ? (imax - 1) / 2 : 0;
        for (int m = 0; m < R.n_odd; ++m) R.odd[m] = A3 * av[2 * m + 2];
        for (int l = 1; l <= R.n_even; ++l) {
            double binom = 1.0;  // C(kk, l)
            for (int kk = l; kk <= R.n_even; ++kk) {
                if (kk > l) binom = binom * kk / (kk - l);
                R.even[l][kk - l] = A3 * av[2 * kk + 1] * binom;
            }
        }
        GenRegionT<float>& F = G.regf[k];
        F.n_odd = R.n_odd;
        F.n_even = R.n_even;
        F.c1 = static_cast<float>(R.c1);
        for (int m = 0; m < GEN_NODD; ++m) F.odd[m] = static_cast<float>(R.odd[m]);
        for (int l = 0; l <= GEN_NEVEN; ++l)
            for (int j = 0; j < GEN_NEVEN; ++j) F.even[l][j] = static_cast<float>(R.even[l][j]);
    }
}

// The six kernels' tables as constants of the code (order M4 M5 M6 WC2 WC4 WC6).
constexpr KiwiKernelTable kGenTableData[KIWI_TAB_NKERNELS] = {KIWI_KERNEL_TABLE_DATA};

constexpr GenTab gen_tab_make(const KiwiKernelTable& K) {
    GenTab G{};
    gen_tab_init(G, K);
    return G;
}

// Kernel KID's GenTab compiled into the code: evaluated through it (every function here inlines), the
// lattice functions fold the kernel's coefficients, loop bounds and zero terms into their instructions
// instead of reading and testing them at run time.
template <int KID>
KIWI_GHD const GenTab& gen_fixed_tab() {
    static constexpr GenTab T = gen_tab_make(kGenTableData[KID]);
    return T;
}

// region of a radius q (units of H), 1-based: region_of, src/kernel_integral_generic.cpp:68-72
template <class Tab>
KIWI_GHD int gen_region_of(const Tab& K, double q) {
    for (int k = 0; k < K.nreg - 1; ++k)
        if (q <= K.c[k]) return k + 1;
    return K.nreg;
}

// Phi_k(q0) for k = 1..nreg: the coefficient of phi in G_k for a wall at q0 (zero below the wall's
// own region j, where no vertex of this wall can lie).  generic_wall_constants,
// src/kernel_integral_generic.cpp:96-117, reduced to the phi terms.
template <class Tab, typename R>
KIWI_GHD void gen_plane_phi(const Tab& K, R q0, R Phi[GEN_NREG]) {
    const int j = gen_region_of(K, static_cast<double>(kval(q0)));
    const R A3 = R(K.A3_pi / GEN_PI);
    for (int k = 1; k <= GEN_NREG; ++k) {
        R v = R(0);
        if (k <= K.nreg && k >= j) {
            R b = R(0);
            for (int p = KIWI_TAB_NB - 1; p >= 0; --p) b = b * q0 + R(K.B[j - 1][k - 1][p]);
            v = A3 * b;
            // I_0 = phi and every even-order basis integral I_-2k (coefficient index 1 - i = 1, 3, 5, ..)
            // hold phi exactly once
            R qp = q0;
            for (int idx = 1; idx < GEN_NI; idx += 2, qp *= q0 * q0) {
                const R av = R(K.a[k - 1][idx]);
                if (av != R(0)) v += A3 * av * qp;
            }
        }
        Phi[k - 1] = v;
    }
}

// Gnon: G_k without its phi terms, for the pyramid (wall da, edge normal db, along dc) at radius r,
// given T = atan(da dc / (db r)), Kh = atanh(dc / r) and 1/r.  da, db, dc > 0.  The basis integrals
// follow basis_integrals (src/kernel_integral_generic.cpp:19-66) with sin/cos of phi taken from the
// lattice, each multiplied by its power of q0 = da and carried in scaled form, so that no intermediate
// grows or shrinks with a = db/da (single precision overflowed there):
//   odd orders, coefficient index 2m + 2: q0^(2m+2) I_-(2m+1) = da^2 (q0^(2m) I_-(2m-1)) + da db Kt_m,
//     Kt_m = L^m K_m = u r^(2m) / (2m) + (2m-1)/(2m) L Kt_(m-1),  Kt_0 = Kh,  L = da^2 + db^2, u = dc/r
//     (the atanh recurrence K_m = u / (2m (1-u^2)^m) + (2m-1)/(2m) K_(m-1) times L^m);
//   even orders, index 2k + 1: sum_l a^2l q0^(2l+1) J_2l P_l(q0^2) = sum_l da Jt_l P_l(da^2),
//     Jt_l = db^(2l) J_2l = db dc S^(l-1) / (2l-1) + (2l-2)/(2l-1) db^2 Jt_(l-1),  Jt_1 = db dc,
//     S = db^2 + dc^2 (the secant recurrence J_2l = sec^(2l-2) t / (2l-1) + (2l-2)/(2l-1) J_2l-2,
//     t = dc/db, times db^(2l)), P_l the compiled binomial sum over k >= l.
// Every scaled term is a sum of non-negative parts.
template <typename R, typename C>
KIWI_GHD R gen_pyramid_nonphi(const GenRegionT<C>& G, R da, R db, R dc, R r, R T, R Kh, R inv_r) {
    R sum = R(G.c1) * T;  // I_1
    if (G.n_odd == 0) return sum;
    const R da2 = da * da, dab = da * db, L = da2 + db * db, r2 = r * r;
    const R u = dc * inv_r;
    R Kt = Kh, r2m = R(1);
    R It = da2 * T + dab * Kh;  // q0^2 I_-1
    if (G.odd[0] != C(0)) sum += R(G.odd[0]) * It;
#pragma unroll
    for (int m = 1; m < GEN_NODD; ++m) {
        if (m >= G.n_odd) break;
        r2m *= r2;
        Kt = u * R(0.5 / m) * r2m + R((2.0 * m - 1.0) / (2.0 * m)) * L * Kt;
        It = da2 * It + dab * Kt;
        if (G.odd[m] != C(0)) sum += R(G.odd[m]) * It;
    }
    if (G.n_even >= 1) {
        const R db2 = db * db, S = db2 + dc * dc, dbc = db * dc;
        R Jt = dbc, Sp = R(1);  // db^(2l) J_2l, S^(l-1)
#pragma unroll
        for (int l = 1; l <= GEN_NEVEN; ++l) {
            if (l > G.n_even) break;
            if (l > 1) {
                Sp *= S;
                Jt = dbc * Sp * R(1.0 / (2.0 * l - 1.0)) + R((2.0 * l - 2.0) / (2.0 * l - 1.0)) * db2 * Jt;
            }
            R P = R(0);
            for (int j = G.n_even - l; j >= 0; --j) P = P * da2 + R(G.even[l][j]);
            sum += da * Jt * P;
        }
    }
    return sum;
}

// The phi-free part of the lattice value at absolute offsets (ax, ay, az) > 0, radius r (units of H)
// in region k: the 6 ordered pyramids (wall a, edge normal b), which share T_b = atan(da dc / (db r)),
// atanh(d_c / r) and six reciprocals.  A region holding only the I_1 term (beyond the support) gives
// c1 times the sum of the six T, twice T0 + T1 + T2 = pi: a constant, returned as such.
template <typename R, typename C>
KIWI_GHD R gen_lattice_free(const GenRegionT<C>& G, R ax, R ay, R az, R r) {
    using M = KMath<R>;
    if (G.n_odd == 0 && G.n_even == 0) return R(static_cast<double>(G.c1) * GEN_PI);
    const R ix = M::rcp(ax), iy = M::rcp(ay), ir = M::rcp(r);
    const R ixy = M::rcp(ax * ax + ay * ay), ixz = M::rcp(ax * ax + az * az), iyz = M::rcp(ay * ay + az * az);
    // T0 + T1 + T2 = pi/2 for positive offsets (m4_lattice.h): one arctangent fewer per point
    const R T0 = M::atan_pos(ay * az * ix * ir), T1 = M::atan_pos(ax * az * iy * ir);
    const R T2 = M4Consts<R>::half_pi - T0 - T1;
    const R K0 = R(0.5) * M::log_((r + ax) * (r + ax) * iyz);
    const R K1 = R(0.5) * M::log_((r + ay) * (r + ay) * ixz);
    const R K2 = R(0.5) * M::log_((r + az) * (r + az) * ixy);
    R s = R(0);
    s += gen_pyramid_nonphi(G, ax, ay, az, r, T1, K2, ir);  // (x, y, z)
    s += gen_pyramid_nonphi(G, ax, az, ay, r, T2, K1, ir);  // (x, z, y)
    s += gen_pyramid_nonphi(G, ay, ax, az, r, T0, K2, ir);  // (y, x, z)
    s += gen_pyramid_nonphi(G, ay, az, ax, r, T2, K0, ir);  // (y, z, x)
    s += gen_pyramid_nonphi(G, az, ax, ay, r, T0, K1, ir);  // (z, x, y)
    s += gen_pyramid_nonphi(G, az, ay, ax, r, T1, K0, ir);  // (z, y, x)
    return s;
}

// gen_lattice_free for kernel KID (gen_fixed_tab) in region k (1-based).  The outermost region holds only
// the I_1 term (every kernel's table), a constant; the inner regions share one evaluation whose loop
// bounds are the largest of theirs and whose coefficients are selected per region among constants.
template <typename R, int KID>
KIWI_GHD R gen_lattice_free_fixed(R ax, R ay, R az, R r, int k) {
    using C = gen_coef_t<R>;
    constexpr int NR = kGenTableData[KID].nreg;
    const GenRegionT<C>* G = gen_regions<C>(gen_fixed_tab<KID>());
    if (k >= NR) return R(static_cast<double>(G[NR - 1].c1) * GEN_PI);
    if constexpr (NR == 2) {
        return gen_lattice_free<R>(G[0], ax, ay, az, r);
    } else {
        auto pick = [&](auto field) {  // field(G[k - 1]) among the inner regions
            auto v = field(G[NR - 2]);
#pragma unroll
            for (int q = NR - 3; q >= 0; --q) v = (k == q + 1) ? field(G[q]) : v;
            return v;
        };
        GenRegionT<C> S{};
        S.n_odd = 0;
        S.n_even = 0;
#pragma unroll
        for (int q = 0; q < NR - 1; ++q) {
            S.n_odd = G[q].n_odd > S.n_odd ? G[q].n_odd : S.n_odd;
            S.n_even = G[q].n_even > S.n_even ? G[q].n_even : S.n_even;
        }
        S.c1 = pick([](const GenRegionT<C>& g) { return g.c1; });
#pragma unroll
        for (int m = 0; m < GEN_NODD; ++m) S.odd[m] = pick([m](const GenRegionT<C>& g) { return g.odd[m]; });
#pragma unroll
        for (int l = 1; l <= GEN_NEVEN; ++l)
#pragma unroll
            for (int j = 0; j < GEN_NEVEN; ++j)
                S.even[l][j] = pick([l, j](const GenRegionT<C>& g) { return g.even[l][j]; });
        return gen_lattice_free<R>(S, ax, ay, az, r);
    }
}

// G_k - G_(k+1) at the pyramid (da, db, dc), phi = atan(dc / db), Phi the wall's phi coefficients:
// the jump of the vertex integral across the boundary between regions k and k + 1 (edge-line offsets).
template <typename R>
KIWI_GHD R gen_pyramid_G_jump(const GenTab& K, int k, R da, R db, R dc, const R* Phi) {
    using M = KMath<R>;
    const R r = M::sqrt_(da * da + db * db + dc * dc);
    const R inv_r = M::rcp(r);
    const R T = M::atan_pos(da * dc * M::rcp(db) * inv_r);
    const R Kh = R(0.5) * M::log_((r + dc) * (r + dc) * M::rcp(da * da + db * db));
    const GenRegionT<gen_coef_t<R>>* G = gen_regions<gen_coef_t<R>>(K);
    const R g0 = gen_pyramid_nonphi(G[k - 1], da, db, dc, r, T, Kh, inv_r);
    const R g1 = gen_pyramid_nonphi(G[k], da, db, dc, r, T, Kh, inv_r);
    return (g0 - g1) + (Phi[k - 1] - Phi[k]) * M::atan2_pos(dc, db);
}

// Region offsets of the edge line with wall distance da and edge distance db (units of H), per
// region k = 1..nreg: generic_edge_constants (src/kernel_integral_generic.cpp:119-148) with the
// segment boundaries as virtual vertices at radius c_k.  Regions below the line's own foot region
// get 0 (no vertex of the line lies there).
template <typename R>
KIWI_GHD void gen_line(const GenTab& K, R da, R db, R E[GEN_NREG]) {
    for (int k = 0; k < GEN_NREG; ++k) E[k] = R(0);
    if (!(da > R(0)) || !(db > R(0))) return;
    const R rs2 = da * da + db * db;
    const int k0 = gen_region_of(K, static_cast<double>(kval(KMath<R>::sqrt_(rs2))));
    R Phi[GEN_NREG];
    gen_plane_phi(K, da, Phi);
    R off = R(0);
#pragma unroll
    for (int k = 1; k < GEN_NREG; ++k) {  // fixed trip count: E stays in registers
        if (k < k0 || k >= K.nreg) continue;
        // boundary between regions k and k + 1 at radius c_k
        const R ck = R(K.c[k - 1]);
        const R dd = ck * ck - rs2;
        const R dc = KMath<R>::sqrt_(dd > R(0) ? dd : R(0));
        if (dc > R(0)) off += gen_pyramid_G_jump(K, k, da, db, dc, Phi);
        E[k] = off;  // region k + 1
    }
}

// W at the lattice point with absolute offsets (ax, ay, az) > 0 (units of H), from scratch (tests).
template <typename R>
KIWI_GHD R gen_lattice_W(const GenTab& K, R ax, R ay, R az) {
    using M = KMath<R>;
    const R d[3] = {ax, ay, az};
    const R r = M::sqrt_(ax * ax + ay * ay + az * az);
    const int k = gen_region_of(K, r);
    const GenRegionT<R>* G = gen_regions<R>(K);
    R W = R(0);
    for (int a = 0; a < 3; ++a) {
        R Phi[GEN_NREG];
        gen_plane_phi(K, d[a], Phi);
        W += R(0.5 * GEN_PI) * Phi[k - 1];
        for (int b = 0; b < 3; ++b) {
            if (b == a) continue;
            const int c = 3 - a - b;
            const R inv_l2 = M::rcp(d[a] * d[a] + d[b] * d[b]);
            const R T = M::atan_pos(d[a] * d[c] * M::rcp(d[b]) * M::rcp(r));
            const R Kh = R(0.5) * M::log_((r + d[c]) * (r + d[c]) * inv_l2);
            W += gen_pyramid_nonphi(G[k - 1], d[a], d[b], d[c], r, T, Kh, M::rcp(r));
            R E[GEN_NREG];
            gen_line(K, d[a], d[b], E);
            W += E[k - 1];
        }
    }
    return W;
}

}  // namespace kiwi_hip
