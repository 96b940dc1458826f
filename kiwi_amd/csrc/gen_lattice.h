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

#ifdef __HIPCC__
#define KIWI_GHD __host__ __device__ inline
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
struct GenRegion {
    int n_odd;                             // odd orders m = 0 .. n_odd - 1
    int n_even;                            // even orders l = 1 .. n_even
    double c1;                             // A3 a_0: the I_1 term
    double odd[GEN_NODD];                  // A3 a_(2m+2)
    double even[GEN_NEVEN + 1][GEN_NEVEN]; // [l][j]: A3 a_(2(l+j)+1) C(l+j, l); row 0 unused
};

// A kernel's table as the device reads it: the KiwiKernelTable fields the evaluation needs plus the
// compiled regions.
struct GenTab {
    int nreg;
    double c[KIWI_TAB_MAXREG];
    double A3_pi;
    double a[KIWI_TAB_MAXREG][KIWI_TAB_NI];
    double B[KIWI_TAB_MAXREG][KIWI_TAB_MAXREG][KIWI_TAB_NB];
    GenRegion reg[KIWI_TAB_MAXREG];
};

inline void gen_tab_init(GenTab& G, const KiwiKernelTable& K) {
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
        if (imax < 2) continue;
        R.n_odd = (imax - 2) / 2 + 1;
        R.n_even = (imax - 1) / 2;
        for (int m = 0; m < R.n_odd; ++m) R.odd[m] = A3 * av[2 * m + 2];
        for (int l = 1; l <= R.n_even; ++l) {
            double binom = 1.0;  // C(kk, l)
            for (int kk = l; kk <= R.n_even; ++kk) {
                if (kk > l) binom = binom * kk / (kk - l);
                R.even[l][kk - l] = A3 * av[2 * kk + 1] * binom;
            }
        }
    }
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
template <class Tab>
KIWI_GHD void gen_plane_phi(const Tab& K, double q0, double Phi[GEN_NREG]) {
    const int j = gen_region_of(K, q0);
    const double A3 = K.A3_pi / GEN_PI;
    for (int k = 1; k <= GEN_NREG; ++k) {
        double v = 0.0;
        if (k <= K.nreg && k >= j) {
            double b = 0.0;
            for (int p = KIWI_TAB_NB - 1; p >= 0; --p) b = b * q0 + K.B[j - 1][k - 1][p];
            v = A3 * b;
            // I_0 = phi and every even-order basis integral I_-2k (coefficient index 1 - i = 1, 3, 5, ..)
            // hold phi exactly once
            double qp = q0;
            for (int idx = 1; idx < GEN_NI; idx += 2, qp *= q0 * q0) {
                const double av = K.a[k - 1][idx];
                if (av != 0.0) v += A3 * av * qp;
            }
        }
        Phi[k - 1] = v;
    }
}

// Gnon: G_k without its phi terms, for the pyramid (wall da, edge normal db, along dc) at radius r,
// given T = atan(da dc / (db r)), Kh = atanh(dc / r) and the reciprocals 1/da, 1/db, 1/r and
// 1/(da^2 + db^2) (a lattice point shares them between its 6 pyramids, so no division is left here).
// da, db, dc > 0.  The basis integrals follow basis_integrals (src/kernel_integral_generic.cpp:19-66)
// with sin/cos of phi taken from the lattice.
KIWI_GHD double gen_pyramid_nonphi(const GenRegion& R, double da, double db, double dc, double r, double T,
                                   double Kh, double inv_da, double inv_db, double inv_r, double inv_l2) {
    double sum = R.c1 * T;  // I_1
    if (R.n_odd == 0) return sum;
    const double a = db * inv_da, a2 = a * a, da2 = da * da;
    // odd orders I_-(2m+1): the atanh recurrence K_m = u / (2m (1 - u^2)^m) + (2m - 1)/(2m) K_(m-1)
    const double u = dc * inv_r;
    const double inv_omu2 = (r * r) * inv_l2;  // 1 / (1 - u^2)
    double Iodd = T, Km = Kh, p1a2 = 1.0, inv_om = 1.0, qp = da2;  // qp = q0^(2m+2)
#pragma unroll
    for (int m = 0; m < GEN_NODD; ++m) {
        if (m >= R.n_odd) break;
        if (m > 0) {
            inv_om *= inv_omu2;
            Km = u * (0.5 / m) * inv_om + ((2.0 * m - 1.0) / (2.0 * m)) * Km;
            p1a2 *= (1.0 + a2);
            qp *= da2;
        }
        Iodd += a * p1a2 * Km;
        if (R.odd[m] != 0.0) sum += R.odd[m] * qp * Iodd;
    }
    // even orders I_-2k without phi: sum_l a^2l J_2l q0^(2l+1) P_l(q0^2), P_l the compiled binomial sum
    // over k >= l, with the secant recurrence J_2l = sec^(2l-2) t / (2l-1) + (2l-2)/(2l-1) J_2l-2
    if (R.n_even >= 1) {
        const double t = dc * inv_db, sec2 = 1.0 + t * t;
        double Jl = t, sp = 1.0, apl = a2, q2l = da2 * da;  // J_2l, sec^(2l-2), a^2l, q0^(2l+1)
#pragma unroll
        for (int l = 1; l <= GEN_NEVEN; ++l) {
            if (l > R.n_even) break;
            if (l > 1) {
                sp *= sec2;
                Jl = sp * t * (1.0 / (2.0 * l - 1.0)) + ((2.0 * l - 2.0) / (2.0 * l - 1.0)) * Jl;
                apl *= a2;
                q2l *= da2;
            }
            double P = 0.0;
            for (int j = R.n_even - l; j >= 0; --j) P = P * da2 + R.even[l][j];
            sum += (q2l * P) * apl * Jl;
        }
    }
    return sum;
}

// G_k - G_(k+1) at the pyramid (da, db, dc), phi = atan(dc / db), Phi the wall's phi coefficients:
// the jump of the vertex integral across the boundary between regions k and k + 1 (edge-line offsets).
KIWI_GHD double gen_pyramid_G_jump(const GenTab& K, int k, double da, double db, double dc, const double* Phi) {
    const double r = std::sqrt(da * da + db * db + dc * dc);
    const double inv_da = 1.0 / da, inv_db = 1.0 / db, inv_r = 1.0 / r, inv_l2 = 1.0 / (da * da + db * db);
    const double T = std::atan(da * dc * inv_db * inv_r);
    const double Kh = 0.5 * std::log((r + dc) * (r + dc) * inv_l2);
    const double g0 = gen_pyramid_nonphi(K.reg[k - 1], da, db, dc, r, T, Kh, inv_da, inv_db, inv_r, inv_l2);
    const double g1 = gen_pyramid_nonphi(K.reg[k], da, db, dc, r, T, Kh, inv_da, inv_db, inv_r, inv_l2);
    return (g0 - g1) + (Phi[k - 1] - Phi[k]) * std::atan2(dc, db);
}

// Region offsets of the edge line with wall distance da and edge distance db (units of H), per
// region k = 1..nreg: generic_edge_constants (src/kernel_integral_generic.cpp:119-148) with the
// segment boundaries as virtual vertices at radius c_k.  Regions below the line's own foot region
// get 0 (no vertex of the line lies there).
KIWI_GHD void gen_line(const GenTab& K, double da, double db, double E[GEN_NREG]) {
    for (int k = 0; k < GEN_NREG; ++k) E[k] = 0.0;
    if (!(da > 0.0) || !(db > 0.0)) return;
    const double rs2 = da * da + db * db;
    const int k0 = gen_region_of(K, std::sqrt(rs2));
    double Phi[GEN_NREG];
    gen_plane_phi(K, da, Phi);
    double off = 0.0;
#pragma unroll
    for (int k = 1; k < GEN_NREG; ++k) {  // fixed trip count: E stays in registers
        if (k < k0 || k >= K.nreg) continue;
        // boundary between regions k and k + 1 at radius c_k
        const double ck = K.c[k - 1];
        const double dc = std::sqrt(std::fmax(ck * ck - rs2, 0.0));
        if (dc > 0.0) off += gen_pyramid_G_jump(K, k, da, db, dc, Phi);
        E[k] = off;  // region k + 1
    }
}

// W at the lattice point with absolute offsets (ax, ay, az) > 0 (units of H), from scratch (tests).
KIWI_GHD double gen_lattice_W(const GenTab& K, double ax, double ay, double az) {
    const double d[3] = {ax, ay, az};
    const double r = std::sqrt(ax * ax + ay * ay + az * az);
    const int k = gen_region_of(K, r);
    double W = 0.0;
    for (int a = 0; a < 3; ++a) {
        double Phi[GEN_NREG];
        gen_plane_phi(K, d[a], Phi);
        W += 0.5 * GEN_PI * Phi[k - 1];
        for (int b = 0; b < 3; ++b) {
            if (b == a) continue;
            const int c = 3 - a - b;
            const double inv_l2 = 1.0 / (d[a] * d[a] + d[b] * d[b]);
            const double T = std::atan(d[a] * d[c] / (d[b] * r));
            const double Kh = 0.5 * std::log((r + d[c]) * (r + d[c]) * inv_l2);
            W += gen_pyramid_nonphi(K.reg[k - 1], d[a], d[b], d[c], r, T, Kh, 1.0 / d[a], 1.0 / d[b], 1.0 / r, inv_l2);
            double E[GEN_NREG];
            gen_line(K, d[a], d[b], E);
            W += E[k - 1];
        }
    }
    return W;
}

}  // namespace kiwi_hip
