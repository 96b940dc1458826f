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
// Written once for host (g++, tests) and device (hipcc); double precision throughout.  The table type
// Tab needs the fields of KiwiKernelTable (nreg, c, A3_pi, a, B).
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
// given T = atan(da dc / (db r)) and Kh = atanh(dc / r).  da, db, dc > 0.  The basis integrals follow
// basis_integrals (src/kernel_integral_generic.cpp:19-66) with sin/cos of phi taken from the lattice.
template <class Tab>
KIWI_GHD double gen_pyramid_nonphi(const Tab& K, int k, double da, double db, double dc, double r, double T,
                                   double Kh) {
    const double A3 = K.A3_pi / GEN_PI;
    const double* av = K.a[k - 1];
    int imax = 0;  // largest nonzero coefficient index
    for (int idx = GEN_NI - 1; idx >= 0; --idx)
        if (av[idx] != 0.0) {
            imax = idx;
            break;
        }
    const double a = db / da, a2 = a * a;
    double sum = (av[0] != 0.0) ? A3 * av[0] * T : 0.0;  // I_1
    if (imax < 2) return sum;
    // odd orders I_-(2m+1), coefficient index 2m + 2
    const double u = dc / r;
    const double omu2 = (da * da + db * db) / (r * r);
    double Iodd = T, Km = Kh, p1a2 = 1.0, om = 1.0, qp = da * da;  // qp = q0^idx
    for (int m = 0; 2 * m + 2 <= imax; ++m) {
        if (m > 0) {
            om *= omu2;  // omu2^m
            Km = u / (2.0 * m * om) + (2.0 * m - 1.0) / (2.0 * m) * Km;
            p1a2 *= (1.0 + a2);
            qp *= da * da;
        }
        Iodd += a * p1a2 * Km;
        const double c = av[2 * m + 2];
        if (c != 0.0) sum += A3 * c * qp * Iodd;
    }
    // even orders I_-2k without phi: sum_l C(k, l) a^2l J_2l, coefficient index 2k + 1, summed over l
    // outermost so that the secant recurrence J_2l = sec^(2l-2) t / (2l-1) + (2l-2)/(2l-1) J_2l-2 runs
    // once (no array: sum_k c_k q0^(2k+1) sum_l C(k,l) a^2l J_2l = sum_l a^2l J_2l sum_{k>=l} ..)
    const int kmax = (imax - 1) / 2;
    if (kmax >= 1) {
        const double t = dc / db, sec2 = 1.0 + t * t;
        double Jl = t, sp = 1.0, apl = a2, q2l = da * da * da;  // J_2l, sec^(2l-2), a^2l, q0^(2l+1)
        for (int l = 1; l <= kmax; ++l) {
            if (l > 1) {
                sp *= sec2;
                Jl = sp * t / (2.0 * l - 1.0) + (2.0 * l - 2.0) / (2.0 * l - 1.0) * Jl;
                apl *= a2;
                q2l *= da * da;
            }
            double w = 0.0, binom = 1.0, qk = q2l;  // C(k, l), q0^(2k+1) for k = l, l+1, ..
            for (int kk = l; kk <= kmax; ++kk) {
                if (kk > l) {
                    binom = binom * kk / (kk - l);
                    qk *= da * da;
                }
                const double c = av[2 * kk + 1];
                if (c != 0.0) w += c * qk * binom;
            }
            sum += A3 * w * apl * Jl;
        }
    }
    return sum;
}

// Full G_k at the pyramid (da, db, dc), phi = atan(dc / db): used for the edge-line offsets.
template <class Tab>
KIWI_GHD double gen_pyramid_G(const Tab& K, int k, double da, double db, double dc) {
    const double r = std::sqrt(da * da + db * db + dc * dc);
    const double T = std::atan(da * dc / (db * r));
    const double Kh = 0.5 * std::log((r + dc) * (r + dc) / (da * da + db * db));
    double Phi[GEN_NREG];
    gen_plane_phi(K, da, Phi);
    return gen_pyramid_nonphi(K, k, da, db, dc, r, T, Kh) + Phi[k - 1] * std::atan2(dc, db);
}

// Region offsets of the edge line with wall distance da and edge distance db (units of H), per
// region k = 1..nreg: generic_edge_constants (src/kernel_integral_generic.cpp:119-148) with the
// segment boundaries as virtual vertices at radius c_k.  Regions below the line's own foot region
// get 0 (no vertex of the line lies there).
template <class Tab>
KIWI_GHD void gen_line(const Tab& K, double da, double db, double E[GEN_NREG]) {
    for (int k = 0; k < GEN_NREG; ++k) E[k] = 0.0;
    if (!(da > 0.0) || !(db > 0.0)) return;
    const double rs2 = da * da + db * db;
    const int k0 = gen_region_of(K, std::sqrt(rs2));
    double off = 0.0;
#pragma unroll
    for (int k = 1; k < GEN_NREG; ++k) {  // fixed trip count: E stays in registers
        if (k < k0 || k >= K.nreg) continue;
        // boundary between regions k and k + 1 at radius c_k
        const double ck = K.c[k - 1];
        const double dc = std::sqrt(std::fmax(ck * ck - rs2, 0.0));
        if (dc > 0.0) off += gen_pyramid_G(K, k, da, db, dc) - gen_pyramid_G(K, k + 1, da, db, dc);
        E[k] = off;  // region k + 1
    }
}

// W at the lattice point with absolute offsets (ax, ay, az) > 0 (units of H), from scratch (tests).
template <class Tab>
KIWI_GHD double gen_lattice_W(const Tab& K, double ax, double ay, double az) {
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
            const double T = std::atan(d[a] * d[c] / (d[b] * r));
            const double Kh = 0.5 * std::log((r + d[c]) * (r + d[c]) / (d[a] * d[a] + d[b] * d[b]));
            W += gen_pyramid_nonphi(K, k, d[a], d[b], d[c], r, T, Kh);
            double E[GEN_NREG];
            gen_line(K, d[a], d[b], E);
            W += E[k - 1];
        }
    }
    return W;
}

}  // namespace kiwi_hip
