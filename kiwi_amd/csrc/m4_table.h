// The tabulated cubic-spline vertex integral (the reference's use_tabulated_integral path): a density
// lookup table over (r0/h, mu0 = r0/sqrt(r0^2 + R0^2), cos phi) with two sub-grids per axis, filled with
// the analytic vertex integral at h = 1, and read back by trilinear interpolation.  Table geometry and
// lookup follow generate_density_table / gridded_integral (src/kernel_integral.cpp:57-190); the values
// come from this build's own FP64 vertex integral (m4_lattice.h: one pyramid's free part, plane and line
// terms).  Host (tests) and device code.
#pragma once

#include <cmath>

#include "m4_lattice.h"

namespace kiwi_hip {

struct M4TableGeom {
    static constexpr int nr01 = 50, nr02 = 199, nR01 = 150, nR02 = 150, nph1 = 150, nph2 = 150;
    static constexpr int nr0 = nr01 + nr02, nR0 = nR01 + nR02, nph = nph1 + nph2;
    static constexpr double rl = 0.1, mul = 0.98, cphil = 0.98;
    // (nr0 + 2) x (nR0 + 1) x (nph + 1) doubles, k fastest
    static constexpr long long size = static_cast<long long>(nr0 + 2) * (nR0 + 1) * (nph + 1);
    static KIWI_HD long long at(int i, int j, int k) {
        return (static_cast<long long>(i) * (nR0 + 1) + j) * (nph + 1) + k;
    }
    // grid coordinates of entry (i, j, k), i = 1..nr0, j, k = 0..n-1 (src/kernel_integral.cpp:80-107)
    static KIWI_HD double r0_of(int i) { return (i <= nr01) ? (rl / nr01) * i : rl + ((2.0 - rl) / nr02) * (i - nr01); }
    static KIWI_HD double mu0_of(int j) {
        return (j < nR01) ? (mul / (nR01 - 1)) * j : mul + ((1.0 - mul) / nR02) * (j - nR01 + 1);
    }
    static KIWI_HD double cphi_of(int k) {
        return (k < nph1) ? (cphil / (nph1 - 1)) * k : cphil + ((1.0 - cphil) / nph2) * (k - nph1 + 1);
    }
};

// The analytic vertex integral at h = 1 for wall distance r0, edge distance R0 (possibly +inf) and angle
// phi: (free part + c_k phi) / pi + D_k of the pyramid (da, db, dc) = (r0, R0, R0 tan phi), region k of
// its radius.  For R0 = +inf the pyramid is the wall's whole region-3 sector: c3(r0) phi / pi.
KIWI_HD double m4_vertex_value(double r0, double R0, double phi) {
    if (r0 == 0.0 || R0 == 0.0 || phi == 0.0) return 0.0;
    const M4Plane<double> pl = m4_plane<double>(r0);
    if (!(R0 < INFINITY)) return pl.c3 * phi * M4Consts<double>::inv_pi;
    const double sn = sin(phi), cs = cos(phi);
    const double da = r0, db = R0, dc = R0 * sn / cs;
    const double r2 = da * da + db * db + dc * dc;
    const double r = sqrt(r2);
    const int k = m4_region(r2);
    const double T = atan(da * dc / (db * r));
    const double L = log((r + dc) * (r + dc) / (da * da + db * db));
    double D = 0.0;
    if (k > 1) {
        double D2, D3;
        m4_line<double>(da, db, D2, D3);
        D = (k == 2) ? D2 : D3;
    }
    return m4_vertex_g<double>(k, da, db, dc, r, T, L, phi, pl) + D;
}

// gridded_integral (src/kernel_integral.cpp:136-190) at h = 1 for wall distance r0, edge distance R0 and
// cos(phi) = cphi (the caller has phi = 0 exactly when cphi = 1): trilinear interpolation in the table.
KIWI_HD double m4_gridded_value(const double* tab, double r0, double R0, double cphi) {
    using G = M4TableGeom;
    if (r0 == 0.0 || R0 == 0.0 || cphi >= 1.0) return 0.0;
    if (r0 < 0.0) r0 = 0.0;
    if (R0 < 0.0) R0 = 0.0;
    const double mu0 = r0 / sqrt(r0 * r0 + R0 * R0);
    if (r0 > 2.0) r0 = 2.0;
    int i, j, k;
    if (r0 < G::rl) i = static_cast<int>(r0 / (G::rl / G::nr01));
    else i = G::nr01 + static_cast<int>((r0 - G::rl) / ((2.0 - G::rl) / G::nr02));
    if (mu0 < G::mul) j = static_cast<int>(mu0 / (G::mul / (G::nR01 - 1)));
    else j = G::nR01 - 1 + static_cast<int>((mu0 - G::mul) / ((1.0 - G::mul) / G::nR02));
    if (cphi < G::cphil) k = static_cast<int>(cphi / (G::cphil / (G::nph1 - 1)));
    else k = G::nph1 - 1 + static_cast<int>((cphi - G::cphil) / ((1.0 - G::cphil) / G::nph2));
    i = i < 0 ? 0 : (i > G::nr0 ? G::nr0 : i);
    j = j < 0 ? 0 : (j > G::nR0 - 1 ? G::nR0 - 1 : j);
    k = k < 0 ? 0 : (k > G::nph - 1 ? G::nph - 1 : k);
    double fr;
    if (r0 < G::rl) fr = (r0 - (G::rl / G::nr01) * i) / (G::rl / G::nr01);
    else fr = (r0 - G::rl - ((2.0 - G::rl) / G::nr02) * (i - G::nr01)) / ((2.0 - G::rl) / G::nr02);
    const double fx1 = fr * tab[G::at(i + 1, j, k)] + (1 - fr) * tab[G::at(i, j, k)];
    const double fx2 = fr * tab[G::at(i + 1, j + 1, k)] + (1 - fr) * tab[G::at(i, j + 1, k)];
    const double fx3 = fr * tab[G::at(i + 1, j, k + 1)] + (1 - fr) * tab[G::at(i, j, k + 1)];
    const double fx4 = fr * tab[G::at(i + 1, j + 1, k + 1)] + (1 - fr) * tab[G::at(i, j + 1, k + 1)];
    if (mu0 < G::mul) fr = (mu0 - G::mul / (G::nR01 - 1) * j) / (G::mul / (G::nR01 - 1));
    else fr = (mu0 - G::mul - (1.0 - G::mul) / G::nR02 * (j - G::nR01 + 1)) / ((1.0 - G::mul) / G::nR02);
    const double fy1 = fr * fx2 + (1 - fr) * fx1;
    const double fy2 = fr * fx4 + (1 - fr) * fx3;
    if (cphi < G::cphil) fr = (cphi - G::cphil / (G::nph1 - 1) * k) / (G::cphil / (G::nph1 - 1));
    else fr = (cphi - G::cphil - (1.0 - G::cphil) / G::nph2 * (k - G::nph1 + 1)) / ((1.0 - G::cphil) / G::nph2);
    const double fz = fr * fy2 + (1 - fr) * fy1;
    if (j == G::nR01 + G::nR02 - 1 || k == G::nph1 + G::nph2 - 1) return 0.0;
    return fz;
}

}  // namespace kiwi_hip
