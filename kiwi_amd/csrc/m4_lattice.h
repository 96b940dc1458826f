// Cubic-spline (M4) kernel integral over Cartesian cells, restructured onto the grid lattice.
//
// The reference integrates a particle's kernel over one cell as 6 faces x 4 edges x 2 vertex
// integrals (src/cartesian_grid.cpp:245-259 -> src/kernel_integral.cpp:471-647), each vertex integral
// being the kernel mass of a "vertex pyramid" (apex at the particle, base a right triangle in the
// face plane; src/kernel_integral.cpp:356-400).  On a Cartesian grid every such pyramid is anchored
// at a lattice point, so the whole cell sum regroups exactly into
//
//     Msum(cell) = sum over the cell's 8 corners  e(corner) * s(corner) * W(corner)
//
//   e = +1 for corners (1,0,0),(0,1,0),(0,0,1),(1,1,1) and -1 for the other four,
//   s = sgn(dx) sgn(dy) sgn(dz) with d = corner - particle (s = 0 if the corner lies in a plane
//       through the particle, where every pyramid anchored there is flat),
//   W(|dx|,|dy|,|dz|) = the sum of the 6 vertex integrals anchored at the corner (one per choice of
//       wall axis a and edge-normal axis b; the third axis c runs along the edge), which is a
//       symmetric function of the three offsets.
//
// W is evaluated once per lattice point and shared by the 8 cells around it, instead of 48 vertex
// integrals per cell.  All quantities are in units of the smoothing length h (h = 1 below), where
// the M4 support is 2.  With r = |d| and, for the pyramid (a,b,c), r0 = da, R_0 = db, tan(phi) =
// dc/db, the reference's I-series closes into polynomials in (da, db, dc, r) plus
//   T_b = atan(da dc / (db r))            (the reference's I1 = atan(u/a), :382)
//   L_c = log((r + dc)^2 / (da^2 + db^2))  (the reference's logs = log(1+u) - log(1-u), :381)
// and the phi terms of the two pyramids sharing a wall axis add up to pi/2, so no atan2 is needed
// per lattice point.  The region matching constants D2/D3 (src/kernel_integral.cpp:253-350) are
// kept per (wall axis, edge line) and evaluated at the virtual vertices where the vertex radius
// reaches 1 and 2 (h and 2h).
//
// The same code compiles for the host (g++) and the device (hipcc); the product only runs it on
// the device.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define KIWI_HD __host__ __device__ inline
#else
#define KIWI_HD inline
#endif

namespace kiwi_hip {

template <typename R>
struct M4Consts {
    static constexpr R pi = R(3.141592653589793238462643383279502884);
    static constexpr R inv_pi = R(0.318309886183790671537767526745028724);
};

// Wall constants B1, B2, B3 (src/kernel_integral.cpp:236-245) with h = 1, multiplied out so that no
// power of 1/r0 remains.  r0 >= 0.
template <typename R>
KIWI_HD void m4_wall_B(R r0, R& B1, R& B2, R& B3) {
    const R r2 = r0 * r0, r3 = r2 * r0;
    if (r0 >= R(2)) {
        B1 = R(0);
        B2 = R(0);
        B3 = R(0.25);
    } else if (r0 > R(1)) {
        B1 = R(0);
        B2 = R(0.25) * r3 * (R(-4) / R(3) + r0 - R(0.3) * r2 + r3 / R(30)) - R(1) / R(60);
        B3 = B2 + R(0.4) * r0;
    } else {
        B1 = R(0.25) * r3 * (R(-2) / R(3) + R(0.3) * r2 - R(0.1) * r3);
        B2 = B1 - R(0.05) * r0;
        B3 = B1 + R(0.35) * r0;
    }
}

// Coefficient of phi/pi in the vertex integral of a vertex in radial region k (1: r < 1, 2: r < 2,
// 3: outside), for wall distance r0.  Region k of a vertex implies r0 < k, so B_k is defined.
template <typename R>
struct M4Plane {
    R c1, c2, c3;
};

template <typename R>
KIWI_HD M4Plane<R> m4_plane(R r0) {
    R B1, B2, B3;
    m4_wall_B(r0, B1, B2, B3);
    const R r2 = r0 * r0, r3 = r2 * r0;
    M4Plane<R> p;
    p.c1 = r3 * (R(1) / R(6) - R(0.075) * r2) + B1;
    p.c2 = r3 * (R(1) / R(3) + R(0.075) * r2) + B2;
    p.c3 = B3;
    return p;
}

// pi times the phi-free part of one vertex pyramid (wall axis a, edge normal b, along-edge c) in
// region k.  da, db, dc > 0; r = |d|; T = atan(da dc / (db r)); L = log((r + dc)^2 / (da^2 + db^2)).
template <typename R>
KIWI_HD R m4_pyramid_free(int k, R da, R db, R dc, R r, R T, R L) {
    if (k == 3) return R(-0.25) * T;
    const R p = da * db * dc;
    const R da2 = da * da, db2 = db * db;
    const R l2 = da2 + db2;
    const R q4 = p * (R(2) * da2 + db2 + dc * dc / R(3));  // r0^5 * (phi-free part of I_-4)
    const R dab = da * db;
    const R X3 = da2 * da2 * T + R(0.25) * dab * (R(2) * dc * r + (l2 + R(2) * da2) * L);  // r0^4 I_-3
    const R X5 = da2 * X3 + R(0.0625) * dab * (R(2) * dc * r * (R(5) * r * r - R(3) * dc * dc) + R(3) * l2 * l2 * L);
    if (k == 1) return p / R(6) - R(0.075) * q4 + X5 / R(40);
    return p / R(3) - R(0.25) * X3 + R(0.075) * q4 - X5 / R(120) + T / R(60);
}

// Full vertex integral of one pyramid at the vertex (da, db, dc) with radius r, excluding the D
// offset: the reference's r0h3/pi * (...) of src/kernel_integral.cpp:389-397 without D2/D3.
template <typename R>
KIWI_HD R m4_vertex_g(int k, R da, R db, R dc, R r, const M4Plane<R>& pl) {
    const R T = atan(da * dc / (db * r));
    const R L = log((r + dc) * (r + dc) / (da * da + db * db));
    const R phi = atan2(dc, db);
    const R ck = (k == 1) ? pl.c1 : (k == 2 ? pl.c2 : pl.c3);
    return (m4_pyramid_free<R>(k, da, db, dc, r, T, L) + ck * phi) * M4Consts<R>::inv_pi;
}

// Region-matching constants of one edge line (wall distance da, edge-line distance db), as offsets
// on the vertex integral: V = g_k + D_k with D_1 = 0, D_2 = D2, D_3 = D3.  This is
// compute_edge_constants (src/kernel_integral.cpp:253-350) with the crossing angles written as
// virtual vertices at radius 1 and 2: D2 = g1 - g2 at r = 1, D3 = g2 + D2 - g3 at r = 2.
template <typename R>
KIWI_HD void m4_line(R da, R db, R& D2, R& D3) {
    D2 = R(0);
    D3 = R(0);
    if (!(da > R(0)) || !(db > R(0))) return;
    const R l2 = da * da + db * db;
    if (!(l2 <= R(4))) return;
    const M4Plane<R> pl = m4_plane(da);
    if (l2 <= R(1)) {
        const R dc = sqrt(R(1) - l2);
        D2 = m4_vertex_g<R>(1, da, db, dc, R(1), pl) - m4_vertex_g<R>(2, da, db, dc, R(1), pl);
    }
    const R dc = sqrt(R(4) - l2);
    D3 = m4_vertex_g<R>(2, da, db, dc, R(2), pl) + D2 - m4_vertex_g<R>(3, da, db, dc, R(2), pl);
}

// Radial region of a lattice point at squared distance r2 (units of h): 1 inside h, 2 inside 2h,
// 3 outside the support (src/kernel_integral.cpp:389-397).
template <typename R>
KIWI_HD int m4_region(R r2) {
    return (r2 < R(1)) ? 1 : (r2 < R(4) ? 2 : 3);
}

// pi times the phi-free part of the 6 pyramids anchored at a lattice point with absolute offsets
// ax, ay, az > 0, radius r, in region k.  The full W is
//     W = free / pi + 0.5 * sum_a c_k(d_a) + sum_{a != b} D_k(d_a, d_b)
// because the phi terms of the pyramids (a,b,c) and (a,c,b) add up to c_k(d_a) (pi/2) / pi.
template <typename R>
KIWI_HD R m4_lattice_free(R ax, R ay, R az, R r, int k) {
    // T_b: b is the edge-normal axis, {a, c} the other two (symmetric in a <-> c)
    const R T0 = atan(ay * az / (ax * r));
    const R T1 = atan(ax * az / (ay * r));
    const R T2 = atan(ax * ay / (az * r));
    if (k == 3) return R(-0.5) * (T0 + T1 + T2);
    // L_c: c is the along-edge axis, log((r + d_c)^2 / (r^2 - d_c^2)) with r^2 - d_c^2 formed exactly
    const R Lx = log((r + ax) * (r + ax) / (ay * ay + az * az));
    const R Ly = log((r + ay) * (r + ay) / (ax * ax + az * az));
    const R Lz = log((r + az) * (r + az) / (ax * ax + ay * ay));
    // the 6 ordered (wall a, edge normal b, along c) pyramids
    return m4_pyramid_free<R>(k, ax, ay, az, r, T1, Lz) + m4_pyramid_free<R>(k, ax, az, ay, r, T2, Ly) +
           m4_pyramid_free<R>(k, ay, ax, az, r, T0, Lz) + m4_pyramid_free<R>(k, ay, az, ax, r, T2, Lx) +
           m4_pyramid_free<R>(k, az, ax, ay, r, T0, Ly) + m4_pyramid_free<R>(k, az, ay, ax, r, T1, Lx);
}

// W at a lattice point: the sum of the 6 vertex integrals anchored there.  ax, ay, az > 0 are the
// absolute offsets (units of h); plane[a] holds m4_plane(|d_a|); D2/D3[a][b] the offsets of the line
// with wall axis a and edge-normal axis b (diagonal unused).
template <typename R>
KIWI_HD R m4_lattice_W(R ax, R ay, R az, const M4Plane<R> plane[3], const R D2[3][3], const R D3[3][3]) {
    const R r2 = ax * ax + ay * ay + az * az;
    const R r = sqrt(r2);
    const int k = m4_region(r2);
    R sum_phi = R(0), sum_D = R(0);
    for (int a = 0; a < 3; ++a) {
        sum_phi += (k == 1) ? plane[a].c1 : (k == 2 ? plane[a].c2 : plane[a].c3);
        for (int b = 0; b < 3; ++b)
            if (a != b) sum_D += (k == 1) ? R(0) : (k == 2 ? D2[a][b] : D3[a][b]);
    }
    return m4_lattice_free<R>(ax, ay, az, r, k) * M4Consts<R>::inv_pi + R(0.5) * sum_phi + sum_D;
}

}  // namespace kiwi_hip
