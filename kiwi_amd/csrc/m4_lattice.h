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
// Arithmetic goes through KMath<R>: for double it is IEEE division and the device libm (the parity
// mode); for float on the device it is the hardware reciprocal, square root and log2 (about 1 ulp)
// and the atan polynomial below.  The same code compiles for the host (g++) and the device (hipcc);
// the product only runs it on the device.
#pragma once

#include <cmath>
#include <type_traits>

#ifdef __HIPCC__
#define KIWI_HD __host__ __device__ inline
#else
#define KIWI_HD inline
#endif

namespace kiwi_hip {

template <typename R>
struct M4Consts {
    static constexpr R pi = R(3.141592653589793238462643383279502884);
    static constexpr R half_pi = R(1.570796326794896619231321691639751442);
    static constexpr R inv_pi = R(0.318309886183790671537767526745028724);
};

// atan on [0, 1] in single precision: t * P(t^2), P of degree 8.  The coefficients are our own fit
// (iteratively reweighted least squares for minimum relative error on 6000 Chebyshev nodes of
// [0, 1], i.e. Lawson's algorithm towards the minimax polynomial), rounded to 10 digits.  With these
// coefficients the polynomial's relative error is 3.6e-8 (largest at t = 1) and the single-precision
// evaluation's 1.3e-7 (tests/test_lattice_math.py checks both against double-precision atan).
// Arguments above 1 use atan(x) = pi/2 - atan(1/x).
template <typename R>
KIWI_HD R atan_poly01(R t) {
    const R u = t * t;
    R p = R(2.850065706e-03);
    p = p * u + R(-1.606937870e-02);
    p = p * u + R(4.269283265e-02);
    p = p * u + R(-7.504416257e-02);
    p = p * u + R(1.064099893e-01);
    p = p * u + R(-1.420366317e-01);
    p = p * u + R(1.999262273e-01);
    p = p * u + R(-3.333307505e-01);
    p = p * u + R(1.0);
    return t * p;
}

template <typename R>
struct KMath;

// the plain value of a scalar (dual numbers, dual.h, add their own overload)
KIWI_HD double kval(double x) { return x; }
KIWI_HD float kval(float x) { return x; }

template <>
struct KMath<double> {
    static KIWI_HD double rcp(double x) { return 1.0 / x; }
    static KIWI_HD double div(double a, double b) { return a / b; }
    static KIWI_HD double sqrt_(double x) { return ::sqrt(x); }
    static KIWI_HD double log_(double x) { return ::log(x); }
    static KIWI_HD double atan_pos(double x) { return ::atan(x); }
    // atan2(y, x) for y, x >= 0
    static KIWI_HD double atan2_pos(double y, double x) { return ::atan2(y, x); }
};

template <>
struct KMath<float> {
#if defined(__HIP_DEVICE_COMPILE__)
    static KIWI_HD float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static KIWI_HD float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    static KIWI_HD float log_(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
#else
    static KIWI_HD float rcp(float x) { return 1.0f / x; }
    static KIWI_HD float sqrt_(float x) { return std::sqrt(x); }
    static KIWI_HD float log_(float x) { return std::log(x); }
#endif
    static KIWI_HD float div(float a, float b) { return a * rcp(b); }
    static KIWI_HD float atan_pos(float x) {
        const bool big = x > 1.0f;
        const float t = big ? rcp(x) : x;
        const float a = atan_poly01(t);
        return big ? M4Consts<float>::half_pi - a : a;
    }
    // one reciprocal: the smaller over the larger argument
    static KIWI_HD float atan2_pos(float y, float x) {
        const float a = atan_poly01(fminf(y, x) * rcp(fmaxf(y, x)));
        return (y > x) ? M4Consts<float>::half_pi - a : a;
    }
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
        B2 = R(0.25) * r3 * (R(-4.0 / 3.0) + r0 - R(0.3) * r2 + r3 * R(1.0 / 30.0)) - R(1.0 / 60.0);
        B3 = B2 + R(0.4) * r0;
    } else {
        B1 = R(0.25) * r3 * (R(-2.0 / 3.0) + R(0.3) * r2 - R(0.1) * r3);
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
    p.c1 = r3 * (R(1.0 / 6.0) - R(0.075) * r2) + B1;
    p.c2 = r3 * (R(1.0 / 3.0) + R(0.075) * r2) + B2;
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
    const R q4 = p * (R(2) * da2 + db2 + dc * dc * R(1.0 / 3.0));  // r0^5 * (phi-free part of I_-4)
    const R dab = da * db;
    const R X3 = da2 * da2 * T + R(0.25) * dab * (R(2) * dc * r + (l2 + R(2) * da2) * L);  // r0^4 I_-3
    const R X5 = da2 * X3 + R(0.0625) * dab * (R(2) * dc * r * (R(5) * r * r - R(3) * dc * dc) + R(3) * l2 * l2 * L);
    if (k == 1) return p * R(1.0 / 6.0) - R(0.075) * q4 + X5 * R(1.0 / 40.0);
    return p * R(1.0 / 3.0) - R(0.25) * X3 + R(0.075) * q4 - X5 * R(1.0 / 120.0) + T * R(1.0 / 60.0);
}

// Full vertex integral of one pyramid at the vertex (da, db, dc) with radius r, excluding the D
// offset: the reference's r0h3/pi * (...) of src/kernel_integral.cpp:389-397 without D2/D3.
template <typename R>
KIWI_HD R m4_vertex_g(int k, R da, R db, R dc, R r, R T, R L, R phi, const M4Plane<R>& pl) {
    const R ck = (k == 1) ? pl.c1 : (k == 2 ? pl.c2 : pl.c3);
    return (m4_pyramid_free<R>(k, da, db, dc, r, T, L) + ck * phi) * M4Consts<R>::inv_pi;
}

// Region-matching constants of one edge line (wall distance da, edge-line distance db), as offsets
// on the vertex integral: V = g_k + D_k with D_1 = 0, D_2 = D2, D_3 = D3.  This is
// compute_edge_constants (src/kernel_integral.cpp:253-350) with the crossing angles written as
// virtual vertices at radius 1 and 2: D2 = g1 - g2 at r = 1, D3 = g2 + D2 - g3 at r = 2.
template <typename R>
KIWI_HD void m4_line(R da, R db, R& D2, R& D3) {
    using M = KMath<R>;
    D2 = R(0);
    D3 = R(0);
    if (!(da > R(0)) || !(db > R(0))) return;
    const R l2 = da * da + db * db;
    // on the boundaries l2 = 4 and l2 = 1 the crossing vertex is the foot itself and the offset is exactly
    // 0 (and so is its derivative, which the adjoint carries through here): take the zero branch there
    if (!(l2 < R(4))) return;
    const M4Plane<R> pl = m4_plane(da);
    const R inv_l2 = M::rcp(l2);
    const R inv_db = M::rcp(db);
    // Both differences share one evaluation of the pyramid polynomials at the virtual vertex:
    //   pi (g1 - g2) = -p/6 - 0.15 q4 + X5/30 + X3/4 - T/60 + (c1 - c2) phi      (r = 1)
    //   pi (g2 - g3) =  p/3 + 0.075 q4 - X5/120 - X3/4 + T/60 + T/4 + (c2 - c3) phi   (r = 2)
    const R da2 = da * da, db2 = db * db, dab = da * db;
    if (l2 < R(1)) {
        const R dc = M::sqrt_(R(1) - l2);
        const R T = M::atan_pos(da * dc * inv_db);
        const R L = M::log_((R(1) + dc) * (R(1) + dc) * inv_l2);
        const R phi = M::atan2_pos(dc, db);
        const R p = dab * dc;
        const R q4 = p * (R(2) * da2 + db2 + dc * dc * R(1.0 / 3.0));
        const R X3 = da2 * da2 * T + R(0.25) * dab * (R(2) * dc + (l2 + R(2) * da2) * L);
        const R X5 = da2 * X3 + R(0.0625) * dab * (R(2) * dc * (R(5) - R(3) * dc * dc) + R(3) * l2 * l2 * L);
        D2 = (-p * R(1.0 / 6.0) - R(0.15) * q4 + X5 * R(1.0 / 30.0) + R(0.25) * X3 - T * R(1.0 / 60.0) +
              (pl.c1 - pl.c2) * phi) * M4Consts<R>::inv_pi;
    }
    const R dc = M::sqrt_(R(4) - l2);
    const R T = M::atan_pos(da * dc * inv_db * R(0.5));
    const R L = M::log_((R(2) + dc) * (R(2) + dc) * inv_l2);
    const R phi = M::atan2_pos(dc, db);
    const R p = dab * dc;
    const R q4 = p * (R(2) * da2 + db2 + dc * dc * R(1.0 / 3.0));
    const R X3 = da2 * da2 * T + R(0.25) * dab * (R(4) * dc + (l2 + R(2) * da2) * L);
    const R X5 = da2 * X3 + R(0.0625) * dab * (R(4) * dc * (R(20) - R(3) * dc * dc) + R(3) * l2 * l2 * L);
    D3 = (p * R(1.0 / 3.0) + R(0.075) * q4 - X5 * R(1.0 / 120.0) - R(0.25) * X3 + T * R(1.0 / 60.0 + 0.25) +
          (pl.c2 - pl.c3) * phi) * M4Consts<R>::inv_pi + D2;
}

// The line-table entry of one pair of lattice coordinates: m4_line(da, db) + m4_line(db, da), the two
// offsets summed, with the terms the two share formed once (l2, the crossing distance dc, the logarithm
// L, p = da db dc).  c12a, c23a are c1 - c2 and c2 - c3 of m4_plane(da), c12b, c23b those of db.
// Summing the two pyramids' polynomials: X3_ab + X3_ba = da^4 T_ab + db^4 T_ba + da db (r dc + l2 L)
// (the (l2 + 2 da^2) and (l2 + 2 db^2) weights of L add up to 4 l2), q4_ab + q4_ba = p (3 l2 + 2 dc^2 / 3),
// and X5 keeps its per-pyramid da^2 X3_ab + db^2 X3_ba with the shared part counted twice.
template <typename R>
KIWI_HD void m4_line_pair(R da, R db, R c12a, R c23a, R c12b, R c23b, R& S2, R& S3) {
    using M = KMath<R>;
    S2 = R(0);
    S3 = R(0);
    if (!(da > R(0)) || !(db > R(0))) return;
    const R l2 = da * da + db * db;
    if (!(l2 < R(4))) return;
    const R inv_l2 = M::rcp(l2);
    const R inv_da = M::rcp(da), inv_db = M::rcp(db);
    const R da2 = da * da, db2 = db * db, dab = da * db;
    const R da4 = da2 * da2, db4 = db2 * db2;
    if (l2 < R(1)) {
        const R dc = M::sqrt_(R(1) - l2);
        const R Tab = M::atan_pos(da * dc * inv_db), Tba = M::atan_pos(db * dc * inv_da);
        const R L = M::log_((R(1) + dc) * (R(1) + dc) * inv_l2);
        const R phab = M::atan2_pos(dc, db), phba = M::atan2_pos(dc, da);
        const R p = dab * dc;
        const R q4 = p * (R(3) * l2 + R(2.0 / 3.0) * dc * dc);
        const R sh = R(0.25) * dab * R(2) * dc;  // the L-free shared part of one X3
        const R X3ab = da4 * Tab + sh + R(0.25) * dab * (l2 + R(2) * da2) * L;
        const R X3ba = db4 * Tba + sh + R(0.25) * dab * (l2 + R(2) * db2) * L;
        const R X5 = da2 * X3ab + db2 * X3ba + R(0.125) * dab * (R(2) * dc * (R(5) - R(3) * dc * dc) + R(3) * l2 * l2 * L);
        S2 = (-p * R(1.0 / 3.0) - R(0.15) * q4 + X5 * R(1.0 / 30.0) + R(0.25) * (X3ab + X3ba) -
              (Tab + Tba) * R(1.0 / 60.0) + c12a * phab + c12b * phba) * M4Consts<R>::inv_pi;
    }
    const R dc = M::sqrt_(R(4) - l2);
    const R Tab = M::atan_pos(da * dc * inv_db * R(0.5)), Tba = M::atan_pos(db * dc * inv_da * R(0.5));
    const R L = M::log_((R(2) + dc) * (R(2) + dc) * inv_l2);
    const R phab = M::atan2_pos(dc, db), phba = M::atan2_pos(dc, da);
    const R p = dab * dc;
    const R q4 = p * (R(3) * l2 + R(2.0 / 3.0) * dc * dc);
    const R sh = R(0.25) * dab * R(4) * dc;
    const R X3ab = da4 * Tab + sh + R(0.25) * dab * (l2 + R(2) * da2) * L;
    const R X3ba = db4 * Tba + sh + R(0.25) * dab * (l2 + R(2) * db2) * L;
    const R X5 = da2 * X3ab + db2 * X3ba + R(0.125) * dab * (R(4) * dc * (R(20) - R(3) * dc * dc) + R(3) * l2 * l2 * L);
    S3 = (p * R(2.0 / 3.0) + R(0.075) * q4 - X5 * R(1.0 / 120.0) - R(0.25) * (X3ab + X3ba) +
          (Tab + Tba) * R(1.0 / 60.0 + 0.25) + c23a * phab + c23b * phba) * M4Consts<R>::inv_pi + S2;
}

// m4_line written as the difference of two full vertex integrals (the literal reading of
// compute_edge_constants); the checked reference for the shared-polynomial form above.
template <typename R>
KIWI_HD void m4_line_by_vertex(R da, R db, R& D2, R& D3) {
    using M = KMath<R>;
    D2 = R(0);
    D3 = R(0);
    if (!(da > R(0)) || !(db > R(0))) return;
    const R l2 = da * da + db * db;
    if (!(l2 < R(4))) return;
    const M4Plane<R> pl = m4_plane(da);
    if (l2 < R(1)) {
        const R dc = M::sqrt_(R(1) - l2);
        const R T = M::atan_pos(M::div(da * dc, db));
        const R L = M::log_(M::div((R(1) + dc) * (R(1) + dc), l2));
        const R phi = M::atan2_pos(dc, db);
        D2 = m4_vertex_g<R>(1, da, db, dc, R(1), T, L, phi, pl) - m4_vertex_g<R>(2, da, db, dc, R(1), T, L, phi, pl);
    }
    const R dc = M::sqrt_(R(4) - l2);
    const R T = M::atan_pos(M::div(da * dc, R(2) * db));
    const R L = M::log_(M::div((R(2) + dc) * (R(2) + dc), l2));
    const R phi = M::atan2_pos(dc, db);
    D3 = m4_vertex_g<R>(2, da, db, dc, R(2), T, L, phi, pl) + D2 - m4_vertex_g<R>(3, da, db, dc, R(2), T, L, phi, pl);
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
//
// The three angles add up to pi/2 for positive offsets: atan(yz/(xr)) + atan(xz/(yr)) + atan(xy/(zr)) =
// pi/2 (the solid angle of the octant box's three far faces seen from the particle is that of the
// whole octant, 4 pi / 8, less nothing: tan of the sum of two is the cotangent of the third).  So T_2 is
// pi/2 - T_0 - T_1, and outside the support, where only -(T_0 + T_1 + T_2)/2 is left, the free part is
// the constant -pi/4: such a point takes no transcendental at all.
template <typename R>
KIWI_HD R m4_lattice_free(R ax, R ay, R az, R r, int k) {
    using M = KMath<R>;
    if (k == 3) return R(-0.25) * M4Consts<R>::pi;
    const R ayz = ay * az, axz = ax * az, axy = ax * ay;
    // T_b: b is the edge-normal axis, {a, c} the other two (symmetric in a <-> c)
    R T0, T1;
    if (std::is_same<R, double>::value) {
        const R inv = M::rcp(ax * ayz * r);  // one division for the two arguments
        T0 = M::atan_pos(ayz * ayz * inv);
        T1 = M::atan_pos(axz * axz * inv);
    } else {
        // single precision: one reciprocal per angle (atan2 of the ratio's two sides)
        T0 = M::atan2_pos(ayz, ax * r);
        T1 = M::atan2_pos(axz, ay * r);
    }
    const R T2 = M4Consts<R>::half_pi - T0 - T1;
    // L_c: c is the along-edge axis, log((r + d_c)^2 / (r^2 - d_c^2)) with r^2 - d_c^2 formed exactly
    const R sx = ay * ay + az * az, sy = ax * ax + az * az, sz = ax * ax + ay * ay;
    // one reciprocal for the three denominators.  In single precision the product is kept above
    // 1e-30 (it is smaller only when all three offsets are below ~1e-5; the L terms then carry weights
    // below 1e-15, so a capped logarithm changes nothing), which keeps the reciprocal finite.
    R ix, iy, iz;
    if (std::is_same<R, double>::value) {
        const R inv = M::rcp(sx * sy * sz);
        ix = sy * sz * inv;
        iy = sx * sz * inv;
        iz = sx * sy * inv;
    } else {
        ix = M::rcp(sx);
        iy = M::rcp(sy);
        iz = M::rcp(sz);
    }
    const R Lx = M::log_((r + ax) * (r + ax) * ix);
    const R Ly = M::log_((r + ay) * (r + ay) * iy);
    const R Lz = M::log_((r + az) * (r + az) * iz);
    // The 6 ordered (wall a, edge normal b, along c) pyramids of m4_pyramid_free, summed in closed
    // form.  With p = ax ay az, and for each T_b the two pyramids (a,b,c), (c,b,a), for each L_c the
    // two pyramids (a,b,c), (b,a,c):
    //   sum r0^4 I_-3 = sum_b (da^4 + dc^4) T_b + 3 p r + sum_c da db (da^2 + db^2) L_c
    //   sum r0^6 I_-5 = sum_b (da^6 + dc^6) T_b + 4 p r^3
    //                   + sum_c da db (9/8 (da^4 + db^4) + 5/4 da^2 db^2) L_c
    //   sum r0^3 I'_-2 = 6 p,   sum r0^5 I'_-4 = p sum_perm (2 da^2 + db^2 + dc^2/3) = (20/3) p r^2
    const R x2 = ax * ax, y2 = ay * ay, z2 = az * az;
    const R x4 = x2 * x2, y4 = y2 * y2, z4 = z2 * z2;
    const R p = axy * az;
    const R pr = p * r;
    const R r2 = r * r;
    // T coefficients (b = x: {y, z}; b = y: {x, z}; b = z: {x, y})
    const R sx3 = ((y4 + z4) * T0 + (x4 + z4) * T1) + (x4 + y4) * T2;
    const R sx5 = ((y4 * y2 + z4 * z2) * T0 + (x4 * x2 + z4 * z2) * T1) + (x4 * x2 + y4 * y2) * T2;
    // L coefficients (c = x: {y, z}; c = y: {x, z}; c = z: {x, y})
    const R kx = ayz * sx, ky = axz * sy, kz = axy * sz;
    const R lx3 = (kx * Lx + ky * Ly) + kz * Lz;
    const R lx5 = (ayz * (R(1.125) * (y4 + z4) + R(1.25) * y2 * z2) * Lx +
                   axz * (R(1.125) * (x4 + z4) + R(1.25) * x2 * z2) * Ly) +
                  axy * (R(1.125) * (x4 + y4) + R(1.25) * x2 * y2) * Lz;
    const R SX3 = sx3 + R(3) * pr + lx3;
    const R SX5 = sx5 + R(4) * pr * r2 + lx5;
    if (k == 1) return p * (R(1) - R(0.5) * r2) + SX5 * R(1.0 / 40.0);
    return p * (R(2) + R(0.5) * r2) - R(0.25) * SX3 - SX5 * R(1.0 / 120.0) + M4Consts<R>::half_pi * R(1.0 / 30.0);
}

// The same sum taken pyramid by pyramid (m4_pyramid_free); kept as the readable form of the closed
// sum above and checked against it in tests/test_lattice_math.py.
template <typename R>
KIWI_HD R m4_lattice_free_by_pyramid(R ax, R ay, R az, R r, int k) {
    using M = KMath<R>;
    const R T0 = M::atan_pos(M::div(ay * az, ax * r));
    const R T1 = M::atan_pos(M::div(ax * az, ay * r));
    const R T2 = M::atan_pos(M::div(ax * ay, az * r));
    const R Lx = M::log_(M::div((r + ax) * (r + ax), ay * ay + az * az));
    const R Ly = M::log_(M::div((r + ay) * (r + ay), ax * ax + az * az));
    const R Lz = M::log_(M::div((r + az) * (r + az), ax * ax + ay * ay));
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
    const R r = KMath<R>::sqrt_(r2);
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
