// The signed face integral over an arbitrary planar polygon: the face walk of the reference
// (face_integral_impl, src/kernel_integral.cpp:471-647) for the Voronoi map, where cells are general
// polyhedra and the lattice restatement does not apply.  The walk drops the perpendicular from the
// particle onto the face plane (r0), and for every edge the perpendicular from that foot onto the edge
// line (R0); each edge contributes the vertex integrals at its two ends, added or subtracted as the
// foot falls inside or beyond the edge, with a sign from the winding.  It also reports whether the
// face comes closer to the particle than the support (the BFS reachability test, :642-645).  A face of
// fewer than three vertices integrates to 0 and reports unreachable here; the caller applies the
// reference's rule for it (:478-481).
//
// The per-edge geometry is restated in well-conditioned form: R0 from a cross product instead of
// r12 sqrt(1 - cos^2), the vertex angle as atan2(t, R0) instead of acos(R0 / r12), where t is the
// vertex's distance along the edge line from the perpendicular foot, and the vertex integrals take t
// itself instead of R0 tan(phi).  In exact arithmetic every quantity, branch and guard equals the
// reference's (R0 = 0 for a degenerate edge, phi = 0 for a vertex at the foot); in double precision the
// reference's acos loses up to half the digits when R0 / r12 is within rounding of 1.
//
// Ev supplies the vertex integral: wall(|r0|) once per face, edge(R0) once per edge, at(phi, t) per
// vertex.  Double precision (the parity mode), or single precision on vertex coordinates taken
// relative to the particle in double and rounded (the mixed mode of the cubic spline); host (tests)
// and device code.
#pragma once

#include <cmath>
#include <type_traits>

#include "gen_lattice.h"
#include "m4_lattice.h"
#include "m4_table.h"

#ifdef __HIPCC__
#define KIWI_FHD __host__ __device__ inline
#else
#define KIWI_FHD inline
#endif

namespace kiwi_hip {

// vertex integral of the analytic cubic spline, at h (physical units in, mass fraction out)
template <typename R>
struct EvM4T {
    R inv_h;
    R da = R(0), db = R(0), D2 = R(0), D3 = R(0);
    M4Plane<R> pl{};
    KIWI_FHD void wall(R ar0) {
        da = ar0 * inv_h;
        pl = m4_plane<R>(da);
    }
    KIWI_FHD void edge(R R0) {
        db = R0 * inv_h;
        m4_line<R>(da, db, D2, D3);
    }
    KIWI_FHD R at(R phi, R t) const {
        using M = KMath<R>;
        if (!(da > R(0)) || !(db > R(0)) || phi == R(0)) return R(0);
        const R dc = t * inv_h;  // = db tan(phi)
        const R r2 = da * da + db * db + dc * dc;
        const R r = M::sqrt_(r2);
        const int k = m4_region(r2);
        const R T = M::atan_pos(M::div(da * dc, db * r));
        const R L = M::log_(M::div((r + dc) * (r + dc), da * da + db * db));
        const R D = (k == 1) ? R(0) : (k == 2 ? D2 : D3);
        return m4_vertex_g<R>(k, da, db, dc, r, T, L, phi, pl) + D;
    }
};
using EvM4 = EvM4T<double>;

// vertex integral of a generic kernel with support H
template <class Tab>
struct EvGen {
    const Tab* K;
    double inv_H;
    double q0 = 0.0, qR = 0.0;
    double Phi[GEN_NREG] = {0.0, 0.0, 0.0, 0.0};
    double E[GEN_NREG] = {0.0, 0.0, 0.0, 0.0};
    KIWI_FHD void wall(double ar0) {
        q0 = ar0 * inv_H;
        gen_plane_phi(*K, q0, Phi);
    }
    KIWI_FHD void edge(double R0) {
        qR = R0 * inv_H;
        gen_line(*K, q0, qR, E);
    }
    KIWI_FHD double at(double phi, double t) const {
        if (!(q0 > 0.0) || !(qR > 0.0) || phi == 0.0) return 0.0;
        const double dc = t * inv_H;  // = qR tan(phi)
        const double r = sqrt(q0 * q0 + qR * qR + dc * dc);
        const int k = gen_region_of(*K, r);
        const double inv_r = 1.0 / r;
        const double T = atan(q0 * dc / qR * inv_r);
        const double Kh = 0.5 * log((r + dc) * (r + dc) / (q0 * q0 + qR * qR));
        double Ek = 0.0, Pk = 0.0;
#pragma unroll
        for (int q = 0; q < GEN_NREG; ++q)
            if (q == k - 1) {
                Ek = E[q];
                Pk = Phi[q];
            }
        return Ek + gen_pyramid_nonphi(K->reg[k - 1], q0, qR, dc, r, T, Kh, inv_r) + Pk * phi;
    }
};

// vertex integral read from the tabulated cubic spline (gridded_integral)
struct EvTab {
    const double* table;
    double inv_h;
    double r0 = 0.0, R0 = 0.0;
    KIWI_FHD void wall(double ar0) { r0 = ar0 * inv_h; }
    KIWI_FHD void edge(double R) { R0 = R * inv_h; }
    KIWI_FHD double at(double phi, double t) const {
        if (phi == 0.0) return 0.0;
        const double tt = t * inv_h;
        return m4_gridded_value(table, r0, R0, R0 / sqrt(R0 * R0 + tt * tt));  // cos(phi)
    }
};

// Vertices of a face: absolute coordinates in double (the parity mode), or coordinates relative to the
// particle, formed in double and rounded to R (the particle then sits at the origin).
struct AbsVerts {
    const double* v;
    KIWI_FHD void get(int j, double& x, double& y, double& z) const {
        x = v[3 * j];
        y = v[3 * j + 1];
        z = v[3 * j + 2];
    }
};
template <typename R>
struct RelVerts {
    const double* v;
    double x0, y0, z0;
    KIWI_FHD void get(int j, R& x, R& y, R& z) const {
        x = static_cast<R>(v[3 * j] - x0);
        y = static_cast<R>(v[3 * j + 1] - y0);
        z = static_cast<R>(v[3 * j + 2] - z0);
    }
};

// small 3-vectors for the walk's geometry
template <typename R>
struct Vec3 {
    R x, y, z;
};
template <typename R>
KIWI_FHD Vec3<R> vsub(const Vec3<R>& a, const Vec3<R>& b) {
    return Vec3<R>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename R>
KIWI_FHD Vec3<R> vcross(const Vec3<R>& a, const Vec3<R>& b) {
    return Vec3<R>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R>
KIWI_FHD R vdot(const Vec3<R>& a, const Vec3<R>& b) {
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
template <typename R, class V>
KIWI_FHD Vec3<R> vertex(const V& vs, int j) {
    Vec3<R> v;
    vs.get(j, v.x, v.y, v.z);
    return v;
}

// The face integral of the polygon of n vertices (accessor vs) for a particle at (x0, y0, z0) with
// kernel support radius `support`; *reachable is set when the face's closest point lies within the
// support.
//
// The double-precision walk takes the face plane through the first three vertices, as the reference
// does (:495-506).  Voronoi faces can have two nearly coincident vertices (edges 1e-4 of the cell), and
// a plane from the first three is then as ill-conditioned as that edge is short: harmless in double
// (1e-13) but 1e-4 of r0 in single precision.  The single-precision walk therefore takes Newell's normal
// (the sum over all edges, which no short edge spoils) through the vertex centroid: the same plane for a
// planar convex polygon in exact arithmetic.  With the normal of the first three vertices the winding
// reference s2 equals -r0 |N| exactly, so the edge sign depends on s1 alone; the single-precision walk
// uses that form.
template <typename R, class V, class Ev>
KIWI_FHD R face_walk_t(R x0, R y0, R z0, R support, const V& vs, int n, Ev& ev, bool* reachable) {
    using M = KMath<R>;
    constexpr bool NEWELL = !std::is_same<R, double>::value;
    *reachable = false;
    if (n < 3) return R(0);
    const Vec3<R> P = {x0, y0, z0};
    const Vec3<R> V0 = vertex<R>(vs, 0), V1 = vertex<R>(vs, 1), V2 = vertex<R>(vs, 2);
    R A, B, C, D;
    if (NEWELL) {
        R gx = R(0), gy = R(0), gz = R(0);
        A = B = C = R(0);
        for (int j = 0; j < n; ++j) {
            R xa, ya, za, xb, yb, zb;
            vs.get(j, xa, ya, za);
            vs.get(j == n - 1 ? 0 : j + 1, xb, yb, zb);
            A += (ya - yb) * (za + zb);
            B += (za - zb) * (xa + xb);
            C += (xa - xb) * (ya + yb);
            gx += xa;
            gy += ya;
            gz += za;
        }
        const R inv_n = R(1) / static_cast<R>(n);
        D = -(A * gx + B * gy + C * gz) * inv_n;
    } else {
        // the plane of the first three vertices (:495-506): normal (V1 - V0) x (V2 - V0) through V0
        const Vec3<R> N = vcross(vsub(V1, V0), vsub(V2, V0));
        A = N.x;
        B = N.y;
        C = N.z;
        D = -((N.x * V0.x + N.y * V0.y) + N.z * V0.z);
    }
    const R nrm = M::sqrt_(A * A + B * B + C * C);
    const R r0 = M::div(A * x0 + B * y0 + C * z0 + D, nrm);
    const R ar0 = fabs(r0);
    const R xp = x0 - M::div(r0 * A, nrm), yp = y0 - M::div(r0 * B, nrm), zp = z0 - M::div(r0 * C, nrm);
    // orientation of the first three vertices as seen from the particle: the triple product
    // (V0 - P) . ((V1 - P) x (V2 - P)) (:525-529), against which every edge's winding is signed
    const R s2 = vdot(vsub(V0, P), vcross(vsub(V1, P), vsub(V2, P)));
    ev.wall(ar0);

    // the foot lies inside the polygon while every edge that is not (numerically) through it sees it on
    // the same side as the first such edge
    bool inside = true, side_known = false, first_positive = false;
    R min_edge = R(INFINITY), total = R(0);
    for (int j = 0; j < n; ++j) {
        const int jn = (j == n - 1) ? 0 : j + 1;
        R x2, y2, z2, x3, y3, z3;
        vs.get(j, x2, y2, z2);
        vs.get(jn, x3, y3, z3);
        // edge geometry (:573-594): R0 is the foot's distance from the edge line, sv the signed distance
        // along the edge from vertex j to the line's own foot point, so vertex j lies |sv| and vertex
        // j+1 lies |r23 - sv| from it; a degenerate edge has R0 = 0 (no contribution)
        const R ex = x3 - x2, ey = y3 - y2, ez = z3 - z2;
        const R tx = xp - x2, ty = yp - y2, tz = zp - z2;
        const R r23 = M::sqrt_(ex * ex + ey * ey + ez * ez);
        const R r12 = M::sqrt_(tx * tx + ty * ty + tz * tz);
        const R r13 = M::sqrt_((xp - x3) * (xp - x3) + (yp - y3) * (yp - y3) + (zp - z3) * (zp - z3));
        const R cxx = ey * tz - ez * ty, cxy = ez * tx - ex * tz, cxz = ex * ty - ey * tx;
        R R0 = R(0), t2 = R(0), t3 = R(0);
        if (r23 > R(0)) {
            R0 = M::div(M::sqrt_(cxx * cxx + cxy * cxy + cxz * cxz), r23);
            const R sv = M::div(ex * tx + ey * ty + ez * tz, r23);
            t2 = fabs(sv);
            t3 = fabs(r23 - sv);
        } else {
            t2 = r12;  // the reference's phi = pi/2 at both ends, with R0 = 0
            t3 = r13;
        }
        // winding sign about the particle (:588-591)
        const R px = xp - x0, py = yp - y0, pz = zp - z0;
        const R ux = x2 - x0, uy = y2 - y0, uz = z2 - z0;
        const R wx = x3 - x0, wy = y3 - y0, wz = z3 - z0;
        const R s1 = px * (uy * wz - uz * wy) + py * (uz * wx - ux * wz) + pz * (ux * wy - uy * wx);
        const R phi1 = (t2 > R(0)) ? M::atan2_pos(t2, R0) : R(0);
        const R phi2 = (t3 > R(0)) ? M::atan2_pos(t3, R0) : R(0);
        // which side of the edge the foot is on: the edge-to-foot cross product against the face normal
        // (:600-611), decided only beyond a relative tolerance of 1e-9 of |N| |edge| max(1, |foot - V_j|_inf)
        const R orient = cxx * A + cxy * B + cxz * C;
        const R tmax = fmax(fmax(fabs(tx), fabs(ty)), fmax(fabs(tz), R(1)));
        const R tol = R(1e-9) * nrm * r23 * tmax;
        if (fabs(orient) > tol) {
            const bool pos = orient > R(0);
            if (!side_known) {
                first_positive = pos;
                side_known = true;
            } else if (pos != first_positive) {
                inside = false;
            }
        }
        // the edge's perpendicular foot beyond the segment: the two vertex integrals subtract (:613-628)
        const bool beyond = (t2 >= r23) || (t3 >= r23);  // r12 sin(phi1) >= r23 || r13 sin(phi2) >= r23
        min_edge = fmin(min_edge, beyond ? fmin(r12, r13) : R0);
        const R Msg = NEWELL ? ((s1 >= R(0) || r0 == R(0)) ? R(-1) : R(1)) : ((s1 * s2 * r0 <= R(0)) ? R(-1) : R(1));
        ev.edge(R0);
        // one call site for the two vertex integrals (the evaluators inline large)
        R I1 = R(0), I2 = R(0);
#pragma unroll 1
        for (int e = 0; e < 2; ++e) {
            const R I = ev.at(e ? phi2 : phi1, e ? t3 : t2);
            if (e) I2 = I;
            else I1 = I;
        }
        if (beyond)
            total += (phi1 >= phi2) ? Msg * (I1 - I2) : Msg * (I2 - I1);
        else
            total += Msg * (I1 + I2);
    }
    // closest point of the face (:642-645); a NaN distance counts as reachable, as there
    const R fd = inside ? ar0 : M::sqrt_(r0 * r0 + min_edge * min_edge);
    *reachable = !(fd >= support);
    return total;
}

// The face integral in double precision of the polygon v[0..n) (absolute xyz triples).
template <class Ev>
KIWI_FHD double face_walk_integral(double x0, double y0, double z0, double support, const double* v, int n, Ev& ev,
                                   bool* reachable) {
    return face_walk_t<double>(x0, y0, z0, support, AbsVerts{v}, n, ev, reachable);
}

}  // namespace kiwi_hip
