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
// vertex.  Double precision; host (tests) and device code.
#pragma once

#include <cmath>

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
struct EvM4 {
    double inv_h;
    double da = 0.0, db = 0.0, D2 = 0.0, D3 = 0.0;
    M4Plane<double> pl{};
    KIWI_FHD void wall(double ar0) {
        da = ar0 * inv_h;
        pl = m4_plane<double>(da);
    }
    KIWI_FHD void edge(double R0) {
        db = R0 * inv_h;
        m4_line<double>(da, db, D2, D3);
    }
    KIWI_FHD double at(double phi, double t) const {
        if (!(da > 0.0) || !(db > 0.0) || phi == 0.0) return 0.0;
        const double dc = t * inv_h;  // = db tan(phi)
        const double r2 = da * da + db * db + dc * dc;
        const double r = sqrt(r2);
        const int k = m4_region(r2);
        const double T = atan(da * dc / (db * r));
        const double L = log((r + dc) * (r + dc) / (da * da + db * db));
        const double D = (k == 1) ? 0.0 : (k == 2 ? D2 : D3);
        return m4_vertex_g<double>(k, da, db, dc, r, T, L, phi, pl) + D;
    }
};

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

// The face integral of the polygon v[0..n) (xyz triples) for a particle at (x0, y0, z0) with kernel
// support radius `support`; *reachable is set when the face's closest point lies within the support.
template <class Ev>
KIWI_FHD double face_walk_integral(double x0, double y0, double z0, double support, const double* v, int n, Ev& ev,
                                   bool* reachable) {
    *reachable = false;
    if (n < 3) return 0.0;
    // plane through the first three vertices (:495-506)
    const double x1 = v[0], y1 = v[1], z1 = v[2];
    const double x2s = v[3], y2s = v[4], z2s = v[5];
    const double x3s = v[6], y3s = v[7], z3s = v[8];
    const double A = (y2s - y1) * (z3s - z1) - (y3s - y1) * (z2s - z1);
    const double B = (z2s - z1) * (x3s - x1) - (z3s - z1) * (x2s - x1);
    const double C = (x2s - x1) * (y3s - y1) - (x3s - x1) * (y2s - y1);
    const double D = -A * x1 - B * y1 - C * z1;
    const double nrm = sqrt(A * A + B * B + C * C);
    const double r0 = (A * x0 + B * y0 + C * z0 + D) / nrm;
    const double ar0 = fabs(r0);
    const double xp = x0 - r0 * A / nrm, yp = y0 - r0 * B / nrm, zp = z0 - r0 * C / nrm;
    // winding sign reference: triple product about the particle (:525-529)
    const double ax = x1 - x0, ay = y1 - y0, az = z1 - z0;
    const double bx = x2s - x0, by = y2s - y0, bz = z2s - z0;
    const double cx = x3s - x0, cy = y3s - y0, cz = z3s - z0;
    const double s2 = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    ev.wall(ar0);

    bool inside = true, have_sign = false;
    double sign_ref = 0.0, min_edge = INFINITY, total = 0.0;
    for (int j = 0; j < n; ++j) {
        const int jn = (j == n - 1) ? 0 : j + 1;
        const double x2 = v[3 * j], y2 = v[3 * j + 1], z2 = v[3 * j + 2];
        const double x3 = v[3 * jn], y3 = v[3 * jn + 1], z3 = v[3 * jn + 2];
        // edge geometry (:573-594): R0 is the foot's distance from the edge line, sv the signed distance
        // along the edge from vertex j to the line's own foot point, so vertex j lies |sv| and vertex
        // j+1 lies |r23 - sv| from it; a degenerate edge has R0 = 0 (no contribution)
        const double ex = x3 - x2, ey = y3 - y2, ez = z3 - z2;
        const double tx = xp - x2, ty = yp - y2, tz = zp - z2;
        const double r23 = sqrt(ex * ex + ey * ey + ez * ez);
        const double r12 = sqrt(tx * tx + ty * ty + tz * tz);
        const double r13 = sqrt((xp - x3) * (xp - x3) + (yp - y3) * (yp - y3) + (zp - z3) * (zp - z3));
        const double cxx = ey * tz - ez * ty, cxy = ez * tx - ex * tz, cxz = ex * ty - ey * tx;
        double R0 = 0.0, t2 = 0.0, t3 = 0.0;
        if (r23 > 0.0) {
            R0 = sqrt(cxx * cxx + cxy * cxy + cxz * cxz) / r23;
            const double sv = (ex * tx + ey * ty + ez * tz) / r23;
            t2 = fabs(sv);
            t3 = fabs(r23 - sv);
        } else {
            t2 = r12;  // the reference's phi = pi/2 at both ends, with R0 = 0
            t3 = r13;
        }
        // winding sign about the particle (:588-591)
        const double px = xp - x0, py = yp - y0, pz = zp - z0;
        const double ux = x2 - x0, uy = y2 - y0, uz = z2 - z0;
        const double wx = x3 - x0, wy = y3 - y0, wz = z3 - z0;
        const double s1 = px * (uy * wz - uz * wy) + py * (uz * wx - ux * wz) + pz * (ux * wy - uy * wx);
        const double phi1 = (t2 > 0.0) ? atan2(t2, R0) : 0.0;
        const double phi2 = (t3 > 0.0) ? atan2(t3, R0) : 0.0;
        // is the foot inside the polygon (:600-611)
        const double side = cxx * A + cxy * B + cxz * C;
        const double side_eps = 1e-9 * nrm * r23 * fmax(fmax(fabs(tx), fabs(ty)), fmax(fabs(tz), 1.0));
        if (fabs(side) > side_eps) {
            if (!have_sign) {
                sign_ref = side;
                have_sign = true;
            } else if ((side > 0.0) != (sign_ref > 0.0)) {
                inside = false;
            }
        }
        // the edge's perpendicular foot beyond the segment: the two vertex integrals subtract (:613-628)
        const bool beyond = (t2 >= r23) || (t3 >= r23);  // r12 sin(phi1) >= r23 || r13 sin(phi2) >= r23
        min_edge = fmin(min_edge, beyond ? fmin(r12, r13) : R0);
        const double M = (s1 * s2 * r0 <= 0.0) ? -1.0 : 1.0;
        ev.edge(R0);
        // one call site for the two vertex integrals (the evaluators inline large)
        double I1 = 0.0, I2 = 0.0;
#pragma unroll 1
        for (int e = 0; e < 2; ++e) {
            const double I = ev.at(e ? phi2 : phi1, e ? t3 : t2);
            if (e) I2 = I;
            else I1 = I;
        }
        if (beyond)
            total += (phi1 >= phi2) ? M * (I1 - I2) : M * (I2 - I1);
        else
            total += M * (I1 + I2);
    }
    // closest point of the face (:642-645); a NaN distance counts as reachable, as there
    const double fd = inside ? ar0 : sqrt(r0 * r0 + min_edge * min_edge);
    *reachable = !(fd >= support);
    return total;
}

}  // namespace kiwi_hip
