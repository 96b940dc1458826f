// The cubic-spline lattice value W (m4_lattice.h) with its gradient, for the adjoint of the Cartesian
// sph-to-grid.  dW/d|d_a| is the kernel's flux through the lattice box's face at offset d_a (W is the
// kernel mass of the box [0, |dx|] x [0, |dy|] x [0, |dz|]); it is obtained exactly, by carrying dual
// numbers (dual.h) through the same closed form the forward evaluates, so the adjoint is the derivative
// of the computed forward and not of an approximation to it.
//
// The region-matching constants of the edge lines are tabulated per particle, as in the forward, with
// their partial derivatives: AdjLine holds, for an unordered axis pair {a, b} at absolute offsets
// (d_a, d_b) and each of the slots D2 and D3, the summed offsets of the two directed lines (wall a with
// edge normal b, and wall b with edge normal a) and their derivatives by d_a and by d_b.  Host (tests)
// and device code.
#pragma once

#include "dual.h"
#include "m4_lattice.h"

namespace kiwi_hip {

template <typename T>
struct AdjLine {
    T v[2][3];  // [slot D2, D3][value, d/d d_a, d/d d_b]
};

template <typename T>
KIWI_HD AdjLine<T> m4_adj_line(T da, T db) {
    using D = Dual<T, 2>;
    const D A = D::var(da, 0), B = D::var(db, 1);
    D d2ab, d3ab, d2ba, d3ba;
    m4_line<D>(A, B, d2ab, d3ab);
    m4_line<D>(B, A, d2ba, d3ba);
    const D s2 = d2ab + d2ba, s3 = d3ab + d3ba;
    AdjLine<T> e;
    e.v[0][0] = s2.v;
    e.v[0][1] = s2.d[0];
    e.v[0][2] = s2.d[1];
    e.v[1][0] = s3.v;
    e.v[1][1] = s3.d[0];
    e.v[1][2] = s3.d[1];
    return e;
}

// plane term c_k(d) of region k with its derivative
template <typename T>
KIWI_HD void m4_adj_plane(T d, int k, T& c, T& dc) {
    using D = Dual<T, 1>;
    const M4Plane<D> pl = m4_plane<D>(D::var(d, 0));
    const D ck = (k == 1) ? pl.c1 : (k == 2 ? pl.c2 : pl.c3);
    c = ck.v;
    dc = ck.d[0];
}

// W at absolute offsets (ax, ay, az) > 0 (units of h) and dW/dax, dW/day, dW/daz; lxy, lxz, lyz are the
// line entries of the three axis pairs at this point
template <typename T>
KIWI_HD void m4_adj_point(T ax, T ay, T az, const AdjLine<T>& lxy, const AdjLine<T>& lxz, const AdjLine<T>& lyz, T& W,
                          T& Wx, T& Wy, T& Wz) {
    using D = Dual<T, 3>;
    const int k = m4_region(ax * ax + ay * ay + az * az);
    const D X = D::var(ax, 0), Y = D::var(ay, 1), Z = D::var(az, 2);
    const D R = KMath<D>::sqrt_(X * X + Y * Y + Z * Z);
    const D f = m4_lattice_free<D>(X, Y, Z, R, k);
    const T ip = M4Consts<T>::inv_pi;
    W = f.v * ip;
    Wx = f.d[0] * ip;
    Wy = f.d[1] * ip;
    Wz = f.d[2] * ip;
    T c, dc;
    m4_adj_plane(ax, k, c, dc);
    W += T(0.5) * c;
    Wx += T(0.5) * dc;
    m4_adj_plane(ay, k, c, dc);
    W += T(0.5) * c;
    Wy += T(0.5) * dc;
    m4_adj_plane(az, k, c, dc);
    W += T(0.5) * c;
    Wz += T(0.5) * dc;
    if (k >= 2) {
        const int s = k - 2;
        W += lxy.v[s][0] + lxz.v[s][0] + lyz.v[s][0];
        Wx += lxy.v[s][1] + lxz.v[s][1];
        Wy += lxy.v[s][2] + lyz.v[s][1];
        Wz += lxz.v[s][2] + lyz.v[s][2];
    }
}

}  // namespace kiwi_hip
