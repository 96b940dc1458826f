// The generic kernels' lattice value W (gen_lattice.h) with its gradient, for the adjoint of the
// Cartesian sph-to-grid on the generic path (M4 forced generic or with another support, M5, M6,
// Wendland C2/C4/C6).  As for the cubic spline (m4_adjoint.h), dual numbers (dual.h) are carried through
// the same closed form the forward evaluates, so the gradient is that of the computed forward.
// Offsets are in units of the support H.
//
// AdjLineGen holds, for an unordered axis pair {a, b} at absolute offsets (d_a, d_b), each region's
// summed offsets of the two directed edge lines (the forward's line-table entry) with their derivatives
// by d_a and by d_b.  Host (tests) and device code.
#pragma once

#include "dual.h"
#include "gen_lattice.h"

namespace kiwi_hip {

template <typename T, int N>
struct GenCoef<Dual<T, N>> {
    using type = T;
};

template <typename T>
struct AdjLineGen {
    T v[GEN_NREG][3];  // [region slot][value, d/d d_a, d/d d_b]
};

template <typename T>
KIWI_GHD AdjLineGen<T> gen_adj_line(const GenTab& K, T da, T db) {
    using D = Dual<T, 2>;
    const D A = D::var(da, 0), B = D::var(db, 1);
    D eab[GEN_NREG], eba[GEN_NREG];
    gen_line<D>(K, A, B, eab);
    gen_line<D>(K, B, A, eba);
    AdjLineGen<T> e;
#pragma unroll
    for (int k = 0; k < GEN_NREG; ++k) {
        const D s = eab[k] + eba[k];
        e.v[k][0] = s.v;
        e.v[k][1] = s.d[0];
        e.v[k][2] = s.d[1];
    }
    return e;
}

// Phi_k(d) of region k (1-based) with its derivative
template <typename T>
KIWI_GHD void gen_adj_plane(const GenTab& K, T d, int k, T& phi, T& dphi) {
    using D = Dual<T, 1>;
    D P[GEN_NREG];
    gen_plane_phi(K, D::var(d, 0), P);
    phi = P[k - 1].v;
    dphi = P[k - 1].d[0];
}

// W at absolute offsets (ax, ay, az) > 0 (units of H) in region k and dW/dax, dW/day, dW/daz; lxy, lxz,
// lyz are the line entries of the three axis pairs at this point.  KID >= 0: K is gen_fixed_tab<KID>(),
// whose free part is evaluated by gen_lattice_free_fixed (coefficients folded into the code).
template <typename T, int KID = -1>
KIWI_GHD void gen_adj_point(const GenTab& K, int k, T ax, T ay, T az, const AdjLineGen<T>& lxy,
                            const AdjLineGen<T>& lxz, const AdjLineGen<T>& lyz, T& W, T& Wx, T& Wy, T& Wz) {
    using D = Dual<T, 3>;
    const D X = D::var(ax, 0), Y = D::var(ay, 1), Z = D::var(az, 2);
    const D R = KMath<D>::sqrt_(X * X + Y * Y + Z * Z);
    D f;
    if constexpr (KID >= 0)
        f = gen_lattice_free_fixed<D, KID>(X, Y, Z, R, k);
    else
        f = gen_lattice_free<D>(gen_regions<T>(K)[k - 1], X, Y, Z, R);
    W = f.v;
    Wx = f.d[0];
    Wy = f.d[1];
    Wz = f.d[2];
    const T half_pi = T(0.5 * GEN_PI);
    T p, dp;
    gen_adj_plane(K, ax, k, p, dp);
    W += half_pi * p;
    Wx += half_pi * dp;
    gen_adj_plane(K, ay, k, p, dp);
    W += half_pi * p;
    Wy += half_pi * dp;
    gen_adj_plane(K, az, k, p, dp);
    W += half_pi * p;
    Wz += half_pi * dp;
    const int s = k - 1;  // the line-table slot of region k (region 1's offsets are zero)
    W += lxy.v[s][0] + lxz.v[s][0] + lyz.v[s][0];
    Wx += lxy.v[s][1] + lxz.v[s][1];
    Wy += lxy.v[s][2] + lyz.v[s][1];
    Wz += lxz.v[s][2] + lyz.v[s][2];
}

}  // namespace kiwi_hip
