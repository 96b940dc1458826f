// Forward-mode dual numbers for the adjoint: a value and N partial derivatives, carried through the
// templated lattice mathematics of m4_lattice.h (KMath<Dual<T, N>> supplies the elementary functions
// with their exact derivatives).  Comparisons look at the value only, so the region and guard branches
// of the lattice code pick the same branch as the plain evaluation.  All loops run over the
// compile-time N and unroll, so a Dual stays in registers.
#pragma once

#include <cmath>

#include "m4_lattice.h"

namespace kiwi_hip {

template <typename T, int N>
struct Dual {
    T v;
    T d[N];
    KIWI_HD constexpr Dual() : v(0), d{} {}
    KIWI_HD constexpr Dual(T x) : v(x), d{} {}  // constants
    // the i-th independent variable at x
    static KIWI_HD Dual var(T x, int i) {
        Dual r(x);
#pragma unroll
        for (int k = 0; k < N; ++k) r.d[k] = (k == i) ? T(1) : T(0);
        return r;
    }
};

#define KIWI_DUAL_LOOP _Pragma("unroll") for (int k = 0; k < N; ++k)

template <typename T, int N>
KIWI_HD Dual<T, N> operator+(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r(a.v + b.v);
    KIWI_DUAL_LOOP r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <typename T, int N>
KIWI_HD Dual<T, N> operator-(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r(a.v - b.v);
    KIWI_DUAL_LOOP r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <typename T, int N>
KIWI_HD Dual<T, N> operator-(const Dual<T, N>& a) {
    Dual<T, N> r(-a.v);
    KIWI_DUAL_LOOP r.d[k] = -a.d[k];
    return r;
}
template <typename T, int N>
KIWI_HD Dual<T, N> operator*(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r(a.v * b.v);
    KIWI_DUAL_LOOP r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <typename T, int N>
KIWI_HD Dual<T, N> operator/(const Dual<T, N>& a, const Dual<T, N>& b) {
    const T inv = T(1) / b.v;
    const T q = a.v * inv;
    Dual<T, N> r(q);
    KIWI_DUAL_LOOP r.d[k] = (a.d[k] - q * b.d[k]) * inv;
    return r;
}
template <typename T, int N>
KIWI_HD Dual<T, N>& operator+=(Dual<T, N>& a, const Dual<T, N>& b) {
    a = a + b;
    return a;
}
template <typename T, int N>
KIWI_HD Dual<T, N>& operator-=(Dual<T, N>& a, const Dual<T, N>& b) {
    a = a - b;
    return a;
}
template <typename T, int N>
KIWI_HD Dual<T, N>& operator*=(Dual<T, N>& a, const Dual<T, N>& b) {
    a = a * b;
    return a;
}
template <typename T, int N>
KIWI_HD bool operator<(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v < b.v; }
template <typename T, int N>
KIWI_HD bool operator>(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v > b.v; }
template <typename T, int N>
KIWI_HD bool operator<=(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v <= b.v; }
template <typename T, int N>
KIWI_HD bool operator>=(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v >= b.v; }

template <typename T, int N>
KIWI_HD T kval(const Dual<T, N>& x) {
    return x.v;
}
template <typename T, int N>
KIWI_HD bool operator!=(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v != b.v; }
template <typename T, int N>
KIWI_HD bool operator==(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v == b.v; }

// unary function with value f and derivative df at the argument
template <typename T, int N>
KIWI_HD Dual<T, N> chain(const Dual<T, N>& x, T f, T df) {
    Dual<T, N> r(f);
    KIWI_DUAL_LOOP r.d[k] = df * x.d[k];
    return r;
}

template <typename T, int N>
struct M4Consts<Dual<T, N>> {
    static constexpr Dual<T, N> pi = Dual<T, N>(M4Consts<T>::pi);
    static constexpr Dual<T, N> half_pi = Dual<T, N>(M4Consts<T>::half_pi);
    static constexpr Dual<T, N> inv_pi = Dual<T, N>(M4Consts<T>::inv_pi);
};

template <typename T, int N>
struct KMath<Dual<T, N>> {
    using D = Dual<T, N>;
    using M = KMath<T>;
    static KIWI_HD D rcp(const D& x) {
        const T i = M::rcp(x.v);
        return chain(x, i, -i * i);
    }
    static KIWI_HD D div(const D& a, const D& b) { return a * rcp(b); }
    // at 0 the derivative is taken as 0 rather than infinity (m4_line keeps its square roots away from
    // 0; this only guards against NaN from 0 * inf)
    static KIWI_HD D sqrt_(const D& x) {
        const T s = M::sqrt_(x.v);
        return chain(x, s, s > T(0) ? T(0.5) * M::rcp(s) : T(0));
    }
    static KIWI_HD D log_(const D& x) { return chain(x, M::log_(x.v), M::rcp(x.v)); }
    static KIWI_HD D atan_pos(const D& x) { return chain(x, M::atan_pos(x.v), M::rcp(T(1) + x.v * x.v)); }
    // atan2(y, x) for y, x >= 0: d = (x dy - y dx) / (x^2 + y^2)
    static KIWI_HD D atan2_pos(const D& y, const D& x) {
        const T inv = M::rcp(x.v * x.v + y.v * y.v);
        D r(M::atan2_pos(y.v, x.v));
        KIWI_DUAL_LOOP r.d[k] = (x.v * y.d[k] - y.v * x.d[k]) * inv;
        return r;
    }
};

#undef KIWI_DUAL_LOOP

}  // namespace kiwi_hip
