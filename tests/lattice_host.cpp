// Host (g++) instantiations of kiwi_amd/csrc/m4_lattice.h for tests/test_lattice_math.py.
// The device kernels compile the same header; here it runs in double and, with the host's libm in
// place of the device's hardware reciprocal / square root / log2, in single precision.
#include <cmath>

#include "../kiwi_amd/csrc/m4_lattice.h"

using namespace kiwi_hip;

namespace {

// signed lattice value U = s * W at the offset (dx, dy, dz) (units of h)
template <typename R>
R lattice_U(R dx, R dy, R dz) {
    if (dx == R(0) || dy == R(0) || dz == R(0)) return R(0);
    const R d[3] = {std::fabs(dx), std::fabs(dy), std::fabs(dz)};
    M4Plane<R> pl[3];
    R D2[3][3] = {}, D3[3][3] = {};
    for (int a = 0; a < 3; ++a) pl[a] = m4_plane<R>(d[a]);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            if (a != b) m4_line<R>(d[a], d[b], D2[a][b], D3[a][b]);
    const R W = m4_lattice_W<R>(d[0], d[1], d[2], pl, D2, D3);
    const bool neg = (dx < R(0)) != ((dy < R(0)) != (dz < R(0)));
    return neg ? -W : W;
}

// Msum of the box [lo, hi] for a particle at p with smoothing length h: the 8-corner stencil
template <typename R>
double lattice_box(const double* p, double h, const double* lo, const double* hi) {
    double m = 0.0;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
                const double cx = di ? hi[0] : lo[0], cy = dj ? hi[1] : lo[1], cz = dk ? hi[2] : lo[2];
                const double e = ((di + dj + dk) % 2 == 1) ? 1.0 : -1.0;
                m += e * static_cast<double>(lattice_U<R>(static_cast<R>((cx - p[0]) / h),
                                                          static_cast<R>((cy - p[1]) / h),
                                                          static_cast<R>((cz - p[2]) / h)));
            }
    return m;
}

}  // namespace

extern "C" {

// closed-form 6-pyramid sum and the literal sum of the 6 pyramids, both in double
void lat_free_both(double ax, double ay, double az, int k, double* out) {
    const double r = std::sqrt(ax * ax + ay * ay + az * az);
    out[0] = m4_lattice_free<double>(ax, ay, az, r, k);
    out[1] = m4_lattice_free_by_pyramid<double>(ax, ay, az, r, k);
}

// shared-polynomial line constants and the difference-of-vertex-integrals form: D2, D3, D2', D3'
void lat_line_both(double da, double db, double* out) {
    m4_line<double>(da, db, out[0], out[1]);
    m4_line_by_vertex<double>(da, db, out[2], out[3]);
}

}  // extern "C"

// the fused pair entry and the two separate offsets it sums, in double (out[0..3]) and float (out[4..7])
template <typename R>
static void line_pair(R da, R db, double* out) {
    const M4Plane<R> pa = m4_plane<R>(da), pb = m4_plane<R>(db);
    R s2, s3, ab2, ab3, ba2, ba3;
    m4_line_pair<R>(da, db, pa.c1 - pa.c2, pa.c2 - pa.c3, pb.c1 - pb.c2, pb.c2 - pb.c3, s2, s3);
    m4_line<R>(da, db, ab2, ab3);
    m4_line<R>(db, da, ba2, ba3);
    out[0] = s2;
    out[1] = s3;
    out[2] = static_cast<double>(ab2) + static_cast<double>(ba2);
    out[3] = static_cast<double>(ab3) + static_cast<double>(ba3);
}
extern "C" {

void lat_line_pair(double da, double db, double* out) {
    line_pair<double>(da, db, out);
    line_pair<float>(static_cast<float>(da), static_cast<float>(db), out + 4);
}

double lat_box(const double* p, double h, const double* lo, const double* hi, int single) {
    return single ? lattice_box<float>(p, h, lo, hi) : lattice_box<double>(p, h, lo, hi);
}

double lat_U(double dx, double dy, double dz) { return lattice_U<double>(dx, dy, dz); }

double lat_atan_poly_d(double t) { return atan_poly01<double>(t); }
float lat_atan_poly_f(float t) { return atan_poly01<float>(t); }
float lat_atan2_pos_f(float y, float x) { return KMath<float>::atan2_pos(y, x); }

}  // extern "C"

// ---- generic kernels (gen_lattice.h) ----
#include "../kiwi_amd/csrc/gen_lattice.h"

namespace {
const KiwiKernelTable kRawTabs[KIWI_TAB_NKERNELS] = {KIWI_KERNEL_TABLE_DATA};

struct GenTabs {
    GenTab t[KIWI_TAB_NKERNELS];
    GenTabs() {
        for (int k = 0; k < KIWI_TAB_NKERNELS; ++k) gen_tab_init(t[k], kRawTabs[k]);
    }
};
const GenTabs kGen;
const GenTab* const kTabs = kGen.t;

double gen_U(const GenTab& K, double dx, double dy, double dz) {
    if (dx == 0.0 || dy == 0.0 || dz == 0.0) return 0.0;
    const double W = gen_lattice_W(K, std::fabs(dx), std::fabs(dy), std::fabs(dz));
    const bool neg = (dx < 0.0) != ((dy < 0.0) != (dz < 0.0));
    return neg ? -W : W;
}
}  // namespace

extern "C" {

// Msum of the box [lo, hi] for kernel kid with support H: the 8-corner stencil of gen_lattice_W
double lat_gen_box(int kid, const double* p, double H, const double* lo, const double* hi) {
    const GenTab& K = kTabs[kid];
    double m = 0.0;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
                const double cx = di ? hi[0] : lo[0], cy = dj ? hi[1] : lo[1], cz = dk ? hi[2] : lo[2];
                const double e = ((di + dj + dk) % 2 == 1) ? 1.0 : -1.0;
                m += e * gen_U(K, (cx - p[0]) / H, (cy - p[1]) / H, (cz - p[2]) / H);
            }
    return m;
}

}  // extern "C"

extern "C" double lat_gen_W(int kid, double ax, double ay, double az) { return gen_lattice_W(kTabs[kid], ax, ay, az); }
// the same lattice value in single precision (the mixed-precision generic path's arithmetic)
extern "C" double lat_gen_W32(int kid, double ax, double ay, double az) {
    return gen_lattice_W<float>(kTabs[kid], static_cast<float>(ax), static_cast<float>(ay), static_cast<float>(az));
}

// ---- tabulated cubic path (m4_table.h) ----
#include <vector>

#include "../kiwi_amd/csrc/m4_table.h"

namespace {
std::vector<double>& host_table() {
    static std::vector<double> t;
    if (t.empty()) {
        using G = M4TableGeom;
        t.assign(static_cast<size_t>(G::size), 0.0);
        for (int i = 1; i <= G::nr0; ++i)
            for (int j = 0; j < G::nR0; ++j) {
                const double r0 = G::r0_of(i), mu0 = G::mu0_of(j);
                const double R0 = r0 * std::sqrt(1.0 - mu0 * mu0) / mu0;  // j = 0: +inf
                for (int k = 0; k < G::nph; ++k) t[G::at(i, j, k)] = m4_vertex_value(r0, R0, std::acos(G::cphi_of(k)));
            }
        for (int j = 0; j < G::nR0; ++j)
            for (int k = 0; k < G::nph; ++k) t[G::at(G::nr0 + 1, j, k)] = t[G::at(G::nr0, j, k)];
    }
    return t;
}
}  // namespace

extern "C" {
double lat_m4_vertex_value(double h, double phi, double r0, double R0) {
    return m4_vertex_value(r0 / h, R0 / h, phi);  // dimensionless: the same at every h
}
double lat_m4_gridded(double h, double phi, double r0, double R0) {
    if (phi == 0.0) return 0.0;
    return m4_gridded_value(host_table().data(), r0 / h, R0 / h, std::cos(phi));
}
}  // extern "C"

// ---- the polygon face walk of the Voronoi map (face_walk.h) ----
#include "../kiwi_amd/csrc/face_walk.h"

extern "C" {
// face integral and reachability of polygon v[0..n) for a particle at p: kind 0 = cubic analytic,
// 1 = generic kernel kid (support sup * h), 2 = cubic tabulated
double lat_face_walk(int kind, int kid, double sup, const double* p, double h, const double* v, int n, int* reach) {
    bool r = false;
    double m = 0.0;
    if (kind == 0) {
        EvM4 ev{1.0 / h};
        m = face_walk_integral(p[0], p[1], p[2], 2.0 * h, v, n, ev, &r);
    } else if (kind == 1) {
        EvGen<GenTab> ev{&kTabs[kid], 1.0 / (sup * h)};
        m = face_walk_integral(p[0], p[1], p[2], sup * h, v, n, ev, &r);
    } else {
        EvTab ev{host_table().data(), 1.0 / h};
        m = face_walk_integral(p[0], p[1], p[2], 2.0 * h, v, n, ev, &r);
    }
    *reach = r ? 1 : 0;
    return m;
}

// the cubic spline's face integral in single precision on vertex coordinates relative to the particle
// (the Voronoi map's mixed mode)
double lat_face_walk32(const double* p, double h, const double* v, int n) {
    bool r = false;
    EvM4T<float> ev{static_cast<float>(1.0 / h)};
    const RelVerts<float> vs{v, p[0], p[1], p[2]};
    return face_walk_t<float>(0.0f, 0.0f, 0.0f, static_cast<float>(2.0 * h), vs, n, ev, &r);
}
}  // extern "C"

// ---- the adjoint's lattice gradient (m4_adjoint.h) ----
#include "../kiwi_amd/csrc/m4_adjoint.h"

extern "C" {
// W, dW/dax, dW/day, dW/daz at absolute offsets (units of h), in double (single = 0) or float
void lat_adj_point(double ax, double ay, double az, int single, double* out) {
    if (single) {
        const float fx = static_cast<float>(ax), fy = static_cast<float>(ay), fz = static_cast<float>(az);
        const AdjLine<float> lxy = m4_adj_line<float>(fx, fy), lxz = m4_adj_line<float>(fx, fz),
                             lyz = m4_adj_line<float>(fy, fz);
        float W, Wx, Wy, Wz;
        m4_adj_point<float>(fx, fy, fz, lxy, lxz, lyz, W, Wx, Wy, Wz);
        out[0] = W;
        out[1] = Wx;
        out[2] = Wy;
        out[3] = Wz;
        return;
    }
    const AdjLine<double> lxy = m4_adj_line(ax, ay), lxz = m4_adj_line(ax, az), lyz = m4_adj_line(ay, az);
    m4_adj_point<double>(ax, ay, az, lxy, lxz, lyz, out[0], out[1], out[2], out[3]);
}
}  // extern "C"

// ---- the generic kernels' lattice gradient (gen_adjoint.h) ----
#include "../kiwi_amd/csrc/gen_adjoint.h"

extern "C" {
// W, dW/dax, dW/day, dW/daz at absolute offsets (units of the support H) for kernel kid, in double
// (single = 0) or float, with the region and line entries as the adjoint kernel forms them
void lat_gen_adj_point(int kid, double ax, double ay, double az, int single, double* out) {
    const GenTab& K = kTabs[kid];
    if (single) {
        const float fx = static_cast<float>(ax), fy = static_cast<float>(ay), fz = static_cast<float>(az);
        const int k = gen_region_of(K, static_cast<double>(std::sqrt(fx * fx + fy * fy + fz * fz)));
        const AdjLineGen<float> lxy = gen_adj_line<float>(K, fx, fy), lxz = gen_adj_line<float>(K, fx, fz),
                                lyz = gen_adj_line<float>(K, fy, fz);
        float W, Wx, Wy, Wz;
        gen_adj_point<float>(K, k, fx, fy, fz, lxy, lxz, lyz, W, Wx, Wy, Wz);
        out[0] = W;
        out[1] = Wx;
        out[2] = Wy;
        out[3] = Wz;
        return;
    }
    const int k = gen_region_of(K, std::sqrt(ax * ax + ay * ay + az * az));
    const AdjLineGen<double> lxy = gen_adj_line(K, ax, ay), lxz = gen_adj_line(K, ax, az), lyz = gen_adj_line(K, ay, az);
    gen_adj_point<double>(K, k, ax, ay, az, lxy, lxz, lyz, out[0], out[1], out[2], out[3]);
}
}  // extern "C"
