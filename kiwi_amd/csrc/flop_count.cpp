// Algorithmic FLOP count of the lattice kernel (m4_lattice.h), per unit of work, by instantiating the
// same templates with a counting scalar.  Convention: every +, -, *, /, sqrt and transcendental call
// (atan, atan2, log) counts 1 FLOP (so an FMA counts 2); comparisons, selects, fabs and integer
// work count 0.  Transcendental calls are also tallied on their own.
//
// Units (one launch of the gather processes a measured number of each; bench.py multiplies):
//   point[k]  one lattice value W in radial region k = 1, 2, 3 (inside h, inside 2h, outside)
//   line[j]   one edge-line constant pair with j = 0, 1, 2 crossing evaluations
//   plane     one wall-plane constant triple
//   cell      one (particle, cell) stencil + accumulation (cells inside the support sphere)
//
// Build + run:  g++ -O1 -std=c++17 kiwi_amd/csrc/flop_count.cpp -o /tmp/fc && /tmp/fc > kiwi_amd/flops_per_unit.json
#include <cmath>
#include <cstdio>

static long g_flops = 0, g_trans = 0;

struct CF {
    double v;
    constexpr CF() : v(0) {}
    constexpr CF(double x) : v(x) {}
};
inline CF operator+(CF a, CF b) { ++g_flops; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { ++g_flops; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { ++g_flops; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { ++g_flops; return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline CF& operator+=(CF& a, CF b) { a = a + b; return a; }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline bool operator<=(CF a, CF b) { return a.v <= b.v; }
inline bool operator>(CF a, CF b) { return a.v > b.v; }
inline bool operator>=(CF a, CF b) { return a.v >= b.v; }
inline bool operator!=(CF a, CF b) { return a.v != b.v; }
inline CF sqrt(CF a) { ++g_flops; ++g_trans; return CF(std::sqrt(a.v)); }
inline CF atan(CF a) { ++g_flops; ++g_trans; return CF(std::atan(a.v)); }
inline CF atan2(CF a, CF b) { ++g_flops; ++g_trans; return CF(std::atan2(a.v, b.v)); }
inline CF log(CF a) { ++g_flops; ++g_trans; return CF(std::log(a.v)); }
inline CF fabs(CF a) { return CF(std::fabs(a.v)); }

#include "m4_lattice.h"

using namespace kiwi_hip;

// the counting scalar follows the double-precision arithmetic (one division per reciprocal)
namespace kiwi_hip {
template <>
struct KMath<CF> {
    static CF rcp(CF x) { return CF(1.0) / x; }
    static CF div(CF a, CF b) { return a / b; }
    static CF sqrt_(CF x) { return sqrt(x); }
    static CF log_(CF x) { return log(x); }
    static CF atan_pos(CF x) { return atan(x); }
    static CF atan2_pos(CF y, CF x) { return atan2(y, x); }
};
}  // namespace kiwi_hip

struct Count {
    long f, t;
};

template <class F>
Count measure(F fn) {
    g_flops = g_trans = 0;
    fn();
    return {g_flops, g_trans};
}

int main() {
    // lattice value at one point in each region: the kernel's phase-B body
    Count pt[3];
    const double pts[3][3] = {{0.3, 0.4, 0.5}, {0.8, 0.9, 0.7}, {1.5, 1.6, 1.7}};
    for (int r = 0; r < 3; ++r) {
        pt[r] = measure([&] {
            CF ax(pts[r][0]), ay(pts[r][1]), az(pts[r][2]);
            CF r2 = ax * ax + ay * ay + az * az;
            CF rr = KMath<CF>::sqrt_(r2);
            int k = m4_region(r2);
            CF csum = CF(0.1) + CF(0.2) + CF(0.3);
            CF dsum = (k == 1) ? CF(0) : CF(0.1) + CF(0.1) + CF(0.1) + CF(0.1) + CF(0.1) + CF(0.1);
            // the kernel's free part (PolicyM4::free_part): a constant outside the support
            CF f = (k == 3) ? CF(-0.25) : m4_lattice_free<CF>(ax, ay, az, rr, k) * M4Consts<CF>::inv_pi;
            CF W = f + CF(0.5) * csum + dsum;
            (void)W;
        });
    }
    // edge-line constants: 0, 1 or 2 crossings (the offset computation (X - p) / h included)
    Count ln[3];
    const double lines[3][2] = {{1.8, 1.9}, {1.2, 0.9}, {0.5, 0.6}};
    for (int j = 0; j < 3; ++j) {
        ln[j] = measure([&] {
            CF da = (CF(1.0) + CF(2.0) * CF(0.5) - CF(0.2)) * CF(0.7);
            CF db = (CF(1.0) + CF(2.0) * CF(0.5) - CF(0.2)) * CF(0.7);
            da = CF(lines[j][0]);
            db = CF(lines[j][1]);
            CF l2 = da * da + db * db;  // branch classification
            CF D2, D3;
            m4_line<CF>(da, db, D2, D3);
            (void)l2;
        });
    }
    Count pl = measure([&] {
        CF d = (CF(1.0) + CF(2.0) * CF(0.5) - CF(0.2)) * CF(0.7);
        M4Plane<CF> p = m4_plane<CF>(fabs(d) * CF(0.5));
        (void)p;
    });
    // per (particle, cell): box-sphere distance, 8-corner stencil, clamp, 4 accumulations
    Count cell = measure([&] {
        CF ex(0.1), ey(0.2), ez(0.3);
        CF dist2 = ex * ex + ey * ey + ez * ez;
        CF Ms = CF(1) + CF(1) + CF(1) + CF(1) - CF(1) - CF(1) - CF(1) - CF(1);
        CF t = Ms * CF(2.0);
        CF a0 = CF(0) + t, a1 = CF(0) + t * CF(1), a2 = CF(0) + t * CF(1), a3 = CF(0) + t * CF(1);
        (void)dist2; (void)a0; (void)a1; (void)a2; (void)a3;
    });
    std::printf("{\n  \"convention\": \"+ - * / sqrt atan atan2 log = 1 FLOP each; FMA = 2\",\n");
    std::printf("  \"point\": [%ld, %ld, %ld],\n  \"point_transcendentals\": [%ld, %ld, %ld],\n", pt[0].f, pt[1].f,
                pt[2].f, pt[0].t, pt[1].t, pt[2].t);
    std::printf("  \"line\": [%ld, %ld, %ld],\n  \"line_transcendentals\": [%ld, %ld, %ld],\n", ln[0].f, ln[1].f,
                ln[2].f, ln[0].t, ln[1].t, ln[2].t);
    std::printf("  \"plane\": %ld,\n  \"cell\": %ld\n}\n", pl.f, cell.f);
    return 0;
}
