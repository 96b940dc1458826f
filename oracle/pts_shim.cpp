// C-ABI shim over the reference's coverage-grid buffer (add_coverage_grid_buffer, src/point_set.cpp:
// 208-273), its candidate lattice (generate_cartesian_lattice, :166-195) and its snapshot readers
// (load_sph_snapshot and detect_snapshot_format, src/snapshot.cpp:495-522), linked with the reference's
// own objects into oracle/_ref/libkiwi_pts.so.  Test infrastructure: it runs the reference function on
// a plain-array point set so the oracle's restatement and the GPU buffer can be checked against it and
// timed beside it.  The library is git-ignored but travels with the working tree to the GPU machine,
// where the coverage-buffer GPU tests load it on the host as their checker; libkiwi_hip never links or
// calls it.
#include <algorithm>
#include <cstring>
#include <exception>
#include <vector>

#include "kiwi/config.hpp"
#include "kiwi/point_set.hpp"
#include "kiwi/snapshot.hpp"

extern "C" {

// Lattice size of the buffer's candidates: nx, ny, nz and the spacing.
int ref_cartesian_lattice(const double* dom, double spacing_fraction, int* n3, double* spacing) {
    kiwi::Domain d;
    d.x_min = dom[0], d.x_max = dom[1], d.y_min = dom[2], d.y_max = dom[3], d.z_min = dom[4], d.z_max = dom[5];
    const kiwi::CartesianLattice lat = kiwi::generate_cartesian_lattice(d, spacing_fraction);
    n3[0] = lat.nx, n3[1] = lat.ny, n3[2] = lat.nz;
    *spacing = lat.spacing;
    return static_cast<int>(lat.points.size());
}

// The buffer sites the reference appends to n particles (ids 1..n): up to cap of them into ox, oy, oz,
// oid; returns their count, or -1 with the exception's message in err.
long long ref_coverage_buffer(long long n, const double* x, const double* y, const double* z, const double* h,
                              const double* dom, double spacing_fraction, double factor, double* ox, double* oy,
                              double* oz, int* oid, long long cap, char* err, int errlen) {
    try {
        kiwi::PointSet p;
        for (long long i = 0; i < n; ++i) p.append(static_cast<int>(i + 1), x[i], y[i], z[i]);
        const std::vector<double> hv(h, h + n);
        kiwi::Domain d;
        d.x_min = dom[0], d.x_max = dom[1], d.y_min = dom[2], d.y_max = dom[3], d.z_min = dom[4], d.z_max = dom[5];
        kiwi::CoverageGridBufferParams prm;
        prm.spacing_fraction = spacing_fraction;
        prm.coverage_radius_factor = factor;
        const kiwi::PointSet out = kiwi::add_coverage_grid_buffer(p, hv, d, prm);
        const long long k = static_cast<long long>(out.size()) - n;
        for (long long j = 0; j < std::min(k, cap); ++j) {
            ox[j] = out.x[n + j];
            oy[j] = out.y[n + j];
            oz[j] = out.z[n + j];
            oid[j] = out.id[n + j];
        }
        return k;
    } catch (const std::exception& e) {
        if (errlen > 0) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return -1;
    }
}

// The reference's snapshot reader (units3: PhantomUnits length_cm, mass_g, velocity_cms, or NULL for
// the defaults): the row count (up to cap rows into cols, 9 columns x cap, and ids), or -1 with the
// exception's message in err.
long long ref_load_snapshot(const char* path, const char* format, double hfact, const double* units3, int allow_missing,
                            double* cols, int* ids, long long cap, char* err, int errlen) {
    try {
        kiwi::PhantomUnits u;
        if (units3) u.length_cm = units3[0], u.mass_g = units3[1], u.velocity_cms = units3[2];
        const kiwi::SPHSnapshot s = kiwi::load_sph_snapshot(path, format, hfact, u, allow_missing != 0);
        const long long n = static_cast<long long>(s.size());
        const std::vector<double>* c[9] = {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.m, &s.h, &s.rho};
        for (long long i = 0; i < std::min(n, cap); ++i) {
            ids[i] = s.id[i];
            for (int k = 0; k < 9; ++k) cols[k * cap + i] = (*c[k])[i];
        }
        return n;
    } catch (const std::exception& e) {
        if (errlen > 0) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return -1;
    }
}

int ref_detect_format(const char* path, char* out, int cap) {
    try {
        const std::string f = kiwi::detect_snapshot_format(path);
        std::strncpy(out, f.c_str(), cap - 1);
        out[cap - 1] = 0;
        return 0;
    } catch (const std::exception& e) {
        std::strncpy(out, e.what(), cap - 1);
        out[cap - 1] = 0;
        return -1;
    }
}

}  // extern "C"
