// One Voronoi cell as a convex polyhedron cut out of the domain box by the bisector planes of the
// site's neighbours: the computation voro++ performs per particle (voronoicell_base::nplane, called by
// container::compute_cell for the reference's tessellation, src/tessellation.cpp:23-61 and
// src/mapping.cpp:222-269), kept as face cycles by one thread.  Host and device code.
//
// Vertices are kept relative to the site.  A cut by the plane of the neighbour at relative position n
// keeps the points p with 2 p.n < |n|^2: a vertex is outside when u = 2 p.n - |n|^2 > tol, inside when
// u < -tol and on the plane otherwise.  This is voro++'s test (voronoicell_base::m_test), whose tol is
// its tolerance, 10 double epsilons, times the cell's max_len_sq, default_length^2 = 1e6 for the
// default-constructed cells of the reference's tessellation (src/mapping.cpp:246,
// src/tessellation.cpp:40): VC_TOL_DEFAULT = 2.22e-9 in squared length units.  Vertices within it of a
// plane neither create slivers nor vanish, so lattice inputs resolve to exact cubes as in voro++.
//
// Faces are vertex cycles, counter-clockwise seen from outside the cell, each with the neighbour it is
// shared with: a site index, or -1 - w for domain wall w = 0..5 (x_min, x_max, y_min, y_max, z_min,
// z_max; voro++'s wall numbering).  A cut keeps the part of every face cycle inside the plane, inserts
// the point where each crossed edge meets the plane (interpolated from the edge's inside end, once per
// edge), and closes the cell with a new face through the kept points on the plane, ordered by angle
// about their mean.
//
// Storage: one cell's arrays are strided (element i at [i * st]), so that a device workspace can lay
// the cells of a wavefront's lanes side by side (st = the number of cells in flight) and lanes that
// touch the same element index read one contiguous run.  The host uses st = 1.
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define KIWI_VHD __host__ __device__ inline
#else
#define KIWI_VHD inline
#endif

namespace kiwi_hip {

constexpr int VC_MAXV = 384;  // vertices of one cell
constexpr int VC_MAXF = 192;  // faces of one cell
constexpr int VC_MAXFV = 48;  // vertices of one face
constexpr int VC_MAXX = 128;  // edges one cut crosses
constexpr double VC_TOL_DEFAULT = 10.0 * 2.220446049250313e-16 * 1e6;

// cut results
constexpr int VC_UNCHANGED = 0, VC_CUT = 1, VC_EMPTY = 2, VC_OVERFLOW = 3;

// bytes of one cell's storage
constexpr int64_t VC_BYTES = static_cast<int64_t>(VC_MAXV) * (4 * 8 + 2 + 1) +
                             static_cast<int64_t>(VC_MAXF) * (4 + VC_MAXFV * 2 + 1) +
                             static_cast<int64_t>(VC_MAXX) * 3 * 2 + static_cast<int64_t>(VC_MAXFV) * (8 + 2 + 2) + 2 * 4;

struct VCell {
    int nv, nf, st;
    double mr2;  // the largest squared distance of a vertex from the site, kept by vc_init_box and vc_cut
    double *x, *y, *z;  // vertices, relative to the site
    uint16_t* fv;       // face f's cycle: fv[(f * VC_MAXFV + k) * st], k < fn[f]
    uint8_t* fn;
    int* nbr;
    // scratch of one cut: per vertex its u, its side (-1 in, 0 on, +1 out) and its new index; per crossed
    // edge (a < b) its new vertex
    double* s;
    int8_t* side;
    uint16_t *remap, *xa, *xb, *xv;
    // scratch of one face: a cycle being rebuilt; the new face's vertices and their angles
    uint16_t *tmp, *cutv;
    double* ang;
    int* meta;  // nv and nf, kept with the storage by vc_save for a later pass

    KIWI_VHD double& X(int v) const { return x[v * st]; }
    KIWI_VHD double& Y(int v) const { return y[v * st]; }
    KIWI_VHD double& Z(int v) const { return z[v * st]; }
    KIWI_VHD double& S(int v) const { return s[v * st]; }
    KIWI_VHD int8_t& SIDE(int v) const { return side[v * st]; }
    KIWI_VHD uint16_t& REMAP(int v) const { return remap[v * st]; }
    // 64-bit: (VC_MAXF * VC_MAXFV) * st passes 2^31 beyond 233,000 cells in flight
    KIWI_VHD uint16_t& FV(int f, int k) const { return fv[static_cast<int64_t>(f * VC_MAXFV + k) * st]; }
    KIWI_VHD uint8_t& FN(int f) const { return fn[f * st]; }
    KIWI_VHD int& NBR(int f) const { return nbr[f * st]; }
    KIWI_VHD uint16_t& XA(int k) const { return xa[k * st]; }
    KIWI_VHD uint16_t& XB(int k) const { return xb[k * st]; }
    KIWI_VHD uint16_t& XV(int k) const { return xv[k * st]; }
    KIWI_VHD uint16_t& TMP(int k) const { return tmp[k * st]; }
    KIWI_VHD uint16_t& CUT(int k) const { return cutv[k * st]; }
    KIWI_VHD double& ANG(int k) const { return ang[k * st]; }
};

KIWI_VHD void vc_save(VCell& c) {
    c.meta[0] = c.nv;
    c.meta[c.st] = c.nf;
}
KIWI_VHD void vc_load(VCell& c) {
    c.nv = c.meta[0];
    c.nf = c.meta[c.st];
}

// Points the cell at storage for `count` cells laid side by side from `base` (VC_BYTES * count bytes,
// 8-byte aligned): cell `slot` of them.
KIWI_VHD void vc_bind(VCell& c, void* base, int64_t count, int64_t slot) {
    char* p = static_cast<char*>(base);
    int64_t off = 0;
    auto take = [&](int64_t elems, int64_t size) {
        char* q = p + off + slot * size;
        off += elems * count * size;
        return q;
    };
    c.st = static_cast<int>(count);
    c.x = reinterpret_cast<double*>(take(VC_MAXV, 8));
    c.y = reinterpret_cast<double*>(take(VC_MAXV, 8));
    c.z = reinterpret_cast<double*>(take(VC_MAXV, 8));
    c.s = reinterpret_cast<double*>(take(VC_MAXV, 8));
    c.ang = reinterpret_cast<double*>(take(VC_MAXFV, 8));
    c.nbr = reinterpret_cast<int*>(take(VC_MAXF, 4));
    c.meta = reinterpret_cast<int*>(take(2, 4));
    c.fv = reinterpret_cast<uint16_t*>(take(static_cast<int64_t>(VC_MAXF) * VC_MAXFV, 2));
    c.remap = reinterpret_cast<uint16_t*>(take(VC_MAXV, 2));
    c.xa = reinterpret_cast<uint16_t*>(take(VC_MAXX, 2));
    c.xb = reinterpret_cast<uint16_t*>(take(VC_MAXX, 2));
    c.xv = reinterpret_cast<uint16_t*>(take(VC_MAXX, 2));
    c.tmp = reinterpret_cast<uint16_t*>(take(VC_MAXFV, 2));
    c.cutv = reinterpret_cast<uint16_t*>(take(VC_MAXFV, 2));
    c.fn = reinterpret_cast<uint8_t*>(take(VC_MAXF, 1));
    c.side = reinterpret_cast<int8_t*>(take(VC_MAXV, 1));
}

// the domain box around the site (sx, sy, sz): 8 vertices, 6 quadrilateral faces
KIWI_VHD void vc_init_box(VCell& c, const double dom[6], double sx, double sy, double sz) {
    const double lo[3] = {dom[0] - sx, dom[2] - sy, dom[4] - sz}, hi[3] = {dom[1] - sx, dom[3] - sy, dom[5] - sz};
    c.nv = 8;
    for (int v = 0; v < 8; ++v) {  // vertex v: bit 0 x, bit 1 y, bit 2 z (0 low, 1 high)
        c.X(v) = (v & 1) ? hi[0] : lo[0];
        c.Y(v) = (v & 2) ? hi[1] : lo[1];
        c.Z(v) = (v & 4) ? hi[2] : lo[2];
    }
    // counter-clockwise seen from outside (right-hand normal outward)
    const uint16_t quads[6][4] = {{0, 4, 6, 2}, {1, 3, 7, 5}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 2, 3, 1}, {4, 5, 7, 6}};
    c.nf = 6;
    for (int f = 0; f < 6; ++f) {
        for (int k = 0; k < 4; ++k) c.FV(f, k) = quads[f][k];
        c.FN(f) = 4;
        c.NBR(f) = -1 - f;
    }
    double m = 0.0;
    for (int v = 0; v < 8; ++v) {
        const double r2 = c.X(v) * c.X(v) + c.Y(v) * c.Y(v) + c.Z(v) * c.Z(v);
        m = r2 > m ? r2 : m;
    }
    c.mr2 = m;
}

// the vertex where the cut's plane crosses edge (a, b), made once per edge; -1 past the limits
KIWI_VHD int vc_edge_vertex(VCell& c, int& nx, int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    for (int k = 0; k < nx; ++k)
        if (c.XA(k) == lo && c.XB(k) == hi) return c.XV(k);
    if (nx >= VC_MAXX || c.nv >= VC_MAXV) return -1;
    // interpolated from the inside end: the same point whichever face asks
    const int in = c.SIDE(lo) < 0 ? lo : hi, out = in == lo ? hi : lo;
    const double t = c.S(in) / (c.S(in) - c.S(out));
    const int v = c.nv++;
    c.X(v) = c.X(in) + t * (c.X(out) - c.X(in));
    c.Y(v) = c.Y(in) + t * (c.Y(out) - c.Y(in));
    c.Z(v) = c.Z(in) + t * (c.Z(out) - c.Z(in));
    c.SIDE(v) = 0;
    c.S(v) = 0.0;
    c.XA(nx) = static_cast<uint16_t>(lo);
    c.XB(nx) = static_cast<uint16_t>(hi);
    c.XV(nx) = static_cast<uint16_t>(v);
    ++nx;
    return v;
}

// Whether the bisector plane of the neighbour at relative position (px, py, pz) leaves the cell as it is
// (no vertex outside it): vc_cut's first test, without storing anything.
KIWI_VHD bool vc_misses(const VCell& c, double px, double py, double pz, double tol) {
    const double rsq = px * px + py * py + pz * pz;
    for (int v = 0; v < c.nv; ++v)
        if (2.0 * (c.X(v) * px + c.Y(v) * py + c.Z(v) * pz) - rsq > tol) return false;
    return true;
}

// Cuts the cell by the bisector plane of the neighbour at relative position (px, py, pz), index nb.
KIWI_VHD int vc_cut(VCell& c, double px, double py, double pz, int nb, double tol) {
    const double rsq = px * px + py * py + pz * pz;
    const int nv0 = c.nv;
    // (callers test vc_misses first: most planes tested miss the cell, and that test stores nothing)
    int nout = 0, nin = 0;
    for (int v = 0; v < nv0; ++v) {
        const double u = 2.0 * (c.X(v) * px + c.Y(v) * py + c.Z(v) * pz) - rsq;
        c.S(v) = u;
        const int8_t sd = u > tol ? 1 : (u < -tol ? -1 : 0);
        c.SIDE(v) = sd;
        nout += sd > 0;
        nin += sd < 0;
    }
    if (nout == 0) return VC_UNCHANGED;
    if (nin == 0) return VC_EMPTY;
    int nx = 0;
    // every face: the kept part of its cycle with the crossing points inserted, staged, then written over
    // the faces already done
    int nf_out = 0;
    for (int f = 0; f < c.nf; ++f) {
        int m = 0;
        const int n = c.FN(f);
        int a = c.FV(f, 0);
        for (int k = 0; k < n; ++k) {
            const int b = c.FV(f, k + 1 < n ? k + 1 : 0);
            const int sa = c.SIDE(a), sb = c.SIDE(b);
            if (sa <= 0) {
                if (m >= VC_MAXFV) return VC_OVERFLOW;
                c.TMP(m++) = static_cast<uint16_t>(a);
            }
            if (sa * sb < 0) {
                const int v = vc_edge_vertex(c, nx, a, b);
                if (v < 0 || m >= VC_MAXFV) return VC_OVERFLOW;
                c.TMP(m++) = static_cast<uint16_t>(v);
            }
            a = b;
        }
        if (m < 3) continue;  // the face was outside (or met the plane in an edge or a point)
        for (int k = 0; k < m; ++k) c.FV(nf_out, k) = c.TMP(k);
        c.FN(nf_out) = static_cast<uint8_t>(m);
        c.NBR(nf_out) = c.NBR(f);
        ++nf_out;
    }
    // the vertices the remaining faces keep
    for (int v = 0; v < c.nv; ++v) c.REMAP(v) = 0;
    for (int f = 0; f < nf_out; ++f)
        for (int k = 0; k < c.FN(f); ++k) c.REMAP(c.FV(f, k)) = 1;
    // the new face: the kept vertices on the plane, in angular order about their mean, counter-clockwise
    // seen from outside (+n)
    int ncut = 0;
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (int v = 0; v < c.nv; ++v) {
        if (c.REMAP(v) && c.SIDE(v) == 0) {
            if (ncut >= VC_MAXFV) return VC_OVERFLOW;
            c.CUT(ncut++) = static_cast<uint16_t>(v);
            mx += c.X(v);
            my += c.Y(v);
            mz += c.Z(v);
        }
    }
    if (ncut >= 3) {
        if (nf_out >= VC_MAXF) return VC_OVERFLOW;
        mx /= ncut;
        my /= ncut;
        mz /= ncut;
        // a right-handed basis (u, w, n) of the plane: u perpendicular to n, w = n x u
        const double inv = 1.0 / sqrt(rsq);
        const double ex = px * inv, ey = py * inv, ez = pz * inv;
        double ux, uy, uz;
        if (fabs(ex) < 0.9) {  // u = (1, 0, 0) x n
            ux = 0.0;
            uy = -ez;
            uz = ey;
        } else {  // u = (0, 1, 0) x n
            ux = ez;
            uy = 0.0;
            uz = -ex;
        }
        const double wx = ey * uz - ez * uy, wy = ez * ux - ex * uz, wz = ex * uy - ey * ux;
        for (int k = 0; k < ncut; ++k) {
            const int v = c.CUT(k);
            const double dx = c.X(v) - mx, dy = c.Y(v) - my, dz = c.Z(v) - mz;
            c.ANG(k) = atan2(dx * wx + dy * wy + dz * wz, dx * ux + dy * uy + dz * uz);
        }
        for (int k = 1; k < ncut; ++k) {  // insertion sort by angle
            const double a = c.ANG(k);
            const uint16_t v = c.CUT(k);
            int j = k - 1;
            while (j >= 0 && c.ANG(j) > a) {
                c.ANG(j + 1) = c.ANG(j);
                c.CUT(j + 1) = c.CUT(j);
                --j;
            }
            c.ANG(j + 1) = a;
            c.CUT(j + 1) = v;
        }
        for (int k = 0; k < ncut; ++k) c.FV(nf_out, k) = c.CUT(k);
        c.FN(nf_out) = static_cast<uint8_t>(ncut);
        c.NBR(nf_out) = nb;
        ++nf_out;
    }
    c.nf = nf_out;
    // drop the vertices no face keeps; the farthest kept vertex on the way
    int nvn = 0;
    double m = 0.0;
    for (int v = 0; v < c.nv; ++v) {
        if (c.REMAP(v)) {
            c.REMAP(v) = static_cast<uint16_t>(nvn);
            const double x = c.X(v), y = c.Y(v), z = c.Z(v);
            c.X(nvn) = x;
            c.Y(nvn) = y;
            c.Z(nvn) = z;
            const double r2 = x * x + y * y + z * z;
            m = r2 > m ? r2 : m;
            ++nvn;
        }
    }
    c.mr2 = m;
    for (int f = 0; f < c.nf; ++f)
        for (int k = 0; k < c.FN(f); ++k) c.FV(f, k) = c.REMAP(c.FV(f, k));
    c.nv = nvn;
    return VC_CUT;
}

// face-vertex entries of the cell
KIWI_VHD int vc_face_vertices(const VCell& c) {
    int t = 0;
    for (int f = 0; f < c.nf; ++f) t += c.FN(f);
    return t;
}


// ---- the tessellation around the cell ----

// A uniform bucket grid of the sites over the domain, for the neighbour search: about VC_SITES_PER_BUCKET
// sites a bucket, at most VC_MAX_BUCKET_AXIS buckets an axis; bucket (i, j, k) is (i * nb[1] + j) * nb[2]
// + k.  tolb bounds how far rounding can put a coordinate across a bucket edge lo + m edge.  Small buckets
// keep the shells close to distance order where sites cluster (measured on the GPU at 2, 1, 0.5 and 0.25
// sites a bucket: 2e5 log-normal sites 420 / 349 / 326 / 329 ms, 1e5 clustered sites 376 (at 1) / 289 /
// 211 ms).
constexpr double VC_SITES_PER_BUCKET = 0.25;
constexpr int VC_MAX_BUCKET_AXIS = 1024;

struct VcGrid {
    double lo[3], edge[3], tolb;
    int nb[3];
};

// a site in bucket order: position and index, one 32-byte record
struct alignas(32) VcSite {
    double x, y, z;
    int idx, pad;
};

inline VcGrid vc_grid_make(const double dom[6], int64_t n) {
    VcGrid g;
    const double L[3] = {dom[1] - dom[0], dom[3] - dom[2], dom[5] - dom[4]};
    const double e = std::cbrt(L[0] * L[1] * L[2] * VC_SITES_PER_BUCKET / static_cast<double>(n > 1 ? n : 1));
    double scale = 0.0;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = dom[2 * a];
        double m = std::ceil(L[a] / e);
        m = m < 1.0 ? 1.0 : (m > VC_MAX_BUCKET_AXIS ? VC_MAX_BUCKET_AXIS : m);
        g.nb[a] = static_cast<int>(m);
        g.edge[a] = L[a] / g.nb[a];
        scale = std::fmax(scale, std::fmax(std::fabs(dom[2 * a]), std::fabs(dom[2 * a + 1])));
    }
    g.tolb = 1e-12 * scale;
    return g;
}

KIWI_VHD int vc_bucket_axis(const VcGrid& g, int a, double v) {
    const double f = floor((v - g.lo[a]) / g.edge[a]);
    return static_cast<int>(fmin(fmax(f, 0.0), static_cast<double>(g.nb[a] - 1)));
}

// voro++'s block grid for the reference's container: pre_container_base::guess_optimal (5.6 sites a block)
// scaled by the default block_scale_factor 3 and rounded (guess_scaled_blocks,
// include/kiwi/voro_grid.hpp:18-26); n_in = the sites pre_container::put keeps (inside [min, max] on
// every axis).  The blocks fix the tessellation's cell order (c_loop_all: blocks x fastest, sites in
// input order within a block) and which sites on a max face the container keeps.
inline void vc_block_dims(const double dom[6], int64_t n_in, int nbk[3]) {
    const double dx = dom[1] - dom[0], dy = dom[3] - dom[2], dz = dom[5] - dom[4];
    const double ilscale = std::pow(static_cast<double>(static_cast<int>(n_in)) / (5.6 * dx * dy * dz), 1 / 3.0);
    const int g[3] = {static_cast<int>(dx * ilscale + 1), static_cast<int>(dy * ilscale + 1),
                      static_cast<int>(dz * ilscale + 1)};
    for (int a = 0; a < 3; ++a) {
        const int m = static_cast<int>(std::round(g[a] * 3.0));
        nbk[a] = m > 1 ? m : 1;
    }
}

// voro++'s block of a site (container_base::put_remap, non-periodic), or -1 where the container drops it
KIWI_VHD int64_t vc_block_index(const double dom[6], const int nbk[3], double x, double y, double z) {
    const double p[3] = {x, y, z};
    int64_t ijk = 0, mul = 1;
    for (int a = 0; a < 3; ++a) {
        if (!(p[a] == p[a])) return -1;
        const double box = (dom[2 * a + 1] - dom[2 * a]) / nbk[a];
        const double sp = 1 / box;
        const double t = (p[a] - dom[2 * a]) * sp;
        if (!(t > -2.0 && t < 2.0 * nbk[a] + 2.0)) return -1;  // well outside (and no overflow of the cast)
        const int i = t < 0 ? static_cast<int>(t) - 1 : static_cast<int>(t);
        if (i < 0 || i >= nbk[a]) return -1;
        ijk += mul * i;
        mul *= nbk[a];
    }
    return ijk;
}

struct alignas(8) int2v {
    int x, y;
};

// A lane's place in its walk over the bucket shells: shell r, bucket offset (dk, dj, di) within it
// (dk slowest; inside the shell's interior slabs only di = -r and r), and the sites [t, t1) of the
// current bucket.
struct VcWalk {
    int r, dk, dj, di, t, t1;
    bool fresh;  // (dk, dj, di) is the next bucket to visit, not the last one visited
};

// Moves the walk to its next bucket inside the grid and loads its site range; false when the walk
// ends: every bucket scanned, or every unscanned bucket farther than twice the farthest vertex (every
// site of a bucket at Chebyshev radius >= r lies beyond one face of the box of radius r - 1 about the
// site's own bucket b).
KIWI_VHD bool vc_next_bucket(VcWalk& w, const VcGrid& g, const int b[3], const double p[3], double mr2,
                             const int* range) {
    for (;;) {
        if (w.fresh) {
            w.fresh = false;
        } else {
            const bool full = w.dk == -w.r || w.dk == w.r || w.dj == -w.r || w.dj == w.r;
            w.di += full ? 1 : 2 * w.r;
            if (w.di > w.r) {
                w.di = -w.r;
                if (++w.dj > w.r) {
                    w.dj = -w.r;
                    if (++w.dk > w.r) {  // the shell is done: the next one, if it can matter
                        const int r = ++w.r;
                        double dmin = INFINITY;
                        for (int a = 0; a < 3; ++a) {
                            if (b[a] - r >= 0) dmin = fmin(dmin, p[a] - (g.lo[a] + (b[a] - r + 1) * g.edge[a]));
                            if (b[a] + r < g.nb[a]) dmin = fmin(dmin, g.lo[a] + (b[a] + r) * g.edge[a] - p[a]);
                        }
                        if (dmin == INFINITY) return false;  // every bucket scanned
                        dmin -= g.tolb;
                        if (dmin > 0.0 && dmin * dmin > 4.0 * mr2) return false;
                        w.dk = w.dj = w.di = -r;
                    }
                }
            }
        }
        const int i = b[0] + w.di, j = b[1] + w.dj, k = b[2] + w.dk;
        if (i < 0 || i >= g.nb[0] || j < 0 || j >= g.nb[1] || k < 0 || k >= g.nb[2]) continue;
        const int64_t bk = (static_cast<int64_t>(i) * g.nb[1] + j) * g.nb[2] + k;
        const int2v rr = reinterpret_cast<const int2v*>(range)[bk];  // [beg, end) in one load
        w.t = rr.x;
        w.t1 = rr.y;
        return true;
    }
}

// Computes the cell of site `self` at (sx, sy, sz): the domain box cut by every site of the buckets (the
// kept sites) closer than twice the cell's farthest vertex, scanned in shells of buckets of growing
// Chebyshev radius until the unscanned buckets are all farther.  range[2 b], range[2 b + 1] (8-byte
// aligned): bucket b's sites in `sites`.
// Each round first advances to the lane's next site whose plane cuts the cell, then cuts: on a wavefront,
// whose lanes walk different buckets, the cuts of all lanes then run together instead of one lane's at
// a time.  Returns VC_CUT, or VC_EMPTY / VC_OVERFLOW.
KIWI_VHD int vc_compute(VCell& c, const double dom[6], const VcGrid& g, const int* range, const VcSite* sites,
                        int self, double sx, double sy, double sz, double tol) {
    vc_init_box(c, dom, sx, sy, sz);
    const double p[3] = {sx, sy, sz};
    const int b[3] = {vc_bucket_axis(g, 0, sx), vc_bucket_axis(g, 1, sy), vc_bucket_axis(g, 2, sz)};
    VcWalk w{0, 0, 0, 0, 0, 0, true};
    for (;;) {
        bool found = false;
        double px = 0.0, py = 0.0, pz = 0.0;
        int nb = 0;
        for (;;) {
            if (w.t < w.t1) {
                const VcSite q = sites[w.t++];
                if (q.idx == self) continue;
                px = q.x - sx;
                py = q.y - sy;
                pz = q.z - sz;
                if (px * px + py * py + pz * pz > 4.0 * c.mr2) continue;  // its plane misses the cell
                if (vc_misses(c, px, py, pz, tol)) continue;
                nb = q.idx;
                found = true;
                break;
            }
            if (!vc_next_bucket(w, g, b, p, c.mr2, range)) break;
        }
        if (!found) break;
        const int rc = vc_cut(c, px, py, pz, nb, tol);  // keeps c.mr2
        if (rc != VC_CUT && rc != VC_UNCHANGED) return rc;
    }
    return VC_CUT;
}

// Writes the cell's faces in CSR form from face `f0` and face-vertex entry `v0` on: face_vert_ptr[f0 + f],
// absolute vertex coordinates (vertex + site), the neighbour across each face as a cell index (-1 for a
// domain wall).  Each cycle is written reversed, clockwise seen from outside: the orientation of voro++'s
// face_vertices (voronoicell_base::face_vertices), which the map's face walk reads.
KIWI_VHD void vc_write(const VCell& c, double sx, double sy, double sz, const int* site_cell, int64_t f0, int64_t v0,
                       int64_t* face_vert_ptr, double* vert_xyz, int* face_nbr) {
    int64_t v = v0;
    for (int f = 0; f < c.nf; ++f) {
        face_vert_ptr[f0 + f] = v;
        const int nb = c.NBR(f);
        face_nbr[f0 + f] = nb < 0 ? -1 : site_cell[nb];
        const int n = c.FN(f);
        for (int k = 0; k < n; ++k, ++v) {
            const int u = c.FV(f, k ? n - k : 0);
            vert_xyz[3 * v] = c.X(u) + sx;
            vert_xyz[3 * v + 1] = c.Y(u) + sy;
            vert_xyz[3 * v + 2] = c.Z(u) + sz;
        }
    }
}

}  // namespace kiwi_hip
