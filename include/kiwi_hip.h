/* libkiwi_hip: MI355X (gfx950) implementation of KIWI's exact kernel-integral mapping.
 *
 * Plain C ABI: no exceptions cross it, every struct is plain old data, pointers and sizes only.
 * It replaces the per-(particle, cell) loops of the reference's two mapping stages:
 *
 *   kiwi_hip_sph_to_grid   <- the particle loop of run_sph_to_grid_stage
 *                             (src/cartesian_grid.cpp:184-301; include/kiwi/cartesian_grid.hpp:88),
 *                             i.e. the 6 x KernelIntegrator::face calls per cell
 *                             (include/kiwi/kernel_integrator.hpp:22-27).
 *   kiwi_hip_voronoi_map   <- the particle loop of run_mapping_stage (src/mapping.cpp:400-460): the BFS
 *                             over the tessellation with compute_cell_integral (:105-206) per cell.
 *
 * and, beside them, the inverse-mapping test, the cartesian-map resample, the coverage-grid buffer, the
 * Lloyd step's centroids and the Voronoi tessellation the map consumes (kiwi_hip_voronoi_tessellate, in
 * place of the voro++ container of src/mapping.cpp:222-269 and src/tessellation.cpp:23-61).
 *
 * The host keeps what the reference keeps outside those loops: configuration, snapshot reading,
 * normalisation (src/cartesian_grid.cpp:294-318, src/mapping.cpp:473-502), file output and caching.
 *
 * Outputs are the UN-normalised per-cell sums the reference accumulates:
 *   mass[c] = sum_p Msum(p,c) m_p,   p*[c] = sum_p Msum(p,c) m_p v*_p
 * Cartesian cells are x-major: c = (ix*ny + iy)*nz + iz (include/kiwi/cartesian_grid.hpp:62-66);
 * Voronoi cells are in the order of the geometry the caller passes (the reference's c_loop_all order).
 *
 * Return codes: KIWI_OK (0) or a negative KIWI_ERR_*; kiwi_hip_last_error() then holds the message.
 * A context owns three HIP streams, all at default priority: the main one; a second one on which the
 * sph-to-grid entry points compute the per-particle line tables beside the item sort; a third one for
 * mixed precision's FP64 fallback, whose line tables and cell kernel run beside the item sort and the
 * FP32 cell kernel.  It owns its device scratch too, and must not be used from two threads at once.
 * Calls are synchronous: outputs are complete when the call returns.  Device inputs must be complete
 * when the call is made: the context's streams do not wait for work the caller queued on other streams
 * (a caller that produces inputs on its own stream synchronises that stream first).
 */
#ifndef KIWI_HIP_H
#define KIWI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KIWI_HIP_ABI_VERSION 12

#define KIWI_OK 0
#define KIWI_ERR_INVALID_ARG (-1)
#define KIWI_ERR_DEVICE (-2)
#define KIWI_ERR_OUT_OF_MEMORY (-3)
#define KIWI_ERR_UNSUPPORTED (-4)
#define KIWI_ERR_LIMIT (-5)

/* arithmetic of the per-pair kernel integrals */
#define KIWI_FP64 0  /* FP64 throughout (parity mode) */
#define KIWI_MIXED 1 /* FP32 bulk, FP64 accumulation */

/* kernels, in the reference's order (src/kernel_spec_data.cpp:12-163) */
#define KIWI_KERNEL_M4 0
#define KIWI_KERNEL_M5 1
#define KIWI_KERNEL_M6 2
#define KIWI_KERNEL_WC2 3
#define KIWI_KERNEL_WC4 4
#define KIWI_KERNEL_WC6 5

typedef struct kiwi_hip_ctx kiwi_hip_ctx;

/* SPHSnapshot's columns (include/kiwi/snapshot.hpp:70-77), structure of arrays, n particles. */
typedef struct kiwi_particles {
    int64_t n;
    const double *x, *y, *z;
    const double *vx, *vy, *vz;
    const double *m, *h;
    int on_device; /* 1: device pointers (already in HBM), 0: host pointers */
} kiwi_particles;

/* Config::domain plus the resolved per-axis cell counts (src/cartesian_grid.cpp:89-97). */
typedef struct kiwi_grid {
    double x_min, x_max, y_min, y_max, z_min, z_max;
    int nx, ny, nz;
} kiwi_grid;

/* KernelIntegrator settings (include/kiwi/kernel_integrator.hpp:118-122, src/config.cpp:267-277). */
typedef struct kiwi_kernel {
    int kernel_id;       /* KIWI_KERNEL_* */
    double support_in_h; /* kernel.support_radius_in_h; 0 = the kernel's default */
    int generic_path;    /* kernel.generic_path */
    int tabulated;       /* *.use_tabulated_integral (cubic default-support path only) */
} kiwi_kernel;

/* n_cells doubles each; on_device selects device or host pointers. */
typedef struct kiwi_cell_sums {
    double *mass, *px, *py, *pz;
    int on_device;
} kiwi_cell_sums;

#define KIWI_MAX_WARN_RECORDS 16

/* One Msum <= -0.01 event (the reference's warning line, src/cartesian_grid.cpp:264-269). */
typedef struct kiwi_warn_record {
    int64_t particle; /* 0-based particle index */
    int64_t cell;     /* flattened cell index */
    double value;
} kiwi_warn_record;

typedef struct kiwi_stats {
    int64_t skipped_outside;  /* particles whose support box misses the grid */
    int64_t n_negative;       /* pairs with Msum < 0, clamped to 0 */
    int64_t n_negative_warn;  /* of which Msum <= -0.01 */
    int64_t pairs_bbox;       /* (particle, cell) pairs in the clamped bounding boxes */
    int64_t pairs_sphere;     /* pairs whose cell intersects the support sphere (the throughput unit) */
    int64_t particle_blocks;  /* (particle, 8^3 cell block) work items */
    double v_mean[3];         /* mean velocity over all particles */
    double device_ms_prep;    /* binning, sorts, block order */
    double device_ms_lines;   /* per-particle edge-line constants: the part not hidden behind the item sort
                                 (they run on other streams beside it) */
    double device_ms_gather;  /* cell integrals: lattice values and per-pair cell sums */
    double device_ms_accum;   /* accumulation of the per-pair sums into the grid */
    double device_ms_lines_kernel; /* the line-table kernels themselves, timed on the streams they run on
                                      (from their start to the end of the last of them) */
    int64_t lattice_points[3]; /* lattice values evaluated, by radial region (inside h, 2h, outside) */
    int64_t line_evals[3];     /* edge-line constants with 0, 1 or 2 crossing evaluations */
    int64_t plane_evals;       /* wall-plane constants evaluated */
    /* mixed precision's large-support FP64 fallback (kiwi_hip_set_fp64_fallback), a part of the totals
       above: its (particle, cell block) items, sphere pairs, lattice values by region, and the time of its
       cell kernel */
    int64_t fp64_fallback_blocks;
    int64_t fp64_fallback_pairs;
    int64_t fp64_fallback_points[3];
    double device_ms_gather_fallback; /* the fallback's cell kernel, on the third stream: it starts when the
                                         items are laid out, before the FP32 one (part of it falls within
                                         device_ms_prep and device_ms_lines) */
    double device_ms_gather_kernel;   /* the main cell kernel alone (the FP32 one in mixed precision) */
    int n_warn_records;
    kiwi_warn_record warn[KIWI_MAX_WARN_RECORDS];
} kiwi_stats;

/* A Voronoi tessellation flattened to plain arrays (the reference's CellFaceGeometry, src/mapping.cpp:49-54,
 * resolved to absolute coordinates and cell indices the way compute_cell_integral reads it, :130-175).
 * Cell c owns faces [cell_face_ptr[c], cell_face_ptr[c+1]); face f owns vertices
 * [face_vert_ptr[f], face_vert_ptr[f+1]), in voro++'s face_vertices() order, as xyz triples in vert_xyz.
 * face_neighbor[f] is the cell index across face f, or -1 for a domain wall or an unresolved neighbour. */
typedef struct kiwi_voronoi_cells {
    int64_t n_cells, n_faces, n_face_vertices;
    const int64_t* cell_face_ptr; /* n_cells + 1 */
    const int64_t* face_vert_ptr; /* n_faces + 1 */
    const double* vert_xyz;       /* 3 * n_face_vertices */
    const int* face_neighbor;     /* n_faces */
    double domain[6];             /* x_min x_max y_min y_max z_min z_max: the tessellated box */
    int on_device;                /* 1: device pointers, 0: host pointers */
} kiwi_voronoi_cells;

typedef struct kiwi_map_stats {
    int64_t unresolved;      /* particles without a home cell (skipped, src/mapping.cpp:436-440) */
    int64_t unmapped;        /* particles whose kernel reached no cell beyond the home cell (:459) */
    int64_t n_negative;      /* (particle, cell) sums < 0, clamped to 0 (:198-203) */
    int64_t n_negative_warn; /* of which <= -0.01 */
    int64_t pairs_candidate; /* (particle, cell) pairs whose bounding boxes meet the support */
    int64_t pairs_visited;   /* pairs the reference's BFS visits (home cell, or a face within the support) */
    int64_t face_evals;      /* face integrals evaluated */
    double v_mean[3];        /* mean velocity over all particles */
    double device_ms_prep;   /* geometry, candidate pairs, visits, face lists */
    double device_ms_faces;  /* face integrals */
    double device_ms_accum;  /* per-pair sums and per-cell accumulation */
    int n_warn_records;
    kiwi_warn_record warn[KIWI_MAX_WARN_RECORDS]; /* cell = the cell index */
} kiwi_map_stats;

int kiwi_hip_abi_version(void);

/* One context per host thread; binds device_ordinal and creates a stream. */
int kiwi_hip_create(kiwi_hip_ctx** out, int device_ordinal);
void kiwi_hip_destroy(kiwi_hip_ctx* ctx);
const char* kiwi_hip_last_error(const kiwi_hip_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that time or order work around the calls. */
void* kiwi_hip_stream(kiwi_hip_ctx* ctx);

/* Direct SPH -> Cartesian grid, replacing src/cartesian_grid.cpp:184-301.
 * precision: KIWI_FP64 or KIWI_MIXED.  KIWI_MIXED computes the lattice values in FP32 (sums in FP64) for
 * the analytic cubic spline, the generic cubic spline and Wendland C2, and for M5 while the FP64 fallback
 * for large supports is on (kiwi_hip_set_fp64_fallback).  M6, Wendland C4 and C6 (whose FP32 lattice
 * values lose more than 1e-6 of the mass of particles cut by the domain boundary) and the tabulated path
 * run in FP64 whatever is asked.  out: n_cells = nx*ny*nz per array.  stats
 * may be NULL. */
int kiwi_hip_sph_to_grid(kiwi_hip_ctx* ctx, const kiwi_particles* particles, const kiwi_grid* grid,
                         const kiwi_kernel* kernel, int precision, const kiwi_cell_sums* out,
                         kiwi_stats* stats);

/* kiwi_hip_sph_to_grid without device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every buffer of the pipeline run is carved from a fixed scratch area
 * of KIWI_NOALLOC_SCRATCH_BYTES that the library holds on each device from its load until the process
 * ends; the call makes no hipMalloc or hipFree.  Particles (n > 0) and outputs must be device pointers;
 * the tabulated integral is not available (KIWI_ERR_UNSUPPORTED).  A problem whose buffers do not fit
 * the area is rejected with KIWI_ERR_LIMIT (pairs of bounding-box cells dominate: about 4 bytes each in
 * KIWI_MIXED, 8 in KIWI_FP64).  Results equal kiwi_hip_sph_to_grid's bit for bit.  All calls on one
 * device share its area, which keeps the last call's contents: calls that overlap on one device, from
 * any context, race and must be serialised by the caller. */
#define KIWI_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int kiwi_hip_sph_to_grid_noalloc(kiwi_hip_ctx* ctx, const kiwi_particles* particles, const kiwi_grid* grid,
                                 const kiwi_kernel* kernel, int precision, const kiwi_cell_sums* out,
                                 kiwi_stats* stats);

/* One problem of a batch: the arguments of kiwi_hip_sph_to_grid. */
typedef struct kiwi_batch_problem {
    kiwi_particles particles;
    kiwi_grid grid;
    kiwi_kernel kernel;
    int precision;
    kiwi_cell_sums out;
    kiwi_stats* stats; /* may be NULL */
} kiwi_batch_problem;

#define KIWI_MAX_BATCH 4096

/* Many independent sph-to-grid problems in one call (no reference counterpart; the reference runs one
 * cutout per process).  Problems that share a kernel path (kernel, support, arithmetic) run as one
 * pipeline: every stage is launched once for all of them, and the cell integrals of all their
 * (particle, cell block) items share one persistent launch.  Each problem's outputs and statistics are
 * those kiwi_hip_sph_to_grid would return for it alone (device times cover its group's run;
 * particle_blocks is -1 when the group holds more than one problem). */
int kiwi_hip_sph_to_grid_batch(kiwi_hip_ctx* ctx, int n_problems, const kiwi_batch_problem* problems);

/* The point set of a Voronoi map with each site's mapping.dat row (the cartesian-map input). */
typedef struct kiwi_sites {
    int64_t n;
    const double *x, *y, *z;              /* site positions; sites outside the domain take no part */
    const int* id;                        /* voro++ particle id of each site */
    const double *density, *vx, *vy, *vz; /* the mapping.dat row of the site's cell */
    const int* mapped;                    /* 0 where the site's cell has no mapping.dat row; NULL: all have */
    int on_device;
} kiwi_sites;

/* cartesian_map.dat fields, res^3 each, x-major ((ix res + iy) res + iz). */
typedef struct kiwi_cartesian_map_out {
    double *density, *vx, *vy, *vz;
    int* cell_id; /* the voro++ id of the Voronoi cell the centre lies in (0 when unresolved) */
    int on_device;
} kiwi_cartesian_map_out;

typedef struct kiwi_cartesian_map_stats {
    int64_t unresolved; /* centres without an enclosing cell (every centre when there are no sites) */
    int64_t unmapped;   /* centres whose cell has no mapping.dat row (left at 0) */
    double device_ms_prep, device_ms_query;
} kiwi_cartesian_map_stats;

/* The cartesian-map resample, replacing the cell loop of run_cartesian_map_stage
 * (src/cartesian_map.cpp:148-183): every cell centre of a res^3 grid over domain[6] takes the row of the
 * Voronoi cell it lies in (its nearest site by squared distance, as voro++ find_voronoi_cell).  Of sites
 * at exactly equal distance the lowest index wins (voro++ keeps the first its block search meets).  Sites
 * outside [min, max] on any axis are dropped, as voro++'s pre_container::put drops them (:67-69).  A site
 * exactly on a max face is kept; voro++ keeps or drops it by the rounding of its block index, so pass the
 * tessellation's own cells as the sites where such points occur. */
int kiwi_hip_cartesian_map(kiwi_hip_ctx* ctx, const kiwi_sites* sites, const double domain[6], int res,
                           const kiwi_cartesian_map_out* out, kiwi_cartesian_map_stats* stats);

/* The coverage buffer's candidate lattice (generate_cartesian_lattice, src/point_set.cpp:166-195):
 * spacing = spacing_fraction x the domain's smallest extent, n_a = max(1, round(extent_a / spacing)) per
 * axis, candidate (i, j, k) at min + spacing (index + 0.5), k fastest.  A host computation, no context. */
typedef struct kiwi_lattice {
    int nx, ny, nz;
    double spacing;
} kiwi_lattice;

int kiwi_hip_cartesian_lattice(const double domain[6], double spacing_fraction, kiwi_lattice* lattice);

/* The buffer sites, in lattice order. */
typedef struct kiwi_buffer_sites {
    double *x, *y, *z;
    int* id;          /* n_particles + 1, n_particles + 2, ... (src/point_set.cpp:262-270) */
    int64_t capacity; /* entries each array holds (the lattice size always suffices) */
    int on_device;
} kiwi_buffer_sites;

typedef struct kiwi_coverage_stats {
    int64_t candidates;       /* lattice size */
    int64_t kept;             /* uncovered candidates (written when they fit the capacity) */
    double h_max;             /* the largest h; the reference's hash bin is factor x h_max */
    int64_t particle_columns; /* work: lattice columns tested over all particles */
    double device_ms;
} kiwi_coverage_stats;

/* The coverage-grid buffer, replacing add_coverage_grid_buffer (src/point_set.cpp:208-273): the
 * lattice candidates that no particle covers (within coverage_radius_factor x its h, among the particles
 * of the 27 hash bins around the candidate as the reference searches them), returned as new sites.
 * Uses the particles' x, y, z and h.  Errors as the reference: spacing_fraction or the factor <= 0 (the
 * configuration's checks, src/config.cpp:109-112) and no h > 0 ("all smoothing lengths are <= 0");
 * beyond it, non-finite particle coordinates.  KIWI_ERR_LIMIT when more sites are kept than
 * out->capacity holds (stats->kept says how many). */
int kiwi_hip_coverage_buffer(kiwi_hip_ctx* ctx, const kiwi_particles* particles, const double domain[6],
                             double spacing_fraction, double coverage_radius_factor, const kiwi_buffer_sites* out,
                             kiwi_coverage_stats* stats);

/* Launch shape of the sph-to-grid cell kernel (no reference counterpart).  Results never depend on it:
 * every (particle, cell block) item is computed and placed the same way whatever the shape. */
typedef struct kiwi_tuning {
    int item_grab; /* items a wavefront takes from a work cursor at once, 1..256 (default 16) */
    int cursors;   /* contiguous item slices with a work cursor each, 1..16 (default 8) */
} kiwi_tuning;

#define KIWI_MAX_CURSORS 16

/* Mixed precision's FP64 fallback (no reference counterpart; the reference computes everything in FP64).
 * A cell's integral is a difference of lattice values of order one, while the cell holds a share of the
 * particle of order (cell / support)^3: in FP32 the relative error of a cell's sum grows with the cube
 * of the support radius in cells.  In KIWI_MIXED, a particle whose support radius (support_radius_in_h x h)
 * spans at least support_cells times the smallest cell edge of its grid has its lattice values, line and
 * plane constants and cell sums computed in FP64 (stored, like the FP32 ones, as float per-pair sums and
 * accumulated in FP64).  support_cells <= 0 switches the fallback off; the default is
 * KIWI_FP64_FALLBACK_CELLS; M5, whose FP32 lattice values cancel more, falls back from 0.75 of it.
 * KIWI_FP64 and the tabulated path are unaffected.  Within a cell, the
 * fallback's particles are added after the others (each group in particle order), so mixed results stay
 * bit-reproducible, and a batch still equals its single calls bit for bit. */
#define KIWI_FP64_FALLBACK_CELLS 16.0
int kiwi_hip_set_fp64_fallback(kiwi_hip_ctx* ctx, double support_cells);
int kiwi_hip_get_fp64_fallback(const kiwi_hip_ctx* ctx, double* support_cells);

/* Graph mode (no reference counterpart).  With on = 1, sph-to-grid calls (single and batched) whose outputs
 * are on the device run their pipeline after the sizing pass (the item layout, sorts, line tables, cell
 * kernels, accumulation and statistics) as a captured HIP graph: captured on the first call with a given
 * signature (problem sizes, kernel path, launch shape and the addresses of every buffer involved), then
 * replayed, one graph launch per kernel path, while the signature holds; a call with another signature
 * captures anew (the context keeps the 8 most recently used).  The replay does all the work of an
 * eager call and gives the same results bit for bit; the library's own stable radix sort replaces
 * hipcub's in the captured pipeline (which would record memset nodes), and a capture that holds a memset
 * node is refused with KIWI_ERR_DEVICE.  Calls with host outputs, the tabulated path and the
 * allocation-free entry point always run eagerly.  Statistics: device_ms_prep is the sizing pass and
 * device_ms_gather the whole replayed pipeline; the other device times are 0 for a replayed call.
 * on = 0 (the default) releases the captured pipelines. */
int kiwi_hip_set_graphs(kiwi_hip_ctx* ctx, int on);
int kiwi_hip_get_graphs(const kiwi_hip_ctx* ctx, int* on);

/* Set (NULL: restore the defaults) or read the context's launch shape, used by every later call. */
int kiwi_hip_set_tuning(kiwi_hip_ctx* ctx, const kiwi_tuning* tuning);
int kiwi_hip_get_tuning(const kiwi_hip_ctx* ctx, kiwi_tuning* tuning);

/* Autotuner: runs the batch (outputs and stats as kiwi_hip_sph_to_grid_batch writes them) once per
 * candidate shape (item_grab 8, 16, 32, 64 x cursors 1, 4, 8, 16), twice each, keeps the shape with the
 * shortest call on the context and returns it with its time (best_ms may be NULL). */
int kiwi_hip_autotune(kiwi_hip_ctx* ctx, int n_problems, const kiwi_batch_problem* problems, kiwi_tuning* best,
                      double* best_ms);

/* Per-particle gradients, n each; on_device selects device or host pointers. */
typedef struct kiwi_particle_grads {
    double *m, *x, *y, *z, *h, *vx, *vy, *vz;
    int on_device;
} kiwi_particle_grads;

typedef struct kiwi_adjoint_stats {
    int64_t skipped_outside; /* particles whose support box misses the grid (zero gradients) */
    int64_t pairs_bbox;      /* (particle, cell) pairs in the bounding boxes */
    int64_t lattice_points;  /* lattice values and gradients evaluated */
    int64_t cells;           /* (particle, cell) pairs inside the support */
    double device_ms_prep;   /* boxes, work order */
    double device_ms;        /* line constants and the gradient pass */
} kiwi_adjoint_stats;

/* The adjoint of kiwi_hip_sph_to_grid (no counterpart in the reference): given dL/d(mass, px, py, pz) per
 * cell (grad_cells, n_cells each, the un-normalised sums the forward returns), the gradients of L with
 * respect to each particle's m, x, y, z, h, vx, vy, vz.  The kernel integrals are recomputed, not stored.
 * The analytic cubic spline and the generic kernels (M4 forced generic or with another support, M5, M6,
 * Wendland C2/C4/C6); the tabulated integral returns KIWI_ERR_UNSUPPORTED.  FP64 differentiates the FP64
 * forward (pairs it clamps at 0 have zero derivative); KIWI_MIXED evaluates the lattice values and
 * gradients in FP32 with FP64 accumulation for the cubic spline, generic M4 and Wendland C2, and in FP64
 * for the others (M5 included: its forward computes in FP32 only beside the FP64 fallback for large
 * supports, which the adjoint does not have). */
int kiwi_hip_sph_to_grid_adjoint(kiwi_hip_ctx* ctx, const kiwi_particles* particles, const kiwi_grid* grid,
                                 const kiwi_kernel* kernel, int precision, const kiwi_cell_sums* grad_cells,
                                 const kiwi_particle_grads* grad_particles, kiwi_adjoint_stats* stats);

/* Voronoi map, replacing the particle loop of run_mapping_stage (src/mapping.cpp:400-460).
 * home_cell[n]: each particle's cell (voro++ find_voronoi_cell resolved through reverse_id_list,
 * :434-440), -1 if unresolved; host or device memory as particles->on_device.  Outputs n_cells each.
 * KIWI_MIXED evaluates the face integrals of the analytic cubic spline in FP32 (vertex coordinates taken
 * relative to the particle), with the visit test and all sums in FP64, and clamps negative cell totals
 * instead of negative (particle, cell) sums (no per-pair warnings); the other kernel paths run in FP64
 * whatever `precision` asks for. */
int kiwi_hip_voronoi_map(kiwi_hip_ctx* ctx, const kiwi_particles* particles, const int* home_cell,
                         const kiwi_voronoi_cells* cells, const kiwi_kernel* kernel, int precision,
                         const kiwi_cell_sums* out, kiwi_map_stats* stats);

/* The Lloyd step (lloyd_step, src/point_set.cpp:275-302, which takes voronoicell_base::centroid of every
 * cell): each cell's centroid, the next position of its site, and its volume, from the flattened
 * tessellation.  site_xyz: 3 * n_cells doubles (x y z of cell c at 3c), the sites the tessellation was
 * built from; centroid_xyz: 3 * n_cells doubles out; volume: n_cells doubles out, or NULL.  All arrays live
 * where cells->on_device says.  A cell of zero volume keeps its site.  Malformed offsets: KIWI_ERR_INVALID_ARG. */
int kiwi_hip_voronoi_centroids(kiwi_hip_ctx* ctx, const kiwi_voronoi_cells* cells, const double* site_xyz,
                               double* centroid_xyz, double* volume);

/* A point set: n positions, host or device memory. */
typedef struct kiwi_point_set {
    int64_t n;
    const double *x, *y, *z;
    int on_device;
} kiwi_point_set;

/* The last tessellation of a context: device arrays owned by the context, valid until its next
 * kiwi_hip_voronoi_tessellate call or kiwi_hip_destroy. */
typedef struct kiwi_tessellation {
    kiwi_voronoi_cells cells; /* on_device = 1; domain = the tessellated box */
    const int* cell_site;     /* n_cells: the input index of each cell's site */
    int64_t n_dropped;        /* sites the container drops (outside the box, or on a max face by voro++'s
                                 rounding of its block index) */
    double device_ms_prep, device_ms_count, device_ms_fill; /* cell order and buckets; count pass; fill pass */
} kiwi_tessellation;

/* The Voronoi tessellation of the sites inside domain[6] (non-periodic), replacing the voro++ container
 * the reference fills and walks in build_cell_geometry (src/mapping.cpp:222-269) and
 * run_tessellation_stage (src/tessellation.cpp:23-61): each cell cut from the domain box by its
 * neighbours' bisector planes with voro++'s plane tolerance (voronoicell_base, default cell size), in
 * FP64.  Cells come in voro++'s c_loop_all order over its block grid (pre_container::guess_optimal scaled
 * by the default block_scale_factor 3, include/kiwi/voro_grid.hpp:18-26): the order of the mapping.dat
 * rows.  Sites are kept as voro++'s container keeps them.  Faces and neighbours as kiwi_voronoi_cells,
 * each face's vertex cycle in voro++'s orientation (it may start at another vertex).  Non-finite
 * coordinates: KIWI_ERR_INVALID_ARG; a cell beyond 384 vertices, 192 faces or 48 vertices on a face:
 * KIWI_ERR_LIMIT. */
int kiwi_hip_voronoi_tessellate(kiwi_hip_ctx* ctx, const kiwi_point_set* sites, const double domain[6],
                                kiwi_tessellation* out);

/* Each point's home cell in the context's last tessellation, replacing voro++'s find_voronoi_cell and
 * the reverse_id_list lookup of the map stage (src/mapping.cpp:434-440): the cell of its nearest site by
 * squared distance (formed as voro++ forms it, each product and sum rounded on its own; of sites at
 * exactly equal distance the lowest input index wins, where voro++ keeps the first its block search
 * meets), or -1 for a point the container does not hold (outside the box, or on a max face by the
 * rounding of its block index) or when there are no cells.  home_cell[n] where points->on_device says;
 * unresolved (may be NULL): the count of -1s. */
int kiwi_hip_voronoi_home_cells(kiwi_hip_ctx* ctx, const kiwi_point_set* points, int* home_cell, int64_t* unresolved);

/* Copies the context's last tessellation into the caller's arrays (sizes from its kiwi_tessellation):
 * cell_face_ptr[n_cells + 1], face_vert_ptr[n_faces + 1], vert_xyz[3 n_face_vertices],
 * face_neighbor[n_faces], cell_site[n_cells]; device memory if on_device, else host memory.  The voro++
 * counterpart is the per-cell face_vertices / vertices / neighbors read-out of build_cell_geometry
 * (src/mapping.cpp:249-266). */
int kiwi_hip_voronoi_tessellation_fetch(kiwi_hip_ctx* ctx, int64_t* cell_face_ptr, int64_t* face_vert_ptr,
                                        double* vert_xyz, int* face_neighbor, int* cell_site, int on_device);

/* Per-particle rows of the inverse-mapping test. */
typedef struct kiwi_inverse_rows {
    double* m_recovered;       /* n: sum over visited cells with M_c > 0 of (Msum m / M_c) M_c */
    double* coverage_fraction; /* n: m_recovered / m (0 for m <= 0) */
    int on_device;             /* outputs and cell_mass in device memory */
} kiwi_inverse_rows;

/* The inverse-mapping test, replacing the particle loop of run_inverse_mapping_test_stage
 * (src/mapping.cpp:540-701): the map's visits and clamped cell integrals per particle, weighted back by
 * the forward map's cell masses.  cell_mass[n_cells]: density x volume of each cell from the forward
 * map (mapping.dat, :553-559), in the cell order of `cells`, where rows->on_device says.  Particles with
 * h <= 0 or no home cell get 0.  Precision as kiwi_hip_voronoi_map (in KIWI_MIXED no per-pair clamp, so
 * that single-precision noise around 0 does not bias the particle's sum). */
int kiwi_hip_inverse_mapping(kiwi_hip_ctx* ctx, const kiwi_particles* particles, const int* home_cell,
                             const kiwi_voronoi_cells* cells, const kiwi_kernel* kernel, int precision,
                             const double* cell_mass, const kiwi_inverse_rows* rows, kiwi_map_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* KIWI_HIP_H */
