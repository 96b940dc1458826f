/* libkiwi_io: KIWI's SPH snapshot readers, run configuration and provenance manifest as a native host
 * library (no GPU).  C ABI, plain pointers.
 *
 * Replaces load_sph_snapshot / detect_snapshot_format (src/snapshot.cpp:495-522) for the plain_ascii,
 * splash_ascii and phantom_binary formats: the same values bit for bit, the same rows kept, the same
 * first error message and the same warnings on stderr; ASCII lines are parsed by a pool of threads.
 * Replaces Config::load_from_file (src/config.cpp:59-206, the same keys, defaults and error messages)
 * and write_manifest / check_manifest_consistency (src/manifest.cpp:147-199, the same bytes). */
#ifndef KIWI_IO_H
#define KIWI_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KIWI_IO_ABI_VERSION 3

enum {
    KIWI_IO_OK = 0,
    KIWI_IO_ERR_INVALID_ARG = 1, /* null argument, unknown format name */
    KIWI_IO_ERR_READ = 2,        /* the file cannot be opened, or host memory ran out */
    KIWI_IO_ERR_FORMAT = 3       /* the reference's reader would throw: the message is its text */
};

/* PhantomUnits (include/kiwi/config.hpp:248-256): the output units of a phantom_binary read. */
typedef struct kiwi_phantom_units {
    double length_cm, mass_g, velocity_cms;
} kiwi_phantom_units;

/* SPHSnapshot (include/kiwi/snapshot.hpp:70-77): n rows; id is the 1-based row number among the rows
 * kept; rho as the file holds it (splash_ascii: g/cm^3, derived from h when the file has no density
 * column, and h from density in the converse case, src/snapshot.cpp:418-444).  Arrays are malloc'ed by
 * the library and released by kiwi_io_free_snapshot. */
typedef struct kiwi_snapshot {
    int64_t n;
    int* id;
    double *x, *y, *z, *vx, *vy, *vz, *m, *h, *rho;
} kiwi_snapshot;

int kiwi_io_abi_version(void);

/* The message of the calling thread's last failed call ("" after a success). */
const char* kiwi_io_last_error(void);

/* Worker threads for the row parse (0: the hardware's count); at most 16.  Returns the count in use. */
int kiwi_io_set_threads(int n);

/* detect_snapshot_format (src/snapshot.cpp:495-509): writes "phantom_binary", "splash_ascii" or
 * "plain_ascii" into format (capacity >= 16). */
int kiwi_io_detect_format(const char* path, char* format, int capacity);

/* load_sph_snapshot (src/snapshot.cpp:511-522) for format "auto", "plain_ascii", "splash_ascii" or
 * "phantom_binary"; hfact as Config::kernel_hfact() gives it (used only by splash_ascii files with one
 * of h and density); phantom_binary with the default PhantomUnits (pc, Msun, km/s) and no missing
 * velocities allowed.  On failure *out is empty and kiwi_io_last_error() holds the reference's
 * message. */
int kiwi_io_load_snapshot(const char* path, const char* format, double hfact, kiwi_snapshot* out);

/* The same with the phantom_binary options: units (NULL: the defaults) and allow_missing_velocity
 * (src/phantom_dump.cpp:270-285). */
int kiwi_io_load_snapshot_ex(const char* path, const char* format, double hfact, const kiwi_phantom_units* units,
                             int allow_missing_velocity, kiwi_snapshot* out);

void kiwi_io_free_snapshot(kiwi_snapshot* snapshot);

/* A stage configuration as the reference's Config holds it (include/kiwi/config.hpp:258-305), with the
 * derived values the stages use resolved: the kernel (Config::kernel_spec, kernel_support_in_h,
 * kernel_hfact, use_generic_kernel_path, src/config.cpp:267-277), the grid's cells per axis
 * (resolution_x/y/z, 0 falling back to resolution, src/cartesian_grid.cpp:89-97) and the run directory
 * (work_folder/<snapshot basename without extension>, :279-287). */
typedef struct kiwi_config {
    char run_name[256];
    char input_snapshot[4096];
    char input_snapshot_format[32];
    char work_folder[4096];
    char run_dir[4096];
    double domain[6]; /* x_min x_max y_min y_max z_min z_max */
    int kernel_id;    /* 0..5: cubic, quartic, quintic spline, Wendland C2, C4, C6 */
    double support_in_h, hfact;
    int generic_path;
    int grid_nx, grid_ny, grid_nz;
    int grid_tabulated, mapping_tabulated;
    int buffer_enabled;
    char buffer_strategy[64];
    double coverage_spacing_fraction, coverage_radius_factor;
    int lloyd_iterations;
    int cartesian_map_resolution;
    double block_scale_factor;
    kiwi_phantom_units phantom_units;
    int allow_missing_velocity;
} kiwi_config;

/* Config::load_from_file.  On failure (KIWI_IO_ERR_FORMAT) kiwi_config_last_error() holds the text the
 * reference prints after "[kiwi] error: ". */
int kiwi_config_load(const char* path, kiwi_config* out);
const char* kiwi_config_last_error(void);

/* Config::to_json as compact JSON (nlohmann::json::dump()): *length set to its size; the text is copied
 * into buf (NUL-terminated, truncated to capacity - 1) when buf is not NULL. */
int kiwi_config_json(const char* path, char* buf, int64_t capacity, int64_t* length);

/* write_manifest: <run_dir>/manifest.json with the reference's schema, key order, number formats and
 * file fingerprints (64-bit FNV-1a, files above 64 MB by 32 MB from each end and the size). */
int kiwi_manifest_write(const char* config_path);

/* check_manifest_consistency: 1 when <run_dir>/manifest.json records a different config (the
 * reference's warning is printed to stderr), 0 otherwise, negative on a config error. */
int kiwi_manifest_check(const char* config_path);

#ifdef __cplusplus
}
#endif

#endif /* KIWI_IO_H */
