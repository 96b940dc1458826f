/* libkiwi_io: KIWI's SPH snapshot readers as a native host library (no GPU).  C ABI, plain pointers.
 *
 * Replaces load_sph_snapshot / detect_snapshot_format (src/snapshot.cpp:495-522) for the plain_ascii,
 * splash_ascii and phantom_binary formats: the same values bit for bit, the same rows kept, the same
 * first error message and the same warnings on stderr; ASCII lines are parsed by a pool of threads. */
#ifndef KIWI_IO_H
#define KIWI_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KIWI_IO_ABI_VERSION 2

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

#ifdef __cplusplus
}
#endif

#endif /* KIWI_IO_H */
