/* libkiwi_io: KIWI's SPH snapshot readers as a native host library (no GPU).  C ABI, plain pointers.
 *
 * Replaces load_sph_snapshot / detect_snapshot_format (src/snapshot.cpp:495-522) for the plain_ascii
 * and splash_ascii formats: the same values bit for bit, the same rows kept, the same first error
 * message and the same warnings on stderr, with the lines parsed by a pool of threads. */
#ifndef KIWI_IO_H
#define KIWI_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KIWI_IO_ABI_VERSION 1

enum {
    KIWI_IO_OK = 0,
    KIWI_IO_ERR_INVALID_ARG = 1, /* null argument, unknown format name */
    KIWI_IO_ERR_READ = 2,        /* the file cannot be opened, or host memory ran out */
    KIWI_IO_ERR_FORMAT = 3,      /* the reference's reader would throw: the message is its text */
    KIWI_IO_ERR_UNSUPPORTED = 4  /* phantom_binary: read with the reference's load_phantom_dump */
};

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

/* load_sph_snapshot (src/snapshot.cpp:511-522) for format "auto", "plain_ascii" or "splash_ascii";
 * hfact as Config::kernel_hfact() gives it (used only by splash_ascii files with one of h and
 * density).  On failure *out is empty and kiwi_io_last_error() holds the reference's message. */
int kiwi_io_load_snapshot(const char* path, const char* format, double hfact, kiwi_snapshot* out);

void kiwi_io_free_snapshot(kiwi_snapshot* snapshot);

#ifdef __cplusplus
}
#endif

#endif /* KIWI_IO_H */
