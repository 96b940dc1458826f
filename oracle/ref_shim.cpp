// C-ABI shim over the reference KIWI kernel math, linked only into oracle/_ref/libkiwi_ref.so.
//
// Test infrastructure: it lets the Python tests (ctypes) call the reference's own
// full_integral_new / compute_face_integral / generic_full_integral / compute_face_integral_generic
// (include/kiwi/kernel_integral.hpp:181,283; include/kiwi/kernel_integral_generic.hpp:56;
// include/kiwi/kernel_integral.hpp:293) on inputs chosen by the test, to pin the oracle's
// restatement of them.  Built from the reference sources in place by oracle/ref.mk; never shipped.
#include <vector>

#include "kiwi/kernel_integral.hpp"
#include "kiwi/kernel_integral_generic.hpp"
#include "kiwi/kernel_spec.hpp"

extern "C" {

double ref_m4_vertex(double h, double phi, double r0, double R0) {
    return kiwi::full_integral_new(h, phi, r0, R0);
}

double ref_m4_face(double x0, double y0, double z0, double h, const double* verts, int nverts, int* out_vert) {
    std::vector<double> v(verts, verts + 3 * nverts);
    return kiwi::compute_face_integral(x0, y0, z0, h, false, nullptr, v, out_vert);
}

double ref_generic_vertex(int kernel_index, double H, double phi, double r0, double R0) {
    const kiwi::KernelSpec& K = *kiwi::all_kernels().at(kernel_index);
    return kiwi::generic_full_integral(K, H, phi, r0, R0);
}

double ref_generic_face(int kernel_index, double x0, double y0, double z0, double H, const double* verts,
                        int nverts, int* out_vert) {
    const kiwi::KernelSpec& K = *kiwi::all_kernels().at(kernel_index);
    std::vector<double> v(verts, verts + 3 * nverts);
    return kiwi::compute_face_integral_generic(x0, y0, z0, K, H, v, out_vert);
}

// Copies one KernelSpec's numbers out (c[4], A2_pi, A3_pi, radkern, hfact, a[4][15], B[4][4][16]) so
// the tests can compare our own exact re-derivation of the tables with the reference's generated ones.
int ref_kernel_spec(int kernel_index, int* nreg, double* c4, double* norm4, double* a60, double* b256) {
    const auto& all = kiwi::all_kernels();
    if (kernel_index < 0 || kernel_index >= static_cast<int>(all.size())) return -1;
    const kiwi::KernelSpec& K = *all[kernel_index];
    *nreg = K.nreg;
    for (int i = 0; i < 4; ++i) c4[i] = K.c[i];
    norm4[0] = K.A2_pi;
    norm4[1] = K.A3_pi;
    norm4[2] = K.radkern;
    norm4[3] = K.hfact;
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 15; ++i) a60[k * 15 + i] = K.a[k][i];
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k)
            for (int p = 0; p < 16; ++p) b256[(j * 4 + k) * 16 + p] = K.B[j][k][p];
    return 0;
}

// Tabulated cubic path: builds the reference's in-memory table once, then gridded_integral().
double ref_m4_gridded(double h, double phi, double r0, double R0) {
    static const kiwi::DensityLookupTable table = kiwi::generate_density_table();
    return kiwi::gridded_integral(table, h, phi, r0, R0);
}

}  // extern "C"
