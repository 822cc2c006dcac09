// ref_matrix_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points over the reference's own `matrix` class
// (include/matrix.h:7-26, src/matrix.cpp:6-155 of EricYJA/2D_PD_GPU), compiled
// from the reference sources where they lie by oracle/Makefile into
// oracle/_ref/libref_matrix.so.  Used by tests/test_oracle.py to check the
// oracle's long-double 2x2 arithmetic (ref_mat_inverse2d, ref_mat_mul)
// bit-for-bit against the reference implementation.
#include <vector>
#include "matrix.h"

static matrix make2(const long double* a)
{
    return matrix(2, std::vector<long double>(a, a + 4));
}

static void store2(const matrix& m, long double* out)
{
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) out[2 * i + j] = m.elements[i][j];
}

extern "C" void refm_inverse2d(const long double* a, long double* out)
{
    matrix m = make2(a);
    store2(m.inverse2D(), out);
}

extern "C" void refm_mul(const long double* a, const long double* b, long double* out)
{
    matrix ma = make2(a), mb = make2(b);
    store2(ma.timeMatrix(mb), out);
}

extern "C" void refm_time_vector3(const long double* c9, const long double* v3, long double* out)
{
    matrix m(3, std::vector<long double>(c9, c9 + 9));
    std::vector<long double> r = m.timeVector(std::vector<long double>(v3, v3 + 3));
    for (int i = 0; i < 3; ++i) out[i] = r[i];
}
