// Harness of oracle/_ref: compiles the reference's own zmath.cpp where it lies
// (the Makefile's REF directory on the include path) and exposes two of its
// functions with C linkage over flat row-major arrays, as a second checker for
// the oracle's restatements:
//   ref_mtm  MTM(a) = a^T a          (zmath.cpp:260-262)
//   ref_inv  Gauss-Jordan inverse    (zmath.cpp:129-140); returns 0 when the
//            reference returned its zero matrix, else 1.
// TEST INFRASTRUCTURE ONLY: nothing in the product links or loads it.
#include "zmath.cpp"

extern "C" void ref_mtm(long rows, int cols, const double* a, double* out) {
    mat m((size_t)rows, vec((size_t)cols));
    for (long i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) m[i][j] = a[i * cols + j];
    const mat r = MTM(m);
    for (int i = 0; i < cols; ++i)
        for (int j = 0; j < cols; ++j) out[(size_t)i * cols + j] = r[i][j];
}

extern "C" int ref_inv(int n, const double* a, double* out) {
    mat m((size_t)n, vec((size_t)n));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) m[i][j] = a[(size_t)i * n + j];
    const mat r = inv(m);
    int nonzero = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            out[(size_t)i * n + j] = r[i][j];
            nonzero |= r[i][j] != 0.0;
        }
    return nonzero;
}
