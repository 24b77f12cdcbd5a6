/* Scalar C restatement of the jax_tinyfluids step -- TEST INFRASTRUCTURE ONLY.
 * Built by oracle/Makefile into oracle/libtforacle.so; loaded only by tests/,
 * the smoke check and the CPU-baseline leg of the benchmark, as the checker.
 * The float64 symbols (tfo_f64_*) are the oracle; the float32 symbols
 * (tfo_f32_*) reproduce the reference's own precision. */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stddef.h>

#define REAL double
#define SUFFIX f64
#define SQRT sqrt
#include "tf_oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef SQRT

#define REAL float
#define SUFFIX f32
#define SQRT sqrtf
#include "tf_oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef SQRT

int tfo_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
