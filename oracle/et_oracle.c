/*
 * et_oracle.c -- CPU restatement of NiftyRec's SPECT projector path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product.  Built with -ffp-contract=off so the single-precision variant
 * keeps the reference's rounding (no fused multiply-adds).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "et_oracle.h"

#define REAL float
#define S f32
#define EXPR expf
#define SINR sinf
#define COSR cosf
#define FLOORR floorf
#include "et_oracle_impl.inc"
#undef REAL
#undef S
#undef EXPR
#undef SINR
#undef COSR
#undef FLOORR

#define REAL double
#define S f64
#define EXPR exp
#define SINR sin
#define COSR cos
#define FLOORR floor
#include "et_oracle_impl.inc"
#undef REAL
#undef S
#undef EXPR
#undef SINR
#undef COSR
#undef FLOORR

/* Camera subset draw of et_array_osem_step, _et_array_interface.cpp:981-1009:
 * two rand() calls per camera, R = base + fine/scale in single precision,
 * index = (unsigned)(V * R).  The reference clamps an index >= V to V (one past
 * the end); this restatement clamps to V - 1. */
int orc_osem_subset(int V, int n_sub, int *idx)
{
    float scale = (float)(RAND_MAX + 1.);
    for (int c = 0; c < n_sub; c++) {
        float base = rand() / scale;
        float fine = rand() / scale;
        float R = base + fine / scale;
        unsigned int i = (unsigned int)(V * R);
        if (i >= (unsigned int)V) i = V - 1;
        idx[c] = (int)i;
    }
    return 0;
}
