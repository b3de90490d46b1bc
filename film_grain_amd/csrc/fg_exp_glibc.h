// fg_exp_glibc.h -- device exponentials that reproduce the C library's
// (glibc) results the reference computes with: exp(-x) exactly on the Knuth
// threshold's domain (f32 x in (0, 12)) through fg_exp_dd and the generated
// table of the inputs where the two differ, and fg_exp_dd elsewhere.
#ifndef FG_EXP_GLIBC_H
#define FG_EXP_GLIBC_H

#include "fg_exp.h"
#include "fg_exp_neg_table.h"  // generated at build time (gen_exp_table.cpp)

namespace fg {

#if defined(__HIPCC__)
static __device__ const uint32_t d_exp_neg_table[FG_EXP_NEG_TABLE_N] = {FG_EXP_NEG_TABLE_VALUES};

// exp(-(double)x) as glibc computes it, for f32 x in (0, 12).
__device__ __forceinline__ double exp_neg_glibc(float x)
{
    return fg_exp_neg_f32(x, d_exp_neg_table, FG_EXP_NEG_TABLE_N);
}
#endif

}  // namespace fg

#endif
