/* ORACLE (test infrastructure only): the fp32 build of orc_core.inc. */
#include "oracle.h"
#define T float
#define FN(x) x##_f32
#define ORC_IS_F32 1
#include "orc_core.inc"
