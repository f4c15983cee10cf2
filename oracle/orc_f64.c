/* ORACLE (test infrastructure only): the fp64 build of orc_core.inc, the
 * reference's own precision. */
#include "oracle.h"
#define T double
#define FN(x) x##_f64
#define ORC_IS_F32 0
#include "orc_core.inc"

/* Sampler entry points for the distribution tests. */
void orc_sample_ggx_f64(const double n[3], double rough, float u1, float u2,
                        double out[3])
{
    V r = ggx_sample(v3(n[0], n[1], n[2]), rough, u1, u2);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void orc_sample_cos_f64(const double n[3], float u1, float u2, double out[3])
{
    V r = cos_hemi(v3(n[0], n[1], n[2]), u1, u2);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void orc_unit_sphere_f64(const float u[4], double out[3])
{
    V r = unit_sphere_pt(u);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
