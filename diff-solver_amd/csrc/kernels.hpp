// kernels.hpp — kernel argument bundles and the launch wrappers shared by the operator,
// step and trajectory translation units.
#pragma once

#include "context.hpp"

namespace ds {

struct F3 {
    float* p[3];
};
struct CF3 {
    const float* p[3];
};
struct Params {   // device pointers to one float each (Parameters, data.ixx:129-134)
    const float* ambient;
    const float* alpha_rho;
    const float* alpha_t;
};

Params params_of(const ds_parameters* p);
CF3 collider_of(const ds_context* c);

void launch_source_buoyancy(ds_context* c, const float* rho, const float* T, const float* s_rho, const float* s_t, CF3 e,
                            const Params& prm, float* rho_s, float* t_s, F3 f);
void launch_integrate(ds_context* c, CF3 u, CF3 f, F3 out);
void launch_advect_velocity(ds_context* c, CF3 u, F3 out, bool constrain);
void launch_project(ds_context* c, CF3 u, const float* ph, const float* pl, F3 out);
void launch_advect_scalars(ds_context* c, int nf, int first_field, const float* src0, const float* src1, CF3 u, float* out0,
                           float* out1);
void copy_pair(ds_context* c, const float* hi, const float* lo, float* out_hi, float* out_lo);

}  // namespace ds
