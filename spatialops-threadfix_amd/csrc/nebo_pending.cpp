// nebo_pending.cpp -- entry points whose kernels are not built yet: they fail
// loudly with NEBO_ERR_UNSUPPORTED instead of computing anything.
#include "nebo_internal.h"

using namespace nebo;

static int pending(const char* what)
{
  clear_error();
  return set_error(NEBO_ERR_UNSUPPORTED, what, "not provided by this build");
}

extern "C" {
int nebo_diffusion_step(const nebo_field_t*, const nebo_field_t*, const nebo_field_t*,
                        const nebo_axis_coefs_t*, double, void*) { return pending("Nebo Adjoint"); }
int nebo_diffusion_adjoint_step(const nebo_field_t*, double*, const nebo_field_t*,
                                const nebo_field_t*, const nebo_field_t*,
                                const nebo_axis_coefs_t*, double, void*) { return pending("Nebo Adjoint"); }
int nebo_diffusion_forward(const nebo_field_t*, const nebo_field_t*, const nebo_field_t*,
                           const nebo_axis_coefs_t*, double, int32_t, int32_t, double*,
                           void*) { return pending("Nebo Adjoint"); }
int nebo_diffusion_backward(const nebo_field_t*, double*, const nebo_field_t*,
                            const nebo_field_t*, const nebo_axis_coefs_t*, double, int32_t,
                            int32_t, const double*, double*, void*) { return pending("Nebo Adjoint"); }
}
