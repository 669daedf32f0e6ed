// Entry points declared in lddmm_amd.h whose kernels are not built yet.
// Each returns LDDMM_EUNSUPPORTED with a message.
#include "common.hpp"

using namespace lddmm;

#define LDDMM_PENDING(name)                                                    \
    do {                                                                       \
        set_error("%s is not built yet", name);                                \
        return LDDMM_EUNSUPPORTED;                                             \
    } while (0)

extern "C" {

int lddmm_surface_varifold(const lddmm_kernel*, const lddmm_surface*, const lddmm_surface*, double*,
                           double*, void*) { LDDMM_PENDING("lddmm_surface_varifold"); }

int lddmm_surface_features(const lddmm_surface*, double*, double*, double*, void*) {
    LDDMM_PENDING("lddmm_surface_features");
}

int lddmm_surface_face_gather(const lddmm_surface*, const double*, const double*, double*, void*) {
    LDDMM_PENDING("lddmm_surface_face_gather");
}

int lddmm_shoot_forward(const lddmm_kernel*, const double*, int32_t, int64_t, int32_t, double*, double,
                        double*, void*) { LDDMM_PENDING("lddmm_shoot_forward"); }

int lddmm_shoot_adjoint(const lddmm_kernel*, const double*, const double*, int32_t, int64_t, int32_t,
                        const double*, double, int32_t, double*, double*, void*) {
    LDDMM_PENDING("lddmm_shoot_adjoint");
}

int lddmm_epdiff_forward(const lddmm_kernel*, int32_t, int64_t, int32_t, double*, double*, void*) {
    LDDMM_PENDING("lddmm_epdiff_forward");
}

int lddmm_epdiff_adjoint(const lddmm_kernel*, int32_t, int64_t, int32_t, const double*, const double*,
                         const double*, const double*, double*, double*, double*, void*) {
    LDDMM_PENDING("lddmm_epdiff_adjoint");
}

}  // extern "C"
