// Launch wrappers of the MDI wrapper's kernels (mdi.hip), for the C ABI.
#pragma once
#include "common.h"

namespace arp {

hipError_t launch_mmse_sic(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                           int per_user, int8_t* xs, hipStream_t st);
hipError_t launch_mdi_radius(const float* H, const float* y, const int8_t* xg, int B, int Mr, int D, int M, int Rmax,
                             int32_t* R, float* Gam, float* dist, hipStream_t st);
hipError_t launch_mdi_build(const float* H, const float* y, const int8_t* xg, const int32_t* list, int n, int Mr, int D,
                            float Q, int32_t* Gi, int32_t* hi, hipStream_t st);

}  // namespace arp
