// kernels.h — host-side launchers of the trace kernels (defined in trace.hip).
#ifndef MCR_KERNELS_H
#define MCR_KERNELS_H

#include <hip/hip_runtime.h>

#include "../../include/mcr.h"
#include "device.h"
#include "wavefront.h"

hipError_t mcr_upload_dfg(const float *lut512);
void launch_raygen(hipStream_t s, const JobDev *jobs, uint32_t njobs, uint32_t count,
                   const Wavefront &w, const mcr_camera &cam, uint32_t seed);
void launch_extend(hipStream_t s, const DevScene &sc, const Wavefront &w, int q, uint32_t grid,
                   int32_t *spill, uint32_t spill_stride, DevCounters *cnt);
void launch_shade(hipStream_t s, const DevScene &sc, const Wavefront &w, int q, uint32_t max_live,
                  const mcr_settings &st, const mcr_camera &cam, const JobDev *jobs, uint32_t njobs,
                  float *out);
void launch_shadow(hipStream_t s, const DevScene &sc, const Wavefront &w, int cast_only,
                   uint32_t grid, int32_t *spill, uint32_t spill_stride, double *seen,
                   int2 *seen_meta, DevCounters *cnt);
void launch_resolve(hipStream_t s, const Wavefront &w, int q, uint32_t max_live,
                    const mcr_settings &st, const DevScene &sc, const JobDev *jobs, uint32_t njobs,
                    float *out, const mcr_camera &cam, uint32_t total);
void launch_probe_closest(hipStream_t s, const DevScene &sc, const double *rays, uint32_t n,
                          double tmin, double tmax, int cull, int32_t *tri, double *t, double *u,
                          double *v, uint8_t *front, uint32_t grid, int32_t *spill,
                          uint32_t spill_stride, DevCounters *cnt);
void launch_probe_any_count(hipStream_t s, const DevScene &sc, const double *rays, uint32_t n,
                            double tmin, double tmax, uint32_t *count, uint32_t grid,
                            int32_t *spill, uint32_t spill_stride, DevCounters *cnt);
void launch_encode(hipStream_t s, uint8_t *image, uint32_t width, const mcr_tile &tile,
                   const float *data, float *sum, uint32_t count, int srgb);

#endif
