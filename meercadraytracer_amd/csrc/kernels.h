// kernels.h — host-side launchers of the trace kernels (defined in trace.hip).
#ifndef MCR_KERNELS_H
#define MCR_KERNELS_H

#include <hip/hip_runtime.h>

#include "../../include/mcr.h"
#include "device.h"
#include "wavefront.h"

hipError_t mcr_upload_dfg(const float *lut512);
hipError_t mcr_upload_srgb8(const float *table256);
void launch_reset_slots(hipStream_t s, const Wavefront &w);
// mix: the run holds samples of more than one view (see with_view in trace.hip).
void launch_refill(hipStream_t s, bool mix, const ViewDev *views, const Wavefront &w, int q,
                   uint32_t count, uint32_t free_top, uint32_t next, uint32_t queued);
// Rewrites queue q (holding `live` slots, at most) in ascending slot order.
void launch_sort_queue(hipStream_t s, const Wavefront &w, int q, uint32_t live);
// wide: the run traces with the precision form's FP32 wide BVH (every view's scene has one).
void launch_extend(hipStream_t s, bool mix, bool wide, const ViewDev *views, const Wavefront &w, int q,
                   uint32_t grid, int32_t *spill, uint32_t spill_stride, DevCounters *cnt);
// Bins the live slots of queue q by shade class into Wavefront::bins (counts in C_BIN words).
void launch_shade_bin(hipStream_t s, bool mix, const ViewDev *views, const Wavefront &w, int q,
                      uint32_t max_live, DevCounters *cnt);
// binned: shade walks the bins of launch_shade_bin instead of queue q.
void launch_shade(hipStream_t s, bool mix, bool binned, const ViewDev *views, const Wavefront &w, int q,
                  uint32_t max_live, DevCounters *cnt);
void launch_shadow(hipStream_t s, bool mix, bool wide, bool tex, const ViewDev *views, const Wavefront &w, uint32_t grid,
                   int32_t *spill, uint32_t spill_stride, double *seen, int2 *seen_meta, double2 *seen_uv,
                   DevCounters *cnt);
// Retraces in the reference's order the shadow rays k_shadow marked LS_DEEP (spill_stride: the
// spill area's thread count, at least MCR_DEEP_THREADS).
void launch_shadow_deep(hipStream_t s, bool mix, bool tex, const ViewDev *views, const Wavefront &w, int32_t *spill,
                        uint32_t spill_stride, DevCounters *cnt);
void launch_resolve(hipStream_t s, bool mix, const ViewDev *views, const Wavefront &w, int q,
                    uint32_t max_live);
void launch_probe_closest(hipStream_t s, const DevScene &sc, const double *rays, uint32_t n,
                          double tmin, double tmax, int cull, int32_t *tri, double *t, double *u,
                          double *v, uint8_t *front, uint32_t grid, int32_t *spill,
                          uint32_t spill_stride, DevCounters *cnt, uint32_t *fetch);
void launch_probe_any_count(hipStream_t s, const DevScene &sc, const double *rays, uint32_t n,
                            double tmin, double tmax, uint32_t *count, uint32_t grid,
                            int32_t *spill, uint32_t spill_stride, DevCounters *cnt, uint32_t *fetch);
void launch_encode_batch(hipStream_t s, const EncodeItem *items, uint32_t n, uint32_t blocks);
void launch_encode(hipStream_t s, uint8_t *image, uint32_t width, const mcr_tile &tile,
                   const float *data, float *sum, double *sum64, uint32_t count, int srgb);

#endif
