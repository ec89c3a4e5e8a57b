// builder.h — GPU binned-SAH BVH builder, bit-identical to the reference's buildBvh
// (src/bvh.ts:16-178). See bvh_build.hip.
#ifndef MCR_BUILDER_H
#define MCR_BUILDER_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

struct BuilderWorkspace {
  void *block = nullptr;   // one allocation for every buffer below
  size_t bytes = 0;
  uint32_t capacity = 0;   // triangles the workspace is sized for
  void *scan_tmp = nullptr;
  size_t scan_bytes = 0;
  uint32_t *host = nullptr;  // pinned readback words
};

void builder_free(BuilderWorkspace &ws);
// Returns MCR_OK or an MCR_E_* code with `err` set. All arrays are device pointers.
int builder_run(BuilderWorkspace &ws, hipStream_t stream, const float *tris, uint32_t n, float *bounds,
                int32_t *nodes, uint32_t *order, uint32_t *node_count, std::string &err);

#endif
