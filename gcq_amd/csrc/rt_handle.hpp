// rt_handle.hpp — the gcq_rt_scene handle of the C ABI: the scene's host
// copy, its lazily uploaded HBM copy and the render context of rt.hip.
#pragma once
#include <vector>

#include "rt_shade.hpp"

struct gcq_rt_scene {
  std::vector<gcq_rt_sphere> sph;
  std::vector<gcq_rt_plane> pl;
  gcq_rt_scene_desc desc;  // raw camera / light, pointers into sph / pl
  // device side (lazy)
  gcq_rt_sphere* d_sph = nullptr;
  gcq_rt_plane* d_pl = nullptr;
  rtk::PrimaryTerms* d_pre = nullptr;  // the camera origin's terms (scene_upload)
  hipStream_t stream = nullptr;
  // render context
  int kind = -1;
  uint32_t ntiles = 0, cap = 0, threads = 0;
  uint64_t qstride = 0, hstride = 0, ymc_cells = 0;
  char* qpool = nullptr;
  char* hpool = nullptr;
  void* adapters = nullptr;
  rtk::RayRec* recs = nullptr;
  uint64_t recs_cap = 0;
  float* img = nullptr;
  uint64_t img_cap = 0;
  uint64_t* counters = nullptr;  // alive, traced per tile
  uint32_t counters_cap = 0;
  uint32_t* err = nullptr;
  // stream-compaction renderer (rt_compact.hip)
  rtk::RayRec* cbuf[2] = {nullptr, nullptr};  // this bounce's rays, the next bounce's
  uint64_t cbuf_cap = 0;                      // records per buffer
  uint64_t* cstatus = nullptr;                // per-block scan status
  uint32_t cstatus_cap = 0;
  uint32_t* cctl = nullptr;                   // block-id ticket, next-bounce count
  float* cimg = nullptr;
  uint64_t cimg_cap = 0;
};

namespace rtk {
// Upload the scene to HBM on first use; returns a gcq status.
int scene_upload(gcq_rt_scene* s);
// The kernel-side view of an uploaded scene (camera normalised on the host,
// rt_final.cpp:851-859).
Scene scene_view(const gcq_rt_scene* s);
// Release the stream-compaction renderer's buffers.
void free_compact(gcq_rt_scene* s);
}  // namespace rtk
