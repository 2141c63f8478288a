// Every template kernel the library launches, listed once. instances.hip is compiled once per
// group (-DYPT_INST_GROUP=n), so the device code builds in parallel: each compilation defines one
// group's instances with YPT_INST_<group>(template), and the host unit (ypt_lib.hip) declares all of
// them with YPT_INST_ALL(extern template), which keeps its own device compilation to the
// non-template kernels.
#pragma once
#include "bdpt_kernels.h"
#include "multijob.h"
#include "wavefront.h"

namespace ypt {

// instance groups compiled separately (instances.hip)
static constexpr int YPT_INST_GROUPS = 11;

#define YPT_TRAV_ARGS_EXT DevScene, PathSoA, Queues, int, uint32_t
#define YPT_TRAV_ARGS_SHADOW DevScene, PathSoA, Queues, uint32_t
#define YPT_TRAV_ARGS_EXT_P DevScene, PathSoA, Queues, int, uint32_t, uint32_t, uint32_t
#define YPT_TRAV_ARGS_SHADOW_P DevScene, PathSoA, Queues, uint32_t, uint32_t, uint32_t

// closest-hit and occlusion traversal of one BVH width
#define YPT_INST_TRAV(EXT, W)                                                                          \
  EXT __global__ void k_extend<W, true>(YPT_TRAV_ARGS_EXT);                                            \
  EXT __global__ void k_extend<W, false>(YPT_TRAV_ARGS_EXT);                                           \
  EXT __global__ void k_shadow<W, true>(YPT_TRAV_ARGS_SHADOW);                                         \
  EXT __global__ void k_shadow<W, false>(YPT_TRAV_ARGS_SHADOW);                                        \
  EXT __global__ void k_extend_p<W, true, true>(YPT_TRAV_ARGS_EXT_P);                                  \
  EXT __global__ void k_extend_p<W, false, true>(YPT_TRAV_ARGS_EXT_P);                                 \
  EXT __global__ void k_extend_p<W, false, false>(YPT_TRAV_ARGS_EXT_P);                                \
  EXT __global__ void k_shadow_p<W, true, true, true>(YPT_TRAV_ARGS_SHADOW_P);                         \
  EXT __global__ void k_shadow_p<W, false, true, true>(YPT_TRAV_ARGS_SHADOW_P);                        \
  EXT __global__ void k_shadow_p<W, true, false, true>(YPT_TRAV_ARGS_SHADOW_P);                        \
  EXT __global__ void k_shadow_p<W, false, false, true>(YPT_TRAV_ARGS_SHADOW_P);                       \
  EXT __global__ void k_shadow_p<W, false, false, false>(YPT_TRAV_ARGS_SHADOW_P);                      \
  EXT __global__ void k_extend_w<W, true, true>(DevScene, PathSoA, Queues, int, uint32_t);                \
  EXT __global__ void k_extend_w<W, false, true>(DevScene, PathSoA, Queues, int, uint32_t);               \
  EXT __global__ void k_extend_w<W, false, false>(DevScene, PathSoA, Queues, int, uint32_t);              \
  EXT __global__ void k_shadow_w<W, true, true>(DevScene, PathSoA, Queues, uint32_t);                     \
  EXT __global__ void k_shadow_w<W, false, true>(DevScene, PathSoA, Queues, uint32_t);                    \
  EXT __global__ void k_shadow_w<W, false, false>(DevScene, PathSoA, Queues, uint32_t);                   \
  EXT __global__ void k_trace_closest<W>(DevScene, const float*, uint32_t, float*, uint32_t*, float*, uint32_t*, uint32_t);  \
  EXT __global__ void k_trace_closest_w<W>(DevScene, const float*, uint32_t, float*, uint32_t*, float*, uint32_t*, uint32_t);

#define YPT_SHADE_ARGS FrameConstants, DevScene, const uint32_t*, uint32_t, PathSoA, Queues, FilmBuffers, int, int
#define YPT_INST_SHADE(EXT, NEE)                                      \
  EXT __global__ void k_shade<NEE, 2, false>(YPT_SHADE_ARGS);         \
  EXT __global__ void k_shade<NEE, 4, false>(YPT_SHADE_ARGS);         \
  EXT __global__ void k_shade<NEE, 8, false>(YPT_SHADE_ARGS);         \
  EXT __global__ void k_shade<NEE, 2, true>(YPT_SHADE_ARGS);

#define YPT_BDPT_ARGS FrameConstants, DevScene, const uint32_t*, BdptBuffers, uint32_t
#define YPT_BDPT_CAM_ARGS FrameConstants, DevScene, const uint32_t*, BdptBuffers, float*, uint32_t
// light subpaths and light tracing of one BVH width
#define YPT_INST_BDPT_LIGHT(EXT, W)                                        \
  EXT __global__ void k_bdpt_light<W, true, false>(YPT_BDPT_ARGS);         \
  EXT __global__ void k_bdpt_light<W, false, false>(YPT_BDPT_ARGS);        \
  EXT __global__ void k_bdpt_light<W, true, true>(YPT_BDPT_ARGS);          \
  EXT __global__ void k_bdpt_light<W, false, true>(YPT_BDPT_ARGS);         \
  EXT __global__ void k_bdpt_splat<W, true>(YPT_BDPT_ARGS);                \
  EXT __global__ void k_bdpt_splat<W, false>(YPT_BDPT_ARGS);

// camera-subpath megakernel of one BVH width and counting mode, at every register budget
#define YPT_INST_BDPT_CAMERA(EXT, W, C)                                    \
  EXT __global__ void k_bdpt_camera<W, C, 1, false>(YPT_BDPT_CAM_ARGS);    \
  EXT __global__ void k_bdpt_camera<W, C, 2, false>(YPT_BDPT_CAM_ARGS);    \
  EXT __global__ void k_bdpt_camera<W, C, 3, false>(YPT_BDPT_CAM_ARGS);    \
  EXT __global__ void k_bdpt_camera<W, C, 4, false>(YPT_BDPT_CAM_ARGS);    \
  EXT __global__ void k_bdpt_camera<W, C, 6, false>(YPT_BDPT_CAM_ARGS);    \
  EXT __global__ void k_bdpt_camera<W, C, 8, false>(YPT_BDPT_CAM_ARGS);    \
  EXT __global__ void k_bdpt_camera<W, C, 1, true>(YPT_BDPT_CAM_ARGS);

#define YPT_CSHADE_ARGS FrameConstants, DevScene, const uint32_t*, BdptBuffers, CamSoA, Queues, int, int, float*
#define YPT_INST_BDPT_CSHADE(EXT)                                      \
  EXT __global__ void k_bdpt_cshade<1, false>(YPT_CSHADE_ARGS);        \
  EXT __global__ void k_bdpt_cshade<2, false>(YPT_CSHADE_ARGS);        \
  EXT __global__ void k_bdpt_cshade<4, false>(YPT_CSHADE_ARGS);        \
  EXT __global__ void k_bdpt_cshade<1, true>(YPT_CSHADE_ARGS);        \
  EXT __global__ void k_bdpt_cshade<1, false, true>(YPT_CSHADE_ARGS);  \
  EXT __global__ void k_bdpt_cshade<2, false, true>(YPT_CSHADE_ARGS);  \
  EXT __global__ void k_bdpt_connect<1>(FrameConstants, DevScene, BdptBuffers, CamSoA, Queues, int); \
  EXT __global__ void k_bdpt_connect<2>(FrameConstants, DevScene, BdptBuffers, CamSoA, Queues, int); \
  EXT __global__ void k_bdpt_connect<4>(FrameConstants, DevScene, BdptBuffers, CamSoA, Queues, int);

// kernels of the server's fused steps (several jobs per launch)
#define YPT_FUSED_TRAV_ARGS FusedArgs, PathSoA, Queues, int, uint32_t, uint32_t, uint32_t
#define YPT_FUSED_SHADOW_ARGS FusedArgs, PathSoA, Queues, uint32_t, uint32_t, uint32_t
#define YPT_INST_FUSED(EXT)                                             \
  EXT __global__ void k_extend_pm<2, true>(YPT_FUSED_TRAV_ARGS);        \
  EXT __global__ void k_extend_pm<2, false>(YPT_FUSED_TRAV_ARGS);       \
  EXT __global__ void k_extend_pm<8, true>(YPT_FUSED_TRAV_ARGS);        \
  EXT __global__ void k_extend_pm<8, false>(YPT_FUSED_TRAV_ARGS);       \
  EXT __global__ void k_shadow_pm<2, true>(YPT_FUSED_SHADOW_ARGS);      \
  EXT __global__ void k_shadow_pm<2, false>(YPT_FUSED_SHADOW_ARGS);     \
  EXT __global__ void k_shadow_pm<8, true>(YPT_FUSED_SHADOW_ARGS);      \
  EXT __global__ void k_shadow_pm<8, false>(YPT_FUSED_SHADOW_ARGS);     \
  EXT __global__ void k_shade_m<2>(FusedArgs, PathSoA, Queues, int);    \
  EXT __global__ void k_shade_m<4>(FusedArgs, PathSoA, Queues, int);    \
  EXT __global__ void k_shade_m<8>(FusedArgs, PathSoA, Queues, int);    \
  EXT __global__ void k_extend_wm<2, true>(FusedArgs, PathSoA, Queues, int, uint32_t);   \
  EXT __global__ void k_extend_wm<2, false>(FusedArgs, PathSoA, Queues, int, uint32_t);  \
  EXT __global__ void k_extend_wm<8, true>(FusedArgs, PathSoA, Queues, int, uint32_t);   \
  EXT __global__ void k_extend_wm<8, false>(FusedArgs, PathSoA, Queues, int, uint32_t);  \
  EXT __global__ void k_shadow_wm<2, true>(FusedArgs, PathSoA, Queues, uint32_t);        \
  EXT __global__ void k_shadow_wm<2, false>(FusedArgs, PathSoA, Queues, uint32_t);       \
  EXT __global__ void k_shadow_wm<8, true>(FusedArgs, PathSoA, Queues, uint32_t);        \
  EXT __global__ void k_shadow_wm<8, false>(FusedArgs, PathSoA, Queues, uint32_t);

#define YPT_INST_ALL(EXT)                                                                          \
  YPT_INST_TRAV(EXT, 2) YPT_INST_TRAV(EXT, 8) YPT_INST_SHADE(EXT, true) YPT_INST_SHADE(EXT, false) \
  YPT_INST_BDPT_LIGHT(EXT, 2) YPT_INST_BDPT_LIGHT(EXT, 8)                                          \
  YPT_INST_BDPT_CAMERA(EXT, 2, false) YPT_INST_BDPT_CAMERA(EXT, 2, true)                           \
  YPT_INST_BDPT_CAMERA(EXT, 8, false) YPT_INST_BDPT_CAMERA(EXT, 8, true) YPT_INST_BDPT_CSHADE(EXT)     \
  YPT_INST_FUSED(EXT)

}  // namespace ypt
