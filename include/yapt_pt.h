/* C ABI of the MI355X spectral path-tracing core.
 *
 * One handle = one path integrator instance bound to one GPU, the role of
 * YAPT::PathIntegratorSubStage (Source/Renderer/inc/Renderer/Shared/RenderPipeline/PathTracer/
 * PathIntegratorSubStage.h:17-72), which PathTraceStage constructs by type
 * (Source/Renderer/src/Shared/RenderPipeline/PathTracer/PathTraceStage.cpp:40-54) and drives
 * through sceneChanged / update / getOutput. Every call returns 0 on success and a negative
 * YPT_E* code on failure; ypt_last_error() describes the last failure of the calling thread.
 * The caller owns host buffers; the library owns device buffers. A handle is used by one thread
 * at a time.
 */
#ifndef YAPT_PT_H
#define YAPT_PT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YPT_OK 0
#define YPT_E_INVALID (-1)   /* bad argument or state */
#define YPT_E_DEVICE (-2)    /* HIP runtime error */
#define YPT_E_MEMORY (-3)    /* device allocation failed */
#define YPT_E_NOSCENE (-4)   /* render before set_scene / set_frame */
#define YPT_E_CANCELLED (-5) /* render server: the job was cancelled */
#define YPT_E_TIMEOUT (-6)   /* render server: the job passed its deadline */

/* integrator types: RendererVarsList.h:24 option order of "RenderPipeline" */
#define YPT_INTEGRATOR_BPT 0
#define YPT_INTEGRATOR_NEE 1  /* BPT_WITH_NEE */
#define YPT_INTEGRATOR_BDPT 2

/* frame flags */
#define YPT_FLAG_DISABLE_JITTER (1u << 1)  /* CONSTANTS_FLAG_DISABLE_TEXEL_JITTER */
#define YPT_FLAG_WHITE_FURNACE (1u << 8)   /* WHITE_FURNACE_TEST as a runtime switch */
#define YPT_FLAG_BDPT_NODE_RANGES (1u << 9) /* use ypt_frame_desc.bdpt_node_ranges (else 0..128 for both) */

/* material_mask bits (MaterialMask_*, MaterialInternal.h / materialsCommon.hlsl:8-10); ALPHA_BLEND
 * turns the albedo texel alpha into coverage */
#define YPT_MATERIAL_TWO_SIDED (1u << 0)
#define YPT_MATERIAL_DISPERSION (1u << 1)
#define YPT_MATERIAL_ALPHA_BLEND (1u << 2)

/* Material parameters as the reference's MaterialParameters (Source/Renderer/inc/Renderer/Shared/
 * MaterialInternal.h:41-125); the library applies the validation of MaterialInternal.cpp:27-43 and
 * the packing clamps of BindlessMaterialManager.cpp:70-100. Colours are linear RGB. */
typedef struct ypt_material {
  float albedo[3];
  float transparency;
  float specular[3];
  float metalness;
  float absorption[3];
  float dielectric_ior;
  float emissive[3];
  float roughness;
  float anisotropy;
  float anisotropy_rotation;
  uint32_t material_mask;  /* YPT_MATERIAL_TWO_SIDED | YPT_MATERIAL_DISPERSION | YPT_MATERIAL_ALPHA_BLEND */
  float specular_amount;
  float clear_coat_amount;
  float clear_coat_ior;
  float clear_coat_roughness;
  float sheen_tint[3];
  float sheen_roughness;
  float sheen_amount;
  float thin_film_thickness_nm;
  float emission_focus;
  float alpha_cutoff;      /* > 0: hits where the albedo texture's alpha is below it are ignored */
  float cauchy_coeffs[2];
  float subsurface_radius[3];
  float subsurface_amount;
  float subsurface_radius_scale;
  /* textures (MaterialParameters albedoTex / normalTex / ormTex / emissiveTex): an index into
   * ypt_scene_desc.textures, or any value outside [0, texture_count) for none (TEX_UNBOUND_INDEX),
   * and the UV scale of each (the texture transform scale) */
  int32_t albedo_tex, normal_tex, orm_tex, emissive_tex;
  float albedo_tex_scale[2], normal_tex_scale[2], orm_tex_scale[2], emissive_tex_scale[2];
  /* Extensions with no counterpart in the reference (0 = off, the reference's behaviour):
   * blackbody_kelvin > 0 makes the emission a Planck spectrum of that temperature, scaled to 1 at
   * 560 nm times the largest emissive component; sellmeier_b[0] > 0 makes the dispersive IOR
   * n^2(lambda) = 1 + sum B_i lambda^2 / (lambda^2 - C_i) (lambda and sqrt(C_i) in micrometres)
   * instead of Cauchy's A + B / lambda^2. */
  float blackbody_kelvin;
  float sellmeier_b[3], sellmeier_c[3];
} ypt_material;

/* An RGBA8 texture, sampled bilinearly with wrapping at mip level 0 (g_colorSampler, SampleLevel 0).
 * srgb: the texels are sRGB encoded and decoded before filtering (R8G8B8A8_SRGB, used for colour);
 * 0 for data (normal, occlusion/roughness/metalness: R8G8B8A8_UNORM). */
typedef struct ypt_texture {
  uint32_t width, height;
  uint32_t srgb;
  const uint8_t* rgba8;  /* width * height * 4, row-major from v = 0 */
} ypt_texture;

/* Environment texture: the renderer variable World.Skycube (CRenderer.h:18), classified as
 * RaytraceCommonResources.cpp:285-313 does (ENVIRONMENT_TYPE_*, RaytraceCommonResources.h:17-19) and
 * looked up by getSkyBoxColor (materialsCommonResources.hlsl:181-223) on every escaping ray, times the
 * frame's env_intensity_scale. Linear RGBA float texels (the reference loads R16G16B16A16_SFLOAT
 * images), rows top to bottom, filtered bilinearly at mip level 0: a long-lat image with wrapping, at
 * (atan2(z, x) / 2pi + 1/2, acos(y) / pi); a cube map as six faces of width x height in the order
 * +X, -X, +Y, -Y, +Z, -Z with the D3D/Vulkan face orientation, clamped at face edges. */
#define YPT_ENVIRONMENT_NONE 0     /* |direction| as the colour, the reference's texture-less sky */
#define YPT_ENVIRONMENT_CUBE 1
#define YPT_ENVIRONMENT_LONGLAT 2
typedef struct ypt_environment {
  uint32_t type;          /* YPT_ENVIRONMENT_* */
  uint32_t width, height; /* of the long-lat image, or of one (square) cube face */
  const float* rgba;      /* width * height * 4 floats, six times that for a cube map */
} ypt_environment;

/* Punctual lights (lightTypes.h:6-8) */
#define YPT_LIGHT_POINT 1
#define YPT_LIGHT_SPOT 2
#define YPT_LIGHT_DIRECTIONAL 3

/* A punctual light: PunctualLightParameters (LightManager.h:15-23), filled as
 * LightManager::fillParameters does (LightManager.cpp:97-107): position and emission direction from
 * the light's transform (direction = transform * (0, 0, -1)), intensity = color * max(0, intensity),
 * and the cosines of the spot cone angles (setSpotConeAngles clamps outer to [0, pi/2] and inner to
 * [0, outer], LightProxy.cpp:49-54). The direction is normalized here; a zero one becomes (0, 0, -1). */
typedef struct ypt_punctual_light {
  uint32_t type;                 /* YPT_LIGHT_* */
  float position[3];             /* point, spot */
  float direction[3];            /* spot, directional */
  float intensity[3];
  float cos_inner, cos_outer;    /* spot */
} ypt_punctual_light;

/* Scene: world-space triangle soup (render-object transforms already applied, the
 * SceneData::objToWorld of PathIntegratorSubStage::sceneChanged), one record per instance
 * (render object x submesh) owning a contiguous triangle range. Mesh lights are the instances
 * whose material is emissive (LightManager.cpp:145). */
typedef struct ypt_scene_desc {
  const float* positions;       /* 3 * vertex_count */
  const float* normals;         /* 3 * vertex_count, or NULL for area-weighted vertex normals */
  const float* tangents;        /* 4 * vertex_count (xyz, handedness w), or NULL: tangent (1, 0, 0) as for
                                   meshes without a tangent buffer (hitShadersCommon.hlsl:198) */
  uint32_t vertex_count;
  const uint32_t* indices;      /* 3 * triangle_count */
  uint32_t triangle_count;
  const uint32_t* instance_first_triangle;
  const uint32_t* instance_triangle_count;
  const uint32_t* instance_material;
  uint32_t instance_count;
  const ypt_material* materials;
  uint32_t material_count;
  const ypt_punctual_light* punctual_lights;  /* may be NULL when punctual_light_count is 0 */
  uint32_t punctual_light_count;
  const float* uvs;              /* 2 * vertex_count, or NULL: UV (0.5, 0.5) as for meshes without UVs */
  const ypt_texture* textures;   /* may be NULL when texture_count is 0 */
  uint32_t texture_count;
  const ypt_environment* environment;  /* NULL: YPT_ENVIRONMENT_NONE. The reference holds it as a renderer
                                          variable; here it travels with the scene, so a resident scene that
                                          several handles or server jobs share carries its sky. */
} ypt_scene_desc;

/* Per-frame constants: the fields of RaytraceConstantData (raytraceCommonResources.hlsl:8-31) the
 * host fills in RaytraceCommonResources::update (RaytraceCommonResources.cpp:64-140), the film
 * merge parameters (mergeNewSamples.comp:7-19) and the integrator parameters. Matrices are
 * row-major and act on column vectors (out = M * v). */
typedef struct ypt_frame_desc {
  uint32_t width, height;
  float uv_to_view[16];
  float view_to_uv[16];
  float world_to_view[16];
  float view_to_world[16];
  float aperture_radius;
  float focus_distance;
  float forward_z_sign;
  float env_intensity_scale;
  uint32_t max_depth;            /* TraceRay segments per path */
  uint32_t regularize_after;     /* RegularizeAfterVertices (default 8) */
  uint32_t flags;                /* YPT_FLAG_* */
  float firefly_clamp_scale;     /* FireflyClampScale (default 12, 0 disables) */
  float firefly_clamp_relax_power;
  uint32_t kill_nan_samples;
  uint32_t light_path_vertices;  /* BDPT: vertices per light subpath incl. the light, 2..16
                                    (maxVerticesPerLightPath, BidirectionalPathIntegratorSubStage.cpp:482);
                                    0 = 16 */
  uint32_t bdpt_node_ranges[4];  /* with YPT_FLAG_BDPT_NODE_RANGES: the BDPT strategies that contribute, light
                                    vertices in [0]..[1] and camera vertices in [2]..[3] (the debug variables
                                    _DEBUG.BDPT.EffectiveLightPathNodes / EffectiveCameraPathNodes,
                                    RendererVarsList.h:46-47, applied by shouldPathContribute,
                                    BDPTCommon.hlsl:1217-1223) */
} ypt_frame_desc;

typedef struct ypt_integrator ypt_integrator;

/* PathTraceStage.cpp:40-54 (construct by type) + PathIntegratorSubStage::initialize */
int ypt_create(int integrator_type, int device, ypt_integrator** out);
/* PathIntegratorSubStage::shutdown + delete */
int ypt_destroy(ypt_integrator* h);
/* PathIntegratorSubStage::sceneChanged (PathIntegratorSubStage.h:62-71) + the acceleration-structure
 * build of PathTraceStage::beforeExecute (PathTraceStage.cpp:215-251). */
int ypt_set_scene(ypt_integrator* h, const ypt_scene_desc* scene);
/* RaytraceCommonResources::update (RaytraceCommonResources.cpp:64-140) + onRenderResolutionChanged.
 * A change of resolution clears the film. */
int ypt_set_frame(ypt_integrator* h, const ypt_frame_desc* frame);
/* clearAccumulatedSamples (clearAccumulatedSamples.comp:1-41; PathTraceStage.cpp:158-161) */
int ypt_clear_film(ypt_integrator* h);
/* PathIntegratorSubStage::update + executeRaytrace + mergeNewSamples, for Sobol sample indices
 * [sample_begin, sample_begin + sample_count): one pass per index, each merged into the film.
 * stream: a hipStream_t or NULL for the library's stream. The kernels of this call run on that stream,
 * after the work already queued on the handle, and every later call on the handle sees their result;
 * the handle keeps no reference to the caller's stream after the call. Asynchronous; see
 * ypt_synchronize. */
int ypt_render(ypt_integrator* h, uint32_t sample_begin, uint32_t sample_count, void* stream);
int ypt_synchronize(ypt_integrator* h);
/* getOutput(COLOR): merged film, W*H float4 (X, Y, Z, 1) into host memory; counts: W*H uint32 or NULL */
int ypt_get_film(ypt_integrator* h, float* xyzw, uint32_t* counts);
/* last pass before merging: W*H float4 (X, Y, Z, hero wavelength) */
int ypt_get_last_samples(ypt_integrator* h, float* xyzl);
/* getOutput(MATERIAL_PARAMS0/1) of the last pass: first-hit features, W*H x 12 floats
 * (normal xyz, albedo xyz, roughness, ior, metalness, transparency, depth, hit) */
int ypt_get_first_hit_features(ypt_integrator* h, float* out);
/* statistics of the last ypt_render, 16 doubles:
 *  [0] extension rays, [1] shadow rays, [6]/[7] BVH nodes visited / triangles tested by extension
 *      rays, [8]/[9] the same for shadow rays   (when ypt_set_option("count_traversal", 1));
 *  [2] extension, [3] shadow, [4] shading, [11] path start, [12] film merge kernel time in ms and
 *      [5] the whole render in ms                (when ypt_set_option("profile", 1));
 *  [10] extension launches.
 * Bidirectional integrator: [0]/[1] rays traced by camera subpaths / by light subpaths and light
 * tracing, [6]/[7] and [8]/[9] their nodes and triangles; [2] camera-subpath, [3] light-subpath,
 * [4] light-tracing, [11] light-subpath compaction, [12] splat sort + sum + film merge times in ms. */
int ypt_get_stats(ypt_integrator* h, double* out16);
/* Launch and diagnostics knobs (no counterpart in the reference, whose traversal is the driver's):
 *  "bvh_width" 2 | 8, "block_size" 64..1024 (multiple of 64), "lds_stack_depth" 1..64,
 *  "passes_in_flight" 0..128 (0 = automatic), "persistent" 0/1 (traversal kernels loop over their
 *  queue), "persistent_waves" 1..64 (waves per CU of those launches), "refill_threshold" 1..64
 *  (idle lanes before a wave fetches new rays), "steps_per_fetch" 1..64 (node steps between those
 *  checks, 8 by default), "wave_ray_below" (queues shorter than this are traced with one wavefront
 *  per ray; 0 = never), "leaf_share" 1..64 (BVH2 persistent traversal: a
 *  wave intersects its waiting leaves once 1/leaf_share of its active lanes wait), "shade_waves" 0..1024 (0 = one workgroup per queue
 *  entry), "shade_reg_waves" 2 | 4 | 8 (register budget of the shading kernel, in waves per SIMD),
 *  "graph" 0/1 (replay each batch of passes as a captured hipGraph), "count_traversal" 0/1,
 *  "first_hit_features" 0/1, "profile" 0/1.
 *  Bidirectional integrator: "bdpt_wavefront" 0/1 (camera subpaths as queued kernels rather than
 *  one kernel per texel), "bdpt_waves" 1 | 2 | 3 | 4 | 6 | 8 (register budget of the camera kernels;
 *  the queued shading kernel takes 1, 2 or 4), "bdpt_split_connections" 0 | 1 | 2 | 4 (queued camera
 *  subpaths: the connections of each shaded vertex in a launch of their own, one thread per light
 *  vertex, with that launch's register budget in waves per SIMD; 0 = in the shading launch's loop;
 *  1 by default; scenes with subsurface materials always use the loop).
 *  None of them changes a result bit: the visible connections of a camera vertex are added in the
 *  order they were made whichever "bdpt_wavefront" form runs. (Scenes with subsurface materials add
 *  the two shadow rays of a subsurface event atomically, which can change the last bits of a sum
 *  from run to run.) */
int ypt_set_option(ypt_integrator* h, const char* name, int64_t value);

/* Closest-hit queries against the scene's BVH on the GPU (the TraceRay / RayQuery the integrators
 * use). rays: n x 8 floats (origin xyz, tmin, direction xyz, tmax). Outputs per ray: t, triangle
 * index (0xFFFFFFFF on a miss) and barycentrics b1, b2 (may be NULL). */
int ypt_trace_closest(ypt_integrator* h, const float* rays, uint32_t n, float* t, uint32_t* tri, float* b12);

/* ypt_trace_closest on device buffers, allocating nothing. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). rays: n x 8 floats in device memory; t, tri, b12 (none may be
 * NULL): device outputs as in ypt_trace_closest. Asynchronous on `stream` (NULL: the handle's stream).
 * The traversal-stack overflow lives in a 16 MiB device array of the library, one per GPU, held until
 * the process ends; rays run in launches of as many rays as it holds (65536 for a stack bound of 64).
 * A BVH whose stack would not fit 64 rays is refused (YPT_E_MEMORY). The array needs no clearing, and
 * every call on one GPU shares it: calls that overlap on one GPU (on different streams, or from
 * different threads) race. */
int ypt_trace_closest_noalloc(ypt_integrator* h, const float* rays, uint32_t n, float* t, uint32_t* tri, float* b12,
                              void* stream);

/* Scrambled sampler values on the GPU: for each (seed, dimension) pair the value of
 * randomSampler.hlsl getRandomSampleFloat for Sobol sample `sample_index`. */
int ypt_sample_dims(ypt_integrator* h, uint32_t sample_index, const uint32_t* seeds, const uint32_t* dims,
                    uint32_t n, float* out);

/* Sobol point of a sample index: 256 uint32 dimensions (RaytraceCommonResources.cpp:623-627). */
int ypt_sobol_point(uint32_t sample_index, uint32_t* out256);

/* The steps after the film (CombineSamplesSubStage + TonemapStage): the edge-avoiding a-trous
 * denoiser (5 passes of denoise.comp over the film, its variance and the prefiltered feature
 * moments of prefilterFeatures.comp; needs "first_hit_features" on while rendering), XYZ -> linear
 * sRGB with the gamut clamp, exposure, the ACES fitted operator (tonemap.frag with TONEMAP_ACES),
 * and the encoding of an RGBA8_SRGB target: the FinalColor readback. Defaults are the renderer
 * variables (RendererVarsList.h:24-54). Auto exposure (the luminance histogram) is not built. */
typedef struct ypt_post_params {
  uint32_t denoise;              /* Denoise.Mode: 0 OFF, 1 ON */
  uint32_t prefilter_features;   /* Denoise.FeaturePrefilter (1) */
  float depth_sigma_scale;       /* 0.02 */
  float normal_sigma;            /* 0.06 */
  float material_sigma;          /* 0.05 */
  float color_sigma;             /* 3 */
  float feature_variance_scale;  /* 1 */
  uint32_t tonemap;              /* Tonemap.Enable: the ACES operator (1); exposure applies regardless */
  float exposure;                /* Tonemap.ManualExposure (1) */
} ypt_post_params;
void ypt_post_params_default(ypt_post_params* p);
/* rgba8: W*H*4 bytes, display encoded; linear_rgba: W*H*4 floats, linear sRGB before exposure (the
 * denoised film); either may be NULL */
int ypt_final_color(ypt_integrator* h, const ypt_post_params* p, uint8_t* rgba8, float* linear_rgba);
/* Feature moments accumulated with the film (mergeNewSamples.comp accumulateFeatures): 4 layers of
 * W*H float4 (albedo, |albedo|^2 / normal, coverage / roughness, metalness, transparency, anisotropy /
 * ior/2, |material|^2, depth, depth^2) */
int ypt_get_feature_moments(ypt_integrator* h, float* out16);
/* Per-pixel variance the film merge estimates from its two accumulation targets (log luminance),
 * the denoiser's colour-weight input: W*H floats */
int ypt_get_film_variance(ypt_integrator* h, float* out);

const char* ypt_last_error(void);
const char* ypt_version(void);

/* ---- Resident scenes: built once, shared by any number of integrator handles and render jobs.
 * The reference keeps one scene per renderer: the bindless mesh/material buffers and the
 * acceleration structures of AccelerationStructureProvider (PathTraceStage.cpp:215-251,
 * PathIntegratorSubStage::sceneChanged, PathIntegratorSubStage.h:62-71). */
typedef struct ypt_scene ypt_scene;
int ypt_scene_create(const ypt_scene_desc* scene, int device, ypt_scene** out);
/* the device memory is released when no handle or job uses the scene any more */
int ypt_scene_destroy(ypt_scene* scene);
/* Moved geometry without a BVH rebuild (the reference rebuilds or updates its TLAS when render-object
 * transforms change, AccelerationStructureHelper.cpp): `scene` has the same triangles, instances and
 * alpha-tested triangles as the handle's scene, with new vertex data (positions, normals, tangents,
 * UVs) and possibly new materials, lights or environment. The scene is re-prepared, its BVH refitted
 * (same tree, new boxes) and its buffers rewritten in place; the frame constants are re-derived and
 * the film cleared. The handle must be the scene's only user (YPT_E_INVALID otherwise). */
int ypt_refit_scene(ypt_integrator* h, const ypt_scene_desc* scene);
/* BVH builds (1) and refits of the handle's scene */
int ypt_scene_bvh_counts(ypt_integrator* h, uint32_t* builds, uint32_t* refits);
int ypt_scene_info(const ypt_scene* scene, uint64_t* device_bytes, uint32_t* triangles);
/* like ypt_set_scene, without a copy or a BVH build */
int ypt_set_resident_scene(ypt_integrator* h, ypt_scene* scene);

/* ---- Render server: continuous batching of progressive render jobs (no counterpart in the
 * reference, whose renderer renders one view; the jobs run the same integrators).
 * Jobs are admitted in arrival order while at most max_active_jobs run and their estimated device
 * memory stays within hbm_budget_bytes. BPT and NEE jobs (scenes without subsurface materials) run
 * their chunks of passes in fused steps: one launch sequence whose queues carry the paths of up to 16
 * jobs and their scenes together. BDPT jobs keep their next chunk in flight on their own stream,
 * beside the fused steps. A job's film depends only on its scene, frame and sample count (Sobol
 * indices 0 .. spp-1), alone or in any mix, fused or not.
 * Server settings through ypt_server_set_option (integrator type ignored): "fuse_jobs" (1 default,
 * 0: every job on its own stream) and "fuse_max_slots" (path slots of one step, 16M default; a larger
 * single chunk runs alone). */
typedef struct ypt_server ypt_server;
int ypt_server_create(int device, uint64_t hbm_budget_bytes, uint32_t max_active_jobs, ypt_server** out);
int ypt_server_destroy(ypt_server* server);
int ypt_server_submit(ypt_server* server, ypt_scene* scene, int integrator_type, const ypt_frame_desc* frame,
                      uint32_t spp, uint64_t* job_id);
/* ypt_set_option for the handles that run jobs of one integrator type (< 0: every type), from the
 * next admitted job on */
int ypt_server_set_option(ypt_server* server, int integrator_type, const char* name, int64_t value);
/* YPT_OK when the job has finished, 1 when timeout_s (< 0: no limit) passed first; YPT_E_CANCELLED or
 * YPT_E_TIMEOUT for a job that ended without its film */
int ypt_server_wait(ypt_server* server, uint64_t job_id, double timeout_s);
/* Cancels a queued or running job; a running one stops before its next chunk of passes, and the jobs
 * around it are untouched. YPT_E_INVALID for a job that has already ended. */
int ypt_server_cancel(ypt_server* server, uint64_t job_id);
/* Gives the job a deadline `seconds` after its submission: past it, the job is dropped from the queue or
 * stopped before its next chunk of passes, and waits on it return YPT_E_TIMEOUT. */
int ypt_server_set_deadline(ypt_server* server, uint64_t job_id, double seconds);
/* film of a finished job (W*H float4 X, Y, Z, 1; may be NULL) and times3 = submit, admit, finish in
 * seconds of ypt_clock (may be NULL); the job is forgotten afterwards (also one that was cancelled or
 * timed out, for which the call returns YPT_E_CANCELLED or YPT_E_TIMEOUT) */
int ypt_server_take(ypt_server* server, uint64_t job_id, float* xyzw, double* times3);
/* [0] queued, [1] active, [2] completed, [3] device bytes of active jobs, [4] peak active, [5] budget */
int ypt_server_stats(ypt_server* server, double* out6);
/* fused steps issued so far and the job chunks they carried (their ratio is the mean number of jobs
 * per step) */
int ypt_server_fused_stats(ypt_server* server, uint64_t* steps, uint64_t* job_chunks);
/* the steady clock of the job times, in seconds */
double ypt_clock(void);

#ifdef __cplusplus
}
#endif
#endif
