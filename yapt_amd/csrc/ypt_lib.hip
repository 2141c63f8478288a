// libyapt_pt: C ABI over the gfx950 wavefront path tracer (see include/yapt_pt.h).
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <map>
#include <tuple>
#include <thread>
#include <unordered_map>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/yapt_pt.h"
#include "core/shade.h"
#include "kernels/bdpt_kernels.h"
#include "kernels/instances.h"
#include "kernels/post.h"
#include "kernels/wavefront.h"
#include "runtime/bvh.h"
#include "runtime/luts.h"
#include "runtime/prepare.h"
#include "runtime/rgb2spec.h"

#include "data/cie_tables.inc"
#include "data/sobol_matrices.inc"

using namespace ypt;

// the template kernels are compiled in instances.hip
namespace ypt {
YPT_INST_ALL(extern template)
}

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

struct HipError : std::runtime_error {
  int code;
  HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      throw HipError(e_ == hipErrorOutOfMemory ? YPT_E_MEMORY : YPT_E_DEVICE,                    \
                     std::string(#expr) + ": " + hipGetErrorString(e_));                          \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  // grow-only: keeps the buffer when it is large enough, so no hipFree happens on a size change
  void ensure(size_t count) {
    if (n >= count) return;
    alloc(count);
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    if (count > n) alloc(count);
    if (count) HIP_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void zero(hipStream_t s) {
    if (n) HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// Process-wide host tables: directional-albedo LUTs and the RGB->spectrum fit (both deterministic).
struct HostTables {
  std::vector<float> msl;
  Rgb2Spec r2s;
  HostTables() {
    msl.resize(MSL_TOTAL);
    generate_ms_luts(msl.data());
    r2s.init(YPT_CIE_X_D, YPT_CIE_Y_D, YPT_CIE_Z_D, YPT_CIE_D65_D, YPT_SRGB_XYZ_TO_RGB_D, YPT_SRGB_RGB_TO_XYZ_D);
  }
};
HostTables& host_tables() {
  static HostTables t;
  return t;
}

uint32_t blocks_for(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// Device scratch of ypt_trace_closest_noalloc: the traversal-stack overflow of its launches, in a
// file-scope array rather than a per-call allocation. Each GPU holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress returns on that device
// (never the array's name, which is a host-side stand-in). Every call on one GPU shares it.
constexpr size_t NOALLOC_SCRATCH_WORDS = (size_t)4 << 20;  // 16 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratchMutex;
void* g_scratchAddr[NOALLOC_MAX_DEVICES] = {};
std::string g_scratchError;

}  // namespace

static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];

namespace {

// the scratch array's address on `device` (current), looked up once per device at handle creation
void init_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratchMutex);
  if (device < 0 || device >= NOALLOC_MAX_DEVICES || g_scratchAddr[device]) return;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratchError = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_scratchAddr[device] = p;
}

}  // namespace

// A scene resident in device memory: host copy, BVH and device buffers. Integrator handles and the
// render server share it (the reference's AccelerationStructureProvider + bindless scene buffers).
struct ResidentScene {
  int device = 0;
  PreparedScene scene;
  BvhResult bvh;
  DevBuf<float> dPos, dNrm, dTan;
  DevBuf<uint32_t> dIdx, dTriInst;
  DevBuf<InstanceGPU> dInst;
  DevBuf<MaterialGPU> dMats;
  DevBuf<LightGPU> dLights;
  DevBuf<Bvh2Node> dNodes2;
  DevBuf<Bvh8Node> dNodes8;
  DevBuf<TriRecord> dTris;
  DevBuf<float> dUv, dSrgbLut, dEnv;
  DevBuf<TextureGPU> dTex;
  DevBuf<uint32_t> dTexels;
  DevBuf<AlphaTri> dAlpha;
  DevBuf<uint32_t> dTriAlpha;
  uint32_t stackBound2 = BVH_MAX_DEPTH, stackBound8 = 8 * BVH_MAX_DEPTH;
  bool subsurface = false;  // some material has subsurfaceAmount > 0
  uint32_t refits = 0;      // ypt_refit_scene calls since the build
  size_t device_bytes() const {
    return dPos.n * 4 + dNrm.n * 4 + dTan.n * 4 + dIdx.n * 4 + dTriInst.n * 4 + dInst.n * sizeof(InstanceGPU) +
           dMats.n * sizeof(MaterialGPU) + dLights.n * sizeof(LightGPU) + dNodes2.n * sizeof(Bvh2Node) +
           dNodes8.n * sizeof(Bvh8Node) + dTris.n * sizeof(TriRecord) + dUv.n * 4 + dTex.n * sizeof(TextureGPU) + dTexels.n * 4 +
           dSrgbLut.n * 4 + dAlpha.n * sizeof(AlphaTri) + dTriAlpha.n * 4 + dEnv.n * 4;
  }
};

struct ypt_integrator {
  int type = YPT_INTEGRATOR_NEE;
  int device = 0;
  // the stream the handle's work runs on: the library's own stream (ownStream), or the caller's
  // stream for the duration of a ypt_render that names one
  hipStream_t stream = nullptr;
  hipStream_t ownStream = nullptr;
  hipEvent_t streamJoin = nullptr;  // orders the caller's stream and the library's stream

  // scene (resident, possibly shared with other handles)
  std::shared_ptr<ResidentScene> res;
  bool hasScene = false;
  DevBuf<float> dMsl, dCie, dD65, dRgb2spec;

  // frame
  ypt_frame_desc frameDesc;
  FrameConstants fc;
  bool hasFrame = false;
  uint32_t width = 0, height = 0;

  // wavefront state
  DevBuf<float4> pOrg, pDir, pThr, pLight, pVolIor, pVolAbs, hit, sOrg, sDir, sContrib;
  DevBuf<uint4> pSt0, pSt1, pVolMat;
  DevBuf<uint32_t> rayQ0, rayQ1, counters, sSlot, ovf, sobolPoints;
  DevBuf<uint32_t> sssStack;  // subsurface probe stacks of the shading launch
  DevBuf<unsigned long long> work;
  uint32_t ovfStride = 0;

  // bidirectional integrator. The light subpaths of pass k+1 are traced on lightStream into one of
  // two sets while the camera stage of pass k reads the other.
  struct LightSet {
    DevBuf<LightNodePacked> nodes;
    DevBuf<LightNodeSurface> surf;  // surface of each stored light vertex, made once per pass
    DevBuf<LightVertexCtx> pre;     // connection context of each stored light vertex, made with it
    DevBuf<uint32_t> counts, headers, headerCount, ovf;
    DevBuf<float> lambdas;
  } lset[2];
  hipStream_t lightStream = nullptr;
  hipEvent_t lightDone[2] = {}, setFree[2] = {}, camReady = nullptr;
  DevBuf<uint32_t> bPerm, bKeys, bVals, bKeysOut, bValsOut, bOvf;
  DevBuf<float4> bCamXYZ, bRecXYZ;
  DevBuf<uint8_t> bSortTemp;
  uint32_t bW = 0, bH = 0, bOvfStride = 0;
  // camera subpaths as a wavefront: per-slot state and the connection-ray queue
  DevBuf<float4> cOrg, cDir, cThr, cTotal, cPrev, cPrevN, cVolIor, cVolAbs, cSOrg, cSDir, cSContrib;
  DevBuf<uint4> cSt0, cSt1, cVolMat;
  DevBuf<float> cDeferred;
  DevBuf<uint32_t> cSSlot, cPending;
  DevBuf<float4> cConn;  // connection records, connCap per texel
  DevBuf<DeferredConnections> cDefer;  // the shaded vertex per texel whose connections the connection launch makes
  DevBuf<uint32_t> cConnList;          // the connection launch's work list
  int bdptWavefront = 1;
  size_t bSortBytes = 0;

  // captured batches (hipGraph): one executable graph per batch size, recaptured when anything that
  // is baked into its kernel arguments changes (fingerprint `key`)
  struct BatchGraph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t key = 0;
  };
  BatchGraph graphs[2];
  DevBuf<uint32_t> batchPoints, mergeDyn;
  int useGraph = 0;

  // film
  DevBuf<float4> accum, merged0, merged1, lastSample;
  DevBuf<uint32_t> counts0, counts1;
  DevBuf<float> variance, features;
  DevBuf<float4> featAccum;  // feature moments accumulated like the film (4 layers per pixel)
  // after the film: packed prefiltered features, a-trous ping-pong colour and variance, 8-bit output
  DevBuf<uint4> postF0, postF1;
  DevBuf<float4> postA, postB;
  DevBuf<float> postVarA, postVarB;
  DevBuf<uchar4> postOut;
  uint32_t filmSamples = 0;  // passes merged since the last clear
  uint32_t filmCur = 0;
  size_t slots = 0;          // path slots allocated
  uint32_t lastPassBase = 0; // slot offset of the last rendered pass (first-hit features)

  // options
  int bvhWidth = 2;
  int blockSize = 128;
  int ldsDepth = 24;
  int countTraversal = 0;
  int writeFeatures = 0;
  int passesInFlight = 0;  // 0 = automatic
  int persistent = 1;      // traversal/shading kernels loop over their queue (see work_grid)
  int persistentWaves = 32;
  int refill = 16;         // idle lanes of a wave before it fetches new rays
  int leafShare = 2;       // BVH2 persistent traversal: leaf phase once 1/leafShare of the active lanes wait
  int shadeWaves = 64;     // waves per CU of the shading launch (0: one workgroup per queue entry)
  int shadeRegWaves = 4;   // register budget of the shading kernel: 2, 4 or 8 waves per SIMD
  int bdptWaves = 4;       // register budget of the camera-subpath kernel: waves per SIMD (1, 2, 3, 4, 6, 8)
  int bdptSplit = 1;       // camera wavefront: connections in their own launch (its waves per SIMD: 1, 2, 4; 0 = in the shading launch)
  int waveBelow = 0;       // unidirectional bounces: ray and shadow queues shorter than this are traced wave-per-ray
  int stepsPerFetch = 8;   // persistent traversal: node steps of a wave between its refill checks
  int cus = 256;           // compute units of the device

  // stats of the last render
  double stats[16] = {0};
  hipEvent_t ev[8] = {};
  bool profile = false;
  std::map<std::tuple<const void*, uint32_t, size_t>, int> occupancy;  // workgroups per CU of a kernel launch shape
  std::vector<hipEvent_t> evPool;  // profile mode: one event per kernel boundary, read after the render
  size_t evUsed = 0;
  hipEvent_t next_event() {
    if (evUsed == evPool.size()) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreate(&e));
      evPool.push_back(e);
    }
    return evPool[evUsed++];
  }

  DevScene dev_scene() const {
    DevScene s;
    std::memset(&s, 0, sizeof(s));
    const ResidentScene& r = *res;
    s.sv.pos = r.dPos.p; s.sv.nrm = r.dNrm.p; s.sv.idx = r.dIdx.p; s.sv.triInst = r.dTriInst.p;
    s.sv.tan = r.scene.tan.empty() ? nullptr : r.dTan.p;
    s.sv.inst = r.dInst.p; s.sv.mats = r.dMats.p; s.sv.lights = r.dLights.p;
    s.sv.rgb2spec = dRgb2spec.n ? dRgb2spec.p : nullptr;
    s.sv.uv = r.dUv.n ? r.dUv.p : nullptr;
    s.sv.textures = r.dTex.n ? r.dTex.p : nullptr;
    s.sv.texels = r.dTexels.n ? r.dTexels.p : nullptr;
    s.sv.srgbLut = r.dSrgbLut.n ? r.dSrgbLut.p : nullptr;
    s.sv.triAlpha = r.dAlpha.n ? r.dTriAlpha.p : nullptr;
    s.sv.alpha = AlphaView{r.dAlpha.n ? r.dAlpha.p : nullptr, r.dTex.p, r.dTexels.p};
    s.sv.env = EnvView{r.scene.envType != ENV_NONE ? r.dEnv.p : nullptr, r.scene.envType, r.scene.envWidth, r.scene.envHeight};
    s.sv.ntris = (uint32_t)r.scene.triInst.size();
    s.sv.ninst = (uint32_t)r.scene.inst.size();
    s.sv.spec.cie_xyz4 = dCie.p;
    s.sv.spec.d65 = dD65.p;
    s.sv.luts.base = dMsl.p;
    s.nodes2 = r.dNodes2.p; s.nodes8 = r.dNodes8.p;
    s.tris.recs = r.dTris.p;
    s.tris.alpha = s.sv.alpha;
    return s;
  }
  PathSoA path_soa() const {
    PathSoA P;
    std::memset(&P, 0, sizeof(P));
    P.orgLambda = pOrg.p; P.dirPdf = pDir.p; P.thr = pThr.p; P.light = pLight.p;
    P.st0 = pSt0.p; P.st1 = pSt1.p; P.volIor = pVolIor.p; P.volMat = pVolMat.p; P.volAbs = pVolAbs.p;
    return P;
  }
  Queues queues() const {
    Queues Q;
    std::memset(&Q, 0, sizeof(Q));
    Q.ray[0] = rayQ0.p; Q.ray[1] = rayQ1.p; Q.counters = counters.p; Q.hit = hit.p;
    Q.sOrgTmax = sOrg.p; Q.sDirTri = sDir.p; Q.sContrib = sContrib.p; Q.sSlot = sSlot.p;
    Q.ovf = ovf.p; Q.ovfStride = ovfStride; Q.workCounters = work.p;
    Q.stepsPerFetch = (uint32_t)stepsPerFetch;
    Q.sssStack = sssStack.p;
    return Q;
  }
  FilmBuffers film() const {
    FilmBuffers F;
    std::memset(&F, 0, sizeof(F));
    F.accum = accum.p; F.merged[0] = merged0.p; F.merged[1] = merged1.p;
    F.counts[0] = counts0.p; F.counts[1] = counts1.p; F.variance = variance.p; F.lastSample = lastSample.p;
    F.features = features.p;
    F.featAccum = writeFeatures ? featAccum.p : nullptr;  // moments only where the features are written
    return F;
  }
};

namespace {

void upload_constant_tables(ypt_integrator* h) {
  HostTables& T = host_tables();
  h->dMsl.upload(T.msl.data(), T.msl.size(), h->stream);
  h->dCie.upload(YPT_CIE_XYZ, 4 * YPT_CIE_RES, h->stream);
  h->dD65.upload(YPT_CIE_D65, YPT_CIE_RES, h->stream);
}

void ensure_frame_buffers(ypt_integrator* h, uint32_t w, uint32_t hgt) {
  size_t npix = (size_t)w * hgt;
  h->counters.ensure(16);
  h->work.ensure(16);
  h->accum.ensure(2 * npix); h->merged0.ensure(npix); h->merged1.ensure(npix); h->lastSample.ensure(npix);
  h->counts0.ensure(npix); h->counts1.ensure(npix); h->variance.ensure(npix);
  h->width = w;
  h->height = hgt;
}

// path-state and queue buffers for `slots` paths in flight
void ensure_slots(ypt_integrator* h, size_t slots) {
  if (h->slots >= slots) return;
  h->pOrg.ensure(slots); h->pDir.ensure(slots); h->pThr.ensure(slots); h->pLight.ensure(slots);
  h->pSt0.ensure(slots); h->pSt1.ensure(slots); h->pVolIor.ensure(slots); h->pVolMat.ensure(slots); h->pVolAbs.ensure(slots * 4);
  h->hit.ensure(slots);
  h->rayQ0.ensure(slots); h->rayQ1.ensure(slots);
  h->sOrg.ensure(slots); h->sDir.ensure(slots); h->sContrib.ensure(slots); h->sSlot.ensure(slots);
  h->features.ensure(slots * FEATURE_STRIDE);
  h->slots = slots;
}

// traversal-stack overflow: one column per thread that can traverse. A thread's id is below the
// queue length (one workgroup per entry, rounded up to a workgroup of at most 1024) or below the
// persistent launch size.
uint32_t sss_shade_grid(const ypt_integrator* h, uint32_t sgrid);

void ensure_overflow(ypt_integrator* h) {
  if (h->res && h->res->subsurface) {
    // two shadow rays per path (vertex and subsurface exit), and the probe stacks of the shading launch
    h->sOrg.ensure(2 * h->slots); h->sDir.ensure(2 * h->slots); h->sContrib.ensure(2 * h->slots); h->sSlot.ensure(2 * h->slots);
    h->sssStack.ensure((size_t)h->res->stackBound2 * sss_shade_grid(h, 0xFFFFFFFFu) * 64);
  }
  size_t stride = h->slots + 1024;
  if (h->persistent) stride = std::min(stride, (size_t)h->cus * h->persistentWaves * 64);
  uint32_t bound = h->bvhWidth == 8 ? h->res->stackBound8 : h->res->stackBound2;
  size_t depth = bound > (uint32_t)h->ldsDepth ? bound - (uint32_t)h->ldsDepth : 1;
  h->ovf.ensure(depth * stride);
  h->ovfStride = (uint32_t)stride;
}

// passes in flight: enough slots (~64M paths, ~22 GB of path state at 1080p) that the deep, nearly
// empty bounce launches of one batch are shared by many passes. The passes of one render are split
// into batches of equal size (40 passes at a limit of 32 run as 20 + 20, not 32 + 8), so no batch is
// left with a short, latency-bound tail.
constexpr uint32_t MAX_PASSES_IN_FLIGHT = 128;
uint32_t passes_in_flight(const ypt_integrator* h, uint32_t count) {
  size_t npix = (size_t)h->width * h->height;
  uint32_t g = h->passesInFlight > 0 ? (uint32_t)h->passesInFlight : (uint32_t)std::max<size_t>(1, (64u << 20) / npix);
  g = std::max(1u, std::min<uint32_t>(g, MAX_PASSES_IN_FLIGHT));
  if (count <= g) return std::max(1u, count);
  uint32_t batches = (count + g - 1) / g;
  return (count + batches - 1) / batches;
}

void clear_film(ypt_integrator* h) {
  h->accum.zero(h->stream); h->merged0.zero(h->stream); h->merged1.zero(h->stream);
  h->counts0.zero(h->stream); h->counts1.zero(h->stream); h->variance.zero(h->stream);
  h->features.zero(h->stream); h->lastSample.zero(h->stream);
  h->filmSamples = 0;
  h->filmCur = 0;
}

// longest root-to-leaf push count of the BVH8 (bounds the traversal stack)
uint32_t bvh8_stack_bound(const std::vector<Bvh8Node>& nodes) {
  if (nodes.empty()) return 1;
  std::vector<uint32_t> depth(nodes.size(), 0);
  uint32_t best = 0;
  for (size_t i = 0; i < nodes.size(); ++i) {  // children have larger indices (breadth-first emission)
    uint32_t inner = 0;
    for (int s = 0; s < 8; ++s)
      if (nodes[i].child[s] != BVH_EMPTY && !(nodes[i].child[s] & BVH_LEAF_BIT)) inner++;
    uint32_t here = depth[i] + (inner > 0 ? inner - 1 : 0);
    best = std::max(best, here);
    for (int s = 0; s < 8; ++s) {
      uint32_t c = nodes[i].child[s];
      if (c != BVH_EMPTY && !(c & BVH_LEAF_BIT)) depth[c] = std::max(depth[c], here);
    }
  }
  return best + 1;
}

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const HipError& e) {
    return fail(e.code, e.what());
  } catch (const std::invalid_argument& e) {
    return fail(YPT_E_INVALID, e.what());
  } catch (const std::bad_alloc&) {
    return fail(YPT_E_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(YPT_E_INVALID, e.what());
  }
}

// launch shape of the traversal and shading kernels: one workgroup per queue entry, or (persistent)
// at most `persistentWaves` waves per CU that loop over the queue
uint32_t work_grid(const ypt_integrator* h, uint32_t nslots) {
  uint32_t B = (uint32_t)h->blockSize;
  uint32_t g = blocks_for(nslots, B);
  if (!h->persistent) return g;
  uint32_t maxb = std::max<uint32_t>(1u, (uint32_t)((size_t)h->cus * h->persistentWaves * 64 / B));
  return std::min(g, maxb);
}

// shading (64-thread workgroups): grid-stride over the queue with `shadeWaves` waves per CU, or one
// workgroup per entry
uint32_t shade_grid(const ypt_integrator* h, uint32_t nslots) {
  uint32_t g = blocks_for(nslots, 64);
  if (h->shadeWaves == 0) return g;
  return std::min(g, std::max<uint32_t>(1u, (uint32_t)h->cus * (uint32_t)h->shadeWaves));
}

// shading with subsurface probes: at most 32 workgroups of 64 per CU, which bounds the probe stacks
uint32_t sss_shade_grid(const ypt_integrator* h, uint32_t sgrid) {
  return std::min<uint32_t>(sgrid, std::max<uint32_t>(1u, (uint32_t)h->cus * 32u));
}

// explicit light connections (shadow rays) exist with NEE, and for BPT when the scene has punctual lights
bool traces_shadow_rays(const ypt_integrator* h) {
  return h->type == YPT_INTEGRATOR_NEE || h->fc.punctualLightCount > 0;
}

void launch_shade(ypt_integrator* h, const DevScene& S, const uint32_t* point, uint32_t npix, const PathSoA& P, const Queues& Q,
                  const FilmBuffers& F, int q, uint32_t sgrid) {
  hipStream_t s = h->stream;
  bool nee = h->type == YPT_INTEGRATOR_NEE;
  int w = h->shadeRegWaves;
#define YPT_SHADE(N, M) hipLaunchKernelGGL((k_shade<N, M>), dim3(sgrid), dim3(64), 0, s, h->fc, S, point, npix, P, Q, F, q, h->writeFeatures)
  if (h->res->subsurface) {
    uint32_t g = sss_shade_grid(h, sgrid);
    if (nee) hipLaunchKernelGGL((k_shade<true, 2, true>), dim3(g), dim3(64), 0, s, h->fc, S, point, npix, P, Q, F, q, h->writeFeatures);
    else hipLaunchKernelGGL((k_shade<false, 2, true>), dim3(g), dim3(64), 0, s, h->fc, S, point, npix, P, Q, F, q, h->writeFeatures);
  } else if (nee) {
    if (w == 2) YPT_SHADE(true, 2); else if (w == 8) YPT_SHADE(true, 8); else YPT_SHADE(true, 4);
  } else {
    if (w == 2) YPT_SHADE(false, 2); else if (w == 8) YPT_SHADE(false, 8); else YPT_SHADE(false, 4);
  }
#undef YPT_SHADE
}

// A persistent launch gets no more workgroups than fit on the device at once (the kernel's
// occupancy at this block size and LDS): a workgroup beyond that would start only after others had
// drained the queue, and its first rays would then run alone at the end of the launch.
template <class K>
uint32_t resident_grid(ypt_integrator* h, K kernel, uint32_t block, size_t lds, uint32_t grid) {
  const void* f = reinterpret_cast<const void*>(kernel);
  auto key = std::make_tuple(f, block, lds);
  auto it = h->occupancy.find(key);
  int nb = 0;
  if (it != h->occupancy.end()) {
    nb = it->second;
  } else {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, (int)block, lds) != hipSuccess || nb < 1) nb = 0;
    h->occupancy[key] = nb;
  }
  if (nb == 0) return grid;
  return std::max<uint32_t>(1u, std::min<uint32_t>(grid, (uint32_t)nb * (uint32_t)h->cus));
}

// persistent traversal launch, its grid clamped to what is resident
#define YPT_PERSISTENT(KERNEL, GRID, BLOCK, LDS, STREAM, ...)                                              \
  hipLaunchKernelGGL(KERNEL, dim3(resident_grid(h, KERNEL, (BLOCK), (LDS), (GRID))), dim3(BLOCK), (LDS), \
                     (STREAM), __VA_ARGS__)

// wave-per-ray launches (k_extend_w, k_shadow_w): one 64-lane workgroup per ray at a time, as many
// as the overflow stacks have columns for, at most 32 per CU; LDS holds the lanes' stacks and the
// wave's node list
uint32_t wave_grid(const ypt_integrator* h) {
  return std::max<uint32_t>(1u, std::min<uint32_t>(h->ovfStride / 64u, (uint32_t)h->cus * 32u));
}
size_t wave_lds(const ypt_integrator* h) { return ((size_t)h->ldsDepth * 64 + 64) * 4; }

template <int WIDTH>
void launch_extend_wave(ypt_integrator* h, const DevScene& S, const PathSoA& P, const Queues& Q, int q) {
  if (Q.waveBelow == 0) return;
  uint32_t d = (uint32_t)h->ldsDepth, g = wave_grid(h);
  size_t lds = wave_lds(h);
  if (h->countTraversal) YPT_PERSISTENT((k_extend_w<WIDTH, true, true>), g, 64, lds, h->stream, S, P, Q, q, d);
  else if (S.tris.alpha.tris != nullptr) YPT_PERSISTENT((k_extend_w<WIDTH, false, true>), g, 64, lds, h->stream, S, P, Q, q, d);
  else YPT_PERSISTENT((k_extend_w<WIDTH, false, false>), g, 64, lds, h->stream, S, P, Q, q, d);
}
template <int WIDTH>
void launch_shadow_wave(ypt_integrator* h, const DevScene& S, const PathSoA& P, const Queues& Q) {
  if (Q.waveBelow == 0) return;
  uint32_t d = (uint32_t)h->ldsDepth, g = wave_grid(h);
  size_t lds = wave_lds(h);
  if (h->countTraversal) YPT_PERSISTENT((k_shadow_w<WIDTH, true, true>), g, 64, lds, h->stream, S, P, Q, d);
  else if (S.tris.alpha.tris != nullptr) YPT_PERSISTENT((k_shadow_w<WIDTH, false, true>), g, 64, lds, h->stream, S, P, Q, d);
  else YPT_PERSISTENT((k_shadow_w<WIDTH, false, false>), g, 64, lds, h->stream, S, P, Q, d);
}

template <int WIDTH>
void launch_extend_thread(ypt_integrator* h, const DevScene& S, const PathSoA& P, const Queues& Q, int q, uint32_t grid) {
  size_t lds = (size_t)h->ldsDepth * h->blockSize * 4;
  uint32_t d = (uint32_t)h->ldsDepth, r = (uint32_t)h->refill;
  if (h->persistent) {
    bool alpha = S.tris.alpha.tris != nullptr;
    if (h->countTraversal)
      YPT_PERSISTENT((k_extend_p<WIDTH, true>), grid, h->blockSize, lds, h->stream, S, P, Q, q, d, r, (uint32_t)h->leafShare);
    else if (alpha)
      YPT_PERSISTENT((k_extend_p<WIDTH, false, true>), grid, h->blockSize, lds, h->stream, S, P, Q, q, d, r, (uint32_t)h->leafShare);
    else
      YPT_PERSISTENT((k_extend_p<WIDTH, false, false>), grid, h->blockSize, lds, h->stream, S, P, Q, q, d, r, (uint32_t)h->leafShare);
  } else if (h->countTraversal) {
    hipLaunchKernelGGL((k_extend<WIDTH, true>), dim3(grid), dim3(h->blockSize), lds, h->stream, S, P, Q, q, d);
  } else {
    hipLaunchKernelGGL((k_extend<WIDTH, false>), dim3(grid), dim3(h->blockSize), lds, h->stream, S, P, Q, q, d);
  }
}
template <int WIDTH>
void launch_shadow_thread(ypt_integrator* h, const DevScene& S, const PathSoA& P, const Queues& Q, uint32_t grid) {
  size_t lds = (size_t)h->ldsDepth * h->blockSize * 4;
  uint32_t d = (uint32_t)h->ldsDepth, r = (uint32_t)h->refill;
  if (h->res->subsurface) {  // two rays may add to one path: the atomic form, persistent at any grid
    if (h->countTraversal)
      YPT_PERSISTENT((k_shadow_p<WIDTH, true, true>), grid, h->blockSize, lds, h->stream, S, P, Q, d, r, (uint32_t)h->leafShare);
    else
      YPT_PERSISTENT((k_shadow_p<WIDTH, false, true>), grid, h->blockSize, lds, h->stream, S, P, Q, d, r, (uint32_t)h->leafShare);
  } else if (h->persistent) {
    bool alpha = S.tris.alpha.tris != nullptr;
    if (h->countTraversal)
      YPT_PERSISTENT((k_shadow_p<WIDTH, true>), grid, h->blockSize, lds, h->stream, S, P, Q, d, r, (uint32_t)h->leafShare);
    else if (alpha)
      YPT_PERSISTENT((k_shadow_p<WIDTH, false, false, true>), grid, h->blockSize, lds, h->stream, S, P, Q, d, r, (uint32_t)h->leafShare);
    else
      YPT_PERSISTENT((k_shadow_p<WIDTH, false, false, false>), grid, h->blockSize, lds, h->stream, S, P, Q, d, r, (uint32_t)h->leafShare);
  } else if (h->countTraversal) {
    hipLaunchKernelGGL((k_shadow<WIDTH, true>), dim3(grid), dim3(h->blockSize), lds, h->stream, S, P, Q, d);
  } else {
    hipLaunchKernelGGL((k_shadow<WIDTH, false>), dim3(grid), dim3(h->blockSize), lds, h->stream, S, P, Q, d);
  }
}


// a bounce's closest-hit (occlusion) launch: the thread-per-ray kernel, then, when short queues go
// wave-per-ray, the wave kernel; the one whose queue-length range does not hold the queue returns
template <int WIDTH>
void launch_extend(ypt_integrator* h, const DevScene& S, const PathSoA& P, const Queues& Q, int q, uint32_t grid) {
  launch_extend_thread<WIDTH>(h, S, P, Q, q, grid);
  launch_extend_wave<WIDTH>(h, S, P, Q, q);
}
template <int WIDTH>
void launch_shadow(ypt_integrator* h, const DevScene& S, const PathSoA& P, const Queues& Q, uint32_t grid) {
  launch_shadow_thread<WIDTH>(h, S, P, Q, grid);
  launch_shadow_wave<WIDTH>(h, S, P, Q);
}


// All launches of one batch of g passes for the unidirectional integrators, reading the Sobol
// points from `point` and (when dyn is set) the film bookkeeping from the device, which the last
// launch advances: the form that is captured into a graph.
void launch_batch(ypt_integrator* h, const DevScene& S, const PathSoA& P, const Queues& Q, const FilmBuffers& F,
                  const uint32_t* point, uint32_t g, uint32_t* dyn) {
  hipStream_t s = h->stream;
  uint32_t npix = h->width * h->height, nslots = npix * g;
  uint32_t grid = work_grid(h, nslots), sgrid = shade_grid(h, nslots), B = (uint32_t)h->blockSize;
  hipLaunchKernelGGL(k_start, dim3(blocks_for(nslots, 256)), dim3(256), 0, s, h->fc, point, P, Q, F, npix, nslots,
                     h->writeFeatures);
  int q = 0;
  for (uint32_t b = 0; b < h->fc.maxRayDepth; ++b) {
    if (h->bvhWidth == 8) launch_extend<8>(h, S, P, Q, q, grid);
    else launch_extend<2>(h, S, P, Q, q, grid);
    if (traces_shadow_rays(h)) {
      launch_shade(h, S, point, npix, P, Q, F, q, sgrid);
      if (h->bvhWidth == 8) launch_shadow<8>(h, S, P, Q, grid);
      else launch_shadow<2>(h, S, P, Q, grid);
    } else {
      launch_shade(h, S, point, npix, P, Q, F, q, sgrid);
    }
    q ^= 1;
  }
  for (uint32_t p = 0; p < g; ++p) {
    MergeParams mp;
    std::memset(&mp, 0, sizeof(mp));
    mp.width = h->width; mp.height = h->height; mp.slotBase = p * npix;
    mp.accumTargets = 2;
    mp.fireflyClampScale = h->frameDesc.firefly_clamp_scale;
    mp.fireflyClampRelaxPower = h->frameDesc.firefly_clamp_relax_power;
    mp.killNaN = h->frameDesc.kill_nan_samples;
    mp.pass = p;
    hipLaunchKernelGGL(k_merge, dim3(blocks_for(npix, 256)), dim3(256), 0, s, S, P, F, mp, (const float4*)nullptr, dyn);
  }
  hipLaunchKernelGGL(k_advance_dyn, dim3(1), dim3(1), 0, s, dyn, g);
}

uint64_t fnv(uint64_t hsh, const void* p, size_t n) {
  const unsigned char* c = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) { hsh ^= c[i]; hsh *= 1099511628211ull; }
  return hsh;
}

// executable graph of one batch of g passes: captured on first use, reused while nothing it bakes
// in (launch shapes, buffer addresses, frame constants) changes
hipGraphExec_t batch_graph(ypt_integrator* h, const DevScene& S, const PathSoA& P, const Queues& Q, const FilmBuffers& F,
                           uint32_t g, uint32_t G) {
  uint64_t k = 1469598103934665603ull;
  k = fnv(k, &h->fc, sizeof(h->fc));
  k = fnv(k, &S, sizeof(S)); k = fnv(k, &P, sizeof(P)); k = fnv(k, &Q, sizeof(Q)); k = fnv(k, &F, sizeof(F));
  // every launch option that changes a kernel choice, a launch shape or a kernel argument
  int opts[16] = {h->type, h->bvhWidth, h->blockSize, h->ldsDepth, h->persistent, h->persistentWaves, h->refill,
                  h->shadeWaves, h->writeFeatures, h->cus, (int)g, (int)h->width, (int)h->height, (int)h->frameDesc.kill_nan_samples,
                  h->shadeRegWaves, h->leafShare};
  k = fnv(k, opts, sizeof(opts));
  float fp[2] = {h->frameDesc.firefly_clamp_scale, h->frameDesc.firefly_clamp_relax_power};
  k = fnv(k, fp, sizeof(fp));
  const void* ptrs[2] = {h->batchPoints.p, h->mergeDyn.p};
  k = fnv(k, ptrs, sizeof(ptrs));
  for (auto& bg : h->graphs)
    if (bg.exec && bg.key == k) return bg.exec;
  ypt_integrator::BatchGraph& bg = h->graphs[g == G ? 0 : 1];
  if (bg.exec) { HIP_CHECK(hipGraphExecDestroy(bg.exec)); bg.exec = nullptr; }
  if (bg.graph) { HIP_CHECK(hipGraphDestroy(bg.graph)); bg.graph = nullptr; }
  HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  launch_batch(h, S, P, Q, F, h->batchPoints.p, g, h->mergeDyn.p);
  hipGraph_t graph = nullptr;
  HIP_CHECK(hipStreamEndCapture(h->stream, &graph));
  bg.graph = graph;
  HIP_CHECK(hipGraphInstantiate(&bg.exec, graph, nullptr, nullptr, 0));
  bg.key = k;
  return bg.exec;
}

// prepare, build the BVH and upload: a scene resident on `device`
std::shared_ptr<ResidentScene> make_resident(const ypt_scene_desc& d, int device, hipStream_t s) {
  auto r = std::make_shared<ResidentScene>();
  r->device = device;
  prepare_scene(d, host_tables().r2s, r->scene);
  BvhBuilder().build(r->scene.pos.data(), r->scene.idx.data(), (uint32_t)r->scene.triInst.size(), r->bvh);
  for (TriRecord& t : r->bvh.tris) t.alpha = r->scene.triAlpha[t.id];
  for (const MaterialGPU& m : r->scene.mats) r->subsurface = r->subsurface || m.subsurfaceAmount > 0.f;
  r->stackBound2 = r->bvh.stats.maxDepth + 1;
  r->stackBound8 = bvh8_stack_bound(r->bvh.nodes8);
  const PreparedScene& ps = r->scene;
  r->dPos.upload(ps.pos.data(), ps.pos.size(), s);
  r->dNrm.upload(ps.nrm.data(), ps.nrm.size(), s);
  if (!ps.tan.empty()) r->dTan.upload(ps.tan.data(), ps.tan.size(), s);
  r->dIdx.upload(ps.idx.data(), ps.idx.size(), s);
  r->dTriInst.upload(ps.triInst.data(), ps.triInst.size(), s);
  r->dInst.upload(ps.inst.data(), ps.inst.size(), s);
  r->dMats.upload(ps.mats.data(), ps.mats.size(), s);
  if (!ps.lights.empty()) r->dLights.upload(ps.lights.data(), ps.lights.size(), s);
  else { LightGPU dummy; std::memset(&dummy, 0, sizeof(dummy)); r->dLights.upload(&dummy, 1, s); }
  r->dNodes2.upload(r->bvh.nodes2.data(), r->bvh.nodes2.size(), s);
  r->dNodes8.upload(r->bvh.nodes8.data(), r->bvh.nodes8.size(), s);
  r->dTris.upload(r->bvh.tris.data(), r->bvh.tris.size(), s);
  if (!ps.uv.empty()) r->dUv.upload(ps.uv.data(), ps.uv.size(), s);
  if (!ps.textures.empty()) {
    r->dTex.upload(ps.textures.data(), ps.textures.size(), s);
    r->dTexels.upload(ps.texels.data(), ps.texels.size(), s);
    r->dSrgbLut.upload(ps.srgbLut.data(), ps.srgbLut.size(), s);
  }
  if (!ps.envTexels.empty()) r->dEnv.upload(ps.envTexels.data(), ps.envTexels.size(), s);
  if (!ps.alphaTris.empty()) {
    r->dAlpha.upload(ps.alphaTris.data(), ps.alphaTris.size(), s);
    r->dTriAlpha.upload(ps.triAlpha.data(), ps.triAlpha.size(), s);
  }
  HIP_CHECK(hipStreamSynchronize(s));
  return r;
}

void attach_scene(ypt_integrator* h, std::shared_ptr<ResidentScene> r) {
  h->res = std::move(r);
  h->hasScene = true;
  if (h->hasFrame) {
    prepare_frame(h->frameDesc, h->res->scene, (uint32_t)h->type, h->fc);
    clear_film(h);
  }
}

// ---------------------------------------------------------------- bidirectional integrator
void ensure_bdpt(ypt_integrator* h) {
  const FrameConstants& fc = h->fc;
  uint32_t W = h->width, H = h->height;
  uint32_t npix = W * H;
  uint32_t npaths = fc.lightPathsPerDim[0] * fc.lightPathsPerDim[1];
  size_t nrec = (size_t)npaths * (BDPT_MAX_LIGHT_VERTICES - 1);
  if (h->bW != W || h->bH != H) {
    for (auto& L : h->lset) {
      L.nodes.ensure((size_t)npaths * BDPT_MAX_LIGHT_VERTICES);
      L.surf.ensure((size_t)npaths * BDPT_MAX_LIGHT_VERTICES);
      L.pre.ensure((size_t)npaths * BDPT_MAX_LIGHT_VERTICES);
      L.counts.ensure(npaths); L.lambdas.ensure(npaths); L.headers.ensure(npaths); L.headerCount.ensure(1);
    }
    h->bCamXYZ.ensure(npix);
    h->bKeys.ensure(nrec); h->bVals.ensure(nrec); h->bKeysOut.ensure(nrec); h->bValsOut.ensure(nrec); h->bRecXYZ.ensure(nrec);
    std::vector<uint32_t> perm(npix);
    texel_permutation(W, H, 0x59415054ull, perm.data());
    h->bPerm.upload(perm.data(), perm.size(), h->stream);
    h->bW = W;
    h->bH = H;
    HIP_CHECK(hipStreamSynchronize(h->stream));
  }
  uint32_t groups = ((W + BDPT_TILE - 1) / BDPT_TILE) * ((H + BDPT_TILE - 1) / BDPT_TILE);
  uint32_t stride = groups * 64u;
  uint32_t bound = h->bvhWidth == 8 ? h->res->stackBound8 : h->res->stackBound2;
  size_t depth = bound > (uint32_t)h->ldsDepth ? bound - (uint32_t)h->ldsDepth : 1;
  h->bOvf.ensure(depth * stride);
  h->bOvfStride = stride;
  for (auto& L : h->lset) L.ovf.ensure(depth * npaths);  // the light kernel's own stack overflow, stride npaths
  if (!h->lightStream) {
    HIP_CHECK(hipStreamCreateWithFlags(&h->lightStream, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
      HIP_CHECK(hipEventCreateWithFlags(&h->lightDone[i], hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&h->setFree[i], hipEventDisableTiming));
    }
    HIP_CHECK(hipEventCreateWithFlags(&h->camReady, hipEventDisableTiming));
  }
  if (h->writeFeatures) h->features.ensure((size_t)npix * FEATURE_STRIDE);
  if (h->bdptWavefront) {
    h->cOrg.ensure(npix); h->cDir.ensure(npix); h->cThr.ensure(npix); h->cTotal.ensure(npix);
    h->cPrev.ensure(npix); h->cPrevN.ensure(npix); h->cDeferred.ensure(npix);
    h->cSt0.ensure(npix); h->cSt1.ensure(npix); h->cVolIor.ensure(npix); h->cVolMat.ensure(npix); h->cVolAbs.ensure((size_t)npix * 4);
    // a subsurface entry and its exit are shaded in one launch: twice the connections per path
    size_t nsh = (size_t)npix * h->fc.lightPathVertices * (h->res->subsurface ? 2u : 1u);
    h->cSOrg.ensure(nsh); h->cSDir.ensure(nsh); h->cSContrib.ensure(nsh); h->cSSlot.ensure(nsh);
    h->cConn.ensure(nsh); h->cPending.ensure(npix);
    if (h->bdptSplit && !h->res->subsurface) {
      h->cDefer.ensure(npix);
      h->cConnList.ensure((size_t)npix * h->fc.lightPathVertices);
    }
    ensure_slots(h, npix);
    ensure_overflow(h);
  }
}

// camera subpaths of one pass as a wavefront: start, then per vertex a persistent closest-hit launch,
// the shading launch (connections into the ray queue) and a persistent occlusion launch
template <int WIDTH, bool COUNT>
void bdpt_camera_wavefront(ypt_integrator* h, const DevScene& S, const uint32_t* point, const BdptBuffers& B) {
  hipStream_t s = h->stream;
  const FrameConstants& fc = h->fc;
  uint32_t npix = h->width * h->height;
  CamSoA C;
  C.orgLambda = h->cOrg.p; C.dirPdf = h->cDir.p; C.thr = h->cThr.p; C.total = h->cTotal.p;
  C.st0 = h->cSt0.p; C.st1 = h->cSt1.p; C.prev = h->cPrev.p; C.prevN = h->cPrevN.p; C.deferred = h->cDeferred.p;
  C.volIor = h->cVolIor.p; C.volMat = h->cVolMat.p; C.volAbs = h->cVolAbs.p;
  C.conn = h->cConn.p; C.pending = h->cPending.p;
  C.connCap = fc.lightPathVertices * (h->res->subsurface ? 2u : 1u);
  // connections in a launch of their own, one thread per (vertex, light vertex) (not with subsurface
  // scenes, whose shading launch can shade two vertices of a path)
  const bool split = h->bdptSplit != 0 && !h->res->subsurface;
  C.defer = split ? h->cDefer.p : nullptr;
  C.connList = split ? h->cConnList.p : nullptr;
  uint32_t xgrid = std::max<uint32_t>(1u, std::min<uint32_t>(blocks_for((size_t)npix * fc.lightPathVertices, 64), (uint32_t)h->cus * 32u));
  PathSoA P;
  std::memset(&P, 0, sizeof(P));
  // the occlusion kernel adds each unoccluded connection to its own record (one ray per record, so
  // the plain read-add-write form); the queue entry's slot is the record index
  P.orgLambda = C.orgLambda; P.dirPdf = C.dirPdf; P.light = C.conn;
  Queues Q = h->queues();
  Q.sOrgTmax = h->cSOrg.p; Q.sDirTri = h->cSDir.p; Q.sContrib = h->cSContrib.p; Q.sSlot = h->cSSlot.p;
  Q.workCounters = h->work.p + 8;
  // short queues wave-per-ray (not with subsurface scenes, whose probes share the stack buffers)
  Q.waveBelow = h->res->subsurface ? 0u : (uint32_t)h->waveBelow;
  float* feats = h->writeFeatures ? h->features.p : (float*)nullptr;
  uint32_t Bk = (uint32_t)h->blockSize;
  uint32_t pgrid = std::max<uint32_t>(1u, std::min<uint32_t>(blocks_for(npix, Bk), (uint32_t)((size_t)h->cus * h->persistentWaves * 64 / Bk)));
  uint32_t sgrid = std::max<uint32_t>(1u, std::min<uint32_t>(blocks_for(npix, Bk), (uint32_t)((size_t)h->cus * h->persistentWaves * 64 / Bk)));
  uint32_t cgrid = std::min<uint32_t>(blocks_for(npix, 64), (uint32_t)h->cus * 32u);
  size_t lds = (size_t)h->ldsDepth * Bk * 4;
  uint32_t d = (uint32_t)h->ldsDepth, r = (uint32_t)h->refill;
  hipLaunchKernelGGL(k_bdpt_cstart, dim3(blocks_for(npix, 256)), dim3(256), 0, s, fc, point, B, C, Q, feats);
  bool alpha = S.tris.alpha.tris != nullptr;
  int q = 0;
  for (uint32_t k = 0; k < fc.maxRayDepth; ++k) {
    if (COUNT || alpha)
      YPT_PERSISTENT((k_extend_p<WIDTH, COUNT>), pgrid, Bk, lds, s, S, P, Q, q, d, r, (uint32_t)h->leafShare);
    else
      YPT_PERSISTENT((k_extend_p<WIDTH, false, false>), pgrid, Bk, lds, s, S, P, Q, q, d, r, (uint32_t)h->leafShare);
    launch_extend_wave<WIDTH>(h, S, P, Q, q);
    if (h->res->subsurface)  // probes in the shading launch, within the stack buffer's 32 workgroups per CU
      hipLaunchKernelGGL((k_bdpt_cshade<1, true>), dim3(cgrid), dim3(64), 0, s, fc, S, point, B, C, Q, q, k == 0 ? 1 : 0, feats);
    else if (split) {
      if (h->bdptWaves == 1)
        hipLaunchKernelGGL((k_bdpt_cshade<1, false, true>), dim3(cgrid), dim3(64), 0, s, fc, S, point, B, C, Q, q, k == 0 ? 1 : 0, feats);
      else
        hipLaunchKernelGGL((k_bdpt_cshade<2, false, true>), dim3(cgrid), dim3(64), 0, s, fc, S, point, B, C, Q, q, k == 0 ? 1 : 0, feats);
      if (h->bdptSplit == 1)
        hipLaunchKernelGGL(k_bdpt_connect<1>, dim3(xgrid), dim3(64), 0, s, fc, S, B, C, Q, q);
      else if (h->bdptSplit == 2)
        hipLaunchKernelGGL(k_bdpt_connect<2>, dim3(xgrid), dim3(64), 0, s, fc, S, B, C, Q, q);
      else
        hipLaunchKernelGGL(k_bdpt_connect<4>, dim3(xgrid), dim3(64), 0, s, fc, S, B, C, Q, q);
    } else if (h->bdptWaves == 1)
      hipLaunchKernelGGL(k_bdpt_cshade<1>, dim3(cgrid), dim3(64), 0, s, fc, S, point, B, C, Q, q, k == 0 ? 1 : 0, feats);
    else if (h->bdptWaves == 2)
      hipLaunchKernelGGL(k_bdpt_cshade<2>, dim3(cgrid), dim3(64), 0, s, fc, S, point, B, C, Q, q, k == 0 ? 1 : 0, feats);
    else
      hipLaunchKernelGGL(k_bdpt_cshade<4>, dim3(cgrid), dim3(64), 0, s, fc, S, point, B, C, Q, q, k == 0 ? 1 : 0, feats);
    if (COUNT || alpha)
      YPT_PERSISTENT((k_shadow_p<WIDTH, COUNT, false, true>), sgrid, Bk, lds, s, S, P, Q, d, r, (uint32_t)h->leafShare);
    else
      YPT_PERSISTENT((k_shadow_p<WIDTH, false, false, false>), sgrid, Bk, lds, s, S, P, Q, d, r, (uint32_t)h->leafShare);
    launch_shadow_wave<WIDTH>(h, S, P, Q);
    q ^= 1;
  }
  hipLaunchKernelGGL(k_bdpt_cfinish, dim3(blocks_for(npix, 256)), dim3(256), 0, s, fc, S, B, C);
}

void mark_span(ypt_integrator* h, hipStream_t s, std::vector<hipEvent_t>* ev) {
  if (!ev) return;
  hipEvent_t e = h->next_event();
  HIP_CHECK(hipEventRecord(e, s));
  ev->push_back(e);
}

// light subpaths of one pass on stream ls, into the set that B points to: trace, vertex surfaces,
// headers (profile spans: light, headers)
template <int WIDTH, bool COUNT>
void bdpt_light_stage(ypt_integrator* h, const DevScene& S, const uint32_t* point, const BdptBuffers& B, hipStream_t ls,
                      std::vector<hipEvent_t>* ev) {
  const FrameConstants& fc = h->fc;
  uint32_t d = (uint32_t)h->ldsDepth;
  size_t lds = (size_t)h->ldsDepth * 64 * 4;
  mark_span(h, ls, ev);
  if (h->res->subsurface)
    hipLaunchKernelGGL((k_bdpt_light<WIDTH, COUNT, true>), dim3(blocks_for(B.npaths, 64)), dim3(64), lds, ls, fc, S, point, B, d);
  else
    hipLaunchKernelGGL((k_bdpt_light<WIDTH, COUNT>), dim3(blocks_for(B.npaths, 64)), dim3(64), lds, ls, fc, S, point, B, d);
  hipLaunchKernelGGL(k_bdpt_node_surfaces, dim3(blocks_for((size_t)B.npaths * BDPT_MAX_LIGHT_VERTICES, 256)), dim3(256), 0, ls, S, B);
  mark_span(h, ls, ev);
  hipLaunchKernelGGL(k_bdpt_headers, dim3(1), dim3(1024), 0, ls, B);
}

// camera subpaths and light tracing of one pass on the handle's stream (profile spans: camera, splat)
template <int WIDTH, bool COUNT>
void bdpt_camera_stage(ypt_integrator* h, const DevScene& S, const uint32_t* point, const BdptBuffers& B, uint32_t groups,
                       std::vector<hipEvent_t>* ev) {
  hipStream_t s = h->stream;
  const FrameConstants& fc = h->fc;
  uint32_t d = (uint32_t)h->ldsDepth;
  size_t lds = (size_t)h->ldsDepth * 64 * 4;
  mark_span(h, s, ev);
  if (h->bdptWavefront) bdpt_camera_wavefront<WIDTH, COUNT>(h, S, point, B);
  else if (h->res->subsurface)
    hipLaunchKernelGGL((k_bdpt_camera<WIDTH, COUNT, 1, true>), dim3(groups), dim3(64), lds, s, fc, S, point, B,
                       h->writeFeatures ? h->features.p : (float*)nullptr, d);
  else if (h->bdptWaves == 2)
    hipLaunchKernelGGL((k_bdpt_camera<WIDTH, COUNT, 2>), dim3(groups), dim3(64), lds, s, fc, S, point, B,
                       h->writeFeatures ? h->features.p : (float*)nullptr, d);
  else if (h->bdptWaves == 3)
    hipLaunchKernelGGL((k_bdpt_camera<WIDTH, COUNT, 3>), dim3(groups), dim3(64), lds, s, fc, S, point, B,
                       h->writeFeatures ? h->features.p : (float*)nullptr, d);
  else if (h->bdptWaves == 4)
    hipLaunchKernelGGL((k_bdpt_camera<WIDTH, COUNT, 4>), dim3(groups), dim3(64), lds, s, fc, S, point, B,
                       h->writeFeatures ? h->features.p : (float*)nullptr, d);
  else if (h->bdptWaves == 6)
    hipLaunchKernelGGL((k_bdpt_camera<WIDTH, COUNT, 6>), dim3(groups), dim3(64), lds, s, fc, S, point, B,
                       h->writeFeatures ? h->features.p : (float*)nullptr, d);
  else if (h->bdptWaves == 8)
    hipLaunchKernelGGL((k_bdpt_camera<WIDTH, COUNT, 8>), dim3(groups), dim3(64), lds, s, fc, S, point, B,
                       h->writeFeatures ? h->features.p : (float*)nullptr, d);
  else
    hipLaunchKernelGGL((k_bdpt_camera<WIDTH, COUNT, 1>), dim3(groups), dim3(64), lds, s, fc, S, point, B,
                       h->writeFeatures ? h->features.p : (float*)nullptr, d);
  mark_span(h, s, ev);
  hipLaunchKernelGGL((k_bdpt_splat<WIDTH, COUNT>), dim3(blocks_for((uint64_t)B.npaths * B.recPerPath, 64)), dim3(64), lds, s, fc, S,
                     point, B, d);
  mark_span(h, s, ev);
}

void render_bdpt(ypt_integrator* h, uint32_t sample_begin, uint32_t sample_count) {
  ensure_bdpt(h);
  hipStream_t s = h->stream;
  const FrameConstants& fc = h->fc;
  uint32_t W = h->width, H = h->height, npix = W * H;
  uint32_t groups = ((W + BDPT_TILE - 1) / BDPT_TILE) * ((H + BDPT_TILE - 1) / BDPT_TILE);
  std::vector<uint32_t> pts((size_t)sample_count * YPT_SOBOL_POINT_DIMS);
  for (uint32_t k = 0; k < sample_count; ++k)
    for (uint32_t dd = 0; dd < YPT_SOBOL_POINT_DIMS; ++dd)
      pts[(size_t)k * YPT_SOBOL_POINT_DIMS + dd] = sobol_uint(YPT_SOBOL_MATRICES, sample_begin + k, dd);
  if (h->sobolPoints.n < pts.size()) h->sobolPoints.alloc(pts.size());
  h->sobolPoints.upload(pts.data(), pts.size(), s);
  BdptBuffers B{};
  B.perm = h->bPerm.p; B.camXYZ = h->bCamXYZ.p;
  B.recKeys = h->bKeys.p; B.recVals = h->bVals.p; B.recXYZ = h->bRecXYZ.p;
  B.ovf = h->bOvf.p; B.ovfStride = h->bOvfStride;
  B.npaths = fc.lightPathsPerDim[0] * fc.lightPathsPerDim[1];
  B.recPerPath = fc.lightPathVertices - 1;
  uint32_t nrec = B.npaths * B.recPerPath;
  int bits = 32 - __builtin_clz(npix);
  {
    size_t need = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, h->bKeys.p, h->bKeysOut.p, h->bVals.p, h->bValsOut.p, (int)nrec, 0,
                                                 bits, s));
    if (need > h->bSortBytes) {
      h->bSortTemp.alloc(need);
      h->bSortBytes = need;
    }
  }
  DevScene S = h->dev_scene();
  PathSoA P = h->path_soa();
  FilmBuffers F = h->film();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  h->evUsed = 0;
  B.work = h->work.p;
  if (h->countTraversal) h->work.zero(s);
  std::vector<std::vector<hipEvent_t>> spans;
  if (h->profile) { e0 = h->next_event(); HIP_CHECK(hipEventRecord(e0, s)); }
  // Light subpaths of pass k+1 are traced on lightStream while the camera stage of pass k runs (not in
  // profile mode, whose spans need one stream). Set k & 1 holds pass k's subpaths; it is free again
  // once pass k's splat kernel has read it.
  const bool overlap = !h->profile;
  hipStream_t ls = overlap ? h->lightStream : s;
  auto with_set = [&](uint32_t k, bool lightKernel) {
    ypt_integrator::LightSet& L = h->lset[k & 1u];
    BdptBuffers b = B;
    b.nodes = L.nodes.p; b.surf = L.surf.p; b.pre = L.pre.p; b.counts = L.counts.p; b.lambdas = L.lambdas.p;
    b.headers = L.headers.p; b.headerCount = L.headerCount.p;
    if (lightKernel) { b.ovf = L.ovf.p; b.ovfStride = b.npaths; }
    return b;
  };
  const bool cnt = h->countTraversal != 0;
  auto light = [&](uint32_t k, std::vector<hipEvent_t>* ev) {
    if (overlap && k >= 2) HIP_CHECK(hipStreamWaitEvent(ls, h->setFree[k & 1u], 0));
    const uint32_t* point = h->sobolPoints.p + (size_t)k * YPT_SOBOL_POINT_DIMS;
    BdptBuffers b = with_set(k, true);
    if (h->bvhWidth == 8) { if (cnt) bdpt_light_stage<8, true>(h, S, point, b, ls, ev); else bdpt_light_stage<8, false>(h, S, point, b, ls, ev); }
    else { if (cnt) bdpt_light_stage<2, true>(h, S, point, b, ls, ev); else bdpt_light_stage<2, false>(h, S, point, b, ls, ev); }
    if (overlap) HIP_CHECK(hipEventRecord(h->lightDone[k & 1u], ls));
  };
  if (overlap) {  // the light stream starts after everything already queued on the handle's stream
    HIP_CHECK(hipEventRecord(h->camReady, s));
    HIP_CHECK(hipStreamWaitEvent(ls, h->camReady, 0));
    if (sample_count > 0) light(0, nullptr);
  }
  for (uint32_t k = 0; k < sample_count; ++k) {
    const uint32_t* point = h->sobolPoints.p + (size_t)k * YPT_SOBOL_POINT_DIMS;
    std::vector<hipEvent_t>* ev = nullptr;
    if (h->profile) { spans.emplace_back(); ev = &spans.back(); }
    if (overlap) {
      if (k + 1 < sample_count) light(k + 1, nullptr);
      HIP_CHECK(hipStreamWaitEvent(s, h->lightDone[k & 1u], 0));
    } else {
      light(k, ev);
    }
    BdptBuffers b = with_set(k, false);
    if (h->bvhWidth == 8) { if (cnt) bdpt_camera_stage<8, true>(h, S, point, b, groups, ev); else bdpt_camera_stage<8, false>(h, S, point, b, groups, ev); }
    else { if (cnt) bdpt_camera_stage<2, true>(h, S, point, b, groups, ev); else bdpt_camera_stage<2, false>(h, S, point, b, groups, ev); }
    if (overlap) HIP_CHECK(hipEventRecord(h->setFree[k & 1u], s));
    size_t bytes = h->bSortBytes;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(h->bSortTemp.p, bytes, h->bKeys.p, h->bKeysOut.p, h->bVals.p, h->bValsOut.p,
                                                 (int)nrec, 0, bits, s));
    hipLaunchKernelGGL(k_bdpt_accum, dim3(blocks_for(nrec, 256)), dim3(256), 0, s, h->bKeysOut.p, h->bValsOut.p, h->bRecXYZ.p,
                       nrec, npix, h->bCamXYZ.p);
    MergeParams mp;
    mp.width = W; mp.height = H; mp.slotBase = 0;
    mp.sampleCount = h->filmSamples; mp.accumTargets = 2;
    mp.fireflyClampScale = h->frameDesc.firefly_clamp_scale;
    mp.fireflyClampRelaxPower = h->frameDesc.firefly_clamp_relax_power;
    mp.killNaN = h->frameDesc.kill_nan_samples;
    mp.cur = h->filmCur;
    mp.pass = 0;
    hipLaunchKernelGGL(k_merge, dim3(blocks_for(npix, 256)), dim3(256), 0, s, S, P, F, mp, (const float4*)h->bCamXYZ.p, (const uint32_t*)nullptr);
    if (ev) {
      hipEvent_t e = h->next_event();
      HIP_CHECK(hipEventRecord(e, s));
      ev->push_back(e);
    }
    h->filmSamples++;
    h->filmCur ^= 1;
  }
  HIP_CHECK(hipGetLastError());
  h->lastPassBase = 0;
  std::memset(h->stats, 0, sizeof(h->stats));
  if (h->profile) {
    e1 = h->next_event();
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    float all = 0.f;
    HIP_CHECK(hipEventElapsedTime(&all, e0, e1));
    h->stats[5] = all;
    // [2] camera subpaths, [3] light subpaths, [4] light tracing splats, [11] headers, [12] sort + sum + merge
    for (const auto& v : spans) {
      float t[5];
      for (int i = 0; i < 5; ++i) HIP_CHECK(hipEventElapsedTime(&t[i], v[i], v[i + 1]));
      h->stats[3] += t[0]; h->stats[11] += t[1]; h->stats[2] += t[2]; h->stats[4] += t[3]; h->stats[12] += t[4];
    }
  }
  h->stats[10] = sample_count;  // camera-subpath launches
  h->stats[13] = 1;
}

// A render on a caller's stream: that stream first waits for the work already queued on the
// library's stream, the render's kernels run on it, and the library's stream then waits for them,
// so every later call on the handle (which synchronizes the library's stream) sees the result. The
// handle keeps no reference to the caller's stream after the call.
struct CallerStream {
  ypt_integrator* h;
  bool active;
  CallerStream(ypt_integrator* hh, hipStream_t caller) : h(hh), active(caller != nullptr && caller != hh->ownStream) {
    if (!active) return;
    HIP_CHECK(hipEventRecord(h->streamJoin, h->ownStream));
    HIP_CHECK(hipStreamWaitEvent(caller, h->streamJoin, 0));
    h->stream = caller;
  }
  ~CallerStream() {
    if (!active) return;
    if (hipEventRecord(h->streamJoin, h->stream) == hipSuccess) (void)hipStreamWaitEvent(h->ownStream, h->streamJoin, 0);
    h->stream = h->ownStream;
  }
};

}  // namespace

extern "C" {

const char* ypt_last_error(void) { return g_error.c_str(); }
const char* ypt_version(void) { return "yapt_pt 0.1 (gfx950 wavefront)"; }

int ypt_sobol_point(uint32_t sample_index, uint32_t* out256) {
  if (!out256) return fail(YPT_E_INVALID, "null output");
  for (uint32_t d = 0; d < YPT_SOBOL_POINT_DIMS; ++d) out256[d] = sobol_uint(YPT_SOBOL_MATRICES, sample_index, d);
  return YPT_OK;
}

int ypt_create(int integrator_type, int device, ypt_integrator** out) {
  return guarded([&]() -> int {
    if (!out) return fail(YPT_E_INVALID, "null handle pointer");
    if (integrator_type < YPT_INTEGRATOR_BPT || integrator_type > YPT_INTEGRATOR_BDPT)
      return fail(YPT_E_INVALID, "unknown integrator type " + std::to_string(integrator_type));
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(YPT_E_INVALID, "device index out of range");
    HIP_CHECK(hipSetDevice(device));
    std::unique_ptr<ypt_integrator> h(new ypt_integrator());
    h->type = integrator_type;
    h->device = device;
    HIP_CHECK(hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, device));
    if (h->cus < 1) h->cus = 1;
    HIP_CHECK(hipStreamCreateWithFlags(&h->ownStream, hipStreamNonBlocking));
    h->stream = h->ownStream;
    HIP_CHECK(hipEventCreateWithFlags(&h->streamJoin, hipEventDisableTiming));
    for (auto& e : h->ev) HIP_CHECK(hipEventCreate(&e));
    init_scratch(device);
    upload_constant_tables(h.get());
    *out = h.release();
    return YPT_OK;
  });
}

int ypt_destroy(ypt_integrator* h) {
  if (!h) return YPT_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto& e : h->ev) if (e) (void)hipEventDestroy(e);
  for (auto& e : h->evPool) (void)hipEventDestroy(e);
  for (auto& g : h->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  if (h->streamJoin) (void)hipEventDestroy(h->streamJoin);
  if (h->ownStream) (void)hipStreamDestroy(h->ownStream);  // never a caller's stream
  if (h->lightStream) {
    (void)hipStreamSynchronize(h->lightStream);
    for (int i = 0; i < 2; ++i) {
      (void)hipEventDestroy(h->lightDone[i]);
      (void)hipEventDestroy(h->setFree[i]);
    }
    (void)hipEventDestroy(h->camReady);
    (void)hipStreamDestroy(h->lightStream);
  }
  delete h;
  return YPT_OK;
}

int ypt_set_option(ypt_integrator* h, const char* name, int64_t value) {
  if (!h || !name) return fail(YPT_E_INVALID, "null argument");
  std::string n(name);
  if (n == "bvh_width") {
    if (value != 2 && value != 8) return fail(YPT_E_INVALID, "bvh_width must be 2 or 8");
    h->bvhWidth = (int)value;
  } else if (n == "block_size") {
    if (value < 64 || value > 1024 || (value % 64) != 0) return fail(YPT_E_INVALID, "block_size must be a multiple of 64 in [64, 1024]");
    h->blockSize = (int)value;
  } else if (n == "lds_stack_depth") {
    if (value < 1 || value > 64) return fail(YPT_E_INVALID, "lds_stack_depth must be in [1, 64]");
    h->ldsDepth = (int)value;
    h->ovf.release();
  } else if (n == "count_traversal") {
    h->countTraversal = value != 0;
  } else if (n == "first_hit_features") {
    h->writeFeatures = value != 0;
  } else if (n == "passes_in_flight") {
    if (value < 0 || value > (int64_t)MAX_PASSES_IN_FLIGHT)
      return fail(YPT_E_INVALID, "passes_in_flight must be in [0, " + std::to_string(MAX_PASSES_IN_FLIGHT) + "] (0 = automatic)");
    h->passesInFlight = (int)value;
  } else if (n == "profile") {
    h->profile = value != 0;
  } else if (n == "persistent") {
    h->persistent = value != 0;
  } else if (n == "persistent_waves") {
    if (value < 1 || value > 64) return fail(YPT_E_INVALID, "persistent_waves must be in [1, 64] (waves per CU)");
    h->persistentWaves = (int)value;
  } else if (n == "graph") {
    h->useGraph = value != 0;
  } else if (n == "bdpt_wavefront") {
    h->bdptWavefront = value != 0;
  } else if (n == "shade_reg_waves") {
    if (value != 2 && value != 4 && value != 8) return fail(YPT_E_INVALID, "shade_reg_waves must be 2, 4 or 8");
    h->shadeRegWaves = (int)value;
  } else if (n == "bdpt_split_connections") {
    if (value != 0 && value != 1 && value != 2 && value != 4) return fail(YPT_E_INVALID, "bdpt_split_connections must be 0, 1, 2 or 4");
    h->bdptSplit = (int)value;
  } else if (n == "bdpt_waves") {
    if (value < 1 || value > 8 || value == 5 || value == 7) return fail(YPT_E_INVALID, "bdpt_waves must be 1, 2, 3, 4, 6 or 8");
    h->bdptWaves = (int)value;
  } else if (n == "shade_waves") {
    if (value < 0 || value > 1024) return fail(YPT_E_INVALID, "shade_waves must be in [0, 1024] (0 = one workgroup per entry)");
    h->shadeWaves = (int)value;
  } else if (n == "leaf_share") {
    if (value < 1 || value > 64) return fail(YPT_E_INVALID, "leaf_share must be in [1, 64]");
    h->leafShare = (int)value;
  } else if (n == "refill_threshold") {
    if (value < 1 || value > 64) return fail(YPT_E_INVALID, "refill_threshold must be in [1, 64] (lanes of a wave)");
    h->refill = (int)value;
  } else if (n == "steps_per_fetch") {
    if (value < 1 || value > 64) return fail(YPT_E_INVALID, "steps_per_fetch must be in [1, 64]");
    h->stepsPerFetch = (int)value;
  } else if (n == "wave_ray_below") {
    if (value < 0 || value > (1 << 30)) return fail(YPT_E_INVALID, "wave_ray_below must be in [0, 2^30] (0: thread-per-ray only)");
    h->waveBelow = (int)value;
  } else {
    return fail(YPT_E_INVALID, "unknown option " + n);
  }
  return YPT_OK;
}

int ypt_set_scene(ypt_integrator* h, const ypt_scene_desc* d) {
  if (!h || !d) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    HIP_CHECK(hipSetDevice(h->device));
    attach_scene(h, make_resident(*d, h->device, h->stream));
    return YPT_OK;
  });
}

int ypt_set_frame(ypt_integrator* h, const ypt_frame_desc* f) {
  if (!h || !f) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasScene) return fail(YPT_E_NOSCENE, "set_scene must come before set_frame");
    HIP_CHECK(hipSetDevice(h->device));
    FrameConstants fc;
    prepare_frame(*f, h->res->scene, (uint32_t)h->type, fc);
    // RGB -> spectrum at run time: environment colours and punctual light intensities
    bool needTable = (f->env_intensity_scale != 0.f && !(f->flags & YPT_FLAG_WHITE_FURNACE)) || fc.punctualLightCount > 0 ||
                     !h->res->scene.textures.empty();
    if (needTable && h->dRgb2spec.n == 0) {
      std::vector<float> table;
      host_tables().r2s.full_table(table, std::max(1u, std::thread::hardware_concurrency()));
      h->dRgb2spec.upload(table.data(), table.size(), h->stream);
    }
    bool resized = f->width != h->width || f->height != h->height;
    ensure_frame_buffers(h, f->width, f->height);
    h->frameDesc = *f;
    h->fc = fc;
    h->hasFrame = true;
    if (resized) clear_film(h);
    HIP_CHECK(hipStreamSynchronize(h->stream));
    return YPT_OK;
  });
}

int ypt_clear_film(ypt_integrator* h) {
  if (!h) return fail(YPT_E_INVALID, "null handle");
  return guarded([&]() -> int {
    if (!h->hasFrame) return fail(YPT_E_NOSCENE, "no frame set");
    HIP_CHECK(hipSetDevice(h->device));
    clear_film(h);
    return YPT_OK;
  });
}

int ypt_render(ypt_integrator* h, uint32_t sample_begin, uint32_t sample_count, void* stream) {
  if (!h) return fail(YPT_E_INVALID, "null handle");
  return guarded([&]() -> int {
    if (!h->hasScene || !h->hasFrame) return fail(YPT_E_NOSCENE, "render needs a scene and a frame");
    HIP_CHECK(hipSetDevice(h->device));
    CallerStream scope(h, (hipStream_t)stream);
    if (h->writeFeatures) h->featAccum.ensure((size_t)4 * h->width * h->height);
    if (h->type == YPT_INTEGRATOR_BDPT) {
      render_bdpt(h, sample_begin, sample_count);
      return YPT_OK;
    }
    hipStream_t s = h->stream;
    uint32_t npix = h->width * h->height;
    uint32_t G = passes_in_flight(h, sample_count);
    ensure_slots(h, (size_t)npix * G);
    ensure_overflow(h);
    // Sobol points of all passes, one upload
    std::vector<uint32_t> pts((size_t)sample_count * YPT_SOBOL_POINT_DIMS);
    for (uint32_t k = 0; k < sample_count; ++k)
      for (uint32_t d = 0; d < YPT_SOBOL_POINT_DIMS; ++d)
        pts[(size_t)k * YPT_SOBOL_POINT_DIMS + d] = sobol_uint(YPT_SOBOL_MATRICES, sample_begin + k, d);
    if (h->sobolPoints.n < pts.size()) h->sobolPoints.alloc(pts.size());
    h->sobolPoints.upload(pts.data(), pts.size(), s);
    if (h->countTraversal) h->work.zero(s);
    DevScene S = h->dev_scene();
    PathSoA P = h->path_soa();
    Queues Q = h->queues();
    // short bounce queues wave-per-ray (not with subsurface scenes, whose shadow launch adds two rays
    // per path atomically)
    Q.waveBelow = h->res->subsurface ? 0u : (uint32_t)h->waveBelow;
    FilmBuffers F = h->film();
    uint32_t B = (uint32_t)h->blockSize;
    // profile mode: events on this stream around every kernel, read once the render has finished
    struct Span { hipEvent_t a, b; int kind; };  // kind 0 extend, 1 shade, 2 shadow, 3 start, 4 merge
    std::vector<Span> spans;
    h->evUsed = 0;
    auto mark = [&]() -> hipEvent_t {
      hipEvent_t e = h->next_event();
      HIP_CHECK(hipEventRecord(e, s));
      return e;
    };
    hipEvent_t eAll0 = nullptr;
    if (h->profile) eAll0 = mark();
    if (h->useGraph && !h->profile && !h->countTraversal) {
      h->batchPoints.ensure((size_t)MAX_PASSES_IN_FLIGHT * YPT_SOBOL_POINT_DIMS);
      h->mergeDyn.ensure(2);
      hipLaunchKernelGGL(k_set_dyn, dim3(1), dim3(1), 0, s, h->mergeDyn.p, h->filmSamples, h->filmCur);
      for (uint32_t k = 0; k < sample_count; k += G) {
        uint32_t g = std::min(G, sample_count - k);
        HIP_CHECK(hipMemcpyAsync(h->batchPoints.p, h->sobolPoints.p + (size_t)k * YPT_SOBOL_POINT_DIMS,
                                 (size_t)g * YPT_SOBOL_POINT_DIMS * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipGraphLaunch(batch_graph(h, S, P, Q, F, g, G), s));
        h->filmSamples += g;
        h->filmCur ^= (g & 1u);
        h->lastPassBase = (g - 1) * npix;
      }
      HIP_CHECK(hipGetLastError());
      std::memset(h->stats, 0, sizeof(h->stats));
      h->stats[10] = (double)h->fc.maxRayDepth * ((sample_count + G - 1) / G);
      h->stats[13] = G;
      return YPT_OK;
    }
    for (uint32_t k = 0; k < sample_count; k += G) {
      uint32_t g = std::min(G, sample_count - k);
      uint32_t nslots = npix * g;
      uint32_t grid = work_grid(h, nslots);
      uint32_t sgrid = shade_grid(h, nslots);
      const uint32_t* point = h->sobolPoints.p + (size_t)k * YPT_SOBOL_POINT_DIMS;
      hipEvent_t e0 = h->profile ? mark() : nullptr;
      hipLaunchKernelGGL(k_start, dim3(blocks_for(nslots, 256)), dim3(256), 0, s, h->fc, point, P, Q, F, npix, nslots,
                         h->writeFeatures);
      if (h->profile) spans.push_back({e0, mark(), 3});
      int q = 0;
      for (uint32_t b = 0; b < h->fc.maxRayDepth; ++b) {
        hipEvent_t a0 = h->profile ? mark() : nullptr;
        if (h->bvhWidth == 8) launch_extend<8>(h, S, P, Q, q, grid);
        else launch_extend<2>(h, S, P, Q, q, grid);
        hipEvent_t a1 = h->profile ? mark() : nullptr;
        launch_shade(h, S, point, npix, P, Q, F, q, sgrid);
        hipEvent_t a2 = h->profile ? mark() : nullptr;
        if (h->profile) { spans.push_back({a0, a1, 0}); spans.push_back({a1, a2, 1}); }
        if (traces_shadow_rays(h)) {
          if (h->bvhWidth == 8) launch_shadow<8>(h, S, P, Q, grid);
          else launch_shadow<2>(h, S, P, Q, grid);
          if (h->profile) spans.push_back({a2, mark(), 2});
        }
        q ^= 1;
      }
      for (uint32_t p = 0; p < g; ++p) {  // passes merge in sample order
        MergeParams mp;
        mp.width = h->width; mp.height = h->height; mp.slotBase = p * npix;
        mp.sampleCount = h->filmSamples; mp.accumTargets = 2;
        mp.fireflyClampScale = h->frameDesc.firefly_clamp_scale;
        mp.fireflyClampRelaxPower = h->frameDesc.firefly_clamp_relax_power;
        mp.killNaN = h->frameDesc.kill_nan_samples;
        mp.cur = h->filmCur;
        mp.pass = p;
        hipEvent_t m0 = h->profile ? mark() : nullptr;
        hipLaunchKernelGGL(k_merge, dim3(blocks_for(npix, 256)), dim3(256), 0, s, S, P, F, mp, (const float4*)nullptr, (const uint32_t*)nullptr);
        if (h->profile) spans.push_back({m0, mark(), 4});
        h->filmSamples++;
        h->filmCur ^= 1;
      }
      h->lastPassBase = (g - 1) * npix;
    }
    HIP_CHECK(hipGetLastError());
    std::memset(h->stats, 0, sizeof(h->stats));
    if (h->profile) {
      hipEvent_t eAll1 = mark();
      HIP_CHECK(hipEventSynchronize(eAll1));
      double t[5] = {0, 0, 0, 0, 0};
      for (const Span& sp : spans) {
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, sp.a, sp.b));
        t[sp.kind] += ms;
      }
      float all = 0.f;
      HIP_CHECK(hipEventElapsedTime(&all, eAll0, eAll1));
      h->stats[2] = t[0]; h->stats[3] = t[2]; h->stats[4] = t[1]; h->stats[5] = all;
      h->stats[11] = t[3]; h->stats[12] = t[4];
    }
    h->stats[10] = (double)h->fc.maxRayDepth * ((sample_count + G - 1) / G);  // extension launches
    h->stats[13] = G;
    return YPT_OK;
  });
}

int ypt_synchronize(ypt_integrator* h) {
  if (!h) return fail(YPT_E_INVALID, "null handle");
  return guarded([&]() -> int {
    HIP_CHECK(hipSetDevice(h->device));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    return YPT_OK;
  });
}

int ypt_get_stats(ypt_integrator* h, double* out16) {
  if (!h || !out16) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    HIP_CHECK(hipSetDevice(h->device));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    if (h->countTraversal && h->work.n) {
      unsigned long long w[8] = {0};
      HIP_CHECK(hipMemcpy(w, h->work.p, sizeof(w), hipMemcpyDeviceToHost));
      if (h->type == YPT_INTEGRATOR_BDPT) {  // camera subpaths | light subpaths and splats
        unsigned long long v[16] = {0};
        HIP_CHECK(hipMemcpy(v, h->work.p, sizeof(v), hipMemcpyDeviceToHost));
        // megakernel [0..2]; wavefront extension [8, 9, 12] and connection [10, 11, 13] launches
        h->stats[6] = (double)(v[0] + v[8] + v[10]); h->stats[7] = (double)(v[1] + v[9] + v[11]);
        h->stats[0] = (double)(v[2] + v[12] + v[13]);
        h->stats[8] = (double)w[3]; h->stats[9] = (double)w[4]; h->stats[1] = (double)w[5];
      } else {
        h->stats[6] = (double)w[0]; h->stats[7] = (double)w[1]; h->stats[8] = (double)w[2]; h->stats[9] = (double)w[3];
        h->stats[0] = (double)w[4]; h->stats[1] = (double)w[5];
      }
    }
    std::memcpy(out16, h->stats, sizeof(h->stats));
    return YPT_OK;
  });
}

int ypt_get_film(ypt_integrator* h, float* xyzw, uint32_t* counts) {
  if (!h || !xyzw) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasFrame) return fail(YPT_E_NOSCENE, "no frame");
    HIP_CHECK(hipSetDevice(h->device));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    size_t npix = (size_t)h->width * h->height;
    const float4* m = (h->filmCur ^ 1) ? h->merged1.p : h->merged0.p;  // last written
    const uint32_t* c = (h->filmCur ^ 1) ? h->counts1.p : h->counts0.p;
    HIP_CHECK(hipMemcpy(xyzw, m, npix * sizeof(float4), hipMemcpyDeviceToHost));
    if (counts) HIP_CHECK(hipMemcpy(counts, c, npix * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return YPT_OK;
  });
}

void ypt_post_params_default(ypt_post_params* p) {
  if (!p) return;
  p->denoise = 0;
  p->prefilter_features = 1;
  p->depth_sigma_scale = 0.02f;
  p->normal_sigma = 0.06f;
  p->material_sigma = 0.05f;
  p->color_sigma = 3.f;
  p->feature_variance_scale = 1.f;
  p->tonemap = 1;
  p->exposure = 1.f;
}

int ypt_final_color(ypt_integrator* h, const ypt_post_params* prm, uint8_t* rgba8, float* linear_rgba) {
  if (!h || !prm || (!rgba8 && !linear_rgba)) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasFrame) return fail(YPT_E_NOSCENE, "no frame");
    HIP_CHECK(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    uint32_t W = h->width, H = h->height;
    size_t npix = (size_t)W * H;
    h->postF0.ensure(npix); h->postF1.ensure(npix);
    h->postA.ensure(npix); h->postB.ensure(npix);
    h->postVarA.ensure(npix); h->postVarB.ensure(npix);
    h->postOut.ensure(npix);
    PostParams pp;
    pp.uvToView = h->fc.uvToView;
    pp.viewToWorld = h->fc.viewToWorld;
    pp.width = W; pp.height = H;
    pp.depthSigmaScale = prm->depth_sigma_scale;
    pp.normalSigma = prm->normal_sigma;
    pp.materialSigma = prm->material_sigma;
    pp.colorSigma = prm->color_sigma;
    pp.featureVarianceScale = prm->feature_variance_scale;
    pp.prefilter = prm->prefilter_features ? 1u : 0u;
    const float4* film = (h->filmCur ^ 1) ? h->merged1.p : h->merged0.p;  // last written
    dim3 tiles((W + 7) / 8, (H + 7) / 8);
    if (prm->denoise) {
      if (h->featAccum.n >= 4 * npix && h->filmSamples > 0) {
        hipLaunchKernelGGL(k_prefilter_features, tiles, dim3(64), 0, s, pp, (const float4*)h->featAccum.p, h->postF0.p, h->postF1.p);
      } else {  // no accumulated features (first_hit_features off): every pixel reads as a miss and passes through
        HIP_CHECK(hipMemsetAsync(h->postF0.p, 0, npix * sizeof(uint4), s));
        HIP_CHECK(hipMemsetAsync(h->postF1.p, 0, npix * sizeof(uint4), s));
      }
      const float4* src = film;
      const float* srcVar = h->variance.p;
      float4* bufs[2] = {h->postA.p, h->postB.p};
      float* vars[2] = {h->postVarA.p, h->postVarB.p};
      const uint32_t passes = 5;  // DENOISE_PASSES_COUNT (CombineSamplesSubStage.cpp:11)
      for (uint32_t i = 0; i < passes; ++i) {
        hipLaunchKernelGGL(k_atrous, tiles, dim3(64), 0, s, pp, i, 1, i + 1 == passes ? 1 : 0, src, srcVar,
                           (const uint4*)h->postF0.p, (const uint4*)h->postF1.p, bufs[i & 1], vars[i & 1]);
        src = bufs[i & 1];
        srcVar = vars[i & 1];
      }
      if (src != h->postA.p) HIP_CHECK(hipMemcpyAsync(h->postA.p, src, npix * sizeof(float4), hipMemcpyDeviceToDevice, s));
    } else {
      hipLaunchKernelGGL(k_atrous, tiles, dim3(64), 0, s, pp, 0u, 0, 1, film, (const float*)h->variance.p,
                         (const uint4*)h->postF0.p, (const uint4*)h->postF1.p, h->postA.p, h->postVarA.p);
    }
    hipLaunchKernelGGL(k_tonemap, dim3(blocks_for(npix, 256)), dim3(256), 0, s, (uint32_t)npix, (const float4*)h->postA.p,
                       prm->exposure, prm->tonemap ? 1 : 0, h->postOut.p);
    HIP_CHECK(hipGetLastError());
    if (linear_rgba) HIP_CHECK(hipMemcpyAsync(linear_rgba, h->postA.p, npix * sizeof(float4), hipMemcpyDeviceToHost, s));
    if (rgba8) HIP_CHECK(hipMemcpyAsync(rgba8, h->postOut.p, npix * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return YPT_OK;
  });
}

int ypt_get_film_variance(ypt_integrator* h, float* out) {
  if (!h || !out) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasFrame) return fail(YPT_E_NOSCENE, "no frame");
    HIP_CHECK(hipSetDevice(h->device));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    HIP_CHECK(hipMemcpy(out, h->variance.p, (size_t)h->width * h->height * sizeof(float), hipMemcpyDeviceToHost));
    return YPT_OK;
  });
}

int ypt_get_feature_moments(ypt_integrator* h, float* out16) {
  if (!h || !out16) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasFrame) return fail(YPT_E_NOSCENE, "no frame");
    size_t npix = (size_t)h->width * h->height;
    if (h->featAccum.n < 4 * npix) return fail(YPT_E_INVALID, "no feature moments: render with first_hit_features on");
    HIP_CHECK(hipSetDevice(h->device));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    HIP_CHECK(hipMemcpy(out16, h->featAccum.p, 4 * npix * sizeof(float4), hipMemcpyDeviceToHost));
    return YPT_OK;
  });
}

int ypt_get_last_samples(ypt_integrator* h, float* xyzl) {
  if (!h || !xyzl) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasFrame) return fail(YPT_E_NOSCENE, "no frame");
    HIP_CHECK(hipSetDevice(h->device));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    HIP_CHECK(hipMemcpy(xyzl, h->lastSample.p, (size_t)h->width * h->height * sizeof(float4), hipMemcpyDeviceToHost));
    return YPT_OK;
  });
}

int ypt_get_first_hit_features(ypt_integrator* h, float* out) {
  if (!h || !out) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasFrame) return fail(YPT_E_NOSCENE, "no frame");
    HIP_CHECK(hipSetDevice(h->device));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    HIP_CHECK(hipMemcpy2D(out, 12 * sizeof(float), h->features.p + (size_t)h->lastPassBase * FEATURE_STRIDE,
                          FEATURE_STRIDE * sizeof(float), 12 * sizeof(float), (size_t)h->width * h->height,
                          hipMemcpyDeviceToHost));
    return YPT_OK;
  });
}

int ypt_trace_closest(ypt_integrator* h, const float* rays, uint32_t n, float* t, uint32_t* tri, float* b12) {
  if (!h || !rays || !t || !tri) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasScene) return fail(YPT_E_NOSCENE, "no scene");
    if (n == 0) return YPT_OK;
    HIP_CHECK(hipSetDevice(h->device));
    DevBuf<float> dr, dt, db;
    DevBuf<uint32_t> dtri, dovf;
    hipStream_t s = h->stream;
    dr.upload(rays, (size_t)n * 8, s);
    dt.alloc(n); db.alloc(2 * (size_t)n); dtri.alloc(n);
    uint32_t bound = h->bvhWidth == 8 ? h->res->stackBound8 : h->res->stackBound2;
    size_t depth = bound > (uint32_t)h->ldsDepth ? bound - (uint32_t)h->ldsDepth : 1;
    DevScene S = h->dev_scene();
    size_t lds = (size_t)h->ldsDepth * h->blockSize * 4;
    uint32_t ld = (uint32_t)h->ldsDepth;
    if (h->waveBelow > 0 && n < (uint32_t)h->waveBelow) {  // wave-per-ray: one 64-lane workgroup per ray
      uint32_t g = std::min<uint32_t>(n, (uint32_t)h->cus * 16u);
      dovf.alloc(depth * g * 64);
      size_t wl = wave_lds(h);
      if (h->bvhWidth == 8)
        hipLaunchKernelGGL(k_trace_closest_w<8>, dim3(g), dim3(64), wl, s, S, dr.p, n, dt.p, dtri.p, db.p, dovf.p, ld);
      else
        hipLaunchKernelGGL(k_trace_closest_w<2>, dim3(g), dim3(64), wl, s, S, dr.p, n, dt.p, dtri.p, db.p, dovf.p, ld);
    } else {
      dovf.alloc(depth * n);
      uint32_t g = blocks_for(n, h->blockSize);
      if (h->bvhWidth == 8)
        hipLaunchKernelGGL(k_trace_closest<8>, dim3(g), dim3(h->blockSize), lds, s, S, dr.p, n, dt.p, dtri.p, db.p, dovf.p, ld);
      else
        hipLaunchKernelGGL(k_trace_closest<2>, dim3(g), dim3(h->blockSize), lds, s, S, dr.p, n, dt.p, dtri.p, db.p, dovf.p, ld);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipMemcpy(t, dt.p, n * sizeof(float), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(tri, dtri.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (b12) HIP_CHECK(hipMemcpy(b12, db.p, 2 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    return YPT_OK;
  });
}

int ypt_trace_closest_noalloc(ypt_integrator* h, const float* rays, uint32_t n, float* t, uint32_t* tri, float* b12,
                              void* stream) {
  if (!h || !rays || !t || !tri || !b12) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasScene) return fail(YPT_E_NOSCENE, "no scene");
    if (n == 0) return YPT_OK;
    if (h->device < 0 || h->device >= NOALLOC_MAX_DEVICES || !g_scratchAddr[h->device])
      return fail(YPT_E_DEVICE, "no scratch address for this device: " + g_scratchError);
    uint32_t bound = h->bvhWidth == 8 ? h->res->stackBound8 : h->res->stackBound2;
    size_t depth = bound > (uint32_t)h->ldsDepth ? bound - (uint32_t)h->ldsDepth : 1;
    // rays per launch: one overflow column of `depth` words each, within the scratch array
    size_t chunk = NOALLOC_SCRATCH_WORDS / depth;
    if (chunk < 64)
      return fail(YPT_E_MEMORY, "the BVH's traversal stack (" + std::to_string(depth) + " words past the LDS) exceeds the " +
                                    std::to_string(NOALLOC_SCRATCH_WORDS * 4 >> 20) + " MiB scratch of ypt_trace_closest_noalloc");
    chunk = std::min<size_t>(chunk & ~(size_t)63, (size_t)1 << 24);
    HIP_CHECK(hipSetDevice(h->device));
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    uint32_t* ovf = (uint32_t*)g_scratchAddr[h->device];
    DevScene S = h->dev_scene();
    size_t lds = (size_t)h->ldsDepth * h->blockSize * 4;
    for (size_t first = 0; first < n; first += chunk) {
      uint32_t m = (uint32_t)std::min<size_t>(chunk, n - first);
      const float* r = rays + first * 8;
      if (h->bvhWidth == 8)
        hipLaunchKernelGGL(k_trace_closest<8>, dim3(blocks_for(m, h->blockSize)), dim3(h->blockSize), lds, s, S, r, m, t + first,
                           tri + first, b12 + 2 * first, ovf, (uint32_t)h->ldsDepth);
      else
        hipLaunchKernelGGL(k_trace_closest<2>, dim3(blocks_for(m, h->blockSize)), dim3(h->blockSize), lds, s, S, r, m, t + first,
                           tri + first, b12 + 2 * first, ovf, (uint32_t)h->ldsDepth);
      HIP_CHECK(hipGetLastError());
    }
    return YPT_OK;
  });
}

int ypt_sample_dims(ypt_integrator* h, uint32_t sample_index, const uint32_t* seeds, const uint32_t* dims, uint32_t n, float* out) {
  if (!h || !seeds || !dims || !out) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (n == 0) return YPT_OK;
    HIP_CHECK(hipSetDevice(h->device));
    uint32_t pt[YPT_SOBOL_POINT_DIMS];
    ypt_sobol_point(sample_index, pt);
    DevBuf<uint32_t> dp, ds, dd;
    DevBuf<float> dout;
    hipStream_t s = h->stream;
    dp.upload(pt, YPT_SOBOL_POINT_DIMS, s);
    ds.upload(seeds, n, s);
    dd.upload(dims, n, s);
    dout.alloc(n);
    hipLaunchKernelGGL(k_sample_dims, dim3(blocks_for(n, 256)), dim3(256), 0, s, dp.p, ds.p, dd.p, n, dout.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipMemcpy(out, dout.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return YPT_OK;
  });
}


// ---------------------------------------------------------------- resident scenes
struct ypt_scene {
  std::shared_ptr<ResidentScene> r;
};

int ypt_scene_create(const ypt_scene_desc* d, int device, ypt_scene** out) {
  if (!d || !out) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(YPT_E_INVALID, "device index out of range");
    HIP_CHECK(hipSetDevice(device));
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<ypt_scene> sc(new ypt_scene());
    try {
      sc->r = make_resident(*d, device, s);
    } catch (...) {
      (void)hipStreamDestroy(s);
      throw;
    }
    HIP_CHECK(hipStreamDestroy(s));
    *out = sc.release();
    return YPT_OK;
  });
}

int ypt_scene_destroy(ypt_scene* sc) {
  delete sc;  // device memory goes when the last handle using the scene lets go of it
  return YPT_OK;
}

int ypt_scene_info(const ypt_scene* sc, uint64_t* device_bytes, uint32_t* triangles) {
  if (!sc) return fail(YPT_E_INVALID, "null scene");
  if (device_bytes) *device_bytes = sc->r->device_bytes();
  if (triangles) *triangles = (uint32_t)sc->r->scene.triInst.size();
  return YPT_OK;
}

// New vertex data for the same triangles: re-prepares the scene, refits its BVH (no rebuild) and
// re-uploads it in place. The handle must be the scene's only user.
int ypt_refit_scene(ypt_integrator* h, const ypt_scene_desc* d) {
  if (!h || !d) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (!h->hasScene) return fail(YPT_E_NOSCENE, "refit needs a scene set before");
    if (h->res.use_count() > 1) return fail(YPT_E_INVALID, "the scene is shared with other handles or jobs; refit a scene of its own");
    ResidentScene& r = *h->res;
    if ((size_t)d->vertex_count * 3 != r.scene.pos.size() || (size_t)d->triangle_count != r.scene.triInst.size())
      return fail(YPT_E_INVALID, "refit: vertex or triangle count differs from the scene's");
    PreparedScene ps;
    prepare_scene(*d, host_tables().r2s, ps);
    if (ps.idx != r.scene.idx || ps.triInst != r.scene.triInst || ps.triAlpha != r.scene.triAlpha || ps.inst.size() != r.scene.inst.size())
      return fail(YPT_E_INVALID, "refit: triangles, instances or alpha-tested triangles differ from the scene's");
    HIP_CHECK(hipSetDevice(h->device));
    HIP_CHECK(hipStreamSynchronize(h->stream));  // nothing in flight reads the buffers being replaced
    r.scene = std::move(ps);
    bvh_refit(r.scene.pos.data(), r.scene.idx.data(), r.bvh);
    r.subsurface = false;
    for (const MaterialGPU& m : r.scene.mats) r.subsurface = r.subsurface || m.subsurfaceAmount > 0.f;
    const PreparedScene& q = r.scene;
    hipStream_t s = h->stream;
    r.dPos.upload(q.pos.data(), q.pos.size(), s);
    r.dNrm.upload(q.nrm.data(), q.nrm.size(), s);
    if (!q.tan.empty()) r.dTan.upload(q.tan.data(), q.tan.size(), s);
    r.dInst.upload(q.inst.data(), q.inst.size(), s);
    r.dMats.upload(q.mats.data(), q.mats.size(), s);
    if (!q.lights.empty()) r.dLights.upload(q.lights.data(), q.lights.size(), s);
    r.dNodes2.upload(r.bvh.nodes2.data(), r.bvh.nodes2.size(), s);
    r.dNodes8.upload(r.bvh.nodes8.data(), r.bvh.nodes8.size(), s);
    r.dTris.upload(r.bvh.tris.data(), r.bvh.tris.size(), s);
    if (!q.uv.empty()) r.dUv.upload(q.uv.data(), q.uv.size(), s);
    if (!q.alphaTris.empty()) r.dAlpha.upload(q.alphaTris.data(), q.alphaTris.size(), s);
    if (!q.envTexels.empty()) r.dEnv.upload(q.envTexels.data(), q.envTexels.size(), s);
    r.refits++;
    HIP_CHECK(hipStreamSynchronize(s));
    if (h->hasFrame) {  // frame constants depend on the scene's bounds and lights
      prepare_frame(h->frameDesc, h->res->scene, (uint32_t)h->type, h->fc);
      clear_film(h);
    }
    return YPT_OK;
  });
}

// BVH builds and refits of the handle's scene since it was created
int ypt_scene_bvh_counts(ypt_integrator* h, uint32_t* builds, uint32_t* refits) {
  if (!h || !builds || !refits) return fail(YPT_E_INVALID, "null argument");
  if (!h->hasScene) return fail(YPT_E_NOSCENE, "no scene");
  *builds = 1;
  *refits = h->res->refits;
  return YPT_OK;
}

int ypt_set_resident_scene(ypt_integrator* h, ypt_scene* sc) {
  if (!h || !sc) return fail(YPT_E_INVALID, "null argument");
  return guarded([&]() -> int {
    if (sc->r->device != h->device) return fail(YPT_E_INVALID, "scene and integrator are on different devices");
    HIP_CHECK(hipSetDevice(h->device));
    attach_scene(h, sc->r);
    return YPT_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------- render server
// Continuous batching of progressive render jobs over resident scenes. A worker thread admits
// queued jobs in arrival order while the number of active jobs and their estimated device memory
// fit, and keeps every active job's next chunk of passes in flight on that job's own stream, so
// the kernels of different jobs overlap on the device. A finished job's film is copied out, its
// handle goes back to a pool (buffers only grow, so no allocation or free runs while others render)
// and the next queued job is admitted. A job's result depends only on its scene, frame and sample
// count: it is the same alone or in any mix.
namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum JobState { JOB_QUEUED = 0, JOB_ACTIVE = 1, JOB_DONE = 2, JOB_FAILED = 3, JOB_CANCELLED = 4, JOB_TIMED_OUT = 5 };

struct ServerJob {
  uint64_t id = 0;
  std::shared_ptr<ResidentScene> scene;
  int type = 0;
  ypt_frame_desc frame;
  uint32_t spp = 0, chunk = 1, passesInFlight = 1;
  uint32_t issued = 0;
  uint64_t mem = 0;
  std::vector<std::pair<std::string, int64_t>> options;  // launch options of its type, copied at admission
  ypt_integrator* h = nullptr;
  hipEvent_t ev = nullptr;
  bool inFlight = false;
  bool fused = false;  // its chunks run in the server's fused steps (not on its own stream)
  std::atomic<int> state{0};  // JOB_QUEUED, JOB_ACTIVE, JOB_DONE, JOB_FAILED, JOB_CANCELLED, JOB_TIMED_OUT
  std::atomic<bool> cancel{false};
  std::atomic<double> deadline{1e300};  // ypt_clock time after which the job is given up
  std::string err;
  double tSubmit = 0, tAdmit = 0, tFinish = 0;
  std::vector<float> film;
};

// passes per batch and the path-state estimate of a job (path slots, queues, stack overflow, film)
void plan_job(ServerJob& j, int cus) {
  uint64_t npix = (uint64_t)j.frame.width * j.frame.height;
  uint64_t bytes = npix * (16 * 2 + 16 * 2 + 4 * 2 + 4 + 16);
  if (j.type == YPT_INTEGRATOR_BDPT) {
    uint64_t npaths = ((j.frame.width + 7) / 8) * (uint64_t)((j.frame.height + 7) / 8);
    bytes += npaths * BDPT_MAX_LIGHT_VERTICES * (sizeof(LightNodePacked) + sizeof(LightNodeSurface)) + npaths * 15 * (4 * 4 + 16) +
             npix * (16 + 4) + npaths * 64 * 48 * 4;
    // camera subpaths as a wavefront: state per texel, and connection rays for every light vertex
    bytes += npix * (16 * 9 + 4 + 16 * 4 + 4) + npix * BDPT_MAX_LIGHT_VERTICES * (16 * 3 + 4 + 16);
    j.passesInFlight = 1;
    j.chunk = std::min<uint32_t>(j.spp, 2);
  } else {
    uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(16, (8ull << 20) / std::max<uint64_t>(npix, 1)));
    g = std::min(g, j.spp);
    uint64_t slots = npix * g;
    bytes += slots * (6 * 16 + 16 + 16 + 64 + 16 + 8 + 52 + 48) + (uint64_t)cus * 32 * 64 * 48 * 4;
    j.passesInFlight = g;
    j.chunk = g;
  }
  j.mem = bytes;
}

// One fused step's bounce loop on pool handle h (its stream, path-state pool, queues and launch
// options): camera rays of every slot, then per bounce the closest-hit, shading and (when a job
// traces them) occlusion launches over the queues that all jobs share.
void launch_fused(ypt_integrator* h, const FusedArgs& A, uint32_t depth, bool alpha, bool shadow) {
  hipStream_t s = h->stream;
  PathSoA P = h->path_soa();
  Queues Q = h->queues();
  Q.waveBelow = (uint32_t)h->waveBelow;
  uint32_t wg = wave_grid(h);
  size_t wl = wave_lds(h);
  uint32_t total = A.slotBase[A.njobs];
  uint32_t grid = work_grid(h, total), sgrid = shade_grid(h, total);
  size_t lds = (size_t)h->ldsDepth * h->blockSize * 4;
  uint32_t d = (uint32_t)h->ldsDepth, r = (uint32_t)h->refill, ls = (uint32_t)h->leafShare;
  uint32_t B = (uint32_t)h->blockSize;
  hipLaunchKernelGGL(k_start_m, dim3(blocks_for(total, 256)), dim3(256), 0, s, A, P, Q);
  int q = 0;
  for (uint32_t b = 0; b < depth; ++b) {
    if (h->bvhWidth == 8) {
      if (alpha) YPT_PERSISTENT((k_extend_pm<8, true>), grid, B, lds, s, A, P, Q, q, d, r, ls);
      else YPT_PERSISTENT((k_extend_pm<8, false>), grid, B, lds, s, A, P, Q, q, d, r, ls);
    } else {
      if (alpha) YPT_PERSISTENT((k_extend_pm<2, true>), grid, B, lds, s, A, P, Q, q, d, r, ls);
      else YPT_PERSISTENT((k_extend_pm<2, false>), grid, B, lds, s, A, P, Q, q, d, r, ls);
    }
    if (Q.waveBelow) {
      if (h->bvhWidth == 8) {
        if (alpha) YPT_PERSISTENT((k_extend_wm<8, true>), wg, 64, wl, s, A, P, Q, q, d);
        else YPT_PERSISTENT((k_extend_wm<8, false>), wg, 64, wl, s, A, P, Q, q, d);
      } else {
        if (alpha) YPT_PERSISTENT((k_extend_wm<2, true>), wg, 64, wl, s, A, P, Q, q, d);
        else YPT_PERSISTENT((k_extend_wm<2, false>), wg, 64, wl, s, A, P, Q, q, d);
      }
    }
    if (h->shadeRegWaves == 2) hipLaunchKernelGGL((k_shade_m<2>), dim3(sgrid), dim3(64), 0, s, A, P, Q, q);
    else if (h->shadeRegWaves == 8) hipLaunchKernelGGL((k_shade_m<8>), dim3(sgrid), dim3(64), 0, s, A, P, Q, q);
    else hipLaunchKernelGGL((k_shade_m<4>), dim3(sgrid), dim3(64), 0, s, A, P, Q, q);
    if (shadow) {
      if (h->bvhWidth == 8) {
        if (alpha) YPT_PERSISTENT((k_shadow_pm<8, true>), grid, B, lds, s, A, P, Q, d, r, ls);
        else YPT_PERSISTENT((k_shadow_pm<8, false>), grid, B, lds, s, A, P, Q, d, r, ls);
      } else {
        if (alpha) YPT_PERSISTENT((k_shadow_pm<2, true>), grid, B, lds, s, A, P, Q, d, r, ls);
        else YPT_PERSISTENT((k_shadow_pm<2, false>), grid, B, lds, s, A, P, Q, d, r, ls);
      }
      if (Q.waveBelow) {
        if (h->bvhWidth == 8) {
          if (alpha) YPT_PERSISTENT((k_shadow_wm<8, true>), wg, 64, wl, s, A, P, Q, d);
          else YPT_PERSISTENT((k_shadow_wm<8, false>), wg, 64, wl, s, A, P, Q, d);
        } else {
          if (alpha) YPT_PERSISTENT((k_shadow_wm<2, true>), wg, 64, wl, s, A, P, Q, d);
          else YPT_PERSISTENT((k_shadow_wm<2, false>), wg, 64, wl, s, A, P, Q, d);
        }
      }
    }
    q ^= 1;
  }
}

}  // namespace

struct ypt_server {
  int device = 0;
  int cus = 256;
  uint64_t budget = 0;
  uint32_t maxActive = 1;
  std::mutex m;
  std::condition_variable cvDone;
  std::deque<std::shared_ptr<ServerJob>> queued;
  std::vector<std::shared_ptr<ServerJob>> active;
  std::unordered_map<uint64_t, std::shared_ptr<ServerJob>> jobs;
  std::vector<ypt_integrator*> pool[3];
  std::vector<std::pair<std::string, int64_t>> options[3];  // launch options per integrator type
  uint64_t nextId = 1, memUsed = 0, completed = 0, peakActive = 0;
  bool stop = false;
  std::thread worker;

  // Fused steps (kernels/multijob.h): the unidirectional jobs whose scenes have no subsurface
  // materials run their chunks together, one launch sequence per step on a pool handle's stream,
  // at most FUSED_MAX_JOBS jobs and fuseMaxSlots path slots per step (a larger single chunk runs
  // alone). Two pools: a step can start while the other is still on the device, so a job that
  // arrives during a long step does not wait for it. BDPT jobs keep their own streams beside them.
  bool fuse = true;
  uint64_t fuseMaxSlots = 16ull << 20;
  struct Fused {
    ypt_integrator* x = nullptr;  // pool: path state, queues, stream and launch options of a step
    DevBuf<FusedJob> jobs;
    DevBuf<uint32_t> points;
    std::vector<FusedJob> hJobs;  // host copies of the job table and Sobol points, kept until the step ends
    std::vector<uint32_t> hPoints;
    std::vector<std::pair<std::string, int64_t>> applied;  // launch options set on the pool handle
    hipEvent_t done = nullptr;
    bool inFlight = false;
    std::vector<std::shared_ptr<ServerJob>> members;
  } fx[2];
  std::atomic<uint64_t> fusedSteps{0}, fusedJobChunks{0};  // read by ypt_server_fused_stats from other threads

  ypt_integrator* take_handle(int type) {
    if (!pool[type].empty()) {
      ypt_integrator* h = pool[type].back();
      pool[type].pop_back();
      return h;
    }
    ypt_integrator* h = nullptr;
    if (ypt_create(type, device, &h) != YPT_OK) throw std::runtime_error(std::string("server: ") + ypt_last_error());
    return h;
  }

  void start_job(ServerJob& j) {
    j.h = take_handle(j.type);
    ypt_integrator* h = j.h;
    if (!j.ev) HIP_CHECK(hipEventCreateWithFlags(&j.ev, hipEventDisableTiming));
    if (h->res != j.scene) attach_scene(h, j.scene);
    for (const auto& o : j.options)
      if (ypt_set_option(h, o.first.c_str(), o.second) != YPT_OK) throw std::runtime_error(ypt_last_error());
    if (ypt_set_frame(h, &j.frame) != YPT_OK) throw std::runtime_error(std::string("server: ") + ypt_last_error());
    if (ypt_set_option(h, "passes_in_flight", j.passesInFlight) != YPT_OK) throw std::runtime_error(ypt_last_error());
    clear_film(h);
    // a fused job's merges run on the pool's stream: its film is cleared before the first step
    if (j.fused) HIP_CHECK(hipStreamSynchronize(h->stream));
  }

  // Issues one fused step over the jobs of `ready` (admission order) that fit the step, each for
  // its next chunk of passes; film merges per job in pass order after the bounce loop.
  void issue_fused(Fused& fx, const std::vector<std::shared_ptr<ServerJob>>& ready, uint64_t maxSlots,
                   const std::vector<std::pair<std::string, int64_t>>& opts) {
    if (!fx.x) {
      if (ypt_create(YPT_INTEGRATOR_NEE, device, &fx.x) != YPT_OK) throw std::runtime_error(std::string("server: ") + ypt_last_error());
      HIP_CHECK(hipEventCreateWithFlags(&fx.done, hipEventDisableTiming));
    }
    ypt_integrator* X = fx.x;
    if (opts != fx.applied) {  // applied once per change: some options drop buffers sized by the old value
      for (const auto& o : opts)
        if (ypt_set_option(X, o.first.c_str(), o.second) != YPT_OK) throw std::runtime_error(ypt_last_error());
      fx.applied = opts;
    }
    std::vector<std::shared_ptr<ServerJob>> mem;
    std::vector<uint32_t> passes;
    uint64_t total = 0;
    for (const auto& j : ready) {
      if (mem.size() == FUSED_MAX_JOBS) break;
      uint32_t g = std::min(j->chunk, j->spp - j->issued);
      uint64_t n = (uint64_t)j->frame.width * j->frame.height * g;
      if (!mem.empty() && total + n > maxSlots) continue;
      mem.push_back(j);
      passes.push_back(g);
      total += n;
    }
    FusedArgs A;
    std::memset(&A, 0, sizeof(A));
    A.njobs = (uint32_t)mem.size();
    fx.hJobs.assign(mem.size(), FusedJob());
    fx.hPoints.clear();
    std::vector<size_t> ptOff(mem.size());
    uint32_t depth = 0, bound = 0;
    bool alpha = false, shadow = false;
    uint64_t slot = 0;
    for (size_t k = 0; k < mem.size(); ++k) {
      ServerJob& j = *mem[k];
      ypt_integrator* h = j.h;
      uint32_t npix = h->width * h->height;
      A.slotBase[k] = (uint32_t)slot;
      slot += (uint64_t)npix * passes[k];
      ptOff[k] = fx.hPoints.size();
      for (uint32_t p = 0; p < passes[k]; ++p)
        for (uint32_t d = 0; d < YPT_SOBOL_POINT_DIMS; ++d) fx.hPoints.push_back(sobol_uint(YPT_SOBOL_MATRICES, j.issued + p, d));
      FusedJob& J = fx.hJobs[k];
      std::memset(&J, 0, sizeof(J));
      J.S = h->dev_scene();
      J.fc = h->fc;
      J.npix = npix;
      J.nee = h->type == YPT_INTEGRATOR_NEE ? 1u : 0u;
      depth = std::max(depth, h->fc.maxRayDepth);
      bound = std::max(bound, X->bvhWidth == 8 ? h->res->stackBound8 : h->res->stackBound2);
      alpha = alpha || J.S.tris.alpha.tris != nullptr;
      shadow = shadow || traces_shadow_rays(h);
    }
    A.slotBase[mem.size()] = (uint32_t)slot;
    // the pool's buffers (grow-only): path state and queues for the step's slots, stack overflow for
    // the deepest BVH among its scenes
    X->counters.ensure(16);
    X->work.ensure(16);
    ensure_slots(X, slot);
    size_t stride = X->slots + 1024;
    if (X->persistent) stride = std::min(stride, (size_t)X->cus * X->persistentWaves * 64);
    size_t ovfDepth = bound > (uint32_t)X->ldsDepth ? bound - (uint32_t)X->ldsDepth : 1;
    X->ovf.ensure(ovfDepth * stride);
    X->ovfStride = (uint32_t)stride;
    fx.points.ensure(fx.hPoints.size());
    fx.jobs.ensure(mem.size());
    for (size_t k = 0; k < mem.size(); ++k) fx.hJobs[k].points = fx.points.p + ptOff[k];
    hipStream_t s = X->stream;
    HIP_CHECK(hipMemcpyAsync(fx.points.p, fx.hPoints.data(), fx.hPoints.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(fx.jobs.p, fx.hJobs.data(), fx.hJobs.size() * sizeof(FusedJob), hipMemcpyHostToDevice, s));
    A.jobs = fx.jobs.p;
    launch_fused(X, A, depth, alpha, shadow);
    PathSoA P = X->path_soa();
    for (size_t k = 0; k < mem.size(); ++k) {
      ServerJob& j = *mem[k];
      ypt_integrator* h = j.h;
      FilmBuffers F = h->film();
      uint32_t npix = h->width * h->height;
      for (uint32_t p = 0; p < passes[k]; ++p) {
        MergeParams mp;
        std::memset(&mp, 0, sizeof(mp));
        mp.width = h->width; mp.height = h->height; mp.slotBase = A.slotBase[k] + p * npix;
        mp.sampleCount = h->filmSamples; mp.accumTargets = 2;
        mp.fireflyClampScale = h->frameDesc.firefly_clamp_scale;
        mp.fireflyClampRelaxPower = h->frameDesc.firefly_clamp_relax_power;
        mp.killNaN = h->frameDesc.kill_nan_samples;
        mp.cur = h->filmCur;
        mp.pass = p;
        hipLaunchKernelGGL(k_merge, dim3(blocks_for(npix, 256)), dim3(256), 0, s, fx.hJobs[k].S, P, F, mp, (const float4*)nullptr,
                           (const uint32_t*)nullptr);
        h->filmSamples++;
        h->filmCur ^= 1;
      }
      j.issued += passes[k];
      j.inFlight = true;
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(fx.done, s));
    fx.inFlight = true;
    fx.members = mem;
    fusedSteps++;
    fusedJobChunks += mem.size();
  }

  void finish_job(ServerJob& j) {
    size_t npix = (size_t)j.frame.width * j.frame.height;
    j.film.resize(npix * 4);
    if (ypt_get_film(j.h, j.film.data(), nullptr) != YPT_OK) throw std::runtime_error(ypt_last_error());
  }

  void release(ServerJob& j) {
    if (j.h) pool[j.type].push_back(j.h);
    j.h = nullptr;
    if (j.ev) (void)hipEventDestroy(j.ev);
    j.ev = nullptr;
  }

  void run() {
    (void)hipSetDevice(device);
    while (true) {
      std::vector<std::shared_ptr<ServerJob>> fresh, act;
      bool dropped = false, useFuse = true;
      uint64_t maxSlots = 0;
      std::vector<std::pair<std::string, int64_t>> fusedOpts;
      {
        std::lock_guard<std::mutex> lk(m);
        if (stop) break;
        double t = now_s();
        for (size_t i = 0; i < queued.size();) {  // cancelled or expired before admission
          auto& q = queued[i];
          if (q->cancel || t > q->deadline) {
            q->state = q->cancel ? JOB_CANCELLED : JOB_TIMED_OUT;
            q->tFinish = t;
            queued.erase(queued.begin() + i);
            dropped = true;
          } else {
            ++i;
          }
        }
        if (dropped) cvDone.notify_all();
        while (!queued.empty() && active.size() < maxActive) {
          auto j = queued.front();
          if (!active.empty() && memUsed + j->mem > budget) break;  // wait for memory to free up
          queued.pop_front();
          memUsed += j->mem;
          j->options = options[j->type];
          j->state = JOB_ACTIVE;
          j->tAdmit = now_s();
          active.push_back(j);
          fresh.push_back(j);
        }
        peakActive = std::max<uint64_t>(peakActive, active.size());
        act = active;
        useFuse = fuse;
        maxSlots = fuseMaxSlots;
        fusedOpts = options[YPT_INTEGRATOR_NEE];
      }
      std::vector<std::shared_ptr<ServerJob>> ended;
      for (auto& j : fresh) {
        try {
          j->fused = useFuse && j->type != YPT_INTEGRATOR_BDPT && !j->scene->subsurface;
          start_job(*j);
        } catch (const std::exception& e) {
          j->err = e.what();
          j->state = JOB_FAILED;
          ended.push_back(j);
        }
      }
      bool progress = !fresh.empty() || dropped;
      // a fused step that has ended hands its jobs back
      for (Fused& fx : this->fx) {
        if (!fx.inFlight) continue;
        hipError_t q = hipEventQuery(fx.done);
        if (q != hipErrorNotReady) {
          fx.inFlight = false;
          for (auto& m : fx.members) {
            m->inFlight = false;
            if (q != hipSuccess && m->state == JOB_ACTIVE) {
              m->err = std::string("fused step: ") + hipGetErrorString(q);
              m->state = JOB_FAILED;
              ended.push_back(m);
            }
          }
          fx.members.clear();
          progress = true;
        }
      }
      std::vector<std::shared_ptr<ServerJob>> ready;  // fused jobs waiting for the next step
      for (auto& j : act) {
        if (j->state != JOB_ACTIVE) continue;
        try {
          if (j->inFlight) {
            if (j->fused) continue;  // released when its step ends
            hipError_t q = hipEventQuery(j->ev);
            if (q == hipErrorNotReady) continue;
            HIP_CHECK(q);
            j->inFlight = false;
          }
          // a cancelled or expired job stops between chunks: nothing of it is in flight any more, so its
          // handle goes back to the pool untouched by the jobs around it
          if (j->cancel || now_s() > j->deadline) {
            j->state = j->cancel ? JOB_CANCELLED : JOB_TIMED_OUT;
            j->tFinish = now_s();
            ended.push_back(j);
            progress = true;
            continue;
          }
          if (j->issued < j->spp && j->fused) {
            ready.push_back(j);
            continue;
          }
          if (j->issued < j->spp) {
            uint32_t n = std::min(j->chunk, j->spp - j->issued);
            if (ypt_render(j->h, j->issued, n, nullptr) != YPT_OK) throw std::runtime_error(ypt_last_error());
            HIP_CHECK(hipEventRecord(j->ev, j->h->stream));
            j->issued += n;
            j->inFlight = true;
          } else {
            finish_job(*j);
            j->tFinish = now_s();
            j->state = JOB_DONE;
            ended.push_back(j);
          }
          progress = true;
        } catch (const std::exception& e) {
          j->err = e.what();
          j->state = JOB_FAILED;
          ended.push_back(j);
        }
      }
      Fused* freePool = !fx[0].inFlight ? &fx[0] : !fx[1].inFlight ? &fx[1] : nullptr;
      if (!ready.empty() && freePool) {
        Fused& fx = *freePool;
        try {
          issue_fused(fx, ready, maxSlots, fusedOpts);
        } catch (const std::exception& e) {
          // the step's stream is drained, and every job it would have carried fails
          if (fx.x) (void)hipStreamSynchronize(fx.x->stream);
          fx.inFlight = false;
          fx.members.clear();
          for (auto& j : ready) {
            j->inFlight = false;
            j->err = e.what();
            j->state = JOB_FAILED;
            ended.push_back(j);
          }
        }
        progress = true;
      }
      if (!ended.empty()) {
        std::lock_guard<std::mutex> lk(m);
        for (auto& j : ended) {
          release(*j);
          memUsed -= j->mem;
          if (j->state == JOB_DONE) completed++;
          for (size_t i = 0; i < active.size(); ++i)
            if (active[i] == j) { active.erase(active.begin() + i); break; }
        }
        cvDone.notify_all();
      }
      if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
  }
};

extern "C" {

int ypt_server_create(int device, uint64_t hbm_budget_bytes, uint32_t max_active_jobs, ypt_server** out) {
  if (!out || max_active_jobs == 0) return fail(YPT_E_INVALID, "null output or zero active jobs");
  return guarded([&]() -> int {
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(YPT_E_INVALID, "device index out of range");
    std::unique_ptr<ypt_server> s(new ypt_server());
    s->device = device;
    HIP_CHECK(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, device));
    s->budget = hbm_budget_bytes;
    s->maxActive = max_active_jobs;
    ypt_server* raw = s.get();
    s->worker = std::thread([raw]() { raw->run(); });
    *out = s.release();
    return YPT_OK;
  });
}

int ypt_server_destroy(ypt_server* s) {
  if (!s) return YPT_OK;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
  }
  if (s->worker.joinable()) s->worker.join();
  for (auto& fx : s->fx) {
    if (!fx.x) continue;
    (void)hipStreamSynchronize(fx.x->stream);
    if (fx.done) (void)hipEventDestroy(fx.done);
    ypt_destroy(fx.x);
    fx.x = nullptr;
  }
  for (auto& kv : s->jobs) {
    ServerJob& j = *kv.second;
    if (j.h) {
      (void)hipStreamSynchronize(j.h->stream);
      s->release(j);
    }
  }
  for (auto& p : s->pool)
    for (ypt_integrator* h : p) ypt_destroy(h);
  delete s;
  return YPT_OK;
}

int ypt_server_submit(ypt_server* s, ypt_scene* sc, int integrator_type, const ypt_frame_desc* frame, uint32_t spp,
                      uint64_t* job_id) {
  if (!s || !sc || !frame || !job_id || spp == 0) return fail(YPT_E_INVALID, "null argument or zero samples");
  if (integrator_type < YPT_INTEGRATOR_BPT || integrator_type > YPT_INTEGRATOR_BDPT) return fail(YPT_E_INVALID, "unknown integrator type");
  if (sc->r->device != s->device) return fail(YPT_E_INVALID, "scene is on another device");
  if (frame->width == 0 || frame->height == 0 || frame->width > 65535 || frame->height > 65535)
    return fail(YPT_E_INVALID, "frame size out of range");
  auto j = std::make_shared<ServerJob>();
  j->scene = sc->r;
  j->type = integrator_type;
  j->frame = *frame;
  j->spp = spp;
  plan_job(*j, s->cus);
  if (j->mem > s->budget) return fail(YPT_E_MEMORY, "job needs more device memory than the server's budget");
  std::lock_guard<std::mutex> lk(s->m);
  j->id = s->nextId++;
  j->tSubmit = now_s();
  s->jobs[j->id] = j;
  s->queued.push_back(j);
  *job_id = j->id;
  return YPT_OK;
}

// launch option for the handles of one integrator type (< 0: all types), applied to jobs admitted later
int ypt_server_set_option(ypt_server* s, int integrator_type, const char* name, int64_t value) {
  if (!s || !name) return fail(YPT_E_INVALID, "null argument");
  if (integrator_type > YPT_INTEGRATOR_BDPT) return fail(YPT_E_INVALID, "unknown integrator type");
  std::string n(name);
  if (n == "fuse_jobs" || n == "fuse_max_slots") {  // server settings, for steps issued from now on
    if (n == "fuse_max_slots" && (value < 1 || value > (int64_t)1 << 31))
      return fail(YPT_E_INVALID, "fuse_max_slots must be in [1, 2^31]");
    std::lock_guard<std::mutex> lk(s->m);
    if (n == "fuse_jobs") s->fuse = value != 0;
    else s->fuseMaxSlots = (uint64_t)value;
    return YPT_OK;
  }
  {
    // validated now on a handle that holds no device state, so a bad value fails here and not in a job
    std::unique_ptr<ypt_integrator> probe(new ypt_integrator());
    int rc = ypt_set_option(probe.get(), name, value);
    if (rc != YPT_OK) return rc;
  }
  std::lock_guard<std::mutex> lk(s->m);
  for (int t = 0; t < 3; ++t) {
    if (integrator_type >= 0 && t != integrator_type) continue;
    bool found = false;
    for (auto& o : s->options[t])
      if (o.first == name) { o.second = value; found = true; }
    if (!found) s->options[t].emplace_back(name, value);
  }
  return YPT_OK;
}

// waits up to timeout_s (< 0: forever); YPT_OK when the job has finished, 1 when it is still running
int ypt_server_wait(ypt_server* s, uint64_t job_id, double timeout_s) {
  if (!s) return fail(YPT_E_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->m);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(YPT_E_INVALID, "unknown job");
  auto j = it->second;
  auto done = [&]() { return j->state >= JOB_DONE; };
  if (timeout_s < 0) s->cvDone.wait(lk, done);
  else if (!s->cvDone.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return 1;
  if (j->state == JOB_FAILED) return fail(YPT_E_DEVICE, "job failed: " + j->err);
  if (j->state == JOB_CANCELLED) return fail(YPT_E_CANCELLED, "job cancelled");
  if (j->state == JOB_TIMED_OUT) return fail(YPT_E_TIMEOUT, "job passed its deadline");
  return YPT_OK;
}

// film of a finished job (W*H float4 X, Y, Z, 1) and its times: submit, admit, finish (seconds,
// steady clock); the job is forgotten afterwards
int ypt_server_take(ypt_server* s, uint64_t job_id, float* xyzw, double* times3) {
  if (!s) return fail(YPT_E_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(YPT_E_INVALID, "unknown job");
  ServerJob& j = *it->second;
  if (j.state < JOB_DONE) return fail(YPT_E_INVALID, "job has not finished");
  if (j.state != JOB_DONE) {
    int st = j.state;
    std::string e = j.err;
    s->jobs.erase(it);
    if (st == JOB_CANCELLED) return fail(YPT_E_CANCELLED, "job cancelled");
    if (st == JOB_TIMED_OUT) return fail(YPT_E_TIMEOUT, "job passed its deadline");
    return fail(YPT_E_DEVICE, "job failed: " + e);
  }
  if (xyzw) std::memcpy(xyzw, j.film.data(), j.film.size() * sizeof(float));
  if (times3) { times3[0] = j.tSubmit; times3[1] = j.tAdmit; times3[2] = j.tFinish; }
  s->jobs.erase(it);
  return YPT_OK;
}

// cancels a queued or running job: it stops before its next chunk of passes
int ypt_server_cancel(ypt_server* s, uint64_t job_id) {
  if (!s) return fail(YPT_E_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(YPT_E_INVALID, "unknown job");
  if (it->second->state >= JOB_DONE) return fail(YPT_E_INVALID, "job has already ended");
  it->second->cancel = true;
  return YPT_OK;
}

// the job is given up (YPT_E_TIMEOUT) if it has not finished `seconds` after its submission
int ypt_server_set_deadline(ypt_server* s, uint64_t job_id, double seconds) {
  if (!s || !(seconds >= 0.0)) return fail(YPT_E_INVALID, "null server or negative deadline");
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(YPT_E_INVALID, "unknown job");
  it->second->deadline = it->second->tSubmit + seconds;
  return YPT_OK;
}

// [0] queued, [1] active, [2] completed, [3] device bytes reserved by active jobs, [4] peak active,
// [5] budget
int ypt_server_stats(ypt_server* s, double* out6) {
  if (!s || !out6) return fail(YPT_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->m);
  out6[0] = (double)s->queued.size();
  out6[1] = (double)s->active.size();
  out6[2] = (double)s->completed;
  out6[3] = (double)s->memUsed;
  out6[4] = (double)s->peakActive;
  out6[5] = (double)s->budget;
  return YPT_OK;
}

int ypt_server_fused_stats(ypt_server* s, uint64_t* steps, uint64_t* job_chunks) {
  if (!s || !steps || !job_chunks) return fail(YPT_E_INVALID, "null argument");
  *steps = s->fusedSteps.load();
  *job_chunks = s->fusedJobChunks.load();
  return YPT_OK;
}

double ypt_clock(void) { return now_s(); }

}  // extern "C"
