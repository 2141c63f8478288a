// The library's host-side code and the CPU checker under the compiler's sanitizers.
//
// Built by tests/test_sanitizers.py twice with g++ (AddressSanitizer + UndefinedBehaviorSanitizer,
// then ThreadSanitizer) and run on a scene the test serialises (scene description arrays and frame
// constants, no textures or environment map). It checks:
//   * the binned-SAH builder (runtime/bvh.h): in both node layouts every triangle record is reached
//     by exactly one leaf, the records are a permutation of the triangles, and every box holds the
//     triangles and boxes below it; a refit to the built positions gives the built tree bit for bit,
//     and a refit after moving an instance holds the same invariants;
//   * the checker's multithreaded renders (NEE, BPT and BDPT on 4 worker threads, with its BVH), whose
//     scene preparation runs the library's own preparation code (runtime/prepare.h, luts.h,
//     rgb2spec.h): every film value finite.
// Prints "host checks ok" and exits 0, or names the failed check and exits 1.
#include "../../oracle/oracle.cpp"
#include "../../yapt_amd/csrc/runtime/bvh.h"

#include <cstdio>
#include <fstream>

namespace {

int failures = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

bool box_holds(const float lo[3], const float hi[3], const float p[3]) {
  for (int a = 0; a < 3; ++a)
    if (!(lo[a] <= p[a] && p[a] <= hi[a])) return false;
  return true;
}

// walks a node layout of WIDTH children from the root; counts how often each record is reached
template <int WIDTH, class Node>
void check_tree(const std::vector<Node>& nodes, const std::vector<ypt::TriRecord>& tris, const char* name) {
  std::vector<uint32_t> seen(tris.size(), 0);
  std::vector<uint32_t> stack{0u};
  size_t visits = 0;
  bool boxes_ok = true, refs_ok = true;
  while (!stack.empty()) {
    uint32_t n = stack.back();
    stack.pop_back();
    if (n >= nodes.size() || ++visits > nodes.size()) { refs_ok = false; break; }
    const Node& nd = nodes[n];
    for (int s = 0; s < WIDTH; ++s) {
      uint32_t c = nd.child[s];
      if (c == ypt::BVH_EMPTY) continue;
      float lo[3] = {nd.lo_x[s], nd.lo_y[s], nd.lo_z[s]}, hi[3] = {nd.hi_x[s], nd.hi_y[s], nd.hi_z[s]};
      if (c & ypt::BVH_LEAF_BIT) {
        uint32_t first = c & 0x07FFFFFFu, count = ((c >> 27) & 15u) + 1u;
        if (first + count > tris.size()) { refs_ok = false; continue; }
        for (uint32_t k = first; k < first + count; ++k) {
          seen[k]++;
          const ypt::TriRecord& r = tris[k];
          boxes_ok = boxes_ok && box_holds(lo, hi, r.v0) && box_holds(lo, hi, r.v1) && box_holds(lo, hi, r.v2);
        }
      } else {
        if (c >= nodes.size()) { refs_ok = false; continue; }
        const Node& ch = nodes[c];
        for (int t = 0; t < WIDTH; ++t) {
          if (ch.child[t] == ypt::BVH_EMPTY) continue;
          float clo[3] = {ch.lo_x[t], ch.lo_y[t], ch.lo_z[t]}, chi[3] = {ch.hi_x[t], ch.hi_y[t], ch.hi_z[t]};
          boxes_ok = boxes_ok && box_holds(lo, hi, clo) && box_holds(lo, hi, chi);
        }
        stack.push_back(c);
      }
    }
  }
  bool once = true;
  for (uint32_t v : seen) once = once && v == 1;
  std::fprintf(stderr, "%s: %zu nodes visited\n", name, visits);
  expect(refs_ok, name);
  expect(once, "every triangle record in exactly one leaf");
  expect(boxes_ok, "boxes hold what lies below them");
}

template <class T>
bool read_array(std::ifstream& f, std::vector<T>& v, size_t n) {
  v.resize(n);
  return n == 0 || (bool)f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: host_checks <scene blob>\n");
    return 2;
  }
  std::ifstream f(argv[1], std::ios::binary);
  uint32_t hdr[12];
  if (!f.read(reinterpret_cast<char*>(hdr), sizeof(hdr)) || hdr[0] != 0x59505448u) {
    std::fprintf(stderr, "bad blob header\n");
    return 2;
  }
  if (hdr[1] != sizeof(ypt_material) || hdr[2] != sizeof(ypt_punctual_light) || hdr[3] != sizeof(ypt_frame_desc)) {
    std::fprintf(stderr, "record sizes differ from the header's\n");
    return 2;
  }
  const uint32_t nv = hdr[4], nt = hdr[5], ni = hdr[6], nm = hdr[7], nl = hdr[8], hasN = hdr[9], hasT = hdr[10], hasUV = hdr[11];
  ypt_frame_desc fd;
  std::vector<ypt_material> mats;
  std::vector<ypt_punctual_light> lights;
  std::vector<float> pos, nrm, tan, uv;
  std::vector<uint32_t> idx, first, count, mat;
  bool ok = (bool)f.read(reinterpret_cast<char*>(&fd), sizeof(fd)) && read_array(f, mats, nm) && read_array(f, lights, nl) &&
            read_array(f, pos, 3 * (size_t)nv) && read_array(f, nrm, hasN ? 3 * (size_t)nv : 0) && read_array(f, tan, hasT ? 4 * (size_t)nv : 0) &&
            read_array(f, uv, hasUV ? 2 * (size_t)nv : 0) && read_array(f, idx, 3 * (size_t)nt) &&
            read_array(f, first, ni) && read_array(f, count, ni) && read_array(f, mat, ni);
  if (!ok) {
    std::fprintf(stderr, "truncated blob\n");
    return 2;
  }

  ypt::BvhResult bvh;
  ypt::BvhBuilder().build(pos.data(), idx.data(), nt, bvh);
  std::vector<uint32_t> ids(nt, 0);
  bool perm = bvh.tris.size() == nt;
  for (const auto& r : bvh.tris) perm = perm && r.id < nt && ids[r.id]++ == 0;
  expect(perm, "triangle records are a permutation of the triangles");
  check_tree<2>(bvh.nodes2, bvh.tris, "BVH2 references");
  check_tree<8>(bvh.nodes8, bvh.tris, "BVH8 references");

  // refit: to unchanged positions it gives the built tree bit for bit; after moving the first
  // instance's vertices the boxes hold the moved triangles again
  {
    ypt::BvhResult same = bvh;
    ypt::bvh_refit(pos.data(), idx.data(), same);
    bool identical = same.nodes2.size() == bvh.nodes2.size() && same.nodes8.size() == bvh.nodes8.size() &&
                     std::memcmp(same.nodes2.data(), bvh.nodes2.data(), bvh.nodes2.size() * sizeof(ypt::Bvh2Node)) == 0 &&
                     std::memcmp(same.nodes8.data(), bvh.nodes8.data(), bvh.nodes8.size() * sizeof(ypt::Bvh8Node)) == 0 &&
                     std::memcmp(same.tris.data(), bvh.tris.data(), bvh.tris.size() * sizeof(ypt::TriRecord)) == 0;
    expect(identical, "refit to the built positions gives the built tree");
    std::vector<float> moved = pos;
    std::vector<uint8_t> touched(nv, 0);
    for (uint32_t t = first[0]; t < first[0] + count[0]; ++t)
      for (int k = 0; k < 3; ++k) touched[idx[3 * t + k]] = 1;
    for (uint32_t v = 0; v < nv; ++v)
      if (touched[v]) { moved[3 * v] += 0.75f; moved[3 * v + 1] -= 0.5f; moved[3 * v + 2] += 2.25f; }
    ypt::BvhResult re = bvh;
    ypt::bvh_refit(moved.data(), idx.data(), re);
    bool records = true;
    for (const auto& r : re.tris)
      for (int a = 0; a < 3; ++a) records = records && r.v0[a] == moved[3 * idx[3 * r.id] + a] && r.v2[a] == moved[3 * idx[3 * r.id + 2] + a];
    expect(records, "refitted triangle records hold the moved vertices");
    check_tree<2>(re.nodes2, re.tris, "refitted BVH2 references");
    check_tree<8>(re.nodes8, re.tris, "refitted BVH8 references");
  }

  ypt_scene_desc sd;
  std::memset(&sd, 0, sizeof(sd));
  sd.positions = pos.data();
  sd.normals = hasN ? nrm.data() : nullptr;
  sd.tangents = hasT ? tan.data() : nullptr;
  sd.uvs = hasUV ? uv.data() : nullptr;
  sd.vertex_count = nv;
  sd.indices = idx.data();
  sd.triangle_count = nt;
  sd.instance_first_triangle = first.data();
  sd.instance_triangle_count = count.data();
  sd.instance_material = mat.data();
  sd.instance_count = ni;
  sd.materials = mats.data();
  sd.material_count = nm;
  sd.punctual_lights = nl ? lights.data() : nullptr;
  sd.punctual_light_count = nl;
  const size_t npix = (size_t)fd.width * fd.height;
  for (int integ : {YPT_INTEGRATOR_NEE, YPT_INTEGRATOR_BPT, YPT_INTEGRATOR_BDPT}) {
    std::vector<float> film(npix * 4), last(npix * 4);
    std::vector<uint32_t> counts(npix);
    double stats[8] = {0};
    int rc = oracle_render(&sd, &fd, integ, 0, 2, 4, 1, 1, film.data(), counts.data(), last.data(), nullptr, stats);
    expect(rc == 0, "checker render");
    bool finite = true;
    for (float v : film) finite = finite && std::isfinite(v);
    expect(finite, "finite film");
    std::fprintf(stderr, "integrator %d: %.0f path samples, %.0f extension rays\n", integ, stats[1], stats[2]);
  }
  if (failures) return 1;
  std::printf("host checks ok\n");
  return 0;
}
