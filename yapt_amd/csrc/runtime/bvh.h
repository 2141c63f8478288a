// Binned-SAH bounding volume hierarchy over a triangle soup (host build, deterministic).
//
// Replaces the driver-built BLAS/TLAS of the reference (AccelerationStructureHelper.cpp:1-262,
// AccelerationStructureBuilderDx12.cpp:65-379, PREFER_FAST_TRACE) with a project-owned structure the
// traversal kernels read from HBM:
//  * BVH2 nodes of 64 bytes, each holding the boxes of both children (one 64-byte fetch per visit);
//  * BVH8 nodes of 256 bytes made by collapsing the binary tree (fewer, wider visits);
//  * triangles re-ordered into leaf order as 48-byte records (three vertices + original index).
// Child references: bit 31 clear = inner node index; bit 31 set = leaf with
// count = ((ref >> 27) & 15) + 1 triangles starting at record (ref & 0x07FFFFFF).
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace ypt {

static constexpr uint32_t BVH_LEAF_BIT = 0x80000000u;
static constexpr uint32_t BVH_EMPTY = 0xFFFFFFFFu;  // unused child slot
static constexpr int BVH_MAX_LEAF = 8;
static constexpr int BVH_MAX_DEPTH = 64;

struct alignas(16) Bvh2Node {
  float lo_x[2], hi_x[2], lo_y[2], hi_y[2];
  float lo_z[2], hi_z[2];
  uint32_t child[2];
  uint32_t pad[2];
};
static_assert(sizeof(Bvh2Node) == 64, "Bvh2Node is one 64-byte fetch");

struct alignas(16) Bvh8Node {
  float lo_x[8], hi_x[8], lo_y[8], hi_y[8], lo_z[8], hi_z[8];
  uint32_t child[8];
  uint32_t pad[8];
};
static_assert(sizeof(Bvh8Node) == 256, "Bvh8Node is 256 bytes");

struct alignas(16) TriRecord {
  float v0[3];
  uint32_t id;
  float v1[3];
  uint32_t alpha;  // 0: opaque; else 1 + the triangle's alpha-test record
  float v2[3];
  uint32_t pad1;
};
static_assert(sizeof(TriRecord) == 48, "TriRecord is 48 bytes");

struct Aabb {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  void grow(const float p[3]) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
  }
  void grow(const Aabb& b) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
  }
  float area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0.f) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
  }
};

struct BvhBuildStats {
  uint32_t nodes = 0, leaves = 0, maxDepth = 0, maxLeafSize = 0;
};

struct BvhResult {
  std::vector<Bvh2Node> nodes2;
  std::vector<Bvh8Node> nodes8;
  std::vector<TriRecord> tris;
  BvhBuildStats stats;
  Aabb bounds;
};

class BvhBuilder {
 public:
  // positions: 3 floats per vertex; indices: 3 per triangle
  void build(const float* pos, const uint32_t* idx, uint32_t ntri, BvhResult& out) {
    if (ntri == 0) throw std::runtime_error("BVH build: scene has no triangles");
    boxes_.resize(ntri);
    cent_.resize(ntri * 3);
    order_.resize(ntri);
    for (uint32_t t = 0; t < ntri; ++t) {
      Aabb b;
      for (int k = 0; k < 3; ++k) b.grow(pos + 3 * idx[3 * t + k]);
      boxes_[t] = b;
      for (int a = 0; a < 3; ++a) cent_[3 * t + a] = 0.5f * (b.lo[a] + b.hi[a]);
      order_[t] = t;
    }
    nodes_.clear();
    nodes_.reserve(2 * ntri);
    BNode root;
    root.begin = 0; root.end = ntri;
    nodes_.push_back(root);
    build_rec(0, 0, out.stats);
    // triangle records in leaf order
    out.tris.resize(ntri);
    for (uint32_t k = 0; k < ntri; ++k) {
      uint32_t t = order_[k];
      TriRecord& r = out.tris[k];
      std::memset(&r, 0, sizeof(r));
      const float* p0 = pos + 3 * idx[3 * t];
      const float* p1 = pos + 3 * idx[3 * t + 1];
      const float* p2 = pos + 3 * idx[3 * t + 2];
      for (int a = 0; a < 3; ++a) { r.v0[a] = p0[a]; r.v1[a] = p1[a]; r.v2[a] = p2[a]; }
      r.id = t;
    }
    out.bounds = nodes_[0].box;
    emit_bvh2(out);
    emit_bvh8(out);
  }

 private:
  struct BNode {
    Aabb box;
    uint32_t begin = 0, end = 0;
    int32_t left = -1, right = -1;
  };
  std::vector<Aabb> boxes_;
  std::vector<float> cent_;
  std::vector<uint32_t> order_;
  std::vector<BNode> nodes_;

  static constexpr int NBINS = 32;

  void build_rec(uint32_t ni, uint32_t depth, BvhBuildStats& st) {
    // explicit stack of (node, depth) to avoid deep recursion
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    stack.push_back({ni, depth});
    while (!stack.empty()) {
      auto [n, d] = stack.back();
      stack.pop_back();
      uint32_t b = nodes_[n].begin, e = nodes_[n].end, cnt = e - b;
      Aabb box, cbox;
      for (uint32_t k = b; k < e; ++k) {
        box.grow(boxes_[order_[k]]);
        cbox.grow(&cent_[3 * order_[k]]);
      }
      nodes_[n].box = box;
      st.maxDepth = std::max(st.maxDepth, d);
      if (cnt <= 2) { st.leaves++; st.maxLeafSize = std::max(st.maxLeafSize, cnt); continue; }
      // SAH over binned centroids, all three axes
      float bestCost = FLT_MAX;
      int bestAxis = -1, bestSplit = -1;
      for (int a = 0; a < 3; ++a) {
        float ext = cbox.hi[a] - cbox.lo[a];
        if (!(ext > 0.f)) continue;
        Aabb bb[NBINS];
        uint32_t bc[NBINS] = {0};
        float scale = NBINS / ext;
        for (uint32_t k = b; k < e; ++k) {
          uint32_t t = order_[k];
          int bi = std::min(NBINS - 1, (int)((cent_[3 * t + a] - cbox.lo[a]) * scale));
          bb[bi].grow(boxes_[t]);
          bc[bi]++;
        }
        float ra[NBINS];
        uint32_t rc[NBINS];
        Aabb acc;
        uint32_t c = 0;
        for (int i = NBINS - 1; i > 0; --i) {
          acc.grow(bb[i]); c += bc[i];
          ra[i] = acc.area(); rc[i] = c;
        }
        acc = Aabb(); c = 0;
        for (int i = 0; i < NBINS - 1; ++i) {
          acc.grow(bb[i]); c += bc[i];
          if (c == 0 || rc[i + 1] == 0) continue;
          float cost = acc.area() * c + ra[i + 1] * rc[i + 1];
          if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = i; }
        }
      }
      float leafCost = box.area() * cnt;
      float splitCost = bestAxis >= 0 ? box.area() * 1.0f + bestCost : FLT_MAX;  // traversal step costs one box test
      uint32_t mid;
      if (bestAxis < 0) {
        if (cnt <= (uint32_t)BVH_MAX_LEAF) { st.leaves++; st.maxLeafSize = std::max(st.maxLeafSize, cnt); continue; }
        mid = b + cnt / 2;  // coincident centroids: split the index range
      } else {
        if (cnt <= (uint32_t)BVH_MAX_LEAF && leafCost <= splitCost) {
          st.leaves++; st.maxLeafSize = std::max(st.maxLeafSize, cnt); continue;
        }
        float ext = cbox.hi[bestAxis] - cbox.lo[bestAxis];
        float scale = NBINS / ext;
        auto it = std::stable_partition(order_.begin() + b, order_.begin() + e, [&](uint32_t t) {
          int bi = std::min(NBINS - 1, (int)((cent_[3 * t + bestAxis] - cbox.lo[bestAxis]) * scale));
          return bi <= bestSplit;
        });
        mid = (uint32_t)(it - order_.begin());
        if (mid == b || mid == e) mid = b + cnt / 2;
      }
      if (d + 1 >= (uint32_t)BVH_MAX_DEPTH) throw std::runtime_error("BVH build: depth limit exceeded");
      BNode L, R;
      L.begin = b; L.end = mid;
      R.begin = mid; R.end = e;
      int32_t li = (int32_t)nodes_.size();
      nodes_.push_back(L);
      nodes_.push_back(R);
      nodes_[n].left = li;
      nodes_[n].right = li + 1;
      st.nodes++;
      stack.push_back({(uint32_t)li + 1, d + 1});
      stack.push_back({(uint32_t)li, d + 1});
    }
  }

  static bool is_leaf(const BNode& n) { return n.left < 0; }
  static uint32_t leaf_ref(const BNode& n) {
    uint32_t cnt = n.end - n.begin;
    return BVH_LEAF_BIT | ((cnt - 1) << 27) | n.begin;
  }

  // BVH2: one GPU node per inner build node, children's boxes stored in the parent.
  void emit_bvh2(BvhResult& out) {
    std::vector<uint32_t> gpuIndex(nodes_.size(), BVH_EMPTY);
    std::vector<uint32_t> inner;
    // breadth-first numbering keeps the top levels together
    std::vector<uint32_t> q;
    q.push_back(0);
    for (size_t h = 0; h < q.size(); ++h) {
      const BNode& n = nodes_[q[h]];
      if (is_leaf(n) && q[h] != 0) continue;
      gpuIndex[q[h]] = (uint32_t)inner.size();
      inner.push_back(q[h]);
      if (!is_leaf(n)) { q.push_back(n.left); q.push_back(n.right); }
    }
    out.nodes2.assign(inner.size(), Bvh2Node());
    for (size_t k = 0; k < inner.size(); ++k) {
      const BNode& n = nodes_[inner[k]];
      Bvh2Node& g = out.nodes2[k];
      std::memset(&g, 0, sizeof(g));
      if (is_leaf(n)) {  // whole scene in one leaf: root with one leaf child and one empty slot
        set_child2(g, 0, n, leaf_ref(n));
        set_empty2(g, 1);
        continue;
      }
      const BNode* ch[2] = {&nodes_[n.left], &nodes_[n.right]};
      for (int c = 0; c < 2; ++c) {
        uint32_t ref = is_leaf(*ch[c]) ? leaf_ref(*ch[c]) : gpuIndex[c == 0 ? n.left : n.right];
        set_child2(g, c, *ch[c], ref);
      }
    }
  }
  static void set_child2(Bvh2Node& g, int c, const BNode& ch, uint32_t ref) {
    g.lo_x[c] = ch.box.lo[0]; g.hi_x[c] = ch.box.hi[0];
    g.lo_y[c] = ch.box.lo[1]; g.hi_y[c] = ch.box.hi[1];
    g.lo_z[c] = ch.box.lo[2]; g.hi_z[c] = ch.box.hi[2];
    g.child[c] = ref;
  }
  static void set_empty2(Bvh2Node& g, int c) {
    g.lo_x[c] = g.lo_y[c] = g.lo_z[c] = FLT_MAX;
    g.hi_x[c] = g.hi_y[c] = g.hi_z[c] = -FLT_MAX;
    g.child[c] = BVH_EMPTY;
  }

  // BVH8: collapse the binary tree, repeatedly opening the child with the largest surface area.
  void emit_bvh8(BvhResult& out) {
    std::vector<uint32_t> work;  // build-node ids that become BVH8 nodes
    std::vector<std::vector<uint32_t>> kids;
    work.push_back(0);
    out.nodes8.clear();
    for (size_t h = 0; h < work.size(); ++h) {
      uint32_t bn = work[h];
      std::vector<uint32_t> c;
      if (is_leaf(nodes_[bn])) c.push_back(bn);
      else { c.push_back(nodes_[bn].left); c.push_back(nodes_[bn].right); }
      while (c.size() < 8) {
        int best = -1;
        float ba = -1.f;
        for (size_t i = 0; i < c.size(); ++i)
          if (!is_leaf(nodes_[c[i]]) && nodes_[c[i]].box.area() > ba) { ba = nodes_[c[i]].box.area(); best = (int)i; }
        if (best < 0) break;
        uint32_t open = c[best];
        c[best] = nodes_[open].left;
        c.insert(c.begin() + best + 1, nodes_[open].right);
      }
      kids.push_back(c);
      for (uint32_t x : c)
        if (!is_leaf(nodes_[x])) work.push_back(x);
    }
    std::vector<uint32_t> gidx(nodes_.size(), BVH_EMPTY);
    for (size_t h = 0; h < work.size(); ++h) gidx[work[h]] = (uint32_t)h;
    out.nodes8.assign(work.size(), Bvh8Node());
    for (size_t h = 0; h < work.size(); ++h) {
      Bvh8Node& g = out.nodes8[h];
      std::memset(&g, 0, sizeof(g));
      for (int s = 0; s < 8; ++s) {
        if (s < (int)kids[h].size()) {
          const BNode& ch = nodes_[kids[h][s]];
          g.lo_x[s] = ch.box.lo[0]; g.hi_x[s] = ch.box.hi[0];
          g.lo_y[s] = ch.box.lo[1]; g.hi_y[s] = ch.box.hi[1];
          g.lo_z[s] = ch.box.lo[2]; g.hi_z[s] = ch.box.hi[2];
          g.child[s] = is_leaf(ch) ? leaf_ref(ch) : gidx[kids[h][s]];
        } else {
          g.lo_x[s] = g.lo_y[s] = g.lo_z[s] = FLT_MAX;
          g.hi_x[s] = g.hi_y[s] = g.hi_z[s] = -FLT_MAX;
          g.child[s] = BVH_EMPTY;
        }
      }
    }
  }
};

// Refit: new vertex positions for the same triangles (a transform change of render objects) keep
// the tree and recompute its boxes. Triangle records take their vertices from `pos` by triangle id;
// nodes are numbered breadth-first, so every child comes after its parent and one pass in reverse
// order sees children first. A box is the union of its triangles' vertices or of its child node's
// boxes, as the build makes it, so refitting to unchanged positions gives the built tree bit for bit.
// The tree's quality decays as objects move away from where it was built; the result stays exact.
inline void bvh_refit(const float* pos, const uint32_t* idx, BvhResult& b) {
  for (TriRecord& r : b.tris) {
    const float* p[3] = {pos + 3 * idx[3 * r.id], pos + 3 * idx[3 * r.id + 1], pos + 3 * idx[3 * r.id + 2]};
    for (int a = 0; a < 3; ++a) { r.v0[a] = p[0][a]; r.v1[a] = p[1][a]; r.v2[a] = p[2][a]; }
  }
  auto leaf_box = [&](uint32_t ref) {
    Aabb box;
    uint32_t first = ref & 0x07FFFFFFu, count = ((ref >> 27) & 15u) + 1u;
    for (uint32_t k = first; k < first + count; ++k) {
      const TriRecord& r = b.tris[k];
      box.grow(r.v0); box.grow(r.v1); box.grow(r.v2);
    }
    return box;
  };
  auto set = [](float* lo_x, float* hi_x, float* lo_y, float* hi_y, float* lo_z, float* hi_z, int c, const Aabb& box) {
    lo_x[c] = box.lo[0]; hi_x[c] = box.hi[0];
    lo_y[c] = box.lo[1]; hi_y[c] = box.hi[1];
    lo_z[c] = box.lo[2]; hi_z[c] = box.hi[2];
  };
  for (size_t k = b.nodes2.size(); k-- > 0;) {
    Bvh2Node& g = b.nodes2[k];
    for (int c = 0; c < 2; ++c) {
      uint32_t ref = g.child[c];
      if (ref == BVH_EMPTY) continue;
      Aabb box;
      if (ref & BVH_LEAF_BIT) {
        box = leaf_box(ref);
      } else {
        const Bvh2Node& ch = b.nodes2[ref];
        for (int d = 0; d < 2; ++d)
          if (ch.child[d] != BVH_EMPTY) {
            float lo[3] = {ch.lo_x[d], ch.lo_y[d], ch.lo_z[d]}, hi[3] = {ch.hi_x[d], ch.hi_y[d], ch.hi_z[d]};
            box.grow(lo); box.grow(hi);
          }
      }
      set(g.lo_x, g.hi_x, g.lo_y, g.hi_y, g.lo_z, g.hi_z, c, box);
    }
  }
  for (size_t k = b.nodes8.size(); k-- > 0;) {
    Bvh8Node& g = b.nodes8[k];
    for (int c = 0; c < 8; ++c) {
      uint32_t ref = g.child[c];
      if (ref == BVH_EMPTY) continue;
      Aabb box;
      if (ref & BVH_LEAF_BIT) {
        box = leaf_box(ref);
      } else {
        const Bvh8Node& ch = b.nodes8[ref];
        for (int d = 0; d < 8; ++d)
          if (ch.child[d] != BVH_EMPTY) {
            float lo[3] = {ch.lo_x[d], ch.lo_y[d], ch.lo_z[d]}, hi[3] = {ch.hi_x[d], ch.hi_y[d], ch.hi_z[d]};
            box.grow(lo); box.grow(hi);
          }
      }
      set(g.lo_x, g.hi_x, g.lo_y, g.hi_y, g.lo_z, g.hi_z, c, box);
    }
  }
  Aabb all;
  for (const TriRecord& r : b.tris) { all.grow(r.v0); all.grow(r.v1); all.grow(r.v2); }
  b.bounds = all;
}

}  // namespace ypt
