// scene.cpp -- assembles the BLAS/TLAS forests and the child-pair layout.
//
// Reference: NaiveSahBvhSource::build_blas / build_tlas / build
// (shader/ray-tracing/src/backend/wavefront_compute/geometry/naive/mod.rs:122-494).
// Compiled with -ffp-contract=off: matrix inverses, instance bounds and the
// triangle pre-pass are single-rounding FP32 in the reference operation order.

#include "scene.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>

namespace rdrt {
namespace {

// ---- Mat4 helpers, column-major fields a1..d4 (math/algebra/src/mat/mat4.rs) ----
struct M4 {
  float a1, a2, a3, a4, b1, b2, b3, b4, c1, c2, c3, c4, d1, d2, d3, d4;
};
static_assert(sizeof(M4) == 64, "M4 layout");

M4 load_m4(const float *m) {
  M4 r;
  std::memcpy(&r, m, 64);
  return r;
}

// mat4.rs det: the 24-term expansion, summed left to right.
float det4(const M4 &m) {
  return m.a1 * m.b2 * m.c3 * m.d4 - m.a1 * m.b2 * m.c4 * m.d3 + m.a1 * m.b3 * m.c4 * m.d2 -
         m.a1 * m.b3 * m.c2 * m.d4 + m.a1 * m.b4 * m.c2 * m.d3 - m.a1 * m.b4 * m.c3 * m.d2 -
         m.a2 * m.b3 * m.c4 * m.d1 + m.a2 * m.b3 * m.c1 * m.d4 - m.a2 * m.b4 * m.c1 * m.d3 +
         m.a2 * m.b4 * m.c3 * m.d1 - m.a2 * m.b1 * m.c3 * m.d4 + m.a2 * m.b1 * m.c4 * m.d3 +
         m.a3 * m.b4 * m.c1 * m.d2 - m.a3 * m.b4 * m.c2 * m.d1 + m.a3 * m.b1 * m.c2 * m.d4 -
         m.a3 * m.b1 * m.c4 * m.d2 + m.a3 * m.b2 * m.c4 * m.d1 - m.a3 * m.b2 * m.c1 * m.d4 -
         m.a4 * m.b1 * m.c2 * m.d3 + m.a4 * m.b1 * m.c3 * m.d2 - m.a4 * m.b2 * m.c3 * m.d1 +
         m.a4 * m.b2 * m.c1 * m.d3 - m.a4 * m.b3 * m.c1 * m.d2 + m.a4 * m.b3 * m.c2 * m.d1;
}

// 3x3 minor sum used by every cofactor of mat4.rs inverse:
// p*(q1*r2 - q2*r1) + s*(...) + w*(...), evaluated in that order.
inline float cof(float p, float q1, float r2, float q2, float r1, float s, float s1, float t2, float s2, float t1,
                 float w, float w1, float u2, float w2, float u1) {
  return p * (q1 * r2 - q2 * r1) + s * (s1 * t2 - s2 * t1) + w * (w1 * u2 - w2 * u1);
}

// mat4.rs inverse, falling back to identity (mat/dimension.rs:18-20).
M4 inverse_or_identity(const M4 &m) {
  float det = det4(m);
  M4 r{};
  if (det == 0.0f) {
    r.a1 = r.b2 = r.c3 = r.d4 = 1.0f;
    return r;
  }
  float pi = 1.0f / det, ni = -pi;
  r.a1 = pi * cof(m.b2, m.c3, m.d4, m.c4, m.d3, m.b3, m.c4, m.d2, m.c2, m.d4, m.b4, m.c2, m.d3, m.c3, m.d2);
  r.a2 = ni * cof(m.a2, m.c3, m.d4, m.c4, m.d3, m.a3, m.c4, m.d2, m.c2, m.d4, m.a4, m.c2, m.d3, m.c3, m.d2);
  r.a3 = pi * cof(m.a2, m.b3, m.d4, m.b4, m.d3, m.a3, m.b4, m.d2, m.b2, m.d4, m.a4, m.b2, m.d3, m.b3, m.d2);
  r.a4 = ni * cof(m.a2, m.b3, m.c4, m.b4, m.c3, m.a3, m.b4, m.c2, m.b2, m.c4, m.a4, m.b2, m.c3, m.b3, m.c2);
  r.b1 = ni * cof(m.b1, m.c3, m.d4, m.c4, m.d3, m.b3, m.c4, m.d1, m.c1, m.d4, m.b4, m.c1, m.d3, m.c3, m.d1);
  r.b2 = pi * cof(m.a1, m.c3, m.d4, m.c4, m.d3, m.a3, m.c4, m.d1, m.c1, m.d4, m.a4, m.c1, m.d3, m.c3, m.d1);
  r.b3 = ni * cof(m.a1, m.b3, m.d4, m.b4, m.d3, m.a3, m.b4, m.d1, m.b1, m.d4, m.a4, m.b1, m.d3, m.b3, m.d1);
  r.b4 = pi * cof(m.a1, m.b3, m.c4, m.b4, m.c3, m.a3, m.b4, m.c1, m.b1, m.c4, m.a4, m.b1, m.c3, m.b3, m.c1);
  r.c1 = pi * cof(m.b1, m.c2, m.d4, m.c4, m.d2, m.b2, m.c4, m.d1, m.c1, m.d4, m.b4, m.c1, m.d2, m.c2, m.d1);
  r.c2 = ni * cof(m.a1, m.c2, m.d4, m.c4, m.d2, m.a2, m.c4, m.d1, m.c1, m.d4, m.a4, m.c1, m.d2, m.c2, m.d1);
  r.c3 = pi * cof(m.a1, m.b2, m.d4, m.b4, m.d2, m.a2, m.b4, m.d1, m.b1, m.d4, m.a4, m.b1, m.d2, m.b2, m.d1);
  r.c4 = ni * cof(m.a1, m.b2, m.c4, m.b4, m.c2, m.a2, m.b4, m.c1, m.b1, m.c4, m.a4, m.b1, m.c2, m.b2, m.c1);
  r.d1 = ni * cof(m.b1, m.c2, m.d3, m.c3, m.d2, m.b2, m.c3, m.d1, m.c1, m.d3, m.b3, m.c1, m.d2, m.c2, m.d1);
  r.d2 = pi * cof(m.a1, m.c2, m.d3, m.c3, m.d2, m.a2, m.c3, m.d1, m.c1, m.d3, m.a3, m.c1, m.d2, m.c2, m.d1);
  r.d3 = ni * cof(m.a1, m.b2, m.d3, m.b3, m.d2, m.a2, m.b3, m.d1, m.b1, m.d3, m.a3, m.b1, m.d2, m.b2, m.d1);
  r.d4 = pi * cof(m.a1, m.b2, m.c3, m.b3, m.c2, m.a2, m.b3, m.c1, m.b1, m.c3, m.a3, m.b1, m.c2, m.b2, m.c1);
  return r;
}

// mat3.rs:37-42 on the upper-left 3x3
float det3(const M4 &m) {
  float t11 = m.c3 * m.b2 - m.b3 * m.c2;
  float t12 = m.b3 * m.c1 - m.c3 * m.b1;
  float t13 = m.c2 * m.b1 - m.b2 * m.c1;
  return m.a1 * t11 + m.a2 * t12 + m.a3 * t13;
}

// mat4.rs:155-169: (M * (p, 1)).xyz / w
void transform_point(const M4 &m, const float p[3], float out[3]) {
  float x = p[0] * m.a1 + p[1] * m.b1 + p[2] * m.c1 + 1.0f * m.d1;
  float y = p[0] * m.a2 + p[1] * m.b2 + p[2] * m.c2 + 1.0f * m.d2;
  float z = p[0] * m.a3 + p[1] * m.b3 + p[2] * m.c3 + 1.0f * m.d3;
  float w = p[0] * m.a4 + p[1] * m.b4 + p[2] * m.c4 + 1.0f * m.d4;
  out[0] = x / w;
  out[1] = y / w;
  out[2] = z / w;
}

inline void grow_point(Box &b, const float p[3]) {
  for (int k = 0; k < 3; k++) {
    b.mn[k] = min_num(b.mn[k], p[k]);
    b.mx[k] = max_num(b.mx[k], p[k]);
  }
}
inline void grow_box(Box &b, const Box &o) {
  for (int k = 0; k < 3; k++) {
    b.mn[k] = min_num(b.mn[k], o.mn[k]);
    b.mx[k] = max_num(b.mx[k], o.mx[k]);
  }
}
inline bool is_empty(const Box &b) { return b.mx[0] < b.mn[0] || b.mx[1] < b.mn[1] || b.mx[2] < b.mn[2]; }

// box3.rs:22-38: the eight corners in 000..111 order.
Box transform_box(const Box &b, const M4 &m) {
  if (is_empty(b)) return b;
  Box r = box_empty();
  for (int c = 0; c < 8; c++) {
    float p[3] = {(c & 4) ? b.mx[0] : b.mn[0], (c & 2) ? b.mx[1] : b.mn[1], (c & 1) ? b.mx[2] : b.mn[2]};
    float q[3];
    transform_point(m, p, q);
    grow_point(r, q);
  }
  return r;
}

// The vertex-only part of the reference ray/triangle test
// (wavefront_compute/geometry/mod.rs:115-117,138-140,144).
TriPre pre_triangle(const float *v0, const float *v1, const float *v2) {
  TriPre t;
  float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  float mag_sq = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  if (mag_sq > 0.0f) {  // vec/dimension.rs:98-105
    float inv = 1.0f / std::sqrt(mag_sq);
    c[0] = c[0] * inv;
    c[1] = c[1] * inv;
    c[2] = c[2] * inv;
  }
  for (int k = 0; k < 3; k++) {
    t.v0[k] = v0[k];
    t.e1[k] = e1[k];
    t.e2[k] = e2[k];
    t.n[k] = c[k];
  }
  t.uu = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
  t.uv = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
  t.vv = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
  t.inverse_d = 1.0f / (t.uv * t.uv - t.uu * t.vv);
  return t;
}

void put_box(float *mn, float *mx, const Box &b) {
  for (int k = 0; k < 3; k++) {
    mn[k] = b.mn[k];
    mx[k] = b.mx[k];
  }
}

// Child-pair conversion of one tree.  `prim_base` offsets leaf ranges into the
// global primitive array; pair nodes are appended to `pairs`.  Returns the
// root ref; `stack_need` receives the deepest push count on any path.
struct PairConv {
  const Tree &t;
  uint32_t prim_base;
  std::vector<PairNode> &pairs;
  std::string &err;
  bool ok = true;

  uint32_t leaf(uint32_t begin, uint32_t count, const Box &b, uint32_t &depth) {
    uint32_t start = prim_base + begin;
    if (start + count > kLeafMaxStart + 1) {
      ok = false;
      err = "scene too large for the child-pair layout (primitive index >= 2^26)";
      return 0;
    }
    depth = 0;
    if (count <= kLeafMaxCount) return leaf_ref(start, count);
    // An oversized leaf (depth-capped build) becomes a chain of pair nodes
    // with the leaf's box on both sides: same triangles, same tests.
    uint32_t idx = (uint32_t)pairs.size();
    pairs.emplace_back();
    uint32_t sub_depth = 0;
    uint32_t rest = leaf(begin + kLeafMaxCount, count - kLeafMaxCount, b, sub_depth);
    PairNode &p = pairs[idx];
    std::memset(&p, 0, sizeof(p));
    put_box(p.lmn, p.lmx, b);
    put_box(p.rmn, p.rmx, b);
    p.lref = leaf_ref(start, kLeafMaxCount);
    p.rref = rest;
    depth = 1 + sub_depth;
    return idx;
  }

  uint32_t ref_of(uint32_t i, uint32_t &depth) {
    const TreeNode &n = t.nodes[i];
    if (!n.inner) return leaf(n.begin, n.end - n.begin, n.box, depth);
    uint32_t idx = (uint32_t)pairs.size();
    pairs.emplace_back();
    uint32_t l = i + 1, r = i + 1 + t.nodes[i + 1].size;
    uint32_t dl = 0, dr = 0;
    uint32_t lref = ref_of(l, dl);
    uint32_t rref = ref_of(r, dr);
    PairNode &p = pairs[idx];
    std::memset(&p, 0, sizeof(p));
    put_box(p.lmn, p.lmx, t.nodes[l].box);
    put_box(p.rmn, p.rmx, t.nodes[r].box);
    p.lref = lref;
    p.rref = rref;
    depth = 1 + std::max(dl, dr);
    return idx;
  }
};

// Collapse of a child-pair tree into child-quad nodes.  A quad starts from
// the two children of one pair node and opens its inner child of largest
// surface area (replacing it by that node's two children) until it holds
// four children or only leaves; leaves keep their refs.  `need` receives the
// deepest push count below `ref`: a quad of n children pushes at most n - 1.
struct QuadConv {
  const std::vector<PairNode> &pairs;
  std::vector<QuadNode> &quads;

  struct Child {
    float mn[3], mx[3];
    uint32_t ref;
  };
  static Child left(const PairNode &p) { return Child{{p.lmn[0], p.lmn[1], p.lmn[2]}, {p.lmx[0], p.lmx[1], p.lmx[2]}, p.lref}; }
  static Child right(const PairNode &p) {
    return Child{{p.rmn[0], p.rmn[1], p.rmn[2]}, {p.rmx[0], p.rmx[1], p.rmx[2]}, p.rref};
  }
  static double area(const Child &c) {
    const double x = (double)c.mx[0] - c.mn[0], y = (double)c.mx[1] - c.mn[1], z = (double)c.mx[2] - c.mn[2];
    return x * y + y * z + z * x;
  }

  uint32_t ref_of(uint32_t ref, uint32_t &need) {
    need = 0;
    if (ref & kLeafBit) return ref;
    Child ch[4] = {left(pairs[ref]), right(pairs[ref])};
    int n = 2;
    while (n < 4) {
      int best = -1;
      double best_area = -1.0;
      for (int k = 0; k < n; k++)
        if (!(ch[k].ref & kLeafBit) && area(ch[k]) > best_area) {
          best = k;
          best_area = area(ch[k]);
        }
      if (best < 0) break;
      const PairNode &q = pairs[ch[best].ref];
      ch[best] = left(q);
      ch[n++] = right(q);
    }
    const uint32_t idx = (uint32_t)quads.size();
    quads.emplace_back();
    uint32_t refs[4] = {kInvalid, kInvalid, kInvalid, kInvalid}, deepest = 0;
    for (int k = 0; k < n; k++) {
      uint32_t d = 0;
      refs[k] = ref_of(ch[k].ref, d);
      deepest = std::max(deepest, d);
    }
    QuadNode &qn = quads[idx];
    std::memset(&qn, 0, sizeof(qn));
    for (int k = 0; k < n; k++) {
      qn.mnx[k] = ch[k].mn[0];
      qn.mny[k] = ch[k].mn[1];
      qn.mnz[k] = ch[k].mn[2];
      qn.mxx[k] = ch[k].mx[0];
      qn.mxy[k] = ch[k].mx[1];
      qn.mxz[k] = ch[k].mx[2];
    }
    for (int k = 0; k < 4; k++) qn.ref[k] = refs[k];
    need = (uint32_t)(n - 1) + deepest;
    return idx;
  }
};

}  // namespace

// Quantized twin of a child-quad node (see QuadNodeQ).  Per axis the step
// is the smallest power of two for which every child's rounded bounds fit in
// 8 bits; each bound is then moved outward until corner + q * step (FP32,
// one rounding, as the kernels decode it) contains the exact bound.
QuadNodeQ quantize_quad(const QuadNode &q) {
  QuadNodeQ out;
  std::memset(&out, 0, sizeof(out));
  int n = 0;
  while (n < 4 && q.ref[n] != kInvalid) n++;
  const float *mn[3] = {q.mnx, q.mny, q.mnz}, *mx[3] = {q.mxx, q.mxy, q.mxz};
  float corner[3], step[3];
  uint32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int a = 0; a < 3; a++) {
    float l = mn[a][0], h = mx[a][0];
    for (int k = 1; k < n; k++) {
      l = std::min(l, mn[a][k]);
      h = std::max(h, mx[a][k]);
    }
    corner[a] = l;
    const double ext = (double)h - (double)l;
    const double min_step = std::ldexp(1.0, -126), max_step = std::ldexp(1.0, 127);
    double s = min_step;
    if (ext > 0.0 && std::isfinite(ext)) s = std::max(min_step, std::ldexp(1.0, (int)std::ceil(std::log2(ext / 255.0))));
    if (!std::isfinite(ext)) s = 2.0 * max_step;
    for (;;) {
      if (s > max_step) {
        // no power-of-two step fits (non-finite bounds): the whole axis
        // decodes to [corner, +inf), which contains every child
        step[a] = FLT_MAX;
        lo[a] = 0u;
        hi[a] = 0u;
        for (int k = 0; k < n; k++) hi[a] |= 255u << (8 * k);
        break;
      }
      const float sf = (float)s;
      bool ok = true;
      uint32_t wl = 0, wh = 0;
      for (int k = 0; k < n && ok; k++) {
        double ql = std::floor(((double)mn[a][k] - (double)l) / s), qh = std::ceil(((double)mx[a][k] - (double)l) / s);
        ql = std::min(std::max(ql, 0.0), 255.0);
        qh = std::max(qh, 0.0);
        uint32_t il = (uint32_t)ql, ih = (uint32_t)std::min(qh, 256.0);
        while (il > 0 && l + (float)il * sf > mn[a][k]) il--;
        while (ih <= 255 && l + (float)ih * sf < mx[a][k]) ih++;
        if (ih > 255) {
          ok = false;
          break;
        }
        wl |= il << (8 * k);
        wh |= ih << (8 * k);
      }
      if (ok) {
        step[a] = sf;
        lo[a] = wl;
        hi[a] = wh;
        break;
      }
      s *= 2.0;
    }
  }
  out.ox = corner[0];
  out.oy = corner[1];
  out.oz = corner[2];
  out.sx = step[0];
  out.sy = step[1];
  out.sz = step[2];
  out.lox = lo[0];
  out.hix = hi[0];
  out.loy = lo[1];
  out.hiy = hi[1];
  out.loz = lo[2];
  out.hiz = hi[2];
  for (int k = 0; k < 4; k++) out.ref[k] = q.ref[k];
  return out;
}

namespace {

// Child-pair conversion of the fast kernel's TLAS tree, whose leaves hold
// single instances addressed by their forest position: map[order index].  A
// depth-capped leaf with several instances becomes a chain of pair nodes.
struct SinglePairConv {
  const Tree &t;
  const std::vector<uint32_t> &map;
  std::vector<PairNode> &pairs;

  uint32_t leaf(uint32_t begin, uint32_t count, const Box &b, uint32_t &depth) {
    depth = 0;
    if (count == 0) return leaf_ref(0, 0);  // empty TLAS
    if (count == 1) return leaf_ref(map[t.order[begin]], 1);
    uint32_t idx = (uint32_t)pairs.size();
    pairs.emplace_back();
    uint32_t sub = 0;
    uint32_t rest = leaf(begin + 1, count - 1, b, sub);
    PairNode &p = pairs[idx];
    std::memset(&p, 0, sizeof(p));
    put_box(p.lmn, p.lmx, b);
    put_box(p.rmn, p.rmx, b);
    p.lref = leaf_ref(map[t.order[begin]], 1);
    p.rref = rest;
    depth = 1 + sub;
    return idx;
  }

  uint32_t ref_of(uint32_t i, uint32_t &depth) {
    const TreeNode &n = t.nodes[i];
    if (!n.inner) return leaf(n.begin, n.end - n.begin, n.box, depth);
    uint32_t idx = (uint32_t)pairs.size();
    pairs.emplace_back();
    uint32_t l = i + 1, r = i + 1 + t.nodes[i + 1].size;
    uint32_t dl = 0, dr = 0;
    uint32_t lref = ref_of(l, dl);
    uint32_t rref = ref_of(r, dr);
    PairNode &p = pairs[idx];
    std::memset(&p, 0, sizeof(p));
    put_box(p.lmn, p.lmx, t.nodes[l].box);
    put_box(p.rmn, p.rmx, t.nodes[r].box);
    p.lref = lref;
    p.rref = rref;
    depth = 1 + std::max(dl, dr);
    return idx;
  }
};

// Reference forest nodes of one tree (naive/mod.rs:338-367).
void append_forest(const Tree &t, uint32_t prim_offset, std::vector<RefNode> &out) {
  uint32_t base = (uint32_t)out.size();
  for (uint32_t i = 0; i < t.nodes.size(); i++) {
    const TreeNode &n = t.nodes[i];
    RefNode g;
    std::memset(&g, 0, sizeof(g));
    put_box(g.mn, g.mx, n.box);
    uint32_t h = t.hit(i), m = t.miss(i);
    g.hit_next = h != kInvalid ? h + base : kInvalid;
    g.miss_next = m != kInvalid ? m + base : kInvalid;
    g.begin = n.begin + prim_offset;
    g.end = n.end + prim_offset;
    out.push_back(g);
  }
}

template <class F>
void parallel_for(size_t n, unsigned threads, F f) {
  if (threads <= 1 || n <= 1) {
    for (size_t i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> pool;
  std::atomic<size_t> next{0};
  for (unsigned w = 0; w < threads; w++)
    pool.emplace_back([&] {
      for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &th : pool) th.join();
}

// Every ref the kernels follow must stay inside the arrays they index: inner
// refs inside the pair array, leaf ranges inside the triangle / instance
// array, forest links and ranges inside theirs.  Checked once per build.
bool ref_ok(uint32_t ref, size_t npairs, size_t nprims) {
  if (ref & kLeafBit) {
    uint32_t start = (ref & ~kLeafBit) >> kLeafCountBits, count = ref & kLeafMaxCount;
    return (size_t)start + count <= nprims;
  }
  return ref < npairs;
}

bool validate_refs(const BuiltScene &s, std::string &err) {
  for (const PairNode &p : s.tri_pairs)
    if (!ref_ok(p.lref, s.tri_pairs.size(), s.tri_ids.size()) || !ref_ok(p.rref, s.tri_pairs.size(), s.tri_ids.size())) {
      err = "internal: BLAS child-pair ref out of range";
      return false;
    }
  for (const PairNode &p : s.tlas_pairs)
    if (!ref_ok(p.lref, s.tlas_pairs.size(), s.insts.size()) ||
        !ref_ok(p.rref, s.tlas_pairs.size(), s.insts.size())) {
      err = "internal: TLAS child-pair ref out of range";
      return false;
    }
  for (const GeoRec &g : s.geos)
    if (!ref_ok(g.root_ref, s.tri_pairs.size(), s.tri_ids.size()) || g.ref_root >= s.tri_forest.size()) {
      err = "internal: geometry root out of range";
      return false;
    }
  for (const TlasRoot &r : s.tlas_roots)
    if (r.ref_root != kInvalid &&
        (!ref_ok(r.root_ref, s.tlas_pairs.size(), s.insts.size()) || r.ref_root >= s.tlas_forest.size())) {
      err = "internal: TLAS root out of range";
      return false;
    }
  auto quad_ok = [](const QuadNode &q, size_t nquads, size_t nprims) {
    for (int k = 0; k < 4; k++)
      if (q.ref[k] != kInvalid && !ref_ok(q.ref[k], nquads, nprims)) return false;
    return q.ref[0] != kInvalid && q.ref[1] != kInvalid;
  };
  for (const QuadNode &q : s.tri_quads)
    if (!quad_ok(q, s.tri_quads.size(), s.tri_ids.size())) {
      err = "internal: BLAS child-quad ref out of range";
      return false;
    }
  for (const QuadNode &q : s.tlas_quads)
    if (!quad_ok(q, s.tlas_quads.size(), s.insts.size())) {
      err = "internal: TLAS child-quad ref out of range";
      return false;
    }
  auto twins = [](const auto &a, const auto &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
      for (int k = 0; k < 4; k++)
        if (a[i].ref[k] != b[i].ref[k]) return false;
    return true;
  };
  if ((!s.device_records && !twins(s.tri_quads, s.tri_qquads)) || !twins(s.tlas_quads, s.tlas_qquads)) {
    err = "internal: quantized quads disagree with the quads";
    return false;
  }
  for (const GeoRec &g : s.geos)
    if (!ref_ok(g.quad_root, s.tri_quads.size(), s.tri_ids.size())) {
      err = "internal: geometry quad root out of range";
      return false;
    }
  if (s.tlas_quad_roots.size() != s.tlas_roots.size()) {
    err = "internal: TLAS quad roots disagree";
    return false;
  }
  for (size_t i = 0; i < s.tlas_roots.size(); i++)
    if (s.tlas_roots[i].ref_root != kInvalid && !ref_ok(s.tlas_quad_roots[i], s.tlas_quads.size(), s.insts.size())) {
      err = "internal: TLAS quad root out of range";
      return false;
    }
  for (const InstRec &in : s.insts)
    if (in.geo_begin > in.geo_end || in.geo_end > s.geos.size()) {
      err = "internal: instance geometry range out of range";
      return false;
    }
  auto forest_ok = [](const auto &f, size_t nprims) {
    for (const RefNode &n : f) {
      if (n.hit_next != kInvalid && n.hit_next >= f.size()) return false;
      if (n.miss_next != kInvalid && n.miss_next >= f.size()) return false;
      if (n.begin > n.end || n.end > nprims) return false;
    }
    return true;
  };
  if (!forest_ok(s.tri_forest, s.redirect.size()) || !forest_ok(s.tlas_forest, s.insts.size())) {
    err = "internal: forest link or range out of range";
    return false;
  }
  if (s.tri_ids.size() != s.redirect.size() || (!s.device_records && s.tris.size() != s.tri_ids.size())) {
    err = "internal: triangle arrays disagree";
    return false;
  }
  for (uint32_t r : s.tri_ids)
    if (r >= s.indices.size() / 3) {
      err = "internal: triangle id out of range";
      return false;
    }
  for (uint32_t r : s.redirect)
    if ((size_t)3 * r + 2 >= s.indices.size()) {
      err = "internal: redirect out of range";
      return false;
    }
  for (uint32_t v : s.indices)
    if ((size_t)3 * v + 2 >= s.vertices.size()) {
      err = "internal: vertex index out of range";
      return false;
    }
  return true;
}

}  // namespace

// The TLAS layer (naive/mod.rs:262-320, 435-454) from the BLAS boxes in
// `out`: instance records, the reference TLAS forest, and the fast kernels'
// single-instance TLAS in pair, quad and quantized-quad layouts.  Clears and
// refills those arrays, so a refit of the BLAS boxes can rebuild it.
int build_tlas_layer(const SceneSrc &src, BuiltScene &out, std::string &err) {
  out.insts.clear();
  out.inst_fwd.clear();
  out.inst_lookup.clear();
  out.tlas_inst_first.clear();
  out.tlas_forest.clear();
  out.tlas_pairs.clear();
  out.tlas_quads.clear();
  out.tlas_qquads.clear();
  out.tlas_roots.clear();
  out.tlas_quad_roots.clear();
  out.max_tlas_depth = 0;
  out.tlas_stack = 0;
  out.tlas_stack4 = 0;
  for (uint32_t ti = 0; ti < src.tlas.size(); ti++) {
    TlasRoot root;
    std::memset(&root, 0, sizeof(root));
    out.tlas_inst_first.push_back((uint32_t)out.insts.size());
    if (!src.tlas_live[ti]) {
      root.ref_root = kInvalid;
      root.root_ref = kInvalid;
      out.tlas_roots.push_back(root);
      out.tlas_quad_roots.push_back(kInvalid);
      continue;
    }
    const auto &inst = src.tlas[ti];
    std::vector<Box> boxes;
    for (const rdrt_instance &in : inst) {
      if (in.blas >= src.blas.size() || !out.blas_has_box[in.blas]) {
        err = "instance references a deleted, empty or unknown BLAS";
        return RDRT_ERR_INVALID_ARG;
      }
      boxes.push_back(transform_box(out.blas_box[in.blas], load_m4(in.transform)));
    }
    Tree tree = build_tree(boxes, src.tlas_opt, 1);
    out.max_tlas_depth = std::max(out.max_tlas_depth, tree.depth);
    uint32_t inst_base = (uint32_t)out.insts.size();
    out.inst_lookup.resize(inst_base + inst.size());
    for (uint32_t j = 0; j < tree.order.size(); j++) out.inst_lookup[inst_base + tree.order[j]] = inst_base + j;
    for (uint32_t k : tree.order) {
      const rdrt_instance &in = inst[k];
      M4 m = load_m4(in.transform);
      uint32_t flags = in.flags;
      if (det3(m) < 0.0f) flags ^= RDRT_INSTANCE_TRIANGLE_FLIP_FACING;
      InstRec r;
      std::memset(&r, 0, sizeof(r));
      M4 inv = inverse_or_identity(m);
      std::memcpy(r.inv, &inv, 64);
      put_box(r.mn, r.mx, boxes[k]);
      r.mask = in.mask;
      r.flags = flags;
      r.geo_begin = out.blas_geo_begin[in.blas];
      r.geo_end = out.blas_geo_end[in.blas];
      r.custom_index = in.instance_custom_index;
      r.order = k;
      r.sbt_offset = in.sbt_offset;
      r.blas = in.blas;
      out.insts.push_back(r);
      out.inst_fwd.insert(out.inst_fwd.end(), in.transform, in.transform + 16);
    }
    root.ref_root = (uint32_t)out.tlas_forest.size();
    append_forest(tree, inst_base, out.tlas_forest);
    put_box(root.mn, root.mx, tree.nodes[0].box);
    // The fast kernel's TLAS: the same binned SAH over the same boxes but with
    // one instance per leaf, so a ray tests only instances whose boxes it
    // enters (the reference's 10-instance leaves are linear scans).  Any tree
    // gives the same closest hit; only its visit order differs.
    if ((uint64_t)inst_base + inst.size() > (uint64_t)kLeafMaxStart + 1) {
      err = "too many instances for the child-pair layout";
      return RDRT_ERR_CAPACITY;
    }
    std::vector<Box> fboxes(inst.size());
    std::vector<uint32_t> fmap(inst.size());
    for (uint32_t j = 0; j < tree.order.size(); j++) {
      fboxes[j] = boxes[tree.order[j]];
      fmap[j] = inst_base + j;
    }
    TreeOptions fine = src.tlas_opt;
    fine.bin_size = 1;
    Tree ftree = build_tree(fboxes, fine, 1);
    SinglePairConv pc{ftree, fmap, out.tlas_pairs};
    uint32_t d = 0;
    root.root_ref = pc.ref_of(0, d);
    out.tlas_stack = std::max(out.tlas_stack, d);
    out.tlas_roots.push_back(root);
    QuadConv qc{out.tlas_pairs, out.tlas_quads};
    out.tlas_quad_roots.push_back(qc.ref_of(root.root_ref, d));
    out.tlas_stack4 = std::max(out.tlas_stack4, d);
  }
  out.tlas_qquads.resize(out.tlas_quads.size());
  for (size_t k = 0; k < out.tlas_quads.size(); k++) out.tlas_qquads[k] = quantize_quad(out.tlas_quads[k]);
  return RDRT_OK;
}

namespace {

// Groups (node, depth) pairs by depth, deepest first.
void by_depth(const std::vector<std::pair<uint32_t, uint32_t>> &nd, std::vector<uint32_t> &order,
              std::vector<uint32_t> &off) {
  uint32_t maxd = 0;
  for (const auto &p : nd) maxd = std::max(maxd, p.second);
  std::vector<uint32_t> count(nd.empty() ? 0 : maxd + 1, 0);
  for (const auto &p : nd) count[p.second]++;
  off.assign(count.size() + 1, 0);
  for (size_t l = 0; l < count.size(); l++) off[l + 1] = off[l] + count[count.size() - 1 - l];
  std::vector<uint32_t> fill(off.begin(), off.end() - 1);
  order.resize(nd.size());
  for (const auto &p : nd) order[fill[count.size() - 1 - p.second]++] = p.first;
}

}  // namespace

RefitLevels refit_levels(const BuiltScene &b) {
  RefitLevels r;
  std::vector<std::pair<uint32_t, uint32_t>> nd, stack;
  // reference forest: every node (leaf <=> hit == miss; right child = miss of the left)
  for (const GeoRec &g : b.geos) {
    stack.assign(1, {g.ref_root, 0u});
    while (!stack.empty()) {
      auto [i, d] = stack.back();
      stack.pop_back();
      nd.push_back({i, d});
      const RefNode &n = b.tri_forest[i];
      if (n.hit_next != n.miss_next) {
        stack.push_back({n.hit_next, d + 1});
        stack.push_back({b.tri_forest[n.hit_next].miss_next, d + 1});
      }
    }
  }
  by_depth(nd, r.forest, r.forest_off);
  nd.clear();
  for (const GeoRec &g : b.geos) {
    if (g.root_ref & kLeafBit) continue;
    stack.assign(1, {g.root_ref, 0u});
    while (!stack.empty()) {
      auto [i, d] = stack.back();
      stack.pop_back();
      nd.push_back({i, d});
      const PairNode &p = b.tri_pairs[i];
      if (!(p.lref & kLeafBit)) stack.push_back({p.lref, d + 1});
      if (!(p.rref & kLeafBit)) stack.push_back({p.rref, d + 1});
    }
  }
  by_depth(nd, r.pairs, r.pairs_off);
  nd.clear();
  for (const GeoRec &g : b.geos) {
    if (g.quad_root & kLeafBit) continue;
    stack.assign(1, {g.quad_root, 0u});
    while (!stack.empty()) {
      auto [i, d] = stack.back();
      stack.pop_back();
      nd.push_back({i, d});
      const QuadNode &q = b.tri_quads[i];
      for (int k = 0; k < 4; k++)
        if (q.ref[k] != kInvalid && !(q.ref[k] & kLeafBit)) stack.push_back({q.ref[k], d + 1});
    }
  }
  by_depth(nd, r.quads, r.quads_off);
  return r;
}

int build_scene(const SceneSrc &src, BuiltScene &out, std::string &err, DeviceTreeBuilder *trees,
                bool device_records) {
  auto t0 = std::chrono::steady_clock::now();
  out = BuiltScene();
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());

  // ---- BLAS: one SAH tree per geometry, built concurrently ----
  struct Job {
    uint32_t blas, geo;
    std::vector<uint32_t> tri_idx;  // 3 per triangle, local vertex ids
    Tree tree;  // the reference's tree (forest, redirect, k_trace_ref)
    Tree fast;  // the fast traversal's tree (child pairs, TriPre order)
    std::vector<Box> boxes;  // triangle boxes, kept for a device build
    bool bad = false;        // an index beyond the geometry's vertices
  };
  std::vector<Job> jobs;
  for (uint32_t b = 0; b < src.blas.size(); b++) {
    if (!src.blas_live[b]) continue;
    for (uint32_t g = 0; g < src.blas[b].size(); g++) jobs.push_back(Job{b, g, {}, {}, {}, {}, false});
  }
  unsigned per_job = jobs.size() >= hw ? 1u : std::max(1u, hw / (unsigned)std::max<size_t>(1, jobs.size()));
  parallel_for(jobs.size(), std::min<unsigned>(hw, (unsigned)jobs.size()), [&](size_t j) {
    Job &job = jobs[j];
    const GeometrySrc &g = src.blas[job.blas][job.geo];
    size_t nv = g.pos.size() / 3;
    if (g.indexed) job.tri_idx = g.idx;
    else {
      job.tri_idx.resize(nv - nv % 3);
      for (size_t v = 0; v < job.tri_idx.size(); v++) job.tri_idx[v] = (uint32_t)v;
    }
    for (uint32_t v : job.tri_idx)
      if (v >= nv) {
        job.bad = true;  // reported below, before any box is read
        return;
      }
    size_t nt = job.tri_idx.size() / 3;
    std::vector<Box> boxes(nt);
    for (size_t t = 0; t < nt; t++) {
      Box bx = box_empty();
      for (int k = 0; k < 3; k++) grow_point(bx, &g.pos[3 * (size_t)job.tri_idx[3 * t + k]]);
      boxes[t] = bx;
    }
    if (trees) {
      job.boxes = std::move(boxes);  // built on the device below
      return;
    }
    job.tree = build_tree(boxes, src.blas_opt, per_job);
    job.fast = build_fast_tree(boxes, FastTreeOptions{}, per_job);
  });
  for (const Job &job : jobs)
    if (job.bad) {
      err = "triangle index out of range of the geometry's vertices";
      return RDRT_ERR_INVALID_ARG;
    }
  if (trees) {
    std::vector<const std::vector<Box> *> geos;
    for (const Job &job : jobs) geos.push_back(&job.boxes);
    std::vector<Tree> ref, fast;
    const int rc = trees->build(geos, src.blas_opt, FastTreeOptions{}, ref, fast, err);
    if (rc != RDRT_OK) return rc;
    for (size_t j = 0; j < jobs.size(); j++) {
      jobs[j].tree = std::move(ref[j]);
      jobs[j].fast = std::move(fast[j]);
      std::vector<Box>().swap(jobs[j].boxes);
    }
  }
  for (const Job &job : jobs) {
    const GeometrySrc &g = src.blas[job.blas][job.geo];
    for (uint32_t v : job.tri_idx)
      if (v >= g.pos.size() / 3) {
        err = "triangle index out of range of the geometry's vertices";
        return RDRT_ERR_INVALID_ARG;
      }
  }

  out.phase_ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();  // trees
  // ---- assemble BLAS arrays in reference order ----
  std::vector<Box> blas_box(src.blas.size(), box_empty());
  std::vector<char> blas_has_box(src.blas.size(), 0);
  std::vector<uint32_t> blas_geo_begin(src.blas.size(), 0), blas_geo_end(src.blas.size(), 0);
  bool any_normals = false, any_material = false, any_uvs = false, any_anyhit = false;
  for (uint32_t b = 0; b < src.blas.size(); b++)
    if (src.blas_live[b])
      for (const GeometrySrc &g : src.blas[b]) {
        any_normals |= !g.normals.empty();
        any_uvs |= !g.uvs.empty();
        any_material |= !g.tri_material.empty();
        any_anyhit |= !g.anyhit.empty();
      }
  // Per geometry, in parallel: forest, child pairs and child quads with
  // local node indices (leaf refs already global), relocated when appended.
  struct Local {
    std::vector<RefNode> forest;
    std::vector<PairNode> pairs;
    std::vector<QuadNode> quads;
    uint32_t root_ref = 0, quad_root = 0, stack = 0, stack4 = 0;
    bool ok = true;
    std::string err;
  };
  std::vector<Local> local(jobs.size());
  {
    std::vector<uint32_t> prim_start(jobs.size(), 0);
    uint64_t acc = 0;
    for (size_t j = 0; j < jobs.size(); j++) {
      prim_start[j] = (uint32_t)acc;
      acc += jobs[j].tri_idx.size() / 3;
    }
    parallel_for(jobs.size(), hw, [&](size_t j) {
      Local &l = local[j];
      append_forest(jobs[j].tree, prim_start[j], l.forest);
      PairConv pc{jobs[j].fast, prim_start[j], l.pairs, l.err};
      l.root_ref = pc.ref_of(0, l.stack);
      l.ok = pc.ok;
      QuadConv qc{l.pairs, l.quads};
      l.quad_root = qc.ref_of(l.root_ref, l.stack4);
    });
  }
  auto inner = [](uint32_t r) { return r != kInvalid && !(r & kLeafBit); };
  // Every geometry's slice of the output arrays, in reference order (one
  // serial pass over the geometries), then the copies and relocations of all
  // geometries in parallel.
  struct Slice {
    uint32_t blas, geo;
    size_t f, p, q, t, v;  // first forest node, pair, quad, triangle, vertex
  };
  std::vector<Slice> slices(jobs.size());
  {
    size_t nf = 0, np = 0, nq = 0, nt = 0, nv = 0;
    size_t ji = 0;
    for (uint32_t b = 0; b < src.blas.size(); b++)
      if (src.blas_live[b])
        for (uint32_t gi = 0; gi < src.blas[b].size(); gi++, ji++) {
          if (!local[ji].ok) {
            err = local[ji].err;
            return RDRT_ERR_CAPACITY;
          }
          slices[ji] = Slice{b, gi, nf, np, nq, nt, nv};
          nf += local[ji].forest.size();
          np += local[ji].pairs.size();
          nq += local[ji].quads.size();
          nt += jobs[ji].tri_idx.size() / 3;
          nv += src.blas[b][gi].pos.size() / 3;
        }
    if (nt > 0xFFFFFFFFull / 3 || nv > 0xFFFFFFFFull || nf >= kInvalid || np >= kLeafBit || nq >= kLeafBit) {
      err = "scene too large for 32-bit node, triangle or vertex indices";
      return RDRT_ERR_CAPACITY;
    }
    out.tri_forest.resize(nf);
    out.tri_pairs.resize(np);
    out.tri_quads.resize(nq);
    out.redirect.resize(nt);
    out.tri_ids.resize(nt);
    out.indices.resize(3 * nt);
    out.vertices.resize(3 * nv);
    if (any_normals) out.normals.resize(3 * nv);
    if (any_uvs) out.uvs.resize(2 * nv);
    if (any_material) out.tri_material.resize(nt);
    if (any_anyhit) out.anyhit.resize(nt);
  }
  std::vector<uint32_t> geo_max_mat(jobs.size(), 0);
  parallel_for(jobs.size(), hw, [&](size_t ji) {
    const Job &job = jobs[ji];
    const Slice &sl = slices[ji];
    const GeometrySrc &g = src.blas[sl.blas][sl.geo];
    const size_t nv = g.pos.size() / 3, nt = job.tri_idx.size() / 3;
    const uint32_t primitive_start = (uint32_t)sl.t, vertex_start = (uint32_t)sl.v;
    std::copy(g.pos.begin(), g.pos.begin() + 3 * nv, out.vertices.begin() + 3 * sl.v);
    for (size_t k = 0; k < job.tree.order.size(); k++) out.redirect[sl.t + k] = primitive_start + job.tree.order[k];
    for (size_t k = 0; k < job.fast.order.size(); k++) out.tri_ids[sl.t + k] = primitive_start + job.fast.order[k];
    for (size_t k = 0; k < job.tri_idx.size(); k++) out.indices[3 * sl.t + k] = vertex_start + job.tri_idx[k];
    if (any_normals) {
      if (g.normals.size() == 3 * nv) std::copy(g.normals.begin(), g.normals.end(), out.normals.begin() + 3 * sl.v);
      else std::fill(out.normals.begin() + 3 * sl.v, out.normals.begin() + 3 * (sl.v + nv), 0.0f);  // geometric normal
    }
    if (any_uvs) {
      if (g.uvs.size() == 2 * nv) std::copy(g.uvs.begin(), g.uvs.end(), out.uvs.begin() + 2 * sl.v);
      else std::fill(out.uvs.begin() + 2 * sl.v, out.uvs.begin() + 2 * (sl.v + nv), 0.0f);
    }
    if (any_material) {
      uint32_t mx = 0;
      for (size_t t = 0; t < nt; t++) {
        const uint32_t m = t < g.tri_material.size() ? g.tri_material[t] : 0u;
        out.tri_material[sl.t + t] = m;
        mx = std::max(mx, m);
      }
      geo_max_mat[ji] = mx;
    }
    if (any_anyhit)
      for (size_t t = 0; t < nt; t++) out.anyhit[sl.t + t] = t < g.anyhit.size() ? g.anyhit[t] : (uint8_t)RDRT_ANYHIT_ACCEPT;
    Local &l = local[ji];
    const uint32_t fb = (uint32_t)sl.f, pb = (uint32_t)sl.p, qb = (uint32_t)sl.q;
    for (size_t k = 0; k < l.forest.size(); k++) {
      RefNode n = l.forest[k];
      if (n.hit_next != kInvalid) n.hit_next += fb;
      if (n.miss_next != kInvalid) n.miss_next += fb;
      out.tri_forest[sl.f + k] = n;
    }
    for (size_t k = 0; k < l.pairs.size(); k++) {
      PairNode pn = l.pairs[k];
      if (inner(pn.lref)) pn.lref += pb;
      if (inner(pn.rref)) pn.rref += pb;
      out.tri_pairs[sl.p + k] = pn;
    }
    for (size_t k = 0; k < l.quads.size(); k++) {
      QuadNode qn = l.quads[k];
      for (int c = 0; c < 4; c++)
        if (inner(qn.ref[c])) qn.ref[c] += qb;
      out.tri_quads[sl.q + k] = qn;
    }
    std::vector<RefNode>().swap(l.forest);
    std::vector<PairNode>().swap(l.pairs);
    std::vector<QuadNode>().swap(l.quads);
  });
  {
    size_t ji = 0;
    for (uint32_t b = 0; b < src.blas.size(); b++) {
      blas_geo_begin[b] = (uint32_t)out.geos.size();
      if (!src.blas_live[b]) {
        blas_geo_end[b] = blas_geo_begin[b];
        continue;
      }
      for (uint32_t gi = 0; gi < src.blas[b].size(); gi++, ji++) {
        const Job &job = jobs[ji];
        const GeometrySrc &g = src.blas[b][gi];
        const Slice &sl = slices[ji];
        const Local &l = local[ji];
        out.geo_max_material.push_back(geo_max_mat[ji]);
        GeoRec gr;
        std::memset(&gr, 0, sizeof(gr));
        put_box(gr.mn, gr.mx, job.tree.nodes[0].box);
        gr.primitive_start = (uint32_t)sl.t;
        gr.geometry_idx = gi;
        gr.flags = g.flags;
        gr.ref_root = (uint32_t)sl.f;
        gr.root_ref = inner(l.root_ref) ? l.root_ref + (uint32_t)sl.p : l.root_ref;
        gr.quad_root = inner(l.quad_root) ? l.quad_root + (uint32_t)sl.q : l.quad_root;
        out.blas_stack = std::max(out.blas_stack, l.stack);
        out.blas_stack4 = std::max(out.blas_stack4, l.stack4);
        out.max_blas_depth = std::max(out.max_blas_depth, job.tree.depth);
        out.geos.push_back(gr);
        grow_box(blas_box[b], job.tree.nodes[0].box);
        blas_has_box[b] = 1;
      }
      blas_geo_end[b] = (uint32_t)out.geos.size();
    }
  }
  out.phase_ms[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();  // forest, pairs, quads
  // triangle pre-pass, in the fast tree's leaf order (here, or on the
  // device at commit)
  out.device_records = device_records;
  if (!device_records) out.tris.resize(out.tri_ids.size());
  parallel_for((out.tris.size() + 65535) / 65536, hw, [&](size_t c) {
    size_t e = std::min(out.tris.size(), (c + 1) * 65536);
    for (size_t k = c * 65536; k < e; k++) {
      size_t t = out.tri_ids[k];
      const float *v0 = &out.vertices[3 * (size_t)out.indices[3 * t]];
      const float *v1 = &out.vertices[3 * (size_t)out.indices[3 * t + 1]];
      const float *v2 = &out.vertices[3 * (size_t)out.indices[3 * t + 2]];
      out.tris[k] = pre_triangle(v0, v1, v2);
    }
  });

  out.phase_ms[2] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();  // triangle records
  // quantized twins of the child-quad nodes (here, or on the device)
  if (!device_records) out.tri_qquads.resize(out.tri_quads.size());
  parallel_for((out.tri_qquads.size() + 65535) / 65536, hw, [&](size_t c) {
    const size_t e = std::min(out.tri_qquads.size(), (c + 1) * 65536);
    for (size_t k = c * 65536; k < e; k++) out.tri_qquads[k] = quantize_quad(out.tri_quads[k]);
  });
  out.blas_box = blas_box;
  out.blas_has_box = blas_has_box;
  out.blas_geo_begin = blas_geo_begin;
  out.blas_geo_end = blas_geo_end;
  {
    out.phase_ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();  // quantized quads
    const int rc = build_tlas_layer(src, out, err);
    if (rc != RDRT_OK) return rc;
  }
  out.binding = src.binding;
  if (!validate_refs(out, err)) return RDRT_ERR_STATE;
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return RDRT_OK;
}

}  // namespace rdrt

namespace rdrt {

// ---- host picking: intersect_nearest_bvh (content/mesh/core/src/feature/
// bvh.rs:57-86) with the host geometry tests of math/geometry/src/dimension3/
// intersection.rs -- single-rounding FP32 in the reference operation order.
namespace {

struct P3 {
  float x, y, z;
};
inline P3 psub(P3 a, P3 b) { return P3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float pdot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 pcross(P3 a, P3 b) { return P3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Ray3 x Box3 as a bool (intersection.rs:128-196 and its is_some() at :203-208).
bool pick_box(P3 o, P3 d, const float *mn, const float *mx) {
  float t_max, t_min, ty_min, ty_max, tz_min, tz_max;
  const float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
  if (ix >= 0.0f) {
    t_min = (mn[0] - o.x) * ix;
    t_max = (mx[0] - o.x) * ix;
  } else {
    t_min = (mx[0] - o.x) * ix;
    t_max = (mn[0] - o.x) * ix;
  }
  if (iy >= 0.0f) {
    ty_min = (mn[1] - o.y) * iy;
    ty_max = (mx[1] - o.y) * iy;
  } else {
    ty_min = (mx[1] - o.y) * iy;
    ty_max = (mn[1] - o.y) * iy;
  }
  if ((t_min > ty_max) || (ty_min > t_max)) return false;
  if (ty_min > t_min || std::isnan(t_min)) t_min = ty_min;
  if (ty_max < t_max || std::isnan(t_max)) t_max = ty_max;
  if (iz >= 0.0f) {
    tz_min = (mn[2] - o.z) * iz;
    tz_max = (mx[2] - o.z) * iz;
  } else {
    tz_min = (mx[2] - o.z) * iz;
    tz_max = (mn[2] - o.z) * iz;
  }
  if ((t_min > tz_max) || (tz_min > t_max)) return false;
  if (tz_min > t_min || std::isnan(t_min)) t_min = tz_min;
  if (tz_max < t_max || std::isnan(t_max)) t_max = tz_max;
  (void)t_min;
  return !(t_max < 0.0f);
}

// Ray3 x Triangle with a face side (intersection.rs:3-76, from GTE's
// IntrRay3Triangle3; Back takes the flipped triangle, triangle.rs:76-82).  Returns the distance, or NaN when there is no hit.
float pick_triangle(P3 o, P3 d, P3 a, P3 b, P3 c, uint32_t side) {
  if (side == RDRT_FACE_BACK) std::swap(a, c);  // triangle.rs:76-82 flip
  const bool culling = side != RDRT_FACE_DOUBLE;
  const P3 e1 = psub(b, a), e2 = psub(c, a), n = pcross(e1, e2);
  float ddn = pdot(d, n), sign;
  if (ddn > 0.0f) {
    if (culling) return NAN;
    sign = 1.0f;
  } else if (ddn < 0.0f) {
    sign = -1.0f;
    ddn = -ddn;
  } else {
    return NAN;
  }
  const P3 diff = psub(o, a);
  const float ddqxe2 = sign * pdot(d, pcross(diff, e2));
  if (ddqxe2 < 0.0f) return NAN;
  const float dde1xq = sign * pdot(d, pcross(e1, diff));
  if (dde1xq < 0.0f) return NAN;
  if (ddqxe2 + dde1xq > ddn) return NAN;
  const float qdn = -sign * pdot(diff, n);
  if (qdn < 0.0f) return NAN;
  return qdn / ddn;
}

// Every hit of one ray in the reference's visit order: visit(triangle, t).
template <class V>
void pick_walk(const BuiltScene &b, const GeoRec &g, P3 o, P3 d, uint32_t side, std::vector<uint32_t> &stack,
               V visit) {
  // traverse_branch_leaf_visitor (content/space/src/bvh/mod.rs:94-114 on
  // utility/abstract-tree/src/lib.rs:33-51): a stack; a branch whose box
  // the ray hits pushes its children left then right (mod.rs:38-45), so
  // the right one is visited first; a leaf is visited without a box test
  stack.assign(1, g.ref_root);
  while (!stack.empty()) {
    const uint32_t i = stack.back();
    stack.pop_back();
    const RefNode &nd = b.tri_forest[i];
    if (nd.hit_next != nd.miss_next) {
      if (pick_box(o, d, nd.mn, nd.mx)) {
        stack.push_back(nd.hit_next);
        stack.push_back(b.tri_forest[nd.hit_next].miss_next);
      }
      continue;
    }
    for (uint32_t s = nd.begin; s < nd.end; s++) {
      const uint32_t tri = b.redirect[s];
      const float *va = &b.vertices[3 * (size_t)b.indices[3 * (size_t)tri]];
      const float *vb = &b.vertices[3 * (size_t)b.indices[3 * (size_t)tri + 1]];
      const float *vc = &b.vertices[3 * (size_t)b.indices[3 * (size_t)tri + 2]];
      const float t =
          pick_triangle(o, d, P3{va[0], va[1], va[2]}, P3{vb[0], vb[1], vb[2]}, P3{vc[0], vc[1], vc[2]}, side);
      if (!std::isnan(t)) visit(tri, t);
    }
  }
}

rdrt_pick_hit make_pick_hit(uint32_t prim, P3 o, P3 d, float t) {
  rdrt_pick_hit h;
  std::memset(&h, 0, sizeof(h));
  h.primitive = prim;
  h.distance = t;
  h.position[0] = o.x + d.x * t;  // hyper_ray.rs:11-13
  h.position[1] = o.y + d.y * t;
  h.position[2] = o.z + d.z * t;
  return h;
}

int pick_check(const BuiltScene &b, uint32_t geo, uint32_t side, std::string &err) {
  if (geo >= b.geos.size()) {
    err = "unknown geometry";
    return RDRT_ERR_INVALID_ARG;
  }
  if (side > RDRT_FACE_DOUBLE) {
    err = "face side is RDRT_FACE_FRONT, RDRT_FACE_BACK or RDRT_FACE_DOUBLE";
    return RDRT_ERR_INVALID_ARG;
  }
  return RDRT_OK;
}

}  // namespace

int pick_all(const BuiltScene &b, uint32_t geo, const float *ray, uint32_t side, rdrt_pick_hit *out, uint32_t capacity,
             uint32_t *count, std::string &err) {
  if (int rc = pick_check(b, geo, side, err)) return rc;
  const P3 o{ray[0], ray[1], ray[2]}, d{ray[3], ray[4], ray[5]};
  std::vector<uint32_t> stack;
  uint32_t k = 0;
  const uint32_t prim0 = b.geos[geo].primitive_start;
  pick_walk(b, b.geos[geo], o, d, side, stack, [&](uint32_t tri, float t) {
    if (k < capacity) out[k] = make_pick_hit(tri - prim0, o, d, t);
    k++;
  });
  *count = k;
  return RDRT_OK;
}

int pick_nearest(const BuiltScene &b, uint32_t geo, const float *rays, uint32_t n, uint32_t side, rdrt_pick_hit *out,
                 std::string &err) {
  if (int rc = pick_check(b, geo, side, err)) return rc;
  const GeoRec &g = b.geos[geo];
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  parallel_for((n + 255) / 256, hw, [&](size_t chunk) {
    std::vector<uint32_t> stack;
    const uint32_t r1 = std::min<uint32_t>(n, (uint32_t)(chunk + 1) * 256);
    for (uint32_t r = (uint32_t)chunk * 256; r < r1; r++) {
      const float *q = rays + 6 * (size_t)r;
      const P3 o{q[0], q[1], q[2]}, d{q[3], q[4], q[5]};
      bool found = false;
      uint32_t best = 0;
      float best_t = 0.0f;
      pick_walk(b, g, o, d, side, stack, [&](uint32_t tri, float t) {
        if (!found || t < best_t) {  // refresh: strictly nearer (intersect_util.rs:15-17,52-61)
          found = true;
          best = tri;
          best_t = t;
        }
      });
      if (found) {
        out[r] = make_pick_hit(best - g.primitive_start, o, d, best_t);
      } else {
        rdrt_pick_hit h;
        std::memset(&h, 0, sizeof(h));
        h.primitive = RDRT_MISS;
        out[r] = h;
      }
    }
  });
  return RDRT_OK;
}

}  // namespace rdrt
