// CPU restatement of the reference's analytic-marching engine.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU engine
// and the timed CPU baseline; the product path never links or calls it.
//
// It restates, in sequential C++ with deterministic tie-breaks, what the
// reference computes on the GPU:
//   * visited set: FNV-1a slot / FNV-1 tag table, probe step 1543, tag-only
//     membership, empty marker = tag of the all-zero key
//     (backend/inc/hash.h:15-67, backend/src/states.cu:10-79,
//      backend/inc/states.h:209-216);
//   * LIFO frontier popped in batches of 1024 (backend/src/cuam_kernel.cu:53-98,
//     backend/inc/states.h:220-282);
//   * constraint planes through the masked affine chain, sign flip, skips and
//     extra planes (backend/inc/process.h:18-193);
//   * distance order, descending and stable (backend/inc/process.h:208-264);
//   * seed start search over constraint pairs (backend/inc/kernel.h:609-856,
//     backend/src/kernel.cu:80-197);
//   * the pivot walk with its accept/close/negate rules
//     (backend/inc/kernel.h:861-1320, backend/src/kernel.cu:203-299);
//   * orientation (backend/inc/kernel.h:1325-1385);
//   * neighbour emission (backend/inc/kernel.h:1390-1506).
//
// Where the reference races (several feasible candidates in one round, the
// order of atomic appends), this file takes the lowest candidate position /
// the lowest thread index.  Arithmetic is written with explicit fma() and
// compiled with -ffp-contract=off so that the GPU's FP64 path can perform the
// identical sequence of IEEE operations.
//
// Besides the faithful LIFO order the library offers a breadth-first order
// (one level at a time, candidates deduplicated in (region, edge) order) that
// the GPU engine follows; the two orders visit the same regions on generic
// inputs, which the tests check.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <unordered_set>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "am_oracle.h"

namespace {

// -------------------------------------------------------------------------
// constants (backend/inc/macro.h:16-44)
// -------------------------------------------------------------------------
template <typename T> struct K;
template <> struct K<double> {
  static constexpr double EPS = 1e-20;        // _FP_EPS_DOUBLE
  static constexpr double SAME_EPS = 3e-14;   // _FP_SAME_EPS_DOUBLE
  static constexpr double POINTHASH_MUL = 1e9;
};
template <> struct K<float> {
  static constexpr float EPS = 1e-12f;        // _FP_EPS_FLOAT
  static constexpr float SAME_EPS = 3e-7f;    // _FP_SAME_EPS_FLOAT
  static constexpr float POINTHASH_MUL = 1e6f;
};
constexpr int BATCH = 1024;           // _VAR_BATCH_SIZE_MAX
constexpr uint32_t TABLE_SIZE = 1u << 25;  // _STATES_TABLE_SIZE
constexpr uint32_t SKIP = 1543;       // _STATES_SKIP_LENGTH
constexpr int XMAX_CAP = 812;         // backend/inc/kernel.h:819

inline double fmaT(double a, double b, double c) { return std::fma(a, b, c); }
inline float fmaT(float a, float b, float c) { return std::fmaf(a, b, c); }

// FNV-1a (slot) and FNV-1 (tag), 32 bit (backend/src/states.cu:11-26).
uint32_t fnv1a(const uint8_t* d, int n) {
  uint32_t h = 0x811C9DC5u;
  while (n--) h = (h ^ *d++) * 0x01000193u;
  return h;
}
uint32_t fnv1(const uint8_t* d, int n) {
  uint32_t h = 0x811C9DC5u;
  while (n--) h = (h * 0x01000193u) ^ *d++;
  return h;
}

// Pair enumeration x -> (i, j), i > j (backend/inc/kernel.h:609-624).  The
// reference evaluates a = ceil(sqrt(2k + 2.25) - 1.5) in float; that is the
// smallest a >= 0 with a(a + 3) >= 2k, computed here in integers.
void get_i_j(int x, int ij_max, int& i, int& j) {
  int k = x / 2;
  int a = 0;
  while (a * (a + 3) < 2 * k) ++a;
  int r = x - a * (a + 1);
  int A = 2 * a + r / (a + 1);
  int B = r % (a + 1);
  i = (A + 1) / 2 + B + 1;
  j = A + 1 - i;
  i = i < ij_max ? i : ij_max - 1;
  j = j < ij_max ? j : ij_max - 1;
}

// -------------------------------------------------------------------------
// network
// -------------------------------------------------------------------------
struct Skip { int from; int tm; };

struct Net {
  std::vector<int> nodes;             // [3, n1..nD, 1]
  int D = 0;                          // hidden layers
  int L = 0, E = 0, C = 0, Wd = 0;    // neurons, extras, constraints, mask words
  std::vector<int> off;               // off[k] = first neuron of hidden node k+1; off[D] = L
  std::vector<std::vector<double>> W, b;   // per linear layer, (out, in) row-major
  std::vector<std::vector<Skip>> skips;    // per linear layer index (1..D-1)
  std::vector<std::vector<double>> tm;     // transform matrices (row-major)
  std::vector<int> tm_h, tm_w;
  std::vector<double> ew, eb;         // extra planes
};

inline bool bit(const uint32_t* m, int g) { return (m[g >> 5] >> (g & 31)) & 1u; }

// Constraint planes of one region, T precision (backend/inc/process.h:18-193).
// A[g*4 + {0,1,2,3}] = (a_x, a_y, a_z, c) of constraint g; eq likewise.
// Canonical order: for each output neuron j, the accumulators start at
// (0, 0, 0, b_j) and take fma(W[j][k], plane_k, acc) over active k ascending.
template <typename T>
void compute_planes(const Net& n, const std::vector<std::vector<T>>& Wt,
                    const std::vector<std::vector<T>>& bt,
                    const std::vector<std::vector<T>>& tmt,
                    const uint32_t* mask, T iso, T* A, T* eq) {
  const int D = n.D;
  // hidden node 1 = rows of W0
  for (int j = 0; j < n.nodes[1]; ++j) {
    A[j * 4 + 0] = Wt[0][j * 3 + 0];
    A[j * 4 + 1] = Wt[0][j * 3 + 1];
    A[j * 4 + 2] = Wt[0][j * 3 + 2];
    A[j * 4 + 3] = bt[0][j];
  }
  for (int li = 1; li <= D; ++li) {   // linear layer li: node li -> node li+1
    const int nin = n.nodes[li], nout = n.nodes[li + 1];
    const int in0 = n.off[li - 1];
    const T* Wl = Wt[li].data();
    for (int j = 0; j < nout; ++j) {
      // +0 turns a -0 bias into +0, so that skipping an inactive input and
      // adding fma(w, +0, acc) give the same bits (no accumulator is ever -0)
      T ax = 0, ay = 0, az = 0, c = bt[li][j] + T(0);
      for (int k = 0; k < nin; ++k) {
        if (!bit(mask, in0 + k)) continue;
        const T w = Wl[(size_t)j * nin + k];
        const T* p = A + (size_t)(in0 + k) * 4;
        ax = fmaT(w, p[0], ax);
        ay = fmaT(w, p[1], ay);
        az = fmaT(w, p[2], az);
        c = fmaT(w, p[3], c);
      }
      T* dst = (li < D) ? A + (size_t)(n.off[li] + j) * 4 : eq;
      dst[0] = ax; dst[1] = ay; dst[2] = az; dst[3] = c;
    }
    if (li < D) {
      // additive skips into node li+1 (backend/inc/process.h:73-146)
      T* dst = A + (size_t)n.off[li] * 4;
      for (const Skip& s : n.skips[li]) {
        const bool ident = n.tm_h[s.tm] == 0 && n.tm_w[s.tm] == 0;
        if (s.from == 0) {
          if (ident) {
            for (int j = 0; j < 3 && j < nout; ++j) dst[j * 4 + j] = dst[j * 4 + j] + T(1);
          } else {
            const T* Tm = tmt[s.tm].data();
            const int h = n.tm_h[s.tm];
            for (int j = 0; j < h; ++j)
              for (int q = 0; q < 3; ++q) dst[j * 4 + q] = dst[j * 4 + q] + Tm[j * 3 + q];
          }
        } else {
          const int os = n.off[s.from - 1], dec = n.nodes[s.from];
          const T* src = A + (size_t)os * 4;
          if (ident) {
            for (int j = 0; j < dec; ++j) {
              if (!bit(mask, os + j)) continue;
              for (int q = 0; q < 4; ++q) dst[j * 4 + q] = dst[j * 4 + q] + src[j * 4 + q];
            }
          } else {
            const T* Tm = tmt[s.tm].data();
            const int h = n.tm_h[s.tm], w = n.tm_w[s.tm];
            for (int j = 0; j < h; ++j) {
              T acc[4] = {dst[j * 4 + 0], dst[j * 4 + 1], dst[j * 4 + 2], dst[j * 4 + 3]};
              for (int k = 0; k < w; ++k) {
                if (!bit(mask, os + k)) continue;
                const T t = Tm[(size_t)j * w + k];
                for (int q = 0; q < 4; ++q) acc[q] = fmaT(t, src[k * 4 + q], acc[q]);
              }
              for (int q = 0; q < 4; ++q) dst[j * 4 + q] = acc[q];
            }
          }
        }
      }
    }
  }
  eq[3] = eq[3] - iso;                 // backend/inc/process.h:198-203
  for (int g = 0; g < n.L; ++g)        // sign flip (backend/inc/process.h:181-186)
    if (bit(mask, g))
      for (int q = 0; q < 4; ++q) A[g * 4 + q] = -A[g * 4 + q];
  for (int e = 0; e < n.E; ++e) {      // extra planes (backend/inc/process.h:189-192)
    A[(n.L + e) * 4 + 0] = T(n.ew[e * 3 + 0]);
    A[(n.L + e) * 4 + 1] = T(n.ew[e * 3 + 1]);
    A[(n.L + e) * 4 + 2] = T(n.ew[e * 3 + 2]);
    A[(n.L + e) * 4 + 3] = T(n.eb[e]);
  }
}

// s = a.v + c and |a|^2 with a fixed fma order.
template <typename T> inline T plane_eval(const T* a, const T* v) {
  return fmaT(a[0], v[0], fmaT(a[1], v[1], fmaT(a[2], v[2], a[3])));
}
template <typename T> inline T plane_nn(const T* a) {
  return fmaT(a[0], a[0], fmaT(a[1], a[1], a[2] * a[2]));
}
// (s / |a| > EPS) <=> s > 0 and s^2 > EPS^2 |a|^2  (backend/inc/kernel.h:1087-1096)
template <typename T> inline bool violates(const T* a, const T* v) {
  const T s = plane_eval(a, v);
  return s > T(0) && s * s > (K<T>::EPS * K<T>::EPS) * plane_nn(a);
}

// Cramer's rule for A x = -b (backend/inc/kernel.h:511,555-573).
#define DET3(a0, a1, a2, a3, a4, a5, a6, a7, a8) \
  ((a0) * ((a4) * (a8) - (a5) * (a7)) - (a1) * ((a3) * (a8) - (a5) * (a6)) + (a2) * ((a3) * (a7) - (a4) * (a6)))
template <typename T>
inline void solve3(const T* p1, const T* p2, const T* p3, T* x) {
  const T A0 = p1[0], A1 = p1[1], A2 = p1[2], A3 = p2[0], A4 = p2[1], A5 = p2[2], A6 = p3[0], A7 = p3[1], A8 = p3[2];
  const T b0 = -p1[3], b1 = -p2[3], b2 = -p3[3];
  const T d0 = DET3(A0, A1, A2, A3, A4, A5, A6, A7, A8);
  const T d1 = DET3(b0, A1, A2, b1, A4, A5, b2, A7, A8);
  const T d2 = DET3(A0, b0, A2, A3, b1, A5, A6, b2, A8);
  const T d3 = DET3(A0, A1, b0, A3, A4, b1, A6, A7, b2);
  x[0] = d1 / d0;
  x[1] = d2 / d0;
  x[2] = d3 / d0;
}

#define DOT3(a, b) ((a)[0] * (b)[0] + (a)[1] * (b)[1] + (a)[2] * (b)[2])
// Turn test of the projected normals (backend/inc/kernel.h:1103-1116).
template <typename T>
inline bool pivot_valid(const T* neq, const T* n1, const T* n2, const T* n3) {
  const T A12 = DOT3(n1, n2), A23 = DOT3(n2, n3), A13 = DOT3(n1, n3), A22 = DOT3(n2, n2);
  const T C1 = DOT3(n1, neq), C2 = DOT3(n2, neq), C3 = DOT3(n3, neq), B = DOT3(neq, neq);
  return ((A12 * A23 - A22 * A13) - (A12 * C2 * C3 + A23 * C1 * C2 - A13 * C2 * C2 - A22 * C1 * C3) / B) > 0;
}

template <typename T>
inline bool same_vertex(const T* a, const T* b) {  // backend/inc/kernel.h:1151-1155
  return std::fabs(a[0] - b[0]) + std::fabs(a[1] - b[1]) + std::fabs(a[2] - b[2]) < K<T>::SAME_EPS;
}

// -------------------------------------------------------------------------
// one region
// -------------------------------------------------------------------------
template <typename T>
struct RegionOut {
  int start_in = -1;     // start index the region was queued with
  int start = -1;        // start plane after the seed search (-1: none)
  int status = 0;        // AMO_ST_*
  int reversed = 0;
  std::vector<int> tab;
  std::vector<T> verts;  // nv * 3 (final orientation)
  std::vector<int> vpair;// nv * 2: (accepted plane, current start) of each vertex
};

// Stable descending order of the distance keys (backend/inc/process.h:244-264).
// key = s|s|/|a|^2 orders like s/|a|; |a| = 0 gives +-inf.
template <typename T>
void distance_order(const T* A, int C, const T* p, std::vector<int>& order, std::vector<T>& key) {
  key.resize(C);
  order.resize(C);
  for (int g = 0; g < C; ++g) {
    const T* a = A + (size_t)g * 4;
    const T s = plane_eval(a, p);
    const T nn = plane_nn(a);
    key[g] = nn > T(0) ? s * std::fabs(s) / nn
                       : (s > T(0) ? INFINITY : -INFINITY);
    if (key[g] != key[g]) key[g] = -INFINITY;  // NaN sorts last
    order[g] = g;
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] > key[y]; });
}

template <typename T>
bool feasible_at(const T* A, int C, const T* v, int ex1, int ex2) {
  for (int g = 0; g < C; ++g) {
    if (g == ex1 || g == ex2) continue;
    if (violates(A + (size_t)g * 4, v)) return false;
  }
  return true;
}

template <typename T>
void process_region(const Net& n, const T* A, const T* eq, const T* p, int start_in, int vm, bool flip,
                    RegionOut<T>& out) {
  const int C = n.C;
  out.start_in = start_in;
  out.tab.clear();
  out.verts.clear();
  out.vpair.clear();
  out.reversed = 0;
  std::vector<int> order;
  std::vector<T> key;
  distance_order(A, C, p, order, key);

  int start = start_in;
  if (start == -1) {
    // seed start search (backend/inc/kernel.h:799-856)
    const int xmax = std::min((C - 1) * (C - 2), XMAX_CAP);
    for (int x = 0; x <= xmax; ++x) {
      int i, j;
      get_i_j(x, C, i, j);
      const int ri = order[i], rj = order[j];
      T v[3];
      solve3(A + (size_t)ri * 4, A + (size_t)rj * 4, eq, v);
      if (feasible_at(A, C, v, ri, rj)) { start = ri; break; }
    }
    if (start == -1) { out.start = -1; out.status = AMO_ST_NO_START; return; }
  }
  out.start = start;

  // pivot walk (backend/inc/kernel.h:861-1320, backend/src/kernel.cu:203-299)
  std::vector<int>& tab = out.tab;
  tab.push_back(start);
  int cur = start;
  auto closed = [&]() {
    const int len = (int)tab.size();
    return (len >= 4 && tab[0] == tab[len - 1]) || len >= vm + 1;
  };
  auto in_tab = [&](int r) {
    for (int t : tab) if ((t >= 0 ? t : -t - 1) == r) return true;
    return false;
  };
  out.status = AMO_ST_OK;
  while (!closed()) {
    if (cur < 0 || cur >= C) { out.status = AMO_ST_BAD_START; break; }  // Q2 position out of range
    int win = -1;
    T vwin[3] = {0, 0, 0};
    for (int w = 0; w < C; ++w) {
      const int r = order[w];
      const int len = (int)tab.size();
      if (!((r == tab[0] && len >= 3) || !in_tab(r))) continue;
      T v[3];
      solve3(A + (size_t)r * 4, A + (size_t)cur * 4, eq, v);
      if (feasible_at(A, C, v, cur, r)) { win = r; vwin[0] = v[0]; vwin[1] = v[1]; vwin[2] = v[2]; break; }
    }
    if (win < 0) { out.status = AMO_ST_GAVE_UP; break; }
    // the last two positive entries (backend/inc/kernel.h:1119-1149)
    int pl = -1, pr = -1, cnt = 0;
    for (int idx = (int)tab.size() - 1; idx >= 0 && cnt < 2; --idx) {
      if (tab[idx] >= 0) { if (cnt == 0) pl = idx; else pr = idx; ++cnt; }
    }
    int valid = -1;
    if (cnt == 2) valid = pivot_valid(eq, A + (size_t)tab[pr] * 4, A + (size_t)tab[pl] * 4, A + (size_t)win * 4) ? 1 : 0;
    const int nv = (int)out.verts.size() / 3;
    if (valid == 0) {
      tab.push_back(-win - 1);
      cur = pl + 1;   // Q2: the reference stores a 1-based tab position here
    } else if (nv > 0 && same_vertex(&out.verts[(nv - 1) * 3], vwin)) {
      tab.push_back(win);
      if (pl >= 0) tab[pl] = -tab[pl] - 1;
      cur = win;
    } else {
      out.verts.push_back(vwin[0]);
      out.verts.push_back(vwin[1]);
      out.verts.push_back(vwin[2]);
      out.vpair.push_back(win);
      out.vpair.push_back(cur);
      tab.push_back(win);
      cur = win;
    }
  }
  // orientation (backend/inc/kernel.h:1325-1371)
  const int nv = (int)out.verts.size() / 3;
  if (nv >= 3) {
    const T* v = out.verts.data();
    const T ax = v[3] - v[0], ay = v[4] - v[1], az = v[5] - v[2];
    const T bx = v[6] - v[3], by = v[7] - v[4], bz = v[8] - v[5];
    const T cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    const T dir = cx * eq[0] + cy * eq[1] + cz * eq[2];
    if ((dir < 0) ^ flip) {
      out.reversed = 1;
      for (int i = 0; i < nv / 2; ++i) {
        for (int q = 0; q < 3; ++q) std::swap(out.verts[i * 3 + q], out.verts[(nv - 1 - i) * 3 + q]);
        std::swap(out.vpair[i * 2 + 0], out.vpair[(nv - 1 - i) * 2 + 0]);
        std::swap(out.vpair[i * 2 + 1], out.vpair[(nv - 1 - i) * 2 + 1]);
      }
    }
  }
}

// Neighbour candidates of one processed region (backend/inc/kernel.h:1390-1474):
// every tab entry but the last that names a hidden neuron; the seed point is
// the midpoint of two vertices picked by counting positive entries, read from
// the vertex list after orientation, as the reference does.
template <typename T>
struct Cand { int edge_pos; int t; T mid[3]; };

template <typename T>
void region_candidates(const Net& n, const RegionOut<T>& r, std::vector<Cand<T>>& cands) {
  cands.clear();
  const int ntab = (int)r.tab.size();
  const int nv = (int)r.verts.size() / 3;
  for (int m = 0; m + 1 < ntab; ++m) {
    const int t = r.tab[m];
    if (!(t >= 0 && t < n.L)) continue;
    const int m_ = (m == 0) ? ntab - 1 : m;
    int cnt = -1, v1 = -1;
    for (int i = 1; i < ntab; ++i) {
      if (r.tab[i] >= 0) ++cnt;
      if (i == m_) v1 = cnt;
    }
    Cand<T> c;
    c.edge_pos = m;
    c.t = t;
    if (nv <= 0 || v1 < 0) {
      c.mid[0] = c.mid[1] = c.mid[2] = T(0);
    } else {
      const int v2 = (v1 + 1) % nv;
      for (int q = 0; q < 3; ++q) c.mid[q] = (r.verts[v1 * 3 + q] + r.verts[v2 * 3 + q]) / T(2);
    }
    cands.push_back(c);
  }
}

struct MaskHash {
  size_t operator()(const std::string& s) const { return std::hash<std::string>()(s); }
};

// -------------------------------------------------------------------------
// run state
// -------------------------------------------------------------------------
struct Run {
  Net net;
  int precision = 0;   // 0 fp64, 1 fp32
  int traversal = 0;   // 0 lifo, 1 bfs
  int vm = 20;
  int threads = 0;
  // results (region id order = processing order)
  std::vector<uint32_t> masks;       // R * Wd
  std::vector<int> info;             // R * AMO_INFO_STRIDE
  std::vector<int> tabs;             // R * (vm + 1)
  std::vector<double> verts;         // R * vm * 3
  std::vector<int> vpairs;           // R * vm * 2
  std::vector<double> points;        // R * 3 (queued seed point)
  long n_regions = 0;
  long tag_collisions = 0;           // tag table said "visited" for an unseen mask
  long n_batches = 0;
  double seconds = 0;
  bool truncated = false;
};

template <typename T>
struct Precast {
  std::vector<std::vector<T>> W, b, tm;
};

template <typename T>
Precast<T> precast(const Net& n) {
  Precast<T> p;
  for (auto& w : n.W) p.W.emplace_back(w.begin(), w.end());
  for (auto& w : n.b) p.b.emplace_back(w.begin(), w.end());
  for (auto& w : n.tm) p.tm.emplace_back(w.begin(), w.end());
  return p;
}

void mask_bytes(const uint32_t* m, int L, uint8_t* out) {
  const int nb = (L + 7) / 8;
  for (int i = 0; i < nb; ++i) out[i] = (uint8_t)(m[i >> 2] >> ((i & 3) * 8));
  // bits beyond L are zero by construction
}

template <typename T>
struct FrontierItem { std::vector<uint32_t> mask; int start; T p[3]; };

template <typename T>
void store_region(Run& run, const FrontierItem<T>& it, const RegionOut<T>& r) {
  const int Wd = run.net.Wd, vm = run.vm;
  run.masks.insert(run.masks.end(), it.mask.begin(), it.mask.end());
  int inf[AMO_INFO_STRIDE] = {0};
  inf[0] = r.start_in;
  inf[1] = r.start;
  inf[2] = (int)r.verts.size() / 3;
  inf[3] = (int)r.tab.size();
  inf[4] = r.reversed;
  inf[5] = r.status;
  run.info.insert(run.info.end(), inf, inf + AMO_INFO_STRIDE);
  std::vector<int> tb(vm + 1, 0);
  for (size_t i = 0; i < r.tab.size() && (int)i < vm + 1; ++i) tb[i] = r.tab[i];
  run.tabs.insert(run.tabs.end(), tb.begin(), tb.end());
  std::vector<double> vv((size_t)vm * 3, 0.0);
  for (size_t i = 0; i < r.verts.size(); ++i) vv[i] = (double)r.verts[i];
  run.verts.insert(run.verts.end(), vv.begin(), vv.end());
  std::vector<int> vp((size_t)vm * 2, 0);
  for (size_t i = 0; i < r.vpair.size(); ++i) vp[i] = r.vpair[i];
  run.vpairs.insert(run.vpairs.end(), vp.begin(), vp.end());
  for (int q = 0; q < 3; ++q) run.points.push_back((double)it.p[q]);
  (void)Wd;
  run.n_regions++;
}

template <typename T>
void process_batch(Run& run, const Precast<T>& pc, const std::vector<FrontierItem<T>>& items,
                   size_t first, size_t count, T iso, bool flip, std::vector<RegionOut<T>>& outs) {
  const Net& n = run.net;
  outs.resize(count);
#pragma omp parallel
  {
    std::vector<T> A((size_t)n.C * 4);
    T eq[4];
#pragma omp for schedule(dynamic, 4)
    for (long q = 0; q < (long)count; ++q) {
      const FrontierItem<T>& it = items[first + q];
      compute_planes<T>(n, pc.W, pc.b, pc.tm, it.mask.data(), iso, A.data(), eq);
      process_region<T>(n, A.data(), eq, it.p, it.start, run.vm, flip, outs[q]);
    }
  }
}

template <typename T>
int run_lifo(Run& run, const std::vector<FrontierItem<T>>& seeds, T iso, bool flip, long max_regions) {
  const Net& n = run.net;
  const Precast<T> pc = precast<T>(n);
  const int nbytes = (n.L + 7) / 8;
  std::vector<uint32_t> table(TABLE_SIZE);
  std::vector<uint8_t> zero(nbytes, 0), kb(nbytes);
  const uint32_t empty = fnv1(zero.data(), nbytes);
  std::fill(table.begin(), table.end(), empty);
  std::unordered_set<std::string, MaskHash> exact;
  // try_insert (backend/inc/hash.h:33-52): tag-only membership
  auto try_insert = [&](const std::vector<uint32_t>& m) -> bool {
    mask_bytes(m.data(), n.L, kb.data());
    uint32_t idx = fnv1a(kb.data(), nbytes) & (TABLE_SIZE - 1);
    const uint32_t tag = fnv1(kb.data(), nbytes);
    while (table[idx] != empty && table[idx] != tag) idx = (idx + SKIP) & (TABLE_SIZE - 1);
    std::string key((const char*)kb.data(), nbytes);
    const bool seen = exact.count(key) > 0;
    if (table[idx] == empty) {
      table[idx] = tag;
      exact.insert(key);
      return true;
    }
    if (!seen) run.tag_collisions++;
    return false;
  };
  std::vector<FrontierItem<T>> stack;
  for (const auto& s : seeds)
    if (try_insert(s.mask)) stack.push_back(s);   // backend/inc/states.h:166-175
  std::vector<RegionOut<T>> outs;
  std::vector<Cand<T>> cands;
  while (!stack.empty()) {
    if (max_regions >= 0 && run.n_regions >= max_regions) { run.truncated = true; break; }
    const size_t bs = std::min<size_t>(stack.size(), BATCH);
    const size_t first = stack.size() - bs;
    process_batch<T>(run, pc, stack, first, bs, iso, flip, outs);
    std::vector<FrontierItem<T>> batch(stack.begin() + first, stack.end());
    stack.resize(first);
    for (size_t q = 0; q < bs; ++q) store_region<T>(run, batch[q], outs[q]);
    // emission order: thread id = edge_pos * bs + region (backend/inc/kernel.h:1402-1404)
    std::vector<std::vector<Cand<T>>> per(bs);
    int maxe = 0;
    for (size_t q = 0; q < bs; ++q) {
      region_candidates<T>(n, outs[q], per[q]);
      for (auto& c : per[q]) maxe = std::max(maxe, c.edge_pos + 1);
    }
    std::vector<FrontierItem<T>> fresh;
    for (int e = 0; e < maxe; ++e) {
      for (size_t q = 0; q < bs; ++q) {
        for (auto& c : per[q]) {
          if (c.edge_pos != e) continue;
          FrontierItem<T> it;
          it.mask = batch[q].mask;
          it.mask[c.t >> 5] ^= (1u << (c.t & 31));
          if (!try_insert(it.mask)) continue;
          it.start = c.t;
          it.p[0] = c.mid[0]; it.p[1] = c.mid[1]; it.p[2] = c.mid[2];
          fresh.push_back(std::move(it));
        }
      }
    }
    for (auto& f : fresh) stack.push_back(std::move(f));  // push_back (backend/inc/states.h:245-282)
    run.n_batches++;
  }
  return 0;
}

template <typename T>
int run_bfs(Run& run, const std::vector<FrontierItem<T>>& seeds, T iso, bool flip, long max_regions) {
  const Net& n = run.net;
  const Precast<T> pc = precast<T>(n);
  const int nbytes = (n.L + 7) / 8;
  std::unordered_set<std::string, MaskHash> visited;
  std::vector<uint8_t> kb(nbytes);
  auto insert = [&](const std::vector<uint32_t>& m) {
    mask_bytes(m.data(), n.L, kb.data());
    return visited.insert(std::string((const char*)kb.data(), nbytes)).second;
  };
  std::vector<FrontierItem<T>> level;
  for (const auto& s : seeds)
    if (insert(s.mask)) level.push_back(s);
  std::vector<RegionOut<T>> outs;
  std::vector<Cand<T>> cands;
  while (!level.empty()) {
    if (max_regions >= 0 && run.n_regions >= max_regions) { run.truncated = true; break; }
    process_batch<T>(run, pc, level, 0, level.size(), iso, flip, outs);
    std::vector<FrontierItem<T>> next;
    for (size_t q = 0; q < level.size(); ++q) {
      store_region<T>(run, level[q], outs[q]);
      region_candidates<T>(n, outs[q], cands);
      for (auto& c : cands) {
        FrontierItem<T> it;
        it.mask = level[q].mask;
        it.mask[c.t >> 5] ^= (1u << (c.t & 31));
        if (!insert(it.mask)) continue;
        it.start = c.t;
        it.p[0] = c.mid[0]; it.p[1] = c.mid[1]; it.p[2] = c.mid[2];
        next.push_back(std::move(it));
      }
    }
    level.swap(next);
    run.n_batches++;
  }
  return 0;
}

}  // namespace

// -------------------------------------------------------------------------
// C API
// -------------------------------------------------------------------------
extern "C" {

uint32_t amo_fnv1a(const uint8_t* d, int n) { return fnv1a(d, n); }
uint32_t amo_fnv1(const uint8_t* d, int n) { return fnv1(d, n); }
void amo_get_i_j(int x, int ij_max, int* i, int* j) { get_i_j(x, ij_max, *i, *j); }

void* amo_create(int n_nodes, const int* nodes, const double* const* W, const double* const* b,
                 const int* arc_table, int arc_rows, int arc_cols,
                 int n_tm, const int* tm_shapes, const double* const* tms,
                 int E, const double* ew, const double* eb,
                 int precision, int traversal, int vm, int threads) {
  if (n_nodes < 3 || nodes[0] != 3 || nodes[n_nodes - 1] != 1) return nullptr;
  if (vm < 3 || vm > 64) return nullptr;
  Run* r = new Run();
  Net& n = r->net;
  n.nodes.assign(nodes, nodes + n_nodes);
  n.D = n_nodes - 2;
  n.off.assign(n.D + 1, 0);
  for (int k = 0; k < n.D; ++k) n.off[k + 1] = n.off[k] + nodes[k + 1];
  n.L = n.off[n.D];
  n.E = E;
  n.C = n.L + E;
  n.Wd = (n.L + 31) / 32;
  for (int li = 0; li <= n.D; ++li) {
    const size_t cnt = (size_t)nodes[li + 1] * nodes[li];
    n.W.emplace_back(W[li], W[li] + cnt);
    n.b.emplace_back(b[li], b[li] + nodes[li + 1]);
  }
  n.skips.assign(n.D + 1, {});
  for (int row = 0; row < arc_rows && row + 1 < n.D; ++row) {   // output-layer row ignored (Q1)
    const int cnum = arc_table[row * arc_cols];
    for (int c = 0; c < cnum; ++c)
      n.skips[row + 1].push_back({arc_table[row * arc_cols + 1 + 2 * c], arc_table[row * arc_cols + 2 + 2 * c]});
  }
  for (int t = 0; t < n_tm; ++t) {
    n.tm_h.push_back(tm_shapes[2 * t]);
    n.tm_w.push_back(tm_shapes[2 * t + 1]);
    const size_t cnt = (size_t)tm_shapes[2 * t] * tm_shapes[2 * t + 1];
    n.tm.emplace_back(tms[t], tms[t] + cnt);
  }
  n.ew.assign(ew, ew + 3 * E);
  n.eb.assign(eb, eb + E);
  r->precision = precision;
  r->traversal = traversal;
  r->vm = vm;
  r->threads = threads;
  return r;
}

int amo_run(void* h, int n_seeds, const double* seed_points, const uint8_t* seed_states,
            double iso, int flip, long max_regions) {
  Run& run = *(Run*)h;
  const Net& n = run.net;
#ifdef _OPENMP
  if (run.threads > 0) omp_set_num_threads(run.threads);
#endif
  run.masks.clear(); run.info.clear(); run.tabs.clear(); run.verts.clear(); run.vpairs.clear();
  run.points.clear();
  run.n_regions = 0; run.tag_collisions = 0; run.n_batches = 0; run.truncated = false;
  auto t0 = std::chrono::steady_clock::now();
  int rc;
  if (run.precision == 0) {
    std::vector<FrontierItem<double>> seeds(n_seeds);
    for (int s = 0; s < n_seeds; ++s) {
      seeds[s].mask.assign(n.Wd, 0u);
      for (int g = 0; g < n.L; ++g)
        if (seed_states[(size_t)s * n.L + g]) seeds[s].mask[g >> 5] |= 1u << (g & 31);
      seeds[s].start = -1;
      for (int q = 0; q < 3; ++q) seeds[s].p[q] = seed_points[s * 3 + q];
    }
    rc = run.traversal == 0 ? run_lifo<double>(run, seeds, iso, flip != 0, max_regions)
                            : run_bfs<double>(run, seeds, iso, flip != 0, max_regions);
  } else {
    std::vector<FrontierItem<float>> seeds(n_seeds);
    for (int s = 0; s < n_seeds; ++s) {
      seeds[s].mask.assign(n.Wd, 0u);
      for (int g = 0; g < n.L; ++g)
        if (seed_states[(size_t)s * n.L + g]) seeds[s].mask[g >> 5] |= 1u << (g & 31);
      seeds[s].start = -1;
      for (int q = 0; q < 3; ++q) seeds[s].p[q] = (float)seed_points[s * 3 + q];
    }
    rc = run.traversal == 0 ? run_lifo<float>(run, seeds, (float)iso, flip != 0, max_regions)
                            : run_bfs<float>(run, seeds, (float)iso, flip != 0, max_regions);
  }
  run.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

long amo_num_regions(void* h) { return ((Run*)h)->n_regions; }
int amo_mask_words(void* h) { return ((Run*)h)->net.Wd; }
int amo_num_constraints(void* h) { return ((Run*)h)->net.C; }
double amo_seconds(void* h) { return ((Run*)h)->seconds; }
long amo_tag_collisions(void* h) { return ((Run*)h)->tag_collisions; }
long amo_num_batches(void* h) { return ((Run*)h)->n_batches; }
int amo_truncated(void* h) { return ((Run*)h)->truncated ? 1 : 0; }

void amo_get(void* h, uint32_t* masks, int* info, int* tabs, double* verts, int* vpairs, double* points) {
  Run& r = *(Run*)h;
  if (masks) std::memcpy(masks, r.masks.data(), r.masks.size() * sizeof(uint32_t));
  if (info) std::memcpy(info, r.info.data(), r.info.size() * sizeof(int));
  if (tabs) std::memcpy(tabs, r.tabs.data(), r.tabs.size() * sizeof(int));
  if (verts) std::memcpy(verts, r.verts.data(), r.verts.size() * sizeof(double));
  if (vpairs) std::memcpy(vpairs, r.vpairs.data(), r.vpairs.size() * sizeof(int));
  if (points) std::memcpy(points, r.points.data(), r.points.size() * sizeof(double));
}

// Planes of one region in FP64 (for tests of the chain itself).
void amo_planes(void* h, const uint32_t* mask, double iso, double* A, double* eq) {
  Run& r = *(Run*)h;
  const Precast<double> pc = precast<double>(r.net);
  compute_planes<double>(r.net, pc.W, pc.b, pc.tm, mask, iso, A, eq);
}

void amo_destroy(void* h) { delete (Run*)h; }

}  // extern "C"
