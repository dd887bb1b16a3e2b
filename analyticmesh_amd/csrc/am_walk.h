// Per-region walk shared by the FP64 kernel and the reduced-precision walk
// kernel: distance order, seed start search, pivot walk, orientation, record
// and neighbour candidates (backend/inc/process.h:208-264,
// backend/inc/kernel.h:799-1506).  One wave64 runs one region; the constraint
// planes are already in LDS as SoA arrays (x, y, z, c).
//
// CHECKED = false: plain evaluation in T (the FP64 path; the operation order
// matches the CPU restatement exactly).
// CHECKED = true: every predicate also carries an error bound derived from the
// planes' relative error bound `eps`; when a sign cannot be decided within its
// bound the walk stops and reports the region as undecided, and the FP64
// kernel redoes it.  Candidate order and which of two feasible candidates is
// taken first only change how a face is written down, never which face it is.
#pragma once
#include "am_device.h"
#include "am_kernels.h"

namespace am {

#define AM_UNDECIDED 0x1u
// A reduced-precision walk decided the region but cannot vouch for one of its
// vertices to 1e-5 (relative to |v|_inf, at least 1e-3): its vertices are
// re-solved from FP64 planes.
#define AM_VERT64 0x40000000u
#define AM_VERT_TOL 1e-5f
// Why a reduced-precision walk left its region to FP64 (flags bit 1 + reason index).
#define AM_WHY_SEED 0       // seed: no plane bounds the seed's shot
#define AM_WHY_START 1      // start edge interval empty or not finite
#define AM_WHY_TAB 2        // proposed plane already in the polygon
#define AM_WHY_UNSURE 3     // a constraint's sign at the new vertex within its bound
#define AM_WHY_VIOL 4       // a constraint certainly violated at the proposed vertex
#define AM_WHY_HMAX 5       // vertex system too ill-conditioned to trust
#define AM_WHY_TURN 6       // turn direction within its bound
#define AM_WHY_TRUNC 7      // polygon reached the vertex cap
#define AM_WHY_NEXT 8       // no plane bounds the next edge
#define AM_WHY_ORIENT 9     // orientation within its bound
#define AM_WHY_SCAN 10      // candidate-scan walk (start search / scan) undecided
#define AM_WHY_SAME 11      // new vertex within the FP64 same-vertex distance of the last one
#define AM_WHY_COUNT 12
#define AM_WHY_BIT(w) (1u << (1 + (w)))

struct WalkParams {
  float eps;        // relative error bound of the planes (CHECKED only); with
                    // per-plane errors, the relative share on top of them
  float eps_turn;   // relative bound for the turn and orientation tests
  float h_max;      // largest Hadamard ratio of a vertex system that is trusted
  float lam = 0.f;  // per-plane errors: a plane's bound is lam sigma (sigma its
                    // error estimate) plus eps of its largest coefficient
  float lam_v = 0.f;  // per-plane errors: the multiple of sigma in the vertex
                      // accuracy estimate; < 0: every vertex is left to FP64
                      // (AM_VERT64 on every decided region), 0: none is
  int vsel = 0;       // with lam_v > 0 and inverse systems kept: a vertex within
                      // the estimate is marked as needing no Newton step (NaN
                      // first entry of its inverse system) instead of sending
                      // its region to FP64
};

template <typename T>
__device__ __forceinline__ T hadamard_ratio(const T* p1, const T* p2, const T* p3) {
  const T d = AM_DET3(p1[0], p1[1], p1[2], p2[0], p2[1], p2[2], p3[0], p3[1], p3[2]);
  const T n1 = __builtin_sqrt(plane_nn(p1[0], p1[1], p1[2]));
  const T n2 = __builtin_sqrt(plane_nn(p2[0], p2[1], p2[2]));
  const T n3 = __builtin_sqrt(plane_nn(p3[0], p3[1], p3[2]));
  return (n1 * n2 * n3) / absT(d);
}

// Error bound of constraint signs at the vertex v of the system A = [p1; p2; p3]
// when plane x's coefficients carry an error of at most m_x: eps times its
// largest coefficient, plus (per-plane mode) an absolute share e_x.
// Perturbing the system by E moves the vertex by dv = -A^-1 (E v + e) to first
// order, so constraint g's sign moves by lambda_g . (E v + e) with
// lambda_g = A^-T a_g; its rows are the cross products of the system's normals
// over the determinant.  The bound for g is
//   m_g (1 + |v|_1) + kappa sum_i |lambda_g,i| m_i (1 + |v|_1),
// where kappa = 1 / (1 - 2 rho), rho = r (1 + H) (H the Hadamard ratio, r the
// largest m_i relative to its plane's largest coefficient), covers the
// second-order term; systems with rho > 1/4 are not trusted.  The vertex
// solve's working-precision rounding is added to every m_i.
// Since |lambda_g,i| <= |n_i|_1 max|a_g| (n_i the rows of A^-T), with plane
// errors relative to the largest coefficient the bound is at most max|P_g| E
// with the per-vertex constant E = eps (1 + |v|_1) + sum_i kappa m_i |n_i|_1.
// FP64 planes (eps ~ 1e-13) use that cheap form; coarse planes keep the
// per-constraint lambda form, which certifies more.
template <typename T>
struct VertexBound {
  T E;                     // FP64: err_g = E max|P_g|
  T n1[3], n2[3], n3[3];   // coarse planes: rows of A^-T (columns of A^-1)
  T w1, w2, w3, e0, eg;    // coarse planes: kappa m_i, eps (1 + |v|_1), lam (1 + |v|_1)
  T s1, s2, s3;            // coarse planes, per-plane mode: kappa m_i with sigma for lam sigma
                           // (the vertex's own error estimate)
  bool ok;
};

template <typename T>
__device__ __forceinline__ T coef_max(const T* p) {
  return fmax(fmax(absT(p[0]), absT(p[1])), fmax(absT(p[2]), absT(p[3])));
}

// The bound is linear in (1 + |v|_1): vertex_bound_unit computes it for
// 1 + |v|_1 = 1 (it does not need the vertex), scale_bound applies the factor.
// e1..e3: absolute coefficient errors of the three planes on top of eps times
// their largest coefficients, lam sigma_i for error estimates sigma_i (0 when
// the planes carry eps only); a constraint's own share is then lam sigma_g
// (checked_verdict takes sigma_g), and the vertex estimate s_i uses lam_v sigma_i.
template <typename T>
__device__ __forceinline__ VertexBound<T> vertex_bound_unit(const T* p1, const T* p2, const T* p3, T e1, T e2, T e3,
                                                            T eps, T h_max, T lam, T lam_v) {
  VertexBound<T> b;
  const T c1[3] = {p2[1] * p3[2] - p2[2] * p3[1], p2[2] * p3[0] - p2[0] * p3[2], p2[0] * p3[1] - p2[1] * p3[0]};
  const T c2[3] = {p3[1] * p1[2] - p3[2] * p1[1], p3[2] * p1[0] - p3[0] * p1[2], p3[0] * p1[1] - p3[1] * p1[0]};
  const T c3[3] = {p1[1] * p2[2] - p1[2] * p2[1], p1[2] * p2[0] - p1[0] * p2[2], p1[0] * p2[1] - p1[1] * p2[0]};
  const T det = p1[0] * c1[0] + p1[1] * c1[1] + p1[2] * c1[2];
  const T adet = absT(det);
  const T iadet = T(1) / adet;
  const T h = __builtin_sqrt(plane_nn(p1[0], p1[1], p1[2]) * plane_nn(p2[0], p2[1], p2[2]) *
                             plane_nn(p3[0], p3[1], p3[2])) * iadet;
  // Cramer's rule in working precision adds an error equivalent to perturbing
  // the system by about 16 u H of its coefficients (each 3x3 expansion's
  // absolute terms sum to at most 5.2 H |det|, with three roundings each)
  // (per-plane mode, lam > 0: the vertex is solved in FP64 and rounded to T,
  // which the constraint's own eps share covers)
  const T u = (sizeof(T) == 8 || lam > T(0)) ? T(1.1102230246251565e-16) : T(5.9604644775390625e-08);
  const T es = eps + T(32) * u * (T(1) + h);
  const T q1 = coef_max(p1), q2 = coef_max(p2), q3 = coef_max(p3);
  // the planes' total coefficient errors
  const T m1 = es * q1 + e1, m2 = es * q2 + e2, m3 = es * q3 + e3;
  const T rel = fmax(fmax(m1 / q1, m2 / q2), m3 / q3);
  const T rho = rel * (T(1) + h);
  b.ok = h <= h_max && rho <= T(0.25);
  // 1.001: slack for the rounding of these bound computations themselves.
  // kappa >= 1.001 / (1 - x) with x = 2 rho <= 1/2 (systems with larger rho
  // are not trusted): 1 / (1 - x) = 1 + x + x^2 / (1 - x) <= 1 + x + 2 x^2
  const T x2 = T(2) * rho;
  const T kappa = T(1.001) * (T(1) + x2 + T(2) * x2 * x2);
  // |n_i|_1 = |c_i|_1 / |det|
  const T l1 = absT(c1[0]) + absT(c1[1]) + absT(c1[2]);
  const T l2 = absT(c2[0]) + absT(c2[1]) + absT(c2[2]);
  const T l3 = absT(c3[0]) + absT(c3[1]) + absT(c3[2]);
  const T w = kappa * (m1 * l1 + m2 * l2 + m3 * l3) * iadet;
  b.E = T(1.001) * (es + w);
  if (sizeof(T) == 4) {
    const T inv = det < T(0) ? -iadet : iadet;   // 1 / det
    for (int i = 0; i < 3; ++i) { b.n1[i] = c1[i] * inv; b.n2[i] = c2[i] * inv; b.n3[i] = c3[i] * inv; }
    b.w1 = kappa * m1;
    b.w2 = kappa * m2;
    b.w3 = kappa * m3;
    b.e0 = T(1.001) * es;
    b.eg = T(1.001) * lam;
    // the vertex estimate (per-plane mode, where the vertex is solved in FP64):
    // the planes' FP32 rounding (u of the largest coefficient) and the FP64
    // solve's, plus lam_v sigma for the error of the sums
    const T il = lam > T(0) ? lam_v / lam : T(0);
    const T ev = T(5.9604644775390625e-08) + T(32) * T(1.1102230246251565e-16) * (T(1) + h);
    b.s1 = kappa * (ev * q1 + e1 * il);
    b.s2 = kappa * (ev * q2 + e2 * il);
    b.s3 = kappa * (ev * q3 + e3 * il);
  }
  return b;
}

template <typename T>
__device__ __forceinline__ VertexBound<T> vertex_bound_unit(const T* p1, const T* p2, const T* p3, T eps, T h_max) {
  return vertex_bound_unit<T>(p1, p2, p3, T(0), T(0), T(0), eps, h_max, T(0), T(0));
}

template <typename T>
__device__ __forceinline__ void scale_bound(VertexBound<T>& b, T vn) {
  b.E *= vn;
  b.w1 *= vn;
  b.w2 *= vn;
  b.w3 *= vn;
  b.e0 *= vn;
  b.eg *= vn;
  b.s1 *= vn;
  b.s2 *= vn;
  b.s3 *= vn;
}

template <typename T>
__device__ __forceinline__ VertexBound<T> vertex_bound(const T* p1, const T* p2, const T* p3, const T* v, T e1, T e2,
                                                       T e3, T eps, T h_max, T lam, T lam_v) {
  VertexBound<T> b = vertex_bound_unit<T>(p1, p2, p3, e1, e2, e3, eps, h_max, lam, lam_v);
  scale_bound<T>(b, T(1) + absT(v[0]) + absT(v[1]) + absT(v[2]));
  return b;
}

// The vertex of planes p1, p2, p3: in FP64 and rounded to T when `wide` (coarse
// planes, per-plane mode), else in T.
template <typename T>
__device__ __forceinline__ void solve_vertex(const T* p1, const T* p2, const T* p3, T* v, bool wide) {
  if (sizeof(T) == 4 && wide) {
    const double a[4] = {p1[0], p1[1], p1[2], p1[3]};
    const double b[4] = {p2[0], p2[1], p2[2], p2[3]};
    const double c[4] = {p3[0], p3[1], p3[2], p3[3]};
    double x[3];
    solve3(a, b, c, x);
    v[0] = T(x[0]); v[1] = T(x[1]); v[2] = T(x[2]);
  } else {
    solve3(p1, p2, p3, v);
  }
}

// Coarse planes, per-plane mode: whether the vertex v (solved from the bound's
// system) is within AM_VERT_TOL of the exact one relative to |v|_inf, by the
// error estimate |dv_x| <= sum_i |A^-1_x,i| s_i (a vertex at the origin never is).
template <typename T>
__device__ __forceinline__ bool vertex_accurate(const VertexBound<T>& b, const T* v) {
  const T tol = T(AM_VERT_TOL) * fmax(fmax(absT(v[0]), absT(v[1])), absT(v[2]));
  bool ok = true;
  for (int x = 0; x < 3; ++x)
    ok = ok && (absT(b.n1[x]) * b.s1 + absT(b.n2[x]) * b.s2 + absT(b.n3[x]) * b.s3 <= tol);
  return ok;
}

// Signed-distance verdict of constraint (ax, ay, az, c) at v under the bound:
// 0 = satisfied, 1 = certainly violated, 2 = undecided.
// eg: the constraint's error estimate sigma_g (coarse planes, per-plane mode;
// its bound share is lam sigma_g), else 0.
// s: the constraint's value at v, plane_eval(ax, ay, az, c, v).
template <typename T>
__device__ __forceinline__ int checked_verdict_at(T s, T ax, T ay, T az, T c, T eg, const VertexBound<T>& b) {
  const T pm = fmax(fmax(absT(ax), absT(ay)), fmax(absT(az), absT(c)));
  T err;
  if (sizeof(T) == 4) {
    // fused multiply-adds: one rounding per term (the bound's slack covers
    // its own evaluation either way)
    const T l1 = fmaT(b.n1[2], az, fmaT(b.n1[1], ay, b.n1[0] * ax));
    const T l2 = fmaT(b.n2[2], az, fmaT(b.n2[1], ay, b.n2[0] * ax));
    const T l3 = fmaT(b.n3[2], az, fmaT(b.n3[1], ay, b.n3[0] * ax));
    err = fmaT(b.w3, absT(l3), fmaT(b.w2, absT(l2), fmaT(b.w1, absT(l1), fmaT(b.eg, eg, b.e0 * pm))));
  } else {
    err = b.E * pm;
  }
  if (s > err) return 1;
  if (s >= -err) return 2;
  return 0;
}

template <typename T>
__device__ __forceinline__ int checked_verdict(T ax, T ay, T az, T c, T eg, const T* v, const VertexBound<T>& b) {
  return checked_verdict_at(plane_eval(ax, ay, az, c, v), ax, ay, az, c, eg, b);
}

// Verdict of one vertex against all constraints owned by this lane.
// 0 = satisfied, 1 = certainly violated, 2 = undecided.
// Pe: per-plane error estimates sigma, or nullptr.
template <typename T, bool CHECKED>
__device__ __forceinline__ int lane_check(const T* Px, const T* Py, const T* Pz, const T* Pc, const T* Pe, int C,
                                          int lane, int ex1, int ex2, const T* v, const VertexBound<T>& vb) {
  int verdict = 0;
#pragma unroll 2
  for (int g = lane; g < C; g += 64) {
    if (g == ex1 || g == ex2) continue;
    const T ax = Px[g], ay = Py[g], az = Pz[g], c = Pc[g];
    if (!CHECKED) {
      if (violates(ax, ay, az, c, v)) return 1;
    } else {
      const int vg = checked_verdict(ax, ay, az, c, Pe ? Pe[g] : T(0), v, vb);
      if (vg == 1) return 1;
      if (vg == 2) verdict = 2;
    }
  }
  return verdict;
}

// Orientation, region record and neighbour candidates of a finished walk
// (backend/inc/kernel.h:1325-1474).  In CHECKED mode an undecided region only
// raises its flag: the FP64 redo writes its record and candidate slots.
// Callers reach it after a barrier that makes the walk's LDS records visible.
template <typename T, bool CHECKED, typename V = T>
__device__ void finish_region(const NetDev& net, const RegionIO& io, const WalkParams& wp, int rid, int slot_region,
                              const uint32_t* smask, const T* eq, int* s_tab, V* s_verts, int* s_vpair, int ntab,
                              int nv, int start_in, int start, int status, bool undecided, unsigned why = 0u,
                              bool vert64 = false, float* s_jinv = nullptr) {
  const int L = net.L, vm = net.vm;
  const int lane = threadIdx.x;     // any block size: strides are blockDim.x
  const int nth = blockDim.x;
  // ---- orientation (backend/inc/kernel.h:1325-1371)
  int reversed = 0;
  if (nv >= 3 && !undecided) {
      const V* q = s_verts;
      const V ax = q[3] - q[0], ay = q[4] - q[1], az = q[5] - q[2];
      const V bx = q[6] - q[3], by = q[7] - q[4], bz = q[8] - q[5];
      const V cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
      const V dir = cx * V(eq[0]) + cy * V(eq[1]) + cz * V(eq[2]);
      if (CHECKED) {
        const V mag = (absT(cx) + absT(cy) + absT(cz)) * (absT(V(eq[0])) + absT(V(eq[1])) + absT(V(eq[2])));
        if (!(absT(dir) > V(wp.eps_turn) * mag)) { undecided = true; why |= AM_WHY_BIT(AM_WHY_ORIENT); }
      }
      reversed = ((dir < V(0)) != (net.flip != 0)) ? 1 : 0;
  }
  // a reversed polygon's vertices are read back to front below (vertex k of
  // the record is vertex nv - 1 - k of the walk), with no reordering in LDS
  const bool rev = nv >= 3 && !undecided && reversed;
  if (CHECKED && (status != AM_ST_OK)) undecided = true;
  if (CHECKED && undecided) {
    if (lane == 0) io.flags[slot_region] = AM_UNDECIDED | why;
    return;   // the FP64 redo writes this region's record and candidate slots
  }
  if (CHECKED && lane == 0) io.flags[slot_region] = vert64 ? AM_VERT64 : 0u;

  // ---- region record
  if (lane == 0) {
    int* inf = io.reg_info + (size_t)rid * AM_INFO_STRIDE;
    inf[0] = start_in; inf[1] = start; inf[2] = nv; inf[3] = ntab;
    inf[4] = reversed; inf[5] = status; inf[6] = CHECKED ? 2 : (io.ids ? 1 : 0);
    inf[7] = (s_jinv && io.reg_jinv) ? 1 : 0;
  }
  if (s_jinv && io.reg_jinv)
    for (int i = lane; i < nv * 9; i += nth) {
      const int k = i / 9, c = i - 9 * k;
      io.reg_jinv[(size_t)rid * vm * 9 + i] = bf16_bits(s_jinv[(rev ? nv - 1 - k : k) * 9 + c]);
    }
  for (int i = lane; i < vm + 1; i += nth) io.reg_tab[(size_t)rid * (vm + 1) + i] = i < ntab ? s_tab[i] : 0;
  for (int i = lane; i < vm * 3; i += nth) {
    const int k = i / 3, c = i - 3 * k;
    io.reg_verts[(size_t)rid * vm * 3 + i] = i < nv * 3 ? (double)s_verts[(rev ? nv - 1 - k : k) * 3 + c] : 0.0;
  }
  for (int i = lane; i < vm * 2; i += nth) {
    const int k = i >> 1, c = i & 1;
    io.reg_vpair[(size_t)rid * vm * 2 + i] = i < nv * 2 ? s_vpair[(rev ? nv - 1 - k : k) * 2 + c] : 0;
  }

  // ---- neighbour candidates (backend/inc/kernel.h:1390-1474)
  const size_t slot0 = (size_t)slot_region * vm;
  if (lane < vm) {
    const int m = lane;
    int t = -1;
    if (m + 1 < ntab) t = s_tab[m];
    const bool valid = t >= 0 && t < L;
    V mid[3] = {V(0), V(0), V(0)};
    if (valid) {
      const int m_ = (m == 0) ? ntab - 1 : m;
      int cnt = -1, v1 = -1;
      for (int i = 1; i < ntab; ++i) {
        if (s_tab[i] >= 0) ++cnt;
        if (i == m_) v1 = cnt;
      }
      if (nv > 0 && v1 >= 0) {
        int v2 = (v1 + 1) % nv;
        const int w1 = rev ? nv - 1 - v1 : v1, w2 = rev ? nv - 1 - v2 : v2;
        for (int c = 0; c < 3; ++c) mid[c] = (s_verts[w1 * 3 + c] + s_verts[w2 * 3 + c]) / V(2);
      }
    }
    io.cand_start[slot0 + m] = valid ? t : -2;
    io.cand_point[(slot0 + m) * 3 + 0] = (double)mid[0];
    io.cand_point[(slot0 + m) * 3 + 1] = (double)mid[1];
    io.cand_point[(slot0 + m) * 3 + 2] = (double)mid[2];
  }
  const int Wd = net.Wd;
  // candidate m's mask: waves take candidates, lanes take words (no division)
  for (int m = lane >> 6; m < vm; m += (nth + 63) >> 6) {
    int t = -1;
    if (m + 1 < ntab) t = s_tab[m];
    if (t >= 0 && t < L)
      for (int w = lane & 63; w < Wd; w += 64) {
        uint32_t word = smask[w];
        if ((t >> 5) == w) word ^= 1u << (t & 31);
        io.cand_mask[(slot0 + m) * Wd + w] = word;
      }
  }
  // each candidate's key hash: the region's sum of word hashes with the
  // flipped word's term swapped (wave 0; Wd <= 64 and vm <= 64)
  if (lane < 64) {
    uint64_t g = 0;
    if (lane < Wd) g = word_hash(smask[lane], lane);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)g, d, 64);
      const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(g >> 32), d, 64);
      g += ((uint64_t)hi << 32) | lo;
    }
    if (lane < vm) {
      int t = -1;
      if (lane + 1 < ntab) t = s_tab[lane];
      if (t >= 0 && t < L) {
        const int wt = t >> 5;
        const uint32_t x = smask[wt];
        io.cand_hash[slot0 + lane] = g - word_hash(x, wt) + word_hash(x ^ (1u << (t & 31)), wt);
      }
    }
  }
}

// Distance key of constraint g at p: s|s|/|a|^2 orders like the signed
// distance s/|a| (backend/inc/process.h:208-239) without a square root.
template <typename T>
__device__ __forceinline__ T distance_key(const T* Px, const T* Py, const T* Pz, const T* Pc, int g, const T* p) {
  const T s = plane_eval(Px[g], Py[g], Pz[g], Pc[g], p);
  const T nn = plane_nn(Px[g], Py[g], Pz[g]);
  T key = nn > T(0) ? s * absT(s) / nn : (s > T(0) ? T(__builtin_inf()) : T(-__builtin_inf()));
  if (key != key) key = -__builtin_inf();  // NaN sorts last
  return key;
}

// (key, index) a comes before b: larger key first, ties to the lower index.
template <typename T>
__device__ __forceinline__ bool key_before(T ka, int ia, T kb, int ib) {
  return ka > kb || (ka == kb && ia < ib);
}

// Full stable descending order of all constraints: keys[0..N2), order[0..N2)
// in LDS, bitonic (backend/inc/process.h:244-264 sorts the same keys).
template <typename T>
__device__ void full_sort(const T* Px, const T* Py, const T* Pz, const T* Pc, int C, int N2, const T* p, T* keys,
                          int* order) {
  const int lane = threadIdx.x;
  __syncthreads();
  for (int g = lane; g < N2; g += 64) {
    keys[g] = g < C ? distance_key<T>(Px, Py, Pz, Pc, g, p) : T(-__builtin_inf());
    order[g] = g;
  }
  __syncthreads();
  for (int k = 2; k <= N2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < N2; i += 64) {
        const int l = i ^ j;
        if (l > i) {
          const T ki = keys[i], kl = keys[l];
          const int oi = order[i], ol = order[l];
          const bool i_first = key_before(ki, oi, kl, ol);
          if (((i & k) == 0) != i_first) {
            keys[i] = kl; keys[l] = ki;
            order[i] = ol; order[l] = oi;
          }
        }
      }
      __syncthreads();
    }
  }
}

// One bitonic step across the wave: the lane keeps the first (keep_first) or
// the second of its element and its partner's (lane ^ j) in the order.
template <typename T>
__device__ __forceinline__ void wave_cas(T& k, int& i, int j, bool keep_first) {
  const T pk = __shfl_xor(k, j);
  const int pi = __shfl_xor(i, j);
  if (keep_first != key_before(k, i, pk, pi)) { k = pk; i = pi; }
}

// The first 64 positions of that order, in registers: lane l ends holding
// position l (sentinel (-inf, INT_MAX) past C).  keys[0, C) and order[0, C)
// are used as scratch.
//  1. keys of all constraints into LDS; each lane's best; T0 = the worst of
//     the 64 lane bests.  Those 64 distinct elements are all >= T0, so the
//     64th of the whole order is >= T0 too: only elements >= T0 can be among
//     the first 64 (about a quarter of them for independent keys).
//  2. the survivors are compacted (ballots, in index order) into order[].
//  3. rounds over 64 survivors at a time: a bitonic sort of the batch across
//     the wave, then the better half of (best, reversed batch) -- a bitonic
//     sequence holding the 64 best of both -- cleaned into order.
// The comparisons are those of full_sort, so the result is exactly its first 64.
template <typename T>
__device__ void wave_top64(const T* Px, const T* Py, const T* Pz, const T* Pc, int C, const T* p, T* keys,
                           int* order, T& bk, int& bi) {
  const int lane = threadIdx.x;
  T mk = T(-__builtin_inf());
  int mi = 0x7fffffff;
  for (int g = lane; g < C; g += 64) {
    const T k = distance_key<T>(Px, Py, Pz, Pc, g, p);
    keys[g] = k;
    if (key_before(k, g, mk, mi)) { mk = k; mi = g; }
  }
  // T0: the worst lane best
  for (int j = 32; j > 0; j >>= 1) {
    const T pk = __shfl_xor(mk, j);
    const int pi = __shfl_xor(mi, j);
    if (key_before(mk, mi, pk, pi)) { mk = pk; mi = pi; }
  }
  int S = 0;
  for (int base = 0; base < C; base += 64) {
    const int g = base + lane;
    const bool keep = g < C && !key_before(mk, mi, keys[g], g);
    const unsigned long long bal = __ballot(keep);
    if (keep) order[S + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] = g;
    S += __popcll(bal);
  }
  __syncthreads();
  bk = T(-__builtin_inf());
  bi = 0x7fffffff;
  for (int base = 0; base < S; base += 64) {
    T ek = T(-__builtin_inf());
    int ei = 0x7fffffff;
    if (base + lane < S) { ei = order[base + lane]; ek = keys[ei]; }
    for (int k = 2; k <= 64; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1)
        wave_cas(ek, ei, j, ((lane & j) == 0) == ((lane & k) == 0));
    if (base == 0) {
      bk = ek;
      bi = ei;
      continue;
    }
    // better of best[l] and batch[63 - l]: a bitonic sequence of the 64 best
    const T rk = __shfl(ek, 63 - lane);
    const int ri = __shfl(ei, 63 - lane);
    if (!key_before(bk, bi, rk, ri)) { bk = rk; bi = ri; }
    for (int j = 32; j > 0; j >>= 1) wave_cas(bk, bi, j, (lane & j) == 0);
  }
  __syncthreads();
}

// Pe, eqe: per-plane error estimates of the constraints and the face plane
// (coarse planes, per-plane mode with wp.lam > 0), as in walk_ray.
template <typename T, bool CHECKED>
__device__ void walk_region(const NetDev& net, const RegionIO& io, const WalkParams& wp, int rid, int slot_region,
                            T* Px, T* Py, T* Pz, T* Pc, T* keys, int* order, uint32_t* used,
                            const uint32_t* smask, const T* eq_in, const T* p, const T* Pe = nullptr,
                            T eqe = T(0)) {
  const bool per_plane = CHECKED && sizeof(T) == 4 && wp.lam > 0.f && Pe != nullptr;
  const T lam = per_plane ? T(wp.lam) : T(0);
  const T* Pg = per_plane ? Pe : nullptr;
  bool vinacc = false;
  __shared__ int s_tab[AM_VM_MAX + 2];
  __shared__ T s_verts[AM_VM_MAX * 3];
  __shared__ int s_vpair[AM_VM_MAX * 2];

  const int C = net.C, vm = net.vm, N2 = io.Csort;
  const int lane = threadIdx.x;
  const T eq[4] = {eq_in[0], eq_in[1], eq_in[2], eq_in[3]};
  const int start_in = io.reg_start[rid];
  bool undecided = false;

  // ---- distance order (keys s|s|/|a|^2, stable descending).  Only the first
  // AM_TOPK positions are built up front (the start search reads positions
  // <= 56 and the walk's winner is nearly always among the first few); the
  // full order is sorted the first time the walk scans past them.
  bool full_order = C <= AM_TOPK;
  {
    T bk;
    int bi;
    wave_top64<T>(Px, Py, Pz, Pc, C, p, keys, order, bk, bi);
    keys[lane] = bk;
    order[lane] = bi;
  }
  __syncthreads();

  int start = start_in;
  int status = AM_ST_OK;
  // ---- seed start search: lowest feasible pair index (backend/inc/kernel.h:799-856)
  if (start == -1) {
    const int xmax = min((C - 1) * (C - 2), 812);
    for (int xb = 0; xb <= xmax; xb += 64) {
      const int x = xb + lane;
      int verdict = x <= xmax ? 0 : 1;
      int ri = 0, rj = 0;
      T v[3] = {T(0), T(0), T(0)};
      VertexBound<T> vb{};
      if (x <= xmax) {
        int i, j;
        get_i_j(x, C, i, j);
        ri = order[i];
        rj = order[j];
        const T pi4[4] = {Px[ri], Py[ri], Pz[ri], Pc[ri]};
        const T pj4[4] = {Px[rj], Py[rj], Pz[rj], Pc[rj]};
        solve_vertex(pi4, pj4, eq, v, per_plane);
        if (CHECKED) {
          vb = vertex_bound<T>(pi4, pj4, eq, v, lam * (Pg ? Pg[ri] : T(0)), lam * (Pg ? Pg[rj] : T(0)), lam * eqe,
                               T(wp.eps), T(wp.h_max), lam, T(wp.lam_v));
          if (!vb.ok) verdict = 2;
        }
      }
      for (int g = 0; g < C; ++g) {
        if (!__any(verdict == 0 || verdict == 2)) break;
        if ((verdict == 0 || (CHECKED && verdict == 2)) && g != ri && g != rj) {
          if (!CHECKED) {
            if (violates(Px[g], Py[g], Pz[g], Pc[g], v)) verdict = 1;
          } else {
            const int vg = checked_verdict(Px[g], Py[g], Pz[g], Pc[g], Pg ? Pg[g] : T(0), v, vb);
            if (vg == 1) verdict = 1;
            else if (vg == 2) verdict = 2;
          }
        }
      }
      const unsigned long long open = __ballot(verdict == 0 || verdict == 2);
      if (open) {
        const int wl = __builtin_ctzll(open);
        const int vw = __shfl(verdict, wl);
        if (CHECKED && vw == 2) { undecided = true; break; }
        start = __shfl(ri, wl);
        break;
      }
    }
    if (start == -1) status = AM_ST_NO_START;
  }

  int ntab = 0, nv = 0;
  if (status == AM_ST_OK && !undecided) {
    // ---- pivot walk (backend/inc/kernel.h:861-1320, backend/src/kernel.cu:203-299)
    if (lane == 0) {
      s_tab[0] = start;
      used[start >> 5] |= 1u << (start & 31);
    }
    __syncthreads();
    ntab = 1;
    int cur = start;
    while (true) {
      if ((ntab >= 4 && s_tab[0] == s_tab[ntab - 1]) || ntab >= vm + 1) break;
      if (cur < 0 || cur >= C) { status = AM_ST_BAD_START; break; }
      int win = -1;
      T v[3] = {T(0), T(0), T(0)};
      const T pc4[4] = {Px[cur], Py[cur], Pz[cur], Pc[cur]};
      const int first_entry = s_tab[0];
      for (int w = 0; w < C; ++w) {
        if (w == AM_TOPK && !full_order) {
          full_sort<T>(Px, Py, Pz, Pc, C, N2, p, keys, order);
          full_order = true;
        }
        const int r = order[w];
        const bool cand = (r == first_entry && ntab >= 3) || !((used[r >> 5] >> (r & 31)) & 1u);
        if (!cand) continue;
        const T pr4[4] = {Px[r], Py[r], Pz[r], Pc[r]};
        T vt[3];
        solve_vertex(pr4, pc4, eq, vt, per_plane);
        if (!CHECKED) {
          // infeasible as soon as one constraint is violated: test them in
          // distance order, a wave-wide chunk at a time, so that a rejected
          // candidate usually stops after the nearest 64 planes (the verdict
          // does not depend on the order of the tests)
          // (first the nearest 64, then every constraint by index; one ballot
          // per four constraints per lane)
          bool viol = false;
          for (int base = -64; base < C; base += 256) {
            bool vl = false;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int pos = base + 64 * u + lane;
              const int g = pos < 0 ? (lane < C ? order[lane] : C) : pos;
              if (g < C && g != cur && g != r) vl = vl || violates(Px[g], Py[g], Pz[g], Pc[g], vt);
            }
            if (__any(vl)) { viol = true; break; }
          }
          if (viol) continue;
          win = r;
          v[0] = vt[0]; v[1] = vt[1]; v[2] = vt[2];
          break;
        }
        const VertexBound<T> vb = vertex_bound<T>(pr4, pc4, eq, vt, lam * (Pg ? Pg[r] : T(0)),
                                                  lam * (Pg ? Pg[cur] : T(0)), lam * eqe, T(wp.eps), T(wp.h_max), lam,
                                                  T(wp.lam_v));
        const int verdict = lane_check<T, CHECKED>(Px, Py, Pz, Pc, Pg, C, lane, cur, r, vt, vb);
        if (__any(verdict == 1)) continue;          // certainly infeasible
        if (CHECKED && (__any(verdict == 2) || !vb.ok)) { undecided = true; break; }
        if (CHECKED && sizeof(T) == 4 &&
            (wp.lam_v < 0.f || (per_plane && wp.lam_v > 0.f && !wp.vsel && !vertex_accurate(vb, vt))))
          vinacc = true;
        win = r;
        v[0] = vt[0]; v[1] = vt[1]; v[2] = vt[2];
        break;
      }
      if (undecided) break;
      if (win < 0) { status = AM_ST_GAVE_UP; break; }
      int pl = -1, pr = -1, cnt = 0;
      for (int idx = ntab - 1; idx >= 0 && cnt < 2; --idx) {
        if (s_tab[idx] >= 0) { if (cnt == 0) pl = idx; else pr = idx; ++cnt; }
      }
      int valid = -1;
      if (cnt == 2) {
        const int a = s_tab[pr], b = s_tab[pl];
        const T n1[3] = {Px[a], Py[a], Pz[a]};
        const T n2[3] = {Px[b], Py[b], Pz[b]};
        const T n3[3] = {Px[win], Py[win], Pz[win]};
        valid = pivot_valid(eq, n1, n2, n3) ? 1 : 0;
        if (CHECKED) {
          const T A12 = AM_DOT3(n1, n2), A23 = AM_DOT3(n2, n3), A13 = AM_DOT3(n1, n3), A22 = AM_DOT3(n2, n2);
          const T C1 = AM_DOT3(n1, eq), C2 = AM_DOT3(n2, eq), C3 = AM_DOT3(n3, eq), B = AM_DOT3(eq, eq);
          const T val = (A12 * A23 - A22 * A13) - (A12 * C2 * C3 + A23 * C1 * C2 - A13 * C2 * C2 - A22 * C1 * C3) / B;
          const T mag = absT(A12 * A23) + absT(A22 * A13) +
                        (absT(A12 * C2 * C3) + absT(A23 * C1 * C2) + absT(A13 * C2 * C2) + absT(A22 * C1 * C3)) / B;
          if (!(absT(val) > T(wp.eps_turn) * mag)) { undecided = true; break; }
        }
      }
      bool same = false;
      if (valid != 0 && nv > 0) {
        const T last[3] = {s_verts[(nv - 1) * 3 + 0], s_verts[(nv - 1) * 3 + 1], s_verts[(nv - 1) * 3 + 2]};
        if (!CHECKED) {
          same = same_vertex(last, v);
        }
        // CHECKED: the accepted vertex was certified strictly inside the plane
        // of the last vertex (the previous tab entry), so it is not the same
        // vertex in the FP64 sense (L1 < 3e-14) either.
      }
      __syncthreads();
      if (valid == 0) {
        if (CHECKED) { undecided = true; break; }   // degenerate turn: leave it to FP64
        if (lane == 0) {
          s_tab[ntab] = -win - 1;
          used[win >> 5] |= 1u << (win & 31);
        }
        ntab += 1;
        cur = pl + 1;   // the reference keeps a 1-based tab position here
      } else if (same) {
        if (lane == 0) {
          s_tab[ntab] = win;
          used[win >> 5] |= 1u << (win & 31);
          if (pl >= 0) s_tab[pl] = -s_tab[pl] - 1;
        }
        ntab += 1;
        cur = win;
      } else {
        if (lane == 0) {
          s_verts[nv * 3 + 0] = v[0];
          s_verts[nv * 3 + 1] = v[1];
          s_verts[nv * 3 + 2] = v[2];
          s_vpair[nv * 2 + 0] = win;
          s_vpair[nv * 2 + 1] = cur;
          s_tab[ntab] = win;
          used[win >> 5] |= 1u << (win & 31);
        }
        nv += 1;
        ntab += 1;
        cur = win;
      }
      __syncthreads();
    }
    if (CHECKED && !undecided) {
      const bool closed = ntab >= 4 && s_tab[0] == s_tab[ntab - 1];
      if (!closed || status != AM_ST_OK) undecided = true;   // truncated or stuck: FP64 decides
    }
  }
  __syncthreads();   // the last records are visible to every thread
  finish_region<T, CHECKED>(net, io, wp, rid, slot_region, smask, eq, s_tab, s_verts, s_vpair, ntab, nv, start_in,
                            start, status, undecided, undecided ? AM_WHY_BIT(AM_WHY_SCAN) : 0u, vinacc);
}

// Minimum of (t, idx) over the wave, ties to the smaller index.
// One DPP step of wave_argmin: the pair from the lane DPP control CTRL names
// (own pair where that lane is outside the row or row mask), kept if it is
// lexicographically smaller.
template <typename T, int CTRL, int ROWS, int BANKS>
__device__ __forceinline__ void dpp_argmin_step(T& t, int& idx) {
  T t2;
  if constexpr (sizeof(T) == 8) {
    const int lo = __double2loint(t), hi = __double2hiint(t);
    t2 = __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, BANKS, false),
                          __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, BANKS, false));
  } else {
    const int b = __float_as_int(t);
    t2 = __int_as_float(__builtin_amdgcn_update_dpp(b, b, CTRL, ROWS, BANKS, false));
  }
  const int i2 = __builtin_amdgcn_update_dpp(idx, idx, CTRL, ROWS, BANKS, false);
  if (t2 < t || (t2 == t && i2 < idx)) { t = t2; idx = i2; }
}

template <typename T>
__device__ __forceinline__ void wave_argmin(T& t, int& idx) {
  // Lexicographic min of (t, idx) over the wave through DPP moves (no LDS
  // round trips): lanes gather rows of 16 (row_shr 1, 2, 4, 8), then rows
  // (row_bcast 15, 31), so lane 63 ends with the minimum; it is then read
  // into every lane.  A lane whose DPP source is out of its row keeps its own
  // pair (bound_ctrl off, `old` = its own value).
  static_assert(sizeof(T) == 8 || sizeof(T) == 4, "float or double");
  dpp_argmin_step<T, 0x111, 0xf, 0xf>(t, idx);   // row_shr:1
  dpp_argmin_step<T, 0x112, 0xf, 0xf>(t, idx);   // row_shr:2
  dpp_argmin_step<T, 0x114, 0xf, 0xf>(t, idx);   // row_shr:4
  dpp_argmin_step<T, 0x118, 0xf, 0xf>(t, idx);   // row_shr:8
  dpp_argmin_step<T, 0x142, 0xa, 0xf>(t, idx);   // row_bcast:15 into rows 1 and 3
  dpp_argmin_step<T, 0x143, 0xc, 0xf>(t, idx);   // row_bcast:31 into rows 2 and 3
  idx = __builtin_amdgcn_readlane(idx, 63);
  if constexpr (sizeof(T) == 8) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(t), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(t), 63);
    t = __hiloint2double(hi, lo);
  } else {
    t = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), 63));
  }
}

// Cramer's rule with one reciprocal (the checked walks' vertex; its extra
// rounding is inside the bound's 32 u (1 + H) term).
template <typename T>
__device__ __forceinline__ void solve3_rcp(const T* p1, const T* p2, const T* p3, T* x) {
  const T A0 = p1[0], A1 = p1[1], A2 = p1[2], A3 = p2[0], A4 = p2[1], A5 = p2[2], A6 = p3[0], A7 = p3[1],
          A8 = p3[2];
  const T b0 = -p1[3], b1 = -p2[3], b2 = -p3[3];
  const T r = T(1) / AM_DET3(A0, A1, A2, A3, A4, A5, A6, A7, A8);
  x[0] = AM_DET3(b0, A1, A2, b1, A4, A5, b2, A7, A8) * r;
  x[1] = AM_DET3(A0, b0, A2, A3, b1, A5, A6, b2, A8) * r;
  x[2] = AM_DET3(A0, A1, b0, A3, A4, b1, A6, A7, b2) * r;
}

// Block-wide reductions for walks run by NW waves (one region per block).
// Each wave reduces with shuffles, then wave results meet in LDS; the combine
// runs in wave order, so the result is the same on every thread.
template <typename T, int NW>
__device__ __forceinline__ void block_argmin(T& t, int& g) {
  wave_argmin(t, g);
  if (NW == 1) return;
  __shared__ T red_t[NW];
  __shared__ int red_g[NW];
  const int wave = threadIdx.x >> 6;
  __syncthreads();                       // earlier readers of red_* are done
  if ((threadIdx.x & 63) == 0) { red_t[wave] = t; red_g[wave] = g; }
  __syncthreads();
  t = red_t[0];
  g = red_g[0];
  for (int w = 1; w < NW; ++w)
    if (red_t[w] < t || (red_t[w] == t && red_g[w] < g)) { t = red_t[w]; g = red_g[w]; }
}

// x with its sign bit xor-ed with sm's top bit (sm: 0 or 0x80000000).
__device__ __forceinline__ double flip_sign(double x, uint32_t sm) {
  uint2 u = __builtin_bit_cast(uint2, x);
  u.y ^= sm;
  return __builtin_bit_cast(double, u);
}
__device__ __forceinline__ float flip_sign(float x, uint32_t sm) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) ^ sm);
}

// Constraint planes of one region as the ray walk reads them: constraints
// g >= g0 from LDS (16-byte groups indexed g - g0); constraints g < g0, which are
// hidden layer 1's, rebuilt from its weights (row g of (W0 | b0), negated where
// the region's neuron g is active), so their LDS is not needed.
template <typename T>
struct PlaneView {
  // LDS planes of constraints g >= g0 at index i = g - g0, in 16-byte groups
  // so that one ds_read_b128 per group serves a lane: FP64, Q[2 i] = (x, y)
  // and Q[2 (Cl + i)] = (z, c) (two arrays of pairs); FP32, Q[4 i] = (x, y, z, c).
  const T* Q;
  int Cl;
  int g0;
  const double* W0;
  const double* b0;
  const uint32_t* mask;
  const T* Pe = nullptr;   // per-plane error estimates sigma (indexed g - g0), or none
  __device__ __forceinline__ T err(int g) const { return (Pe && g >= g0) ? Pe[g - g0] : T(0); }
  __device__ __forceinline__ void get(int g, T* p) const {
    if (g < g0) {
      // an active neuron's plane is negated: flip the sign bits (one xor each)
      const uint32_t sm = mbit(mask, g) ? 0x80000000u : 0u;
      const T v[4] = {T(W0[g * 3 + 0]), T(W0[g * 3 + 1]), T(W0[g * 3 + 2]), T(b0[g])};
      for (int q = 0; q < 4; ++q) p[q] = flip_sign(v[q], sm);
    } else {
      const int i = g - g0;
      if constexpr (sizeof(T) == 8) {
        const double2 a = reinterpret_cast<const double2*>(Q)[i];
        const double2 b = reinterpret_cast<const double2*>(Q)[Cl + i];
        p[0] = a.x;
        p[1] = a.y;
        p[2] = b.x;
        p[3] = b.y;
      } else {
        const float4 a = reinterpret_cast<const float4*>(Q)[i];
        p[0] = a.x;
        p[1] = a.y;
        p[2] = a.z;
        p[3] = a.w;
      }
    }
  }
};

// Nearest plane crossed by the ray x + t d (t may be negative when x lies
// outside a half-space), skipping ex1/ex2; INT_MAX when no plane bounds it.
// Each thread keeps its best crossing t = num / den (den > 0) as a fraction
// and compares by cross-multiplication, dividing once: the result is a
// proposal that the checked test certifies, so ties may break either way.
template <typename T, int NW>
__device__ __forceinline__ void shoot(const PlaneView<T>& P, int C, const T* x, const T* d, int ex1, int ex2,
                                      T& tmin, int& gmin) {
  T nb = T(1), db = T(0);   // no crossing yet
  int gb = 0x7fffffff;
#pragma unroll 2
  for (int g = threadIdx.x; g < C; g += 64 * NW) {
    if (g == ex1 || g == ex2) continue;
    T pg[4];
    P.get(g, pg);
    const T al = fmaT(pg[0], d[0], fmaT(pg[1], d[1], pg[2] * d[2]));
    if (al > T(0)) {
      const T num = -plane_eval(pg[0], pg[1], pg[2], pg[3], x);
      // num / al < nb / db  <=>  num db < nb al  (al, db > 0); db == 0: none yet
      if (db == T(0) || num * db < nb * al) { nb = num; db = al; gb = g; }
    }
  }
  T tb = db > T(0) ? nb / db : T(__builtin_inff());
  block_argmin<T, NW>(tb, gb);
  tmin = tb;
  gmin = gb;
}

// argmin (t, g) and OR of verdict bits over the block, with one LDS exchange
// and one barrier.  The exchange slots alternate between calls (`phase`, the
// same count on every thread): a slot is written again two calls later, after
// every wave has passed the barrier of the call in between, so its last
// readers are done.
template <typename T, int NW>
__device__ __forceinline__ void block_argmin_or(T& t, int& g, int& bits, int& phase) {
  wave_argmin(t, g);
  int b = (__any(bits & 1) ? 1 : 0) | (__any(bits & 2) ? 2 : 0);
  if (NW == 1) { bits = b; return; }
  __shared__ T red_t[2][NW];
  __shared__ int red_g[2][NW];
  __shared__ int red_b[2][NW];
  const int wave = threadIdx.x >> 6, sl = phase & 1;
  ++phase;
  if ((threadIdx.x & 63) == 0) { red_t[sl][wave] = t; red_g[sl][wave] = g; red_b[sl][wave] = b; }
  __syncthreads();
  t = red_t[sl][0];
  g = red_g[sl][0];
  b = red_b[sl][0];
  for (int w = 1; w < NW; ++w) {
    if (red_t[sl][w] < t || (red_t[sl][w] == t && red_g[sl][w] < g)) { t = red_t[sl][w]; g = red_g[sl][w]; }
    b |= red_b[sl][w];
  }
  bits = b;
}

// Two argmins of (t, g) over the block with one exchange and one barrier
// (slots alternate with `phase` as in block_argmin_or).
template <typename T, int NW>
__device__ __forceinline__ void block_argmin2(T& t1, int& g1, T& t2, int& g2, int& phase) {
  wave_argmin(t1, g1);
  wave_argmin(t2, g2);
  if (NW == 1) return;
  __shared__ T red_t[2][2][NW];
  __shared__ int red_g[2][2][NW];
  const int wave = threadIdx.x >> 6, sl = phase & 1;
  ++phase;
  if ((threadIdx.x & 63) == 0) {
    red_t[sl][0][wave] = t1;
    red_g[sl][0][wave] = g1;
    red_t[sl][1][wave] = t2;
    red_g[sl][1][wave] = g2;
  }
  __syncthreads();
  t1 = red_t[sl][0][0];
  g1 = red_g[sl][0][0];
  t2 = red_t[sl][1][0];
  g2 = red_g[sl][1][0];
  for (int w = 1; w < NW; ++w) {
    if (red_t[sl][0][w] < t1 || (red_t[sl][0][w] == t1 && red_g[sl][0][w] < g1)) {
      t1 = red_t[sl][0][w];
      g1 = red_g[sl][0][w];
    }
    if (red_t[sl][1][w] < t2 || (red_t[sl][1][w] == t2 && red_g[sl][1][w] < g2)) {
      t2 = red_t[sl][1][w];
      g2 = red_g[sl][1][w];
    }
  }
}

// One pass over the constraints for a proposed vertex v: its checked verdict
// bits over all constraints but c1/c2 (bit 0 certainly violated, bit 1
// undecided), and the nearest plane crossed by the ray v + t d, skipping
// s1/s2 (the next step's proposal; meaningful only when the bits are 0).
template <typename T, int NW>
__device__ __forceinline__ int check_and_shoot(const PlaneView<T>& P, int C, const T* v, const VertexBound<T>& vb,
                                               int c1, int c2, const T* d, int s1, int s2, T& tmin, int& gmin,
                                               int& phase) {
  int bits = 0;
  T nb = T(1), db = T(0);
  int gb = 0x7fffffff;
#pragma unroll 2
  for (int g = threadIdx.x; g < C; g += 64 * NW) {
    T pg[4];
    P.get(g, pg);
    const T ax = pg[0], ay = pg[1], az = pg[2], c = pg[3];
    // the constraint at the vertex, once for both uses
    const T sv = plane_eval(ax, ay, az, c, v);
    if (g != c1 && g != c2) {
      const int vg = checked_verdict_at(sv, ax, ay, az, c, P.err(g), vb);
      bits |= vg == 1 ? 1 : (vg == 2 ? 2 : 0);
    }
    if (g != s1 && g != s2) {
      const T al = fmaT(ax, d[0], fmaT(ay, d[1], az * d[2]));
      if (al > T(0)) {
        const T num = -sv;
        if (db == T(0) || num * db < nb * al) { nb = num; db = al; gb = g; }
      }
    }
  }
  T tb = db > T(0) ? nb / db : T(__builtin_inff());
  block_argmin_or<T, NW>(tb, gb, bits, phase);
  tmin = tb;
  gmin = gb;
  return bits;
}

// Nearest planes crossed by x + t d for t > 0 and for t < 0 (the ray along
// -d), skipping ex1, in one pass: (tp, gp) along +d, (tm, gm) along -d.
template <typename T, int NW>
__device__ __forceinline__ void shoot_both(const PlaneView<T>& P, int C, const T* x, const T* d, int ex1, T& tp,
                                           int& gp, T& tm, int& gm, int& phase) {
  T np = T(1), dp = T(0), nm = T(1), dmn = T(0);
  int bp = 0x7fffffff, bm = 0x7fffffff;
#pragma unroll 2
  for (int g = threadIdx.x; g < C; g += 64 * NW) {
    if (g == ex1) continue;
    T pg[4];
    P.get(g, pg);
    const T al = fmaT(pg[0], d[0], fmaT(pg[1], d[1], pg[2] * d[2]));
    if (al == T(0)) continue;
    const T num = -plane_eval(pg[0], pg[1], pg[2], pg[3], x);
    if (al > T(0)) {
      if (dp == T(0) || num * dp < np * al) { np = num; dp = al; bp = g; }
    } else {
      // along -d the crossing parameter is num / (-al)
      const T a2 = -al;
      if (dmn == T(0) || num * dmn < nm * a2) { nm = num; dmn = a2; bm = g; }
    }
  }
  T t1 = dp > T(0) ? np / dp : T(__builtin_inff());
  T t2 = dmn > T(0) ? nm / dmn : T(__builtin_inff());
  block_argmin2<T, NW>(t1, bp, t2, bm, phase);
  tp = t1; gp = bp; tm = t2; gm = bm;
}

// The ray walk's per-region records in dynamic LDS, sized by the polygon
// vertex cap vm (not by AM_VM_MAX): the tab, the vertices, their constraint
// pairs and (reduced planes, per-plane mode) their inverse systems.  Kernels
// carve them after their planes; walk_rec_bytes is the space they take.
struct WalkRecs {
  int* tab;        // [vm + 2]
  double* verts;   // [vm][3]
  int* vpair;      // [vm][2]
  float* jinv;     // [vm][9], or nullptr
};
__host__ __device__ inline size_t walk_rec_bytes(int vm, bool jinv) {
  auto a16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
  return a16((size_t)vm * 3 * 8) + a16((size_t)vm * 2 * 4) + a16((size_t)(vm + 2) * 4) +
         (jinv ? a16((size_t)vm * 9 * 4) : 0);
}
__device__ inline WalkRecs carve_walk_recs(unsigned char* p, int vm, bool jinv) {
  auto a16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
  WalkRecs r;
  r.verts = reinterpret_cast<double*>(p);
  p += a16((size_t)vm * 3 * 8);
  r.vpair = reinterpret_cast<int*>(p);
  p += a16((size_t)vm * 2 * 4);
  r.tab = reinterpret_cast<int*>(p);
  p += a16((size_t)(vm + 2) * 4);
  r.jinv = jinv ? reinterpret_cast<float*>(p) : nullptr;
  return r;
}

// Reduced-precision walk by ray shooting.  From the start plane's edge each
// step proposes the next edge plane as the nearest plane crossed along the
// current edge line (one pass over the constraints), then certifies the new
// vertex with the checked feasibility test.  On generic input the proposal
// is the unique feasible candidate the reference's scan would accept, so the
// face is the same; anything the bounds cannot decide goes to FP64.
// The per-vertex serial work is split over the waves, none of it depending on
// another's: the vertex solve, its error bound per unit of 1 + |v|_1, the next
// edge direction, and the turn test (with, for coarse planes, the FP64 vertex)
// run side by side and meet in LDS at one barrier; the passes over the
// constraints use every wave.
// eqe: the face plane's error estimate (per-plane mode: coarse planes with
// P.Pe and wp.lam > 0; the vertices are then solved in FP64 from the coarse
// planes, and a region whose vertices the estimate cannot put within
// AM_VERT_TOL gets the AM_VERT64 flag).
// start_in: the region's start plane (io.reg_start[rid]), read by the caller
// before it stages the planes, so its latency hides behind theirs.
template <typename T, int NW>
__device__ void walk_ray(const NetDev& net, const RegionIO& io, const WalkParams& wp, int rid, int slot_region,
                         const PlaneView<T>& P, const uint32_t* smask, const T* eq, const T* p, const WalkRecs& rec,
                         int start_in, T eqe = T(0)) {
  const bool per_plane = sizeof(T) == 4 && wp.lam > 0.f && P.Pe != nullptr;
  bool vinacc = false;
  int* const s_tab = rec.tab;
  double* const s_verts = rec.verts;     // vertices (re-solved in FP64 from the planes)
  int* const s_vpair = rec.vpair;
  float* const s_jinv = rec.jinv;        // per-plane mode: inverse systems for the Newton step
  __shared__ T s_v[6];                 // the vertex and the next edge direction
  __shared__ VertexBound<T> s_vb;      // its bound per unit of 1 + |v|_1
  __shared__ unsigned s_turn;          // turn test: 0 passed, else its reason bit
  __shared__ double s_vd[3];           // coarse planes: the vertex solved in FP64
  const int C = net.C, vm = net.vm;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int NONE = 0x7fffffff;
  bool undecided = false;
  unsigned why = 0u;
  int start = start_in;
  T tdummy;
  int rph = 0;   // block_argmin_or's exchange slot count

  if (start == -1) {
    // seed: project p onto the f = iso plane and shoot inside it
    const T nn = plane_nn(eq[0], eq[1], eq[2]);
    const T se = plane_eval(eq[0], eq[1], eq[2], eq[3], p) / nn;
    const T q[3] = {p[0] - se * eq[0], p[1] - se * eq[1], p[2] - se * eq[2]};
    const T ax = absT(eq[0]), ay = absT(eq[1]), az = absT(eq[2]);
    const T e[3] = {T(ax <= ay && ax <= az ? 1 : 0), T(ay < ax && ay <= az ? 1 : 0), T(az < ax && az < ay ? 1 : 0)};
    const T d0[3] = {eq[1] * e[2] - eq[2] * e[1], eq[2] * e[0] - eq[0] * e[2], eq[0] * e[1] - eq[1] * e[0]};
    int g0;
    shoot<T, NW>(P, C, q, d0, -1, -1, tdummy, g0);
    if (g0 == NONE) { undecided = true; why |= AM_WHY_BIT(AM_WHY_SEED); }
    start = g0;
  }
  int ntab = 0, nv = 0;
  int r = NONE;
  if (!undecided && (start < 0 || start >= C)) { undecided = true; why |= AM_WHY_BIT(AM_WHY_SEED); }
  if (!undecided) {
    // edge on the start plane: its line's interval inside every other half-space
    T ps[4];
    P.get(start, ps);
    const T d[3] = {ps[1] * eq[2] - ps[2] * eq[1], ps[2] * eq[0] - ps[0] * eq[2], ps[0] * eq[1] - ps[1] * eq[0]};
    const T pd[4] = {d[0], d[1], d[2], -(d[0] * p[0] + d[1] * p[1] + d[2] * p[2])};
    T x0[3];
    solve3(ps, eq, pd, x0);
    T th, tl;
    int rh, rl;
    shoot_both<T, NW>(P, C, x0, d, start, th, rh, tl, rl, rph);
    if (rh == NONE || rl == NONE || !(th + tl > T(0)) || !(x0[0] == x0[0])) {
      undecided = true;
      why |= AM_WHY_BIT(AM_WHY_START);
    }
    r = rh;
  }
  if (!undecided) {
    if (threadIdx.x == 0) s_tab[0] = start;
    ntab = 1;
    int cur = start;
    int tab1 = start, tab2 = -1;   // the last two tab entries
    while (true) {
      // r is the proposed next edge plane.  The serial work of its vertex,
      // one task per wave (all on wave 0 when the block has one):
      if (wave == 0) {   // the vertex
        T pr4[4], pc4[4];
        P.get(r, pr4);
        P.get(cur, pc4);
        T v[3];
        if (per_plane) {
          const double a4[4] = {pr4[0], pr4[1], pr4[2], pr4[3]};
          const double b4[4] = {pc4[0], pc4[1], pc4[2], pc4[3]};
          const double e4[4] = {eq[0], eq[1], eq[2], eq[3]};
          double vd[3];
          solve3(a4, b4, e4, vd);
          v[0] = T(vd[0]); v[1] = T(vd[1]); v[2] = T(vd[2]);
          if (lane == 0) { s_vd[0] = vd[0]; s_vd[1] = vd[1]; s_vd[2] = vd[2]; }
        } else {
          solve3_rcp(pr4, pc4, eq, v);
        }
        if (lane == 0) { s_v[0] = v[0]; s_v[1] = v[1]; s_v[2] = v[2]; }
      }
      if (wave == 1 % NW) {   // its error bound
        T pr4[4], pc4[4];
        P.get(r, pr4);
        P.get(cur, pc4);
        const T lam = per_plane ? T(wp.lam) : T(0);
        const VertexBound<T> vb = vertex_bound_unit<T>(pr4, pc4, eq, lam * P.err(r), lam * P.err(cur), lam * eqe,
                                                       T(wp.eps), T(wp.h_max), lam, T(wp.lam_v));
        if (lane == 0) s_vb = vb;
      }
      if (wave == 2 % NW) {   // the next edge runs along plane r, away from the plane we came along
        T pr4[4], pc4[4];
        P.get(r, pr4);
        P.get(cur, pc4);
        T dn[3] = {pr4[1] * eq[2] - pr4[2] * eq[1], pr4[2] * eq[0] - pr4[0] * eq[2], pr4[0] * eq[1] - pr4[1] * eq[0]};
        if (pc4[0] * dn[0] + pc4[1] * dn[1] + pc4[2] * dn[2] > T(0)) { dn[0] = -dn[0]; dn[1] = -dn[1]; dn[2] = -dn[2]; }
        if (lane == 0) { s_v[3] = dn[0]; s_v[4] = dn[1]; s_v[5] = dn[2]; }
      }
      if (wave == 3 % NW) {   // the turn test (the walk must certainly turn the right way)
        unsigned w = 0u;
        if (ntab >= 2) {
          T n1[4], n2[4], n3[4];
          P.get(tab2, n1);
          P.get(tab1, n2);
          P.get(r, n3);
          const T A12 = AM_DOT3(n1, n2), A23 = AM_DOT3(n2, n3), A13 = AM_DOT3(n1, n3), A22 = AM_DOT3(n2, n2);
          const T C1 = AM_DOT3(n1, eq), C2 = AM_DOT3(n2, eq), C3 = AM_DOT3(n3, eq), B = AM_DOT3(eq, eq);
          // (A12 A23 - A22 A13) - X / B > eps_turn (|A12 A23| + |A22 A13| + Y / B), times B > 0
          const T val = (A12 * A23 - A22 * A13) * B - (A12 * C2 * C3 + A23 * C1 * C2 - A13 * C2 * C2 - A22 * C1 * C3);
          const T mag = (absT(A12 * A23) + absT(A22 * A13)) * B +
                        (absT(A12 * C2 * C3) + absT(A23 * C1 * C2) + absT(A13 * C2 * C2) + absT(A22 * C1 * C3));
          if (!(val > T(wp.eps_turn) * mag)) w = AM_WHY_BIT(AM_WHY_TURN);
        }
        if (sizeof(T) == 4 && !per_plane) {
          T ar[4], ac[4];
          P.get(r, ar);
          P.get(cur, ac);
          const double a4[4] = {ar[0], ar[1], ar[2], ar[3]};
          const double b4[4] = {ac[0], ac[1], ac[2], ac[3]};
          const double e4[4] = {eq[0], eq[1], eq[2], eq[3]};
          double vd[3];
          solve3(a4, b4, e4, vd);
          if (lane == 0) { s_vd[0] = vd[0]; s_vd[1] = vd[1]; s_vd[2] = vd[2]; }
        }
        if (lane == 0) s_turn = w;
      }
      __syncthreads();
      // (the tab, the turn verdict, the vertex and its bound; LDS written
      // before this barrier is not rewritten until the pass's barriers below)
      bool in_tab = false;
      for (int i = 0; i < ntab; ++i) in_tab = in_tab || (s_tab[i] == r);
      if (in_tab && !(r == start && ntab >= 3)) { undecided = true; why |= AM_WHY_BIT(AM_WHY_TAB); break; }
      const unsigned turn = s_turn;
      if (turn != 0u) { undecided = true; why |= turn; break; }
      const T v[3] = {s_v[0], s_v[1], s_v[2]};
      const T dn[3] = {s_v[3], s_v[4], s_v[5]};
      VertexBound<T> vb = s_vb;
      scale_bound<T>(vb, T(1) + absT(v[0]) + absT(v[1]) + absT(v[2]));
      double vd[3];
      if (sizeof(T) == 8) {
        vd[0] = v[0]; vd[1] = v[1]; vd[2] = v[2];
      } else {
        vd[0] = s_vd[0]; vd[1] = s_vd[1]; vd[2] = s_vd[2];
      }
      int nx;
      const int bits = check_and_shoot<T, NW>(P, C, v, vb, cur, r, dn, r, cur, tdummy, nx, rph);
      if (bits != 0 || !vb.ok) {
        undecided = true;
        if (bits & 1) why |= AM_WHY_BIT(AM_WHY_VIOL);
        else if (bits & 2) why |= AM_WHY_BIT(AM_WHY_UNSURE);
        else why |= AM_WHY_BIT(AM_WHY_HMAX);
        break;
      }
      const bool vacc = sizeof(T) == 4 && per_plane && wp.lam_v > 0.f && vertex_accurate(vb, v);
      if (sizeof(T) == 4 && (wp.lam_v < 0.f || (per_plane && wp.lam_v > 0.f && !wp.vsel && !vacc))) vinacc = true;
      // The reference's "same vertex" test (L1 < 3e-14, FP64): with coarse
      // planes the certified check above already puts the new vertex farther
      // than that from the last one; with FP64 planes it may not, so vertices
      // that close are left to the scan.
      if (nv > 0) {
        const double l1 = fabs(vd[0] - s_verts[(nv - 1) * 3 + 0]) + fabs(vd[1] - s_verts[(nv - 1) * 3 + 1]) +
                          fabs(vd[2] - s_verts[(nv - 1) * 3 + 2]);
        if (!(l1 > 4.0 * Tol<double>::SAME_EPS)) { undecided = true; why |= AM_WHY_BIT(AM_WHY_SAME); break; }
      }
      // record the vertex (read after the next barrier)
      if (threadIdx.x == 0) {
        s_verts[nv * 3 + 0] = vd[0];
        s_verts[nv * 3 + 1] = vd[1];
        s_verts[nv * 3 + 2] = vd[2];
        s_vpair[nv * 2 + 0] = r;
        s_vpair[nv * 2 + 1] = cur;
        s_tab[ntab] = r;
        if (per_plane) {
          for (int x = 0; x < 3; ++x) {
            s_jinv[nv * 9 + x] = float(vb.n1[x]);
            s_jinv[nv * 9 + 3 + x] = float(vb.n2[x]);
            s_jinv[nv * 9 + 6 + x] = float(vb.n3[x]);
          }
          if (wp.vsel && vacc) s_jinv[nv * 9] = __builtin_nanf("");   // no Newton step needed
        }
      }
      nv += 1;
      ntab += 1;
      tab2 = tab1;
      tab1 = r;
      if (ntab >= 4 && r == start) break;          // closed
      if (ntab >= vm + 1) { undecided = true; why |= AM_WHY_BIT(AM_WHY_TRUNC); break; }   // truncated: FP64 decides
      // walk along the new edge line (found by the fused pass above)
      cur = r;
      if (nx == NONE) { undecided = true; why |= AM_WHY_BIT(AM_WHY_NEXT); break; }
      r = nx;
    }
  }
  __syncthreads();   // the last records are visible to every thread
  finish_region<T, true, double>(net, io, wp, rid, slot_region, smask, eq, s_tab, s_verts, s_vpair, ntab, nv,
                                 start_in, start, AM_ST_OK, undecided, why, vinacc,
                                 per_plane ? s_jinv : nullptr);
}

}  // namespace am
