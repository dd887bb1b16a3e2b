// FP64 region kernel: one wave64 per visited region.
//
// For one activation region it builds every constraint plane through the
// masked affine chain, orders the constraints by their signed distance to the
// region's seed point, finds the start plane (seed regions only), walks the
// polygon of f = iso inside the region edge by edge, orients it, writes the
// region's record and emits one neighbour candidate per crossed hidden-neuron
// edge.  The semantics are the reference's (backend/inc/process.h:18-264,
// backend/inc/kernel.h:799-1506) with deterministic choices where the
// reference races: the lowest feasible pair index for the start search and
// the lowest feasible candidate position in the walk.
//
// LDS holds the planes as four SoA arrays (x, y, z, c) so that a lane-strided
// sweep over constraints is bank-conflict free and a uniform read of one
// plane is a broadcast.
#include "am_device.h"
#include "am_kernels.h"

namespace am {

__global__ __launch_bounds__(64) void k_region_fp64(NetDev net, RegionIO io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = net.C, L = net.L, vm = net.vm;
  const int Cs = io.Cs;        // stride of one SoA plane array (>= C)
  const int N2 = io.Csort;     // power of two >= C
  double* Px = reinterpret_cast<double*>(smem_raw);
  double* Py = Px + Cs;
  double* Pz = Py + Cs;
  double* Pc = Pz + Cs;
  double* keys = Pc + Cs;
  int* order = reinterpret_cast<int*>(keys + N2);
  uint32_t* used = reinterpret_cast<uint32_t*>(order + N2);
  uint32_t* smask = used + ((C + 31) >> 5);

  __shared__ int s_tab[AM_VM_MAX + 2];
  __shared__ double s_verts[AM_VM_MAX * 3];
  __shared__ int s_vpair[AM_VM_MAX * 2];
  __shared__ double s_eq[4];

  const int lane = threadIdx.x;
  const int rid = io.first + blockIdx.x;
  if (blockIdx.x >= io.count) return;

  for (int w = lane; w < net.Wd; w += 64) smask[w] = io.reg_mask[(size_t)rid * net.Wd + w];
  for (int w = lane; w < ((C + 31) >> 5); w += 64) used[w] = 0u;
  const int start_in = io.reg_start[rid];
  double p[3];
  p[0] = io.reg_point[(size_t)rid * 3 + 0];
  p[1] = io.reg_point[(size_t)rid * 3 + 1];
  p[2] = io.reg_point[(size_t)rid * 3 + 2];

  // ---- planes: hidden node 1 is W0 itself
  for (int j = lane; j < net.nodes[1]; j += 64) {
    Px[j] = net.W0[j * 3 + 0];
    Py[j] = net.W0[j * 3 + 1];
    Pz[j] = net.W0[j * 3 + 2];
    Pc[j] = net.b[0][j];
  }
  __syncthreads();
  // ---- masked affine chain, one output neuron per lane and sub-step
  for (int li = 1; li <= net.D; ++li) {
    const int nin = net.nodes[li], nout = net.nodes[li + 1], in0 = net.off[li - 1];
    const double* WT = net.WT[li];
    const double* bb = net.b[li];
    const bool last = (li == net.D);
    const int dst0 = last ? 0 : net.off[li];
    for (int jb = 0; jb < nout; jb += 256) {
      double ax[4], ay[4], az[4], ac[4];
      int jj[4];
      bool ok[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        jj[t] = jb + lane + 64 * t;
        ok[t] = jj[t] < nout;
        ax[t] = 0.0; ay[t] = 0.0; az[t] = 0.0;
        ac[t] = ok[t] ? bb[jj[t]] : 0.0;
      }
      for (int k = 0; k < nin; ++k) {
        if (!mbit(smask, in0 + k)) continue;
        const double px = Px[in0 + k], py = Py[in0 + k], pz = Pz[in0 + k], pc = Pc[in0 + k];
        const double* wrow = WT + (size_t)k * nout;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (ok[t]) {
            const double w = wrow[jj[t]];
            ax[t] = fmaT(w, px, ax[t]);
            ay[t] = fmaT(w, py, ay[t]);
            az[t] = fmaT(w, pz, az[t]);
            ac[t] = fmaT(w, pc, ac[t]);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (!ok[t]) continue;
        if (last) {
          s_eq[0] = ax[t]; s_eq[1] = ay[t]; s_eq[2] = az[t]; s_eq[3] = ac[t];
        } else {
          const int g = dst0 + jj[t];
          Px[g] = ax[t]; Py[g] = ay[t]; Pz[g] = az[t]; Pc[g] = ac[t];
        }
      }
    }
    __syncthreads();
    if (!last) {
      // additive skips into this node (backend/inc/process.h:73-146)
      for (int s = 0; s < net.nskip[li]; ++s) {
        const int from = net.skip_from[li][s], tm = net.skip_tm[li][s];
        const bool ident = net.tm_h[tm] == 0 && net.tm_w[tm] == 0;
        if (from == 0) {
          if (ident) {
            if (lane < 3 && lane < nout) {
              const int g = dst0 + lane;
              if (lane == 0) Px[g] = Px[g] + 1.0;
              if (lane == 1) Py[g] = Py[g] + 1.0;
              if (lane == 2) Pz[g] = Pz[g] + 1.0;
            }
          } else {
            const double* T = net.tm[tm];
            for (int j = lane; j < net.tm_h[tm]; j += 64) {
              const int g = dst0 + j;
              Px[g] = Px[g] + T[j * 3 + 0];
              Py[g] = Py[g] + T[j * 3 + 1];
              Pz[g] = Pz[g] + T[j * 3 + 2];
            }
          }
        } else {
          const int os = net.off[from - 1], dec = net.nodes[from];
          if (ident) {
            for (int j = lane; j < dec; j += 64) {
              if (!mbit(smask, os + j)) continue;
              const int g = dst0 + j;
              Px[g] = Px[g] + Px[os + j];
              Py[g] = Py[g] + Py[os + j];
              Pz[g] = Pz[g] + Pz[os + j];
              Pc[g] = Pc[g] + Pc[os + j];
            }
          } else {
            const double* TT = net.tmT[tm];
            const int h = net.tm_h[tm], w = net.tm_w[tm];
            for (int j = lane; j < h; j += 64) {
              const int g = dst0 + j;
              double a0 = Px[g], a1 = Py[g], a2 = Pz[g], a3 = Pc[g];
              for (int k = 0; k < w; ++k) {
                if (!mbit(smask, os + k)) continue;
                const double t = TT[(size_t)k * h + j];
                a0 = fmaT(t, Px[os + k], a0);
                a1 = fmaT(t, Py[os + k], a1);
                a2 = fmaT(t, Pz[os + k], a2);
                a3 = fmaT(t, Pc[os + k], a3);
              }
              Px[g] = a0; Py[g] = a1; Pz[g] = a2; Pc[g] = a3;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  if (lane == 0) s_eq[3] = s_eq[3] - net.iso;
  // sign flip (backend/inc/process.h:181-186) and extra planes (:189-192)
  for (int g = lane; g < L; g += 64) {
    if (mbit(smask, g)) { Px[g] = -Px[g]; Py[g] = -Py[g]; Pz[g] = -Pz[g]; Pc[g] = -Pc[g]; }
  }
  for (int e = lane; e < net.E; e += 64) {
    Px[L + e] = net.ew[e * 3 + 0];
    Py[L + e] = net.ew[e * 3 + 1];
    Pz[L + e] = net.ew[e * 3 + 2];
    Pc[L + e] = net.eb[e];
  }
  __syncthreads();
  const double eq[4] = {s_eq[0], s_eq[1], s_eq[2], s_eq[3]};

  // ---- distance keys s|s|/|a|^2 (orders like s/|a|), stable descending sort
  for (int g = lane; g < N2; g += 64) {
    double key = -__builtin_inf();
    if (g < C) {
      const double s = plane_eval(Px[g], Py[g], Pz[g], Pc[g], p);
      const double nn = plane_nn(Px[g], Py[g], Pz[g]);
      key = nn > 0.0 ? s * absT(s) / nn : (s > 0.0 ? __builtin_inf() : -__builtin_inf());
      if (key != key) key = -__builtin_inf();  // NaN sorts last
    }
    keys[g] = key;
    order[g] = g;
  }
  __syncthreads();
  for (int k = 2; k <= N2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < N2; i += 64) {
        const int l = i ^ j;
        if (l > i) {
          const double ki = keys[i], kl = keys[l];
          const int oi = order[i], ol = order[l];
          const bool i_first = (ki > kl) || (ki == kl && oi < ol);
          if (((i & k) == 0) != i_first) {
            keys[i] = kl; keys[l] = ki;
            order[i] = ol; order[l] = oi;
          }
        }
      }
      __syncthreads();
    }
  }

  int start = start_in;
  int status = AM_ST_OK;
  // ---- seed start search: lowest feasible pair index (backend/inc/kernel.h:799-856)
  if (start == -1) {
    const int xmax = min((C - 1) * (C - 2), 812);
    for (int xb = 0; xb <= xmax; xb += 64) {
      const int x = xb + lane;
      bool okx = x <= xmax;
      int ri = 0, rj = 0;
      double v[3] = {0.0, 0.0, 0.0};
      if (okx) {
        int i, j;
        get_i_j(x, C, i, j);
        ri = order[i];
        rj = order[j];
        const double pi4[4] = {Px[ri], Py[ri], Pz[ri], Pc[ri]};
        const double pj4[4] = {Px[rj], Py[rj], Pz[rj], Pc[rj]};
        solve3(pi4, pj4, eq, v);
      }
      for (int g = 0; g < C; ++g) {
        if (!__any(okx)) break;
        if (okx && g != ri && g != rj && violates(Px[g], Py[g], Pz[g], Pc[g], v)) okx = false;
      }
      const unsigned long long bal = __ballot(okx);
      if (bal) {
        const int wl = __builtin_ctzll(bal);
        start = __shfl(ri, wl);
        break;
      }
    }
    if (start == -1) status = AM_ST_NO_START;
  }

  int ntab = 0, nv = 0;
  int reversed = 0;
  if (status == AM_ST_OK) {
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
      double v[3] = {0.0, 0.0, 0.0};
      const double pc4[4] = {Px[cur], Py[cur], Pz[cur], Pc[cur]};
      const int first_entry = s_tab[0];
      for (int w = 0; w < C; ++w) {
        const int r = order[w];
        const bool cand = (r == first_entry && ntab >= 3) || !((used[r >> 5] >> (r & 31)) & 1u);
        if (!cand) continue;
        const double pr4[4] = {Px[r], Py[r], Pz[r], Pc[r]};
        double vt[3];
        solve3(pr4, pc4, eq, vt);
        bool bad = false;
        for (int g = lane; g < C; g += 64) {
          if (g == cur || g == r) continue;
          if (violates(Px[g], Py[g], Pz[g], Pc[g], vt)) { bad = true; break; }
        }
        if (!__any(bad)) {
          win = r;
          v[0] = vt[0]; v[1] = vt[1]; v[2] = vt[2];
          break;
        }
      }
      if (win < 0) { status = AM_ST_GAVE_UP; break; }
      int pl = -1, pr = -1, cnt = 0;
      for (int idx = ntab - 1; idx >= 0 && cnt < 2; --idx) {
        if (s_tab[idx] >= 0) { if (cnt == 0) pl = idx; else pr = idx; ++cnt; }
      }
      int valid = -1;
      if (cnt == 2) {
        const int a = s_tab[pr], b = s_tab[pl];
        const double n1[3] = {Px[a], Py[a], Pz[a]};
        const double n2[3] = {Px[b], Py[b], Pz[b]};
        const double n3[3] = {Px[win], Py[win], Pz[win]};
        valid = pivot_valid(eq, n1, n2, n3) ? 1 : 0;
      }
      bool same = false;
      if (valid != 0 && nv > 0) {
        const double last[3] = {s_verts[(nv - 1) * 3 + 0], s_verts[(nv - 1) * 3 + 1], s_verts[(nv - 1) * 3 + 2]};
        same = same_vertex(last, v);
      }
      __syncthreads();
      if (valid == 0) {
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
    // ---- orientation (backend/inc/kernel.h:1325-1371)
    if (nv >= 3) {
      const double* q = s_verts;
      const double ax = q[3] - q[0], ay = q[4] - q[1], az = q[5] - q[2];
      const double bx = q[6] - q[3], by = q[7] - q[4], bz = q[8] - q[5];
      const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
      const double dir = cx * eq[0] + cy * eq[1] + cz * eq[2];
      reversed = ((dir < 0.0) != (net.flip != 0)) ? 1 : 0;
      __syncthreads();
      if (reversed && lane < nv / 2) {
        const int a = lane, b = nv - 1 - lane;
        for (int c = 0; c < 3; ++c) {
          const double t = s_verts[a * 3 + c];
          s_verts[a * 3 + c] = s_verts[b * 3 + c];
          s_verts[b * 3 + c] = t;
        }
        for (int c = 0; c < 2; ++c) {
          const int t = s_vpair[a * 2 + c];
          s_vpair[a * 2 + c] = s_vpair[b * 2 + c];
          s_vpair[b * 2 + c] = t;
        }
      }
      __syncthreads();
    }
  }

  // ---- region record
  if (lane == 0) {
    int* inf = io.reg_info + (size_t)rid * AM_INFO_STRIDE;
    inf[0] = start_in; inf[1] = start; inf[2] = nv; inf[3] = ntab;
    inf[4] = reversed; inf[5] = status; inf[6] = 0; inf[7] = 0;
  }
  for (int i = lane; i < vm + 1; i += 64) io.reg_tab[(size_t)rid * (vm + 1) + i] = i < ntab ? s_tab[i] : 0;
  for (int i = lane; i < vm * 3; i += 64) io.reg_verts[(size_t)rid * vm * 3 + i] = i < nv * 3 ? s_verts[i] : 0.0;
  for (int i = lane; i < vm * 2; i += 64) io.reg_vpair[(size_t)rid * vm * 2 + i] = i < nv * 2 ? s_vpair[i] : 0;

  // ---- neighbour candidates (backend/inc/kernel.h:1390-1474)
  const size_t slot0 = (size_t)blockIdx.x * vm;
  if (lane < vm) {
    const int m = lane;
    int t = -1;
    if (m + 1 < ntab) t = s_tab[m];
    const bool valid = t >= 0 && t < L;
    double mid[3] = {0.0, 0.0, 0.0};
    if (valid) {
      const int m_ = (m == 0) ? ntab - 1 : m;
      int cnt = -1, v1 = -1;
      for (int i = 1; i < ntab; ++i) {
        if (s_tab[i] >= 0) ++cnt;
        if (i == m_) v1 = cnt;
      }
      if (nv > 0 && v1 >= 0) {
        const int v2 = (v1 + 1) % nv;
        for (int c = 0; c < 3; ++c) mid[c] = (s_verts[v1 * 3 + c] + s_verts[v2 * 3 + c]) / 2.0;
      }
    }
    io.cand_start[slot0 + m] = valid ? t : -2;
    io.cand_point[(slot0 + m) * 3 + 0] = mid[0];
    io.cand_point[(slot0 + m) * 3 + 1] = mid[1];
    io.cand_point[(slot0 + m) * 3 + 2] = mid[2];
  }
  const int Wd = net.Wd;
  for (int idx = lane; idx < vm * Wd; idx += 64) {
    const int m = idx / Wd, w = idx - m * Wd;
    int t = -1;
    if (m + 1 < ntab) t = s_tab[m];
    if (t >= 0 && t < L) {
      uint32_t word = smask[w];
      if ((t >> 5) == w) word ^= 1u << (t & 31);
      io.cand_mask[(slot0 + m) * Wd + w] = word;
    }
  }
}

}  // namespace am
