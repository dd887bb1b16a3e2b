// Host side of the engine: the C-ABI of include/am.h.
//
// The marching loop is breadth first: a level is the set of regions found by
// the previous level.  Per level the engine launches the region kernel (one
// wave64 per region), inserts the level's neighbour candidates into the
// visited set, and compacts the new regions into the region store, which
// becomes the next level.  The reference instead pops LIFO batches of 1024
// (backend/src/cuam_kernel.cu:53-98); both orders reach the same regions.
#include <cstring>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <thread>
#include <climits>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <deque>
#include <vector>

#include "../../include/am.h"
#include "am_hash.h"
#include "am_kernels.h"
#include "am_fast.h"
#include "am_adjoint.h"
#include "am_newton.h"
#include "am_combine.h"
#include "am_scan.h"

using namespace am;

namespace {

std::string g_create_error;

// One device allocation, owned: moves hand it over, copies are not allowed
// (a copy would free the same pointer twice).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct HipErr {
  hipError_t e;
  const char* what;
};

#define HIPCHK(x)                                 \
  do {                                            \
    hipError_t e_ = (x);                          \
    if (e_ != hipSuccess) throw HipErr{e_, #x};   \
  } while (0)

struct ArgErr {
  std::string msg;
};
struct CapErr {
  std::string msg;
};

template <typename T>
void ensure(DevBuf<T>& b, size_t n, bool keep, hipStream_t st) {
  if (b.n >= n) return;
  size_t cap = std::max<size_t>(n, b.n ? b.n * 2 : 1024);
  T* q = nullptr;
  HIPCHK(hipMalloc(&q, cap * sizeof(T)));
  if (keep && b.p && b.n) HIPCHK(hipMemcpyAsync(q, b.p, b.n * sizeof(T), hipMemcpyDeviceToDevice, st));
  if (b.p) {
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipFree(b.p));
  }
  b.p = q;
  b.n = cap;
}

inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// As ensure, for state that must read zero before its first use: a grown
// buffer is cleared by a fill kernel (no memset node, should a caller capture
// the stream later on).
void ensure_zeroed(DevBuf<unsigned long long>& b, size_t n, hipStream_t st) {
  if (b.n >= n) return;
  ensure(b, n, false, st);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(b.n * 2, 256)), dim3(256), 0, st, reinterpret_cast<uint32_t*>(b.p),
                     b.n * 2, 0u);
  HIPCHK(hipGetLastError());
}

}  // namespace

struct am_ctx {
  // architecture
  std::vector<int> nodes, off;
  int D = 0, L = 0, E = 0, C = 0, Wd = 0, vm = 20, precision = AM_PREC_FP64;
  std::vector<std::vector<std::pair<int, int>>> skips;  // per linear layer index
  std::string err;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;   // used when the caller passes no stream
  int dev = 0;
  size_t max_lds = 65536;
  long long max_regions = -1;
  long long region_limit = -1;   // AM_ERR_CAPACITY past this many regions (-1: the engine's own limit)
  int ray_waves = 4;             // waves per region of the FP64 ray walk (2, 4, 8)
  bool graph_levels = false;     // each level's launches captured into a graph and launched as one
  hipGraphExec_t gexec[2] = {nullptr, nullptr};   // alternate levels: one may still run while the other is updated
  long long graph_instantiations = 0, graph_updates = 0;
  int scan_blocks = 2048;        // grid of the candidate scan over undecided regions
  bool adjoint_tiled = true;   // am_adjoint: tiled chain + GEMM pull-back where the tiled chain runs
  bool adjoint_mfma = true;    // am_adjoint's GEMMs on the FP64 matrix cores (else VALU fmas)
  int table_log2 = 22;

  // network (device)
  DevBuf<double> W0, ew, eb;
  std::vector<DevBuf<double>> WT, Wrow, b, tm, tmT;
  std::vector<int> tm_h, tm_w;
  NetDev net{};
  FastNet fast{};
  FastBuffers fbuf;

  // region store
  DevBuf<uint32_t> reg_mask;
  DevBuf<int> reg_start, reg_info, reg_tab, reg_vpair;
  DevBuf<double> reg_point, reg_verts;
  size_t reg_cap = 0;
  long long R = 0;

  // level candidates
  DevBuf<uint32_t> cand_mask;
  DevBuf<int> cand_start, slot, rank;
  DevBuf<double> cand_point;
  DevBuf<uint64_t> cand_hash;
  DevBuf<int> scan_part;
  DevBuf<unsigned long long> scan_state;   // tile states of the visited-set stage's one-pass scan
  DevBuf<unsigned> rflags;

  // visited table
  DevBuf<uint32_t> table;
  size_t table_cap = 0;

  // mesh
  bool has_run = false, has_mesh = false;
  // am_combine's workspace (kept between calls: the adjoint reads its item
  // offsets and first-occurrence flags) and the mesh in device memory
  DevBuf<char> cw_reg, cw_item;
  CombineWork cwork;
  // cwork holds the vertex dedup of the current region store (the FP32 form's
  // Newton step makes it; am_combine reuses it): faces and vertex uses
  bool dedup_ready = false;
  long long dedup_F = 0, dedup_N = 0;
  DevBuf<int> m_face_off, m_face_idx;
  DevBuf<double> m_xyz;
  long long mesh_V = 0, mesh_F = 0, mesh_NI = 0;
  // host copy of the mesh, made on first need (am_export_ply)
  bool mesh_on_host = false;
  std::vector<double> mesh_v;
  std::vector<int> mesh_off, mesh_idx;

  // staging for callers' arrays in host memory (dev_view)
  std::deque<DevBuf<char>> stage;
  size_t stage_used = 0;

  // pinned, coherent host word the device writes scan totals into (arm_host_total)
  int* host_total = nullptr;
  int* host_total_dev = nullptr;
  // pinned staging for am_get_mesh into pageable host arrays (grown, kept)
  unsigned char* mesh_pin = nullptr;
  size_t mesh_pin_bytes = 0;
  std::vector<hipEvent_t> mesh_ev;   // one per 2 MB chunk of the staged copy (created once, kept)

  // tiled FP64 chain (nets without skips, layers up to AM_CHAIN_MAXW wide)
  bool tile_ok = false;
  bool chain_mfma = true;     // the certified forms' FP64 chain on the matrix cores ("chain_mfma")
  DevBuf<double> tplanes, teq;
  // am_adjoint's buffers, kept from call to call: dL/dv, dW, db, G and P per layer
  DevBuf<double> adj_gv, adj_dv;
  DevBuf<double> adj_part;   // per-block partial sums of dW / db, added in a fixed order
  DevBuf<int> adj_dn, adj_dg;
  std::vector<DevBuf<double>> adj_dW, adj_db, adj_G, adj_P;
  DevBuf<int> tslots, tcnt;               // undecided slots of a chunk and their count
  DevBuf<int> redo_ids, redo_cnt;         // regions a reduced-precision walk left to FP64, their count
  DevBuf<unsigned long long> redo_state;  // tile state of the one-pass scan listing them (zero between uses)
  DevBuf<uint16_t> reg_jinv;              // [R][vm][9] bf16: the reduced walk's inverse vertex systems
  bool jinv_on = false;                   // this run keeps them (reduced form, Newton vertices)
  DevBuf<int> runacc;                     // per run: FP64 ray walk undecided count, then per reason
  StageTimer timer;
  std::vector<DevBuf<double>> WTp, bpad;   // W_li^T and b_li with n_out padded to AM_CR (zeros)

  double stats[AM_STAT_COUNT] = {0};
};

namespace {

void grow_regions(am_ctx* c, size_t need) {
  if (need <= c->reg_cap) return;
  size_t cap = std::max<size_t>(need, c->reg_cap ? c->reg_cap * 2 : (1u << 14));
  const hipStream_t st = c->stream;
  ensure(c->reg_mask, cap * c->Wd, true, st);
  ensure(c->reg_start, cap, true, st);
  ensure(c->reg_point, cap * 3, true, st);
  ensure(c->reg_info, cap * AM_INFO_STRIDE, true, st);
  ensure(c->reg_tab, cap * (c->vm + 1), true, st);
  ensure(c->reg_verts, cap * c->vm * 3, true, st);
  ensure(c->reg_vpair, cap * c->vm * 2, true, st);
  if (c->jinv_on) ensure(c->reg_jinv, cap * c->vm * 9, true, st);
  c->reg_cap = cap;
}

void grow_cands(am_ctx* c, size_t n) {
  const hipStream_t st = c->stream;
  ensure(c->cand_mask, n * c->Wd, false, st);
  ensure(c->cand_start, n, false, st);
  ensure(c->cand_point, n * 3, false, st);
  ensure(c->cand_hash, n, false, st);
  ensure(c->slot, n, false, st);
  ensure(c->rank, n, false, st);
}

// The level's scan (hand-written, am_scan.h) over the new-region flags: the
// total is written by the device straight into a pinned, coherent host word,
// which the host then polls (wait_total): a short tight spin, then a yield
// between polls so that host threads driving other contexts (a batch of
// networks) keep their cores, and a stream sync after 100 ms, which also
// reports a failed kernel.  Kernels enqueued after the scan may still run when
// the word is read; later work on the stream stays ordered behind them.
void arm_host_total(am_ctx* c) {
  if (!c->host_total) {
    HIPCHK(hipHostMalloc((void**)&c->host_total, 256, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void**)&c->host_total_dev, c->host_total, 0));
  }
  *reinterpret_cast<volatile int*>(c->host_total) = -1;
}

int wait_total(am_ctx* c) {
  volatile int* h = reinterpret_cast<volatile int*>(c->host_total);
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned spins = 0; *h < 0; ++spins) {
    if (spins < 4096) continue;
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(100)) {
      HIPCHK(hipStreamSynchronize(c->stream));
      break;
    }
    std::this_thread::yield();
  }
  return *h;
}

void ensure_table(am_ctx* c, size_t entries) {
  size_t need = c->table_cap ? c->table_cap : (size_t(1) << c->table_log2);
  while (entries * 2 > need) need *= 2;
  if (need > (size_t(1) << 31)) throw CapErr{"visited set would exceed 2^31 slots"};
  if (need == c->table_cap) return;
  const hipStream_t st = c->stream;
  c->table.release();
  ensure(c->table, need, false, st);
  c->table_cap = need;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(need, 256)), dim3(256), 0, st, c->table.p, need, 0xFFFFFFFFu);
  if (c->R > 0)
    hipLaunchKernelGGL(k_rehash, dim3(blocks(c->R, 256)), dim3(256), 0, st, c->reg_mask.p, (int)c->R, c->Wd,
                       c->table.p, (uint32_t)(need - 1));
  HIPCHK(hipGetLastError());
}

// Visited-set stage of a level, in three parts so that the launches can be
// captured into a graph: insert_prepare grows the table and the region store
// (host work: allocations, the rehash), insert_launch enqueues the kernels
// (nothing else), insert_finish waits for the count and commits late where the
// store had to grow after it.
struct InsertPlan {
  size_t n = 0;
  bool early = false;
  int R0 = 0;
};

InsertPlan insert_prepare(am_ctx* c, size_t n) {
  InsertPlan p;
  p.n = n;
  p.R0 = (int)c->R;
  if (n == 0) return p;
  if (n > (size_t)INT_MAX / 2) throw CapErr{"too many candidates in one level"};
  ensure_table(c, (size_t)c->R + n);
  // Where the region store holds the worst case (every candidate new; grown
  // to it for levels up to 4M candidates), the new regions are committed
  // before the host learns how many there are: one wait per level, no host
  // round trip between the count and the commit.
  p.early = (size_t)c->R + n <= c->reg_cap || n <= ((size_t)1 << 22);
  if (p.early) grow_regions(c, (size_t)c->R + n);
  ensure(c->scan_state, scan_commit_tiles(n) + 1, false, c->stream);
  arm_host_total(c);
  return p;
}

CommitIO commit_io(const am_ctx* c, int R0) {
  CommitIO cio;
  cio.slot = c->slot.p;
  cio.cand_mask = c->cand_mask.p;
  cio.cand_start = c->cand_start.p;
  cio.cand_point = c->cand_point.p;
  cio.reg_mask = c->reg_mask.p;
  cio.reg_start = c->reg_start.p;
  cio.reg_point = c->reg_point.p;
  cio.table = c->table.p;
  cio.Wd = c->Wd;
  cio.R0 = R0;
  return cio;
}

void insert_launch(am_ctx* c, const InsertPlan& p) {
  if (p.n == 0) return;
  const hipStream_t st = c->stream;
  const size_t n = p.n;
  InsertIO io;
  io.cand_mask = c->cand_mask.p;
  io.cand_hash = c->cand_hash.p;
  io.cand_start = c->cand_start.p;
  io.reg_mask = c->reg_mask.p;
  io.table = c->table.p;
  io.slot = c->slot.p;
  io.n = (int)n;
  io.Wd = c->Wd;
  io.cap_mask = (uint32_t)(c->table_cap - 1);
  const size_t tiles = scan_commit_tiles(n);
  io.scan_state = c->scan_state.p;
  io.scan_words = (int)tiles + 1;
  hipLaunchKernelGGL(k_insert, dim3(blocks(n, 256)), dim3(256), 0, st, io);
  HIPCHK(hipGetLastError());
  // the new-region flags' scan in one pass; with the store already large
  // enough it commits the new regions too, before the host learns the count
  ScanCommitIO sio;
  sio.cio = commit_io(c, p.R0);
  sio.state = c->scan_state.p;
  sio.n = (int)n;
  sio.tiles = (int)tiles;
  sio.rank = p.early ? nullptr : c->rank.p;
  sio.total = c->host_total_dev;
  hipLaunchKernelGGL(k_scan_commit, dim3((unsigned)tiles), dim3(kScanCommitThreads), 0, st, sio);
  HIPCHK(hipGetLastError());
}

int insert_finish(am_ctx* c, const InsertPlan& p) {
  if (p.n == 0) return 0;
  const int total = wait_total(c);
  if (total == 0) return 0;
  // am_combine indexes (region, vertex slot) items with 32-bit ints: R * vm must stay below 2^31
  long long limit = std::min<long long>((long long)INT_MAX / 2, (long long)INT_MAX / c->vm);
  if (c->region_limit >= 0) limit = std::min(limit, c->region_limit);
  if ((long long)c->R + total > limit)
    throw CapErr{"region count would exceed the limit of " + std::to_string(limit) + " regions"};
  if (!p.early) {
    grow_regions(c, (size_t)c->R + total);
    const CommitIO cio = commit_io(c, p.R0);
    hipLaunchKernelGGL(k_commit, dim3(blocks(p.n, 256)), dim3(256), 0, c->stream, cio, (const int*)c->rank.p,
                       (int)p.n);
    HIPCHK(hipGetLastError());
  }
  c->R += total;
  return total;
}

// Capture the launches `enqueue` makes on the context's stream into a graph
// and launch it.  Levels alternate between two executable graphs: the one of
// two levels back is updated in place when the kernels are the same (only
// arguments and grids differ), else instantiated anew.  (When a level's count
// reaches the host, that level's last kernels may still run; the graph two
// levels back has finished.)  The timer is off meanwhile (no event nodes) and
// nothing in the capture allocates, copies synchronously or sets memory.
template <typename F>
void launch_as_graph(am_ctx* c, int parity, F&& enqueue) {
  hipGraphExec_t& exec = c->gexec[parity & 1];
  const hipStream_t st = c->stream;
  hipGraph_t g = nullptr;
  c->timer.enabled = false;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    enqueue();
  } catch (...) {
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    c->timer.enabled = true;
    throw;
  }
  HIPCHK(hipStreamEndCapture(st, &g));
  c->timer.enabled = true;
  if (exec) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(exec, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
      c->graph_updates += 1;
    } else {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(exec);
      exec = nullptr;
    }
  }
  if (!exec) {
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      throw HipErr{e, "hipGraphInstantiate"};
    }
    c->graph_instantiations += 1;
  }
  const hipError_t e = hipGraphLaunch(exec, st);
  (void)hipGraphDestroy(g);
  HIPCHK(e);
}

// Insert n candidates (already in cand_*) and append the new ones as regions.
int insert_level(am_ctx* c, size_t n) {
  const InsertPlan p = insert_prepare(c, n);
  insert_launch(c, p);
  return insert_finish(c, p);
}

size_t region_lds_bytes(const am_ctx* c, int& Cs, int& Csort) {
  Cs = c->C;
  Csort = AM_TOPK;   // the walk's top-64 order lives in keys/order[0, 64)
  while (Csort < c->C) Csort <<= 1;
  return (size_t)4 * Cs * sizeof(double) + (size_t)Csort * (sizeof(double) + sizeof(int)) +
         (size_t)((c->C + 31) / 32) * 4 + (size_t)c->Wd * 4 + 64;
}

void launch_region_fp64(am_ctx* c, int first, int count, const int* ids = nullptr) {
  int Cs, Csort;
  const size_t lds = region_lds_bytes(c, Cs, Csort);
  RegionIO io;
  io.reg_mask = c->reg_mask.p;
  io.reg_start = c->reg_start.p;
  io.reg_point = c->reg_point.p;
  io.reg_info = c->reg_info.p;
  io.reg_tab = c->reg_tab.p;
  io.reg_verts = c->reg_verts.p;
  io.reg_vpair = c->reg_vpair.p;
  io.cand_mask = c->cand_mask.p;
  io.cand_start = c->cand_start.p;
  io.cand_point = c->cand_point.p;
  io.cand_hash = c->cand_hash.p;
  io.first = first;
  io.count = count;
  io.Cs = Cs;
  io.Csort = Csort;
  io.flags = nullptr;
  io.ids = ids;
  const size_t m0 = c->timer.mark(c->stream);
  hipLaunchKernelGGL(k_region_fp64, dim3(count), dim3(64), lds, c->stream, c->net, io);
  HIPCHK(hipGetLastError());
  c->timer.span(m0, c->timer.mark(c->stream), T_REGION64);
}

size_t chain_lds_bytes(const am_ctx* c) {
  return ((size_t)AM_CXR * AM_CXS + 1) * sizeof(double) + (size_t)c->D * AM_KLS * sizeof(uint16_t) +
         (size_t)AM_CT * c->Wd * 4 + 16;
}

// LDS of k_chain_tile_mfma (half tiles: AM_HT regions)
size_t chain_half_lds_bytes(const am_ctx* c) {
  return ((size_t)AM_CXR * AM_HXS + 1) * sizeof(double) + (size_t)c->D * AM_KLS * sizeof(uint16_t) +
         (size_t)AM_HT * c->Wd * 4 + 16;
}

// Per-run counters of the FP64 stages (kernel times go to the context's timer;
// the ray walk's undecided count and reasons accumulate on the device).
struct Fp64Counts {
  long long scanned_host = 0, chain_launches = 0;
};

// Checked FP64 ray walk: planes are exact FP64, so eps only covers the walk's
// own rounding (the vertex solve's share is added per system).
const WalkParams kRay64{1e-13f, 1e-12f, 1e6f};
// The same walk on planes from the matrix-core chain (another summation
// order: up to a few 1e-14 of a plane's largest coefficient from the
// sequential-fma planes on the 4x256 net); the bound keeps a margin of 10x.
const WalkParams kRay64Mf{1e-12f, 1e-11f, 1e6f};

// planes past hidden layer 1 (the walk rebuilds layer 1's) and the mask
size_t ray64_lds_bytes(const am_ctx* c) {
  return (size_t)4 * (c->C - c->nodes[1]) * sizeof(double) + (((size_t)c->Wd * 4 + 15) & ~(size_t)15) +
         walk_rec_bytes(c->vm, false);
}

// FP64 planes + walk of `count` regions (region ids first.., or ids[0..count)),
// through the tiled chain in chunks; candidate slots and flags are numbered
// from `lvl_first`.  mode 0: the candidate scan on every region (the FP64
// records of the CPU restatement, bit for bit); mode 1: the checked FP64 ray
// walk, and the scan on the regions it leaves undecided, reusing their planes
// (the scan runs a block-stride loop up to the device-side count); mode 2:
// after a reduced-precision walk that flagged the level in rflags
// (`undecided` of them, -1 when the count is not known on the host): the
// regions it decided get FP64 vertices
// (k_vertex_fp64, their other records stay as written), the others go
// through mode 1.
// Nothing here waits for the device.  Falls back to the one-region kernel
// where the tiled chain does not apply.
// The tiled chain's arguments for regions first .. first + n - 1 (or ids),
// planes into the context's tile buffers with component stride Ch.
ChainIO chain_io(const am_ctx* c, const int* ids, int first, int n, int Ch) {
  ChainIO cio;
  cio.reg_mask = c->reg_mask.p;
  cio.ids = ids;
  cio.first = first;
  cio.count = n;
  cio.Cs = Ch;
  cio.planes = c->tplanes.p;
  cio.eq = c->teq.p;
  for (int li = 0; li < AM_MAX_LINEAR; ++li) {
    cio.Wp[li] = li >= 1 && li <= c->D ? c->WTp[li].p : nullptr;
    cio.bp[li] = li >= 1 && li <= c->D ? c->bpad[li].p : nullptr;
    cio.noutp[li] = li >= 1 && li <= c->D ? (c->nodes[li + 1] + AM_CR - 1) / AM_CR * AM_CR : 0;
  }
  return cio;
}

// The FP64 ray walk with the context's waves per region.
void launch_ray64(am_ctx* c, int n, size_t lds, const RegionIO& io, int rid0, int Ch, int only_undecided,
                  const WalkParams& kRay64, const SlotList& left = SlotList{}) {
  const double* P = c->tplanes.p;
  const double* E = c->teq.p;
  switch (c->ray_waves) {
    case 2:
      hipLaunchKernelGGL(k_walk_ray64<2>, dim3(n), dim3(128), lds, c->stream, c->net, io, rid0, kRay64, P, E, Ch,
                         only_undecided, left);
      break;
    case 8:
      hipLaunchKernelGGL(k_walk_ray64<8>, dim3(n), dim3(512), lds, c->stream, c->net, io, rid0, kRay64, P, E, Ch,
                         only_undecided, left);
      break;
    default:
      hipLaunchKernelGGL(k_walk_ray64<4>, dim3(n), dim3(256), lds, c->stream, c->net, io, rid0, kRay64, P, E, Ch,
                         only_undecided, left);
  }
}

void set_ray64_lds(size_t lds) {
  HIPCHK(hipFuncSetAttribute((const void*)k_walk_ray64<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  HIPCHK(hipFuncSetAttribute((const void*)k_walk_ray64<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  HIPCHK(hipFuncSetAttribute((const void*)k_walk_ray64<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

// Buffers of run_fp64_regions for a level of `count` regions (planes of one
// chunk, the undecided-slot list); returns the chunk size.  Idempotent, so it
// can run before a capture and again, as a no-op, inside it.
size_t fp64_prepare(am_ctx* c, int count) {
  const int Ch = (c->C + 63) & ~63;
  size_t chunk = ((size_t)1 << 30) / ((size_t)32 * Ch);
  chunk = std::max<size_t>(1024, std::min<size_t>(chunk, 65536)) & ~(size_t)(AM_CT - 1);
  const size_t cap = std::min<size_t>(chunk, (size_t)count);
  ensure(c->tplanes, cap * 4 * Ch, false, c->stream);
  ensure(c->teq, cap * 4, false, c->stream);
  ensure(c->tslots, cap, false, c->stream);
  ensure(c->tcnt, 1, false, c->stream);
  return chunk;
}

void run_fp64_regions(am_ctx* c, int lvl_first, int first, int count, const int* ids, int mode, Fp64Counts& t,
                      long long undecided = -1, const int* dcount = nullptr) {
  if (!c->tile_ok) {
    // one region per workgroup: the chain and the candidate scan in one kernel
    if (ids) {
      launch_region_fp64(c, lvl_first, count, ids);
    } else {
      launch_region_fp64(c, first, count);
    }
    t.scanned_host += count;
    return;
  }
  const bool after_walk = mode == 2;
  if (after_walk) mode = 1;
  if (mode == 1 && ray64_lds_bytes(c) > c->max_lds) mode = 0;
  int Cs, Csort;
  const size_t wlds = region_lds_bytes(c, Cs, Csort);
  const size_t rlds = ray64_lds_bytes(c);
  const int Ch = (c->C + 63) & ~63;
  const size_t chunk = fp64_prepare(c, count);
  const size_t clds = chain_lds_bytes(c), hlds = chain_half_lds_bytes(c);
  StageTimer& tm = c->timer;
  // regions the reduced walk may have decided with a vertex left to FP64 (the
  // FP32 form's Newton modes leave none)
  const WalkParams& rw = c->fast.wp;
  const bool vert64 = rw.lam_v < 0.f || (rw.lam_v > 0.f && !rw.vsel);
  // the FP64 ray walk lists what it leaves undecided for the candidate scan
  // itself (the chain zeroes the list's counter)
  const bool fused = mode == 1;
  for (int off = 0; off < count; off += (int)chunk) {
    const int n = std::min<int>((int)chunk, count - off);
    ChainIO cio = chain_io(c, ids ? ids + off : nullptr, first + off, n, Ch);
    cio.dcount = dcount;
    cio.dbase = off;
    if (fused) cio.zero = c->tcnt.p;
    RegionIO io;
    io.reg_mask = c->reg_mask.p;
    io.reg_start = c->reg_start.p;
    io.reg_point = c->reg_point.p;
    io.reg_info = c->reg_info.p;
    io.reg_tab = c->reg_tab.p;
    io.reg_verts = c->reg_verts.p;
    io.reg_vpair = c->reg_vpair.p;
    io.cand_mask = c->cand_mask.p;
    io.cand_start = c->cand_start.p;
    io.cand_point = c->cand_point.p;
    io.cand_hash = c->cand_hash.p;
    io.first = lvl_first;
    io.count = n;
    io.Cs = Cs;
    io.Csort = Csort;
    io.flags = (mode == 1 || after_walk) ? c->rflags.p : nullptr;
    io.ids = ids ? ids + off : nullptr;
    io.dcount = dcount;
    io.dbase = off;
    const size_t m0 = tm.mark(c->stream);
    // the records of the candidate scan on every region (mode 0) are the
    // oracle's bit for bit: its planes come from the sequential-fma chain
    const bool mf = c->chain_mfma && (mode != 0 || after_walk);
    if (mf)
      hipLaunchKernelGGL(k_chain_tile_mfma, dim3((n + AM_HT - 1) / AM_HT), dim3(AM_HALF_THREADS), hlds, c->stream,
                         c->net, cio);
    else
      hipLaunchKernelGGL(k_chain_tile_fp64, dim3((n + AM_CT - 1) / AM_CT), dim3(AM_CHAIN_THREADS), clds, c->stream,
                         c->net, cio);
    HIPCHK(hipGetLastError());
    t.chain_launches += 1;
    const size_t m1 = tm.mark(c->stream);
    tm.span(m0, m1, T_CHAIN64);
    size_t m1v = m1;
    if (after_walk && undecided != count && vert64) {
      hipLaunchKernelGGL(k_vertex_fp64, dim3(n), dim3(64), 0, c->stream, c->net, io, first + off,
                         (const double*)c->tplanes.p, (const double*)c->teq.p, Ch);
      HIPCHK(hipGetLastError());
      m1v = tm.mark(c->stream);
      tm.span(m1, m1v, T_VERT64);
    }
    if (after_walk && undecided == 0) continue;
    if (mode == 0 && !after_walk) {
      hipLaunchKernelGGL(k_walk_fp64, dim3(n), dim3(64), wlds, c->stream, c->net, io, first + off, (const int*)nullptr,
                         (const int*)nullptr, (const double*)c->tplanes.p, (const double*)c->teq.p, Ch);
      HIPCHK(hipGetLastError());
      tm.span(m1, tm.mark(c->stream), T_SCAN64);
      t.scanned_host += n;
      continue;
    }
    if (mode == 1) {
      launch_ray64(c, n, rlds, io, first + off, Ch, after_walk ? 1 : 0, mf ? kRay64Mf : kRay64,
                   SlotList{c->tcnt.p, c->tslots.p, c->runacc.p});
      HIPCHK(hipGetLastError());
      const size_t m2 = tm.mark(c->stream);
      tm.span(m1v, m2, T_RAY64);
    } else {
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->tcnt.p, (size_t)1, 0);
      hipLaunchKernelGGL(k_collect_slots, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                         (const unsigned*)c->rflags.p, n, dcount, off, (const int*)(ids ? ids + off : nullptr),
                         first + off, lvl_first, c->tcnt.p, c->tslots.p, c->runacc.p);
      HIPCHK(hipGetLastError());
    }
    io.flags = nullptr;
    const size_t m3 = tm.mark(c->stream);
    hipLaunchKernelGGL(k_walk_fp64, dim3(std::min(n, c->scan_blocks)), dim3(64), wlds, c->stream, c->net, io, first + off,
                       (const int*)c->tcnt.p, (const int*)c->tslots.p, (const double*)c->tplanes.p,
                       (const double*)c->teq.p, Ch);
    HIPCHK(hipGetLastError());
    tm.span(m3, tm.mark(c->stream), T_SCAN64);
  }
}

// A device-readable view of a caller's array: the pointer itself when it is
// device memory, else a copy in one of the context's staging buffers (the
// copy is ordered on st before the kernels that read it; the buffers are
// reused from run to run, c->stage_used counts this run's).
const void* dev_view(am_ctx* c, const void* src, size_t bytes, hipStream_t st) {
  if (!src || bytes == 0) return src;
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, src) == hipSuccess &&
      (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged))
    return src;
  (void)hipGetLastError();   // plain host memory: the query's error is not a failure
  if (c->stage_used == c->stage.size()) c->stage.emplace_back();
  DevBuf<char>& b = c->stage[c->stage_used++];
  ensure(b, bytes, false, st);
  HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
  return b.p;
}

// n elements (float or double) -> FP64 copy
void upload(am_ctx* c, DevBuf<double>& dst, const void* src, size_t n, int dtype, hipStream_t st) {
  ensure(dst, std::max<size_t>(n, 1), false, st);
  if (n == 0) return;
  const int f64 = dtype == AM_DTYPE_F64;
  const void* v = dev_view(c, src, n * (f64 ? 8 : 4), st);
  launch_transpose_pad(v, f64, (int)n, 1, dst.p, (int)n, 1, st);
  HIPCHK(hipGetLastError());
}

// (rows, cols) row-major -> FP64 (cols, rows) row-major
void upload_T(am_ctx* c, DevBuf<double>& dst, const void* src, int rows, int cols, int dtype, hipStream_t st) {
  const size_t n = (size_t)rows * cols;
  ensure(dst, std::max<size_t>(n, 1), false, st);
  if (n == 0) return;
  const int f64 = dtype == AM_DTYPE_F64;
  const void* v = dev_view(c, src, n * (f64 ? 8 : 4), st);
  launch_transpose_pad(v, f64, rows, cols, dst.p, rows, cols, st);
  HIPCHK(hipGetLastError());
}

// W (rows x cols, row-major) -> W^T padded: [max(cols, AM_CHAIN_MAXW) + 1][round_up(rows, AM_CR)],
// zeros past `rows` and `cols`; b -> b padded to round_up(rows, AM_CR).  The tiled
// chain reads whole AM_CR-row blocks, and row AM_CHAIN_MAXW (+0) for its padding.
void upload_pad(am_ctx* c, DevBuf<double>& wdst, DevBuf<double>& bdst, const void* w, const void* b, int rows,
                int cols, int dtype, hipStream_t st) {
  const int rp = (rows + AM_CR - 1) / AM_CR * AM_CR;
  const int cp = std::max(cols, AM_CHAIN_MAXW) + 1;
  const int f64 = dtype == AM_DTYPE_F64;
  const size_t esz = f64 ? 8 : 4;
  ensure(wdst, (size_t)cp * rp, false, st);
  ensure(bdst, (size_t)rp, false, st);
  launch_transpose_pad(dev_view(c, w, (size_t)rows * cols * esz, st), f64, rows, cols, wdst.p, rp, cp, st);
  launch_transpose_pad(dev_view(c, b, (size_t)rows * esz, st), f64, rows, 1, bdst.p, rp, 1, st);
  HIPCHK(hipGetLastError());
}

am_status fail(am_ctx* c, am_status s, const std::string& m) {
  if (c) c->err = m;
  return s;
}

template <typename F>
am_status guarded(am_ctx* c, F&& f) {
  try {
    return f();
  } catch (const HipErr& e) {
    return fail(c, AM_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(e.e) + " in " + e.what);
  } catch (const ArgErr& e) {
    return fail(c, AM_ERR_ARG, e.msg);
  } catch (const CapErr& e) {
    return fail(c, AM_ERR_CAPACITY, e.msg);
  } catch (const std::exception& e) {
    return fail(c, AM_ERR_HIP, std::string("error: ") + e.what());
  }
}

// ---------------------------------------------------------------------------
// A batch of networks marched in one job.  Every network's regions share one
// region store, visited set and candidate buffer: masks carry one extra word
// holding the network id (so equal masks of different networks stay distinct
// keys, and the visited-set and mesh kernels need no change), and each level
// runs the tiled FP64 chain and the walks for all networks together -- a chain
// tile holds regions of one network, every other kernel finds a region's
// network in its mask's last word.  The breadth-first level of each network is
// the one it has alone, so per network the records are those of its own march.
// ---------------------------------------------------------------------------
struct BatchNet {
  std::vector<int> nodes, off;
  int D = 0, L = 0, Wn = 0;
  DevBuf<double> W0;
  std::vector<DevBuf<double>> WT, Wrow, b, WTp, bpad;
};

}  // namespace

struct am_batch {
  am_ctx* c = nullptr;
  int precision = AM_PREC_FP64_RAY;
  std::vector<BatchNet> nets;
  int Wmax = 0, Dmax = 0, Cmax = 0, Lmax = 0;
  std::vector<NetDev> nets_h;
  std::vector<ChainIO> cios_h;
  DevBuf<NetDev> nets_d;
  DevBuf<ChainIO> cios_d;
  DevBuf<int4> tiles_d;
  int4* tiles_h = nullptr;       // pinned staging of a level's tile list
  size_t tiles_cap = 0;
  DevBuf<int> hist_d, sel_flags, sel_rank, sel_ids;
  int* hist_h = nullptr;         // pinned: regions per network of the level
};

namespace {

// LDS of the candidate scan / the FP64 ray walk / the chain for the widest network of the batch.
size_t batch_scan_lds(const am_batch* b) {
  size_t m = 0;
  for (const BatchNet& n : b->nets) {
    int Csort = AM_TOPK;
    while (Csort < n.L) Csort <<= 1;
    m = std::max(m, (size_t)4 * n.L * 8 + (size_t)Csort * 12 + (size_t)((n.L + 31) / 32) * 4 + (size_t)b->c->Wd * 4 + 64);
  }
  return m;
}
size_t batch_ray_lds(const am_batch* b) {
  size_t m = 0;
  for (const BatchNet& n : b->nets)
    m = std::max(m, (size_t)4 * (n.L - n.nodes[1]) * 8 + (((size_t)b->c->Wd * 4 + 15) & ~(size_t)15) +
                        walk_rec_bytes(b->c->vm, false));
  return m;
}
size_t batch_chain_lds(const am_batch* b) {
  return ((size_t)AM_CXR * AM_CXS + 1) * sizeof(double) + (size_t)b->Dmax * AM_KLS * sizeof(uint16_t) +
         (size_t)AM_CT * b->c->Wd * 4 + 16;
}

// Planes and walks of the level's regions (first .. first + count - 1, grouped
// by network in ascending order; hist_h has the per-network counts).
struct BatchLevelPlan {
  long long first = 0, count = 0;
  size_t chunk = 0, nt = 0;
  std::vector<size_t> chunk_tile0;
};

void batch_level_launch(am_batch* b, const BatchLevelPlan& pl, Fp64Counts& t);

// Host part of a level: buffers and the tile list (in pinned memory).
BatchLevelPlan batch_level_prepare(am_batch* b, long long first, long long count) {
  am_ctx* c = b->c;
  const hipStream_t st = c->stream;
  BatchLevelPlan pl;
  pl.first = first;
  pl.count = count;
  const int Ch = (b->Cmax + 63) & ~63;
  size_t chunk = ((size_t)1 << 30) / ((size_t)32 * Ch);
  chunk = std::max<size_t>(1024, std::min<size_t>(chunk, 65536)) & ~(size_t)(AM_CT - 1);
  const size_t cap = std::min<size_t>(chunk, (size_t)count);
  ensure(c->tplanes, cap * 4 * Ch, false, st);
  ensure(c->teq, cap * 4, false, st);
  ensure(c->tslots, cap, false, st);
  ensure(c->tcnt, 1, false, st);
  // tiles: per chunk, per network segment, AM_CT regions (chunk-relative r0)
  const int nn = (int)b->nets.size();
  std::vector<long long> seg(nn + 1, 0);
  for (int i = 0; i < nn; ++i) seg[i + 1] = seg[i] + b->hist_h[i];
  if (seg[nn] != count) throw std::runtime_error("batch level: network counts do not add up");
  const size_t max_tiles = (size_t)(count + AM_CT - 1) / AM_CT + (size_t)nn * ((count + chunk - 1) / chunk + 1);
  if (max_tiles > b->tiles_cap) {
    if (b->tiles_h) HIPCHK(hipHostFree(b->tiles_h));
    b->tiles_h = nullptr;
    b->tiles_cap = std::max(max_tiles, b->tiles_cap * 2);
    HIPCHK(hipHostMalloc((void**)&b->tiles_h, b->tiles_cap * sizeof(int4), hipHostMallocDefault));
    ensure(b->tiles_d, b->tiles_cap, false, st);
  }
  std::vector<size_t> chunk_tile0;
  size_t nt = 0;
  for (long long cs = 0; cs < count; cs += (long long)chunk) {
    const long long ce = std::min<long long>(cs + (long long)chunk, count);
    chunk_tile0.push_back(nt);
    for (int i = 0; i < nn; ++i) {
      const long long a = std::max(seg[i], cs), e = std::min(seg[i + 1], ce);
      for (long long r = a; r < e; r += AM_CT) b->tiles_h[nt++] = make_int4(i, (int)(r - cs), (int)std::min<long long>(AM_CT, e - r), 0);
    }
  }
  chunk_tile0.push_back(nt);
  pl.chunk = chunk;
  pl.nt = nt;
  pl.chunk_tile0 = std::move(chunk_tile0);
  // stream-ordered after the previous level's kernels (which read the old
  // list) and kept out of graph captures (no memcpy nodes)
  HIPCHK(hipMemcpyAsync(b->tiles_d.p, b->tiles_h, nt * sizeof(int4), hipMemcpyHostToDevice, st));
  return pl;
}

// The level's launches (kernels only: capturable).
void batch_level_launch(am_batch* b, const BatchLevelPlan& pl, Fp64Counts& t) {
  am_ctx* c = b->c;
  const hipStream_t st = c->stream;
  const long long first = pl.first, count = pl.count;
  const size_t chunk = pl.chunk;
  const std::vector<size_t>& chunk_tile0 = pl.chunk_tile0;
  const int Ch = (b->Cmax + 63) & ~63;
  const size_t clds = batch_chain_lds(b), slds = batch_scan_lds(b), rlds = batch_ray_lds(b);
  const bool ray = b->precision == AM_PREC_FP64_RAY && rlds <= c->max_lds;
  // the ray walk certifies its decisions: its planes may come from the matrix cores
  const bool mf = ray && c->chain_mfma;
  const WalkParams& kr = mf ? kRay64Mf : kRay64;
  StageTimer& tm = c->timer;
  size_t k = 0;
  for (long long cs = 0; cs < count; cs += (long long)chunk, ++k) {
    const int n = (int)std::min<long long>((long long)chunk, count - cs);
    const int rid0 = (int)(first + cs);
    const ChainLevel lv{rid0, n, nullptr, c->tplanes.p, c->teq.p};
    const int ntile = (int)(chunk_tile0[k + 1] - chunk_tile0[k]);
    const size_t m0 = tm.mark(st);
    hipLaunchKernelGGL((mf ? k_chain_tile_fp64_multi<true> : k_chain_tile_fp64_multi<false>), dim3(ntile),
                       dim3(AM_CHAIN_THREADS), clds, st, (const NetDev*)b->nets_d.p, (const ChainIO*)b->cios_d.p, lv,
                       (const int4*)(b->tiles_d.p + chunk_tile0[k]));
    HIPCHK(hipGetLastError());
    t.chain_launches += 1;
    const size_t m1 = tm.mark(st);
    tm.span(m0, m1, T_CHAIN64);
    RegionIO io;
    io.reg_mask = c->reg_mask.p;
    io.reg_start = c->reg_start.p;
    io.reg_point = c->reg_point.p;
    io.reg_info = c->reg_info.p;
    io.reg_tab = c->reg_tab.p;
    io.reg_verts = c->reg_verts.p;
    io.reg_vpair = c->reg_vpair.p;
    io.cand_mask = c->cand_mask.p;
    io.cand_start = c->cand_start.p;
    io.cand_point = c->cand_point.p;
    io.cand_hash = c->cand_hash.p;
    io.first = (int)first;
    io.count = n;
    io.Cs = 0;
    io.Csort = 0;
    io.flags = ray ? c->rflags.p : nullptr;
    io.ids = nullptr;
    if (!ray) {
      hipLaunchKernelGGL(k_walk_fp64_multi, dim3(n), dim3(64), slds, st, (const NetDev*)b->nets_d.p, io, c->Wd,
                         rid0, (const int*)nullptr, (const int*)nullptr, (const double*)c->tplanes.p,
                         (const double*)c->teq.p, Ch);
      HIPCHK(hipGetLastError());
      tm.span(m1, tm.mark(st), T_SCAN64);
      t.scanned_host += n;
      continue;
    }
    {
      const NetDev* nd = (const NetDev*)b->nets_d.p;
      const double* P = c->tplanes.p;
      const double* E = c->teq.p;
      switch (c->ray_waves) {
        case 2:
          hipLaunchKernelGGL(k_walk_ray64_multi<2>, dim3(n), dim3(128), rlds, st, nd, io, c->Wd, rid0, kr, P, E, Ch);
          break;
        case 8:
          hipLaunchKernelGGL(k_walk_ray64_multi<8>, dim3(n), dim3(512), rlds, st, nd, io, c->Wd, rid0, kr, P, E, Ch);
          break;
        default:
          hipLaunchKernelGGL(k_walk_ray64_multi<4>, dim3(n), dim3(256), rlds, st, nd, io, c->Wd, rid0, kr, P, E, Ch);
      }
    }
    HIPCHK(hipGetLastError());
    const size_t m2 = tm.mark(st);
    tm.span(m1, m2, T_RAY64);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, c->tcnt.p, (size_t)1, 0);
    hipLaunchKernelGGL(k_collect_slots, dim3((n + 255) / 256), dim3(256), 0, st, (const unsigned*)c->rflags.p, n,
                       (const int*)nullptr, 0, (const int*)nullptr, rid0, (int)first, c->tcnt.p, c->tslots.p,
                       c->runacc.p);
    io.flags = nullptr;
    hipLaunchKernelGGL(k_walk_fp64_multi, dim3(std::min(n, c->scan_blocks)), dim3(64), slds, st, (const NetDev*)b->nets_d.p,
                       io, c->Wd, rid0, (const int*)c->tcnt.p, (const int*)c->tslots.p,
                       (const double*)c->tplanes.p, (const double*)c->teq.p, Ch);
    HIPCHK(hipGetLastError());
    tm.span(m2, tm.mark(st), T_SCAN64);
  }
}

// Regions per network of level first .. first + count - 1 into b->hist_h (waits for the stream).
void batch_hist(am_batch* b, long long first, long long count) {
  am_ctx* c = b->c;
  const hipStream_t st = c->stream;
  const int nn = (int)b->nets.size();
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks(nn, 256)), dim3(256), 0, st, b->hist_d.p, (size_t)nn, 0);
  if (count > 0)
    hipLaunchKernelGGL(k_net_hist, dim3(blocks((size_t)count, 256)), dim3(256), 0, st, c->reg_mask.p, c->Wd,
                       (int)first, (int)count, b->hist_d.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->hist_h, b->hist_d.p, nn * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
namespace {

// Whether p is ordinary (pageable) host memory: not device, pinned or managed.
bool pageable_host(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return true;
  }
  return a.type == hipMemoryTypeUnregistered;
}

}  // namespace

extern "C" {

am_status am_create(am_ctx** out, int n_nodes, const int* nodes, const int* arc_table, int arc_rows,
                    int arc_cols, int num_extra_constraints, int precision) {
  if (!out) return AM_ERR_ARG;
  *out = nullptr;
  if (n_nodes < 3 || !nodes || nodes[0] != 3 || nodes[n_nodes - 1] != 1) {
    g_create_error = "nodes must start with 3, end with 1 and have at least 3 entries";
    return AM_ERR_ARG;
  }
  if (n_nodes - 1 > AM_MAX_LINEAR) {
    g_create_error = "too many layers";
    return AM_ERR_ARG;
  }
  if (arc_rows != n_nodes - 2 || arc_cols < 1 || (arc_cols + 1) % 2 != 0 || !arc_table) {
    g_create_error = "arc_table must be int32 [len(nodes)-2, 2k+1]";
    return AM_ERR_ARG;
  }
  if (num_extra_constraints < 0) {
    g_create_error = "num_extra_constraints must be >= 0";
    return AM_ERR_ARG;
  }
  if (precision != AM_PREC_FP32 && precision != AM_PREC_SPLIT_BF16X3 && precision != AM_PREC_FP64 &&
      precision != AM_PREC_FP64_RAY) {
    g_create_error = "unknown precision";
    return AM_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    g_create_error = "no HIP device available";
    return AM_ERR_NOGPU;
  }
  am_ctx* c = new am_ctx();
  c->nodes.assign(nodes, nodes + n_nodes);
  c->D = n_nodes - 2;
  c->off.assign(c->D + 1, 0);
  for (int k = 0; k < c->D; ++k) c->off[k + 1] = c->off[k] + nodes[k + 1];
  c->L = c->off[c->D];
  c->E = num_extra_constraints;
  c->C = c->L + c->E;
  c->Wd = (c->L + 31) / 32;
  c->precision = precision;
  if (c->Wd > AM_MAX_WORDS) {
    g_create_error = "more than 2048 hidden neurons";
    delete c;
    return AM_ERR_ARG;
  }
  c->skips.assign(c->D + 1, {});
  for (int row = 0; row < arc_rows; ++row) {
    const int cnt = arc_table[row * arc_cols];
    if (cnt < 0 || 1 + 2 * cnt > arc_cols) {
      g_create_error = "arc_table row declares more skips than it holds";
      delete c;
      return AM_ERR_ARG;
    }
    if (row + 1 >= c->D) continue;  // output-layer skips: ignored, as the reference does
    if (cnt > AM_MAX_SKIPS) {
      g_create_error = "too many skips into one layer";
      delete c;
      return AM_ERR_ARG;
    }
    for (int k = 0; k < cnt; ++k)
      c->skips[row + 1].push_back({arc_table[row * arc_cols + 1 + 2 * k], arc_table[row * arc_cols + 2 + 2 * k]});
  }
  am_status s = guarded(c, [&]() -> am_status {
    HIPCHK(hipGetDevice(&c->dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->dev));
    c->max_lds = prop.sharedMemPerBlock;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, c->dev) == hipSuccess && optin > 0)
      c->max_lds = std::max<size_t>(c->max_lds, (size_t)optin);
    HIPCHK(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
    c->stream = c->own_stream;
    return AM_OK;
  });
  if (s != AM_OK) {
    g_create_error = c->err;
    delete c;
    return s;
  }
  *out = c;
  return AM_OK;
}

am_status am_set_option(am_ctx* c, const char* key, long long value) {
  if (!c || !key) return AM_ERR_ARG;
  const std::string k(key);
  if (k == "vm") {
    if (value < 3 || value > AM_VM_MAX) return fail(c, AM_ERR_ARG, "vm must be in [3, 64]");
    if (c->has_run) return fail(c, AM_ERR_STATE, "vm must be set before the first am_run");
    if (value != c->vm) {
      // the region store's tab, vertex and pair arrays are strided by vm: drop
      // them (a failed am_run may have sized them already)
      c->reg_mask.release(); c->reg_start.release(); c->reg_point.release(); c->reg_info.release();
      c->reg_tab.release(); c->reg_verts.release(); c->reg_vpair.release(); c->reg_jinv.release();
      c->reg_cap = 0;
      c->R = 0;
      c->dedup_ready = false;
    }
    c->vm = (int)value;
  } else if (k == "table_log2") {
    if (value < 10 || value > 31) return fail(c, AM_ERR_ARG, "table_log2 must be in [10, 31]");
    c->table_log2 = (int)value;
  } else if (k == "max_regions") {
    c->max_regions = value;
  } else if (k == "region_limit") {
    c->region_limit = value;
  } else if (k == "graph_levels") {
    if (value != 0 && value != 1) return fail(c, AM_ERR_ARG, "graph_levels must be 0 or 1");
    c->graph_levels = value != 0;
  } else if (k == "ray_waves") {
    if (value != 2 && value != 4 && value != 8) return fail(c, AM_ERR_ARG, "ray_waves must be 2, 4 or 8");
    c->ray_waves = (int)value;
  } else if (k == "scan_blocks") {
    if (value < 64 || value > (1 << 20)) return fail(c, AM_ERR_ARG, "scan_blocks must be in [64, 2^20]");
    c->scan_blocks = (int)value;
  } else if (k == "walk") {
    if (value != 0 && value != 1) return fail(c, AM_ERR_ARG, "walk must be 0 (ray) or 1 (scan)");
    c->fast.walk_mode = (int)value;
  } else if (k == "ray_waves_reduced") {
    if (value != 2 && value != 4 && value != 8) return fail(c, AM_ERR_ARG, "ray_waves_reduced must be 2, 4 or 8");
    c->fast.ray_waves = (int)value;
  } else if (k == "chain_mfma") {
    if (value != 0 && value != 1) return fail(c, AM_ERR_ARG, "chain_mfma must be 0 or 1");
    c->chain_mfma = value != 0;
  } else if (k == "reduced_vertices") {
    if (value < 0 || value > 3)
      return fail(c, AM_ERR_ARG,
                  "reduced_vertices must be 0 (FP64 re-solve), 1 (estimate), 2 (Newton) or 3 (Newton where the "
                  "estimate is not within 1e-5)");
    c->fast.vertex_mode = (int)value;
  } else if (k == "adjoint_mfma") {
    if (value != 0 && value != 1) return fail(c, AM_ERR_ARG, "adjoint_mfma must be 0 or 1");
    c->adjoint_mfma = value != 0;
  } else if (k == "adjoint_tiled") {
    if (value != 0 && value != 1) return fail(c, AM_ERR_ARG, "adjoint_tiled must be 0 or 1");
    c->adjoint_tiled = value != 0;
  } else {
    return fail(c, AM_ERR_ARG, "unknown option " + k);
  }
  return AM_OK;
}

// FP64 Newton step for the vertices the reduced walk solved from its planes
// (defined with the mesh assembly, whose vertex dedup it uses); returns the
// number of vertices refined.
long long newton_refine(am_ctx* c);

am_status am_run(am_ctx* c, const void* const* weights, const void* const* biases, int n_tm,
                 const int* tm_shapes, const void* const* tms, int dtype, const void* points,
                 const uint8_t* states, int n_seeds, const void* extra_w, const void* extra_b, double iso,
                 int flip_insideout, void* stream) {
  if (!c) return AM_ERR_ARG;
  return guarded(c, [&]() -> am_status {
    if (!weights || !biases) throw ArgErr{"weights and biases are required"};
    if (dtype != AM_DTYPE_F32 && dtype != AM_DTYPE_F64) throw ArgErr{"dtype must be AM_DTYPE_F32 or AM_DTYPE_F64"};
    if (n_seeds < 1 || !points || !states) throw ArgErr{"at least one seed (point + state) is required"};
    if (c->E > 0 && (!extra_w || !extra_b)) throw ArgErr{"extra constraints declared but not given"};
    if (n_tm < 0 || n_tm > AM_MAX_TMS || (n_tm > 0 && (!tm_shapes || !tms))) throw ArgErr{"bad transform matrices"};
    for (auto& row : c->skips)
      for (auto& s : row) {
        if (s.second < 0 || s.second >= n_tm) throw ArgErr{"arc_table names a missing transform matrix"};
        if (s.first < 0 || s.first > c->D) throw ArgErr{"arc_table names a missing source layer"};
      }
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    const hipStream_t st = c->stream;
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < AM_STAT_COUNT; ++i) c->stats[i] = 0;
    c->has_run = false;
    c->has_mesh = false;
    c->dedup_ready = false;

    // ---- network upload (fp64, weights transposed for the chain)
    const int nl = c->D + 1;
    c->WT.resize(nl);
    c->Wrow.resize(nl);
    c->b.resize(nl);
    // device-readable views of the weights (host arrays are staged once)
    c->stage_used = 0;
    const size_t esz = dtype == AM_DTYPE_F64 ? 8 : 4;
    std::vector<const void*> wv(nl), bv(nl);
    wv[0] = dev_view(c, weights[0], (size_t)c->nodes[1] * 3 * esz, st);
    for (int li = 1; li < nl; ++li)
      wv[li] = dev_view(c, weights[li], (size_t)c->nodes[li + 1] * c->nodes[li] * esz, st);
    for (int li = 0; li < nl; ++li) bv[li] = dev_view(c, biases[li], (size_t)c->nodes[li + 1] * esz, st);
    upload(c, c->W0, wv[0], (size_t)c->nodes[1] * 3, dtype, st);
    for (int li = 0; li < nl; ++li) {
      upload(c, c->b[li], bv[li], (size_t)c->nodes[li + 1], dtype, st);
      if (li >= 1) upload_T(c, c->WT[li], wv[li], c->nodes[li + 1], c->nodes[li], dtype, st);
      if (li >= 1) upload(c, c->Wrow[li], wv[li], (size_t)c->nodes[li + 1] * c->nodes[li], dtype, st);
    }
    c->WTp.resize(nl);
    c->bpad.resize(nl);
    for (int li = 1; li < nl; ++li)
      upload_pad(c, c->WTp[li], c->bpad[li], wv[li], bv[li], c->nodes[li + 1], c->nodes[li], dtype, st);
    c->tm.resize(n_tm);
    c->tmT.resize(n_tm);
    c->tm_h.assign(n_tm, 0);
    c->tm_w.assign(n_tm, 0);
    for (int t = 0; t < n_tm; ++t) {
      const int h = tm_shapes[2 * t], w = tm_shapes[2 * t + 1];
      c->tm_h[t] = h;
      c->tm_w[t] = w;
      if (h > 0 && w > 0) {
        upload(c, c->tm[t], tms[t], (size_t)h * w, dtype, st);
        upload_T(c, c->tmT[t], tms[t], h, w, dtype, st);
      }
    }
    // skip shape checks (backend/src/cuam.cpp:160-172)
    for (int li = 1; li < c->D; ++li)
      for (auto& s : c->skips[li]) {
        const int h = c->tm_h[s.second], w = c->tm_w[s.second];
        if (h == 0 && w == 0) {
          if (s.first > 0 && c->nodes[s.first] != c->nodes[li + 1])
            throw ArgErr{"identity skip between layers of different widths"};
        } else {
          if (h != c->nodes[li + 1]) throw ArgErr{"transform matrix rows must match the destination width"};
          if (w != c->nodes[s.first]) throw ArgErr{"transform matrix columns must match the source width"};
        }
      }
    upload(c, c->ew, extra_w, (size_t)c->E * 3, dtype, st);
    upload(c, c->eb, extra_b, (size_t)c->E, dtype, st);

    NetDev& n = c->net;
    std::memset(&n, 0, sizeof(n));
    n.D = c->D; n.L = c->L; n.E = c->E; n.C = c->C; n.Wd = c->Wd; n.vm = c->vm;
    for (int i = 0; i < (int)c->nodes.size(); ++i) n.nodes[i] = c->nodes[i];
    for (int i = 0; i <= c->D; ++i) n.off[i] = c->off[i];
    n.W0 = c->W0.p;
    for (int li = 0; li < nl; ++li) {
      n.WT[li] = li >= 1 ? c->WT[li].p : nullptr;
      n.Wrow[li] = li >= 1 ? c->Wrow[li].p : nullptr;
      n.b[li] = c->b[li].p;
    }
    for (int li = 0; li < nl; ++li) {
      n.nskip[li] = (int)c->skips[li].size();
      for (size_t s = 0; s < c->skips[li].size(); ++s) {
        n.skip_from[li][s] = c->skips[li][s].first;
        n.skip_tm[li][s] = c->skips[li][s].second;
      }
    }
    for (int t = 0; t < n_tm; ++t) {
      n.tm[t] = c->tm[t].p;
      n.tmT[t] = c->tmT[t].p;
      n.tm_h[t] = c->tm_h[t];
      n.tm_w[t] = c->tm_w[t];
    }
    n.ew = c->ew.p;
    n.eb = c->eb.p;
    n.iso = iso;
    n.flip = flip_insideout ? 1 : 0;

    int Cs, Csort;
    const size_t lds = region_lds_bytes(c, Cs, Csort);
    if (lds > c->max_lds)
      throw CapErr{"network too wide for the FP64 region kernel's LDS (" + std::to_string(lds) + " bytes)"};
    if (lds > 65536) {
      HIPCHK(hipFuncSetAttribute((const void*)k_region_fp64, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      HIPCHK(hipFuncSetAttribute((const void*)k_walk_fp64, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    c->tile_ok = true;
    for (int li = 0; li <= c->D; ++li)
      if (!c->skips[li].empty()) c->tile_ok = false;
    for (int i = 1; i <= c->D + 1; ++i)
      if (c->nodes[i] > AM_CHAIN_MAXW) c->tile_ok = false;
    if (c->tile_ok) {
      const size_t clds = chain_lds_bytes(c);
      if (clds > c->max_lds) c->tile_ok = false;
      else if (clds > 65536) {
        HIPCHK(hipFuncSetAttribute((const void*)k_chain_tile_fp64, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)clds));
        const size_t hlds = chain_half_lds_bytes(c);
        if (hlds > 65536)
          HIPCHK(hipFuncSetAttribute((const void*)k_chain_tile_mfma, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)hlds));
      }
    }
    // The reduced forms need the tiled FP64 chain for their vertices and
    // re-decisions; nets it does not take (skips, layers over 256 wide) run
    // those forms on the FP64 region kernel, with the same topology.
    const bool reduced_asked = c->precision == AM_PREC_FP32 || c->precision == AM_PREC_SPLIT_BF16X3;
    const bool fp64_form = !reduced_asked || !c->tile_ok;
    // FP32 form, Newton vertices: the walk keeps each vertex's inverse system
    // (the Newton kernel takes layers up to AM_CHAIN_MAXW wide, which the
    // reduced forms require anyway, and the ray walk)
    c->jinv_on = !fp64_form && c->precision == AM_PREC_FP32 && c->fast.vertex_mode >= 2 &&
                 c->fast.walk_mode == 0 && newton_lds_bytes(c->Wd) <= c->max_lds;
    const int vmode = c->fast.vertex_mode;
    if (!fp64_form && c->precision == AM_PREC_FP32 && vmode >= 2 && !c->jinv_on) c->fast.vertex_mode = 0;
    if (!fp64_form) fast_prepare(c->fast, c->fbuf, n, c->precision, wv.data(), bv.data(), dtype, st);
    c->fast.vertex_mode = vmode;
    if (c->jinv_on) {
      ensure(c->reg_jinv, c->reg_cap * c->vm * 9, true, st);
      HIPCHK(hipFuncSetAttribute((const void*)k_newton_fp64, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)newton_lds_bytes(c->Wd)));
    }
    if (c->tile_ok && ray64_lds_bytes(c) > 65536 && ray64_lds_bytes(c) <= c->max_lds)
      set_ray64_lds(ray64_lds_bytes(c));

    const auto t_setup = std::chrono::steady_clock::now();
    // ---- seeds -> first level (backend/inc/states.h:147-176)
    c->R = 0;
    c->table_cap = 0;
    c->table.release();
    grow_cands(c, (size_t)n_seeds);
    {
      std::vector<uint32_t> m((size_t)n_seeds * c->Wd, 0u);
      std::vector<uint8_t> hs((size_t)n_seeds * c->L);
      HIPCHK(hipMemcpy(hs.data(), states, hs.size(), hipMemcpyDefault));
      for (int s = 0; s < n_seeds; ++s)
        for (int g = 0; g < c->L; ++g)
          if (hs[(size_t)s * c->L + g]) m[(size_t)s * c->Wd + (g >> 5)] |= 1u << (g & 31);
      std::vector<double> pts((size_t)n_seeds * 3);
      if (dtype == AM_DTYPE_F64) {
        HIPCHK(hipMemcpy(pts.data(), points, pts.size() * sizeof(double), hipMemcpyDefault));
      } else {
        std::vector<float> f(pts.size());
        HIPCHK(hipMemcpy(f.data(), points, f.size() * sizeof(float), hipMemcpyDefault));
        for (size_t i = 0; i < f.size(); ++i) pts[i] = f[i];
      }
      std::vector<int> startv(n_seeds, -1);
      HIPCHK(hipMemcpy(c->cand_mask.p, m.data(), m.size() * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->cand_point.p, pts.data(), pts.size() * 8, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->cand_start.p, startv.data(), startv.size() * 4, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_hash_keys, dim3(blocks(n_seeds, 256)), dim3(256), 0, st, (const uint32_t*)c->cand_mask.p,
                       n_seeds, c->Wd, c->cand_hash.p);
    long long lvl_first = 0;
    long long lvl_count = insert_level(c, (size_t)n_seeds);
    long long levels = 0, cands = 0, launches = 0;
    long long why[AM_WHY_COUNT] = {0};
    Fp64Counts f64;
    c->timer.reset();
    // run counters on the device: [0, 1 + W) the FP64 ray walk's undecided
    // regions and reasons, [1 + W, 2 + 2W) the reduced-precision walk's, then
    // the regions it decided with a vertex left to FP64
    constexpr int NACC = 2 * (1 + AM_WHY_COUNT) + 1;
    ensure(c->runacc, NACC, false, st);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, c->runacc.p, (size_t)NACC, 0);
    while (lvl_count > 0) {
      if (c->max_regions >= 0 && lvl_first >= c->max_regions) break;
      const size_t ncand = (size_t)lvl_count * c->vm;
      grow_cands(c, ncand);
      ensure(c->rflags, (size_t)lvl_count, false, st);
      if (fp64_form && c->graph_levels) {
        // the level's launches (planes, walks, visited-set insert and commit)
        // as one graph; the host work (growth) happens before the capture
        if (c->tile_ok) fp64_prepare(c, (int)lvl_count);
        const InsertPlan plan = insert_prepare(c, ncand);
        launch_as_graph(c, (int)levels, [&]() {
          run_fp64_regions(c, (int)lvl_first, (int)lvl_first, (int)lvl_count, nullptr,
                           c->precision == AM_PREC_FP64 ? 0 : 1, f64);
          insert_launch(c, plan);
        });
        const long long before = c->R;
        const int added = insert_finish(c, plan);
        launches += 1;
        cands += (long long)ncand;
        levels += 1;
        lvl_first = before;
        lvl_count = added;
        continue;
      }
      // the reduced forms' level: face kernel and checked walk, then the FP64
      // redo of the regions the walk left to FP64
      auto fast_level_of = [&]() {
        FastLevel fl;
        fl.first = (int)lvl_first;
        fl.count = (int)lvl_count;
        fl.reg_mask = c->reg_mask.p; fl.reg_start = c->reg_start.p; fl.reg_point = c->reg_point.p;
        fl.reg_info = c->reg_info.p; fl.reg_tab = c->reg_tab.p; fl.reg_verts = c->reg_verts.p;
        fl.reg_vpair = c->reg_vpair.p; fl.cand_mask = c->cand_mask.p; fl.cand_start = c->cand_start.p;
        fl.cand_point = c->cand_point.p;
        fl.cand_hash = c->cand_hash.p;
        fl.reg_jinv = c->jinv_on ? c->reg_jinv.p : nullptr;
        return fl;
      };
      if (!fp64_form && c->graph_levels) {
        // as the FP64 forms' graph above: the host work first (buffers, the
        // region store's growth), then every launch of the level in one graph
        if (c->tile_ok) fp64_prepare(c, (int)lvl_count);
        fast_level_prepare(c->fast, c->fbuf, n, (int)lvl_count);
        ensure(c->redo_ids, (size_t)lvl_count, false, st);
        ensure(c->redo_cnt, 1, false, st);
        ensure_zeroed(c->redo_state, redo_state_words((int)lvl_count), st);
        const InsertPlan plan = insert_prepare(c, ncand);
        const FastLevel fl = fast_level_of();
        const RedoList redo{c->redo_ids.p, c->redo_cnt.p, c->redo_state.p};
        FastTimes ft{};
        launch_as_graph(c, (int)levels, [&]() {
          fast_level(c->fast, c->fbuf, n, fl, st, c->timer, ft, c->rflags.p, c->runacc.p + 1 + AM_WHY_COUNT, redo,
                     true);
          run_fp64_regions(c, (int)lvl_first, (int)lvl_first, (int)lvl_count, c->redo_ids.p, 2, f64, -1,
                           c->redo_cnt.p);
          insert_launch(c, plan);
        });
        const long long before = c->R;
        const int added = insert_finish(c, plan);
        launches += ft.launches;
        cands += (long long)ncand;
        levels += 1;
        lvl_first = before;
        lvl_count = added;
        continue;
      }
      if (fp64_form) {
        run_fp64_regions(c, (int)lvl_first, (int)lvl_first, (int)lvl_count, nullptr,
                         c->precision == AM_PREC_FP64 ? 0 : 1, f64);
        launches += 1;
      } else {
        const FastLevel fl = fast_level_of();
        ensure(c->redo_ids, (size_t)lvl_count, false, st);
        ensure(c->redo_cnt, 1, false, st);
        ensure_zeroed(c->redo_state, redo_state_words((int)lvl_count), st);
        const RedoList redo{c->redo_ids.p, c->redo_cnt.p, c->redo_state.p};
        FastTimes ft{};
        fast_level(c->fast, c->fbuf, n, fl, st, c->timer, ft, c->rflags.p, c->runacc.p + 1 + AM_WHY_COUNT, redo);
        // FP64 planes of the regions the walk left to FP64 only (their list,
        // made by fast_level's last launch, and count stay on the device;
        // the launches are sized for the whole
        // level and their blocks past the count exit): the ones it decided
        // with a vertex it could not vouch for get their vertices re-solved
        // from them; the undecided ones are re-decided in place (records and
        // candidate slots rewritten) by the checked FP64 ray walk, and by the
        // candidate scan where that is undecided too.
        run_fp64_regions(c, (int)lvl_first, (int)lvl_first, (int)lvl_count, c->redo_ids.p, 2, f64, -1,
                         c->redo_cnt.p);
        launches += ft.launches;
      }
      const size_t mi = c->timer.mark(st);
      const long long before = c->R;
      const int added = insert_level(c, ncand);   // waits once: the next level's size
      c->timer.span(mi, c->timer.mark(st), T_INSERT);
      cands += (long long)ncand;
      levels += 1;
      lvl_first = before;
      lvl_count = added;
    }
    if (lvl_count > 0 && c->max_regions >= 0) {
      // regions found but not processed: drop them from the store
      c->R = lvl_first;
    }
    long long newton_n = 0;
    if (c->jinv_on && c->R > 0) {
      const size_t mn = c->timer.mark(st);
      newton_n = newton_refine(c);
      c->timer.span(mn, c->timer.mark(st), T_NEWTON);
    }
    int acc[NACC] = {0};
    HIPCHK(hipMemcpyAsync(acc, c->runacc.p, sizeof(acc), hipMemcpyDeviceToHost, st));
    const long long redo = fp64_form ? 0 : acc[1 + AM_WHY_COUNT];
    const long long vert64 = fp64_form ? 0 : acc[2 + 2 * AM_WHY_COUNT];
    HIPCHK(hipStreamSynchronize(st));
    c->timer.resolve();
    const auto t1 = std::chrono::steady_clock::now();
    const double* T = c->timer.acc;
    const double fp64_ms = T[T_CHAIN64] + T[T_RAY64] + T[T_SCAN64] + T[T_REGION64] + T[T_VERT64];
    c->stats[AM_STAT_RUN_MS] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    c->stats[AM_STAT_INSERT_MS] = T[T_INSERT];
    if (fp64_form) {
      c->stats[AM_STAT_REGION_MS] = fp64_ms;
      c->stats[AM_STAT_GEMM_MS] = T[T_CHAIN64];
      c->stats[AM_STAT_WALK_MS] = T[T_RAY64] + T[T_SCAN64] + T[T_REGION64];
      c->stats[AM_STAT_REDO_MS] = 0;
      for (int w = 0; w < AM_WHY_COUNT; ++w) why[w] = acc[1 + w];
    } else {
      c->stats[AM_STAT_REGION_MS] = T[T_GEMM] + T[T_WALK];
      c->stats[AM_STAT_GEMM_MS] = T[T_GEMM];
      c->stats[AM_STAT_WALK_MS] = T[T_WALK];
      c->stats[AM_STAT_REDO_MS] = fp64_ms;
      for (int w = 0; w < AM_WHY_COUNT; ++w) why[w] = acc[2 + AM_WHY_COUNT + w];
    }
    c->stats[AM_STAT_LEVELS] = (double)levels;
    c->stats[AM_STAT_REGIONS] = (double)c->R;
    c->stats[AM_STAT_CANDIDATES] = (double)cands;
    c->stats[AM_STAT_REDO_REGIONS] = (double)redo;
    c->stats[AM_STAT_REGION_LAUNCHES] = (double)launches;
    c->stats[AM_STAT_CHAIN64_MS] = T[T_CHAIN64];
    c->stats[AM_STAT_WALK64_MS] = T[T_SCAN64] + T[T_REGION64];
    c->stats[AM_STAT_RAY64_MS] = T[T_RAY64];
    c->stats[AM_STAT_SCAN64_REGIONS] = (double)(f64.scanned_host + acc[0]);
    c->stats[AM_STAT_RAY64_UNDECIDED] = (double)acc[0];
    c->stats[AM_STAT_CHAIN64_LAUNCHES] = (double)f64.chain_launches;
    c->stats[AM_STAT_VERT64_MS] = T[T_VERT64];
    c->stats[AM_STAT_VERT64_REGIONS] = (double)vert64;
    c->stats[AM_STAT_SETUP_MS] = std::chrono::duration<double, std::milli>(t_setup - t0).count();
    c->stats[AM_STAT_FP64_ONLY] = reduced_asked && fp64_form ? 1.0 : 0.0;
    c->stats[AM_STAT_HOST_WAITS] = (double)(levels + 1);
    c->stats[AM_STAT_GRAPH_LAUNCHES] = c->graph_levels ? (double)levels : 0.0;
    c->stats[AM_STAT_NEWTON_MS] = T[T_NEWTON];
    c->stats[AM_STAT_NEWTON_VERTICES] = (double)newton_n;
    for (int w = 0; w < AM_WHY_COUNT && AM_STAT_WHY0 + w < AM_STAT_COUNT; ++w) c->stats[AM_STAT_WHY0 + w] = (double)why[w];
    c->has_run = true;
    return AM_OK;
  });
}

// dst row i = the first copy_words words of src row ids[i] (row strides in words).
__global__ void k_gather_words(const uint32_t* src, int src_stride, int copy_words, const int* ids, int n,
                               uint32_t* dst, int dst_stride) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * copy_words) return;
  const int i = (int)(t / copy_words), w = (int)(t - (size_t)i * copy_words);
  dst[(size_t)i * dst_stride + w] = src[(size_t)ids[i] * src_stride + w];
}

// Region scan of the mesh assembly; returns (faces, items).
static void combine_counts(am_ctx* c, const CombineIn& in, CombineWork& w, long long& F, long long& N) {
  combine_count(in, w, c->stream);
  unsigned long long tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, w.totals, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  F = (long long)(tot >> 32);
  N = (long long)(tot & 0xffffffffu);
}

static long long combine_vertices(am_ctx* c, const CombineWork& w) {
  unsigned v = 0;
  HIPCHK(hipMemcpyAsync(&v, (const unsigned*)(w.totals + 1), sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return (long long)v;
}

static CombineIn combine_in(const am_ctx* c) {
  CombineIn in;
  in.reg_mask = c->reg_mask.p;
  in.info = c->reg_info.p;
  in.vpair = c->reg_vpair.p;
  in.verts = c->reg_verts.p;
  in.R = (int)c->R;
  in.Wd = c->Wd;
  in.vm = c->vm;
  in.L = c->L;
  return in;
}

// The first use of each vertex whose record carries the reduced walk's
// inverse system (info[7] = 1), compacted in use order.
// Vertices of the Newton step: the first use of each vertex, in a region whose
// walk kept inverse systems, unless the walk marked it as needing no step (a
// NaN first entry of its inverse system).
struct NewtonSrc {
  const unsigned* vflag;
  const int* item_region;
  const int* info;
  const unsigned long long* roff;
  const uint16_t* jinv;
  int vm;
  __device__ int operator()(size_t i) const {
    if (!vflag[i]) return 0;
    const int r = item_region[i];
    if (info[(size_t)r * AM_INFO_STRIDE + 7] != 1) return 0;
    const int k = (int)i - (int)(roff[r] & 0xffffffffu);
    const uint16_t h = jinv[((size_t)r * vm + k) * 9];
    return (h & 0x7fffu) > 0x7f80u ? 0 : 1;
  }
};
struct NewtonSink {
  int* items;
  __device__ void operator()(size_t i, int prefix, int v) const {
    if (v) items[prefix] = (int)i;
  }
};

long long newton_refine(am_ctx* c) {
  const hipStream_t st = c->stream;
  const size_t R = (size_t)c->R;
  if (R * c->vm > (size_t)INT_MAX / 2) throw CapErr{"too many (region, vertex) slots for the vertex refinement"};
  const CombineIn in = combine_in(c);
  CombineWork& w = c->cwork;
  w = CombineWork{};
  ensure(c->cw_reg, combine_region_bytes((long long)R), false, st);
  combine_carve_regions(c->cw_reg.p, (long long)R, w);
  long long F = 0, N = 0;
  combine_counts(c, in, w, F, N);
  if (N == 0) return 0;
  ensure(c->cw_item, combine_item_bytes(N), false, st);
  combine_carve_items(c->cw_item.p, N, w);
  combine_dedup(in, w, N, st);
  c->dedup_ready = true;
  c->dedup_F = F;
  c->dedup_N = N;
  ensure(c->redo_ids, (size_t)N, false, st);
  ensure(c->redo_cnt, 1, false, st);
  ensure(c->scan_part, scan_blocks((size_t)N) + 1, false, st);
  scan_launch<int>(NewtonSrc{w.vflag, w.item_region, c->reg_info.p, w.roff, c->reg_jinv.p, c->vm}, (size_t)N,
                   NewtonSink{c->redo_ids.p},
                   c->scan_part.p, c->redo_cnt.p, st);
  int n = 0;
  HIPCHK(hipMemcpyAsync(&n, c->redo_cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  NewtonIO io;
  io.reg_mask = c->reg_mask.p;
  io.vpair = c->reg_vpair.p;
  io.verts = c->reg_verts.p;
  io.jinv = c->reg_jinv.p;
  io.items = c->redo_ids.p;
  io.count = c->redo_cnt.p;
  io.item_region = w.item_region;
  io.roff = w.roff;
  if (n > 0) {
    const int Ch = (c->C + 63) & ~63;
    hipLaunchKernelGGL(k_newton_fp64, dim3((unsigned)((n + AM_NT - 1) / AM_NT)), dim3(AM_NEWTON_THREADS),
                       newton_lds_bytes(c->Wd), st, c->net, chain_io(c, nullptr, 0, 0, Ch), io);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_newton_scatter, dim3(blocks((size_t)N, 256)), dim3(256), 0, st, io, (const uint32_t*)w.table,
                     (const int*)w.slot, (const unsigned*)w.vflag, (const int*)c->reg_info.p, (int)N, c->vm);
  HIPCHK(hipGetLastError());
  return n;
}

am_status am_combine(am_ctx* c, double scale, const double* center3) {
  if (!c) return AM_ERR_ARG;
  if (!c->has_run) return fail(c, AM_ERR_STATE, "am_combine called before am_run");
  return guarded(c, [&]() -> am_status {
    const auto t0 = std::chrono::steady_clock::now();
    const hipStream_t st = c->stream;
    const double c3[3] = {center3 ? center3[0] : 0.0, center3 ? center3[1] : 0.0, center3 ? center3[2] : 0.0};
    const size_t R = (size_t)c->R;
    if (R * c->vm > (size_t)INT_MAX / 2) throw CapErr{"too many (region, vertex) slots for am_combine"};
    c->mesh_v.clear();
    c->mesh_off.assign(1, 0);
    c->mesh_idx.clear();
    c->mesh_V = c->mesh_F = c->mesh_NI = 0;
    c->mesh_on_host = false;
    c->has_mesh = false;
    if (R == 0) {
      c->has_mesh = true;
      c->mesh_on_host = true;
      return AM_OK;
    }
    const CombineIn in = combine_in(c);
    CombineWork& w = c->cwork;
    long long F = 0, N = 0;
    if (c->dedup_ready) {
      // the dedup of this region store is already in cwork (vertex
      // coordinates, which the Newton step refined, are read by the assign)
      F = c->dedup_F;
      N = c->dedup_N;
    } else {
      w = CombineWork{};
      ensure(c->cw_reg, combine_region_bytes((long long)R), false, st);
      combine_carve_regions(c->cw_reg.p, (long long)R, w);
      combine_counts(c, in, w, F, N);
    }
    long long V = 0;
    if (N > 0) {
      if (!c->dedup_ready) {
        ensure(c->cw_item, combine_item_bytes(N), false, st);
        combine_carve_items(c->cw_item.p, N, w);
        combine_dedup(in, w, N, st);
        c->dedup_ready = true;
        c->dedup_F = F;
        c->dedup_N = N;
      }
      V = combine_vertices(c, w);
      ensure(c->m_xyz, (size_t)std::max<long long>(V * 3, 1), false, st);
      ensure(c->m_face_off, (size_t)F + 1, false, st);
      ensure(c->m_face_idx, (size_t)N, false, st);
      CombineOut out{c->m_xyz.p, c->m_face_off.p, c->m_face_idx.p, (int)F};
      combine_assign(in, w, out, N, scale, c3, st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));
    }
    c->mesh_V = V;
    c->mesh_F = F;
    c->mesh_NI = N;
    c->has_mesh = true;
    c->stats[AM_STAT_COMBINE_MS] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return AM_OK;
  });
}

am_status am_combine_noalloc(am_ctx* c, double scale, const double* center3, double* vertices,
                             long long vertex_cap, int* face_offsets, long long face_cap, int* face_indices,
                             long long index_cap, long long* counts3) {
  if (!c) return AM_ERR_ARG;
  if (!c->has_run) return fail(c, AM_ERR_STATE, "am_combine_noalloc called before am_run");
  return guarded(c, [&]() -> am_status {
    if (!vertices || !face_offsets || !face_indices || !counts3) throw ArgErr{"null output buffer"};
    for (const void* p : {(const void*)vertices, (const void*)face_offsets, (const void*)face_indices}) {
      hipPointerAttribute_t a{};
      if (hipPointerGetAttributes(&a, p) != hipSuccess ||
          (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged)) {
        (void)hipGetLastError();
        throw ArgErr{"am_combine_noalloc writes its outputs from the device: pass device memory"};
      }
    }
    const hipStream_t st = c->stream;
    const double c3[3] = {center3 ? center3[0] : 0.0, center3 ? center3[1] : 0.0, center3 ? center3[2] : 0.0};
    const long long R = c->R;
    counts3[0] = counts3[1] = counts3[2] = 0;
    if (R > AM_NOALLOC_MAX_REGIONS)
      throw CapErr{"am_combine_noalloc takes at most " + std::to_string(AM_NOALLOC_MAX_REGIONS) + " regions (" +
                   std::to_string(R) + " given)"};
    if (face_cap < 1) throw CapErr{"face_offsets needs room for at least one entry"};
    if (R == 0) {
      const int zero = 0;
      HIPCHK(hipMemcpyAsync(face_offsets, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
      return AM_OK;
    }
    void* base = nullptr;
    size_t bytes = 0;
    HIPCHK(noalloc_scratch(&base, &bytes));
    const size_t rb = combine_region_bytes(R);
    const CombineIn in = combine_in(c);
    CombineWork w;
    combine_carve_regions((char*)base, R, w);
    long long F = 0, N = 0;
    combine_counts(c, in, w, F, N);
    if (N > AM_NOALLOC_MAX_ITEMS || rb + combine_item_bytes(N) > bytes)
      throw CapErr{"am_combine_noalloc takes at most " + std::to_string(AM_NOALLOC_MAX_ITEMS) +
                   " vertex uses (" + std::to_string(N) + " in this mesh)"};
    if (F + 1 > face_cap || N > index_cap)
      throw CapErr{"output too small: the mesh has " + std::to_string(F) + " faces and " + std::to_string(N) +
                   " face indices"};
    long long V = 0;
    if (N > 0) {
      combine_carve_items((char*)base + rb, N, w);
      combine_dedup(in, w, N, st);
      V = combine_vertices(c, w);
      if (V > vertex_cap)
        throw CapErr{"output too small: the mesh has " + std::to_string(V) + " vertices"};
      CombineOut out{vertices, face_offsets, face_indices, (int)F};
      combine_assign(in, w, out, N, scale, c3, st);
      HIPCHK(hipGetLastError());
    } else {
      const int zero = 0;
      HIPCHK(hipMemcpyAsync(face_offsets, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    counts3[0] = V;
    counts3[1] = F;
    counts3[2] = N;
    return AM_OK;
  });
}

am_status am_get_counts(am_ctx* c, long long* nr, long long* nf, long long* nv, long long* ni) {
  if (!c) return AM_ERR_ARG;
  if (nr) *nr = c->R;
  if (nf) *nf = c->has_mesh ? c->mesh_F : 0;
  if (nv) *nv = c->has_mesh ? c->mesh_V : 0;
  if (ni) *ni = c->has_mesh ? c->mesh_NI : 0;
  return AM_OK;
}

// Host copy of the combined mesh (for writers that walk it on the host).
static void mesh_to_host(am_ctx* c) {
  if (c->mesh_on_host) return;
  c->mesh_v.resize((size_t)c->mesh_V * 3);
  c->mesh_off.resize((size_t)c->mesh_F + 1);
  c->mesh_idx.resize((size_t)c->mesh_NI);
  if (c->mesh_V) HIPCHK(hipMemcpy(c->mesh_v.data(), c->m_xyz.p, (size_t)c->mesh_V * 3 * 8, hipMemcpyDeviceToHost));
  if (c->mesh_F) HIPCHK(hipMemcpy(c->mesh_off.data(), c->m_face_off.p, (size_t)c->mesh_F * 4, hipMemcpyDeviceToHost));
  if (c->mesh_NI)
    HIPCHK(hipMemcpy(c->mesh_idx.data(), c->m_face_idx.p, (size_t)c->mesh_NI * 4, hipMemcpyDeviceToHost));
  c->mesh_off[c->mesh_F] = (int)c->mesh_NI;
  c->mesh_on_host = true;
}

am_status am_get_mesh(am_ctx* c, double* vertices, int* face_offsets, int* face_indices) {
  if (!c) return AM_ERR_ARG;
  if (!c->has_mesh) return fail(c, AM_ERR_STATE, "am_get_mesh called before am_combine");
  return guarded(c, [&]() -> am_status {
    if (c->mesh_on_host) {
      if (vertices) std::memcpy(vertices, c->mesh_v.data(), c->mesh_v.size() * 8);
      if (face_offsets) std::memcpy(face_offsets, c->mesh_off.data(), c->mesh_off.size() * 4);
      if (face_indices) std::memcpy(face_indices, c->mesh_idx.data(), c->mesh_idx.size() * 4);
      return AM_OK;
    }
    // Device, pinned or managed caller arrays get the arrays straight from
    // device memory.  Pageable ones: each array goes by one DMA into a pinned
    // staging buffer the context keeps, then by a host copy over several
    // threads (a plain copy into pageable memory stages through small chunks
    // and one host thread: about 10 GB/s).
    struct Part {
      void* dst;
      const void* src;
      size_t bytes;
    };
    Part parts[3] = {{vertices, c->m_xyz.p, vertices ? (size_t)c->mesh_V * 3 * 8 : 0},
                     {face_offsets, c->m_face_off.p, face_offsets ? (size_t)c->mesh_F * 4 : 0},
                     {face_indices, c->m_face_idx.p, face_indices ? (size_t)c->mesh_NI * 4 : 0}};
    size_t staged = 0;
    bool stage[3] = {false, false, false};
    for (int i = 0; i < 3; ++i) {
      if (!parts[i].bytes) continue;
      stage[i] = pageable_host(parts[i].dst);
      if (stage[i]) staged += (parts[i].bytes + 255) & ~(size_t)255;
      else HIPCHK(hipMemcpy(parts[i].dst, parts[i].src, parts[i].bytes, hipMemcpyDefault));
    }
    if (staged) {
      if (staged > c->mesh_pin_bytes) {
        if (c->mesh_pin) HIPCHK(hipHostFree(c->mesh_pin));
        c->mesh_pin = nullptr;
        c->mesh_pin_bytes = 0;
        HIPCHK(hipHostMalloc((void**)&c->mesh_pin, staged, hipHostMallocDefault));
        c->mesh_pin_bytes = staged;
      }
      // chunks of 2 MB, each DMA followed by an event; host threads copy a
      // chunk out as soon as its event has passed, so the copies overlap the
      // later DMAs
      struct Chunk {
        void* dst;
        const unsigned char* src;
        size_t len;
      };
      std::vector<Chunk> chunks;
      const size_t kChunk = (size_t)2 << 20;
      size_t o = 0;
      for (int i = 0; i < 3; ++i)
        if (stage[i]) {
          for (size_t a = 0; a < parts[i].bytes; a += kChunk) {
            const size_t len = std::min(kChunk, parts[i].bytes - a);
            if (chunks.size() == c->mesh_ev.size()) {
              hipEvent_t e;
              HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
              c->mesh_ev.push_back(e);
            }
            HIPCHK(hipMemcpyAsync(c->mesh_pin + o + a, (const unsigned char*)parts[i].src + a, len,
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipEventRecord(c->mesh_ev[chunks.size()], c->stream));
            chunks.push_back(Chunk{(unsigned char*)parts[i].dst + a, c->mesh_pin + o + a, len});
          }
          o += (parts[i].bytes + 255) & ~(size_t)255;
        }
      const int nt = (int)std::min<size_t>(8, chunks.size());
      std::atomic<int> bad{0};
      auto work = [&](int id) {
        for (size_t k = id; k < chunks.size(); k += nt) {
          if (hipEventSynchronize(c->mesh_ev[k]) != hipSuccess) {
            bad = 1;
            return;
          }
          std::memcpy(chunks[k].dst, chunks[k].src, chunks[k].len);
        }
      };
      std::vector<std::thread> th;
      for (int id = 1; id < nt; ++id) th.emplace_back(work, id);
      work(0);
      for (auto& t : th) t.join();
      if (bad) HIPCHK(hipStreamSynchronize(c->stream));   // reports the failed copy
    }
    if (face_offsets) {
      const int ni = (int)c->mesh_NI;
      HIPCHK(hipMemcpy(face_offsets + c->mesh_F, &ni, 4, hipMemcpyDefault));
    }
    return AM_OK;
  });
}

am_status am_get_regions(am_ctx* c, uint32_t* masks, int* info, int* tabs, double* verts, int* vpairs,
                         double* points) {
  if (!c) return AM_ERR_ARG;
  if (!c->has_run) return fail(c, AM_ERR_STATE, "am_get_regions called before am_run");
  return guarded(c, [&]() -> am_status {
    const size_t R = (size_t)c->R;
    if (R == 0) return AM_OK;
    if (masks) HIPCHK(hipMemcpy(masks, c->reg_mask.p, R * c->Wd * 4, hipMemcpyDeviceToHost));
    if (info) HIPCHK(hipMemcpy(info, c->reg_info.p, R * AM_INFO_STRIDE * 4, hipMemcpyDeviceToHost));
    if (tabs) HIPCHK(hipMemcpy(tabs, c->reg_tab.p, R * (c->vm + 1) * 4, hipMemcpyDeviceToHost));
    if (verts) HIPCHK(hipMemcpy(verts, c->reg_verts.p, R * c->vm * 3 * 8, hipMemcpyDeviceToHost));
    if (vpairs) HIPCHK(hipMemcpy(vpairs, c->reg_vpair.p, R * c->vm * 2 * 4, hipMemcpyDeviceToHost));
    if (points) HIPCHK(hipMemcpy(points, c->reg_point.p, R * 3 * 8, hipMemcpyDeviceToHost));
    return AM_OK;
  });
}

am_status am_planes(am_ctx* c, const uint32_t* masks, int n, float* planes, float* eq, float* sigma) {
  if (!c || !masks || !planes || !eq || n < 1) return AM_ERR_ARG;
  if (!c->has_run || c->precision == AM_PREC_FP64 || c->precision == AM_PREC_FP64_RAY || !c->tile_ok)
    return fail(c, AM_ERR_STATE, "am_planes needs a reduced-precision context after am_run (on a net the "
                                 "reduced forms take: no skips, layers up to 256 wide)");
  return guarded(c, [&]() -> am_status {
    const hipStream_t st = c->stream;
    const int Cs = c->fast.dev.Cs;
    DevBuf<uint32_t> m;
    DevBuf<float> p, e;
    ensure(m, (size_t)n * c->Wd, false, st);
    ensure(p, (size_t)n * AM_PCOMP * Cs, false, st);
    ensure(e, (size_t)n * AM_EQS, false, st);
    HIPCHK(hipMemcpy(m.p, masks, (size_t)n * c->Wd * 4, hipMemcpyDefault));
    fast_planes(c->fast, m.p, n, p.p, e.p, st);
    std::vector<float> hp((size_t)n * AM_PCOMP * Cs), he((size_t)n * AM_EQS);
    HIPCHK(hipMemcpy(hp.data(), p.p, hp.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(he.data(), e.p, he.size() * 4, hipMemcpyDeviceToHost));
    for (int r = 0; r < n; ++r) {
      for (int q = 0; q < 4; ++q) {
        std::memcpy(planes + ((size_t)r * 4 + q) * c->C, hp.data() + ((size_t)r * AM_PCOMP + q) * Cs,
                    (size_t)c->C * 4);
        eq[(size_t)r * 4 + q] = he[(size_t)r * AM_EQS + q];
      }
      if (sigma && c->fast.dev.sig) {
        std::memcpy(sigma + (size_t)r * (c->C + 1), hp.data() + ((size_t)r * AM_PCOMP + 4) * Cs, (size_t)c->C * 4);
        sigma[(size_t)r * (c->C + 1) + c->C] = he[(size_t)r * AM_EQS + 4];
      } else if (sigma) {
        std::fill(sigma + (size_t)r * (c->C + 1), sigma + (size_t)(r + 1) * (c->C + 1), 0.f);   // no estimates
      }
    }
    return AM_OK;
  });
}

am_status am_topology_fingerprint(am_ctx* c, unsigned long long* out5) {
  if (!c || !out5) return AM_ERR_ARG;
  if (!c->has_run) return fail(c, AM_ERR_STATE, "am_topology_fingerprint called before am_run");
  return guarded(c, [&]() -> am_status {
    const hipStream_t st = c->stream;
    DevBuf<unsigned long long> acc;
    ensure(acc, 6, false, st);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t*)acc.p, (size_t)12, 0u);
    if (c->R > 0)
      hipLaunchKernelGGL(k_fingerprint, dim3(blocks((size_t)c->R, 256)), dim3(256), 0, st, c->reg_mask.p,
                         c->reg_info.p, c->reg_vpair.p, (int)c->R, c->Wd, c->Wd, c->vm, c->L, -1, acc.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out5, acc.p, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return AM_OK;
  });
}

am_status am_get_stats(am_ctx* c, double* out, int n) {
  if (!c || !out) return AM_ERR_ARG;
  for (int i = 0; i < n && i < AM_STAT_COUNT; ++i) out[i] = c->stats[i];
  return AM_OK;
}

am_status am_adjoint(am_ctx* c, const void* grad_vertices, int dtype, void* const* grad_weights,
                     void* const* grad_biases) {
  if (!c) return AM_ERR_ARG;
  if (!c->has_mesh) return fail(c, AM_ERR_STATE, "am_adjoint called before am_combine");
  return guarded(c, [&]() -> am_status {
    if (!grad_vertices || !grad_weights || !grad_biases) throw ArgErr{"null gradient buffer"};
    if (dtype != AM_DTYPE_F32 && dtype != AM_DTYPE_F64) throw ArgErr{"bad dtype"};
    for (auto& row : c->skips)
      if (!row.empty()) throw ArgErr{"the adjoint does not support skip connections"};
    const hipStream_t st = c->stream;
    const size_t V = (size_t)c->mesh_V;
    const int nl = c->D + 1;
    c->adj_dW.resize(nl);
    c->adj_db.resize(nl);
    c->adj_G.resize(nl);
    c->adj_P.resize(nl);
    DevBuf<double>& gv = c->adj_gv;
    c->stage_used = 0;
    upload(c, gv, grad_vertices, V * 3, dtype, st);
    std::vector<DevBuf<double>>&dW = c->adj_dW, &db = c->adj_db, &Gst = c->adj_G, &Pst = c->adj_P;
    for (int li = 0; li < nl; ++li) {
      const size_t nw = (size_t)c->nodes[li + 1] * c->nodes[li];
      ensure(dW[li], nw, false, st);
      ensure(db[li], (size_t)c->nodes[li + 1], false, st);
      HIPCHK(hipMemsetAsync(dW[li].p, 0, nw * 8, st));
      HIPCHK(hipMemsetAsync(db[li].p, 0, (size_t)c->nodes[li + 1] * 8, st));
    }
    const int R = (int)c->R;
    // tiled: planes from the tiled chain, the pull-back as GEMMs over all
    // regions of a chunk; else one kernel per region does chain and pull-back
    const bool tiled = c->tile_ok && c->adjoint_tiled;
    // regions per chunk: the chunk's planes (4 x Ch doubles per region) and
    // its G and P arrays (4 x the layer widths, twice) each stay within 1 GiB
    const int Chp = (c->C + 63) & ~63;
    long long widths = 0;
    for (int li = 0; li <= c->D + 1; ++li) widths += c->nodes[li];
    long long chl = std::min<long long>(65536, ((long long)1 << 30) / (32LL * Chp));
    chl = std::min<long long>(chl, ((long long)1 << 30) / (64LL * widths));
    const int CH = (int)std::max<long long>(AM_CT, chl & ~(long long)(AM_CT - 1));
    const int chunk = std::min(R, CH);
    for (int li = 0; li < nl; ++li) {
      ensure(Gst[li], (size_t)chunk * 4 * c->nodes[li + 1], false, st);
      if (li >= 1) ensure(Pst[li], (size_t)chunk * 4 * c->nodes[li], false, st);
    }
    const int Ch = (c->C + 63) & ~63;
    size_t lds;
    const int dcap = 2 * c->vm;
    if (tiled) {
      ensure(c->tplanes, (size_t)chunk * 4 * Ch, false, st);
      ensure(c->teq, (size_t)chunk * 4, false, st);
      ensure(c->adj_dn, (size_t)chunk, false, st);
      ensure(c->adj_dg, (size_t)chunk * dcap, false, st);
      ensure(c->adj_dv, (size_t)chunk * dcap * 4, false, st);
      lds = (size_t)4 * c->L * 8 + (size_t)c->L * 4;
      if (lds > c->max_lds) throw CapErr{"network too wide for the adjoint kernel's LDS"};
      if (lds > 65536)
        HIPCHK(hipFuncSetAttribute((const void*)k_adjoint_direct, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    } else {
      lds = (size_t)8 * c->L * 8 + (size_t)c->Wd * 4 + 64;
      if (lds > c->max_lds) throw CapErr{"network too wide for the adjoint kernel's LDS"};
      if (lds > 65536)
        HIPCHK(hipFuncSetAttribute((const void*)k_adjoint_fp64, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    }
    // dW_li = G_li+1^T P_li with K = 4 x regions split into nz slices (at
    // least 512 K each, at most 128 slices within 32 MiB per layer), each
    // stored whole and added in a fixed order: the gradients do not depend on
    // the order in which blocks finish
    auto k_slices = [&](int K, int M, int N, int& kchunk) {
      const long long cap = std::min<long long>(128, std::max<long long>(8, ((long long)1 << 22) / ((long long)M * N)));
      const int nz = (int)std::min<long long>((K + 511) / 512, cap);
      kchunk = ((K + nz - 1) / nz + 15) & ~15;
      return (K + kchunk - 1) / kchunk;
    };
    {
      size_t need = 0;
      const int ny = (chunk + 255) / 256;
      for (int li = 1; li < nl; ++li) {
        int kc;
        const int M = c->nodes[li + 1], N = c->nodes[li];
        need = std::max(need, (size_t)k_slices(chunk * 4, M, N, kc) * M * N);
        need = std::max(need, (size_t)ny * M);
      }
      need = std::max(need, (size_t)ny * 4 * c->nodes[1]);
      ensure(c->adj_part, need, false, st);
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int r0 = 0; r0 < R; r0 += CH) {
      const int cnt = std::min(CH, R - r0);
      AdjointIO io;
      io.reg_mask = c->reg_mask.p;
      io.info = c->reg_info.p;
      io.vpair = c->reg_vpair.p;
      io.item_off = c->cwork.roff;
      io.vfirst = c->cwork.vflag;
      io.vid = c->m_face_idx.p;
      io.gradv = gv.p;
      io.r0 = r0;
      io.count = cnt;
      io.vm = c->vm;
      for (int li = 0; li < nl; ++li) {
        io.Gst[li] = Gst[li].p;
        io.Pst[li] = li >= 1 ? Pst[li].p : nullptr;
      }
      io.dn = c->adj_dn.p;
      io.dg = c->adj_dg.p;
      io.dv = c->adj_dv.p;
      io.dcap = dcap;
      const uint32_t* mask0 = c->reg_mask.p + (size_t)r0 * c->Wd;
      if (tiled) {
        hipLaunchKernelGGL(k_chain_tile_fp64, dim3((cnt + AM_CT - 1) / AM_CT), dim3(AM_CHAIN_THREADS),
                           chain_lds_bytes(c), st, c->net, chain_io(c, nullptr, r0, cnt, Ch));
        hipLaunchKernelGGL(k_adjoint_direct, dim3(cnt), dim3(64), lds, st, c->net, io, (const double*)c->tplanes.p,
                           (const double*)c->teq.p, Ch);
        HIPCHK(hipGetLastError());
        // G of hidden layer li = mask (G_li+1 W_li) + its direct terms, li = D .. 1
        // (rows: region x component)
        for (int li = nl - 1; li >= 1; --li) {
          const int nout = c->nodes[li + 1], nin = c->nodes[li], M = cnt * 4;
          hipLaunchKernelGGL((c->adjoint_mfma ? k_adj_mfma<false, 0> : k_adj_gemm<false, 0>),
                             dim3((nin + 63) / 64, (M + 63) / 64, 1), dim3(256), 0, st, (const double*)Gst[li].p,
                             nout, c->net.Wrow[li], nin, Gst[li - 1].p, nin, M, nin, nout, nout, mask0, c->Wd,
                             c->net.off[li - 1]);
          hipLaunchKernelGGL(k_adj_scatter, dim3((unsigned)(((size_t)cnt * dcap + 255) / 256)), dim3(256), 0, st,
                             (const int*)c->adj_dn.p, (const int*)c->adj_dg.p, (const double*)c->adj_dv.p, cnt, dcap,
                             c->net.off[li - 1], nin, Gst[li - 1].p);
        }
        HIPCHK(hipGetLastError());
      } else {
        hipLaunchKernelGGL(k_adjoint_fp64, dim3(cnt), dim3(64), lds, st, c->net, io);
        HIPCHK(hipGetLastError());
      }
      // dW_li = sum over (region, component) of G_li+1^T P_li; db_li the c
      // components; partials of blocks added in a fixed order (k_slice_sum)
      const int K = cnt * 4, ny = (cnt + 255) / 256;
      double* part = c->adj_part.p;
      auto slice_sum = [&](const double* src, int nslice, size_t sstride, int n, double* out, int ostride) {
        hipLaunchKernelGGL(k_slice_sum, dim3((n + 255) / 256), dim3(256), 0, st, src, nslice, sstride, n, out,
                           ostride);
      };
      for (int li = 1; li < nl; ++li) {
        const int M = c->nodes[li + 1], N = c->nodes[li];
        int kchunk;
        const int nz = k_slices(K, M, N, kchunk);
        hipLaunchKernelGGL((c->adjoint_mfma ? k_adj_mfma<true, 1> : k_adj_gemm<true, 1>),
                           dim3((N + 63) / 64, (M + 63) / 64, nz), dim3(256), 0, st, (const double*)Gst[li].p, M,
                           (const double*)Pst[li].p, N, part, N, M, N, K, kchunk, (const uint32_t*)nullptr, 0, 0);
        slice_sum(part, nz, (size_t)M * N, M * N, dW[li].p, 1);
        hipLaunchKernelGGL(k_col_sum, dim3((M + 63) / 64, ny), dim3(256), 0, st, (const double*)Gst[li].p, cnt, M, 3,
                           part);
        slice_sum(part, ny, (size_t)M, M, db[li].p, 1);
      }
      const int n1 = c->nodes[1];
      hipLaunchKernelGGL(k_col_sum4, dim3((n1 + 63) / 64, ny), dim3(256), 0, st, (const double*)Gst[0].p, cnt, n1,
                         part);
      for (int q = 0; q < 4; ++q)
        slice_sum(part + (size_t)q * n1, ny, (size_t)4 * n1, n1, q < 3 ? dW[0].p + q : db[0].p, q < 3 ? 3 : 1);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    c->stats[12] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int li = 0; li < nl; ++li) {
      const size_t nw = (size_t)c->nodes[li + 1] * c->nodes[li], nb = (size_t)c->nodes[li + 1];
      std::vector<double> hw(nw), hb(nb);
      HIPCHK(hipMemcpy(hw.data(), dW[li].p, nw * 8, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(hb.data(), db[li].p, nb * 8, hipMemcpyDeviceToHost));
      if (dtype == AM_DTYPE_F64) {
        HIPCHK(hipMemcpy(grad_weights[li], hw.data(), nw * 8, hipMemcpyDefault));
        HIPCHK(hipMemcpy(grad_biases[li], hb.data(), nb * 8, hipMemcpyDefault));
      } else {
        std::vector<float> fw(hw.begin(), hw.end()), fb(hb.begin(), hb.end());
        HIPCHK(hipMemcpy(grad_weights[li], fw.data(), nw * 4, hipMemcpyDefault));
        HIPCHK(hipMemcpy(grad_biases[li], fb.data(), nb * 4, hipMemcpyDefault));
      }
    }
    return AM_OK;
  });
}

am_status am_export_ply(am_ctx* c, const char* path, int is_polymesh, int is_float32) {
  if (!c || !path) return AM_ERR_ARG;
  if (!c->has_mesh) return fail(c, AM_ERR_STATE, "am_export_ply called before am_combine");
  try {
    mesh_to_host(c);
  } catch (const std::exception& e) {
    return fail(c, AM_ERR_HIP, e.what());
  }
  const long long V = (long long)c->mesh_v.size() / 3;
  const long long F = (long long)c->mesh_off.size() - 1;
  long long T = 0;
  for (long long f = 0; f < F; ++f) T += c->mesh_off[f + 1] - c->mesh_off[f] - 2;
  std::ofstream o(path, std::ios::out | std::ios::binary);
  if (!o) return fail(c, AM_ERR_IO, std::string("cannot open ") + path);
  // header layout of backend/inc/polymesh.h:368-377
  o << "ply\n";
  o << "format binary_little_endian 1.0\n";
  o << "element vertex " << V << "\n";
  const char* ft = is_float32 ? "float" : "double";
  o << "property " << ft << " x\n";
  o << "property " << ft << " y\n";
  o << "property " << ft << " z\n";
  o << "element face " << (is_polymesh ? F : T) << "\n";
  o << "property list uchar int vertex_index\n";
  o << "end_header\n";
  if (is_float32) {
    std::vector<float> f(c->mesh_v.size());
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)c->mesh_v[i];
    o.write((const char*)f.data(), f.size() * 4);
  } else {
    o.write((const char*)c->mesh_v.data(), c->mesh_v.size() * 8);
  }
  std::vector<char> buf;
  buf.reserve((size_t)(is_polymesh ? F * 5 + c->mesh_idx.size() * 4 : T * 13));
  auto put_i = [&](int v) {
    const char* p = (const char*)&v;
    buf.insert(buf.end(), p, p + 4);
  };
  for (long long f = 0; f < F; ++f) {
    const int a = c->mesh_off[f], n = c->mesh_off[f + 1] - a;
    if (is_polymesh) {
      buf.push_back((char)(unsigned char)n);
      for (int j = 0; j < n; ++j) put_i(c->mesh_idx[a + j]);
    } else {
      for (int j = 0; j < n - 2; ++j) {  // fan (backend/inc/polymesh.h:405-419)
        buf.push_back((char)3);
        put_i(c->mesh_idx[a]);
        put_i(c->mesh_idx[a + j + 1]);
        put_i(c->mesh_idx[a + j + 2]);
      }
    }
  }
  o.write(buf.data(), buf.size());
  o.close();
  if (!o) return fail(c, AM_ERR_IO, std::string("write failed: ") + path);
  return AM_OK;
}

am_status am_destroy(am_ctx* c) {
  if (!c) return AM_ERR_ARG;
  if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
  for (hipGraphExec_t& e : c->gexec)
    if (e) (void)hipGraphExecDestroy(e);
  c->timer.release();
  fast_release(c->fbuf);
  if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
  if (c->host_total) (void)hipHostFree(c->host_total);   // every level waited for its total
  if (c->mesh_pin) (void)hipHostFree(c->mesh_pin);
  for (hipEvent_t e : c->mesh_ev) (void)hipEventDestroy(e);
  hipStream_t own = c->own_stream;
  delete c;
  if (own) (void)hipStreamDestroy(own);
  return AM_OK;
}

const char* am_last_error(const am_ctx* c) {
  if (!c) return g_create_error.c_str();
  return c->err.c_str();
}


// ---------------------------------------------------------------------------
// batch of networks
// ---------------------------------------------------------------------------
am_status am_batch_create(am_batch** out, int n_nets, const int* n_nodes, const int* const* nodes, int precision) {
  if (!out) return AM_ERR_ARG;
  *out = nullptr;
  if (n_nets < 1 || !n_nodes || !nodes) {
    g_create_error = "a batch needs at least one network";
    return AM_ERR_ARG;
  }
  if (precision != AM_PREC_FP64 && precision != AM_PREC_FP64_RAY) {
    g_create_error = "a batch runs the FP64 forms (AM_PREC_FP64, AM_PREC_FP64_RAY)";
    return AM_ERR_ARG;
  }
  if (n_nets > (1 << 20)) {
    g_create_error = "too many networks in one batch";
    return AM_ERR_ARG;
  }
  am_batch* b = new am_batch();
  b->precision = precision;
  int Lmax = 0, Dmax = 0, Wmax = 0;
  for (int i = 0; i < n_nets; ++i) {
    const int nn = n_nodes[i];
    const int* nd = nodes[i];
    if (nn < 3 || !nd || nd[0] != 3 || nd[nn - 1] != 1 || nn - 1 > AM_MAX_LINEAR) {
      g_create_error = "network " + std::to_string(i) + ": nodes must start with 3, end with 1, 3..17 entries";
      delete b;
      return AM_ERR_ARG;
    }
    BatchNet bn;
    bn.nodes.assign(nd, nd + nn);
    bn.D = nn - 2;
    bn.off.assign(bn.D + 1, 0);
    for (int k = 0; k < bn.D; ++k) bn.off[k + 1] = bn.off[k] + nd[k + 1];
    bn.L = bn.off[bn.D];
    bn.Wn = (bn.L + 31) / 32;
    for (int k = 1; k < nn; ++k)
      if (nd[k] > AM_CHAIN_MAXW) {
        g_create_error = "network " + std::to_string(i) + ": a batch takes layers up to 256 wide";
        delete b;
        return AM_ERR_ARG;
      }
    Lmax = std::max(Lmax, bn.L);
    Dmax = std::max(Dmax, bn.D);
    Wmax = std::max(Wmax, bn.Wn);
    b->nets.push_back(std::move(bn));
  }
  if (Wmax + 1 > AM_MAX_WORDS) {
    g_create_error = "networks too wide for a batch";
    delete b;
    return AM_ERR_ARG;
  }
  // the shared context: masks of Wmax words + the network word
  std::vector<int> n0 = {3, std::max(1, Lmax), 1};
  const int arc = 0;
  am_ctx* c = nullptr;
  const am_status s = am_create(&c, 3, n0.data(), &arc, 1, 1, 0, AM_PREC_FP64);
  if (s != AM_OK) {
    delete b;
    return s;
  }
  c->Wd = Wmax + 1;
  c->L = Lmax;
  c->C = Lmax;
  b->c = c;
  b->Wmax = Wmax;
  b->Dmax = Dmax;
  b->Lmax = Lmax;
  b->Cmax = Lmax;
  *out = b;
  return AM_OK;
}

am_status am_batch_set_option(am_batch* b, const char* key, long long value) {
  if (!b || !key) return AM_ERR_ARG;
  const std::string k(key);
  if (k == "vm" || k == "table_log2" || k == "region_limit" || k == "ray_waves" || k == "scan_blocks" ||
      k == "graph_levels")
    return am_set_option(b->c, key, value);
  return fail(b->c, AM_ERR_ARG, "unknown batch option " + k);
}

am_status am_batch_run(am_batch* b, const void* const* weights, const void* const* biases, int dtype,
                       const void* const* points, const uint8_t* const* states, const int* n_seeds, double iso,
                       int flip_insideout, void* stream) {
  if (!b) return AM_ERR_ARG;
  am_ctx* c = b->c;
  return guarded(c, [&]() -> am_status {
    if (!weights || !biases || !points || !states || !n_seeds) throw ArgErr{"null argument"};
    if (dtype != AM_DTYPE_F32 && dtype != AM_DTYPE_F64) throw ArgErr{"dtype must be AM_DTYPE_F32 or AM_DTYPE_F64"};
    const int nn = (int)b->nets.size();
    long long total_seeds = 0;
    for (int i = 0; i < nn; ++i) {
      if (n_seeds[i] < 1 || !points[i] || !states[i]) throw ArgErr{"every network needs at least one seed"};
      total_seeds += n_seeds[i];
    }
    if (total_seeds > INT_MAX / 2) throw CapErr{"too many seeds"};
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    const hipStream_t st = c->stream;
    const auto t0 = std::chrono::steady_clock::now();
    for (double& x : c->stats) x = 0;
    c->has_run = false;
    c->has_mesh = false;
    c->dedup_ready = false;
    c->stage_used = 0;
    const size_t esz = dtype == AM_DTYPE_F64 ? 8 : 4;
    const int Ch = (b->Cmax + 63) & ~63;
    b->nets_h.assign(nn, NetDev{});
    b->cios_h.assign(nn, ChainIO{});
    size_t wi = 0;
    for (int i = 0; i < nn; ++i) {
      BatchNet& bn = b->nets[i];
      const int nl = bn.D + 1;
      std::vector<const void*> wv(nl), bv(nl);
      for (int li = 0; li < nl; ++li) {
        wv[li] = dev_view(c, weights[wi + li], (size_t)bn.nodes[li + 1] * bn.nodes[li] * esz, st);
        bv[li] = dev_view(c, biases[wi + li], (size_t)bn.nodes[li + 1] * esz, st);
      }
      wi += nl;
      bn.WT.resize(nl);
      bn.b.resize(nl);
      bn.WTp.resize(nl);
      bn.bpad.resize(nl);
      upload(c, bn.W0, wv[0], (size_t)bn.nodes[1] * 3, dtype, st);
      for (int li = 0; li < nl; ++li) {
        upload(c, bn.b[li], bv[li], (size_t)bn.nodes[li + 1], dtype, st);
        if (li >= 1) {
          upload_T(c, bn.WT[li], wv[li], bn.nodes[li + 1], bn.nodes[li], dtype, st);
          upload_pad(c, bn.WTp[li], bn.bpad[li], wv[li], bv[li], bn.nodes[li + 1], bn.nodes[li], dtype, st);
        }
      }
      NetDev& n = b->nets_h[i];
      std::memset(&n, 0, sizeof(n));
      n.D = bn.D; n.L = bn.L; n.E = 0; n.C = bn.L; n.Wd = c->Wd; n.vm = c->vm;
      for (int k = 0; k < (int)bn.nodes.size(); ++k) n.nodes[k] = bn.nodes[k];
      for (int k = 0; k <= bn.D; ++k) n.off[k] = bn.off[k];
      n.W0 = bn.W0.p;
      for (int li = 0; li < nl; ++li) {
        n.WT[li] = li >= 1 ? bn.WT[li].p : nullptr;
        n.b[li] = bn.b[li].p;
      }
      n.iso = iso;
      n.flip = flip_insideout ? 1 : 0;
      ChainIO& cio = b->cios_h[i];
      std::memset(&cio, 0, sizeof(cio));
      cio.Cs = Ch;
      for (int li = 1; li <= bn.D; ++li) {
        cio.Wp[li] = bn.WTp[li].p;
        cio.bp[li] = bn.bpad[li].p;
        cio.noutp[li] = (bn.nodes[li + 1] + AM_CR - 1) / AM_CR * AM_CR;
      }
    }
    ensure(b->nets_d, (size_t)nn, false, st);
    HIPCHK(hipMemcpyAsync(b->nets_d.p, b->nets_h.data(), nn * sizeof(NetDev), hipMemcpyHostToDevice, st));
    ensure(b->cios_d, (size_t)nn, false, st);
    ensure(b->hist_d, (size_t)nn, false, st);
    if (!b->hist_h) HIPCHK(hipHostMalloc((void**)&b->hist_h, (size_t)1 << 22, hipHostMallocDefault));
    if ((size_t)nn * sizeof(int) > ((size_t)1 << 22)) throw CapErr{"too many networks"};
    const size_t clds = batch_chain_lds(b), slds = batch_scan_lds(b), rlds = batch_ray_lds(b);
    if (clds > c->max_lds || slds > c->max_lds) throw CapErr{"networks too wide for the batch kernels' LDS"};
    if (clds > 65536) {
      HIPCHK(hipFuncSetAttribute((const void*)k_chain_tile_fp64_multi<false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)clds));
      HIPCHK(hipFuncSetAttribute((const void*)k_chain_tile_fp64_multi<true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)clds));
    }
    if (slds > 65536)
      HIPCHK(hipFuncSetAttribute((const void*)k_walk_fp64_multi, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)slds));
    if (rlds > 65536 && rlds <= c->max_lds)
      for (const void* f : {(const void*)k_walk_ray64_multi<2>, (const void*)k_walk_ray64_multi<4>,
                            (const void*)k_walk_ray64_multi<8>})
        HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)rlds));
    const auto t_setup = std::chrono::steady_clock::now();

    // seeds of every network -> the first level (network word appended)
    c->R = 0;
    c->table_cap = 0;
    c->table.release();
    grow_cands(c, (size_t)total_seeds);
    {
      std::vector<uint32_t> m((size_t)total_seeds * c->Wd, 0u);
      std::vector<double> pts((size_t)total_seeds * 3);
      size_t q = 0;
      for (int i = 0; i < nn; ++i) {
        const int L = b->nets[i].L, ns = n_seeds[i];
        std::vector<uint8_t> hs((size_t)ns * L);
        HIPCHK(hipMemcpy(hs.data(), states[i], hs.size(), hipMemcpyDefault));
        std::vector<double> p((size_t)ns * 3);
        if (dtype == AM_DTYPE_F64) {
          HIPCHK(hipMemcpy(p.data(), points[i], p.size() * 8, hipMemcpyDefault));
        } else {
          std::vector<float> f(p.size());
          HIPCHK(hipMemcpy(f.data(), points[i], f.size() * 4, hipMemcpyDefault));
          for (size_t k = 0; k < f.size(); ++k) p[k] = f[k];
        }
        for (int s2 = 0; s2 < ns; ++s2, ++q) {
          uint32_t* mw = m.data() + q * c->Wd;
          for (int g = 0; g < L; ++g)
            if (hs[(size_t)s2 * L + g]) mw[g >> 5] |= 1u << (g & 31);
          mw[c->Wd - 1] = (uint32_t)i;
          for (int k = 0; k < 3; ++k) pts[q * 3 + k] = p[(size_t)s2 * 3 + k];
        }
      }
      std::vector<int> startv((size_t)total_seeds, -1);
      HIPCHK(hipMemcpy(c->cand_mask.p, m.data(), m.size() * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->cand_point.p, pts.data(), pts.size() * 8, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->cand_start.p, startv.data(), startv.size() * 4, hipMemcpyHostToDevice));
    }
    c->timer.reset();
    constexpr int NACC = 2 * (1 + AM_WHY_COUNT);
    ensure(c->runacc, NACC, false, st);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, c->runacc.p, (size_t)NACC, 0);
    hipLaunchKernelGGL(k_hash_keys, dim3(blocks((size_t)total_seeds, 256)), dim3(256), 0, st,
                       (const uint32_t*)c->cand_mask.p, (int)total_seeds, c->Wd, c->cand_hash.p);
    long long lvl_first = 0;
    long long lvl_count = insert_level(c, (size_t)total_seeds);
    batch_hist(b, lvl_first, lvl_count);
    long long levels = 0, cands = 0, launches = 0, waits = 1;
    const uint32_t* mask_at_upload = nullptr;
    Fp64Counts f64;
    while (lvl_count > 0) {
      const size_t ncand = (size_t)lvl_count * c->vm;
      grow_cands(c, ncand);
      ensure(c->rflags, (size_t)lvl_count, false, st);
      // with graph-captured levels the store grows before the level's kernels
      // are captured, so the chain must see the grown store's masks
      InsertPlan ip;
      if (c->graph_levels) ip = insert_prepare(c, ncand);
      if (c->reg_mask.p != mask_at_upload) {
        // the chain reads region masks through each network's ChainIO
        for (ChainIO& cio : b->cios_h) cio.reg_mask = c->reg_mask.p;
        HIPCHK(hipMemcpyAsync(b->cios_d.p, b->cios_h.data(), nn * sizeof(ChainIO), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        mask_at_upload = c->reg_mask.p;
      }
      const BatchLevelPlan pl = batch_level_prepare(b, lvl_first, lvl_count);
      launches += 1;
      const long long before = c->R;
      int added = 0;
      if (c->graph_levels) {
        launch_as_graph(c, (int)levels, [&]() {
          batch_level_launch(b, pl, f64);
          insert_launch(c, ip);
        });
        added = insert_finish(c, ip);
      } else {
        batch_level_launch(b, pl, f64);
        const size_t mi = c->timer.mark(st);
        added = insert_level(c, ncand);
        c->timer.span(mi, c->timer.mark(st), T_INSERT);
      }
      cands += (long long)ncand;
      levels += 1;
      lvl_first = before;
      lvl_count = added;
      batch_hist(b, lvl_first, lvl_count);
      waits += 2;
    }
    int acc[NACC] = {0};
    HIPCHK(hipMemcpyAsync(acc, c->runacc.p, sizeof(acc), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    c->timer.resolve();
    const auto t1 = std::chrono::steady_clock::now();
    const double* T = c->timer.acc;
    c->stats[AM_STAT_RUN_MS] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    c->stats[AM_STAT_INSERT_MS] = T[T_INSERT];
    c->stats[AM_STAT_REGION_MS] = T[T_CHAIN64] + T[T_RAY64] + T[T_SCAN64];
    c->stats[AM_STAT_GEMM_MS] = T[T_CHAIN64];
    c->stats[AM_STAT_WALK_MS] = T[T_RAY64] + T[T_SCAN64];
    c->stats[AM_STAT_LEVELS] = (double)levels;
    c->stats[AM_STAT_REGIONS] = (double)c->R;
    c->stats[AM_STAT_CANDIDATES] = (double)cands;
    c->stats[AM_STAT_REGION_LAUNCHES] = (double)launches;
    c->stats[AM_STAT_CHAIN64_MS] = T[T_CHAIN64];
    c->stats[AM_STAT_WALK64_MS] = T[T_SCAN64];
    c->stats[AM_STAT_RAY64_MS] = T[T_RAY64];
    c->stats[AM_STAT_SCAN64_REGIONS] = (double)(f64.scanned_host + acc[0]);
    c->stats[AM_STAT_RAY64_UNDECIDED] = (double)acc[0];
    c->stats[AM_STAT_CHAIN64_LAUNCHES] = (double)f64.chain_launches;
    c->stats[AM_STAT_SETUP_MS] = std::chrono::duration<double, std::milli>(t_setup - t0).count();
    c->stats[AM_STAT_HOST_WAITS] = (double)waits;
    c->stats[AM_STAT_GRAPH_LAUNCHES] = c->graph_levels ? (double)levels : 0.0;
    c->has_run = true;
    return AM_OK;
  });
}

am_status am_batch_fingerprint(am_batch* b, int net, unsigned long long* out6) {
  if (!b || !out6) return AM_ERR_ARG;
  am_ctx* c = b->c;
  if (!c->has_run) return fail(c, AM_ERR_STATE, "am_batch_fingerprint called before am_batch_run");
  if (net < 0 || net >= (int)b->nets.size()) return fail(c, AM_ERR_ARG, "no such network");
  return guarded(c, [&]() -> am_status {
    const hipStream_t st = c->stream;
    DevBuf<unsigned long long> acc;
    ensure(acc, 6, false, st);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t*)acc.p, (size_t)12, 0u);
    if (c->R > 0)
      hipLaunchKernelGGL(k_fingerprint, dim3(blocks((size_t)c->R, 256)), dim3(256), 0, st, c->reg_mask.p,
                         c->reg_info.p, c->reg_vpair.p, (int)c->R, c->Wd, b->nets[net].Wn, c->vm, b->nets[net].L, net,
                         acc.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out6, acc.p, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return AM_OK;
  });
}

am_status am_batch_get_regions(am_batch* b, int net, long long* n_regions, uint32_t* masks, int* info, int* tabs,
                               double* verts, int* vpairs, double* points) {
  if (!b || !n_regions) return AM_ERR_ARG;
  am_ctx* c = b->c;
  if (!c->has_run) return fail(c, AM_ERR_STATE, "am_batch_get_regions called before am_batch_run");
  if (net < 0 || net >= (int)b->nets.size()) return fail(c, AM_ERR_ARG, "no such network");
  return guarded(c, [&]() -> am_status {
    const hipStream_t st = c->stream;
    const int R = (int)c->R;
    *n_regions = 0;
    if (R == 0) return AM_OK;
    ensure(b->sel_flags, (size_t)R, false, st);
    ensure(b->sel_rank, (size_t)R + 1, false, st);
    ensure(b->sel_ids, (size_t)R, false, st);
    hipLaunchKernelGGL(k_net_flags, dim3(blocks(R, 256)), dim3(256), 0, st, c->reg_mask.p, c->Wd, R, net,
                       b->sel_flags.p);
    ensure(c->scan_part, scan_blocks((size_t)R) + 1, false, st);
    scan_launch<int>(ArraySrc<int>{b->sel_flags.p}, (size_t)R, ArraySink<int>{b->sel_rank.p}, c->scan_part.p,
                     b->sel_rank.p + R, st);
    hipLaunchKernelGGL(k_net_scatter, dim3(blocks(R, 256)), dim3(256), 0, st, (const int*)b->sel_flags.p,
                       (const int*)b->sel_rank.p, R, b->sel_ids.p);
    HIPCHK(hipGetLastError());
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, b->sel_rank.p + R, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *n_regions = n;
    if (n == 0 || (!masks && !info && !tabs && !verts && !vpairs && !points)) return AM_OK;
    const int vm = c->vm, Wn = b->nets[net].Wn;
    DevBuf<uint32_t> tmp;
    auto gather = [&](const void* src, int stride_words, int copy_words, void* dst) {
      if (!dst) return;
      ensure(tmp, (size_t)n * copy_words, false, st);
      const size_t tot = (size_t)n * copy_words;
      hipLaunchKernelGGL(k_gather_words, dim3(blocks(tot, 256)), dim3(256), 0, st, (const uint32_t*)src,
                         stride_words, copy_words, (const int*)b->sel_ids.p, n, tmp.p, copy_words);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(dst, tmp.p, tot * 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    };
    gather(c->reg_mask.p, c->Wd, Wn, masks);
    gather(c->reg_info.p, AM_INFO_STRIDE, AM_INFO_STRIDE, info);
    gather(c->reg_tab.p, vm + 1, vm + 1, tabs);
    gather(c->reg_verts.p, vm * 6, vm * 6, verts);
    gather(c->reg_vpair.p, vm * 2, vm * 2, vpairs);
    gather(c->reg_point.p, 6, 6, points);
    return AM_OK;
  });
}

am_status am_batch_combine(am_batch* b, double scale, const double* center3) {
  if (!b) return AM_ERR_ARG;
  return am_combine(b->c, scale, center3);
}

am_status am_batch_get_counts(am_batch* b, long long* n_regions, long long* n_faces, long long* n_vertices,
                              long long* n_face_indices) {
  if (!b) return AM_ERR_ARG;
  return am_get_counts(b->c, n_regions, n_faces, n_vertices, n_face_indices);
}

am_status am_batch_get_stats(am_batch* b, double* out, int n) {
  if (!b) return AM_ERR_ARG;
  return am_get_stats(b->c, out, n);
}

const char* am_batch_last_error(const am_batch* b) { return b ? am_last_error(b->c) : am_last_error(nullptr); }

am_status am_batch_destroy(am_batch* b) {
  if (!b) return AM_ERR_ARG;
  if (b->c) (void)hipStreamSynchronize(b->c->stream);
  if (b->tiles_h) (void)hipHostFree(b->tiles_h);
  if (b->hist_h) (void)hipHostFree(b->hist_h);
  am_ctx* c = b->c;
  b->nets.clear();
  b->nets_d.release();
  b->cios_d.release();
  b->tiles_d.release();
  b->hist_d.release();
  b->sel_flags.release();
  b->sel_rank.release();
  b->sel_ids.release();
  delete b;
  if (c) am_destroy(c);
  return AM_OK;
}

}  // extern "C"
