// Compile-time limits and the network descriptor passed to every kernel.
#pragma once
#include <stdint.h>

#define AM_MAX_LINEAR 16        // linear layers (hidden layers + 1)
#define AM_MAX_SKIPS 4          // additive skips into one layer
#define AM_MAX_TMS 16           // transform matrices
#define AM_MAX_WORDS 64         // packed mask words: up to 2048 hidden neurons
#define AM_VM_MAX 64            // polygon vertex cap (one lane per vertex)
#define AM_TOPK 64              // distance-order positions the walk builds up front (sort width >= this)
#define AM_INFO_STRIDE 8

// Region status codes (shared with the host API).
#define AM_ST_OK 0
#define AM_ST_NO_START 1
#define AM_ST_GAVE_UP 2
#define AM_ST_BAD_START 3

namespace am {

// Network as the kernels see it.  Hidden node k (1-based) owns mask bits
// [off[k-1], off[k]).  WT[li] is linear layer li transposed to (in, out) so
// that lanes owning consecutive output neurons read consecutive addresses.
struct NetDev {
  int D;            // hidden layers
  int L;            // hidden neurons
  int E;            // extra planes
  int C;            // constraints = L + E
  int Wd;           // mask words
  int vm;           // polygon vertex cap (tab holds vm + 1 entries)
  int nodes[AM_MAX_LINEAR + 1];
  int off[AM_MAX_LINEAR + 1];
  const double* W0;                 // (n1, 3) row-major, fp64
  const double* WT[AM_MAX_LINEAR];  // li >= 1: (n_li, n_li+1) fp64
  const double* Wrow[AM_MAX_LINEAR];  // li >= 1: (n_li+1, n_li) fp64, as given (adjoint)
  const double* b[AM_MAX_LINEAR];
  int nskip[AM_MAX_LINEAR];
  int skip_from[AM_MAX_LINEAR][AM_MAX_SKIPS];
  int skip_tm[AM_MAX_LINEAR][AM_MAX_SKIPS];
  const double* tm[AM_MAX_TMS];     // row-major (h, w) as given
  const double* tmT[AM_MAX_TMS];    // transposed (w, h)
  int tm_h[AM_MAX_TMS];
  int tm_w[AM_MAX_TMS];
  const double* ew;                 // (E, 3)
  const double* eb;                 // (E,)
  double iso;
  int flip;
};

}  // namespace am
