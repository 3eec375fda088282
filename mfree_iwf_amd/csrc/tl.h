// Total-Lagrangian (Randles-Libersky corrected) SPH on gfx950: state and
// launch interface (internal).
#pragma once

#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mfree_iwf_amd.h"
#include "kernels.h"

namespace mfree {

// Buffers owned by the context that the TL path shares (storage is idx
// order and never reordered for TL bodies).
struct TlShared {
  hipStream_t st;
  int precision;
  uint32_t S, N;
  SceneConst* sc;
  SceneGeom* geom;
  Fields8 y, K, acc;
  double *m, *h;
  uint32_t *scene, *gid;
  const Chunk* chunks;
  const uint32_t* chunk_first;
  uint32_t nchunks;
  double* part;
  uint32_t *key, *rank, *count, *tsum, *cell_start, *perm;
  uint32_t* err;
  double h_search;  // <= 0: per-particle h
  const double *fcx, *fcy;  // contact forces by global idx (= storage index), or null
};

// Action lists of TL bodies: tensile_tl (Randles-Libersky), tensile_corot
// (co-rotated, plain kernel) and tensile_potential (elastic potential,
// Randles-Libersky kernel of both pair directions).
enum TlVariant : int { kTlRl = 0, kTlCorot = 1, kTlPotential = 2 };

struct TlState {
  bool ready = false;
  int variant = kTlRl;
  std::string msg;
  uint32_t N = 0;
  double *X = nullptr, *Y = nullptr;  // reference configuration
  uint8_t* bc = nullptr;              // 0 none, 1 left strip, 2 right strip
  uint32_t* rowp = nullptr;           // N+1, storage (= idx) space
  uint32_t* col = nullptr;
  double* w64 = nullptr;              // corrected (w, w_x, w_y) per pair
  float* w32 = nullptr;               // FP32 copy for the mixed path
  uint64_t npairs = 0;
  double* qw = nullptr;               // quadrature weight m/rho at precompute
  double* rec = nullptr;              // pass-1 records for pass 2 (96 B per particle)
  double* part8 = nullptr;            // pass-1 derivative parts (4N: rho_t, S_t)
  double* diag = nullptr;             // per-particle diagnostics (16N): H, Fdot, P (RL); H, R, P, S
                                      // (co-rotated); H, Fe, -, S (potential)
  double2* rot = nullptr;             // co-rotated: (cos t, sin t) of each particle's rotation
  uint32_t* rev = nullptr;            // potential: index of the reverse pair (j, i) of pair q, or ~0u
  Fields8 Kb{};                       // second stage buffer: pass 1 reads the
                                      // stage state, pass 2 writes the next one
  std::vector<void*> allocs;
  std::vector<double> hX, hY;
  std::vector<uint8_t> hbc;
};

using EventPairs = std::vector<std::pair<hipEvent_t, hipEvent_t>>;
using TakeEvent = std::function<hipEvent_t()>;

mfree_status tl_alloc(TlState& t, int variant, uint32_t N, const double* X, const double* Y,
                      const std::vector<std::vector<uint32_t>>& left,
                      const std::vector<std::vector<uint32_t>>& right, const std::vector<uint32_t>& off);
void tl_free(TlState& t);
mfree_status tl_ensure_precomputed(TlState& t, const TlShared& s);
void tl_enqueue_step(TlState& t, const TlShared& s, EventPairs* evs, const TakeEvent& take);
void tl_enqueue_eval(TlState& t, const TlShared& s, double* d8, double* aux);
mfree_status tl_get_pairs(TlState& t, const TlShared& s, uint32_t body, uint32_t o, uint32_t n, uint32_t* row_ptr,
                          uint32_t* col, uint64_t cap, uint64_t* npairs);

}  // namespace mfree
