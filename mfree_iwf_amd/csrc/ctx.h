// The context behind the C ABI (one set of bodies on one device) and the
// host helpers shared by capi.hip and surrogate.hip.
#pragma once

#include <algorithm>
#include <atomic>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mfree_iwf_amd.h"
#include "kernels.h"
#include "tl.h"

using mfree::Chunk;
using mfree::Fields8;
using mfree::PairRec32;
using mfree::PairRec64;
using mfree::SceneConst;
using mfree::SceneGeom;
using mfree::TlState;

namespace mfree {
struct RefPipe;   // reference-structured FP64 pipeline (baseline.hip)
struct VtkState;  // VTK writer: staging buffers, copy stream, worker (vtk.hip)
}

struct mfree_ctx {
  mfree_options opt;
  std::string err;
  bool committed = false;
  hipStream_t stream = nullptr;

  // body registry (host)
  std::vector<mfree_params> prm;
  std::vector<uint32_t> off{0};
  std::vector<double> time;
  std::vector<std::vector<uint32_t>> bc_left, bc_right;
  std::vector<double> v_bc;
  // initial host state in idx order, uploaded at commit
  std::vector<double> hx, hy, hX, hY, hm, hrho, hh, hvx, hvy, hSxx, hSxy, hSyy;

  // device
  uint32_t N = 0, S = 0, nchunks = 0;
  std::vector<void*> allocs;
  Fields8 y[2]{}, K{}, acc{};
  double *m[2]{}, *h[2]{};
  uint32_t *scene[2]{}, *gid[2]{};
  int cur = 0;
  uint32_t *key = nullptr, *rank = nullptr, *count = nullptr, *cell_start = nullptr, *perm = nullptr;
  uint32_t *tsum = nullptr, *errb = nullptr;
  uint32_t* berr = nullptr;       // errb + 1: the bodies' error words
  uint32_t* perm_keep = nullptr;  // isolated bodies: the step's stage-1 permutation
  uint32_t* perm_tmp = nullptr;   // UL binning: the scatter's order (and ids, keys at its slots), 3N
  std::vector<uint32_t> body_bits;  // isolated bodies: error bits of each body (sticky)
  uint2* meta = nullptr;
  double2* pos64 = nullptr;
  PairRec64* rec64 = nullptr;
  float2* pos32 = nullptr;
  PairRec32* rec32 = nullptr;
  double* part = nullptr;
  Chunk* chunks = nullptr;
  uint32_t* chunk_first = nullptr;
  SceneConst* sc = nullptr;
  SceneGeom* geom = nullptr;
  std::vector<SceneConst> h_sc;
  uint32_t* tile_off = nullptr;  // nscene+1
  uint32_t tile_cells = 0;       // 0: per-particle pair kernel
  double cell_scale = 1.0;       // cell side = 2 h_max * cell_scale (mixed UL: autotuned; else 1)
  uint32_t tile_grid = 0;
  // neighbour-list path (mixed UL RK4 steps)
  bool use_list = false;
  int list_lanes = 1;
  bool list_staged = true;  // list steps: neighbours staged in LDS (stage 1, and stages with one lane)
  uint32_t nl_cap = 0;
  uint32_t *nl_col = nullptr, *nl_deg = nullptr;
  uint16_t* nl16 = nullptr;  // staged list stages: the list as band-slot words
  float2* pos32b = nullptr;      // second record buffer (records of the next stage)
  PairRec32* rec32b = nullptr;
  double* vpart = nullptr;       // max |v|^2 per binning chunk

  // penalty contact between bodies (contact.hip)
  std::vector<std::vector<uint32_t>> bnd;  // per body: outline particle idx (body-local), in outline order
  std::vector<double> hfcx, hfcy;          // contact forces given with the bodies (idx order), or empty
  bool fc_active = false;                  // kernels read fcx / fcy (set once forces can be non-zero)
  double *fcx = nullptr, *fcy = nullptr;   // contact forces by global idx
  uint32_t* inv_gid = nullptr;             // global idx -> storage index (refreshed by the boundary update)
  std::vector<uint32_t> bnd_off;           // first outline point of each body (S + 1)
  uint32_t bnd_cap = 0;                    // outline points the device arrays hold
  bool bnd_dirty = false;                  // outlines changed since the last upload
  uint32_t* bnd_gid = nullptr;             // outline point -> global idx
  uint2* bnd_rng = nullptr;                // outline point -> (first point of its body, point count)
  double4* bnd_pts = nullptr;              // outline points: x, y, nx, ny

  TlState tl;  // Total-Lagrangian frozen lists (method TL)
  mfree::RefPipe* ref = nullptr;  // pipeline MFREE_PIPELINE_REFERENCE
  mfree::VtkState* vtk = nullptr;  // VTK writer (created by the first write)
  double* energy_buf = nullptr;   // mfree_energy partials (in allocs), then the 3 results
  size_t energy_out = 0;          // offset of the results in energy_buf
  // process-unique id: a learned stepper bound to a context compares ids, not
  // addresses (a new context may reuse a destroyed one's address)
  uint64_t serial = next_serial();
  uint64_t state_version = 0;  // advanced by mfree_set_state (h may change)
  static uint64_t next_serial() {
    static std::atomic<uint64_t> n{0};
    return ++n;
  }

  hipGraphExec_t gexec[2] = {nullptr, nullptr};

  bool prof = false;
  double pair_ms = 0., step_ms = 0.;
  uint64_t pair_n = 0, steps = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
};

namespace mfree {

// Total-Lagrangian methods (frozen neighbour lists, tl.hip) and their action list
inline bool tl_method(int32_t m) {
  return m == MFREE_METHOD_TL_RL || m == MFREE_METHOD_TL_COROTATIONAL || m == MFREE_METHOD_TL_POTENTIAL;
}
inline int tl_variant(int32_t m) {
  return m == MFREE_METHOD_TL_COROTATIONAL ? kTlCorot : m == MFREE_METHOD_TL_POTENTIAL ? kTlPotential : kTlRl;
}

inline mfree_status fail(mfree_ctx* c, mfree_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}

#define HIPCHK(ctx, call)                                                                          \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return fail(ctx, MFREE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

// Device allocation owned by the context (freed by mfree_destroy).
template <typename T>
hipError_t dalloc(mfree_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e == hipSuccess) {
    c->allocs.push_back(q);
    *p = (T*)q;
  }
  return e;
}

// Allocates and uploads everything at the first compute call.
mfree_status commit(mfree_ctx* c);
// Per-body device constants from the reference's formulas (src/simulation_data.cpp:32-47).
void scene_consts(const mfree_params& p, uint32_t offset, uint32_t n, SceneConst& s);
// Reads and clears the device error bits; maps them to a status.
mfree_status check_err_bits(mfree_ctx* c);
// Binning / UL-stage arguments over storage buffer b.
BinArgs bin_args(mfree_ctx* c, const double* x, const double* y, int b);
UlArgs ul_args(mfree_ctx* c, int b);
// nsteps RK4 steps of every body (graphs when enabled).
mfree_status run_steps(mfree_ctx* c, uint32_t nsteps);
// Idx-order gather of one storage-order field of every body: dst[gid[s]] = src[s].
void launch_to_idx_order(hipStream_t st, uint32_t n, const uint32_t* gid, const double* src, double* dst);
// Joins a pending VTK write and frees the writer.
void vtk_free(mfree_ctx* c);
// Status of a body of a context that isolates its bodies (MFREE_OK: not failed).
mfree_status body_status(const mfree_ctx* c, uint32_t body);

// One explicit RK step of every UL body, enqueued on the context's stream
// (kernels only, capturable).  Storage moves to the other buffer.
struct RkPlan {
  const SceneConst* sc = nullptr;  // constants of the step (dt); null: the context's
  int order = 4;                   // 4: RK4; 3: Kutta's third-order method
  bool rate_out = false;           // acc <- mean rate of the step
  bool advance_time = true;        // advance the bodies' time by dt
  Fields8 k1s{}, k2s{}, k3s{};     // keep the stage rates k1, k2 (and k3 of RK4), storage order of the new buffer
  // after the stage-1 binning, before the stage-1 prepass (argument: the
  // storage buffer of the state)
  std::function<void(int)> after_bin1;
  // adjusts the stage-1 pair-pass arguments
  std::function<void(UlArgs&)> stage1;
};
void enqueue_ul_rk(mfree_ctx* c, const RkPlan& p, std::vector<std::pair<hipEvent_t, hipEvent_t>>* evs);

// Penalty contact (contact.hip): device arrays at commit, and the switch
// that makes the stepping kernels read the contact forces (drops captured
// graphs, whose kernel arguments do not read them).
mfree_status contact_alloc(mfree_ctx* c);
mfree_status contact_activate(mfree_ctx* c);
void drop_graphs(mfree_ctx* c);

// Reference-structured FP64 pipeline (baseline.hip).
mfree_status ref_alloc(mfree_ctx* c);
void ref_free(mfree_ctx* c);
void ref_enqueue_step(mfree_ctx* c);
void ref_enqueue_eval(mfree_ctx* c, double* d8);

}  // namespace mfree
