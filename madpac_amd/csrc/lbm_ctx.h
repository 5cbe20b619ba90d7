// The dense solver context behind the opaque lbm_ctx handle of include/lbm_capi.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "lbm_capi.h"
#include "lbm_internal.h"

namespace lbm {
struct AdjointState;  // lbm_adjoint.hip
}

// What a captured pair of steps bakes into its launches (besides fixed buffers and parameters):
// a change of any of these re-captures the graph (lbm_step_graph).
struct lbm_graph_key {
  int bricked = -1, remap = -1, by = -1, walls = 0, force = 0, spec = 0, couple_on = 0, couple_mode = 0,
      couple_rho = 0;
  double couple_dt = 0.0;
  const void *alpha = nullptr, *bcu = nullptr, *ccells = nullptr, *cu = nullptr, *crho = nullptr, *csave = nullptr;
  int64_t ncouple = 0;
  bool operator==(const lbm_graph_key& o) const {
    return bricked == o.bricked && remap == o.remap && by == o.by && walls == o.walls && force == o.force && spec == o.spec &&
           couple_on == o.couple_on && couple_mode == o.couple_mode && couple_rho == o.couple_rho &&
           couple_dt == o.couple_dt && alpha == o.alpha && bcu == o.bcu && ccells == o.ccells && cu == o.cu &&
           crho == o.crho && csave == o.csave && ncouple == o.ncouple;
  }
};

struct lbm_ctx {
  lbm_config cfg{};
  lbm::DenseGeom g{};                  // geometry of the cell order in use (one of the two below)
  lbm::DenseGeom g_rows{}, g_bricks{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  bool fp32 = false;                   // FP32-shifted storage
  bool f32math = false;                // ... with FP32 collision arithmetic (LBM_FP32_SHIFTED_F32)
  size_t esize = 8;
  void* buf[2] = {nullptr, nullptr};  // populations (double or shifted float), ping-pong
  int cur = 0;                         // buffer holding the current stored state
  bool aa = false;                     // LBM_LAYOUT_AA: one buffer (buf[0]) updated in place
  int parity = 0;                      // AA: 0 = NATURAL layout (next step even), 1 = SWAPPED
  uint32_t* words = nullptr;           // per-cell fluid / bounce / wall-class word
  uint8_t* d_inner = nullptr;
  uint8_t* d_cls = nullptr;
  bool has_cls = false;
  void* alpha = nullptr;               // Brinkman porosity (same element type as the populations)
  bool flags_set = false, state_set = false;
  bool walls = false, force = false;
  bool spec = false;                   // step-kernel load schedule (lbm_tune)
  lbm::BoundaryClass bct[lbm::MAX_CLASSES]{};  // host copy of the boundary-class table
  lbm::BoundaryClass* d_bct = nullptr;          // device copy (read by the step kernels)
  double* d_bcu = nullptr;                      // per-cell target velocities [3*N] or null
  // MD coupling (lbm_couple.hip)
  int64_t* d_ccells = nullptr;                  // coupled cells
  int64_t ncouple = 0;
  double *d_cu = nullptr, *d_crho = nullptr, *d_cout = nullptr;
  double* d_csave = nullptr;                    // in-place layout: coupled cells' populations before a step
  bool couple_on = false, couple_rho = false;
  int couple_mode = 0;
  double couple_dt = 1.0;
  lbm::CollideParams p64{};
  lbm::CollideParams32 p32{};
  double* partials = nullptr;
  int64_t npartials = 0;
  double* small = nullptr;             // 8 doubles of reduction results
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  int64_t nfluid = 0;
  lbm::AdjointState* adj = nullptr;
  // lbm_step_graph: two steps captured per starting phase (buffer `cur`, or the AA parity)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  lbm_graph_key gkey[2];
};

double lbm_tau_effective(double tau);
int lbm_enqueue_steps(lbm_ctx* c, int nsteps);
// the geometry with the layout of the current stored state (for the reading/writing kernels)
inline lbm::DenseGeom lbm_state_geom(const lbm_ctx* c) {
  lbm::DenseGeom g = c->g;
  g.layout = !c->aa ? lbm::LAYOUT_PULL : (c->parity ? lbm::LAYOUT_AA_SWAPPED : lbm::LAYOUT_AA_NATURAL);
  return g;
}
// one step from src into dst (any two state-sized device buffers), leaving buf/cur alone
int lbm_enqueue_step_from(lbm_ctx* c, const void* src, void* dst);
// re-collision of the MD-coupled cells for the step src -> dst (no-op without coupling)
int lbm_enqueue_coupling(lbm_ctx* c, const void* src, void* dst);
// in-place layout: save the coupled cells' populations before a step, re-collide them after it
int lbm_enqueue_coupling_save(lbm_ctx* c);
int lbm_enqueue_coupling_aa(lbm_ctx* c, int parity);
void lbm_coupling_release(lbm_ctx* c);
void lbm_graphs_release(lbm_ctx* c);
void lbm_adjoint_release(lbm_ctx* c);
// once per device, with the device current: the device address of the fixed array that
// lbm_reduce_moments_noalloc uses (lbm_noalloc.hip)
hipError_t lbm_noalloc_setup(int device);
