// Adjoint (reverse-mode) collide+stream and revolve checkpointing for dJ/d(porosity).
//
// No reference counterpart: the reference has no adjoint.  The forward step being reversed is
// k_dense_step (pull form): f_k = P(s_k) gathers the reference state from the stored state,
// s_{k+1} = C(f_k) collides per cell (BGK, Guo forcing, Brinkman scaling of the equilibrium
// velocity by (1 - alpha), moving-wall term).  The reverse sweep keeps lambda_k = dJ/df_k (the
// adjoint of the reference state) and runs, per cell,
//     mu_q      = dJ/ds_{k+1}(q, x) = lambda_{k+1,q}(x + c_q)  (or lambda_{k+1,18-q}(x) when
//                 x + c_q is not fluid)                       -- the transpose of the pull, a gather
//     lambda_k  = (dC/df)^T mu ,   dJ/dalpha(x) += (dC/dalpha)^T mu
// so one adjoint step reads the stored state (19 gathers, to rebuild rho and u), lambda_{k+1}
// (19 gathers) and writes lambda_k (19 aligned stores): 57 values per cell, plus alpha and the
// gradient.  The per-cell algebra is the same as the oracle's CPU tape (lbo_step_adjoint).
//
// Checkpointing: the states s_k needed in reverse order are recomputed from checkpoints placed by
// an optimal binomial ("revolve") schedule, found by dynamic programming over
//     F(l, c) = min_m  m + F(l - m, c - 1) + F(m, c),   F(1, c) = 0,   F(l, 0) = l(l-1)/2,
// the fewest forward steps that reverse l steps with c free checkpoint slots (Griewank & Walther,
// ACM TOMS 26(1), 2000).
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_geom.h"

namespace lbm {

struct AdjointState {
  std::vector<void*> slots;  // checkpoint slots (stored state), slot 0 = initial state
  void* final_state = nullptr;
  void* lam[2] = {nullptr, nullptr};
  double* galpha = nullptr;
  int64_t N = 0;
  size_t esize = 0;
};

// ---------------------------------------------------------------------------------------------
// dJ/df of J = cs^2 (<rho>_xin - <rho>_xout): uniform per population on the two planes.  The adjoint
// states lambda use the state's arrangement of populations (planes or interleaved), so one set of
// addresses serves the state and the adjoint in k_adj_step.
template <typename L, bool AOS>
__global__ __launch_bounds__(BX * BY_MAX) void k_adj_terminal(DenseGeom g, const uint32_t* __restrict__ words, int xin,
                                                          int xout, double vin, double vout, L* __restrict__ lam) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const bool fl = words[c.i] & FLUID;
  double v = 0.0;
  if (fl && c.x == xin) v += vin;
  if (fl && c.x == xout) v += vout;
  const int64_t b = AOS ? pop_at<true>(g.N, 0, c.i) : c.i;
#pragma unroll
  for (int q = 0; q < Q; q++) lam[AOS ? b + 64 * q : (int64_t)q * g.N + b] = (L)v;
}

// The adjoint step.  T: stored-state type; L: adjoint type (double or float); gradient in double.
template <typename T, bool FORCE, bool ALPHA, bool WALLS, bool AOS>
__global__ __launch_bounds__(BX * BY_MAX) void k_adj_step(DenseGeom g, const T* __restrict__ s,
                                                      const T* __restrict__ lam_out, T* __restrict__ lam_in,
                                                      const uint32_t* __restrict__ words, const T* __restrict__ alpha,
                                                      double* __restrict__ galpha, CollideParams p,
                                                      const BoundaryClass* __restrict__ bct) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const uint32_t w = words[c.i];
  const CellAddr<AOS> Ad = cell_addr<AOS>(g, c);
  if (!(w & FLUID)) {
#pragma unroll
    for (int q = 0; q < Q; q++) lam_in[Ad.own(q)] = T(0);
    return;
  }
  T f[Q], mu[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    f[q] = s[((w >> (Q - 1 - q)) & 1u) ? Ad.own(Q - 1 - q) : Ad.pull(q)];
    const bool bounce = (w >> q) & 1u;  // x + c_q is not fluid
    mu[q] = bounce ? lam_out[Ad.own(Q - 1 - q)] : lam_out[Ad.push(q)];
  }
  constexpr bool D = sizeof(T) == 8;
  using R = T;  // arithmetic type
  const R one = R(1), half = R(0.5), three = R(3), nine = R(9), four5 = R(4.5), one5 = R(1.5);
  R drho_wall = R(0);
  if (WALLS) {
    // moving walls only (the host rejects other boundary types)
    const uint32_t k = (w >> CLASS_SHIFT) & CLASS_BITS;
    if (k) {
      const uint32_t m = w & bct[k].mask;
      const double uw[3] = {bct[k].u[0], bct[k].u[1], bct[k].u[2]};
#pragma unroll
      for (int q = 0; q < Q; q++)
        if ((m >> q) & 1u) drho_wall -= mu[q] * (R)moving_wall_f64(q, uw, 1.0);
    }
  }
  // rho, j (the FP32 state holds f - w; the rest part adds to rho only)
  R rho = R(0), j0 = R(0), j1 = R(0), j2 = R(0);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    rho += f[q];
    j0 += (R)CX(q) * f[q];
    j1 += (R)CY(q) * f[q];
    j2 += (R)CZ(q) * f[q];
  }
  if (!D) rho += one;
  const R om = (R)p.omega;
  const R g0 = FORCE ? (R)p.g[0] : R(0), g1 = FORCE ? (R)p.g[1] : R(0), g2 = FORCE ? (R)p.g[2] : R(0);
  const R a = ALPHA ? alpha[c.i] : R(0);
  const R sc = ALPHA ? (one - a) : one;
  const R u0 = j0 / rho, u1 = j1 / rho, u2 = j2 / rho;
  const R v0 = sc * (u0 + half * g0), v1 = sc * (u1 + half * g1), v2 = sc * (u2 + half * g2);
  const R vv = v0 * v0 + v1 * v1 + v2 * v2;
  const R fpref = FORCE ? (one - half * om) : R(0);
  // Moment form.  With M0 = sum mu w, M1 = sum mu w c, M2 = sum mu w c c^T (w_q the lattice
  // weights), the per-population sums of the tape collapse to
  //   dPhi/drho = om [M0 (1 - 1.5 v.v) + 3 v.M1 + 4.5 v^T M2 v] + fpref [3 (M1.g - M0 v.g) + 9 v^T M2 g]
  //   dPhi/dv   = rho { om [3 M1 + 9 M2 v - 3 M0 v] + fpref [9 M2 g - 3 M0 g] }
  // and lambda_q = (1 - om) mu_q + A + B.c_q is affine in c_q.
  R M0 = R(0), m1x = R(0), m1y = R(0), m1z = R(0);
  R mxx = R(0), myy = R(0), mzz = R(0), mxy = R(0), mxz = R(0), myz = R(0);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const R wm = (R)W64(q) * mu[q];
    M0 += wm;
    if (CX(q) == 1) m1x += wm;
    if (CX(q) == -1) m1x -= wm;
    if (CY(q) == 1) m1y += wm;
    if (CY(q) == -1) m1y -= wm;
    if (CZ(q) == 1) m1z += wm;
    if (CZ(q) == -1) m1z -= wm;
    if (CX(q) != 0) mxx += wm;
    if (CY(q) != 0) myy += wm;
    if (CZ(q) != 0) mzz += wm;
    if (CX(q) * CY(q) == 1) mxy += wm;
    if (CX(q) * CY(q) == -1) mxy -= wm;
    if (CX(q) * CZ(q) == 1) mxz += wm;
    if (CX(q) * CZ(q) == -1) mxz -= wm;
    if (CY(q) * CZ(q) == 1) myz += wm;
    if (CY(q) * CZ(q) == -1) myz -= wm;
  }
  // M2 v and M2 g
  const R m2v0 = mxx * v0 + mxy * v1 + mxz * v2;
  const R m2v1 = mxy * v0 + myy * v1 + myz * v2;
  const R m2v2 = mxz * v0 + myz * v1 + mzz * v2;
  const R m2g0 = mxx * g0 + mxy * g1 + mxz * g2;
  const R m2g1 = mxy * g0 + myy * g1 + myz * g2;
  const R m2g2 = mxz * g0 + myz * g1 + mzz * g2;
  const R vm1 = v0 * m1x + v1 * m1y + v2 * m1z;
  const R vm2v = v0 * m2v0 + v1 * m2v1 + v2 * m2v2;
  const R gm1 = g0 * m1x + g1 * m1y + g2 * m1z;
  const R vg = v0 * g0 + v1 * g1 + v2 * g2;
  const R vm2g = v0 * m2g0 + v1 * m2g1 + v2 * m2g2;
  const R dPr = om * (M0 * (one - one5 * vv) + three * vm1 + four5 * vm2v) +
                fpref * (three * (gm1 - M0 * vg) + nine * vm2g);
  const R dv0 = rho * (om * (three * m1x + nine * m2v0 - three * M0 * v0) + fpref * (nine * m2g0 - three * M0 * g0));
  const R dv1 = rho * (om * (three * m1y + nine * m2v1 - three * M0 * v1) + fpref * (nine * m2g1 - three * M0 * g1));
  const R dv2 = rho * (om * (three * m1z + nine * m2v2 - three * M0 * v2) + fpref * (nine * m2g2 - three * M0 * g2));
  const R sr = sc / rho;
  const R A = dPr + drho_wall - sr * (dv0 * u0 + dv1 * u1 + dv2 * u2);
  const R B0 = sr * dv0, B1 = sr * dv1, B2 = sr * dv2;
  const R omc = one - om;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    R l = omc * mu[q] + A;
    if (CX(q) == 1) l += B0;
    if (CX(q) == -1) l -= B0;
    if (CY(q) == 1) l += B1;
    if (CY(q) == -1) l -= B1;
    if (CZ(q) == 1) l += B2;
    if (CZ(q) == -1) l -= B2;
    lam_in[Ad.own(q)] = l;
  }
  if (ALPHA) {
    R ga = R(0);
    ga -= dv0 * (u0 + half * g0);
    ga -= dv1 * (u1 + half * g1);
    ga -= dv2 * (u2 + half * g2);
    galpha[c.i] += (double)ga;
  }
}

template <typename T, bool AOS>
static hipError_t adj_step_t(const DenseGeom& g, const void* s, const void* lo, void* li, const uint32_t* words,
                             const void* alpha, double* galpha, bool force, bool walls, const CollideParams& p,
                             const BoundaryClass* bct, hipStream_t st) {
  const dim3 grid((unsigned)g.nblocks_padded), block(BX, g.by);
  const T* sp = (const T*)s;
  const T* lop = (const T*)lo;
  T* lip = (T*)li;
  const T* ap = (const T*)alpha;
#define LBM_ADJ(F, A, W) \
  hipLaunchKernelGGL((k_adj_step<T, F, A, W, AOS>), grid, block, 0, st, g, sp, lop, lip, words, ap, galpha, p, bct)
  const int key = (force ? 4 : 0) | (alpha ? 2 : 0) | (walls ? 1 : 0);
  switch (key) {
    case 0: LBM_ADJ(false, false, false); break;
    case 1: LBM_ADJ(false, false, true); break;
    case 2: LBM_ADJ(false, true, false); break;
    case 3: LBM_ADJ(false, true, true); break;
    case 4: LBM_ADJ(true, false, false); break;
    case 5: LBM_ADJ(true, false, true); break;
    case 6: LBM_ADJ(true, true, false); break;
    default: LBM_ADJ(true, true, true); break;
  }
#undef LBM_ADJ
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Optimal binomial checkpoint schedule (dynamic programming, memoised).
struct Revolve {
  int cmax;
  std::vector<std::vector<int64_t>> F;  // F[l][c]
  std::vector<std::vector<int>> M;      // best first split
  Revolve(int lmax, int c) : cmax(c), F(lmax + 1, std::vector<int64_t>(c + 1, 0)), M(lmax + 1, std::vector<int>(c + 1, 0)) {
    for (int l = 1; l <= lmax; l++) {
      F[l][0] = (int64_t)l * (l - 1) / 2;
      for (int cc = 1; cc <= c; cc++) {
        if (l == 1) {
          F[l][cc] = 0;
          continue;
        }
        int64_t best = INT64_MAX;
        int bm = 1;
        for (int m = 1; m < l; m++) {
          const int64_t v = m + F[l - m][cc - 1] + F[m][cc];
          if (v < best) {
            best = v;
            bm = m;
          }
        }
        F[l][cc] = best;
        M[l][cc] = bm;
      }
    }
  }
};

}  // namespace lbm

using namespace lbm;

namespace {
// Device time of the reverse sweep's two phases: a pair of events brackets each recompute segment
// and each adjoint step on the context's stream; after the sweep the pairs are summed per phase.
// Events are created as the sweep needs them (a few per step) and destroyed with the timer.
struct PhaseTimer {
  enum Phase { RECOMPUTE = 0, ADJOINT = 1 };
  std::vector<hipEvent_t> ev;
  std::vector<int> phase;
  size_t used = 0;
  hipError_t init(size_t pairs) {
    ev.reserve(2 * pairs);
    phase.reserve(pairs);
    return hipSuccess;
  }
  hipError_t begin(hipStream_t s, Phase p) {
    while (ev.size() < 2 * used + 2) {
      hipEvent_t e = nullptr;
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      ev.push_back(e);
    }
    if (phase.size() <= used) phase.resize(used + 1);
    phase[used] = p;
    return hipEventRecord(ev[2 * used], s);
  }
  hipError_t end(hipStream_t s) { return hipEventRecord(ev[2 * used++ + 1], s); }
  hipError_t sum(float out[2]) {
    for (size_t k = 0; k < used; k++) {
      float ms = 0.f;
      const hipError_t r = hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]);
      if (r != hipSuccess) return r;
      out[phase[k]] += ms;
    }
    return hipSuccess;
  }
  ~PhaseTimer() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};
}  // namespace

void lbm_adjoint_release(lbm_ctx* c) {
  if (!c || !c->adj) return;
  AdjointState* a = c->adj;
  for (void* p : a->slots)
    if (p) (void)hipFree(p);
  if (a->final_state) (void)hipFree(a->final_state);
  if (a->lam[0]) (void)hipFree(a->lam[0]);
  if (a->lam[1]) (void)hipFree(a->lam[1]);
  if (a->galpha) (void)hipFree(a->galpha);
  delete a;
  c->adj = nullptr;
}

namespace {
int adj_fail(lbm_ctx* c, int code, const std::string& m) {
  c->err = m;
  return code;
}
}  // namespace

static int adjoint_two_buffer(lbm_ctx* c, int nsteps, int snapshots, int xin, int xout, double* J,
                              double* grad_alpha, lbm_adjoint_stats* stats);

static int adjoint_args(lbm_ctx* c, int nsteps, int snapshots, int xin, int xout, const double* J,
                        const double* grad_alpha) {
  if (!J || !grad_alpha) return adj_fail(c, LBM_ERR_ARG, "output is null");
  // the schedule search is O(nsteps^2 * snapshots) on the host
  if (nsteps < 1 || nsteps > 2048) return adj_fail(c, LBM_ERR_ARG, "nsteps must be in 1..2048");
  if (snapshots < 0 || snapshots > 64) return adj_fail(c, LBM_ERR_ARG, "snapshots must be in 0..64");
  if (xin < 0 || xin >= c->g.nx || xout < 0 || xout >= c->g.nx) return adj_fail(c, LBM_ERR_ARG, "plane out of range");
  if (!c->flags_set || !c->state_set) return adj_fail(c, LBM_ERR_STATE, "flags and state must be set first");
  if (c->couple_on) return adj_fail(c, LBM_ERR_STATE, "the adjoint does not cover MD-coupled cells");
  for (int k = 1; k < MAX_CLASSES; k++)
    if (c->bct[k].mask && (c->bct[k].type & BC_TYPE_MASK) != BC_MOVING_WALL)
      return adj_fail(c, LBM_ERR_STATE, "the adjoint supports halfway bounce-back and moving walls only");
  return LBM_OK;
}

// Re-store the current state in another layout of the same cell order, value by value into a second
// buffer (no staging memory, so it cannot fail for lack of memory once both buffers exist).
static int restore_layout(lbm_ctx* c, const lbm::DenseGeom& from, const void* src, const lbm::DenseGeom& to,
                          void* dst) {
  hipError_t e = lbm::launch_state_relayout(from, to, c->fp32, src, dst, c->words, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? LBM_OK : adj_fail(c, LBM_ERR_HIP, std::string("layout conversion: ") + hipGetErrorString(e));
}

// On an in-place (AA) context the adjoint borrows a second state buffer for its duration: the state
// is re-stored in the two-buffer pull form, the adjoint runs as on a two-buffer context, and the
// final state goes back to the in-place layout (NATURAL, parity 0).
extern "C" int lbm_adjoint_pressure_drop(lbm_ctx* c, int nsteps, int snapshots, int xin, int xout, double* J,
                                         double* grad_alpha, lbm_adjoint_stats* stats) {
  if (!c) return LBM_ERR_ARG;
  int rc = adjoint_args(c, nsteps, snapshots, xin, xout, J, grad_alpha);
  if (rc) return rc;
  if (!c->aa) return adjoint_two_buffer(c, nsteps, snapshots, xin, xout, J, grad_alpha, stats);
  if (hipSetDevice(c->cfg.device) != hipSuccess) return adj_fail(c, LBM_ERR_HIP, "hipSetDevice failed");
  const size_t pbytes = (size_t)Q * c->g.N * c->esize;
  if (lbm::device_alloc(&c->buf[1], pbytes) != hipSuccess) {
    c->buf[1] = nullptr;
    return adj_fail(c, LBM_ERR_NOMEM, "the adjoint on an in-place context needs a second state buffer");
  }
  rc = restore_layout(c, lbm_state_geom(c), c->buf[0], [&] {
    lbm::DenseGeom g = c->g;
    g.layout = lbm::LAYOUT_PULL;
    return g;
  }(), c->buf[1]);
  if (rc == LBM_OK) {
    c->aa = false;
    c->cur = 1;
    rc = adjoint_two_buffer(c, nsteps, snapshots, xin, xout, J, grad_alpha, stats);
    // back in place, whatever the adjoint returned (the state is the final one on success)
    // The in-place layout is chosen for capacity: its adjoint workspace (checkpoints, adjoint states)
    // is not kept between calls.
    lbm_adjoint_release(c);
    lbm::DenseGeom natural = c->g;
    natural.layout = lbm::LAYOUT_AA_NATURAL;
    // the relayout needs distinct buffers: a final state left in buf[0] moves to buf[1] first
    int rc2 = LBM_OK;
    if (c->cur == 0) {
      hipError_t e = hipMemcpyAsync(c->buf[1], c->buf[0], pbytes, hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess) rc2 = adj_fail(c, LBM_ERR_HIP, std::string("final state copy: ") + hipGetErrorString(e));
      c->cur = 1;
    }
    if (!rc2) rc2 = restore_layout(c, lbm_state_geom(c), c->buf[1], natural, c->buf[0]);
    c->aa = true;
    c->cur = 0;
    c->parity = 0;
    if (rc2) c->state_set = false;  // buf[0] holds no valid state: it must be set again
    if (!rc) rc = rc2;
  }
  (void)hipFree(c->buf[1]);
  c->buf[1] = nullptr;
  return rc;
}

static int adjoint_two_buffer(lbm_ctx* c, int nsteps, int snapshots, int xin, int xout, double* J,
                              double* grad_alpha, lbm_adjoint_stats* stats) {
  if (hipSetDevice(c->cfg.device) != hipSuccess) return adj_fail(c, LBM_ERR_HIP, "hipSetDevice failed");

  const int64_t N = c->g.N;
  const size_t sbytes = (size_t)Q * N * c->esize;
  const int nslots = std::min(snapshots, nsteps) + 1;
  // (re)allocate workspace
  if (c->adj && ((int)c->adj->slots.size() != nslots || c->adj->N != N || c->adj->esize != c->esize))
    lbm_adjoint_release(c);
  if (!c->adj) {
    AdjointState* a = new AdjointState();
    a->N = N;
    a->esize = c->esize;
    c->adj = a;
    a->slots.assign(nslots, nullptr);
    bool ok = true;
    for (int k = 0; k < nslots && ok; k++) ok = lbm::device_alloc(&a->slots[k], sbytes) == hipSuccess;
    ok = ok && lbm::device_alloc(&a->final_state, sbytes) == hipSuccess;
    ok = ok && lbm::device_alloc(&a->lam[0], sbytes) == hipSuccess && lbm::device_alloc(&a->lam[1], sbytes) == hipSuccess;
    ok = ok && lbm::device_alloc(&a->galpha, (size_t)N * sizeof(double)) == hipSuccess;
    if (!ok) {
      lbm_adjoint_release(c);
      return adj_fail(c, LBM_ERR_NOMEM, "adjoint workspace allocation failed");
    }
  }
  AdjointState* a = c->adj;
  hipStream_t st = c->stream;
#define ACHK(call)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return adj_fail(c, LBM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

  Revolve rv(nsteps, nslots - 1);
  // First forward sweep from s_0, storing the checkpoints of the schedule's first descent.
  std::map<int, int> held;  // time step -> slot holding s_t
  std::vector<bool> busy(nslots, false);
  ACHK(hipMemcpyAsync(a->slots[0], c->buf[c->cur], sbytes, hipMemcpyDeviceToDevice, st));
  held[0] = 0;
  busy[0] = true;
  {
    int t = 0, l = nsteps, cc = nslots - 1;
    std::vector<int> descent;
    while (l > 1 && cc > 0) {
      const int m = rv.M[l][cc];
      t += m;
      l -= m;
      cc--;
      descent.push_back(t);
    }
    ACHK(hipEventRecord(c->ev0, st));
    int now = 0;
    for (int pos : descent) {
      int rc = lbm_enqueue_steps(c, pos - now);
      if (rc) return rc;
      now = pos;
      int slot = -1;
      for (int k = 1; k < nslots; k++)
        if (!busy[k]) {
          slot = k;
          break;
        }
      ACHK(hipMemcpyAsync(a->slots[slot], c->buf[c->cur], sbytes, hipMemcpyDeviceToDevice, st));
      busy[slot] = true;
      held[pos] = slot;
    }
    int rc = lbm_enqueue_steps(c, nsteps - now);
    if (rc) return rc;
    ACHK(hipEventRecord(c->ev1, st));
    ACHK(hipMemcpyAsync(a->final_state, c->buf[c->cur], sbytes, hipMemcpyDeviceToDevice, st));
  }
  float fwd_ms = 0.f;
  ACHK(hipEventSynchronize(c->ev1));
  ACHK(hipEventElapsedTime(&fwd_ms, c->ev0, c->ev1));

  // objective and its gradient w.r.t. the final reference state
  double Jv = 0.0;
  {
    int rc = lbm_pressure_drop(c, xin, xout, &Jv);
    if (rc) return rc;
  }
  double hsum[4];
  ACHK(launch_dense_plane_rho(c->g, c->fp32, c->buf[c->cur], c->words, xin, xout, c->partials, c->small, st));
  ACHK(hipMemcpyAsync(hsum, c->small, sizeof(hsum), hipMemcpyDeviceToHost, st));
  ACHK(hipStreamSynchronize(st));
  const double vin = (1.0 / 3.0) / hsum[1], vout = -(1.0 / 3.0) / hsum[3];
  int lc = 0;
  const dim3 grid((unsigned)c->g.nblocks_padded), block(BX, c->g.by);
  if (c->fp32) {
    if (c->g.aos)
      hipLaunchKernelGGL((k_adj_terminal<float, true>), grid, block, 0, st, c->g, c->words, xin, xout, vin, vout,
                         (float*)a->lam[lc]);
    else
      hipLaunchKernelGGL((k_adj_terminal<float, false>), grid, block, 0, st, c->g, c->words, xin, xout, vin, vout,
                         (float*)a->lam[lc]);
  } else {
    if (c->g.aos)
      hipLaunchKernelGGL((k_adj_terminal<double, true>), grid, block, 0, st, c->g, c->words, xin, xout, vin, vout,
                         (double*)a->lam[lc]);
    else
      hipLaunchKernelGGL((k_adj_terminal<double, false>), grid, block, 0, st, c->g, c->words, xin, xout, vin, vout,
                         (double*)a->lam[lc]);
  }
  ACHK(hipGetLastError());
  ACHK(hipMemsetAsync(a->galpha, 0, (size_t)N * sizeof(double), st));

  // reverse sweep.  Each recompute segment and each adjoint step is bracketed by a pair of events
  // from a pool, so the two phases are timed separately on the device (summed after the sweep).
  int64_t recompute = 0, reversed = 0;
  int rc_err = 0;
  PhaseTimer timer;
  ACHK(timer.init(4 * (size_t)nsteps + 8));
  ACHK(hipEventRecord(c->ev0, st));
  // Checkpoints are read where they lie: the state s_{t0+j} is the slot itself for j = 0, else the
  // first recomputed step reads the slot and writes the context's ping-pong buffers.
  auto state_from = [&](int t0, int j, const void** out) -> int {
    auto it = held.find(t0);
    if (it == held.end()) return adj_fail(c, LBM_ERR_STATE, "checkpoint schedule lost a state");
    const void* s0 = a->slots[it->second];
    if (j == 0) {
      *out = s0;
      return 0;
    }
    hipError_t e = timer.begin(st, PhaseTimer::RECOMPUTE);
    if (e != hipSuccess) return adj_fail(c, LBM_ERR_HIP, std::string("recompute timing: ") + hipGetErrorString(e));
    int r = lbm_enqueue_step_from(c, s0, c->buf[c->cur]);
    if (!r) r = lbm_enqueue_steps(c, j - 1);
    e = timer.end(st);
    if (!r && e != hipSuccess) return adj_fail(c, LBM_ERR_HIP, std::string("recompute timing: ") + hipGetErrorString(e));
    *out = c->buf[c->cur];
    return r;
  };
  auto adj_step = [&](const void* sp) -> int {
    hipError_t e = timer.begin(st, PhaseTimer::ADJOINT);
    if (e == hipSuccess)
      e = c->fp32 ? (c->g.aos ? adj_step_t<float, true>(c->g, sp, a->lam[lc], a->lam[lc ^ 1], c->words, c->alpha,
                                                         a->galpha, c->force, c->walls, c->p64, c->d_bct, st)
                               : adj_step_t<float, false>(c->g, sp, a->lam[lc], a->lam[lc ^ 1], c->words, c->alpha,
                                                          a->galpha, c->force, c->walls, c->p64, c->d_bct, st))
                  : (c->g.aos ? adj_step_t<double, true>(c->g, sp, a->lam[lc], a->lam[lc ^ 1], c->words, c->alpha,
                                                          a->galpha, c->force, c->walls, c->p64, c->d_bct, st)
                               : adj_step_t<double, false>(c->g, sp, a->lam[lc], a->lam[lc ^ 1], c->words, c->alpha,
                                                           a->galpha, c->force, c->walls, c->p64, c->d_bct, st));
    if (e == hipSuccess) e = timer.end(st);
    if (e != hipSuccess) return adj_fail(c, LBM_ERR_HIP, std::string("adjoint step: ") + hipGetErrorString(e));
    lc ^= 1;
    reversed++;
    return 0;
  };
  // reverse steps t0 .. t0+l-1 with s_t0 held and cc free slots
  std::function<int(int, int, int)> rev = [&](int t0, int l, int cc) -> int {
    if (rc_err) return rc_err;
    const void* sp = nullptr;
    if (l == 1) {
      int r = state_from(t0, 0, &sp);
      return r ? r : adj_step(sp);
    }
    if (cc == 0) {
      for (int j = l - 1; j >= 0; j--) {
        int r = state_from(t0, j, &sp);
        if (r) return r;
        recompute += j;
        r = adj_step(sp);
        if (r) return r;
      }
      return 0;
    }
    const int m = rv.M[l][cc];
    const int tm = t0 + m;
    bool placed_here = false;
    if (!held.count(tm)) {
      int slot = -1;
      for (int k = 1; k < nslots; k++)
        if (!busy[k]) {
          slot = k;
          break;
        }
      if (slot < 0) return adj_fail(c, LBM_ERR_STATE, "no free checkpoint slot");
      // s_tm = m steps from the checkpoint at t0, the last one written straight into the new slot
      auto it = held.find(t0);
      if (it == held.end()) return adj_fail(c, LBM_ERR_STATE, "checkpoint schedule lost a state");
      int r;
      if (m == 1) {
        hipError_t e = timer.begin(st, PhaseTimer::RECOMPUTE);
        r = e == hipSuccess ? lbm_enqueue_step_from(c, a->slots[it->second], a->slots[slot])
                            : adj_fail(c, LBM_ERR_HIP, "recompute timing failed");
        if (!r && timer.end(st) != hipSuccess) r = adj_fail(c, LBM_ERR_HIP, "recompute timing failed");
      } else {
        r = state_from(t0, m - 1, &sp);
        if (!r) {
          hipError_t e = timer.begin(st, PhaseTimer::RECOMPUTE);
          r = e == hipSuccess ? lbm_enqueue_step_from(c, sp, a->slots[slot])
                              : adj_fail(c, LBM_ERR_HIP, "recompute timing failed");
          if (!r && timer.end(st) != hipSuccess) r = adj_fail(c, LBM_ERR_HIP, "recompute timing failed");
        }
      }
      if (r) return r;
      recompute += m;
      busy[slot] = true;
      held[tm] = slot;
      placed_here = true;
    }
    (void)placed_here;
    int r = rev(tm, l - m, cc - 1);
    if (r) return r;
    busy[held[tm]] = false;
    held.erase(tm);
    return rev(t0, m, cc);
  };
  rc_err = rev(0, nsteps, nslots - 1);
  if (rc_err) return rc_err;
  ACHK(hipEventRecord(c->ev1, st));
  float rev_ms = 0.f;
  ACHK(hipEventSynchronize(c->ev1));
  ACHK(hipEventElapsedTime(&rev_ms, c->ev0, c->ev1));
  float phase_ms[2] = {0.f, 0.f};
  ACHK(timer.sum(phase_ms));

  std::vector<double> gstored((size_t)N);
  ACHK(hipMemcpyAsync(gstored.data(), a->galpha, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, st));
  // leave the context at the state after nsteps
  ACHK(hipMemcpyAsync(c->buf[c->cur], a->final_state, sbytes, hipMemcpyDeviceToDevice, st));
  ACHK(hipStreamSynchronize(st));
  // stored (bricked) order -> the caller's row-major order
  const DenseGeom& g = c->g;
  for (int z = 0; z < g.nz; z++)
    for (int y = 0; y < g.ny; y++)
      for (int x = 0; x < g.nx; x++)
        grad_alpha[(int64_t)x + (int64_t)g.nx * ((int64_t)y + (int64_t)g.ny * z)] = gstored[(size_t)stored_index(g, x, y, z)];
  *J = Jv;
  if (stats) {
    stats->snapshots_used = nslots - 1;
    stats->forward_steps = nsteps;
    stats->recompute_steps = recompute;
    stats->reverse_steps = reversed;
    stats->forward_ms = fwd_ms;
    stats->reverse_ms = rev_ms;
    stats->recompute_ms = phase_ms[PhaseTimer::RECOMPUTE];
    stats->adjoint_ms = phase_ms[PhaseTimer::ADJOINT];
  }
  for (int64_t i = 0; i < g.Nu; i++)
    if (!std::isfinite(grad_alpha[i])) return adj_fail(c, LBM_ERR_NUMERIC, "non-finite gradient");
#undef ACHK
  return LBM_OK;
}
