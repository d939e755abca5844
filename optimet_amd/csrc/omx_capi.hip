// omx_capi.hip -- the C ABI (include/optimet_mi355x.h): context, device
// buffers, the update()/solve() sequence of solver::PreconditionedMatrix
// (PreconditionedMatrixSolver.h:45-100) and the iterative solvers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <functional>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/optimet_mi355x.h"
#include "omx_kernels.hpp"
#include "omx_sh.hpp"

using omx::cplx;

namespace {

thread_local std::string g_err;

struct OmxError : std::runtime_error {
  int code;
  OmxError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if(e_ != hipSuccess)                                                                            \
      throw OmxError(e_ == hipErrorOutOfMemory ? OMX_ERR_MEMORY : OMX_ERR_DEVICE,                   \
                     std::string(#x) + ": " + hipGetErrorString(e_));                               \
  } while(0)

// Device buffer owned by the context.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b, size_t *total) {
    if(b <= bytes && p) return;
    release(total);
    if(b == 0) return;
    HIPCHK(hipMalloc(&p, b));
    bytes = b;
    *total += b;
  }
  void release(size_t *total) {
    if(p) {
      (void)hipFree(p);
      *total -= bytes;
    }
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Normwise backward error accepted from a refinement that stopped above its
// residual band: 1000 units of FP64 roundoff.
const double kBackwardErrorBound = 1e3 * 1.1e-16;

// The inner solve of OMX_SOLVER_GMRES32_IR: one GMRES cycle of at most
// kInnerSteps steps on the complex-FP32 copy, stopped at kInnerTol relative
// residual (the FP32 matrix holds S to about 6e-8, so the inner solve can
// gain about six digits per outer step).
constexpr int kInnerSteps = 30;
constexpr double kInnerTol = 1e-6;

// OMX_SOLVER_AUTO: systems of order N >= kAutoKrylovN start with the mixed
// Krylov, within a budget of outer steps whose matvecs cost what the mixed LU
// would (8/3 N^3 flop at about 140 TFLOP/s against an 8 N^2-byte product at
// about 6 TB/s: 0.0143 N products, kInnerSteps + 1 of them per outer step),
// and fall back to the mixed LU + refinement on the same matrices when it
// does not converge; smaller systems go to the LU directly.
constexpr int kAutoKrylovN = 16384;
inline int auto_krylov_steps(int N) { return std::max(4, (int)(0.0143 * N / (kInnerSteps + 1))); }
// Below kAutoKrylovN, with the second harmonic: V's mixed LU starts at once on
// the second stream pair (V does not depend on the fundamental solution), and
// S is first tried with FP64 GMRES beside it, so that S's solve is hidden
// while it takes less than V's factorisation.  Its budget is the products
// that cost what an LU would (8/3 N^3 flop at about 125 TFLOP/s against a
// 16 N^2-byte FP64 product at about 6 TB/s: N / 125); S falls back to its own
// mixed LU when GMRES has not converged by then.
inline int auto_gmres64_steps(int N) { return std::max(20, N / 125); }
// Batched sweeps take the same rule (S by lockstep FP64 GMRES beside V's
// batched factorisation, within twice that budget) from this order on; below
// it the budget (2 N / 125 < 33 products) is under what such systems need.
constexpr int kSweepKrylovN = 2048;

const double kMu0 = 4.0 * omx::kPi * 1e-7;
const double kEps0 = 1.0 / (kMu0 * omx::kC0 * omx::kC0);

}  // namespace

struct omx_context {
  int device = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev[8];
  size_t total = 0;
  // geometry
  int P = 0, nmax = 0, nf = 0, N = 0, Np = 0;
  std::vector<double> xyz, radius;
  std::vector<cplx> eps, mu, eps_sh, ksippp, ksiparppar, gamma;  // absolute eps, mu; relative ones kept below
  cplx bg_eps_r = omx::mk(1.0), bg_mu_r = omx::mk(1.0);
  bool have_geom = false, have_exc = false;
  // excitation
  double lambda = 0, theta = 0, phi = 0;
  cplx eth = omx::mk(1.0), eph = omx::mk(0.0);
  int sh = 0;
  // solver
  int kind = OMX_SOLVER_LU64;
  int kind_req = OMX_SOLVER_LU64;  // as set; OMX_SOLVER_AUTO resolves into kind at each update
  int auto_budget = 0;             // outer Krylov steps of OMX_SOLVER_AUTO (0: the cost model)
  bool krylov_s = false;           // OMX_SOLVER_AUTO: S tried with FP64 GMRES beside V's factorisation
  bool autotune = true;            // sweeps and solves pick their bulk product and triangular solves by timing
  bool tuned = false, tuned_split = true;  // the single-system choice of the last update
  int pivot = OMX_PIVOT_BLOCK;
  // fraction of S's columns factored before V's factorisation starts
  // (OMX_SH_OVERLAP overrides; 0 starts both together)
  double sh_overlap = 0.5;
  // OMX_SOLVER_AUTO with S by GMRES: fraction of V's columns factored before
  // S's GMRES starts (OMX_KRYLOV_MARK overrides; 0 starts both together)
  double krylov_mark = 0.5;
  int lu32_gemm = OMX_GEMM_BF16X3;
  int fallbacks = 0;
  double tol = 1e-16;
  int maxit = 300, restart = 40;
  // state
  bool updated = false, solved = false;
  bool factored = false, factored_sh = false;
  bool dirty32 = false, dirty32_sh = false;  // the FP32 matrix holds factors, not S (V)
  int cg_nmax = 0;
  // device buffers
  DBuf d_xyz, d_radius, d_eps, d_mu, d_eps_sh, d_kppp, d_kpar, d_gam, d_sph, d_ov, d_inc, d_incloc, d_Q, d_w,
      d_coef, d_S64, d_S32, d_V64, d_V32, d_x, d_xsh, d_xint, d_xintsh, d_K, d_K1, d_cross, d_cg, d_abscoef, d_abssh,
      d_absbes,
      d_fpts, d_fout;
  DBuf d_piv, d_perm, d_cand, d_cnt, d_map, d_status, d_wlu;
  hipStream_t aux = nullptr;
  hipEvent_t lu_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // the second-harmonic system's own streams, events, status and maps, so
  // its factorisation can run beside the fundamental solve
  hipStream_t st2 = nullptr, aux2 = nullptr;
  hipEvent_t lu_ev2[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t xev[4] = {nullptr, nullptr, nullptr, nullptr};
  DBuf d_status2, d_map2;
  std::vector<hipEvent_t> pev2;
  omx::LuProf prof2{};
  DBuf d_basis, d_basisZ, d_w1, d_w2, d_r, d_part, d_h, d_nrm, d_gpart;
  // pinned host staging of one Arnoldi step's projections and norm
  cplx *h_pin = nullptr;
  int h_pin_n = 0;
  omx::LuWork lu{}, lu_sh{};
  bool profile = false;
  std::vector<hipEvent_t> pev;
  omx::LuProf prof{};
  double prof_gemm_ms = 0, prof_panel_ms = 0, prof_flop = 0, prof_total_ms = 0;
  int prof_launches = 0;
  DBuf d_piv2, d_perm2;
  DBuf d_dinv, d_dinv2, d_F64, d_F64sh;  // diagonal-block inverses; FP64 factors (the FP64 matrices stay intact)
  // wavelength sweeps: per-system buffers of one batch ([system][...])
  struct Sweep {
    DBuf S64, S32, V64, V32, F64, F64sh, eps, eps_sh, ov, incloc, Q, w, x, xint, xsh, xintsh, K, K1, cross, abscoef,
        abssh, absbes;
    DBuf perm, perm2, dinv, dinv2, map, status, basis, basisZ, r, w1, w2, gpart, part, hd, nrm, sc,
        act, kz, omz, sys, map2, status2;
  } sw;
  // ACA-compressed operators S (0) and V (1) of OMX_SOLVER_ACA_GMRES
  struct AcaMat {
    int n = 0, P = 0, G = 0, naca = 0, ndense = 0, rmax = 0;
    double rsum = 0, bytes = 0, ref_mb = 0;
    std::vector<int> rank;           // [i P + j]: r, -1 dense, 0 diagonal
    std::vector<const cplx *> where;  // [i P + j]: U (compressed) or the block (dense)
    std::vector<DBuf> pools;          // factor pools, one per compression chunk
    DBuf dense, desc, rowptr, acalist, t, part;
    bool built = false;
    void release(size_t *T) {
      for(auto &p : pools) p.release(T);
      pools.clear();
      DBuf *b[] = {&dense, &desc, &rowptr, &acalist, &t, &part};
      for(DBuf *x : b) x->release(T);
      rank.clear();
      where.clear();
      built = false;
      bytes = ref_mb = rsum = 0;
      naca = ndense = rmax = 0;
    }
  } aca[2];
  DBuf d_acaA, d_acaU, d_acaV, d_acaR, d_acaOff, d_plist;
  // the inner GMRES of OMX_SOLVER_GMRES32_IR: its own basis, vectors and pinned staging
  struct GmresWs {
    DBuf basis, r, w1, w2, h, part, nrm, gpart;
    cplx *h_pin = nullptr;
    int h_pin_n = 0;
    void release(size_t *T) {
      DBuf *b[] = {&basis, &r, &w1, &w2, &h, &part, &nrm, &gpart};
      for(DBuf *x : b) x->release(T);
      if(h_pin) (void)hipHostFree(h_pin);
      h_pin = nullptr;
      h_pin_n = 0;
    }
  } ws_in;
  // Captured factorisations (omx_set_graphs): executable graphs of S's
  // factorisation (s1, then s2 after V's start mark) and V's, and the state
  // they were captured for; one set for single solves, one for sweep chunks
  struct FactorGraphs {
    hipGraphExec_t s1 = nullptr, s2 = nullptr, v = nullptr;
    std::vector<uintptr_t> key;
    void reset() {
      hipGraphExec_t *g[] = {&s1, &s2, &v};
      for(hipGraphExec_t *x : g) {
        if(*x) (void)hipGraphExecDestroy(*x);
        *x = nullptr;
      }
      key.clear();
    }
  } fg, fg_sweep;
  bool graphs = false;
  int graph_captures = 0, graph_launches = 0;
  omx_stats stats{};
  void *scratch = nullptr;  // this device's address of the no-allocation solve's scratch array
  long long mv_calls = 0;  // FP64 matrix-vector products since the last omx_set_profile
  double mv_ms = 0, mv_bytes = 0;
  double omega() const { return omx::kC0 * (2.0 * omx::kPi / lambda); }
  cplx waveK() const {
    std::complex<double> b = std::sqrt(std::complex<double>(bg_eps_r.x, bg_eps_r.y) *
                                       std::complex<double>(bg_mu_r.x, bg_mu_r.y));
    double k0 = 2.0 * omx::kPi / lambda;
    return omx::mk(k0 * b.real(), k0 * b.imag());
  }
};

namespace {

template <typename F> int guard(F &&f) {
  try {
    f();
    return OMX_OK;
  } catch(OmxError &e) {
    g_err = e.what();
    return e.code;
  } catch(std::exception &e) {
    g_err = e.what();
    return OMX_ERR_ARG;
  }
}

cplx C(const double *p, int i) { return omx::mk(p[2 * i], p[2 * i + 1]); }

bool use32(int kind) { return kind == OMX_SOLVER_LU32_IR || kind == OMX_SOLVER_GMRES32_IR; }
bool is_lu(int kind) { return kind == OMX_SOLVER_LU64 || kind == OMX_SOLVER_LU32_IR; }

void alloc_lu(omx_context *c, omx::LuWork &w, DBuf &piv, DBuf &perm, DBuf &dinv, bool second) {
  const int Np = c->Np;
  const int gmax = (Np + omx::kPanelRows - 1) / omx::kPanelRows + 1;
  piv.ensure(sizeof(int) * Np, &c->total);
  dinv.ensure(sizeof(cplx) * omx::dinv_elems(Np), &c->total);
  perm.ensure(sizeof(int) * Np, &c->total);
  // tournament scratch: one half per system, as S and V may be factored
  // at the same time
  c->d_cand.ensure(sizeof(int) * 4 * gmax * omx::kNB, &c->total);
  c->d_cnt.ensure(sizeof(int) * 4 * gmax, &c->total);
  c->d_wlu.ensure(sizeof(cplx) * 6 * omx::kNB * omx::kNB, &c->total);
  if(!second) {
    c->d_map.ensure(sizeof(int) * (3 * ((size_t)Np + 2 * omx::kNB)), &c->total);
    c->d_status.ensure(64, &c->total);
  } else {
    c->d_map2.ensure(sizeof(int) * (3 * ((size_t)Np + 2 * omx::kNB)), &c->total);
    c->d_status2.ensure(64, &c->total);
  }
  DBuf &map = second ? c->d_map2 : c->d_map;
  DBuf &status = second ? c->d_status2 : c->d_status;
  HIPCHK(hipMemsetAsync(status.p, 0, 64, c->st));
  w = omx::LuWork{};
  w.piv = piv.as<int>();
  w.perm = perm.as<int>();
  const int half = second ? 1 : 0;
  w.cand[0] = c->d_cand.as<int>() + (2 * half) * gmax * omx::kNB;
  w.cand[1] = c->d_cand.as<int>() + (2 * half + 1) * gmax * omx::kNB;
  w.cnt[0] = c->d_cnt.as<int>() + (2 * half) * gmax;
  w.cnt[1] = c->d_cnt.as<int>() + (2 * half + 1) * gmax;
  w.mdst = map.as<int>();
  w.msrc = map.as<int>() + (Np + 2 * omx::kNB);
  w.mcnt = map.as<int>() + 2 * (Np + 2 * omx::kNB);
  w.status = status.as<int>();
  w.wlu = c->d_wlu.as<cplx>() + (size_t)half * 3 * omx::kNB * omx::kNB;
  w.dinv = dinv.p;
  w.pivot = c->pivot;
  hipEvent_t *ev = second ? c->lu_ev2 : c->lu_ev;
  w.aux = second ? c->aux2 : c->aux;
  w.ev_next = ev[0];
  w.ev_piv = ev[1];
  w.ev_bulk = ev[2];
  w.ev_pan = ev[3];
  w.gmax = gmax;
}

// (Re)allocates every size-dependent buffer.
void alloc_all(omx_context *c) {
  const int P = c->P, nf = c->nf, N = c->N, Np = c->Np, S = c->nmax + 1;
  size_t *T = &c->total;
  c->d_xyz.ensure(sizeof(double) * 3 * P, T);
  c->d_radius.ensure(sizeof(double) * P, T);
  c->d_eps.ensure(sizeof(cplx) * P, T);
  c->d_mu.ensure(sizeof(cplx) * P, T);
  c->d_eps_sh.ensure(sizeof(cplx) * P, T);
  c->d_kppp.ensure(sizeof(cplx) * P, T);
  c->d_kpar.ensure(sizeof(cplx) * P, T);
  c->d_gam.ensure(sizeof(cplx) * P, T);
  c->d_sph.ensure(sizeof(double) * 3 * P, T);
  c->d_ov.ensure(sizeof(cplx) * omx::OV_COUNT * P * 2 * S, T);
  c->d_inc.ensure(sizeof(cplx) * 2 * nf, T);
  c->d_incloc.ensure(sizeof(cplx) * N, T);
  c->d_Q.ensure(sizeof(cplx) * Np, T);
  c->d_w.ensure(sizeof(double) * 2 * P * nf, T);
  c->d_x.ensure(sizeof(cplx) * Np, T);
  c->d_xsh.ensure(sizeof(cplx) * Np, T);
  c->d_xint.ensure(sizeof(cplx) * Np, T);
  c->d_xintsh.ensure(sizeof(cplx) * Np, T);
  c->d_K.ensure(sizeof(cplx) * Np, T);
  c->d_K1.ensure(sizeof(cplx) * Np, T);
  c->d_cross.ensure(sizeof(double) * 4 * P, T);
  c->d_abscoef.ensure(sizeof(cplx) * P * nf, T);
  c->d_abssh.ensure(sizeof(double) * P, T);
  c->d_absbes.ensure(sizeof(cplx) * P * omx::kAbsBes, T);
  c->d_r.ensure(sizeof(cplx) * Np, T);
  c->d_w1.ensure(sizeof(cplx) * Np, T);
  c->d_w2.ensure(sizeof(cplx) * Np, T);
  c->d_gpart.ensure(sizeof(cplx) * Np * omx::kGemvSplitMax, T);
  c->d_part.ensure(sizeof(cplx) * 128 * (c->restart + 2), T);
  c->d_h.ensure(sizeof(cplx) * 2 * (c->restart + 2), T);
  c->d_nrm.ensure(sizeof(double) * 256, T);
}

void alloc_matrices(omx_context *c) {
  const size_t n2 = (size_t)c->Np * c->Np;
  size_t *T = &c->total;
  // FP64 matrix: factorised (LU64), streamed (GMRES) or the residual operator
  // (IR); the compressed operators of ACA replace it
  bool need64 = c->kind != OMX_SOLVER_ACA_GMRES;
  bool need32 = use32(c->kind);
  if(need64) c->d_S64.ensure(sizeof(cplx) * n2, T); else c->d_S64.release(T);
  if(need32) c->d_S32.ensure(sizeof(float2) * n2, T); else c->d_S32.release(T);
  if(c->sh) {
    if(need64) c->d_V64.ensure(sizeof(cplx) * n2, T); else c->d_V64.release(T);
    if(need32) c->d_V32.ensure(sizeof(float2) * n2, T); else c->d_V32.release(T);
  } else {
    c->d_V64.release(T);
    c->d_V32.release(T);
  }
  const bool f64 = c->kind == OMX_SOLVER_LU64;
  if(f64) c->d_F64.ensure(sizeof(cplx) * n2, T); else c->d_F64.release(T);
  if(f64 && c->sh) c->d_F64sh.ensure(sizeof(cplx) * n2, T); else c->d_F64sh.release(T);
  c->d_basis.ensure(sizeof(cplx) * (size_t)c->Np * (c->restart + 1), T);
  // preconditioned (flexible) GMRES keeps M^{-1} v_j: the LU kinds and the mixed Krylov
  const bool flex = is_lu(c->kind) || c->kind == OMX_SOLVER_GMRES32_IR;
  if(flex) c->d_basisZ.ensure(sizeof(cplx) * (size_t)c->Np * (c->restart + 1), T);
  else c->d_basisZ.release(T);
  if(c->kind == OMX_SOLVER_GMRES32_IR) {
    auto &w = c->ws_in;
    const size_t Np = (size_t)c->Np;
    w.basis.ensure(sizeof(cplx) * Np * (kInnerSteps + 1), T);
    w.r.ensure(sizeof(cplx) * Np, T);
    w.w1.ensure(sizeof(cplx) * Np, T);
    w.w2.ensure(sizeof(cplx) * Np, T);
    w.h.ensure(sizeof(cplx) * 2 * (kInnerSteps + 2), T);
    w.part.ensure(sizeof(cplx) * 128 * (kInnerSteps + 2), T);
    w.nrm.ensure(sizeof(double) * 256, T);
    w.gpart.ensure(sizeof(cplx) * Np * omx::kGemvSplitMax, T);
  } else {
    c->ws_in.release(T);
  }
  alloc_lu(c, c->lu, c->d_piv, c->d_perm, c->d_dinv, false);
  if(c->sh) alloc_lu(c, c->lu_sh, c->d_piv2, c->d_perm2, c->d_dinv2, true);
}

float elapsed(omx_context *c, int a, int b) {
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev[a], c->ev[b]);
  return ms;
}

// ------------------------------------------------------------------ ACA
// Scattering_matrix_ACA_FF / _SH (PreconditionedMatrix.cpp:489-551,
// 696-759) on the device: block (i, j != i) is compressed when
// |c_i - c_j| >= 2 (r_i + r_j) (:526; the distance of Tools::findDistance on
// the reader's spherical coordinates, Tools.cpp:30-44) and kept dense
// otherwise.  Dense blocks are assembled by the pair kernel straight into
// their pool; compressed blocks are assembled chunk by chunk into n x n
// staging slots, compressed there (aca_kernel, at most n terms each) and
// their factors packed into a pool of exactly sum 2 n r entries per chunk,
// so the full matrix never exists on the device.
constexpr double kAcaEps = 1e-3;            // PreconditionedMatrix.cpp:774
constexpr double kAcaStagingBytes = 1.5e9;  // assembly + factor staging of one chunk

void aca_build(omx_context *c, bool shf) {
  omx_context::AcaMat &m = c->aca[shf ? 1 : 0];
  size_t *T = &c->total;
  m.release(T);
  const int P = c->P, nmax = c->nmax, n = 2 * c->nf;
  const size_t n2 = (size_t)n * n;
  hipStream_t st = c->st;
  m.n = n;
  m.P = P;
  m.rank.assign((size_t)P * P, 0);
  m.where.assign((size_t)P * P, nullptr);
  // centres through the reader's Cartesian -> spherical -> Cartesian round trip
  std::vector<double> ctr(3 * (size_t)P);
  for(int i = 0; i < P; ++i) {
    const double x = c->xyz[3 * i], y = c->xyz[3 * i + 1], z = c->xyz[3 * i + 2];
    const double r = std::sqrt(x * x + y * y + z * z);
    const double th = r > 0.0 ? std::acos(z / r) : 0.0, ph = r > 0.0 ? std::atan2(y, x) : 0.0;
    ctr[3 * i] = r * std::sin(th) * std::cos(ph);
    ctr[3 * i + 1] = r * std::sin(th) * std::sin(ph);
    ctr[3 * i + 2] = r * std::cos(th);
  }
  std::vector<int> dl, al;
  for(int i = 0; i < P; ++i)
    for(int j = 0; j < P; ++j) {
      if(i == j) continue;
      const double d = std::sqrt(std::pow(ctr[3 * j] - ctr[3 * i], 2.0) + std::pow(ctr[3 * j + 1] - ctr[3 * i + 1], 2.0) +
                                 std::pow(ctr[3 * j + 2] - ctr[3 * i + 2], 2.0));
      if(d >= 2.0 * (c->radius[i] + c->radius[j])) {
        al.push_back(i * P + j);
      } else {
        dl.push_back(i * P + j);
        m.rank[(size_t)i * P + j] = -1;
      }
    }
  m.ndense = (int)dl.size();
  m.naca = (int)al.size();
  cplx k = c->waveK();
  if(shf) k = omx::mk(2.0 * k.x, 2.0 * k.y);  // PreconditionedMatrix.cpp:725
  omx::PairArgs base{P, nmax, shf ? omx::OV_TSH : omx::OV_T, k, c->d_sph.as<double>(), c->d_ov.as<cplx>(),
                     c->d_coef.as<double>(), (size_t)n, nullptr, nullptr};
  base.zblk = n2;
  const size_t listn = std::max(dl.size(), (size_t)1);
  if(m.ndense) {
    m.dense.ensure(sizeof(cplx) * n2 * dl.size(), T);
    c->d_plist.ensure(sizeof(int) * listn, T);
    HIPCHK(hipMemcpyAsync(c->d_plist.p, dl.data(), sizeof(int) * dl.size(), hipMemcpyHostToDevice, st));
    omx::PairArgs pa = base;
    pa.S64 = m.dense.as<cplx>();
    pa.plist = c->d_plist.as<int>();
    pa.nlist = m.ndense;
    omx::run_pairs(pa, st);
    for(size_t b = 0; b < dl.size(); ++b) m.where[dl[b]] = m.dense.as<cplx>() + b * n2;
    HIPCHK(hipStreamSynchronize(st));  // dl and the list buffer are reused below
  }
  if(m.naca) {
    const int C = (int)std::max<double>(1.0, std::min<double>(m.naca, kAcaStagingBytes / (48.0 * n2)));
    c->d_acaA.ensure(sizeof(cplx) * n2 * C, T);
    c->d_acaU.ensure(sizeof(cplx) * n2 * C, T);
    c->d_acaV.ensure(sizeof(cplx) * n2 * C, T);
    c->d_acaR.ensure(sizeof(int) * C, T);
    c->d_acaOff.ensure(sizeof(size_t) * C, T);
    c->d_plist.ensure(sizeof(int) * std::max((size_t)C, listn), T);
    std::vector<int> rk(C);
    std::vector<size_t> off(C);
    for(int b0 = 0; b0 < m.naca; b0 += C) {
      const int nb = std::min(C, m.naca - b0);
      HIPCHK(hipMemcpyAsync(c->d_plist.p, al.data() + b0, sizeof(int) * nb, hipMemcpyHostToDevice, st));
      omx::PairArgs pa = base;
      pa.S64 = c->d_acaA.as<cplx>();
      pa.plist = c->d_plist.as<int>();
      pa.nlist = nb;
      omx::run_pairs(pa, st);
      omx::AcaArgs aa{n, nb, kAcaEps, c->d_acaA.as<cplx>(), c->d_acaU.as<cplx>(), c->d_acaV.as<cplx>(),
                      c->d_acaR.as<int>()};
      omx::run_aca(aa, st);
      HIPCHK(hipMemcpyAsync(rk.data(), c->d_acaR.p, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      size_t tot = 0;
      for(int b = 0; b < nb; ++b) {
        if(rk[b] < 1 || rk[b] > n) throw OmxError(OMX_ERR_DEVICE, "ACA returned rank " + std::to_string(rk[b]));
        off[b] = tot;
        tot += 2 * (size_t)n * rk[b];
      }
      m.pools.emplace_back();
      DBuf &pool = m.pools.back();
      pool.ensure(sizeof(cplx) * tot, T);
      HIPCHK(hipMemcpyAsync(c->d_acaOff.p, off.data(), sizeof(size_t) * nb, hipMemcpyHostToDevice, st));
      omx::run_aca_pack(c->d_acaU.as<cplx>(), c->d_acaV.as<cplx>(), n, nb, c->d_acaR.as<int>(), pool.as<cplx>(),
                        c->d_acaOff.as<size_t>(), st);
      for(int b = 0; b < nb; ++b) {
        m.rank[al[b0 + b]] = rk[b];
        m.where[al[b0 + b]] = pool.as<cplx>() + off[b];
      }
      HIPCHK(hipStreamSynchronize(st));  // off, rk and the list buffer are reused
    }
    // the staging is only needed while compressing
    DBuf *stg[] = {&c->d_acaA, &c->d_acaU, &c->d_acaV};
    for(DBuf *b : stg) b->release(T);
  }
  // descriptors, block row by block row
  std::vector<omx::AcaDesc> desc;
  std::vector<int> rowptr(P + 1, 0), acalist;
  int toff = 0;
  m.bytes = 0;
  m.ref_mb = (double)P * n2 * 16.0 / 1e6;  // identity blocks
  for(int i = 0; i < P; ++i) {
    rowptr[i] = (int)desc.size();
    for(int j = 0; j < P; ++j) {
      const int r = m.rank[(size_t)i * P + j];
      if(i == j) continue;
      omx::AcaDesc d{m.where[(size_t)i * P + j], j, r, 0};
      if(r > 0) {
        d.toff = toff;
        toff += r;
        acalist.push_back((int)desc.size());
        m.bytes += 32.0 * n * r;
        m.ref_mb += 2.0 * n * r * 16.0 / 1e6;
        m.rmax = std::max(m.rmax, r);
        m.rsum += r;
      } else {
        m.bytes += 16.0 * n2;
        m.ref_mb += n2 * 16.0 / 1e6;
      }
      desc.push_back(d);
    }
  }
  rowptr[P] = (int)desc.size();
  m.G = P > 1 ? std::max(1, std::min(P - 1, (2048 + P - 1) / P)) : 0;
  m.desc.ensure(sizeof(omx::AcaDesc) * std::max<size_t>(desc.size(), 1), T);
  m.rowptr.ensure(sizeof(int) * (P + 1), T);
  m.acalist.ensure(sizeof(int) * std::max<size_t>(acalist.size(), 1), T);
  m.t.ensure(sizeof(cplx) * std::max(toff, 1), T);
  m.part.ensure(sizeof(cplx) * (size_t)std::max(m.G, 1) * P * n, T);
  if(!desc.empty())
    HIPCHK(hipMemcpyAsync(m.desc.p, desc.data(), sizeof(omx::AcaDesc) * desc.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(m.rowptr.p, rowptr.data(), sizeof(int) * (P + 1), hipMemcpyHostToDevice, st));
  if(!acalist.empty())
    HIPCHK(hipMemcpyAsync(m.acalist.p, acalist.data(), sizeof(int) * acalist.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the host vectors end here
  m.built = true;
}

// y = S x (V x) with the compressed operator, on stream st.
void aca_apply(omx_context *c, bool shf, const cplx *x, cplx *y, hipStream_t st) {
  omx_context::AcaMat &m = c->aca[shf ? 1 : 0];
  if(!m.built) throw OmxError(OMX_ERR_STATE, "no compressed operator");
  omx::AcaOp op{c->P, m.n, c->N, c->Np, m.G, m.desc.as<omx::AcaDesc>(), m.rowptr.as<int>(), m.acalist.as<int>(),
                m.naca, m.t.as<cplx>(), m.part.as<cplx>()};
  omx::aca_matvec(op, x, y, st);
}

// ------------------------------------------------------------------ GMRES
// Right-preconditioned restarted GMRES in FP64 with classical Gram-Schmidt
// (two passes).  prec(in, out) applies M^{-1} (identity when null).  The
// Arnoldi process follows Gmres_Zcomp (PreconditionedMatrix.cpp:892-985):
// Givens rotations on the Hessenberg column, convergence on |g_{j+1}|.
struct Gmres {
  omx_context *c;
  const cplx *A;  // FP64 operator
  std::function<void(const cplx *, cplx *)> prec;
  int N;
  hipStream_t s;  // the stream every step runs on
  // an operator given as a product (set: used instead of A), and its bytes
  // read per product
  std::function<void(const cplx *, cplx *)> op = nullptr;
  double op_bytes = 0.0;
  // a workspace of its own (null: the context's); one_cycle: x starts at 0
  // (the caller zeroes it) and a single cycle runs, with no residual product
  // before or after it (a fixed-cost preconditioner)
  omx_context::GmresWs *ws = nullptr;
  bool one_cycle = false;
  double nrm(const cplx *x) {
    DBuf &dn = ws ? ws->nrm : c->d_nrm;
    omx::nrm2sq(x, N, dn.as<double>() + 1, dn.as<double>(), s);
    double v;
    HIPCHK(hipMemcpyAsync(&v, dn.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return std::sqrt(v);
  }
  void matvec(const cplx *x, cplx *y) {
    if(c->profile) HIPCHK(hipEventRecord(c->ev[6], s));
    if(op) {
      op(x, y);
      c->mv_bytes += op_bytes;
    } else {
      omx::gemv64(A, (size_t)c->Np, N, x, y, c->d_gpart.as<cplx>(), s);
      c->mv_bytes += 16.0 * c->Np * (double)c->Np;
    }
    ++c->mv_calls;
    if(c->profile) {
      HIPCHK(hipEventRecord(c->ev[7], s));
      HIPCHK(hipEventSynchronize(c->ev[7]));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
      c->mv_ms += ms;
    }
  }
  // Solves A x = b; x holds the initial guess.  Returns iterations; *res the
  // final true relative residual.
  int solve(const cplx *b, cplx *x, double tol, int maxit, int m, double *res) {
    // flexible form: the preconditioned vectors M^{-1} v_j are kept (Zb), so
    // the end of a cycle needs no further preconditioner application
    cplx *V = (ws ? ws->basis : c->d_basis).as<cplx>(), *Zb = c->d_basisZ.as<cplx>();
    cplx *r = (ws ? ws->r : c->d_r).as<cplx>(), *w = (ws ? ws->w1 : c->d_w1).as<cplx>();
    cplx *z = (ws ? ws->w2 : c->d_w2).as<cplx>();
    cplx *hd = (ws ? ws->h : c->d_h).as<cplx>(), *part = (ws ? ws->part : c->d_part).as<cplx>();
    DBuf &dnrm = ws ? ws->nrm : c->d_nrm;
    cplx *&h_pin = ws ? ws->h_pin : c->h_pin;
    int &h_pin_n = ws ? ws->h_pin_n : c->h_pin_n;
    const size_t ld = (size_t)c->Np;
    double bn = nrm(b);
    if(bn == 0.0) {
      omx::vaxpby(x, omx::mk(0.0), x, omx::mk(0.0), N, s);
      *res = 0.0;
      return 0;
    }
    if(h_pin_n < 2 * (m + 2) + 1) {
      if(h_pin) (void)hipHostFree(h_pin);
      h_pin = nullptr;
      h_pin_n = 0;
      HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&h_pin), sizeof(cplx) * (2 * (m + 2) + 1)));
      h_pin_n = 2 * (m + 2) + 1;
    }
    cplx *hpin = h_pin;
    int it = 0, stall = 0;
    double prev = 0.0;
    bool tiny_step = false;
    std::vector<std::complex<double>> H((size_t)(m + 1) * m), g(m + 1), cs(m), sn(m);
    int cycles = 0;
    while(true) {
      if(one_cycle && cycles == 1) break;
      if(one_cycle) {  // x = 0: r = b
        omx::vcopy(r, b, N, s);
      } else {
        matvec(x, r);
        omx::vsub(r, b, r, N, s);
      }
      ++cycles;
      double beta = one_cycle ? bn : nrm(r);
      *res = beta / bn;
      if(*res <= tol || it >= maxit || tiny_step) break;
      // stagnation at the FP64 floor: a restart cycle without halving once the
      // residual is within the acceptance band (<= 1e3 tol), two once it is
      // down to 1e-10 (further up a slow phase of the restarted iteration is
      // waited out, up to maxit)
      if(prev > 0.0 && beta > 0.5 * prev && *res < 1e-10) {
        if(++stall >= 2 || *res <= 1e3 * tol) break;
      } else {
        stall = 0;
      }
      prev = beta;
      omx::vaxpby(V, omx::mk(1.0 / beta), r, omx::mk(0.0), N, s);
      std::fill(g.begin(), g.end(), 0.0);
      g[0] = beta;
      int j = 0;
      for(; j < m && it < maxit; ++j, ++it) {
        const cplx *vj = V + (size_t)j * ld;
        if(prec) {
          prec(vj, Zb + (size_t)j * ld);
          matvec(Zb + (size_t)j * ld, w);
        } else {
          matvec(vj, w);
        }
        // two Gram-Schmidt passes and the norm, copied to pinned memory with
        // one wait for all three
        for(int pass = 0; pass < 2; ++pass) {
          cplx *hp = hd + (size_t)pass * (m + 2);
          omx::dotc_multi(V, ld, j + 1, w, N, part, hp, s);
          omx::axpy_multi(V, ld, j + 1, hp, w, N, s);
          HIPCHK(hipMemcpyAsync(hpin + (size_t)pass * (m + 2), hp, sizeof(cplx) * (j + 1), hipMemcpyDeviceToHost, s));
        }
        omx::nrm2sq(w, N, dnrm.as<double>() + 1, dnrm.as<double>(), s);
        HIPCHK(hipMemcpyAsync(hpin + 2 * (m + 2), dnrm.p, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const cplx *h1 = hpin, *h2 = hpin + (m + 2);
        double hn = std::sqrt(reinterpret_cast<const double *>(hpin + 2 * (m + 2))[0]);
        for(int i = 0; i <= j; ++i)
          H[(size_t)j * (m + 1) + i] = std::complex<double>(h1[i].x + h2[i].x, h1[i].y + h2[i].y);
        H[(size_t)j * (m + 1) + j + 1] = hn;
        if(hn > 0.0) omx::vaxpby(V + (size_t)(j + 1) * ld, omx::mk(1.0 / hn), w, omx::mk(0.0), N, s);
        auto *hc = &H[(size_t)j * (m + 1)];
        for(int i = 0; i < j; ++i) {
          std::complex<double> t = cs[i] * hc[i] + sn[i] * hc[i + 1];
          hc[i + 1] = -std::conj(sn[i]) * hc[i] + cs[i] * hc[i + 1];
          hc[i] = t;
        }
        double den = std::sqrt(std::norm(hc[j]) + std::norm(hc[j + 1]));
        if(den == 0.0) {
          cs[j] = 1.0;
          sn[j] = 0.0;
        } else {
          std::complex<double> ph = std::abs(hc[j]) > 0 ? hc[j] / std::abs(hc[j]) : 1.0;
          cs[j] = std::abs(hc[j]) / den;
          sn[j] = ph * std::conj(hc[j + 1]) / den;
        }
        hc[j] = cs[j] * hc[j] + sn[j] * hc[j + 1];
        hc[j + 1] = 0.0;
        g[j + 1] = -std::conj(sn[j]) * g[j];
        g[j] = cs[j] * g[j];
        if(std::abs(g[j + 1]) <= 0.5 * tol * bn || hn == 0.0) {
          ++j;
          ++it;
          break;
        }
      }
      // y = H^{-1} g (upper triangular j x j), x += M^{-1} V y
      std::vector<std::complex<double>> y(j);
      for(int i = j - 1; i >= 0; --i) {
        std::complex<double> s = g[i];
        for(int k = i + 1; k < j; ++k) s -= H[(size_t)k * (m + 1) + i] * y[k];
        y[i] = s / H[(size_t)i * (m + 1) + i];
      }
      std::vector<cplx> yh(j);
      for(int i = 0; i < j; ++i) yh[i] = omx::mk(y[i].real(), y[i].imag());
      HIPCHK(hipMemcpyAsync(hd, yh.data(), sizeof(cplx) * j, hipMemcpyHostToDevice, s));
      omx::gemv_multi(prec ? Zb : V, ld, j, hd, z, N, s);
      const cplx *dx = z;
      omx::vaxpby(x, omx::mk(1.0), dx, omx::mk(1.0), N, s);
      // a correction below the FP64 resolution of x ends the refinement
      if(!one_cycle && nrm(dx) <= 1e-16 * nrm(x)) tiny_step = true;
    }
    return it;
  }
};

hipStream_t sys_stream(const omx_context *c, bool shf) { return shf ? c->st2 : c->st; }

void check_status(omx_context *c, bool shf = false) {
  int s[2] = {0, 0};
  hipStream_t st = sys_stream(c, shf);
  HIPCHK(hipMemcpyAsync(s, (shf ? c->d_status2 : c->d_status).p, sizeof(s), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if(s[1]) throw OmxError(OMX_ERR_DEVICE, "device wait timed out (code " + std::to_string(s[1]) + ")");
  if(s[0]) throw OmxError(OMX_ERR_SINGULAR, "zero pivot at column " + std::to_string(s[0] - 1));
}

omx::LuProf *prof_begin(omx_context *c, bool shf) {
  if(!c->profile) return nullptr;
  std::vector<hipEvent_t> &pev = shf ? c->pev2 : c->pev;
  omx::LuProf &prof = shf ? c->prof2 : c->prof;
  int cap = c->Np / omx::kNB + 1;
  if((int)pev.size() < 4 * cap + 2) {
    for(auto e : pev) (void)hipEventDestroy(e);
    pev.assign(4 * cap + 2, nullptr);
    for(auto &e : pev) HIPCHK(hipEventCreate(&e));
  }
  prof.ev = pev.data() + 2;
  prof.cap = cap;
  HIPCHK(hipEventRecord(pev[0], sys_stream(c, shf)));
  return &prof;
}
void prof_end(omx_context *c, bool shf) {
  if(!c->profile) return;
  std::vector<hipEvent_t> &pev = shf ? c->pev2 : c->pev;
  omx::LuProf &prof = shf ? c->prof2 : c->prof;
  HIPCHK(hipEventRecord(pev[1], sys_stream(c, shf)));
  HIPCHK(hipEventSynchronize(pev[1]));
  float ms = 0;
  for(int i = 0; i < prof.used; ++i) {
    hipEvent_t *e = prof.ev + 4 * i;
    HIPCHK(hipEventElapsedTime(&ms, e[0], e[1]));
    c->prof_panel_ms += ms;
    HIPCHK(hipEventElapsedTime(&ms, e[2], e[3]));
    c->prof_gemm_ms += ms;
  }
  HIPCHK(hipEventElapsedTime(&ms, pev[0], pev[1]));
  c->prof_total_ms += ms;
  c->prof_flop += prof.gemm_flop;
  c->prof_launches += prof.gemm_launches;
}

// Enqueues the factorisation of S (V) with pivot mode pv on the system's
// stream; the status is checked by the caller.
void factor_enqueue(omx_context *c, bool shf, int pv) {
  omx::LuWork &lu = shf ? c->lu_sh : c->lu;
  cplx *A64 = (shf ? c->d_V64 : c->d_S64).as<cplx>();
  const int Np = c->Np;
  const size_t lda = (size_t)Np, n2 = lda * Np;
  hipStream_t st = sys_stream(c, shf);
  omx::zero_ints((shf ? c->d_status2 : c->d_status).as<int>(), 4, st);  // a kernel: this may be captured
  lu.pivot = pv;
  lu.split = c->tuned ? c->tuned_split : c->lu32_gemm == OMX_GEMM_BF16X3;
  omx::LuProf *prof = prof_begin(c, shf);
  int rc;
  if(c->kind == OMX_SOLVER_LU64) {
    cplx *F64 = (shf ? c->d_F64sh : c->d_F64).as<cplx>();
    HIPCHK(hipMemcpyAsync(F64, A64, sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
    rc = omx::lu_factor<double2>(reinterpret_cast<double2 *>(F64), Np, lda, lu, st, prof);
  } else {
    float2 *A32 = (shf ? c->d_V32 : c->d_S32).as<float2>();
    bool &dirty = shf ? c->dirty32_sh : c->dirty32;
    if(dirty) omx::demote(A32, A64, n2, st);  // an earlier factorisation overwrote it
    dirty = true;
    rc = omx::lu_factor<float2>(A32, Np, lda, lu, st, prof);
  }
  if(rc) throw OmxError(OMX_ERR_DEVICE, "lu_factor launch failed (" + std::to_string(rc) + ")");
  prof_end(c, shf);
  (shf ? c->factored_sh : c->factored) = true;
}

// ------------------------------------------------------------------ graphs
// With omx_set_graphs on, the block-pivoting factorisations are captured once
// per state (the key: order, kind, pivoting, product, whether the FP32 copy
// must be refreshed first, and every buffer address the kernels take) and
// replayed afterwards.  S's capture is cut at V's start mark into two graphs
// and V's factorisation is captured on its own stream pair, so a replay keeps
// the eager overlap: S's first part, then V on the second stream after it
// beside S's rest.  Inside the captures every clear is a kernel.

void end_capture_quietly(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if(hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(st, &g);
    if(g) (void)hipGraphDestroy(g);
  }
  (void)hipGetLastError();
}

hipGraphExec_t instantiate(hipGraph_t g) {
  if(!g) return nullptr;
  size_t n = 0;
  HIPCHK(hipGraphGetNodes(g, nullptr, &n));
  hipGraphExec_t e = nullptr;
  if(n) HIPCHK(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
  return e;
}

std::vector<uintptr_t> single_key(omx_context *c, bool with_v, int mark) {
  const bool split = c->tuned ? c->tuned_split : c->lu32_gemm == OMX_GEMM_BF16X3;
  std::vector<uintptr_t> k{1,
                           (uintptr_t)with_v,
                           (uintptr_t)c->Np,
                           (uintptr_t)c->kind,
                           (uintptr_t)c->pivot,
                           (uintptr_t)split,
                           (uintptr_t)mark,
                           (uintptr_t)c->dirty32,
                           (uintptr_t)(with_v && c->dirty32_sh)};
  const omx::LuWork *lw[] = {&c->lu, &c->lu_sh};
  for(const omx::LuWork *w : lw) {
    const void *p[] = {w->perm, w->mdst, w->msrc, w->mcnt, w->status, w->dinv, w->aux};
    for(const void *x : p) k.push_back((uintptr_t)x);
  }
  const void *b[] = {c->d_S64.p, c->d_S32.p, c->d_V64.p, c->d_V32.p, c->d_F64.p, c->d_F64sh.p, c->d_status.p,
                     c->d_status2.p};
  for(const void *x : b) k.push_back((uintptr_t)x);
  return k;
}

struct CaptureMark {
  omx_context *c;
  hipGraph_t g1 = nullptr, gv = nullptr;
};

// At V's start mark inside S's capture: close S's first graph, capture V's
// factorisation on the second stream pair, and reopen S's capture (the
// look-ahead's next waits are on events recorded in the new capture).
void capture_at_mark(void *arg) {
  auto *cm = static_cast<CaptureMark *>(arg);
  omx_context *c = cm->c;
  HIPCHK(hipStreamEndCapture(c->st, &cm->g1));
  HIPCHK(hipStreamBeginCapture(c->st2, hipStreamCaptureModeThreadLocal));
  factor_enqueue(c, true, c->pivot);
  HIPCHK(hipStreamEndCapture(c->st2, &cm->gv));
  HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  HIPCHK(hipEventRecord(c->lu.ev_next, c->st));
  HIPCHK(hipEventRecord(c->lu.ev_bulk, c->st));
}

// Enqueues S's factorisation (and V's from S's mark column on, with_v) as
// graph launches, capturing them first when the state has changed.
void factor_graphed(omx_context *c, bool with_v, int mark) {
  auto &g = c->fg;
  std::vector<uintptr_t> key = single_key(c, with_v, mark);
  if(g.key != key) {
    g.reset();
    CaptureMark cm{c};
    hipGraph_t g2 = nullptr;
    omx::LuWork &lu = c->lu;
    auto unmark = [&] {
      lu.ev_mark = nullptr;
      lu.on_mark = nullptr;
      lu.mark_arg = nullptr;
    };
    const bool d32 = c->dirty32, d32sh = c->dirty32_sh;
    try {
      HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
      if(with_v) {
        lu.ev_mark = c->xev[0];
        lu.mark_col = mark;
        lu.on_mark = capture_at_mark;
        lu.mark_arg = &cm;
      }
      factor_enqueue(c, false, c->pivot);
      unmark();
      HIPCHK(hipStreamEndCapture(c->st, with_v ? &g2 : &cm.g1));
      if(with_v && !cm.gv) throw OmxError(OMX_ERR_DEVICE, "factorisation capture: V's start mark was not reached");
      g.s1 = instantiate(cm.g1);
      g.s2 = instantiate(g2);
      g.v = instantiate(cm.gv);
    } catch(...) {
      unmark();
      end_capture_quietly(c->st);
      end_capture_quietly(c->st2);
      hipGraph_t gs[] = {cm.g1, cm.gv, g2};
      for(hipGraph_t x : gs)
        if(x) (void)hipGraphDestroy(x);
      g.reset();
      c->dirty32 = d32;
      c->dirty32_sh = d32sh;
      c->factored = c->factored_sh = false;
      throw;
    }
    hipGraph_t gs[] = {cm.g1, cm.gv, g2};
    for(hipGraph_t x : gs)
      if(x) (void)hipGraphDestroy(x);
    g.key = key;
    ++c->graph_captures;
  }
  // what factor_enqueue leaves behind
  const bool split = c->tuned ? c->tuned_split : c->lu32_gemm == OMX_GEMM_BF16X3;
  c->lu.pivot = c->pivot;
  c->lu.split = split;
  if(c->kind != OMX_SOLVER_LU64) c->dirty32 = true;
  c->factored = true;
  if(g.s1) HIPCHK(hipGraphLaunch(g.s1, c->st));
  if(with_v) {
    c->lu_sh.pivot = c->pivot;
    c->lu_sh.split = split;
    if(c->kind != OMX_SOLVER_LU64) c->dirty32_sh = true;
    c->factored_sh = true;
    HIPCHK(hipEventRecord(c->xev[0], c->st));
    HIPCHK(hipStreamWaitEvent(c->st2, c->xev[0], 0));
    if(g.v) HIPCHK(hipGraphLaunch(g.v, c->st2));
  }
  if(g.s2) HIPCHK(hipGraphLaunch(g.s2, c->st));
  ++c->graph_launches;
}

// The target may sit below the FP64 residual floor of an ill-conditioned
// system: a solution whose normwise backward error ||b - A x|| / (||A||_F ||x||
// + ||b||) is within 1000 units of FP64 roundoff is accepted (the relative
// residual's floor is u ||A|| ||x|| / ||b||, far above u when x is large).  The
// bound is a constant: it does not grow with tol, so a caller's loose tol
// never lets an FP32-quality answer through.
bool backward_stable(omx_context *c, Gmres &gm, const cplx *A64, const cplx *rhs, const cplx *x, double res,
                     hipStream_t st) {
  const size_t Np = (size_t)c->Np;
  const double bn = gm.nrm(rhs), xn = gm.nrm(x);
  double a2 = 0.0;
  const size_t n2 = Np * Np, chunk = (size_t)1 << 28;
  for(size_t off = 0; off < n2; off += chunk) {
    omx::nrm2sq(A64 + off, (int)std::min(chunk, n2 - off), c->d_nrm.as<double>() + 1, c->d_nrm.as<double>(), st);
    double v = 0.0;
    HIPCHK(hipMemcpyAsync(&v, c->d_nrm.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    a2 += v;
  }
  const double eta = (res * bn) / (std::sqrt(a2) * xn + bn);
  return eta <= kBackwardErrorBound;
}

// Solves (S or V) x = rhs with the context's solver kind, on the system's
// stream (a factorisation already enqueued by factor_enqueue is used).
void solve_system(omx_context *c, bool shf, const cplx *rhs, cplx *x, int *iters, double *res) {
  omx::LuWork &lu = shf ? c->lu_sh : c->lu;
  cplx *A64 = (shf ? c->d_V64 : c->d_S64).as<cplx>();
  bool &fact = shf ? c->factored_sh : c->factored;
  const int Np = c->Np;
  const size_t lda = (size_t)Np;
  hipStream_t st = sys_stream(c, shf);
  DBuf &status = shf ? c->d_status2 : c->d_status;
  Gmres gm{c, A64, nullptr, Np, st};
  if(c->kind == OMX_SOLVER_ACA_GMRES) {  // Gmres_Zcomp on the compressed operator
    gm.op = [c, shf, st](const cplx *in, cplx *out) { aca_apply(c, shf, in, out, st); };
    gm.op_bytes = c->aca[shf ? 1 : 0].bytes;
  }
  // mixed Krylov: the preconditioner is one GMRES cycle on the FP32 copy
  Gmres inner{c, nullptr, nullptr, Np, st};
  if(c->kind == OMX_SOLVER_GMRES32_IR) {
    float2 *A32w = (shf ? c->d_V32 : c->d_S32).as<float2>();
    bool &dirty = shf ? c->dirty32_sh : c->dirty32;
    if(dirty) omx::demote(A32w, A64, (size_t)Np * lda, st);  // an earlier factorisation overwrote it
    dirty = false;
    const float2 *A32c = A32w;
    cplx *gpart = c->ws_in.gpart.as<cplx>();
    inner.op = [A32c, lda, Np, gpart, st](const cplx *in, cplx *out) {
      omx::gemv32(A32c, lda, Np, in, out, gpart, st);
    };
    inner.op_bytes = 8.0 * Np * (double)Np;
    inner.ws = &c->ws_in;
    inner.one_cycle = true;
    gm.prec = [&inner, Np, st](const cplx *in, cplx *out) {
      HIPCHK(hipMemsetAsync(out, 0, sizeof(cplx) * Np, st));
      double r = 0.0;
      inner.solve(in, out, kInnerTol, kInnerSteps, kInnerSteps, &r);
    };
  }
  if(!is_lu(c->kind)) {
    const bool try_first = c->kind_req == OMX_SOLVER_AUTO;
    try {
      HIPCHK(hipMemsetAsync(status.p, 0, 16, st));
      HIPCHK(hipMemsetAsync(x, 0, sizeof(cplx) * Np, st));
      const int maxit =
          try_first ? std::min(c->maxit, c->auto_budget > 0 ? c->auto_budget : auto_krylov_steps(c->N)) : c->maxit;
      *iters = gm.solve(rhs, x, c->tol, maxit, c->restart, res);
      check_status(c, shf);
      if(!(*res <= 1e3 * c->tol) && (c->kind == OMX_SOLVER_ACA_GMRES || !backward_stable(c, gm, A64, rhs, x, *res, st)))
        throw OmxError(OMX_ERR_NOCONV, "GMRES stopped at relative residual " + std::to_string(*res));
      return;
    } catch(OmxError &e) {
      if(!(try_first && e.code == OMX_ERR_NOCONV)) throw;
      c->kind = OMX_SOLVER_LU32_IR;  // the same S and S32: on to the mixed LU
      ++c->fallbacks;
    }
  }
  if(!shf && c->krylov_s && !fact) {
    // FP64 GMRES on S while V factorises; S's own mixed LU if it falls short
    omx::zero_ints(status.as<int>(), 4, st);
    HIPCHK(hipMemsetAsync(x, 0, sizeof(cplx) * Np, st));
    const int maxit = std::min(c->maxit, c->auto_budget > 0 ? c->auto_budget : auto_gmres64_steps(c->N));
    *iters = gm.solve(rhs, x, c->tol, maxit, c->restart, res);
    if(*res <= 1e3 * c->tol || backward_stable(c, gm, A64, rhs, x, *res, st)) return;
    ++c->fallbacks;
  }
  const bool f64 = c->kind == OMX_SOLVER_LU64;
  cplx *F64 = (shf ? c->d_F64sh : c->d_F64).as<cplx>();
  float2 *A32 = (shf ? c->d_V32 : c->d_S32).as<float2>();
  // Block pivoting first (when selected); a zero pivot or a refinement that
  // does not converge falls back to tournament (partial) pivoting.
  const int first = fact ? lu.pivot : c->pivot;
  for(int pv = first; pv <= OMX_PIVOT_PARTIAL; ++pv) {
    const bool last = pv == OMX_PIVOT_PARTIAL;
    try {
      if(!fact || lu.pivot != pv) factor_enqueue(c, shf, pv);
      check_status(c, shf);  // zero pivots of this factorisation
      if(f64) {
        gm.prec = [F64, &lu, Np, lda, st](const cplx *in, cplx *out) {
          omx::lu_solve<double2>(reinterpret_cast<const double2 *>(F64), Np, lda, lu, in, out, st);
        };
        omx::lu_solve<double2>(reinterpret_cast<const double2 *>(F64), Np, lda, lu, rhs, x, st);
        if(pv == OMX_PIVOT_PARTIAL) {  // plain FP64 LU with partial pivoting
          check_status(c, shf);
          *iters = 0;
          *res = -1.0;
          return;
        }
      } else {
        gm.prec = [A32, &lu, Np, lda, st](const cplx *in, cplx *out) {
          omx::lu_solve<float2>(A32, Np, lda, lu, in, out, st);
        };
        omx::lu_solve<float2>(A32, Np, lda, lu, rhs, x, st);  // x0 = M^{-1} b
      }
      *iters = gm.solve(rhs, x, c->tol, c->maxit, c->restart, res);
      check_status(c, shf);
      // the target may sit below the FP64 residual floor of an ill-conditioned
      // system; a stagnated refinement within 1000x of it is accepted, and so
      // is one whose normwise backward error ||b - A x|| / (||A||_F ||x|| + ||b||)
      // is within 1000x of the FP64 unit roundoff (the relative residual's
      // floor is u ||A|| ||x|| / ||b||, far above u when x is large).  That
      // bound is a constant: it does not grow with tol, so a caller's loose
      // tol never lets an FP32-quality answer through.
      if(!(*res <= 1e3 * c->tol) && !backward_stable(c, gm, A64, rhs, x, *res, st))
        throw OmxError(OMX_ERR_NOCONV, "refinement stopped at relative residual " + std::to_string(*res) +
                                           ", backward error above " + std::to_string(kBackwardErrorBound));
      return;
    } catch(OmxError &e) {
      if(last || (e.code != OMX_ERR_SINGULAR && e.code != OMX_ERR_NOCONV)) throw;
      ++c->fallbacks;
      fact = false;
    }
  }
}

// ------------------------------------------------------------------ batches
// Restarted GMRES for nb systems in lockstep (the Gmres above, with every
// vector operation batched): each system keeps its own Hessenberg matrix,
// rotations, residual and stopping state; a system that has converged (or
// stalled) in a cycle has its further basis vectors scaled to zero and its
// solution left alone.  prec(in, out) applies the batched preconditioner.
struct GmresB {
  omx_context *c;
  const cplx *A;  // FP64 operators, sA apart
  size_t sA;
  int Np, nb;
  std::function<void(const cplx *, cplx *)> prec;
  omx_context::Sweep &w;
  omx::LuWork *lw = nullptr;  // the preconditioner's factorisation: its solves may take a subset
  size_t sv() const { return (size_t)Np; }
  // The systems with mask[z] set: their count, and (when some are left out)
  // their indices in w.sys, so the matrix-sized work runs on those only.
  // Converged systems thus drop out of the lockstep batch's products and
  // triangular solves; their vectors are left as they are and ignored.
  int subset(const std::vector<int> &mask, const int **sys) {
    std::vector<int> idx;
    for(int z = 0; z < nb; ++z)
      if(mask[z]) idx.push_back(z);
    *sys = nullptr;
    if((int)idx.size() == nb || idx.empty()) return (int)idx.size();
    HIPCHK(hipMemcpyAsync(w.sys.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, c->st));
    *sys = w.sys.as<int>();
    return (int)idx.size();
  }
  void precond(const cplx *in, cplx *out, const std::vector<int> &mask) {
    if(!lw) {
      prec(in, out);
      return;
    }
    const int *sys;
    const int n = subset(mask, &sys);
    if(n == 0) return;
    lw->sys = sys;
    lw->nsys = n;
    prec(in, out);
    lw->sys = nullptr;
    lw->nsys = 0;
  }
  std::vector<double> nrms(const cplx *x) {
    omx::nrm2sq_b(x, sv(), Np, nb, w.nrm.as<double>() + nb, w.nrm.as<double>(), c->st);
    std::vector<double> v(nb);
    HIPCHK(hipMemcpyAsync(v.data(), w.nrm.p, sizeof(double) * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for(double &e : v) e = std::sqrt(e);
    return v;
  }
  void matvec(const cplx *x, cplx *y, const std::vector<int> &mask) {
    const int *sys;
    const int n = subset(mask, &sys);
    if(n == 0) return;
    if(c->profile) HIPCHK(hipEventRecord(c->ev[4], c->st));
    omx::gemv64(A, (size_t)Np, Np, x, y, w.gpart.as<cplx>(), c->st, nb, sA, sv(), sys, n);
    ++c->mv_calls;
    c->mv_bytes += 16.0 * Np * (double)Np * n;
    if(c->profile) {
      HIPCHK(hipEventRecord(c->ev[5], c->st));
      HIPCHK(hipEventSynchronize(c->ev[5]));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
      c->mv_ms += ms;
    }
  }
  // per-system scalars a[z], b[z] in device memory
  void scale(cplx *y, size_t sy, const std::vector<cplx> &a, const cplx *x, size_t sx, const std::vector<cplx> &b,
             const int *act) {
    cplx *sc = w.sc.as<cplx>();
    std::vector<cplx> h(2 * nb);
    for(int z = 0; z < nb; ++z) {
      h[z] = a[z];
      h[nb + z] = b[z];
    }
    HIPCHK(hipMemcpyAsync(sc, h.data(), sizeof(cplx) * 2 * nb, hipMemcpyHostToDevice, c->st));
    omx::vaxpby_b(y, sy, sc, x, sx, sc + nb, Np, nb, act, c->st);
  }
  // Solves A_z x_z = b_z (x holds the initial guesses); res[z] the final
  // relative residual, its[z] the iterations spent on system z.
  void solve(const cplx *b, cplx *x, double tol, int maxit, int m, std::vector<double> &res, std::vector<int> &its) {
    // basis vector k of system z at V + (k nb + z) Np: the systems' k-th
    // vectors are contiguous, as the batched preconditioner expects
    cplx *V = w.basis.as<cplx>(), *Zb = w.basisZ.as<cplx>();  // Zb: M^{-1} V (flexible form)
    const size_t sV = (size_t)Np, ld = (size_t)Np * nb;
    cplx *r = w.r.as<cplx>(), *wv = w.w1.as<cplx>();
    cplx *hd = w.hd.as<cplx>(), *part = w.part.as<cplx>();
    const int sh = m + 2;
    int *act_d = w.act.as<int>();
    res.assign(nb, 0.0);
    its.assign(nb, 0);
    const std::vector<double> bn = nrms(b);
    std::vector<int> active(nb), stall(nb, 0);
    std::vector<double> prev(nb, 0.0);
    for(int z = 0; z < nb; ++z) active[z] = bn[z] > 0.0;
    const cplx zero = omx::mk(0.0), one = omx::mk(1.0);
    {  // zero right-hand side: zero solution
      std::vector<cplx> a(nb, zero), bb(nb);
      bool any = false;
      for(int z = 0; z < nb; ++z) {
        bb[z] = active[z] ? one : zero;
        any |= !active[z];
      }
      if(any) scale(x, sv(), a, x, sv(), bb, nullptr);
    }
    using Z = std::complex<double>;
    std::vector<Z> H((size_t)nb * (m + 1) * m), g((size_t)nb * (m + 1)), cs((size_t)nb * m), sn((size_t)nb * m);
    std::vector<int> jz(nb);
    int it = 0;
    while(true) {
      matvec(x, r, active);
      omx::vsub_b(r, b, r, Np, nb, sv(), c->st);
      const std::vector<double> beta = nrms(r);
      bool any = false;
      for(int z = 0; z < nb; ++z) {
        if(!active[z]) continue;
        res[z] = beta[z] / bn[z];
        if(res[z] <= tol || its[z] >= maxit || !(res[z] == res[z])) {
          active[z] = 0;
          continue;
        }
        if(prev[z] > 0.0 && beta[z] > 0.5 * prev[z]) {
          // within the acceptance band one cycle without halving is the
          // FP64 floor; further up, two
          if(++stall[z] >= 2 || res[z] <= 1e3 * tol) {
            active[z] = 0;
            continue;
          }
        } else {
          stall[z] = 0;
        }
        prev[z] = beta[z];
        any = true;
      }
      if(!any || it >= maxit) break;
      std::vector<int> live(active);  // systems still building this cycle's basis
      {
        std::vector<cplx> a(nb, zero), bb(nb, zero);
        for(int z = 0; z < nb; ++z) a[z] = live[z] ? omx::mk(1.0 / beta[z]) : zero;
        scale(V, sV, a, r, sv(), bb, nullptr);  // V_0 of every system
      }
      for(int z = 0; z < nb; ++z) {
        std::fill(g.begin() + (size_t)z * (m + 1), g.begin() + (size_t)(z + 1) * (m + 1), Z(0.0));
        g[(size_t)z * (m + 1)] = beta[z];
        jz[z] = 0;
      }
      for(int j = 0; j < m && it < maxit; ++j, ++it) {
        const cplx *vj = V + (size_t)j * ld;
        cplx *zj = Zb + (size_t)j * ld;
        precond(vj, zj, live);
        matvec(zj, wv, live);
        std::vector<cplx> h1((size_t)nb * sh), h2((size_t)nb * sh);
        for(int pass = 0; pass < 2; ++pass) {
          omx::dotc_multi_b(V, ld, sV, j + 1, wv, sv(), Np, nb, part, hd, sh, c->st);
          omx::axpy_multi_b(V, ld, sV, j + 1, hd, sh, wv, sv(), Np, nb, c->st);
          HIPCHK(hipMemcpyAsync(pass ? h2.data() : h1.data(), hd, sizeof(cplx) * nb * sh, hipMemcpyDeviceToHost,
                                c->st));
        }
        const std::vector<double> hn = nrms(wv);
        {
          std::vector<cplx> a(nb, zero), bb(nb, zero);
          for(int z = 0; z < nb; ++z) a[z] = (live[z] && hn[z] > 0.0) ? omx::mk(1.0 / hn[z]) : zero;
          scale(V + (size_t)(j + 1) * ld, sV, a, wv, sv(), bb, nullptr);
        }
        bool any_live = false;
        for(int z = 0; z < nb; ++z) {
          if(!live[z]) continue;
          Z *hc = &H[((size_t)z * m + j) * (m + 1)];
          for(int i = 0; i <= j; ++i)
            hc[i] = Z(h1[(size_t)z * sh + i].x + h2[(size_t)z * sh + i].x, h1[(size_t)z * sh + i].y + h2[(size_t)z * sh + i].y);
          hc[j + 1] = hn[z];
          Z *csz = &cs[(size_t)z * m], *snz = &sn[(size_t)z * m], *gz = &g[(size_t)z * (m + 1)];
          for(int i = 0; i < j; ++i) {
            Z t = csz[i] * hc[i] + snz[i] * hc[i + 1];
            hc[i + 1] = -std::conj(snz[i]) * hc[i] + csz[i] * hc[i + 1];
            hc[i] = t;
          }
          const double den = std::sqrt(std::norm(hc[j]) + std::norm(hc[j + 1]));
          if(den == 0.0) {
            csz[j] = 1.0;
            snz[j] = 0.0;
          } else {
            const Z ph = std::abs(hc[j]) > 0 ? hc[j] / std::abs(hc[j]) : Z(1.0);
            csz[j] = std::abs(hc[j]) / den;
            snz[j] = ph * std::conj(hc[j + 1]) / den;
          }
          hc[j] = csz[j] * hc[j] + snz[j] * hc[j + 1];
          hc[j + 1] = 0.0;
          gz[j + 1] = -std::conj(snz[j]) * gz[j];
          gz[j] = csz[j] * gz[j];
          jz[z] = j + 1;
          ++its[z];
          if(std::abs(gz[j + 1]) <= 0.5 * tol * bn[z] || hn[z] == 0.0 || its[z] >= maxit)
            live[z] = 0;  // its later basis vectors are zero
          else
            any_live = true;
        }
        if(!any_live) {
          ++it;
          break;
        }
      }
      // y_z = H_z^{-1} g_z, x_z += Z_z y_z (= M^{-1} V_z y_z)
      int jmax = 0;
      std::vector<cplx> yh((size_t)nb * sh, zero);
      for(int z = 0; z < nb; ++z) {
        if(!active[z]) continue;
        const int jj = jz[z];
        jmax = std::max(jmax, jj);
        std::vector<Z> y(jj);
        for(int i = jj - 1; i >= 0; --i) {
          Z sacc = g[(size_t)z * (m + 1) + i];
          for(int k = i + 1; k < jj; ++k) sacc -= H[((size_t)z * m + k) * (m + 1) + i] * y[k];
          y[i] = sacc / H[((size_t)z * m + i) * (m + 1) + i];
        }
        for(int i = 0; i < jj; ++i) yh[(size_t)z * sh + i] = omx::mk(y[i].real(), y[i].imag());
      }
      if(jmax == 0) break;
      HIPCHK(hipMemcpyAsync(hd, yh.data(), sizeof(cplx) * nb * sh, hipMemcpyHostToDevice, c->st));
      omx::gemv_multi_b(Zb, ld, sV, jmax, hd, sh, wv, sv(), Np, nb, c->st);
      std::vector<int> act_h(active);
      HIPCHK(hipMemcpyAsync(act_d, act_h.data(), sizeof(int) * nb, hipMemcpyHostToDevice, c->st));
      scale(x, sv(), std::vector<cplx>(nb, one), wv, sv(), std::vector<cplx>(nb, one), act_d);
      // a correction below the FP64 resolution of x ends that system's refinement
      const std::vector<double> dn = nrms(wv), xn = nrms(x);
      for(int z = 0; z < nb; ++z)
        if(active[z] && dn[z] <= 1e-16 * xn[z]) active[z] = 0;
    }
  }
};

// Buffers of a batch of nb systems of order Np (mixed: FP32 copies too).
// Enqueues V's factorisation from a host thread of its own: the enqueueing
// of a factorisation proceeds at the device's pace, so enqueued from the
// calling thread it would hold back S's iterative solve (which needs the host
// between its products) until V's factorisation had run.  join() rethrows an
// error of the thread; the destructor joins without rethrowing.
struct SideEnqueue {
  std::thread t;
  std::exception_ptr err;
  void start(omx_context *c, void (*f)(void *)) {
    run(c->device, [c, f] { f(c); });
  }
  void run(int device, std::function<void()> job) {
    t = std::thread([this, device, job] {
      try {
        HIPCHK(hipSetDevice(device));
        job();
      } catch(...) {
        err = std::current_exception();
      }
    });
  }
  void join() {
    if(t.joinable()) t.join();
    if(err) {
      std::exception_ptr e = err;
      err = nullptr;
      std::rethrow_exception(e);
    }
  }
  ~SideEnqueue() {
    if(t.joinable()) t.join();
  }
};

void sweep_alloc(omx_context *c, int nb, bool f64, bool sh) {
  auto &w = c->sw;
  size_t *T = &c->total;
  const size_t Np = c->Np, n2 = Np * Np, P = c->P, S = c->nmax + 1;
  const int m = c->restart;
  w.S64.ensure(sizeof(cplx) * n2 * nb, T);
  if(f64) {
    w.F64.ensure(sizeof(cplx) * n2 * nb, T);
    w.S32.release(T);
  } else {
    w.S32.ensure(sizeof(float2) * n2 * nb, T);
    w.F64.release(T);
  }
  if(sh) {
    w.V64.ensure(sizeof(cplx) * n2 * nb, T);
    if(f64) w.F64sh.ensure(sizeof(cplx) * n2 * nb, T); else w.V32.ensure(sizeof(float2) * n2 * nb, T);
  }
  w.eps.ensure(sizeof(cplx) * P * nb, T);
  w.eps_sh.ensure(sizeof(cplx) * P * nb, T);
  w.ov.ensure(sizeof(cplx) * omx::OV_COUNT * P * 2 * S * nb, T);
  for(DBuf *b : {&w.incloc, &w.Q, &w.x, &w.xint, &w.xsh, &w.xintsh, &w.K, &w.K1, &w.r, &w.w1, &w.w2})
    b->ensure(sizeof(cplx) * Np * nb, T);
  w.w.ensure(sizeof(double) * 2 * P * c->nf * nb, T);
  w.cross.ensure(sizeof(double) * 4 * P * nb, T);
  w.abscoef.ensure(sizeof(cplx) * P * c->nf * nb, T);
  w.abssh.ensure(sizeof(double) * P * nb, T);
  w.absbes.ensure(sizeof(cplx) * P * omx::kAbsBes * nb, T);
  for(DBuf *b : {&w.perm, &w.perm2}) b->ensure(sizeof(int) * Np * nb, T);
  for(DBuf *b : {&w.dinv, &w.dinv2}) b->ensure(sizeof(cplx) * omx::dinv_elems(Np) * nb, T);
  w.map.ensure(sizeof(int) * 3 * (Np + 2 * omx::kNB) * nb, T);
  w.status.ensure(sizeof(int) * 4 * nb, T);
  w.sys.ensure(sizeof(int) * nb, T);
  if(sh) {  // V's factorisation runs beside S's: maps and status words of its own
    w.map2.ensure(sizeof(int) * 3 * (Np + 2 * omx::kNB) * nb, T);
    w.status2.ensure(sizeof(int) * 4 * nb, T);
  }
  w.basis.ensure(sizeof(cplx) * Np * (m + 1) * nb, T);
  w.basisZ.ensure(sizeof(cplx) * Np * (m + 1) * nb, T);
  w.gpart.ensure(sizeof(cplx) * Np * omx::kGemvSplitMax * nb, T);
  w.part.ensure(sizeof(cplx) * 128 * (m + 2) * nb, T);
  w.hd.ensure(sizeof(cplx) * (m + 2) * nb, T);
  w.nrm.ensure(sizeof(double) * 160 * nb, T);
  w.sc.ensure(sizeof(cplx) * 2 * nb, T);
  w.act.ensure(sizeof(int) * nb, T);
  w.kz.ensure(sizeof(cplx) * 2 * nb, T);
  w.omz.ensure(sizeof(double) * nb, T);
}

omx::LuWork sweep_lu(omx_context *c, int nb, bool second) {
  auto &w = c->sw;
  const int Np = c->Np;
  omx::LuWork lw{};
  lw.perm = (second ? w.perm2 : w.perm).as<int>();
  lw.piv = lw.perm;  // unused by block pivoting
  lw.dinv = (second ? w.dinv2 : w.dinv).p;
  const int ms = Np + 2 * omx::kNB;
  int *map = (second ? w.map2 : w.map).as<int>();
  lw.mdst = map;
  lw.msrc = map + (size_t)ms * nb;
  lw.mcnt = map + 2 * (size_t)ms * nb;
  lw.status = (second ? w.status2 : w.status).as<int>();
  lw.wlu = c->d_wlu.p ? c->d_wlu.as<cplx>() + (second ? 3 * omx::kNB * omx::kNB : 0) : nullptr;
  // V's factorisation runs on the second stream pair, beside S's work
  lw.aux = second ? c->aux2 : c->aux;
  const hipEvent_t *ev = second ? c->lu_ev2 : c->lu_ev;
  lw.ev_next = ev[0];
  lw.ev_piv = ev[1];
  lw.ev_bulk = ev[2];
  lw.ev_pan = ev[3];
  lw.gmax = 0;
  lw.pivot = omx::kPivotBlock;
  lw.split = c->lu32_gemm == OMX_GEMM_BF16X3;
  lw.nb = nb;
  lw.bs.a = (size_t)Np * Np;
  lw.bs.inv = omx::dinv_elems(Np);
  lw.bs.vec = (size_t)Np;
  lw.bs.perm = Np;
  lw.bs.map = ms;
  lw.bs.cnt = ms;
  lw.bs.status = 4;
  return lw;
}

// ------------------------------------------------------------------ autotuner
// Per shape (device, padded order, batch, factor type) the batched sweeps
// pick the bulk-update product of FP32 factors (split-bf16 or FP32 MFMA) and
// the triangular-solve variant by timing every candidate on the batch's own
// buffers before its first chunk (zeroed matrices, identity permutations:
// the kernels have no value-dependent branches, so the times are those of the
// real work), once per process.  A candidate replaces the default only when
// it is at least 5% faster, so tuning again returns the same choice unless
// two candidates sit at that margin.  The choices are kept per process and
// can be read back (omx_tune_query).
struct TuneChoice {
  bool split = true;
  int trsv = 0;
  double t_prod[2] = {0, 0};                   // ms: split-bf16, FP32 MFMA (FP32 factors only)
  double t_trsv[omx::kTrsvVariants] = {0, 0, 0};  // ms per batched LU solve
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, TuneChoice> g_tune;  // (device, Np, nb, fp64)
constexpr double kTuneMargin = 0.95;

double time_median(hipStream_t st, const std::function<void()> &f, int reps = 5) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  std::vector<float> ts;
  try {
    f();  // warm-up
    for(int r = 0; r < reps; ++r) {
      HIPCHK(hipEventRecord(a, st));
      f();
      HIPCHK(hipEventRecord(b, st));
      HIPCHK(hipEventSynchronize(b));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, a, b));
      ts.push_back(ms);
    }
  } catch(...) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  std::sort(ts.begin(), ts.end());
  return ts[ts.size() / 2];
}

omx::LuWork sweep_lu(omx_context *c, int nb, bool second);

// The tuned choice for this batch (tuning it first if needed); runs on the
// chunk's buffers before the chunk fills them.
TuneChoice sweep_tune(omx_context *c, int nb, bool f64) {
  const auto key = std::make_tuple(c->device, c->Np, nb, f64 ? 1 : 0);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if(it != g_tune.end()) return it->second;
  }
  auto &w = c->sw;
  const int Np = c->Np;
  const size_t n2 = (size_t)Np * Np, lda = (size_t)Np;
  hipStream_t st = c->st;
  TuneChoice ch;
  ch.split = c->lu32_gemm == OMX_GEMM_BF16X3;
  void *F = f64 ? w.F64.p : w.S32.p;
  const size_t fbytes = (f64 ? sizeof(double2) : sizeof(float2)) * n2 * nb;
  HIPCHK(hipMemsetAsync(F, 0, fbytes, st));
  if(!f64) {
    float2 *A = static_cast<float2 *>(F);
    ch.t_prod[0] = time_median(st, [&] { omx::bulk_update_probe<float2>(A, Np, lda, nb, n2, true, st); });
    ch.t_prod[1] = time_median(st, [&] { omx::bulk_update_probe<float2>(A, Np, lda, nb, n2, false, st); });
    const int dflt = ch.split ? 0 : 1, other = 1 - dflt;
    if(ch.t_prod[other] < kTuneMargin * ch.t_prod[dflt]) ch.split = other == 0;
    HIPCHK(hipMemsetAsync(F, 0, fbytes, st));
  }
  // triangular solves on zeroed factors and inverses, identity permutations
  std::vector<int> ident((size_t)Np * nb);
  for(int z = 0; z < nb; ++z)
    for(int i = 0; i < Np; ++i) ident[(size_t)z * Np + i] = i;
  HIPCHK(hipMemcpyAsync(w.perm.p, ident.data(), sizeof(int) * ident.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(w.dinv.p, 0, w.dinv.bytes, st));
  HIPCHK(hipMemsetAsync(w.r.p, 0, w.r.bytes, st));
  omx::LuWork lw = sweep_lu(c, nb, false);
  for(int v = 0; v < omx::kTrsvVariants; ++v) {
    lw.trsv = v;
    ch.t_trsv[v] = time_median(st, [&] {
      if(f64)
        omx::lu_solve<double2>(static_cast<const double2 *>(F), Np, lda, lw, w.r.as<cplx>(), w.w2.as<cplx>(), st);
      else
        omx::lu_solve<float2>(static_cast<const float2 *>(F), Np, lda, lw, w.r.as<cplx>(), w.w2.as<cplx>(), st);
    });
  }
  int best = 0;
  for(int v = 1; v < omx::kTrsvVariants; ++v)
    if(ch.t_trsv[v] < ch.t_trsv[best]) best = v;
  ch.trsv = ch.t_trsv[best] < kTuneMargin * ch.t_trsv[0] ? best : 0;
  HIPCHK(hipStreamSynchronize(st));  // ident ends here
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[key] = ch;
  return ch;
}

// The single-system LU solves (omx_update / omx_solve) tune the same way,
// with batch 1 and the context's own buffers (before the assembly fills
// them).
TuneChoice single_tune(omx_context *c, bool f64) {
  const auto key = std::make_tuple(c->device, c->Np, 1, f64 ? 1 : 0);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if(it != g_tune.end()) return it->second;
  }
  const int Np = c->Np;
  const size_t n2 = (size_t)Np * Np, lda = (size_t)Np;
  hipStream_t st = c->st;
  TuneChoice ch;
  ch.split = c->lu32_gemm == OMX_GEMM_BF16X3;
  void *F = f64 ? c->d_F64.p : c->d_S32.p;
  const size_t fbytes = (f64 ? sizeof(double2) : sizeof(float2)) * n2;
  HIPCHK(hipMemsetAsync(F, 0, fbytes, st));
  if(!f64) {
    float2 *A = static_cast<float2 *>(F);
    ch.t_prod[0] = time_median(st, [&] { omx::bulk_update_probe<float2>(A, Np, lda, 1, 0, true, st); });
    ch.t_prod[1] = time_median(st, [&] { omx::bulk_update_probe<float2>(A, Np, lda, 1, 0, false, st); });
    const int dflt = ch.split ? 0 : 1, other = 1 - dflt;
    if(ch.t_prod[other] < kTuneMargin * ch.t_prod[dflt]) ch.split = other == 0;
    HIPCHK(hipMemsetAsync(F, 0, fbytes, st));
  }
  std::vector<int> ident(Np);
  for(int i = 0; i < Np; ++i) ident[i] = i;
  HIPCHK(hipMemcpyAsync(c->d_perm.p, ident.data(), sizeof(int) * Np, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(c->d_dinv.p, 0, c->d_dinv.bytes, st));
  HIPCHK(hipMemsetAsync(c->d_r.p, 0, c->d_r.bytes, st));
  omx::LuWork &lw = c->lu;
  for(int v = 0; v < omx::kTrsvVariants; ++v) {
    lw.trsv = v;
    ch.t_trsv[v] = time_median(st, [&] {
      if(f64)
        omx::lu_solve<double2>(static_cast<const double2 *>(F), Np, lda, lw, c->d_r.as<cplx>(), c->d_w2.as<cplx>(),
                               st);
      else
        omx::lu_solve<float2>(static_cast<const float2 *>(F), Np, lda, lw, c->d_r.as<cplx>(), c->d_w2.as<cplx>(), st);
    });
  }
  lw.trsv = 0;
  int best = 0;
  for(int v = 1; v < omx::kTrsvVariants; ++v)
    if(ch.t_trsv[v] < ch.t_trsv[best]) best = v;
  ch.trsv = ch.t_trsv[best] < kTuneMargin * ch.t_trsv[0] ? best : 0;
  HIPCHK(hipStreamSynchronize(st));  // ident ends here
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[key] = ch;
  return ch;
}

// One batch of a wavelength sweep: nb systems (wavelengths lam[z], absolute
// permittivities eps[z * P + i], eps_sh[z * P + i]) assembled, factorised and
// refined side by side.  Systems whose factorisation hit a zero pivot or whose
// refinement did not converge are flagged in bad[z] for the caller to redo
// one at a time with partial pivoting.
void sweep_chunk(omx_context *c, int nb, const double *lam, const cplx *eps, const cplx *eps_sh, double *x_sca,
                 double *x_int, double *x_sca_sh, double *x_int_sh, double *xs, std::vector<int> &bad,
                 std::vector<int> &iters, std::vector<double> &resid, bool krylov_s) {
  const bool f64 = c->kind == OMX_SOLVER_LU64, sh = c->sh;
  sweep_alloc(c, nb, f64, sh);
  const TuneChoice tune = c->autotune ? sweep_tune(c, nb, f64) : TuneChoice{c->lu32_gemm == OMX_GEMM_BF16X3, 0};
  auto &w = c->sw;
  const int P = c->P, nmax = c->nmax, N = c->N, Np = c->Np;
  const size_t n2 = (size_t)Np * Np, lda = (size_t)Np;
  const size_t ovs = (size_t)omx::OV_COUNT * P * 2 * (nmax + 1), ws = (size_t)2 * P * c->nf;
  hipStream_t st = c->st;
  HIPCHK(hipMemcpyAsync(w.eps.p, eps, sizeof(cplx) * P * nb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(w.eps_sh.p, eps_sh, sizeof(cplx) * P * nb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(w.status.p, 0, w.status.bytes, st));
  if(sh) HIPCHK(hipMemsetAsync(w.status2.p, 0, w.status2.bytes, st));
  // the lockstep refinement leaves the vectors of systems that dropped out
  // as they are: start them finite
  HIPCHK(hipMemsetAsync(w.basisZ.p, 0, w.basisZ.bytes, st));
  HIPCHK(hipMemsetAsync(w.w1.p, 0, w.w1.bytes, st));
  HIPCHK(hipMemsetAsync(w.Q.p, 0, w.Q.bytes, st));
  HIPCHK(hipMemsetAsync(c->d_status.p, 0, 16, st));
  const cplx bge = kEps0 * c->bg_eps_r, bgm = kMu0 * c->bg_mu_r;
  std::complex<double> bgn = std::sqrt(std::complex<double>(c->bg_eps_r.x, c->bg_eps_r.y) *
                                       std::complex<double>(c->bg_mu_r.x, c->bg_mu_r.y));
  // plane-wave coefficients do not depend on the wavelength (Excitation.cpp:48-75)
  omx::IncArgs ia{nmax, c->theta, c->phi, c->eth, c->eph, c->d_inc.as<cplx>()};
  omx::run_incident(ia, st);
  cplx *cS64 = w.S64.as<cplx>(), *cV64 = w.V64.as<cplx>();
  float2 *cS32 = w.S32.as<float2>(), *cV32 = w.V32.as<float2>();
  // per-system wavenumbers (k at omega, 2k at 2 omega) and angular frequencies;
  // every assembly kernel then covers the whole batch in one launch (system in blockIdx.z)
  std::vector<cplx> hk(2 * nb);
  std::vector<double> hom(nb);
  for(int z = 0; z < nb; ++z) {
    const double k0 = 2.0 * omx::kPi / lam[z];
    hom[z] = omx::kC0 * k0;
    hk[z] = omx::mk(k0 * bgn.real(), k0 * bgn.imag());
    hk[nb + z] = omx::mk(2.0 * hk[z].x, 2.0 * hk[z].y);  // PreconditionedMatrix.cpp:592
  }
  HIPCHK(hipMemcpyAsync(w.kz.p, hk.data(), sizeof(cplx) * 2 * nb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(w.omz.p, hom.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st));
  const cplx *kz = w.kz.as<cplx>();
  omx::ObjArgs oa{P, nmax, hom[0], bge, bgm, c->d_xyz.as<double>(), c->d_radius.as<double>(), w.eps.as<cplx>(),
                  c->d_mu.as<cplx>(), w.eps_sh.as<cplx>(), c->d_sph.as<double>(), w.ov.as<cplx>()};
  oa.omz = w.omz.as<double>();
  oa.zov = ovs;
  omx::run_objects(oa, st, nb);
  omx::OriginArgs og{P, nmax, hk[0], c->d_sph.as<double>(), w.ov.as<cplx>(), c->d_coef.as<double>(),
                     c->d_inc.as<cplx>(), w.incloc.as<cplx>(), w.Q.as<cplx>(), w.w.as<double>()};
  og.kz = kz;
  og.zov = ovs;
  og.zvec = lda;
  og.zw = ws;
  omx::run_origin(og, st, nb);
  omx::PairArgs pa{P, nmax, omx::OV_T, hk[0], c->d_sph.as<double>(), w.ov.as<cplx>(), c->d_coef.as<double>(), lda,
                   cS64, f64 ? nullptr : cS32};
  pa.kz = kz;
  pa.zov = ovs;
  pa.zS = n2;
  omx::run_pairs(pa, st, nb);
  omx::run_pad64(cS64, N, Np, lda, st, nb, n2);
  if(!f64) omx::run_pad32(cS32, N, Np, lda, st, nb, n2);
  if(sh) {
    omx::PairArgs pv = pa;
    pv.tkind = omx::OV_TSH;
    pv.k = hk[nb];
    pv.kz = kz + nb;
    pv.S64 = cV64;
    pv.S32 = f64 ? nullptr : cV32;
    omx::run_pairs(pv, st, nb);
    omx::run_pad64(cV64, N, Np, lda, st, nb, n2);
    if(!f64) omx::run_pad32(cV32, N, Np, lda, st, nb, n2);
  }
  HIPCHK(hipGetLastError());
  auto status_scan = [&](std::vector<int> &flag, bool second) {
    std::vector<int> h(4 * nb);
    HIPCHK(hipMemcpyAsync(h.data(), (second ? w.status2 : w.status).p, sizeof(int) * 4 * nb, hipMemcpyDeviceToHost,
                          st));
    int g[2];
    HIPCHK(hipMemcpyAsync(g, c->d_status.p, sizeof(g), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if(h[1] || g[1]) throw OmxError(OMX_ERR_DEVICE, "device wait timed out in a batched solve");
    for(int z = 0; z < nb; ++z)
      if(h[4 * z]) flag[z] = 1;
  };
  // factorisation of S (V) for every system, enqueued on stream fst
  auto factor = [&](bool second, omx::LuWork &lw, hipStream_t fst) -> void * {
    cplx *A64 = second ? cV64 : cS64;
    void *F;
    int rc;
    if(f64) {
      F = (second ? w.F64sh : w.F64).p;
      HIPCHK(hipMemcpyAsync(F, A64, sizeof(cplx) * n2 * nb, hipMemcpyDeviceToDevice, fst));
      rc = omx::lu_factor<double2>(static_cast<double2 *>(F), Np, lda, lw, fst, nullptr);
    } else {
      F = (second ? w.V32 : w.S32).p;
      rc = omx::lu_factor<float2>(static_cast<float2 *>(F), Np, lda, lw, fst, nullptr);
    }
    if(rc) throw OmxError(OMX_ERR_DEVICE, "batched lu_factor launch failed (" + std::to_string(rc) + ")");
    return F;
  };
  // V does not depend on the fundamental solution: its factorisation starts
  // on the second stream pair right after the assembly and runs beside S's
  // factorisation and refinement; V's refinement waits for it
  omx::LuWork lwS = sweep_lu(c, nb, false), lwV{};
  lwS.split = tune.split;
  lwS.trsv = tune.trsv;
  if(sh) {
    lwV = sweep_lu(c, nb, true);
    lwV.split = tune.split;
    lwV.trsv = tune.trsv;
  }
  void *FS = f64 ? w.F64.p : w.S32.p, *FV = sh ? (f64 ? w.F64sh.p : w.V32.p) : nullptr;
  SideEnqueue vq;
  if(krylov_s) {
    // S is solved by GMRES (no factorisation); V's factorisation is enqueued
    // from a host thread of its own so that S's iterations start at once
    HIPCHK(hipEventRecord(c->xev[0], st));
    vq.run(c->device, [&] {
      HIPCHK(hipStreamWaitEvent(c->st2, c->xev[0], 0));
      factor(true, lwV, c->st2);
      HIPCHK(hipEventRecord(c->xev[1], c->st2));
    });
  } else if(c->graphs) {
    // captured once per batch shape and buffers, then replayed chunk after chunk
    auto &g = c->fg_sweep;
    std::vector<uintptr_t> key{2, (uintptr_t)nb, (uintptr_t)Np, (uintptr_t)f64, (uintptr_t)sh, (uintptr_t)tune.split};
    const void *p[] = {w.S64.p, w.S32.p, w.V64.p, w.V32.p, w.F64.p, w.F64sh.p, w.perm.p, w.perm2.p, w.map.p,
                       w.map2.p, w.status.p, w.status2.p, w.dinv.p, w.dinv2.p};
    for(const void *x : p) key.push_back((uintptr_t)x);
    if(g.key != key) {
      g.reset();
      hipGraph_t gs = nullptr, gv = nullptr;
      try {
        if(sh) {
          HIPCHK(hipStreamBeginCapture(c->st2, hipStreamCaptureModeThreadLocal));
          factor(true, lwV, c->st2);
          HIPCHK(hipStreamEndCapture(c->st2, &gv));
        }
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        factor(false, lwS, st);
        HIPCHK(hipStreamEndCapture(st, &gs));
        g.s1 = instantiate(gs);
        g.v = instantiate(gv);
      } catch(...) {
        end_capture_quietly(c->st2);
        end_capture_quietly(st);
        if(gs) (void)hipGraphDestroy(gs);
        if(gv) (void)hipGraphDestroy(gv);
        g.reset();
        throw;
      }
      if(gs) (void)hipGraphDestroy(gs);
      if(gv) (void)hipGraphDestroy(gv);
      g.key = key;
      ++c->graph_captures;
    }
    if(sh) {
      HIPCHK(hipEventRecord(c->xev[0], st));
      HIPCHK(hipStreamWaitEvent(c->st2, c->xev[0], 0));
      if(g.v) HIPCHK(hipGraphLaunch(g.v, c->st2));
      HIPCHK(hipEventRecord(c->xev[1], c->st2));
    }
    if(g.s1) HIPCHK(hipGraphLaunch(g.s1, st));
    ++c->graph_launches;
  } else {
    if(sh) {
      HIPCHK(hipEventRecord(c->xev[0], st));
      HIPCHK(hipStreamWaitEvent(c->st2, c->xev[0], 0));
      factor(true, lwV, c->st2);
      HIPCHK(hipEventRecord(c->xev[1], c->st2));
    }
    factor(false, lwS, st);
  }
  // refinement of every system of S (V); per-system convergence flags
  auto refine = [&](bool second, omx::LuWork &lw, void *F, const cplx *rhs, cplx *x, bool from_x = false) {
    status_scan(bad, second);
    cplx *A64 = second ? cV64 : cS64;
    omx::LuWork *lwp = &lw;
    GmresB gm{c, A64, n2, Np, nb, nullptr, w};
    gm.lw = lwp;
    if(f64)
      gm.prec = [c, F, lwp, Np, lda](const cplx *in, cplx *out) {
        omx::lu_solve<double2>(static_cast<const double2 *>(F), Np, lda, *lwp, in, out, c->st);
      };
    else
      gm.prec = [c, F, lwp, Np, lda](const cplx *in, cplx *out) {
        omx::lu_solve<float2>(static_cast<const float2 *>(F), Np, lda, *lwp, in, out, c->st);
      };
    if(!from_x) gm.prec(rhs, x);  // else: start from the iterate x holds
    std::vector<double> res;
    std::vector<int> its;
    // the batch runs in lockstep until its slowest system stops: a system
    // still short of the target after four restart cycles is left to the
    // one-at-a-time path (partial pivoting) instead of holding the batch
    const int bmax = nb > 1 ? std::min(c->maxit, 4 * c->restart) : c->maxit;
    gm.solve(rhs, x, c->tol, bmax, c->restart, res, its);
    status_scan(bad, second);
    for(int z = 0; z < nb; ++z) {
      if(!(res[z] <= 1e3 * c->tol)) bad[z] = 1;
      iters[2 * z + (second ? 1 : 0)] = its[z];
      resid[2 * z + (second ? 1 : 0)] = res[z];
    }
  };
  if(krylov_s) {
    // lockstep FP64 GMRES on every system of S, from zero, within twice the
    // single-system budget (S's solve runs beside V's factorisation); a
    // system short of the target is left to the one-at-a-time path
    status_scan(bad, false);
    GmresB gm{c, cS64, n2, Np, nb, nullptr, w};
    cplx *Sx = w.x.as<cplx>();
    gm.prec = [c, Np, nb](const cplx *in, cplx *out) { omx::vcopy(out, in, Np * nb, c->st); };
    HIPCHK(hipMemsetAsync(Sx, 0, sizeof(cplx) * lda * nb, st));
    std::vector<double> res;
    std::vector<int> its;
    const int budget = std::min(c->maxit, c->auto_budget > 0 ? c->auto_budget : 2 * auto_gmres64_steps(c->N));
    gm.solve(w.Q.as<cplx>(), Sx, c->tol, budget, c->restart, res, its);
    bool short_any = false;
    for(int z = 0; z < nb; ++z) {
      short_any |= !(res[z] <= 1e3 * c->tol);
      iters[2 * z] = its[z];
      resid[2 * z] = res[z];
    }
    if(short_any) {
      // S factored for the batch after all; every system refined from its
      // GMRES iterate (those already at the target stop at once)
      factor(false, lwS, st);
      refine(false, lwS, FS, w.Q.as<cplx>(), Sx, true);
      ++c->fallbacks;
    }
  } else {
    refine(false, lwS, FS, w.Q.as<cplx>(), w.x.as<cplx>());
  }
  for(int z = 0; z < nb; ++z)
    omx::vmul_ov(w.xint.as<cplx>() + z * lda, w.x.as<cplx>() + z * lda, w.ov.as<cplx>() + z * ovs, omx::OV_IAUX, P,
                 nmax, nullptr, N, st);
  if(sh) {
    if(c->cg_nmax != nmax) {
      size_t nf = (size_t)c->nf, sz = nf * nf * nf;
      c->d_cg.ensure(sizeof(double) * 9 * sz, &c->total);
      omx::run_cg(nmax, c->d_cg.as<double>(), st);
      c->cg_nmax = nmax;
    }
    HIPCHK(hipMemsetAsync(w.K.p, 0, w.K.bytes, st));
    for(int z = 0; z < nb; ++z) {
      const double omega = omx::kC0 * (2.0 * omx::kPi / lam[z]);
      omx::ShArgs sa{P, nmax, omega, bge, bgm, c->d_radius.as<double>(), w.eps.as<cplx>() + (size_t)z * P,
                     c->d_mu.as<cplx>(), w.eps_sh.as<cplx>() + (size_t)z * P, c->d_kppp.as<cplx>(),
                     c->d_kpar.as<cplx>(), c->d_gam.as<cplx>(), w.ov.as<cplx>() + z * ovs, c->d_cg.as<double>(),
                     w.xint.as<cplx>() + z * lda, w.K.as<cplx>() + z * lda, w.K1.as<cplx>() + z * lda};
      omx::run_shsrc(sa, st);
    }
    vq.join();  // V's factorisation enqueued (and its event recorded) before the wait below
    HIPCHK(hipStreamWaitEvent(st, c->xev[1], 0));  // V's factorisation
    refine(true, lwV, FV, w.K.as<cplx>(), w.xsh.as<cplx>());
    for(int z = 0; z < nb; ++z)
      omx::vmul_ov(w.xintsh.as<cplx>() + z * lda, w.xsh.as<cplx>() + z * lda, w.ov.as<cplx>() + z * ovs,
                   omx::OV_IAUXSH1, P, nmax, w.K1.as<cplx>() + z * lda, N, st);
  }
  // cross sections (Result.cpp:557-761)
  for(int z = 0; z < nb; ++z) {
    omx::CrossArgs ca{P, nmax, w.x.as<cplx>() + z * lda, sh ? w.xsh.as<cplx>() + z * lda : nullptr,
                      w.incloc.as<cplx>() + z * lda, w.w.as<double>() + z * ws, w.ov.as<cplx>() + z * ovs,
                      w.cross.as<double>() + (size_t)z * 4 * P};
    omx::run_cross(ca, st);
  }
  // SH absorption (Result.cpp:763-794), every system in one batch
  if(sh) {
    omx::AbsShArgs aa{P, nmax, hom[0], bge, bgm, c->d_radius.as<double>(), w.eps.as<cplx>(), c->d_mu.as<cplx>(),
                      w.eps_sh.as<cplx>(), c->d_gam.as<cplx>(), c->d_cg.as<double>(), w.xint.as<cplx>(),
                      w.xintsh.as<cplx>(), w.abscoef.as<cplx>(), w.abssh.as<double>()};
    aa.bes = w.absbes.as<cplx>();
    aa.omz = w.omz.as<double>();
    aa.nb = nb;
    aa.zvec = lda;
    omx::run_abssh(aa, st);
  }
  std::vector<double> part((size_t)4 * P * nb), pabs((size_t)P * nb, 0.0);
  HIPCHK(hipMemcpyAsync(part.data(), w.cross.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
  if(sh) HIPCHK(hipMemcpyAsync(pabs.data(), w.abssh.p, sizeof(double) * pabs.size(), hipMemcpyDeviceToHost, st));
  auto cp = [&](double *dst, const DBuf &b) {
    if(dst)
      HIPCHK(hipMemcpy2DAsync(dst, sizeof(cplx) * N, b.p, sizeof(cplx) * Np, sizeof(cplx) * N, nb,
                              hipMemcpyDeviceToHost, st));
  };
  cp(x_sca, w.x);
  cp(x_int, w.xint);
  if(sh) {
    cp(x_sca_sh, w.xsh);
    cp(x_int_sh, w.xintsh);
  }
  HIPCHK(hipStreamSynchronize(st));
  if(xs)
    for(int z = 0; z < nb; ++z) {
      double sm[4] = {0, 0, 0, 0};
      for(int j = 0; j < P; ++j)
        for(int e = 0; e < 4; ++e) sm[e] += part[(size_t)z * 4 * P + 4 * j + e];
      const double rk = 2.0 * omx::kPi / lam[z] * bgn.real();
      double sa = 0.0;
      for(int j = 0; j < P; ++j) sa += pabs[(size_t)z * P + j];
      xs[5 * z + 0] = (-1. / (rk * rk)) * sm[0];  // Result.cpp:576
      xs[5 * z + 1] = (1. / (rk * rk)) * sm[1];   // Result.cpp:646
      xs[5 * z + 2] = (1. / (rk * rk)) * sm[2];   // Result.cpp:675
      xs[5 * z + 3] = sh ? (1.0 / (4.0 * c->bg_eps_r.x * c->bg_mu_r.x)) * sm[3] : 0.0;  // Result.cpp:760
      xs[5 * z + 4] = sa;                                                               // Result.cpp:763-794
    }
}

// ------------------------------------------------------------------ no-allocation solve
// omx_dense_solve_noalloc keeps every work array of its factorisation and
// triangular solves in one file-scope device array, so a call allocates
// nothing.  Each device holds its own copy for the life of the process; the
// host learns that copy's address once per device (hipGetSymbolAddress in
// omx_create, with the device current) and passes that address to kernels:
// the array's name in host code is not a device address.
constexpr int kNoAllocMaxN = 4096;  // largest order (and padded order) served
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
struct NoAllocLayout {
  static constexpr int Np = kNoAllocMaxN;
  static constexpr int gmax = Np / omx::kPanelRows + 1;
  static constexpr size_t status = 0;  // 16 ints
  static constexpr size_t piv = status + align256(16 * sizeof(int));  // Np ints
  static constexpr size_t perm = piv + align256(Np * sizeof(int));  // Np ints
  static constexpr size_t cand = perm + align256(Np * sizeof(int));  // 2 x gmax x 64 ints
  static constexpr size_t cnt = cand + align256(2 * gmax * omx::kNB * sizeof(int));  // 2 x gmax ints
  static constexpr size_t map = cnt + align256(2 * gmax * sizeof(int));  // 3 x (Np + 128) ints
  static constexpr size_t wlu = map + align256(3 * (Np + 2 * omx::kNB) * sizeof(int));  // 3 x 64 x 64 complex
  static constexpr size_t dinv = wlu + align256(3 * omx::kNB * omx::kNB * sizeof(cplx));  // 2 x 64 x Np complex
  static constexpr size_t rhs = dinv + align256(2 * omx::kNB * (size_t)Np * sizeof(cplx));  // Np complex
  static constexpr size_t x = rhs + align256(Np * sizeof(cplx));  // Np complex
  static constexpr size_t bytes = x + align256(Np * sizeof(cplx));
};
static_assert(NoAllocLayout::bytes < (9u << 20), "no-allocation scratch stays under 9 MiB");

static __device__ __align__(256) unsigned char scratch[NoAllocLayout::bytes];

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_addr;  // per device, filled once

// Device address of this device's scratch array (the device must be current).
void *scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if((int)g_scratch_addr.size() <= device) g_scratch_addr.resize(device + 1, nullptr);
  if(!g_scratch_addr[device]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if(!p) throw OmxError(OMX_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_addr[device] = p;
  }
  return g_scratch_addr[device];
}

// Clears n ints (a kernel rather than a memset, so the call also suits code
// that records it into a graph).
__global__ void zero_ints_kernel(int *p, int n) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}
// rb[0, N) = b, rb[N, Np) = 0: the right-hand side of the padded system.
__global__ void stage_rhs_kernel(const cplx *b, cplx *rb, int N, int Np) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < Np) rb[i] = i < N ? b[i] : omx::mk(0.0);
}

}  // namespace

extern "C" {

const char *omx_last_error(void) { return g_err.c_str(); }

int omx_create(omx_context **out, int device) {
  return guard([&] {
    if(!out) throw OmxError(OMX_ERR_ARG, "null context pointer");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if(device < 0 || device >= n) throw OmxError(OMX_ERR_DEVICE, "no HIP device " + std::to_string(device));
    auto *c = new omx_context();
    if(const char *e = std::getenv("OMX_SH_OVERLAP")) c->sh_overlap = std::atof(e);
    if(const char *e = std::getenv("OMX_KRYLOV_MARK")) c->krylov_mark = std::atof(e);
    c->device = device;
    try {
      HIPCHK(hipSetDevice(device));
      HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
      HIPCHK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
      HIPCHK(hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking));
      HIPCHK(hipStreamCreateWithFlags(&c->aux2, hipStreamNonBlocking));
      for(auto &e : c->lu_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      for(auto &e : c->lu_ev2) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      for(auto &e : c->xev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      for(auto &e : c->ev) HIPCHK(hipEventCreate(&e));
      c->d_coef.ensure(sizeof(double) * 6 * omx::kCoefStride, &c->total);
      std::vector<double> h(6 * omx::kCoefStride, 0.0);
      omx::fill_coef(h.data());
      HIPCHK(hipMemcpy(c->d_coef.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
      omx::upload_factorials();
      c->scratch = scratch_address(device);
    } catch(...) {
      omx_destroy(c);
      throw;
    }
    *out = c;
  });
}

void omx_destroy(omx_context *c) {
  if(!c) return;
  (void)hipSetDevice(c->device);
  if(c->st) (void)hipStreamSynchronize(c->st);
  if(c->st2) (void)hipStreamSynchronize(c->st2);
  c->fg.reset();
  c->fg_sweep.reset();
  c->d_status2.release(&c->total);
  c->d_map2.release(&c->total);
  DBuf *bufs[] = {&c->d_xyz, &c->d_radius, &c->d_eps, &c->d_mu, &c->d_eps_sh, &c->d_kppp, &c->d_kpar, &c->d_gam,
                  &c->d_sph, &c->d_ov, &c->d_inc, &c->d_incloc, &c->d_Q, &c->d_w, &c->d_coef, &c->d_S64,
                  &c->d_S32, &c->d_V64, &c->d_V32, &c->d_x, &c->d_xsh, &c->d_xint, &c->d_xintsh, &c->d_K,
                  &c->d_K1, &c->d_cross, &c->d_cg, &c->d_piv, &c->d_perm, &c->d_cand, &c->d_cnt, &c->d_map,
                  &c->d_wlu, &c->d_status, &c->d_basis, &c->d_basisZ, &c->d_w1, &c->d_w2, &c->d_r,
                  &c->d_part, &c->d_h, &c->d_nrm, &c->d_gpart, &c->d_piv2, &c->d_perm2,
                  &c->d_dinv, &c->d_dinv2, &c->d_F64, &c->d_F64sh, &c->d_abscoef, &c->d_abssh, &c->d_absbes,
                  &c->d_fpts, &c->d_fout};
  for(DBuf *b : bufs) b->release(&c->total);
  auto &w = c->sw;
  DBuf *sbufs[] = {&w.S64, &w.S32, &w.V64, &w.V32, &w.F64, &w.F64sh, &w.eps, &w.eps_sh, &w.ov, &w.incloc, &w.Q,
                   &w.w, &w.x, &w.xint, &w.xsh, &w.xintsh, &w.K, &w.K1, &w.cross, &w.perm, &w.perm2,
                   &w.dinv, &w.dinv2, &w.map, &w.status, &w.basis, &w.basisZ, &w.r, &w.w1, &w.w2, &w.gpart,
                   &w.part, &w.hd, &w.nrm, &w.sc, &w.act, &w.abscoef, &w.abssh, &w.absbes, &w.kz, &w.omz};
  for(DBuf *b : sbufs) b->release(&c->total);
  c->aca[0].release(&c->total);
  c->aca[1].release(&c->total);
  c->ws_in.release(&c->total);
  DBuf *abufs[] = {&c->d_acaA, &c->d_acaU, &c->d_acaV, &c->d_acaR, &c->d_acaOff, &c->d_plist};
  for(DBuf *b : abufs) b->release(&c->total);
  for(auto &e : c->ev)
    if(e) (void)hipEventDestroy(e);
  for(auto e : c->pev)
    if(e) (void)hipEventDestroy(e);
  for(auto e : c->lu_ev)
    if(e) (void)hipEventDestroy(e);
  for(auto e : c->lu_ev2)
    if(e) (void)hipEventDestroy(e);
  for(auto e : c->xev)
    if(e) (void)hipEventDestroy(e);

  for(auto e : c->pev2)
    if(e) (void)hipEventDestroy(e);
  if(c->aux2) (void)hipStreamDestroy(c->aux2);
  if(c->st2) (void)hipStreamDestroy(c->st2);
  if(c->aux) (void)hipStreamDestroy(c->aux);
  if(c->st) (void)hipStreamDestroy(c->st);
  if(c->h_pin) (void)hipHostFree(c->h_pin);
  delete c;
}

int omx_set_geometry(omx_context *c, int nobj, int nmax, const double *xyz, const double *radius,
                     const double *eps_r, const double *mu_r, const double *eps_r_sh, const double *ksippp,
                     const double *ksiparppar, const double *gamma, const double *bg_eps_r,
                     const double *bg_mu_r) {
  return guard([&] {
    if(!c || nobj < 1 || nmax < 1 || nmax > omx::kMaxNmax || !xyz || !radius || !eps_r || !mu_r)
      throw OmxError(OMX_ERR_ARG, "invalid geometry (1 <= nmax <= 10, nobj >= 1, arrays non-null)");
    HIPCHK(hipSetDevice(c->device));
    // Geometry::pushObject rejects overlapping spheres (Geometry.cpp:39-51)
    for(int i = 0; i < nobj; ++i)
      for(int j = 0; j < i; ++j) {
        double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
        if(std::sqrt(dx * dx + dy * dy + dz * dz) <= radius[i] + radius[j])
          throw OmxError(OMX_ERR_OVERLAP, "spheres " + std::to_string(j) + " and " + std::to_string(i) + " overlap");
      }
    for(int i = 0; i < nobj; ++i)
      if(!(radius[i] > 0.0) || !std::isfinite(eps_r[2 * i]) || !std::isfinite(eps_r[2 * i + 1]))
        throw OmxError(OMX_ERR_ARG, "non-positive radius or non-finite permittivity at sphere " + std::to_string(i));
    c->P = nobj;
    c->nmax = nmax;
    c->nf = nmax * (nmax + 2);
    c->N = 2 * c->nf * nobj;
    c->Np = (c->N + omx::kNB - 1) / omx::kNB * omx::kNB;
    c->xyz.assign(xyz, xyz + 3 * nobj);
    c->radius.assign(radius, radius + nobj);
    c->eps.resize(nobj);
    c->mu.resize(nobj);
    c->eps_sh.resize(nobj);
    c->ksippp.resize(nobj);
    c->ksiparppar.resize(nobj);
    c->gamma.resize(nobj);
    for(int i = 0; i < nobj; ++i) {
      c->eps[i] = kEps0 * C(eps_r, i);
      c->mu[i] = kMu0 * C(mu_r, i);
      c->eps_sh[i] = eps_r_sh ? kEps0 * C(eps_r_sh, i) : omx::mk(kEps0);
      c->ksippp[i] = ksippp ? C(ksippp, i) : omx::mk(0.0);
      c->ksiparppar[i] = ksiparppar ? C(ksiparppar, i) : omx::mk(0.0);
      c->gamma[i] = gamma ? C(gamma, i) : omx::mk(0.0);
    }
    c->bg_eps_r = bg_eps_r ? C(bg_eps_r, 0) : omx::mk(1.0);
    c->bg_mu_r = bg_mu_r ? C(bg_mu_r, 0) : omx::mk(1.0);
    c->have_geom = true;
    c->updated = c->solved = false;
  });
}

int omx_set_excitation(omx_context *c, double wavelength, double theta, double phi, const double *etheta,
                       const double *ephi, int sh) {
  return guard([&] {
    if(!c || !(wavelength > 0.0)) throw OmxError(OMX_ERR_ARG, "invalid excitation");
    c->lambda = wavelength;
    c->theta = theta;
    c->phi = phi;
    c->eth = etheta ? C(etheta, 0) : omx::mk(1.0);
    c->eph = ephi ? C(ephi, 0) : omx::mk(0.0);
    c->sh = sh ? 1 : 0;
    c->have_exc = true;
    c->updated = c->solved = false;
  });
}

int omx_set_solver(omx_context *c, int kind, double tol, int maxit, int restart) {
  return guard([&] {
    if(!c || kind < 0 || kind > 5 || !(tol > 0.0) || maxit < 1 || restart < 1 || restart > 256)
      throw OmxError(OMX_ERR_ARG, "invalid solver settings");
    c->kind = c->kind_req = kind;
    c->tol = tol;
    c->maxit = maxit;
    c->restart = restart;
    c->updated = c->solved = false;
  });
}

int omx_set_auto_budget(omx_context *c, int steps) {
  return guard([&] {
    if(!c || steps < 0) throw OmxError(OMX_ERR_ARG, "invalid budget");
    c->auto_budget = steps;
  });
}

int omx_set_graphs(omx_context *c, int on) {
  return guard([&] {
    if(!c) throw OmxError(OMX_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    c->graphs = on != 0;
    if(!c->graphs) {
      HIPCHK(hipStreamSynchronize(c->st));
      HIPCHK(hipStreamSynchronize(c->st2));
      c->fg.reset();
      c->fg_sweep.reset();
    }
  });
}

int omx_set_autotune(omx_context *c, int on) {
  return guard([&] {
    if(!c) throw OmxError(OMX_ERR_ARG, "null context");
    c->autotune = on != 0;
  });
}

int omx_tune_query(int device, int Np, int nb, int fp64, int *split, int *trsv, double *times) {
  return guard([&] {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(std::make_tuple(device, Np, nb, fp64 ? 1 : 0));
    if(it == g_tune.end()) throw OmxError(OMX_ERR_STATE, "shape not tuned");
    const TuneChoice &ch = it->second;
    if(split) *split = ch.split ? 1 : 0;
    if(trsv) *trsv = ch.trsv;
    if(times) {
      times[0] = ch.t_prod[0];
      times[1] = ch.t_prod[1];
      for(int v = 0; v < omx::kTrsvVariants; ++v) times[2 + v] = ch.t_trsv[v];
    }
  });
}

int omx_tune_reset(void) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune.clear();
  return OMX_OK;
}

int omx_set_pivoting(omx_context *c, int mode) {
  return guard([&] {
    if(!c || (mode != OMX_PIVOT_BLOCK && mode != OMX_PIVOT_PARTIAL)) throw OmxError(OMX_ERR_ARG, "invalid pivoting");
    c->pivot = mode;
    c->factored = c->factored_sh = false;
    c->solved = false;
  });
}

int omx_set_lu32_gemm(omx_context *c, int mode) {
  return guard([&] {
    if(!c || (mode != OMX_GEMM_BF16X3 && mode != OMX_GEMM_FP32)) throw OmxError(OMX_ERR_ARG, "invalid gemm mode");
    c->lu32_gemm = mode;
    c->factored = c->factored_sh = false;
    c->solved = false;
  });
}

int omx_set_cg_tables(omx_context *c, int nmax, const double *const *tables) {
  return guard([&] {
    if(!c || nmax < 1 || nmax > omx::kMaxNmax || !tables) throw OmxError(OMX_ERR_ARG, "invalid tables");
    HIPCHK(hipSetDevice(c->device));
    size_t nf = (size_t)nmax * (nmax + 2), sz = nf * nf * nf;
    c->d_cg.ensure(sizeof(double) * 9 * sz, &c->total);
    for(int t = 0; t < 9; ++t) {
      if(!tables[t]) throw OmxError(OMX_ERR_ARG, "null table");
      HIPCHK(hipMemcpyAsync(c->d_cg.as<double>() + t * sz, tables[t], sizeof(double) * sz, hipMemcpyHostToDevice,
                            c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->cg_nmax = nmax;
  });
}

int omx_build_cg_tables(omx_context *c, int nmax) {
  return guard([&] {
    if(!c || nmax < 1 || nmax > omx::kMaxNmax) throw OmxError(OMX_ERR_ARG, "invalid nmax");
    HIPCHK(hipSetDevice(c->device));
    size_t nf = (size_t)nmax * (nmax + 2), sz = nf * nf * nf;
    c->d_cg.ensure(sizeof(double) * 9 * sz, &c->total);
    omx::run_cg(nmax, c->d_cg.as<double>(), c->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    c->cg_nmax = nmax;
  });
}

int omx_update(omx_context *c) {
  return guard([&] {
    if(!c || !c->have_geom || !c->have_exc) throw OmxError(OMX_ERR_STATE, "geometry and excitation must be set");
    HIPCHK(hipSetDevice(c->device));
    if(c->kind_req == OMX_SOLVER_AUTO)
      c->kind = c->N >= kAutoKrylovN ? OMX_SOLVER_GMRES32_IR : OMX_SOLVER_LU32_IR;
    c->krylov_s = c->kind_req == OMX_SOLVER_AUTO && c->kind == OMX_SOLVER_LU32_IR && c->sh;
    alloc_all(c);
    alloc_matrices(c);
    c->tuned = false;
    if(c->autotune && is_lu(c->kind)) {
      const TuneChoice t = single_tune(c, c->kind == OMX_SOLVER_LU64);
      c->lu.trsv = t.trsv;
      if(c->sh) c->lu_sh.trsv = t.trsv;
      c->tuned = true;
      c->tuned_split = t.split;
    }
    const int P = c->P, nmax = c->nmax;
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    HIPCHK(hipMemcpyAsync(c->d_xyz.p, c->xyz.data(), sizeof(double) * 3 * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_radius.p, c->radius.data(), sizeof(double) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_eps.p, c->eps.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_mu.p, c->mu.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_eps_sh.p, c->eps_sh.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_kppp.p, c->ksippp.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_kpar.p, c->ksiparppar.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_gam.p, c->gamma.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    const double omega = c->omega();
    const cplx bge = kEps0 * c->bg_eps_r, bgm = kMu0 * c->bg_mu_r;
    omx::ObjArgs oa{P, nmax, omega, bge, bgm, c->d_xyz.as<double>(), c->d_radius.as<double>(),
                    c->d_eps.as<cplx>(), c->d_mu.as<cplx>(), c->d_eps_sh.as<cplx>(), c->d_sph.as<double>(),
                    c->d_ov.as<cplx>()};
    omx::run_objects(oa, c->st);
    omx::IncArgs ia{nmax, c->theta, c->phi, c->eth, c->eph, c->d_inc.as<cplx>()};
    omx::run_incident(ia, c->st);
    HIPCHK(hipMemsetAsync(c->d_Q.p, 0, sizeof(cplx) * c->Np, c->st));
    omx::OriginArgs og{P, nmax, c->waveK(), c->d_sph.as<double>(), c->d_ov.as<cplx>(), c->d_coef.as<double>(),
                       c->d_inc.as<cplx>(), c->d_incloc.as<cplx>(), c->d_Q.as<cplx>(), c->d_w.as<double>()};
    omx::run_origin(og, c->st);
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    const size_t lda = (size_t)c->Np;
    if(c->kind == OMX_SOLVER_ACA_GMRES) {
      aca_build(c, false);
      if(c->sh) aca_build(c, true); else c->aca[1].release(&c->total);
    } else {
      c->aca[0].release(&c->total);
      c->aca[1].release(&c->total);
    }
    omx::PairArgs pa{P, nmax, omx::OV_T, c->waveK(), c->d_sph.as<double>(), c->d_ov.as<cplx>(),
                     c->d_coef.as<double>(), lda, c->d_S64.as<cplx>(), c->d_S32.as<float2>()};
    if(c->kind != OMX_SOLVER_ACA_GMRES) omx::run_pairs(pa, c->st);
    if(c->d_S64.p) omx::run_pad64(c->d_S64.as<cplx>(), c->N, c->Np, lda, c->st);
    if(c->d_S32.p) omx::run_pad32(c->d_S32.as<float2>(), c->N, c->Np, lda, c->st);
    if(c->sh && c->kind != OMX_SOLVER_ACA_GMRES) {
      omx::PairArgs pv = pa;
      pv.tkind = omx::OV_TSH;
      cplx k = c->waveK();
      pv.k = omx::mk(2.0 * k.x, 2.0 * k.y);  // PreconditionedMatrix.cpp:592
      pv.S64 = c->d_V64.as<cplx>();
      pv.S32 = c->d_V32.as<float2>();
      omx::run_pairs(pv, c->st);
      if(c->d_V64.p) omx::run_pad64(c->d_V64.as<cplx>(), c->N, c->Np, lda, c->st);
      if(c->d_V32.p) omx::run_pad32(c->d_V32.as<float2>(), c->N, c->Np, lda, c->st);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[2], c->st));
    HIPCHK(hipEventSynchronize(c->ev[2]));
    c->stats.t_update_ms = elapsed(c, 0, 2);
    c->stats.t_assemble_ms = elapsed(c, 1, 2);
    c->stats.N = c->N;
    c->stats.Np = c->Np;
    c->stats.device_bytes = c->total;
    c->updated = true;
    c->solved = false;
    c->factored = c->factored_sh = false;
    c->dirty32 = c->dirty32_sh = false;
  });
}

// The nine Clebsch-Gordan-type tables of the SH source for the context's
// nmax (Symbol.cpp:1036-1451), built once per nmax.
void ensure_cg(omx_context *c, hipStream_t st) {
  if(c->cg_nmax == c->nmax) return;
  size_t nf = (size_t)c->nf, sz = nf * nf * nf;
  c->d_cg.ensure(sizeof(double) * 9 * sz, &c->total);
  omx::run_cg(c->nmax, c->d_cg.as<double>(), st);
  c->cg_nmax = c->nmax;
}

// The second-harmonic source K and K1 of the context's x_int
// (PreconditionedMatrix.cpp:1347-1436, Geometry.cpp:250-310), on stream st.
void sh_source(omx_context *c, hipStream_t st) {
  HIPCHK(hipMemsetAsync(c->d_K.p, 0, sizeof(cplx) * c->Np, st));
  omx::ShArgs sa{c->P, c->nmax, c->omega(), kEps0 * c->bg_eps_r, kMu0 * c->bg_mu_r, c->d_radius.as<double>(),
                 c->d_eps.as<cplx>(), c->d_mu.as<cplx>(), c->d_eps_sh.as<cplx>(), c->d_kppp.as<cplx>(),
                 c->d_kpar.as<cplx>(), c->d_gam.as<cplx>(), c->d_ov.as<cplx>(), c->d_cg.as<double>(),
                 c->d_xint.as<cplx>(), c->d_K.as<cplx>(), c->d_K1.as<cplx>()};
  omx::run_shsrc(sa, st);
}

int omx_solve(omx_context *c, double *x_sca, double *x_int, double *x_sca_sh, double *x_int_sh) {
  return guard([&] {
    if(!c || !c->updated) throw OmxError(OMX_ERR_STATE, "omx_update must precede omx_solve");
    HIPCHK(hipSetDevice(c->device));
    const int N = c->N, P = c->P, nmax = c->nmax;
    double tf = 0, ts = 0, tsh = 0;
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    // V does not depend on the fundamental solution, only its right-hand
    // side does: its factorisation runs on the second stream pair beside the
    // fundamental solve (not while profiling, which times each factorisation
    // alone)
    const bool lu_kind = is_lu(c->kind);
    const bool graphed =
        c->graphs && lu_kind && !c->profile && c->pivot == OMX_PIVOT_BLOCK && !c->factored && !c->krylov_s;
    SideEnqueue vq;
    if(c->sh && lu_kind && !c->factored_sh && !c->profile) {
      // V's factorisation starts once S's has passed a fraction of its
      // columns (by then S's trailing updates no longer fill the device),
      // and overlaps the rest of it and the fundamental refinement
      // (V's work is enqueued right at the mark, in the middle of S's
      // enqueueing: a wait on the mark event enqueued only after all of S's
      // work was seen to hold V back until S's end)
      const int mark = c->krylov_s ? 0 : (int)(c->sh_overlap * c->Np);
      auto start_v = [](void *arg) {
        omx_context *cc = static_cast<omx_context *>(arg);
        HIPCHK(hipStreamWaitEvent(cc->st2, cc->xev[0], 0));
        factor_enqueue(cc, true, cc->pivot);
      };
      if(mark > 0 && graphed) {
        factor_graphed(c, true, mark);
      } else if(mark > 0 && !c->factored) {
        c->lu.ev_mark = c->xev[0];
        c->lu.mark_col = mark;
        c->lu.on_mark = start_v;
        c->lu.mark_arg = c;
        try {
          factor_enqueue(c, false, c->pivot);
        } catch(...) {
          c->lu.ev_mark = nullptr;
          c->lu.on_mark = nullptr;
          throw;
        }
        c->lu.ev_mark = nullptr;
        c->lu.on_mark = nullptr;
      } else if(c->krylov_s && !c->factored) {
        // S's GMRES starts once V's factorisation has passed krylov_mark of
        // its columns: the first columns carry most of the bulk-update work,
        // which the GMRES products would only slow down, and the panel-bound
        // rest leaves the device room for them
        HIPCHK(hipEventRecord(c->xev[0], c->st));
        const int kmark = (int)(c->krylov_mark * c->Np);
        if(kmark > 0) {
          c->lu_sh.ev_mark = c->xev[3];
          c->lu_sh.mark_col = kmark;
          c->lu_sh.on_mark = nullptr;
        }
        vq.start(c, start_v);
        if(kmark > 0) {
          try {
            vq.join();  // V's factorisation enqueued, with its mark recorded
          } catch(...) {
            c->lu_sh.ev_mark = nullptr;
            throw;
          }
          c->lu_sh.ev_mark = nullptr;
          HIPCHK(hipStreamWaitEvent(c->st, c->xev[3], 0));
        }
      } else {
        HIPCHK(hipEventRecord(c->xev[0], c->st));
        start_v(c);
      }
    }
    if(graphed && !c->factored) factor_graphed(c, false, 0);
    // X_sca = S^{-1} Q (PreconditionedMatrixSolver.h:58)
    solve_system(c, false, c->d_Q.as<cplx>(), c->d_x.as<cplx>(), &c->stats.iterations, &c->stats.residual);
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    // X_int = X_sca (.) Iaux (Solver.cpp:57-77)
    omx::vmul_ov(c->d_xint.as<cplx>(), c->d_x.as<cplx>(), c->d_ov.as<cplx>(), omx::OV_IAUX, P, nmax, nullptr, N, c->st);
    HIPCHK(hipEventRecord(c->ev[2], c->st));
    HIPCHK(hipEventSynchronize(c->ev[2]));
    ts += elapsed(c, 0, 2);
    if(c->sh) {
      ensure_cg(c, c->st);
      HIPCHK(hipEventRecord(c->ev[3], c->st));
      sh_source(c, c->st);
      HIPCHK(hipEventRecord(c->ev[4], c->st));
      // the second-harmonic solve runs on the second stream, after the source
      // (and after V's factorisation has been enqueued there)
      HIPCHK(hipEventRecord(c->xev[1], c->st));
      vq.join();
      HIPCHK(hipStreamWaitEvent(c->st2, c->xev[1], 0));
      solve_system(c, true, c->d_K.as<cplx>(), c->d_xsh.as<cplx>(), &c->stats.iterations_sh, &c->stats.residual_sh);
      // X_int_SH = X_sca_SH (.) IauxSH1 - K1ana (Solver.cpp:95-116)
      omx::vmul_ov(c->d_xintsh.as<cplx>(), c->d_xsh.as<cplx>(), c->d_ov.as<cplx>(), omx::OV_IAUXSH1, P, nmax,
                   c->d_K1.as<cplx>(), N, c->st2);
      HIPCHK(hipEventRecord(c->xev[2], c->st2));
      HIPCHK(hipStreamWaitEvent(c->st, c->xev[2], 0));
      HIPCHK(hipEventRecord(c->ev[5], c->st));
      HIPCHK(hipEventSynchronize(c->ev[5]));
      tsh = elapsed(c, 3, 4);
      ts += elapsed(c, 4, 5);
    }
    HIPCHK(hipGetLastError());
    c->stats.t_factor_ms = tf;
    c->stats.t_solve_ms = ts;
    c->stats.t_sh_source_ms = tsh;
    auto cp = [&](double *dst, const DBuf &b) {
      if(dst) HIPCHK(hipMemcpyAsync(dst, b.p, sizeof(cplx) * N, hipMemcpyDeviceToHost, c->st));
    };
    cp(x_sca, c->d_x);
    cp(x_int, c->d_xint);
    if(c->sh) {
      cp(x_sca_sh, c->d_xsh);
      cp(x_int_sh, c->d_xintsh);
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->solved = true;
  });
}

// ------------------------------------------------------------------ groups
// Batch buffers of a group solve in the leader context: the sweep's buffers
// that the batched factorisation, triangular solves and lockstep refinement
// use, at the group's padded order L->Np.
void group_alloc(omx_context *L, int nb, bool f64, bool sh) {
  auto &w = L->sw;
  size_t *T = &L->total;
  const size_t Np = L->Np, n2 = Np * Np;
  const int m = L->restart;
  w.S64.ensure(sizeof(cplx) * n2 * nb, T);
  if(f64) w.F64.ensure(sizeof(cplx) * n2 * nb, T); else w.S32.ensure(sizeof(float2) * n2 * nb, T);
  if(sh) {
    w.V64.ensure(sizeof(cplx) * n2 * nb, T);
    if(f64) w.F64sh.ensure(sizeof(cplx) * n2 * nb, T); else w.V32.ensure(sizeof(float2) * n2 * nb, T);
  }
  for(DBuf *b : {&w.Q, &w.x, &w.xsh, &w.K, &w.r, &w.w1, &w.w2}) b->ensure(sizeof(cplx) * Np * nb, T);
  for(DBuf *b : {&w.perm, &w.perm2}) b->ensure(sizeof(int) * Np * nb, T);
  for(DBuf *b : {&w.dinv, &w.dinv2}) b->ensure(sizeof(cplx) * omx::dinv_elems(Np) * nb, T);
  for(DBuf *b : {&w.map, &w.map2}) b->ensure(sizeof(int) * 3 * (Np + 2 * omx::kNB) * nb, T);
  for(DBuf *b : {&w.status, &w.status2}) b->ensure(sizeof(int) * 4 * nb, T);
  w.sys.ensure(sizeof(int) * nb, T);
  w.basis.ensure(sizeof(cplx) * Np * (m + 1) * nb, T);
  w.basisZ.ensure(sizeof(cplx) * Np * (m + 1) * nb, T);
  w.gpart.ensure(sizeof(cplx) * Np * omx::kGemvSplitMax * nb, T);
  w.part.ensure(sizeof(cplx) * 128 * (m + 2) * nb, T);
  w.hd.ensure(sizeof(cplx) * (m + 2) * nb, T);
  w.nrm.ensure(sizeof(double) * 160 * nb, T);
  w.sc.ensure(sizeof(cplx) * 2 * nb, T);
  w.act.ensure(sizeof(int) * nb, T);
  L->d_status.ensure(64, T);
}

int omx_solve_group(omx_context *L, int n, omx_context *const *mem) {
  return guard([&] {
    if(!L || n < 1 || !mem) throw OmxError(OMX_ERR_ARG, "omx_solve_group: leader, count and members required");
    for(int i = 0; i < n; ++i) {
      omx_context *m = mem[i];
      if(!m || m == L) throw OmxError(OMX_ERR_ARG, "omx_solve_group: member " + std::to_string(i) + " invalid");
      if(!m->updated) throw OmxError(OMX_ERR_STATE, "omx_solve_group: member " + std::to_string(i) + " not updated");
      if(m->device != L->device) throw OmxError(OMX_ERR_ARG, "omx_solve_group: members must share the leader's device");
      if(m->kind != mem[0]->kind || (m->kind != OMX_SOLVER_LU64 && m->kind != OMX_SOLVER_LU32_IR))
        throw OmxError(OMX_ERR_ARG, "omx_solve_group: members need one LU kind (lu64 or lu32ir)");
      if((m->sh != 0) != (mem[0]->sh != 0)) throw OmxError(OMX_ERR_ARG, "omx_solve_group: mixed SH flags");
      for(int j = 0; j < i; ++j)
        if(mem[j] == m) throw OmxError(OMX_ERR_ARG, "omx_solve_group: a member appears twice");
    }
    HIPCHK(hipSetDevice(L->device));
    const bool f64 = mem[0]->kind == OMX_SOLVER_LU64, sh = mem[0]->sh != 0;
    int Np = 0;
    for(int i = 0; i < n; ++i) Np = std::max(Np, mem[i]->Np);
    // the leader runs the batch at the largest padded order with the first
    // member's refinement settings, in its sweep buffers only; its own
    // sizes and settings come back on every exit
    struct Restore {
      omx_context &c;
      int N, Np, kind, sh, maxit, restart, gemm;
      double tol;
      ~Restore() {
        c.N = N;
        c.Np = Np;
        c.kind = kind;
        c.sh = sh;
        c.maxit = maxit;
        c.restart = restart;
        c.lu32_gemm = gemm;
        c.tol = tol;
      }
    } restore{*L, L->N, L->Np, L->kind, L->sh, L->maxit, L->restart, L->lu32_gemm, L->tol};
    L->Np = Np;
    L->N = Np;
    L->kind = mem[0]->kind;
    L->sh = sh;
    L->tol = mem[0]->tol;
    L->maxit = mem[0]->maxit;
    L->restart = mem[0]->restart;
    L->lu32_gemm = mem[0]->lu32_gemm;
    const int nb = n;
    const size_t n2 = (size_t)Np * Np, lda = (size_t)Np;
    group_alloc(L, nb, f64, sh);
    const TuneChoice tune = L->autotune ? sweep_tune(L, nb, f64) : TuneChoice{L->lu32_gemm == OMX_GEMM_BF16X3, 0};
    auto &w = L->sw;
    hipStream_t st = L->st;
    for(DBuf *b : {&w.status, &w.status2, &w.basisZ, &w.w1}) HIPCHK(hipMemsetAsync(b->p, 0, b->bytes, st));
    HIPCHK(hipMemsetAsync(L->d_status.p, 0, 16, st));
    // the members' matrices and right-hand sides, bordered by the identity
    cplx *S64 = w.S64.as<cplx>(), *V64 = sh ? w.V64.as<cplx>() : nullptr;
    for(int z = 0; z < nb; ++z) {
      omx_context *m = mem[z];
      omx::pad_matrix64(S64 + z * n2, Np, m->d_S64.as<cplx>(), m->Np, st);
      if(f64) omx::pad_matrix64(w.F64.as<cplx>() + z * n2, Np, m->d_S64.as<cplx>(), m->Np, st);
      else omx::pad_matrix32(w.S32.as<float2>() + z * n2, Np, m->d_S64.as<cplx>(), m->Np, st);
      if(sh) {
        omx::pad_matrix64(V64 + z * n2, Np, m->d_V64.as<cplx>(), m->Np, st);
        if(f64) omx::pad_matrix64(w.F64sh.as<cplx>() + z * n2, Np, m->d_V64.as<cplx>(), m->Np, st);
        else omx::pad_matrix32(w.V32.as<float2>() + z * n2, Np, m->d_V64.as<cplx>(), m->Np, st);
      }
      omx::pad_vector(w.Q.as<cplx>() + z * lda, Np, m->d_Q.as<cplx>(), m->N, st);
    }
    HIPCHK(hipGetLastError());
    void *FS = f64 ? w.F64.p : w.S32.p, *FV = sh ? (f64 ? w.F64sh.p : w.V32.p) : nullptr;
    auto factor = [&](void *F, omx::LuWork &lw, hipStream_t fst) {
      const int rc = f64 ? omx::lu_factor<double2>(static_cast<double2 *>(F), Np, lda, lw, fst, nullptr)
                         : omx::lu_factor<float2>(static_cast<float2 *>(F), Np, lda, lw, fst, nullptr);
      if(rc) throw OmxError(OMX_ERR_DEVICE, "batched lu_factor launch failed (" + std::to_string(rc) + ")");
    };
    omx::LuWork lwS = sweep_lu(L, nb, false), lwV{};
    lwS.split = tune.split;
    lwS.trsv = tune.trsv;
    if(sh) {
      lwV = sweep_lu(L, nb, true);
      lwV.split = tune.split;
      lwV.trsv = tune.trsv;
      HIPCHK(hipEventRecord(L->xev[0], st));
      HIPCHK(hipStreamWaitEvent(L->st2, L->xev[0], 0));
      factor(FV, lwV, L->st2);
      HIPCHK(hipEventRecord(L->xev[1], L->st2));
    }
    factor(FS, lwS, st);
    std::vector<int> bad(nb, 0), iters(2 * nb, 0);
    std::vector<double> resid(2 * nb, 0.0);
    auto status_scan = [&](bool second) {
      std::vector<int> h(4 * nb);
      HIPCHK(hipMemcpyAsync(h.data(), (second ? w.status2 : w.status).p, sizeof(int) * 4 * nb,
                            hipMemcpyDeviceToHost, st));
      int g[2];
      HIPCHK(hipMemcpyAsync(g, L->d_status.p, sizeof(g), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      if(h[1] || g[1]) throw OmxError(OMX_ERR_DEVICE, "device wait timed out in a group solve");
      for(int z = 0; z < nb; ++z)
        if(h[4 * z]) bad[z] = 1;
    };
    auto refine = [&](bool second, omx::LuWork &lw, void *F, const cplx *A64, const cplx *rhs, cplx *x) {
      status_scan(second);
      omx::LuWork *lwp = &lw;
      GmresB gm{L, A64, n2, Np, nb, nullptr, w};
      gm.lw = lwp;
      if(f64)
        gm.prec = [L, F, lwp, Np, lda](const cplx *in, cplx *out) {
          omx::lu_solve<double2>(static_cast<const double2 *>(F), Np, lda, *lwp, in, out, L->st);
        };
      else
        gm.prec = [L, F, lwp, Np, lda](const cplx *in, cplx *out) {
          omx::lu_solve<float2>(static_cast<const float2 *>(F), Np, lda, *lwp, in, out, L->st);
        };
      gm.prec(rhs, x);
      std::vector<double> res;
      std::vector<int> its;
      const int bmax = nb > 1 ? std::min(L->maxit, 4 * L->restart) : L->maxit;
      gm.solve(rhs, x, L->tol, bmax, L->restart, res, its);
      status_scan(second);
      for(int z = 0; z < nb; ++z) {
        if(!(res[z] <= 1e3 * L->tol)) bad[z] = 1;
        iters[2 * z + (second ? 1 : 0)] = its[z];
        resid[2 * z + (second ? 1 : 0)] = res[z];
      }
    };
    // fundamental: S x = Q for every member, then each member's x and x_int
    refine(false, lwS, FS, S64, w.Q.as<cplx>(), w.x.as<cplx>());
    for(int z = 0; z < nb; ++z) {
      omx_context *m = mem[z];
      HIPCHK(hipMemcpyAsync(m->d_x.p, w.x.as<cplx>() + z * lda, sizeof(cplx) * m->Np, hipMemcpyDeviceToDevice, st));
      omx::vmul_ov(m->d_xint.as<cplx>(), m->d_x.as<cplx>(), m->d_ov.as<cplx>(), omx::OV_IAUX, m->P, m->nmax, nullptr,
                   m->N, st);
    }
    if(sh) {
      // each member's SH source from its own x_int, then V x = K in one batch
      for(int z = 0; z < nb; ++z) {
        omx_context *m = mem[z];
        ensure_cg(m, st);
        sh_source(m, st);
        omx::pad_vector(w.K.as<cplx>() + z * lda, Np, m->d_K.as<cplx>(), m->N, st);
      }
      HIPCHK(hipStreamWaitEvent(st, L->xev[1], 0));  // V's factorisation
      refine(true, lwV, FV, V64, w.K.as<cplx>(), w.xsh.as<cplx>());
      for(int z = 0; z < nb; ++z) {
        omx_context *m = mem[z];
        HIPCHK(hipMemcpyAsync(m->d_xsh.p, w.xsh.as<cplx>() + z * lda, sizeof(cplx) * m->Np, hipMemcpyDeviceToDevice,
                              st));
        omx::vmul_ov(m->d_xintsh.as<cplx>(), m->d_xsh.as<cplx>(), m->d_ov.as<cplx>(), omx::OV_IAUXSH1, m->P, m->nmax,
                     m->d_K1.as<cplx>(), m->N, st);
      }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipStreamSynchronize(L->st2));
    for(int z = 0; z < nb; ++z) {
      omx_context *m = mem[z];
      if(bad[z]) {
        // zero pivot in the block or a refinement short of the target: this
        // member alone, with the single-system path and its fallbacks
        if(omx_solve(m, nullptr, nullptr, nullptr, nullptr) != OMX_OK)
          throw OmxError(OMX_ERR_NOCONV, "omx_solve_group: member " + std::to_string(z) + ": " + omx_last_error());
        ++L->fallbacks;
        continue;
      }
      m->stats.iterations = iters[2 * z];
      m->stats.iterations_sh = iters[2 * z + 1];
      m->stats.residual = resid[2 * z];
      m->stats.residual_sh = resid[2 * z + 1];
      m->solved = true;
    }
    L->stats.sweep_batch = nb;
  });
}

int omx_read_solution(omx_context *c, double *x_sca, double *x_int, double *x_sca_sh, double *x_int_sh) {
  return guard([&] {
    if(!c || !c->solved) throw OmxError(OMX_ERR_STATE, "no solution: omx_solve or omx_solve_group first");
    HIPCHK(hipSetDevice(c->device));
    auto cp = [&](double *dst, const DBuf &b) {
      if(dst) HIPCHK(hipMemcpyAsync(dst, b.p, sizeof(cplx) * c->N, hipMemcpyDeviceToHost, c->st));
    };
    cp(x_sca, c->d_x);
    cp(x_int, c->d_xint);
    if(c->sh) {
      cp(x_sca_sh, c->d_xsh);
      cp(x_int_sh, c->d_xintsh);
    }
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int omx_cross_sections(omx_context *c, double *out4) {
  return guard([&] {
    if(!c || !c->solved || !out4) throw OmxError(OMX_ERR_STATE, "omx_solve must precede omx_cross_sections");
    HIPCHK(hipSetDevice(c->device));
    omx::CrossArgs ca{c->P, c->nmax, c->d_x.as<cplx>(), c->sh ? c->d_xsh.as<cplx>() : nullptr,
                      c->d_incloc.as<cplx>(), c->d_w.as<double>(), c->d_ov.as<cplx>(), c->d_cross.as<double>()};
    omx::run_cross(ca, c->st);
    std::vector<double> part(4 * c->P);
    HIPCHK(hipMemcpyAsync(part.data(), c->d_cross.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double s[4] = {0, 0, 0, 0};
    for(int j = 0; j < c->P; ++j)
      for(int e = 0; e < 4; ++e) s[e] += part[4 * j + e];
    double rk = c->waveK().x;
    out4[0] = (-1. / (rk * rk)) * s[0];  // Result.cpp:576
    out4[1] = (1. / (rk * rk)) * s[1];   // Result.cpp:646
    out4[2] = (1. / (rk * rk)) * s[2];   // Result.cpp:675
    out4[3] = c->sh ? (1.0 / (4.0 * c->bg_eps_r.x * c->bg_mu_r.x)) * s[3] : 0.0;  // Result.cpp:760
  });
}

int omx_absorption_cross_section_sh(omx_context *c, double *out) {
  return guard([&] {
    if(!c || !c->solved || !out) throw OmxError(OMX_ERR_STATE, "omx_solve must precede omx_absorption_cross_section_sh");
    if(!c->sh || c->cg_nmax != c->nmax) throw OmxError(OMX_ERR_STATE, "no second-harmonic solution");
    HIPCHK(hipSetDevice(c->device));
    omx::AbsShArgs aa{c->P, c->nmax, c->omega(), kEps0 * c->bg_eps_r, kMu0 * c->bg_mu_r, c->d_radius.as<double>(),
                      c->d_eps.as<cplx>(), c->d_mu.as<cplx>(), c->d_eps_sh.as<cplx>(), c->d_gam.as<cplx>(),
                      c->d_cg.as<double>(), c->d_xint.as<cplx>(), c->d_xintsh.as<cplx>(), c->d_abscoef.as<cplx>(),
                      c->d_abssh.as<double>()};
    aa.bes = c->d_absbes.as<cplx>();
    omx::run_abssh(aa, c->st);
    HIPCHK(hipGetLastError());
    std::vector<double> part(c->P);
    HIPCHK(hipMemcpyAsync(part.data(), c->d_abssh.p, sizeof(double) * c->P, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double s = 0.0;
    for(double v : part) s += v;
    *out = s;
  });
}

int omx_fields(omx_context *c, int npts, const double *xyz, double *E, double *H, double *E_sh, double *H_sh) {
  return guard([&] {
    if(!c || !c->solved) throw OmxError(OMX_ERR_STATE, "omx_solve must precede omx_fields");
    if(npts < 1 || !xyz || !E || !H) throw OmxError(OMX_ERR_ARG, "omx_fields: points and E, H outputs required");
    const bool sh = E_sh || H_sh;
    if(sh && (!E_sh || !H_sh)) throw OmxError(OMX_ERR_ARG, "omx_fields: E_sh and H_sh go together");
    if(sh && (!c->sh || c->cg_nmax != c->nmax)) throw OmxError(OMX_ERR_STATE, "no second-harmonic solution");
    HIPCHK(hipSetDevice(c->device));
    const int chunk = 1 << 16;
    const int cap = std::min(npts, chunk);
    c->d_fpts.ensure(sizeof(double) * 3 * cap, &c->total);
    c->d_fout.ensure(sizeof(cplx) * 12 * cap, &c->total);
    for(int i0 = 0; i0 < npts; i0 += chunk) {
      const int n = std::min(chunk, npts - i0);
      cplx *o = c->d_fout.as<cplx>();
      HIPCHK(hipMemcpyAsync(c->d_fpts.p, xyz + 3 * (size_t)i0, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->st));
      omx::FieldArgs fa{c->P, c->nmax, sh ? 1 : 0, n, c->omega(), c->waveK(), kEps0 * c->bg_eps_r, kMu0 * c->bg_mu_r,
                        c->d_xyz.as<double>(), c->d_radius.as<double>(), c->d_eps.as<cplx>(), c->d_mu.as<cplx>(),
                        c->d_eps_sh.as<cplx>(), c->d_gam.as<cplx>(), sh ? c->d_cg.as<double>() : nullptr,
                        c->d_inc.as<cplx>(), c->d_x.as<cplx>(), c->d_xint.as<cplx>(), c->d_xsh.as<cplx>(),
                        c->d_xintsh.as<cplx>(), c->d_fpts.as<double>(), o, o + 3 * (size_t)cap, o + 6 * (size_t)cap,
                        o + 9 * (size_t)cap};
      omx::run_fields(fa, c->st);
      HIPCHK(hipGetLastError());
      const size_t b = sizeof(cplx) * 3 * n, off = 6 * (size_t)i0;
      HIPCHK(hipMemcpyAsync(E + off, o, b, hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipMemcpyAsync(H + off, o + 3 * (size_t)cap, b, hipMemcpyDeviceToHost, c->st));
      if(sh) {
        HIPCHK(hipMemcpyAsync(E_sh + off, o + 6 * (size_t)cap, b, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(H_sh + off, o + 9 * (size_t)cap, b, hipMemcpyDeviceToHost, c->st));
      }
      HIPCHK(hipStreamSynchronize(c->st));
    }
  });
}

int omx_scattering_size(const omx_context *c) { return c ? c->N : 0; }

int omx_get_stats(const omx_context *c, omx_stats *out) {
  if(!c || !out) return OMX_ERR_ARG;
  *out = c->stats;
  out->device_bytes = c->total;
  out->t_lu_gemm_ms = c->prof_gemm_ms;
  out->lu_gemm_flop = c->prof_flop;
  out->lu_gemm_launches = c->prof_launches;
  out->t_lu_panel_ms = c->prof_panel_ms;
  out->t_factor_total_ms = c->prof_total_ms;
  out->pivot_fallbacks = c->fallbacks;
  out->matvec_calls = (int)c->mv_calls;
  out->t_matvec_ms = c->mv_ms;
  out->matvec_bytes = c->mv_bytes;
  out->graph_captures = c->graph_captures;
  out->graph_launches = c->graph_launches;
  return OMX_OK;
}

int omx_set_profile(omx_context *c, int on) {
  return guard([&] {
    if(!c) throw OmxError(OMX_ERR_ARG, "null context");
    c->profile = on != 0;
    c->prof_gemm_ms = c->prof_panel_ms = c->prof_flop = c->prof_total_ms = 0;
    c->prof_launches = 0;
    c->mv_calls = 0;
    c->mv_ms = 0;
    c->mv_bytes = 0;
  });
}

int omx_read_matrix(omx_context *c, int which, double *out) {
  return guard([&] {
    if(!c || !c->updated || !out) throw OmxError(OMX_ERR_STATE, "no assembled matrix");
    const DBuf &b = which == 0 ? c->d_S64 : c->d_V64;
    if(!b.p) throw OmxError(OMX_ERR_STATE, "matrix not held in FP64");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpy2DAsync(out, sizeof(cplx) * c->N, b.p, sizeof(cplx) * c->Np, sizeof(cplx) * c->N, c->N,
                            hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int omx_read_matrix_block(omx_context *c, int which, int row0, int col0, int nrows, int ncols, double *out) {
  return guard([&] {
    if(!c || !c->updated || !out) throw OmxError(OMX_ERR_STATE, "no assembled matrix");
    if(row0 < 0 || col0 < 0 || nrows < 1 || ncols < 1 || row0 + nrows > c->N || col0 + ncols > c->N)
      throw OmxError(OMX_ERR_ARG, "block outside the N x N matrix");
    const DBuf &b = which == 0 ? c->d_S64 : c->d_V64;
    if(!b.p) throw OmxError(OMX_ERR_STATE, "matrix not held in FP64");
    HIPCHK(hipSetDevice(c->device));
    const cplx *src = b.as<cplx>() + (size_t)col0 * c->Np + row0;
    HIPCHK(hipMemcpy2DAsync(out, sizeof(cplx) * nrows, src, sizeof(cplx) * c->Np, sizeof(cplx) * nrows, ncols,
                            hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int omx_aca_info(omx_context *c, int which, int *ranks, double *info) {
  return guard([&] {
    if(!c || which < 0 || which > 1) throw OmxError(OMX_ERR_ARG, "invalid arguments");
    const omx_context::AcaMat &m = c->aca[which];
    if(!m.built) throw OmxError(OMX_ERR_STATE, "no compressed operator (OMX_SOLVER_ACA_GMRES and omx_update first)");
    if(ranks) std::copy(m.rank.begin(), m.rank.end(), ranks);
    if(info) {
      info[0] = m.bytes;
      info[1] = 16.0 * (double)c->N * c->N;
      info[2] = m.ref_mb;
      info[3] = m.naca;
      info[4] = m.ndense;
      info[5] = m.rmax;
      info[6] = m.naca ? m.rsum / m.naca : 0.0;
    }
  });
}

int omx_aca_block(omx_context *c, int which, int i, int j, int *rank, double *U, double *V) {
  return guard([&] {
    if(!c || which < 0 || which > 1 || !rank || !U || !V) throw OmxError(OMX_ERR_ARG, "invalid arguments");
    const omx_context::AcaMat &m = c->aca[which];
    if(!m.built) throw OmxError(OMX_ERR_STATE, "no compressed operator (OMX_SOLVER_ACA_GMRES and omx_update first)");
    if(i < 0 || j < 0 || i >= m.P || j >= m.P) throw OmxError(OMX_ERR_ARG, "block outside the cluster");
    HIPCHK(hipSetDevice(c->device));
    const size_t b = (size_t)i * m.P + j, n = m.n;
    const int r = m.rank[b];
    *rank = r;
    if(r < 0) {
      HIPCHK(hipMemcpyAsync(U, m.where[b], sizeof(cplx) * n * n, hipMemcpyDeviceToHost, c->st));
    } else if(r > 0) {
      HIPCHK(hipMemcpyAsync(U, m.where[b], sizeof(cplx) * n * r, hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipMemcpyAsync(V, m.where[b] + n * r, sizeof(cplx) * n * r, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int omx_read_vector(omx_context *c, int which, double *out) {
  return guard([&] {
    if(!c || !c->updated || !out) throw OmxError(OMX_ERR_STATE, "nothing assembled");
    const DBuf *b = which == 0 ? &c->d_Q : which == 1 ? &c->d_K : which == 2 ? &c->d_K1 : &c->d_incloc;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(out, b->p, sizeof(cplx) * c->N, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

int omx_dense_solve(omx_context *c, int N, const double *A, const double *b, double *x, int kind) {
  return guard([&] {
    if(!c || N < 1 || !A || !b || !x || (kind != OMX_SOLVER_LU64 && kind != OMX_SOLVER_LU32_IR))
      throw OmxError(OMX_ERR_ARG, "invalid dense solve");
    HIPCHK(hipSetDevice(c->device));
    // reuse the context machinery with a synthetic "geometry" of the right
    // size; the geometry's sizes, solver kind and SH flag come back on every
    // exit (a failed solve must not leave the dense sizes behind: the next
    // omx_update would size its buffers from them)
    omx_context &ctx = *c;
    struct Restore {
      omx_context &c;
      int N, Np, kind, sh;
      ~Restore() {
        c.N = N;
        c.Np = Np;
        c.kind = kind;
        c.sh = sh;
        c.updated = false;
        c.solved = false;
        c.factored = c.factored_sh = false;
      }
    } restore{ctx, ctx.N, ctx.Np, ctx.kind, ctx.sh};
    ctx.N = N;
    ctx.Np = (N + omx::kNB - 1) / omx::kNB * omx::kNB;
    ctx.kind = kind;
    ctx.sh = 0;
    ctx.updated = false;
    const size_t Np = ctx.Np;
    ctx.d_x.ensure(sizeof(cplx) * Np, &ctx.total);
    ctx.d_Q.ensure(sizeof(cplx) * Np, &ctx.total);
    ctx.d_r.ensure(sizeof(cplx) * Np, &ctx.total);
    ctx.d_w1.ensure(sizeof(cplx) * Np, &ctx.total);
    ctx.d_w2.ensure(sizeof(cplx) * Np, &ctx.total);
    ctx.d_gpart.ensure(sizeof(cplx) * Np * omx::kGemvSplitMax, &ctx.total);
    ctx.d_part.ensure(sizeof(cplx) * 128 * (ctx.restart + 2), &ctx.total);
    ctx.d_h.ensure(sizeof(cplx) * 2 * (ctx.restart + 2), &ctx.total);
    ctx.d_nrm.ensure(sizeof(double) * 256, &ctx.total);
    alloc_matrices(&ctx);
    HIPCHK(hipMemsetAsync(ctx.d_Q.p, 0, sizeof(cplx) * Np, ctx.st));
    HIPCHK(hipMemcpyAsync(ctx.d_Q.p, b, sizeof(cplx) * N, hipMemcpyHostToDevice, ctx.st));
    HIPCHK(hipMemcpy2DAsync(ctx.d_S64.p, sizeof(cplx) * Np, A, sizeof(cplx) * N, sizeof(cplx) * N, N,
                            hipMemcpyHostToDevice, ctx.st));
    omx::run_pad64(ctx.d_S64.as<cplx>(), N, (int)Np, Np, ctx.st);
    if(kind == OMX_SOLVER_LU32_IR) {
      std::vector<float2> a32((size_t)N * N);
      for(size_t i = 0; i < a32.size(); ++i) a32[i] = make_float2((float)A[2 * i], (float)A[2 * i + 1]);
      HIPCHK(hipMemcpy2DAsync(ctx.d_S32.p, sizeof(float2) * Np, a32.data(), sizeof(float2) * N, sizeof(float2) * N, N,
                              hipMemcpyHostToDevice, ctx.st));
      omx::run_pad32(ctx.d_S32.as<float2>(), N, (int)Np, Np, ctx.st);
      HIPCHK(hipStreamSynchronize(ctx.st));
    }
    ctx.factored = false;
    ctx.dirty32 = false;
    int it = 0;
    double res = 0;
    solve_system(&ctx, false, ctx.d_Q.as<cplx>(), ctx.d_x.as<cplx>(), &it, &res);
    HIPCHK(hipMemcpyAsync(x, ctx.d_x.p, sizeof(cplx) * N, hipMemcpyDeviceToHost, ctx.st));
    HIPCHK(hipStreamSynchronize(ctx.st));
    ctx.stats.iterations = it;
    ctx.stats.residual = res;
  });
}

int omx_solve_sweep(omx_context *c, int nl, const double *lambdas, const double *eps_r, const double *eps_r_sh,
                    int batch, double *x_sca, double *x_int, double *x_sca_sh, double *x_int_sh, double *xs,
                    int *iterations) {
  return guard([&] {
    if(!c || !c->have_geom || !c->have_exc || nl < 1 || !lambdas || batch < 0)
      throw OmxError(OMX_ERR_ARG, "invalid sweep (geometry, excitation and wavelengths required)");
    // the automatic choice in a batched sweep: V's mixed LU, and with the
    // second harmonic S by lockstep FP64 GMRES beside V's factorisation
    if(c->kind_req == OMX_SOLVER_AUTO) c->kind = OMX_SOLVER_LU32_IR;
    const bool krylov_s = c->kind_req == OMX_SOLVER_AUTO && c->sh && c->N >= kSweepKrylovN;
    if(!is_lu(c->kind)) throw OmxError(OMX_ERR_ARG, "sweeps need an LU solver kind");
    for(int l = 0; l < nl; ++l)
      if(!(lambdas[l] > 0.0)) throw OmxError(OMX_ERR_ARG, "non-positive wavelength");
    HIPCHK(hipSetDevice(c->device));
    alloc_all(c);
    c->d_status.ensure(64, &c->total);
    const int P = c->P, N = c->N;
    // geometry and material constants on the device (as omx_update)
    HIPCHK(hipMemcpyAsync(c->d_xyz.p, c->xyz.data(), sizeof(double) * 3 * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_radius.p, c->radius.data(), sizeof(double) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_mu.p, c->mu.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_kppp.p, c->ksippp.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_kpar.p, c->ksiparppar.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_gam.p, c->gamma.data(), sizeof(cplx) * P, hipMemcpyHostToDevice, c->st));
    // batch size: at most 64 systems and half of the free device memory
    const bool f64 = c->kind == OMX_SOLVER_LU64;
    const double per = (double)c->Np * c->Np * ((c->sh ? 2 : 1) * (f64 ? 32.0 : 24.0));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    int nb = batch > 0 ? batch : 64;
    nb = std::max(1, std::min({nb, nl, (int)std::max(1.0, 0.5 * (double)(fr + c->sw.S64.bytes) / per)}));
    std::vector<cplx> eps((size_t)nl * P), eps_sh((size_t)nl * P);
    for(int l = 0; l < nl; ++l)
      for(int i = 0; i < P; ++i) {
        eps[(size_t)l * P + i] = eps_r ? kEps0 * C(eps_r, l * P + i) : c->eps[i];
        eps_sh[(size_t)l * P + i] = eps_r_sh ? kEps0 * C(eps_r_sh, l * P + i) : c->eps_sh[i];
      }
    auto off = [&](double *p, int l, size_t n) { return p ? p + 2 * n * l : nullptr; };
    std::vector<int> redo;
    int itmax[2] = {0, 0};
    double resmax[2] = {0, 0};
    HIPCHK(hipEventRecord(c->ev[6], c->st));
    for(int l0 = 0; l0 < nl; l0 += nb) {
      const int n = std::min(nb, nl - l0);
      std::vector<int> bad(n, 0), its(2 * n, 0);
      std::vector<double> res(2 * n, 0.0);
      sweep_chunk(c, n, lambdas + l0, eps.data() + (size_t)l0 * P, eps_sh.data() + (size_t)l0 * P,
                  off(x_sca, l0, N), off(x_int, l0, N), off(x_sca_sh, l0, N), off(x_int_sh, l0, N),
                  xs ? xs + 5 * l0 : nullptr, bad, its, res, krylov_s);
      for(int z = 0; z < n; ++z) {
        if(bad[z]) redo.push_back(l0 + z);
        if(iterations) {
          iterations[2 * (l0 + z)] = its[2 * z];
          iterations[2 * (l0 + z) + 1] = its[2 * z + 1];
        }
        for(int e = 0; e < 2; ++e) {
          itmax[e] = std::max(itmax[e], its[2 * z + e]);
          if(!bad[z]) resmax[e] = std::max(resmax[e], res[2 * z + e]);
        }
      }
    }
    HIPCHK(hipEventRecord(c->ev[7], c->st));
    HIPCHK(hipEventSynchronize(c->ev[7]));
    const float t_batch = elapsed(c, 6, 7);
    // systems the batch could not solve: one at a time with partial pivoting
    if(!redo.empty()) {
      // the batch's matrix-sized buffers are released first, so the
      // one-at-a-time solves (which allocate a single system's matrices,
      // copies and bases) never hold memory beside a whole batch: the
      // context's peak stays the batch footprint a scheduler admitted
      {
        auto &w = c->sw;
        DBuf *big[] = {&w.S64, &w.S32, &w.V64, &w.V32, &w.F64, &w.F64sh, &w.basis, &w.basisZ, &w.gpart};
        for(DBuf *b : big) b->release(&c->total);
      }
      const double lam0 = c->lambda;
      const std::vector<cplx> e0 = c->eps, es0 = c->eps_sh;
      const int piv0 = c->pivot;
      c->pivot = OMX_PIVOT_PARTIAL;
      try {
        for(int l : redo) {
          c->lambda = lambdas[l];
          for(int i = 0; i < P; ++i) {
            c->eps[i] = eps[(size_t)l * P + i];
            c->eps_sh[i] = eps_sh[(size_t)l * P + i];
          }
          int rc = omx_update(c);
          if(rc) throw OmxError(rc, omx_last_error());
          rc = omx_solve(c, off(x_sca, l, N), off(x_int, l, N), off(x_sca_sh, l, N), off(x_int_sh, l, N));
          if(rc) throw OmxError(rc, omx_last_error());
          if(xs) {
            rc = omx_cross_sections(c, xs + 5 * l);
            if(rc) throw OmxError(rc, omx_last_error());
            xs[5 * l + 4] = 0.0;
            if(c->sh) {
              rc = omx_absorption_cross_section_sh(c, xs + 5 * l + 4);
              if(rc) throw OmxError(rc, omx_last_error());
            }
          }
          if(iterations) {
            iterations[2 * l] = c->stats.iterations;
            iterations[2 * l + 1] = c->stats.iterations_sh;
          }
          ++c->fallbacks;
        }
      } catch(...) {
        c->lambda = lam0;
        c->eps = e0;
        c->eps_sh = es0;
        c->pivot = piv0;
        throw;
      }
      c->lambda = lam0;
      c->eps = e0;
      c->eps_sh = es0;
      c->pivot = piv0;
      c->updated = c->solved = false;
    }
    c->stats.N = c->N;
    c->stats.Np = c->Np;
    c->stats.iterations = itmax[0];
    c->stats.iterations_sh = itmax[1];
    c->stats.residual = resmax[0];
    c->stats.residual_sh = resmax[1];
    c->stats.t_solve_ms = t_batch;
    c->stats.sweep_batch = nb;
    c->stats.device_bytes = c->total;
  });
}

int omx_mem_info(omx_context *c, size_t *free_bytes, size_t *total_bytes, size_t *held_bytes) {
  return guard([&] {
    if(!c) throw OmxError(OMX_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    if(free_bytes) *free_bytes = fr;
    if(total_bytes) *total_bytes = tot;
    if(held_bytes) *held_bytes = c->total;
  });
}

int omx_trim(omx_context *c) {
  return guard([&] {
    if(!c) throw OmxError(OMX_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamSynchronize(c->st2));
    c->fg.reset();
    c->fg_sweep.reset();
    auto &w = c->sw;
    DBuf *big[] = {&c->d_S64, &c->d_S32, &c->d_V64, &c->d_V32, &c->d_F64, &c->d_F64sh, &w.S64, &w.S32, &w.V64,
                   &w.V32, &w.F64, &w.F64sh, &w.basis, &w.basisZ, &w.gpart, &c->d_basis, &c->d_basisZ,
                   &c->d_gpart};
    for(DBuf *b : big) b->release(&c->total);
    c->aca[0].release(&c->total);
    c->aca[1].release(&c->total);
    c->ws_in.release(&c->total);
    c->updated = c->solved = false;
    c->factored = c->factored_sh = false;
  });
}

int omx_dense_solve_noalloc(omx_context *c, int N, double *A, size_t lda, const double *b, double *x) {
  return guard([&] {
    if(!c || !A || !b || !x) throw OmxError(OMX_ERR_ARG, "omx_dense_solve_noalloc: null argument");
    if(N < 1 || N > kNoAllocMaxN)
      throw OmxError(OMX_ERR_ARG, "omx_dense_solve_noalloc: N = " + std::to_string(N) + " is outside [1, " +
                                      std::to_string(kNoAllocMaxN) + "], the limit of its scratch array");
    const int Np = (N + omx::kNB - 1) / omx::kNB * omx::kNB;
    if(lda < (size_t)Np)
      throw OmxError(OMX_ERR_ARG, "omx_dense_solve_noalloc: lda = " + std::to_string(lda) +
                                      " is below N rounded up to 64 (" + std::to_string(Np) + ")");
    if(!c->scratch) throw OmxError(OMX_ERR_STATE, "omx_dense_solve_noalloc: context has no scratch address");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->st;
    using L = NoAllocLayout;
    unsigned char *base = static_cast<unsigned char *>(c->scratch);
    auto ip = [&](size_t off) { return reinterpret_cast<int *>(base + off); };
    omx::LuWork w{};
    w.status = ip(L::status);
    w.piv = ip(L::piv);
    w.perm = ip(L::perm);
    w.cand[0] = ip(L::cand);
    w.cand[1] = ip(L::cand) + L::gmax * omx::kNB;
    w.cnt[0] = ip(L::cnt);
    w.cnt[1] = ip(L::cnt) + L::gmax;
    w.mdst = ip(L::map);
    w.msrc = ip(L::map) + (L::Np + 2 * omx::kNB);
    w.mcnt = ip(L::map) + 2 * (L::Np + 2 * omx::kNB);
    w.wlu = base + L::wlu;
    w.dinv = base + L::dinv;
    w.aux = c->aux;
    w.ev_next = c->lu_ev[0];
    w.ev_piv = c->lu_ev[1];
    w.ev_bulk = c->lu_ev[2];
    w.ev_pan = c->lu_ev[3];
    w.gmax = L::gmax;
    w.pivot = omx::kPivotPartial;
    w.nb = 1;
    cplx *rb = reinterpret_cast<cplx *>(base + L::rhs), *xs = reinterpret_cast<cplx *>(base + L::x);
    // the array keeps the last call's contents: the status words start at zero
    const int nz = (int)((L::piv - L::status) / sizeof(int));
    hipLaunchKernelGGL(zero_ints_kernel, dim3((nz + 255) / 256), dim3(256), 0, st, ip(L::status), nz);
    hipLaunchKernelGGL(stage_rhs_kernel, dim3((Np + 255) / 256), dim3(256), 0, st, reinterpret_cast<const cplx *>(b),
                       rb, N, Np);
    omx::run_pad64(reinterpret_cast<cplx *>(A), N, Np, lda, st);
    HIPCHK(hipGetLastError());
    const int rc = omx::lu_factor<double2>(reinterpret_cast<double2 *>(A), Np, lda, w, st, nullptr);
    if(rc) throw OmxError(OMX_ERR_DEVICE, "lu_factor launch failed (" + std::to_string(rc) + ")");
    auto status = [&]() {
      int s[2] = {0, 0};
      HIPCHK(hipMemcpyAsync(s, w.status, sizeof(s), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      if(s[1]) throw OmxError(OMX_ERR_DEVICE, "device wait timed out (code " + std::to_string(s[1]) + ")");
      if(s[0]) throw OmxError(OMX_ERR_SINGULAR, "zero pivot at column " + std::to_string(s[0] - 1));
    };
    status();
    omx::lu_solve<double2>(reinterpret_cast<const double2 *>(A), Np, lda, w, rb, xs, st);
    HIPCHK(hipMemcpyAsync(x, xs, sizeof(cplx) * N, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipGetLastError());
    status();
  });
}

}  // extern "C"
