// Host runtime and C ABI of the engine (include/skx.h).
//
// A handle owns the device state of one lattice:
//   F[2], V[2]   field / velocity, ping-ponged by the stage-4 kernel
//   l[3]         RK4 velocity slopes l1, l2, l3 (l4 never leaves registers)
//   st           DevStatus (energies, arrest flag, convergence, counters)
//   partial      per-block reduction partials
// and one HIP stream.  Stage kernels are launched directly or replayed from
// a captured hipGraph of a chunk of steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/skx.h"
#include "skx_internal.h"

using namespace skx;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                                    \
  do {                                                                                                   \
    hipError_t _e = (expr);                                                                              \
    if (_e != hipSuccess) return fail(SKX_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

struct Lattice {
  int theory = 0, prec = 0, nf = 0, nsk = 4;
  int X = 0, Y = 0, Z = 0, halo = 2;
  long long N = 0;
  double lsx = 0, lsy = 0, lsz = 0;
  PhysParams P{};
  int killkinen = 1, unit_length = 1;
};

int parse_desc(const skx_desc* d, Lattice& L) {
  if (!d || !d->p_i || !d->p_f) return fail(SKX_EINVAL, "null descriptor or parameter array");
  if (d->n_i < 11 || d->n_f < 8) return fail(SKX_EINVAL, "packed parameter arrays too short (need p_i[11], p_f[8])");
  if (d->precision != SKX_FP64 && d->precision != SKX_FP32) return fail(SKX_EINVAL, "unknown precision");
  const int32_t* pi = d->p_i;
  const double* pf = d->p_f;
  L.theory = d->theory;
  L.prec = d->precision;
  L.X = pi[0];
  L.Y = pi[1];
  L.Z = pi[2];
  L.halo = pi[3];
  L.nf = pi[5];
  if (L.X < 1 || L.Y < 1 || L.Z < 1) return fail(SKX_EINVAL, "lattice dimensions must be positive");
  if (L.halo < 2) return fail(SKX_EINVAL, "halo (p_i[3]) must be >= 2 for the radius-2 stencils");
  if (pi[4] != 3) return fail(SKX_EINVAL, "number_coordinates (p_i[4]) must be 3");
  L.N = (long long)L.X * L.Y * L.Z;
  if (pi[6] != L.N) return fail(SKX_EINVAL, "p_i[6] (dim_grid) != X*Y*Z");
  if ((long long)pi[7] != L.N * L.nf) return fail(SKX_EINVAL, "p_i[7] (dim_fields) != nf*N");
  if (L.N * L.nf >= (1LL << 31)) return fail(SKX_EINVAL, "lattice too large for one handle (nf*N >= 2^31)");
  L.killkinen = pi[8] ? 1 : 0;
  L.unit_length = pi[10] ? 1 : 0;
  L.lsx = pf[3];
  L.lsy = pf[4];
  L.lsz = pf[5];
  if (!(L.lsx > 0) || !(L.lsy > 0) || !(L.lsz > 0)) return fail(SKX_EINVAL, "lattice spacings must be positive");
  L.P.vol = pf[6];
  L.P.dt = pf[7];
  switch (L.theory) {
    case SKX_THEORY_PION:
      if (L.nf != 4) return fail(SKX_EINVAL, "pion theory needs number_total_fields == 4");
      if (d->n_f < 10) return fail(SKX_EINVAL, "pion theory needs p_f[8] (mpi) and p_f[9] (kappa)");
      L.P.m2 = pf[8] * pf[8];
      L.P.k2 = pf[9] * pf[9];
      break;
    case SKX_THEORY_OMEGA:
    case SKX_THEORY_COULOMB:
      if (L.nf != 5) return fail(SKX_EINVAL, "omega / Coulomb theories need number_total_fields == 5");
      if (d->n_i < 12 || pi[11] != 4) return fail(SKX_EINVAL, "omega / Coulomb theories need p_i[11] == 4");
      if (d->n_f < 10) return fail(SKX_EINVAL, "omega / Coulomb theories need p_f[8] and p_f[9]");
      L.P.m2 = pf[8] * pf[8];
      L.P.coup = pf[9];
      break;
    case SKX_THEORY_RHO:
      if (L.nf != 13) return fail(SKX_EINVAL, "rho theory needs number_total_fields == 13");
      if (d->n_f < 11) return fail(SKX_EINVAL, "rho theory needs p_f[8..10]");
      L.P.m2 = pf[8] * pf[8];
      L.P.coup = pf[9] * pf[9];
      L.P.calpha = pf[10];
      break;
    case SKX_THEORY_BERGER:
      if (L.nf != 4) return fail(SKX_EINVAL, "Berger theory needs number_total_fields == 4");
      if (d->n_f < 13) return fail(SKX_EINVAL, "Berger theory needs p_f[8] (mpi), p_f[9] (kappa), p_f[12] (alpha)");
      L.P.m2 = pf[8] * pf[8];
      L.P.k2 = pf[9] * pf[9];
      L.P.coup = pf[12] * pf[12] - 1.0;  // beta (theories/berger_skyrme/kernels.py:602)
      break;
    case SKX_THEORY_LIGHTLY_BOUND:
      if (L.nf != 4) return fail(SKX_EINVAL, "lightly bound theory needs number_total_fields == 4");
      if (d->n_f < 10) return fail(SKX_EINVAL, "lightly bound theory needs p_f[8] (mpi) and p_f[9] (alpha)");
      L.P.m2 = pf[8] * pf[8];
      L.P.coup = pf[9];
      break;
    default:
      return fail(SKX_EINVAL, "unknown theory id");
  }
  return SKX_OK;
}

LaunchShape choose_tile(const Lattice& L, const skx_desc* d) {
  LaunchShape s{32, 8, 1};
  if (d->tile_z == 64) s = {64, 4, 1};
  if (d->tile_z == 16) s = {16, 16, 1};
  if (d->tile_z == 0 && L.Z <= 16) s = {16, 16, 1};
  return s;
}

// Default tile of one stage.  In FP32 some stages run on a 32x16 tile, chosen
// per theory and stage from measurement (wide_tile_stage).  The taller tile
// cuts the y-halo share of every plane from 1.69 to 1.41 loads per site, but
// has half the blocks to spread over the CUs.  A forced tile_z and lattices
// with Y < 32 keep the base tile for every stage; a batch makes its choice for
// all its lattices at creation (skx_batch_create).
//
// Pion: stage 4 (FP32 B=4, us) 56.7 -> 53.7 at 128^3 and 133.2 -> 125.9 at
// 160^3 on 32x16; stages 0-3 are 1-5% faster on 32x8 (stage 1: 37.5 against
// 38.6, and 62.0 against 64.4).  Berger's stage 1 needs 256 VGPRs and runs 1
// wave per SIMD on 32x8, while a 32x16 block holds 2 (128^3: 75 -> 56 us;
// lightly bound stage 1 54 -> 49), so both run stages 0, 1 and 4 on 32x16.
//
// Omega and Coulomb (five components): their stage 4 holds 1 wave per SIMD
// on 32x8 (Coulomb: 242 VGPRs and 88 KB of LDS per 4-wave block) and 2 on
// 32x16, and runs faster there at both sizes measured (FP32, B=8, us):
// omega 67.7 -> 62.1 at 128^3, 161.8 -> 151.4 at 160^3; Coulomb 100.6 -> 73.5
// and 258.3 -> 176.7.  Their other stages are as fast or faster on 32x8
// (Coulomb stages 2 and 3 at 160^3: 89.6 and 108.9 against 99.3 and 125.7).
// In a batch, omega keeps 32x8 for every stage: in batches of 64^3-96^3
// lattices its 32x16 stages 1 and 4 ran 11-15% slower (half the blocks to
// spread over the CUs).  Rho keeps 32x8: its stage 1 ran 218 -> 283 us on
// 32x16 at 128^3.
bool wide_tile_stage(int theory, int stage, bool batch) {
  switch (theory) {
    case SKX_THEORY_RHO: return false;
    case SKX_THEORY_OMEGA: return !batch && stage == 4;
    case SKX_THEORY_COULOMB:
    case SKX_THEORY_PION: return stage == 4;
    default: return stage <= 1 || stage == 4;
  }
}
constexpr LaunchShape kWideTile{32, 16, 1};

LaunchShape stage_tile(const Lattice& L, const LaunchShape& base, int stage, bool per_stage) {
  if (!per_stage || L.prec != SKX_FP32 || base.tz != 32 || base.ty != 8 || L.Y < 32) return base;
  return wide_tile_stage(L.theory, stage, false) ? kWideTile : base;
}

// Planes per block along x for one stage kernel.  A block's run time grows
// with its (chunk + 4) halo-inclusive planes, and the CUs hold `capacity` =
// CUs x resident blocks per CU blocks at once, so a launch of nb blocks takes
// about ceil(nb / capacity) x (chunk + 4) plane times: the chunk count with
// the least of that (the fewest chunks among equals) is taken.  Filling the
// resident capacity in one round is that optimum whenever the columns divide
// it (128^3 on every tile).  Where they do not, more and shorter chunks in
// two or more rounds beat one round that leaves CUs idle: a 192^3 lattice on
// 32x8 tiles at one block per CU has 144 columns, and one 192-plane chunk per
// column runs on 144 CUs, against seven 28-plane chunks in 4 rounds.
int choose_xchunk(const Lattice& L, const LaunchShape& tile, int blocks_per_cu, int n_cu, int forced) {
  if (forced > 0) return forced;
  const long long cols = std::max(1LL, (long long)((L.Z + tile.tz - 1) / tile.tz) * ((L.Y + tile.ty - 1) / tile.ty));
  const long long capacity = (long long)std::max(blocks_per_cu, 1) * std::max(n_cu, 1);
  long long best_chunk = L.X, best_cost = -1;
  for (long long n = 1; n <= std::min<long long>(L.X, 64); ++n) {
    const long long chunk = (L.X + n - 1) / n;
    const long long nb = cols * ((L.X + chunk - 1) / chunk);
    const long long cost = ((nb + capacity - 1) / capacity) * (chunk + 4);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best_chunk = chunk;
    }
  }
  return (int)std::max(1LL, best_chunk);
}

size_t elem_size(int prec) { return prec == SKX_FP64 ? sizeof(double) : sizeof(float); }

// The FP64 pion RK stages 2-4 run two threads per site (skx_stage_pair.h): two
// waves per SIMD where the one-lane kernels hold one, and no register spills.
bool pair_theory(int theory, int prec) { return prec == SKX_FP64 && theory == SKX_THEORY_PION; }
bool use_lane_pairs(const Lattice& L, uint32_t flags) {
  return pair_theory(L.theory, L.prec) && !(flags & SKX_FLAG_ONE_LANE) &&
         lane_pair_fits(L.N, (long long)L.Y * L.Z, elem_size(L.prec));
}

}  // namespace

struct skx_engine {
  Lattice L;
  int device = 0;
  bool owns_stream = true;  // false for lattices of a batch (they share its stream)
  bool owns_status = true;  // false for lattices of a batch (status lives in the batch's array)
  uint32_t flags = 0;
  LaunchShape shape[5]{};  // per stage kind: 0 energy pass, 1..4 RK4 stages
  hipStream_t stream = nullptr;
  void* F[2] = {nullptr, nullptr};
  void* V[2] = {nullptr, nullptr};
  void* l[3] = {nullptr, nullptr, nullptr};
  void* aux = nullptr;  // rho: per-site D^a_j and R^a_ij of the curl pass (18 components)
  int cur = 0;
  DevStatus* st = nullptr;
  DevStatus* st_host = nullptr;  // pinned mirror
  double* partial = nullptr;
  double* eg = nullptr;     // lazily allocated probe / density buffer (double, nf*N)
  double* dens2 = nullptr;  // lazily allocated second density buffer (double, N)
  double* obs = nullptr;    // lazily allocated observable partials (double, kObsMaxBlocks * 9)
  StepRecord* hist = nullptr;
  int hist_cap = 0;
  // captured chunks of `graph_steps` steps, one per starting parity
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_steps = 0;
  int graph_flags = 0;
  // tracing
  double trace_ms[5] = {0, 0, 0, 0, 0};
  long long trace_n[5] = {0, 0, 0, 0, 0};
  struct TracedLaunch {
    int stage;
    hipEvent_t e0, e1;
  };
  std::vector<hipEvent_t> ev_pool;
  std::vector<TracedLaunch> ev_pending;
};

namespace {

// upper bound of the blocks of any stage launch (x-chunks of one plane)
size_t max_blocks(const Lattice& L, const LaunchShape& tile) {
  LaunchShape s = tile;
  s.xchunk = 1;
  return (size_t)stage_blocks(L.X, L.Y, L.Z, s);
}

size_t aux_bytes(const Lattice& L) {
  return L.theory == SKX_THEORY_RHO ? (size_t)L.N * 18 * elem_size(L.prec) : 0;
}

size_t footprint(const Lattice& L, const LaunchShape& tile) {
  const size_t arr = (size_t)L.N * L.nf * elem_size(L.prec);
  return 7 * arr + aux_bytes(L) + 2 * max_blocks(L, tile) * sizeof(double) + sizeof(DevStatus);
}

template <typename T>
StageArgs<T> make_args(skx_engine* e, int stage, int flags, int c, const void* field_override = nullptr) {
  const Lattice& L = e->L;
  StageArgs<T> a{};
  a.F = static_cast<const T*>(field_override ? field_override : e->F[c]);
  a.V = static_cast<const T*>(e->V[c]);
  a.la = nullptr;
  a.lb = nullptr;
  a.lc = nullptr;
  a.lout = nullptr;
  a.Fout = nullptr;
  a.Vout = nullptr;
  a.aux = static_cast<T*>(e->aux);
  if (stage == 1) a.lout = static_cast<T*>(e->l[0]);
  if (stage == 2) a.lout = static_cast<T*>(e->l[1]);
  if (stage == 3) {
    a.la = static_cast<const T*>(e->l[0]);
    a.lout = static_cast<T*>(e->l[2]);
  }
  if (stage == 4) {
    a.la = static_cast<const T*>(e->l[1]);
    a.lb = static_cast<const T*>(e->l[0]);
    a.lc = static_cast<const T*>(e->l[2]);
    a.Fout = static_cast<T*>(e->F[c ^ 1]);
    a.Vout = static_cast<T*>(e->V[c ^ 1]);
  }
  a.eg_out = nullptr;
  a.dens_e = nullptr;
  a.dens_b = nullptr;
  a.st = e->st;
  a.partial = e->partial;
  a.hist = e->hist;
  a.X = L.X;
  a.Y = L.Y;
  a.Z = L.Z;
  a.halo = L.halo;
  a.xchunk = e->shape[stage].xchunk;
  a.N = L.N;
  a.ilsx = T(1.0 / L.lsx);
  a.ilsy = T(1.0 / L.lsy);
  a.ilsz = T(1.0 / L.lsz);
  a.P = L.P;
  a.unit_length = L.unit_length;
  a.flags = flags;
  return a;
}

hipEvent_t pooled_event(skx_engine* e) {
  if (!e->ev_pool.empty()) {
    hipEvent_t ev = e->ev_pool.back();
    e->ev_pool.pop_back();
    return ev;
  }
  // timing-only events: no system-scope fence (cache writeback and
  // invalidate) when they are recorded, so that bracketing a kernel does not
  // flush the L2 that the next kernel would find warm in an untraced run
  hipEvent_t ev = nullptr;
  (void)hipEventCreateWithFlags(&ev, hipEventDisableSystemFence);
  return ev;
}

// Fold the timings of completed traced launches into the per-stage totals.
void resolve_trace(skx_engine* e) {
  for (auto& t : e->ev_pending) {
    (void)hipEventSynchronize(t.e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t.e0, t.e1);
    e->trace_ms[t.stage] += ms;
    e->trace_n[t.stage] += 1;
    e->ev_pool.push_back(t.e0);
    e->ev_pool.push_back(t.e1);
  }
  e->ev_pending.clear();
}

hipError_t launch(skx_engine* e, int stage, int flags, int c, double* eg_out = nullptr, double* dens_e = nullptr,
                  double* dens_b = nullptr, const void* field_override = nullptr) {
  const bool trace = (e->flags & SKX_FLAG_TRACE) != 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (trace) {
    e0 = pooled_event(e);
    e1 = pooled_event(e);
    (void)hipEventRecord(e0, e->stream);
  }
  hipError_t r;
  if (e->L.prec == SKX_FP64) {
    StageArgs<double> a = make_args<double>(e, stage, flags, c, field_override);
    a.eg_out = eg_out;
    a.dens_e = dens_e;
    a.dens_b = dens_b;
    r = launch_stage_f64(e->L.theory, stage, a, e->shape[stage], e->stream);
  } else {
    StageArgs<float> a = make_args<float>(e, stage, flags, c, field_override);
    a.eg_out = eg_out;
    a.dens_e = dens_e;
    a.dens_b = dens_b;
    r = launch_stage_f32(e->L.theory, stage, a, e->shape[stage], e->stream);
  }
  if (trace) {
    (void)hipEventRecord(e1, e->stream);
    e->ev_pending.push_back({stage, e0, e1});
  }
  return r;
}

int push_status(skx_engine* e) {
  HIP_TRY(hipMemcpyAsync(e->st, e->st_host, sizeof(DevStatus), hipMemcpyHostToDevice, e->stream));
  return SKX_OK;
}
int pull_status(skx_engine* e) {
  HIP_TRY(hipMemcpyAsync(e->st_host, e->st, sizeof(DevStatus), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  resolve_trace(e);
  return SKX_OK;
}

// One ANF step of the device-resident loop from parity c (no host sync).
hipError_t enqueue_run_step(skx_engine* e, int c, int fuse_baryon) {
  const int s1 = FUSE_ENERGY | DECIDE_ARREST | CHECK_DONE | (fuse_baryon ? FUSE_BARYON : 0);
  hipError_t r;
  if ((r = launch(e, 1, s1, c)) != hipSuccess) return r;
  if ((r = launch(e, 2, CHECK_DONE, c)) != hipSuccess) return r;
  if ((r = launch(e, 3, CHECK_DONE, c)) != hipSuccess) return r;
  return launch(e, 4, RUN_MODE | CHECK_DONE, c);
}

void drop_graphs(skx_engine* e) {
  for (int i = 0; i < 2; ++i)
    if (e->graph[i]) {
      (void)hipGraphExecDestroy(e->graph[i]);
      e->graph[i] = nullptr;
    }
  e->graph_steps = 0;
}

int build_graph(skx_engine* e, int parity, int steps, int fuse_baryon) {
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  hipError_t r = hipSuccess;
  for (int s = 0; s < steps && r == hipSuccess; ++s) r = enqueue_run_step(e, (parity + s) & 1, fuse_baryon);
  hipError_t r2 = hipStreamEndCapture(e->stream, &g);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("stage launch during capture: ") + hipGetErrorString(r));
  if (r2 != hipSuccess) return fail(SKX_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(r2));
  hipGraphExec_t ex = nullptr;
  hipError_t r3 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (r3 != hipSuccess) return fail(SKX_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(r3));
  e->graph[parity] = ex;
  // upload the executable graph now, so that its first launch (inside a run,
  // perhaps a timed one) does not pay for the upload of its kernel nodes
  HIP_TRY(hipGraphUpload(ex, e->stream));
  return SKX_OK;
}

// Step graphs of `chunk` steps for both starting parities (captured once per
// chunk size and history setting; a run replays them for whole chunks only).
int ensure_graphs(skx_engine* e, int chunk, int fuse_baryon) {
  if (e->graph_steps != chunk || e->graph_flags != fuse_baryon) {
    drop_graphs(e);
    e->graph_steps = chunk;
    e->graph_flags = fuse_baryon;
  }
  for (int parity = 0; parity < 2; ++parity)
    if (!e->graph[parity])
      if (int rc = build_graph(e, parity, chunk, fuse_baryon)) return rc;
  return SKX_OK;
}

bool graphs_enabled(const skx_engine* e) { return !(e->flags & (SKX_FLAG_NO_GRAPH | SKX_FLAG_TRACE)); }

int run_loop(skx_engine* e, int32_t max_steps, double tol, int32_t check_every, int killkinen, skx_run_stats* out) {
  auto t0 = std::chrono::steady_clock::now();
  DevStatus* h = e->st_host;
  std::memset(h, 0, sizeof(DevStatus));
  h->keep = 1.0;
  h->tol = tol;
  h->max_steps = max_steps;
  h->killkinen = killkinen;
  h->hist_cap = e->hist ? e->hist_cap : 0;
  if (int rc = push_status(e)) return rc;
  const int c0 = e->cur;
  const int chunk = check_every > 0 ? check_every : 32;
  // traced runs launch directly: every kernel is bracketed by its own event pair
  const bool use_graph = graphs_enabled(e);
  const int fuse_baryon = e->hist ? 1 : 0;
  // both parities' graphs are captured before the first launch, so that no
  // capture happens between launches of a run
  if (use_graph && max_steps >= chunk)
    if (int rc = ensure_graphs(e, chunk, fuse_baryon)) return rc;
  int issued = 0, graph_launches = 0;
  while (issued < max_steps) {
    // the launches of a chunk start from the parity the device will be at
    // if no step is skipped; skipped steps (done flag) stop the run anyway
    const int parity = (c0 + issued) & 1;
    const int remaining = max_steps - issued;
    if (use_graph && remaining >= chunk) {
      HIP_TRY(hipGraphLaunch(e->graph[parity], e->stream));
      issued += chunk;
      ++graph_launches;
    } else {  // a partial chunk: direct launches of exactly the remaining steps
      const int n = std::min(chunk, remaining);
      for (int s = 0; s < n; ++s) {
        hipError_t r = enqueue_run_step(e, (parity + s) & 1, fuse_baryon);
        if (r != hipSuccess) return fail(SKX_EHIP, std::string("stage launch: ") + hipGetErrorString(r));
      }
      issued += n;
    }
    // Without a tolerance (tol < 0) a run stops early only on a non-finite
    // value: after its last chunk the final energy pass is enqueued at once,
    // on the parity of all steps taken, and checked after the one read below.
    if (issued >= max_steps && !(tol >= 0.0)) break;
    if (int rc = pull_status(e)) return rc;
    if (h->done) break;
  }
  const bool speculative = issued >= max_steps && !(tol >= 0.0);
  int steps = speculative ? issued : h->steps;
  e->cur = (c0 + steps) & 1;
  // energy of the final field, the last step's arrest decision, and its application
  {
    hipError_t r = launch(e, 0, FUSE_ENERGY | FUSE_BARYON | DECIDE_ARREST, e->cur);
    if (r != hipSuccess) return fail(SKX_EHIP, std::string("energy pass: ") + hipGetErrorString(r));
    const long long n = e->L.N * e->L.nf;
    r = e->L.prec == SKX_FP64 ? launch_apply_arrest_f64(static_cast<double*>(e->V[e->cur]), n, e->st, e->stream)
                              : launch_apply_arrest_f32(static_cast<float*>(e->V[e->cur]), n, e->st, e->stream);
    if (r != hipSuccess) return fail(SKX_EHIP, std::string("arrest: ") + hipGetErrorString(r));
  }
  if (int rc = pull_status(e)) return rc;
  if (h->steps != steps) {
    // a non-finite value stopped the run early: report the field it stopped at
    steps = h->steps;
    e->cur = (c0 + steps) & 1;
    hipError_t r = launch(e, 0, FUSE_ENERGY | FUSE_BARYON, e->cur);
    if (r != hipSuccess) return fail(SKX_EHIP, std::string("energy pass: ") + hipGetErrorString(r));
    if (int rc = pull_status(e)) return rc;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (out) {
    out->steps = steps;
    out->converged = (steps > 0 && h->err_abs < tol) ? 1 : 0;
    out->energy = h->energy;
    out->err_signed = h->err_signed;
    out->err_abs = h->err_abs;
    out->baryon = h->baryon;
    out->seconds = std::chrono::duration<double>(t1 - t0).count();
    out->issued = issued;
    out->graph_launches = graph_launches;
  }
  if (h->nonfinite) return fail(SKX_ENONFINITE, "non-finite energy or gradient during the run");
  return SKX_OK;
}

int ensure_eg(skx_engine* e) {
  if (e->eg) return SKX_OK;
  if (hipMalloc(&e->eg, (size_t)e->L.N * e->L.nf * sizeof(double)) != hipSuccess) {
    e->eg = nullptr;
    return fail(SKX_ENOMEM, "probe buffer allocation failed");
  }
  return SKX_OK;
}
int ensure_dens2(skx_engine* e) {
  if (e->dens2) return SKX_OK;
  if (hipMalloc(&e->dens2, (size_t)e->L.N * sizeof(double)) != hipSuccess) {
    e->dens2 = nullptr;
    return fail(SKX_ENOMEM, "density buffer allocation failed");
  }
  return SKX_OK;
}

// Device initial data (skx_init_field): the ansatz into F[cur], V[cur] = 0,
// and for the omega theory omega = -c CS from the stage-0 baryon density.
int init_field(skx_engine* e, const skx_init_desc* d) {
  if (!d) return fail(SKX_EINVAL, "null init descriptor");
  if (d->n_lumps < 0 || d->n_lumps > SKX_MAX_LUMPS) return fail(SKX_EINVAL, "n_lumps must be 0..SKX_MAX_LUMPS");
  for (int l = 0; l < d->n_lumps; ++l) {
    const skx_lump& q = d->lumps[l];
    if (q.kind != SKX_LUMP_HEDGEHOG && q.kind != SKX_LUMP_RATIONAL_MAP) return fail(SKX_EINVAL, "unknown lump kind");
    if (q.profile != SKX_PROFILE_EXP && q.profile != SKX_PROFILE_GAUSS && q.profile != SKX_PROFILE_FIT)
      return fail(SKX_EINVAL, "unknown profile kind");
    if (q.profile != SKX_PROFILE_FIT && !(q.profile_param > 0.0))
      return fail(SKX_EINVAL, "exp / gauss profile length must be positive");
  }
  if (!(d->noise >= 0.0)) return fail(SKX_EINVAL, "noise must be >= 0");
  if (d->omega_cs != 0.0 && e->L.theory != SKX_THEORY_OMEGA)
    return fail(SKX_EINVAL, "omega_cs applies to the omega theory only");
  const Lattice& L = e->L;
  InitArgs a{};
  a.d = *d;
  a.X = L.X;
  a.Y = L.Y;
  a.Z = L.Z;
  a.nf = L.nf;
  a.halo = L.halo;
  a.lsx = L.lsx;
  a.lsy = L.lsy;
  a.lsz = L.lsz;
  hipError_t r = L.prec == SKX_FP64
                     ? launch_init_field_f64(static_cast<double*>(e->F[e->cur]), static_cast<double*>(e->V[e->cur]), a,
                                             e->stream)
                     : launch_init_field_f32(static_cast<float*>(e->F[e->cur]), static_cast<float*>(e->V[e->cur]), a,
                                             e->stream);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("init kernel: ") + hipGetErrorString(r));
  if (d->omega_cs != 0.0) {
    if (int rc = ensure_dens2(e)) return rc;
    r = launch(e, 0, FUSE_BARYON, e->cur, nullptr, nullptr, e->dens2);  // CS * vol per site
    if (r != hipSuccess) return fail(SKX_EHIP, std::string("Chern-Simons pass: ") + hipGetErrorString(r));
    const double scale = -d->omega_cs / L.P.vol;
    const size_t off = (size_t)4 * L.N;
    r = L.prec == SKX_FP64
            ? launch_set_component_f64(static_cast<double*>(e->F[e->cur]) + off, e->dens2, L.N, scale, e->stream)
            : launch_set_component_f32(static_cast<float*>(e->F[e->cur]) + off, e->dens2, L.N, scale, e->stream);
    if (r != hipSuccess) return fail(SKX_EHIP, std::string("omega init: ") + hipGetErrorString(r));
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  return SKX_OK;
}

int copy_in(skx_engine* e, void* dst, const double* host, size_t n) {
  if (e->L.prec == SKX_FP64) {
    HIP_TRY(hipMemcpy(dst, host, n * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = static_cast<float>(host[i]);
    HIP_TRY(hipMemcpy(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  return SKX_OK;
}
int copy_out(skx_engine* e, double* host, const void* src, size_t n) {
  if (e->L.prec == SKX_FP64) {
    HIP_TRY(hipMemcpy(host, src, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(n);
    HIP_TRY(hipMemcpy(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) host[i] = static_cast<double>(tmp[i]);
  }
  return SKX_OK;
}

void release(skx_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graphs(e);
  resolve_trace(e);
  for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
  for (void* p : {e->F[0], e->F[1], e->V[0], e->V[1], e->l[0], e->l[1], e->l[2], e->aux})
    if (p) (void)hipFree(p);
  if (e->st && e->owns_status) (void)hipFree(e->st);
  if (e->st_host && e->owns_status) (void)hipHostFree(e->st_host);
  if (e->partial) (void)hipFree(e->partial);
  if (e->eg) (void)hipFree(e->eg);
  if (e->dens2) (void)hipFree(e->dens2);
  if (e->obs) (void)hipFree(e->obs);
  if (e->hist) (void)hipFree(e->hist);
  if (e->stream && e->owns_stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

bool observable_needs_com(int32_t kind) {
  return kind == SKX_OBS_RMS || (kind >= SKX_OBS_QUADRUPOLE && kind != SKX_OBS_U);
}

// Argument checks of the observable entry points.
int check_observable(const skx_engine* h, int32_t kind, const double* com, const double* out, int32_t n) {
  static const int kCount[] = {4, 1, 9, 9, 9, 9, 1};
  if (kind < SKX_OBS_COM || kind > SKX_OBS_D) return fail(SKX_EINVAL, "unknown observable kind");
  if (kind >= SKX_OBS_QUADRUPOLE && h->L.theory != SKX_THEORY_PION)
    return fail(SKX_EINVAL, "quadrupole, U, V, W and D exist for the nuclear theory only");
  if (!out || n < kCount[kind]) return fail(SKX_EINVAL, "output holds fewer values than the observable has");
  if (observable_needs_com(kind) && !com) return fail(SKX_EINVAL, "null centre of mass");
  return SKX_OK;
}

// One observable pass: per-block partials into `partial` (device, `cap` bytes),
// copied to `host` (at least as large) and folded there in block order with
// Neumaier compensation, so the result is deterministic.
int observable_pass(skx_engine* h, int32_t kind, const double* com, double* out, double* partial, size_t cap,
                    double* host) {
  static const int kCount[] = {4, 1, 9, 9, 9, 9, 1};
  const int nv = kCount[kind];
  const bool needs_com = observable_needs_com(kind);
  const Lattice& L = h->L;
  const int blocks = observable_blocks(L.N);
  const size_t need = (size_t)blocks * nv * sizeof(double);
  if (need > cap)
    return fail(SKX_EINVAL, "observable pass needs " + std::to_string(need) + " bytes of partials, more than the " +
                                std::to_string(cap) + " available");
  ObsArgs a{};
  a.kind = kind;
  a.nv = nv;
  a.X = L.X;
  a.Y = L.Y;
  a.Z = L.Z;
  a.halo = L.halo;
  a.lsx = L.lsx;
  a.lsy = L.lsy;
  a.lsz = L.lsz;
  for (int i = 0; i < 3; ++i) a.com[i] = needs_com ? com[i] : 0.0;
  a.P = L.P;
  hipError_t r = L.prec == SKX_FP64
                     ? launch_observable_f64(static_cast<const double*>(h->F[h->cur]), partial, a, blocks, h->stream)
                     : launch_observable_f32(static_cast<const float*>(h->F[h->cur]), partial, a, blocks, h->stream);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("observable pass: ") + hipGetErrorString(r));
  HIP_TRY(hipMemcpyAsync(host, partial, need, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int q = 0; q < nv; ++q) {
    double s = 0.0, c = 0.0;
    for (int b = 0; b < blocks; ++b) {
      const double v = host[(size_t)b * nv + q];
      const double t = s + v;
      c += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
      s = t;
    }
    out[q] = s + c;
  }
  return SKX_OK;
}

#define CHECK_HANDLE(h)                                                        \
  do {                                                                         \
    if (!(h)) return fail(SKX_EINVAL, "null handle");                          \
    if (hipSetDevice((h)->device) != hipSuccess) return fail(SKX_EHIP, "hipSetDevice failed"); \
  } while (0)

// Build a handle.  `shared` / `ext_st` / `ext_st_host` are non-null for a
// lattice of a batch (shared stream, status in the batch's arrays);
// `xchunk_all` > 0 forces one x-chunk for every stage.  `per_stage`: a tile
// per stage (a solo handle asks for it with tile_z 0; a batch lattice takes
// its batch's choice, so its own energy passes partition the lattice as a
// solo handle's do).
int create_engine(const skx_desc* desc, hipStream_t shared, DevStatus* ext_st, DevStatus* ext_st_host,
                  int xchunk_all, bool per_stage, skx_engine** out) {
  *out = nullptr;
  Lattice L;
  if (int rc = parse_desc(desc, L)) return rc;
  if (desc->tile_z != 0 && desc->tile_z != 16 && desc->tile_z != 32 && desc->tile_z != 64)
    return fail(SKX_EINVAL, "tile_z must be 0, 16, 32 or 64");
  skx_engine* e = new skx_engine();
  e->L = L;
  e->device = desc->device;
  e->flags = desc->flags;
  const LaunchShape tile = choose_tile(L, desc);
  if (!shape_supported(tile)) {
    delete e;
    return fail(SKX_EINVAL, "unsupported launch shape");
  }
  hipError_t r = hipSetDevice(e->device);
  if (r != hipSuccess) {
    delete e;
    return fail(SKX_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(r));
  }
  int n_cu = 0;
  if ((r = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, e->device)) != hipSuccess) {
    delete e;
    return fail(SKX_EHIP, std::string("hipDeviceGetAttribute: ") + hipGetErrorString(r));
  }
  for (int stg = 0; stg < 5; ++stg) {
    LaunchShape st = stage_tile(L, tile, stg, per_stage);
    if (stg >= 2 && use_lane_pairs(L, desc->flags)) st.lanes = 2;
    int bpc = 0;
    if ((r = stage_blocks_per_cu(L.prec == SKX_FP32, L.theory, stg, st, &bpc)) != hipSuccess) {
      delete e;
      return fail(SKX_EHIP, std::string("occupancy query: ") + hipGetErrorString(r));
    }
    e->shape[stg] = st;
    e->shape[stg].xchunk = choose_xchunk(L, st, bpc, n_cu, xchunk_all > 0 ? xchunk_all : desc->x_chunk);
  }
  if (shared) {
    e->stream = shared;
    e->owns_stream = false;
  } else if ((r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) {
    e->stream = nullptr;
    release(e);
    return fail(SKX_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(r));
  }
  if (ext_st) {
    e->st = ext_st;
    e->st_host = ext_st_host;
    e->owns_status = false;
  }
  const size_t arr = (size_t)L.N * L.nf * elem_size(L.prec);
  void** bufs[7] = {&e->F[0], &e->F[1], &e->V[0], &e->V[1], &e->l[0], &e->l[1], &e->l[2]};
  for (void** b : bufs) {
    if (hipMalloc(b, arr) != hipSuccess) {
      *b = nullptr;
      release(e);
      return fail(SKX_ENOMEM, "device allocation of lattice arrays failed");
    }
  }
  if (aux_bytes(L) && hipMalloc(&e->aux, aux_bytes(L)) != hipSuccess) {
    e->aux = nullptr;
    release(e);
    return fail(SKX_ENOMEM, "device allocation of the rho curl buffer failed");
  }
  size_t nblk = 0;
  for (int stg = 0; stg < 5; ++stg) nblk = std::max(nblk, (size_t)stage_blocks(L.X, L.Y, L.Z, e->shape[stg]));
  if (hipMalloc(&e->partial, 2 * nblk * sizeof(double)) != hipSuccess ||
      (!ext_st && hipMalloc(&e->st, sizeof(DevStatus)) != hipSuccess) ||
      (!ext_st && hipHostMalloc(&e->st_host, sizeof(DevStatus), hipHostMallocDefault) != hipSuccess)) {
    release(e);
    return fail(SKX_ENOMEM, "device allocation of reduction state failed");
  }
  // zero everything with a fill kernel (no memset nodes anywhere)
  for (void** b : bufs) (void)launch_fill(*b, arr, e->stream);
  if (e->aux) (void)launch_fill(e->aux, aux_bytes(L), e->stream);
  (void)launch_fill(e->partial, 2 * nblk * sizeof(double), e->stream);
  std::memset(e->st_host, 0, sizeof(DevStatus));
  e->st_host->keep = 1.0;
  if (int rc = push_status(e)) {
    release(e);
    return rc;
  }
  if ((r = hipStreamSynchronize(e->stream)) != hipSuccess) {
    release(e);
    return fail(SKX_EHIP, std::string("initialisation: ") + hipGetErrorString(r));
  }
  *out = e;
  return SKX_OK;
}

}  // namespace

// Lets the host-only sources (skx_io.cpp) report through skx_last_error.
namespace skx {
int record_error(int code, const char* msg) { return fail(code, msg); }
}  // namespace skx

extern "C" {

int skx_version(void) { return SKX_VERSION; }

const char* skx_last_error(void) { return g_last_error.c_str(); }

int skx_footprint_bytes(const skx_desc* desc, size_t* bytes) {
  Lattice L;
  if (int rc = parse_desc(desc, L)) return rc;
  if (!bytes) return fail(SKX_EINVAL, "null output");
  *bytes = footprint(L, choose_tile(L, desc));
  return SKX_OK;
}

int skx_create(const skx_desc* desc, skx_handle* out) {
  if (!out) return fail(SKX_EINVAL, "null output handle");
  return create_engine(desc, nullptr, nullptr, nullptr, 0, desc->tile_z == 0, out);
}

int skx_destroy(skx_handle h) {
  if (!h) return SKX_OK;
  release(h);
  return SKX_OK;
}

int skx_upload(skx_handle h, int32_t which, const double* host, size_t n) {
  CHECK_HANDLE(h);
  if (!host) return fail(SKX_EINVAL, "null host array");
  if (n != (size_t)(h->L.N * h->L.nf)) return fail(SKX_EINVAL, "upload size must be nf*N");
  if (which != SKX_FIELD && which != SKX_VELOCITY) return fail(SKX_EINVAL, "unknown buffer");
  HIP_TRY(hipStreamSynchronize(h->stream));
  return copy_in(h, which == SKX_FIELD ? h->F[h->cur] : h->V[h->cur], host, n);
}

int skx_download(skx_handle h, int32_t which, double* host, size_t n) {
  CHECK_HANDLE(h);
  if (!host) return fail(SKX_EINVAL, "null host array");
  if (n != (size_t)(h->L.N * h->L.nf)) return fail(SKX_EINVAL, "download size must be nf*N");
  if (which != SKX_FIELD && which != SKX_VELOCITY) return fail(SKX_EINVAL, "unknown buffer");
  HIP_TRY(hipStreamSynchronize(h->stream));
  return copy_out(h, host, which == SKX_FIELD ? h->F[h->cur] : h->V[h->cur], n);
}

int skx_anf_step(skx_handle h, double e_prev, double* e_new, double* err_signed, double* err_abs) {
  CHECK_HANDLE(h);
  DevStatus* s = h->st_host;
  std::memset(s, 0, sizeof(DevStatus));
  s->keep = 1.0;
  if (int rc = push_status(h)) return rc;
  const int c = h->cur;
  hipError_t r = hipSuccess;
  for (int stage = 1; stage <= 4 && r == hipSuccess; ++stage) r = launch(h, stage, 0, c);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("stage launch: ") + hipGetErrorString(r));
  h->cur = c ^ 1;
  if ((r = launch(h, 0, FUSE_ENERGY, h->cur)) != hipSuccess)
    return fail(SKX_EHIP, std::string("energy pass: ") + hipGetErrorString(r));
  if (int rc = pull_status(h)) return rc;
  const double E = s->energy;
  // arrest: core/integrator.py:403-407 (strict >, FP64 host energies)
  if (h->L.killkinen && (E > e_prev)) {
    const size_t arr = (size_t)h->L.N * h->L.nf * elem_size(h->L.prec);
    HIP_TRY(launch_fill(h->V[h->cur], arr, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  if (e_new) *e_new = E;
  if (err_signed) *err_signed = s->err_signed;
  if (err_abs) *err_abs = s->err_abs;
  if (s->nonfinite || !std::isfinite(E)) return fail(SKX_ENONFINITE, "non-finite energy or gradient");
  return SKX_OK;
}

int skx_anf_run(skx_handle h, int32_t max_steps, double tol, int32_t check_every, skx_run_stats* stats) {
  CHECK_HANDLE(h);
  if (max_steps < 0) return fail(SKX_EINVAL, "max_steps must be >= 0");
  if (check_every < 0) return fail(SKX_EINVAL, "check_every must be >= 0");
  return run_loop(h, max_steps, tol, check_every, h->L.killkinen, stats);
}

int skx_anf_prepare(skx_handle h, int32_t check_every) {
  CHECK_HANDLE(h);
  if (check_every < 0) return fail(SKX_EINVAL, "check_every must be >= 0");
  if (!graphs_enabled(h)) return SKX_OK;
  return ensure_graphs(h, check_every > 0 ? check_every : 32, h->hist ? 1 : 0);
}

int skx_rk4_steps(skx_handle h, int32_t n, skx_run_stats* stats) {
  CHECK_HANDLE(h);
  if (n < 0) return fail(SKX_EINVAL, "n must be >= 0");
  return run_loop(h, n, -1.0, 0, 0, stats);
}

int skx_energy(skx_handle h, double* energy) {
  CHECK_HANDLE(h);
  if (!energy) return fail(SKX_EINVAL, "null output");
  hipError_t r = launch(h, 0, FUSE_ENERGY, h->cur);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("energy pass: ") + hipGetErrorString(r));
  if (int rc = pull_status(h)) return rc;
  *energy = h->st_host->energy;
  return SKX_OK;
}

int skx_baryon(skx_handle h, double* baryon) {
  CHECK_HANDLE(h);
  if (!baryon) return fail(SKX_EINVAL, "null output");
  hipError_t r = launch(h, 0, FUSE_BARYON, h->cur);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("baryon pass: ") + hipGetErrorString(r));
  if (int rc = pull_status(h)) return rc;
  *baryon = h->st_host->baryon;
  return SKX_OK;
}

int skx_density(skx_handle h, int32_t kind, double* host, size_t n) {
  CHECK_HANDLE(h);
  if (!host) return fail(SKX_EINVAL, "null host array");
  if (n != (size_t)h->L.N) return fail(SKX_EINVAL, "density size must be N");
  if (kind == SKX_DENSITY_ENERGY || kind == SKX_DENSITY_BARYON) {
    if (int rc = ensure_dens2(h)) return rc;
    const bool en = kind == SKX_DENSITY_ENERGY;
    hipError_t r = launch(h, 0, en ? FUSE_ENERGY : FUSE_BARYON, h->cur, nullptr, en ? h->dens2 : nullptr,
                          en ? nullptr : h->dens2);
    if (r != hipSuccess) return fail(SKX_EHIP, std::string("density pass: ") + hipGetErrorString(r));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(host, h->dens2, n * sizeof(double), hipMemcpyDeviceToHost));
    return SKX_OK;
  }
  if (kind == SKX_DENSITY_POTENTIAL || kind == SKX_DENSITY_RHO) {
    const int nf = h->L.nf;
    if (kind == SKX_DENSITY_POTENTIAL && nf < 5) return fail(SKX_EINVAL, "theory has no potential component");
    if (kind == SKX_DENSITY_RHO && nf < 13) return fail(SKX_EINVAL, "theory has no rho components");
    std::vector<double> f((size_t)h->L.N * nf);
    if (int rc = skx_download(h, SKX_FIELD, f.data(), f.size())) return rc;
    const size_t N = (size_t)h->L.N;
    if (kind == SKX_DENSITY_POTENTIAL) {
      std::memcpy(host, f.data() + 4 * N, N * sizeof(double));
    } else {
      for (size_t i = 0; i < N; ++i) {
        double s = 0.0;
        for (int a = 4; a < nf; ++a) s += f[a * N + i] * f[a * N + i];
        host[i] = s;
      }
    }
    return SKX_OK;
  }
  return fail(SKX_EINVAL, "unknown density kind");
}

int skx_init_field(skx_handle h, const skx_init_desc* desc) {
  CHECK_HANDLE(h);
  return init_field(h, desc);
}

int skx_observable(skx_handle h, int32_t kind, const double* com, double* out, int32_t n) {
  CHECK_HANDLE(h);
  if (int rc = check_observable(h, kind, com, out, n)) return rc;
  if (!h->obs) {
    if (hipMalloc(&h->obs, (size_t)kObsMaxBlocks * 9 * sizeof(double)) != hipSuccess) {
      h->obs = nullptr;
      return fail(SKX_ENOMEM, "observable partials allocation failed");
    }
  }
  std::vector<double> part((size_t)kObsMaxBlocks * 9);
  return observable_pass(h, kind, com, out, h->obs, (size_t)kObsMaxBlocks * 9 * sizeof(double), part.data());
}

int skx_observable_noalloc(skx_handle h, int32_t kind, const double* com, double* out, int32_t n) {
  CHECK_HANDLE(h);
  if (int rc = check_observable(h, kind, com, out, n)) return rc;
  double* partial = nullptr;
  size_t bytes = 0;
  hipError_t r = observable_scratch(h->device, &partial, &bytes);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
  // host copy of the partials: a fixed per-thread buffer of the scratch's size, no allocation per call
  static thread_local double part[kObsMaxBlocks * 9];
  return observable_pass(h, kind, com, out, partial, bytes, part);
}

int skx_eval_gradient(skx_handle h, const double* temp, double* eg, double* l, size_t n) {
  CHECK_HANDLE(h);
  if (!temp) return fail(SKX_EINVAL, "null stage field");
  if (n != (size_t)(h->L.N * h->L.nf)) return fail(SKX_EINVAL, "probe size must be nf*N");
  if (int rc = ensure_eg(h)) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  // the stage field goes to the l2 slot, the probe writes l1: solver F / V stay untouched
  if (int rc = copy_in(h, h->l[1], temp, n)) return rc;
  hipError_t r = launch(h, 1, 0, h->cur, h->eg, nullptr, nullptr, h->l[1]);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("probe: ") + hipGetErrorString(r));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (eg) HIP_TRY(hipMemcpy(eg, h->eg, n * sizeof(double), hipMemcpyDeviceToHost));
  if (l) {
    if (int rc = copy_out(h, l, h->l[0], n)) return rc;
  }
  return SKX_OK;
}

int skx_history_attach(skx_handle h, int32_t capacity) {
  CHECK_HANDLE(h);
  if (capacity < 0) return fail(SKX_EINVAL, "capacity must be >= 0");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->hist) {
    (void)hipFree(h->hist);
    h->hist = nullptr;
    h->hist_cap = 0;
  }
  drop_graphs(h);  // captured launches hold the old history pointer
  if (capacity == 0) return SKX_OK;
  if (hipMalloc(&h->hist, (size_t)capacity * sizeof(StepRecord)) != hipSuccess) {
    h->hist = nullptr;
    return fail(SKX_ENOMEM, "history allocation failed");
  }
  h->hist_cap = capacity;
  HIP_TRY(launch_fill(h->hist, (size_t)capacity * sizeof(StepRecord), h->stream));
  return SKX_OK;
}

int skx_history_read(skx_handle h, skx_step_record* out, int32_t n, int32_t* n_read) {
  CHECK_HANDLE(h);
  if (!out || !n_read) return fail(SKX_EINVAL, "null output");
  HIP_TRY(hipStreamSynchronize(h->stream));
  int k = std::min(n, std::min(h->hist_cap, h->st_host->steps));
  if (k < 0) k = 0;
  if (k > 0) HIP_TRY(hipMemcpy(out, h->hist, (size_t)k * sizeof(StepRecord), hipMemcpyDeviceToHost));
  *n_read = k;
  return SKX_OK;
}

int skx_trace_read(skx_handle h, double ms[5], int64_t counts[5]) {
  CHECK_HANDLE(h);
  HIP_TRY(hipStreamSynchronize(h->stream));
  resolve_trace(h);
  for (int i = 0; i < 5; ++i) {
    if (ms) ms[i] = h->trace_ms[i];
    if (counts) counts[i] = h->trace_n[i];
  }
  return SKX_OK;
}

int skx_stage_shape(skx_handle h, int32_t stage, int32_t shape[4]) {
  CHECK_HANDLE(h);
  if (stage < 0 || stage > 4 || !shape) return fail(SKX_EINVAL, "stage must be 0..4 and shape non-null");
  const LaunchShape& s = h->shape[stage];
  shape[0] = s.tz;
  shape[1] = s.ty;
  shape[2] = s.xchunk;
  shape[3] = s.lanes;
  return SKX_OK;
}

int skx_trace_reset(skx_handle h) {
  CHECK_HANDLE(h);
  HIP_TRY(hipStreamSynchronize(h->stream));
  resolve_trace(h);
  for (int i = 0; i < 5; ++i) {
    h->trace_ms[i] = 0;
    h->trace_n[i] = 0;
  }
  return SKX_OK;
}

int skx_sync(skx_handle h) {
  CHECK_HANDLE(h);
  HIP_TRY(hipStreamSynchronize(h->stream));
  return SKX_OK;
}

}  // extern "C"

// ====================================================================== batches
// A batch runs many lattices of one theory / precision / tile shape with one
// launch per RK4 stage.  Every lattice (slot) is a full handle (its own
// buffers, partials and status) that shares the batch's stream; its status
// lives in one contiguous device array so a poll is a single copy.  A block
// table maps each block of the fixed-size grid to a slot, so lattices join
// and leave between chunks without re-capturing the step graph: only the
// tables are re-uploaded.  Chunks have an even number of steps and every
// lattice joins at a chunk boundary, so all lattices share the ping-pong
// parity of the graph's launches.

struct skx_batch {
  int theory = 0, prec = 0, device = 0;
  uint32_t flags = 0;
  LaunchShape tile{};       // base tile: block map 0
  LaunchShape stile[5]{};   // tile of RK4 stage 1..4
  int smap[5]{};            // block map of RK4 stage 1..4 (1: the 32 x 16 tile)
  int nmaps = 1;
  int max_slots = 0, max_blocks = 0;
  hipStream_t stream = nullptr;
  std::vector<skx_engine*> slot;
  std::vector<int> first_block[2], nblocks[2];  // per map, per slot
  std::vector<int> h_blk[2];                    // per map: block -> slot (-1 idle)
  void* d_slots = nullptr;  // BatchSlot<T>[max_slots]
  void* h_slots = nullptr;  // pinned mirror
  int* d_blk[2] = {nullptr, nullptr};
  DevStatus* d_status = nullptr;
  DevStatus* h_status = nullptr;
  bool dirty = true;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
};

namespace {

size_t slot_bytes(int prec) { return prec == SKX_FP64 ? sizeof(BatchSlot<double>) : sizeof(BatchSlot<float>); }

template <typename T>
void fill_slot(skx_batch* b, int i) {
  BatchSlot<T>* tab = static_cast<BatchSlot<T>*>(b->h_slots);
  BatchSlot<T>& sl = tab[i];
  std::memset(&sl, 0, sizeof(sl));
  skx_engine* e = b->slot[i];
  if (!e) return;
  const int fl[4] = {FUSE_ENERGY | DECIDE_ARREST | CHECK_DONE, CHECK_DONE, CHECK_DONE, RUN_MODE | CHECK_DONE};
  for (int stage = 1; stage <= 4; ++stage)
    for (int par = 0; par < 2; ++par) {
      StageArgs<T> a = make_args<T>(e, stage, fl[stage - 1], par);
      a.nbz = (e->L.Z + b->stile[stage].tz - 1) / b->stile[stage].tz;
      a.nby = (e->L.Y + b->stile[stage].ty - 1) / b->stile[stage].ty;
      a.nbx = (e->L.X + a.xchunk - 1) / a.xchunk;
      sl.args[stage - 1][par] = a;
    }
  for (int m = 0; m < b->nmaps; ++m) {
    sl.first_block[m] = b->first_block[m][i];
    sl.nblocks[m] = b->nblocks[m][i];
  }
  sl.active = 1;
}

int upload_tables(skx_batch* b) {
  if (!b->dirty) return SKX_OK;
  for (int i = 0; i < b->max_slots; ++i) {
    if (b->prec == SKX_FP64) fill_slot<double>(b, i);
    else fill_slot<float>(b, i);
  }
  HIP_TRY(hipMemcpyAsync(b->d_slots, b->h_slots, slot_bytes(b->prec) * b->max_slots, hipMemcpyHostToDevice,
                         b->stream));
  for (int m = 0; m < b->nmaps; ++m)
    HIP_TRY(hipMemcpyAsync(b->d_blk[m], b->h_blk[m].data(), sizeof(int) * b->max_blocks, hipMemcpyHostToDevice,
                           b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->dirty = false;
  return SKX_OK;
}

hipError_t enqueue_batch_step(skx_batch* b, int parity) {
  for (int stage = 1; stage <= 4; ++stage) {
    const int m = b->smap[stage];
    hipError_t r = b->prec == SKX_FP64
                       ? launch_stage_batch_f64(b->theory, stage, b->stile[stage],
                                                static_cast<const BatchSlot<double>*>(b->d_slots), b->d_blk[m], m,
                                                parity, b->max_blocks, b->stream)
                       : launch_stage_batch_f32(b->theory, stage, b->stile[stage],
                                                static_cast<const BatchSlot<float>*>(b->d_slots), b->d_blk[m], m,
                                                parity, b->max_blocks, b->stream);
    if (r != hipSuccess) return r;
  }
  return hipSuccess;
}

int batch_graph(skx_batch* b, int steps) {
  if (b->graph && b->graph_steps == steps) return SKX_OK;
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t r = hipSuccess;
  for (int s = 0; s < steps && r == hipSuccess; ++s) r = enqueue_batch_step(b, s & 1);
  hipError_t r2 = hipStreamEndCapture(b->stream, &g);
  if (r != hipSuccess) return fail(SKX_EHIP, std::string("batched launch during capture: ") + hipGetErrorString(r));
  if (r2 != hipSuccess) return fail(SKX_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(r2));
  hipError_t r3 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (r3 != hipSuccess) {
    b->graph = nullptr;
    return fail(SKX_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(r3));
  }
  b->graph_steps = steps;
  HIP_TRY(hipGraphUpload(b->graph, b->stream));
  return SKX_OK;
}

void release_batch(skx_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (skx_engine* e : b->slot)
    if (e) release(e);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->d_slots) (void)hipFree(b->d_slots);
  if (b->h_slots) (void)hipHostFree(b->h_slots);
  for (int m = 0; m < 2; ++m)
    if (b->d_blk[m]) (void)hipFree(b->d_blk[m]);
  if (b->d_status) (void)hipFree(b->d_status);
  if (b->h_status) (void)hipHostFree(b->h_status);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

void slot_status(skx_batch* b, int i, skx_slot_status* o) {
  const DevStatus& s = b->h_status[i];
  o->slot = i;
  o->active = b->slot[i] ? 1 : 0;
  o->steps = s.steps;
  o->done = s.done;
  o->converged = (s.steps > 0 && s.err_abs < s.tol) ? 1 : 0;
  o->nonfinite = s.nonfinite;
  o->energy = s.energy;
  o->err_signed = s.err_signed;
  o->err_abs = s.err_abs;
  o->baryon = s.baryon;
}

#define CHECK_BATCH(b)                                                         \
  do {                                                                         \
    if (!(b)) return fail(SKX_EINVAL, "null batch");                           \
    if (hipSetDevice((b)->device) != hipSuccess) return fail(SKX_EHIP, "hipSetDevice failed"); \
  } while (0)

}  // namespace

extern "C" {

int skx_batch_create(const skx_batch_desc* d, skx_batch_handle* out) {
  if (!d || !out) return fail(SKX_EINVAL, "null argument");
  *out = nullptr;
  if (d->theory < SKX_THEORY_PION || d->theory > SKX_THEORY_LIGHTLY_BOUND)
    return fail(SKX_EINVAL, "unknown theory id");
  if (d->precision != SKX_FP64 && d->precision != SKX_FP32) return fail(SKX_EINVAL, "unknown precision");
  if (d->max_slots < 1 || d->max_blocks < 1) return fail(SKX_EINVAL, "max_slots and max_blocks must be positive");
  skx_batch* b = new skx_batch();
  b->theory = d->theory;
  b->prec = d->precision;
  b->device = d->device;
  b->flags = d->flags;
  b->tile = d->tile_z == 64 ? LaunchShape{64, 4, 1} : (d->tile_z == 16 ? LaunchShape{16, 16, 1} : LaunchShape{32, 8, 1});
  // tile_z 0: the FP32 per-stage defaults of a single lattice (stage_tile), for every lattice of the batch
  for (int stage = 1; stage <= 4; ++stage) {
    const bool wide = d->tile_z == 0 && d->precision == SKX_FP32 && wide_tile_stage(d->theory, stage, true);
    b->stile[stage] = wide ? kWideTile : b->tile;
    if (stage >= 2 && pair_theory(d->theory, d->precision) && !(d->flags & SKX_FLAG_ONE_LANE))
      b->stile[stage].lanes = 2;
    b->smap[stage] = wide ? 1 : 0;
    if (wide) b->nmaps = 2;
  }
  b->max_slots = d->max_slots;
  b->max_blocks = d->max_blocks;
  b->slot.assign(b->max_slots, nullptr);
  for (int m = 0; m < 2; ++m) {
    b->first_block[m].assign(b->max_slots, 0);
    b->nblocks[m].assign(b->max_slots, 0);
    b->h_blk[m].assign(b->max_blocks, -1);
  }
  hipError_t r;
  if ((r = hipSetDevice(b->device)) != hipSuccess) {
    delete b;
    return fail(SKX_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(r));
  }
  if ((r = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) {
    b->stream = nullptr;
    release_batch(b);
    return fail(SKX_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(r));
  }
  const size_t sb = slot_bytes(b->prec) * b->max_slots;
  if (hipMalloc(&b->d_slots, sb) != hipSuccess || hipHostMalloc(&b->h_slots, sb, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&b->d_blk[0], sizeof(int) * b->max_blocks) != hipSuccess ||
      (b->nmaps > 1 && hipMalloc(&b->d_blk[1], sizeof(int) * b->max_blocks) != hipSuccess) ||
      hipMalloc(&b->d_status, sizeof(DevStatus) * b->max_slots) != hipSuccess ||
      hipHostMalloc(&b->h_status, sizeof(DevStatus) * b->max_slots, hipHostMallocDefault) != hipSuccess) {
    release_batch(b);
    return fail(SKX_ENOMEM, "batch table allocation failed");
  }
  std::memset(b->h_status, 0, sizeof(DevStatus) * b->max_slots);
  (void)launch_fill(b->d_status, sizeof(DevStatus) * b->max_slots, b->stream);
  *out = b;
  return SKX_OK;
}

int skx_batch_destroy(skx_batch_handle b) {
  if (!b) return SKX_OK;
  release_batch(b);
  return SKX_OK;
}

int skx_batch_add(skx_batch_handle b, const skx_desc* lattice, const double* F0, const double* V0, int32_t max_steps,
                  double tol, int32_t* slot_out) {
  CHECK_BATCH(b);
  if (!lattice || !slot_out) return fail(SKX_EINVAL, "null argument");
  if (lattice->theory != b->theory || lattice->precision != b->prec)
    return fail(SKX_EINVAL, "lattice theory / precision differ from the batch's");
  if (max_steps < 0) return fail(SKX_EINVAL, "max_steps must be >= 0");
  int i = 0;
  while (i < b->max_slots && b->slot[i]) ++i;
  if (i == b->max_slots) return fail(SKX_EBUSY, "no free slot in the batch");
  Lattice L;
  if (int rc = parse_desc(lattice, L)) return rc;
  if (b->stile[4].lanes == 2 && !lane_pair_fits(L.N, (long long)L.Y * L.Z, elem_size(L.prec)))
    return fail(SKX_EINVAL, "the lattice is too large for the two-lanes-per-site FP64 kernels of this batch "
                            "(N + Y*Z must stay below 2^28 sites); create the batch with SKX_FLAG_ONE_LANE");
  // x-chunk: the whole lattice unless the caller asks for shorter marches
  // (the batch's other lattices supply the parallelism)
  const int xc = (lattice->x_chunk > 0 && lattice->x_chunk < L.X) ? lattice->x_chunk : L.X;
  int first[2] = {-1, -1}, nblk[2] = {0, 0};
  for (int m = 0; m < b->nmaps; ++m) {  // first fit in each block map
    const LaunchShape& t = m == 0 ? b->tile : kWideTile;
    const long long nb = (long long)((L.Z + t.tz - 1) / t.tz) * ((L.Y + t.ty - 1) / t.ty) * ((L.X + xc - 1) / xc);
    if (nb > b->max_blocks)  // can never fit, unlike a grid that is busy now
      return fail(SKX_EINVAL, "the lattice needs " + std::to_string(nb) + " blocks, more than the batch's max_blocks (" +
                                  std::to_string(b->max_blocks) + ")");
    nblk[m] = (int)nb;
    for (int s = 0, run = 0; s < b->max_blocks; ++s) {
      run = (b->h_blk[m][s] < 0) ? run + 1 : 0;
      if (run == nblk[m]) {
        first[m] = s - nblk[m] + 1;
        break;
      }
    }
    if (first[m] < 0) return fail(SKX_EBUSY, "not enough free blocks in the batch grid");
  }
  skx_desc d = *lattice;
  d.tile_z = b->tile.tz;
  d.flags = (d.flags & ~(uint32_t)SKX_FLAG_ONE_LANE) | (b->flags & SKX_FLAG_ONE_LANE);
  skx_engine* e = nullptr;
  if (int rc = create_engine(&d, b->stream, b->d_status + i, b->h_status + i, xc, b->nmaps > 1, &e)) return rc;
  const size_t n = (size_t)L.N * L.nf;
  if (F0) {
    if (int rc = copy_in(e, e->F[0], F0, n)) {
      release(e);
      return rc;
    }
  } else {  // the vacuum (skx_batch_init may set a configuration on the device)
    skx_init_desc vac{};
    if (int rc = init_field(e, &vac)) {
      release(e);
      return rc;
    }
  }
  if (V0) {
    if (int rc = copy_in(e, e->V[0], V0, n)) {
      release(e);
      return rc;
    }
  }
  DevStatus* h = b->h_status + i;
  std::memset(h, 0, sizeof(DevStatus));
  h->keep = 1.0;
  h->tol = tol;
  h->max_steps = max_steps;
  h->killkinen = L.killkinen;
  if (int rc = push_status(e)) {
    release(e);
    return rc;
  }
  b->slot[i] = e;
  for (int m = 0; m < b->nmaps; ++m) {
    b->first_block[m][i] = first[m];
    b->nblocks[m][i] = nblk[m];
    for (int s = first[m]; s < first[m] + nblk[m]; ++s) b->h_blk[m][s] = i;
  }
  b->dirty = true;
  *slot_out = i;
  return SKX_OK;
}

int skx_batch_launch(skx_batch_handle b, int32_t steps) {
  CHECK_BATCH(b);
  if (steps < 0) return fail(SKX_EINVAL, "steps must be >= 0");
  if (int rc = upload_tables(b)) return rc;
  const int chunk = 2 * std::max(1, (int)((steps + 1) / 2));
  if (!(b->flags & SKX_FLAG_NO_GRAPH)) {
    if (int rc = batch_graph(b, chunk)) return rc;
    HIP_TRY(hipGraphLaunch(b->graph, b->stream));
  } else {
    for (int s = 0; s < chunk; ++s) {
      hipError_t r = enqueue_batch_step(b, s & 1);
      if (r != hipSuccess) return fail(SKX_EHIP, std::string("batched launch: ") + hipGetErrorString(r));
    }
  }
  HIP_TRY(hipMemcpyAsync(b->h_status, b->d_status, sizeof(DevStatus) * b->max_slots, hipMemcpyDeviceToHost,
                         b->stream));
  return SKX_OK;
}

int skx_batch_wait(skx_batch_handle b) {
  CHECK_BATCH(b);
  HIP_TRY(hipStreamSynchronize(b->stream));
  return SKX_OK;
}

int skx_batch_run(skx_batch_handle b, int32_t steps) {
  if (int rc = skx_batch_launch(b, steps)) return rc;
  return skx_batch_wait(b);
}

int skx_batch_poll(skx_batch_handle b, skx_slot_status* out, int32_t n, int32_t* n_out) {
  CHECK_BATCH(b);
  if (!out || !n_out) return fail(SKX_EINVAL, "null argument");
  int k = 0;
  for (int i = 0; i < b->max_slots && k < n; ++i)
    if (b->slot[i]) slot_status(b, i, &out[k++]);
  *n_out = k;
  return SKX_OK;
}

int skx_batch_init(skx_batch_handle b, int32_t slot, const skx_init_desc* desc) {
  CHECK_BATCH(b);
  if (slot < 0 || slot >= b->max_slots || !b->slot[slot]) return fail(SKX_EINVAL, "no lattice in that slot");
  if (b->h_status[slot].steps != 0) return fail(SKX_EINVAL, "the lattice has already run; initialise it right after skx_batch_add");
  // lattices of a batch are at ping-pong parity 0 between chunks
  return init_field(b->slot[slot], desc);
}

int skx_batch_remove(skx_batch_handle b, int32_t slot, double* F_out, double* V_out, skx_slot_status* final_status) {
  CHECK_BATCH(b);
  if (slot < 0 || slot >= b->max_slots || !b->slot[slot]) return fail(SKX_EINVAL, "no lattice in that slot");
  skx_engine* e = b->slot[slot];
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (int rc = pull_status(e)) return rc;
  // the lattice joined at parity 0 and has taken `steps` steps
  e->cur = e->st_host->steps & 1;
  if (e->st_host->steps > 0) {
    hipError_t r = launch(e, 0, FUSE_ENERGY | FUSE_BARYON | DECIDE_ARREST, e->cur);
    if (r != hipSuccess) return fail(SKX_EHIP, std::string("energy pass: ") + hipGetErrorString(r));
    const long long n = e->L.N * e->L.nf;
    r = e->L.prec == SKX_FP64 ? launch_apply_arrest_f64(static_cast<double*>(e->V[e->cur]), n, e->st, e->stream)
                              : launch_apply_arrest_f32(static_cast<float*>(e->V[e->cur]), n, e->st, e->stream);
    if (r != hipSuccess) return fail(SKX_EHIP, std::string("arrest: ") + hipGetErrorString(r));
    if (int rc = pull_status(e)) return rc;
  }
  if (final_status) slot_status(b, slot, final_status);
  const size_t n = (size_t)e->L.N * e->L.nf;
  if (F_out) {
    if (int rc = copy_out(e, F_out, e->F[e->cur], n)) return rc;
  }
  if (V_out) {
    if (int rc = copy_out(e, V_out, e->V[e->cur], n)) return rc;
  }
  for (int m = 0; m < b->nmaps; ++m)
    for (int s = b->first_block[m][slot]; s < b->first_block[m][slot] + b->nblocks[m][slot]; ++s) b->h_blk[m][s] = -1;
  b->slot[slot] = nullptr;
  release(e);
  std::memset(b->h_status + slot, 0, sizeof(DevStatus));
  b->dirty = true;
  return SKX_OK;
}

}  // extern "C"
