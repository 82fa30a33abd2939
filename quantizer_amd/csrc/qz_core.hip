// libqzhip.so: MI355X-native per-step particle-mesh gravity cycle.
//
// Host-side driver of the device kernels in qz_kernels.h / qz_branch_kernels.h
// and the native C ABI declared in include/qz.h.  One HIP stream per context;
// every entry point returns after its device work has finished, because the
// reference's host tree code reads results between the stages of its step
// loop (main.c:369-535).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qz.h"
#include "qz_engine.h"
#include "qz_internal.h"
#include "qz_kernels.h"

namespace qz
{

thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

bool hip_ok(hipError_t e, const char *what)
{
  if (e != hipSuccess)
  {
    set_error("%s: %s", what, hipGetErrorString(e));
    return false;
  }
  return true;
}

static inline unsigned nblk(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
// blocks of the grid-stride reduction kernels (one partial per block)
static inline unsigned nred(size_t n) { return std::min<unsigned>(nblk(n), 2048u); }

// ------------------------------------------------------------------ timing
void Timer::begin(hipStream_t s)
{
  if (!on)
    return;
  hipEvent_t a;
  (void)hipEventCreate(&a);
  (void)hipEventRecord(a, s);
  pending.push_back(a);
}

void Timer::end(hipStream_t s, const char *name, double bytes)
{
  if (!on || pending.empty())
    return;
  hipEvent_t b;
  (void)hipEventCreate(&b);
  (void)hipEventRecord(b, s);
  recs.push_back({name, pending.back(), b, bytes});
  pending.pop_back();
}

void Timer::host(const char *name, double ms)
{
  std::lock_guard<std::mutex> lock(host_mu);
  std::string nm = std::string("host:") + name;
  auto it = std::find_if(acc.begin(), acc.end(), [&](const Acc &x) { return x.name == nm; });
  if (it == acc.end())
    acc.push_back({nm, ms, 1, 0.0});
  else
  {
    it->ms += ms;
    it->count++;
  }
}

void Timer::collect()
{
  for (auto &r : recs)
  {
    (void)hipEventSynchronize(r.b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto it = std::find_if(acc.begin(), acc.end(), [&](const Acc &x) { return x.name == r.name; });
    if (it == acc.end())
      acc.push_back({r.name, (double)ms, 1, r.bytes});
    else
    {
      it->ms += ms;
      it->count++;
      it->bytes += r.bytes;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  recs.clear();
  for (hipEvent_t e : pending)
    (void)hipEventDestroy(e);
  pending.clear();
}

#define QZ_LAUNCH(name, kern, grid, ...) QZ_LAUNCHB(name, 0.0, kern, grid, __VA_ARGS__)
#define QZ_LAUNCH_T(name, threads, kern, grid, ...) QZ_LAUNCH_TB(name, threads, 0.0, kern, grid, __VA_ARGS__)
#define QZ_LAUNCH_TB(name, threads, bytes, kern, grid, ...)                                                           \
  do                                                                                                                  \
  {                                                                                                                   \
    timer.begin(stream);                                                                                              \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, stream, __VA_ARGS__);                                      \
    timer.end(stream, name, (double)(bytes));                                                                         \
  } while (0)
#define QZ_LAUNCHB(name, bytes, kern, grid, ...)                                                                      \
  do                                                                                                                  \
  {                                                                                                                   \
    timer.begin(stream);                                                                                              \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, __VA_ARGS__);                                       \
    timer.end(stream, name, (double)(bytes));                                                                         \
  } while (0)

// ------------------------------------------------------------------ engine
template <typename T> static T *dmalloc(size_t n, size_t &total)
{
  T *p = nullptr;
  if (n == 0)
    n = 1;
  if (hipMalloc((void **)&p, n * sizeof(T)) != hipSuccess)
    return nullptr;
  total += n * sizeof(T);
  return p;
}

template <typename P, typename G> struct Engine final : EngineBase
{
  // particles: pos[0..2] FP64 positions; pa[3..9] = vx vy vz ax ay az m in P
  double *pos[3] = {}, *posb[3] = {};
  P *pa[10] = {};
  P *pb[10] = {};
  int *id_a = nullptr, *id_b = nullptr;
  unsigned long long *key_in = nullptr, *key_out = nullptr;
  int *slot_in = nullptr, *slot_out = nullptr;
  int *cell_start = nullptr;
  G *tails = nullptr; // 8 corner sums per run tail (deposit stage 1)
  int *heavy = nullptr, *nheavy = nullptr; // cells of many windows (k_cell_totals) and their count
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  // head grid
  G *rho = nullptr;
  double *phi = nullptr;
  G *gax = nullptr, *gay = nullptr, *gaz = nullptr;
  std::vector<G *> lev_e, lev_r;
  std::vector<G *> lev_t; // levels below the top: the pre-smoothed correction between k_mg_down and k_mg_up
  std::vector<int> lev_n;
  bool mg_fused = true; // fused tile kernels on the levels below the top (QZ_HEAD_MG_FUSED=0 turns them off)
  bool node_mg_fused = true; // fused node levels k_bmg_down / k_bmg_up (QZ_NODE_MG_FUSED=0 turns them off)
  bool setup_ahead = true;   // first tree level's node set-up during the head solve (QZ_NODE_SETUP_AHEAD=0: in place)
  bool head_graphs = true;   // head outer iterations replayed as graphs (QZ_HEAD_GRAPH=0: launched one by one)
  bool batch_graphs = true;  // node PCG batches replayed as graphs (QZ_BATCH_GRAPH=0: launched one by one)
  bool host_refine = false;  // marks, zones and child boxes below the head on the host (QZ_HOST_TREE=1)
  double *part = nullptr; // reduction partials
  double *red = nullptr;  // device scalars
  double *hred = nullptr; // pinned host mirror
  size_t part_cap = 0;
  // asynchronous snapshots: a device copy of the sorted state, the pinned host
  // buffer it is downloaded into, and the host thread that writes the file
  struct SnapStage
  {
    double *pos[3] = {nullptr, nullptr, nullptr};
    P *vel[3] = {nullptr, nullptr, nullptr};
    int *id = nullptr;
    unsigned char *host = nullptr; // pinned: x, y, z, vx, vy, vz, id
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;
    std::thread writer;
    bool ok = true;
    std::string err;
  } snap;
  bool keys_valid = false;  // the particles are sorted and cell_start is current
  bool keys_fresh = false;  // the unsorted keys and slots of the current positions are in the sort's input
  double face_gm = 0.0; // G M of the head's monopole faces, fixed when the head is initialised
  bool acc_live = false;         // particle accelerations hold values (false: all zero, as reset leaves them)
  bool head_acc_queued = false;  // this pass's head accelerations are queued on the set-up stream
  // The next CFL time step folded into the kick: plev (one byte per sorted
  // particle) holds the finest level of each particle from the last gather,
  // valid (lev_ready) until the particles move, are reordered or reloaded;
  // dtpart holds the kick's per-block minima for CFL number dt_cfl, valid
  // (dt_ready) until the particles change or a time step consumes them.
  unsigned char *plev = nullptr;
  double *dtpart = nullptr;
  bool lev_ready = false, lev_head = false, dt_ready = false;
  double dt_cfl = 0.5;
  void particles_changed() { lev_ready = dt_ready = false; }
  // Re-sort of near-sorted keys (QZ_RESORT=0: a full radix sort every time):
  // the drift from a sorted state flags the particles whose key changed
  // (moved; moved_valid records that the last drift wrote the flags, which
  // describe the keys while keys_fresh holds), the sort then partitions, sorts
  // the movers and merges.  Scratch: block counts and offsets, partitioned and
  // sorted keys and slots, merge splits, a pinned count.
  bool resort = true, moved_valid = false;
  unsigned char *moved = nullptr;
  int *rs_cnt = nullptr, *rs_off = nullptr, *rs_split = nullptr, *rs_hcount = nullptr;
  int *rs_idx[3] = {nullptr, nullptr, nullptr};
  void *rs_key[3] = {nullptr, nullptr, nullptr};
  void *rs_scan = nullptr;
  size_t rs_scan_bytes = 0;
  int rs_movers = 0; // movers of the last re-sort
  bool resort_keys();
  std::vector<double> host_mass; // the masses as given (snapshots write them unchanged)
  double *phi_prev = nullptr; // the previous converged head potential
  int phi_hist = 0;           // converged head potentials since the last reset of the history
  // captured outer iteration of the head solve, per (phi, rhs) pair: [0] the
  // potential, [1] the adjoint solve
  // and each V-cycle count of an outer iteration (1 .. kMaxInner)
  static constexpr int kMaxInner = 4;
  hipGraphExec_t head_graph[2][kMaxInner] = {};
  unsigned char *refl_rec = nullptr; // when set, update_A records reflected axes (its slot order)
  // adjoint state (qz_adjoint.inl)
  double *adj_xb = nullptr, *adj_vb = nullptr, *adj_ab = nullptr, *adj_abA = nullptr, *adj_psi = nullptr;
  double *adj_psi_prev = nullptr; // the previous adjoint head solution of a reverse sweep
  int adj_hist = 0;               // adjoint head solutions of this reverse sweep (up to 2)
  // The cotangents (adj_xb, adj_vb, adj_ab, adj_abA) and the masses adj_m are
  // kept in the slot order of the state they belong to; adj_xb2, adj_vb2 and
  // adj_m2 receive them in the previous step's order.  adj_abP: abar as the
  // deposit's mass type (unused when P is double).
  double *adj_xb2 = nullptr, *adj_vb2 = nullptr;
  G *adj_tails[2] = {nullptr, nullptr}; // run tails of abar's y and z components (x: tails)
  P *adj_abP = nullptr, *adj_m = nullptr, *adj_m2 = nullptr;
  G *adj_gbar[3] = {nullptr, nullptr, nullptr}, *adj_rhs = nullptr;
  // refinement levels, in the grid-pool layout: Abar (3), phibar and psi in
  // FP64, the node solves' right-hand side, point kinds; the head phibar sent up
  // by level 1; transfer records (keys, values, sources) and their sort buffer
  double *adj_nab[3] = {nullptr, nullptr, nullptr}, *adj_npb = nullptr, *adj_npsi = nullptr, *adj_hpb = nullptr;
  G *adj_nrhs = nullptr;
  unsigned char *adj_kind = nullptr;
  long long adj_ncap = 0;
  unsigned *adj_rkey[2] = {nullptr, nullptr};
  int *adj_rval[2] = {nullptr, nullptr}, *adj_bpre = nullptr;
  long long *adj_rsrc = nullptr;
  void *adj_rtmp = nullptr;
  size_t adj_rtmp_bytes = 0;
  long long adj_rcap = 0, adj_bpre_cap = 0;
  struct Ckpt
  {
    double *x[3];
    P *a[7];
    int *id;
    double *phi, *phi_prev; // head potential and the previous one (the solve's starting guess)
    int hist;
    bool prev; // phi_prev holds the previous potential
  };
  struct StepRec
  {
    double *x[3];
    int *id;
    unsigned long long *key;
    int *cstart;
    int *perm; // the step's sort: slot s holds the particle of slot perm[s] before it
    G *g[3];
    unsigned char *refl;
    TreeSnap *tsnap; // with refinement levels: the tree and node fields of the step
  };
  std::vector<Ckpt> ckpts; // checkpoint buffers (kept for later runs); the first adj_nckpt are current
  int adj_nckpt = 0;
  Ckpt adj_fin{}; // the final state during a reverse sweep
  std::vector<StepRec> recs, recs2; // two record sets: one reversed while the other is recomputed
  // The reverse sweep's helper: a second engine that recomputes the segment
  // before the one being reversed, on a host thread and its own stream
  // (QZ_ADJ_OVERLAP=0: every segment is recomputed in turn on this engine).
  Engine *helper = nullptr;
  bool adj_overlap = true;
  int adj_nsteps = 0, adj_every = 0;
  bool adj_sweep = false; // a reverse sweep is running (its head solves warm-start)
  bool adj_warm = false;  // adj_psi holds the previous adjoint head solution of this reverse sweep
  bool adj_last_recorded = false; // adjoint_forward recorded the last segment's steps
  double adj_dt = 0.0;

  ~Engine() override { release(); }

  // ---------------------------------------------------------------- measurement
  // The head's finest red half sweep replayed `reps` times back to back between
  // two events on the engine stream (the level-0 correction array is scratch
  // between solves).  Average ms per launch and algorithmic bytes per launch.
  bool bench_sweep(int reps, double *ms, double *bytes) override
  {
    const size_t p1 = (size_t)n + 1, np = p1 * p1 * p1, nhalf = rbhalf(np);
    const Cube cf = make_cube(n + 1);
    const G h2 = (G)(1.0 / ((double)n * n));
    hipEvent_t a, b;
    if (!hip_ok(hipEventCreate(&a), "events") || !hip_ok(hipEventCreate(&b), "events"))
      return false;
    hipLaunchKernelGGL(k_rbgs<G>, dim3(nblk(nhalf)), dim3(kBlock), 0, stream, lev_e[0], lev_r[0], cf, h2, 0, 0);
    (void)hipEventRecord(a, stream);
    for (int i = 0; i < reps; i++)
      hipLaunchKernelGGL(k_rbgs<G>, dim3(nblk(nhalf)), dim3(kBlock), 0, stream, lev_e[0], lev_r[0], cf, h2, 0, 0);
    (void)hipEventRecord(b, stream);
    bool ok = hip_ok(hipEventSynchronize(b), "bench sweep");
    float t = 0.f;
    (void)hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = (double)t / std::max(1, reps);
    *bytes = 2.0 * sizeof(G) * (double)np; // own colour r, e read and written; the other colour's e once
    return ok;
  }

  // One copy of `bytes` bytes (rounded to doubles) through k_copy_probe.
  bool copy_probe(size_t bytes, double *ms) override
  {
    const size_t nd = bytes / sizeof(double);
    double *src = nullptr, *dst = nullptr;
    if (!hip_ok(hipMalloc((void **)&src, nd * sizeof(double)), "copy probe") ||
        !hip_ok(hipMalloc((void **)&dst, nd * sizeof(double)), "copy probe"))
    {
      (void)hipFree(src);
      return false;
    }
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(nd)), dim3(kBlock), 0, stream, src, 1.0, nd);
    (void)hipEventRecord(a, stream);
    hipLaunchKernelGGL(k_copy_probe, dim3(nblk(nd)), dim3(kBlock), 0, stream, src, dst, nd);
    (void)hipEventRecord(b, stream);
    bool ok = hip_ok(hipEventSynchronize(b), "copy probe");
    float t = 0.f;
    (void)hipEventElapsedTime(&t, a, b);
    *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(src);
    (void)hipFree(dst);
    return ok;
  }

  void release()
  {
    if (snap.writer.joinable())
      snap.writer.join();
    for (int a = 0; a < 3; a++)
    {
      (void)hipFree(snap.pos[a]);
      (void)hipFree(snap.vel[a]);
    }
    (void)hipFree(snap.id);
    if (snap.host)
      (void)hipHostFree(snap.host);
    if (snap.ready)
      (void)hipEventDestroy(snap.ready);
    if (snap.stream)
      (void)hipStreamDestroy(snap.stream);
    for (int a = 0; a < 3; a++)
    {
      (void)hipFree(pos[a]);
      (void)hipFree(posb[a]);
    }
    for (int a = 3; a < 10; a++)
    {
      (void)hipFree(pa[a]);
      (void)hipFree(pb[a]);
    }
    (void)hipFree(id_a);
    (void)hipFree(id_b);
    (void)hipFree(key_in);
    (void)hipFree(key_out);
    (void)hipFree(slot_in);
    (void)hipFree(slot_out);
    void *rs[] = {moved, rs_cnt, rs_off, rs_split, rs_idx[0], rs_idx[1], rs_idx[2], rs_key[0], rs_key[1], rs_key[2],
                  rs_scan};
    for (void *q : rs)
      (void)hipFree(q);
    if (rs_hcount)
      (void)hipHostFree(rs_hcount);
    (void)hipFree(cell_start);
    (void)hipFree(tails);
    (void)hipFree(heavy);
    (void)hipFree(nheavy);
    (void)hipFree(sort_tmp);
    (void)hipFree(rho);
    (void)hipFree(phi);
    (void)hipFree(gax);
    (void)hipFree(gay);
    (void)hipFree(gaz);
    for (size_t l = 1; l < lev_e.size(); l++)
    {
      (void)hipFree(lev_e[l]);
      (void)hipFree(lev_r[l]);
      (void)hipFree(lev_t[l]);
    }
    if (!lev_e.empty())
    {
      (void)hipFree(lev_e[0]);
      (void)hipFree(lev_r[0]);
    }
    (void)hipFree(part);
    (void)hipFree(red);
    (void)hipFree(phi_prev);
    if (hred)
      (void)hipHostFree(hred);
    for (int q = 0; q < 2; q++)
      for (int c = 0; c < kMaxInner; c++)
        if (head_graph[q][c])
          (void)hipGraphExecDestroy(head_graph[q][c]);
    join_ahead();
    release_adjoint();
    release_tree();
  }

  bool init() override
  {
    N = cfg.n_particles;
    n = 1 << cfg.lmin;
    D = std::min(cfg.lmax - cfg.lmin, 7);
    if (const char *f = std::getenv("QZ_HEAD_MG_FUSED"))
      mg_fused = f[0] != '0';
    if (const char *f = std::getenv("QZ_NODE_MG_FUSED"))
      node_mg_fused = f[0] != '0';
    if (const char *f = std::getenv("QZ_NODE_SETUP_AHEAD"))
      setup_ahead = f[0] != '0';
    if (const char *f = std::getenv("QZ_HEAD_GRAPH"))
      head_graphs = f[0] != '0';
    if (const char *f = std::getenv("QZ_BATCH_GRAPH"))
      batch_graphs = f[0] != '0';
    if (const char *f = std::getenv("QZ_HOST_TREE"))
      host_refine = f[0] == '1';
    if (const char *f = std::getenv("QZ_ADJ_OVERLAP"))
      adj_overlap = f[0] != '0';
    if (const char *f = std::getenv("QZ_RESORT"))
      resort = f[0] != '0';
    size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    for (int a = 0; a < 3; a++)
    {
      pos[a] = dmalloc<double>(N, bytes);
      posb[a] = dmalloc<double>(N, bytes);
      if (!pos[a] || !posb[a])
        return hip_ok(hipErrorOutOfMemory, "particle arrays");
    }
    for (int a = 3; a < 10; a++)
    {
      pa[a] = dmalloc<P>(N, bytes);
      pb[a] = dmalloc<P>(N, bytes);
      if (!pa[a] || !pb[a])
        return hip_ok(hipErrorOutOfMemory, "particle arrays");
    }
    id_a = dmalloc<int>(N, bytes);
    id_b = dmalloc<int>(N, bytes);
    key_in = dmalloc<unsigned long long>(N, bytes);
    key_out = dmalloc<unsigned long long>(N, bytes);
    slot_in = dmalloc<int>(N, bytes);
    slot_out = dmalloc<int>(N, bytes);
    cell_start = dmalloc<int>((size_t)n * n * n + 1, bytes);
    tails = dmalloc<G>(8 * (size_t)N, bytes);
    heavy = dmalloc<int>((size_t)N / (64 * (kTotalsSerial - 1)) + 16, bytes);
    nheavy = dmalloc<int>(1, bytes);
    end_bit = 3 * (cfg.lmin + D);
    sort_tmp_bytes = end_bit <= 32 ? qz_sort_temp_bytes(N, end_bit) : qz_sort64_temp_bytes(N, end_bit);
    sort_tmp = dmalloc<char>(sort_tmp_bytes, bytes);
    rho = dmalloc<G>(npts, bytes);
    phi = dmalloc<double>(npts, bytes);
    gax = dmalloc<G>(npts, bytes);
    gay = dmalloc<G>(npts, bytes);
    gaz = dmalloc<G>(npts, bytes);
    if (!id_a || !id_b || !key_in || !key_out || !slot_in || !slot_out || !cell_start || !tails || !heavy || !nheavy ||
        !sort_tmp || !rho || !phi || !gax || !gay || !gaz)
      return hip_ok(hipErrorOutOfMemory, "head arrays");
    // multigrid hierarchy: level 0 = head (n cells), halve down to 4 cells
    for (int m = n; m >= 4; m /= 2)
    {
      size_t p1 = (size_t)m + 1;
      lev_n.push_back(m);
      lev_e.push_back(dmalloc<G>(p1 * p1 * p1, bytes));
      lev_r.push_back(dmalloc<G>(p1 * p1 * p1, bytes));
      const bool fusable = m <= kMgFusedCells && m >= 16;
      lev_t.push_back(fusable ? dmalloc<G>(p1 * p1 * p1, bytes) : nullptr);
      if (!lev_e.back() || !lev_r.back() || (fusable && !lev_t.back()))
        return hip_ok(hipErrorOutOfMemory, "multigrid levels");
    }
    if (lev_n.back() > 8)
    {
      set_error("head grid too small for the multigrid hierarchy");
      return false;
    }
    for (size_t l = 0; l < lev_n.size(); l++)
    {
      size_t p1 = (size_t)lev_n[l] + 1, np = p1 * p1 * p1;
      QZ_LAUNCH("fill", k_fill<G>, nblk(np), lev_e[l], (G)0, np);
      QZ_LAUNCH("fill", k_fill<G>, nblk(np), lev_r[l], (G)0, np);
    }
    part_cap = std::max<size_t>(nblk(npts), 4 * (size_t)nblk(N)) + 16;
    part = dmalloc<double>(part_cap, bytes);
    plev = dmalloc<unsigned char>(N, bytes);
    dtpart = dmalloc<double>(nblk(N), bytes);
    if (!part || !plev || !dtpart)
      return hip_ok(hipErrorOutOfMemory, "reduction partials");
    red = dmalloc<double>(16, bytes);
    if (!hip_ok(hipHostMalloc((void **)&hred, 16 * sizeof(double)), "pinned scalars"))
      return false;
    QZ_LAUNCH("fill", k_fill<G>, nblk(npts), rho, (G)0, npts);
    return hip_ok(hipStreamSynchronize(stream), "init");
  }

  // init_head: also set the head potential and acceleration to their initial
  // monopole values (initialization.c:258); otherwise the head potential is kept
  // as the next solve's starting guess, as the reference never resets it.
  bool set_particles(const double *const *src, bool init_head) override
  {
    std::vector<P> tmp(N);
    for (int a = 0; a < 3; a++)
      if (!hip_ok(hipMemcpy(pos[a], src[a], sizeof(double) * N, hipMemcpyHostToDevice), "upload positions"))
        return false;
    for (int a = 3; a < 7; a++)
    {
      int dst = a < 6 ? a : 9;
      for (int i = 0; i < N; i++)
        tmp[i] = (P)src[a][i];
      if (!hip_ok(hipMemcpy(pa[dst], tmp.data(), sizeof(P) * N, hipMemcpyHostToDevice), "upload particles"))
        return false;
    }
    std::vector<int> ids(N);
    double mt = 0.0;
    host_mass.assign(src[6], src[6] + N);
    for (int i = 0; i < N; i++)
    {
      ids[i] = i;
      mt += src[6][i];
    }
    total_mass = mt;
    if (!hip_ok(hipMemcpy(id_a, ids.data(), sizeof(int) * N, hipMemcpyHostToDevice), "upload ids"))
      return false;
    for (int a = 6; a < 9; a++)
      QZ_LAUNCH("fill", k_fill<P>, nblk(N), pa[a], (P)0, (size_t)N);
    acc_live = false;
    particles_changed();
    keys_fresh = false;
    if (init_head)
    {
      phi_hist = 0;
      size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
      face_gm = cfg.G * total_mass;
      QZ_LAUNCH("init_head", k_init_head<G>, nblk(npts), phi, gax, gay, gaz, make_cube(n + 1), face_gm,
                cfg.head_bc == QZ_BC_ZERO ? 1 : 0);
    }
    keys_valid = false;
    return hip_ok(hipStreamSynchronize(stream), "set_particles");
  }

  bool get_particles(double *const *dst) override
  {
    std::vector<P> tmp(N);
    std::vector<int> ids(N);
    if (!hip_ok(hipMemcpy(ids.data(), id_a, sizeof(int) * N, hipMemcpyDeviceToHost), "download ids"))
      return false;
    std::vector<double> tx(N);
    for (int q = 0; q < 3; q++)
    {
      if (!dst[q])
        continue;
      if (!hip_ok(hipMemcpy(tx.data(), pos[q], sizeof(double) * N, hipMemcpyDeviceToHost), "download positions"))
        return false;
      for (int i = 0; i < N; i++)
        dst[q][ids[i]] = tx[i];
    }
    for (int q = 3; q < 9; q++)
    {
      if (!dst[q])
        continue;
      if (!hip_ok(hipMemcpy(tmp.data(), pa[q], sizeof(P) * N, hipMemcpyDeviceToHost), "download particles"))
        return false;
      for (int i = 0; i < N; i++)
        dst[q][ids[i]] = (double)tmp[i];
    }
    return true;
  }

  // ---------------------------------------------------------------- stages
  bool update_A(double dt) override
  {
    // the drift also writes the new positions' keys in the sort's input form,
    // and, from a sorted state, which particles changed key
    const bool narrow = end_bit <= 32;
    void *key = narrow ? (void *)key_out : (void *)key_in;
    const bool track = resort && keys_valid;
    if (track && !moved && !(moved = dmalloc<unsigned char>(N, bytes)))
      return hip_ok(hipErrorOutOfMemory, "moved flags");
    QZ_LAUNCHB("update_A", (6.0 * 8 + 9.0 * sizeof(P) + (narrow ? 8.0 : 12.0) + (track ? 1.0 : 0.0)) * N,
               k_update_A<P>, nblk(N), pos[0], pos[1], pos[2], pa[3], pa[4], pa[5], pa[6], pa[7], pa[8], N, dt,
               refl_rec, key, slot_in, cfg.lmin, D, narrow ? 1 : 0, track ? moved : nullptr);
    keys_valid = false;
    keys_fresh = true;
    particles_changed();
    moved_valid = track;
    return hip_ok(hipGetLastError(), "particle_updating_A"); // stream-ordered; readers synchronise
  }

  bool update_B(double dt) override
  {
    // with the levels of the last gather at hand, the kick also folds the next
    // CFL time step into per-block minima (one more byte read per particle)
    const bool fold = lev_ready;
    QZ_LAUNCHB("update_B", (9.0 * sizeof(P) + (fold && !lev_head ? 1.0 : 0.0)) * N, k_update_B<P>, nblk(N), pa[3],
               pa[4], pa[5], pa[6], pa[7], pa[8], N, (P)dt, lev_head ? nullptr : plev, cfg.lmin, dt_cfl,
               fold ? dtpart : nullptr);
    dt_ready = fold;
    return hip_ok(hipGetLastError(), "particle_updating_B"); // stream-ordered; readers synchronise
  }

  // reset's bookkeeping without its fills (see force_overlapped)
  bool reset_marks()
  {
    acc_live = false;
    particles_changed();
    return true;
  }

  bool reset() override
  {
    size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    for (int a = 6; a < 9; a++)
      QZ_LAUNCHB("reset", (double)sizeof(P) * N, k_fill<P>, nblk(N), pa[a], (P)0, (size_t)N);
    acc_live = false;
    particles_changed();
    QZ_LAUNCHB("reset", (double)sizeof(G) * npts, k_fill<G>, nblk(npts), rho, (G)0, npts);
    return hip_ok(hipGetLastError(), "reset"); // stream-ordered; readers synchronise
  }

  // Window tails of the cells [cb, ce) (ce null: cb[c + 1]) summed into each
  // cell's last slot, after k_deposit_runs and before any fold reads them.
  void cell_totals(const int *cb, const int *ce, long long ncell)
  {
    if (ncell <= 0)
      return;
    QZ_LAUNCH("cell_totals", k_fill<int>, 1, nheavy, 0, (size_t)1);
    QZ_LAUNCHB("cell_totals", 8.0 * ncell, k_cell_totals<G>, nblk((size_t)ncell), cb, ce, ncell, tails, heavy,
               nheavy); // the cell ranges; multi-window cells add their tails
    QZ_LAUNCH("cell_totals", k_cell_totals_heavy<G>, 1024, cb, ce, heavy, nheavy, tails);
  }

  // Sort particles by octree key and rebuild the head cell ranges.
  bool sort_particles()
  {
    if (keys_valid)
      return true;
    // keys fitting 32 bits are sorted as such (fewer bytes per pass), then widened
    if (!keys_fresh)
    {
      if (end_bit <= 32)
        QZ_LAUNCHB("keys", (3.0 * 8 + 8.0) * N, k_keys32<double>, nblk(N), pos[0], pos[1], pos[2],
                   (unsigned *)key_out, slot_in, N, cfg.lmin, D);
      else
        QZ_LAUNCHB("keys", (3.0 * 8 + 12.0) * N, k_keys<double>, nblk(N), pos[0], pos[1], pos[2], key_in, slot_in, N,
                   cfg.lmin, D);
    }
    const bool near = keys_fresh && moved_valid; // the drift's flags describe these keys
    keys_fresh = false;
    timer.begin(stream);
    if (near)
    {
      if (!resort_keys()) // leaves 64-bit sorted keys in key_out
        return false;
    }
    else if (end_bit <= 32)
    {
      if (!hip_ok(qz_sort_pairs(sort_tmp, sort_tmp_bytes, (const unsigned *)key_out, (unsigned *)key_in, slot_in,
                                slot_out, N, end_bit, stream),
                  "radix sort"))
        return false;
      QZ_LAUNCH("keys", k_widen_keys, nblk(N), (const unsigned *)key_in, key_out, N);
    }
    else if (!hip_ok(qz_sort64_pairs(sort_tmp, sort_tmp_bytes, key_in, key_out, slot_in, slot_out, N, end_bit, stream),
                     "radix sort"))
      return false;
    const double kb = end_bit <= 32 ? 4.0 : 8.0; // key bytes
    // full sort: keys and slots read and written once per pass (three passes
    // of 9 bits for 32-bit keys); re-sort: flags and keys read, partition
    // written, the movers sorted, the merge read and written
    if (near)
      timer.end(stream, "resort", (kb + 1.0 + 2.0 * (kb + 4.0)) * N + 6.0 * (kb + 4.0) * rs_movers);
    else
      timer.end(stream, "sort", 2.0 * (kb + 4.0) * N * (end_bit <= 32 ? 3.0 : (end_bit + 7) / 8));
    PtclArrays<P> src, dst;
    for (int a = 0; a < 3; a++)
    {
      src.x[a] = pos[a];
      dst.x[a] = posb[a];
    }
    for (int a = 0; a < 7; a++)
    {
      src.a[a] = pa[3 + a];
      dst.a[a] = pb[3 + a];
    }
    src.id = id_a;
    dst.id = id_b;
    const int move_acc = acc_live ? 1 : 0;
    QZ_LAUNCHB("permute", (6.0 * 8 + (move_acc ? 14.0 : 8.0) * sizeof(P) + 12.0) * N, k_permute<P>, nblk(N), src, dst,
               slot_out, N, move_acc);
    for (int a = 0; a < 3; a++)
      std::swap(pos[a], posb[a]);
    for (int a = 3; a < 10; a++)
      if (move_acc || a < 6 || a == 9)
        std::swap(pa[a], pb[a]);
    std::swap(id_a, id_b);
    particles_changed(); // reordered
    const size_t ncs = (size_t)n * n * n + 1;
    QZ_LAUNCH("cell_start", k_fill<int>, nblk(ncs), cell_start, -1, ncs);
    QZ_LAUNCHB("cell_start", 8.0 * N + 4.0 * n * n * n, k_cell_start, nblk((size_t)N + 1), key_out, N, D, n * n * n,
               cell_start);
    QZ_LAUNCH("cell_start", k_cell_start_gaps, nblk(ncs), key_out, N, D, n * n * n, cell_start);
    keys_valid = true;
    return true;
  }

  bool grid_density() override
  {
    if (!deposit_head())
      return false;
    if (has_tree() && !branch_density())
      return false;
    return hip_ok(hipGetLastError(), "grid_density"); // stream-ordered; readers synchronise
  }

  // The head's density (grid_density.c:35-111): CIC weights of every particle
  // summed per cell in slot order, then per grid point over its 8 cells.
  bool deposit_head()
  {
    if (!sort_particles())
      return false;
    size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    double H = 1.0 / n;
    double coeff = 4.0 * cfg.G * kPi / (H * H * H);
    QZ_LAUNCHB("deposit_runs", (24.0 + sizeof(P) + 8.0) * N, (k_deposit_runs<P, G>), nblk(N), key_out, pos[0], pos[1],
               pos[2], pa[9], N, cfg.lmin, 3 * D, n - 1, tails);
    cell_totals(cell_start, nullptr, (long long)n * n * n);
    QZ_LAUNCHB("deposit_fold", 4.0 * n * n * n + (double)sizeof(G) * npts, (k_deposit_head_fold<G, G>), nblk(npts),
               cell_start, tails, rho, make_cube(n + 1), coeff);
    return true;
  }

  // V-cycle on the error equation of level l; the correction starts from zero
  // (the first red half sweep ignores the array), so no fill is needed.
  void vcycle(size_t l, bool zero_start)
  {
    int m = lev_n[l];
    size_t p1 = (size_t)m + 1, np = p1 * p1 * p1;
    G h2 = (G)(1.0 / ((double)m * m));
    if (l + 1 == lev_n.size())
    {
      timer.begin(stream);
      hipLaunchKernelGGL(k_coarse_cg<G>, dim3(1), dim3(kBlock), 0, stream, lev_e[l], lev_r[l], m, 64);
      timer.end(stream, "coarse_cg");
      return;
    }
    // Below the top level the coarsest levels run in one workgroup (zero start
    // only), and the levels of 16 to kMgFusedCells cells (the top one too) as
    // two fused tile kernels each with the default sweep counts; larger levels
    // move enough data to keep the per-level kernels.
    if (l > 0 && zero_start && m <= Tail<G>::cells)
    {
      timer.begin(stream);
      if (m == 16)
        hipLaunchKernelGGL((k_vcycle_tail<G, 16>), dim3(1), dim3(kTailThreads), 0, stream, lev_e[l], lev_r[l],
                           cfg.pre_sweeps, cfg.post_sweeps);
      else
        hipLaunchKernelGGL((k_vcycle_tail<G, 8>), dim3(1), dim3(kTailThreads), 0, stream, lev_e[l], lev_r[l],
                           cfg.pre_sweeps, cfg.post_sweeps);
      timer.end(stream, "vcycle_tail", 2.0 * sizeof(G) * (m + 1.0) * (m + 1.0) * (m + 1.0)); // r in, e out
      return;
    }
    if (lev_t[l] && mg_fused && cfg.pre_sweeps == 2 && cfg.post_sweeps == 2)
    {
      const int T = m / (2 * kMgCT); // tiles per axis (m >= 16)
      const int mc = lev_n[l + 1];
      const size_t c1 = (size_t)mc + 1, nc = c1 * c1 * c1;
      const G h2 = (G)(1.0 / ((double)m * m));
      timer.begin(stream);
      if (zero_start)
        hipLaunchKernelGGL((k_mg_down<G, true>), dim3(T * T * T), dim3(kMgThreads), 0, stream, lev_e[l], lev_t[l],
                           lev_r[l], lev_r[l + 1], m, T, h2, (G)((double)m * m));
      else
        hipLaunchKernelGGL((k_mg_down<G, false>), dim3(T * T * T), dim3(kMgThreads), 0, stream, lev_e[l], lev_t[l],
                           lev_r[l], lev_r[l + 1], m, T, h2, (G)((double)m * m));
      timer.end(stream, "mg_down", (double)sizeof(G) * ((zero_start ? 2.0 : 3.0) * np + nc));
      vcycle(l + 1, true);
      timer.begin(stream);
      hipLaunchKernelGGL(k_mg_up<G>, dim3(T * T * T), dim3(kMgThreads), 0, stream, lev_e[l], lev_r[l], lev_t[l],
                         lev_e[l + 1], m, T, h2);
      timer.end(stream, "mg_up", (double)sizeof(G) * (3.0 * np + nc));
      return;
    }
    const size_t nhalf = rbhalf(np);
    const Cube cf = make_cube(m + 1);
    for (int sw = 0; sw < cfg.pre_sweeps; sw++)
    {
      int z = (zero_start && sw == 0) ? 1 : 0;
      QZ_LAUNCHB("rbgs", (z ? 1.0 : 2.0) * sizeof(G) * np, k_rbgs<G>, nblk(nhalf), lev_e[l], lev_r[l], cf, h2, 0, z);
      QZ_LAUNCHB("rbgs", 1.5 * sizeof(G) * np, k_rbgs<G>, nblk(np - nhalf), lev_e[l], lev_r[l], cf, h2, 1, 0);
    }
    int mc = lev_n[l + 1];
    size_t c1 = (size_t)mc + 1, nc = c1 * c1 * c1;
    const RestrictShape rs = restrict_shape(mc);
    timer.begin(stream);
    hipLaunchKernelGGL(k_restrict_residual<G>, dim3(rs.xt * rs.yt * rs.zt), dim3(rs.tx, rs.ty), 0, stream, lev_e[l],
                       lev_r[l], lev_r[l + 1], mc, rs.yt, rs.zt, rs.L, (G)((double)m * m));
    timer.end(stream, "restrict", (double)sizeof(G) * (1.5 * np + nc));
    vcycle(l + 1, true);
    QZ_LAUNCHB("prolong", (double)sizeof(G) * (2.0 * np + nc), k_prolong_correct<G>, nblk((size_t)mc * mc * mc),
               lev_e[l], lev_e[l + 1], make_cube(mc));
    for (int sw = 0; sw < cfg.post_sweeps; sw++)
    {
      QZ_LAUNCHB("rbgs", 2.0 * sizeof(G) * np, k_rbgs<G>, nblk(nhalf), lev_e[l], lev_r[l], cf, h2, 0, 0);
      QZ_LAUNCHB("rbgs", 1.5 * sizeof(G) * np, k_rbgs<G>, nblk(np - nhalf), lev_e[l], lev_r[l], cf, h2, 1, 0);
    }
  }

  // FP64 residual of Lap_h ph = rh into lev_r[0] and its sum of squares into red[0]
  void outer_residual(double *ph, const G *rh, double inv_h2)
  {
    const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    const unsigned nbr = outer_residual_blocks(n + 1);
    QZ_LAUNCH_TB("outer_residual", dim3(kResTX, kResTY), (8.0 + 2.0 * sizeof(G)) * npts, k_outer_residual<G>, nbr, ph, rh,
                 lev_r[0], n + 1, inv_h2, part);
    hipLaunchKernelGGL(k_reduce_final<0>, dim3(1), dim3(kBlock), 0, stream, part, (int)nbr, red);
  }

  // One outer iteration of a head solve Lap_h ph = rh after the first residual:
  // `inner` V-cycles in G on the correction, FP64 correction with the pointwise
  // change test, then the new FP64 residual and its sum of squares.
  void head_outer_body(double *ph, const G *rh, int inner)
  {
    size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    double H = 1.0 / n;
    unsigned nb = nred(npts);
    for (int c = 0; c < inner; c++)
      vcycle(0, c == 0);
    QZ_LAUNCHB("outer_correct", (16.0 + sizeof(G)) * npts, k_outer_correct<G>, nb, ph, lev_e[0], make_cube(n + 1), cfg.tol_change,
               part);
    hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kBlock), 0, stream, part, (int)nb, red + 1);
    outer_residual(ph, rh, 1.0 / (H * H));
  }

  bool launch_head_body(double *ph, const G *rh, int inner, int slot)
  {
    if (timer.on || !head_graphs)
    {
      head_outer_body(ph, rh, inner); // eager launches (so that each kernel is timed)
      return true;
    }
    hipGraphExec_t &ge = head_graph[slot][inner - 1];
    if (!ge)
    {
      hipGraph_t g = nullptr;
      if (!hip_ok(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "graph capture"))
        return false;
      head_outer_body(ph, rh, inner);
      const bool made = hip_ok(hipStreamEndCapture(stream, &g), "graph capture") &&
                        hip_ok(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "graph instantiate");
      if (g)
        (void)hipGraphDestroy(g);
      if (!made)
      {
        ge = nullptr;
        return false;
      }
    }
    return hip_ok(hipGraphLaunch(ge, stream), "graph launch");
  }

  // V-cycles of the next outer iteration (cfg.inner_cycles when set).  From the
  // relative residual res of the last one: once it is below tol, one V-cycle,
  // whose correction the pointwise change test measures (the potential solve
  // needs that pass; the adjoint solve has then converged); otherwise enough
  // V-cycles, at an assumed reduction of kVcycleRate each, to bring it to
  // kVerifyRes (the potential: where the change over one more V-cycle passes the
  // test) or below tol (the adjoint), and at most kMaxInner, which is where the
  // FP32 correction of the mixed path stops gaining.
  int next_inner(double res, bool adjoint) const
  {
    if (cfg.inner_cycles > 0)
      return std::min(cfg.inner_cycles, kMaxInner);
    if (res < cfg.tol)
      return 1;
    constexpr double kVcycleRate = 0.05, kVerifyRes = 1e-9;
    const double target = adjoint ? 0.5 * cfg.tol : kVerifyRes;
    if (!(res > target))
      return 1;
    const int c = (int)std::ceil(std::log(target / res) / std::log(kVcycleRate));
    return std::max(1, std::min(c, kMaxInner));
  }

  // Head solve: FP64 outer residual/correction around inner V-cycles in G.
  // Potential (adjoint = false): convergence follows poisson_error.c:290-347
  // (method 0): the RMS residual over interior points divided by 4 pi G
  // rho_mean, and the pointwise change of phi over the last outer iteration (the
  // reference compares with phi from iter_between_check_potential_solution = 5
  // V-cycles earlier).  Adjoint solve (zero faces): the RMS residual relative to
  // the RMS right-hand side; ph starts at 0, or (warm) at the previous adjoint
  // solution of the reverse sweep.
  bool solve_head(double *ph, const G *rh, bool adjoint, bool warm = false)
  {
    size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    double H = 1.0 / n;
    double inv_h2 = 1.0 / (H * H);
    double rhomean = 4.0 * cfg.G * kPi * total_mass / ((double)n * n * n * H * H * H);
    double ninterior = (double)(n - 1) * (n - 1) * (n - 1);
    unsigned nb = nred(npts);
    int outers = 0, vcycles = 0;
    double change = 1e300, res = 0.0, scale = rhomean;
    if (adjoint && warm)
    {
      QZ_LAUNCHB("outer_residual", (double)sizeof(G) * npts, k_interior_sumsq<G>, nb, rh, make_cube(n + 1), part);
      hipLaunchKernelGGL(k_reduce_final<0>, dim3(1), dim3(kBlock), 0, stream, part, (int)nb, red);
      if (!hip_ok(hipMemcpyAsync(hred, red, sizeof(double), hipMemcpyDeviceToHost, stream), "residual readback") ||
          !hip_ok(hipStreamSynchronize(stream), "residual"))
        return false;
      scale = std::sqrt(hred[0] / ninterior);
      if (scale == 0.0)
        scale = 1.0;
    }
    outer_residual(ph, rh, inv_h2);
    bool ok = false;
    for (int outer = 0;; outer++)
    {
      if (!hip_ok(hipMemcpyAsync(hred, red, 2 * sizeof(double), hipMemcpyDeviceToHost, stream), "residual readback") ||
          !hip_ok(hipStreamSynchronize(stream), "residual"))
        return false;
      double rms = std::sqrt(hred[0] / ninterior);
      if (adjoint && !warm && outer == 0)
        scale = rms > 0.0 ? rms : 1.0;
      res = rms / scale;
      if (outer > 0)
        change = hred[1];
      if (adjoint ? (res < cfg.tol || rms == 0.0) : (outer > 0 && change <= 1.0 && res < cfg.tol))
      {
        ok = true; // change flag values > 1 mark points outside [1-tol, 1+tol]
        break;
      }
      if (vcycles >= cfg.max_iterations)
        break;
      const int inner = next_inner(res, adjoint);
      if (!launch_head_body(ph, rh, inner, adjoint ? 1 : 0))
        return false;
      vcycles += inner;
      outers++;
    }
    if (!adjoint)
    {
      stats.head_outer = outers;
      stats.head_vcycles = vcycles;
      stats.head_residual = res;
      stats.head_change = change;
    }
    if (!ok)
      set_error("The precision was not reached in the %s solve after %d V-cycles (residual %.3e, change %.3e)",
                adjoint ? "adjoint head" : "parent node", vcycles, res, change);
    return ok;
  }

  // Head potential.  The solve starts from the linear extrapolation of the last
  // two converged potentials once two exist (the reference starts from the last
  // one, reset.c:54-57); the stopping criteria are the same either way.
  bool potential_head() override
  {
    HostSpan span(timer, "potential_head");
    const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    if (phi_hist == 2)
      QZ_LAUNCHB("extrapolate", 32.0 * npts, k_extrapolate, nblk(npts), phi, phi_prev, npts);
    const bool ok = solve_head(phi, rho, false);
    if (ok && phi_hist == 0)
    {
      if (!phi_prev)
        phi_prev = dmalloc<double>(npts, bytes);
      if (!phi_prev || !hip_ok(hipMemcpyAsync(phi_prev, phi, sizeof(double) * npts, hipMemcpyDeviceToDevice, stream),
                               "potential history"))
        return false;
    }
    if (ok && phi_hist < 2)
      phi_hist++;
    return ok;
  }

  void grid_acc_head(hipStream_t stream)
  {
    size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    double H = 1.0 / n;
    QZ_LAUNCHB("grid_acc_head", (8.0 + 3.0 * sizeof(G)) * npts, k_grid_acc_head<G>, nblk(npts), phi, gax, gay, gaz,
               make_cube(n + 1), 1.0 / (2.0 * H), face_gm, cfg.head_bc == QZ_BC_ZERO ? 1 : 0);
  }

  bool grid_acceleration() override
  {
    if (!head_acc_queued)
      grid_acc_head(stream);
    else if (!wait_head_acc())
      return false;
    if (has_tree() && !branch_grid_acceleration())
      return false;
    return hip_ok(hipGetLastError(), "grid_acceleration"); // stream-ordered; readers synchronise
  }

  bool particle_acceleration() override
  {
    if (!sort_particles())
      return false;
    const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    if (has_tree())
    {
      if (!branch_particle_acceleration())
        return false;
    }
    else
    {
      QZ_LAUNCHB("gather_head", (24.0 + 3.0 * sizeof(P)) * N + 3.0 * sizeof(G) * npts, (k_gather_head<P, G>), nblk(N), pos[0], pos[1], pos[2], gax, gay, gaz, pa[6], pa[7],
                pa[8], N, n);
      lev_head = true; // every particle is on the head level
    }
    acc_live = true;
    lev_ready = true;
    dt_ready = false; // new accelerations
    return hip_ok(hipGetLastError(), "particle_acceleration"); // stream-ordered; readers synchronise
  }

  bool head_fields(double *out5) override
  {
    size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    std::vector<G> tmp(npts);
    G *src[5] = {rho, nullptr, gax, gay, gaz};
    for (int a = 0; a < 5; a++)
    {
      if (a == 1)
      {
        if (!hip_ok(hipMemcpy(out5 + npts, phi, sizeof(double) * npts, hipMemcpyDeviceToHost), "download phi"))
          return false;
        continue;
      }
      if (!hip_ok(hipMemcpy(tmp.data(), src[a], sizeof(G) * npts, hipMemcpyDeviceToHost), "download field"))
        return false;
      for (size_t i = 0; i < npts; i++)
        out5[a * npts + i] = (double)tmp[i];
    }
    return true;
  }

  bool set_head_potential(const double *pot) override
  {
    size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
    phi_hist = 0;
    return hip_ok(hipMemcpy(phi, pot, sizeof(double) * npts, hipMemcpyHostToDevice), "upload phi");
  }

  bool observables(int exact_pe, double *out6) override
  {
    unsigned nb = nblk(N);
    QZ_LAUNCH("kin_mom", k_kin_mom<P>, nb, pa[3], pa[4], pa[5], pa[9], N, part);
    for (int q = 0; q < 4; q++)
      hipLaunchKernelGGL(k_reduce_final<0>, dim3(1), dim3(kBlock), 0, stream, part + q * nb, (int)nb, red + q);
    if (!hip_ok(hipMemcpyAsync(hred, red, 4 * sizeof(double), hipMemcpyDeviceToHost, stream), "readback") ||
        !hip_ok(hipStreamSynchronize(stream), "observables"))
      return false;
    double ek = 0.5 * hred[0], px = hred[1], py = hred[2], pz = hred[3];
    if (exact_pe)
      QZ_LAUNCH("pairwise_pe", k_pairwise_pe<P>, nb, pos[0], pos[1], pos[2], pa[9], N, part);
    else if (has_tree())
      branch_grid_pe(nb);
    else
      QZ_LAUNCH("grid_pe", k_grid_pe_head<P>, nb, pos[0], pos[1], pos[2], pa[9], phi, N, n, part);
    hipLaunchKernelGGL(k_reduce_final<0>, dim3(1), dim3(kBlock), 0, stream, part, (int)nb, red);
    if (!hip_ok(hipMemcpyAsync(hred, red, sizeof(double), hipMemcpyDeviceToHost, stream), "readback") ||
        !hip_ok(hipStreamSynchronize(stream), "observables"))
      return false;
    double ep = exact_pe ? -cfg.G * hred[0] : 0.5 * cfg.G * hred[0];
    out6[0] = ek;
    out6[1] = ep;
    out6[2] = ek + ep;
    out6[3] = px;
    out6[4] = py;
    out6[5] = pz;
    return true;
  }

  bool timestep(double cfl, double dt_max, double *dt) override;

  // Snapshot file in the reference layout (output_snapshots.c:141-226).  The
  // kinetic energy and momentum are FP64 sums on the host over the downloaded
  // velocities; the exact pairwise potential energy comes from the device.
  bool write_snapshot(const char *path, const qz_units &u, double actualtime, bool with_obs) override
  {
    std::vector<double> x(N), y(N), z(N), vx(N), vy(N), vz(N);
    double *dst[9] = {x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), nullptr, nullptr, nullptr};
    if (!get_particles(dst))
      return false;
    double en[3] = {0, 0, 0}, mom[3] = {0, 0, 0};
    if (with_obs)
    {
      double o6[6];
      if (!observables(1, o6))
        return false;
      double ek = 0.0; // observables.c:38-75 with the unit conversion
      for (int i = 0; i < N; i++)
      {
        const double cx = vx[i] / u.velocity + u.cm_v[0], cy = vy[i] / u.velocity + u.cm_v[1],
                     cz = vz[i] / u.velocity + u.cm_v[2];
        ek += host_mass[i] * (cx * cx + cy * cy + cz * cz);
        mom[0] += host_mass[i] * cx;
        mom[1] += host_mass[i] * cy;
        mom[2] += host_mass[i] * cz;
      }
      en[0] = 0.5 * ek;
      en[1] = o6[1] * (u.time * u.time) / (u.dist * u.dist); // observables.c:110
      en[2] = en[0] + en[1];
    }
    const double t = actualtime / u.time;
    for (int i = 0; i < N; i++)
    {
      x[i] = (x[i] - 0.5) / u.dist + u.cm[0] + u.cm_v[0] * t;
      y[i] = (y[i] - 0.5) / u.dist + u.cm[1] + u.cm_v[1] * t;
      z[i] = (z[i] - 0.5) / u.dist + u.cm[2] + u.cm_v[2] * t;
      vx[i] = vx[i] / u.velocity + u.cm_v[0];
      vy[i] = vy[i] / u.velocity + u.cm_v[1];
      vz[i] = vz[i] / u.velocity + u.cm_v[2];
    }
    FILE *f = fopen(path, "wb");
    if (!f)
    {
      set_error("Error opening snapshot file %s", path);
      return false;
    }
    const double *arr[7] = {host_mass.data(), x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data()};
    bool ok = true;
    for (int a = 0; a < 7; a++)
      ok = ok && fwrite(arr[a], sizeof(double), (size_t)N, f) == (size_t)N;
    if (with_obs)
      ok = ok && fwrite(en, sizeof(double), 3, f) == 3 && fwrite(mom, sizeof(double), 3, f) == 3;
    ok = ok && fwrite(&t, sizeof(double), 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    if (!ok)
      set_error("writing snapshot %s failed", path);
    return ok;
  }

  // The snapshot file of write_snapshot from a host copy of the sorted state
  // (positions, velocities, IDs), with the observables block given by the caller.
  bool write_snapshot_file(const char *path, const qz_units &u, double actualtime, const double *obs6,
                           const unsigned char *host)
  {
    const double *hx[3], *hdbl = (const double *)host;
    for (int a = 0; a < 3; a++)
      hx[a] = hdbl + (size_t)a * N;
    const P *hv = (const P *)(host + 24 * (size_t)N);
    const int *hid = (const int *)(host + (24 + 3 * sizeof(P)) * (size_t)N);
    std::vector<double> out(6 * (size_t)N);
    const double t = actualtime / u.time;
    for (int i = 0; i < N; i++)
    {
      const size_t q = (size_t)hid[i];
      for (int a = 0; a < 3; a++)
      {
        out[(size_t)a * N + q] = (hx[a][i] - 0.5) / u.dist + u.cm[a] + u.cm_v[a] * t;
        out[(size_t)(3 + a) * N + q] = (double)hv[(size_t)a * N + i] / u.velocity + u.cm_v[a];
      }
    }
    FILE *f = fopen(path, "wb");
    if (!f)
    {
      set_error("Error opening snapshot file %s", path);
      return false;
    }
    bool ok = fwrite(host_mass.data(), sizeof(double), (size_t)N, f) == (size_t)N &&
              fwrite(out.data(), sizeof(double), 6 * (size_t)N, f) == 6 * (size_t)N;
    if (obs6)
      ok = ok && fwrite(obs6, sizeof(double), 6, f) == 6;
    ok = ok && fwrite(&t, sizeof(double), 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    if (!ok)
      set_error("writing snapshot %s failed", path);
    return ok;
  }

  // Snapshot without stalling the step loop: the state is copied device to
  // device on the engine's stream (ordered after the work queued so far), then
  // a host thread downloads the copy on a stream of its own and writes the
  // file while later steps run.  One snapshot is in flight at a time.
  bool write_snapshot_async(const char *path, const qz_units &u, double actualtime, const double *obs6) override
  {
    if (!snapshot_wait())
      return false;
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "snapshot device"))
      return false;
    const size_t host_bytes = (24 + 3 * sizeof(P) + 4) * (size_t)N;
    if (!snap.id)
    {
      size_t b = 0;
      for (int a = 0; a < 3; a++)
      {
        snap.pos[a] = dmalloc<double>(N, b);
        snap.vel[a] = dmalloc<P>(N, b);
      }
      snap.id = dmalloc<int>(N, b);
      bytes += b;
      if (!snap.id || !snap.pos[2] || !snap.vel[2] ||
          !hip_ok(hipHostMalloc((void **)&snap.host, host_bytes), "snapshot staging") ||
          !hip_ok(hipStreamCreateWithFlags(&snap.stream, hipStreamNonBlocking), "snapshot stream") ||
          !hip_ok(hipEventCreateWithFlags(&snap.ready, hipEventDisableTiming), "snapshot event"))
        return hip_ok(hipErrorOutOfMemory, "snapshot staging");
    }
    for (int a = 0; a < 3; a++)
      if (!hip_ok(hipMemcpyAsync(snap.pos[a], pos[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream), "snapshot") ||
          !hip_ok(hipMemcpyAsync(snap.vel[a], pa[3 + a], sizeof(P) * N, hipMemcpyDeviceToDevice, stream), "snapshot"))
        return false;
    if (!hip_ok(hipMemcpyAsync(snap.id, id_a, sizeof(int) * N, hipMemcpyDeviceToDevice, stream), "snapshot") ||
        !hip_ok(hipEventRecord(snap.ready, stream), "snapshot"))
      return false;
    std::string file(path);
    std::vector<double> obs(obs6 ? obs6 : nullptr, obs6 ? obs6 + 6 : nullptr);
    snap.ok = true;
    snap.err.clear();
    snap.writer = std::thread([this, dev, file, u, actualtime, obs]() {
      unsigned char *h = snap.host;
      bool ok = hip_ok(hipSetDevice(dev), "snapshot device") &&
                hip_ok(hipStreamWaitEvent(snap.stream, snap.ready, 0), "snapshot");
      for (int a = 0; a < 3 && ok; a++)
        ok = hip_ok(hipMemcpyAsync(h + 8 * (size_t)a * N, snap.pos[a], sizeof(double) * N, hipMemcpyDeviceToHost,
                                   snap.stream),
                    "snapshot download") &&
             hip_ok(hipMemcpyAsync(h + (24 + sizeof(P) * a) * (size_t)N, snap.vel[a], sizeof(P) * N,
                                   hipMemcpyDeviceToHost, snap.stream),
                    "snapshot download");
      ok = ok &&
           hip_ok(hipMemcpyAsync(h + (24 + 3 * sizeof(P)) * (size_t)N, snap.id, sizeof(int) * N,
                                 hipMemcpyDeviceToHost, snap.stream),
                  "snapshot download") &&
           hip_ok(hipStreamSynchronize(snap.stream), "snapshot download") &&
           write_snapshot_file(file.c_str(), u, actualtime, obs.empty() ? nullptr : obs.data(), h);
      snap.ok = ok;
      if (!ok)
        snap.err = g_last_error; // errors are kept per thread
    });
    return true;
  }

  bool snapshot_wait() override
  {
    if (snap.writer.joinable())
      snap.writer.join();
    if (!snap.ok)
    {
      set_error("%s", snap.err.c_str());
      snap.ok = true;
      return false;
    }
    return true;
  }

  // ---------------------------------------------------------------- adjoint (qz_adjoint.inl)
  void release_adjoint();
  bool ensure_adjoint();
  bool ensure_adjoint_tree();
  bool force_fields();
  bool vjp_tree(const double *ab);
  bool slot_cotangents(const double *in, double *out, bool to_slots);
  bool save_ckpt(Ckpt &c);
  void restore_ckpt(const Ckpt &c);
  void free_ckpt(Ckpt &c);
  bool ensure_recs(std::vector<StepRec> &recs);
  void free_recs(std::vector<StepRec> &v);
  bool recompute_segment(const Ckpt &ck, std::vector<StepRec> &rs, int c0, int c1, double dt);
  bool make_helper();
  bool snap_tree(TreeSnap *&t);
  bool restore_tree(const TreeSnap *t);
  void free_tree_snap(TreeSnap *&t);
  bool save_rec(StepRec &r, bool handover = false);
  bool load_rec(StepRec &r);
  bool adjoint_head_solve();
  const P *abar_as_mass(const double *ab);
  void scatter_runs3(const P *abm, int lv);
  void cell_totals3(const int *cb, const int *ce, long long ncell);
  bool vjp_at_current(const double *ab); // xbar (adj_xb) += J_A(x)^T ab, both in the current slot order
  bool force_vjp(const double *abar, double *xbar) override;
  bool adjoint_forward(int nsteps, double dt, int every) override;
  bool adjoint_backward(const double *xbN, const double *vbN, double *xb0, double *vb0) override;

  // ---------------------------------------------------------------- branches (qz_branch.inl)
  bool has_tree() const { return tree != nullptr && tree->n_nodes > 0; }
  bool tree_adaptation() override;
  bool refine_head_device(int &zone_cells, hipStream_t s);
  bool zones_in_order(int M, int nz, bool unsorted, hipStream_t s);
  bool refine_level_device(long long c0, long long ncell, long long b0, long long nb, int &zone_cells, hipStream_t s);
  bool tree_state();
  bool build_tree(hipStream_t s);
  bool force_overlapped(bool with_particles) override;
  bool potential_branches() override;
  bool branch_cg_level(size_t l);
  bool branch_mg_level(size_t l, bool &solved, double *phi_pool, const G *rho_pool, bool adjoint);
  bool bmg_prepare(size_t l, hipStream_t stream);
  bool join_ahead() override;
  bool start_ahead(bool build);
  bool ensure_bmg_buffers(int n_itab, long long n_chunks, long long pool, hipStream_t stream);
  template <typename F>
  bool iterate_batches(F &&batch, int per_batch, int max_it, int count, int &it, bool &all_done,
                       hipGraphExec_t given = nullptr);
  void bmg_batch(size_t l, const BmgPlan &pl, hipStream_t stream, double *phi_pool, bool adjoint);
  bool branch_density();
  bool branch_grid_acceleration();
  bool wait_head_acc();
  bool branch_particle_acceleration();
  void branch_grid_pe(unsigned nb); // partial sums of m phi at the finest node into part
  void release_tree();
  int num_nodes() override { return tree ? tree->n_nodes : 0; }
  bool node_header(int i, int *hdr) override;
  bool node_lists(int i, int *ixyz, int *bxyz, int *sxyz) override;
  bool node_fields(int i, double *out5) override;
};

} // namespace qz

#include "qz_branch.inl"
#include "qz_adjoint.inl"

namespace qz
{

EngineBase *make_engine(const qz_config &cfg)
{
  EngineBase *e = nullptr;
  if (cfg.precision == QZ_PREC_MIXED)
    e = new Engine<float, float>();
  else
    e = new Engine<double, double>();
  e->cfg = cfg;
  return e;
}

} // namespace qz

// ======================================================================== C ABI
using qz::EngineBase;

struct qz_ctx
{
  EngineBase *eng;
};

// ------------------------------------------------- diagnostics without allocation
// qz_observables_noalloc keeps its partial sums in this fixed array instead of
// memory it allocates: kNoAllocBlocks x 4 block partials (E_kin and momentum;
// the pairwise energy reuses the first kNoAllocBlocks), then 16 doubles for the
// reduced sums and the six results.  Each device holds its own copy from module
// load until the process ends.  Host code reaches it only through the device
// address hipGetSymbolAddress returns (the symbol's host-side name is not a
// device address).
namespace
{
constexpr int kNoAllocBlocks = QZ_NOALLOC_MAX_PARTICLES / qz::kBlock;
constexpr size_t kNoAllocDoubles = 4 * (size_t)kNoAllocBlocks + 16;
__device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
double *scratch_for_current_device()
{
  int dev = -1;
  if (!qz::hip_ok(hipGetDevice(&dev), "qz_observables_noalloc: device"))
    return nullptr;
  if (dev < 0 || dev >= kMaxDevices)
  {
    qz::set_error("qz_observables_noalloc: device %d outside the supported range", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev])
  {
    void *p = nullptr;
    if (!qz::hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "qz_observables_noalloc: scratch address"))
      return nullptr;
    g_scratch_ptr[dev] = (double *)p;
  }
  return g_scratch_ptr[dev];
}

// out6 = E_kin, E_pot, E_tot, px, py, pz from the reduced sums (observables.c:38-112):
// sums[0] = sum m v^2, sums[1..3] = sum m v, sums[4] = pairwise sum m_i m_j / r_ij.
// Rounded operations keep the arithmetic identical to qz_observables' host code.
__global__ void k_observables_finish(const double *__restrict__ sums, double G, int with_pe,
                                     double *__restrict__ out6)
{
  if (threadIdx.x != 0)
    return;
  const double ek = 0.5 * sums[0];
  const double ep = with_pe ? __dmul_rn(-G, sums[4]) : 0.0;
  out6[0] = ek;
  out6[1] = ep;
  out6[2] = __dadd_rn(ek, ep);
  out6[3] = sums[1];
  out6[4] = sums[2];
  out6[5] = sums[3];
}
} // namespace

extern "C" {

void qz_default_config(qz_config *c)
{
  c->lmin = 5;
  c->lmax = 5;
  c->n_particles = 0;
  c->precision = QZ_PREC_FP64;
  c->head_bc = QZ_BC_MONOPOLE;
  c->ref_criterion_ptcl = 4;
  c->n_exp = 1;
  c->G = 1.0;
  c->tol = 1e-10;
  c->tol_change = 1e-10;
  c->max_iterations = 5000;
  c->inner_cycles = 0;
  c->pre_sweeps = 2;
  c->post_sweeps = 2;
  c->branch_solver = QZ_BRANCH_MG;
}

const char *qz_last_error(void) { return qz::g_last_error.c_str(); }

qz_ctx *qz_create(const qz_config *cfg)
{
  if (!cfg || cfg->lmin < 2 || cfg->lmin > 10 || cfg->lmax < cfg->lmin || cfg->n_particles <= 0)
  {
    qz::set_error("qz_create: invalid configuration");
    return nullptr;
  }
  if (cfg->n_exp != 1)
  {
    qz::set_error("qz_create: only n_exp = 1 is supported");
    return nullptr;
  }
  if (cfg->lmax - cfg->lmin > 7)
  {
    qz::set_error("qz_create: at most 7 refinement levels are supported");
    return nullptr;
  }
  int dev = 0;
  if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0)
  {
    qz::set_error("qz_create: no HIP device available");
    return nullptr;
  }
  qz_config c = *cfg;
  if (c.inner_cycles < 0)
    c.inner_cycles = 0; // 0: V-cycles per outer iteration chosen from its residual
  if (c.pre_sweeps <= 0)
    c.pre_sweeps = 2;
  if (c.post_sweeps <= 0)
    c.post_sweeps = 2;
  EngineBase *e = qz::make_engine(c);
  if (!qz::hip_ok(hipStreamCreate(&e->stream), "stream") || !e->init())
  {
    delete e;
    return nullptr;
  }
  qz_ctx *ctx = new qz_ctx{e};
  return ctx;
}

void qz_destroy(qz_ctx *ctx)
{
  if (!ctx)
    return;
  if (ctx->eng)
  {
    (void)hipStreamSynchronize(ctx->eng->stream);
    hipStream_t s = ctx->eng->stream;
    ctx->eng->timer.collect();
    delete ctx->eng;
    (void)hipStreamDestroy(s);
  }
  delete ctx;
}

#define QZ_CHECK_CTX(ctx)                                                                                             \
  if (!(ctx) || !(ctx)->eng)                                                                                          \
  {                                                                                                                   \
    qz::set_error("null context");                                                                                    \
    return 1;                                                                                                         \
  }

static int load_particles(qz_ctx *ctx, const double *x, const double *y, const double *z, const double *vx,
                          const double *vy, const double *vz, const double *m, bool init_head, const char *who)
{
  QZ_CHECK_CTX(ctx);
  const double *src[7] = {x, y, z, vx, vy, vz, m};
  for (int a = 0; a < 7; a++)
    if (!src[a])
    {
      qz::set_error("%s: null array", who);
      return 1;
    }
  for (int i = 0; i < ctx->eng->N; i++)
  {
    if (!(x[i] >= 0.0 && x[i] < 1.0 && y[i] >= 0.0 && y[i] < 1.0 && z[i] >= 0.0 && z[i] < 1.0))
    {
      qz::set_error("The input particle ID = %d is outside of the simulation box", i);
      return 1;
    }
  }
  return ctx->eng->set_particles(src, init_head) ? 0 : 1;
}

int qz_set_particles(qz_ctx *ctx, const double *x, const double *y, const double *z, const double *vx,
                     const double *vy, const double *vz, const double *m)
{
  return load_particles(ctx, x, y, z, vx, vy, vz, m, true, "qz_set_particles");
}

int qz_load_particles(qz_ctx *ctx, const double *x, const double *y, const double *z, const double *vx,
                      const double *vy, const double *vz, const double *m)
{
  return load_particles(ctx, x, y, z, vx, vy, vz, m, false, "qz_load_particles");
}

int qz_get_particles(qz_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy, double *vz, double *ax,
                     double *ay, double *az)
{
  QZ_CHECK_CTX(ctx);
  double *dst[9] = {x, y, z, vx, vy, vz, ax, ay, az};
  return ctx->eng->get_particles(dst) ? 0 : 1;
}

int qz_particle_updating_A(qz_ctx *ctx, double dt)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->update_A(dt) ? 0 : 1;
}
int qz_particle_updating_B(qz_ctx *ctx, double dt)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->update_B(dt) ? 0 : 1;
}
int qz_reset(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->reset() ? 0 : 1;
}
int qz_tree_adaptation(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  if (ctx->eng->cfg.lmin == ctx->eng->cfg.lmax)
    return 0;
  return ctx->eng->tree_adaptation() ? 0 : 1;
}
int qz_grid_density(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->grid_density() ? 0 : 1;
}
int qz_potential_head_node(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->potential_head() ? 0 : 1;
}
int qz_potential_branches(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->potential_branches() ? 0 : 1;
}
int qz_potential(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  if (!ctx->eng->potential_head())
    return 1;
  if (ctx->eng->cfg.lmin < ctx->eng->cfg.lmax && !ctx->eng->potential_branches())
    return 1;
  return 0;
}
int qz_grid_acceleration(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->grid_acceleration() ? 0 : 1;
}
int qz_particle_acceleration(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->particle_acceleration() ? 0 : 1;
}

int qz_force(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  // the stages of main.c:296-338 in order; the tree below the head is built
  // while the head density and potential are computed (same results)
  return ctx->eng->force_overlapped(true) ? 0 : 1;
}

int qz_step(qz_ctx *ctx, double dt)
{
  QZ_CHECK_CTX(ctx);
  if (qz_particle_updating_A(ctx, dt) || qz_force(ctx) || qz_particle_updating_B(ctx, dt))
    return 1;
  return 0;
}

int qz_get_head_fields(qz_ctx *ctx, double *out5)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->head_fields(out5) ? 0 : 1;
}

int qz_set_head_potential(qz_ctx *ctx, const double *pot)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->set_head_potential(pot) ? 0 : 1;
}

int qz_num_nodes(qz_ctx *ctx)
{
  if (!ctx || !ctx->eng)
    return 0;
  return ctx->eng->num_nodes();
}

int qz_node_header(qz_ctx *ctx, int i, int *hdr24)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->node_header(i, hdr24) ? 0 : 1;
}

int qz_node_grid_lists(qz_ctx *ctx, int i, int *intr_xyz, int *bdry_xyz, int *sim_xyz)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->node_lists(i, intr_xyz, bdry_xyz, sim_xyz) ? 0 : 1;
}

int qz_node_fields(qz_ctx *ctx, int i, double *out5)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->node_fields(i, out5) ? 0 : 1;
}

int qz_get_stats(qz_ctx *ctx, qz_stats *st)
{
  QZ_CHECK_CTX(ctx);
  *st = ctx->eng->stats;
  return 0;
}

int qz_observables(qz_ctx *ctx, int exact_pe, double *out6)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->observables(exact_pe, out6) ? 0 : 1;
}

int qz_timestep(qz_ctx *ctx, double cfl, double dt_max, double *dt)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->timestep(cfl, dt_max, dt) ? 0 : 1;
}

int qz_force_vjp(qz_ctx *ctx, const double *abar, double *xbar)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->force_vjp(abar, xbar) ? 0 : 1;
}

int qz_adjoint_forward(qz_ctx *ctx, int nsteps, double dt, int checkpoint_every)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->adjoint_forward(nsteps, dt, checkpoint_every) ? 0 : 1;
}

int qz_adjoint_backward(qz_ctx *ctx, const double *xbar_n, const double *vbar_n, double *xbar_0, double *vbar_0)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->adjoint_backward(xbar_n, vbar_n, xbar_0, vbar_0) ? 0 : 1;
}

int qz_keys(const double *x, const double *y, const double *z, int n, int lv, const int *ts, const int *rd, int *out)
{
  if (n <= 0)
    return 0;
  double *dx = nullptr, *dy = nullptr, *dz = nullptr;
  int *dout = nullptr;
  size_t b = sizeof(double) * (size_t)n;
  bool ok = qz::hip_ok(hipMalloc((void **)&dx, b), "alloc") && qz::hip_ok(hipMalloc((void **)&dy, b), "alloc") &&
            qz::hip_ok(hipMalloc((void **)&dz, b), "alloc") &&
            qz::hip_ok(hipMalloc((void **)&dout, sizeof(int) * (size_t)n), "alloc") &&
            qz::hip_ok(hipMemcpy(dx, x, b, hipMemcpyHostToDevice), "copy") &&
            qz::hip_ok(hipMemcpy(dy, y, b, hipMemcpyHostToDevice), "copy") &&
            qz::hip_ok(hipMemcpy(dz, z, b, hipMemcpyHostToDevice), "copy");
  if (ok)
  {
    hipLaunchKernelGGL(qz::k_box_keys, dim3((n + qz::kBlock - 1) / qz::kBlock), dim3(qz::kBlock), 0, 0, dx, dy, dz, n,
                       lv, ts[0], ts[1], ts[2], rd[0], rd[1], dout);
    ok = qz::hip_ok(hipGetLastError(), "launch") &&
         qz::hip_ok(hipMemcpy(out, dout, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost), "copy back");
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  (void)hipFree(dz);
  (void)hipFree(dout);
  return ok ? 0 : 1;
}

int qz_timing_enable(qz_ctx *ctx, int on)
{
  QZ_CHECK_CTX(ctx);
  if (!ctx->eng->join_ahead())
    return 1;
  ctx->eng->timer.collect();
  ctx->eng->timer.acc.clear();
  ctx->eng->timer.on = on == 1;
  ctx->eng->timer.host_on = on == 2;
  return 0;
}

int qz_timing_report(qz_ctx *ctx, char *names, int name_len, double *ms, int *launches, double *bytes, int max)
{
  if (!ctx || !ctx->eng)
    return 0;
  (void)hipStreamSynchronize(ctx->eng->stream);
  ctx->eng->timer.collect();
  int k = 0;
  for (auto &a : ctx->eng->timer.acc)
  {
    if (k >= max)
      break;
    snprintf(names + (size_t)k * name_len, name_len, "%s", a.name.c_str());
    ms[k] = a.ms;
    launches[k] = a.count;
    if (bytes)
      bytes[k] = a.bytes;
    k++;
  }
  return k;
}

int qz_write_snapshot(qz_ctx *ctx, const char *path, const qz_units *units, double actualtime, int with_observables)
{
  QZ_CHECK_CTX(ctx);
  if (!path || !units || units->dist == 0.0 || units->velocity == 0.0 || units->time == 0.0)
  {
    qz::set_error("qz_write_snapshot: bad arguments");
    return 1;
  }
  return ctx->eng->write_snapshot(path, *units, actualtime, with_observables != 0) ? 0 : 1;
}

int qz_write_snapshot_async(qz_ctx *ctx, const char *path, const qz_units *units, double actualtime,
                            const double *obs6)
{
  QZ_CHECK_CTX(ctx);
  if (!path || !units)
  {
    qz::set_error("qz_write_snapshot_async: null argument");
    return 1;
  }
  return ctx->eng->write_snapshot_async(path, *units, actualtime, obs6) ? 0 : 1;
}

int qz_snapshot_wait(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  return ctx->eng->snapshot_wait() ? 0 : 1;
}

// Parameters.dat in the layout of output_main_parameters.c:167-230: the 38
// "name value" lines (example_plot.py reads these), then the wall times in
// the reference's categories.  The solver lines describe this library's solve
// under the reference's names; the reference's own tree-adaptation, potential
// sub-timers and memory tables describe its host data structures and are not
// written.
int qz_write_parameters(qz_ctx *ctx, const char *path, const qz_units *units, const qz_run_info *info)
{
  QZ_CHECK_CTX(ctx);
  if (!path || !units || !info)
  {
    qz::set_error("qz_write_parameters: null argument");
    return 1;
  }
  const qz_config &c = ctx->eng->cfg;
  FILE *f = fopen(path, "w");
  if (!f)
  {
    qz::set_error("Error opening file %s", path);
    return 1;
  }
  const int pre = c.pre_sweeps, post = c.post_sweeps;
  fprintf(f, "lmin %d\n", c.lmin);
  fprintf(f, "lmax %d\n", c.lmax);
  fprintf(f, "_MAX_NUMBER_OF_ITERATIONS_IN_POISSON_EQUATION_ %d\n", c.max_iterations);
  fprintf(f, "_ERROR_THRESHOLD_IN_THE_POISSON_EQUATION_ %1.5e\n", c.tol);
  fprintf(f, "_ERROR_THRESHOLD_IN_THE_POISSON_EQUATION_2 %1.5e\n", c.tol_change);
  fprintf(f, "check_poisson_error_method %d\n", 0); // RMS residual + pointwise change (poisson_error.c:36-133)
  fprintf(f, "multigrid_cycle %d\n", 0);            // V-cycles
  fprintf(f, "solverPreS %d\n", 0);                 // Gauss-Seidel family (red-black here)
  fprintf(f, "solverfinddphic %d\n", 0);
  fprintf(f, "solverPostS %d\n", 0);
  fprintf(f, "_NiterPreS_ %d\n", pre);
  fprintf(f, "_NiterPostS_ %d\n", post);
  fprintf(f, "_Niterfinddphic_ %d\n", 2);
  fprintf(f, "_Iter_branches_solver_ %d\n", c.max_iterations);
  fprintf(f, "_w_SOR_ %1.5e\n", 1.3);      // node red half sweeps
  fprintf(f, "_w_SOR_HEAD_ %1.5e\n", 1.4); // head red half sweeps
  fprintf(f, "head_pot_method %d\n", 0);
  fprintf(f, "_CFL_ %1.5e\n", info->cfl);
  fprintf(f, "_MAX_dt_ %1.5e\n", info->max_dt);
  fprintf(f, "ref_criterion_ptcl %d\n", c.ref_criterion_ptcl);
  fprintf(f, "n_exp %d\n", c.n_exp);
  fprintf(f, "force_stencil %d\n", 0); // 3-point central differences
  fprintf(f, "potential_energy_type %d\n", info->potential_energy_type);
  fprintf(f, "GC_iter %d\n", info->gc_iter);
  fprintf(f, "Ngrid %d\n", (1 << c.lmin) + 1);
  fprintf(f, "Nparticles_initial %d\n", info->n_particles_initial > 0 ? info->n_particles_initial : c.n_particles);
  fprintf(f, "Nparticles_final %d\n", c.n_particles);
  fprintf(f, "Number_timesteps\t%d\n", info->number_timesteps);
  fprintf(f, "Number_outputs %d\n", info->number_outputs);
  fprintf(f, "fr_output %d\n", info->fr_output);
  fprintf(f, "MaxIterations %d\n", info->max_iterations);
  fprintf(f, "BoxSize %1.5e\n", info->box_size);
  fprintf(f, "vtype %d\n", 2); // double
  fprintf(f, "Maxdt %1.5e years\n", info->final_time / units->time);
  fprintf(f, "G %1.5e\n", c.G);
  fprintf(f, "boundary_type %d\n", 1); // reflexive
  fprintf(f, "GL_total_mass_initial %f\n", info->total_mass_initial);
  fprintf(f, "GL_total_mass_final %f\n", info->total_mass_final);
  static const char *names[16] = {"Global variables",    "Input",
                                  "Initialization",      "Tree Construction",
                                  "Grid Density",        "Total potential",
                                  "Grid Acceleration",   "Particle acceleration",
                                  "Time-step computing", "Particle Updating A",
                                  "Tree Adaptation",     "Reset",
                                  "Particle Updating B", "Observables",
                                  "Garbage Collector",   "Output Main Parameters"};
  double times[20] = {0.0};
  if (info->times)
    for (int i = 0; i < 20; i++)
      times[i] = info->times[i];
  double total = 0.0;
  for (int i = 0; i < 20; i++)
    total += times[i];
  const double pct = total > 0.0 ? 100.0 / total : 0.0;
  fprintf(f, "\n\nTIMES in seconds\n\n");
  fprintf(f, "TOTAL %.3e\n", total);
  for (int i = 0; i < 15; i++)
    fprintf(f, "%s = %1.2e ~ %.1f %%\n", names[i], times[i], times[i] * pct);
  fprintf(f, "Output Snapshots = %.2e ~ %.1f %%\n", times[19], times[19] * pct);
  if (fclose(f) != 0)
  {
    qz::set_error("writing %s failed", path);
    return 1;
  }
  return 0;
}

int qz_bench_sweep(qz_ctx *ctx, int reps, double *ms_per_launch, double *bytes_per_launch)
{
  QZ_CHECK_CTX(ctx);
  if (reps < 1 || !ms_per_launch || !bytes_per_launch)
  {
    qz::set_error("qz_bench_sweep: bad arguments");
    return 1;
  }
  return ctx->eng->bench_sweep(reps, ms_per_launch, bytes_per_launch) ? 0 : 1;
}

int qz_copy_probe(qz_ctx *ctx, size_t bytes, double *ms)
{
  QZ_CHECK_CTX(ctx);
  if (!ms || bytes < sizeof(double))
  {
    qz::set_error("qz_copy_probe: bad arguments");
    return 1;
  }
  return ctx->eng->copy_probe(bytes, ms) ? 0 : 1;
}

int qz_sync(qz_ctx *ctx)
{
  QZ_CHECK_CTX(ctx);
  return qz::hip_ok(hipStreamSynchronize(ctx->eng->stream), "sync") ? 0 : 1;
}

int qz_observables_noalloc(const double *x, const double *y, const double *z, const double *vx, const double *vy,
                           const double *vz, const double *m, int n, double G, int with_pe, double *out6)
{
  if (n < 0 || n > QZ_NOALLOC_MAX_PARTICLES)
  {
    qz::set_error("qz_observables_noalloc: n = %d outside [0, %d], the size of its fixed scratch array", n,
                  QZ_NOALLOC_MAX_PARTICLES);
    return 1;
  }
  if (!out6 || (n > 0 && (!vx || !vy || !vz || !m || (with_pe && (!x || !y || !z)))))
  {
    qz::set_error("qz_observables_noalloc: null array");
    return 1;
  }
  double *scr = scratch_for_current_device();
  if (!scr)
    return 1;
  double *sums = scr + 4 * (size_t)kNoAllocBlocks, *res = sums + 8;
  // Every value read below is written first in the same call, so nothing left
  // by an earlier call needs clearing; concurrent calls on one device race.
  const int nb = n > 0 ? (n + qz::kBlock - 1) / qz::kBlock : 0;
  if (nb > 0)
    hipLaunchKernelGGL(qz::k_kin_mom<double>, dim3(nb), dim3(qz::kBlock), 0, 0, vx, vy, vz, m, n, scr);
  for (int q = 0; q < 4; q++)
    hipLaunchKernelGGL(qz::k_reduce_final<0>, dim3(1), dim3(qz::kBlock), 0, 0, scr + (size_t)q * nb, nb, sums + q);
  if (with_pe)
  {
    if (nb > 0)
      hipLaunchKernelGGL(qz::k_pairwise_pe<double>, dim3(nb), dim3(qz::kBlock), 0, 0, x, y, z, m, n, scr);
    hipLaunchKernelGGL(qz::k_reduce_final<0>, dim3(1), dim3(qz::kBlock), 0, 0, scr, nb, sums + 4);
  }
  hipLaunchKernelGGL(k_observables_finish, dim3(1), dim3(64), 0, 0, sums, G, with_pe ? 1 : 0, res);
  if (!qz::hip_ok(hipGetLastError(), "qz_observables_noalloc: launch"))
    return 1;
  return qz::hip_ok(hipMemcpy(out6, res, 6 * sizeof(double), hipMemcpyDefault), "qz_observables_noalloc: result") ? 0
                                                                                                                   : 1;
}

} // extern "C"
