// Adjoint of the PM step, head grid and refinement levels (Engine<P,G> methods).
// Included by qz_core.hip.
//
// Step j of the fixed-dt KDK loop (particle_updating.c:92-147, 2677-2691) with
// h = dt / 2 and a_j = A(x_j) the force pass at x_j:
//   u = v_{j-1} + h a_{j-1};  y = x_{j-1} + dt u;  x_j = s y + c;  v_j = s u + h a_j
// (s = -1, c in {0, 2} on reflected axes).  Reverse sweep, j = N .. 1:
//   abar_j = h vbar_j + abarA;  xbar_j += J_A(x_j)^T abar_j
//   ybar = s xbar_j;  ubar = s vbar_j + dt ybar;  xbar <- ybar, vbar <- ubar, abarA <- h ubar
// and finally xbar_0 += J_A(x_0)^T abarA.  Forward states are recomputed from
// checkpoints every `every` steps.  Within a segment the sorted positions, keys,
// head cell ranges, grid accelerations and reflection flags of each step are
// kept, with refinement levels also the tree and node accelerations, so the
// reverse sweep needs no forward solve.
#include <type_traits>

#include "qz_adjoint_kernels.h"

namespace qz
{

template <typename P, typename G> void Engine<P, G>::release_adjoint()
{
  void *p1[] = {adj_xb,     adj_vb,     adj_ab,     adj_abA,    adj_psi,    adj_abP,    adj_m,      adj_m2,
                adj_xb2,    adj_vb2,    adj_tails[0], adj_tails[1],
                adj_gbar[0], adj_gbar[1], adj_gbar[2], adj_rhs,    adj_nab[0], adj_nab[1], adj_nab[2],
                adj_npb,    adj_npsi,   adj_hpb,    adj_nrhs,   adj_kind,   adj_rkey[0], adj_rkey[1],
                adj_rval[0], adj_rval[1], adj_bpre,  adj_rsrc,   adj_rtmp,   adj_psi_prev};
  for (void *q : p1)
    (void)hipFree(q);
  adj_xb = adj_vb = adj_ab = adj_abA = adj_psi = adj_psi_prev = adj_npb = adj_npsi = adj_hpb = nullptr;
  adj_hist = 0;
  adj_abP = adj_m = adj_m2 = nullptr;
  adj_xb2 = adj_vb2 = nullptr;
  adj_tails[0] = adj_tails[1] = nullptr;
  adj_gbar[0] = adj_gbar[1] = adj_gbar[2] = adj_rhs = adj_nrhs = nullptr;
  adj_nab[0] = adj_nab[1] = adj_nab[2] = nullptr;
  adj_kind = nullptr;
  adj_rkey[0] = adj_rkey[1] = nullptr;
  adj_rval[0] = adj_rval[1] = adj_bpre = nullptr;
  adj_rsrc = nullptr;
  adj_rtmp = nullptr;
  adj_ncap = adj_rcap = adj_bpre_cap = 0;
  adj_rtmp_bytes = 0;
  for (auto &c : ckpts)
    free_ckpt(c);
  ckpts.clear();
  adj_nckpt = 0;
  free_ckpt(adj_fin);
  free_recs(recs);
  free_recs(recs2);
  if (helper) // its reverse-sweep thread is always joined before a call returns
  {
    const hipStream_t st = helper->stream;
    (void)hipStreamSynchronize(st);
    delete helper;
    (void)hipStreamDestroy(st);
    helper = nullptr;
  }
}

template <typename P, typename G> void Engine<P, G>::free_recs(std::vector<StepRec> &v)
{
  for (auto &r : v)
  {
    for (int a = 0; a < 3; a++)
    {
      (void)hipFree(r.x[a]);
      (void)hipFree(r.g[a]);
    }
    (void)hipFree(r.id);
    (void)hipFree(r.key);
    (void)hipFree(r.cstart);
    (void)hipFree(r.perm);
    (void)hipFree(r.refl);
    free_tree_snap(r.tsnap);
  }
  v.clear();
}

template <typename P, typename G> bool Engine<P, G>::ensure_adjoint()
{
  if (adj_xb)
    return true;
  size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  size_t b = 0;
  adj_m = dmalloc<P>(N, b);
  adj_m2 = dmalloc<P>(N, b);
  if (cfg.lmin < cfg.lmax)
    adj_hpb = dmalloc<double>(npts, b);
  if (!adj_m || !adj_m2 || (cfg.lmin < cfg.lmax && !adj_hpb))
    return hip_ok(hipErrorOutOfMemory, "adjoint buffers");
  adj_xb2 = dmalloc<double>(3 * (size_t)N, b);
  adj_vb2 = dmalloc<double>(3 * (size_t)N, b);
  for (int a = 0; a < 2; a++)
    if (!(adj_tails[a] = dmalloc<G>(8 * (size_t)N, b)))
      return hip_ok(hipErrorOutOfMemory, "adjoint buffers");
  if (!std::is_same<P, double>::value)
    adj_abP = dmalloc<P>(3 * (size_t)N, b);
  if (!adj_xb2 || !adj_vb2 || (!std::is_same<P, double>::value && !adj_abP))
    return hip_ok(hipErrorOutOfMemory, "adjoint buffers");
  adj_xb = dmalloc<double>(3 * (size_t)N, b);
  adj_vb = dmalloc<double>(3 * (size_t)N, b);
  adj_ab = dmalloc<double>(3 * (size_t)N, b);
  adj_abA = dmalloc<double>(3 * (size_t)N, b);
  adj_psi = dmalloc<double>(npts, b);
  adj_psi_prev = dmalloc<double>(npts, b);
  for (int a = 0; a < 3; a++)
    adj_gbar[a] = dmalloc<G>(npts, b);
  adj_rhs = dmalloc<G>(npts, b);
  if (!adj_xb || !adj_vb || !adj_ab || !adj_abA || !adj_psi || !adj_psi_prev || !adj_gbar[0] ||
      !adj_gbar[1] || !adj_gbar[2] || !adj_rhs)
    return hip_ok(hipErrorOutOfMemory, "adjoint buffers");
  bytes += b;
  return true;
}

// The adjoint head solve Lap psi = phibar with zero faces: the same multigrid
// with the adjoint convergence test.  Within a reverse sweep it starts from the
// linear extrapolation of the last two adjoint solutions (from the last one
// after the first step), as the potential solve does with its history.
template <typename P, typename G> bool Engine<P, G>::adjoint_head_solve()
{
  const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  if (!adj_warm)
  {
    adj_hist = 0;
    QZ_LAUNCHB("adj_fill", 8.0 * npts, k_fill<double>, nblk(npts), adj_psi, 0.0, npts);
  }
  else if (adj_hist == 2)
    QZ_LAUNCHB("extrapolate", 32.0 * npts, k_extrapolate, nblk(npts), adj_psi, adj_psi_prev, npts);
  if (!solve_head(adj_psi, adj_rhs, true, adj_warm))
    return false;
  if (adj_sweep)
  {
    if (adj_hist == 0 &&
        !hip_ok(hipMemcpyAsync(adj_psi_prev, adj_psi, sizeof(double) * npts, hipMemcpyDeviceToDevice, stream),
                "adjoint potential history"))
      return false;
    if (adj_hist < 2)
      adj_hist++;
  }
  adj_warm = adj_sweep;
  return true;
}

// abar (3 N, slot order) as the deposit's mass type: the array itself when P
// is double, else a converted copy.
template <typename P, typename G> const P *Engine<P, G>::abar_as_mass(const double *ab)
{
  if constexpr (std::is_same<P, double>::value)
    return ab;
  else
  {
    const size_t n3 = 3 * (size_t)N;
    QZ_LAUNCHB("adj_abar_mass", (8.0 + sizeof(P)) * n3, k_to_mass_type<P>, nblk(n3), ab, adj_abP, n3);
    return adj_abP;
  }
}

// Run tails of the three components of abar (stride N, mass type) at level lv:
// tails (x), adj_tails[0] (y), adj_tails[1] (z).
template <typename P, typename G> void Engine<P, G>::scatter_runs3(const P *abm, int lv)
{
  const int l = lv - cfg.lmin;
  QZ_LAUNCHB("adj_scatter_runs", (32.0 + 3.0 * sizeof(P) + 24.0) * N, (k_deposit_runs3<P, G>), nblk(N), key_out,
             pos[0], pos[1], pos[2], abm, N, lv, 3 * (D - l), (1 << lv) - 1, tails, adj_tails[0], adj_tails[1]);
}

// cell_totals over the three tail arrays of scatter_runs3.
template <typename P, typename G> void Engine<P, G>::cell_totals3(const int *cb, const int *ce, long long ncell)
{
  if (ncell <= 0)
    return;
  QZ_LAUNCH("cell_totals", k_fill<int>, 1, nheavy, 0, (size_t)1);
  QZ_LAUNCHB("cell_totals", 8.0 * ncell, k_cell_totals3<G>, nblk((size_t)ncell), cb, ce, ncell, tails, adj_tails[0],
             adj_tails[1], heavy, nheavy);
  QZ_LAUNCH("cell_totals", k_cell_totals_heavy3<G>, 1024, cb, ce, heavy, nheavy, tails, adj_tails[0], adj_tails[1]);
}

// adj_xb += J_A(x)^T ab at the current sorted positions, both in the current
// slot order (the masses adj_m too), with the grid accelerations of A(x) in
// gax/gay/gaz and the head cell ranges current.
template <typename P, typename G> bool Engine<P, G>::vjp_at_current(const double *ab)
{
  if (has_tree())
    return vjp_tree(ab);
  size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  double H = 1.0 / n;
  // Gbar = scatter of abar (the deposit with abar as the mass, coefficient 1),
  // the three components in one pass
  scatter_runs3(abar_as_mass(ab), cfg.lmin);
  cell_totals3(cell_start, nullptr, (long long)n * n * n);
  QZ_LAUNCHB("adj_scatter_fold", 4.0 * n * n * n + 3.0 * sizeof(G) * npts, (k_adj_head_scatter3<G, G>), nblk(npts),
             cell_start, (const int *)nullptr, tails, adj_tails[0], adj_tails[1], adj_gbar[0], adj_gbar[1],
             adj_gbar[2], make_cube(n + 1));
  QZ_LAUNCHB("adj_grad_t", 4.0 * sizeof(G) * npts, k_grad_transpose<G>, nblk(npts), adj_gbar[0], adj_gbar[1],
             adj_gbar[2], adj_rhs, make_cube(n + 1), 1.0 / (2.0 * H));
  if (!adjoint_head_solve())
    return false;
  const double coeff = 4.0 * cfg.G * kPi / (H * H * H);
  QZ_LAUNCHB("adj_gather", (24.0 + sizeof(P) + 4.0 + 48.0) * N + (8.0 + 3.0 * sizeof(G)) * npts,
             (k_adjoint_gather<P, G>), nblk(N), pos[0], pos[1], pos[2], adj_m, adj_psi, gax, gay, gaz, ab,
             ab + N, ab + 2 * (size_t)N, adj_xb, adj_xb + N, adj_xb + 2 * (size_t)N, N, n, coeff);
  return true;
}

// The force pass up to the grid accelerations: tree, density, head and node
// potentials, grid accelerations (no-ops for the tree stages on the head grid).
template <typename P, typename G> bool Engine<P, G>::force_fields() { return force_overlapped(false); }

// Grow-only adjoint pools of the refinement levels (grid-pool layout) and the
// transfer records (8 per boundary point of the largest level).
template <typename P, typename G> bool Engine<P, G>::ensure_adjoint_tree()
{
  TreeDev &d = *tree->dev;
  const long long ng = std::max<long long>(d.used_grid, 1);
  if (ng > adj_ncap)
  {
    void *old[] = {adj_nab[0], adj_nab[1], adj_nab[2], adj_npb, adj_npsi, adj_nrhs, adj_kind};
    for (void *q : old)
      (void)hipFree(q);
    adj_ncap = next_cap(adj_ncap, ng);
    size_t b = 0;
    for (int a = 0; a < 3; a++)
      adj_nab[a] = dmalloc<double>(adj_ncap, b);
    adj_npb = dmalloc<double>(adj_ncap, b);
    adj_npsi = dmalloc<double>(adj_ncap, b);
    adj_nrhs = dmalloc<G>(adj_ncap, b);
    adj_kind = dmalloc<unsigned char>(adj_ncap, b);
    if (!adj_nab[0] || !adj_nab[1] || !adj_nab[2] || !adj_npb || !adj_npsi || !adj_nrhs || !adj_kind)
      return hip_ok(hipErrorOutOfMemory, "adjoint node pools");
  }
  long long nbmax = 1;
  int cmax = 1;
  for (size_t l = 1; l < d.lfirst.size(); l++)
  {
    long long nb = 0;
    for (int q = d.lfirst[l]; q < d.lfirst[l] + d.lcount[l]; q++)
      nb += d.hnodes[q].nbdry;
    nbmax = std::max(nbmax, nb);
    cmax = std::max(cmax, d.lcount[l]);
  }
  if (8 * nbmax > adj_rcap)
  {
    void *old[] = {adj_rkey[0], adj_rkey[1], adj_rval[0], adj_rval[1], adj_rsrc, adj_rtmp};
    for (void *q : old)
      (void)hipFree(q);
    adj_rcap = next_cap(adj_rcap, 8 * nbmax);
    if (adj_rcap > 0x7fffffffLL)
    {
      set_error("adjoint transfer records exceed 2^31 entries");
      return false;
    }
    adj_rtmp_bytes = qz_sort_temp_bytes((int)adj_rcap, 32);
    size_t b = 0;
    for (int k = 0; k < 2; k++)
    {
      adj_rkey[k] = dmalloc<unsigned>(adj_rcap, b);
      adj_rval[k] = dmalloc<int>(adj_rcap, b);
    }
    adj_rsrc = dmalloc<long long>(adj_rcap / 8 + 1, b);
    if (!hip_ok(hipMalloc(&adj_rtmp, std::max<size_t>(adj_rtmp_bytes, 1)), "adjoint records") || !adj_rkey[0] ||
        !adj_rkey[1] || !adj_rval[0] || !adj_rval[1] || !adj_rsrc)
      return hip_ok(hipErrorOutOfMemory, "adjoint records");
  }
  if (cmax > adj_bpre_cap)
  {
    (void)hipFree(adj_bpre);
    adj_bpre_cap = next_cap(adj_bpre_cap, cmax);
    if (!hip_ok(hipMalloc((void **)&adj_bpre, sizeof(int) * adj_bpre_cap), "adjoint records"))
      return false;
  }
  return true;
}

// adj_xb += J_A(x)^T ab with refinement levels, after a force pass at the
// current sorted positions (tree, node and head fields current); slot order.
template <typename P, typename G> bool Engine<P, G>::vjp_tree(const double *ab)
{
  if (!ensure_adjoint_tree())
    return false;
  TreeDev &d = *tree->dev;
  const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  const double H = 1.0 / n;
  const size_t ng = (size_t)d.used_grid;
  const int nl = (int)d.lfirst.size();
  const Cube cb = make_cube(n + 1);
  QZ_LAUNCH("adj_fill", k_fill<double>, nblk(ng), adj_npb, 0.0, ng);
  QZ_LAUNCH("adj_fill", k_fill<unsigned char>, nblk(ng), adj_kind, (unsigned char)0, ng);
  QZ_LAUNCH("adj_fill", k_fill<double>, nblk(npts), adj_hpb, 0.0, npts);
  for (int l = 1; l < nl; l++)
  {
    dim3 grid(blocks_per_node(d.lpmax[l]), d.lcount[l]);
    QZ_LAUNCH("adj_kinds", k_adj_kinds, grid, d.nodes, d.lfirst[l], d.sim_g, d.bdry_g, d.intr_g, adj_kind);
  }
  // Abar: scatter of abar over the unrefined cells of the head and of every level
  const P *abm = abar_as_mass(ab);
  scatter_runs3(abm, cfg.lmin);
  cell_totals3(cell_start, nullptr, (long long)n * n * n);
  QZ_LAUNCHB("adj_scatter_fold", 4.0 * n * n * n + 3.0 * sizeof(G) * npts, (k_adj_head_scatter3<G, G>), nblk(npts),
             cell_start, d.head_child, tails, adj_tails[0], adj_tails[1], adj_gbar[0], adj_gbar[1], adj_gbar[2], cb);
  for (int l = 1; l < nl; l++)
  {
    scatter_runs3(abm, cfg.lmin + l);
    cell_totals3(d.cstart + d.lc0[l], d.cend + d.lc0[l], d.lncell[l]);
    dim3 grid(blocks_per_node(d.lpmax[l]), d.lcount[l]);
    QZ_LAUNCH("adj_scatter_branch", (k_adj_branch_scatter3<G>), grid, d.nodes, d.lfirst[l], d.bdry_g, d.bdry_xyz,
              d.intr_g, d.intr_xyz, d.box, d.bstart, d.bend, tails, adj_tails[0], adj_tails[1], adj_nab[0],
              adj_nab[1], adj_nab[2]);
  }
  // levels, finest first: every child is done before its parent
  HostSpan span_lv(timer, "vjp_levels");
  std::vector<int> bpre;
  for (int l = nl - 1; l >= 1; l--)
  {
    const int first = d.lfirst[l], count = d.lcount[l];
    dim3 grid(blocks_per_node(d.lpmax[l]), count);
    QZ_LAUNCH("adj_grad_t", k_adj_node_grad_t, grid, d.nodes, first, d.bdry_g, d.bdry_xyz, d.intr_g, d.intr_xyz,
              adj_kind, adj_nab[0], adj_nab[1], adj_nab[2], adj_npb);
    QZ_LAUNCH("adj_fill", k_fill<double>, nblk(d.lng[l]), adj_npsi + d.lg0[l], 0.0, (size_t)d.lng[l]);
    QZ_LAUNCH("adj_rhs", (k_adj_node_rhs<G>), grid, d.nodes, first, d.intr_g, adj_npb, adj_nrhs);
    bool solved = false;
    if (!branch_mg_level((size_t)l, solved, adj_npsi, adj_nrhs, true))
      return false;
    if (!solved)
    {
      set_error("The precision was not reached in the adjoint solve of level %d", cfg.lmin + l);
      return false;
    }
    QZ_LAUNCH("adj_bdry", k_adj_node_bdry, grid, d.nodes, first, d.bdry_g, d.bdry_xyz, adj_kind, adj_npsi, adj_npb);
    // boundary Abar and phibar to the parents (transposed 8-point average)
    bpre.assign(count, 0);
    long long nb = 0;
    for (int q = 0; q < count; q++)
    {
      bpre[q] = (int)nb;
      nb += d.hnodes[first + q].nbdry;
    }
    if (nb == 0)
      continue;
    const int R = (int)(8 * nb);
    const bool to_head = l == 1;
    const unsigned long long top = to_head ? (unsigned long long)npts : (unsigned long long)ng;
    int bits = 1;
    while (bits < 32 && (1ull << bits) < top)
      bits++;
    if (!hip_ok(hipMemcpyAsync(adj_bpre, bpre.data(), sizeof(int) * count, hipMemcpyHostToDevice, stream),
                "adjoint records"))
      return false;
    QZ_LAUNCH("adj_xfer", k_adj_xfer_records, grid, d.nodes, first, adj_bpre, d.bdry_g, d.bdry_xyz, n, adj_rkey[0],
              adj_rval[0], adj_rsrc);
    timer.begin(stream);
    if (!hip_ok(qz_sort_pairs(adj_rtmp, adj_rtmp_bytes, adj_rkey[0], adj_rkey[1], adj_rval[0], adj_rval[1], R, bits,
                              stream),
                "adjoint records"))
      return false;
    timer.end(stream, "adj_xfer_sort");
    if (to_head)
      QZ_LAUNCH("adj_xfer", (k_adj_xfer_apply<G>), nblk(R), adj_rkey[1], adj_rval[1], R, adj_rsrc, adj_nab[0],
                adj_nab[1], adj_nab[2], adj_npb, adj_gbar[0], adj_gbar[1], adj_gbar[2], adj_hpb);
    else
      QZ_LAUNCH("adj_xfer", (k_adj_xfer_apply<double>), nblk(R), adj_rkey[1], adj_rval[1], R, adj_rsrc, adj_nab[0],
                adj_nab[1], adj_nab[2], adj_npb, adj_nab[0], adj_nab[1], adj_nab[2], adj_npb);
  }
  // head: transposed central difference, the level-1 phibar, the adjoint solve
  QZ_LAUNCHB("adj_grad_t", 4.0 * sizeof(G) * npts, k_grad_transpose<G>, nblk(npts), adj_gbar[0], adj_gbar[1],
             adj_gbar[2], adj_rhs, cb, 1.0 / (2.0 * H));
  QZ_LAUNCH("adj_head_add", k_adj_head_add<G>, nblk(npts), adj_rhs, adj_hpb, cb);
  if (!adjoint_head_solve())
    return false;
  QZ_LAUNCH("adj_gather_amr", (k_adjoint_gather_amr<P, G>), nblk(N), pos[0], pos[1], pos[2], adj_m, adj_psi,
            gax, gay, gaz, d.head_child, d.nodes, d.box, adj_npsi, (const G *)d.ax, (const G *)d.ay, (const G *)d.az,
            ab, ab + N, ab + 2 * (size_t)N, adj_xb, adj_xb + N, adj_xb + 2 * (size_t)N, N, cfg.lmin,
            4.0 * cfg.G * kPi);
  return hip_ok(hipGetLastError(), "adjoint");
}

// Three cotangent components (stride N) between ID order and the current slot
// order of the particles: out[s] = in[id[s]] (to_slots) or out[id[s]] = in[s].
template <typename P, typename G> bool Engine<P, G>::slot_cotangents(const double *in, double *out, bool to_slots)
{
  for (int a = 0; a < 3; a++)
  {
    const size_t o = (size_t)a * N;
    if (to_slots)
      QZ_LAUNCHB("adj_order", 20.0 * N, k_id_to_slot<double>, nblk(N), in + o, id_a, out + o, N);
    else
      QZ_LAUNCHB("adj_order", 20.0 * N, k_slot_to_id<double>, nblk(N), in + o, id_a, out + o, N);
  }
  return hip_ok(hipGetLastError(), "adjoint cotangent order");
}

template <typename P, typename G> bool Engine<P, G>::force_vjp(const double *abar, double *xbar)
{
  if (!ensure_adjoint())
    return false;
  // force pass at the current positions (tree, grid accelerations, cell ranges);
  // it leaves the particles sorted, the masses with them
  if (!force_fields())
    return false;
  size_t n3 = 3 * (size_t)N;
  if (!hip_ok(hipMemcpyAsync(adj_m, pa[9], sizeof(P) * N, hipMemcpyDeviceToDevice, stream), "adjoint masses") ||
      !hip_ok(hipMemcpy(adj_abA, abar, sizeof(double) * n3, hipMemcpyDefault), "upload abar") ||
      !slot_cotangents(adj_abA, adj_ab, true))
    return false;
  QZ_LAUNCH("adj_fill", k_fill<double>, nblk(n3), adj_xb, 0.0, n3);
  adj_sweep = adj_warm = false;
  if (!vjp_at_current(adj_ab) || !slot_cotangents(adj_xb, adj_abA, false))
    return false;
  return hip_ok(hipMemcpy(xbar, adj_abA, sizeof(double) * n3, hipMemcpyDefault), "download xbar");
}

// A checkpoint: the particle state in its sorted order plus the head potential
// and the potential history the next head solve starts from, so that a
// recomputation from it repeats the forward steps exactly.
// Copy the state into checkpoint c, allocating its buffers on first use (a
// checkpoint is reused by later runs of the same context).
template <typename P, typename G> bool Engine<P, G>::save_ckpt(Ckpt &c)
{
  const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  size_t b = 0;
  auto get = [&](auto *&buf, size_t count) {
    using T = std::remove_reference_t<decltype(*buf)>;
    if (!buf)
      buf = dmalloc<T>(count, b);
    return buf != nullptr;
  };
  bool ok = true;
  for (int a = 0; a < 3; a++)
    ok = ok && get(c.x[a], N) &&
         hip_ok(hipMemcpyAsync(c.x[a], pos[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream), "checkpoint");
  for (int a = 0; a < 7; a++)
    ok = ok && get(c.a[a], N) &&
         hip_ok(hipMemcpyAsync(c.a[a], pa[3 + a], sizeof(P) * N, hipMemcpyDeviceToDevice, stream), "checkpoint");
  ok = ok && get(c.id, N) &&
       hip_ok(hipMemcpyAsync(c.id, id_a, sizeof(int) * N, hipMemcpyDeviceToDevice, stream), "checkpoint");
  ok = ok && get(c.phi, npts) &&
       hip_ok(hipMemcpyAsync(c.phi, phi, sizeof(double) * npts, hipMemcpyDeviceToDevice, stream), "checkpoint");
  c.prev = phi_prev != nullptr;
  if (c.prev)
    ok = ok && get(c.phi_prev, npts) &&
         hip_ok(hipMemcpyAsync(c.phi_prev, phi_prev, sizeof(double) * npts, hipMemcpyDeviceToDevice, stream),
                "checkpoint");
  c.hist = phi_hist;
  bytes += b;
  return ok || hip_ok(hipErrorOutOfMemory, "checkpoint");
}

template <typename P, typename G> void Engine<P, G>::restore_ckpt(const Ckpt &c)
{
  const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  timer.begin(stream);
  for (int a = 0; a < 3; a++)
    (void)hipMemcpyAsync(pos[a], c.x[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream);
  for (int a = 0; a < 7; a++)
    (void)hipMemcpyAsync(pa[3 + a], c.a[a], sizeof(P) * N, hipMemcpyDeviceToDevice, stream);
  (void)hipMemcpyAsync(id_a, c.id, sizeof(int) * N, hipMemcpyDeviceToDevice, stream);
  (void)hipMemcpyAsync(phi, c.phi, sizeof(double) * npts, hipMemcpyDeviceToDevice, stream);
  if (c.prev && phi_prev)
    (void)hipMemcpyAsync(phi_prev, c.phi_prev, sizeof(double) * npts, hipMemcpyDeviceToDevice, stream);
  timer.end(stream, "ckpt_restore");
  phi_hist = c.prev && phi_prev ? c.hist : 0;
  keys_valid = false;
  keys_fresh = false;
  acc_live = true;
  particles_changed();
}

template <typename P, typename G> void Engine<P, G>::free_ckpt(Ckpt &c)
{
  for (int a = 0; a < 3; a++)
    (void)hipFree(c.x[a]);
  for (int a = 0; a < 7; a++)
    (void)hipFree(c.a[a]);
  (void)hipFree(c.id);
  (void)hipFree(c.phi);
  (void)hipFree(c.phi_prev);
  c = Ckpt{};
}

// Per-step records for one segment: positions, IDs and keys in sorted order,
// head cell ranges, head grid accelerations and the reflected axes of the step;
// with refinement levels also the tree and node fields (TreeSnap), so the
// reverse sweep needs no forward solve.
template <typename P, typename G> bool Engine<P, G>::ensure_recs(std::vector<StepRec> &recs)
{
  const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1, ncell = (size_t)n * n * n;
  while ((int)recs.size() < adj_every + 1)
  {
    StepRec r{};
    size_t bb = 0;
    for (int a = 0; a < 3; a++)
    {
      r.x[a] = dmalloc<double>(N, bb);
      r.g[a] = dmalloc<G>(npts, bb);
    }
    r.id = dmalloc<int>(N, bb);
    r.key = dmalloc<unsigned long long>(N, bb);
    r.cstart = dmalloc<int>(ncell + 1, bb);
    r.perm = dmalloc<int>(N, bb);
    r.refl = dmalloc<unsigned char>(N, bb);
    const bool ok = r.x[0] && r.x[1] && r.x[2] && r.id && r.refl && r.g[0] && r.g[1] && r.g[2] && r.key &&
                    r.cstart && r.perm;
    recs.push_back(r);
    if (!ok)
      return hip_ok(hipErrorOutOfMemory, "adjoint step records");
  }
  return true;
}

// Record the sorted state after a force pass.
template <typename P, typename G> bool Engine<P, G>::save_rec(StepRec &r, bool handover)
{
  const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1, ncell = (size_t)n * n * n;
  // positions and IDs carry on into the next step (copied).  With handover
  // (the reverse sweep's recomputation, where the next force pass rebuilds
  // keys, cell ranges and grid accelerations before anything reads them) the
  // record takes those buffers and gives the engine its own free ones;
  // otherwise (the forward run, whose state must stay whole) they are copied.
  timer.begin(stream);
  for (int a = 0; a < 3; a++)
    (void)hipMemcpyAsync(r.x[a], pos[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream);
  (void)hipMemcpyAsync(r.id, id_a, sizeof(int) * N, hipMemcpyDeviceToDevice, stream);
  G **gs[3] = {&gax, &gay, &gaz};
  if (handover)
  {
    std::swap(r.key, key_out);
    std::swap(r.cstart, cell_start);
    std::swap(r.perm, slot_out); // the sort's output, rewritten by the next sort
    for (int a = 0; a < 3; a++)
      std::swap(r.g[a], *gs[a]);
  }
  else
  {
    (void)hipMemcpyAsync(r.key, key_out, sizeof(unsigned long long) * N, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(r.cstart, cell_start, sizeof(int) * (ncell + 1), hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(r.perm, slot_out, sizeof(int) * N, hipMemcpyDeviceToDevice, stream);
    for (int a = 0; a < 3; a++)
      (void)hipMemcpyAsync(r.g[a], *gs[a], sizeof(G) * npts, hipMemcpyDeviceToDevice, stream);
  }
  timer.end(stream, "rec_save");
  if (has_tree())
    return snap_tree(r.tsnap);
  return true;
}

// Make a recorded state current for its VJP.
template <typename P, typename G> bool Engine<P, G>::load_rec(StepRec &r)
{
  // the reverse step reads a record once: the engine takes its buffers by
  // exchange (the record is refilled before it is read again)
  for (int a = 0; a < 3; a++)
    std::swap(pos[a], r.x[a]);
  std::swap(id_a, r.id);
  particles_changed();
  keys_fresh = false;
  std::swap(key_out, r.key);
  std::swap(cell_start, r.cstart);
  std::swap(gax, r.g[0]);
  std::swap(gay, r.g[1]);
  std::swap(gaz, r.g[2]);
  if (cfg.lmin < cfg.lmax)
  {
    if (!r.tsnap)
    {
      set_error("adjoint: a step record holds no tree");
      return false;
    }
    if (!restore_tree(r.tsnap))
      return false;
  }
  keys_valid = true;
  return true;
}

// Checkpointed forward: from the current state (a force pass done), run nsteps
// fixed-dt steps, keeping the full particle state every `every` steps.
template <typename P, typename G> bool Engine<P, G>::adjoint_forward(int nsteps, double dt, int every)
{
  if (!ensure_adjoint())
    return false;
  if (nsteps <= 0 || every <= 0)
  {
    set_error("adjoint_forward: nsteps and every must be positive");
    return false;
  }
  adj_nckpt = 0; // the checkpoint buffers of earlier runs are reused
  adj_nsteps = nsteps;
  adj_every = every;
  adj_dt = dt;
  adj_last_recorded = false;
  if (!ensure_recs(recs))
    return false;
  // the last segment [last, nsteps] is recorded on the way (the reverse sweep
  // starts there, so it needs no recomputation)
  const int last = ((nsteps - 1) / every) * every;
  for (int s = 0; s <= nsteps; s++)
  {
    if (s % every == 0 && s < nsteps)
    {
      if (adj_nckpt == (int)ckpts.size())
        ckpts.push_back(Ckpt{});
      if (!save_ckpt(ckpts[adj_nckpt++]))
        return false;
    }
    if (s == 0 && last == 0 && !save_rec(recs[0])) // the force pass of state 0 is current
      return false;
    if (s == nsteps)
      break;
    const bool rec = s + 1 > last;
    refl_rec = rec ? recs[s + 1 - last].refl : nullptr;
    const bool ok = update_A(dt);
    refl_rec = nullptr;
    if (!ok || !force_fields() || !particle_acceleration() || !update_B(dt))
      return false;
    if (rec && !save_rec(recs[s + 1 - last]))
      return false;
  }
  adj_last_recorded = true;
  return hip_ok(hipStreamSynchronize(stream), "adjoint_forward");
}

// Recompute segment [c0, c1] from its checkpoint into the records rs: rs[0] =
// state c0 (for the first segment only, whose a_0 term closes the sweep),
// rs[k] = state c0 + k.  The checkpoint carries a_c0, so no force pass is
// needed to step from it.  Ends with the stream drained.
template <typename P, typename G>
bool Engine<P, G>::recompute_segment(const Ckpt &ck, std::vector<StepRec> &rs, int c0, int c1, double dt)
{
  restore_ckpt(ck);
  if (c0 == 0)
  {
    // the record of state 0 holds its fields (and tree); the force pass clears
    // the accelerations, which the checkpoint then restores exactly
    if (!force_fields())
      return false;
    restore_ckpt(ck);
    if (!save_rec(rs[0], true))
      return false;
  }
  for (int j = c0 + 1; j <= c1; j++)
  {
    refl_rec = rs[j - c0].refl;
    bool ok = update_A(dt);
    refl_rec = nullptr;
    if (!ok || !force_fields() || !particle_acceleration() || !update_B(dt) || !save_rec(rs[j - c0], true))
      return false;
  }
  return hip_ok(hipStreamSynchronize(stream), "recompute");
}

// The helper engine of the reverse sweep: the same configuration, its own
// buffers and a non-blocking stream, the same total mass and head faces, and a
// potential history buffer so that a checkpoint restores the same starting
// guess.  It launches eagerly (no graph capture while this engine captures).
template <typename P, typename G> bool Engine<P, G>::make_helper()
{
  if (helper)
    return true;
  auto *hp = new Engine<P, G>();
  hp->cfg = cfg;
  const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  if (!hip_ok(hipStreamCreateWithFlags(&hp->stream, hipStreamNonBlocking), "helper stream"))
  {
    delete hp;
    return false;
  }
  if (!hp->init() || !(hp->phi_prev = dmalloc<double>(npts, hp->bytes)))
  {
    const hipStream_t st = hp->stream;
    delete hp;
    (void)hipStreamDestroy(st);
    set_error("adjoint_backward: the helper engine could not be created (device memory)");
    return false;
  }
  hp->head_graphs = hp->batch_graphs = false;
  hp->total_mass = total_mass;
  hp->face_gm = face_gm;
  hp->host_mass = host_mass;
  helper = hp;
  return true;
}

template <typename P, typename G> bool Engine<P, G>::adjoint_backward(const double *xbN, const double *vbN,
                                                                    double *xb0, double *vb0)
{
  if (adj_nckpt == 0)
  {
    set_error("adjoint_backward: run qz_adjoint_forward first");
    return false;
  }
  const size_t n3 = 3 * (size_t)N;
  const double dt = adj_dt, h = 0.5 * dt;
  // keep the final state to restore it afterwards (the flags of the sweep are
  // cleared on every return)
  Ckpt &fin = adj_fin;
  struct FinGuard
  {
    Engine *e;
    ~FinGuard() { e->adj_sweep = e->adj_warm = e->adj_last_recorded = false; }
  } fin_guard{this};
  if (!save_ckpt(fin))
    return false;
  adj_sweep = true;
  adj_warm = false;
  // with more than one segment, the recomputation of the segment before the one
  // being reversed runs on the helper engine (its own buffers and stream) on a
  // host thread: A holds the records being reversed, B the ones being filled
  const bool overlap = adj_overlap && adj_nckpt > 1;
  if (!ensure_recs(recs) || (overlap && (!ensure_recs(recs2) || !make_helper())))
    return false;
  std::vector<StepRec> *A = &recs, *B = &recs2;
  // while the helper works, neither engine captures graphs (no capture can then
  // see the other thread's allocations or synchronisations); launched one by
  // one the solves take the same time
  struct LaunchMode
  {
    Engine *e;
    bool hg, bg;
    ~LaunchMode()
    {
      e->head_graphs = hg;
      e->batch_graphs = bg;
    }
  } launch_mode{this, head_graphs, batch_graphs};
  if (overlap)
    head_graphs = batch_graphs = false;
  int dev = 0;
  if (overlap && !hip_ok(hipGetDevice(&dev), "adjoint device"))
    return false;
  struct Ahead
  {
    std::thread th;
    bool ok = true;
    std::string err;
    ~Ahead()
    {
      if (th.joinable())
        th.join();
    }
    bool join()
    {
      if (th.joinable())
        th.join();
      if (!ok)
        set_error("%s", err.c_str());
      return ok;
    }
  } ahead;
  // the cotangents may be host or device memory (the copies infer the
  // direction); they arrive in ID order and are kept in the slot order of the
  // state they belong to, which is the particles' current order at the final
  // state (the masses likewise)
  if (!hip_ok(hipMemcpy(adj_ab, xbN, sizeof(double) * n3, hipMemcpyDefault), "upload xbar") ||
      !hip_ok(hipMemcpy(adj_abA, vbN, sizeof(double) * n3, hipMemcpyDefault), "upload vbar") ||
      !slot_cotangents(adj_ab, adj_xb, true) || !slot_cotangents(adj_abA, adj_vb, true) ||
      !hip_ok(hipMemcpyAsync(adj_m, pa[9], sizeof(P) * N, hipMemcpyDeviceToDevice, stream), "adjoint masses"))
    return false;
  QZ_LAUNCH("adj_fill", k_fill<double>, nblk(n3), adj_abA, 0.0, n3);
  bool have = adj_last_recorded; // A holds the records of segment adj_nckpt - 1
  for (int sc = adj_nckpt - 1; sc >= 0; sc--)
  {
    const int c0 = sc * adj_every;
    const int c1 = std::min(c0 + adj_every, adj_nsteps);
    if (!have)
    {
      HostSpan span_rc(timer, "adj_recompute");
      if (!recompute_segment(ckpts[sc], *A, c0, c1, dt))
        return false;
    }
    if (overlap && sc > 0)
    {
      // B is free: the reverse sweep that read it ended with a stream drain
      const int p0 = (sc - 1) * adj_every;
      Engine *hp = helper;
      std::vector<StepRec> *rb = B;
      const Ckpt *ck = &ckpts[sc - 1];
      ahead.ok = true;
      ahead.th = std::thread([hp, rb, ck, p0, dt, dev, this, &ahead]() {
        ahead.ok = hip_ok(hipSetDevice(dev), "adjoint device") &&
                   hp->recompute_segment(*ck, *rb, p0, std::min(p0 + adj_every, adj_nsteps), dt);
        if (!ahead.ok)
          ahead.err = g_last_error;
      });
    }
    // reverse sweep over the segment
    for (int j = c1; j > c0; j--)
    {
      StepRec &r = (*A)[j - c0];
      if (!load_rec(r))
        return false;
      QZ_LAUNCH("adj_abar", k_adj_abar, nblk(n3), adj_vb, adj_abA, adj_ab, (int)n3, h);
      HostSpan span_vjp(timer, "adj_vjp");
      if (!vjp_at_current(adj_ab))
        return false;
      // to the slot order before this step's sort, where its drift ran
      QZ_LAUNCHB("adj_unsort", (2.0 * (48.0 + sizeof(P)) + 4.0) * N, k_adj_unsort<P>, nblk(N), r.perm, adj_xb,
                 adj_vb, adj_m, adj_xb2, adj_vb2, adj_m2, N);
      std::swap(adj_xb, adj_xb2);
      std::swap(adj_vb, adj_vb2);
      std::swap(adj_m, adj_m2);
      QZ_LAUNCH("adj_kick_drift", k_adj_kick_drift, nblk(N), adj_xb, adj_vb, adj_abA, r.refl, N, dt);
    }
    have = false;
    if (overlap && sc > 0)
    {
      HostSpan span_w(timer, "adj_recompute_wait");
      // A's buffers are free once the sweep has drained; B is complete once the
      // helper's thread (which drains the helper's stream) has ended
      if (!hip_ok(hipStreamSynchronize(stream), "reverse sweep") || !ahead.join())
        return false;
      std::swap(A, B);
      have = true;
    }
  }
  // a_0 = A(x_0) fed the first half kick; then back to ID order
  if (!load_rec((*A)[0]) || !vjp_at_current(adj_abA) || !slot_cotangents(adj_xb, adj_ab, false) ||
      !slot_cotangents(adj_vb, adj_abA, false))
    return false;
  if (!hip_ok(hipMemcpy(xb0, adj_ab, sizeof(double) * n3, hipMemcpyDefault), "download xbar") ||
      !hip_ok(hipMemcpy(vb0, adj_abA, sizeof(double) * n3, hipMemcpyDefault), "download vbar"))
    return false;
  // back to the final forward state: its fields (and tree) are rebuilt, then the
  // state is restored once more for the exact accelerations, potential and
  // potential history (the state is in key order, so the sort of the force
  // pass leaves the order unchanged)
  HostSpan span_fin(timer, "adj_final_force");
  restore_ckpt(fin);
  if (!force_fields())
    return false;
  restore_ckpt(fin);
  return hip_ok(hipStreamSynchronize(stream), "adjoint_backward");
}

} // namespace qz
