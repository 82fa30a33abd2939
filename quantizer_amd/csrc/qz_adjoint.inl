// Adjoint of the head-grid PM step (Engine<P,G> methods).  Included by qz_core.hip.
//
// Step j of the fixed-dt KDK loop (particle_updating.c:92-147, 2677-2691) with
// h = dt / 2 and a_j = A(x_j) the force pass at x_j:
//   u = v_{j-1} + h a_{j-1};  y = x_{j-1} + dt u;  x_j = s y + c;  v_j = s u + h a_j
// (s = -1, c in {0, 2} on reflected axes).  Reverse sweep, j = N .. 1:
//   abar_j = h vbar_j + abarA;  xbar_j += J_A(x_j)^T abar_j
//   ybar = s xbar_j;  ubar = s vbar_j + dt ybar;  xbar <- ybar, vbar <- ubar, abarA <- h ubar
// and finally xbar_0 += J_A(x_0)^T abarA.  Forward states are recomputed from
// checkpoints every `every` steps; within a segment the sorted positions, keys,
// head cell ranges, grid accelerations and reflection flags of each step are
// kept so the reverse sweep needs no forward Poisson solve.
#include "qz_adjoint_kernels.h"

namespace qz
{

template <typename P, typename G> void Engine<P, G>::release_adjoint()
{
  void *p1[] = {adj_xb, adj_vb, adj_ab, adj_abA, adj_psi, adj_slot, adj_gbar[0], adj_gbar[1], adj_gbar[2], adj_rhs};
  for (void *q : p1)
    (void)hipFree(q);
  adj_xb = adj_vb = adj_ab = adj_abA = adj_psi = nullptr;
  adj_slot = nullptr;
  adj_gbar[0] = adj_gbar[1] = adj_gbar[2] = adj_rhs = nullptr;
  for (auto &c : ckpts)
  {
    for (int a = 0; a < 3; a++)
      (void)hipFree(c.x[a]);
    for (int a = 0; a < 7; a++)
      (void)hipFree(c.a[a]);
    (void)hipFree(c.id);
  }
  ckpts.clear();
  for (auto &r : recs)
  {
    for (int a = 0; a < 3; a++)
    {
      (void)hipFree(r.x[a]);
      (void)hipFree(r.g[a]);
    }
    (void)hipFree(r.id);
    (void)hipFree(r.key);
    (void)hipFree(r.cstart);
    (void)hipFree(r.refl);
  }
  recs.clear();
}

template <typename P, typename G> bool Engine<P, G>::ensure_adjoint()
{
  if (adj_xb)
    return true;
  if (cfg.lmin != cfg.lmax)
  {
    set_error("the adjoint covers the head-grid step (lmin == lmax)");
    return false;
  }
  size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  size_t b = 0;
  adj_xb = dmalloc<double>(3 * (size_t)N, b);
  adj_vb = dmalloc<double>(3 * (size_t)N, b);
  adj_ab = dmalloc<double>(3 * (size_t)N, b);
  adj_abA = dmalloc<double>(3 * (size_t)N, b);
  adj_psi = dmalloc<double>(npts, b);
  adj_slot = dmalloc<P>(N, b);
  for (int a = 0; a < 3; a++)
    adj_gbar[a] = dmalloc<G>(npts, b);
  adj_rhs = dmalloc<G>(npts, b);
  if (!adj_xb || !adj_vb || !adj_ab || !adj_abA || !adj_psi || !adj_slot || !adj_gbar[0] || !adj_gbar[1] ||
      !adj_gbar[2] || !adj_rhs)
    return hip_ok(hipErrorOutOfMemory, "adjoint buffers");
  bytes += b;
  return true;
}

// adj_xb (ID order) += J_A(x)^T ab at the current sorted positions, with the
// grid accelerations of A(x) in gax/gay/gaz and the head cell ranges current.
template <typename P, typename G> bool Engine<P, G>::vjp_at_current(const double *ab)
{
  size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1;
  double H = 1.0 / n;
  // Gbar = scatter of abar (the deposit with abar as the mass, coefficient 1)
  for (int a = 0; a < 3; a++)
  {
    QZ_LAUNCHB("adj_slot", 12.0 * N, k_by_id_to_slot<P>, nblk(N), ab + (size_t)a * N, id_a, adj_slot, N);
    QZ_LAUNCHB("adj_scatter_runs", (24.0 + sizeof(P) + 8.0) * N, (k_deposit_runs<P, G>), nblk(N), key_out, pos[0],
               pos[1], pos[2], adj_slot, N, cfg.lmin, 3 * D, n - 1, tails);
    QZ_LAUNCHB("adj_scatter_fold", 4.0 * n * n * n + (double)sizeof(G) * npts, (k_deposit_head_fold<G, G>),
               nblk(npts), cell_start, tails, adj_gbar[a], make_cube(n + 1), 1.0);
  }
  QZ_LAUNCHB("adj_grad_t", 4.0 * sizeof(G) * npts, k_grad_transpose<G>, nblk(npts), adj_gbar[0], adj_gbar[1],
             adj_gbar[2], adj_rhs, make_cube(n + 1), 1.0 / (2.0 * H));
  // Lap psi = phibar with zero faces: the same multigrid, adjoint convergence test
  QZ_LAUNCHB("adj_fill", 8.0 * npts, k_fill<double>, nblk(npts), adj_psi, 0.0, npts);
  if (!solve_head(adj_psi, adj_rhs, true))
    return false;
  const double coeff = 4.0 * cfg.G * kPi / (H * H * H);
  QZ_LAUNCHB("adj_gather", (24.0 + sizeof(P) + 4.0 + 48.0) * N + (8.0 + 3.0 * sizeof(G)) * npts,
             (k_adjoint_gather<P, G>), nblk(N), pos[0], pos[1], pos[2], pa[9], id_a, adj_psi, gax, gay, gaz, ab,
             ab + N, ab + 2 * (size_t)N, adj_xb, adj_xb + N, adj_xb + 2 * (size_t)N, N, n, coeff);
  return true;
}

template <typename P, typename G> bool Engine<P, G>::force_vjp(const double *abar, double *xbar)
{
  if (!ensure_adjoint())
    return false;
  // force pass at the current positions (grid accelerations and cell ranges)
  if (!reset() || !grid_density() || !potential_head() || !grid_acceleration())
    return false;
  size_t n3 = 3 * (size_t)N;
  if (!hip_ok(hipMemcpy(adj_ab, abar, sizeof(double) * n3, hipMemcpyHostToDevice), "upload abar"))
    return false;
  QZ_LAUNCH("adj_fill", k_fill<double>, nblk(n3), adj_xb, 0.0, n3);
  if (!vjp_at_current(adj_ab))
    return false;
  return hip_ok(hipMemcpy(xbar, adj_xb, sizeof(double) * n3, hipMemcpyDeviceToHost), "download xbar");
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
  for (auto &c : ckpts)
  {
    for (int a = 0; a < 3; a++)
      (void)hipFree(c.x[a]);
    for (int a = 0; a < 7; a++)
      (void)hipFree(c.a[a]);
    (void)hipFree(c.id);
  }
  ckpts.clear();
  adj_nsteps = nsteps;
  adj_every = every;
  adj_dt = dt;
  for (int s = 0; s <= nsteps; s++)
  {
    if (s % every == 0 && s < nsteps)
    {
      Ckpt c{};
      size_t b = 0;
      for (int a = 0; a < 3; a++)
      {
        c.x[a] = dmalloc<double>(N, b);
        if (!c.x[a] || !hip_ok(hipMemcpyAsync(c.x[a], pos[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream),
                               "checkpoint"))
          return false;
      }
      for (int a = 0; a < 7; a++)
      {
        c.a[a] = dmalloc<P>(N, b);
        if (!c.a[a] || !hip_ok(hipMemcpyAsync(c.a[a], pa[3 + a], sizeof(P) * N, hipMemcpyDeviceToDevice, stream),
                               "checkpoint"))
          return false;
      }
      c.id = dmalloc<int>(N, b);
      if (!c.id || !hip_ok(hipMemcpyAsync(c.id, id_a, sizeof(int) * N, hipMemcpyDeviceToDevice, stream), "checkpoint"))
        return false;
      ckpts.push_back(c);
    }
    if (s == nsteps)
      break;
    if (!update_A(dt) || !reset() || !grid_density() || !potential_head() || !grid_acceleration() ||
        !particle_acceleration() || !update_B(dt))
      return false;
  }
  return hip_ok(hipStreamSynchronize(stream), "adjoint_forward");
}

template <typename P, typename G> bool Engine<P, G>::adjoint_backward(const double *xbN, const double *vbN,
                                                                    double *xb0, double *vb0)
{
  if (ckpts.empty())
  {
    set_error("adjoint_backward: run qz_adjoint_forward first");
    return false;
  }
  const size_t n1 = (size_t)n + 1, npts = n1 * n1 * n1, n3 = 3 * (size_t)N;
  const size_t ncell = (size_t)n * n * n;
  const double dt = adj_dt, h = 0.5 * dt;
  // keep the final state to restore it afterwards
  Ckpt fin{};
  size_t b = 0;
  for (int a = 0; a < 3; a++)
  {
    fin.x[a] = dmalloc<double>(N, b);
    (void)hipMemcpyAsync(fin.x[a], pos[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream);
  }
  for (int a = 0; a < 7; a++)
  {
    fin.a[a] = dmalloc<P>(N, b);
    (void)hipMemcpyAsync(fin.a[a], pa[3 + a], sizeof(P) * N, hipMemcpyDeviceToDevice, stream);
  }
  fin.id = dmalloc<int>(N, b);
  (void)hipMemcpyAsync(fin.id, id_a, sizeof(int) * N, hipMemcpyDeviceToDevice, stream);
  // per-step records for one segment
  if ((int)recs.size() < adj_every + 1)
  {
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
      r.refl = dmalloc<unsigned char>(N, bb);
      if (!r.x[0] || !r.x[1] || !r.x[2] || !r.g[0] || !r.g[1] || !r.g[2] || !r.id || !r.key || !r.cstart || !r.refl)
        return hip_ok(hipErrorOutOfMemory, "adjoint step records");
      recs.push_back(r);
    }
  }
  auto restore_ckpt = [&](const Ckpt &c) {
    for (int a = 0; a < 3; a++)
      (void)hipMemcpyAsync(pos[a], c.x[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream);
    for (int a = 0; a < 7; a++)
      (void)hipMemcpyAsync(pa[3 + a], c.a[a], sizeof(P) * N, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(id_a, c.id, sizeof(int) * N, hipMemcpyDeviceToDevice, stream);
    keys_valid = false;
  };
  // record the sorted state after a force pass (positions, ids, keys, ranges, grid a)
  auto save_rec = [&](StepRec &r) {
    for (int a = 0; a < 3; a++)
      (void)hipMemcpyAsync(r.x[a], pos[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(r.id, id_a, sizeof(int) * N, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(r.key, key_out, sizeof(unsigned long long) * N, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(r.cstart, cell_start, sizeof(int) * (ncell + 1), hipMemcpyDeviceToDevice, stream);
    G *src[3] = {gax, gay, gaz};
    for (int a = 0; a < 3; a++)
      (void)hipMemcpyAsync(r.g[a], src[a], sizeof(G) * npts, hipMemcpyDeviceToDevice, stream);
  };
  auto load_rec = [&](const StepRec &r) {
    for (int a = 0; a < 3; a++)
      (void)hipMemcpyAsync(pos[a], r.x[a], sizeof(double) * N, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(id_a, r.id, sizeof(int) * N, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(key_out, r.key, sizeof(unsigned long long) * N, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(cell_start, r.cstart, sizeof(int) * (ncell + 1), hipMemcpyDeviceToDevice, stream);
    G *dst[3] = {gax, gay, gaz};
    for (int a = 0; a < 3; a++)
      (void)hipMemcpyAsync(dst[a], r.g[a], sizeof(G) * npts, hipMemcpyDeviceToDevice, stream);
    keys_valid = true;
  };
  if (!hip_ok(hipMemcpy(adj_xb, xbN, sizeof(double) * n3, hipMemcpyHostToDevice), "upload xbar") ||
      !hip_ok(hipMemcpy(adj_vb, vbN, sizeof(double) * n3, hipMemcpyHostToDevice), "upload vbar"))
    return false;
  QZ_LAUNCH("adj_fill", k_fill<double>, nblk(n3), adj_abA, 0.0, n3);
  for (int sc = (int)ckpts.size() - 1; sc >= 0; sc--)
  {
    int c0 = sc * adj_every;
    int c1 = std::min(c0 + adj_every, adj_nsteps);
    // recompute the segment: rec[0] = state c0, rec[k] = state c0 + k
    restore_ckpt(ckpts[sc]);
    if (!reset() || !grid_density() || !potential_head() || !grid_acceleration())
      return false;
    save_rec(recs[0]);
    for (int j = c0 + 1; j <= c1; j++)
    {
      refl_rec = recs[j - c0].refl;
      bool ok = update_A(dt);
      refl_rec = nullptr;
      if (!ok || !reset() || !grid_density() || !potential_head() || !grid_acceleration() ||
          !particle_acceleration() || !update_B(dt))
        return false;
      save_rec(recs[j - c0]);
    }
    // reverse sweep over the segment
    for (int j = c1; j > c0; j--)
    {
      load_rec(recs[j - c0]);
      QZ_LAUNCH("adj_abar", k_adj_abar, nblk(n3), adj_vb, adj_abA, adj_ab, (int)n3, h);
      if (!vjp_at_current(adj_ab))
        return false;
      QZ_LAUNCH("adj_kick_drift", k_adj_kick_drift, nblk(N), adj_xb, adj_vb, adj_abA, recs[j - c0].refl, N, dt);
    }
  }
  // a_0 = A(x_0) fed the first half kick
  load_rec(recs[0]);
  if (!vjp_at_current(adj_abA))
    return false;
  if (!hip_ok(hipMemcpy(xb0, adj_xb, sizeof(double) * n3, hipMemcpyDeviceToHost), "download xbar") ||
      !hip_ok(hipMemcpy(vb0, adj_vb, sizeof(double) * n3, hipMemcpyDeviceToHost), "download vbar"))
    return false;
  // back to the final forward state
  restore_ckpt(fin);
  if (!hip_ok(hipStreamSynchronize(stream), "adjoint_backward"))
    return false;
  for (int a = 0; a < 3; a++)
    (void)hipFree(fin.x[a]);
  for (int a = 0; a < 7; a++)
    (void)hipFree(fin.a[a]);
  (void)hipFree(fin.id);
  return hip_ok(hipStreamSynchronize(stream), "adjoint_backward");
}

} // namespace qz
