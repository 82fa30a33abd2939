// Batched masked multigrid for the refinement nodes of one tree level.
//
// Each node's box grid ((rd+1)^3 points, rd even) is the finest multigrid
// level, padded on the high side so every axis has 2^(M-1) * c + 1 points.
// Unknowns ("active" points) are the node's interior grid points; boundary,
// simulation-boundary and absent points are Dirichlet (zero correction).  A
// coarse point is active when the fine point it injects from is active.  All
// level arrays use the colour-split layout per node (every axis length is odd,
// so the linear-index parity is the red/black colour).
//
// Work is handed out through chunk tables: one entry {node, begin, end} per
// workgroup, over a node's interior points, over the linear points of one of
// its multigrid levels (mask set-up), or over the list of active points of a
// level and colour (sweeps, restriction, prolongation: the inactive points of
// the padded boxes are never visited).  A launch covers every node of the tree
// level without idle workgroups, and converged nodes drop out per chunk.
#pragma once
#include <hip/hip_runtime.h>

#include "qz_branch_kernels.h"

namespace qz
{

constexpr int kMaxMgLevels = 8;
constexpr int kBmgCoarseCells = 8;  // cells per axis of a node's coarsest grid, at most
constexpr int kBmgMaxOuter = 60;    // outer iterations before the level is handed to CG
constexpr int kBmgChunkIntr = 1024; // interior points per chunk (4 per thread)
constexpr int kBmgChunkGrid = 512;  // linear grid points per chunk (mask set-up)
constexpr int kBmgChunkAct = 512;   // active points per chunk (sweeps, restriction, prolongation)
// Over-relaxation of the red half sweeps of the node V-cycle above the tail
// (red-black SOR, red points only: the black half sweep stays plain Gauss-
// Seidel, so the black residual is zero and the red-point restriction holds).
// 1.3 takes about 8 % fewer PCG iterations than 1 on a 2^24-particle clumps
// workload; the tail keeps plain Gauss-Seidel.
constexpr double kBmgRedOmega = 1.3;

struct BMg
{
  int X[kMaxMgLevels], Y[kMaxMgLevels], Z[kMaxMgLevels];
  long long off[kMaxMgLevels]; // offset of the node's level-m block in the pools
  FDiv dx[kMaxMgLevels], dxy[kMaxMgLevels]; // divide by X[m], by X[m] Y[m]
  FDiv gx, gxy;                              // divide by the box grid's rd_x + 1, (rd_x + 1)(rd_y + 1)
  int nlev;
  int ts;          // first level of the node's LDS tail (k_bmg_tail): levels ts .. nlev - 1
  int done;        // converged: outer iterations used
  int ic0, ic1;    // the node's interior chunks (entries of the interior table)
  double rz[2];    // (-r).z of the last two iterations
  double alpha[2]; // step lengths of the last two iterations
  double rs0;      // adjoint solves: r.r of the starting residual
};

struct Chunk
{
  int node; // position in the tree level
  int begin, end;
  int color; // colour of the active-list chunks (unused otherwise)
};

// Start of every region of the active list: rstart[r] = first list entry with
// pool index >= rbeg[r] (the list is sorted by pool index).
__global__ void k_bmg_region_starts(const int *__restrict__ alist, const int *__restrict__ acount,
                                    const int *__restrict__ rbeg, int nreg, int *__restrict__ rstart)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nreg)
    return;
  int lo = 0, hi = *acount;
  const int v = rbeg[r];
  while (lo < hi)
  {
    const int mid = (lo + hi) >> 1;
    if (alist[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  rstart[r] = lo;
}

__device__ inline long long bmg_np(const BMg &b, int m) { return (long long)b.X[m] * b.Y[m] * b.Z[m]; }

// this workgroup's chunk (XCD-contiguous numbering: a node's chunks share an L2)
__device__ inline const Chunk &bmg_chunk(const Chunk *tab) { return tab[xcd_block(blockIdx.x, gridDim.x)]; }

// padded colour-split position at MG level 0 of box-grid point g
__device__ inline long long bmg_pos0(const DNode &d, const BMg &b, int g)
{
  const int gX = d.rd[0] + 1, gY = d.rd[1] + 1;
  const int q = (int)fdiv((unsigned)g, b.gx), k = (int)fdiv((unsigned)g, b.gxy);
  const int i = g - q * gX, j = q - k * gY;
  const long long idx = i + (long long)b.X[0] * (j + (long long)b.Y[0] * k);
  const long long hf = (bmg_np(b, 0) + 1) >> 1;
  return b.off[0] + (idx & 1) * hf + (idx >> 1);
}

// fold per-chunk partials over the node's chunk range [c0, c1)
__device__ inline double fold_range(const double *part, int c0, int c1, bool max_op)
{
  __shared__ double s_tot;
  double v = max_op ? -1.0e308 : 0.0;
  for (int i = c0 + threadIdx.x; i < c1; i += blockDim.x)
    v = max_op ? fmax(v, part[i]) : v + part[i];
  v = max_op ? block_reduce<1>(v) : block_reduce<0>(v);
  if (threadIdx.x == 0)
    s_tot = v;
  __syncthreads();
  const double r = s_tot;
  __syncthreads();
  return r;
}

// level-0 mask: 1 at the node's interior points (box grid index -> padded index)
__global__ void k_bmg_mask0(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm,
                            const Chunk *__restrict__ itab, const int *__restrict__ intr_g,
                            unsigned char *__restrict__ mask)
{
  const Chunk &c = bmg_chunk(itab);
  const DNode &d = nodes[first + c.node];
  const BMg &b = bm[c.node];
  for (int t = c.begin + threadIdx.x; t < c.end; t += blockDim.x)
    mask[bmg_pos0(d, b, intr_g[d.poff + t])] = 1;
}

// coarse mask by injection: active(I) = active_fine(2I); table over level m+1
__global__ void k_bmg_mask_coarsen(const BMg *__restrict__ bm, const Chunk *__restrict__ tab, int m,
                                   const unsigned char *__restrict__ mf, unsigned char *__restrict__ mc)
{
  const Chunk &c = bmg_chunk(tab);
  const BMg &b = bm[c.node];
  const long long npc = bmg_np(b, m + 1), hc = (npc + 1) >> 1;
  const long long npf = bmg_np(b, m), hf = (npf + 1) >> 1;
  for (int t = c.begin + threadIdx.x; t < c.end; t += blockDim.x)
  {
    const int JK = (int)fdiv((unsigned)t, b.dx[m + 1]), K = (int)fdiv((unsigned)t, b.dxy[m + 1]);
    const int I = t - JK * b.X[m + 1], J = JK - K * b.Y[m + 1];
    const long long f = 2 * I + (long long)b.X[m] * (2 * J + (long long)b.Y[m] * 2 * K);
    mc[b.off[m + 1] + (t & 1) * hc + (t >> 1)] = mf[b.off[m] + (f & 1) * hf + (f >> 1)];
  }
}

// The outer solver is a flexible preconditioned CG in FP64 on A = -Lap with the
// V-cycle (in G precision) as preconditioner.  Signs: r below is rho - Lap(phi)
// (the reference's residual, conjugate_gradient.c:359-368); the CG residual of
// A phi = -rho is -r, and the V-cycle applied to r returns z ~ A^-1 (-r).
// Interior kernels write one partial per chunk (table entry).

// r = rho - Lap(phi) on interior points: FP64 copy, G copy for the V-cycle,
// partial sums of r^2.
template <typename G>
__global__ void k_bmg_residual(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm,
                               const Chunk *__restrict__ itab, const int *__restrict__ intr_g,
                               const double *__restrict__ phi, const G *__restrict__ rho, double *__restrict__ r,
                               G *__restrict__ r0, double *__restrict__ part)
{
  const unsigned e = xcd_block(blockIdx.x, gridDim.x);
  const Chunk &c = itab[e];
  const DNode &d = nodes[first + c.node];
  const BMg &b = bm[c.node];
  double acc = 0.0;
  if (!b.done)
  {
    const double inv_h2 = (double)(1 << d.lv) * (double)(1 << d.lv);
    const int gX = d.rd[0] + 1, gXY = (d.rd[0] + 1) * (d.rd[1] + 1);
    const double *ph = phi + d.goff;
    for (int t = c.begin + threadIdx.x; t < c.end; t += blockDim.x)
    {
      int g = intr_g[d.poff + t];
      double v = (double)rho[d.goff + g] +
                 inv_h2 * (6.0 * ph[g] - ph[g + 1] - ph[g - 1] - ph[g + gX] - ph[g - gX] - ph[g + gXY] - ph[g - gXY]);
      acc += v * v;
      r[d.goff + g] = v;
      r0[bmg_pos0(d, b, g)] = (G)v;
    }
  }
  acc = block_reduce<0>(acc);
  if (threadIdx.x == 0)
    part[e] = acc;
}

// partial dot products of the new preconditioned residual z (V-cycle output):
// a = sum r z, c = sum z q (q = A p of the previous direction)
template <typename G>
__global__ void k_bmg_zdots(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm,
                            const Chunk *__restrict__ itab, const int *__restrict__ intr_g, const G *__restrict__ z0,
                            const double *__restrict__ r, const double *__restrict__ q, double *__restrict__ part_a,
                            double *__restrict__ part_c)
{
  const unsigned e = xcd_block(blockIdx.x, gridDim.x);
  const Chunk &c = itab[e];
  const DNode &d = nodes[first + c.node];
  const BMg &b = bm[c.node];
  double a = 0.0, cc = 0.0;
  if (!b.done)
  {
    for (int t = c.begin + threadIdx.x; t < c.end; t += blockDim.x)
    {
      int g = intr_g[d.poff + t];
      double z = (double)z0[bmg_pos0(d, b, g)];
      a += r[d.goff + g] * z;
      cc += z * q[d.goff + g];
    }
  }
  a = block_reduce<0>(a);
  cc = block_reduce<0>(cc);
  if (threadIdx.x == 0)
  {
    part_a[e] = a;
    part_c[e] = cc;
  }
}

// p = z + beta p with the flexible (Polak-Ribiere) beta = -alpha (z.q) / (z_old.(-r_old));
// beta = 0 on the first iteration.  rz of this iteration goes to slot it & 1
// (written by the node's first chunk).
template <typename G>
__global__ void k_bmg_direction(const DNode *__restrict__ nodes, int first, BMg *__restrict__ bm,
                                const Chunk *__restrict__ itab, const int *__restrict__ it_base, int it_local,
                                const int *__restrict__ intr_g, const G *__restrict__ z0, double *__restrict__ p,
                                const double *__restrict__ part_a, const double *__restrict__ part_c)
{
  const unsigned e = xcd_block(blockIdx.x, gridDim.x);
  const Chunk &c = itab[e];
  const DNode &d = nodes[first + c.node];
  BMg &b = bm[c.node];
  if (b.done)
    return;
  const int it = *it_base + it_local;
  const double rz = -fold_range(part_a, b.ic0, b.ic1, false); // (-r).z
  double beta = 0.0;
  if (it > 0)
  {
    const double zq = fold_range(part_c, b.ic0, b.ic1, false);
    const double rz_old = b.rz[(it - 1) & 1];
    beta = rz_old != 0.0 ? -b.alpha[(it - 1) & 1] * zq / rz_old : 0.0;
  }
  for (int t = c.begin + threadIdx.x; t < c.end; t += blockDim.x)
  {
    int g = intr_g[d.poff + t];
    long long o = d.goff + g;
    p[o] = (double)z0[bmg_pos0(d, b, g)] + beta * p[o];
  }
  if ((int)e == b.ic0 && threadIdx.x == 0)
    b.rz[it & 1] = rz;
}

// q = A p = -Lap(p) on interior points (p = 0 elsewhere), partial p.q
__global__ void k_bmg_apply(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm,
                            const Chunk *__restrict__ itab, const int *__restrict__ intr_g,
                            const double *__restrict__ p, double *__restrict__ q, double *__restrict__ part)
{
  const unsigned e = xcd_block(blockIdx.x, gridDim.x);
  const Chunk &c = itab[e];
  const DNode &d = nodes[first + c.node];
  const BMg &b = bm[c.node];
  double acc = 0.0;
  if (!b.done)
  {
    const double inv_h2 = (double)(1 << d.lv) * (double)(1 << d.lv);
    const int X = dnode_gX(d), XY = dnode_gXY(d);
    const double *pp = p + d.goff;
    for (int t = c.begin + threadIdx.x; t < c.end; t += blockDim.x)
    {
      int g = intr_g[d.poff + t];
      double v = inv_h2 * (6.0 * pp[g] - pp[g + 1] - pp[g - 1] - pp[g + X] - pp[g - X] - pp[g + XY] - pp[g - XY]);
      q[d.goff + g] = v;
      acc += pp[g] * v;
    }
  }
  acc = block_reduce<0>(acc);
  if (threadIdx.x == 0)
    part[e] = acc;
}

// alpha = rz / p.q; phi += alpha p; r += alpha q (r = rho - Lap phi); G copy of r
// for the next V-cycle; partial r^2 and the pointwise change flag of the update
// (poisson_error.c:36-81).
template <typename G>
__global__ void k_bmg_update(const DNode *__restrict__ nodes, int first, BMg *__restrict__ bm,
                             const Chunk *__restrict__ itab, const int *__restrict__ it_base, int it_local,
                             const int *__restrict__ intr_g, double *__restrict__ phi, const double *__restrict__ p,
                             const double *__restrict__ q, double *__restrict__ r, G *__restrict__ r0,
                             const double *__restrict__ part_pq, double *__restrict__ part_rs,
                             double *__restrict__ part_viol, double tol_change)
{
  const unsigned e = xcd_block(blockIdx.x, gridDim.x);
  const Chunk &c = itab[e];
  const DNode &d = nodes[first + c.node];
  BMg &b = bm[c.node];
  double acc = 0.0, viol = 0.0;
  if (!b.done)
  {
    const int it = *it_base + it_local;
    const double pq = fold_range(part_pq, b.ic0, b.ic1, false);
    const double alpha = pq != 0.0 ? b.rz[it & 1] / pq : 0.0;
    for (int t = c.begin + threadIdx.x; t < c.end; t += blockDim.x)
    {
      int g = intr_g[d.poff + t];
      long long o = d.goff + g;
      double old = phi[o];
      double nw = old + alpha * p[o];
      phi[o] = nw;
      double a = nw / old;
      if ((1.0 - tol_change > a) || (1.0 + tol_change < a) || a != a)
        viol = 1.0;
      double rr = r[o] + alpha * q[o];
      r[o] = rr;
      acc += rr * rr;
      r0[bmg_pos0(d, b, g)] = (G)rr;
    }
    if ((int)e == b.ic0 && threadIdx.x == 0)
      b.alpha[it & 1] = alpha;
  }
  acc = block_reduce<0>(acc);
  viol = block_reduce<1>(viol);
  if (threadIdx.x == 0)
  {
    part_rs[e] = acc;
    part_viol[e] = viol;
  }
}

// Per node: converged when the last correction kept every point within
// [1-tol_change, 1+tol_change] of its previous value (poisson_error.c:36-81)
// and the residual meets the CG criterion of conj_grad_error
// (poisson_error.c:236-288): sum r^2 < nintr (tol 4 pi G rho_mean)^2.
// Adjoint solves (zero start, zero Dirichlet values) stop on the residual
// alone: sum r^2 <= tol^2 times its starting value.  One workgroup per node.
__global__ void k_bmg_check(BMg *__restrict__ bm, DNode *__restrict__ nodes, int first, const int *__restrict__ it_base,
                            int it_local, const double *__restrict__ part_rs, const double *__restrict__ part_viol,
                            double tol, double Gc, int adjoint, int *__restrict__ done_count)
{
  BMg &b = bm[blockIdx.x];
  if (b.done)
    return;
  const int outer = *it_base + it_local;
  const int have_change = outer > 0;
  DNode &d = nodes[first + blockIdx.x];
  const double rs = fold_range(part_rs, b.ic0, b.ic1, false);
  const double viol = have_change && !adjoint ? fold_range(part_viol, b.ic0, b.ic1, true) : 0.0;
  if (threadIdx.x == 0)
  {
    bool ok;
    if (adjoint)
    {
      if (outer == 0)
        b.rs0 = rs;
      ok = rs <= tol * tol * b.rs0;
    }
    else
    {
      const double H = 1.0 / (1 << d.lv);
      const double rhomean = 4.0 * Gc * kPiD * d.mass / ((double)d.ncell * H * H * H);
      const double aux_tol = (double)d.nintr * (tol * rhomean) * (tol * rhomean);
      ok = have_change && viol == 0.0 && rs < aux_tol;
    }
    if (ok || d.nintr == 0)
    {
      b.done = outer + 1;
      d.done = outer + 1;
      atomicAdd(done_count, 1);
    }
  }
}

// red half sweep: e <- e + w (gs - e), from a zero start e <- w gs
template <typename G> __device__ inline G bmg_red_relax(const G *e, long long p, G gs, int zero_start)
{
  const G w = (G)kBmgRedOmega;
  return zero_start ? w * gs : e[p] + w * (gs - e[p]);
}

// red-black half sweep on Lap e = r at MG level m over the active points of one
// colour (chunks of the active list: pool indices of the level's points); the
// red half sweep over-relaxed by kBmgRedOmega
template <typename G>
__global__ void k_bmg_rbgs(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm,
                           const Chunk *__restrict__ tab, const int *__restrict__ alist, int m, G *__restrict__ e,
                           const G *__restrict__ r, int color, int zero_start)
{
  const Chunk &c = bmg_chunk(tab);
  const BMg &b = bm[c.node];
  if (b.done)
    return;
  const double H = (double)(1 << m) / (double)(1 << nodes[first + c.node].lv);
  const G h2 = (G)(H * H);
  const int X = b.X[m], Y = b.Y[m];
  const long long XY = (long long)X * Y, np = XY * b.Z[m], hf = (np + 1) >> 1;
  const long long base = b.off[m] + color * hf; // this colour's half
  const G *o = e + b.off[m] + (color == 0 ? hf : 0);
  for (int i = c.begin + threadIdx.x; i < c.end; i += blockDim.x)
  {
    const long long p = alist[i];
    const G rv = r[p]; // issued ahead of the neighbour loads
    const long long idx = 2 * (p - base) + color;
    G sum = zero_start ? (G)0
                       : o[(idx + 1) >> 1] + o[(idx - 1) >> 1] + o[(idx + X) >> 1] + o[(idx - X) >> 1] +
                             o[(idx + XY) >> 1] + o[(idx - XY) >> 1];
    const G gs = (sum - h2 * rv) * (G)(1.0 / 6.0);
    e[p] = color != 0 ? gs : bmg_red_relax(e, p, gs, zero_start);
  }
}

// Full-weighting restriction of the masked fine residual to the active coarse
// points (active-list chunks of level m+1).  It follows a red-black sweep pair that
// ended on black, so the residual at black fine points is zero and only the 13
// red points of the 27-point stencil contribute: the centre (weight 8/64) and
// the 12 points two offsets away (weight 2/64).
template <typename G>
__global__ void k_bmg_restrict(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm,
                               const Chunk *__restrict__ tab, const int *__restrict__ alist, int m,
                               const G *__restrict__ ef, const G *__restrict__ rf, const unsigned char *__restrict__ mf,
                               G *__restrict__ rc)
{
  const Chunk &c = bmg_chunk(tab);
  const BMg &b = bm[c.node];
  if (b.done)
    return;
  const double Hf = (double)(1 << m) / (double)(1 << nodes[first + c.node].lv);
  const G inv_h2 = (G)(1.0 / (Hf * Hf));
  const int Xf = b.X[m], Yf = b.Y[m];
  const long long XYf = (long long)Xf * Yf, npf = XYf * b.Z[m], hf = (npf + 1) >> 1;
  const int Xc = b.X[m + 1], Yc = b.Y[m + 1];
  const long long npc = (long long)Xc * Yc * b.Z[m + 1], hc = (npc + 1) >> 1;
  // red fine points live in the first half, black ones in the second
  const G *eR = ef + b.off[m], *eB = ef + b.off[m] + hf;
  const G *rR = rf + b.off[m];
  const unsigned char *mR = mf + b.off[m];
  // residual r - Lap e at the red fine point with linear index f (f even), zero
  // where the mask is clear.  An active coarse point has I in [1, X_c - 2] (the
  // masks come from interior points by injection), so the 13 fine points and
  // their neighbours stay inside the node's level-m block: the loads are safe
  // unconditionally and the mask selects the value, which lets all loads issue
  // together.
  auto red_res = [&](long long f) -> G {
    const long long h = f >> 1;
    const bool on = mR[h] != 0;
    G lap = (eB[(f + 1) >> 1] + eB[(f - 1) >> 1] + eB[(f + Xf) >> 1] + eB[(f - Xf) >> 1] + eB[(f + XYf) >> 1] +
             eB[(f - XYf) >> 1] - (G)6 * eR[h]) *
            inv_h2;
    const G v = rR[h] - lap;
    return on ? v : (G)0;
  };
  for (int i = c.begin + threadIdx.x; i < c.end; i += blockDim.x)
  {
    const long long p = alist[i], cp = p - b.off[m + 1];
    const int t = (int)(cp < hc ? 2 * cp : 2 * (cp - hc) + 1); // linear index on the coarse level
    const int JK = (int)fdiv((unsigned)t, b.dx[m + 1]), K = (int)fdiv((unsigned)t, b.dxy[m + 1]);
    const int I = t - JK * Xc, J = JK - K * Yc;
    const long long f = 2 * I + (long long)Xf * (2 * J + (long long)Yf * 2 * K);
    G edges = red_res(f + 1 + Xf) + red_res(f + 1 - Xf) + red_res(f - 1 + Xf) + red_res(f - 1 - Xf) +
              red_res(f + 1 + XYf) + red_res(f + 1 - XYf) + red_res(f - 1 + XYf) + red_res(f - 1 - XYf) +
              red_res(f + Xf + XYf) + red_res(f + Xf - XYf) + red_res(f - Xf + XYf) + red_res(f - Xf - XYf);
    rc[p] = (G)0.125 * red_res(f) + (G)(2.0 / 64.0) * edges;
  }
}

// trilinear prolongation of the coarse correction into the active fine points
// (active-list chunks of level m)
template <typename G>
__global__ void k_bmg_prolong(const BMg *__restrict__ bm, const Chunk *__restrict__ tab, const int *__restrict__ alist,
                              int m, G *__restrict__ ef, const G *__restrict__ ec)
{
  const Chunk &ch = bmg_chunk(tab);
  const BMg &b = bm[ch.node];
  if (b.done)
    return;
  const int Xf = b.X[m], Yf = b.Y[m];
  const long long npf = (long long)Xf * Yf * b.Z[m], hf = (npf + 1) >> 1;
  const int Xc = b.X[m + 1], Yc = b.Y[m + 1];
  const long long npc = (long long)Xc * Yc * b.Z[m + 1], hc = (npc + 1) >> 1;
  const G *c = ec + b.off[m + 1];
  auto PC = [&](int i, int j, int k) {
    long long idx = i + (long long)Xc * (j + (long long)Yc * k);
    return (idx & 1) * hc + (idx >> 1);
  };
  for (int q = ch.begin + threadIdx.x; q < ch.end; q += blockDim.x)
  {
    const long long p = alist[q], fp = p - b.off[m];
    const int t = (int)(fp < hf ? 2 * fp : 2 * (fp - hf) + 1); // linear index on the fine level
    const int JK = (int)fdiv((unsigned)t, b.dx[m]), k = (int)fdiv((unsigned)t, b.dxy[m]);
    const int i = t - JK * Xf, j = JK - k * Yf;
    int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
    int i1 = i0 + (i & 1), j1 = j0 + (j & 1), k1 = k0 + (k & 1);
    G v = (c[PC(i0, j0, k0)] + c[PC(i1, j0, k0)] + c[PC(i0, j1, k0)] + c[PC(i1, j1, k0)] + c[PC(i0, j0, k1)] +
           c[PC(i1, j0, k1)] + c[PC(i0, j1, k1)] + c[PC(i1, j1, k1)]) *
          (G)0.125;
    ef[p] += v;
  }
}

// coarsest level of a node: at most this many points (kBmgCoarseCells + 1 per axis)
constexpr int kBmgCoarseMax = (kBmgCoarseCells + 1) * (kBmgCoarseCells + 1) * (kBmgCoarseCells + 1);

// ---------------------------------------------------------------- fused node levels
// A node level m >= 1 that the tail does not hold runs, for all nodes that
// smooth there, as two launches instead of ten: k_bmg_down (two masked
// pre-smoothing pairs from a zero start and the masked restriction into level
// m + 1) and k_bmg_up (the prolongation from m + 1 and two post-smoothing
// pairs).  The tiling is that of the head's k_mg_down / k_mg_up: a tile owns
// kNmCT^3 coarse points (I0 .. I0 + 3 per axis, coarse interior only) and the
// fine points [2 I0 - 1, 2 I0 + 6], stages the fine region [2 I0 - 5, 2 I0 + 11]
// (17 per axis) of the node's level in LDS with r, e and the active mask, and
// narrows the region by one point per sweep phase, so the owned points end as
// k_bmg_rbgs, k_bmg_restrict and k_bmg_prolong leave them (same formulas and
// summation orders).  Points outside the node's padded box count as inactive.
constexpr int kNmCT = 4, kNmLF = 17, kNmNP = kNmLF * kNmLF * kNmLF, kNmHF = (kNmNP + 1) / 2;
constexpr int kNmThreads = 512;

struct NTile
{
  int node, I0, J0, K0; // node position in the tree level, first coarse point of the tile
};

// masked red-black half sweep of the staged region [S, kNmLF - 1 - S]^3 (RED:
// the red half sweep, over-relaxed as in k_bmg_rbgs)
template <typename G, int S, bool ZS, bool RED>
__device__ inline void nm_sweep(G *sE, const G *sR, const unsigned char *sM, int lp, G h2)
{
  constexpr int LF = kNmLF, LF2 = LF * LF, L = LF - 2 * S, HX = (L + 1) / 2, SLOTS = HX * L * L;
  const G *o = sE + (1 - lp) * kNmHF;
  for (int q = threadIdx.x; q < SLOTS; q += blockDim.x)
  {
    const int row = q / HX, u = q - row * HX;
    const int y = S + row % L, z = S + row / L;
    const int x = S + ((S + y + z + lp) & 1) + 2 * u;
    if (x > LF - 1 - S)
      continue;
    const int idx = x + LF * y + LF2 * z;
    const int p = lp * kNmHF + (idx >> 1);
    if (!sM[p])
      continue;
    const G rv = sR[p];
    G sum = ZS ? (G)0
               : o[(idx + 1) >> 1] + o[(idx - 1) >> 1] + o[(idx + LF) >> 1] + o[(idx - LF) >> 1] +
                     o[(idx + LF2) >> 1] + o[(idx - LF2) >> 1];
    const G gs = (sum - h2 * rv) * (G)(1.0 / 6.0);
    sE[p] = RED ? bmg_red_relax(sE, p, gs, ZS ? 1 : 0) : gs;
  }
  __syncthreads();
}

// stage r and the mask (and, for k_bmg_up, the pre-smoothed correction) of the
// tile's fine region of node level m; all global loads before the LDS stores
template <typename G, bool WITH_E>
__device__ inline void nm_stage(const BMg &b, int m, int fx, int fy, int fz, const G *r, const unsigned char *mask,
                                const G *et, G *sR, unsigned char *sM, G *sT)
{
  constexpr int LF = kNmLF, LF2 = LF * LF, SU = (kNmNP + kNmThreads - 1) / kNmThreads;
  const int X = b.X[m], Y = b.Y[m], Z = b.Z[m];
  const long long np = (long long)X * Y * Z, hf = (np + 1) >> 1;
  const G *rb = r + b.off[m];
  const unsigned char *mb = mask + b.off[m];
  const G *tb = et + b.off[m];
  G rv[SU], tv[SU];
  unsigned char mv[SU];
#pragma unroll
  for (int q = 0; q < SU; q++)
  {
    const int t = threadIdx.x + q * kNmThreads;
    const int x = t % LF, y = (t / LF) % LF, z = t / LF2;
    const int gx = fx + x, gy = fy + y, gz = fz + z;
    const bool in = t < kNmNP && gx >= 0 && gx < X && gy >= 0 && gy < Y && gz >= 0 && gz < Z;
    const long long g = in ? gx + (long long)X * (gy + (long long)Y * gz) : 0;
    const long long gp = (g & 1) * hf + (g >> 1);
    const G a = rb[gp];
    const unsigned char c = mb[gp];
    rv[q] = in ? a : (G)0;
    mv[q] = in ? c : (unsigned char)0;
    if (WITH_E)
    {
      // inactive points are Dirichlet zeros, whatever et holds there (points
      // on the box faces belong to no tile and are never written)
      const G ev = tb[gp];
      tv[q] = in && c ? ev : (G)0;
    }
  }
#pragma unroll
  for (int q = 0; q < SU; q++)
  {
    const int t = threadIdx.x + q * kNmThreads;
    if (t < kNmNP)
    {
      const int p = (t & 1) * kNmHF + (t >> 1);
      sR[p] = rv[q];
      sM[p] = mv[q];
      sT[p] = WITH_E ? tv[q] : (G)0;
    }
  }
}

template <typename G>
__global__ void __launch_bounds__(kNmThreads)
    k_bmg_down(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm, const NTile *__restrict__ tiles,
               int m, G *__restrict__ r, G *__restrict__ et, const unsigned char *__restrict__ mask)
{
  constexpr int LF = kNmLF, LF2 = LF * LF, HF = kNmHF;
  __shared__ G sE[kNmNP], sR[kNmNP];
  __shared__ unsigned char sM[kNmNP];
  const NTile tl = tiles[xcd_block(blockIdx.x, gridDim.x)];
  const BMg &b = bm[tl.node];
  if (b.done)
    return;
  const double H = (double)(1 << m) / (double)(1 << nodes[first + tl.node].lv);
  const G h2 = (G)(H * H), inv_h2 = (G)(1.0 / (H * H));
  const int fx = 2 * tl.I0 - 5, fy = 2 * tl.J0 - 5, fz = 2 * tl.K0 - 5;
  const int red = (fx + fy + fz) & 1, black = red ^ 1; // local parity of the red points
  nm_stage<G, false>(b, m, fx, fy, fz, r, mask, (const G *)nullptr, sR, sM, sE); // sE = 0
  __syncthreads();
  nm_sweep<G, 0, true, true>(sE, sR, sM, red, h2);
  nm_sweep<G, 1, false, false>(sE, sR, sM, black, h2);
  nm_sweep<G, 2, false, true>(sE, sR, sM, red, h2);
  nm_sweep<G, 3, false, false>(sE, sR, sM, black, h2);
  const int X = b.X[m], Y = b.Y[m], Z = b.Z[m];
  const long long np = (long long)X * Y * Z, hf = (np + 1) >> 1;
  {
    // masked residual at the red points of [4, 12]^3, in place of r (k_bmg_restrict's red_res)
    constexpr int S = 4, L = LF - 2 * S, HX = (L + 1) / 2, SLOTS = HX * L * L;
    const G *o = sE + black * HF;
    for (int q = threadIdx.x; q < SLOTS; q += blockDim.x)
    {
      const int row = q / HX, u = q - row * HX;
      const int y = S + row % L, z = S + row / L;
      const int x = S + ((S + y + z + red) & 1) + 2 * u;
      if (x > LF - 1 - S)
        continue;
      const int idx = x + LF * y + LF2 * z;
      const int p = red * HF + (idx >> 1);
      const G lap = (o[(idx + 1) >> 1] + o[(idx - 1) >> 1] + o[(idx + LF) >> 1] + o[(idx - LF) >> 1] +
                     o[(idx + LF2) >> 1] + o[(idx - LF2) >> 1] - (G)6 * sE[p]) *
                    inv_h2;
      const G v = sR[p] - lap;
      sR[p] = sM[p] ? v : (G)0;
    }
    // the smoothed correction of the owned points [4, 11]^3 inside the box
    G *tb = et + b.off[m];
    for (int q = threadIdx.x; q < 512; q += blockDim.x)
    {
      const int x = 4 + (q & 7), y = 4 + ((q >> 3) & 7), z = 4 + (q >> 6);
      const int gx = fx + x, gy = fy + y, gz = fz + z;
      if (gx < 0 || gx >= X || gy < 0 || gy >= Y || gz < 0 || gz >= Z)
        continue;
      const long long g = gx + (long long)X * (gy + (long long)Y * gz);
      const int idx = x + LF * y + LF2 * z;
      tb[(g & 1) * hf + (g >> 1)] = sE[(idx & 1) * HF + (idx >> 1)];
    }
  }
  __syncthreads();
  // restriction to the tile's active coarse points (k_bmg_restrict's order)
  const int Xc = b.X[m + 1], Yc = b.Y[m + 1], Zc = b.Z[m + 1];
  const long long npc = (long long)Xc * Yc * Zc, hc = (npc + 1) >> 1;
  const G *res = sR + red * HF;
  for (int q = threadIdx.x; q < kNmCT * kNmCT * kNmCT; q += blockDim.x)
  {
    const int I = tl.I0 + (q & 3), J = tl.J0 + ((q >> 2) & 3), K = tl.K0 + (q >> 4);
    if (I > Xc - 2 || J > Yc - 2 || K > Zc - 2)
      continue;
    const long long c = I + (long long)Xc * (J + (long long)Yc * K);
    const long long cp = b.off[m + 1] + (c & 1) * hc + (c >> 1);
    if (!mask[cp])
      continue;
    const int idx = (2 * (q & 3) + 5) + LF * (2 * ((q >> 2) & 3) + 5) + LF2 * (2 * (q >> 4) + 5);
    auto R = [&](int d) { return res[(idx + d) >> 1]; };
    const G edges = R(1 + LF) + R(1 - LF) + R(-1 + LF) + R(-1 - LF) + R(1 + LF2) + R(1 - LF2) + R(-1 + LF2) +
                    R(-1 - LF2) + R(LF + LF2) + R(LF - LF2) + R(-LF + LF2) + R(-LF - LF2);
    r[cp] = (G)0.125 * R(0) + (G)(2.0 / 64.0) * edges;
  }
}

template <typename G>
__global__ void __launch_bounds__(kNmThreads)
    k_bmg_up(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm, const NTile *__restrict__ tiles,
             int m, G *__restrict__ e, const G *__restrict__ et, const G *__restrict__ r,
             const unsigned char *__restrict__ mask)
{
  constexpr int LF = kNmLF, LF2 = LF * LF, HF = kNmHF;
  constexpr int LC = kNmCT + 6; // coarse points per axis under the region
  __shared__ G sE[kNmNP], sR[kNmNP], sC[LC * LC * LC];
  __shared__ unsigned char sM[kNmNP];
  const NTile tl = tiles[xcd_block(blockIdx.x, gridDim.x)];
  const BMg &b = bm[tl.node];
  if (b.done)
    return;
  const double H = (double)(1 << m) / (double)(1 << nodes[first + tl.node].lv);
  const G h2 = (G)(H * H);
  const int fx = 2 * tl.I0 - 5, fy = 2 * tl.J0 - 5, fz = 2 * tl.K0 - 5;
  const int cx = tl.I0 - 3, cy = tl.J0 - 3, cz = tl.K0 - 3;
  const int red = (fx + fy + fz) & 1, black = red ^ 1;
  {
    const int Xc = b.X[m + 1], Yc = b.Y[m + 1], Zc = b.Z[m + 1];
    const long long npc = (long long)Xc * Yc * Zc, hc = (npc + 1) >> 1;
    const G *cb = e + b.off[m + 1];
    constexpr int CU = (LC * LC * LC + kNmThreads - 1) / kNmThreads;
    G cv[CU];
#pragma unroll
    for (int q = 0; q < CU; q++)
    {
      const int t = threadIdx.x + q * kNmThreads;
      const int gx = cx + t % LC, gy = cy + (t / LC) % LC, gz = cz + t / (LC * LC);
      const bool in = t < LC * LC * LC && gx >= 0 && gx < Xc && gy >= 0 && gy < Yc && gz >= 0 && gz < Zc;
      const long long g = in ? gx + (long long)Xc * (gy + (long long)Yc * gz) : 0;
      const G v = cb[(g & 1) * hc + (g >> 1)];
      cv[q] = in ? v : (G)0;
    }
#pragma unroll
    for (int q = 0; q < CU; q++)
    {
      const int t = threadIdx.x + q * kNmThreads;
      if (t < LC * LC * LC)
        sC[t] = cv[q];
    }
  }
  nm_stage<G, true>(b, m, fx, fy, fz, r, mask, et, sR, sM, sE); // sE = the pre-smoothed correction
  __syncthreads();
  // prolongation into the active fine points (k_bmg_prolong's order)
  for (int t = threadIdx.x; t < kNmNP; t += blockDim.x)
  {
    const int p = (t & 1) * HF + (t >> 1);
    if (!sM[p])
      continue;
    const int gx = fx + t % LF, gy = fy + (t / LF) % LF, gz = fz + t / LF2;
    const int i0 = gx >> 1, j0 = gy >> 1, k0 = gz >> 1;
    const int i1 = i0 + (gx & 1), j1 = j0 + (gy & 1), k1 = k0 + (gz & 1);
    auto C = [&](int i, int j, int k) { return sC[(i - cx) + LC * ((j - cy) + LC * (k - cz))]; };
    const G v = (C(i0, j0, k0) + C(i1, j0, k0) + C(i0, j1, k0) + C(i1, j1, k0) + C(i0, j0, k1) + C(i1, j0, k1) +
                 C(i0, j1, k1) + C(i1, j1, k1)) *
                (G)0.125;
    sE[p] += v;
  }
  __syncthreads();
  nm_sweep<G, 1, false, true>(sE, sR, sM, red, h2);
  nm_sweep<G, 2, false, false>(sE, sR, sM, black, h2);
  nm_sweep<G, 3, false, true>(sE, sR, sM, red, h2);
  nm_sweep<G, 4, false, false>(sE, sR, sM, black, h2);
  const int X = b.X[m], Y = b.Y[m], Z = b.Z[m];
  const long long np = (long long)X * Y * Z, hf = (np + 1) >> 1;
  G *ebase = e + b.off[m];
  for (int q = threadIdx.x; q < 512; q += blockDim.x)
  {
    const int x = 4 + (q & 7), y = 4 + ((q >> 3) & 7), z = 4 + (q >> 6);
    const int gx = fx + x, gy = fy + y, gz = fz + z;
    if (gx < 0 || gx >= X || gy < 0 || gy >= Y || gz < 0 || gz >= Z)
      continue;
    const long long g = gx + (long long)X * (gy + (long long)Y * gz);
    const int idx = x + LF * y + LF2 * z;
    ebase[(g & 1) * hf + (g >> 1)] = sE[(idx & 1) * HF + (idx >> 1)];
  }
}

// ---------------------------------------------------------------- node tail in one workgroup
// The coarse end of a node's V-cycle, levels ts .. nlev - 1, in one workgroup
// per node with every level staged in LDS: the pre-smoothing sweeps and the
// restriction on the way down, the masked CG of the coarsest level, the
// prolongation and post-smoothing sweeps on the way up, separated by workgroup
// barriers instead of launches.  Same update formulas and summation orders as
// k_bmg_rbgs, k_bmg_restrict and k_bmg_prolong; the coarsest level by masked
// CG in FP64 to a 1e4 residual reduction.
// The restriction into level ts (and the prolongation from it) stay with the
// per-level kernels.  LDS: the coarsest level's FP64 CG vectors (x, r, p),
// then per level e and r (G, colour-split as in the pools) and the active mask
// as bits; ts is the smallest level whose tail fits kBmgTailLds.
constexpr int kBmgTailThreads = 1024;
constexpr size_t kBmgTailLds = 163840 - 2048; // the workgroup's LDS less the static reduction arrays

template <typename G> __host__ __device__ inline size_t bmg_tail_bytes(const BMg &b, int ts)
{
  const int mc = b.nlev - 1;
  size_t s = 3 * sizeof(double) * (size_t)b.X[mc] * b.Y[mc] * b.Z[mc];
  for (int m = ts; m < b.nlev; m++)
  {
    const size_t np = (size_t)b.X[m] * b.Y[m] * b.Z[m];
    s += 2 * sizeof(G) * np + 4 * ((np + 31) / 32);
    s = (s + 7) & ~(size_t)7;
  }
  return s;
}

// Workgroup sum with one barrier: each wave folds its lanes, lane 0 stores the
// wave's partial in slot set `buf` of sh, and after the barrier every thread
// adds the partials in wave order (the same value everywhere).  Alternating
// buf between consecutive calls keeps a call's stores clear of the previous
// call's reads (a thread reaches them only past the intervening barrier).
__device__ inline double tail_sum(double v, double (*sh)[kBmgTailThreads / 64], int buf)
{
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0)
    sh[buf][threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  const int nw = (int)((blockDim.x + 63) >> 6);
  for (int w = 0; w < nw; w++)
    r += sh[buf][w];
  return r;
}

template <typename G>
__global__ void __launch_bounds__(kBmgTailThreads)
    k_bmg_tail(const DNode *__restrict__ nodes, int first, const BMg *__restrict__ bm, G *__restrict__ e,
               const G *__restrict__ r, const unsigned char *__restrict__ mask, int pre, int post, int iters)
{
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ unsigned s_off[kMaxMgLevels][3]; // byte offsets of e, r, mask bits per level
  __shared__ double s_red[2][kBmgTailThreads / 64]; // tail_sum partials
  const BMg &b = bm[blockIdx.x];
  if (b.done)
    return;
  const int ts = b.ts, M = b.nlev, mc = M - 1;
  const int tid = threadIdx.x, NT = blockDim.x;
  const double lvs = (double)(1 << nodes[first + blockIdx.x].lv);
  const int npc = b.X[mc] * b.Y[mc] * b.Z[mc];
  double *xs = (double *)smem, *rs = xs + npc, *ps = rs + npc;
  if (tid == 0)
  {
    size_t o = 3 * sizeof(double) * (size_t)npc;
    for (int m = ts; m < M; m++)
    {
      const size_t np = (size_t)b.X[m] * b.Y[m] * b.Z[m];
      s_off[m][0] = (unsigned)o;
      s_off[m][1] = (unsigned)(o + sizeof(G) * np);
      s_off[m][2] = (unsigned)(o + 2 * sizeof(G) * np);
      o += 2 * sizeof(G) * np + 4 * ((np + 31) / 32);
      o = (o + 7) & ~(size_t)7;
    }
  }
  __syncthreads();
  auto E = [&](int m) { return (G *)(smem + s_off[m][0]); };
  auto R = [&](int m) { return (G *)(smem + s_off[m][1]); };
  auto Bt = [&](int m) { return (unsigned *)(smem + s_off[m][2]); };
  auto on = [&](int m, int pos) { return (Bt(m)[pos >> 5] >> (pos & 31)) & 1u; };
  // stage: e = 0 and the mask bits on every tail level, r of level ts.  The
  // global loads go kStageU per thread at a time, all issued before their
  // LDS stores; a wave's 64 mask bytes become two mask words by ballot.
  constexpr int kStageU = 8;
  for (int m = ts; m < M; m++)
  {
    const int np = b.X[m] * b.Y[m] * b.Z[m], nw = (np + 31) >> 5;
    G *Em = E(m);
    for (int t = tid; t < np; t += NT)
      Em[t] = (G)0;
    const unsigned char *mk = mask + b.off[m];
    unsigned *Bm = Bt(m);
    for (int t0 = 0; t0 < np; t0 += kStageU * NT)
    {
      unsigned char v[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; u++)
      {
        const int p = t0 + u * NT + tid;
        v[u] = mk[p < np ? p : 0];
      }
#pragma unroll
      for (int u = 0; u < kStageU; u++)
      {
        const int p = t0 + u * NT + tid;
        const unsigned long long bal = __ballot(p < np && v[u] != 0);
        if ((tid & 63) == 0)
        {
          const int w = (t0 + u * NT + tid) >> 5;
          if (w < nw)
            Bm[w] = (unsigned)bal;
          if (w + 1 < nw)
            Bm[w + 1] = (unsigned)(bal >> 32);
        }
      }
    }
  }
  {
    const int np = b.X[ts] * b.Y[ts] * b.Z[ts];
    const G *rg = r + b.off[ts];
    G *Rt = R(ts);
    for (int t0 = 0; t0 < np; t0 += kStageU * NT)
    {
      G v[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; u++)
      {
        const int p = t0 + u * NT + tid;
        v[u] = rg[p < np ? p : 0];
      }
#pragma unroll
      for (int u = 0; u < kStageU; u++)
      {
        const int p = t0 + u * NT + tid;
        if (p < np)
          Rt[p] = v[u];
      }
    }
  }
  __syncthreads();
  auto sweep = [&](int m, int color, int zs) {
    const double H = (double)(1 << m) / lvs;
    const G h2 = (G)(H * H);
    const int X = b.X[m], XY = X * b.Y[m], np = XY * b.Z[m], hf = (np + 1) >> 1;
    const int cnt = color == 0 ? hf : np - hf;
    G *Em = E(m);
    const G *Rm = R(m);
    const G *o = Em + (color == 0 ? hf : 0);
    for (int h = tid; h < cnt; h += NT)
    {
      const int pos = color * hf + h;
      if (!on(m, pos))
        continue;
      const G rv = Rm[pos];
      const int idx = 2 * h + color;
      G sum = zs ? (G)0
                 : o[(idx + 1) >> 1] + o[(idx - 1) >> 1] + o[(idx + X) >> 1] + o[(idx - X) >> 1] +
                       o[(idx + XY) >> 1] + o[(idx - XY) >> 1];
      Em[pos] = (sum - h2 * rv) * (G)(1.0 / 6.0);
    }
    __syncthreads();
  };
  auto restrict_ = [&](int m) { // level m -> m + 1 after a sweep pair ending on black
    const double Hf = (double)(1 << m) / lvs;
    const G inv_h2 = (G)(1.0 / (Hf * Hf));
    const int Xf = b.X[m], Yf = b.Y[m], XYf = Xf * Yf, hf = (XYf * b.Z[m] + 1) >> 1;
    const int Xc = b.X[m + 1], npc_ = Xc * b.Y[m + 1] * b.Z[m + 1], hc = (npc_ + 1) >> 1;
    const G *eR = E(m), *eB = E(m) + hf, *rR = R(m);
    G *Rc = R(m + 1);
    auto red_res = [&](int f) -> G {
      const int h = f >> 1;
      const bool act = on(m, h) != 0;
      G lap = (eB[(f + 1) >> 1] + eB[(f - 1) >> 1] + eB[(f + Xf) >> 1] + eB[(f - Xf) >> 1] + eB[(f + XYf) >> 1] +
               eB[(f - XYf) >> 1] - (G)6 * eR[h]) *
              inv_h2;
      const G v = rR[h] - lap;
      return act ? v : (G)0;
    };
    for (int pos = tid; pos < npc_; pos += NT)
    {
      if (!on(m + 1, pos))
        continue;
      const int t = pos < hc ? 2 * pos : 2 * (pos - hc) + 1;
      const int JK = (int)fdiv((unsigned)t, b.dx[m + 1]), K = (int)fdiv((unsigned)t, b.dxy[m + 1]);
      const int I = t - JK * Xc, J = JK - K * b.Y[m + 1];
      const int f = 2 * I + Xf * (2 * J + Yf * 2 * K);
      G edges = red_res(f + 1 + Xf) + red_res(f + 1 - Xf) + red_res(f - 1 + Xf) + red_res(f - 1 - Xf) +
                red_res(f + 1 + XYf) + red_res(f + 1 - XYf) + red_res(f - 1 + XYf) + red_res(f - 1 - XYf) +
                red_res(f + Xf + XYf) + red_res(f + Xf - XYf) + red_res(f - Xf + XYf) + red_res(f - Xf - XYf);
      Rc[pos] = (G)0.125 * red_res(f) + (G)(2.0 / 64.0) * edges;
    }
    __syncthreads();
  };
  auto prolong = [&](int m) { // level m + 1 -> m, e_m += P e_{m+1} on the active points
    const int Xf = b.X[m], Yf = b.Y[m], npf = Xf * Yf * b.Z[m], hf = (npf + 1) >> 1;
    const int Xc = b.X[m + 1], Yc = b.Y[m + 1], npc_ = Xc * Yc * b.Z[m + 1], hc = (npc_ + 1) >> 1;
    const G *c = E(m + 1);
    G *Em = E(m);
    auto PC = [&](int i, int j, int k) {
      const int idx = i + Xc * (j + Yc * k);
      return (idx & 1) * hc + (idx >> 1);
    };
    for (int pos = tid; pos < npf; pos += NT)
    {
      if (!on(m, pos))
        continue;
      const int t = pos < hf ? 2 * pos : 2 * (pos - hf) + 1;
      const int JK = (int)fdiv((unsigned)t, b.dx[m]), k = (int)fdiv((unsigned)t, b.dxy[m]);
      const int i = t - JK * Xf, j = JK - k * Yf;
      const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
      const int i1 = i0 + (i & 1), j1 = j0 + (j & 1), k1 = k0 + (k & 1);
      const G v = (c[PC(i0, j0, k0)] + c[PC(i1, j0, k0)] + c[PC(i0, j1, k0)] + c[PC(i1, j1, k0)] +
                   c[PC(i0, j0, k1)] + c[PC(i1, j0, k1)] + c[PC(i0, j1, k1)] + c[PC(i1, j1, k1)]) *
                  (G)0.125;
      Em[pos] += v;
    }
    __syncthreads();
  };
  auto coarse = [&]() { // masked CG on level mc in FP64
    const int X = b.X[mc], XY = X * b.Y[mc], np = npc, hf = (np + 1) >> 1;
    const double H = (double)(1 << mc) / lvs;
    const double inv_h2 = 1.0 / (H * H);
    G *Em = E(mc);
    const G *Rm = R(mc);
    auto P = [&](int idx) { return (idx & 1) * hf + (idx >> 1); };
    for (int t = tid; t < np; t += NT)
    {
      const bool in = on(mc, P(t)) != 0;
      xs[t] = 0.0;
      rs[t] = in ? -(double)Rm[P(t)] : 0.0; // A = -Lap (SPD): solve A x = -r
      ps[t] = rs[t];
    }
    __syncthreads();
    double rr = 0.0;
    for (int t = tid; t < np; t += NT)
      rr += rs[t] * rs[t];
    int buf = 0;
    rr = tail_sum(rr, s_red, buf);
    buf ^= 1;
    const double rr0 = rr;
    constexpr int QL = (kBmgCoarseMax + kBmgTailThreads - 1) / kBmgTailThreads;
    for (int it = 0; it < iters && rr > 1e-8 * rr0 && rr > 0.0; it++)
    {
      double pap = 0.0, ap_local[QL];
      int q = 0;
      for (int t = tid; t < np; t += NT, q++)
      {
        double v = 0.0;
        if (on(mc, P(t)))
          v = inv_h2 * (6.0 * ps[t] - ps[t + 1] - ps[t - 1] - ps[t + X] - ps[t - X] - ps[t + XY] - ps[t - XY]);
        ap_local[q] = v;
        pap += ps[t] * v;
      }
      pap = tail_sum(pap, s_red, buf);
      buf ^= 1;
      const double alpha = rr / pap;
      double rn = 0.0;
      q = 0;
      for (int t = tid; t < np; t += NT, q++)
      {
        xs[t] += alpha * ps[t];
        rs[t] -= alpha * ap_local[q];
        rn += rs[t] * rs[t];
      }
      rn = tail_sum(rn, s_red, buf);
      buf ^= 1;
      const double beta = rn / rr;
      for (int t = tid; t < np; t += NT)
        ps[t] = rs[t] + beta * ps[t];
      rr = rn;
      __syncthreads();
    }
    for (int t = tid; t < np; t += NT)
      Em[P(t)] = (G)xs[t];
    __syncthreads();
  };
  for (int m = ts; m < mc; m++)
  {
    for (int sw = 0; sw < pre; sw++)
    {
      sweep(m, 0, sw == 0);
      sweep(m, 1, 0);
    }
    restrict_(m);
  }
  coarse();
  for (int m = mc - 1; m >= ts; m--)
  {
    prolong(m);
    for (int sw = 0; sw < post; sw++)
    {
      sweep(m, 0, 0);
      sweep(m, 1, 0);
    }
  }
  {
    const int np = b.X[ts] * b.Y[ts] * b.Z[ts];
    G *eg = e + b.off[ts];
    const G *Et = E(ts);
    for (int t = tid; t < np; t += NT)
      eg[t] = Et[t];
  }
}

} // namespace qz
