// Adjoint (reverse-mode) kernels of the head-grid PM step.
//
// Forward map of one force pass at positions x (head only):
//   rho_g = c sum_i m_i W(x_i, g)                    (CIC deposit, c = 4 pi G / H^3)
//   Lap_h phi = rho on interior points, faces fixed   (multigrid)
//   G_g = (phi[g - e] - phi[g + e]) / 2H per axis     (interior; faces fixed)
//   a_i = sum_g W(x_i, g) G_g                         (CIC gather)
// Its vector-Jacobian product for a cotangent abar:
//   Gbar_g = sum_i W(x_i, g) abar_i                   (a scatter: the deposit with abar as mass)
//   phibar = transpose of the central difference applied to Gbar
//   Lap_h psi = phibar, zero faces                    (the same self-adjoint solve)
//   xbar_i += sum_g grad W(x_i, g) (c m_i psi_g + abar_i . G_g)
#pragma once
#include <hip/hip_runtime.h>

#include "qz_branch_kernels.h"
#include "qz_kernels.h"

namespace qz
{

// dst[id[p]] = src[p] (sorted slot order -> ID order) and back
template <typename T>
__global__ void k_slot_to_id(const T *__restrict__ src, const int *__restrict__ id, T *__restrict__ dst, int N)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N)
    dst[id[p]] = src[p];
}

template <typename T>
__global__ void k_id_to_slot(const T *__restrict__ src, const int *__restrict__ id, T *__restrict__ dst, int N)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N)
    dst[p] = src[id[p]];
}

// dst[i] = (P) src[i] over n entries (a cotangent as the deposit's mass type)
template <typename P>
__global__ void k_to_mass_type(const double *__restrict__ src, P *__restrict__ dst, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    dst[i] = (P)src[i];
}

// The cotangents of positions and velocities (3 components each, stride N)
// and the masses, from the slot order of a step's force pass to the slot order
// before that step's sort: the sort put the particle of old slot perm[s] at
// slot s, so dst[perm[s]] = src[s].  The permutation is mostly local (a
// particle moves less than a cell per step), so the stores stay near-coalesced.
template <typename P>
__global__ void k_adj_unsort(const int *__restrict__ perm, const double *__restrict__ xb,
                             const double *__restrict__ vb, const P *__restrict__ m, double *__restrict__ xb_out,
                             double *__restrict__ vb_out, P *__restrict__ m_out, int N)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N)
    return;
  const int d = perm[s];
  double x[3], v[3];
  for (int a = 0; a < 3; a++)
  {
    x[a] = xb[(size_t)a * N + s];
    v[a] = vb[(size_t)a * N + s];
  }
  const P mm = m[s];
  for (int a = 0; a < 3; a++)
  {
    xb_out[(size_t)a * N + d] = x[a];
    vb_out[(size_t)a * N + d] = v[a];
  }
  m_out[d] = mm;
}

// The scatter of the three components of abar (stride N, the deposit's mass
// type) in one pass over the keys and positions: k_deposit_runs with three
// masses, each component's run tails in its own array (t0, t1, t2).
template <typename P, typename S>
__global__ void k_deposit_runs3(const unsigned long long *__restrict__ key, const double *__restrict__ x,
                                const double *__restrict__ y, const double *__restrict__ z, const P *__restrict__ m,
                                int N, int lv, int seg_shift, int clamp_top, S *__restrict__ t0, S *__restrict__ t1,
                                S *__restrict__ t2)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = p < N;
  const unsigned long long c = live ? (key[p] >> seg_shift) : ~0ull;
  double v[3][8];
  if (live)
  {
    const double s = (double)(1 << lv);
    const double px = x[p] * s, py = y[p] * s, pz = z[p] * s;
    const int fx = min((int)px, clamp_top), fy = min((int)py, clamp_top), fz = min((int)pz, clamp_top);
    const double wx1 = px - fx, wy1 = py - fy, wz1 = pz - fz;
    const double wx2 = 1.0 - wx1, wy2 = 1.0 - wy1, wz2 = 1.0 - wz1;
    const double w[8] = {wx2 * wy2 * wz2, wx1 * wy2 * wz2, wx2 * wy1 * wz2, wx1 * wy1 * wz2,
                         wx2 * wy2 * wz1, wx1 * wy2 * wz1, wx2 * wy1 * wz1, wx1 * wy1 * wz1};
#pragma unroll
    for (int a = 0; a < 3; a++)
    {
      const double mm = (double)m[(size_t)a * N + p];
#pragma unroll
      for (int q = 0; q < 8; q++)
        v[a][q] = mm * w[q];
    }
  }
  else
  {
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int q = 0; q < 8; q++)
        v[a][q] = 0.0;
  }
  wave_seg_scan(v[0], c);
  wave_seg_scan(v[1], c);
  wave_seg_scan(v[2], c);
  const unsigned long long cn = shfl_down_u64(c, 1);
  if (live && (lane == 63 || p == N - 1 || cn != c))
  {
    S *const t[3] = {t0, t1, t2};
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int q = 0; q < 8; q++)
        t[a][8 * (size_t)p + q] = (S)v[a][q];
  }
}

// k_cell_totals and k_cell_totals_heavy over three tail arrays (one heavy list).
template <typename S>
__global__ void k_cell_totals3(const int *__restrict__ cb, const int *__restrict__ ce, long long ncell,
                               S *__restrict__ t0, S *__restrict__ t1, S *__restrict__ t2, int *__restrict__ heavy,
                               int *__restrict__ nheavy)
{
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell)
    return;
  const int a = cb[c], e = ce ? ce[c] : cb[c + 1];
  if (e - a <= 0 || (a >> 6) == ((e - 1) >> 6))
    return;
  if (((e - 1) >> 6) - (a >> 6) + 1 > kTotalsSerial)
  {
    heavy[atomicAdd(nheavy, 1)] = (int)c;
    return;
  }
  S *const t[3] = {t0, t1, t2};
  for (int k = 0; k < 3; k++)
  {
    double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int u = a; u < e;)
    {
      const int tl = min(e, ((u >> 6) + 1) << 6) - 1;
#pragma unroll
      for (int q = 0; q < 8; q++)
        sum[q] += (double)t[k][8 * (size_t)tl + q];
      u = tl + 1;
    }
#pragma unroll
    for (int q = 0; q < 8; q++)
      t[k][8 * (size_t)(e - 1) + q] = (S)sum[q];
  }
}

template <typename S>
__global__ void k_cell_totals_heavy3(const int *__restrict__ cb, const int *__restrict__ ce,
                                     const int *__restrict__ heavy, const int *__restrict__ nheavy, S *__restrict__ t0,
                                     S *__restrict__ t1, S *__restrict__ t2)
{
  const int nh = *nheavy;
  S *const t[3] = {t0, t1, t2};
  for (int h = blockIdx.x; h < nh; h += gridDim.x)
  {
    const int c = heavy[h];
    const int a = cb[c], e = ce ? ce[c] : cb[c + 1];
    for (int k = 0; k < 3; k++)
    {
      double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int w = (a >> 6) + (int)threadIdx.x; w <= ((e - 1) >> 6); w += blockDim.x)
      {
        const int tl = min(e, (w + 1) << 6) - 1;
#pragma unroll
        for (int q = 0; q < 8; q++)
          sum[q] += (double)t[k][8 * (size_t)tl + q];
      }
#pragma unroll
      for (int q = 0; q < 8; q++)
        sum[q] = block_reduce<0>(sum[q]); // ends with a barrier: every window is read before the write
      if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 8; q++)
          t[k][8 * (size_t)(e - 1) + q] = (S)sum[q];
      __syncthreads();
    }
  }
}

// Transpose of the interior central difference: phibar[q] = sum_a (Gbar_a[q + e_a] - Gbar_a[q - e_a]) / 2H,
// each term present when that neighbour is an interior point; zero on the faces.
template <typename G>
__global__ void k_grad_transpose(const G *__restrict__ gx, const G *__restrict__ gy, const G *__restrict__ gz,
                                 G *__restrict__ out, Cube cb, double inv_2h)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= (unsigned)n1 * n1 * n1)
    return;
  int i, j, k;
  cube_ijk(cb, t, i, j, k);
  if (i == 0 || i == n || j == 0 || j == n || k == 0 || k == n)
  {
    out[t] = (G)0;
    return;
  }
  const size_t s1 = n1, s2 = (size_t)n1 * n1;
  double v = 0.0;
  if (i + 1 < n)
    v += (double)gx[t + 1];
  if (i - 1 > 0)
    v -= (double)gx[t - 1];
  if (j + 1 < n)
    v += (double)gy[t + s1];
  if (j - 1 > 0)
    v -= (double)gy[t - s1];
  if (k + 1 < n)
    v += (double)gz[t + s2];
  if (k - 1 > 0)
    v -= (double)gz[t - s2];
  out[t] = (G)(v * inv_2h);
}

// xbar_i += sum_g grad W(x_i, g) (coeff m_i psi_g + abar_i . G_g); m, abar and xbar in the
// slot order of the positions.
template <typename P, typename G>
__global__ void k_adjoint_gather(const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, const P *__restrict__ m, const double *__restrict__ psi, const G *__restrict__ gax, const G *__restrict__ gay,
                                 const G *__restrict__ gaz, const double *__restrict__ abx,
                                 const double *__restrict__ aby, const double *__restrict__ abz,
                                 double *__restrict__ xbx, double *__restrict__ xby, double *__restrict__ xbz, int N,
                                 int n, double coeff)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N)
    return;
  const int n1 = n + 1;
  const double s = (double)n;
  double pxs = x[p] * s, pys = y[p] * s, pzs = z[p] * s;
  int fx = min((int)pxs, n - 1), fy = min((int)pys, n - 1), fz = min((int)pzs, n - 1);
  double tx = pxs - fx, ty = pys - fy, tz = pzs - fz;
  double wx[2] = {1.0 - tx, tx}, wy[2] = {1.0 - ty, ty}, wz[2] = {1.0 - tz, tz};
  double dw[2] = {-s, s};
  const int q = p;
  double ab[3] = {abx[q], aby[q], abz[q]};
  double cm = coeff * (double)m[q];
  double gxs = 0, gys = 0, gzs = 0;
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++)
      {
        size_t g = pidx(fx + ii, fy + jj, fz + kk, n1);
        double val = cm * psi[g] + ab[0] * (double)gax[g] + ab[1] * (double)gay[g] + ab[2] * (double)gaz[g];
        gxs += dw[ii] * wy[jj] * wz[kk] * val;
        gys += wx[ii] * dw[jj] * wz[kk] * val;
        gzs += wx[ii] * wy[jj] * dw[kk] * val;
      }
  xbx[q] += gxs;
  xby[q] += gys;
  xbz[q] += gzs;
}

// abar_j = h vbar_j + abarA (the two uses of a_j: the closing half kick of step
// j and the opening half kick of step j+1).
__global__ void k_adj_abar(const double *__restrict__ vb, const double *__restrict__ abA, double *__restrict__ ab,
                           int n3, double h)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3)
    ab[i] = h * vb[i] + abA[i];
}

// Reverse of the kick-drift of step j (particle_updating.c:92-147 and
// 2677-2691): with s = -1 on reflected axes,
//   ybar = s xbar, ubar = s vbar + dt ybar; xbar <- ybar, vbar <- ubar, abarA <- h ubar.
__global__ void k_adj_kick_drift(double *__restrict__ xb, double *__restrict__ vb, double *__restrict__ abA,
                                 const unsigned char *__restrict__ refl, int N, double dt)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N)
    return;
  unsigned char f = refl ? refl[i] : 0;
  for (int a = 0; a < 3; a++)
  {
    size_t k = (size_t)a * N + i;
    double sg = ((f >> a) & 1) ? -1.0 : 1.0;
    double yb = sg * xb[k];
    double ub = sg * vb[k] + dt * yb;
    xb[k] = yb;
    vb[k] = ub;
    abA[k] = 0.5 * dt * ub;
  }
}

// ---------------------------------------------------------------- with refinement levels
// Forward map per node, in level order (the tree held fixed: cell membership is
// piecewise constant in the positions): rho from the particles of its cells,
// boundary phi and a from the 8-point parent average, interior solve, interior
// central differences, gather from the finest node.  Its VJP, finest level first:
//   Abar(node)   = scatter of abar from the particles of its unrefined cells,
//                  plus the children's boundary Abar through the transposed average
//   phibar(node) = transposed interior central difference of Abar, plus the
//                  children's boundary phibar through the transposed average
//   L psi = phibar on the interior points, zero elsewhere (rhobar = psi)
//   boundary phibar -= (1/H^2) sum of psi over its interior neighbours
//   xbar_i += sum over the nodes holding i of c_l m_i grad W_l psi_l
//             + grad W_f (abar_i . a_f) at its finest node f.

enum : unsigned char
{
  kPtNone = 0,
  kPtSim = 1,
  kPtBdry = 2,
  kPtIntr = 3
};

// Point kinds of the level's nodes in the grid-pool layout (pool zero-filled first).
__global__ void k_adj_kinds(const DNode *__restrict__ nodes, int first, const int *__restrict__ sim_g,
                            const int *__restrict__ bdry_g, const int *__restrict__ intr_g,
                            unsigned char *__restrict__ kind)
{
  const DNode &d = nodes[first + blockIdx.y];
  const int tot = d.nsim + d.nbdry + d.nintr;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x)
  {
    if (t < d.nsim)
      kind[d.goff + sim_g[d.poff + t]] = kPtSim;
    else if (t < d.nsim + d.nbdry)
      kind[d.goff + bdry_g[d.poff + t - d.nsim]] = kPtBdry;
    else
      kind[d.goff + intr_g[d.poff + t - d.nsim - d.nbdry]] = kPtIntr;
  }
}

// Boundary (t < nbdry) or interior point t of node d: slot in the lists, local
// grid index and global coordinates.
__device__ inline void adj_point(const DNode &d, int t, const int *bdry_g, const int *bdry_xyz, const int *intr_g,
                                 const int *intr_xyz, int &gl, int &gx, int &gy, int &gz)
{
  const bool bd = t < d.nbdry;
  const int sl = d.poff + (bd ? t : t - d.nbdry);
  const int *xyz = bd ? bdry_xyz : intr_xyz;
  gl = bd ? bdry_g[sl] : intr_g[sl];
  gx = xyz[3 * sl];
  gy = xyz[3 * sl + 1];
  gz = xyz[3 * sl + 2];
}

// Abar at the boundary and interior points of the level's nodes: the corner
// sums (from k_deposit_runs over abar) of the particles in the point's 8 box
// cells that are not refined (their particles gather from the child).
template <typename S>
__global__ void k_adj_branch_scatter(const DNode *__restrict__ nodes, int first, const int *__restrict__ bdry_g,
                                     const int *__restrict__ bdry_xyz, const int *__restrict__ intr_g,
                                     const int *__restrict__ intr_xyz, const int *__restrict__ box,
                                     const int *__restrict__ bstart, const int *__restrict__ bend,
                                     const S *__restrict__ tails, double *__restrict__ out)
{
  const DNode &d = nodes[first + blockIdx.y];
  const int tot = d.nbdry + d.nintr;
  const int rX = d.rd[0], rXY = d.rd[0] * d.rd[1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x)
  {
    int gl, gx, gy, gz;
    adj_point(d, t, bdry_g, bdry_xyz, intr_g, intr_xyz, gl, gx, gy, gz);
    const long long b0 = d.boff + (gx - 1 - d.ts[0]) + (long long)(gy - 1 - d.ts[1]) * rX +
                         (long long)(gz - 1 - d.ts[2]) * rXY;
    int a[8], e[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
    {
      const long long b = b0 + (u & 1) + ((u >> 1) & 1) * (long long)rX + (u >> 2) * (long long)rXY;
      const int s0 = bstart[b], s1 = bend[b], st = box[b];
      a[u] = st < 0 ? s0 : 0; // refined cells: their particles gather from the child
      e[u] = st < 0 ? s1 : 0;
    }
    out[d.goff + gl] = fold8(tails, a, e);
  }
}

// k_adj_branch_scatter for the three components at once.
template <typename S>
__global__ void k_adj_branch_scatter3(const DNode *__restrict__ nodes, int first, const int *__restrict__ bdry_g,
                                     const int *__restrict__ bdry_xyz, const int *__restrict__ intr_g,
                                     const int *__restrict__ intr_xyz, const int *__restrict__ box,
                                     const int *__restrict__ bstart, const int *__restrict__ bend,
                                     const S *__restrict__ t0, const S *__restrict__ t1, const S *__restrict__ t2,
                                      double *__restrict__ o0, double *__restrict__ o1, double *__restrict__ o2)
{
  const DNode &d = nodes[first + blockIdx.y];
  const int tot = d.nbdry + d.nintr;
  const int rX = d.rd[0], rXY = d.rd[0] * d.rd[1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x)
  {
    int gl, gx, gy, gz;
    adj_point(d, t, bdry_g, bdry_xyz, intr_g, intr_xyz, gl, gx, gy, gz);
    const long long b0 = d.boff + (gx - 1 - d.ts[0]) + (long long)(gy - 1 - d.ts[1]) * rX +
                         (long long)(gz - 1 - d.ts[2]) * rXY;
    int a[8], e[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
    {
      const long long b = b0 + (u & 1) + ((u >> 1) & 1) * (long long)rX + (u >> 2) * (long long)rXY;
      const int s0 = bstart[b], s1 = bend[b], st = box[b];
      a[u] = st < 0 ? s0 : 0; // refined cells: their particles gather from the child
      e[u] = st < 0 ? s1 : 0;
    }
    o0[d.goff + gl] = fold8(t0, a, e);
    o1[d.goff + gl] = fold8(t1, a, e);
    o2[d.goff + gl] = fold8(t2, a, e);
  }
}

// Head Abar: the head deposit fold with coefficient 1 over the unrefined head cells.
template <typename S, typename G>
__global__ void k_adj_head_scatter(const int *__restrict__ start, const int *__restrict__ head_child,
                                   const S *__restrict__ tails, G *__restrict__ out, Cube cb)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= (unsigned)n1 * n1 * n1)
    return;
  int i, j, k;
  cube_ijk(cb, t, i, j, k);
  int a[8], e[8];
#pragma unroll
  for (int u = 0; u < 8; u++)
  {
    const int ci = i - 1 + (u & 1), cj = j - 1 + ((u >> 1) & 1), ck = k - 1 + (u >> 2);
    const bool in = ci >= 0 && ci < n && cj >= 0 && cj < n && ck >= 0 && ck < n;
    const int c = in ? ci + n * (cj + n * ck) : 0;
    const int s0 = start[c], s1 = start[c + 1], hc = head_child[c];
    const bool ok = in && hc < 0;
    a[u] = ok ? s0 : 0;
    e[u] = ok ? s1 : 0;
  }
  out[t] = (G)fold8(tails, a, e);
}


// k_adj_head_scatter for the three components at once (head_child null: every
// head cell counts, the head grid without levels).
template <typename S, typename G>
__global__ void k_adj_head_scatter3(const int *__restrict__ start, const int *__restrict__ head_child,
                                    const S *__restrict__ t0, const S *__restrict__ t1, const S *__restrict__ t2,
                                    G *__restrict__ o0, G *__restrict__ o1, G *__restrict__ o2, Cube cb)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= (unsigned)n1 * n1 * n1)
    return;
  int i, j, k;
  cube_ijk(cb, t, i, j, k);
  int a[8], e[8];
#pragma unroll
  for (int u = 0; u < 8; u++)
  {
    const int ci = i - 1 + (u & 1), cj = j - 1 + ((u >> 1) & 1), ck = k - 1 + (u >> 2);
    const bool in = ci >= 0 && ci < n && cj >= 0 && cj < n && ck >= 0 && ck < n;
    const int c = in ? ci + n * (cj + n * ck) : 0;
    const int s0 = start[c], s1 = start[c + 1];
    const bool ok = in && (!head_child || head_child[c] < 0);
    a[u] = ok ? s0 : 0;
    e[u] = ok ? s1 : 0;
  }
  o0[t] = (G)fold8(t0, a, e);
  o1[t] = (G)fold8(t1, a, e);
  o2[t] = (G)fold8(t2, a, e);
}

// phibar += transposed interior central difference of Abar at the boundary and
// interior points of the level's nodes (grid_acceleration.c:36-76 computes
// a(g) = (phi(g - e) - phi(g + e)) / 2H at interior points g only).
__global__ void k_adj_node_grad_t(const DNode *__restrict__ nodes, int first, const int *__restrict__ bdry_g,
                                  const int *__restrict__ bdry_xyz, const int *__restrict__ intr_g,
                                  const int *__restrict__ intr_xyz, const unsigned char *__restrict__ kind,
                                  const double *__restrict__ abx, const double *__restrict__ aby,
                                  const double *__restrict__ abz, double *__restrict__ pb)
{
  const DNode &d = nodes[first + blockIdx.y];
  const int tot = d.nbdry + d.nintr;
  const long long X = d.rd[0] + 1, XY = X * (d.rd[1] + 1);
  const double o2H = 0.5 * (double)(1 << d.lv);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x)
  {
    int gl, gx, gy, gz;
    adj_point(d, t, bdry_g, bdry_xyz, intr_g, intr_xyz, gl, gx, gy, gz);
    const int lx = gx - d.ts[0], ly = gy - d.ts[1], lz = gz - d.ts[2];
    const long long g = d.goff + gl;
    double v = 0.0;
    if (lx < d.rd[0] && kind[g + 1] == kPtIntr)
      v += abx[g + 1];
    if (lx > 0 && kind[g - 1] == kPtIntr)
      v -= abx[g - 1];
    if (ly < d.rd[1] && kind[g + X] == kPtIntr)
      v += aby[g + X];
    if (ly > 0 && kind[g - X] == kPtIntr)
      v -= aby[g - X];
    if (lz < d.rd[2] && kind[g + XY] == kPtIntr)
      v += abz[g + XY];
    if (lz > 0 && kind[g - XY] == kPtIntr)
      v -= abz[g - XY];
    pb[g] += o2H * v;
  }
}

// Right-hand side of the adjoint node solve: rhs = phibar on the interior points.
template <typename G>
__global__ void k_adj_node_rhs(const DNode *__restrict__ nodes, int first, const int *__restrict__ intr_g,
                               const double *__restrict__ pb, G *__restrict__ rhs)
{
  const DNode &d = nodes[first + blockIdx.y];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nintr; t += gridDim.x * blockDim.x)
  {
    const long long g = d.goff + intr_g[d.poff + t];
    rhs[g] = (G)pb[g];
  }
}

// Boundary phibar -= (1/H^2) sum of psi over the interior neighbours (the
// boundary values enter the interior equations through the Laplacian).
__global__ void k_adj_node_bdry(const DNode *__restrict__ nodes, int first, const int *__restrict__ bdry_g,
                                const int *__restrict__ bdry_xyz, const unsigned char *__restrict__ kind,
                                const double *__restrict__ psi, double *__restrict__ pb)
{
  const DNode &d = nodes[first + blockIdx.y];
  const long long X = d.rd[0] + 1, XY = X * (d.rd[1] + 1);
  const double s = (double)(1 << d.lv);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nbdry; t += gridDim.x * blockDim.x)
  {
    const int sl = d.poff + t;
    const long long g = d.goff + bdry_g[sl];
    const int lx = bdry_xyz[3 * sl] - d.ts[0], ly = bdry_xyz[3 * sl + 1] - d.ts[1], lz = bdry_xyz[3 * sl + 2] - d.ts[2];
    double v = 0.0;
    if (lx < d.rd[0] && kind[g + 1] == kPtIntr)
      v += psi[g + 1];
    if (lx > 0 && kind[g - 1] == kPtIntr)
      v += psi[g - 1];
    if (ly < d.rd[1] && kind[g + X] == kPtIntr)
      v += psi[g + X];
    if (ly > 0 && kind[g - X] == kPtIntr)
      v += psi[g - X];
    if (lz < d.rd[2] && kind[g + XY] == kPtIntr)
      v += psi[g + XY];
    if (lz > 0 && kind[g - XY] == kPtIntr)
      v += psi[g - XY];
    pb[g] -= s * s * v;
  }
}

// Transfer records of the transposed parent average: boundary point t of node
// q (record r = bpre[q] + t) sends 1/8 of its Abar and phibar to each of the 8
// parent points ((c >> 1) or ((c + 1) >> 1) per axis).  key = target index
// (head: packed cube; else the parent's grid-pool index), val = 8 r + k, src[r]
// = the boundary point's pool index.  Sorting the keys (stably) orders the
// contributions to each target by record, so the sums are deterministic.
__global__ void k_adj_xfer_records(const DNode *__restrict__ nodes, int first, const int *__restrict__ bpre,
                                   const int *__restrict__ bdry_g, const int *__restrict__ bdry_xyz, int n,
                                   unsigned *__restrict__ key, int *__restrict__ val, long long *__restrict__ src)
{
  const DNode &d = nodes[first + blockIdx.y];
  long long X, XY, base;
  int tx, ty, tz;
  if (d.parent < 0)
  {
    X = n + 1;
    XY = X * X;
    base = 0;
    tx = ty = tz = 0;
  }
  else
  {
    const DNode &p = nodes[d.parent];
    X = p.rd[0] + 1;
    XY = X * (p.rd[1] + 1);
    base = p.goff;
    tx = p.ts[0];
    ty = p.ts[1];
    tz = p.ts[2];
  }
  const int r0 = bpre[blockIdx.y];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nbdry; t += gridDim.x * blockDim.x)
  {
    const int sl = d.poff + t, r = r0 + t;
    src[r] = d.goff + bdry_g[sl];
    const int cx = bdry_xyz[3 * sl], cy = bdry_xyz[3 * sl + 1], cz = bdry_xyz[3 * sl + 2];
    const long long ax[2] = {(cx >> 1) - tx, ((cx + 1) >> 1) - tx};
    const long long ay[2] = {(cy >> 1) - ty, ((cy + 1) >> 1) - ty};
    const long long az[2] = {(cz >> 1) - tz, ((cz + 1) >> 1) - tz};
#pragma unroll
    for (int k = 0; k < 8; k++)
    {
      key[8 * (long long)r + k] = (unsigned)(base + ax[k & 1] + ay[(k >> 1) & 1] * X + az[k >> 2] * XY);
      val[8 * (long long)r + k] = 8 * r + k;
    }
  }
}

// Sum of each run of equal (sorted) keys, in record order, added to the target:
// tab_a[key] += sum / 8, tpb[key] += sum / 8.  One thread per run head.
template <typename T>
__global__ void k_adj_xfer_apply(const unsigned *__restrict__ skey, const int *__restrict__ sval, int R,
                                 const long long *__restrict__ src, const double *__restrict__ abx,
                                 const double *__restrict__ aby, const double *__restrict__ abz,
                                 const double *__restrict__ pb, T *__restrict__ tax, T *__restrict__ tay,
                                 T *__restrict__ taz, double *__restrict__ tpb)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R)
    return;
  const unsigned k = skey[i];
  if (i > 0 && skey[i - 1] == k)
    return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int j = i; j < R && skey[j] == k; j++)
  {
    const long long s = src[sval[j] >> 3];
    s0 += abx[s];
    s1 += aby[s];
    s2 += abz[s];
    s3 += pb[s];
  }
  tax[k] = (T)((double)tax[k] + 0.125 * s0);
  tay[k] = (T)((double)tay[k] + 0.125 * s1);
  taz[k] = (T)((double)taz[k] + 0.125 * s2);
  tpb[k] += 0.125 * s3;
}

// Head right-hand side: the children's phibar added on the interior points.
template <typename G>
__global__ void k_adj_head_add(G *__restrict__ rhs, const double *__restrict__ pb, Cube cb)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (unsigned)n1 * n1 * n1)
    return;
  int i, j, k;
  cube_ijk(cb, t, i, j, k);
  if (i == 0 || i == n || j == 0 || j == n || k == 0 || k == n)
    return;
  rhs[t] = (G)((double)rhs[t] + pb[t]);
}

// One level's CIC weights and derivatives at position (px, py, pz) scaled to
// that level: the cell corner (fx, fy, fz), weights w and d w / d x per axis.
struct CicGrad
{
  int fx, fy, fz;
  double wx[2], wy[2], wz[2], ds;
};

__device__ inline CicGrad cic_grad(double x, double y, double z, int lv, int clamp_top)
{
  CicGrad c;
  const double s = (double)(1 << lv);
  const double px = x * s, py = y * s, pz = z * s;
  c.fx = min((int)px, clamp_top);
  c.fy = min((int)py, clamp_top);
  c.fz = min((int)pz, clamp_top);
  const double tx = px - c.fx, ty = py - c.fy, tz = pz - c.fz;
  c.wx[0] = 1.0 - tx;
  c.wx[1] = tx;
  c.wy[0] = 1.0 - ty;
  c.wy[1] = ty;
  c.wz[0] = 1.0 - tz;
  c.wz[1] = tz;
  c.ds = s;
  return c;
}

// xbar_i += sum over the nodes holding i (head included) of grad W (c_l m psi)
// plus grad W (abar . a) at the finest one; m, abar and xbar in the slot order of the positions.
template <typename P, typename G>
__global__ void k_adjoint_gather_amr(const double *__restrict__ x, const double *__restrict__ y,
                                     const double *__restrict__ z, const P *__restrict__ m,
                                     const double *__restrict__ hpsi, const G *__restrict__ hax,
                                     const G *__restrict__ hay, const G *__restrict__ haz,
                                     const int *__restrict__ head_child, const DNode *__restrict__ nodes,
                                     const int *__restrict__ box, const double *__restrict__ npsi,
                                     const G *__restrict__ bax, const G *__restrict__ bay, const G *__restrict__ baz,
                                     const double *__restrict__ abx, const double *__restrict__ aby,
                                     const double *__restrict__ abz, double *__restrict__ xbx,
                                     double *__restrict__ xby, double *__restrict__ xbz, int N, int lmin, double G4pi)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N)
    return;
  const int n = 1 << lmin;
  const double px = x[p], py = y[p], pz = z[p];
  const int q = p;
  const double ab[3] = {abx[q], aby[q], abz[q]};
  const double mm = (double)m[q];
  double gxs = 0.0, gys = 0.0, gzs = 0.0;
  // head
  CicGrad c = cic_grad(px, py, pz, lmin, n - 1);
  int nq = head_child[c.fx + n * (c.fy + n * c.fz)];
  {
    const double cm = G4pi * c.ds * c.ds * c.ds * mm;
    const bool fin = nq < 0;
    const long long X = n + 1, XY = X * X;
    const long long g0 = c.fx + c.fy * X + c.fz * XY;
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
      for (int jj = 0; jj < 2; jj++)
#pragma unroll
        for (int ii = 0; ii < 2; ii++)
        {
          const long long g = g0 + ii + jj * X + kk * XY;
          double val = cm * hpsi[g];
          if (fin)
            val += ab[0] * (double)hax[g] + ab[1] * (double)hay[g] + ab[2] * (double)haz[g];
          const double dx = ii ? c.ds : -c.ds, dy = jj ? c.ds : -c.ds, dz = kk ? c.ds : -c.ds;
          gxs += dx * c.wy[jj] * c.wz[kk] * val;
          gys += c.wx[ii] * dy * c.wz[kk] * val;
          gzs += c.wx[ii] * c.wy[jj] * dz * val;
        }
  }
  while (nq >= 0)
  {
    const DNode &d = nodes[nq];
    c = cic_grad(px, py, pz, d.lv, (1 << d.lv) - 1);
    const int b = (c.fx - d.ts[0]) + (c.fy - d.ts[1]) * d.rd[0] + (c.fz - d.ts[2]) * d.rd[0] * d.rd[1];
    const int st = box[d.boff + b];
    const bool fin = st < 0;
    const double cm = G4pi * c.ds * c.ds * c.ds * mm;
    const long long X = d.rd[0] + 1, XY = X * (d.rd[1] + 1);
    const long long g0 = d.goff + (c.fx - d.ts[0]) + (c.fy - d.ts[1]) * X + (c.fz - d.ts[2]) * XY;
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
      for (int jj = 0; jj < 2; jj++)
#pragma unroll
        for (int ii = 0; ii < 2; ii++)
        {
          const long long g = g0 + ii + jj * X + kk * XY;
          double val = cm * npsi[g];
          if (fin)
            val += ab[0] * (double)bax[g] + ab[1] * (double)bay[g] + ab[2] * (double)baz[g];
          const double dx = ii ? c.ds : -c.ds, dy = jj ? c.ds : -c.ds, dz = kk ? c.ds : -c.ds;
          gxs += dx * c.wy[jj] * c.wz[kk] * val;
          gys += c.wx[ii] * dy * c.wz[kk] * val;
          gzs += c.wx[ii] * c.wy[jj] * dz * val;
        }
    if (fin)
      break;
    nq = st;
  }
  xbx[q] += gxs;
  xby[q] += gys;
  xbz[q] += gzs;
}

} // namespace qz
