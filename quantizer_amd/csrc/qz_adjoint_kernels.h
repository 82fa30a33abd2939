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

#include "qz_kernels.h"

namespace qz
{

// dst[p] = src[id[p]] (ID order -> sorted slot order)
template <typename P>
__global__ void k_by_id_to_slot(const double *__restrict__ src, const int *__restrict__ id, P *__restrict__ dst,
                                int N)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N)
    dst[p] = (P)src[id[p]];
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

// xbar_i += sum_g grad W(x_i, g) (coeff m_i psi_g + abar_i . G_g), written in ID order.
template <typename P, typename G>
__global__ void k_adjoint_gather(const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, const P *__restrict__ m, const int *__restrict__ id,
                                 const double *__restrict__ psi, const G *__restrict__ gax, const G *__restrict__ gay,
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
  int q = id[p];
  double ab[3] = {abx[q], aby[q], abz[q]};
  double cm = coeff * (double)m[p];
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

} // namespace qz
