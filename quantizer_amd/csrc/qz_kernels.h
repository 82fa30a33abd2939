// Device kernels of the PM gravity step for gfx950 (CDNA4, 64-wide wavefronts).
//
// Every kernel here is memory-bound integer/stencil/scatter-gather work; none
// is a dense contraction, so nothing is mapped to MFMA.  Layout conventions:
//   * particles: structure of arrays, permanently kept sorted by the octree
//     key (head cell, then child-octant path), so every cell at every level is
//     a contiguous slot range;
//   * grids: (n+1)^3 points of a cube with n cells per side, x fastest.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace qz
{

constexpr int kBlock = 256;

__host__ __device__ inline size_t pidx(int i, int j, int k, int n1) { return (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n1 * k); }

// Colour-split layout of the multigrid level arrays.  Every level has an odd
// number of points per side (2^k + 1), so the parity of the linear index
// i + n1 (j + n1 k) equals the red/black colour (i + j + k) mod 2: red points
// (even index) are stored first, black points after them.  All six stencil
// neighbours of a point have the other colour.
__host__ __device__ inline size_t rbpos(size_t idx, size_t half) { return (idx & 1) * half + (idx >> 1); }
// Division by a fixed divisor d >= 1 for dividends x < 2^31 without an integer
// divide: q = (x m) >> (31 + s) with s = ceil(log2 d), m = ceil(2^(31+s) / d).
// Exact: m d - 2^(31+s) = e < d <= 2^s, so x m / 2^(31+s) exceeds x / d by
// less than x e / (d 2^(31+s)) < 1/d, which never reaches the next integer.
struct FDiv
{
  unsigned m;
  int s;
};
inline FDiv make_fdiv(unsigned d)
{
  int s = 0;
  while ((1ull << s) < d)
    s++;
  const unsigned long long p = 1ull << (31 + s);
  return FDiv{(unsigned)((p + d - 1) / d), s};
}
__host__ __device__ inline unsigned fdiv(unsigned x, FDiv f)
{
  return (unsigned)(((unsigned long long)x * f.m) >> (31 + f.s));
}

// Workgroups of a 1-D grid go to the 8 XCDs round-robin by id.  Renumbered
// this way each XCD takes one contiguous eighth of the range, so the +-x, y, z
// neighbours a stencil reads were mostly fetched by the same XCD's L2.
__device__ inline unsigned xcd_block(unsigned b, unsigned nb)
{
  const unsigned per = nb >> 3;
  return b < (per << 3) ? (b & 7u) * per + (b >> 3) : b;
}

// A cube of n1^3 grid points, x fastest, with the divisors its linear index needs.
struct Cube
{
  int n1;
  FDiv d1, d2; // by n1 and by n1^2
};
inline Cube make_cube(int n1) { return Cube{n1, make_fdiv((unsigned)n1), make_fdiv((unsigned)n1 * (unsigned)n1)}; }
__host__ __device__ inline void cube_ijk(const Cube &c, unsigned t, int &i, int &j, int &k)
{
  const unsigned q = fdiv(t, c.d1);
  i = (int)(t - q * (unsigned)c.n1);
  k = (int)fdiv(t, c.d2);
  j = (int)(q - (unsigned)k * (unsigned)c.n1);
}

__host__ __device__ inline size_t rbhalf(size_t npts) { return (npts + 1) >> 1; }

// ---------------------------------------------------------------- generic
template <typename T> __global__ void k_fill(T *__restrict__ a, T v, size_t n)
{
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n)
    a[t] = v;
}

// Deterministic block reduction of doubles: sum (op=0) or max (op=1).
template <int OP> __device__ inline double block_reduce(double v)
{
  __shared__ double sh[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1)
  {
    double o = __shfl_down(v, off, 64);
    v = OP == 0 ? v + o : fmax(v, o);
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    double r = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++)
      r = OP == 0 ? r + sh[w] : fmax(r, sh[w]);
    sh[0] = r;
  }
  __syncthreads();
  double r = sh[0];
  __syncthreads();
  return r;
}

// The same reduction for workgroups of up to 1024 threads.
template <int OP> __device__ inline double block_reduce_n(double v)
{
  __shared__ double sh[16];
  for (int off = 32; off > 0; off >>= 1)
  {
    double o = __shfl_down(v, off, 64);
    v = OP == 0 ? v + o : fmax(v, o);
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    double r = sh[0];
    for (int w = 1; w < (int)((blockDim.x + 63) >> 6); w++)
      r = OP == 0 ? r + sh[w] : fmax(r, sh[w]);
    sh[0] = r;
  }
  __syncthreads();
  double r = sh[0];
  __syncthreads();
  return r;
}

// Second stage: one block folds the per-block partials in a fixed order.
template <int OP> __global__ void k_reduce_final(const double *__restrict__ part, int np, double *__restrict__ out)
{
  double v = OP == 0 ? 0.0 : -1.0e308;
  // unrolled so that eight loads are in flight; the values are still folded
  // one after another in index order
#pragma unroll 8
  for (int i = threadIdx.x; i < np; i += blockDim.x)
    v = OP == 0 ? v + part[i] : fmax(v, part[i]);
  v = block_reduce<OP>(v);
  if (threadIdx.x == 0)
    *out = v;
}

// First stage of the max fold of many per-block minima: block b folds -part
// over [b chunk, (b + 1) chunk) into out[b] (one load per thread per chunk of
// kBlock, instead of a single block walking every partial).
__global__ void k_reduce_neg_max_stage(const double *__restrict__ part, int np, int chunk, double *__restrict__ out)
{
  double v = -1.0e308;
  const int i1 = min(np, (int)blockIdx.x * chunk + chunk);
#pragma unroll 4
  for (int i = blockIdx.x * chunk + threadIdx.x; i < i1; i += blockDim.x)
    v = fmax(v, -part[i]);
  v = block_reduce<1>(v);
  if (threadIdx.x == 0)
    out[blockIdx.x] = v;
}

// ---------------------------------------------------------------- particles
// Octree key: finest coordinate F = floor(x 2^(lmin+D)); head cell from F >> D,
// then D octant digits (x + 2y + 4z) from the most to the least significant
// level.  floor(x 2^l) = floor(x 2^(lmin+D)) >> (lmin+D-l) exactly for x in [0,1),
// so one key reproduces ptcl_idx_to_box_idx (ptcl_idx_to_box_idx.c:62-69) at
// every level.
template <typename P>
__device__ inline unsigned long long octree_key(P x, P y, P z, int lmin, int D)
{
  const P s = (P)(1 << (lmin + D));
  const int top = (1 << (lmin + D)) - 1;
  // a reflected particle can land on exactly 1.0; it stays in the last cell
  int fx = min(max((int)(x * s), 0), top), fy = min(max((int)(y * s), 0), top), fz = min(max((int)(z * s), 0), top);
  int n = 1 << lmin;
  unsigned long long hc = (unsigned long long)(fx >> D) +
                          (unsigned long long)n * ((unsigned long long)(fy >> D) + (unsigned long long)n * (fz >> D));
  unsigned long long k = hc;
  for (int l = D - 1; l >= 0; l--)
    k = (k << 3) | (unsigned long long)(((fx >> l) & 1) | (((fy >> l) & 1) << 1) | (((fz >> l) & 1) << 2));
  return k;
}

template <typename P>
__global__ void k_keys(const P *__restrict__ x, const P *__restrict__ y, const P *__restrict__ z,
                       unsigned long long *__restrict__ key, int *__restrict__ slot, int N, int lmin, int D)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N)
    return;
  key[i] = octree_key<P>(x[i], y[i], z[i], lmin, D);
  slot[i] = i;
}

// The same keys written as 32 bits (keys of at most 32 significant bits), the
// sort's input form.
template <typename P>
__global__ void k_keys32(const P *__restrict__ x, const P *__restrict__ y, const P *__restrict__ z,
                         unsigned *__restrict__ key, int *__restrict__ slot, int N, int lmin, int D)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N)
    return;
  key[i] = (unsigned)octree_key<P>(x[i], y[i], z[i], lmin, D);
  slot[i] = i;
}

__global__ void k_narrow_keys(const unsigned long long *__restrict__ in, unsigned *__restrict__ out, int N)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N)
    out[i] = (unsigned)in[i];
}

__global__ void k_widen_keys(const unsigned *__restrict__ in, unsigned long long *__restrict__ out, int N)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N)
    out[i] = in[i];
}

// Gather every per-particle array into the sorted order (out[i] = in[perm[i]]).
// Positions are FP64 in every precision mode: the CIC weights of a level-lv
// cell carry the position error times 2^lv, so FP32 positions would cost
// 2^lv * 6e-8 relative density error at the refined levels.
template <typename P> struct PtclArrays
{
  double *x[3];
  P *a[7]; // vx vy vz ax ay az m
  int *id;
};

template <typename P>
__global__ void k_permute(PtclArrays<P> src, PtclArrays<P> dst, const int *__restrict__ perm, int N, int move_acc)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N)
    return;
  int s = perm[i];
  // all loads first, then all stores: the arrays are separate buffers, and
  // interleaved (the compiler cannot rule out aliasing) every copy waited for
  // its own load
  double xv[3];
  P av[7];
#pragma unroll
  for (int a = 0; a < 3; a++)
    xv[a] = src.x[a][s];
#pragma unroll
  for (int a = 0; a < 7; a++)
    if (move_acc || a < 3 || a == 6) // accelerations that are all zero stay where they are
      av[a] = src.a[a][s];
  const int idv = src.id[s];
#pragma unroll
  for (int a = 0; a < 3; a++)
    dst.x[a][i] = xv[a];
#pragma unroll
  for (int a = 0; a < 7; a++)
    if (move_acc || a < 3 || a == 6)
      dst.a[a][i] = av[a];
  dst.id[i] = idv;
}

// start[c] = first sorted slot whose head cell is >= c, for c in [0, ncell].
// Slot i owns the cells after the previous slot's cell up to its own; it writes
// at most kCellFill of them (the last ones), so clustered particles (long runs
// of empty cells) cost no long serial loops.  The rest keep the -1 of the fill
// before and are found by k_cell_start_gaps.
constexpr int kCellFill = 64;

__global__ void k_cell_start(const unsigned long long *__restrict__ key, int N, int D, int ncell,
                             int *__restrict__ start)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > N)
    return;
  long long prev = i == 0 ? -1 : (long long)(key[i - 1] >> (3 * D));
  long long cur = i == N ? (long long)ncell : (long long)(key[i] >> (3 * D));
  for (long long c = max(prev + 1, cur - (kCellFill - 1)); c <= cur; c++)
    start[c] = i;
}

// The cells k_cell_start left at -1: binary search for the first slot whose
// head cell is >= c.
__global__ void k_cell_start_gaps(const unsigned long long *__restrict__ key, int N, int D, int ncell,
                                  int *__restrict__ start)
{
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncell || start[c] >= 0)
    return;
  int lo = 0, hi = N;
  while (lo < hi)
  {
    const int mid = (lo + hi) >> 1;
    if ((long long)(key[mid] >> (3 * D)) < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  start[c] = lo;
}

// A2 deposit, stage 1 (grid_density.c:35-267): every particle computes its 8
// CIC corner weights times its mass at level lv; one wavefront (64 consecutive
// sorted slots) then runs a segmented inclusive scan over runs of equal cell
// (key >> seg_shift).  The last lane of each run stores the run's 8 corner sums
// at its slot.  A cell's total is the ordered sum of its runs' tails, one per
// 64-slot window it overlaps.  Deterministic: fixed shuffle order, no atomics.
//
// The scan moves values between lanes with DPP (data-parallel primitives: a
// lane reads another lane's register as an instruction operand, no LDS
// traffic): row_shr 1, 2, 4, 8 scan each row of 16 lanes, then row_bcast 15
// adds lane 15's sum to the next row (rows 1 and 3) and row_bcast 31 adds lane
// 31's sum to rows 2 and 3.  Each step adds the source's sum only where the
// source lane holds the same cell: the keys are sorted, so the lanes between
// them hold it too and every lane ends with its run's sum up to itself.
template <int CTRL, int ROWS> __device__ inline unsigned dpp_u32(unsigned old, unsigned v)
{
  return (unsigned)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWS, 0xf, false);
}

// one step of the segmented scan over K values per lane (lanes without a
// source read the key ~0, which no live cell has, and the value 0)
template <int CTRL, int ROWS, int K>
__device__ inline void seg_scan_step(double (&v)[K], unsigned klo, unsigned khi)
{
  const unsigned slo = dpp_u32<CTRL, ROWS>(~0u, klo), shi = dpp_u32<CTRL, ROWS>(~0u, khi);
  const bool same = slo == klo && shi == khi;
#pragma unroll
  for (int q = 0; q < K; q++)
  {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v[q]);
    const unsigned lo = dpp_u32<CTRL, ROWS>(0u, (unsigned)b), hi = dpp_u32<CTRL, ROWS>(0u, (unsigned)(b >> 32));
    const double o = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    if (same)
      v[q] += o;
  }
}

// inclusive segmented scan of v over the wavefront, segments = equal keys c
// (non-decreasing along the lanes); every lane of the wavefront takes part
template <int K> __device__ inline void wave_seg_scan(double (&v)[K], unsigned long long c)
{
  const unsigned klo = (unsigned)c, khi = (unsigned)(c >> 32);
  // a wavefront whose lanes all hold different cells (sparse regions: one
  // particle per cell) has nothing to scan
  {
    const unsigned plo = dpp_u32<0x138, 0xf>(~0u, klo), phi = dpp_u32<0x138, 0xf>(~0u, khi); // wave_shr:1
    if (__ballot(plo == klo && phi == khi) == 0ull)
      return;
  }
  seg_scan_step<0x111, 0xf>(v, klo, khi); // row_shr:1
  seg_scan_step<0x112, 0xf>(v, klo, khi); // row_shr:2
  seg_scan_step<0x114, 0xf>(v, klo, khi); // row_shr:4
  seg_scan_step<0x118, 0xf>(v, klo, khi); // row_shr:8
  seg_scan_step<0x142, 0xa>(v, klo, khi); // row_bcast:15, rows 1 and 3
  seg_scan_step<0x143, 0xc>(v, klo, khi); // row_bcast:31, rows 2 and 3
}
__device__ inline unsigned long long shfl_up_u64(unsigned long long v, int d)
{
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo = __shfl_up(lo, d, 64);
  hi = __shfl_up(hi, d, 64);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ inline unsigned long long shfl_down_u64(unsigned long long v, int d)
{
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo = __shfl_down(lo, d, 64);
  hi = __shfl_down(hi, d, 64);
  return ((unsigned long long)hi << 32) | lo;
}

template <typename P, typename S>
__global__ void k_deposit_runs(const unsigned long long *__restrict__ key, const double *__restrict__ x,
                               const double *__restrict__ y, const double *__restrict__ z, const P *__restrict__ m,
                               int N, int lv, int seg_shift, int clamp_top, S *__restrict__ tails)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = p < N;
  unsigned long long c = live ? (key[p] >> seg_shift) : ~0ull;
  double v[8];
  if (live)
  {
    const double s = (double)(1 << lv);
    double px = x[p] * s, py = y[p] * s, pz = z[p] * s;
    int fx = min((int)px, clamp_top), fy = min((int)py, clamp_top), fz = min((int)pz, clamp_top);
    double wx1 = px - fx, wy1 = py - fy, wz1 = pz - fz;
    double wx2 = 1.0 - wx1, wy2 = 1.0 - wy1, wz2 = 1.0 - wz1;
    double mm = (double)m[p];
    v[0] = mm * (wx2 * wy2 * wz2);
    v[1] = mm * (wx1 * wy2 * wz2);
    v[2] = mm * (wx2 * wy1 * wz2);
    v[3] = mm * (wx1 * wy1 * wz2);
    v[4] = mm * (wx2 * wy2 * wz1);
    v[5] = mm * (wx1 * wy2 * wz1);
    v[6] = mm * (wx2 * wy1 * wz1);
    v[7] = mm * (wx1 * wy1 * wz1);
  }
  else
  {
#pragma unroll
    for (int q = 0; q < 8; q++)
      v[q] = 0.0;
  }
  wave_seg_scan(v, c);
  unsigned long long cn = shfl_down_u64(c, 1);
  bool tail = live && (lane == 63 || p == N - 1 || cn != c);
  if (tail)
  {
#pragma unroll
    for (int q = 0; q < 8; q++)
      tails[8 * (size_t)p + q] = (S)v[q];
  }
}

// Sum of corner q of the cell whose sorted slot range is [a, e): its runs end at
// min(e, next 64-slot boundary) - 1.
// Cells whose slots span several 64-slot windows hold one run tail per window.
// k_cell_totals / k_cell_totals_heavy replace the tail at such a cell's last
// slot with the sum of all its window tails (windows in order, FP64 sums), so
// every cell's 8 corner sums are then the tail at its last slot: a fold reads
// one tail per cell.  No other run reads that slot.  Cells of up to
// kTotalsSerial windows are summed by one thread; longer ones are listed and
// summed by a whole workgroup (fixed reduction order).  Deterministic.
constexpr int kTotalsSerial = 8;

template <typename S>
__global__ void k_cell_totals(const int *__restrict__ cb, const int *__restrict__ ce, long long ncell,
                              S *__restrict__ tails, int *__restrict__ heavy, int *__restrict__ nheavy)
{
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell)
    return;
  const int a = cb[c], e = ce ? ce[c] : cb[c + 1];
  if (e - a <= 0 || (a >> 6) == ((e - 1) >> 6))
    return;
  if (((e - 1) >> 6) - (a >> 6) + 1 > kTotalsSerial)
  {
    heavy[atomicAdd(nheavy, 1)] = (int)c; // the list holds at most N / (64 (kTotalsSerial - 1)) + 1 cells
    return;
  }
  double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = a; t < e;)
  {
    const int tl = min(e, ((t >> 6) + 1) << 6) - 1;
#pragma unroll
    for (int q = 0; q < 8; q++)
      sum[q] += (double)tails[8 * (size_t)tl + q];
    t = tl + 1;
  }
#pragma unroll
  for (int q = 0; q < 8; q++)
    tails[8 * (size_t)(e - 1) + q] = (S)sum[q];
}

template <typename S>
__global__ void k_cell_totals_heavy(const int *__restrict__ cb, const int *__restrict__ ce,
                                    const int *__restrict__ heavy, const int *__restrict__ nheavy,
                                    S *__restrict__ tails)
{
  const int nh = *nheavy;
  for (int h = blockIdx.x; h < nh; h += gridDim.x)
  {
    const int c = heavy[h];
    const int a = cb[c], e = ce ? ce[c] : cb[c + 1];
    double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int w = (a >> 6) + (int)threadIdx.x; w <= ((e - 1) >> 6); w += blockDim.x)
    {
      const int tl = min(e, (w + 1) << 6) - 1;
#pragma unroll
      for (int q = 0; q < 8; q++)
        sum[q] += (double)tails[8 * (size_t)tl + q];
    }
#pragma unroll
    for (int q = 0; q < 8; q++)
      sum[q] = block_reduce<0>(sum[q]); // ends with a barrier: every window is read before the write
    if (threadIdx.x == 0)
#pragma unroll
      for (int q = 0; q < 8; q++)
        tails[8 * (size_t)(e - 1) + q] = (S)sum[q];
    __syncthreads();
  }
}

// A grid point's fold over its 8 adjacent cells u = (ci - i + 1) + 2 (cj - j + 1)
// + 4 (ck - k + 1), slot ranges [a[u], e[u]) (empty for cells that do not
// count), each contributing its corner 7 - u from the tail at its last slot
// (k_cell_totals has run).  The 8 loads issue together.  Needs N >= 1.
template <typename S> __device__ inline double fold8(const S *__restrict__ tails, const int (&a)[8], const int (&e)[8])
{
  S f[8];
#pragma unroll
  for (int u = 0; u < 8; u++)
    f[u] = tails[8 * (size_t)max(e[u] - 1, 0) + (7 - u)];
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < 8; u++)
    if (a[u] < e[u])
      acc += (double)f[u];
  return acc;
}

// A2 head deposit, stage 2: each grid point adds corner (i-ci, j-cj, k-ck) of
// its 8 adjacent cells; coeff = 4 pi G / H^3 (grid_density.c:62).
template <typename S, typename G>
__global__ void k_deposit_head_fold(const int *__restrict__ start, const S *__restrict__ tails, G *__restrict__ rho,
                                    Cube cb, double coeff)
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
    const bool ok = ci >= 0 && ci < n && cj >= 0 && cj < n && ck >= 0 && ck < n;
    const int c = ok ? ci + n * (cj + n * ck) : 0;
    const int s0 = start[c], s1 = start[c + 1];
    a[u] = ok ? s0 : 0;
    e[u] = ok ? s1 : 0;
  }
  rho[t] = (G)(coeff * fold8(tails, a, e));
}

// A12 head gather (particle_acceleration.c:33-132).
template <typename P, typename G>
__global__ void k_gather_head(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                              const G *__restrict__ gax, const G *__restrict__ gay, const G *__restrict__ gaz,
                              P *__restrict__ ax, P *__restrict__ ay, P *__restrict__ az, int N, int n)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N)
    return;
  const int n1 = n + 1;
  const double s = (double)n;
  double pxs = x[p] * s, pys = y[p] * s, pzs = z[p] * s;
  int fx = min((int)pxs, n - 1), fy = min((int)pys, n - 1), fz = min((int)pzs, n - 1);
  double wx1 = pxs - fx, wy1 = pys - fy, wz1 = pzs - fz;
  double wx2 = 1.0 - wx1, wy2 = 1.0 - wy1, wz2 = 1.0 - wz1;
  double w[8] = {wx2 * wy2 * wz2, wx1 * wy2 * wz2, wx2 * wy1 * wz2, wx1 * wy1 * wz2,
                 wx2 * wy2 * wz1, wx1 * wy2 * wz1, wx2 * wy1 * wz1, wx1 * wy1 * wz1};
  size_t g = pidx(fx, fy, fz, n1);
  double sx = 0, sy = 0, sz = 0;
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++)
      {
        size_t gn = g + ii + (size_t)jj * n1 + (size_t)kk * n1 * n1;
        double wq = w[ii + 2 * jj + 4 * kk];
        sx += (double)gax[gn] * wq;
        sy += (double)gay[gn] * wq;
        sz += (double)gaz[gn] * wq;
      }
  ax[p] = (P)sx;
  ay[p] = (P)sy;
  az[p] = (P)sz;
}

// A14 kick + drift + reflexive boundary (particle_updating.c:92-147).
template <typename P>
__global__ void k_update_A(double *__restrict__ x, double *__restrict__ y, double *__restrict__ z, P *__restrict__ vx,
                           P *__restrict__ vy, P *__restrict__ vz, P *__restrict__ ax, P *__restrict__ ay,
                           P *__restrict__ az, int N, double dt, unsigned char *__restrict__ refl,
                           void *__restrict__ key, int *__restrict__ slot, int lmin, int D, int narrow,
                           unsigned char *__restrict__ moved)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N)
    return;
  const P h = (P)0.5;
  const P dtp = (P)dt;
  unsigned char flags = 0;
  double pos[3] = {x[i], y[i], z[i]};
  // the key before the drift, for the moved flag of the re-sort
  const unsigned long long k_old = moved ? octree_key<double>(pos[0], pos[1], pos[2], lmin, D) : 0ull;
  P vel[3] = {vx[i], vy[i], vz[i]};
  P acc[3] = {ax[i], ay[i], az[i]};
#pragma unroll
  for (int a = 0; a < 3; a++)
  {
    vel[a] += acc[a] * dtp * h;
    pos[a] += (double)vel[a] * dt;
    if (pos[a] < 0.0 || pos[a] >= 1.0)
    {
      pos[a] = pos[a] >= 1.0 ? 2.0 - pos[a] : -pos[a];
      vel[a] = -vel[a];
      acc[a] = -acc[a];
      flags |= (unsigned char)(1 << a);
    }
  }
  if (refl)
    refl[i] = flags; // reflected axes in this slot order, for the adjoint of this step
  x[i] = pos[0];
  y[i] = pos[1];
  z[i] = pos[2];
  vx[i] = vel[0];
  vy[i] = vel[1];
  vz[i] = vel[2];
  // the accelerations change only on a reflected axis
  if (flags & 1)
    ax[i] = acc[0];
  if (flags & 2)
    ay[i] = acc[1];
  if (flags & 4)
    az[i] = acc[2];
  // the octree key of the new position for the next sort (k_keys / k_keys32
  // on the same position), while it is in registers
  if (key)
  {
    const unsigned long long k = octree_key<double>(pos[0], pos[1], pos[2], lmin, D);
    if (narrow)
      ((unsigned *)key)[i] = (unsigned)k;
    else
      ((unsigned long long *)key)[i] = k;
    slot[i] = i;
    if (moved)
      moved[i] = k != k_old ? 1 : 0;
  }
}

// The CFL time step of one axis with velocity and acceleration
// (timestep.c:73-195, 370-587): the dt that moves a particle CFL * H of its
// finest node (H its cell size), folded into dt by fmin.
__device__ inline double cfl_axis(double v, double a, double H, double cfl, double dt)
{
  double av = fabs(v), aa = fabs(a);
  if (aa < 1.0e-16)
  {
    if (av > 1.0e-16)
      dt = fmin(dt, cfl * H / av);
    return dt;
  }
  double c = v * v + 2.0 * aa * H * cfl;
  double t;
  if (1.0 - v * v / c > 1.0e-8)
    t = (-av + sqrt(c)) / aa;
  else
  {
    double eps = fabs(a * H * cfl / (v * v));
    t = fabs(v / a) * (eps - eps * eps * 0.5);
  }
  return fmin(dt, t);
}

// A15 kick (particle_updating.c:2677-2691).  With `part`, the kick also folds
// the next CFL time step of each particle (main.c calls timestep() right after
// this update, on these velocities and accelerations) into per-block minima:
// plev holds each particle's finest level from the last gather (null: the
// head level lmin for every particle).  The time step then needs only the
// second stage of the reduction.
template <typename P>
__global__ void k_update_B(P *__restrict__ vx, P *__restrict__ vy, P *__restrict__ vz, const P *__restrict__ ax,
                           const P *__restrict__ ay, const P *__restrict__ az, int N, P dt,
                           const unsigned char *__restrict__ plev, int lmin, double cfl, double *__restrict__ part)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double dtm = 1e300;
  if (i < N)
  {
    const P h = (P)0.5;
    const P a0 = ax[i], a1 = ay[i], a2 = az[i];
    const P v0 = vx[i] + a0 * dt * h, v1 = vy[i] + a1 * dt * h, v2 = vz[i] + a2 * dt * h;
    vx[i] = v0;
    vy[i] = v1;
    vz[i] = v2;
    if (part)
    {
      const int lv = plev ? (int)plev[i] : lmin;
      const double H = 1.0 / (1 << lv);
      dtm = cfl_axis((double)v0, (double)a0, H, cfl, dtm);
      dtm = cfl_axis((double)v1, (double)a1, H, cfl, dtm);
      dtm = cfl_axis((double)v2, (double)a2, H, cfl, dtm);
    }
  }
  if (part)
  {
    const double r = -block_reduce<1>(-dtm); // the block's minimum
    if (threadIdx.x == 0)
      part[blockIdx.x] = r;
  }
}

// A16 sums: m v^2, m v (observables.c:38-75, 622-659); 4 partials per block.
template <typename P>
__global__ void k_kin_mom(const P *__restrict__ vx, const P *__restrict__ vy, const P *__restrict__ vz,
                          const P *__restrict__ m, int N, double *__restrict__ part)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ek = 0, px = 0, py = 0, pz = 0;
  if (i < N)
  {
    double a = vx[i], b = vy[i], c = vz[i], mm = m[i];
    ek = mm * (a * a + b * b + c * c);
    px = mm * a;
    py = mm * b;
    pz = mm * c;
  }
  ek = block_reduce<0>(ek);
  px = block_reduce<0>(px);
  py = block_reduce<0>(py);
  pz = block_reduce<0>(pz);
  if (threadIdx.x == 0)
  {
    part[blockIdx.x] = ek;
    part[gridDim.x + blockIdx.x] = px;
    part[2 * gridDim.x + blockIdx.x] = py;
    part[3 * gridDim.x + blockIdx.x] = pz;
  }
}

// Exact pairwise potential energy (observables.c:77-112), FP64, tiled through LDS.
template <typename P>
__global__ void k_pairwise_pe(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                              const P *__restrict__ m, int N, double *__restrict__ part)
{
  __shared__ double tx[kBlock], ty[kBlock], tz[kBlock], tm[kBlock];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double xi = 0, yi = 0, zi = 0, mi = 0;
  if (i < N)
  {
    xi = x[i];
    yi = y[i];
    zi = z[i];
    mi = m[i];
  }
  double acc = 0.0;
  // only j < i contributes: tiles past the block's last i are skipped (the
  // bound is the same for every thread of the block)
  const int jend = min(N, (int)(blockIdx.x + 1) * (int)blockDim.x);
  for (int base = 0; base < jend; base += kBlock)
  {
    int j = base + threadIdx.x;
    tx[threadIdx.x] = j < N ? (double)x[j] : 0.0;
    ty[threadIdx.x] = j < N ? (double)y[j] : 0.0;
    tz[threadIdx.x] = j < N ? (double)z[j] : 0.0;
    tm[threadIdx.x] = j < N ? (double)m[j] : 0.0;
    __syncthreads();
    if (i < N)
    {
      int lim = min(kBlock, N - base);
      for (int q = 0; q < lim; q++)
      {
        int jj = base + q;
        if (jj < i)
        {
          double dx = xi - tx[q], dy = yi - ty[q], dz = zi - tz[q];
          acc += mi * tm[q] / sqrt(dx * dx + dy * dy + dz * dz);
        }
      }
    }
    __syncthreads();
  }
  acc = block_reduce<0>(acc);
  if (threadIdx.x == 0)
    part[blockIdx.x] = acc;
}

// Grid-interpolated potential energy 1/2 sum m phi(x) (observables.c:594-619), head only.
template <typename P>
__global__ void k_grid_pe_head(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                               const P *__restrict__ m, const double *__restrict__ phi, int N, int n,
                               double *__restrict__ part)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (p < N)
  {
    const int n1 = n + 1;
    const double s = (double)n;
    double pxs = x[p] * s, pys = y[p] * s, pzs = z[p] * s;
    int fx = min((int)pxs, n - 1), fy = min((int)pys, n - 1), fz = min((int)pzs, n - 1);
    double wx1 = pxs - fx, wy1 = pys - fy, wz1 = pzs - fz;
    double wx[2] = {1.0 - wx1, wx1}, wy[2] = {1.0 - wy1, wy1}, wz[2] = {1.0 - wz1, wz1};
    for (int kk = 0; kk < 2; kk++)
      for (int jj = 0; jj < 2; jj++)
        for (int ii = 0; ii < 2; ii++)
          acc += phi[pidx(fx + ii, fy + jj, fz + kk, n1)] * wx[ii] * wy[jj] * wz[kk];
    acc *= (double)m[p];
  }
  acc = block_reduce<0>(acc);
  if (threadIdx.x == 0)
    part[blockIdx.x] = acc;
}

// ---------------------------------------------------------------- head grid
// The monopole about the box centre at grid point (i, j, k) of an n-cell grid
// (initial_potential_and_acceleration_head.c:297-335; dist = 0 -> 1e-3 H), or
// zeros for zero faces.
template <typename G>
__device__ inline void monopole_point(int i, int j, int k, int n, double GM, int zero_bc, double &phi, G &ax, G &ay,
                                      G &az)
{
  if (zero_bc)
  {
    phi = 0.0;
    ax = ay = az = (G)0;
    return;
  }
  double H = 1.0 / n;
  double ai = i * H - 0.5, aj = j * H - 0.5, ak = k * H - 0.5;
  double dist = sqrt(ai * ai + aj * aj + ak * ak);
  if (dist == 0.0)
    dist = H * 1.0e-3;
  double c = -GM / (dist * dist * dist);
  phi = -GM / dist;
  ax = (G)(c * ai);
  ay = (G)(c * aj);
  az = (G)(c * ak);
}

// A10: monopole -GM/r and its gradient about the box centre on every point of
// the solve domain.
template <typename G>
__global__ void k_init_head(double *__restrict__ phi, G *__restrict__ ax, G *__restrict__ ay, G *__restrict__ az,
                            Cube cb, double GM, int zero_bc)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= (unsigned)n1 * n1 * n1)
    return;
  int i, j, k;
  cube_ijk(cb, t, i, j, k);
  monopole_point(i, j, k, n, GM, zero_bc, phi[t], ax[t], ay[t], az[t]);
}

// Outer FP64 residual r = rho - Lap(phi) on interior points (multigrid.c:140-150,
// poisson_error.c:116-117), stored in the V-cycle type, with per-block sum of r^2.
// Only interior points are written: r stays 0 on the faces from its allocation.
//
// A block is kResTX lanes by kResTY rows of interior columns and marches up
// kResKC interior planes: each lane keeps phi of its column at k - 1, k, k + 1
// in registers (loading k + 3 while plane k is computed), and the in-plane
// neighbours come from one LDS copy of the plane (double-buffered by plane
// parity, one barrier per plane).  The first and last lanes of a row also load
// the x neighbours outside the tile, and the first and last rows the rows below
// and above it, so every phi value is read from memory about once.  A row of
// 256 lanes spans the 255 interior columns of a 257^3 grid; 4 rows by 32
// planes make 512 workgroups of 1024 lanes there, two per CU, and took 62 us
// against 73 us for 64 x 4 x 16 and 66-70 us for the other shapes measured.
// The Laplacian is summed in a fixed order per point.
constexpr int kResTX = 256, kResTY = 4, kResKC = 32;
inline unsigned outer_residual_blocks(int n1)
{
  const unsigned m = (unsigned)(n1 - 2); // interior points per axis
  return ((m + kResTX - 1) / kResTX) * ((m + kResTY - 1) / kResTY) * ((m + kResKC - 1) / kResKC);
}

template <typename G>
__global__ void __launch_bounds__(kResTX * kResTY) k_outer_residual(const double *__restrict__ phi, const G *__restrict__ rho,
                                                                    G *__restrict__ r, int n1, double inv_h2,
                                                                    double *__restrict__ part)
{
  constexpr int TX = kResTX, TY = kResTY, KC = kResKC;
  static_assert(TX % 64 == 0 && TY >= 2, "rows of whole wavefronts, at least two rows");
  __shared__ double S[2][TY + 2][TX + 2];
  const int n = n1 - 1, m = n1 - 2;
  const unsigned npts = (unsigned)n1 * n1 * n1, hf = (npts + 1) >> 1, s1 = n1, s2 = (unsigned)n1 * n1;
  const int gx = (m + TX - 1) / TX, gy = (m + TY - 1) / TY;
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int bx = (int)(lb % (unsigned)gx), by = (int)((lb / (unsigned)gx) % (unsigned)gy), bz = (int)(lb / (unsigned)(gx * gy));
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = 1 + bx * TX + tx, j = 1 + by * TY + ty;
  const int kb = 1 + bz * KC, ke = min(kb + KC, n); // output planes [kb, ke)
  const bool col = i <= n && j <= n; // a grid column: its values are neighbours of the outputs
  const bool out = i < n && j < n;   // an interior column: one output per plane
  // y neighbours outside the tile (row 0: j - 1, last row: j + 1) and x
  // neighbours outside it (lane 0: i - 1, the last lane: i + 1), for output columns
  const bool hcol = out && (ty == 0 || ty == TY - 1);
  const int jh = ty == 0 ? j - 1 : j + 1, hrow = ty == 0 ? 0 : TY + 1;
  const bool xcol = out && (tx == 0 || tx == TX - 1);
  const int ih = tx == 0 ? i - 1 : i + 1, xslot = tx == 0 ? 0 : TX + 1;
  const unsigned ic = (unsigned)min(i, n), jc = (unsigned)min(j, n);
  unsigned t = ic + s1 * jc + s2 * (unsigned)kb;
  unsigned th = ic + s1 * (unsigned)min(jh, n) + s2 * (unsigned)kb;
  unsigned tq = (unsigned)min(ih, n) + s1 * jc + s2 * (unsigned)kb;
  // loads run two planes ahead of the plane being computed (rho one plane)
  const bool two = kb + 1 < ke; // a second plane (then kb + 2 <= n)
  double cm = out ? phi[t - s2] : 0.0;
  double c0 = col ? phi[t] : 0.0;
  double cp = col ? phi[t + s2] : 0.0; // kb + 1 <= n
  double cpp = (col && two) ? phi[t + 2 * s2] : 0.0;
  double h0 = hcol ? phi[th] : 0.0, h1 = (hcol && two) ? phi[th + s2] : 0.0;
  double x0 = xcol ? phi[tq] : 0.0, x1 = (xcol && two) ? phi[tq + s2] : 0.0;
  G r0 = out ? rho[t] : (G)0;
  double sq = 0.0;
  for (int k = kb; k < ke; k++, t += s2, th += s2, tq += s2)
  {
    const int buf = k & 1;
    const bool next = k + 1 < ke, third = k + 2 < ke; // then k + 3 <= n
    const double c3 = (col && third) ? phi[t + 3 * s2] : 0.0;
    const double h2 = (hcol && third) ? phi[th + 2 * s2] : 0.0;
    const double x2 = (xcol && third) ? phi[tq + 2 * s2] : 0.0;
    const G r1 = (out && next) ? rho[t + s2] : (G)0;
    S[buf][ty + 1][tx + 1] = c0;
    if (hcol)
      S[buf][hrow][tx + 1] = h0;
    if (xcol)
      S[buf][ty + 1][xslot] = x0;
    __syncthreads();
    if (out)
    {
      const double res = (double)r0 + inv_h2 * (6.0 * c0 - S[buf][ty + 1][tx + 2] - S[buf][ty + 1][tx] -
                                                S[buf][ty + 2][tx + 1] - S[buf][ty][tx + 1] - cp - cm);
      sq += res * res;
      r[(t & 1) * hf + (t >> 1)] = (G)res;
    }
    cm = c0;
    c0 = cp;
    cp = cpp;
    cpp = c3;
    h0 = h1;
    h1 = h2;
    x0 = x1;
    x1 = x2;
    r0 = r1;
  }
  // block sum over the TX x TY lanes
  constexpr int W = TX * TY / 64;
  __shared__ double ws[W];
  for (int off = 32; off > 0; off >>= 1)
    sq += __shfl_down(sq, off, 64);
  const int wv = (ty * TX + tx) >> 6;
  if ((tx & 63) == 0)
    ws[wv] = sq;
  __syncthreads();
  if (tx == 0 && ty == 0)
  {
    double v = 0.0;
    for (int w = 0; w < W; w++)
      v += ws[w];
    part[blockIdx.x] = v;
  }
}

// Partial sums of rho^2 over the interior points (the residual of phi = 0).
template <typename G> __global__ void k_interior_sumsq(const G *__restrict__ rho, Cube cb, double *__restrict__ part)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned npts = (unsigned)n1 * n1 * n1;
  double sq = 0.0;
  const unsigned chunk = ((npts + gridDim.x - 1) / gridDim.x + blockDim.x - 1) / blockDim.x * blockDim.x;
  const unsigned t0 = xcd_block(blockIdx.x, gridDim.x) * chunk, t1 = min(npts, t0 + chunk);
  for (unsigned t = t0 + threadIdx.x; t < t1; t += blockDim.x)
  {
    int i, j, k;
    cube_ijk(cb, t, i, j, k);
    if (i > 0 && i < n && j > 0 && j < n && k > 0 && k < n)
    {
      const double v = (double)rho[t];
      sq += v * v;
    }
  }
  sq = block_reduce<0>(sq);
  if (threadIdx.x == 0)
    part[blockIdx.x] = sq;
}

// Outer correction phi += e (multigrid.c:199-209) with the pointwise change
// test of poisson_error.c:36-81 folded in: per-block max of a flag
// (>= 2 where phi/phi_old leaves [1-tol, 1+tol]) and of |phi/phi_old - 1|.
template <typename G>
__global__ void k_outer_correct(double *__restrict__ phi, const G *__restrict__ e, Cube cb, double tol,
                                double *__restrict__ part)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned npts = (unsigned)n1 * n1 * n1, hf = (npts + 1) >> 1;
  double dev = 0.0;
  const unsigned chunk = ((npts + gridDim.x - 1) / gridDim.x + blockDim.x - 1) / blockDim.x * blockDim.x;
  const unsigned t0 = xcd_block(blockIdx.x, gridDim.x) * chunk, t1 = min(npts, t0 + chunk);
  for (unsigned t = t0 + threadIdx.x; t < t1; t += blockDim.x)
  {
    int i, j, k;
    cube_ijk(cb, t, i, j, k);
    if (i > 0 && i < n && j > 0 && j < n && k > 0 && k < n)
    {
      double old = phi[t];
      double nw = old + (double)e[(t & 1) * hf + (t >> 1)];
      phi[t] = nw;
      double ratio = nw / old;
      double d = (1.0 - tol > ratio || 1.0 + tol < ratio) ? 2.0 + fabs(ratio - 1.0) : fabs(ratio - 1.0);
      if (d != d)
        d = 1e300; // 0/0: treat as not converged
      dev = fmax(dev, d);
    }
  }
  dev = block_reduce<1>(dev);
  if (threadIdx.x == 0)
    part[blockIdx.x] = dev;
}

// Starting guess of a head solve from the last two converged potentials:
// phi <- 2 phi - phi_prev, phi_prev <- phi (faces are fixed and stay equal).
__global__ void k_extrapolate(double *__restrict__ phi, double *__restrict__ phi_prev, size_t n)
{
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n)
    return;
  const double c = phi[t];
  phi[t] = 2.0 * c - phi_prev[t];
  phi_prev[t] = c;
}

// Red-black Gauss-Seidel half sweep on the error equation Lap e = r
// (gauss_saidel.c:40-53 restated with red-black ordering, omega = 1), on the
// colour-split layout: one thread per point of the colour, neighbours read
// from the other colour's half.
// Over-relaxation of the red half sweeps of the head V-cycle (red-black SOR,
// red points only: the black half sweep stays plain Gauss-Seidel, so the black
// residual is zero and the red-point restriction holds).  1.4 takes 96 against
// 108 V-cycles over ten steps of a 2^24-particle clumps workload on 256^3.
constexpr double kHeadRedOmega = 1.4;

// red half sweep update: e <- e + w (gs - e), from a zero start e <- w gs
template <typename G> __device__ inline G head_red_relax(const G *e, unsigned p, G gs, bool zero_start)
{
  const G w = (G)kHeadRedOmega;
  return zero_start ? w * gs : e[p] + w * (gs - e[p]);
}

template <typename G>
__global__ void k_rbgs(G *__restrict__ e, const G *__restrict__ r, Cube cb, G h2, int color, int zero_start)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned npts = (unsigned)n1 * n1 * n1;
  const unsigned half = (npts + 1) >> 1;
  const unsigned count = color == 0 ? half : npts - half;
  const unsigned h = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (h >= count)
    return;
  // r is loaded first: issued ahead of the neighbour loads, its latency overlaps
  // theirs (loaded after them, it cost a second memory round trip per wave)
  const unsigned me = color * half + h;
  const G rv = r[me];
  const unsigned idx = 2 * h + color;
  int i, j, k;
  cube_ijk(cb, idx, i, j, k);
  if (i == 0 || i == n || j == 0 || j == n || k == 0 || k == n)
    return;
  const unsigned s1 = n1, s2 = (unsigned)n1 * n1;
  const G *o = e + (color == 0 ? half : 0); // the other colour
  // zero_start: first half sweep from a zero initial guess (the other colour
  // counts as 0 whatever the array holds), which replaces a separate fill
  G sum = zero_start ? (G)0
                     : o[(idx + 1) >> 1] + o[(idx - 1) >> 1] + o[(idx + s1) >> 1] + o[(idx - s1) >> 1] +
                           o[(idx + s2) >> 1] + o[(idx - s2) >> 1];
  const G gs = (sum - h2 * rv) * (G)(1.0 / 6.0);
  e[me] = color != 0 ? gs : head_red_relax(e, me, gs, zero_start != 0);
}

// Residual of the fine error equation, full-weighting restricted to the coarse
// grid (27-point stencil, weights 1/64 * (1,2,1)^3), written at the coarse
// interior (the faces of rc are never written and stay zero).  It follows a
// red-black sweep pair that ended on black, which leaves a zero residual at
// black fine points: only red points count, the centre (8/64) and the 12
// points with two odd offsets (2/64).  Each of those 12 is shared by four
// coarse points, so coarse cell (I, J, K) owns three of them and computes each
// once: xy at fine (2I+1, 2J+1, 2K), xz at (2I+1, 2J, 2K+1) and yz at
// (2I, 2J+1, 2K+1).  Coarse point (I, J, K) sums xy of cells (I-1..I, J-1..J,
// K), xz of (I-1..I, J, K-1..K) and yz of (I, J-1..J, K-1..K).
//
// A block of TX x TY threads holds cells (bx (TX-1) + tx, by (TY-1) + ty) and
// marches up a chunk [K0, K1) of coarse planes: per plane it computes the
// owned residuals (four of the black neighbours come from the previous plane
// in registers), shares xy and the K-pair sums of xz and yz through LDS
// (double-buffered by plane parity: one barrier per plane) and writes the
// (TX-1) x (TY-1) coarse points of its tile; one plane below K0 supplies xz
// and yz at K0 - 1.  The blocks are numbered chunk fastest, then y tile, then
// x tile, in XCD-aware order.
template <typename G>
__global__ void __launch_bounds__(1024) k_restrict_residual(const G *__restrict__ e, const G *__restrict__ r,
                                                           G *__restrict__ rc, int nc, int yt, int zt, int L,
                                                           G inv_h2)
{
  __shared__ G s_xy[2][1024], s_xz[2][1024], s_yz[2][1024];
  const int TX = blockDim.x, TY = blockDim.y, tx = threadIdx.x, ty = threadIdx.y;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int bz = (int)(b % (unsigned)zt), byx = (int)(b / (unsigned)zt);
  const int by = byx % yt, bx = byx / yt;
  const int Ic = bx * (TX - 1) + tx, Jc = by * (TY - 1) + ty;
  const int K0 = 1 + bz * L, K1 = min(nc, K0 + L);
  const bool cell = Ic < nc && Jc < nc;
  const bool out = cell && tx > 0 && ty > 0;
  const int me = ty * TX + tx;
  const unsigned n1 = 2 * (unsigned)nc + 1, s1 = n1, s2 = n1 * n1, hf = (s2 * n1 + 1) >> 1;
  const unsigned c1 = (unsigned)nc + 1, c2 = c1 * c1, hc = (c2 * c1 + 1) >> 1;
  const G *eB = e + hf; // black half
  auto B = [&](unsigned x) { return eB[x >> 1]; };
  // residual at red fine point x from the sum of its six black neighbours
  auto res = [&](unsigned x, G nb) {
    const unsigned hp = x >> 1;
    return r[hp] - (nb - (G)6 * e[hp]) * inv_h2;
  };
  // fine index of (2 Ic, 2 Jc, 2 (K0 - 1)).  For a cell (Ic, Jc < nc) every
  // point read below has coordinates in [-1, 2 nc] with at most one -1, at a
  // plane >= 0 (or >= 1 where the -1 is in z), so its index lies in the grid;
  // the values read at a -1 feed residuals that no coarse point uses (xz of
  // Jc = 0, yz of Ic = 0, the centre of the halo cells).
  unsigned f = 2u * (unsigned)Ic + s1 * (2u * (unsigned)Jc) + s2 * (2u * (unsigned)(K0 - 1));
  G xz0 = 0, yz0 = 0;   // xz, yz of the plane below
  G cm2 = 0, cxy = 0, cp1 = 0, cps1 = 0; // black values carried to the next plane
  if (cell)
  {
    const G bs2 = B(f + s2), bxy = B(f + 1 + s1 + s2), bxzk = B(f + 1 + 2 * s2), byzk = B(f + s1 + 2 * s2);
    xz0 = res(f + 1 + s2, bs2 + B(f + 2 + s2) + B(f + 1 + s2 - s1) + bxy + B(f + 1) + bxzk);
    yz0 = res(f + s1 + s2, B(f + s1 + s2 - 1) + bxy + bs2 + B(f + 2 * s1 + s2) + B(f + s1) + byzk);
    cm2 = bs2;
    cxy = bxy;
    cp1 = bxzk;
    cps1 = byzk;
  }
  for (int K = K0; K < K1; K++)
  {
    f += 2 * s2;
    G ctr = 0, xy = 0, xz = 0, yz = 0;
    if (cell)
    {
      // carried: cm2 = B(f - s2), cp1 = B(f + 1), cps1 = B(f + s1), cxy = B(f + 1 + s1 - s2)
      const G bs2 = B(f + s2), bxy = B(f + 1 + s1 + s2), bxzk = B(f + 1 + 2 * s2), byzk = B(f + s1 + 2 * s2);
      ctr = res(f, B(f - 1) + cp1 + B(f - s1) + cps1 + cm2 + bs2);
      xy = res(f + 1 + s1, cps1 + B(f + 2 + s1) + cp1 + B(f + 1 + 2 * s1) + cxy + bxy);
      xz = res(f + 1 + s2, bs2 + B(f + 2 + s2) + B(f + 1 + s2 - s1) + bxy + cp1 + bxzk);
      yz = res(f + s1 + s2, B(f + s1 + s2 - 1) + bxy + bs2 + B(f + 2 * s1 + s2) + cps1 + byzk);
      cm2 = bs2;
      cxy = bxy;
      cp1 = bxzk;
      cps1 = byzk;
    }
    const int buf = K & 1;
    s_xy[buf][me] = xy;
    s_xz[buf][me] = xz0 + xz;
    s_yz[buf][me] = yz0 + yz;
    xz0 = xz;
    yz0 = yz;
    __syncthreads();
    if (out)
    {
      const G edges = (s_xy[buf][me] + s_xy[buf][me - 1]) + (s_xy[buf][me - TX] + s_xy[buf][me - TX - 1]) +
                      (s_xz[buf][me] + s_xz[buf][me - 1]) + (s_yz[buf][me] + s_yz[buf][me - TX]);
      const unsigned q = (unsigned)Ic + c1 * (unsigned)Jc + c2 * (unsigned)K;
      rc[(q & 1) * hc + (q >> 1)] = (G)0.125 * ctr + (G)(2.0 / 64.0) * edges;
    }
  }
}

// Launch shape of k_restrict_residual onto nc coarse cells per side: a tile
// spans the whole x range where it can (TX = nc up to 256), TY rows make up to
// 1024 threads, and the K chunk length L gives about 512 blocks.
struct RestrictShape
{
  int tx, ty, xt, yt, zt, L;
};
inline RestrictShape restrict_shape(int nc)
{
  RestrictShape s;
  s.tx = nc < 256 ? nc : 256;
  s.ty = 1024 / s.tx < nc ? 1024 / s.tx : nc;
  if (s.ty < 2)
    s.ty = 2;
  s.xt = (nc - 1 + s.tx - 2) / (s.tx - 1);
  s.yt = (nc - 1 + s.ty - 2) / (s.ty - 1);
  const long long cols = (long long)(nc - 1) * s.xt * s.yt;
  s.L = (int)((cols + 511) / 512);
  if (s.L < 2)
    s.L = 2;
  s.zt = (nc - 1 + s.L - 1) / s.L;
  return s;
}

// Trilinear prolongation of the coarse correction into the fine interior
// (multigrid.c:51-104) fused with the correction e += P ec (:199-209).  One
// thread per coarse cell (I, J, K) in [0, nc)^3 (the Cube `cells` has nc per
// side) updates the cell's 8 fine points (2I + a, 2J + b, 2K + c), a, b, c in
// {0, 1}, from the cell's 8 corners: one coarse load per fine point where a
// thread per fine point needs eight.  Fine points on the low faces (a
// coordinate 0) are left alone; the high faces (2 nc) lie outside every cell.
// For a fixed (a, b, c) the fine points of consecutive I are consecutive
// entries of one colour half, so each of the 8 updates is a coalesced access.
template <typename G>
__global__ void k_prolong_correct(G *__restrict__ e, const G *__restrict__ ec, Cube cells)
{
  const int nc = cells.n1;
  const unsigned ncell = (unsigned)nc * nc * nc;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= ncell)
    return;
  int I, J, K;
  cube_ijk(cells, t, I, J, K);
  const unsigned c1 = (unsigned)nc + 1, c2 = c1 * c1, hc = (c2 * c1 + 1) >> 1;
  const unsigned q = (unsigned)I + c1 * (unsigned)J + c2 * (unsigned)K;
  auto C = [&](unsigned d) {
    const unsigned v = q + d;
    return ec[(v & 1) * hc + (v >> 1)];
  };
  const G c000 = C(0), c100 = C(1), c010 = C(c1), c110 = C(1 + c1);
  const G c001 = C(c2), c101 = C(1 + c2), c011 = C(c1 + c2), c111 = C(1 + c1 + c2);
  const unsigned n1 = 2 * (unsigned)nc + 1, s1 = n1, s2 = n1 * n1, hf = (s2 * n1 + 1) >> 1;
  const unsigned f = 2 * (unsigned)I + s1 * (2 * (unsigned)J) + s2 * (2 * (unsigned)K); // even: red
  auto add = [&](unsigned d, G v) {
    const unsigned w = f + d;
    e[(w & 1) * hf + (w >> 1)] += v;
  };
  // sums along x of the four x edges of the cell
  const G x00 = c000 + c100, x10 = c010 + c110, x01 = c001 + c101, x11 = c011 + c111;
  const bool i0 = I > 0, j0 = J > 0, k0 = K > 0;
  if (i0 && j0 && k0)
    add(0, c000);
  if (j0 && k0)
    add(1, (G)0.5 * x00);
  if (i0 && k0)
    add(s1, (G)0.5 * (c000 + c010));
  if (k0)
    add(1 + s1, (G)0.25 * (x00 + x10));
  if (i0 && j0)
    add(s2, (G)0.5 * (c000 + c001));
  if (j0)
    add(1 + s2, (G)0.25 * (x00 + x01));
  if (i0)
    add(s1 + s2, (G)0.25 * ((c000 + c010) + (c001 + c011)));
  add(1 + s1 + s2, (G)0.125 * ((x00 + x10) + (x01 + x11)));
}

// ------------------------------------------------------------------ V-cycle tail in one workgroup
// The coarse end of a head V-cycle (levels of NTOP, NTOP / 2, ..., 4 cells per
// side) in a single launch: e and r of every tail level staged in LDS
// (colour-split, like the global levels), the same red-black sweeps, the
// 13-point restriction of the red residuals, trilinear prolongation, and on
// the 4-cell level an exact solve, separated by workgroup barriers instead of
// launches.  Level sizes are compile-time constants.
template <typename G> struct Tail
{
  static constexpr int cells = sizeof(G) == 4 ? 16 : 8; // finest tail level
};
constexpr int kTailThreads = 1024;

template <int M> constexpr int tail_points()
{
  if constexpr (M < 4)
    return 0;
  else
    return (M + 1) * (M + 1) * (M + 1) + tail_points<M / 2>();
}

// red-black half sweep of Lap e = r on the M-cell level (faces stay zero)
template <typename G, int M> __device__ inline void tail_sweep(G *E, const G *R, int color, bool zs)
{
  constexpr int P = M + 1, PP = P * P, NP = PP * P, HF = (NP + 1) / 2;
  const G h2 = (G)(1.0 / ((double)M * M));
  const int cnt = color == 0 ? HF : NP - HF;
  const G *o = E + (color == 0 ? HF : 0);
  for (int h = threadIdx.x; h < cnt; h += blockDim.x)
  {
    const int idx = 2 * h + color;
    const int i = idx % P, j = (idx / P) % P, k = idx / PP;
    if (i == 0 || i == M || j == 0 || j == M || k == 0 || k == M)
      continue;
    const G rv = R[color * HF + h];
    G sum = zs ? (G)0
               : o[(idx + 1) >> 1] + o[(idx - 1) >> 1] + o[(idx + P) >> 1] + o[(idx - P) >> 1] + o[(idx + PP) >> 1] +
                     o[(idx - PP) >> 1];
    E[color * HF + h] = (sum - h2 * rv) * (G)(1.0 / 6.0);
  }
  __syncthreads();
}

// full-weighting restriction M -> M/2 after a sweep pair ending on black: the
// red residuals once into `res`, then the 13-point sums (zero on the faces)
template <typename G, int M> __device__ inline void tail_restrict(const G *E, const G *R, G *RC, G *res)
{
  constexpr int P = M + 1, PP = P * P, NP = PP * P, HF = (NP + 1) / 2;
  constexpr int MC = M / 2, PC = MC + 1, PPC = PC * PC, NPC = PPC * PC, HC = (NPC + 1) / 2;
  const G inv_h2 = (G)((double)M * M);
  const G *EB = E + HF;
  for (int h = threadIdx.x; h < HF; h += blockDim.x)
  {
    const int f = 2 * h;
    const int i = f % P, j = (f / P) % P, k = f / PP;
    G v = (G)0;
    if (!(i == 0 || i == M || j == 0 || j == M || k == 0 || k == M))
    {
      const G lap = (EB[(f + 1) >> 1] + EB[(f - 1) >> 1] + EB[(f + P) >> 1] + EB[(f - P) >> 1] + EB[(f + PP) >> 1] +
                     EB[(f - PP) >> 1] - (G)6 * E[h]) *
                    inv_h2;
      v = R[h] - lap;
    }
    res[h] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NPC; t += blockDim.x)
  {
    const int I = t % PC, J = (t / PC) % PC, K = t / PPC;
    G v = (G)0;
    if (!(I == 0 || I == MC || J == 0 || J == MC || K == 0 || K == MC))
    {
      const int f = 2 * I + P * 2 * J + PP * 2 * K;
      auto S = [&](int d) { return res[(f + d) >> 1]; };
      const G edges = S(1 + P) + S(1 - P) + S(-1 + P) + S(-1 - P) + S(1 + PP) + S(1 - PP) + S(-1 + PP) + S(-1 - PP) +
                      S(P + PP) + S(P - PP) + S(-P + PP) + S(-P - PP);
      v = (G)0.125 * S(0) + (G)(2.0 / 64.0) * edges;
    }
    RC[(t & 1) * HC + (t >> 1)] = v;
  }
  __syncthreads();
}

// trilinear prolongation M/2 -> M fused with the correction, one thread per
// coarse cell (the formulas of k_prolong_correct)
template <typename G, int M> __device__ inline void tail_prolong(G *E, const G *EC)
{
  constexpr int P = M + 1, PP = P * P, NP = PP * P, HF = (NP + 1) / 2;
  constexpr int MC = M / 2, C1 = MC + 1, C2 = C1 * C1, HC = (C2 * C1 + 1) / 2;
  for (int t = threadIdx.x; t < MC * MC * MC; t += blockDim.x)
  {
    const int I = t % MC, J = (t / MC) % MC, K = t / (MC * MC);
    const int q = I + C1 * J + C2 * K;
    auto C = [&](int d) { return EC[((q + d) & 1) * HC + ((q + d) >> 1)]; };
    const G c000 = C(0), c100 = C(1), c010 = C(C1), c110 = C(1 + C1);
    const G c001 = C(C2), c101 = C(1 + C2), c011 = C(C1 + C2), c111 = C(1 + C1 + C2);
    const int f = 2 * I + P * 2 * J + PP * 2 * K;
    auto add = [&](int d, G v) { E[((f + d) & 1) * HF + ((f + d) >> 1)] += v; };
    const G x00 = c000 + c100, x10 = c010 + c110, x01 = c001 + c101, x11 = c011 + c111;
    const bool i0 = I > 0, j0 = J > 0, k0 = K > 0;
    if (i0 && j0 && k0)
      add(0, c000);
    if (j0 && k0)
      add(1, (G)0.5 * x00);
    if (i0 && k0)
      add(P, (G)0.5 * (c000 + c010));
    if (k0)
      add(1 + P, (G)0.25 * (x00 + x10));
    if (i0 && j0)
      add(PP, (G)0.5 * (c000 + c001));
    if (j0)
      add(1 + PP, (G)0.25 * (x00 + x01));
    if (i0)
      add(P + PP, (G)0.25 * ((c000 + c010) + (c001 + c011)));
    add(1 + P + PP, (G)0.125 * ((x00 + x10) + (x01 + x11)));
  }
  __syncthreads();
}

// The 4-cell level solved exactly: Lap e = r with zero faces on its 27
// interior points is e = -(1/16) T^-1 r, T = 6 I - adjacency, and T is
// diagonalised by the 3-point sine transform S (S_ai = sin(pi a i / 4) / sqrt 2,
// a, i = 1..3, symmetric and orthogonal) with eigenvalues mu_a + mu_b + mu_c,
// mu_a = 2 - 2 cos(pi a / 4).  Two passes of 27 threads, in FP64.
template <typename G> __device__ inline void tail_coarse4(G *E, const G *R, double *cb)
{
  constexpr int P = 5, PP = 25, HF = 63;
  const double s = 0.70710678118654752440; // 1 / sqrt 2 = cos(pi / 4)
  const double S[3][3] = {{0.5, s, 0.5}, {s, 0.0, -s}, {0.5, -s, 0.5}};
  const double mu[3] = {2.0 - 2.0 * s, 2.0, 2.0 + 2.0 * s};
  const int tid = threadIdx.x;
  if (tid < 27)
  {
    const int a = tid % 3, b = (tid / 3) % 3, c = tid / 9;
    double B = 0.0;
    for (int v = 0; v < 27; v++)
    {
      const int i = v % 3, j = (v / 3) % 3, k = v / 9;
      const int g = (i + 1) + P * (j + 1) + PP * (k + 1);
      B += S[a][i] * S[b][j] * S[c][k] * (double)R[(g & 1) * HF + (g >> 1)];
    }
    cb[tid] = B / (mu[a] + mu[b] + mu[c]);
  }
  __syncthreads();
  if (tid < 27)
  {
    const int i = tid % 3, j = (tid / 3) % 3, k = tid / 9;
    double x = 0.0;
    for (int w = 0; w < 27; w++)
    {
      const int a = w % 3, b = (w / 3) % 3, c = w / 9;
      x += S[a][i] * S[b][j] * S[c][k] * cb[w];
    }
    const int g = (i + 1) + P * (j + 1) + PP * (k + 1);
    E[(g & 1) * HF + (g >> 1)] = (G)(-x / 16.0);
  }
  __syncthreads();
}

template <typename G, int M> __device__ inline void tail_level(G *se, G *sr, G *res, double *cb, int pre, int post)
{
  if constexpr (M == 4)
    tail_coarse4<G>(se, sr, cb);
  else
  {
    constexpr int NP = (M + 1) * (M + 1) * (M + 1);
    for (int s = 0; s < pre; s++)
    {
      tail_sweep<G, M>(se, sr, 0, s == 0);
      tail_sweep<G, M>(se, sr, 1, false);
    }
    tail_restrict<G, M>(se, sr, sr + NP, res);
    tail_level<G, M / 2>(se + NP, sr + NP, res, cb, pre, post);
    tail_prolong<G, M>(se, se + NP);
    for (int s = 0; s < post; s++)
    {
      tail_sweep<G, M>(se, sr, 0, false);
      tail_sweep<G, M>(se, sr, 1, false);
    }
  }
}

template <typename G, int NTOP>
__global__ void __launch_bounds__(kTailThreads)
    k_vcycle_tail(G *__restrict__ e_top, const G *__restrict__ r_top, int pre, int post)
{
  constexpr int NPT = tail_points<NTOP>(), NP0 = (NTOP + 1) * (NTOP + 1) * (NTOP + 1), HF0 = (NP0 + 1) / 2;
  __shared__ G se[NPT], sr[NPT], res[HF0];
  __shared__ double cb[27];
  // the top level's r: every load of the thread issued before its LDS stores
  constexpr int SU = (NP0 + kTailThreads - 1) / kTailThreads;
  G rv[SU];
#pragma unroll
  for (int q = 0; q < SU; q++)
  {
    const int t = threadIdx.x + q * kTailThreads;
    rv[q] = r_top[t < NP0 ? t : 0];
  }
  for (int t = threadIdx.x; t < NPT; t += blockDim.x)
  {
    se[t] = (G)0;
    sr[t] = (G)0;
  }
#pragma unroll
  for (int q = 0; q < SU; q++)
  {
    const int t = threadIdx.x + q * kTailThreads;
    if (t < NP0)
      sr[t] = rv[q];
  }
  __syncthreads();
  tail_level<G, NTOP>(se, sr, res, cb, pre, post);
  for (int t = threadIdx.x; t < NP0; t += blockDim.x)
    e_top[t] = se[t];
}

// Coarsest level: conjugate gradients in one workgroup (FP64 internally) on a
// grid of at most 9^3 points, zero Dirichlet faces, solving Lap e = r.
template <typename G> __global__ void k_coarse_cg(G *__restrict__ e, const G *__restrict__ r, int n, int iters)
{
  constexpr int MAXP = 729;
  __shared__ double xs[MAXP], rs[MAXP], ps[MAXP], aps[MAXP];
  const int n1 = n + 1, np = n1 * n1 * n1;
  const double inv_h2 = (double)n * n;
  for (int t = threadIdx.x; t < np; t += blockDim.x)
  {
    int i = t % n1, j = (t / n1) % n1, k = t / (n1 * n1);
    bool in = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
    xs[t] = 0.0;
    // A = -Lap (SPD); solve A x = -r
    rs[t] = in ? -(double)r[rbpos(t, rbhalf(np))] : 0.0;
    ps[t] = rs[t];
  }
  __syncthreads();
  double rr = 0.0;
  for (int t = threadIdx.x; t < np; t += blockDim.x)
    rr += rs[t] * rs[t];
  rr = block_reduce<0>(rr);
  const double rr0 = rr;
  for (int it = 0; it < iters && rr > 1e-28 * rr0 && rr > 0.0; it++)
  {
    double pap = 0.0;
    for (int t = threadIdx.x; t < np; t += blockDim.x)
    {
      int i = t % n1, j = (t / n1) % n1, k = t / (n1 * n1);
      bool in = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
      double v = 0.0;
      if (in)
        v = inv_h2 * (6.0 * ps[t] - ps[t + 1] - ps[t - 1] - ps[t + n1] - ps[t - n1] - ps[t + n1 * n1] -
                      ps[t - n1 * n1]);
      aps[t] = v;
      pap += ps[t] * v;
    }
    pap = block_reduce<0>(pap);
    double alpha = rr / pap;
    double rn = 0.0;
    for (int t = threadIdx.x; t < np; t += blockDim.x)
    {
      xs[t] += alpha * ps[t];
      rs[t] -= alpha * aps[t];
      rn += rs[t] * rs[t];
    }
    rn = block_reduce<0>(rn);
    double beta = rn / rr;
    for (int t = threadIdx.x; t < np; t += blockDim.x)
      ps[t] = rs[t] + beta * ps[t];
    rr = rn;
    __syncthreads();
  }
  for (int t = threadIdx.x; t < np; t += blockDim.x)
    e[rbpos(t, rbhalf(np))] = (G)xs[t];
}

// ------------------------------------------------------------------ fused levels below the top
// Below the finest head level a grid is small enough that a launch costs more
// than the data it moves.  There the two pre-smoothing red-black pairs and the
// residual restriction of a level run as one kernel (k_mg_down), and the
// prolongation with the two post-smoothing pairs as another (k_mg_up).  Each
// workgroup owns a tile of kMgCT^3 coarse points and the fine points
// [2 I0 - 1, 2 I0 + 2 kMgCT - 2] per axis (I0 the tile's first coarse index),
// and stages the fine region [2 I0 - 5, 2 I0 + 2 kMgCT + 3] (kMgLF points per
// axis, odd) in LDS.  Every sweep phase runs on a region one point narrower
// per side than the phase before it, so the owned points end exactly as the
// per-level kernels leave them (same update formulas and summation orders);
// the halo is recomputed by the neighbouring tiles.  LDS holds e and r of the
// region colour-split by local parity (kMgLF odd: local parity is the
// coordinate-sum parity, and the global colour is it XOR that of the origin).
constexpr int kMgCT = 4;                          // coarse points per tile side
constexpr int kMgLF = 2 * kMgCT + 9;              // fine points per LDS region side
constexpr int kMgNP = kMgLF * kMgLF * kMgLF;      // LDS region points
constexpr int kMgHF = (kMgNP + 1) / 2;            // first (even local parity) half
constexpr int kMgThreads = 512;
constexpr int kMgFusedCells = 64; // largest level (cells per side) run fused

__device__ inline bool mg_interior(int X, int Y, int Z, int m)
{
  return X >= 1 && X <= m - 1 && Y >= 1 && Y <= m - 1 && Z >= 1 && Z <= m - 1;
}

// One red-black half sweep on the local points of parity lp in the region
// [S, LF - 1 - S]^3 of an LF^3 LDS region (LF odd) that are interior points of
// the level (m cells); ZS is the zero-start form of the first sweep (k_rbgs
// with zero_start).  Slots are laid out row by row, (LF - 2S + 1) / 2 per row.
template <typename G, int LF, int S, bool ZS>
__device__ inline void mg_tile_sweep(G *sE, const G *sR, int lp, int fx, int fy, int fz, int m, G h2, bool red)
{
  constexpr int LF2 = LF * LF, HFL = (LF * LF2 + 1) / 2, L = LF - 2 * S, HX = (L + 1) / 2, SLOTS = HX * L * L;
  const G *o = sE + (1 - lp) * HFL;
  for (int q = threadIdx.x; q < SLOTS; q += blockDim.x)
  {
    const int row = q / HX, u = q - row * HX;
    const int y = S + row % L, z = S + row / L;
    const int x = S + ((S + y + z + lp) & 1) + 2 * u;
    if (x > LF - 1 - S || !mg_interior(fx + x, fy + y, fz + z, m))
      continue;
    const int idx = x + LF * y + LF2 * z;
    const int p = lp * HFL + (idx >> 1);
    const G rv = sR[p];
    G sum = ZS ? (G)0
               : o[(idx + 1) >> 1] + o[(idx - 1) >> 1] + o[(idx + LF) >> 1] + o[(idx - LF) >> 1] +
                     o[(idx + LF2) >> 1] + o[(idx - LF2) >> 1];
    const G gs = (sum - h2 * rv) * (G)(1.0 / 6.0);
    sE[p] = red ? head_red_relax(sE, (unsigned)p, gs, ZS) : gs;
  }
  __syncthreads();
}

// Pre-smoothing (two pairs) and full-weighting restriction of the residual of
// level l (m cells) into rc of level l + 1, for the tiles of T^3 (T = m / (2
// kMgCT)); writes the smoothed correction of the owned fine points to et (for
// k_mg_up).  ZS: the correction starts from zero (e is not read); otherwise
// from e, whose first red sweep then needs one more halo point (the region is
// kMgLF + 2 per side).  Matches k_rbgs x4 followed by k_restrict_residual point
// for point.
template <typename G, bool ZS>
__global__ void __launch_bounds__(kMgThreads)
    k_mg_down(const G *__restrict__ e, G *__restrict__ et, const G *__restrict__ r, G *__restrict__ rc, int m, int T,
              G h2, G inv_h2)
{
  constexpr int H = ZS ? 4 : 5;                // halo: phases before the residual
  constexpr int LF = 2 * kMgCT + 1 + 2 * H, LF2 = LF * LF, NP = LF * LF2, HF = (NP + 1) / 2;
  __shared__ G sE[NP], sR[NP];
  const int tid = threadIdx.x, NT = blockDim.x;
  const int b = (int)xcd_block(blockIdx.x, gridDim.x);
  const int I0 = 1 + kMgCT * (b % T), J0 = 1 + kMgCT * ((b / T) % T), K0 = 1 + kMgCT * (b / (T * T));
  const int fx = 2 * I0 - 1 - H, fy = 2 * J0 - 1 - H, fz = 2 * K0 - 1 - H; // level coordinates of local (0, 0, 0)
  const int goff = (fx + fy + fz) & 1;                                  // local parity of the red points
  const unsigned n1 = (unsigned)m + 1, s1 = n1, s2 = n1 * n1, hf = (s2 * n1 + 1) >> 1;
  // staging: every global load of the thread is issued before the first LDS
  // store (a load feeding a store per iteration costs a round trip each);
  // points outside the grid read index 0 and are replaced by zero
  {
    constexpr int SU = (NP + kMgThreads - 1) / kMgThreads;
    G rv[SU], ev[SU];
#pragma unroll
    for (int q = 0; q < SU; q++)
    {
      const int t = tid + q * kMgThreads;
      const int x = t % LF, y = (t / LF) % LF, z = t / LF2;
      const int X = fx + x, Y = fy + y, Z = fz + z;
      const bool in = t < NP && X >= 0 && X <= m && Y >= 0 && Y <= m && Z >= 0 && Z <= m;
      const unsigned g = in ? (unsigned)X + s1 * (unsigned)Y + s2 * (unsigned)Z : 0u;
      const unsigned gp = (g & 1) * hf + (g >> 1);
      const G a = r[gp];
      rv[q] = in ? a : (G)0;
      if (!ZS)
      {
        const G c = e[gp];
        ev[q] = in && mg_interior(X, Y, Z, m) ? c : (G)0;
      }
      else
        ev[q] = (G)0;
    }
#pragma unroll
    for (int q = 0; q < SU; q++)
    {
      const int t = tid + q * kMgThreads;
      if (t < NP)
      {
        const int p = (t & 1) * HF + (t >> 1);
        sR[p] = rv[q];
        sE[p] = ev[q];
      }
    }
  }
  __syncthreads();
  const int red = goff, black = goff ^ 1;
  if constexpr (ZS)
  {
    mg_tile_sweep<G, LF, 0, true>(sE, sR, red, fx, fy, fz, m, h2, true);
    mg_tile_sweep<G, LF, 1, false>(sE, sR, black, fx, fy, fz, m, h2, false);
    mg_tile_sweep<G, LF, 2, false>(sE, sR, red, fx, fy, fz, m, h2, true);
    mg_tile_sweep<G, LF, 3, false>(sE, sR, black, fx, fy, fz, m, h2, false);
  }
  else
  {
    mg_tile_sweep<G, LF, 1, false>(sE, sR, red, fx, fy, fz, m, h2, true);
    mg_tile_sweep<G, LF, 2, false>(sE, sR, black, fx, fy, fz, m, h2, false);
    mg_tile_sweep<G, LF, 3, false>(sE, sR, red, fx, fy, fz, m, h2, true);
    mg_tile_sweep<G, LF, 4, false>(sE, sR, black, fx, fy, fz, m, h2, false);
  }
  // residual r - Lap e at the red points of [H, LF - 1 - H]^3, in place of r
  {
    constexpr int S = H, L = LF - 2 * S, HX = (L + 1) / 2, SLOTS = HX * L * L;
    const G *o = sE + black * HF;
    for (int q = tid; q < SLOTS; q += NT)
    {
      const int row = q / HX, u = q - row * HX;
      const int y = S + row % L, z = S + row / L;
      const int x = S + ((S + y + z + red) & 1) + 2 * u;
      if (x > LF - 1 - S || !mg_interior(fx + x, fy + y, fz + z, m))
        continue;
      const int idx = x + LF * y + LF2 * z;
      const int p = red * HF + (idx >> 1);
      const G nb = o[(idx - 1) >> 1] + o[(idx + 1) >> 1] + o[(idx - LF) >> 1] + o[(idx + LF) >> 1] +
                   o[(idx - LF2) >> 1] + o[(idx + LF2) >> 1];
      sR[p] = sR[p] - (nb - (G)6 * sE[p]) * inv_h2;
    }
    // the smoothed correction of the owned points [H, H + 2 kMgCT - 1]^3
    constexpr int OW = 2 * kMgCT;
    for (int q = tid; q < OW * OW * OW; q += NT)
    {
      const int x = H + q % OW, y = H + (q / OW) % OW, z = H + q / (OW * OW);
      const int X = fx + x, Y = fy + y, Z = fz + z;
      if (!mg_interior(X, Y, Z, m))
        continue;
      const int idx = x + LF * y + LF2 * z;
      const unsigned g = (unsigned)X + s1 * (unsigned)Y + s2 * (unsigned)Z;
      et[(g & 1) * hf + (g >> 1)] = sE[(idx & 1) * HF + (idx >> 1)];
    }
  }
  __syncthreads();
  // restriction to the tile's coarse interior points (summation order of k_restrict_residual)
  const int mc = m / 2;
  const unsigned c1 = (unsigned)mc + 1, c2 = c1 * c1, hc = (c2 * c1 + 1) >> 1;
  const G *res = sR + red * HF;
  for (int q = tid; q < kMgCT * kMgCT * kMgCT; q += NT)
  {
    const int a = q % kMgCT, bb = (q / kMgCT) % kMgCT, c = q / (kMgCT * kMgCT);
    const int I = I0 + a, J = J0 + bb, K = K0 + c;
    if (I > mc - 1 || J > mc - 1 || K > mc - 1)
      continue;
    const int idx = (2 * a + 1 + H) + LF * (2 * bb + 1 + H) + LF2 * (2 * c + 1 + H);
    auto R = [&](int d) { return res[(idx + d) >> 1]; };
    const G ctr = R(0);
    const G xy00 = R(1 + LF), xy10 = R(-1 + LF), xy01 = R(1 - LF), xy11 = R(-1 - LF);
    const G xz0 = R(1 - LF2) + R(1 + LF2), xz1 = R(-1 - LF2) + R(-1 + LF2);
    const G yz0 = R(LF - LF2) + R(LF + LF2), yz1 = R(-LF - LF2) + R(-LF + LF2);
    const G edges = (xy00 + xy10) + (xy01 + xy11) + (xz0 + xz1) + (yz0 + yz1);
    const unsigned qc = (unsigned)I + c1 * (unsigned)J + c2 * (unsigned)K;
    rc[(qc & 1) * hc + (qc >> 1)] = (G)0.125 * ctr + (G)(2.0 / 64.0) * edges;
  }
}

// Prolongation of the coarse correction ec (level l + 1) into level l's
// pre-smoothed correction et (written by k_mg_down) and the two post-smoothing
// pairs, for the same tiles as k_mg_down; the result goes to e (a separate
// array: the tiles read et in their halos, which other tiles own).  Matches
// k_prolong_correct followed by k_rbgs x4.
template <typename G>
__global__ void __launch_bounds__(kMgThreads)
    k_mg_up(G *__restrict__ e, const G *__restrict__ r, const G *__restrict__ et, const G *__restrict__ ec, int m,
            int T, G h2)
{
  constexpr int LF = kMgLF, LF2 = LF * LF, HF = kMgHF;
  constexpr int LC = kMgCT + 6; // coarse points per axis under the region
  __shared__ G sE[kMgNP], sR[kMgNP], sC[LC * LC * LC];
  const int tid = threadIdx.x, NT = blockDim.x;
  const int b = (int)xcd_block(blockIdx.x, gridDim.x);
  const int I0 = 1 + kMgCT * (b % T), J0 = 1 + kMgCT * ((b / T) % T), K0 = 1 + kMgCT * (b / (T * T));
  const int fx = 2 * I0 - 5, fy = 2 * J0 - 5, fz = 2 * K0 - 5;
  const int cx = I0 - 3, cy = J0 - 3, cz = K0 - 3; // coarse coordinates of sC (0, 0, 0)
  const int goff = (fx + fy + fz) & 1;
  const int mc = m / 2;
  const unsigned c1 = (unsigned)mc + 1, c2 = c1 * c1, hc = (c2 * c1 + 1) >> 1;
  const unsigned n1 = (unsigned)m + 1, s1 = n1, s2 = n1 * n1, hf = (s2 * n1 + 1) >> 1;
  // staging: all global loads (coarse correction, r and the pre-smoothed
  // correction of the region) are issued before the first LDS store; points
  // outside the grid read index 0 and are replaced by zero
  constexpr int CU = (LC * LC * LC + kMgThreads - 1) / kMgThreads, SU = (kMgNP + kMgThreads - 1) / kMgThreads;
  G cv[CU], rv[SU], tv[SU];
#pragma unroll
  for (int q = 0; q < CU; q++)
  {
    const int t = tid + q * kMgThreads;
    const int X = cx + t % LC, Y = cy + (t / LC) % LC, Z = cz + t / (LC * LC);
    const bool in = t < LC * LC * LC && X >= 0 && X <= mc && Y >= 0 && Y <= mc && Z >= 0 && Z <= mc;
    const unsigned g = in ? (unsigned)X + c1 * (unsigned)Y + c2 * (unsigned)Z : 0u;
    const G v = ec[(g & 1) * hc + (g >> 1)];
    cv[q] = in ? v : (G)0;
  }
#pragma unroll
  for (int q = 0; q < SU; q++)
  {
    const int t = tid + q * kMgThreads;
    const int x = t % LF, y = (t / LF) % LF, z = t / LF2;
    const int X = fx + x, Y = fy + y, Z = fz + z;
    const bool in = t < kMgNP && X >= 0 && X <= m && Y >= 0 && Y <= m && Z >= 0 && Z <= m;
    const unsigned g = in ? (unsigned)X + s1 * (unsigned)Y + s2 * (unsigned)Z : 0u;
    const unsigned gp = (g & 1) * hf + (g >> 1);
    const G a = r[gp], c = et[gp];
    rv[q] = in ? a : (G)0;
    tv[q] = in ? c : (G)0;
  }
#pragma unroll
  for (int q = 0; q < CU; q++)
  {
    const int t = tid + q * kMgThreads;
    if (t < LC * LC * LC)
      sC[t] = cv[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SU; q++)
  {
    const int t = tid + q * kMgThreads;
    if (t >= kMgNP)
      continue;
    const int x = t % LF, y = (t / LF) % LF, z = t / LF2;
    const int X = fx + x, Y = fy + y, Z = fz + z;
    G ev = (G)0;
    if (mg_interior(X, Y, Z, m))
    {
      // the trilinear value of k_prolong_correct for corner offsets (X & 1, Y & 1, Z & 1)
      const int q0 = ((X >> 1) - cx) + LC * (((Y >> 1) - cy) + LC * ((Z >> 1) - cz));
      const G c000 = sC[q0], c100 = sC[q0 + 1], c010 = sC[q0 + LC], c110 = sC[q0 + 1 + LC];
      const G c001 = sC[q0 + LC * LC], c101 = sC[q0 + 1 + LC * LC], c011 = sC[q0 + LC + LC * LC],
              c111 = sC[q0 + 1 + LC + LC * LC];
      const G x00 = c000 + c100, x10 = c010 + c110, x01 = c001 + c101, x11 = c011 + c111;
      G v;
      switch ((X & 1) | ((Y & 1) << 1) | ((Z & 1) << 2))
      {
      case 0: v = c000; break;
      case 1: v = (G)0.5 * x00; break;
      case 2: v = (G)0.5 * (c000 + c010); break;
      case 3: v = (G)0.25 * (x00 + x10); break;
      case 4: v = (G)0.5 * (c000 + c001); break;
      case 5: v = (G)0.25 * (x00 + x01); break;
      case 6: v = (G)0.25 * ((c000 + c010) + (c001 + c011)); break;
      default: v = (G)0.125 * ((x00 + x10) + (x01 + x11)); break;
      }
      ev = tv[q] + v;
    }
    const int p = (t & 1) * HF + (t >> 1);
    sR[p] = rv[q];
    sE[p] = ev;
  }
  __syncthreads();
  const int red = goff, black = goff ^ 1;
  mg_tile_sweep<G, kMgLF, 1, false>(sE, sR, red, fx, fy, fz, m, h2, true);
  mg_tile_sweep<G, kMgLF, 2, false>(sE, sR, black, fx, fy, fz, m, h2, false);
  mg_tile_sweep<G, kMgLF, 3, false>(sE, sR, red, fx, fy, fz, m, h2, true);
  mg_tile_sweep<G, kMgLF, 4, false>(sE, sR, black, fx, fy, fz, m, h2, false);
  constexpr int OW = 2 * kMgCT;
  for (int q = tid; q < OW * OW * OW; q += NT)
  {
    const int x = 4 + q % OW, y = 4 + (q / OW) % OW, z = 4 + q / (OW * OW);
    const int X = fx + x, Y = fy + y, Z = fz + z;
    if (!mg_interior(X, Y, Z, m))
      continue;
    const int idx = x + LF * y + LF2 * z;
    const unsigned g = (unsigned)X + s1 * (unsigned)Y + s2 * (unsigned)Z;
    e[(g & 1) * hf + (g >> 1)] = sE[(idx & 1) * HF + (idx >> 1)];
  }
}

// A11 head gradient on interior points (grid_acceleration.c:36-76); the faces
// keep their A10 values.
template <typename G>
__global__ void k_grid_acc_head(const double *__restrict__ phi, G *__restrict__ ax, G *__restrict__ ay,
                                G *__restrict__ az, Cube cb, double inv_2h, double GM, int zero_bc)
{
  const int n1 = cb.n1, n = n1 - 1;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= (unsigned)n1 * n1 * n1)
    return;
  int i, j, k;
  cube_ijk(cb, t, i, j, k);
  if (i == 0 || i == n || j == 0 || j == n || k == 0 || k == n)
  {
    // the faces keep the initial monopole values (the reference never updates
    // them); they are rewritten here so the output depends on no earlier
    // contents of the arrays, which the adjoint's step records exchange
    double unused;
    monopole_point(i, j, k, n, GM, zero_bc, unused, ax[t], ay[t], az[t]);
    return;
  }
  size_t s1 = n1, s2 = (size_t)n1 * n1;
  ax[t] = (G)(inv_2h * (phi[t - 1] - phi[t + 1]));
  ay[t] = (G)(inv_2h * (phi[t - s1] - phi[t + s1]));
  az[t] = (G)(inv_2h * (phi[t - s2] - phi[t + s2]));
}

// Plain copy of n doubles (the reference traffic for hardware byte counters).
__global__ void k_copy_probe(const double *__restrict__ src, double *__restrict__ dst, size_t n)
{
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n)
    dst[t] = src[t];
}

// A1 for arbitrary node geometry (host-facing utility).
__global__ void k_box_keys(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                           int n, int lv, int tsx, int tsy, int tsz, int rdx, int rdy, int *__restrict__ out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  double s = (double)(1 << lv);
  double px = x[i] * s, py = y[i] * s, pz = z[i] * s;
  int fx = px < 0 ? -1 : (int)px, fy = py < 0 ? -1 : (int)py, fz = pz < 0 ? -1 : (int)pz;
  out[i] = (fx - tsx) + (fy - tsy) * rdx + (fz - tsz) * rdx * rdy;
}

} // namespace qz
