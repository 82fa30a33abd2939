// Device kernels for refinement nodes ("branches").  All nodes of one level are
// processed by one launch: a 2-D grid whose y index is the node and whose x
// blocks stride over that node's points (or one block per node where an
// ordered compaction is needed).  Deterministic: no float atomics; per-node
// reductions fold fixed per-block partials in a fixed order.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "qz_kernels.h"

namespace qz
{

constexpr double kPiD = 3.14159265358979323846;

// x-blocks per node for the 2-D node launches: enough to cover the largest
// node's point list with ~8 points per thread, at most 64.
inline int blocks_per_node(long long pmax)
{
  long long b = (pmax + 8 * kBlock - 1) / (8 * kBlock);
  return (int)std::min<long long>(64, std::max<long long>(1, b));
}

struct DNode
{
  int lv;     // absolute level
  int parent; // global node index, -1 = head
  int ts[3], rd[3], bmin[3], bmax[3];
  int simc;   // bit a set: node touches the simulation boundary on axis a
  int ncell, coff, zoff;
  int pcap, poff; // point-list slot range (upper bound of points)
  int nintr, nbdry, nsim;
  long long goff, boff; // grid pool and box pool offsets
  double mass;
  int nptcl;
  int done; // CG converged
  double rs[2];
};

struct ParentView
{
  const double *phi; // parent potential (head: head array)
  const void *ax, *ay, *az;
};

__device__ inline int dnode_gX(const DNode &d) { return d.rd[0] + 1; }
__device__ inline int dnode_gXY(const DNode &d) { return (d.rd[0] + 1) * (d.rd[1] + 1); }

// Key range of a cell at relative level lr (1..D) with global coordinates c.
__device__ inline void cell_key_range(int cx, int cy, int cz, int lr, int lmin, int D, unsigned long long &lo,
                                      unsigned long long &hi)
{
  const int n = 1 << lmin;
  unsigned long long k = (unsigned long long)(cx >> lr) +
                         (unsigned long long)n * ((unsigned long long)(cy >> lr) + (unsigned long long)n * (cz >> lr));
  for (int t = lr - 1; t >= 0; t--)
    k = (k << 3) | (unsigned long long)(((cx >> t) & 1) | (((cy >> t) & 1) << 1) | (((cz >> t) & 1) << 2));
  lo = k << (3 * (D - lr));
  hi = (k + 1) << (3 * (D - lr));
}

__device__ inline int lower_bound_u64(const unsigned long long *a, int n, unsigned long long v)
{
  int lo = 0, hi = n;
  while (lo < hi)
  {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Which node of the level owns a slot of a concatenated pool (offsets ascending).
__device__ inline int find_node_ll(const DNode *nodes, int first, int count, long long slot, int which)
{
  int lo = 0, hi = count - 1;
  while (lo < hi)
  {
    int mid = (lo + hi + 1) >> 1;
    long long off = which == 0 ? nodes[first + mid].boff : nodes[first + mid].goff;
    if (off <= slot)
      lo = mid;
    else
      hi = mid - 1;
  }
  return first + lo;
}

// Box status of the level's nodes: -4 everywhere, -5 on simulation-boundary
// slabs (tree_construction.c:1718-1935, non-pbc).
__global__ void k_init_boxes(const DNode *__restrict__ nodes, int first, int count, long long b0, long long nb,
                             int *__restrict__ box, int *__restrict__ bstart, int *__restrict__ bend)
{
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb)
    return;
  long long slot = b0 + s;
  int q = find_node_ll(nodes, first, count, slot, 0);
  const DNode &d = nodes[q];
  long long b = slot - d.boff;
  int bcap = d.rd[0] * d.rd[1] * d.rd[2];
  if (b >= bcap)
    return;
  int i = (int)(b % d.rd[0]), j = (int)((b / d.rd[0]) % d.rd[1]), k = (int)(b / ((long long)d.rd[0] * d.rd[1]));
  int v = -4;
  const int top = (1 << d.lv) - 1;
  const int c[3] = {i, j, k};
  for (int a = 0; a < 3; a++)
  {
    if (!((d.simc >> a) & 1))
      continue;
    int lo_end = d.bmin[a] == 0 ? d.bmin[a] - d.ts[a] : 0;
    int hi_beg = d.bmax[a] == top ? d.bmax[a] + 1 - d.ts[a] : d.rd[a];
    if (c[a] < lo_end || c[a] >= hi_beg)
      v = -5;
  }
  box[slot] = v;
  bstart[slot] = 0;
  bend[slot] = 0;
}

// Child cells (tree_construction.c:1663-1716): cell s of the level is octant
// (s & 7) of zone cell s >> 3; mark it existing (-3) and find its particle
// slot range in the sorted octree keys.  The keys of a cell lie inside its
// parent cell's range (head: head_start, the head cells' first slots; a node:
// the slot range of its box entry), so the searches run there.
__global__ void k_expand_cells(const DNode *__restrict__ nodes, const int *__restrict__ zone_node,
                               const int *__restrict__ zone_xyz, int c0, int ncell, int lmin, int D,
                               const unsigned long long *__restrict__ keys, const int *__restrict__ head_start,
                               int *__restrict__ cell_xyz, int *__restrict__ cell_box, int *__restrict__ cstart,
                               int *__restrict__ cend, int *__restrict__ box, int *bstart, int *bend)
{
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ncell)
    return;
  int cs = c0 + s;
  int z = cs >> 3, o = cs & 7;
  int q = zone_node[z];
  const DNode &d = nodes[q];
  int gx = 2 * zone_xyz[3 * z] + (o & 1), gy = 2 * zone_xyz[3 * z + 1] + ((o >> 1) & 1),
      gz = 2 * zone_xyz[3 * z + 2] + ((o >> 2) & 1);
  int b = (gx - d.ts[0]) + (gy - d.ts[1]) * d.rd[0] + (gz - d.ts[2]) * d.rd[0] * d.rd[1];
  cell_xyz[3 * cs] = gx;
  cell_xyz[3 * cs + 1] = gy;
  cell_xyz[3 * cs + 2] = gz;
  cell_box[cs] = b;
  int pa, pe; // the parent cell's slot range
  const int zx = zone_xyz[3 * z], zy = zone_xyz[3 * z + 1], zz = zone_xyz[3 * z + 2];
  if (d.parent < 0)
  {
    const int n = 1 << lmin;
    const int hc = zx + n * (zy + n * zz);
    pa = head_start[hc];
    pe = head_start[hc + 1];
  }
  else
  {
    const DNode &p = nodes[d.parent];
    const long long pb = p.boff + (zx - p.ts[0]) + (zy - p.ts[1]) * p.rd[0] + (long long)(zz - p.ts[2]) * p.rd[0] * p.rd[1];
    pa = bstart[pb];
    pe = bend[pb];
  }
  unsigned long long lo, hi;
  cell_key_range(gx, gy, gz, d.lv - lmin, lmin, D, lo, hi);
  int a = pa + lower_bound_u64(keys + pa, pe - pa, lo), e = pa + lower_bound_u64(keys + pa, pe - pa, hi);
  cstart[cs] = a;
  cend[cs] = e;
  box[d.boff + b] = -3;
  bstart[d.boff + b] = a;
  bend[d.boff + b] = e;
}

// Per-node particle count and mass (tree_construction.c:2020-2022): x blocks
// of each node stride over its cells; a cell's mass is the sum of its 8 CIC
// corner sums (the weights of a particle add up to 1), read from the deposit
// runs of this level after k_cell_totals.  Per-block partials folded in order.
template <typename S>
__global__ void k_node_mass(const DNode *__restrict__ nodes, int first, const int *__restrict__ cstart,
                            const int *__restrict__ cend, const S *__restrict__ tails, double *__restrict__ part)
{
  const DNode &d = nodes[first + blockIdx.y];
  double mass = 0.0, cnt = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < d.ncell; c += gridDim.x * blockDim.x)
  {
    int a = cstart[d.coff + c], e = cend[d.coff + c];
    cnt += (double)(e - a);
    if (a < e)
    {
      // the cell's 8 corner sums: the tail at its last slot (k_cell_totals)
      const S *w = tails + 8 * (size_t)(e - 1);
      mass += (((double)w[0] + (double)w[1]) + ((double)w[2] + (double)w[3])) +
              (((double)w[4] + (double)w[5]) + ((double)w[6] + (double)w[7]));
    }
  }
  mass = block_reduce<0>(mass);
  cnt = block_reduce<0>(cnt);
  if (threadIdx.x == 0)
  {
    size_t row = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
    part[2 * row] = mass;
    part[2 * row + 1] = cnt;
  }
}

__global__ void k_node_mass_fold(DNode *__restrict__ nodes, int first, const double *__restrict__ part, int nb)
{
  DNode &d = nodes[first + blockIdx.x];
  double mass = 0.0, cnt = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
  {
    size_t row = (size_t)blockIdx.x * nb + b;
    mass += part[2 * row];
    cnt += part[2 * row + 1];
  }
  mass = block_reduce<0>(mass);
  cnt = block_reduce<0>(cnt);
  if (threadIdx.x == 0)
  {
    d.mass = mass;
    d.nptcl = (int)cnt;
  }
}

// Grid-point classification of create_child_nodes (tree_construction.c:2038-2268)
// from the existence pattern of the 8 cells around a point: bit q of the index
// is cell q in the order i0j0k0, im1j0k0, i0jm1k0, im1jm1k0, i0j0km1, im1j0km1,
// i0jm1km1, im1jm1km1.  Table value: bit0 exists, bit1 boundary.
__host__ __device__ inline int classify_pattern(int e)
{
  auto X = [&](int q) { return (e >> q) & 1; };
  auto NX = [&](int q) { return !((e >> q) & 1); };
  int exist = 0, bdry = 0;
  if (X(0))
  {
    exist = 1;
    bdry = (NX(1) && NX(3) && NX(5) && NX(7)) || (NX(2) && NX(3) && NX(6) && NX(7)) ||
           (NX(4) && NX(5) && NX(6) && NX(7));
  }
  else if (X(1))
  {
    exist = 1;
    bdry = (NX(2) && NX(4) && NX(6)) || (NX(2) && NX(3) && NX(6) && NX(7)) || (NX(4) && NX(5) && NX(6) && NX(7));
  }
  else if (X(2))
  {
    exist = 1;
    bdry = (NX(3) && NX(5) && NX(7)) || (NX(4) && NX(5)) || (NX(4) && NX(5) && NX(6) && NX(7));
  }
  else if (X(3))
  {
    exist = 1;
    bdry = (NX(4) && NX(6)) || (NX(4) && NX(5)) || (NX(4) && NX(5) && NX(6) && NX(7));
  }
  else if (X(4) || X(5) || X(6) || X(7))
  {
    exist = 1;
    bdry = 1;
  }
  return exist | (bdry << 1);
}

// Classification of point t (loop order kk, jj, ii over the smallest box grid)
// of a node: -1 no point, 0 interior, 1 boundary, 2 simulation boundary.
__device__ inline int classify_point(const DNode &d, const int *__restrict__ bx, int t, int &gidx, int &gx, int &gy,
                                     int &gz)
{
  const int lx = d.bmax[0] - d.bmin[0] + 2, ly = d.bmax[1] - d.bmin[1] + 2;
  const int rX = d.rd[0], rXY = d.rd[0] * d.rd[1];
  const int ii = d.bmin[0] - d.ts[0] + t % lx;
  const int jj = d.bmin[1] - d.ts[1] + (t / lx) % ly;
  const int kk = d.bmin[2] - d.ts[2] + t / (lx * ly);
  const int b = ii + jj * rX + kk * rXY;
  const int nb[8] = {b, b - 1, b - rX, b - 1 - rX, b - rXY, b - 1 - rXY, b - rX - rXY, b - 1 - rX - rXY};
  int e = 0;
  for (int qq = 0; qq < 8; qq++)
    e |= (bx[nb[qq]] > -4 ? 1 : 0) << qq;
  const int cls = classify_pattern(e);
  if (!(cls & 1))
    return -1;
  const int bd = (cls >> 1) & 1, top = 1 << d.lv;
  gx = ii + d.ts[0];
  gy = jj + d.ts[1];
  gz = kk + d.ts[2];
  gidx = ii + jj * (d.rd[0] + 1) + kk * (d.rd[0] + 1) * (d.rd[1] + 1);
  const bool simb = d.simc != 0 && bd && (gx == 0 || gx == top || gy == 0 || gy == top || gz == 0 || gz == top);
  return simb ? 2 : (bd ? 1 : 0);
}

// The node's points are split into chunks of kClassBlock consecutive points, one
// workgroup each (ctab: node position in the level, first point).  Pass 1
// counts each kind per chunk; an exclusive scan over the chunks of the level
// gives every chunk its offset; pass 2 writes the points, in order, into the
// node's interior / boundary / simulation-boundary slots.
constexpr int kClassBlock = 1024;
__global__ void __launch_bounds__(kClassBlock)
    k_classify_count(const DNode *__restrict__ nodes, int first, const int *__restrict__ box,
                     const int2 *__restrict__ ctab, int nch, int *__restrict__ cnt)
{
  __shared__ int wsum[3][kClassBlock / 64];
  const int2 c = ctab[blockIdx.x];
  const DNode &d = nodes[first + c.x];
  const int total = (d.bmax[0] - d.bmin[0] + 2) * (d.bmax[1] - d.bmin[1] + 2) * (d.bmax[2] - d.bmin[2] + 2);
  const int t = c.y + threadIdx.x;
  int kind = -1, gidx, gx, gy, gz;
  if (t < total)
    kind = classify_point(d, box + d.boff, t, gidx, gx, gy, gz);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = 0; k < 3; k++)
  {
    const unsigned long long m = __ballot(kind == k);
    if (lane == 0)
      wsum[k][wid] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < 3)
  {
    int acc = 0;
    for (int w = 0; w < kClassBlock / 64; w++)
      acc += wsum[threadIdx.x][w];
    cnt[threadIdx.x * nch + blockIdx.x] = acc;
  }
}

__global__ void __launch_bounds__(kClassBlock)
    k_classify_write(const DNode *__restrict__ nodes, int first, const int *__restrict__ box,
                     const int2 *__restrict__ ctab, int nch, const int *__restrict__ off,
                     const int *__restrict__ chunk0, int *__restrict__ intr_g, int *__restrict__ intr_xyz,
                     int *__restrict__ bdry_g, int *__restrict__ bdry_xyz, int *__restrict__ sim_g,
                     int *__restrict__ sim_xyz)
{
  __shared__ int wsum[3][kClassBlock / 64];
  const int2 c = ctab[blockIdx.x];
  const DNode &d = nodes[first + c.x];
  const int total = (d.bmax[0] - d.bmin[0] + 2) * (d.bmax[1] - d.bmin[1] + 2) * (d.bmax[2] - d.bmin[2] + 2);
  const int t = c.y + threadIdx.x;
  int kind = -1, gidx = 0, gx = 0, gy = 0, gz = 0;
  if (t < total)
    kind = classify_point(d, box + d.boff, t, gidx, gx, gy, gz);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int my_off[3];
  for (int k = 0; k < 3; k++)
  {
    const unsigned long long m = __ballot(kind == k);
    my_off[k] = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0)
      wsum[k][wid] = __popcll(m);
  }
  __syncthreads();
  if (kind >= 0)
  {
    int wpre = 0;
    for (int w = 0; w < wid; w++)
      wpre += wsum[kind][w];
    const int f = chunk0[c.x];
    const int slot = d.poff + off[kind * nch + blockIdx.x] - off[kind * nch + f] + wpre + my_off[kind];
    int *g = kind == 0 ? intr_g : (kind == 1 ? bdry_g : sim_g);
    int *xyz = kind == 0 ? intr_xyz : (kind == 1 ? bdry_xyz : sim_xyz);
    g[slot] = gidx;
    xyz[3 * slot] = gx;
    xyz[3 * slot + 1] = gy;
    xyz[3 * slot + 2] = gz;
  }
}

// per node: totals of each kind from the chunk counts and offsets
__global__ void k_classify_totals(DNode *__restrict__ nodes, int first, int count, int nch,
                                  const int *__restrict__ cnt, const int *__restrict__ off,
                                  const int *__restrict__ chunk0)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count)
    return;
  const int f = chunk0[q], l = chunk0[q + 1];
  int tot[3] = {0, 0, 0};
  if (l > f)
    for (int k = 0; k < 3; k++)
      tot[k] = off[k * nch + l - 1] + cnt[k * nch + l - 1] - off[k * nch + f];
  DNode &d = nodes[first + q];
  d.nintr = tot[0];
  d.nbdry = tot[1];
  d.nsim = tot[2];
}

// Refinement criterion counts (tree_construction.c:529): cells of the level
// holding >= crit particles, appended as (node position in level, cell position).
__global__ void k_flag_cells(const DNode *__restrict__ nodes, const int *__restrict__ zone_node, int first, int c0,
                             int ncell, const int *__restrict__ cstart, const int *__restrict__ cend, int crit,
                             int *__restrict__ count, int *__restrict__ out, int cap)
{
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ncell)
    return;
  int cs = c0 + s;
  if (cend[cs] - cstart[cs] >= crit)
  {
    int q = zone_node[cs >> 3];
    int at = atomicAdd(count, 1);
    if (at < cap)
    {
      out[2 * at] = q - first;
      out[2 * at + 1] = cs - nodes[q].coff;
    }
  }
}

// ------------------------------------------------------------------ ordered compaction of byte flags
// The indices i < n with f[i] != 0, ascending.  A block takes kSelItems
// consecutive flags, kSelPer per thread (one 16-byte load where the run is
// whole): k_select_count stores each block's count (and a zero after the last
// block, so that the exclusive scan of the counts ends with the total),
// k_select_write recounts, scans the block's threads and writes the indices
// from the block's first slot.
constexpr int kSelPer = 16;
constexpr int kSelItems = kSelPer * kBlock;

__device__ inline int nonzero_bytes(unsigned w)
{
  return __popc((((w & 0x7f7f7f7fu) + 0x7f7f7f7fu) | w) & 0x80808080u);
}

__device__ inline void load_flags16(const unsigned char *__restrict__ f, long long n, long long i0, unsigned w[4])
{
  if (i0 + kSelPer <= n)
  {
    const uint4 v = *(const uint4 *)(f + i0);
    w[0] = v.x;
    w[1] = v.y;
    w[2] = v.z;
    w[3] = v.w;
    return;
  }
  for (int q = 0; q < 4; q++)
  {
    w[q] = 0;
    for (int k = 0; k < 4; k++)
      if (i0 + 4 * q + k < n)
        w[q] |= (unsigned)f[i0 + 4 * q + k] << (8 * k);
  }
}

__global__ void __launch_bounds__(kBlock) k_select_count(const unsigned char *__restrict__ f, long long n,
                                                         int *__restrict__ bcount)
{
  __shared__ int ws[kBlock / 64];
  const long long i0 = ((long long)blockIdx.x * kBlock + threadIdx.x) * kSelPer;
  unsigned w[4] = {0, 0, 0, 0};
  if (i0 < n)
    load_flags16(f, n, i0, w);
  int c = nonzero_bytes(w[0]) + nonzero_bytes(w[1]) + nonzero_bytes(w[2]) + nonzero_bytes(w[3]);
  for (int off = 32; off > 0; off >>= 1)
    c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0)
    ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    int t = 0;
    for (int q = 0; q < kBlock / 64; q++)
      t += ws[q];
    bcount[blockIdx.x] = t;
    if (blockIdx.x == 0)
      bcount[gridDim.x] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) k_select_write(const unsigned char *__restrict__ f, long long n,
                                                         const int *__restrict__ boff, int *__restrict__ out,
                                                         int *__restrict__ num)
{
  __shared__ int ws[kBlock / 64];
  const long long i0 = ((long long)blockIdx.x * kBlock + threadIdx.x) * kSelPer;
  unsigned w[4] = {0, 0, 0, 0};
  if (i0 < n)
    load_flags16(f, n, i0, w);
  const int c = nonzero_bytes(w[0]) + nonzero_bytes(w[1]) + nonzero_bytes(w[2]) + nonzero_bytes(w[3]);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int incl = c;
  for (int off = 1; off < 64; off <<= 1)
  {
    const int o = __shfl_up(incl, off, 64);
    if (lane >= off)
      incl += o;
  }
  if (lane == 63)
    ws[wv] = incl;
  __syncthreads();
  int pos = boff[blockIdx.x] + incl - c;
  for (int q = 0; q < wv; q++)
    pos += ws[q];
  for (int q = 0; q < 4; q++)
    for (int k = 0; k < 4; k++)
      if ((w[q] >> (8 * k)) & 0xffu)
        out[pos++] = (int)(i0 + 4 * q + k);
  if (blockIdx.x == 0 && threadIdx.x == 0)
    *num = boff[gridDim.x];
}

// ------------------------------------------------------------------ re-sort of near-sorted keys
// After a drift from a sorted state most particles keep their key.  Those that
// keep it ("stayers") are still in key order, and in slot order within a key;
// the others ("movers", flagged by the drift) are radix-sorted by key alone
// (stable, so in slot order within a key) and the two runs are merged by
// (key, old slot).  The result is the permutation of a stable radix sort of all
// keys, exactly.

// Stable partition of the keys by the moved flags: movers to mk/mi at their
// rank among the movers, stayers to sk/si at theirs (slot minus the movers
// before it).  A workgroup takes the kBlock * kSelPer entries that
// k_select_count counted for it (moff: the exclusive scan of those counts), in
// kSelPer rounds of kBlock consecutive entries, so that loads and stores are
// coalesced; ranks within a round come from wavefront ballots.
template <typename K>
__global__ void __launch_bounds__(kBlock) k_part_write(const unsigned char *__restrict__ moved,
                                                       const K *__restrict__ key, long long n,
                                                       const int *__restrict__ moff, K *__restrict__ sk,
                                                       int *__restrict__ si, K *__restrict__ mk,
                                                       int *__restrict__ mi)
{
  __shared__ int wcnt[kBlock / 64];
  const long long base = (long long)blockIdx.x * kBlock * kSelPer;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int mbase = moff[blockIdx.x]; // movers before this round
  for (int q = 0; q < kSelPer; q++)
  {
    const long long i = base + (long long)q * kBlock + threadIdx.x;
    const bool live = i < n;
    const bool mv = live && moved[i] != 0;
    const unsigned long long bm = __ballot(mv);
    const int wrank = __popcll(bm & ((1ull << lane) - 1ull));
    if (lane == 0)
      wcnt[wv] = __popcll(bm);
    __syncthreads();
    int before = 0, tot = 0;
    for (int w = 0; w < kBlock / 64; w++)
    {
      const int c = wcnt[w];
      before += w < wv ? c : 0;
      tot += c;
    }
    __syncthreads(); // wcnt is rewritten next round
    if (live)
    {
      const K kv = key[i];
      const int r = mbase + before + wrank; // movers before entry i
      if (mv)
      {
        mk[r] = kv;
        mi[r] = (int)i;
      }
      else
      {
        sk[i - r] = kv;
        si[i - r] = (int)i;
      }
    }
    mbase += tot;
  }
}

template <typename K> __device__ inline bool key_slot_less(K ka, int ia, K kb, int ib)
{
  return ka < kb || (ka == kb && ia < ib);
}

// Entries of A among the first d of the merge of A[0, na) and B[0, nb), both
// ordered by (key, slot) (merge path: the first a with A[a] not before
// B[d - 1 - a]).
template <typename K>
__device__ inline int merge_split(const K *ak, const int *ai, int na, const K *bk, const int *bi, int nb, int d)
{
  int lo = max(0, d - nb), hi = min(d, na);
  while (lo < hi)
  {
    const int mid = (lo + hi) >> 1;
    if (key_slot_less(ak[mid], ai[mid], bk[d - 1 - mid], bi[d - 1 - mid]))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

constexpr int kMergeThreads = 256, kMergeItems = 8, kMergeTile = kMergeThreads * kMergeItems;

// split[t] = entries of A among the first t kMergeTile outputs, t = 0 .. tiles
template <typename K>
__global__ void k_merge_splits(const K *__restrict__ ak, const int *__restrict__ ai, int na,
                               const K *__restrict__ bk, const int *__restrict__ bi, int nb, int tiles,
                               int *__restrict__ split)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > tiles)
    return;
  split[t] = merge_split(ak, ai, na, bk, bi, nb, min(na + nb, t * kMergeTile));
}

// One tile of kMergeTile outputs per workgroup: its A and B ranges staged in
// LDS, each thread merges kMergeItems consecutive outputs from its own merge
// path split, the results go back through LDS for coalesced stores (keys
// written as KO: 32-bit keys leave widened to 64 bits).
template <typename K, typename KO>
__global__ void __launch_bounds__(kMergeThreads) k_merge(const K *__restrict__ ak, const int *__restrict__ ai, int na,
                                                         const K *__restrict__ bk, const int *__restrict__ bi,
                                                         int nb, const int *__restrict__ split,
                                                         KO *__restrict__ ok, int *__restrict__ oi)
{
  __shared__ K sk[kMergeTile];
  __shared__ int sx[kMergeTile];
  const int n = na + nb;
  const int d0 = blockIdx.x * kMergeTile, d1 = min(n, d0 + kMergeTile);
  const int a0 = split[blockIdx.x], a1 = split[blockIdx.x + 1];
  const int b0 = d0 - a0;
  const int la = a1 - a0, len = d1 - d0, lb = len - la;
  for (int t = threadIdx.x; t < len; t += kMergeThreads)
  {
    if (t < la)
    {
      sk[t] = ak[a0 + t];
      sx[t] = ai[a0 + t];
    }
    else
    {
      sk[t] = bk[b0 + t - la];
      sx[t] = bi[b0 + t - la];
    }
  }
  __syncthreads();
  const int t0 = threadIdx.x * kMergeItems;
  K rk[kMergeItems];
  int rx[kMergeItems];
  int cnt = 0;
  if (t0 < len)
  {
    int i = merge_split(sk, sx, la, sk + la, sx + la, lb, t0);
    int j = t0 - i;
    cnt = min(len - t0, kMergeItems);
#pragma unroll
    for (int q = 0; q < kMergeItems; q++)
      if (q < cnt)
      {
        const bool take_a = j >= lb || (i < la && key_slot_less(sk[i], sx[i], sk[la + j], sx[la + j]));
        const int p = take_a ? i : la + j;
        rk[q] = sk[p];
        rx[q] = sx[p];
        i += take_a ? 1 : 0;
        j += take_a ? 0 : 1;
      }
  }
  __syncthreads(); // every thread has read its inputs from LDS
#pragma unroll
  for (int q = 0; q < kMergeItems; q++)
    if (q < cnt)
    {
      sk[t0 + q] = rk[q];
      sx[t0 + q] = rx[q];
    }
  __syncthreads();
  for (int t = threadIdx.x; t < len; t += kMergeThreads)
  {
    ok[d0 + t] = (KO)sk[t];
    oi[d0 + t] = sx[t];
  }
}

// ------------------------------------------------------------------ head level marks and zones
// fill_cell_ref for the head (tree_construction.c:527-575, n_exp = 1): a head
// cell is marked when some cell of its 3x3x3 block inside the box holds at
// least crit particles.  The block is a product of three intervals, so the mark
// is a flag per cell followed by three one-axis dilations (k_head_dilate).
__global__ void k_head_flag(const int *__restrict__ start, long long ncell, int crit, unsigned char *__restrict__ f)
{
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncell)
    f[c] = start[c + 1] - start[c] >= crit ? 1 : 0;
}

// out = in OR its two neighbours along `axis` (inside the n^3 box of cells,
// n = 2^lg).
__global__ void k_head_dilate(const unsigned char *__restrict__ in, unsigned char *__restrict__ out, int lg,
                              int axis)
{
  const long long n = 1LL << lg, ncell = n * n * n;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell)
    return;
  const int x = (int)((c >> (lg * axis)) & (n - 1));
  const long long st = 1LL << (lg * axis);
  unsigned char m = in[c];
  if (x > 0)
    m |= in[c - st];
  if (x < n - 1)
    m |= in[c + st];
  out[c] = m;
}

// The same marks 16 cells per thread (rows of n >= 16 cells are whole 16-byte
// vectors): k_head_flag_x writes the flags ORed along x (the count of a cell
// is start[c + 1] - start[c]), k_head_dilate_yz ORs the nine x-dilated rows at
// y - 1 .. y + 1, z - 1 .. z + 1 inside the box.
__global__ void k_head_flag_x(const int *__restrict__ start, long long ncell, int lg, int crit, uint4 *__restrict__ fx)
{
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long c0 = v * 16;
  if (c0 >= ncell)
    return;
  const int n = 1 << lg, x0 = (int)(c0 & (n - 1));
  const bool left = x0 > 0, right = x0 + 16 < n;
  int st[19]; // start[c0 - 1 .. c0 + 17]
  for (int k = 0; k < 19; k++)
  {
    const long long c = c0 - 1 + k;
    st[k] = ((k == 0 && !left) || (k == 18 && !right)) ? 0 : start[c];
  }
  unsigned fl[18]; // flags of cells c0 - 1 .. c0 + 16
  for (int k = 0; k < 18; k++)
    fl[k] = st[k + 1] - st[k] >= crit ? 1u : 0u;
  if (!left)
    fl[0] = 0;
  if (!right)
    fl[17] = 0;
  unsigned w[4];
  for (int q = 0; q < 4; q++)
  {
    w[q] = 0;
    for (int k = 0; k < 4; k++)
    {
      const int j = 4 * q + k + 1; // cell c0 + 4 q + k
      w[q] |= (fl[j - 1] | fl[j] | fl[j + 1]) << (8 * k);
    }
  }
  fx[v] = make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ void k_head_dilate_yz(const uint4 *__restrict__ fx, long long nvec, int lg, uint4 *__restrict__ mark)
{
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvec)
    return;
  const int n = 1 << lg;
  const long long c0 = v * 16;
  const int y = (int)((c0 >> lg) & (n - 1)), z = (int)(c0 >> (2 * lg));
  const long long sy = (long long)n / 16, sz = (long long)n * n / 16;
  uint4 m = make_uint4(0, 0, 0, 0);
  for (int dz = -1; dz <= 1; dz++)
  {
    if (z + dz < 0 || z + dz >= n)
      continue;
    for (int dy = -1; dy <= 1; dy++)
    {
      if (y + dy < 0 || y + dy >= n)
        continue;
      const uint4 o = fx[v + dy * sy + dz * sz];
      m.x |= o.x;
      m.y |= o.y;
      m.z |= o.z;
      m.w |= o.w;
    }
  }
  mark[v] = m;
}

// map[cell] = position of the marked cell in the ordered list
__global__ void k_head_index(const int *__restrict__ list, int M, int *__restrict__ map)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M)
    map[list[i]] = i;
}

// Listed cells of a map back to -1 (the map is -1 everywhere else).
__global__ void k_head_unset(const int *__restrict__ list, int M, int *__restrict__ map)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M)
    map[list[i]] = -1;
}

// Union-find with parent[x] <= x at all times: roots link only under smaller
// roots (atomic compare-and-swap on the root), path halving only replaces a
// parent by one of its ancestors, so racing finds and links are harmless and
// the final root of every set is its smallest member.
__device__ inline int uf_find(int *par, int x)
{
  while (true)
  {
    const int p = __atomic_load_n(&par[x], __ATOMIC_RELAXED);
    if (p == x)
      return x;
    const int gp = __atomic_load_n(&par[p], __ATOMIC_RELAXED);
    if (gp != p)
      __atomic_store_n(&par[x], gp, __ATOMIC_RELAXED);
    x = gp;
  }
}

__device__ inline void uf_union(int *par, int a, int b)
{
  while (true)
  {
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b)
      return;
    if (a > b)
    {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicCAS(&par[b], b, a);
    if (old == b)
      return;
    b = old;
  }
}

// 6-connected zones of the marked cells (fill_zones_ref, tree_construction.c:748-891):
// each marked cell joins its marked +x, +y and +z neighbours.  The zone's root is
// its first cell in cell order, which is the cell the reference's flood fill
// starts the zone from; zones are numbered by their roots in cell order.  The
// joins run in two passes (k_zone_hook_local, k_zone_hook_global), below.

// root[i] = the root of entry i; rootflag[i] = 1 for roots, rootflag[M] = 0
// (for the count).  The roots go to their own array: a find's path halving in
// one thread can store a non-root ancestor into par[i] after thread i stored
// the root there, so par is not read for results after the joins; every find
// still returns the true root (parents only ever move to ancestors).
// *unsorted (zeroed before) becomes nonzero when some entry's root is smaller
// than its predecessor's: the zones are then not runs of consecutive entries
// and need the sort by zone.
__global__ void k_head_compress(int M, int *par, int *__restrict__ root, int *__restrict__ rootflag,
                                int *__restrict__ unsorted)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > M)
    return;
  if (i == M)
  {
    rootflag[M] = 0;
    return;
  }
  const int r = uf_find(par, i);
  if (i > 0 && r < uf_find(par, i - 1))
    atomicOr(unsorted, 1);
  root[i] = r;
  rootflag[i] = r == i;
}

// zone[i] = rank of entry i's root; idx[i] = i.  idx may be the root array
// (each thread reads its root before it writes its index).
__global__ void k_head_zone_of(int M, const int *root, const int *__restrict__ rank, unsigned *__restrict__ zone,
                               int *idx)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M)
    return;
  const int r = root[i];
  zone[i] = (unsigned)rank[r];
  idx[i] = i;
}

__global__ void k_bbox_init(int nz, int *__restrict__ bbox)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 6 * nz)
    bbox[t] = (t % 6) < 3 ? 0x7fffffff : -1;
}

// After the stable sort by zone: zone start offsets, bounding boxes (head cell
// coordinates) and the zone-cell arrays the node upload uses (zone order, cell
// order within a zone, as zones_by_ref lists them: tree_construction.c:906-914).
__global__ void k_head_zone_cells(int M, const unsigned *__restrict__ zk, const int *__restrict__ ival,
                                  const int *__restrict__ list, int n, int first, int *__restrict__ zstart,
                                  int *__restrict__ bbox, int *__restrict__ zone_node, int *__restrict__ zone_xyz)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < M;
  const unsigned z = live ? zk[t] : 0xffffffffu;
  int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
  if (live)
  {
    if (t == 0 || zk[t - 1] != z)
      zstart[z] = t;
    const int c = list[ival[t]];
    const int xyz[3] = {c % n, (c / n) % n, c / (n * n)};
    zone_node[t] = first + (int)z;
    for (int a = 0; a < 3; a++)
    {
      zone_xyz[3 * t + a] = xyz[a];
      lo[a] = hi[a] = xyz[a];
    }
  }
  // segmented min/max over the wavefront's runs of equal zone (the list is
  // sorted by zone), then one atomic per run and axis instead of one per cell
  const int lane = threadIdx.x & 63;
  for (int off = 1; off < 64; off <<= 1)
  {
    const unsigned oz = __shfl_down(z, off, 64);
    int olo[3], ohi[3];
    for (int a = 0; a < 3; a++)
    {
      olo[a] = __shfl_down(lo[a], off, 64);
      ohi[a] = __shfl_down(hi[a], off, 64);
    }
    if (lane + off < 64 && oz == z)
      for (int a = 0; a < 3; a++)
      {
        lo[a] = min(lo[a], olo[a]);
        hi[a] = max(hi[a], ohi[a]);
      }
  }
  const unsigned pz = __shfl_up(z, 1, 64);
  if (live && (lane == 0 || pz != z))
    for (int a = 0; a < 3; a++)
    {
      atomicMin(&bbox[6 * z + a], lo[a]);
      atomicMax(&bbox[6 * z + 3 + a], hi[a]);
    }
}

// ------------------------------------------------------------------ node level marks and zones
// fill_cell_ref for the nodes of one level (tree_construction.c:527-575, n_exp = 1):
// a cell of a node is marked when a cell of its 3x3x3 block holds at least crit
// particles.  The flagged cells are first written into a byte map over the
// level's box pool (bflag, zero elsewhere), then every cell ORs its 27 box
// neighbours.  Cells lie at least extra / 2 >= 1 box entries inside their box
// (child box geometry), so the block never leaves it; entries outside the node
// or beyond the simulation box are never flagged.
__global__ void k_level_flag(const DNode *__restrict__ nodes, const int *__restrict__ zone_node, int c0, int ncell,
                             long long b0, const int *__restrict__ cell_box, const int *__restrict__ cstart,
                             const int *__restrict__ cend, int crit, unsigned char *__restrict__ bflag)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ncell)
    return;
  const int cs = c0 + s;
  if (cend[cs] - cstart[cs] >= crit)
    bflag[nodes[zone_node[cs >> 3]].boff - b0 + cell_box[cs]] = 1;
}

__global__ void k_level_mark(const DNode *__restrict__ nodes, const int *__restrict__ zone_node, int c0, int ncell,
                             long long b0, const int *__restrict__ cell_box, const unsigned char *__restrict__ bflag,
                             unsigned char *__restrict__ mark)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ncell)
    return;
  const int cs = c0 + s;
  const DNode &d = nodes[zone_node[cs >> 3]];
  const long long rX = d.rd[0], rXY = (long long)d.rd[0] * d.rd[1];
  const unsigned char *f = bflag + (d.boff - b0 + cell_box[cs]);
  unsigned char m = 0;
  for (int kk = -1; kk <= 1; kk++)
    for (int jj = -1; jj <= 1; jj++)
      for (int ii = -1; ii <= 1; ii++)
        m |= f[ii + jj * rX + kk * rXY];
  mark[s] = m;
}

// map[box entry of marked cell i] = i.  The list holds level cell positions
// (cell c0 + list[i]) in cell order.
__global__ void k_level_index(const int *__restrict__ list, int M, int c0, const DNode *__restrict__ nodes,
                              const int *__restrict__ zone_node, const int *__restrict__ cell_box, long long b0,
                              int *__restrict__ map)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M)
    return;
  const int cs = c0 + list[i];
  map[nodes[zone_node[cs >> 3]].boff - b0 + cell_box[cs]] = i;
}

// 6-connected zones of each node's marked cells (fill_zones_ref,
// tree_construction.c:748-891): a marked cell joins its marked +x, +y and +z
// box neighbours, which belong to the same node.  As for the head, the root of
// a zone is its first cell in cell order and the zones are numbered by their
// roots; nodes are in level order, so zone numbers run over the parents in
// order and over each parent's zones in the order the reference creates them.

// The +x, +y, +z neighbours of marked cell i as list entries (-1: not marked),
// for the head (n > 0: list holds head cell indices, map is the n^3 cell map)
// or for the nodes of a level (n = 0: list holds level cell positions, map
// covers the level's box pool).
struct ZoneNbr
{
  const int *list, *map;
  int n;
  int c0;
  const DNode *nodes;
  const int *zone_node, *cell_box;
  long long b0;
  __device__ void get(int i, int nb[3]) const
  {
    if (n > 0)
    {
      const int c = list[i];
      const int x = c % n, y = (c / n) % n, z = c / (n * n);
      nb[0] = x + 1 < n ? map[c + 1] : -1;
      nb[1] = y + 1 < n ? map[c + n] : -1;
      nb[2] = z + 1 < n ? map[c + n * n] : -1;
      return;
    }
    // node cells lie at least one box entry inside their box (child box geometry)
    const int cs = c0 + list[i];
    const DNode &d = nodes[zone_node[cs >> 3]];
    const long long b = d.boff - b0 + cell_box[cs];
    nb[0] = map[b + 1];
    nb[1] = map[b + d.rd[0]];
    nb[2] = map[b + (long long)d.rd[0] * d.rd[1]];
  }
};

// Two passes.  k_zone_hook_local joins, in LDS, the pairs whose cells both lie
// in the block's kBlock consecutive list entries (for nodes these are 2x2x2
// octant groups of consecutive parent cells, a compact piece of a node; for the
// head a few rows) and stores, as parent and as lroot, the smallest entry of
// each entry's block-local set (so parent <= entry holds).  k_zone_hook_global
// then joins the pairs that cross blocks, by their local roots, once per
// distinct pair of local roots in a block (an LDS hash set): a global pass over
// every pair builds long parent chains on large zones and walks them from
// every cell.
__global__ void __launch_bounds__(kBlock) k_zone_hook_local(ZoneNbr g, int M, int *__restrict__ par,
                                                            int *__restrict__ lroot)
{
  __shared__ int lp[kBlock];
  const int t = threadIdx.x, i0 = blockIdx.x * kBlock, i = i0 + t;
  const int cnt = min(kBlock, M - i0);
  lp[t] = t;
  __syncthreads();
  if (i < M)
  {
    int nb[3];
    g.get(i, nb);
    for (int a = 0; a < 3; a++)
      if (nb[a] >= i0 && nb[a] < i0 + cnt)
        uf_union(lp, t, nb[a] - i0);
  }
  __syncthreads();
  const int r = uf_find(lp, t);
  if (i < M)
  {
    par[i] = i0 + r;
    lroot[i] = i0 + r;
  }
}

constexpr int kZoneHashBits = 10; // > 3 kBlock entries: the set never fills
__global__ void __launch_bounds__(kBlock) k_zone_hook_global(ZoneNbr g, int M, const int *__restrict__ lroot,
                                                             int *par)
{
  constexpr unsigned kSize = 1u << kZoneHashBits;
  constexpr unsigned long long kEmpty = ~0ull;
  __shared__ unsigned long long set[kSize];
  for (unsigned t = threadIdx.x; t < kSize; t += blockDim.x)
    set[t] = kEmpty;
  __syncthreads();
  const int i0 = blockIdx.x * kBlock, i = i0 + threadIdx.x;
  if (i >= M)
    return;
  int nb[3];
  g.get(i, nb);
  const int ri = lroot[i];
  for (int a = 0; a < 3; a++)
  {
    const int j = nb[a];
    if (j < 0 || (j >= i0 && j < i0 + kBlock))
      continue;
    const int rj = lroot[j];
    const unsigned lo = (unsigned)min(ri, rj), hi = (unsigned)max(ri, rj);
    const unsigned long long key = ((unsigned long long)lo << 32) | hi;
    unsigned h = ((lo * 0x9E3779B1u) ^ (hi * 0x85EBCA77u)) >> (32 - kZoneHashBits);
    while (true)
    {
      const unsigned long long old = atomicCAS(&set[h], kEmpty, key);
      if (old == kEmpty)
      {
        uf_union(par, (int)lo, (int)hi);
        break;
      }
      if (old == key)
        break;
      h = (h + 1) & (kSize - 1);
    }
  }
}

// After the stable sort by zone: zone starts, parents (global node index of
// the zone's node), bounding boxes (level cell coordinates), and the zone-cell
// arrays of the next level's upload (zone order, cell order within a zone).
// Each thread takes kZoneRun consecutive entries and folds their bounding box
// while the zone stays the same (a zone ending inside the run is flushed with
// atomics); the runs still open at the threads' last entries are combined over
// the wavefront, so a zone of a million cells costs a few thousand atomics.
constexpr int kZoneRun = 8;
__global__ void k_level_zone_cells(int M, const unsigned *__restrict__ zk, const int *__restrict__ ival,
                                   const int *__restrict__ list, int c0, const int *__restrict__ cell_xyz,
                                   const int *__restrict__ zone_node, int first, int *__restrict__ zstart,
                                   int *__restrict__ zpar, int *__restrict__ bbox, int *__restrict__ zone_node_out,
                                   int *__restrict__ zone_xyz_out)
{
  const long long t0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * kZoneRun;
  unsigned z = 0xffffffffu;
  int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
  for (int k = 0; k < kZoneRun && t0 + k < M; k++)
  {
    const int t = (int)t0 + k;
    const unsigned zt = zk[t];
    const int cs = c0 + list[ival[t]];
    if (t == 0 || zk[t - 1] != zt)
    {
      zstart[zt] = t;
      zpar[zt] = zone_node[cs >> 3];
    }
    if (zt != z)
    {
      if (z != 0xffffffffu)
        for (int a = 0; a < 3; a++)
        {
          atomicMin(&bbox[6 * z + a], lo[a]);
          atomicMax(&bbox[6 * z + 3 + a], hi[a]);
        }
      z = zt;
      for (int a = 0; a < 3; a++)
      {
        lo[a] = 0x7fffffff;
        hi[a] = -1;
      }
    }
    zone_node_out[t] = first + (int)zt;
    for (int a = 0; a < 3; a++)
    {
      const int v = cell_xyz[3 * cs + a];
      zone_xyz_out[3 * t + a] = v;
      lo[a] = min(lo[a], v);
      hi[a] = max(hi[a], v);
    }
  }
  // the open runs, combined over lanes of equal zone (entries are sorted by
  // zone, so equal zones at two lanes hold every lane between them)
  const int lane = threadIdx.x & 63;
  for (int off = 1; off < 64; off <<= 1)
  {
    const unsigned oz = __shfl_down(z, off, 64);
    int olo[3], ohi[3];
    for (int a = 0; a < 3; a++)
    {
      olo[a] = __shfl_down(lo[a], off, 64);
      ohi[a] = __shfl_down(hi[a], off, 64);
    }
    if (lane + off < 64 && oz == z)
      for (int a = 0; a < 3; a++)
      {
        lo[a] = min(lo[a], olo[a]);
        hi[a] = max(hi[a], ohi[a]);
      }
  }
  const unsigned pz = __shfl_up(z, 1, 64);
  if (z != 0xffffffffu && (lane == 0 || pz != z))
    for (int a = 0; a < 3; a++)
    {
      atomicMin(&bbox[6 * z + a], lo[a]);
      atomicMax(&bbox[6 * z + 3 + a], hi[a]);
    }
}

// Refined cells point at their child node: head_child[c] / box[parent] = child index.
__global__ void k_mark_children(const DNode *__restrict__ nodes, const int *__restrict__ zone_node,
                                const int *__restrict__ zone_xyz, int z0, int nz, int n, int *__restrict__ head_child,
                                int *__restrict__ box)
{
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nz)
    return;
  int z = z0 + s;
  int q = zone_node[z];
  int pq = nodes[q].parent;
  int cx = zone_xyz[3 * z], cy = zone_xyz[3 * z + 1], cz = zone_xyz[3 * z + 2];
  if (pq < 0)
    head_child[cx + n * (cy + n * cz)] = q;
  else
  {
    const DNode &p = nodes[pq];
    int b = (cx - p.ts[0]) + (cy - p.ts[1]) * p.rd[0] + (cz - p.ts[2]) * p.rd[0] * p.rd[1];
    box[p.boff + b] = q;
  }
}

// 8-point parent average at ((c >> 1), (c + 1) >> 1) per axis
// (potential_branches.c:204-331, grid_acceleration.c:316-383).
template <typename T>
__device__ inline double parent_avg(const DNode *nodes, int pq, const T *pool, const T *head, int n, int cx, int cy,
                                    int cz)
{
  int a1x = cx >> 1, a1y = cy >> 1, a1z = cz >> 1;
  int a2x = (cx + 1) >> 1, a2y = (cy + 1) >> 1, a2z = (cz + 1) >> 1;
  const T *f;
  long long X, XY;
  int tx, ty, tz;
  if (pq < 0)
  {
    f = head;
    X = n + 1;
    XY = (long long)(n + 1) * (n + 1);
    tx = ty = tz = 0;
  }
  else
  {
    const DNode &p = nodes[pq];
    f = pool + p.goff;
    X = p.rd[0] + 1;
    XY = (long long)(p.rd[0] + 1) * (p.rd[1] + 1);
    tx = p.ts[0];
    ty = p.ts[1];
    tz = p.ts[2];
  }
  long long b1x = a1x - tx, b1y = a1y - ty, b1z = a1z - tz;
  long long b2x = a2x - tx, b2y = a2y - ty, b2z = a2z - tz;
  double s = (double)f[b1x + b1y * X + b1z * XY] + (double)f[b2x + b1y * X + b1z * XY] +
             (double)f[b1x + b2y * X + b1z * XY] + (double)f[b2x + b2y * X + b1z * XY] +
             (double)f[b1x + b1y * X + b2z * XY] + (double)f[b2x + b1y * X + b2z * XY] +
             (double)f[b1x + b2y * X + b2z * XY] + (double)f[b2x + b2y * X + b2z * XY];
  return 0.125 * s;
}

// Simulation-boundary coordinates as the reference computes them: an index of
// 0 or 2^lv maps to 0.0 (potential_branches.c:114-145); for accelerations the
// y coordinate lacks the factor H (grid_acceleration.c:240).
__device__ inline void simb_coords(int lv, int x, int y, int z, bool accel_quirk, double &ax, double &ay, double &az)
{
  double H = 1.0 / (1 << lv);
  int top = 1 << lv;
  ax = (x == 0 || x == top) ? 0.0 : x * H;
  ay = (y == 0 || y == top) ? 0.0 : (accel_quirk ? (double)y : y * H);
  az = (z == 0 || z == top) ? 0.0 : z * H;
}

// A8 initial potential of the level's nodes (potential_branches.c:153-332) and pot_old := pot.
__global__ void k_branch_init_pot(const DNode *__restrict__ nodes, int first, const int *__restrict__ sim_g,
                                  const int *__restrict__ sim_xyz, const int *__restrict__ bdry_g,
                                  const int *__restrict__ bdry_xyz, const int *__restrict__ intr_g,
                                  const int *__restrict__ intr_xyz, double *__restrict__ phi,
                                  double *__restrict__ phi_old, const double *__restrict__ head_phi, int n, double GM,
                                  int zero_bc)
{
  const DNode &d = nodes[first + blockIdx.y];
  double *ph = phi + d.goff;
  double *po = phi_old + d.goff;
  const int tot = d.nsim + d.nbdry + d.nintr;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x)
  {
    if (t < d.nsim)
    {
      int s = d.poff + t;
      double ax, ay, az;
      simb_coords(d.lv, sim_xyz[3 * s], sim_xyz[3 * s + 1], sim_xyz[3 * s + 2], false, ax, ay, az);
      double dist = sqrt((ax - 0.5) * (ax - 0.5) + (ay - 0.5) * (ay - 0.5) + (az - 0.5) * (az - 0.5));
      double v = zero_bc ? 0.0 : -GM / dist;
      ph[sim_g[s]] = v;
      po[sim_g[s]] = v;
    }
    else if (t < d.nsim + d.nbdry)
    {
      int s = d.poff + t - d.nsim;
      double v = parent_avg<double>(nodes, d.parent, phi, head_phi, n, bdry_xyz[3 * s], bdry_xyz[3 * s + 1],
                                    bdry_xyz[3 * s + 2]);
      ph[bdry_g[s]] = v;
      po[bdry_g[s]] = v;
    }
    else
    {
      int s = d.poff + t - d.nsim - d.nbdry;
      double v = parent_avg<double>(nodes, d.parent, phi, head_phi, n, intr_xyz[3 * s], intr_xyz[3 * s + 1],
                                    intr_xyz[3 * s + 2]);
      ph[intr_g[s]] = v;
      po[intr_g[s]] = v;
    }
  }
}

// ---------------------------------------------------------------- batched CG (conjugate_gradient.c:322-463)
// Partial-sum layout: part[(node - first) * gridDim.x + blockIdx.x].
__device__ inline double fold_partials(const double *part, int row, int nb)
{
  __shared__ double s_tot;
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    v += part[row * nb + i];
  v = block_reduce<0>(v);
  if (threadIdx.x == 0)
    s_tot = v;
  __syncthreads();
  double r = s_tot;
  __syncthreads();
  return r;
}

__device__ inline double fold_max(const double *part, int row, int nb)
{
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    v = fmax(v, part[row * nb + i]);
  return block_reduce<1>(v);
}

// r0 = rho - Lap(phi), p = r, partial r.r (conjugate_gradient.c:359-375)
template <typename G>
__global__ void k_cg_init(DNode *__restrict__ nodes, int first, const int *__restrict__ intr_g,
                          const double *__restrict__ phi, const G *__restrict__ rho, double *__restrict__ r,
                          double *__restrict__ p, double *__restrict__ part)
{
  const DNode &d = nodes[first + blockIdx.y];
  const double inv_h2 = (double)(1 << d.lv) * (double)(1 << d.lv);
  const int X = dnode_gX(d), XY = dnode_gXY(d);
  const double *ph = phi + d.goff;
  double acc = 0.0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nintr; t += gridDim.x * blockDim.x)
  {
    int g = intr_g[d.poff + t];
    double v = (double)rho[d.goff + g] +
               inv_h2 * (6.0 * ph[g] - ph[g + 1] - ph[g - 1] - ph[g + X] - ph[g - X] - ph[g + XY] - ph[g - XY]);
    r[d.goff + g] = v;
    p[d.goff + g] = v;
    acc += v * v;
  }
  acc = block_reduce<0>(acc);
  if (threadIdx.x == 0)
    part[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

__global__ void k_cg_init_fold(DNode *__restrict__ nodes, int first, const double *__restrict__ part, int nb)
{
  DNode &d = nodes[first + blockIdx.x];
  double v = fold_partials(part, blockIdx.x, nb);
  if (threadIdx.x == 0)
  {
    d.rs[0] = v;
    d.done = 0;
  }
}

// Ap = Lap(p) on interior points (p = 0 elsewhere), partial p.Ap (:393-400)
__global__ void k_cg_ap(const DNode *__restrict__ nodes, int first, const int *__restrict__ intr_g,
                        const double *__restrict__ p, double *__restrict__ Ap, double *__restrict__ part)
{
  const DNode &d = nodes[first + blockIdx.y];
  if (d.done)
    return;
  const double inv_h2 = (double)(1 << d.lv) * (double)(1 << d.lv);
  const int X = dnode_gX(d), XY = dnode_gXY(d);
  const double *pp = p + d.goff;
  double acc = 0.0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nintr; t += gridDim.x * blockDim.x)
  {
    int g = intr_g[d.poff + t];
    double v = -inv_h2 * (6.0 * pp[g] - pp[g + 1] - pp[g - 1] - pp[g + X] - pp[g - X] - pp[g + XY] - pp[g - XY]);
    Ap[d.goff + g] = v;
    acc += pp[g] * v;
  }
  acc = block_reduce<0>(acc);
  if (threadIdx.x == 0)
    part[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// phi += alpha p, r -= alpha Ap, partial r.r; on check iterations also the
// pointwise change flag against phi_old (poisson_error.c:36-81) (:401-417)
__global__ void k_cg_update(const DNode *__restrict__ nodes, int first, const int *__restrict__ it_base, int it_local,
                            const int *__restrict__ intr_g, double *__restrict__ phi,
                            const double *__restrict__ phi_old, double *__restrict__ r, const double *__restrict__ p,
                            const double *__restrict__ Ap, const double *__restrict__ part_pap,
                            double *__restrict__ part_rs, double *__restrict__ part_viol, double tol2)
{
  const DNode &d = nodes[first + blockIdx.y];
  if (d.done)
    return;
  const int it = *it_base + it_local;
  const int check = (it + 1) % 5 == 0;
  const int nb = gridDim.x;
  double pap = fold_partials(part_pap, blockIdx.y, nb);
  double alpha = d.rs[it & 1] / pap;
  double acc = 0.0, viol = 0.0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nintr; t += gridDim.x * blockDim.x)
  {
    long long g = d.goff + intr_g[d.poff + t];
    double ph = phi[g] + alpha * p[g];
    phi[g] = ph;
    double rr = r[g] - alpha * Ap[g];
    r[g] = rr;
    acc += rr * rr;
    if (check)
    {
      double a = ph / phi_old[g];
      if ((1.0 - tol2 > a) || (1.0 + tol2 < a) || a != a)
        viol = 1.0;
    }
  }
  acc = block_reduce<0>(acc);
  if (check)
    viol = block_reduce<1>(viol);
  if (threadIdx.x == 0)
  {
    part_rs[blockIdx.y * gridDim.x + blockIdx.x] = acc;
    if (check)
      part_viol[blockIdx.y * gridDim.x + blockIdx.x] = viol;
  }
}

// Convergence (conjugate_gradient.c:426-442 with conj_grad_error method 0,
// poisson_error.c:236-288), pot_old refresh on failure, p = r + beta p (:447-455).
__global__ void k_cg_direction(DNode *__restrict__ nodes, int first, const int *__restrict__ it_base, int it_local,
                               const int *__restrict__ intr_g, const double *__restrict__ phi,
                               double *__restrict__ phi_old, const double *__restrict__ r, double *__restrict__ p,
                               const double *__restrict__ part_rs, const double *__restrict__ part_viol,
                               double tol, double G, int *__restrict__ done_count)
{
  DNode &d = nodes[first + blockIdx.y];
  if (d.done)
    return;
  const int it = *it_base + it_local;
  const int check = (it + 1) % 5 == 0;
  const int nb = gridDim.x;
  double rsnew = fold_partials(part_rs, blockIdx.y, nb);
  double rsold = d.rs[it & 1];
  if (check)
  {
    double viol = fold_max(part_viol, blockIdx.y, nb);
    double H = 1.0 / (1 << d.lv);
    double rhomean = 4.0 * G * kPiD * d.mass / ((double)d.ncell * H * H * H);
    double aux_tol = (double)d.nintr * (tol * rhomean) * (tol * rhomean);
    bool ok = viol == 0.0 && rsnew < aux_tol;
    __syncthreads();
    if (ok)
    {
      // every block of the node reaches the same decision; block 0 records it
      if (blockIdx.x == 0 && threadIdx.x == 0)
      {
        d.done = it + 1;
        atomicAdd(done_count, 1);
      }
      return;
    }
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nintr; t += gridDim.x * blockDim.x)
    {
      long long g = d.goff + intr_g[d.poff + t];
      phi_old[g] = phi[g];
    }
  }
  double beta = rsnew / rsold;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nintr; t += gridDim.x * blockDim.x)
  {
    long long g = d.goff + intr_g[d.poff + t];
    p[g] = r[g] + beta * p[g];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    d.rs[(it + 1) & 1] = rsnew;
}

__global__ void k_add_int(int *__restrict__ a, int v)
{
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *a += v;
}

// A11 branch accelerations (grid_acceleration.c:199-394).
template <typename G>
__global__ void k_branch_grid_acc(const DNode *__restrict__ nodes, int first, const int *__restrict__ sim_g,
                                  const int *__restrict__ sim_xyz, const int *__restrict__ bdry_g,
                                  const int *__restrict__ bdry_xyz, const int *__restrict__ intr_g,
                                  const double *__restrict__ phi, G *__restrict__ ax, G *__restrict__ ay,
                                  G *__restrict__ az, const G *__restrict__ hax, const G *__restrict__ hay,
                                  const G *__restrict__ haz, int n, double GM, int zero_bc)
{
  const DNode &d = nodes[first + blockIdx.y];
  const int tot = d.nsim + d.nbdry + d.nintr;
  const double H = 1.0 / (1 << d.lv);
  const double o2H = 1.0 / (2.0 * H);
  const int X = dnode_gX(d), XY = dnode_gXY(d);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x)
  {
    if (t < d.nsim)
    {
      int s = d.poff + t;
      double qx, qy, qz;
      simb_coords(d.lv, sim_xyz[3 * s], sim_xyz[3 * s + 1], sim_xyz[3 * s + 2], true, qx, qy, qz);
      double dist = sqrt((qx - 0.5) * (qx - 0.5) + (qy - 0.5) * (qy - 0.5) + (qz - 0.5) * (qz - 0.5));
      double c2 = zero_bc ? 0.0 : -GM / (dist * dist * dist);
      long long g = d.goff + sim_g[s];
      ax[g] = (G)(c2 * (qx - 0.5));
      ay[g] = (G)(c2 * (qy - 0.5));
      az[g] = (G)(c2 * (qz - 0.5));
    }
    else if (t < d.nsim + d.nbdry)
    {
      int s = d.poff + t - d.nsim;
      int cx = bdry_xyz[3 * s], cy = bdry_xyz[3 * s + 1], cz = bdry_xyz[3 * s + 2];
      long long g = d.goff + bdry_g[s];
      ax[g] = (G)parent_avg<G>(nodes, d.parent, ax, hax, n, cx, cy, cz);
      ay[g] = (G)parent_avg<G>(nodes, d.parent, ay, hay, n, cx, cy, cz);
      az[g] = (G)parent_avg<G>(nodes, d.parent, az, haz, n, cx, cy, cz);
    }
    else
    {
      int s = d.poff + t - d.nsim - d.nbdry;
      int gl = intr_g[s];
      const double *ph = phi + d.goff;
      long long g = d.goff + gl;
      ax[g] = (G)(o2H * (ph[gl - 1] - ph[gl + 1]));
      ay[g] = (G)(o2H * (ph[gl - X] - ph[gl + X]));
      az[g] = (G)(o2H * (ph[gl - XY] - ph[gl + XY]));
    }
  }
}

// A2 branch deposit (grid_density.c:113-267) as a gather over the 8 box cells
// around each listed grid point, in slot order.  coeff = 4 pi G / H^3.
template <typename P, typename G>
__global__ void k_branch_deposit(const DNode *__restrict__ nodes, int first, const int *__restrict__ sim_g,
                                 const int *__restrict__ sim_xyz, const int *__restrict__ bdry_g,
                                 const int *__restrict__ bdry_xyz, const int *__restrict__ intr_g,
                                 const int *__restrict__ intr_xyz, const int *__restrict__ bstart,
                                 const int *__restrict__ bend, const double *__restrict__ x,
                                 const double *__restrict__ y, const double *__restrict__ z, const P *__restrict__ m,
                                 G *__restrict__ rho, double G4pi)
{
  const DNode &d = nodes[first + blockIdx.y];
  const int tot = d.nsim + d.nbdry + d.nintr;
  const double s = (double)(1 << d.lv);
  const double H = 1.0 / s;
  const double coeff = G4pi / (H * H * H);
  const int rX = d.rd[0], rXY = d.rd[0] * d.rd[1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x)
  {
    int sl, gl;
    const int *xyz;
    if (t < d.nsim)
    {
      sl = d.poff + t;
      gl = sim_g[sl];
      xyz = sim_xyz;
    }
    else if (t < d.nsim + d.nbdry)
    {
      sl = d.poff + t - d.nsim;
      gl = bdry_g[sl];
      xyz = bdry_xyz;
    }
    else
    {
      sl = d.poff + t - d.nsim - d.nbdry;
      gl = intr_g[sl];
      xyz = intr_xyz;
    }
    int gx = xyz[3 * sl], gy = xyz[3 * sl + 1], gz = xyz[3 * sl + 2];
    double acc = 0.0;
    for (int ck = gz - 1; ck <= gz; ck++)
      for (int cj = gy - 1; cj <= gy; cj++)
        for (int ci = gx - 1; ci <= gx; ci++)
        {
          long long b = d.boff + (ci - d.ts[0]) + (long long)(cj - d.ts[1]) * rX + (long long)(ck - d.ts[2]) * rXY;
          int a = bstart[b], e = bend[b];
          for (int p = a; p < e; p++)
          {
            double px = x[p] * s, py = y[p] * s, pz = z[p] * s;
            double wx1 = px - ci, wy1 = py - cj, wz1 = pz - ck;
            double wx = ci == gx ? 1.0 - wx1 : wx1;
            double wy = cj == gy ? 1.0 - wy1 : wy1;
            double wz = ck == gz ? 1.0 - wz1 : wz1;
            acc += (double)m[p] * (wx * wy * wz);
          }
        }
    rho[d.goff + gl] = (G)(coeff * acc);
  }
}

// A2 branch deposit, stage 2 (grid_density.c:113-267): each listed grid point
// of the level's nodes adds the corner sums of its 8 box cells.
template <typename S, typename G>
__global__ void k_branch_deposit_fold(const DNode *__restrict__ nodes, int first, const int *__restrict__ sim_g,
                                      const int *__restrict__ sim_xyz, const int *__restrict__ bdry_g,
                                      const int *__restrict__ bdry_xyz, const int *__restrict__ intr_g,
                                      const int *__restrict__ intr_xyz, const int *__restrict__ bstart,
                                      const int *__restrict__ bend, const S *__restrict__ tails, G *__restrict__ rho,
                                      double G4pi)
{
  const DNode &d = nodes[first + blockIdx.y];
  const int tot = d.nsim + d.nbdry + d.nintr;
  const double H = 1.0 / (double)(1 << d.lv);
  const double coeff = G4pi / (H * H * H);
  const int rX = d.rd[0], rXY = d.rd[0] * d.rd[1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x)
  {
    int sl, gl;
    const int *xyz;
    if (t < d.nsim)
    {
      sl = d.poff + t;
      gl = sim_g[sl];
      xyz = sim_xyz;
    }
    else if (t < d.nsim + d.nbdry)
    {
      sl = d.poff + t - d.nsim;
      gl = bdry_g[sl];
      xyz = bdry_xyz;
    }
    else
    {
      sl = d.poff + t - d.nsim - d.nbdry;
      gl = intr_g[sl];
      xyz = intr_xyz;
    }
    int gx = xyz[3 * sl], gy = xyz[3 * sl + 1], gz = xyz[3 * sl + 2];
    // the 8 box cells around the point (all inside the box: node cells keep a
    // padding of empty box cells)
    const long long b0 = d.boff + (gx - 1 - d.ts[0]) + (long long)(gy - 1 - d.ts[1]) * rX +
                         (long long)(gz - 1 - d.ts[2]) * rXY;
    int a[8], e[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
    {
      const long long b = b0 + (u & 1) + ((u >> 1) & 1) * (long long)rX + (u >> 2) * (long long)rXY;
      a[u] = bstart[b];
      e[u] = bend[b];
    }
    rho[d.goff + gl] = (G)(coeff * fold8(tails, a, e));
  }
}

// A12 with refinement (particle_acceleration.c:134-355): each particle takes
// the trilinear acceleration of the finest node holding it.
template <typename P, typename G>
__global__ void k_gather_amr(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                             const G *__restrict__ hax, const G *__restrict__ hay, const G *__restrict__ haz,
                             const int *__restrict__ head_child, const DNode *__restrict__ nodes,
                             const int *__restrict__ box, const G *__restrict__ bax, const G *__restrict__ bay,
                             const G *__restrict__ baz, P *__restrict__ ax, P *__restrict__ ay, P *__restrict__ az,
                             int N, int lmin, unsigned char *__restrict__ plev)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N)
    return;
  const int n = 1 << lmin;
  const double px0 = (double)x[p], py0 = (double)y[p], pz0 = (double)z[p];
  int hx = min((int)(px0 * n), n - 1), hy = min((int)(py0 * n), n - 1), hz = min((int)(pz0 * n), n - 1);
  int q = head_child[hx + n * (hy + n * hz)];
  int lv = lmin;
  while (q >= 0)
  {
    const DNode &d = nodes[q];
    double s = (double)(1 << d.lv);
    int cx = (int)(px0 * s), cy = (int)(py0 * s), cz = (int)(pz0 * s);
    int b = (cx - d.ts[0]) + (cy - d.ts[1]) * d.rd[0] + (cz - d.ts[2]) * d.rd[0] * d.rd[1];
    int st = box[d.boff + b];
    lv = d.lv;
    if (st < 0)
      break;
    q = st;
  }
  double s = (double)(1 << lv);
  double pxs = px0 * s, pys = py0 * s, pzs = pz0 * s;
  int fx = (int)pxs, fy = (int)pys, fz = (int)pzs;
  const G *gx, *gy, *gz;
  long long X, XY, g;
  if (q < 0)
  {
    fx = min(fx, n - 1);
    fy = min(fy, n - 1);
    fz = min(fz, n - 1);
    gx = hax;
    gy = hay;
    gz = haz;
    X = n + 1;
    XY = (long long)(n + 1) * (n + 1);
    g = fx + fy * X + fz * XY;
  }
  else
  {
    const DNode &d = nodes[q];
    gx = bax + d.goff;
    gy = bay + d.goff;
    gz = baz + d.goff;
    X = d.rd[0] + 1;
    XY = (long long)(d.rd[0] + 1) * (d.rd[1] + 1);
    g = (fx - d.ts[0]) + (fy - d.ts[1]) * X + (fz - d.ts[2]) * XY;
  }
  double wx1 = pxs - fx, wy1 = pys - fy, wz1 = pzs - fz;
  double wx2 = 1.0 - wx1, wy2 = 1.0 - wy1, wz2 = 1.0 - wz1;
  double w[8] = {wx2 * wy2 * wz2, wx1 * wy2 * wz2, wx2 * wy1 * wz2, wx1 * wy1 * wz2,
                 wx2 * wy2 * wz1, wx1 * wy2 * wz1, wx2 * wy1 * wz1, wx1 * wy1 * wz1};
  double sx = 0, sy = 0, sz = 0;
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++)
      {
        long long gn = g + ii + jj * X + kk * XY;
        double wq = w[ii + 2 * jj + 4 * kk];
        sx += (double)gx[gn] * wq;
        sy += (double)gy[gn] * wq;
        sz += (double)gz[gn] * wq;
      }
  ax[p] = (P)sx;
  ay[p] = (P)sy;
  az[p] = (P)sz;
  if (plev)
    plev[p] = (unsigned char)lv; // the finest level holding the particle, for the next time step
}

// Finest node holding a position (the walk of k_gather_amr): returns the node
// index (-1 = head) and sets its level.
__device__ inline int finest_node(double px0, double py0, double pz0, const int *head_child, const DNode *nodes,
                                  const int *box, int lmin, int &lv)
{
  const int n = 1 << lmin;
  int hx = min((int)(px0 * n), n - 1), hy = min((int)(py0 * n), n - 1), hz = min((int)(pz0 * n), n - 1);
  int q = head_child ? head_child[hx + n * (hy + n * hz)] : -1;
  lv = lmin;
  int found = -1;
  while (q >= 0)
  {
    const DNode &d = nodes[q];
    double s = (double)(1 << d.lv);
    int cx = (int)(px0 * s), cy = (int)(py0 * s), cz = (int)(pz0 * s);
    int b = (cx - d.ts[0]) + (cy - d.ts[1]) * d.rd[0] + (cz - d.ts[2]) * d.rd[0] * d.rd[1];
    int st = box[d.boff + b];
    lv = d.lv;
    found = q;
    if (st < 0)
      break;
    q = st;
  }
  return found;
}

// Grid potential energy with refinement (observables.c:114-591): partial sums
// of m phi(x), phi interpolated at the finest node holding each particle.
template <typename P>
__global__ void k_grid_pe_amr(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, const P *__restrict__ m, const double *__restrict__ hphi,
                              const int *__restrict__ head_child, const DNode *__restrict__ nodes,
                              const int *__restrict__ box, const double *__restrict__ bphi, int N, int lmin,
                              double *__restrict__ part)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (p < N)
  {
    const int n = 1 << lmin;
    const double px0 = x[p], py0 = y[p], pz0 = z[p];
    int lv = lmin;
    const int q = finest_node(px0, py0, pz0, head_child, nodes, box, lmin, lv);
    const double s = (double)(1 << lv);
    const double pxs = px0 * s, pys = py0 * s, pzs = pz0 * s;
    int fx = (int)pxs, fy = (int)pys, fz = (int)pzs;
    const double *f;
    long long X, XY, g;
    if (q < 0)
    {
      fx = min(fx, n - 1);
      fy = min(fy, n - 1);
      fz = min(fz, n - 1);
      f = hphi;
      X = n + 1;
      XY = X * X;
      g = fx + fy * X + fz * XY;
    }
    else
    {
      const DNode &d = nodes[q];
      f = bphi + d.goff;
      X = d.rd[0] + 1;
      XY = X * (d.rd[1] + 1);
      g = (fx - d.ts[0]) + (fy - d.ts[1]) * X + (fz - d.ts[2]) * XY;
    }
    const double wx1 = pxs - fx, wy1 = pys - fy, wz1 = pzs - fz;
    const double wx[2] = {1.0 - wx1, wx1}, wy[2] = {1.0 - wy1, wy1}, wz[2] = {1.0 - wz1, wz1};
    for (int kk = 0; kk < 2; kk++)
      for (int jj = 0; jj < 2; jj++)
        for (int ii = 0; ii < 2; ii++)
          acc += f[g + ii + jj * X + kk * XY] * wx[ii] * wy[jj] * wz[kk];
    acc *= (double)m[p];
  }
  acc = block_reduce<0>(acc);
  if (threadIdx.x == 0)
    part[blockIdx.x] = acc;
}


// CFL time step with velocity and acceleration (timestep.c:73-195, 370-587):
// per axis the dt that moves a particle CFL * H of its finest node, min over
// particles and axes (cfl_axis, qz_kernels.h); per-block minima for a
// deterministic second stage.
template <typename P>
__global__ void k_cfl_dt(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                         const P *__restrict__ vx, const P *__restrict__ vy, const P *__restrict__ vz,
                         const P *__restrict__ ax, const P *__restrict__ ay, const P *__restrict__ az, int N,
                         const int *__restrict__ head_child, const DNode *__restrict__ nodes,
                         const int *__restrict__ box, int lmin, double cfl, double dt_max, double *__restrict__ part)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  double dt = dt_max;
  if (p < N)
  {
    int lv;
    finest_node((double)x[p], (double)y[p], (double)z[p], head_child, nodes, box, lmin, lv);
    double H = 1.0 / (1 << lv);
    dt = cfl_axis((double)vx[p], (double)ax[p], H, cfl, dt);
    dt = cfl_axis((double)vy[p], (double)ay[p], H, cfl, dt);
    dt = cfl_axis((double)vz[p], (double)az[p], H, cfl, dt);
  }
  // min through the max reduction of -dt
  double r = -block_reduce<1>(-dt);
  if (threadIdx.x == 0)
    part[blockIdx.x] = r;
}

} // namespace qz
