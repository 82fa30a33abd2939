// Refinement-node (branch) stages of Engine<P,G>.  Included by qz_core.hip.
//
// Tree build per step (the reference adapts its tree in tree_adaptation.c; this
// rebuilds it the way tree_construction.c:2521-2577 builds it):
//   device: refinement counts per cell from the sorted octree keys, dilated
//           marks, 6-connected zones (union-find), zone bounding boxes
//   host:   child box geometry from each zone's bounding box (qz_tree.cpp)
//   device: child cells, box status, particle ranges, masses, grid-point lists
// Per level only the zones' parents, bounding boxes and sizes cross to the
// host, and only node descriptors cross back.  With QZ_HOST_TREE=1 the marks
// and zones below the head are found on the host instead (qz_tree.cpp), from
// the compact list of flagged cells.
#include "qz_branch_kernels.h"
#include "qz_branch_mg_kernels.h"
#include "qz_tree.h"

namespace qz
{

// The node multigrid set-up of one tree level (bmg_prepare), kept for the
// solve of that level, which uses it once.
struct BmgPlan
{
  int level = -1;       // tree level the set-up belongs to, -1 when none is held
  bool cg_only = false; // a box too large for the hierarchy: CG solves the level
  bool ahead = false;   // built on the set-up stream (the solve waits for ahead_done)
  bool nfuse = false;
  int n_itab = 0, mmax = 0;
  size_t tail_lds = 0;
  int sw0[kMaxMgLevels][2], swn[kMaxMgLevels][2], rs0[kMaxMgLevels], rsn[kMaxMgLevels];
  int nt0[kMaxMgLevels], ntn[kMaxMgLevels];
  double act_s[kMaxMgLevels], act_r[kMaxMgLevels];
  double act_f[kMaxMgLevels]; // active points of the levels run by the fused tile kernels
  double tail_act = 0.0;      // active points of the nodes' tail top levels
  hipGraphExec_t gexec = nullptr; // the PCG batch captured with the set-up (run ahead only)
};

struct TreeDev
{
  DNode *nodes = nullptr;
  int node_cap = 0;
  int *zone_node = nullptr, *zone_xyz = nullptr;
  long long zone_cap = 0;
  int *cell_xyz = nullptr, *cell_box = nullptr, *cstart = nullptr, *cend = nullptr;
  long long cell_cap = 0;
  int *box = nullptr, *bstart = nullptr, *bend = nullptr;
  long long box_cap = 0;
  int *intr_g = nullptr, *intr_xyz = nullptr, *bdry_g = nullptr, *bdry_xyz = nullptr, *sim_g = nullptr,
      *sim_xyz = nullptr;
  long long pt_cap = 0;
  double *phi = nullptr, *phi_old = nullptr, *cg_r = nullptr, *cg_p = nullptr, *cg_Ap = nullptr;
  void *rho = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr;
  long long grid_cap = 0;
  size_t gsize = 0; // sizeof(G)
  int *head_child = nullptr;
  // head cells whose head_child entry the last build set, listed in hlist[0,
  // hc_set) (every other entry is -1); -1 when unknown (new map, restored tree)
  int hc_set = -1;
  int *flags = nullptr, *flag_count = nullptr;
  int flag_cap = 0;
  // head level marks and zones (device): mark per head cell, ordered list of
  // marked cells, union-find and zone arrays sized by the marked count
  unsigned char *hmark = nullptr;
  int *hlist = nullptr;
  // ordered compaction of byte flags (select_flags): per-block counts, their
  // exclusive scan and the scan's scratch
  int *sel_cnt = nullptr, *sel_off = nullptr;
  long long sel_cap = 0;
  void *sel_scan = nullptr;
  size_t sel_scan_bytes = 0;
  int *hpar = nullptr, *hflag = nullptr, *hrank = nullptr, *hidx = nullptr, *hidx_s = nullptr, *hz_node = nullptr,
      *hz_xyz = nullptr, *hzstart = nullptr, *hbbox = nullptr;
  unsigned *hzone = nullptr, *hzone_s = nullptr;
  int *hzpar = nullptr; // levels below the head: the parent of each zone
  long long hcap = 0;
  // levels below the head: flagged-cell map over the level's box pool, marks
  // and ordered list of the level's marked cells, box-entry -> list map
  unsigned char *lbflag = nullptr, *lmark = nullptr;
  int *lbmap = nullptr, *llist = nullptr;
  long long lbox_cap = 0, lcell_cap = 0;
  void *htmp = nullptr;
  size_t htmp_bytes = 0;
  double *part1 = nullptr, *part2 = nullptr, *part3 = nullptr;
  long long part_cap = 0;
  double *mpart = nullptr;
  long long mpart_cap = 0;
  int *done_count = nullptr;
  int *hflag_count = nullptr; // pinned, 4 ints
  int *hflags = nullptr;      // pinned mirror of flags (2 flag_cap ints): the refinement flags read by the host
  hipEvent_t done_ev[2] = {nullptr, nullptr}; // pipelined batch loop
  std::vector<DNode> hnodes;
  // per relative level (index 1..): node range and pool ranges
  std::vector<int> lfirst, lcount;
  std::vector<long long> lc0, lncell, lb0, lnb, lz0, lnz, lg0, lng, lpmax;
  std::vector<int> lnbm; // x blocks of the node-mass reduction per level
  std::vector<long long> lnintr; // interior points per level
  // batched multigrid: per-node hierarchy table and level pools (shared by all tree levels)
  std::vector<BMg> hbmg;
  BMg *bmg = nullptr;
  int bmg_cap = 0;
  void *mg_e = nullptr, *mg_r = nullptr;
  void *mg_t = nullptr; // fused node levels: the pre-smoothed correction (k_bmg_down -> k_bmg_up)
  NTile *ntiles = nullptr; // fused node levels: tile table
  long long ntile_cap = 0;
  std::vector<NTile> hntiles;
  BmgPlan plan;
  void drop_plan()
  {
    if (plan.gexec)
      (void)hipGraphExecDestroy(plan.gexec);
    plan.gexec = nullptr;
    plan.level = -1;
  }
  // The first tree level's set-up runs ahead on a host thread and its own
  // stream (after ahead_tree, the end of the tree build) while the head density
  // and potential are computed.
  std::thread ahead;
  bool ahead_ok = true;
  std::string ahead_err; // the set-up thread's error (errors are kept per thread)
  hipStream_t ahead_stream = nullptr;
  hipEvent_t ahead_tree = nullptr, ahead_done = nullptr;
  hipEvent_t acc_go = nullptr, acc_done = nullptr; // head accelerations on the set-up stream
  unsigned char *mg_mask = nullptr;
  long long mg_cap = 0;
  // grid-point classification: chunk table, per-chunk counts and offsets
  std::vector<int2> hctab;
  std::vector<int> hchunk0;
  int2 *ctab = nullptr;
  int *chunk0 = nullptr, *ccnt = nullptr, *coff = nullptr;
  long long ctab_cap = 0, chunk0_cap = 0;
  void *cscan_tmp = nullptr;
  size_t cscan_bytes = 0;
  std::vector<Chunk> hchunks; // work tables of the level being solved
  Chunk *chunks = nullptr;
  long long chunk_cap = 0;
  int *alist = nullptr, *acount = nullptr; // active points (pool indices) and their count
  long long alist_cap = 0;
  std::vector<int> hrbeg; // region begins in the pools
  int *rbeg = nullptr, *rstart = nullptr;
  long long reg_cap = 0;
  long long used_zone = 0, used_cell = 0, used_box = 0, used_pt = 0, used_grid = 0;
};

// A copy of everything the adjoint of a force pass reads from the tree: the node
// table, head children, box statuses and slot ranges, cell ranges, grid-point
// lists, node grid accelerations, and the host-side level tables.  Device parts
// are grow-only buffers, so a record is refilled without reallocation.
struct TreeSnap
{
  struct Buf
  {
    void *p = nullptr;
    size_t cap = 0;
  };
  Buf nodes, head_child, box, bstart, bend, cstart, cend, intr_g, bdry_g, sim_g, intr_xyz, bdry_xyz, sim_xyz, ax, ay, az;
  int n_nodes = 0;
  long long used_box = 0, used_cell = 0, used_pt = 0, used_grid = 0;
  std::vector<DNode> hnodes;
  std::vector<int> lfirst, lcount, lnbm;
  std::vector<long long> lc0, lncell, lb0, lnb, lz0, lnz, lg0, lng, lpmax, lnintr;
  void release()
  {
    Buf *all[] = {&nodes, &head_child, &box, &bstart, &bend, &cstart, &cend, &intr_g,
                  &bdry_g, &sim_g, &intr_xyz, &bdry_xyz, &sim_xyz, &ax, &ay, &az};
    for (Buf *b : all)
    {
      (void)hipFree(b->p);
      b->p = nullptr;
      b->cap = 0;
    }
  }
};

// Grow a device buffer keeping its first `used` entries.  The copy runs on the
// stream that writes the buffer (the tree may be built on a non-blocking
// stream, which a null-stream copy would not wait for), and that stream is
// drained before the old buffer is freed.
template <typename T> static bool regrow(T *&p, long long newcap, long long used, const char *what, hipStream_t s)
{
  T *np = nullptr;
  if (!hip_ok(hipMalloc((void **)&np, sizeof(T) * (size_t)std::max<long long>(newcap, 1)), what))
    return false;
  if (p && used > 0 &&
      !hip_ok(hipMemcpyAsync(np, p, sizeof(T) * (size_t)used, hipMemcpyDeviceToDevice, s), what))
    return false;
  if (!hip_ok(hipStreamSynchronize(s), what))
    return false;
  (void)hipFree(p);
  p = np;
  return true;
}

static long long next_cap(long long cap, long long need) { return need <= cap ? cap : std::max(need, cap + cap / 2 + 1024); }

// Grow-only buffers of the grid-point classification.
static bool reserve_classify(TreeDev &d, int nch, int count, hipStream_t s)
{
  if ((nch > d.ctab_cap || count + 1 > d.chunk0_cap) && !hip_ok(hipStreamSynchronize(s), "classify"))
    return false; // the buffers about to be freed may still be read by the stream
  if (nch > d.ctab_cap)
  {
    (void)hipFree(d.ctab);
    (void)hipFree(d.ccnt);
    (void)hipFree(d.coff);
    (void)hipFree(d.cscan_tmp);
    d.ctab = nullptr;
    d.ccnt = d.coff = nullptr;
    d.cscan_tmp = nullptr;
    d.ctab_cap = next_cap(d.ctab_cap, nch);
    d.cscan_bytes = qz_scan_temp_bytes((int)d.ctab_cap);
    if (!hip_ok(hipMalloc((void **)&d.ctab, sizeof(int2) * d.ctab_cap), "classify") ||
        !hip_ok(hipMalloc((void **)&d.ccnt, 3 * sizeof(int) * d.ctab_cap), "classify") ||
        !hip_ok(hipMalloc((void **)&d.coff, 3 * sizeof(int) * d.ctab_cap), "classify") ||
        !hip_ok(hipMalloc(&d.cscan_tmp, std::max<size_t>(d.cscan_bytes, 1)), "classify"))
      return false;
  }
  if (count + 1 > d.chunk0_cap)
  {
    (void)hipFree(d.chunk0);
    d.chunk0 = nullptr;
    d.chunk0_cap = next_cap(d.chunk0_cap, count + 1);
    if (!hip_ok(hipMalloc((void **)&d.chunk0, sizeof(int) * d.chunk0_cap), "classify"))
      return false;
  }
  return true;
}

template <typename P, typename G> void Engine<P, G>::release_tree()
{
  if (!tree)
    return;
  TreeDev *d = tree->dev;
  if (d)
  {
    if (d->ahead.joinable())
      d->ahead.join();
    d->drop_plan();
    if (d->ahead_stream)
      (void)hipStreamDestroy(d->ahead_stream);
    for (hipEvent_t e : {d->ahead_tree, d->ahead_done, d->acc_go, d->acc_done})
      if (e)
        (void)hipEventDestroy(e);
    void *ptrs[] = {d->nodes, d->zone_node, d->zone_xyz, d->cell_xyz, d->cell_box, d->cstart, d->cend, d->box,
                    d->bstart, d->bend, d->intr_g, d->intr_xyz, d->bdry_g, d->bdry_xyz, d->sim_g, d->sim_xyz,
                    d->phi, d->phi_old, d->cg_r, d->cg_p, d->cg_Ap, d->rho, d->ax,
                    d->ay, d->az, d->head_child, d->flags, d->flag_count, d->part1, d->part2, d->part3,
                    d->done_count, d->mpart, d->bmg, d->mg_e, d->mg_r, d->mg_t, d->ntiles, d->mg_mask, d->hmark, d->hlist,
                    d->sel_cnt, d->sel_off, d->sel_scan, d->hpar, d->hflag, d->hrank, d->hidx, d->hidx_s, d->hz_node, d->hz_xyz,
                    d->hzstart, d->hbbox, d->hzone, d->hzone_s, d->htmp, d->hzpar, d->lbflag, d->lmark,
                    d->lbmap, d->llist, d->chunks, d->alist, d->acount, d->rbeg, d->rstart, d->ctab, d->chunk0, d->ccnt,
                    d->coff, d->cscan_tmp};
    for (void *q : ptrs)
      (void)hipFree(q);
    if (d->hflag_count)
      (void)hipHostFree(d->hflag_count);
    if (d->hflags)
      (void)hipHostFree(d->hflags);
    for (hipEvent_t e : d->done_ev)
      if (e)
        (void)hipEventDestroy(e);
    delete d;
  }
  delete tree->host;
  delete tree;
  tree = nullptr;
}

// Reserve pools for the level about to be uploaded (grow-only, contents kept).
template <typename G> static bool reserve_level(TreeDev *d, hipStream_t s, long long nz, long long ncell, long long nb, long long npt,
                                                long long ng)
{
  long long c = next_cap(d->zone_cap, d->used_zone + nz);
  if (c != d->zone_cap)
  {
    if (!regrow(d->zone_node, c, d->used_zone, "zone pool", s) || !regrow(d->zone_xyz, 3 * c, 3 * d->used_zone, "zone pool", s))
      return false;
    d->zone_cap = c;
  }
  c = next_cap(d->cell_cap, d->used_cell + ncell);
  if (c != d->cell_cap)
  {
    long long u = d->used_cell;
    if (!regrow(d->cell_box, c, u, "cell pool", s) || !regrow(d->cstart, c, u, "cell pool", s) ||
        !regrow(d->cend, c, u, "cell pool", s) || !regrow(d->cell_xyz, 3 * c, 3 * u, "cell pool", s))
      return false;
    d->cell_cap = c;
  }
  c = next_cap(d->box_cap, d->used_box + nb);
  if (c != d->box_cap)
  {
    long long u = d->used_box;
    if (!regrow(d->box, c, u, "box pool", s) || !regrow(d->bstart, c, u, "box pool", s) || !regrow(d->bend, c, u, "box pool", s))
      return false;
    d->box_cap = c;
  }
  c = next_cap(d->pt_cap, d->used_pt + npt);
  if (c != d->pt_cap)
  {
    long long u = d->used_pt;
    if (!regrow(d->intr_g, c, u, "point pool", s) || !regrow(d->bdry_g, c, u, "point pool", s) ||
        !regrow(d->sim_g, c, u, "point pool", s) ||
        !regrow(d->intr_xyz, 3 * c, 3 * u, "point pool", s) || !regrow(d->bdry_xyz, 3 * c, 3 * u, "point pool", s) ||
        !regrow(d->sim_xyz, 3 * c, 3 * u, "point pool", s))
      return false;
    d->pt_cap = c;
  }
  c = next_cap(d->grid_cap, d->used_grid + ng);
  if (c != d->grid_cap)
  {
    long long u = d->used_grid;
    G *rho = (G *)d->rho, *ax = (G *)d->ax, *ay = (G *)d->ay, *az = (G *)d->az;
    if (!regrow(d->phi, c, u, "grid pool", s) || !regrow(d->phi_old, c, u, "grid pool", s) ||
        !regrow(d->cg_r, c, u, "grid pool", s) || !regrow(d->cg_p, c, u, "grid pool", s) ||
        !regrow(d->cg_Ap, c, u, "grid pool", s) || !regrow(rho, c, u, "grid pool", s) || !regrow(ax, c, u, "grid pool", s) ||
        !regrow(ay, c, u, "grid pool", s) || !regrow(az, c, u, "grid pool", s))
      return false;
    d->rho = rho;
    d->ax = ax;
    d->ay = ay;
    d->az = az;
    d->grid_cap = c;
  }
  return true;
}

// The indices i < n with f[i] != 0, ascending, into out, and their count into
// *num (device), on `stream` (k_select_count, a scan of the block counts,
// k_select_write).
static bool select_flags(TreeDev &d, const unsigned char *f, long long n, int *out, int *num, hipStream_t stream)
{
  const long long nb = std::max<long long>(1, (n + kSelItems - 1) / kSelItems);
  if (nb + 1 > d.sel_cap)
  {
    if (!hip_ok(hipStreamSynchronize(stream), "selection"))
      return false; // the buffers about to be replaced may still be read by the stream
    (void)hipFree(d.sel_cnt);
    (void)hipFree(d.sel_off);
    (void)hipFree(d.sel_scan);
    d.sel_cnt = d.sel_off = nullptr;
    d.sel_scan = nullptr;
    d.sel_cap = 0;
    const long long cap = next_cap(0, nb + 1);
    d.sel_scan_bytes = std::max<size_t>(qz_scan_temp_bytes((int)cap), 1);
    if (!hip_ok(hipMalloc((void **)&d.sel_cnt, sizeof(int) * (size_t)cap), "selection") ||
        !hip_ok(hipMalloc((void **)&d.sel_off, sizeof(int) * (size_t)cap), "selection") ||
        !hip_ok(hipMalloc(&d.sel_scan, d.sel_scan_bytes), "selection"))
      return false;
    d.sel_cap = cap;
  }
  hipLaunchKernelGGL(k_select_count, dim3((unsigned)nb), dim3(kBlock), 0, stream, f, n, d.sel_cnt);
  if (!hip_ok(qz_exclusive_scan(d.sel_scan, d.sel_scan_bytes, d.sel_cnt, d.sel_off, (int)nb + 1, stream), "selection"))
    return false;
  hipLaunchKernelGGL(k_select_write, dim3((unsigned)nb), dim3(kBlock), 0, stream, f, n, (const int *)d.sel_off, out,
                     num);
  return hip_ok(hipGetLastError(), "selection");
}

// Re-sort after a drift from a sorted state (the moved flags current): the
// unsorted keys (narrow: key_out as 32 bits, else key_in) partitioned into
// stayers and movers, the movers radix-sorted, then merged by (key, slot) into
// the sorted 64-bit keys key_out and slot_out, the permutation a stable radix
// sort of all keys gives.  One host read of the mover count, queued behind the
// partition so that the device works while the host waits.
template <typename P, typename G> bool Engine<P, G>::resort_keys()
{
  const bool narrow = end_bit <= 32;
  const size_t ks = narrow ? sizeof(unsigned) : sizeof(unsigned long long);
  const long long per = (long long)kBlock * kSelPer;
  const int nb = (int)((N + per - 1) / per);
  const int tiles = (N + kMergeTile - 1) / kMergeTile;
  if (!rs_cnt)
  {
    size_t b = 0;
    rs_cnt = dmalloc<int>(nb + 1, b);
    rs_off = dmalloc<int>(nb + 1, b);
    rs_split = dmalloc<int>(tiles + 1, b);
    for (int q = 0; q < 3; q++)
    {
      rs_idx[q] = dmalloc<int>(N, b);
      rs_key[q] = dmalloc<char>(ks * (size_t)N, b);
    }
    rs_scan_bytes = std::max<size_t>(qz_scan_temp_bytes(nb + 1), 1);
    rs_scan = dmalloc<char>(rs_scan_bytes, b);
    bytes += b;
    if (!rs_cnt || !rs_off || !rs_split || !rs_idx[0] || !rs_idx[1] || !rs_idx[2] || !rs_key[0] || !rs_key[1] ||
        !rs_key[2] || !rs_scan || !hip_ok(hipHostMalloc((void **)&rs_hcount, sizeof(int)), "re-sort count"))
      return hip_ok(hipErrorOutOfMemory, "re-sort buffers");
  }
  using U = unsigned long long;
  hipLaunchKernelGGL(k_select_count, dim3(nb), dim3(kBlock), 0, stream, moved, (long long)N, rs_cnt);
  if (!hip_ok(qz_exclusive_scan(rs_scan, rs_scan_bytes, rs_cnt, rs_off, nb + 1, stream), "re-sort scan") ||
      !hip_ok(hipMemcpyAsync(rs_hcount, rs_off + nb, sizeof(int), hipMemcpyDeviceToHost, stream), "re-sort count"))
    return false;
  if (narrow)
    hipLaunchKernelGGL(k_part_write<unsigned>, dim3(nb), dim3(kBlock), 0, stream, moved, (const unsigned *)key_out,
                       (long long)N, rs_off, (unsigned *)rs_key[0], rs_idx[0], (unsigned *)rs_key[1], rs_idx[1]);
  else
    hipLaunchKernelGGL(k_part_write<U>, dim3(nb), dim3(kBlock), 0, stream, moved, (const U *)key_in, (long long)N,
                       rs_off, (U *)rs_key[0], rs_idx[0], (U *)rs_key[1], rs_idx[1]);
  if (!hip_ok(hipStreamSynchronize(stream), "re-sort count"))
    return false;
  const int M = *rs_hcount, S = N - M; // movers, stayers
  rs_movers = M;
  if (narrow)
  {
    const unsigned *sk = (const unsigned *)rs_key[0], *mk = (const unsigned *)rs_key[1];
    unsigned *mks = (unsigned *)rs_key[2];
    if (M > 0 && !hip_ok(qz_sort_pairs(sort_tmp, sort_tmp_bytes, mk, mks, rs_idx[1], rs_idx[2], M, end_bit, stream),
                         "re-sort movers"))
      return false;
    hipLaunchKernelGGL(k_merge_splits<unsigned>, dim3(nblk(tiles + 1)), dim3(kBlock), 0, stream, sk, rs_idx[0], S,
                       (const unsigned *)mks, rs_idx[2], M, tiles, rs_split);
    hipLaunchKernelGGL((k_merge<unsigned, U>), dim3(tiles), dim3(kMergeThreads), 0, stream, sk, rs_idx[0], S,
                       (const unsigned *)mks, rs_idx[2], M, rs_split, key_out, slot_out);
  }
  else
  {
    const U *sk = (const U *)rs_key[0], *mk = (const U *)rs_key[1];
    U *mks = (U *)rs_key[2];
    if (M > 0 && !hip_ok(qz_sort64_pairs(sort_tmp, sort_tmp_bytes, mk, mks, rs_idx[1], rs_idx[2], M, end_bit, stream),
                         "re-sort movers"))
      return false;
    hipLaunchKernelGGL(k_merge_splits<U>, dim3(nblk(tiles + 1)), dim3(kBlock), 0, stream, sk, rs_idx[0], S,
                       (const U *)mks, rs_idx[2], M, tiles, rs_split);
    hipLaunchKernelGGL((k_merge<U, U>), dim3(tiles), dim3(kMergeThreads), 0, stream, sk, rs_idx[0], S, (const U *)mks,
                       rs_idx[2], M, rs_split, key_out, slot_out);
  }
  return hip_ok(hipGetLastError(), "re-sort");
}

// Zone of every marked cell (the rank of its root) into d.hzone_s and the list
// entries in zone order into d.hidx_s: a stable sort by zone, or, when the
// zones already are runs of consecutive entries (`unsorted` false: every root
// at least its predecessor's), the entries as they are.
template <typename P, typename G> bool Engine<P, G>::zones_in_order(int M, int nz, bool unsorted, hipStream_t s)
{
  hipStream_t stream = s;
  TreeDev &d = *tree->dev;
  if (!unsorted)
  {
    QZ_LAUNCHB("tree_zone_order", 16.0 * M, k_head_zone_of, nblk(M), M, (const int *)d.hidx, d.hrank, d.hzone_s,
               d.hidx_s);
    return true;
  }
  int end_bit = 1;
  while ((1LL << end_bit) < nz)
    end_bit++;
  QZ_LAUNCHB("tree_zone_order", 16.0 * M, k_head_zone_of, nblk(M), M, (const int *)d.hidx, d.hrank, d.hzone,
             d.hidx);
  timer.begin(stream);
  if (!hip_ok(qz_sort_pairs(d.htmp, d.htmp_bytes, d.hzone, d.hzone_s, d.hidx, d.hidx_s, M, end_bit, stream), "zones"))
    return false;
  timer.end(stream, "tree_zone_sort", 16.0 * M);
  return true;
}

// The union-find and zone arrays of the marked cells of one level (head or
// nodes), for `need` entries (grow-only, contents not kept).
static bool reserve_zone_scratch(TreeDev &d, long long need, hipStream_t stream)
{
  if (need <= d.hcap)
    return true;
  if (!hip_ok(hipStreamSynchronize(stream), "zones"))
    return false; // the buffers about to be replaced may still be read by the stream
  long long cap = next_cap(d.hcap, need);
  int **ints[] = {&d.hpar, &d.hflag, &d.hrank, &d.hidx, &d.hidx_s, &d.hz_node, &d.hzstart, &d.hzpar};
  for (int **q : ints)
  {
    (void)hipFree(*q);
    *q = nullptr;
    if (!hip_ok(hipMalloc((void **)q, sizeof(int) * (size_t)cap), "zones"))
      return false;
  }
  (void)hipFree(d.hz_xyz);
  (void)hipFree(d.hbbox);
  (void)hipFree(d.hzone);
  (void)hipFree(d.hzone_s);
  (void)hipFree(d.htmp);
  d.hz_xyz = d.hbbox = nullptr;
  d.hzone = d.hzone_s = nullptr;
  d.htmp = nullptr;
  d.hcap = 0;
  d.htmp_bytes = std::max(qz_sort_temp_bytes((int)cap, 32), qz_scan_temp_bytes((int)cap));
  if (!hip_ok(hipMalloc((void **)&d.hz_xyz, sizeof(int) * 3 * (size_t)cap), "zones") ||
      !hip_ok(hipMalloc((void **)&d.hbbox, sizeof(int) * 6 * (size_t)cap), "zones") ||
      !hip_ok(hipMalloc((void **)&d.hzone, sizeof(unsigned) * (size_t)cap), "zones") ||
      !hip_ok(hipMalloc((void **)&d.hzone_s, sizeof(unsigned) * (size_t)cap), "zones") ||
      !hip_ok(hipMalloc(&d.htmp, d.htmp_bytes), "zones"))
    return false;
  d.hcap = cap;
  return true;
}

// Head level marks and zones on the device (fill_cell_ref + fill_zones_ref for
// the head, tree_construction.c:527-926): marks from the particle counts of the
// sorted keys, ordered compaction, union-find zones, a stable sort by zone.  The
// host receives zone bounding boxes and sizes, plus the zone cells only when the
// children can refine further.  d.head_child serves as the cell -> list map
// until the node upload overwrites the zone cells with node indices.
template <typename P, typename G> bool Engine<P, G>::refine_head_device(int &zone_cells, hipStream_t s)
{
  hipStream_t stream = s; // every launch and copy below goes to the tree build's stream
  TreeDev &d = *tree->dev;
  TreeHost &th = *tree->host;
  const long long ncell = (long long)n * n * n;
  if (!d.hmark)
  {
    if (!hip_ok(hipMalloc((void **)&d.hmark, (size_t)ncell), "head marks") ||
        !hip_ok(hipMalloc((void **)&d.hlist, sizeof(int) * (size_t)ncell), "head marks"))
      return false;
  }
  // flags and their x, y, z dilations (the head list is free scratch until the
  // selection below fills it)
  unsigned char *tmp = (unsigned char *)d.hlist;
  if (n >= 16)
  {
    const long long nvec = ncell / 16;
    QZ_LAUNCHB("tree_head_mark", 4.0 * ncell + ncell, k_head_flag_x, nblk(nvec), cell_start, ncell, cfg.lmin,
               cfg.ref_criterion_ptcl, (uint4 *)tmp);
    QZ_LAUNCHB("tree_head_mark", 2.0 * ncell, k_head_dilate_yz, nblk(nvec), (const uint4 *)tmp, nvec, cfg.lmin,
               (uint4 *)d.hmark);
  }
  else
  {
    QZ_LAUNCHB("tree_head_mark", 5.0 * ncell, k_head_flag, nblk(ncell), cell_start, ncell, cfg.ref_criterion_ptcl, tmp);
    QZ_LAUNCHB("tree_head_mark", 2.0 * ncell, k_head_dilate, nblk(ncell), tmp, d.hmark, cfg.lmin, 0);
    QZ_LAUNCHB("tree_head_mark", 2.0 * ncell, k_head_dilate, nblk(ncell), d.hmark, tmp, cfg.lmin, 1);
    QZ_LAUNCHB("tree_head_mark", 2.0 * ncell, k_head_dilate, nblk(ncell), tmp, d.hmark, cfg.lmin, 2);
  }
  timer.begin(stream);
  if (!select_flags(d, d.hmark, ncell, d.hlist, d.flag_count, stream))
    return false;
  timer.end(stream, "tree_head_select", 2.0 * ncell);
  if (!hip_ok(hipMemcpyAsync(d.hflag_count, d.flag_count, sizeof(int), hipMemcpyDeviceToHost, stream), "marks") ||
      !hip_ok(hipStreamSynchronize(stream), "marks"))
    return false;
  const int M = *d.hflag_count;
  zone_cells = M;
  if (M == 0)
  {
    th.add_head_zones(0, nullptr, nullptr, nullptr);
    return true;
  }
  if (!reserve_zone_scratch(d, M + 1, stream))
    return false;
  // per marked cell: the list entry, the cell map, the six neighbours' map
  // entries and the union-find parents
  QZ_LAUNCHB("tree_head_zones", 12.0 * M, k_head_index, nblk(M), d.hlist, M, d.head_child);
  {
    const ZoneNbr g{d.hlist, d.head_child, n, 0, nullptr, nullptr, nullptr, 0};
    QZ_LAUNCHB("tree_head_hook", 28.0 * M, k_zone_hook_local, nblk(M), g, M, d.hpar, d.hidx);
    QZ_LAUNCHB("tree_head_hook", 28.0 * M, k_zone_hook_global, nblk(M), g, M, (const int *)d.hidx, d.hpar);
  }
  QZ_LAUNCH("tree_head_zones", k_fill<int>, 1, d.flag_count + 1, 0, (size_t)1);
  QZ_LAUNCHB("tree_head_compress", 8.0 * (M + 1), k_head_compress, nblk(M + 1), M, d.hpar, d.hidx, d.hflag,
             d.flag_count + 1);
  timer.begin(stream);
  if (!hip_ok(qz_exclusive_scan(d.htmp, d.htmp_bytes, d.hflag, d.hrank, M + 1, stream), "head zones"))
    return false;
  timer.end(stream, "tree_head_zones", 8.0 * (M + 1));
  if (!hip_ok(hipMemcpyAsync(d.hflag_count, d.hrank + M, sizeof(int), hipMemcpyDeviceToHost, stream), "zones") ||
      !hip_ok(hipMemcpyAsync(d.hflag_count + 1, d.flag_count + 1, sizeof(int), hipMemcpyDeviceToHost, stream),
              "zones") ||
      !hip_ok(hipStreamSynchronize(stream), "zones"))
    return false;
  const int nz = d.hflag_count[0];
  if (!zones_in_order(M, nz, d.hflag_count[1] != 0, stream))
    return false;
  QZ_LAUNCHB("tree_head_zones", 24.0 * nz, k_bbox_init, nblk(6LL * nz), nz, d.hbbox);
  QZ_LAUNCHB("tree_head_zones", 40.0 * M, k_head_zone_cells, nblk(M), M, d.hzone_s, d.hidx_s, d.hlist, n, 0, d.hzstart,
             d.hbbox,
            d.hz_node, d.hz_xyz);
  std::vector<int> zstart(nz + 1), bbox(6 * (size_t)nz), zsize(nz);
  const bool need_cells = host_refine && cfg.lmin + 1 < cfg.lmax;
  std::vector<int> cells(need_cells ? 3 * (size_t)M : 0);
  if (!hip_ok(hipMemcpyAsync(zstart.data(), d.hzstart, sizeof(int) * nz, hipMemcpyDeviceToHost, stream), "zones") ||
      !hip_ok(hipMemcpyAsync(bbox.data(), d.hbbox, sizeof(int) * 6 * nz, hipMemcpyDeviceToHost, stream), "zones") ||
      (need_cells &&
       !hip_ok(hipMemcpyAsync(cells.data(), d.hz_xyz, sizeof(int) * 3 * (size_t)M, hipMemcpyDeviceToHost, stream),
               "zones")) ||
      !hip_ok(hipStreamSynchronize(stream), "zones"))
    return false;
  zstart[nz] = M;
  for (int z = 0; z < nz; z++)
    zsize[z] = zstart[z + 1] - zstart[z];
  th.add_head_zones(nz, bbox.data(), zsize.data(), need_cells ? cells.data() : nullptr);
  return true;
}

// Marks and zones of the nodes of tree level lr on the device, and the next
// level's nodes from them (fill_cell_ref + fill_zones_ref + create_child_nodes
// per parent, tree_construction.c:508-926, 1092-1661): the level's cells are
// [c0, c0 + ncell) and its box pool entries [b0, b0 + nb).  The marks, the
// union-find and the zone sort are those of the head level run over all
// parents at once; the host receives per zone its parent, bounding box and
// size, and the zone cells of the next level stay in d.hz_node / d.hz_xyz for
// its upload.  zone_cells: the number of zone cells (marked cells).
template <typename P, typename G>
bool Engine<P, G>::refine_level_device(long long c0, long long ncell, long long b0, long long nb, int &zone_cells,
                                       hipStream_t s)
{
  hipStream_t stream = s; // every launch and copy below goes to the tree build's stream
  TreeDev &d = *tree->dev;
  TreeHost &th = *tree->host;
  zone_cells = 0;
  if (ncell > d.lcell_cap)
  {
    if (!hip_ok(hipStreamSynchronize(stream), "level marks"))
      return false;
    (void)hipFree(d.lmark);
    (void)hipFree(d.llist);
    d.lmark = nullptr;
    d.llist = nullptr;
    d.lcell_cap = 0;
    const long long cap = next_cap(0, ncell);
    if (!hip_ok(hipMalloc((void **)&d.lmark, (size_t)cap), "level marks") ||
        !hip_ok(hipMalloc((void **)&d.llist, sizeof(int) * (size_t)cap), "level marks"))
      return false;
    d.lcell_cap = cap;
  }
  if (nb > d.lbox_cap)
  {
    if (!hip_ok(hipStreamSynchronize(stream), "level marks"))
      return false;
    (void)hipFree(d.lbflag);
    (void)hipFree(d.lbmap);
    d.lbflag = nullptr;
    d.lbmap = nullptr;
    d.lbox_cap = 0;
    const long long cap = next_cap(0, nb);
    if (!hip_ok(hipMalloc((void **)&d.lbflag, (size_t)cap), "level marks") ||
        !hip_ok(hipMalloc((void **)&d.lbmap, sizeof(int) * (size_t)cap), "level marks"))
      return false;
    d.lbox_cap = cap;
  }
  if (th.extra < 2)
  {
    set_error("tree_adaptation: child boxes need a margin of at least one cell");
    return false;
  }
  QZ_LAUNCHB("tree_level_mark", 1.0 * nb, k_fill<unsigned char>, nblk(nb), d.lbflag, (unsigned char)0, (size_t)nb);
  QZ_LAUNCHB("tree_level_mark", 16.0 * ncell, k_level_flag, nblk(ncell), d.nodes, d.zone_node, (int)c0, (int)ncell, b0,
             d.cell_box, d.cstart, d.cend, cfg.ref_criterion_ptcl, d.lbflag);
  QZ_LAUNCHB("tree_level_mark", 12.0 * ncell, k_level_mark, nblk(ncell), d.nodes, d.zone_node, (int)c0, (int)ncell, b0,
             d.cell_box, d.lbflag, d.lmark);
  QZ_LAUNCHB("tree_level_mark", 4.0 * nb, k_fill<int>, nblk(nb), d.lbmap, -1, (size_t)nb);
  timer.begin(stream);
  if (!select_flags(d, d.lmark, ncell, d.llist, d.flag_count, stream))
    return false;
  timer.end(stream, "tree_level_select", 2.0 * ncell);
  if (!hip_ok(hipMemcpyAsync(d.hflag_count, d.flag_count, sizeof(int), hipMemcpyDeviceToHost, stream), "marks") ||
      !hip_ok(hipStreamSynchronize(stream), "marks"))
    return false;
  const int M = *d.hflag_count;
  zone_cells = M;
  if (M == 0)
    return true;
  if (!reserve_zone_scratch(d, M + 1, stream))
    return false;
  QZ_LAUNCHB("tree_level_index", 16.0 * M, k_level_index, nblk(M), d.llist, M, (int)c0, d.nodes, d.zone_node,
             d.cell_box, b0, d.lbmap);
  {
    const ZoneNbr g{d.llist, d.lbmap, 0, (int)c0, d.nodes, d.zone_node, d.cell_box, b0};
    QZ_LAUNCHB("tree_level_hook", 28.0 * M, k_zone_hook_local, nblk(M), g, M, d.hpar, d.hidx);
    QZ_LAUNCHB("tree_level_hook", 28.0 * M, k_zone_hook_global, nblk(M), g, M, (const int *)d.hidx, d.hpar);
  }
  QZ_LAUNCH("tree_level_zones", k_fill<int>, 1, d.flag_count + 1, 0, (size_t)1);
  QZ_LAUNCHB("tree_level_compress", 8.0 * (M + 1), k_head_compress, nblk(M + 1), M, d.hpar, d.hidx, d.hflag,
             d.flag_count + 1);
  timer.begin(stream);
  if (!hip_ok(qz_exclusive_scan(d.htmp, d.htmp_bytes, d.hflag, d.hrank, M + 1, stream), "level zones"))
    return false;
  timer.end(stream, "tree_level_scan", 8.0 * (M + 1));
  if (!hip_ok(hipMemcpyAsync(d.hflag_count, d.hrank + M, sizeof(int), hipMemcpyDeviceToHost, stream), "zones") ||
      !hip_ok(hipMemcpyAsync(d.hflag_count + 1, d.flag_count + 1, sizeof(int), hipMemcpyDeviceToHost, stream),
              "zones") ||
      !hip_ok(hipStreamSynchronize(stream), "zones"))
    return false;
  const int nz = d.hflag_count[0];
  if (!zones_in_order(M, nz, d.hflag_count[1] != 0, stream))
    return false;
  QZ_LAUNCHB("tree_level_zones", 24.0 * nz, k_bbox_init, nblk(6LL * nz), nz, d.hbbox);
  QZ_LAUNCHB("tree_level_zone_cells", 44.0 * M, k_level_zone_cells, nblk((M + kZoneRun - 1) / kZoneRun), M, d.hzone_s, d.hidx_s, d.llist, (int)c0,
             d.cell_xyz, d.zone_node, (int)th.nodes.size(), d.hzstart, d.hzpar, d.hbbox, d.hz_node, d.hz_xyz);
  std::vector<int> zstart(nz + 1), zpar(nz), bbox(6 * (size_t)nz), zsize(nz);
  if (!hip_ok(hipMemcpyAsync(zstart.data(), d.hzstart, sizeof(int) * nz, hipMemcpyDeviceToHost, stream), "zones") ||
      !hip_ok(hipMemcpyAsync(zpar.data(), d.hzpar, sizeof(int) * nz, hipMemcpyDeviceToHost, stream), "zones") ||
      !hip_ok(hipMemcpyAsync(bbox.data(), d.hbbox, sizeof(int) * 6 * nz, hipMemcpyDeviceToHost, stream), "zones") ||
      !hip_ok(hipStreamSynchronize(stream), "zones"))
    return false;
  zstart[nz] = M;
  for (int z = 0; z < nz; z++)
    zsize[z] = zstart[z + 1] - zstart[z];
  th.add_level_zones(nz, zpar.data(), bbox.data(), zsize.data());
  return true;
}

// The tree's host and device state, created on first use.
template <typename P, typename G> bool Engine<P, G>::tree_state()
{
  if (!tree)
  {
    tree = new TreeState();
    tree->host = new TreeHost();
    tree->host->init(cfg.lmin, cfg.lmax, cfg.ref_criterion_ptcl);
    tree->dev = new TreeDev();
    TreeDev *d = tree->dev;
    d->gsize = sizeof(G);
    size_t ncell = (size_t)n * n * n;
    d->flag_cap = std::min<long long>((long long)ncell, (long long)N / std::max(1, cfg.ref_criterion_ptcl) + 16);
    if (!hip_ok(hipMalloc((void **)&d->head_child, sizeof(int) * ncell), "head_child") ||
        !hip_ok(hipMalloc((void **)&d->flags, sizeof(int) * 2 * (size_t)d->flag_cap), "flags") ||
        !hip_ok(hipMalloc((void **)&d->flag_count, sizeof(int) * 4), "flag count") ||
        !hip_ok(hipMalloc((void **)&d->done_count, sizeof(int) * 4), "done count") ||
        !hip_ok(hipHostMalloc((void **)&d->hflag_count, sizeof(int) * 4), "pinned") ||
        !hip_ok(hipHostMalloc((void **)&d->hflags, sizeof(int) * 2 * (size_t)d->flag_cap), "pinned"))
      return false;
  }
  return true;
}

template <typename P, typename G> bool Engine<P, G>::tree_adaptation()
{
  if (cfg.lmin == cfg.lmax)
    return true;
  if (!tree_state() || !join_ahead())
    return false;
  tree->dev->drop_plan();
  HostSpan span_all(timer, "tree_adaptation");
  if (!sort_particles() || !build_tree(stream))
    return false;
  // the first level's node multigrid set-up is host work that needs only the
  // tree: run it during the head density and solve
  TreeDev &d = *tree->dev;
  if (setup_ahead && cfg.branch_solver == QZ_BRANCH_MG && d.lfirst.size() > 1 && !timer.on && !timer.host_on)
    return start_ahead(false);
  return true;
}

// The tree below the head from the sorted particles (tree_construction.c rules):
// marks and zones of every level on the device, the child box geometry on the
// host, the node tables, cell and point lists on the device.  Everything is
// issued on stream s (the engine's stream, or the set-up stream when the tree
// is built while the head density and potential are computed).
template <typename P, typename G> bool Engine<P, G>::build_tree(hipStream_t s)
{
  hipStream_t stream = s; // every launch and copy below goes to s
  TreeHost &th = *tree->host;
  TreeDev &d = *tree->dev;
  const size_t ncell_head = (size_t)n * n * n;
  // back to all -1: only the cells the last build set, when they are known
  if (d.hc_set < 0)
    QZ_LAUNCHB("tree_fill", 4.0 * ncell_head, k_fill<int>, nblk(ncell_head), d.head_child, -1, ncell_head);
  else if (d.hc_set > 0)
    QZ_LAUNCHB("tree_fill", 8.0 * d.hc_set, k_head_unset, nblk(d.hc_set), d.hlist, d.hc_set, d.head_child);
  d.hc_set = -1;
  int zone_cells = 0; // zone cells of the next level, when its zones were found on the device
  {
    HostSpan span(timer, "tree_head_zones");
    if (!refine_head_device(zone_cells, stream))
      return false;
  }
  d.hc_set = zone_cells; // the marked head cells, all of them zone cells
  d.hnodes.clear();
  d.lfirst.assign(1, 0);
  d.lcount.assign(1, 1);
  d.lc0.assign(1, 0);
  d.lncell.assign(1, 0);
  d.lb0.assign(1, 0);
  d.lnb.assign(1, 0);
  d.lz0.assign(1, 0);
  d.lnz.assign(1, 0);
  d.lg0.assign(1, 0);
  d.lng.assign(1, 0);
  d.lpmax.assign(1, 0);
  d.lnbm.assign(1, 0);
  d.used_zone = d.used_cell = d.used_box = d.used_pt = d.used_grid = 0;
  int lr = 1;
  while (lr < (int)th.lev_start.size())
  {
    int first = th.lev_start[lr], count = th.lev_count[lr];
    HostSpan span_lv(timer, "tree_level_upload");
    // ---- descriptors and zone cells of this level
    long long nz = 0, ncell = 0, nb = 0, npt = 0, ng = 0, pmax = 0;
    for (int q = first; q < first + count; q++)
    {
      const HNode &h = th.nodes[q];
      nz += h.ncell / 8;
      ncell += h.ncell;
      nb += h.bcap;
      npt += h.pcap;
      ng += h.gcap;
      pmax = std::max<long long>(pmax, h.pcap);
    }
    if (!reserve_level<G>(&d, stream, nz, ncell, nb, npt, ng))
      return false;
    const bool dev_zones = lr == 1 || !host_refine; // zones found on the device
    std::vector<int> zn(dev_zones ? 0 : nz), zx(dev_zones ? 0 : 3 * nz);
    long long zo = 0;
    for (int q = first; q < first + count; q++)
    {
      HNode &h = th.nodes[q];
      DNode dn{};
      dn.lv = h.lv;
      dn.parent = h.parent;
      for (int a = 0; a < 3; a++)
      {
        dn.ts[a] = h.ts[a];
        dn.rd[a] = h.rd[a];
        dn.bmin[a] = h.bmin[a];
        dn.bmax[a] = h.bmax[a];
        if (h.sc[a])
          dn.simc |= 1 << a;
      }
      dn.ncell = h.ncell;
      dn.coff = (int)d.used_cell;
      dn.zoff = (int)d.used_zone;
      dn.pcap = h.pcap;
      dn.poff = (int)d.used_pt;
      dn.goff = d.used_grid;
      dn.boff = d.used_box;
      h.goff = dn.goff;
      h.boff = dn.boff;
      h.coff = dn.coff;
      h.poff = dn.poff;
      for (int j = 0; !dev_zones && j < h.ncell / 8; j++)
      {
        // zone cell j (parent-level coordinates) holds child cells 8j .. 8j + 7
        zn[zo] = q;
        for (int a = 0; a < 3; a++)
          zx[3 * zo + a] = h.zxyz[3 * j + a];
        zo++;
      }
      d.used_zone += h.ncell / 8;
      d.used_cell += h.ncell;
      d.used_box += h.bcap;
      d.used_pt += h.pcap;
      d.used_grid += h.gcap;
      d.hnodes.push_back(dn);
    }
    long long z0 = d.used_zone - nz, c0 = d.used_cell - ncell, b0 = d.used_box - nb, g0 = d.used_grid - ng;
    d.lfirst.push_back(first);
    d.lcount.push_back(count);
    d.lc0.push_back(c0);
    d.lncell.push_back(ncell);
    d.lb0.push_back(b0);
    d.lnb.push_back(nb);
    d.lz0.push_back(z0);
    d.lnz.push_back(nz);
    d.lg0.push_back(g0);
    d.lng.push_back(ng);
    d.lpmax.push_back(pmax);
    if ((long long)d.hnodes.size() > d.node_cap)
    {
      long long cap = next_cap(d.node_cap, (long long)d.hnodes.size() + 64);
      if (!regrow(d.nodes, cap, (long long)first, "node table", stream))
        return false;
      d.node_cap = (int)cap;
    }
    if (dev_zones && nz != zone_cells)
    {
      set_error("tree_adaptation: level %d zone cells %lld != %d", lr, nz, zone_cells);
      return false;
    }
    if (!hip_ok(hipMemcpyAsync(d.nodes + first, d.hnodes.data() + first, sizeof(DNode) * count,
                               hipMemcpyHostToDevice, stream),
                "node table") ||
        !hip_ok(hipMemcpyAsync(d.zone_node + z0, dev_zones ? d.hz_node : zn.data(), sizeof(int) * nz,
                               dev_zones ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream),
                "zones") ||
        !hip_ok(hipMemcpyAsync(d.zone_xyz + 3 * z0, dev_zones ? d.hz_xyz : zx.data(), sizeof(int) * 3 * nz,
                               dev_zones ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream),
                "zones"))
      return false;
    // ---- device: boxes, cells, ranges, masses, grid-point lists
    QZ_LAUNCHB("tree_boxes", 12.0 * nb, k_init_boxes, nblk(nb), d.nodes, first, count, b0, nb, d.box, d.bstart,
               d.bend); // box status and its particle range
    // per cell: the zone cell in, its coordinates, box index and particle range out
    QZ_LAUNCHB("tree_cells", 36.0 * ncell, k_expand_cells, nblk(ncell), d.nodes, d.zone_node, d.zone_xyz, (int)c0,
               (int)ncell,
              cfg.lmin, D, key_out, cell_start, d.cell_xyz, d.cell_box, d.cstart, d.cend, d.box, d.bstart, d.bend);
    {
      long long maxcell = 0;
      for (int q = first; q < first + count; q++)
        maxcell = std::max<long long>(maxcell, th.nodes[q].ncell);
      int nbm = blocks_per_node(maxcell);
      d.lnbm.push_back(nbm);
      long long need = 2LL * count * nbm;
      if (need > d.mpart_cap)
      {
        long long cap = next_cap(d.mpart_cap, need);
        if (!regrow(d.mpart, cap, 0, "mass partials", stream))
          return false;
        d.mpart_cap = cap;
      }
    }
    // grid-point classification: chunks of kClassBlock points per node
    {
      std::vector<int2> &ct = d.hctab;
      std::vector<int> &c0v = d.hchunk0;
      ct.clear();
      c0v.assign(count + 1, 0);
      for (int q = 0; q < count; q++)
      {
        c0v[q] = (int)ct.size();
        const long long pc = th.nodes[first + q].pcap;
        for (long long t0 = 0; t0 < pc; t0 += kClassBlock)
          ct.push_back(make_int2(q, (int)t0));
      }
      c0v[count] = (int)ct.size();
      const int nch = (int)ct.size();
      if (nch > 0)
      {
        if (!reserve_classify(d, nch, count, stream))
          return false;
        if (!hip_ok(hipMemcpyAsync(d.ctab, ct.data(), sizeof(int2) * nch, hipMemcpyHostToDevice, stream), "classify") ||
            !hip_ok(hipMemcpyAsync(d.chunk0, c0v.data(), sizeof(int) * (count + 1), hipMemcpyHostToDevice, stream),
                    "classify"))
          return false;
        timer.begin(stream);
        hipLaunchKernelGGL(k_classify_count, dim3(nch), dim3(kClassBlock), 0, stream, d.nodes, first, d.box, d.ctab,
                           nch, d.ccnt);
        for (int k = 0; k < 3; k++)
          if (!hip_ok(qz_exclusive_scan(d.cscan_tmp, d.cscan_bytes, d.ccnt + (size_t)k * nch, d.coff + (size_t)k * nch,
                                        nch, stream),
                      "classify"))
            return false;
        hipLaunchKernelGGL(k_classify_write, dim3(nch), dim3(kClassBlock), 0, stream, d.nodes, first, d.box, d.ctab,
                           nch, d.coff, d.chunk0, d.intr_g, d.intr_xyz, d.bdry_g, d.bdry_xyz, d.sim_g, d.sim_xyz);
        hipLaunchKernelGGL(k_classify_totals, dim3(nblk(count)), dim3(kBlock), 0, stream, d.nodes, first, count, nch,
                           d.ccnt, d.coff, d.chunk0);
        timer.end(stream, "tree_classify", 8.0 * npt); // box statuses in, listed points (index + xyz) out
      }
    }
    QZ_LAUNCHB("tree_children", 20.0 * nz, k_mark_children, nblk(nz), d.nodes, d.zone_node, d.zone_xyz, (int)z0,
               (int)nz, n,
              d.head_child, d.box);
    // zero the level's grid fields (points that are never listed stay 0, as calloc leaves them)
    QZ_LAUNCHB("tree_fill", 8.0 * ng, k_fill<double>, nblk(ng), d.phi + g0, 0.0, (size_t)ng);
    QZ_LAUNCHB("tree_fill", (double)sizeof(G) * ng, k_fill<G>, nblk(ng), (G *)d.ax + g0, (G)0, (size_t)ng);
    QZ_LAUNCHB("tree_fill", (double)sizeof(G) * ng, k_fill<G>, nblk(ng), (G *)d.ay + g0, (G)0, (size_t)ng);
    QZ_LAUNCHB("tree_fill", (double)sizeof(G) * ng, k_fill<G>, nblk(ng), (G *)d.az + g0, (G)0, (size_t)ng);
    // ---- next level's refinement counts
    int lv = cfg.lmin + lr;
    if (lv < cfg.lmax && !host_refine)
    {
      HostSpan span(timer, "tree_refine_level");
      if (!refine_level_device(c0, ncell, b0, nb, zone_cells, stream))
        return false;
    }
    else if (lv < cfg.lmax)
    {
      QZ_LAUNCH("tree_fill", k_fill<int>, 1, d.flag_count, 0, (size_t)1);
      QZ_LAUNCH("tree_flags", k_flag_cells, nblk(ncell), d.nodes, d.zone_node, first, (int)c0, (int)ncell, d.cstart,
                d.cend, cfg.ref_criterion_ptcl, d.flag_count, d.flags, d.flag_cap);
      std::vector<int> fn, fc;
      {
        HostSpan span_f(timer, "tree_flags");
        if (!hip_ok(hipMemcpyAsync(d.hflag_count, d.flag_count, sizeof(int), hipMemcpyDeviceToHost, stream), "flags") ||
            !hip_ok(hipStreamSynchronize(stream), "flags"))
          return false;
        int nf = std::min(*d.hflag_count, d.flag_cap);
        // into pinned memory: a pageable copy of these megabytes stages through
        // the runtime's bounce buffer at a fraction of the link's rate
        const int *fl = d.hflags;
        if (nf > 0 &&
            (!hip_ok(hipMemcpyAsync(d.hflags, d.flags, sizeof(int) * 2 * nf, hipMemcpyDeviceToHost, stream), "flags") ||
             !hip_ok(hipStreamSynchronize(stream), "flags")))
          return false;
        fn.resize(nf);
        fc.resize(nf);
        for (int i = 0; i < nf; i++)
        {
          fn[i] = fl[2 * i];
          fc[i] = fl[2 * i + 1];
        }
      }
      HostSpan span(timer, "tree_refine_level");
      th.refine_level(lr, fn, fc);
      if (timer.on || timer.host_on)
      {
        timer.host("tree_refine_group", th.ms_mark);
        timer.host("tree_refine_nodes", th.ms_zones);
        timer.host("tree_refine_append", th.ms_children);
      }
    }
    lr++;
  }
  tree->n_nodes = (int)th.nodes.size();
  // fetch the classified counts and masses
  if (tree->n_nodes > 0 &&
      !hip_ok(hipMemcpyAsync(d.hnodes.data(), d.nodes, sizeof(DNode) * tree->n_nodes, hipMemcpyDeviceToHost, stream),
              "node table"))
    return false;
  if (!hip_ok(hipStreamSynchronize(stream), "tree_adaptation"))
    return false;
  stats.n_nodes = tree->n_nodes;
  stats.n_levels = (int)d.lfirst.size();
  stats.branch_points = 0;
  stats.branch_intr = 0;
  d.lnintr.assign(d.lfirst.size(), 0);
  for (size_t l = 1; l < d.lfirst.size(); l++)
    for (int q = d.lfirst[l]; q < d.lfirst[l] + d.lcount[l]; q++)
      d.lnintr[l] += d.hnodes[q].nintr;
  for (auto &dn : d.hnodes)
  {
    stats.branch_points += dn.nintr + dn.nbdry + dn.nsim;
    stats.branch_intr += dn.nintr;
  }
  // the partial-sum buffers: nodes x blocks per node
  long long need = 0;
  for (size_t l = 1; l < d.lfirst.size(); l++)
    need = std::max<long long>(need, (long long)d.lcount[l] * blocks_per_node(d.lpmax[l]));
  if (need > d.part_cap)
  {
    if (!hip_ok(hipStreamSynchronize(stream), "partials"))
      return false;
    (void)hipFree(d.part1);
    (void)hipFree(d.part2);
    (void)hipFree(d.part3);
    d.part_cap = need + 1024;
    if (!hip_ok(hipMalloc((void **)&d.part1, sizeof(double) * d.part_cap), "partials") ||
        !hip_ok(hipMalloc((void **)&d.part2, sizeof(double) * d.part_cap), "partials") ||
        !hip_ok(hipMalloc((void **)&d.part3, sizeof(double) * d.part_cap), "partials"))
      return false;
  }
  return true;
}

// A force pass (reset, tree, density, potential, grid accelerations and, with
// with_particles, the particle accelerations) in which the tree below the head
// is built on the set-up thread and stream while the engine's stream deposits
// the head density and solves the head potential, which do not depend on the
// tree.  The node stages then run as in the stage sequence, with the same
// kernels on the same data, so the results are those of the stages called one
// by one.  With timing on, or QZ_NODE_SETUP_AHEAD=0, the stages run in order.
template <typename P, typename G> bool Engine<P, G>::force_overlapped(bool with_particles)
{
  if (cfg.lmin == cfg.lmax || !setup_ahead || timer.on)
    return (with_particles ? reset_marks() : reset()) && tree_adaptation() && grid_density() && potential_head() &&
           potential_branches() &&
           grid_acceleration() && (!with_particles || particle_acceleration());
  // with the particles' accelerations gathered at the end of this pass, reset
  // need not zero them (nothing reads them before the gather writes every
  // particle's), nor the head density (its fold writes every grid point)
  if (!(with_particles ? reset_marks() : reset()) || !tree_state() || !join_ahead())
    return false;
  TreeDev &d = *tree->dev;
  d.drop_plan();
  if (!sort_particles() || !start_ahead(true))
    return false;
  bool head_ok;
  {
    HostSpan span(timer, "overlap_head");
    head_ok = deposit_head() && potential_head();
  }
  {
    HostSpan span(timer, "overlap_join_wait");
    if (!join_ahead() || !head_ok)
      return false;
  }
  if (!hip_ok(hipStreamWaitEvent(stream, d.ahead_done, 0), "set-up stream"))
    return false;
  // the head accelerations need only the head potential: they run on the
  // set-up stream beside the node solves, which leave most CUs idle, and the
  // engine's stream waits for them before the node accelerations
  if (!d.acc_go && (!hip_ok(hipEventCreateWithFlags(&d.acc_go, hipEventDisableTiming), "set-up event") ||
                    !hip_ok(hipEventCreateWithFlags(&d.acc_done, hipEventDisableTiming), "set-up event")))
    return false;
  if (!hip_ok(hipEventRecord(d.acc_go, stream), "set-up event") ||
      !hip_ok(hipStreamWaitEvent(d.ahead_stream, d.acc_go, 0), "set-up stream"))
    return false;
  grid_acc_head(d.ahead_stream);
  if (!hip_ok(hipEventRecord(d.acc_done, d.ahead_stream), "set-up event"))
    return false;
  head_acc_queued = true;
  const bool ok = (!has_tree() || branch_density()) && potential_branches() && grid_acceleration() &&
                  (!with_particles || particle_acceleration()) &&
                  hip_ok(hipGetLastError(), "force pass"); // stream-ordered; readers synchronise
  if (head_acc_queued) // a stage failed before the wait: keep later passes in order
    (void)wait_head_acc();
  return ok;
}

// The engine's stream waits for the head accelerations queued on the set-up stream.
template <typename P, typename G> bool Engine<P, G>::wait_head_acc()
{
  head_acc_queued = false;
  return hip_ok(hipStreamWaitEvent(stream, tree->dev->acc_done, 0), "set-up stream");
}

template <typename P, typename G> bool Engine<P, G>::branch_density()
{
  TreeDev &d = *tree->dev;
  for (size_t l = 1; l < d.lfirst.size(); l++)
  {
    QZ_LAUNCHB("branch_fill", (double)sizeof(G) * d.lng[l], k_fill<G>, nblk(d.lng[l]), (G *)d.rho + d.lg0[l], (G)0,
               (size_t)d.lng[l]);
    const int lv = cfg.lmin + (int)l;
    QZ_LAUNCHB("deposit_runs", (24.0 + sizeof(P) + 8.0) * N, (k_deposit_runs<P, G>), nblk(N), key_out, pos[0], pos[1],
               pos[2], pa[9], N, lv, 3 * (D - (int)l), (1 << lv) - 1, tails);
    cell_totals(d.cstart + d.lc0[l], d.cend + d.lc0[l], d.lncell[l]);
    dim3 grid(blocks_per_node(d.lpmax[l]), d.lcount[l]);
    timer.begin(stream);
    hipLaunchKernelGGL((k_branch_deposit_fold<G, G>), grid, dim3(kBlock), 0, stream, d.nodes, d.lfirst[l], d.sim_g,
                       d.sim_xyz, d.bdry_g, d.bdry_xyz, d.intr_g, d.intr_xyz, d.bstart, d.bend, tails, (G *)d.rho,
                       4.0 * cfg.G * kPi);
    timer.end(stream, "deposit_branch_fold", 4.0 * d.lncell[l] + (double)sizeof(G) * d.lng[l]);
    // node masses for the CG threshold (poisson_error.c:240) from this level's runs
    timer.begin(stream);
    hipLaunchKernelGGL(k_node_mass<G>, dim3(d.lnbm[l], d.lcount[l]), dim3(kBlock), 0, stream, d.nodes, d.lfirst[l],
                       d.cstart, d.cend, tails, d.mpart);
    hipLaunchKernelGGL(k_node_mass_fold, dim3(d.lcount[l]), dim3(kBlock), 0, stream, d.nodes, d.lfirst[l], d.mpart,
                       d.lnbm[l]);
    timer.end(stream, "node_mass", 8.0 * d.lncell[l]);
  }
  return true;
}

// Potential of the refinement nodes, level by level (potential_branches.c:401-470):
// boundary and starting values from the parent, then either the reference's
// per-node CG or the batched masked multigrid over all nodes of the level.
template <typename P, typename G> bool Engine<P, G>::potential_branches()
{
  if (!has_tree())
    return true;
  HostSpan span(timer, "potential_branches");
  TreeDev &d = *tree->dev;
  stats.branch_iterations = 0;
  stats.branch_iter_points = 0;
  stats.branch_cg_levels = 0;
  const int zero_bc = cfg.head_bc == QZ_BC_ZERO ? 1 : 0;
  for (size_t l = 1; l < d.lfirst.size(); l++)
  {
    const int first = d.lfirst[l], count = d.lcount[l];
    dim3 grid(blocks_per_node(d.lpmax[l]), count);
    timer.begin(stream);
    hipLaunchKernelGGL(k_branch_init_pot, grid, dim3(kBlock), 0, stream, d.nodes, first, d.sim_g, d.sim_xyz, d.bdry_g,
                       d.bdry_xyz, d.intr_g, d.intr_xyz, d.phi, d.phi_old, phi, n, cfg.G * total_mass, zero_bc);
    timer.end(stream, "branch_init_pot", 24.0 * d.lng[l]); // phi and phi_old out, parent averages
    bool solved = false;
    if (cfg.branch_solver == QZ_BRANCH_MG && !branch_mg_level(l, solved, d.phi, (const G *)d.rho, false))
      return false;
    if (!solved)
    {
      stats.branch_cg_levels++;
      if (!branch_cg_level(l))
        return false;
    }
    if (!hip_ok(hipMemcpyAsync(d.hnodes.data() + first, d.nodes + first, sizeof(DNode) * count,
                               hipMemcpyDeviceToHost, stream),
                "node table") ||
        !hip_ok(hipStreamSynchronize(stream), "potential_branches"))
      return false;
    for (int q = first; q < first + count; q++)
    {
      stats.branch_iterations = std::max(stats.branch_iterations, d.hnodes[q].done);
      stats.branch_iter_points += (long long)d.hnodes[q].done * d.hnodes[q].nintr;
    }
  }
  return hip_ok(hipStreamSynchronize(stream), "potential_branches");
}

// Replay a captured launch sequence (or run it eagerly while kernels are being
// timed) until done_count reaches `count` or `max_it` iterations have run.
template <typename P, typename G>
template <typename F>
bool Engine<P, G>::iterate_batches(F &&batch, int per_batch, int max_it, int count, int &it, bool &all_done,
                                   hipGraphExec_t given)
{
  TreeDev &d = *tree->dev;
  hipGraphExec_t gexec = nullptr;
  if (given && (timer.on || !batch_graphs))
    (void)hipGraphExecDestroy(given); // timed runs launch eagerly
  else if (given)
    gexec = given; // captured with the set-up (owned from here on)
  else if (!timer.on && batch_graphs)
  {
    HostSpan span(timer, "graph_instantiate");
    hipGraph_t g = nullptr;
    if (!hip_ok(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "graph capture"))
      return false;
    batch();
    const bool made = hip_ok(hipStreamEndCapture(stream, &g), "graph capture") &&
                      hip_ok(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0), "graph instantiate");
    if (g)
      (void)hipGraphDestroy(g);
    if (!made)
      return false;
  }
  all_done = false;
  // Pipelined: batch k + 1 is queued before the done count after batch k is
  // read, so the device does not idle while the host looks.  A batch queued
  // after convergence finds every node done and returns at once.
  if (!d.done_ev[0] && (!hip_ok(hipEventCreateWithFlags(&d.done_ev[0], hipEventDisableTiming), "event") ||
                        !hip_ok(hipEventCreateWithFlags(&d.done_ev[1], hipEventDisableTiming), "event")))
    return false;
  int *hc = d.hflag_count + 2; // two pinned slots, one per batch in flight
  auto launch = [&](int slot) -> bool {
    bool r = true;
    if (gexec)
      r = hip_ok(hipGraphLaunch(gexec, stream), "graph launch");
    else
      batch();
    return r && hip_ok(hipMemcpyAsync(hc + slot, d.done_count, sizeof(int), hipMemcpyDeviceToHost, stream), "done") &&
           hip_ok(hipEventRecord(d.done_ev[slot], stream), "done");
  };
  bool ok = max_it > 0 && launch(0);
  int launched = max_it > 0 ? per_batch : 0, slot = 0;
  for (it = 0; ok && it < max_it; it += per_batch, slot ^= 1)
  {
    if (launched < max_it)
    {
      ok = launch(slot ^ 1);
      launched += per_batch;
    }
    ok = ok && hip_ok(hipEventSynchronize(d.done_ev[slot]), "done");
    if (ok && hc[slot] >= count)
    {
      all_done = true;
      it += per_batch;
      break;
    }
  }
  // The batch queued ahead (a replay of gexec and its done-count copy) may still
  // be in flight here, on the error path too: drain the stream before the
  // executable graph, whose kernel arguments those launches read, is destroyed.
  const bool drained = hip_ok(hipStreamSynchronize(stream), "iterate");
  if (gexec)
    (void)hipGraphExecDestroy(gexec);
  return ok && drained;
}

// Unpreconditioned CG per node, all nodes of the level batched
// (conjugate_gradient.c:327-460, interior points only).
template <typename P, typename G> bool Engine<P, G>::branch_cg_level(size_t l)
{
  TreeDev &d = *tree->dev;
  const int first = d.lfirst[l], count = d.lcount[l];
  const int nbx = blocks_per_node(d.lpmax[l]);
  dim3 grid(nbx, count);
  const double ni = (double)d.lnintr[l];
  QZ_LAUNCHB("branch_fill", 8.0 * d.lng[l], k_fill<double>, nblk(d.lng[l]), d.cg_p + d.lg0[l], 0.0,
             (size_t)d.lng[l]);
  timer.begin(stream);
  hipLaunchKernelGGL((k_cg_init<G>), grid, dim3(kBlock), 0, stream, d.nodes, first, d.intr_g, d.phi, (G *)d.rho,
                     d.cg_r, d.cg_p, d.part2);
  hipLaunchKernelGGL(k_cg_init_fold, dim3(count), dim3(kBlock), 0, stream, d.nodes, first, d.part2, nbx);
  timer.end(stream, "cg_init", ni * (4.0 + 8.0 + sizeof(G) + 16.0));
  QZ_LAUNCH("tree_fill", k_fill<int>, 1, d.done_count, 0, (size_t)2);
  int *it_base = d.done_count + 1;
  // one batch = 5 CG iterations (conjugate_gradient.c:379-456), the last one
  // with the convergence check (iter_between_check_potential_solution = 5).
  // Algorithmic bytes per interior point: Ap reads idx + p, writes Ap; update
  // reads idx, phi, r, p, Ap, writes phi, r (+ phi_old on checks); direction
  // reads idx, r, p, writes p.
  auto batch = [&]() {
    for (int il = 0; il < 5; il++)
    {
      QZ_LAUNCHB("branch_cg_ap", ni * 20.0, k_cg_ap, grid, d.nodes, first, d.intr_g, d.cg_p, d.cg_Ap, d.part1);
      QZ_LAUNCHB("branch_cg_update", ni * 52.0, k_cg_update, grid, d.nodes, first, it_base, il, d.intr_g, d.phi,
                 d.phi_old, d.cg_r, d.cg_p, d.cg_Ap, d.part1, d.part2, d.part3, cfg.tol_change);
      QZ_LAUNCHB("branch_cg_direction", ni * 28.0, k_cg_direction, grid, d.nodes, first, it_base, il, d.intr_g, d.phi,
                 d.phi_old, d.cg_r, d.cg_p, d.part2, d.part3, cfg.tol, cfg.G, d.done_count);
    }
    hipLaunchKernelGGL(k_add_int, dim3(1), dim3(64), 0, stream, it_base, 5);
  };
  int it = 0;
  bool all_done = false;
  // cycle_size = min(size, 5000) per node (conjugate_gradient.c:377)
  if (!iterate_batches(batch, 5, cfg.max_iterations, count, it, all_done))
    return false;
  if (!all_done)
  {
    set_error("conjugate_gradient did not converge on level %d after %d iterations", cfg.lmin + (int)l, it);
    return false;
  }
  return true;
}

// Grow-only buffers of the batched node multigrid: level pools (pool points),
// node table (count), chunk table, per-chunk partials, active list, regions.
template <typename P, typename G>
bool Engine<P, G>::ensure_bmg_buffers(int n_itab, long long n_chunks, long long pool, hipStream_t stream)
{
  TreeDev &d = *tree->dev;
  const int count = (int)d.hbmg.size();
  if (pool > d.mg_cap)
  {
    (void)hipFree(d.mg_e);
    (void)hipFree(d.mg_r);
    (void)hipFree(d.mg_t);
    (void)hipFree(d.mg_mask);
    (void)hipFree(d.alist);
    d.mg_e = d.mg_r = d.mg_t = nullptr;
    d.mg_mask = nullptr;
    d.alist = nullptr;
    d.mg_cap = next_cap(d.mg_cap, pool);
    if (!hip_ok(hipMalloc(&d.mg_e, sizeof(G) * d.mg_cap), "multigrid pool") ||
        !hip_ok(hipMalloc(&d.mg_r, sizeof(G) * d.mg_cap), "multigrid pool") ||
        !hip_ok(hipMalloc(&d.mg_t, sizeof(G) * d.mg_cap), "multigrid pool") ||
        !hip_ok(hipMalloc((void **)&d.mg_mask, d.mg_cap), "multigrid pool") ||
        !hip_ok(hipMalloc((void **)&d.alist, sizeof(int) * d.mg_cap), "multigrid pool"))
      return false;
  }
  if (!d.acount && !hip_ok(hipMalloc((void **)&d.acount, sizeof(int) * 4), "multigrid pool"))
    return false;
  if (count > d.bmg_cap)
  {
    (void)hipFree(d.bmg);
    d.bmg_cap = std::max(count, 2 * d.bmg_cap);
    if (!hip_ok(hipMalloc((void **)&d.bmg, sizeof(BMg) * d.bmg_cap), "multigrid table"))
      return false;
  }
  const long long nreg = 2LL * count * kMaxMgLevels + 1;
  if (nreg > d.reg_cap)
  {
    (void)hipFree(d.rbeg);
    (void)hipFree(d.rstart);
    d.rbeg = d.rstart = nullptr;
    d.reg_cap = next_cap(d.reg_cap, nreg);
    if (!hip_ok(hipMalloc((void **)&d.rbeg, sizeof(int) * d.reg_cap), "regions") ||
        !hip_ok(hipMalloc((void **)&d.rstart, sizeof(int) * d.reg_cap), "regions"))
      return false;
  }
  if (n_chunks > d.chunk_cap)
  {
    // keep what is already uploaded (the dense tables precede the list tables)
    Chunk *nc = nullptr;
    long long cap = next_cap(d.chunk_cap, n_chunks);
    if (!hip_ok(hipMalloc((void **)&nc, sizeof(Chunk) * cap), "chunk table"))
      return false;
    if (d.chunks)
    {
      (void)hipMemcpyAsync(nc, d.chunks, sizeof(Chunk) * d.chunk_cap, hipMemcpyDeviceToDevice, stream);
      (void)hipStreamSynchronize(stream);
      (void)hipFree(d.chunks);
    }
    d.chunks = nc;
    d.chunk_cap = cap;
  }
  if (n_itab + 1024 > d.part_cap)
  {
    (void)hipFree(d.part1);
    (void)hipFree(d.part2);
    (void)hipFree(d.part3);
    d.part1 = d.part2 = d.part3 = nullptr;
    d.part_cap = n_itab + 1024;
    if (!hip_ok(hipMalloc((void **)&d.part1, sizeof(double) * d.part_cap), "partials") ||
        !hip_ok(hipMalloc((void **)&d.part2, sizeof(double) * d.part_cap), "partials") ||
        !hip_ok(hipMalloc((void **)&d.part3, sizeof(double) * d.part_cap), "partials"))
      return false;
  }
  return true;
}

// Batched masked multigrid for all nodes of a tree level.  Per node the box
// grid is the finest level, padded on the high side to 2^(M-1) c + 1 points per
// axis with c <= kBmgCoarseCells; unknowns are the interior points, everything
// else is Dirichlet.  The outer solver is a flexible preconditioned CG in FP64
// (convergence test of the reference's CG on every iteration), preconditioned
// by V-cycle(s) in G precision.  `solved` stays false when a box is too large
// for the hierarchy or the iteration cap is reached; the caller then runs CG
// from the current phi.  `phi_pool` / `rho_pool` are the unknowns and the
// right-hand side in the grid-pool layout; `adjoint` solves start from zero
// with zero Dirichlet values and stop on the relative residual.
// Set-up of the batched node multigrid of tree level l, on `stream` (the
// engine's stream, or the set-up stream of a run ahead): hierarchy table,
// masks, active lists, chunk and tile tables, kept in d.plan for the solve.
template <typename P, typename G> bool Engine<P, G>::bmg_prepare(size_t l, hipStream_t stream)
{
  HostSpan prep(timer, "bmg_prepare");
  TreeDev &d = *tree->dev;
  d.drop_plan();
  const int first = d.lfirst[l], count = d.lcount[l];
  std::vector<BMg> &hb = d.hbmg;
  hb.assign(count, BMg{});
  int mmax = 0;
  size_t tail_lds = 0; // LDS of the node tails (the largest over the nodes)
  long long pool = 0;
  long long npm[kMaxMgLevels] = {0};   // largest node grid at MG level m
  double ptsm[kMaxMgLevels] = {0.0};   // points over nodes at MG level m
  for (int q = 0; q < count; q++)
  {
    const DNode &hn = d.hnodes[first + q]; // box dimensions (the device table's copy)
    int M = 0, c[3] = {0, 0, 0};
    for (int t = 1; t <= kMaxMgLevels && !M; t++)
    {
      const int s = 1 << (t - 1);
      bool fit = true;
      for (int a = 0; a < 3; a++)
      {
        c[a] = (hn.rd[a] + s - 1) / s;
        fit = fit && c[a] <= kBmgCoarseCells;
      }
      if (fit)
        M = t;
    }
    if (!M)
    {
      // a box too large for the coarsest workgroup: CG solves this level
      d.plan = BmgPlan{};
      d.plan.level = (int)l;
      d.plan.cg_only = true;
      return true;
    }
    BMg &b = hb[q];
    b.nlev = M;
    b.done = 0;
    b.gx = make_fdiv((unsigned)(hn.rd[0] + 1));
    b.gxy = make_fdiv((unsigned)((hn.rd[0] + 1) * (hn.rd[1] + 1)));
    for (int m = 0; m < M; m++)
    {
      const int f = 1 << (M - 1 - m);
      b.X[m] = f * c[0] + 1;
      b.Y[m] = f * c[1] + 1;
      b.Z[m] = f * c[2] + 1;
      b.off[m] = pool;
      b.dx[m] = make_fdiv((unsigned)b.X[m]);
      b.dxy[m] = make_fdiv((unsigned)b.X[m] * (unsigned)b.Y[m]);
      const long long np = (long long)b.X[m] * b.Y[m] * b.Z[m];
      pool += np;
      npm[m] = std::max(npm[m], np);
      ptsm[m] += (double)np;
    }
    mmax = std::max(mmax, M);
    // the node's LDS tail starts at the finest level whose tail fits
    b.ts = M - 1;
    for (int t = 0; t < M; t++)
      if (bmg_tail_bytes<G>(b, t) <= kBmgTailLds)
      {
        b.ts = t;
        break;
      }
    tail_lds = std::max(tail_lds, bmg_tail_bytes<G>(b, b.ts));
  }
  // ---- chunk tables: interior points per node; per MG level the linear points
  // of every node that has the level (mask set-up)
  std::vector<Chunk> &tab = d.hchunks;
  tab.clear();
  auto add_chunks = [&](int q, long long b0, long long n, int step, int color) {
    for (long long x = b0; x < b0 + n; x += step)
      tab.push_back(Chunk{q, (int)x, (int)std::min<long long>(b0 + n, x + step), color});
  };
  for (int q = 0; q < count; q++)
  {
    hb[q].ic0 = (int)tab.size();
    add_chunks(q, 0, d.hnodes[first + q].nintr, kBmgChunkIntr, 0);
    hb[q].ic1 = (int)tab.size();
  }
  const int n_itab = (int)tab.size();
  int c0[kMaxMgLevels], cn[kMaxMgLevels];
  for (int m = 0; m < mmax; m++)
  {
    c0[m] = (int)tab.size();
    for (int q = 0; q < count; q++)
      if (hb[q].nlev > m)
        add_chunks(q, 0, (long long)hb[q].X[m] * hb[q].Y[m] * hb[q].Z[m], kBmgChunkGrid, 0);
    cn[m] = (int)tab.size() - c0[m];
  }
  const size_t dense_end = tab.size();
  if (!ensure_bmg_buffers(n_itab, (long long)dense_end + 16LL * count * mmax + 64, pool, stream))
    return false;
  G *e = (G *)d.mg_e, *r = (G *)d.mg_r;
  unsigned char *mk = d.mg_mask;
  BMg *bm = d.bmg;
  const Chunk *itab = d.chunks;
  auto C = [&](int m) { return (const Chunk *)d.chunks + c0[m]; };
  if (!hip_ok(hipMemcpyAsync(d.chunks, tab.data(), sizeof(Chunk) * dense_end, hipMemcpyHostToDevice, stream),
              "chunk table"))
    return false;
  QZ_LAUNCHB("bmg_setup", (double)pool, k_fill<unsigned char>, nblk(pool), mk, (unsigned char)0, (size_t)pool);
  QZ_LAUNCHB("bmg_setup", (double)sizeof(G) * pool, k_fill<G>, nblk(pool), e, (G)0, (size_t)pool);
  QZ_LAUNCH("bmg_setup", k_fill<int>, 1, d.done_count, 0, (size_t)2);
  if (!hip_ok(hipMemcpyAsync(bm, hb.data(), sizeof(BMg) * count, hipMemcpyHostToDevice, stream), "multigrid table"))
    return false;
  if (n_itab > 0)
    QZ_LAUNCHB("bmg_setup", 5.0 * d.lnintr[l], k_bmg_mask0, n_itab, d.nodes, first, bm, itab, d.intr_g, mk);
  for (int m = 0; m + 1 < mmax; m++)
    if (cn[m + 1] > 0)
      QZ_LAUNCHB("bmg_setup", 2.0 * ptsm[m + 1], k_bmg_mask_coarsen, cn[m + 1], bm, C(m + 1), m, mk, mk);

  // ---- active points: the pool indices of all mask bytes, in pool order, and
  // where each (node, level, colour) region starts in that list
  std::vector<int> &rbeg = d.hrbeg;
  rbeg.clear();
  for (int q = 0; q < count; q++)
    for (int m = 0; m < hb[q].nlev; m++)
    {
      const long long np = (long long)hb[q].X[m] * hb[q].Y[m] * hb[q].Z[m];
      rbeg.push_back((int)hb[q].off[m]);
      rbeg.push_back((int)(hb[q].off[m] + ((np + 1) >> 1)));
    }
  const int nreg = (int)rbeg.size();
  rbeg.push_back((int)pool);
  timer.begin(stream);
  if (!select_flags(d, mk, pool, d.alist, d.acount, stream) ||
      !hip_ok(hipMemcpyAsync(d.rbeg, rbeg.data(), sizeof(int) * (nreg + 1), hipMemcpyHostToDevice, stream),
              "regions"))
    return false;
  hipLaunchKernelGGL(k_bmg_region_starts, dim3(nblk(nreg + 1)), dim3(kBlock), 0, stream, d.alist, d.acount, d.rbeg,
                     nreg + 1, d.rstart);
  timer.end(stream, "bmg_setup", 5.0 * pool); // mask bytes in, active point indices out
  std::vector<int> rstart(nreg + 1);
  {
    HostSpan span_w(timer, "bmg_prepare_wait");
    if (!hip_ok(hipMemcpyAsync(rstart.data(), d.rstart, sizeof(int) * (nreg + 1), hipMemcpyDeviceToHost, stream),
                "regions") ||
        !hip_ok(hipStreamSynchronize(stream), "regions"))
      return false;
  }
  // active-list tables: sweeps per (level, colour) and prolongation per level over
  // the nodes that smooth there with the per-level kernels (m < ts: the levels
  // from ts on run in the node's tail); restriction into level m over the nodes
  // whose tail starts at or below it
  int sw0[kMaxMgLevels][2], swn[kMaxMgLevels][2], rs0[kMaxMgLevels], rsn[kMaxMgLevels];
  double act_s[kMaxMgLevels] = {0.0}, act_r[kMaxMgLevels] = {0.0}; // active points per table
  double act_f[kMaxMgLevels] = {0.0}, tail_act = 0.0; // fused levels and tail top levels (byte accounting)
  std::vector<int> reg0(count); // first region of each node
  for (int q = 0, rr = 0; q < count; q++)
  {
    reg0[q] = rr;
    rr += 2 * hb[q].nlev;
  }
  // node levels m >= 1 that the per-level kernels would smooth run as the fused
  // pair k_bmg_down / k_bmg_up (default sweep counts): tiles per (level, node).
  // Level 0 stays with the active-list kernels: its tiles stage about ten
  // times the points they own, which costs more there than the launches save.
  const bool nfuse = node_mg_fused && cfg.pre_sweeps == 2 && cfg.post_sweeps == 2;
  auto fused = [&](int m) { return nfuse && m >= 1; };
  int nt0[kMaxMgLevels] = {0}, ntn[kMaxMgLevels] = {0};
  std::vector<NTile> &ntab = d.hntiles;
  ntab.clear();
  for (int q = 0, rr = 0; q < count; rr += 2 * hb[q].nlev, q++)
  {
    tail_act += rstart[rr + 2 * hb[q].ts + 2] - rstart[rr + 2 * hb[q].ts];
    for (int m = 1; m < hb[q].ts && m + 1 < hb[q].nlev; m++)
      if (fused(m))
        act_f[m] += rstart[rr + 2 * m + 2] - rstart[rr + 2 * m];
  }
  for (int m = 1; m < mmax; m++)
  {
    nt0[m] = (int)ntab.size();
    for (int q = 0; fused(m) && q < count; q++)
      if (hb[q].nlev > m + 1 && m < hb[q].ts)
      {
        // tiles of kNmCT coarse points cover the fine points [1, xm] per axis
        // that can be active: x 2^m is an interior box point, so x <= (rd - 1) >> m
        const DNode &hn = d.hnodes[first + q];
        auto tiles = [&](int a, int Xm) { return (std::min(Xm - 2, (hn.rd[a] - 1) >> m) + 7) / 8; };
        const int tx = tiles(0, hb[q].X[m]), ty = tiles(1, hb[q].Y[m]), tz = tiles(2, hb[q].Z[m]);
        for (int c = 0; c < tz; c++)
          for (int b2 = 0; b2 < ty; b2++)
            for (int a = 0; a < tx; a++)
              ntab.push_back(NTile{q, 1 + kNmCT * a, 1 + kNmCT * b2, 1 + kNmCT * c});
      }
    ntn[m] = (int)ntab.size() - nt0[m];
  }
  for (int m = 0; m < mmax; m++)
  {
    for (int c = 0; c < 2; c++)
    {
      sw0[m][c] = (int)tab.size();
      for (int q = 0; q < count; q++)
        if (hb[q].nlev > m + 1 && m < hb[q].ts && !fused(m))
        {
          const int rg = reg0[q] + 2 * m + c;
          add_chunks(q, rstart[rg], rstart[rg + 1] - rstart[rg], kBmgChunkAct, c);
          act_s[m] += rstart[rg + 1] - rstart[rg];
        }
      swn[m][c] = (int)tab.size() - sw0[m][c];
    }
    rs0[m] = (int)tab.size();
    for (int q = 0; q < count; q++)
      if (hb[q].nlev > m && m <= hb[q].ts && !(m >= 1 && fused(m - 1)))
      {
        const int rg = reg0[q] + 2 * m;
        add_chunks(q, rstart[rg], rstart[rg + 2] - rstart[rg], kBmgChunkAct, 0);
        act_r[m] += rstart[rg + 2] - rstart[rg];
      }
    rsn[m] = (int)tab.size() - rs0[m];
  }
  if (!ensure_bmg_buffers(n_itab, (long long)tab.size(), pool, stream) ||
      !hip_ok(hipMemcpyAsync(d.chunks + dense_end, tab.data() + dense_end, sizeof(Chunk) * (tab.size() - dense_end),
                             hipMemcpyHostToDevice, stream),
              "chunk table"))
    return false;
  itab = d.chunks;
  if ((long long)ntab.size() > d.ntile_cap)
  {
    (void)hipFree(d.ntiles);
    d.ntile_cap = next_cap(d.ntile_cap, (long long)ntab.size());
    if (!hip_ok(hipMalloc((void **)&d.ntiles, sizeof(NTile) * d.ntile_cap), "node tiles"))
      return false;
  }
  if (!ntab.empty() &&
      !hip_ok(hipMemcpyAsync(d.ntiles, ntab.data(), sizeof(NTile) * ntab.size(), hipMemcpyHostToDevice, stream),
              "node tiles"))
    return false;
  // the tail's LDS exceeds the default dynamic limit: raised once per process
  // (a function-local static is initialised once, also with set-ups on two threads)
  static const hipError_t tail_attr =
      hipFuncSetAttribute((const void *)k_bmg_tail<G>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBmgTailLds);
  if (!hip_ok(tail_attr, "node tail LDS"))
    return false;
  BmgPlan &pl = d.plan;
  pl.cg_only = false;
  pl.ahead = stream != this->stream;
  pl.nfuse = nfuse;
  pl.n_itab = n_itab;
  pl.mmax = mmax;
  pl.tail_lds = tail_lds;
  std::memcpy(pl.sw0, sw0, sizeof(sw0));
  std::memcpy(pl.swn, swn, sizeof(swn));
  std::memcpy(pl.rs0, rs0, sizeof(rs0));
  std::memcpy(pl.rsn, rsn, sizeof(rsn));
  std::memcpy(pl.nt0, nt0, sizeof(nt0));
  std::memcpy(pl.ntn, ntn, sizeof(ntn));
  std::memcpy(pl.act_s, act_s, sizeof(act_s));
  std::memcpy(pl.act_r, act_r, sizeof(act_r));
  std::memcpy(pl.act_f, act_f, sizeof(act_f));
  pl.tail_act = tail_act;
  pl.level = (int)l;
  return true;
}

// Wait for a set-up run ahead (if one is running); false when it failed.
template <typename P, typename G> bool Engine<P, G>::join_ahead()
{
  if (!tree || !tree->dev || !tree->dev->ahead.joinable())
    return true;
  TreeDev &d = *tree->dev;
  d.ahead.join();
  if (!d.ahead_ok)
  {
    d.drop_plan();
    set_error("%s", d.ahead_err.c_str());
  }
  return d.ahead_ok;
}

// Start work on the set-up stream and a host thread, ordered after the work
// queued so far on the engine's stream: with build, the tree below the head
// (build_tree), then, when the tree has levels, the first level's node
// multigrid set-up and its captured PCG batch.  ahead_done marks the end of the
// stream work the engine's stream must wait for.
//
// Until join_ahead() the thread owns the host tree and every TreeDev buffer:
// the head child map, flag and zone lists, node table, zone, cell, box and
// point pools, the node field pools it zeroes (phi, ax, ay, az), the partial
// sums part1-3 and mass partials it may reallocate, the classification tables,
// and the node multigrid pools and plan.  The engine's stream meanwhile uses
// only the particle arrays, the head grids and the engine's own reduction
// buffers (part, red); every stage that reads the tree joins first.
template <typename P, typename G> bool Engine<P, G>::start_ahead(bool build)
{
  TreeDev &d = *tree->dev;
  // non-blocking: a capture on it is not disturbed by legacy-stream calls of other threads
  if (!d.ahead_stream && (!hip_ok(hipStreamCreateWithFlags(&d.ahead_stream, hipStreamNonBlocking), "set-up stream") ||
                          !hip_ok(hipEventCreateWithFlags(&d.ahead_tree, hipEventDisableTiming), "set-up event") ||
                          !hip_ok(hipEventCreateWithFlags(&d.ahead_done, hipEventDisableTiming), "set-up event")))
    return false;
  // the set-up stream's wait is queued here, before this thread can begin a
  // graph capture on the engine's stream (the head solve): a wait issued during
  // that capture, on an event of the capturing stream, is refused
  if (!hip_ok(hipEventRecord(d.ahead_tree, stream), "set-up event") ||
      !hip_ok(hipStreamWaitEvent(d.ahead_stream, d.ahead_tree, 0), "set-up stream"))
    return false;
  d.ahead_ok = true;
  d.ahead = std::thread([this, &d, build]() {
    HostSpan span(timer, build ? "ahead_tree_and_setup" : "ahead_setup");
    d.ahead_ok = !build || build_tree(d.ahead_stream);
    const bool levels = d.ahead_ok && d.lfirst.size() > 1 && cfg.branch_solver == QZ_BRANCH_MG;
    d.ahead_ok = d.ahead_ok && (!levels || bmg_prepare(1, d.ahead_stream)) &&
                 hip_ok(hipEventRecord(d.ahead_done, d.ahead_stream), "set-up event");
    // the PCG batch of the level's potential solve, captured on the set-up stream
    // (after ahead_done: nothing recorded during the capture is work of its own)
    if (d.ahead_ok && levels && !d.plan.cg_only && batch_graphs)
    {
      hipGraph_t g = nullptr;
      d.ahead_ok = hip_ok(hipStreamBeginCapture(d.ahead_stream, hipStreamCaptureModeRelaxed), "graph capture");
      if (d.ahead_ok)
      {
        bmg_batch(1, d.plan, d.ahead_stream, d.phi, false);
        d.ahead_ok = hip_ok(hipStreamEndCapture(d.ahead_stream, &g), "graph capture") &&
                     hip_ok(hipGraphInstantiate(&d.plan.gexec, g, nullptr, nullptr, 0), "graph instantiate");
      }
      if (g)
        (void)hipGraphDestroy(g);
    }
    if (!d.ahead_ok)
      d.ahead_err = g_last_error;
  });
  return true;
}

// One batch of the node PCG of tree level l with set-up pl, issued on `stream`:
// two iterations of check, V-cycle, dots, direction, apply and update, then the
// iteration counter.  Captured into a graph (on the engine's stream, or ahead on
// the set-up stream) or launched eagerly while kernels are timed.
template <typename P, typename G>
void Engine<P, G>::bmg_batch(size_t l, const BmgPlan &pl, hipStream_t stream, double *phi_pool, bool adjoint)
{
  TreeDev &d = *tree->dev;
  const int first = d.lfirst[l], count = d.lcount[l];
  const int n_itab = pl.n_itab, mmax = pl.mmax;
  const size_t tail_lds = pl.tail_lds;
  const auto &sw0 = pl.sw0;
  const auto &swn = pl.swn;
  const auto &rs0 = pl.rs0;
  const auto &rsn = pl.rsn;
  const auto &nt0 = pl.nt0;
  const auto &ntn = pl.ntn;
  const auto &act_s = pl.act_s;
  const auto &act_r = pl.act_r;
  const auto &act_f = pl.act_f;
  auto fused = [&](int m) { return pl.nfuse && m >= 1; };
  const int inner = 1; // one V-cycle per PCG iteration (a second one costs more than the iterations it saves)
  G *e = (G *)d.mg_e, *r = (G *)d.mg_r, *et = (G *)d.mg_t;
  unsigned char *mk = d.mg_mask;
  BMg *bm = d.bmg;
  const Chunk *itab = d.chunks;
  auto SW = [&](int m, int c) { return (const Chunk *)d.chunks + sw0[m][c]; };
  auto RS = [&](int m) { return (const Chunk *)d.chunks + rs0[m]; };

  const double ni = (double)d.lnintr[l], gs = (double)sizeof(G);
  const int pre = std::max(1, cfg.pre_sweeps), post = std::max(0, cfg.post_sweeps);
  int *it_base = d.done_count + 1;
  auto sweep = [&](int m, int color, int zs) {
    // per active point of the colour: list entry, r, e written, 6 neighbours of
    // the other colour (read once over the sweep)
    if (swn[m][color] > 0)
      QZ_LAUNCHB("bmg_rbgs", 0.5 * act_s[m] * (4.0 + 3.0 * gs), (k_bmg_rbgs<G>), swn[m][color], d.nodes, first, bm,
                 SW(m, color), d.alist, m, e, r, color, zs);
  };
  auto vcycle = [&]() {
    for (int c = 0; c < inner; c++)
    {
      for (int m = 0; m + 1 < mmax; m++)
      {
        if (fused(m))
        {
          if (ntn[m] > 0)
            // reads r and the mask bits of the level, writes the pre-smoothed
            // correction and the next level's residual
            QZ_LAUNCH_TB("bmg_down", kNmThreads, act_f[m] * (2.0 * gs + 0.125) + act_r[m + 1] * gs, (k_bmg_down<G>),
                         ntn[m], d.nodes, first, bm,
                         (const NTile *)d.ntiles + nt0[m], m, r, et, (const unsigned char *)mk);
          continue;
        }
        for (int sw = 0; sw < pre; sw++)
        {
          sweep(m, 0, sw == 0 && (m > 0 || c == 0));
          sweep(m, 1, 0);
        }
        if (rsn[m + 1] > 0)
          QZ_LAUNCHB("bmg_restrict", act_s[m] * (1.0 + 2.0 * gs) + act_r[m + 1] * (4.0 + gs), (k_bmg_restrict<G>),
                     rsn[m + 1], d.nodes, first, bm, RS(m + 1), d.alist, m, e, r, mk, r);
      }
      timer.begin(stream);
      hipLaunchKernelGGL(k_bmg_tail<G>, dim3(count), dim3(kBmgTailThreads), tail_lds, stream, d.nodes, first, bm, e,
                         r, mk, pre, post, 100);
      timer.end(stream, "bmg_tail", pl.tail_act * 2.0 * gs); // r in, e out at each tail's top level
      for (int m = mmax - 2; m >= 0; m--)
      {
        if (fused(m))
        {
          if (ntn[m] > 0)
            // reads the coarse correction, the pre-smoothed correction, r and the
            // mask bits, writes the correction
            QZ_LAUNCH_TB("bmg_up", kNmThreads, act_r[m + 1] * gs + act_f[m] * (3.0 * gs + 0.125), (k_bmg_up<G>),
                         ntn[m], d.nodes, first, bm,
                         (const NTile *)d.ntiles + nt0[m], m, e, (const G *)et, (const G *)r,
                         (const unsigned char *)mk);
          continue;
        }
        const int np_ = swn[m][0] + swn[m][1];
        if (np_ > 0)
          QZ_LAUNCHB("bmg_prolong", act_s[m] * (4.0 + 2.0 * gs) + act_r[m + 1] * gs, (k_bmg_prolong<G>), np_, bm,
                     SW(m, 0), d.alist, m, e, e);
        for (int sw = 0; sw < post; sw++)
        {
          sweep(m, 0, 0);
          sweep(m, 1, 0);
        }
      }
    }
  };
  for (int il = 0; il < 2; il++)
  {
    QZ_LAUNCHB("bmg_check", (double)count * (sizeof(BMg) + sizeof(DNode) + 32.0), k_bmg_check, count, bm, d.nodes,
               first, it_base, il, d.part2, d.part3, cfg.tol, cfg.G,
              adjoint ? 1 : 0, d.done_count);
    if (n_itab == 0)
      continue;
    vcycle();
    QZ_LAUNCHB("bmg_zdots", ni * (4.0 + gs + 16.0), (k_bmg_zdots<G>), n_itab, d.nodes, first, bm, itab, d.intr_g,
               (const G *)e, d.cg_r, d.cg_Ap, d.part1, d.part3);
    QZ_LAUNCHB("bmg_direction", ni * (4.0 + gs + 16.0), (k_bmg_direction<G>), n_itab, d.nodes, first, bm, itab,
               it_base, il, d.intr_g, (const G *)e, d.cg_p, d.part1, d.part3);
    QZ_LAUNCHB("bmg_apply", ni * 20.0, k_bmg_apply, n_itab, d.nodes, first, bm, itab, d.intr_g, d.cg_p, d.cg_Ap,
               d.part1);
    QZ_LAUNCHB("bmg_update", ni * (4.0 + 48.0 + gs), (k_bmg_update<G>), n_itab, d.nodes, first, bm, itab, it_base,
               il, d.intr_g, phi_pool, d.cg_p, d.cg_Ap, d.cg_r, r, d.part1, d.part2, d.part3, cfg.tol_change);
  }
  hipLaunchKernelGGL(k_add_int, dim3(1), dim3(64), 0, stream, it_base, 2);
}

template <typename P, typename G>
bool Engine<P, G>::branch_mg_level(size_t l, bool &solved, double *phi_pool, const G *rho_pool, bool adjoint)
{
  solved = false;
  if (!join_ahead())
    return false;
  TreeDev &d = *tree->dev;
  if (d.plan.level != (int)l && !bmg_prepare(l, stream))
    return false;
  const BmgPlan pl = d.plan; // a set-up serves one solve (its graph moves to this solve)
  d.plan.gexec = nullptr;
  d.plan.level = -1;
  if (pl.cg_only)
    return true;
  if (pl.ahead && !hip_ok(hipStreamWaitEvent(stream, d.ahead_done, 0), "set-up stream"))
  {
    if (pl.gexec)
      (void)hipGraphExecDestroy(pl.gexec);
    return false;
  }
  const int first = d.lfirst[l], count = d.lcount[l];
  const int n_itab = pl.n_itab;
  G *r = (G *)d.mg_r;
  BMg *bm = d.bmg;
  const Chunk *itab = d.chunks;
  const double ni = (double)d.lnintr[l], gs = (double)sizeof(G);
  // Flexible PCG in FP64 with the V-cycle as preconditioner.  Algorithmic bytes
  // per interior point: residual reads idx, phi, rho and writes r (FP64 + G);
  // zdots reads idx, z, r, q; direction reads idx, z, p, writes p; apply reads
  // idx, p, writes q; update reads idx, phi, p, q, r, writes phi, r (FP64 + G).
  QZ_LAUNCHB("bmg_setup", 8.0 * d.lng[l], k_fill<double>, nblk(d.lng[l]), d.cg_p + d.lg0[l], 0.0, (size_t)d.lng[l]);
  if (n_itab > 0)
    QZ_LAUNCHB("bmg_residual", ni * (4.0 + 8.0 + gs + 8.0 + gs), (k_bmg_residual<G>), n_itab, d.nodes, first, bm,
               itab, d.intr_g, phi_pool, rho_pool, d.cg_r, r, d.part2);
  // the batch captured ahead was built for this level's potential solve
  hipGraphExec_t given = pl.gexec;
  if (given && (phi_pool != d.phi || adjoint))
  {
    (void)hipGraphExecDestroy(given);
    given = nullptr;
  }
  auto batch = [&]() { bmg_batch(l, pl, stream, phi_pool, adjoint); };
  int it = 0;
  bool all_done = false;
  HostSpan span_it(timer, "bmg_iterate");
  if (!iterate_batches(batch, 2, kBmgMaxOuter, count, it, all_done, given))
    return false;
  solved = all_done;
  return true;
}

template <typename P, typename G> bool Engine<P, G>::branch_grid_acceleration()
{
  TreeDev &d = *tree->dev;
  const int zero_bc = cfg.head_bc == QZ_BC_ZERO ? 1 : 0;
  for (size_t l = 1; l < d.lfirst.size(); l++)
  {
    dim3 grid(blocks_per_node(d.lpmax[l]), d.lcount[l]);
    timer.begin(stream);
    hipLaunchKernelGGL((k_branch_grid_acc<G>), grid, dim3(kBlock), 0, stream, d.nodes, d.lfirst[l], d.sim_g,
                       d.sim_xyz, d.bdry_g, d.bdry_xyz, d.intr_g, d.phi, (G *)d.ax, (G *)d.ay, (G *)d.az, gax, gay, gaz,
                       n, cfg.G * total_mass, zero_bc);
    timer.end(stream, "grid_acc_branch", (8.0 + 3.0 * sizeof(G)) * d.lng[l]);
  }
  return true;
}

template <typename P, typename G> bool Engine<P, G>::branch_particle_acceleration()
{
  TreeDev &d = *tree->dev;
  // per particle: position in, its head cell's map entry, acceleration out;
  // per grid point of the head and the nodes: the three accelerations
  const double npts = (double)(n + 1) * (n + 1) * (n + 1);
  QZ_LAUNCHB("gather_amr", (28.0 + 3.0 * sizeof(P)) * N + 3.0 * sizeof(G) * (npts + (double)d.used_grid),
             (k_gather_amr<P, G>), nblk(N), pos[0], pos[1], pos[2], gax, gay, gaz, d.head_child, d.nodes,
            d.box, (G *)d.ax, (G *)d.ay, (G *)d.az, pa[6], pa[7], pa[8], N, cfg.lmin, plev);
  lev_head = false;
  return true;
}

template <typename P, typename G> void Engine<P, G>::branch_grid_pe(unsigned nb)
{
  TreeDev &d = *tree->dev;
  QZ_LAUNCH("grid_pe", k_grid_pe_amr<P>, nb, pos[0], pos[1], pos[2], pa[9], phi, d.head_child, d.nodes, d.box, d.phi,
            N, cfg.lmin, part);
}

template <typename P, typename G> bool Engine<P, G>::node_header(int i, int *h)
{
  if (!has_tree() || i < 0 || i >= tree->n_nodes)
  {
    set_error("node index out of range");
    return false;
  }
  if (!join_ahead())
    return false;
  // refresh the host mirror: counts and masses are filled on the device after the build
  if (i == 0 && !hip_ok(hipMemcpy(tree->dev->hnodes.data(), tree->dev->nodes, sizeof(DNode) * tree->n_nodes,
                                  hipMemcpyDeviceToHost),
                        "node table"))
    return false;
  const HNode &hn = tree->host->nodes[i];
  const DNode &dn = tree->dev->hnodes[i];
  int lvrel = hn.lv - cfg.lmin;
  const auto &ls = tree->host->lev_start;
  int parent_pos = hn.parent < 0 ? 0 : hn.parent - ls[lvrel - 1];
  int nch = 0;
  for (const auto &o : tree->host->nodes)
    if (o.parent == i)
      nch++;
  int hdr[24] = {hn.lv, lvrel, i - ls[lvrel], parent_pos, hn.ts[0], hn.ts[1], hn.ts[2], hn.rd[0], hn.rd[1], hn.rd[2],
                 hn.bmin[0], hn.bmin[1], hn.bmin[2], hn.bmax[0], hn.bmax[1], hn.bmax[2], hn.ncell, dn.nintr, dn.nbdry,
                 dn.nsim, dn.nptcl, 0, nch, hn.gcap};
  // particles outside the refinement zones (tree_construction.c:2286-2287)
  int inside = 0;
  for (const auto &o : tree->host->nodes)
    if (o.parent == i)
      inside += tree->dev->hnodes[&o - &tree->host->nodes[0]].nptcl;
  hdr[21] = dn.nptcl - inside;
  memcpy(h, hdr, sizeof(hdr));
  return true;
}

template <typename P, typename G> bool Engine<P, G>::node_lists(int i, int *ixyz, int *bxyz, int *sxyz)
{
  if (!has_tree() || i < 0 || i >= tree->n_nodes)
  {
    set_error("node index out of range");
    return false;
  }
  const DNode &dn = tree->dev->hnodes[i];
  TreeDev &d = *tree->dev;
  return hip_ok(hipMemcpy(ixyz, d.intr_xyz + 3 * (size_t)dn.poff, sizeof(int) * 3 * dn.nintr, hipMemcpyDeviceToHost),
                "lists") &&
         hip_ok(hipMemcpy(bxyz, d.bdry_xyz + 3 * (size_t)dn.poff, sizeof(int) * 3 * dn.nbdry, hipMemcpyDeviceToHost),
                "lists") &&
         hip_ok(hipMemcpy(sxyz, d.sim_xyz + 3 * (size_t)dn.poff, sizeof(int) * 3 * dn.nsim, hipMemcpyDeviceToHost),
                "lists");
}

template <typename P, typename G> bool Engine<P, G>::node_fields(int i, double *out5)
{
  if (!has_tree() || i < 0 || i >= tree->n_nodes)
  {
    set_error("node index out of range");
    return false;
  }
  const DNode &dn = tree->dev->hnodes[i];
  const HNode &hn = tree->host->nodes[i];
  TreeDev &d = *tree->dev;
  size_t g = hn.gcap;
  std::vector<G> tmp(g);
  G *src[5] = {(G *)d.rho, nullptr, (G *)d.ax, (G *)d.ay, (G *)d.az};
  for (int a = 0; a < 5; a++)
  {
    if (a == 1)
    {
      if (!hip_ok(hipMemcpy(out5 + g, d.phi + dn.goff, sizeof(double) * g, hipMemcpyDeviceToHost), "fields"))
        return false;
      continue;
    }
    if (!hip_ok(hipMemcpy(tmp.data(), src[a] + dn.goff, sizeof(G) * g, hipMemcpyDeviceToHost), "fields"))
      return false;
    for (size_t t = 0; t < g; t++)
      out5[a * g + t] = (double)tmp[t];
  }
  return true;
}

template <typename P, typename G> bool Engine<P, G>::timestep(double cfl, double dt_max, double *dt)
{
  unsigned nb = nblk(N);
  // the minima the last kick folded in, when they are for this CFL number and
  // nothing has changed since; otherwise one pass over the particles
  const bool folded = dt_ready && cfl == dt_cfl;
  double *src = folded ? dtpart : part;
  if (!folded)
  {
    const int *hc = has_tree() ? tree->dev->head_child : nullptr;
    const DNode *nd = has_tree() ? tree->dev->nodes : nullptr;
    const int *bx = has_tree() ? tree->dev->box : nullptr;
    QZ_LAUNCHB("cfl_dt", 9.0 * sizeof(P) * N, k_cfl_dt<P>, nb, pos[0], pos[1], pos[2], pa[3], pa[4], pa[5], pa[6],
               pa[7], pa[8], N, hc, nd, bx, cfg.lmin, cfl, dt_max, part);
  }
  // fold the per-block minima through max(-dt) in two stages, the first into
  // part past the minima (part holds at least 4 nb values) or into part when
  // the minima are the kick's
  const unsigned ns = std::min<unsigned>(kBlock, (nb + kBlock - 1) / kBlock);
  const int chunk = (int)((nb + ns - 1) / ns);
  double *stage = src == part ? part + nb : part;
  hipLaunchKernelGGL(k_reduce_neg_max_stage, dim3(ns), dim3(kBlock), 0, stream, src, (int)nb, chunk, stage);
  hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kBlock), 0, stream, stage, (int)ns, red);
  dt_ready = false;
  dt_cfl = cfl; // the CFL number the next kick folds in
  if (!hip_ok(hipMemcpyAsync(hred, red, sizeof(double), hipMemcpyDeviceToHost, stream), "readback") ||
      !hip_ok(hipStreamSynchronize(stream), "timestep"))
    return false;
  // min(dt_max, every particle's dt), as k_cfl_dt starts each particle from dt_max
  *dt = folded ? std::min(dt_max, -hred[0]) : -hred[0];
  return true;
}


// Copy the tree state of the current force pass into `t` (allocates only when
// it grows).
template <typename P, typename G> bool Engine<P, G>::snap_tree(TreeSnap *&t)
{
  if (!t)
    t = new TreeSnap();
  TreeDev &d = *tree->dev;
  const size_t ncell_head = (size_t)n * n * n, gs = sizeof(G);
  auto put = [&](TreeSnap::Buf &b, const void *src, size_t bytes) -> bool {
    if (bytes > b.cap)
    {
      (void)hipFree(b.p);
      b.p = nullptr;
      b.cap = 0;
      if (!hip_ok(hipMalloc(&b.p, std::max<size_t>(bytes + bytes / 4, 256)), "tree record"))
        return false;
      b.cap = bytes + bytes / 4;
    }
    return bytes == 0 || hip_ok(hipMemcpyAsync(b.p, src, bytes, hipMemcpyDeviceToDevice, stream), "tree record");
  };
  const size_t nb = (size_t)d.used_box, nc = (size_t)d.used_cell, np = (size_t)d.used_pt, ng = (size_t)d.used_grid;
  timer.begin(stream);
  struct End
  {
    Timer &t;
    hipStream_t s;
    ~End() { t.end(s, "tree_snap"); }
  } end_{timer, stream};
  t->n_nodes = tree->n_nodes;
  t->used_box = d.used_box;
  t->used_cell = d.used_cell;
  t->used_pt = d.used_pt;
  t->used_grid = d.used_grid;
  t->hnodes = d.hnodes;
  t->lfirst = d.lfirst;
  t->lcount = d.lcount;
  t->lnbm = d.lnbm;
  t->lc0 = d.lc0;
  t->lncell = d.lncell;
  t->lb0 = d.lb0;
  t->lnb = d.lnb;
  t->lz0 = d.lz0;
  t->lnz = d.lnz;
  t->lg0 = d.lg0;
  t->lng = d.lng;
  t->lpmax = d.lpmax;
  t->lnintr = d.lnintr;
  return put(t->nodes, d.nodes, sizeof(DNode) * (size_t)tree->n_nodes) &&
         put(t->head_child, d.head_child, sizeof(int) * ncell_head) && put(t->box, d.box, sizeof(int) * nb) &&
         put(t->bstart, d.bstart, sizeof(int) * nb) && put(t->bend, d.bend, sizeof(int) * nb) &&
         put(t->cstart, d.cstart, sizeof(int) * nc) && put(t->cend, d.cend, sizeof(int) * nc) &&
         put(t->intr_g, d.intr_g, sizeof(int) * np) && put(t->bdry_g, d.bdry_g, sizeof(int) * np) &&
         put(t->sim_g, d.sim_g, sizeof(int) * np) && put(t->intr_xyz, d.intr_xyz, 3 * sizeof(int) * np) &&
         put(t->bdry_xyz, d.bdry_xyz, 3 * sizeof(int) * np) && put(t->sim_xyz, d.sim_xyz, 3 * sizeof(int) * np) &&
         put(t->ax, d.ax, gs * ng) && put(t->ay, d.ay, gs * ng) && put(t->az, d.az, gs * ng);
}

// Make a recorded tree current (into the live grow-only pools).
template <typename P, typename G> bool Engine<P, G>::restore_tree(const TreeSnap *t)
{
  if (!join_ahead())
    return false;
  TreeDev &d = *tree->dev;
  d.drop_plan();
  if (t->n_nodes > d.node_cap || t->used_box > d.box_cap || t->used_cell > d.cell_cap || t->used_pt > d.pt_cap ||
      t->used_grid > d.grid_cap)
  {
    set_error("restore_tree: a recorded tree exceeds the live pools");
    return false;
  }
  const size_t ncell_head = (size_t)n * n * n, gs = sizeof(G);
  auto get = [&](void *dst, const TreeSnap::Buf &b, size_t bytes) -> bool {
    return bytes == 0 || hip_ok(hipMemcpyAsync(dst, b.p, bytes, hipMemcpyDeviceToDevice, stream), "tree record");
  };
  const size_t nb = (size_t)t->used_box, nc = (size_t)t->used_cell, np = (size_t)t->used_pt, ng = (size_t)t->used_grid;
  timer.begin(stream);
  struct End
  {
    Timer &t;
    hipStream_t s;
    ~End() { t.end(s, "tree_restore"); }
  } end_{timer, stream};
  d.hc_set = -1; // the recorded map's cells are not those in hlist
  if (!get(d.nodes, t->nodes, sizeof(DNode) * (size_t)t->n_nodes) ||
      !get(d.head_child, t->head_child, sizeof(int) * ncell_head) || !get(d.box, t->box, sizeof(int) * nb) ||
      !get(d.bstart, t->bstart, sizeof(int) * nb) || !get(d.bend, t->bend, sizeof(int) * nb) ||
      !get(d.cstart, t->cstart, sizeof(int) * nc) || !get(d.cend, t->cend, sizeof(int) * nc) ||
      !get(d.intr_g, t->intr_g, sizeof(int) * np) || !get(d.bdry_g, t->bdry_g, sizeof(int) * np) ||
      !get(d.sim_g, t->sim_g, sizeof(int) * np) || !get(d.intr_xyz, t->intr_xyz, 3 * sizeof(int) * np) ||
      !get(d.bdry_xyz, t->bdry_xyz, 3 * sizeof(int) * np) || !get(d.sim_xyz, t->sim_xyz, 3 * sizeof(int) * np) ||
      !get(d.ax, t->ax, gs * ng) || !get(d.ay, t->ay, gs * ng) || !get(d.az, t->az, gs * ng))
    return false;
  tree->n_nodes = t->n_nodes;
  d.used_box = t->used_box;
  d.used_cell = t->used_cell;
  d.used_pt = t->used_pt;
  d.used_grid = t->used_grid;
  d.hnodes = t->hnodes;
  d.lfirst = t->lfirst;
  d.lcount = t->lcount;
  d.lnbm = t->lnbm;
  d.lc0 = t->lc0;
  d.lncell = t->lncell;
  d.lb0 = t->lb0;
  d.lnb = t->lnb;
  d.lz0 = t->lz0;
  d.lnz = t->lnz;
  d.lg0 = t->lg0;
  d.lng = t->lng;
  d.lpmax = t->lpmax;
  d.lnintr = t->lnintr;
  return true;
}

template <typename P, typename G> void Engine<P, G>::free_tree_snap(TreeSnap *&t)
{
  if (!t)
    return;
  t->release();
  delete t;
  t = nullptr;
}

} // namespace qz
