// Host side of the refinement tree (see qz_tree.h).  Restates the marking,
// zoning and child-box steps of tree_construction.c (non-periodic boundaries)
// for nodes below the head; the head level's marks and zones are found on the
// device and arrive here as zone bounding boxes:
//   fill_cell_ref       tree_construction.c:508-598
//   fill_zones_ref      tree_construction.c:685-926
//   create_child_nodes  tree_construction.c:1032-1661 (box geometry part)
// The node is rebuilt from scratch every step instead of adapted
// (tree_adaptation.c); fields are compared at global grid coordinates.
#include "qz_tree.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace qz
{

void TreeHost::init(int lmin_, int lmax_, int crit)
{
  lmin = lmin_;
  lmax = lmax_;
  n = 1 << lmin;
  ref_crit = crit;
  extra = 10; // min_box_extra_size_per_side_real for n_exp = 1 (global_variables.c:294-295)
  nodes.clear();
  lev_start.assign(1, 0);
  lev_count.assign(1, 1);
}

void TreeHost::clear()
{
  nodes.clear();
  lev_start.assign(1, 0);
  lev_count.assign(1, 1);
}

// 6-connected flood fill of the -1 cells (fill_zones_ref, tree_construction.c:748-891).
// box: statuses, rX/rXY strides, cell_ref: box indices of marked cells in cell order.
static void zone_fill(int *box, int rX, int rXY, const std::vector<int> &ref_box,
                      std::vector<int> &zone_size, std::vector<std::vector<int>> &zones_by_ref)
{
  int nref = (int)ref_box.size();
  std::vector<int> aux(nref > 0 ? nref : 1);
  int added_all = 0, ref_i = 0, zone = 0;
  zone_size.clear();
  while (nref > added_all)
  {
    int b = ref_box[ref_i];
    if (box[b] == -1)
    {
      int zs = 0;
      aux[0] = b;
      box[b] = zone;
      zs++;
      added_all++;
      int insp = 0;
      while (zs > insp && nref > added_all)
      {
        int add = 0;
        int c = aux[insp];
        const int nbs[6] = {c + 1, c - 1, c + rX, c - rX, c + rXY, c - rXY};
        for (int q = 0; q < 6; q++)
        {
          if (box[nbs[q]] == -1)
          {
            aux[zs + add] = nbs[q];
            box[nbs[q]] = zone;
            add++;
          }
        }
        zs += add;
        added_all += add;
        insp++;
      }
      zone_size.push_back(zs);
      zone++;
    }
    ref_i++;
  }
  // zone member lists in cell_ref order (tree_construction.c:906-914)
  zones_by_ref.assign(zone, {});
  for (int z = 0; z < zone; z++)
    zones_by_ref[z].reserve(zone_size[z]);
  for (int i = 0; i < nref; i++)
    zones_by_ref[box[ref_box[i]]].push_back(i);
}

// Persistent host workers for the tree passes (thread start-up per pass cost
// more than the passes of the small levels): kTreeThreads - 1 workers and the
// calling thread take the indices of one job at a time; callers on several
// threads are served in turn.
class TreePool
{
public:
  static TreePool &get()
  {
    static TreePool pool;
    return pool;
  }
  int workers() const { return (int)th_.size(); }
  void run(int count, const std::function<void(int)> &f)
  {
    std::lock_guard<std::mutex> one(job_);
    {
      std::lock_guard<std::mutex> lk(mx_);
      f_ = &f;
      count_ = count;
      next_.store(0);
      active_ = (int)th_.size();
      gen_++;
    }
    wake_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mx_);
    idle_.wait(lk, [&] { return active_ == 0; }); // every worker has left this job
    f_ = nullptr;
  }
  ~TreePool()
  {
    {
      std::lock_guard<std::mutex> lk(mx_);
      stop_ = true;
      gen_++;
    }
    wake_.notify_all();
    for (auto &t : th_)
      t.join();
  }

private:
  TreePool()
  {
    const int hw = std::max(1u, std::thread::hardware_concurrency());
    // QZ_TREE_THREADS caps the threads (the caller's included), for hosts that
    // keep their own cores busy meanwhile (an OpenMP program around the library)
    int cap = kTreeThreads;
    if (const char *f = std::getenv("QZ_TREE_THREADS"))
      cap = std::max(1, std::min(kTreeThreads, std::atoi(f)));
    const int nw = std::max(0, std::min(cap, hw) - 1);
    for (int t = 0; t < nw; t++)
      th_.emplace_back([this]() { loop(); });
  }
  void work()
  {
    for (int q = next_++; q < count_; q = next_++)
      (*f_)(q);
  }
  void loop()
  {
    unsigned long seen = 0;
    for (;;)
    {
      std::unique_lock<std::mutex> lk(mx_);
      wake_.wait(lk, [&] { return gen_ != seen; });
      seen = gen_;
      if (stop_)
        return;
      lk.unlock();
      work();
      lk.lock();
      if (--active_ == 0)
        idle_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex job_, mx_;
  std::condition_variable wake_, idle_;
  const std::function<void(int)> *f_ = nullptr;
  int count_ = 0, active_ = 0;
  std::atomic<int> next_{0};
  unsigned long gen_ = 0;
  bool stop_ = false;
};

// f(0) .. f(count - 1) on the tree workers, indices handed out one at a time
// (the calls must be independent of each other)
template <typename F> static void parallel_for(int count, F &&f)
{
  if (count <= 1 || TreePool::get().workers() == 0)
  {
    for (int q = 0; q < count; q++)
      f(q);
    return;
  }
  const std::function<void(int)> fn = [&](int q) { f(q); };
  TreePool::get().run(count, fn);
}

// f(begin, end) over consecutive chunks of [0, n) on the tree workers, at
// least `grain` indices per chunk (one chunk, on the caller, when n is small).
// Not to be called from inside a parallel_for job.
template <typename F> static void parallel_range(size_t n, size_t grain, F &&f)
{
  const size_t workers = (size_t)TreePool::get().workers() + 1;
  if (workers <= 1 || n < 2 * grain)
  {
    f((size_t)0, n);
    return;
  }
  const int chunks = (int)std::min<size_t>((n + grain - 1) / grain, 4 * workers);
  parallel_for(chunks, [&](int q) { f(n * (size_t)q / (size_t)chunks, n * (size_t)(q + 1) / (size_t)chunks); });
}

// The zones of zone_fill found on all tree threads: union-find over the marked
// cells with 6-neighbour links, each link hooking the larger root under the
// smaller (lock-free compare-and-swap), so a zone's root is its first marked
// cell in cell_ref order.  Numbering the zones by root then gives zone_fill's
// numbering (its flood fills start at the first unvisited marked cell), and the
// member lists are built in cell_ref order as there.  The marked cells' box
// entries hold their cell_ref index afterwards (zone_fill leaves the zone id;
// neither is read again).
static int uf_find(int *par, int x)
{
  for (;;)
  {
    const int p = __atomic_load_n(par + x, __ATOMIC_RELAXED);
    if (p == x)
      return x;
    const int g = __atomic_load_n(par + p, __ATOMIC_RELAXED);
    if (g != p)
      __atomic_store_n(par + x, g, __ATOMIC_RELAXED); // path halving: g is in x's tree and <= p
    x = p;
  }
}

static void uf_unite(int *par, int a, int b)
{
  for (;;)
  {
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b)
      return;
    if (a < b)
      std::swap(a, b);
    int expect = a; // a is a root: hook it under the smaller root b
    if (__atomic_compare_exchange_n(par + a, &expect, b, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      return;
  }
}

static void zone_union(int *box, int rX, int rXY, const std::vector<int> &ref_box, std::vector<int> &zone_size,
                       std::vector<std::vector<int>> &zones_by_ref, bool &one_zone)
{
  const int nref = (int)ref_box.size();
  std::vector<int> par(nref);
  parallel_range((size_t)nref, 1 << 14, [&](size_t a, size_t b) {
    for (size_t t = a; t < b; t++)
    {
      par[t] = (int)t;
      box[ref_box[t]] = (int)t; // distinct cells: distinct entries
    }
  });
  parallel_range((size_t)nref, 1 << 12, [&](size_t a, size_t b) {
    for (size_t t = a; t < b; t++)
    {
      const int c = ref_box[t];
      const int nbs[3] = {c + 1, c + rX, c + rXY};
      for (int q = 0; q < 3; q++)
      {
        const int v = box[nbs[q]]; // >= 0: a marked cell's cell_ref index
        if (v >= 0)
          uf_unite(par.data(), (int)t, v);
      }
    }
  });
  std::vector<int> root(nref);
  std::atomic<int> others{0}; // members whose root is not the first marked cell
  parallel_range((size_t)nref, 1 << 14, [&](size_t a, size_t b) {
    int o = 0;
    for (size_t t = a; t < b; t++)
    {
      root[t] = uf_find(par.data(), (int)t);
      o += root[t] != 0;
    }
    if (o)
      others.fetch_add(o, std::memory_order_relaxed);
  });
  // one zone holding every marked cell (the usual case for a large parent):
  // its members are the cells in cell_ref order, which the caller has
  one_zone = others.load() == 0;
  if (one_zone)
  {
    zone_size.assign(1, nref);
    zones_by_ref.clear();
    return;
  }
  // zones numbered by their first member; members in cell_ref order
  std::vector<int> zone_of(nref, -1);
  int nz = 0;
  for (int t = 0; t < nref; t++)
    if (root[t] == t)
      zone_of[t] = nz++;
  zone_size.assign(nz, 0);
  for (int t = 0; t < nref; t++)
    zone_size[zone_of[root[t]]]++;
  zones_by_ref.assign(nz, {});
  for (int z = 0; z < nz; z++)
    zones_by_ref[z].reserve(zone_size[z]);
  for (int t = 0; t < nref; t++)
    zones_by_ref[zone_of[root[t]]].push_back(t);
}

// Box statuses of a node that can refine further (tree_construction.c:1262-1310):
// -4 outside the node, -3 at its cells, -5 beyond the simulation box.  `wide`:
// the fill, the cells and the outside planes are spread over the tree threads
// (the three passes run one after the other, as in the serial form).
void TreeHost::build_node_box(HNode &nd, bool wide) const
{
  const int rX = nd.rd[0], rXY = nd.rd[0] * nd.rd[1];
  const size_t cap = (size_t)nd.rd[0] * nd.rd[1] * nd.rd[2];
  nd.box.reset(new int[cap]);
  int *box = nd.box.get();
  auto fill = [&](size_t a, size_t b) { std::fill(box + a, box + b, -4); };
  auto cells = [&](size_t a, size_t b) {
    for (size_t c = a; c < b; c++)
      box[nd.cell_box((int)c)] = -3;
  };
  if (wide)
  {
    parallel_range(cap, (size_t)1 << 16, fill);
    parallel_range((size_t)nd.ncell, (size_t)1 << 14, cells);
  }
  else
  {
    fill(0, cap);
    cells(0, (size_t)nd.ncell);
  }
  if (!nd.sim_contact)
    return;
  const int top = (1 << nd.lv) - 1;
  for (int a = 0; a < 3; a++)
  {
    if (!nd.sc[a])
      continue;
    const int lo_end = nd.bmin[a] == 0 ? nd.bmin[a] - nd.ts[a] : 0;
    const int hi_beg = nd.bmax[a] == top ? nd.bmax[a] + 1 - nd.ts[a] : nd.rd[a];
    auto planes = [&](size_t k0, size_t k1) {
      for (int k = (int)k0; k < (int)k1; k++)
        for (int j = 0; j < nd.rd[1]; j++)
          for (int i = 0; i < nd.rd[0]; i++)
          {
            const int q = a == 0 ? i : (a == 1 ? j : k);
            if (q < lo_end || q >= hi_beg)
              box[i + j * rX + k * rXY] = -5;
          }
    };
    if (wide)
      parallel_range((size_t)nd.rd[2], 1, planes);
    else
      planes(0, (size_t)nd.rd[2]);
  }
}

void TreeHost::build_boxes(int first, int count)
{
  std::vector<int> small, big;
  for (int q = first; q < first + count; q++)
    if (nodes[q].lv < lmax)
      ((long long)nodes[q].bcap >= kTreeBigNode ? big : small).push_back(q);
  parallel_for((int)small.size(), [&](int t) { build_node_box(nodes[small[t]], false); });
  for (int q : big)
    build_node_box(nodes[q], true);
}

// create_child_nodes box geometry (tree_construction.c:1092-1661, non-pbc paths):
// the child box spans the zone's cells refined once, plus `extra` cells split
// over both sides.
HNode TreeHost::child_geometry(int parent_global, int parent_lv, int zone, const int *bmin, const int *bmax,
                               int nzcell) const
{
  HNode ch;
  ch.lv = parent_lv + 1;
  ch.parent = parent_global;
  ch.zone = zone;
  for (int a = 0; a < 3; a++)
    if (bmin[a] == 0 || bmax[a] == (1 << parent_lv) - 1)
    {
      ch.sim_contact = true;
      ch.sc[a] = true;
    }
  for (int a = 0; a < 3; a++)
  {
    ch.bmin[a] = 2 * bmin[a];
    ch.bmax[a] = 2 * bmax[a] + 1;
    int dim = ch.bmax[a] - ch.bmin[a] + 1;
    ch.rd[a] = dim + extra;
    ch.ts[a] = ch.bmin[a] - (ch.rd[a] - dim) / 2;
  }
  ch.ncell = 8 * nzcell;
  ch.bcap = ch.rd[0] * ch.rd[1] * ch.rd[2];
  ch.gcap = (ch.rd[0] + 1) * (ch.rd[1] + 1) * (ch.rd[2] + 1);
  ch.pcap = (ch.bmax[0] - ch.bmin[0] + 2) * (ch.bmax[1] - ch.bmin[1] + 2) * (ch.bmax[2] - ch.bmin[2] + 2);
  return ch;
}

std::vector<HNode> TreeHost::make_children(int parent_global, int parent_lv, const int *pcxyz,
                                          const std::vector<int> &zone_sizes,
                                          const std::vector<std::vector<int>> &zone_cells, bool wide) const
{
  int nz = (int)zone_sizes.size();
  std::vector<HNode> kids;
  kids.reserve(nz);
  for (int z = 0; z < nz; z++)
  {
    const std::vector<int> &mem = zone_cells[z];
    const int m = (int)mem.size();
    // child cells: 8 per zone cell in zone order, octant order x fastest
    // (tree_construction.c:1200-1260), kept as the zone cells
    std::vector<int> zx(3 * (size_t)m);
    auto gather = [&](size_t t0, size_t t1, int *lo, int *hi) {
      int l[3] = {lo[0], lo[1], lo[2]}, h[3] = {hi[0], hi[1], hi[2]}; // in registers, stored once
      for (size_t t = t0; t < t1; t++)
        for (int a = 0; a < 3; a++)
        {
          const int v = pcxyz[3 * (size_t)mem[t] + a];
          zx[3 * t + a] = v;
          l[a] = std::min(l[a], v);
          h[a] = std::max(h[a], v);
        }
      for (int a = 0; a < 3; a++)
      {
        lo[a] = l[a];
        hi[a] = h[a];
      }
    };
    int bmin[3] = {1 << 30, 1 << 30, 1 << 30}, bmax[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    if (wide && m >= (1 << 15))
    {
      const int nck = 64;
      std::vector<int> lo(3 * nck, 1 << 30), hi(3 * nck, -(1 << 30));
      parallel_for(nck, [&](int q) {
        gather((size_t)m * q / nck, (size_t)m * (q + 1) / nck, &lo[3 * q], &hi[3 * q]);
      });
      for (int q = 0; q < nck; q++)
        for (int a = 0; a < 3; a++)
        {
          bmin[a] = std::min(bmin[a], lo[3 * q + a]);
          bmax[a] = std::max(bmax[a], hi[3 * q + a]);
        }
    }
    else
      gather(0, (size_t)m, bmin, bmax);
    HNode ch = child_geometry(parent_global, parent_lv, z, bmin, bmax, m);
    ch.zxyz = std::move(zx);
    kids.push_back(std::move(ch));
  }
  return kids;
}

int TreeHost::add_head_zones(int nz, const int *bbox6, const int *zsize, const int *cells)
{
  clear();
  const int first = (int)nodes.size();
  std::vector<size_t> at(nz + 1, 0);
  for (int z = 0; z < nz; z++)
    at[z + 1] = at[z] + zsize[z];
  nodes.resize(first + nz);
  // the zones' nodes are independent: built on up to kTreeThreads threads
  auto build = [&](int z) {
    HNode ch = child_geometry(-1, lmin, z, bbox6 + 6 * z, bbox6 + 6 * z + 3, zsize[z]);
    if (cells)
      ch.zxyz.assign(cells + 3 * at[z], cells + 3 * at[z + 1]);
    nodes[first + z] = std::move(ch);
  };
  parallel_for(nz, build);
  if (cells)
    build_boxes(first, nz);
  if (nz > 0)
  {
    lev_start.push_back(first);
    lev_count.push_back(nz);
  }
  return nz;
}

int TreeHost::add_level_zones(int nz, const int *parent, const int *bbox6, const int *zsize)
{
  const int first = (int)nodes.size();
  nodes.resize(first + nz);
  int zone = 0;
  for (int z = 0; z < nz; z++)
  {
    zone = z > 0 && parent[z] == parent[z - 1] ? zone + 1 : 0;
    HNode &p = nodes[parent[z]];
    nodes[first + z] = child_geometry(parent[z], p.lv, zone, bbox6 + 6 * z, bbox6 + 6 * z + 3, zsize[z]);
    if (zone == 0)
      p.child.clear();
    p.child.push_back(first + z);
  }
  if (nz > 0)
  {
    lev_start.push_back(first);
    lev_count.push_back(nz);
  }
  return nz;
}

// Marks, zones and children of one parent (tree_construction.c:508-926).
// fill_cell_ref with n_exp = 1: each flagged cell marks the existing cells of
// its 3x3x3 block (box status -3 -> -1); the marked cells are then read in cell
// position order (the reference's cell_ref order) and flood-filled into zones.
// With `wide` the marking and the scan run on all tree threads: marks of one
// entry from several threads store the same value (relaxed atomic accesses),
// and the scan's chunks are joined in cell order, so the result is the serial
// one.  The flood fill stays serial.
std::vector<HNode> TreeHost::refine_parent(int global, const int *fcell, int nflag, bool wide)
{
  HNode &nd = nodes[global];
  const int rX = nd.rd[0], rXY = nd.rd[0] * nd.rd[1];
  int *box = nd.box.get();
  auto mark = [&](size_t f0, size_t f1) {
    for (size_t f = f0; f < f1; f++)
    {
      const int b = nd.cell_box(fcell[f]);
      for (int kk = -1; kk <= 1; kk++)
        for (int jj = -1; jj <= 1; jj++)
          for (int ii = -1; ii <= 1; ii++)
          {
            int *e = box + (b + ii + jj * rX + kk * rXY);
            if (__atomic_load_n(e, __ATOMIC_RELAXED) == -3)
              __atomic_store_n(e, -1, __ATOMIC_RELAXED);
          }
    }
  };
  std::vector<int> marked_pos;
  if (wide)
  {
    parallel_range((size_t)nflag, 4096, mark);
    const int nck = 64;
    std::vector<std::vector<int>> part(nck);
    parallel_for(nck, [&](int q) {
      const int c0 = (int)((long long)nd.ncell * q / nck), c1 = (int)((long long)nd.ncell * (q + 1) / nck);
      std::vector<int> mine; // local: the vector headers in `part` share cache lines
      for (int c = c0; c < c1; c++)
        if (box[nd.cell_box(c)] == -1)
          mine.push_back(c);
      part[q] = std::move(mine);
    });
    size_t tot = 0;
    for (auto &v : part)
      tot += v.size();
    marked_pos.reserve(tot);
    for (auto &v : part)
      marked_pos.insert(marked_pos.end(), v.begin(), v.end());
  }
  else
  {
    mark(0, (size_t)nflag);
    for (int c = 0; c < nd.ncell; c++)
      if (box[nd.cell_box(c)] == -1)
        marked_pos.push_back(c);
  }
  if (marked_pos.empty())
    return {};
  const size_t nm = marked_pos.size();
  std::vector<int> ref_box(nm), pc(3 * nm);
  auto coords = [&](size_t t0, size_t t1) {
    for (size_t t = t0; t < t1; t++)
    {
      ref_box[t] = nd.cell_box(marked_pos[t]);
      nd.cell_xyz(marked_pos[t], &pc[3 * t]);
    }
  };
  if (wide)
    parallel_range(nm, 1 << 14, coords);
  else
    coords(0, nm);
  std::vector<int> zone_sizes;
  std::vector<std::vector<int>> zones;
  bool one_zone = false;
  if (wide)
    zone_union(box, rX, rXY, ref_box, zone_sizes, zones, one_zone);
  else
    zone_fill(box, rX, rXY, ref_box, zone_sizes, zones);
  if (!one_zone)
    return make_children(global, nd.lv, pc.data(), zone_sizes, zones, wide);
  // the single child of make_children for one zone of every marked cell: the
  // zone cells are the marked cells' coordinates in order, the box their span
  const int nck = 64;
  std::vector<int> lo(3 * nck, 1 << 30), hi(3 * nck, -(1 << 30));
  parallel_for(nck, [&](int q) {
    int l[3] = {1 << 30, 1 << 30, 1 << 30}, h[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (size_t t = nm * q / nck; t < nm * (q + 1) / nck; t++)
      for (int a = 0; a < 3; a++)
      {
        l[a] = std::min(l[a], pc[3 * t + a]);
        h[a] = std::max(h[a], pc[3 * t + a]);
      }
    for (int a = 0; a < 3; a++)
    {
      lo[3 * q + a] = l[a];
      hi[3 * q + a] = h[a];
    }
  });
  int bmin[3] = {1 << 30, 1 << 30, 1 << 30}, bmax[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  for (int q = 0; q < nck; q++)
    for (int a = 0; a < 3; a++)
    {
      bmin[a] = std::min(bmin[a], lo[3 * q + a]);
      bmax[a] = std::max(bmax[a], hi[3 * q + a]);
    }
  std::vector<HNode> kids(1);
  kids[0] = child_geometry(global, nd.lv, 0, bmin, bmax, (int)nm);
  kids[0].zxyz = std::move(pc);
  return kids;
}

// Refinement of one level: per parent with flagged cells, the marks, the zones
// and the child nodes.  Parents are independent (each owns its box statuses
// and makes its own children): the small ones are spread over up to
// kTreeThreads host threads, one parent at a time each, and the big ones run
// one after the other with their loops spread over the threads.  The children
// are appended in parent order, so the tree is the same as a serial pass
// builds, and then get their box statuses.
int TreeHost::refine_level(int lr, const std::vector<int> &flag_node, const std::vector<int> &flag_cell)
{
  int start = lev_start[lr], cnt = lev_count[lr];
  int next_first = (int)nodes.size();
  using clk = std::chrono::steady_clock;
  auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  ms_mark = ms_zones = ms_children = 0.0;
  // flags grouped by node (any order within a node: marking is a set operation)
  const clk::time_point tg = clk::now();
  std::vector<int> fbeg(cnt + 1, 0), fcell(flag_cell.size());
  for (size_t f = 0; f < flag_node.size(); f++)
    fbeg[flag_node[f] + 1]++;
  for (int q = 0; q < cnt; q++)
    fbeg[q + 1] += fbeg[q];
  {
    std::vector<int> at(fbeg.begin(), fbeg.end() - 1);
    for (size_t f = 0; f < flag_node.size(); f++)
      fcell[at[flag_node[f]]++] = flag_cell[f];
  }
  const clk::time_point t0 = clk::now();
  ms_mark = ms(tg, t0);
  std::vector<std::vector<HNode>> kids(cnt);
  std::vector<int> small, big;
  for (int q = 0; q < cnt; q++)
    if (fbeg[q] < fbeg[q + 1])
      ((long long)nodes[start + q].ncell >= kTreeBigNode ? big : small).push_back(q);
  parallel_for((int)small.size(), [&](int t) {
    const int q = small[t];
    kids[q] = refine_parent(start + q, fcell.data() + fbeg[q], fbeg[q + 1] - fbeg[q], false);
  });
  for (int q : big)
    kids[q] = refine_parent(start + q, fcell.data() + fbeg[q], fbeg[q + 1] - fbeg[q], true);
  const clk::time_point t1 = clk::now();
  ms_zones = ms(t0, t1);
  for (int q = 0; q < cnt; q++)
  {
    HNode &nd = nodes[start + q];
    nd.child.clear();
    for (size_t z = 0; z < kids[q].size(); z++)
      nd.child.push_back((int)nodes.size() + (int)z);
    for (auto &k : kids[q])
      nodes.push_back(std::move(k));
  }
  int made = (int)nodes.size() - next_first;
  build_boxes(next_first, made);
  ms_children = ms(t1, clk::now());
  if (made > 0)
  {
    lev_start.push_back(next_first);
    lev_count.push_back(made);
  }
  return made;
}

} // namespace qz
