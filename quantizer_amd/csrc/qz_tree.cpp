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
static void zone_fill(std::vector<int> &box, int rX, int rXY, const std::vector<int> &ref_box,
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
    const int nw = std::max(0, std::min(kTreeThreads, hw) - 1);
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

void TreeHost::build_node_box(HNode &nd) const
{
  int rX = nd.rd[0], rXY = nd.rd[0] * nd.rd[1];
  size_t cap = (size_t)nd.rd[0] * nd.rd[1] * nd.rd[2];
  nd.box.assign(cap, -4);
  for (int c = 0; c < nd.ncell; c++)
    nd.box[nd.cell_box(c)] = -3;
  if (nd.sim_contact)
  {
    int top = (1 << nd.lv) - 1;
    for (int a = 0; a < 3; a++)
    {
      if (!nd.sc[a])
        continue;
      int lo_end = nd.bmin[a] == 0 ? nd.bmin[a] - nd.ts[a] : 0;
      int hi_beg = nd.bmax[a] == top ? nd.bmax[a] + 1 - nd.ts[a] : nd.rd[a];
      for (int k = 0; k < nd.rd[2]; k++)
        for (int j = 0; j < nd.rd[1]; j++)
          for (int i = 0; i < nd.rd[0]; i++)
          {
            int q = a == 0 ? i : (a == 1 ? j : k);
            if (q < lo_end || q >= hi_beg)
              nd.box[i + j * rX + k * rXY] = -5;
          }
    }
  }
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

// Child cells: 8 per zone cell in zone order, octant order x fastest
// (tree_construction.c:1200-1260), kept as the zone cells (HNode::cell_xyz and
// cell_box expand them); the box status of nodes that can refine further.
void TreeHost::child_cells(HNode &ch, const int *zxyz, int nzcell) const
{
  ch.zxyz.assign(zxyz, zxyz + 3 * (size_t)nzcell);
  if (ch.lv < lmax)
    build_node_box(ch);
}

std::vector<HNode> TreeHost::make_children(int parent_global, int parent_lv, const int *pcxyz,
                                          const std::vector<int> &zone_sizes,
                                          const std::vector<std::vector<int>> &zone_cells) const
{
  int nz = (int)zone_sizes.size();
  std::vector<HNode> kids;
  kids.reserve(nz);
  std::vector<int> zx;
  for (int z = 0; z < nz; z++)
  {
    int bmin[3] = {1 << 30, 1 << 30, 1 << 30}, bmax[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    const int m = (int)zone_cells[z].size();
    zx.resize(3 * (size_t)m);
    for (int t = 0; t < m; t++)
      for (int a = 0; a < 3; a++)
      {
        int v = pcxyz[3 * zone_cells[z][t] + a];
        zx[3 * t + a] = v;
        bmin[a] = std::min(bmin[a], v);
        bmax[a] = std::max(bmax[a], v);
      }
    HNode ch = child_geometry(parent_global, parent_lv, z, bmin, bmax, m);
    child_cells(ch, zx.data(), m);
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
      child_cells(ch, cells + 3 * at[z], zsize[z]);
    nodes[first + z] = std::move(ch);
  };
  parallel_for(nz, build);
  if (nz > 0)
  {
    lev_start.push_back(first);
    lev_count.push_back(nz);
  }
  return nz;
}

// Refinement of one level: per parent with flagged cells, the marks, the zones
// and the child nodes.  Parents are independent (each owns its box statuses
// and makes its own children), so they are spread over up to kTreeThreads
// host threads, handed out one at a time; the children are then appended in
// parent order, so the tree is the same as a serial pass builds.
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
  auto refine_one = [&](int q) {
    if (fbeg[q] == fbeg[q + 1])
      return;
    HNode &nd = nodes[start + q];
    int rX = nd.rd[0], rXY = nd.rd[0] * nd.rd[1];
    // fill_cell_ref with n_exp = 1: each flagged cell marks the existing cells
    // of its 3x3x3 block (box status -3 -> -1); the marked cells are then read
    // in cell position order (the reference's cell_ref order)
    for (int f = fbeg[q]; f < fbeg[q + 1]; f++)
    {
      const int b = nd.cell_box(fcell[f]);
      for (int kk = -1; kk <= 1; kk++)
        for (int jj = -1; jj <= 1; jj++)
          for (int ii = -1; ii <= 1; ii++)
          {
            const int nb = b + ii + jj * rX + kk * rXY;
            if (nd.box[nb] == -3)
              nd.box[nb] = -1;
          }
    }
    std::vector<int> marked_pos;
    for (int c = 0; c < nd.ncell; c++)
      if (nd.box[nd.cell_box(c)] == -1)
        marked_pos.push_back(c);
    if (marked_pos.empty())
      return;
    std::vector<int> ref_box(marked_pos.size());
    std::vector<int> pc(3 * marked_pos.size());
    for (size_t t = 0; t < marked_pos.size(); t++)
    {
      ref_box[t] = nd.cell_box(marked_pos[t]);
      nd.cell_xyz(marked_pos[t], &pc[3 * t]);
    }
    std::vector<int> zone_sizes;
    std::vector<std::vector<int>> zones;
    zone_fill(nd.box, rX, rXY, ref_box, zone_sizes, zones);
    kids[q] = make_children(start + q, nd.lv, pc.data(), zone_sizes, zones);
  };
  parallel_for(cnt, refine_one);
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
  ms_children = ms(t1, clk::now());
  int made = (int)nodes.size() - next_first;
  if (made > 0)
  {
    lev_start.push_back(next_first);
    lev_count.push_back(made);
  }
  return made;
}

} // namespace qz
