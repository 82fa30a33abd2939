// Host side of the refinement tree: the node list and the child box geometry
// of tree_construction.c.  Marks and zones are found on the device and arrive
// as zone bounding boxes (add_head_zones, add_level_zones); refine_level finds
// them on the host instead (dilated refinement marks, 6-connected zones), from
// the flagged cells.  Bulk per-cell and per-grid-point work (child cell
// expansion, box status, grid-point classification, particle ranges,
// refinement counts) runs on the device from what this produces.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

namespace qz
{

// One refinement node as the host knows it.  Geometry fields follow struct node
// (initialize_node.h:63-216): box_ts_*, box_real_dim_*, box_min_*, box_max_*.
struct HNode
{
  int lv = 0;       // absolute level
  int parent = -1;  // global node index, -1 = head
  int zone = 0;     // zone id in the parent (= ID, tree_construction.c:1090)
  int ts[3] = {0, 0, 0};
  int rd[3] = {0, 0, 0};
  int bmin[3] = {0, 0, 0};
  int bmax[3] = {0, 0, 0};
  bool sc[3] = {false, false, false};
  bool sim_contact = false;
  // cells: 8 per zone cell of the parent, in zone order then octant order (x
  // fastest); cell c is octant c & 7 of zone cell c >> 3
  std::vector<int> zxyz;      // 3 per zone cell, parent-level (lv - 1) coordinates
  std::unique_ptr<int[]> box; // box status, rd[0] rd[1] rd[2] entries (built for refinable nodes)
  std::vector<int> child;     // global node indices by zone id
  // pools on the device (filled by the device mirror)
  long long goff = 0, boff = 0, coff = 0, poff = 0;
  int gcap = 0, bcap = 0, ncell = 0, pcap = 0;

  // global level-lv coordinates and box index of cell c
  void cell_xyz(int c, int *g) const
  {
    const int j = c >> 3, o = c & 7;
    g[0] = 2 * zxyz[3 * j] + (o & 1);
    g[1] = 2 * zxyz[3 * j + 1] + ((o >> 1) & 1);
    g[2] = 2 * zxyz[3 * j + 2] + (o >> 2);
  }
  int cell_box(int c) const
  {
    int g[3];
    cell_xyz(c, g);
    return (g[0] - ts[0]) + (g[1] - ts[1]) * rd[0] + (g[2] - ts[2]) * rd[0] * rd[1];
  }
};

constexpr int kTreeThreads = 16; // host threads of refine_level
// Nodes with at least this many cells (or box entries) are worked on by all
// tree threads at once; smaller ones are spread over the threads one node each.
constexpr long long kTreeBigNode = 1 << 16;

struct TreeHost
{
  int lmin = 0, lmax = 0, n = 0, ref_crit = 8, extra = 10;
  // wall time of refine_level's parts (last call): grouping the flags, the
  // per-parent marks, zones and children (threaded), appending the children and
  // building their box statuses (threaded)
  double ms_mark = 0.0, ms_zones = 0.0, ms_children = 0.0;
  std::vector<HNode> nodes;     // refinement nodes, level order (tentacles order, tree_construction.c:2465-2495)
  std::vector<int> lev_start, lev_count; // relative level 1.. -> ranges in nodes

  void init(int lmin_, int lmax_, int crit);
  // Head level, zones found on the device: per zone the bounding box of its
  // head cells (min xyz, max xyz) and its cell count; `cells` holds the head
  // cell coordinates of every zone in zone order (needed only when the
  // children can refine further, else null).  Returns the level-1 node count.
  int add_head_zones(int nz, const int *bbox6, const int *zsize, const int *cells);
  // Level lr >= 1: flagged cells as (node position in level, cell position), any order.
  int refine_level(int lr, const std::vector<int> &flag_node, const std::vector<int> &flag_cell);
  // The next level from zones found on the device: per zone its parent (global
  // node index), the bounding box of its cells in parent-level coordinates (min
  // xyz, max xyz) and its cell count; zones in parent order and, within a
  // parent, in the order of their first cells.  The zone cells stay on the
  // device, so these nodes cannot be refined by refine_level.  Returns the
  // number of nodes made.
  int add_level_zones(int nz, const int *parent, const int *bbox6, const int *zsize);
  void clear();

private:
  HNode child_geometry(int parent_global, int parent_lv, int zone, const int *bmin, const int *bmax, int nzcell) const;
  std::vector<HNode> make_children(int parent_global, int parent_lv, const int *cell_xyz_parent,
                                   const std::vector<int> &zone_sizes, const std::vector<std::vector<int>> &zones,
                                   bool wide) const;
  // marks, zones and children of one parent; `wide`: spread its loops over the tree threads
  std::vector<HNode> refine_parent(int global, const int *fcell, int nflag, bool wide);
  // box statuses of the nodes [first, first + count) that can refine further
  void build_boxes(int first, int count);
  void build_node_box(HNode &nd, bool wide) const;
};

} // namespace qz
