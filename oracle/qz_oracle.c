/*
 * qz_oracle.c -- CPU restatement of the QUANTIZER per-step PM gravity cycle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the "port" CPU baseline).  The
 * product path never links or calls this file.
 *
 * Each routine restates the reference algorithm and cites the file:line it
 * follows (relative to the QUANTIZER source tree).  Everything is FP64 and
 * sequential, like the reference with NUMBER_OF_THEADS = 1
 * (global_variables.c:250).  Non-periodic boundaries only (bdry_cond_type = 1).
 *
 * Pinning (see tests/test_oracle.py): golden vectors from the compiled
 * reference (oracle/_ref), closed forms (key brute force, deposit mass
 * conservation, exact trilinear interpolation, discrete sine eigenmodes for
 * the Poisson operator), and invariants.  Known deviation: the tree is rebuilt
 * from scratch every step (tree_construction.c) instead of adapted
 * incrementally (tree_adaptation.c); fields are compared at global grid
 * coordinates, so box padding differences do not matter.
 */
#include "qz_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OQ_PI 3.14159265358979323846 /* global_variables.c:221 */

typedef struct
{
  int lv;     /* absolute level */
  int parent; /* index in sim->nodes, -1 for the head */
  int ID;
  int ts[3], rd[3], bmin[3], bmax[3];
  int sim_contact, sc[3];
  int ncell;
  int *cx, *cy, *cz, *cbox;
  int boxcap;
  int *box; /* -5 outside the simulation, -4 not in node, -3 exists, >=0 zone id */
  int nintr, nbdry, nsim;
  int *ixyz, *iidx, *bxyz, *bidx, *sxyz, *sidx;
  int gcap;
  double *d, *pot, *pot_old, *ax, *ay, *az;
  double mass;
  int nptcl_full, nptcl_outs;
  int nchild;
  int *child; /* indices in sim->nodes, by zone id */
  /* particles per box cell (cell_struct.ptr_ptcl restated as CSR) */
  int *cell_cnt;   /* boxcap */
  int *cell_start; /* boxcap + 1 */
  int *cell_list;  /* nptcl_full */
  double *cell_mass;
  /* zone scratch */
  int nzones;
  int *zone_size;
  int **zones;
} onode;

struct oq_sim
{
  int lmin, lmax, n, ref_crit, n_exp, extra, head_bc;
  double G, tol, tol2, total_mass;
  int max_iter, check_every;
  int pre, post, coarse;
  double w_head;
  double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *m;
  int nnodes, capnodes;
  onode *nodes;
  int nlev;           /* number of levels present (head = level index 0) */
  int lev_start[64];  /* node index ranges per relative level */
  int lev_count[64];
  int stat_vcycles, stat_cg_iters, stat_cg_solves, stat_fail;
};

/* ------------------------------------------------------------------------ */
/* A1: ptcl_idx_to_box_idx.c:30-139 (non-periodic part)                      */
/* ------------------------------------------------------------------------ */
static inline int key_floor(double p)
{
  /* ptcl_idx_to_box_idx.c:62-64: x<0 -> -1, else truncation */
  return p < 0 ? -1 : (int)p;
}

int oq_key(int lv, const int *ts, const int *rd, double x, double y, double z)
{
  double s = (double)(1 << lv); /* :53-55 */
  int fx = key_floor(x * s), fy = key_floor(y * s), fz = key_floor(z * s);
  int bx = fx - ts[0], by = fy - ts[1], bz = fz - ts[2]; /* :67-69 */
  return bx + by * rd[0] + bz * rd[0] * rd[1];           /* :136 */
}

static inline int node_key(const onode *nd, double x, double y, double z)
{
  return oq_key(nd->lv, nd->ts, nd->rd, x, y, z);
}

/* ------------------------------------------------------------------------ */
/* node bookkeeping                                                          */
/* ------------------------------------------------------------------------ */
static void node_free(onode *nd)
{
  free(nd->cx);
  free(nd->cy);
  free(nd->cz);
  free(nd->cbox);
  free(nd->box);
  free(nd->ixyz);
  free(nd->iidx);
  free(nd->bxyz);
  free(nd->bidx);
  free(nd->sxyz);
  free(nd->sidx);
  free(nd->d);
  free(nd->pot);
  free(nd->pot_old);
  free(nd->ax);
  free(nd->ay);
  free(nd->az);
  free(nd->child);
  free(nd->cell_cnt);
  free(nd->cell_start);
  free(nd->cell_list);
  free(nd->cell_mass);
  if (nd->zones)
  {
    for (int i = 0; i < nd->nzones; i++)
      free(nd->zones[i]);
    free(nd->zones);
  }
  free(nd->zone_size);
  memset(nd, 0, sizeof(*nd));
}

static onode *push_node(oq_sim *s)
{
  if (s->nnodes == s->capnodes)
  {
    s->capnodes = s->capnodes ? 2 * s->capnodes : 64;
    s->nodes = (onode *)realloc(s->nodes, sizeof(onode) * (size_t)s->capnodes);
  }
  onode *nd = &s->nodes[s->nnodes++];
  memset(nd, 0, sizeof(*nd));
  return nd;
}

static inline int gdx(const onode *nd) { return nd->rd[0] + 1; }
static inline int gdxy(const onode *nd) { return (nd->rd[0] + 1) * (nd->rd[1] + 1); }

/* Per-box-cell particle lists (the reference's cell_struct lists,
 * initialization.c:160-183 for the head), built from a list of particles. */
static void node_bin_particles(oq_sim *s, onode *nd, const int *plist, int np)
{
  free(nd->cell_cnt);
  free(nd->cell_start);
  free(nd->cell_list);
  free(nd->cell_mass);
  nd->cell_cnt = (int *)calloc((size_t)nd->boxcap, sizeof(int));
  nd->cell_start = (int *)calloc((size_t)nd->boxcap + 1, sizeof(int));
  nd->cell_list = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  nd->cell_mass = (double *)calloc((size_t)nd->boxcap, sizeof(double));
  int *keys = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  for (int t = 0; t < np; t++)
  {
    int p = plist[t];
    keys[t] = node_key(nd, s->x[p], s->y[p], s->z[p]);
    nd->cell_cnt[keys[t]]++;
    nd->cell_mass[keys[t]] += s->m[p];
  }
  for (int b = 0; b < nd->boxcap; b++)
    nd->cell_start[b + 1] = nd->cell_start[b] + nd->cell_cnt[b];
  int *fill = (int *)calloc((size_t)nd->boxcap, sizeof(int));
  for (int t = 0; t < np; t++)
  {
    int b = keys[t];
    nd->cell_list[nd->cell_start[b] + fill[b]++] = plist[t];
  }
  free(fill);
  free(keys);
}

/* Head node: initialization.c:75-258 */
static void build_head(oq_sim *s)
{
  onode *h = push_node(s);
  int n = 1 << s->lmin;
  int bder = 1; /* bder_os_sim = max(1, n_exp), global_variables.c:398 */
  int side = n + 2 * bder;
  h->lv = s->lmin;
  h->parent = -1;
  for (int a = 0; a < 3; a++)
  {
    h->ts[a] = -bder;
    h->rd[a] = side;
    h->bmin[a] = 0;
    h->bmax[a] = n - 1;
    h->sc[a] = 1;
  }
  h->sim_contact = 1;
  h->boxcap = side * side * side;
  h->box = (int *)malloc(sizeof(int) * (size_t)h->boxcap);
  for (int k = 0; k < side; k++)
    for (int j = 0; j < side; j++)
      for (int i = 0; i < side; i++)
      {
        int b = i + side * (j + side * k);
        int border = i <= bder - 1 || i >= side - bder || j <= bder - 1 || j >= side - bder || k <= bder - 1 ||
                     k >= side - bder;
        h->box[b] = border ? -5 : -3; /* initialization.c:96-113 */
      }
  h->ncell = n * n * n;
  h->cx = (int *)malloc(sizeof(int) * (size_t)h->ncell);
  h->cy = (int *)malloc(sizeof(int) * (size_t)h->ncell);
  h->cz = (int *)malloc(sizeof(int) * (size_t)h->ncell);
  h->cbox = (int *)malloc(sizeof(int) * (size_t)h->ncell);
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++)
      {
        int c = i + n * (j + n * k);
        h->cx[c] = i;
        h->cy[c] = j;
        h->cz[c] = k;
        h->cbox[c] = (i + bder) + side * ((j + bder) + side * (k + bder)); /* :135-141 */
      }
  /* grid points: initialization.c:215-242 */
  int g = side + 1;
  h->gcap = g * g * g;
  int nint = (n - 1) * (n - 1) * (n - 1);
  int nsim = (n + 1) * (n + 1) * (n + 1) - nint;
  h->ixyz = (int *)malloc(sizeof(int) * 3 * (size_t)nint);
  h->iidx = (int *)malloc(sizeof(int) * (size_t)nint);
  h->sxyz = (int *)malloc(sizeof(int) * 3 * (size_t)nsim);
  h->sidx = (int *)malloc(sizeof(int) * (size_t)nsim);
  for (int k = bder; k < side + 1 - bder; k++)
    for (int j = bder; j < side + 1 - bder; j++)
      for (int i = bder; i < side + 1 - bder; i++)
      {
        int gi = i + g * (j + g * k);
        int face = i == bder || i == side - bder || j == bder || j == side - bder || k == bder || k == side - bder;
        if (face)
        {
          h->sxyz[3 * h->nsim] = i - bder;
          h->sxyz[3 * h->nsim + 1] = j - bder;
          h->sxyz[3 * h->nsim + 2] = k - bder;
          h->sidx[h->nsim++] = gi;
        }
        else
        {
          h->ixyz[3 * h->nintr] = i - bder;
          h->ixyz[3 * h->nintr + 1] = j - bder;
          h->ixyz[3 * h->nintr + 2] = k - bder;
          h->iidx[h->nintr++] = gi;
        }
      }
  h->d = (double *)calloc((size_t)h->gcap, sizeof(double));
  h->pot = (double *)calloc((size_t)h->gcap, sizeof(double));
  h->pot_old = (double *)calloc((size_t)h->gcap, sizeof(double));
  h->ax = (double *)calloc((size_t)h->gcap, sizeof(double));
  h->ay = (double *)calloc((size_t)h->gcap, sizeof(double));
  h->az = (double *)calloc((size_t)h->gcap, sizeof(double));
  h->mass = s->total_mass;
  h->nptcl_full = s->n;
  h->nptcl_outs = s->n;
  /* A10: initial_potential_and_acceleration_head.c:297-335 (monopole on all
   * box points except the last plane, which the reference loop never reaches) */
  double H = 1.0 / n;
  for (int k = 0; k < side; k++)
  {
    double ak = (k - bder) * H - 0.5;
    for (int j = 0; j < side; j++)
    {
      double aj = (j - bder) * H - 0.5;
      for (int i = 0; i < side; i++)
      {
        double ai = (i - bder) * H - 0.5;
        double dist = sqrt(ai * ai + aj * aj + ak * ak);
        if (dist == 0.0)
          dist = 1.0 / n * 1.0e-3; /* :314-317 */
        int gi = i + g * (j + g * k);
        if (s->head_bc == 0)
        {
          double coeff = -s->G * s->total_mass / (dist * dist * dist);
          h->pot[gi] = -s->G * s->total_mass / dist;
          h->ax[gi] = coeff * ai;
          h->ay[gi] = coeff * aj;
          h->az[gi] = coeff * ak;
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Tree construction (tree_construction.c, non-periodic paths)               */
/* ------------------------------------------------------------------------ */

/* fill_cell_ref, tree_construction.c:508-598 */
static int *fill_cell_ref(oq_sim *s, onode *nd, int *nref)
{
  int rX = nd->rd[0], rXY = nd->rd[0] * nd->rd[1];
  int cnt = 0;
  for (int i = 0; i < nd->ncell; i++)
  {
    int b = nd->cbox[i];
    if (nd->cell_mass[b] >= 1.0e100 || nd->cell_cnt[b] >= s->ref_crit)
    {
      if (nd->box[b] == -3)
      {
        cnt++;
        nd->box[b] = -1;
      }
      for (int kk = -s->n_exp; kk < s->n_exp + 1; kk++)
        for (int jj = -s->n_exp; jj < s->n_exp + 1; jj++)
          for (int ii = -s->n_exp; ii < s->n_exp + 1; ii++)
          {
            int nb = b + ii + jj * rX + kk * rXY;
            if (nd->box[nb] == -3)
            {
              cnt++;
              nd->box[nb] = -1;
            }
          }
    }
  }
  int *ref = (int *)malloc(sizeof(int) * (size_t)(cnt > 0 ? cnt : 1));
  int c = 0;
  for (int i = 0; i < nd->ncell; i++)
    if (nd->box[nd->cbox[i]] == -1)
      ref[c++] = i;
  *nref = c;
  return ref;
}

/* fill_zones_ref, tree_construction.c:685-926 (6-connected flood fill) */
static void fill_zones_ref(onode *nd, const int *ref, int nref)
{
  int rX = nd->rd[0], rXY = nd->rd[0] * nd->rd[1];
  int *aux = (int *)malloc(sizeof(int) * (size_t)(nref > 0 ? nref : 1));
  int added_all = 0, ref_i = 0, zone = 0;
  int zcap = 8;
  int *zsize = (int *)malloc(sizeof(int) * (size_t)zcap);
  while (nref > added_all)
  {
    int b = nd->cbox[ref[ref_i]];
    if (nd->box[b] == -1)
    {
      int zs = 0;
      aux[0] = b;
      nd->box[b] = zone;
      zs++;
      added_all++;
      int insp = 0;
      while (zs > insp && nref > added_all)
      {
        int add = 0;
        int c = aux[insp];
        int nbs[6] = {c + 1, c - 1, c + rX, c - rX, c + rXY, c - rXY};
        for (int q = 0; q < 6; q++)
        {
          if (nd->box[nbs[q]] == -1)
          {
            aux[zs + add] = nbs[q];
            nd->box[nbs[q]] = zone;
            add++;
          }
        }
        zs += add;
        added_all += add;
        insp++;
      }
      if (zone + 1 > zcap)
      {
        zcap *= 2;
        zsize = (int *)realloc(zsize, sizeof(int) * (size_t)zcap);
      }
      zsize[zone] = zs;
      zone++;
    }
    ref_i++;
  }
  nd->nzones = zone;
  nd->zone_size = zsize;
  nd->zones = (int **)malloc(sizeof(int *) * (size_t)(zone > 0 ? zone : 1));
  int *fill = (int *)calloc((size_t)(zone > 0 ? zone : 1), sizeof(int));
  for (int z = 0; z < zone; z++)
    nd->zones[z] = (int *)malloc(sizeof(int) * (size_t)zsize[z]);
  for (int i = 0; i < nref; i++)
  {
    int cell = ref[i];
    int z = nd->box[nd->cbox[cell]];
    nd->zones[z][fill[z]++] = cell;
  }
  free(fill);
  free(aux);
}

/* Grid-point classification of create_child_nodes (tree_construction.c:2038-2202):
 * returns bit0 = exists, bit1 = boundary, from the existence of the 8 cells
 * around the point; bit order c[i0j0k0, im1j0k0, i0jm1k0, im1jm1k0, i0j0km1,
 * im1j0km1, i0jm1km1, im1jm1km1]. */
static int classify_point(const int *e)
{
  /* e[q] = 1 if that cell exists in the node (box > -4) */
  enum { A = 0, Bm = 1, Cm = 2, D = 3, E = 4, F = 5, Gq = 6, Hq = 7 };
  /* names: A=i0j0k0 Bm=im1j0k0 Cm=i0jm1k0 D=im1jm1k0 E=i0j0km1 F=im1j0km1 Gq=i0jm1km1 Hq=im1jm1km1 */
#define NX(q) (!e[q])
  int exist = 0, bdry = 0;
  if (e[A])
  {
    exist = 1;
    if (NX(Bm) && NX(D) && NX(F) && NX(Hq))
      bdry = 1;
    else if (NX(Cm) && NX(D) && NX(Gq) && NX(Hq))
      bdry = 1;
    else if (NX(E) && NX(F) && NX(Gq) && NX(Hq))
      bdry = 1;
  }
  else if (e[Bm])
  {
    exist = 1;
    if (NX(Cm) && NX(E) && NX(Gq))
      bdry = 1;
    else if (NX(Cm) && NX(D) && NX(Gq) && NX(Hq))
      bdry = 1;
    else if (NX(E) && NX(F) && NX(Gq) && NX(Hq))
      bdry = 1;
  }
  else if (e[Cm])
  {
    exist = 1;
    if (NX(D) && NX(F) && NX(Hq))
      bdry = 1;
    else if (NX(E) && NX(F))
      bdry = 1;
    else if (NX(E) && NX(F) && NX(Gq) && NX(Hq))
      bdry = 1;
  }
  else if (e[D])
  {
    exist = 1;
    if (NX(E) && NX(Gq))
      bdry = 1;
    else if (NX(E) && NX(F))
      bdry = 1;
    else if (NX(E) && NX(F) && NX(Gq) && NX(Hq))
      bdry = 1;
  }
  else if (e[E] || e[F] || e[Gq] || e[Hq])
  {
    exist = 1;
    bdry = 1;
  }
#undef NX
  return exist | (bdry << 1);
}

/* create_child_nodes, tree_construction.c:1032-2291 (non-periodic) */
static void create_children(oq_sim *s, int pidx)
{
  onode *nd = &s->nodes[pidx];
  int lv = nd->lv;
  int nz = nd->nzones;
  int *kids = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
  for (int z = 0; z < nz; z++)
  {
    onode *ch = push_node(s);
    nd = &s->nodes[pidx]; /* push may move the array */
    kids[z] = s->nnodes - 1;
    ch->ID = z;
    ch->lv = lv + 1;
    ch->parent = pidx;
    /* box min/max over zone cells, sim-boundary contact (:1404-1600 non-pbc branch) */
    for (int a = 0; a < 3; a++)
    {
      ch->bmin[a] = 1 << 30;
      ch->bmax[a] = -(1 << 30);
    }
    for (int j = 0; j < nd->zone_size[z]; j++)
    {
      int c = nd->zones[z][j];
      int ci[3] = {nd->cx[c], nd->cy[c], nd->cz[c]};
      for (int a = 0; a < 3; a++)
      {
        if (ch->bmin[a] > ci[a])
          ch->bmin[a] = ci[a];
        if (ch->bmax[a] < ci[a])
          ch->bmax[a] = ci[a];
      }
    }
    for (int a = 0; a < 3; a++)
      if (ch->bmin[a] == 0 || ch->bmax[a] == (1 << lv) - 1)
      {
        ch->sim_contact = 1;
        ch->sc[a] = 1;
      }
    /* :1623-1661 */
    int dim[3];
    for (int a = 0; a < 3; a++)
    {
      ch->bmin[a] = 2 * ch->bmin[a];
      ch->bmax[a] = 2 * ch->bmax[a] + 1;
      dim[a] = ch->bmax[a] - ch->bmin[a] + 1;
      ch->rd[a] = dim[a] + s->extra;
      int pos = (ch->rd[a] - dim[a]) / 2;
      ch->ts[a] = ch->bmin[a] - pos;
    }
    int rX = ch->rd[0], rXY = ch->rd[0] * ch->rd[1];
    /* child cells, 8 per zone cell (:1663-1716) */
    ch->ncell = 8 * nd->zone_size[z];
    ch->cx = (int *)malloc(sizeof(int) * (size_t)ch->ncell);
    ch->cy = (int *)malloc(sizeof(int) * (size_t)ch->ncell);
    ch->cz = (int *)malloc(sizeof(int) * (size_t)ch->ncell);
    ch->cbox = (int *)malloc(sizeof(int) * (size_t)ch->ncell);
    for (int j = 0; j < nd->zone_size[z]; j++)
    {
      int c = nd->zones[z][j];
      int axx = nd->cx[c] * 2, ayy = nd->cy[c] * 2, azz = nd->cz[c] * 2;
      int bx = axx - ch->ts[0], by = ayy - ch->ts[1], bz = azz - ch->ts[2];
      int bb = bx + by * rX + bz * rXY;
      for (int kk = 0; kk < 2; kk++)
        for (int jj = 0; jj < 2; jj++)
          for (int ii = 0; ii < 2; ii++)
          {
            int ci = ii + jj * 2 + kk * 4 + j * 8;
            ch->cx[ci] = axx + ii;
            ch->cy[ci] = ayy + jj;
            ch->cz[ci] = azz + kk;
            ch->cbox[ci] = bb + ii + jj * rX + kk * rXY;
          }
    }
    /* box status (:1718-1735, sim boundary -5 :1838-1935) */
    ch->boxcap = ch->rd[0] * ch->rd[1] * ch->rd[2];
    ch->box = (int *)malloc(sizeof(int) * (size_t)ch->boxcap);
    for (int b = 0; b < ch->boxcap; b++)
      ch->box[b] = -4;
    for (int c = 0; c < ch->ncell; c++)
      ch->box[ch->cbox[c]] = -3;
    if (ch->sim_contact)
    {
      for (int a = 0; a < 3; a++)
      {
        if (!ch->sc[a])
          continue;
        int lo_end = ch->bmin[a] == 0 ? ch->bmin[a] - ch->ts[a] : 0;
        int hi_beg = ch->bmax[a] == (1 << (lv + 1)) - 1 ? ch->bmax[a] + 1 - ch->ts[a] : ch->rd[a];
        for (int k = 0; k < ch->rd[2]; k++)
          for (int j = 0; j < ch->rd[1]; j++)
            for (int i = 0; i < ch->rd[0]; i++)
            {
              int q = a == 0 ? i : (a == 1 ? j : k);
              if (q < lo_end || q >= hi_beg)
                ch->box[i + j * rX + k * rXY] = -5;
            }
      }
    }
    /* particles of the zone cells, in parent list order (:1937-2030) */
    int np = 0;
    for (int j = 0; j < nd->zone_size[z]; j++)
      np += nd->cell_cnt[nd->cbox[nd->zones[z][j]]];
    int *plist = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    int t = 0;
    ch->mass = 0.0;
    for (int j = 0; j < nd->zone_size[z]; j++)
    {
      int b = nd->cbox[nd->zones[z][j]];
      for (int q = nd->cell_start[b]; q < nd->cell_start[b + 1]; q++)
        plist[t++] = nd->cell_list[q];
      ch->mass += nd->cell_mass[b];
    }
    ch->nptcl_full = np;
    node_bin_particles(s, ch, plist, np);
    free(plist);
    /* grid points (:2038-2268) */
    int gX = ch->rd[0] + 1, gXY = (ch->rd[0] + 1) * (ch->rd[1] + 1);
    int maxpts = (ch->bmax[0] - ch->bmin[0] + 2) * (ch->bmax[1] - ch->bmin[1] + 2) * (ch->bmax[2] - ch->bmin[2] + 2);
    ch->ixyz = (int *)malloc(sizeof(int) * 3 * (size_t)maxpts);
    ch->iidx = (int *)malloc(sizeof(int) * (size_t)maxpts);
    ch->bxyz = (int *)malloc(sizeof(int) * 3 * (size_t)maxpts);
    ch->bidx = (int *)malloc(sizeof(int) * (size_t)maxpts);
    ch->sxyz = (int *)malloc(sizeof(int) * 3 * (size_t)maxpts);
    ch->sidx = (int *)malloc(sizeof(int) * (size_t)maxpts);
    for (int kk = ch->bmin[2] - ch->ts[2]; kk < ch->bmax[2] - ch->ts[2] + 2; kk++)
      for (int jj = ch->bmin[1] - ch->ts[1]; jj < ch->bmax[1] - ch->ts[1] + 2; jj++)
        for (int ii = ch->bmin[0] - ch->ts[0]; ii < ch->bmax[0] - ch->ts[0] + 2; ii++)
        {
          int b = ii + jj * rX + kk * rXY;
          int nb[8] = {b, b - 1, b - rX, b - 1 - rX, b - rXY, b - 1 - rXY, b - rX - rXY, b - 1 - rX - rXY};
          int e[8];
          for (int q = 0; q < 8; q++)
            e[q] = ch->box[nb[q]] > -4;
          int cls = classify_point(e);
          int exist = cls & 1, bdry = (cls >> 1) & 1;
          if (!exist)
            continue;
          int gxx = ii + ch->ts[0], gyy = jj + ch->ts[1], gzz = kk + ch->ts[2];
          int simb = 0;
          if (ch->sim_contact && bdry)
          {
            int top = 1 << (lv + 1);
            if (gxx == 0 || gxx == top || gyy == 0 || gyy == top || gzz == 0 || gzz == top)
              simb = 1;
          }
          int gi = ii + jj * gX + kk * gXY;
          if (simb)
          {
            ch->sxyz[3 * ch->nsim] = gxx;
            ch->sxyz[3 * ch->nsim + 1] = gyy;
            ch->sxyz[3 * ch->nsim + 2] = gzz;
            ch->sidx[ch->nsim++] = gi;
          }
          else if (bdry)
          {
            ch->bxyz[3 * ch->nbdry] = gxx;
            ch->bxyz[3 * ch->nbdry + 1] = gyy;
            ch->bxyz[3 * ch->nbdry + 2] = gzz;
            ch->bidx[ch->nbdry++] = gi;
          }
          else
          {
            ch->ixyz[3 * ch->nintr] = gxx;
            ch->ixyz[3 * ch->nintr + 1] = gyy;
            ch->ixyz[3 * ch->nintr + 2] = gzz;
            ch->iidx[ch->nintr++] = gi;
          }
        }
    ch->gcap = (ch->rd[0] + 1) * (ch->rd[1] + 1) * (ch->rd[2] + 1);
    ch->d = (double *)calloc((size_t)ch->gcap, sizeof(double));
    ch->pot = (double *)calloc((size_t)ch->gcap, sizeof(double));
    ch->pot_old = (double *)calloc((size_t)ch->gcap, sizeof(double));
    ch->ax = (double *)calloc((size_t)ch->gcap, sizeof(double));
    ch->ay = (double *)calloc((size_t)ch->gcap, sizeof(double));
    ch->az = (double *)calloc((size_t)ch->gcap, sizeof(double));
    ch->nptcl_outs = ch->nptcl_full;
    nd->nptcl_outs -= ch->nptcl_full; /* :2286-2287 */
  }
  nd->nchild = nz;
  nd->child = kids;
  /* transfer_tiny_child_nodes_to_memory_pool, :2350-2409: zones with fewer
   * particles than the criterion are dropped (never happens when the mass
   * criterion is off, since every zone holds a cell meeting the criterion) */
  for (int c = 0; c < nd->nchild; c++)
  {
    onode *ch = &s->nodes[nd->child[c]];
    if (ch->nptcl_full < s->ref_crit)
    {
      int last = nd->nchild - 1;
      s->nodes[nd->child[last]].ID = c;
      for (int q = 0; q < nd->zone_size[last]; q++)
        nd->box[nd->cbox[nd->zones[last][q]]] = c;
      for (int q = 0; q < nd->zone_size[c]; q++)
        nd->box[nd->cbox[nd->zones[c][q]]] = -3;
      ch->nptcl_full = -1; /* tombstone, compacted below */
      nd->child[c] = nd->child[last];
      int tmp = nd->zone_size[c];
      nd->zone_size[c] = nd->zone_size[last];
      nd->zone_size[last] = tmp;
      int *tz = nd->zones[c];
      nd->zones[c] = nd->zones[last];
      nd->zones[last] = tz;
      nd->nchild--;
      c--;
    }
  }
}

static void clear_tree(oq_sim *s)
{
  for (int i = 1; i < s->nnodes; i++)
    node_free(&s->nodes[i]);
  s->nnodes = 1;
  onode *h = &s->nodes[0];
  free(h->child);
  h->child = NULL;
  h->nchild = 0;
  if (h->zones)
  {
    for (int i = 0; i < h->nzones; i++)
      free(h->zones[i]);
    free(h->zones);
    h->zones = NULL;
  }
  free(h->zone_size);
  h->zone_size = NULL;
  h->nzones = 0;
  for (int b = 0; b < h->boxcap; b++)
    if (h->box[b] != -5)
      h->box[b] = -3;
  h->nptcl_outs = s->n;
  s->nlev = 1;
  s->lev_start[0] = 0;
  s->lev_count[0] = 1;
}

/* tree_construction(), tree_construction.c:2521-2577 */
int oq_build_tree(oq_sim *s)
{
  clear_tree(s);
  int *all = (int *)malloc(sizeof(int) * (size_t)s->n);
  for (int i = 0; i < s->n; i++)
    all[i] = i;
  node_bin_particles(s, &s->nodes[0], all, s->n);
  free(all);
  int lvrel = 0;
  while (lvrel < s->lmax - s->lmin && lvrel < s->nlev)
  {
    int start = s->lev_start[lvrel], cnt = s->lev_count[lvrel];
    int next_start = s->nnodes, next_cnt = 0;
    for (int q = 0; q < cnt; q++)
    {
      int pidx = start + q;
      int nref;
      int *ref = fill_cell_ref(s, &s->nodes[pidx], &nref);
      fill_zones_ref(&s->nodes[pidx], ref, nref);
      free(ref);
      if (s->nodes[pidx].nzones > 0)
      {
        create_children(s, pidx);
        next_cnt += s->nodes[pidx].nchild;
      }
    }
    /* drop tombstoned tiny children by compacting this level's new nodes */
    if (s->nnodes - next_start != next_cnt)
    {
      int *remap = (int *)malloc(sizeof(int) * (size_t)s->nnodes);
      int w = next_start;
      for (int r = next_start; r < s->nnodes; r++)
      {
        if (s->nodes[r].nptcl_full < 0)
        {
          node_free(&s->nodes[r]);
          remap[r] = -1;
          continue;
        }
        remap[r] = w;
        if (w != r)
          s->nodes[w] = s->nodes[r];
        w++;
      }
      s->nnodes = w;
      for (int q = 0; q < cnt; q++)
      {
        onode *p = &s->nodes[start + q];
        for (int c = 0; c < p->nchild; c++)
          p->child[c] = remap[p->child[c]];
      }
      free(remap);
    }
    /* fill_tentacles (:2465-2495): children of each parent appended in zone order.
     * create_children already appended them contiguously per parent, but a
     * swap in the tiny-node removal may have reordered one parent's list. */
    if (next_cnt > 0)
    {
      onode *tmp = (onode *)malloc(sizeof(onode) * (size_t)next_cnt);
      int w = 0;
      for (int q = 0; q < cnt; q++)
      {
        onode *p = &s->nodes[start + q];
        for (int c = 0; c < p->nchild; c++)
          tmp[w++] = s->nodes[p->child[c]];
      }
      w = 0;
      for (int q = 0; q < cnt; q++)
      {
        onode *p = &s->nodes[start + q];
        for (int c = 0; c < p->nchild; c++)
        {
          s->nodes[next_start + w] = tmp[w];
          p->child[c] = next_start + w;
          w++;
        }
      }
      free(tmp);
      s->lev_start[lvrel + 1] = next_start;
      s->lev_count[lvrel + 1] = next_cnt;
      s->nlev = lvrel + 2;
    }
    lvrel++;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* A13 reset.c:29-64                                                          */
/* ------------------------------------------------------------------------ */
void oq_reset(oq_sim *s)
{
  for (int i = 0; i < s->n; i++)
    s->ax[i] = s->ay[i] = s->az[i] = 0.0;
  onode *h = &s->nodes[0];
  memset(h->d, 0, sizeof(double) * (size_t)h->gcap);
}

/* ------------------------------------------------------------------------ */
/* A2 grid_density.c:35-306                                                   */
/* ------------------------------------------------------------------------ */
static void density_head(oq_sim *s)
{
  onode *h = &s->nodes[0];
  int lv = h->lv;
  double H = 1.0 / (1 << lv);
  double coeff = 4.0 * s->G * OQ_PI / (H * H * H); /* :62 */
  int gX = gdx(h), gXY = gdxy(h);
  for (int i = 0; i < s->n; i++)
  {
    double px = s->x[i] * (1 << lv), py = s->y[i] * (1 << lv), pz = s->z[i] * (1 << lv);
    int fx = (int)px, fy = (int)py, fz = (int)pz;
    double wx1 = px - fx, wy1 = py - fy, wz1 = pz - fz;
    double wx2 = 1.0 - wx1, wy2 = 1.0 - wy1, wz2 = 1.0 - wz1;
    double w[8] = {wx2 * wy2 * wz2, wx1 * wy2 * wz2, wx2 * wy1 * wz2, wx1 * wy1 * wz2,
                   wx2 * wy2 * wz1, wx1 * wy2 * wz1, wx2 * wy1 * wz1, wx1 * wy1 * wz1};
    int g = (fx - h->ts[0]) + (fy - h->ts[1]) * gX + (fz - h->ts[2]) * gXY;
    for (int kk = 0; kk < 2; kk++)
      for (int jj = 0; jj < 2; jj++)
        for (int ii = 0; ii < 2; ii++)
          h->d[g + ii + jj * gX + kk * gXY] += coeff * s->m[i] * w[ii + 2 * jj + 4 * kk];
  }
}

static void density_branch(oq_sim *s, onode *nd)
{
  int lv = nd->lv;
  double H = 1.0 / (1 << lv);
  double oneH = (double)(1 << lv);
  double coeff = 4.0 * s->G * OQ_PI / (H * H * H * H * H * H); /* :147 */
  int gX = gdx(nd), gXY = gdxy(nd);
  memset(nd->d, 0, sizeof(double) * (size_t)nd->gcap);
  for (int c = 0; c < nd->ncell; c++)
  {
    int b = nd->cbox[c];
    int cnt = nd->cell_cnt[b];
    if (cnt <= 0)
      continue;
    int p0 = nd->cell_list[nd->cell_start[b]];
    double px = s->x[p0] * oneH, py = s->y[p0] * oneH, pz = s->z[p0] * oneH;
    int fx = (int)px, fy = (int)py, fz = (int)pz;
    int g = (fx - nd->ts[0]) + (fy - nd->ts[1]) * gX + (fz - nd->ts[2]) * gXY;
    double rx = H * fx, ry = H * fy, rz = H * fz;
    for (int q = nd->cell_start[b]; q < nd->cell_start[b + 1]; q++)
    {
      int p = nd->cell_list[q];
      double wx1 = s->x[p] - rx, wy1 = s->y[p] - ry, wz1 = s->z[p] - rz;
      double wx2 = H - wx1, wy2 = H - wy1, wz2 = H - wz1;
      double w[8] = {wx2 * wy2 * wz2, wx1 * wy2 * wz2, wx2 * wy1 * wz2, wx1 * wy1 * wz2,
                     wx2 * wy2 * wz1, wx1 * wy2 * wz1, wx2 * wy1 * wz1, wx1 * wy1 * wz1};
      for (int kk = 0; kk < 2; kk++)
        for (int jj = 0; jj < 2; jj++)
          for (int ii = 0; ii < 2; ii++)
            nd->d[g + ii + jj * gX + kk * gXY] += coeff * s->m[p] * w[ii + 2 * jj + 4 * kk];
    }
  }
}

void oq_grid_density(oq_sim *s)
{
  density_head(s);
  for (int i = 1; i < s->nnodes; i++)
    density_branch(s, &s->nodes[i]);
}

/* ------------------------------------------------------------------------ */
/* A4-A7 head Poisson: multigrid.c, gauss_saidel.c, poisson_error.c           */
/* ------------------------------------------------------------------------ */
typedef struct
{
  int pre, post, coarse;
  double w;
} mgpar;

/* gauss_saidel.c:29-55 */
static void gs_sweeps(double *phi, const double *rho, int gs, int iters, double w)
{
  double H2 = (1.0 / (gs - 1)) * (1.0 / (gs - 1));
  int g2 = gs * gs, C1 = gs - 1, C2 = gs * (gs - 1), C3 = g2 * (gs - 1);
  for (int it = 0; it < iters; it++)
    for (int k = g2; k < C3; k += g2)
      for (int j = gs; j < C2; j += gs)
        for (int i = 1; i < C1; i++)
        {
          int g = k + j + i;
          double a = (1.0 / 6.0) * (-H2 * rho[g] + phi[g + 1] + phi[g - 1] + phi[g + gs] + phi[g - gs] +
                                    phi[g + g2] + phi[g - g2]);
          phi[g] = w * a + (1.0 - w) * phi[g];
        }
}

/* multigrid.c:106-229, V-cycle (type 0) with GS smoothing, injection
 * restriction (:29-49) and trilinear prolongation (:51-104). */
static void vcycle(double *phi, const double *rho, int gs, const mgpar *mp)
{
  int g2 = gs * gs, g3 = g2 * gs;
  int cs = (gs + 1) / 2, c2 = cs * cs, c3 = c2 * cs;
  double oneH2 = (double)(gs - 1) * (gs - 1);
  int C1 = gs - 1, C2 = gs * (gs - 1), C3 = g2 * (gs - 1);
  double *dphic = (double *)calloc((size_t)c3, sizeof(double));
  double *rest = (double *)calloc((size_t)g3, sizeof(double));
  gs_sweeps(phi, rho, gs, mp->pre, mp->w);
  for (int k = g2; k < C3; k += g2)
    for (int j = gs; j < C2; j += gs)
      for (int i = 1; i < C1; i++)
      {
        int g = k + j + i;
        rest[g] = rho[g] + oneH2 * (6 * phi[g] - phi[g + 1] - phi[g - 1] - phi[g + gs] - phi[g - gs] -
                                    phi[g + g2] - phi[g - g2]);
      }
  double *restc = (double *)calloc((size_t)c3, sizeof(double));
  for (int lz = 1; lz < cs - 1; lz++)
    for (int ly = 1; ly < cs - 1; ly++)
      for (int lx = 1; lx < cs - 1; lx++)
        restc[lz * c2 + ly * cs + lx] = rest[2 * lz * g2 + 2 * ly * gs + 2 * lx];
  if (cs < 4)
    gs_sweeps(phi, rho, gs, mp->coarse, mp->w); /* :158-173 quirk: smooths the fine phi */
  else
    vcycle(dphic, restc, cs, mp);
  double *dphi = (double *)calloc((size_t)g3, sizeof(double));
  for (int lz = 1; lz < cs - 1; lz++)
    for (int ly = 1; ly < cs - 1; ly++)
      for (int lx = 1; lx < cs - 1; lx++)
      {
        int A = 2 * lz * g2 + 2 * ly * gs + 2 * lx;
        double v = dphic[lz * c2 + ly * cs + lx];
        dphi[A] = v;
        int o1[6] = {1, -1, gs, -gs, g2, -g2};
        for (int q = 0; q < 6; q++)
          dphi[A + o1[q]] += 0.5 * v;
        int o2[12] = {gs + 1, gs - 1, -gs + 1, -gs - 1, g2 + 1, g2 - 1, -g2 + 1, -g2 - 1,
                      g2 + gs, g2 - gs, -g2 + gs, -g2 - gs};
        for (int q = 0; q < 12; q++)
          dphi[A + o2[q]] += 0.25 * v;
        int o3[8] = {g2 + gs + 1, g2 + gs - 1, g2 - gs + 1, g2 - gs - 1,
                     -g2 + gs + 1, -g2 + gs - 1, -g2 - gs + 1, -g2 - gs - 1};
        for (int q = 0; q < 8; q++)
          dphi[A + o3[q]] += 0.125 * v;
      }
  for (int k = g2; k < C3; k += g2)
    for (int j = gs; j < C2; j += gs)
      for (int i = 1; i < C1; i++)
        phi[k + j + i] += dphi[k + j + i];
  gs_sweeps(phi, rho, gs, mp->post, mp->w);
  free(dphic);
  free(rest);
  free(restc);
  free(dphi);
}

/* poisson_error.c:36-81 */
static int cmp_prev(const onode *nd, double tol2)
{
  double lo = 1.0 - tol2, hi = 1.0 + tol2;
  for (int i = 0; i < nd->nintr; i++)
  {
    int g = nd->iidx[i];
    double a = nd->pot[g] / nd->pot_old[g];
    if (lo > a || hi < a)
      return 0;
  }
  return 1;
}

/* poisson_error.c:83-133 (method 0, RMS residual over interior points) */
static int err_method0(const oq_sim *s, const onode *nd, double tol)
{
  double H = 1.0 / (1 << nd->lv);
  double oneH2 = 1.0 / (H * H);
  double rhomean = 4.0 * s->G * OQ_PI * nd->mass / (nd->ncell * H * H * H);
  int gX = gdx(nd), gXY = gdxy(nd);
  double err = 0.0;
  for (int i = 0; i < nd->nintr; i++)
  {
    int g = nd->iidx[i];
    const double *p = nd->pot;
    double diff = nd->d[g] + oneH2 * (6.0 * p[g] - p[g + 1] - p[g - 1] - p[g + gX] - p[g - gX] - p[g + gXY] - p[g - gXY]);
    err += diff * diff;
  }
  err = sqrt(err / nd->nintr) / rhomean;
  return err < tol;
}

int oq_head_solve(int lmin, double *phi, const double *rho, double node_mass, double tol, int max_iter)
{
  /* standalone version of potential_head_node on a packed cube, head node
   * geometry: interior = packed points 1..n-1, cell_size = n^3 */
  int n = 1 << lmin, gs = n + 1;
  size_t g3 = (size_t)gs * gs * gs;
  mgpar mp = {2, 2, 2, 1.0};
  double *old = (double *)malloc(sizeof(double) * g3);
  memcpy(old, phi, sizeof(double) * g3);
  double H = 1.0 / n, oneH2 = 1.0 / (H * H);
  double rhomean = 4.0 * OQ_PI * node_mass / ((double)n * n * n * H * H * H);
  int iter = 0, ok = 0;
  while (iter < max_iter && !ok)
  {
    iter++;
    vcycle(phi, rho, gs, &mp);
    if (iter % 5 == 0)
    {
      ok = 1;
      for (int k = 1; k < n && ok; k++)
        for (int j = 1; j < n && ok; j++)
          for (int i = 1; i < n; i++)
          {
            size_t g = i + (size_t)gs * (j + (size_t)gs * k);
            double a = phi[g] / old[g];
            if (1.0 - tol > a || 1.0 + tol < a)
            {
              ok = 0;
              break;
            }
          }
      if (ok)
      {
        double err = 0.0;
        size_t cnt = 0;
        for (int k = 1; k < n; k++)
          for (int j = 1; j < n; j++)
            for (int i = 1; i < n; i++)
            {
              size_t g = i + (size_t)gs * (j + (size_t)gs * k);
              double diff = rho[g] + oneH2 * (6.0 * phi[g] - phi[g + 1] - phi[g - 1] - phi[g + gs] - phi[g - gs] -
                                              phi[g + (size_t)gs * gs] - phi[g - (size_t)gs * gs]);
              err += diff * diff;
              cnt++;
            }
        err = sqrt(err / cnt) / rhomean;
        ok = err < tol;
      }
      if (!ok)
        memcpy(old, phi, sizeof(double) * g3);
    }
  }
  free(old);
  return iter == max_iter ? -1 : iter;
}

/* potential_head_node.c:32-155 */
static int potential_head(oq_sim *s)
{
  onode *h = &s->nodes[0];
  int n = 1 << s->lmin;
  int size = n + 1, size2 = size * size;
  int gsd = h->rd[0] + 1, gsd2 = gsd * gsd;
  double *ap = (double *)malloc(sizeof(double) * (size_t)size2 * size);
  double *ad = (double *)malloc(sizeof(double) * (size_t)size2 * size);
  mgpar mp = {s->pre, s->post, s->coarse, s->w_head};
  memcpy(h->pot_old, h->pot, sizeof(double) * (size_t)h->gcap); /* :110 */
  int iter = 0, check = 0;
  while (iter < s->max_iter && !check)
  {
    iter++;
    for (int k = 1; k < gsd - 1; k++)
      for (int j = 1; j < gsd - 1; j++)
        for (int i = 1; i < gsd - 1; i++)
        {
          int bg = i + j * gsd + k * gsd2;
          int ai = (i - 1) + (j - 1) * size + (k - 1) * size2;
          ap[ai] = h->pot[bg];
          ad[ai] = h->d[bg];
        }
    vcycle(ap, ad, size, &mp);
    for (int k = 1; k < gsd - 1; k++)
      for (int j = 1; j < gsd - 1; j++)
        for (int i = 1; i < gsd - 1; i++)
        {
          int bg = i + j * gsd + k * gsd2;
          int ai = (i - 1) + (j - 1) * size + (k - 1) * size2;
          h->pot[bg] = ap[ai];
        }
    if (iter % s->check_every == 0)
    {
      check = cmp_prev(h, s->tol2) && err_method0(s, h, s->tol); /* poisson_error.c:290-347 */
      if (!check)
        memcpy(h->pot_old, h->pot, sizeof(double) * (size_t)h->gcap);
    }
  }
  free(ap);
  free(ad);
  s->stat_vcycles = iter;
  if (iter == s->max_iter)
    return 1; /* :151-155 */
  return 0;
}

/* ------------------------------------------------------------------------ */
/* A8 potential_branches.c:93-332, A9 conjugate_gradient.c:322-463            */
/* ------------------------------------------------------------------------ */

/* 8-point parent average at ((c>>1),(c+1)>>1) per axis (potential_branches.c:204-331,
 * grid_acceleration.c:316-383). */
static double parent_avg(const onode *pt, const double *f, int cx, int cy, int cz)
{
  int a1x = cx >> 1, a1y = cy >> 1, a1z = cz >> 1;
  int a2x = (cx + 1) >> 1, a2y = (cy + 1) >> 1, a2z = (cz + 1) >> 1;
  int b1x = a1x - pt->ts[0], b1y = a1y - pt->ts[1], b1z = a1z - pt->ts[2];
  int b2x = a2x - pt->ts[0], b2y = a2y - pt->ts[1], b2z = a2z - pt->ts[2];
  int X = gdx(pt), XY = gdxy(pt);
  return 0.125 * (f[b1x + b1y * X + b1z * XY] + f[b2x + b1y * X + b1z * XY] + f[b1x + b2y * X + b1z * XY] +
                  f[b2x + b2y * X + b1z * XY] + f[b1x + b1y * X + b2z * XY] + f[b2x + b1y * X + b2z * XY] +
                  f[b1x + b2y * X + b2z * XY] + f[b2x + b2y * X + b2z * XY]);
}

/* Simulation-boundary coordinates as the reference computes them: an index
 * equal to 0 or 2^lv maps to 0.0 (potential_branches.c:114-145); for the
 * acceleration the y coordinate lacks the factor H (grid_acceleration.c:240). */
static void simb_coords(int lv, const int *xyz, double *ax, double *ay, double *az, int accel_quirk)
{
  double H = 1.0 / (1 << lv);
  int top = 1 << lv;
  *ax = (xyz[0] == 0 || xyz[0] == top) ? 0.0 : xyz[0] * H;
  if (xyz[1] == 0 || xyz[1] == top)
    *ay = 0.0;
  else
    *ay = accel_quirk ? (double)xyz[1] : xyz[1] * H;
  *az = (xyz[2] == 0 || xyz[2] == top) ? 0.0 : xyz[2] * H;
}

static void branch_initial_potential(oq_sim *s, onode *nd)
{
  const onode *pt = &s->nodes[nd->parent];
  /* sim boundary: potential_branches.c:93-151 */
  for (int i = 0; i < nd->nsim; i++)
  {
    double ax, ay, az;
    simb_coords(nd->lv, &nd->sxyz[3 * i], &ax, &ay, &az, 0);
    double dist = sqrt((ax - 0.5) * (ax - 0.5) + (ay - 0.5) * (ay - 0.5) + (az - 0.5) * (az - 0.5));
    nd->pot[nd->sidx[i]] = s->head_bc == 0 ? -s->G * s->total_mass / dist : 0.0;
  }
  for (int i = 0; i < nd->nbdry; i++)
    nd->pot[nd->bidx[i]] = parent_avg(pt, pt->pot, nd->bxyz[3 * i], nd->bxyz[3 * i + 1], nd->bxyz[3 * i + 2]);
  for (int i = 0; i < nd->nintr; i++)
    nd->pot[nd->iidx[i]] = parent_avg(pt, pt->pot, nd->ixyz[3 * i], nd->ixyz[3 * i + 1], nd->ixyz[3 * i + 2]);
}

/* conjugate_gradient.c:322-463 */
static int branch_cg(oq_sim *s, onode *nd)
{
  int ni = nd->nintr;
  double H = 1.0 / (1 << nd->lv);
  double oneH2 = 1.0 / (H * H);
  int size = nd->gcap;
  int X = gdx(nd), XY = gdxy(nd);
  double *r = (double *)calloc((size_t)size, sizeof(double));
  double *p = (double *)calloc((size_t)size, sizeof(double));
  double *Ap = (double *)malloc(sizeof(double) * (size_t)size);
  double rsold = 0.0, rsnew = 0.0;
  double *phi = nd->pot;
  for (int i = 0; i < ni; i++)
  {
    int g = nd->iidx[i];
    r[g] = nd->d[g] + oneH2 * (6.0 * phi[g] - phi[g + 1] - phi[g - 1] - phi[g + X] - phi[g - X] - phi[g + XY] - phi[g - XY]);
    p[g] = r[g];
    rsold += r[g] * r[g];
  }
  double rhomean = 4.0 * s->G * OQ_PI * nd->mass / (nd->ncell * H * H * H);
  double aux_tol = ni * (s->tol * rhomean) * (s->tol * rhomean); /* poisson_error.c:242-253 */
  int cycle = size < s->max_iter ? size : s->max_iter;
  for (int it = 0; it < cycle; it++)
  {
    double pAp = 0.0;
    for (int j = 0; j < ni; j++)
    {
      int g = nd->iidx[j];
      Ap[g] = -oneH2 * (6.0 * p[g] - p[g + 1] - p[g - 1] - p[g + X] - p[g - X] - p[g + XY] - p[g - XY]);
      pAp += p[g] * Ap[g];
    }
    double alpha = rsold / pAp;
    rsnew = 0.0;
    for (int j = 0; j < ni; j++)
    {
      int g = nd->iidx[j];
      phi[g] += alpha * p[g];
      r[g] -= alpha * Ap[g];
      rsnew += r[g] * r[g];
    }
    s->stat_cg_iters++;
    if ((it + 1) % s->check_every == 0)
    {
      int ok = cmp_prev(nd, s->tol2) && rsnew < aux_tol;
      if (!ok)
        memcpy(nd->pot_old, nd->pot, sizeof(double) * (size_t)nd->gcap);
      if (ok)
      {
        free(r);
        free(p);
        free(Ap);
        return 0;
      }
    }
    for (int j = 0; j < ni; j++)
    {
      int g = nd->iidx[j];
      p[g] = r[g] + rsnew / rsold * p[g];
    }
    rsold = rsnew;
  }
  free(r);
  free(p);
  free(Ap);
  return 1;
}

int oq_potential(oq_sim *s)
{
  s->stat_cg_iters = 0;
  s->stat_cg_solves = 0;
  if (potential_head(s))
    return 1;
  for (int i = 1; i < s->nnodes; i++) /* level order (potential_branches.c:441-516) */
  {
    onode *nd = &s->nodes[i];
    branch_initial_potential(s, nd);
    memcpy(nd->pot_old, nd->pot, sizeof(double) * (size_t)nd->gcap);
    s->stat_cg_solves++;
    if (branch_cg(s, nd))
      return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* A11 grid_acceleration.c:36-76, 199-394                                    */
/* ------------------------------------------------------------------------ */
static void grid_acc_central(onode *nd)
{
  double H = 1.0 / (1 << nd->lv);
  double o2H = 1.0 / (2.0 * H);
  int X = gdx(nd), XY = gdxy(nd);
  for (int i = 0; i < nd->nintr; i++)
  {
    int g = nd->iidx[i];
    nd->ax[g] = o2H * (nd->pot[g - 1] - nd->pot[g + 1]);
    nd->ay[g] = o2H * (nd->pot[g - X] - nd->pot[g + X]);
    nd->az[g] = o2H * (nd->pot[g - XY] - nd->pot[g + XY]);
  }
}

void oq_grid_acceleration(oq_sim *s)
{
  grid_acc_central(&s->nodes[0]);
  for (int i = 1; i < s->nnodes; i++)
  {
    onode *nd = &s->nodes[i];
    const onode *pt = &s->nodes[nd->parent];
    /* sim boundary (:199-268) */
    for (int q = 0; q < nd->nsim; q++)
    {
      double ax, ay, az;
      simb_coords(nd->lv, &nd->sxyz[3 * q], &ax, &ay, &az, 1);
      double dist = sqrt((ax - 0.5) * (ax - 0.5) + (ay - 0.5) * (ay - 0.5) + (az - 0.5) * (az - 0.5));
      double c2 = s->head_bc == 0 ? -s->G * s->total_mass / (dist * dist * dist) : 0.0;
      int g = nd->sidx[q];
      nd->ax[g] = c2 * (ax - 0.5);
      nd->ay[g] = c2 * (ay - 0.5);
      nd->az[g] = c2 * (az - 0.5);
    }
    /* boundary: parent accelerations (:316-383) */
    for (int q = 0; q < nd->nbdry; q++)
    {
      int g = nd->bidx[q];
      const int *c = &nd->bxyz[3 * q];
      nd->ax[g] = parent_avg(pt, pt->ax, c[0], c[1], c[2]);
      nd->ay[g] = parent_avg(pt, pt->ay, c[0], c[1], c[2]);
      nd->az[g] = parent_avg(pt, pt->az, c[0], c[1], c[2]);
    }
    grid_acc_central(nd);
  }
}

/* ------------------------------------------------------------------------ */
/* A12 particle_acceleration.c:33-355                                         */
/* ------------------------------------------------------------------------ */
static void gather_one(oq_sim *s, const onode *nd, int p)
{
  int lv = nd->lv;
  double px = s->x[p] * (1 << lv), py = s->y[p] * (1 << lv), pz = s->z[p] * (1 << lv);
  int fx = (int)px, fy = (int)py, fz = (int)pz;
  double wx1 = px - fx, wy1 = py - fy, wz1 = pz - fz;
  double wx2 = 1.0 - wx1, wy2 = 1.0 - wy1, wz2 = 1.0 - wz1;
  double w[8] = {wx2 * wy2 * wz2, wx1 * wy2 * wz2, wx2 * wy1 * wz2, wx1 * wy1 * wz2,
                 wx2 * wy2 * wz1, wx1 * wy2 * wz1, wx2 * wy1 * wz1, wx1 * wy1 * wz1};
  int X = gdx(nd), XY = gdxy(nd);
  int g = (fx - nd->ts[0]) + (fy - nd->ts[1]) * X + (fz - nd->ts[2]) * XY;
  for (int kk = 0; kk < 2; kk++)
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++)
      {
        int q = ii + 2 * jj + 4 * kk;
        int gn = g + ii + jj * X + kk * XY;
        s->ax[p] += nd->ax[gn] * w[q];
        s->ay[p] += nd->ay[gn] * w[q];
        s->az[p] += nd->az[gn] * w[q];
      }
}

void oq_particle_acceleration(oq_sim *s)
{
  if (s->lmin == s->lmax)
  {
    for (int p = 0; p < s->n; p++)
      gather_one(s, &s->nodes[0], p);
    return;
  }
  for (int i = 0; i < s->nnodes; i++)
  {
    const onode *nd = &s->nodes[i];
    for (int c = 0; c < nd->ncell; c++)
    {
      int b = nd->cbox[c];
      if (nd->nchild > 0 && nd->box[b] >= 0)
        continue; /* refined cell: its particles take a from the child */
      for (int q = nd->cell_start[b]; q < nd->cell_start[b + 1]; q++)
        gather_one(s, nd, nd->cell_list[q]);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* A14/A15 particle_updating.c:39-207 (reflexive) and 2677-2836               */
/* ------------------------------------------------------------------------ */
int oq_update_A(oq_sim *s, double dt)
{
  for (int i = 0; i < s->n; i++)
  {
    s->vx[i] += s->ax[i] * dt * 0.5;
    s->vy[i] += s->ay[i] * dt * 0.5;
    s->vz[i] += s->az[i] * dt * 0.5;
    s->x[i] += s->vx[i] * dt;
    s->y[i] += s->vy[i] * dt;
    s->z[i] += s->vz[i] * dt;
    if (s->x[i] < 0.0 || s->x[i] >= 1.0)
    {
      s->x[i] = s->x[i] >= 1.0 ? 2.0 - s->x[i] : -s->x[i];
      s->vx[i] *= -1.0;
      s->ax[i] *= -1.0;
    }
    if (s->y[i] < 0.0 || s->y[i] >= 1.0)
    {
      s->y[i] = s->y[i] >= 1.0 ? 2.0 - s->y[i] : -s->y[i];
      s->vy[i] *= -1.0;
      s->ay[i] *= -1.0;
    }
    if (s->z[i] < 0.0 || s->z[i] >= 1.0)
    {
      s->z[i] = s->z[i] >= 1.0 ? 2.0 - s->z[i] : -s->z[i];
      s->vz[i] *= -1.0;
      s->az[i] *= -1.0;
    }
  }
  return 0;
}

void oq_update_B(oq_sim *s, double dt)
{
  for (int i = 0; i < s->n; i++)
  {
    s->vx[i] += s->ax[i] * dt * 0.5;
    s->vy[i] += s->ay[i] * dt * 0.5;
    s->vz[i] += s->az[i] * dt * 0.5;
  }
}

int oq_force(oq_sim *s)
{
  if (s->lmin < s->lmax)
    oq_build_tree(s);
  oq_reset(s);
  oq_grid_density(s);
  if (oq_potential(s))
  {
    s->stat_fail = 1;
    return 1;
  }
  oq_grid_acceleration(s);
  oq_particle_acceleration(s);
  return 0;
}

int oq_step(oq_sim *s, double dt)
{
  if (oq_update_A(s, dt))
    return 1;
  if (oq_force(s))
    return 1;
  oq_update_B(s, dt);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* A16 grid potential energy (potential_energy_type = 1, observables.c:114-619): */
/* 1/2 G sum m_i phi(x_i), phi interpolated with CIC weights at each particle's  */
/* finest node (the head alone without levels).                                 */
/* ------------------------------------------------------------------------ */
static double pot_one(const oq_sim *s, const onode *nd, int p)
{
  int lv = nd->lv;
  double px = s->x[p] * (1 << lv), py = s->y[p] * (1 << lv), pz = s->z[p] * (1 << lv);
  int fx = (int)px, fy = (int)py, fz = (int)pz;
  double wx1 = px - fx, wy1 = py - fy, wz1 = pz - fz;
  double w[2][3] = {{1.0 - wx1, 1.0 - wy1, 1.0 - wz1}, {wx1, wy1, wz1}};
  int X = gdx(nd), XY = gdxy(nd);
  int g = (fx - nd->ts[0]) + (fy - nd->ts[1]) * X + (fz - nd->ts[2]) * XY;
  double v = 0.0;
  for (int kk = 0; kk < 2; kk++)
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++)
        v += nd->pot[g + ii + jj * X + kk * XY] * w[ii][0] * w[jj][1] * w[kk][2];
  return v;
}

double oq_grid_pe(const oq_sim *s)
{
  double e = 0.0;
  if (s->lmin == s->lmax)
  {
    for (int p = 0; p < s->n; p++)
      e += s->m[p] * pot_one(s, &s->nodes[0], p);
  }
  else
    for (int q = 0; q < s->nnodes; q++)
    {
      const onode *nd = &s->nodes[q];
      for (int c = 0; c < nd->ncell; c++)
      {
        int b = nd->cbox[c];
        if (nd->nchild > 0 && nd->box[b] >= 0)
          continue;
        for (int t = nd->cell_start[b]; t < nd->cell_start[b + 1]; t++)
          e += s->m[nd->cell_list[t]] * pot_one(s, nd, nd->cell_list[t]);
      }
    }
  return 0.5 * s->G * e; /* observables.c:617 */
}

/* ------------------------------------------------------------------------ */
/* 8(f).2 CFL time step, time_step_method = 1 (timestep.c:73-195, 370-587,      */
/* 589-699): per particle and axis the dt that moves it CFL * H of its finest    */
/* node under constant velocity and acceleration, min over all, capped at      */
/* dt_max.                                                                     */
/* ------------------------------------------------------------------------ */
static double cfl_axis(double v, double a, double H, double cfl, double dt)
{
  double av = fabs(v), aa = fabs(a), t;
  if (aa < 1.0e-16)
  {
    if (av > 1.0e-16)
    {
      t = cfl * H / av; /* timestep.c:86-93 */
      dt = dt < t ? dt : t;
    }
    return dt;
  }
  double c = v * v + 2.0 * aa * H * cfl; /* :96-106 */
  if (1.0 - v * v / c > 1.0e-8)
    t = (-av + sqrt(c)) / aa;
  else
  {
    double eps = fabs(a * H * cfl / (v * v));
    t = fabs(v / a) * (eps - eps * eps * 0.5);
  }
  return dt < t ? dt : t;
}

double oq_cfl_dt(int n, const double *v3, const double *a3, const double *H, double cfl, double dt_max)
{
  double dt = dt_max;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++)
      dt = cfl_axis(v3[(size_t)k * n + i], a3[(size_t)k * n + i], H[i], cfl, dt);
  return dt;
}

double oq_timestep(const oq_sim *s, double cfl, double dt_max)
{
  double *H = (double *)malloc(sizeof(double) * (size_t)s->n);
  if (s->lmin == s->lmax)
    for (int i = 0; i < s->n; i++)
      H[i] = 1.0 / (1 << s->lmin); /* timestep.c:73-195 */
  else
    for (int q = 0; q < s->nnodes; q++) /* :370-587: unrefined cells of every node */
    {
      const onode *nd = &s->nodes[q];
      for (int c = 0; c < nd->ncell; c++)
      {
        int b = nd->cbox[c];
        if (nd->nchild > 0 && nd->box[b] >= 0)
          continue;
        for (int t = nd->cell_start[b]; t < nd->cell_start[b + 1]; t++)
          H[nd->cell_list[t]] = 1.0 / (1 << nd->lv);
      }
    }
  double *v3 = (double *)malloc(sizeof(double) * 3 * (size_t)s->n), *a3 = v3 ? (double *)malloc(sizeof(double) * 3 * (size_t)s->n) : NULL;
  double dt = dt_max;
  if (v3 && a3)
  {
    memcpy(v3, s->vx, sizeof(double) * (size_t)s->n);
    memcpy(v3 + s->n, s->vy, sizeof(double) * (size_t)s->n);
    memcpy(v3 + 2 * (size_t)s->n, s->vz, sizeof(double) * (size_t)s->n);
    memcpy(a3, s->ax, sizeof(double) * (size_t)s->n);
    memcpy(a3 + s->n, s->ay, sizeof(double) * (size_t)s->n);
    memcpy(a3 + 2 * (size_t)s->n, s->az, sizeof(double) * (size_t)s->n);
    dt = oq_cfl_dt(s->n, v3, a3, H, cfl, dt_max);
  }
  free(v3);
  free(a3);
  free(H);
  return dt;
}

/* ------------------------------------------------------------------------ */
/* A16 observables.c:38-112, 622-659                                          */
/* ------------------------------------------------------------------------ */
void oq_observables(const oq_sim *s, double *out6)
{
  double ek = 0.0, ep = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  for (int i = 0; i < s->n; i++)
  {
    double v2 = s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i];
    ek += s->m[i] * v2;
    px += s->m[i] * s->vx[i];
    py += s->m[i] * s->vy[i];
    pz += s->m[i] * s->vz[i];
  }
  for (int i = 0; i < s->n; i++)
  {
    double sum = 0.0;
    for (int j = 0; j < i; j++)
    {
      double dx = s->x[i] - s->x[j], dy = s->y[i] - s->y[j], dz = s->z[i] - s->z[j];
      sum += s->m[i] * s->m[j] / sqrt(dx * dx + dy * dy + dz * dz);
    }
    ep += sum;
  }
  out6[0] = 0.5 * ek;
  out6[1] = -ep * s->G;
  out6[2] = out6[0] + out6[1];
  out6[3] = px;
  out6[4] = py;
  out6[5] = pz;
}

/* ------------------------------------------------------------------------ */
/* construction and accessors                                                 */
/* ------------------------------------------------------------------------ */
oq_sim *oq_create(int lmin, int lmax, int ref_crit, int n, const double *x, const double *y, const double *z,
                  const double *vx, const double *vy, const double *vz, const double *m, int head_bc)
{
  oq_sim *s = (oq_sim *)calloc(1, sizeof(oq_sim));
  s->lmin = lmin;
  s->lmax = lmax;
  s->n = n;
  s->ref_crit = ref_crit;
  s->n_exp = 1;
  s->extra = 10; /* min_box_extra_size_per_side_real, global_variables.c:294-295 */
  s->head_bc = head_bc;
  s->G = 1.0;
  s->tol = 1e-10; /* global_variables.c:318-319 */
  s->tol2 = 1e-10;
  s->max_iter = 5000; /* :313 */
  s->check_every = 5; /* :332 */
  s->pre = s->post = s->coarse = 2;
  s->w_head = 1.0;
  double **dst[10] = {&s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz, &s->ax, &s->ay, &s->az, &s->m};
  for (int a = 0; a < 10; a++)
    *dst[a] = (double *)calloc((size_t)n, sizeof(double));
  memcpy(s->x, x, sizeof(double) * (size_t)n);
  memcpy(s->y, y, sizeof(double) * (size_t)n);
  memcpy(s->z, z, sizeof(double) * (size_t)n);
  memcpy(s->vx, vx, sizeof(double) * (size_t)n);
  memcpy(s->vy, vy, sizeof(double) * (size_t)n);
  memcpy(s->vz, vz, sizeof(double) * (size_t)n);
  memcpy(s->m, m, sizeof(double) * (size_t)n);
  s->total_mass = 0.0;
  for (int i = 0; i < n; i++)
    s->total_mass += m[i];
  build_head(s);
  clear_tree(s);
  return s;
}

void oq_destroy(oq_sim *s)
{
  for (int i = 0; i < s->nnodes; i++)
    node_free(&s->nodes[i]);
  free(s->nodes);
  free(s->x);
  free(s->y);
  free(s->z);
  free(s->vx);
  free(s->vy);
  free(s->vz);
  free(s->ax);
  free(s->ay);
  free(s->az);
  free(s->m);
  free(s);
}

void oq_set_tolerance(oq_sim *s, double tol)
{
  s->tol = tol;
  s->tol2 = tol;
}

int oq_num_particles(const oq_sim *s) { return s->n; }

void oq_get_particles(const oq_sim *s, double *out9)
{
  const double *src[9] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az};
  for (int a = 0; a < 9; a++)
    memcpy(out9 + (size_t)a * s->n, src[a], sizeof(double) * (size_t)s->n);
}

void oq_set_particles(oq_sim *s, const double *in9)
{
  double *dst[9] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az};
  for (int a = 0; a < 9; a++)
    memcpy(dst[a], in9 + (size_t)a * s->n, sizeof(double) * (size_t)s->n);
}

void oq_get_head(const oq_sim *s, double *out5)
{
  const onode *h = &s->nodes[0];
  int n1 = (1 << s->lmin) + 1, g = h->rd[0] + 1;
  const double *src[5] = {h->d, h->pot, h->ax, h->ay, h->az};
  size_t n3 = (size_t)n1 * n1 * n1;
  for (int a = 0; a < 5; a++)
    for (int k = 0; k < n1; k++)
      for (int j = 0; j < n1; j++)
        for (int i = 0; i < n1; i++)
          out5[a * n3 + i + (size_t)n1 * (j + (size_t)n1 * k)] = src[a][(i + 1) + g * ((j + 1) + g * (k + 1))];
}

void oq_set_head_pot(oq_sim *s, const double *pot)
{
  onode *h = &s->nodes[0];
  int n1 = (1 << s->lmin) + 1, g = h->rd[0] + 1;
  for (int k = 0; k < n1; k++)
    for (int j = 0; j < n1; j++)
      for (int i = 0; i < n1; i++)
        h->pot[(i + 1) + g * ((j + 1) + g * (k + 1))] = pot[i + (size_t)n1 * (j + (size_t)n1 * k)];
}

int oq_num_nodes(const oq_sim *s) { return s->nnodes - 1; }

void oq_node_header(const oq_sim *s, int i, int *h)
{
  const onode *nd = &s->nodes[i + 1];
  int lvrel = nd->lv - s->lmin;
  int pidx = nd->parent - s->lev_start[lvrel - 1];
  int hdr[24] = {nd->lv, lvrel, (i + 1) - s->lev_start[lvrel], pidx,
                 nd->ts[0], nd->ts[1], nd->ts[2], nd->rd[0], nd->rd[1], nd->rd[2],
                 nd->bmin[0], nd->bmin[1], nd->bmin[2], nd->bmax[0], nd->bmax[1], nd->bmax[2],
                 nd->ncell, nd->nintr, nd->nbdry, nd->nsim, nd->nptcl_full, nd->nptcl_outs, nd->nchild, nd->gcap};
  memcpy(h, hdr, sizeof(hdr));
}

void oq_node_ints(const oq_sim *s, int i, int which, int *out)
{
  const onode *nd = &s->nodes[i + 1];
  switch (which)
  {
  case 0:
    for (int c = 0; c < nd->ncell; c++)
    {
      out[3 * c] = nd->cx[c];
      out[3 * c + 1] = nd->cy[c];
      out[3 * c + 2] = nd->cz[c];
    }
    break;
  case 1:
    memcpy(out, nd->cbox, sizeof(int) * (size_t)nd->ncell);
    break;
  case 2:
    memcpy(out, nd->ixyz, sizeof(int) * 3 * (size_t)nd->nintr);
    break;
  case 3:
    memcpy(out, nd->iidx, sizeof(int) * (size_t)nd->nintr);
    break;
  case 4:
    memcpy(out, nd->bxyz, sizeof(int) * 3 * (size_t)nd->nbdry);
    break;
  case 5:
    memcpy(out, nd->bidx, sizeof(int) * (size_t)nd->nbdry);
    break;
  case 6:
    memcpy(out, nd->sxyz, sizeof(int) * 3 * (size_t)nd->nsim);
    break;
  case 7:
    memcpy(out, nd->sidx, sizeof(int) * (size_t)nd->nsim);
    break;
  case 8:
    memcpy(out, nd->box, sizeof(int) * (size_t)nd->boxcap);
    break;
  default:
    break;
  }
}

void oq_node_fields(const oq_sim *s, int i, double *out5)
{
  const onode *nd = &s->nodes[i + 1];
  const double *src[5] = {nd->d, nd->pot, nd->ax, nd->ay, nd->az};
  for (int a = 0; a < 5; a++)
    memcpy(out5 + (size_t)a * nd->gcap, src[a], sizeof(double) * (size_t)nd->gcap);
}

void oq_stats(const oq_sim *s, int *out4)
{
  out4[0] = s->stat_vcycles;
  out4[1] = s->stat_cg_iters;
  out4[2] = s->stat_cg_solves;
  out4[3] = s->nnodes - 1;
}

/* ------------------------------------------------------------------------ */
/* FP64 reverse mode of the head-grid step (checker for the GPU adjoint).      */
/* Forward force pass a = A(x): CIC deposit (grid_density.c:35-111), Poisson   */
/* solve (potential_head_node.c), central-difference gradient                 */
/* (grid_acceleration.c:36-76), CIC gather (particle_acceleration.c:33-132).   */
/* ------------------------------------------------------------------------ */

/* solve Lap psi = rhs on the packed cube with zero faces until the RMS residual
 * falls below tol times the RMS right-hand side (V-cycles of multigrid.c) */
static int adj_solve(int lmin, double *psi, const double *rhs, double tol)
{
  int n = 1 << lmin, gs = n + 1;
  size_t g3 = (size_t)gs * gs * gs;
  mgpar mp = {2, 2, 2, 1.0};
  double oneH2 = (double)n * n;
  double r0 = 0.0;
  size_t cnt = 0;
  for (int k = 1; k < n; k++)
    for (int j = 1; j < n; j++)
      for (int i = 1; i < n; i++)
      {
        size_t g = i + (size_t)gs * (j + (size_t)gs * k);
        r0 += rhs[g] * rhs[g];
        cnt++;
      }
  r0 = sqrt(r0 / cnt);
  memset(psi, 0, sizeof(double) * g3);
  if (r0 == 0.0)
    return 0;
  for (int it = 1; it <= 500; it++)
  {
    vcycle(psi, rhs, gs, &mp);
    double err = 0.0;
    for (int k = 1; k < n; k++)
      for (int j = 1; j < n; j++)
        for (int i = 1; i < n; i++)
        {
          size_t g = i + (size_t)gs * (j + (size_t)gs * k);
          double d = rhs[g] - oneH2 * (psi[g + 1] + psi[g - 1] + psi[g + gs] + psi[g - gs] + psi[g + (size_t)gs * gs] +
                                       psi[g - (size_t)gs * gs] - 6.0 * psi[g]);
          err += d * d;
        }
    if (sqrt(err / cnt) < tol * r0)
      return it;
  }
  return -1;
}

/* ------------------------------------------------------------------------ */
/* FP64 reverse mode of the force pass with refinement levels.  The tree is   */
/* held fixed (cell membership is piecewise constant in the positions).       */
/* Forward, per node in level order: rho from the particles of its cells      */
/* (grid_density.c:113-267), boundary potential and accelerations from the    */
/* parent average (potential_branches.c:204-331, grid_acceleration.c:316-383),*/
/* interior solve (conjugate_gradient.c:322-463), interior central            */
/* differences (grid_acceleration.c:36-76), gather from the finest node       */
/* (particle_acceleration.c:134-355).  Reverse, finest level first:           */
/*   Abar  = scatter of abar from the particles whose finest node it is,      */
/*           plus the children's boundary Abar through the transposed average */
/*   phibar = transposed central difference of Abar, plus the children's      */
/*           boundary phibar through the transposed average                  */
/*   L psi = phibar on the interior (zero on the rest): rhobar = psi          */
/*   boundary phibar -= (1/H^2) sum of psi over its interior neighbours       */
/* and x_p collects c_l m grad W_l . psi_l over every node holding p plus      */
/* grad W_f . (abar . a_f) at its finest node f.                              */
/* ------------------------------------------------------------------------ */
typedef struct
{
  double *ab[3], *pb, *psi;
  unsigned char *kind; /* 0 none, 1 simulation boundary, 2 boundary, 3 interior */
} oadj;

/* add v / 8 to the 8 parent points of child grid point c (transpose of parent_avg) */
static void parent_avg_t(const onode *pt, double *f, const int *c, double v)
{
  int a[2][3];
  for (int q = 0; q < 3; q++)
  {
    a[0][q] = (c[q] >> 1) - pt->ts[q];
    a[1][q] = ((c[q] + 1) >> 1) - pt->ts[q];
  }
  int X = gdx(pt), XY = gdxy(pt);
  for (int kk = 0; kk < 2; kk++)
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++)
        f[a[ii][0] + a[jj][1] * X + a[kk][2] * XY] += 0.125 * v;
}

/* CIC weights and their x, y, z derivatives at the cell of p at level lv */
static int cic_grad(const onode *nd, double x, double y, double z, double w[8], double dw[3][8])
{
  int lv = nd->lv;
  double sc = (double)(1 << lv);
  double px = x * sc, py = y * sc, pz = z * sc;
  int fx = (int)px, fy = (int)py, fz = (int)pz;
  double t[3] = {px - fx, py - fy, pz - fz};
  for (int kk = 0; kk < 2; kk++)
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++)
      {
        int q = ii + 2 * jj + 4 * kk;
        double wx = ii ? t[0] : 1.0 - t[0], wy = jj ? t[1] : 1.0 - t[1], wz = kk ? t[2] : 1.0 - t[2];
        double dx = ii ? sc : -sc, dy = jj ? sc : -sc, dz = kk ? sc : -sc;
        w[q] = wx * wy * wz;
        dw[0][q] = dx * wy * wz;
        dw[1][q] = wx * dy * wz;
        dw[2][q] = wx * wy * dz;
      }
  return (fx - nd->ts[0]) + (fy - nd->ts[1]) * gdx(nd) + (fz - nd->ts[2]) * gdxy(nd);
}

/* L psi = b on the interior points of a node (the operator of
 * conjugate_gradient.c:322-463), psi zero elsewhere; CG to rr <= 1e-28 rr0 */
static int node_adj_solve(const onode *nd, const unsigned char *kind, double *psi, const double *b)
{
  int ni = nd->nintr, X = gdx(nd), XY = gdxy(nd);
  double oneH2 = (double)(1 << nd->lv) * (double)(1 << nd->lv);
  double *r = (double *)calloc((size_t)nd->gcap, sizeof(double));
  double *p = (double *)calloc((size_t)nd->gcap, sizeof(double));
  double *Ap = (double *)calloc((size_t)nd->gcap, sizeof(double));
  double rs = 0.0;
  memset(psi, 0, sizeof(double) * (size_t)nd->gcap);
  for (int i = 0; i < ni; i++)
  {
    int g = nd->iidx[i];
    r[g] = p[g] = b[g];
    rs += b[g] * b[g];
  }
  double rs0 = rs;
  int rc = 1;
  for (int it = 0; it < 20 * ni + 100 && rs > 0.0; it++)
  {
    if (rs <= 1e-28 * rs0)
    {
      rc = 0;
      break;
    }
    double pAp = 0.0;
    for (int i = 0; i < ni; i++)
    {
      int g = nd->iidx[i];
      Ap[g] = oneH2 * (p[g + 1] + p[g - 1] + p[g + X] + p[g - X] + p[g + XY] + p[g - XY] - 6.0 * p[g]);
      pAp += p[g] * Ap[g];
    }
    double al = rs / pAp, rn = 0.0;
    for (int i = 0; i < ni; i++)
    {
      int g = nd->iidx[i];
      psi[g] += al * p[g];
      r[g] -= al * Ap[g];
      rn += r[g] * r[g];
    }
    for (int i = 0; i < ni; i++)
    {
      int g = nd->iidx[i];
      p[g] = r[g] + rn / rs * p[g];
    }
    rs = rn;
  }
  if (rs == 0.0 || rs <= 1e-28 * rs0)
    rc = 0;
  (void)kind;
  free(r);
  free(p);
  free(Ap);
  return rc;
}

static int amr_force_vjp(oq_sim *s, const double *abar, double *xbar)
{
  int nn = s->nnodes, rc = 0;
  size_t N = (size_t)s->n;
  oadj *A = (oadj *)calloc((size_t)nn, sizeof(oadj));
  for (int i = 0; i < nn; i++)
  {
    const onode *nd = &s->nodes[i];
    for (int a = 0; a < 3; a++)
      A[i].ab[a] = (double *)calloc((size_t)nd->gcap, sizeof(double));
    A[i].pb = (double *)calloc((size_t)nd->gcap, sizeof(double));
    A[i].psi = (double *)calloc((size_t)nd->gcap, sizeof(double));
    A[i].kind = (unsigned char *)calloc((size_t)nd->gcap, 1);
    for (int q = 0; q < nd->nsim; q++)
      A[i].kind[nd->sidx[q]] = 1;
    for (int q = 0; q < nd->nbdry; q++)
      A[i].kind[nd->bidx[q]] = 2;
    for (int q = 0; q < nd->nintr; q++)
      A[i].kind[nd->iidx[q]] = 3;
  }
  /* transpose of the gather: abar to the finest node's points */
  for (int i = 0; i < nn; i++)
  {
    const onode *nd = &s->nodes[i];
    int X = gdx(nd), XY = gdxy(nd);
    for (int c = 0; c < nd->ncell; c++)
    {
      int b = nd->cbox[c];
      if (nd->nchild > 0 && nd->box[b] >= 0)
        continue;
      for (int q = nd->cell_start[b]; q < nd->cell_start[b + 1]; q++)
      {
        int p = nd->cell_list[q];
        double w[8], dw[3][8];
        int g = cic_grad(nd, s->x[p], s->y[p], s->z[p], w, dw);
        for (int kk = 0; kk < 2; kk++)
          for (int jj = 0; jj < 2; jj++)
            for (int ii = 0; ii < 2; ii++)
              for (int a = 0; a < 3; a++)
                A[i].ab[a][g + ii + jj * X + kk * XY] += w[ii + 2 * jj + 4 * kk] * abar[a * N + p];
      }
    }
  }
  /* nodes are stored parents first: the reverse order has every child done
   * before its parent */
  for (int i = nn - 1; i >= 0 && !rc; i--)
  {
    onode *nd = &s->nodes[i];
    oadj *J = &A[i];
    double sc = (double)(1 << nd->lv), o2H = 0.5 * sc, oneH2 = sc * sc;
    int X = gdx(nd), XY = gdxy(nd);
    for (int q = 0; q < nd->nintr; q++)
    {
      int g = nd->iidx[q];
      J->pb[g - 1] += o2H * J->ab[0][g];
      J->pb[g + 1] -= o2H * J->ab[0][g];
      J->pb[g - X] += o2H * J->ab[1][g];
      J->pb[g + X] -= o2H * J->ab[1][g];
      J->pb[g - XY] += o2H * J->ab[2][g];
      J->pb[g + XY] -= o2H * J->ab[2][g];
    }
    if (i == 0)
      break;
    if (node_adj_solve(nd, J->kind, J->psi, J->pb))
      rc = 1;
    const int off[6] = {1, -1, X, -X, XY, -XY};
    for (int q = 0; q < nd->nintr; q++)
    {
      int g = nd->iidx[q];
      for (int e = 0; e < 6; e++)
        if (J->kind[g + off[e]] == 2)
          J->pb[g + off[e]] -= oneH2 * J->psi[g];
    }
    const onode *pt = &s->nodes[nd->parent];
    oadj *Pj = &A[nd->parent];
    for (int q = 0; q < nd->nbdry; q++)
    {
      int g = nd->bidx[q];
      const int *c = &nd->bxyz[3 * q];
      for (int a = 0; a < 3; a++)
        parent_avg_t(pt, Pj->ab[a], c, J->ab[a][g]);
      parent_avg_t(pt, Pj->pb, c, J->pb[g]);
    }
  }
  /* the head solve on the packed cube with zero faces */
  if (!rc)
  {
    onode *h = &s->nodes[0];
    int n = 1 << s->lmin, n1 = n + 1;
    size_t p3 = (size_t)n1 * n1 * n1;
    double *rhs = (double *)calloc(p3, sizeof(double));
    double *psi = (double *)calloc(p3, sizeof(double));
    for (int q = 0; q < h->nintr; q++)
    {
      const int *c = &h->ixyz[3 * q];
      rhs[c[0] + (size_t)n1 * (c[1] + (size_t)n1 * c[2])] = A[0].pb[h->iidx[q]];
    }
    if (adj_solve(s->lmin, psi, rhs, 1e-12) < 0)
      rc = 1;
    for (int q = 0; q < h->nintr; q++)
    {
      const int *c = &h->ixyz[3 * q];
      A[0].psi[h->iidx[q]] = psi[c[0] + (size_t)n1 * (c[1] + (size_t)n1 * c[2])];
    }
    free(rhs);
    free(psi);
  }
  /* particles: deposit term at every node holding p, gather term at the finest */
  for (int i = 0; i < nn && !rc; i++)
  {
    const onode *nd = &s->nodes[i];
    const oadj *J = &A[i];
    double H = 1.0 / (1 << nd->lv);
    double coeff = 4.0 * s->G * OQ_PI / (H * H * H);
    int X = gdx(nd), XY = gdxy(nd);
    for (int c = 0; c < nd->ncell; c++)
    {
      int b = nd->cbox[c];
      int finest = !(nd->nchild > 0 && nd->box[b] >= 0);
      for (int q = nd->cell_start[b]; q < nd->cell_start[b + 1]; q++)
      {
        int p = nd->cell_list[q];
        double w[8], dw[3][8];
        int g = cic_grad(nd, s->x[p], s->y[p], s->z[p], w, dw);
        double ab0 = abar[p], ab1 = abar[N + p], ab2 = abar[2 * N + p];
        for (int kk = 0; kk < 2; kk++)
          for (int jj = 0; jj < 2; jj++)
            for (int ii = 0; ii < 2; ii++)
            {
              int k = ii + 2 * jj + 4 * kk, gn = g + ii + jj * X + kk * XY;
              double val = coeff * s->m[p] * J->psi[gn];
              if (finest)
                val += ab0 * nd->ax[gn] + ab1 * nd->ay[gn] + ab2 * nd->az[gn];
              for (int a = 0; a < 3; a++)
                xbar[a * N + p] += dw[a][k] * val;
            }
      }
    }
  }
  for (int i = 0; i < nn; i++)
  {
    for (int a = 0; a < 3; a++)
      free(A[i].ab[a]);
    free(A[i].pb);
    free(A[i].psi);
    free(A[i].kind);
  }
  free(A);
  return rc;
}

/* xbar += J_A(x)^T abar at the current positions.  abar, xbar: 3N. */
int oq_force_vjp(oq_sim *s, const double *abar, double *xbar)
{
  if (s->lmin < s->lmax)
  {
    oq_build_tree(s);
    oq_reset(s);
    oq_grid_density(s);
    if (oq_potential(s))
      return 1;
    oq_grid_acceleration(s);
    return amr_force_vjp(s, abar, xbar);
  }
  oq_reset(s);
  oq_grid_density(s);
  if (oq_potential(s))
    return 1;
  oq_grid_acceleration(s);
  onode *h = &s->nodes[0];
  int n = 1 << s->lmin, n1 = n + 1;
  int gX = gdx(h), gXY = gdxy(h);
  size_t p3 = (size_t)n1 * n1 * n1;
  double *gb = (double *)calloc(3 * p3, sizeof(double));
  double *rhs = (double *)calloc(p3, sizeof(double));
  double *psi = (double *)calloc(p3, sizeof(double));
  const double H = 1.0 / n;
  /* Gbar = sum_i W(x_i, g) abar_i on the packed solve domain */
  for (int i = 0; i < s->n; i++)
  {
    double px = s->x[i] * n, py = s->y[i] * n, pz = s->z[i] * n;
    int fx = (int)px, fy = (int)py, fz = (int)pz;
    double tx = px - fx, ty = py - fy, tz = pz - fz;
    double wx[2] = {1.0 - tx, tx}, wy[2] = {1.0 - ty, ty}, wz[2] = {1.0 - tz, tz};
    for (int kk = 0; kk < 2; kk++)
      for (int jj = 0; jj < 2; jj++)
        for (int ii = 0; ii < 2; ii++)
        {
          size_t g = (fx + ii) + (size_t)n1 * ((fy + jj) + (size_t)n1 * (fz + kk));
          double w = wx[ii] * wy[jj] * wz[kk];
          for (int a = 0; a < 3; a++)
            gb[a * p3 + g] += w * abar[(size_t)a * s->n + i];
        }
  }
  /* transpose of the interior central difference */
  for (int k = 1; k < n; k++)
    for (int j = 1; j < n; j++)
      for (int i = 1; i < n; i++)
      {
        size_t g = i + (size_t)n1 * (j + (size_t)n1 * k);
        double v = 0.0;
        if (i + 1 < n)
          v += gb[g + 1];
        if (i - 1 > 0)
          v -= gb[g - 1];
        if (j + 1 < n)
          v += gb[p3 + g + n1];
        if (j - 1 > 0)
          v -= gb[p3 + g - n1];
        if (k + 1 < n)
          v += gb[2 * p3 + g + (size_t)n1 * n1];
        if (k - 1 > 0)
          v -= gb[2 * p3 + g - (size_t)n1 * n1];
        rhs[g] = v / (2.0 * H);
      }
  int rc = adj_solve(s->lmin, psi, rhs, 1e-12) < 0;
  /* gather of weight gradients: grad W (c m psi + abar . G) */
  const double coeff = 4.0 * s->G * OQ_PI / (H * H * H);
  for (int i = 0; i < s->n; i++)
  {
    double px = s->x[i] * n, py = s->y[i] * n, pz = s->z[i] * n;
    int fx = (int)px, fy = (int)py, fz = (int)pz;
    double tx = px - fx, ty = py - fy, tz = pz - fz;
    double wx[2] = {1.0 - tx, tx}, wy[2] = {1.0 - ty, ty}, wz[2] = {1.0 - tz, tz};
    double dw[2] = {-(double)n, (double)n};
    double ab[3] = {abar[i], abar[(size_t)s->n + i], abar[2 * (size_t)s->n + i]};
    for (int kk = 0; kk < 2; kk++)
      for (int jj = 0; jj < 2; jj++)
        for (int ii = 0; ii < 2; ii++)
        {
          size_t g = (fx + ii) + (size_t)n1 * ((fy + jj) + (size_t)n1 * (fz + kk));
          int gb_ = (fx + ii + 1) + gX * (fy + jj + 1) + gXY * (fz + kk + 1); /* reference box grid index */
          double val = coeff * s->m[i] * psi[g] + ab[0] * h->ax[gb_] + ab[1] * h->ay[gb_] + ab[2] * h->az[gb_];
          xbar[i] += dw[ii] * wy[jj] * wz[kk] * val;
          xbar[(size_t)s->n + i] += wx[ii] * dw[jj] * wz[kk] * val;
          xbar[2 * (size_t)s->n + i] += wx[ii] * wy[jj] * dw[kk] * val;
        }
  }
  free(gb);
  free(rhs);
  free(psi);
  return rc;
}

/* Reverse mode of nsteps fixed-dt KDK steps from the current state (a force
 * pass done).  Stores every state; on return the sim holds the final state. */
int oq_adjoint(oq_sim *s, int nsteps, double dt, const double *xbN, const double *vbN, double *xb0, double *vb0)
{
  size_t N = (size_t)s->n, n9 = 9 * N, n3 = 3 * N;
  double *st = (double *)malloc(sizeof(double) * n9 * (size_t)(nsteps + 1));
  unsigned char *refl = (unsigned char *)calloc(N * (size_t)(nsteps + 1), 1);
  oq_get_particles(s, st);
  for (int j = 1; j <= nsteps; j++)
  {
    /* reflection flags of step j from the drift of step j */
    double *prev = st + n9 * (size_t)(j - 1);
    for (size_t i = 0; i < N; i++)
      for (int a = 0; a < 3; a++)
      {
        double u = prev[(3 + a) * N + i] + prev[(6 + a) * N + i] * dt * 0.5;
        double y = prev[a * N + i] + u * dt;
        if (y < 0.0 || y >= 1.0)
          refl[N * j + i] |= (unsigned char)(1 << a);
      }
    if (oq_step(s, dt))
      return 1;
    oq_get_particles(s, st + n9 * (size_t)j);
  }
  double *xb = (double *)malloc(sizeof(double) * n3), *vb = (double *)malloc(sizeof(double) * n3);
  double *abA = (double *)calloc(n3, sizeof(double)), *ab = (double *)malloc(sizeof(double) * n3);
  memcpy(xb, xbN, sizeof(double) * n3);
  memcpy(vb, vbN, sizeof(double) * n3);
  const double h = 0.5 * dt;
  for (int j = nsteps; j >= 1; j--)
  {
    oq_set_particles(s, st + n9 * (size_t)j);
    for (size_t k = 0; k < n3; k++)
      ab[k] = h * vb[k] + abA[k];
    if (oq_force_vjp(s, ab, xb))
      return 1;
    for (size_t i = 0; i < N; i++)
      for (int a = 0; a < 3; a++)
      {
        size_t k = a * N + i;
        double sg = ((refl[N * j + i] >> a) & 1) ? -1.0 : 1.0;
        double yb = sg * xb[k];
        double ub = sg * vb[k] + dt * yb;
        xb[k] = yb;
        vb[k] = ub;
        abA[k] = h * ub;
      }
  }
  oq_set_particles(s, st);
  if (oq_force_vjp(s, abA, xb))
    return 1;
  memcpy(xb0, xb, sizeof(double) * n3);
  memcpy(vb0, vb, sizeof(double) * n3);
  oq_set_particles(s, st + n9 * (size_t)nsteps);
  free(st);
  free(refl);
  free(xb);
  free(vb);
  free(abA);
  free(ab);
  return 0;
}
