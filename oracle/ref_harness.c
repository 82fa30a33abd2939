/*
 * Driver for the compiled reference (QUANTIZER, built from its own sources by
 * oracle/Makefile into oracle/_ref/).  This file is test infrastructure: it
 * replaces the reference's main() (source/main.c:169-590) with a fixed-dt
 * loop over the same step sequence (main.c:369-535) so that the reference can
 * be run on seeded synthetic inputs and its fields dumped as golden vectors,
 * or timed as the CPU baseline.
 *
 * Usage:
 *   ref_harness <particles.bin> <outdir> <lmin> <lmax> <ref_crit> <nsteps> <dt>
 *               [--time] [--threads T] [--noenergy] [--adaptive FR] [--cfl-steps]
 *
 * --cfl-steps (with --time) takes each timed step's dt from the reference's
 * timestep() (timestep.c:589-699, CFL 0.5, capped at _MAX_dt_ = dt), as
 * main.c:369-380 does, instead of holding dt fixed.  A fixed dt chosen at t = 0
 * can let particles of a collapsing clump cross more than a node's box margin
 * in one step; particle_updating.c:740-805 then files them in cells outside the
 * node's box and grid_density.c:160-164 reads past ptr_box_idx.
 *
 * --adaptive FR runs main.c:341-573's loop instead of the fixed-dt one: the
 * time step from timestep() at every step (dt is then _MAX_dt_), observables()
 * and output_snapshots() before the loop, every FR steps and after the last,
 * garbage_collector() every Garbage_Collector_iter steps, and
 * output_main_parameters() at the end; the snapshots and Parameters.dat go to
 * ../output relative to the working directory and dts.txt (one time step per
 * line) to outdir.
 *
 * particles.bin: int32 N, then x,y,z,vx,vy,vz,m as N float64 each, already in
 * code units (box [0,1)^3, G = 1).
 *
 * Outputs (raw little-endian, see tests/golden/make_golden.py for readers):
 *   s0_head.bin   head rho, pot, ax, ay, az on the (2^lmin+1)^3 solve domain
 *                 after the first force computation (main.c:296-338)
 *   s0_acc.bin    particle ax, ay, az after the first force computation
 *   s0_nodes.bin  branch-node geometry and fields after the first force pass
 *   sK_ptcl.bin   x,y,z,vx,vy,vz,ax,ay,az after nsteps fixed-dt steps
 * and with --snapshot dist velocity time cmx cmy cmz cmvx cmvy cmvz the
 * reference's own snapshot file of the final state (output_snapshots.c) under
 * ../output/ relative to the working directory
 *   sK_head.bin   head fields after the last step
 *   energies.txt  E_kin E_pot E_tot px py pz at step 0 and step K
 */
#include "main.h"

#include <stdint.h>

/* The resident drop-in keeps the particles on the device and exports this to
 * copy them into GL_ptcl_*; the reference's own objects have no such function. */
#pragma weak qz_dropin_sync_host
int qz_dropin_sync_host(void);

static void sync_host(void)
{
  if (qz_dropin_sync_host)
    qz_dropin_sync_host();
}

static double now_s(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void set_parameters(int lmin_, int lmax_, int ref_crit, int N)
{
  /* global_variables() (global_variables.c:467-516) has already run and set
   * the reference defaults; here the level/particle-count dependent globals
   * are recomputed the way init_global_user_params / init_global_ref_crit /
   * init_global_border_sim_box do it, for the requested configuration. */
  _G_ = 1.0;
  _conversion_dist_ = 1.0;
  _conversion_time_ = 1.0;
  _conversion_velocity_ = 1.0;
  lmin = lmin_;
  lmax = lmax_;
  no_lmin_cell = 1 << lmin;
  no_lmin_cell_pow2 = no_lmin_cell * no_lmin_cell;
  no_lmin_cell_pow3 = no_lmin_cell_pow2 * no_lmin_cell;
  no_grid = no_lmin_cell + 1;
  no_grid_pow2 = no_grid * no_grid;
  no_grid_pow3 = no_grid_pow2 * no_grid;
  GL_no_ptcl_initial = N;
  GL_no_ptcl_final = N;
  ref_criterion_ptcl = ref_crit;
  ref_criterion_mass = 1.0e100;
  n_exp = 1;
  min_box_extra_size_per_side_real =
      min_box_extra_size_per_side_user > (n_exp - 1) ? 2 * min_box_extra_size_per_side_user : 2 * n_exp - 2;
  bder_os_sim = 1 > n_exp ? 1 : n_exp;
  box_side_lmin = no_lmin_cell + 2 * bder_os_sim;
  box_side_lmin_pow2 = box_side_lmin * box_side_lmin;
  box_side_lmin_pow3 = box_side_lmin_pow2 * box_side_lmin;
  bdry_cond_type = 1;

  GL_ptcl_mass = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_x = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_y = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_z = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_vx = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_vy = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_vz = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_ax = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_ay = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl_az = (vtype *)calloc(N, sizeof(vtype));
  GL_ptcl[0] = GL_ptcl_mass;
  GL_ptcl[1] = GL_ptcl_x;
  GL_ptcl[2] = GL_ptcl_y;
  GL_ptcl[3] = GL_ptcl_z;
  GL_ptcl[4] = GL_ptcl_vx;
  GL_ptcl[5] = GL_ptcl_vy;
  GL_ptcl[6] = GL_ptcl_vz;
  GL_ptcl[7] = GL_ptcl_ax;
  GL_ptcl[8] = GL_ptcl_ay;
  GL_ptcl[9] = GL_ptcl_az;
}

static int read_particles(const char *path)
{
  FILE *f = fopen(path, "rb");
  if (!f)
    return -1;
  int32_t n;
  if (fread(&n, 4, 1, f) != 1)
    return -1;
  double *buf = (double *)malloc(sizeof(double) * (size_t)n);
  vtype *dst[7] = {GL_ptcl_x, GL_ptcl_y, GL_ptcl_z, GL_ptcl_vx, GL_ptcl_vy, GL_ptcl_vz, GL_ptcl_mass};
  for (int a = 0; a < 7; a++)
  {
    if (fread(buf, sizeof(double), (size_t)n, f) != (size_t)n)
      return -1;
    for (int i = 0; i < n; i++)
      dst[a][i] = buf[i];
  }
  free(buf);
  fclose(f);
  /* input_code_units bookkeeping (input.c:252-277): total mass and the
   * centre-of-mass velocity used by observables(); inputs are already centred */
  GL_total_mass_initial = 0.0;
  for (int i = 0; i < n; i++)
    GL_total_mass_initial += GL_ptcl_mass[i];
  GL_cm_x = GL_cm_y = GL_cm_z = 0.0;
  GL_cm_vx = GL_cm_vy = GL_cm_vz = 0.0;
  return n;
}

static FILE *open_out(const char *dir, const char *name)
{
  char path[2048];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "wb");
  if (!f)
  {
    fprintf(stderr, "cannot open %s\n", path);
    exit(2);
  }
  return f;
}

static void dump_head(const char *dir, const char *name)
{
  struct node *h = GL_ptr_tree;
  int gs = box_side_lmin + 1; /* grid side of the head box (initialization.c:248) */
  int n1 = no_lmin_cell + 1;  /* solve domain side (potential_head_node.c:46) */
  FILE *f = open_out(dir, name);
  vtype *arrs[5] = {h->ptr_d, h->ptr_pot, h->ptr_ax, h->ptr_ay, h->ptr_az};
  double *buf = (double *)malloc(sizeof(double) * (size_t)n1 * n1 * n1);
  for (int a = 0; a < 5; a++)
  {
    for (int k = 0; k < n1; k++)
      for (int j = 0; j < n1; j++)
        for (int i = 0; i < n1; i++)
          buf[i + n1 * (j + n1 * k)] =
              arrs[a][(i + bder_os_sim) + gs * ((j + bder_os_sim) + gs * (k + bder_os_sim))];
    fwrite(buf, sizeof(double), (size_t)n1 * n1 * n1, f);
  }
  free(buf);
  fclose(f);
}

static void dump_particles(const char *dir, const char *name, int with_xv)
{
  sync_host();
  FILE *f = open_out(dir, name);
  int n = GL_no_ptcl_final;
  vtype *arrs[9] = {GL_ptcl_x, GL_ptcl_y, GL_ptcl_z, GL_ptcl_vx, GL_ptcl_vy, GL_ptcl_vz,
                    GL_ptcl_ax, GL_ptcl_ay, GL_ptcl_az};
  int a0 = with_xv ? 0 : 6;
  for (int a = a0; a < 9; a++)
  {
    for (int i = 0; i < n; i++)
    {
      double v = (double)arrs[a][i];
      fwrite(&v, sizeof(double), 1, f);
    }
  }
  fclose(f);
}

static void wr_i(FILE *f, int v)
{
  int32_t x = v;
  fwrite(&x, 4, 1, f);
}

static void wr_ia(FILE *f, const int *a, int n)
{
  for (int i = 0; i < n; i++)
    wr_i(f, a[i]);
}

static void wr_da(FILE *f, const vtype *a, int n)
{
  for (int i = 0; i < n; i++)
  {
    double v = (double)a[i];
    fwrite(&v, sizeof(double), 1, f);
  }
}

/* Node record: 24 int32 header, then int32 arrays and float64 fields.  The
 * parent is named by its index in GL_tentacles[lv-1]. */
static void dump_nodes(const char *dir, const char *name)
{
  FILE *f = open_out(dir, name);
  int total = 0;
  for (int lv = 1; lv <= GL_tentacles_level_max; lv++)
    total += GL_tentacles_size[lv];
  wr_i(f, total);
  for (int lv = 1; lv <= GL_tentacles_level_max; lv++)
  {
    for (int n = 0; n < GL_tentacles_size[lv]; n++)
    {
      struct node *p = GL_tentacles[lv][n];
      int parent = -1;
      for (int q = 0; q < GL_tentacles_size[lv - 1]; q++)
        if (GL_tentacles[lv - 1][q] == p->ptr_pt)
          parent = q;
      int hdr[24] = {p->lv, lv, n, parent,
                     p->box_ts_x, p->box_ts_y, p->box_ts_z,
                     p->box_real_dim_x, p->box_real_dim_y, p->box_real_dim_z,
                     p->box_min_x, p->box_min_y, p->box_min_z,
                     p->box_max_x, p->box_max_y, p->box_max_z,
                     p->cell_size, p->grid_intr_size, p->grid_bdry_size, p->grid_sim_bdry_size,
                     p->no_ptcl_full_node, p->no_ptcl_outs_ref_zones, p->chn_size,
                     (p->box_real_dim_x + 1) * (p->box_real_dim_y + 1) * (p->box_real_dim_z + 1)};
      wr_ia(f, hdr, 24);
      wr_ia(f, p->ptr_cell_idx_x, p->cell_size);
      wr_ia(f, p->ptr_cell_idx_y, p->cell_size);
      wr_ia(f, p->ptr_cell_idx_z, p->cell_size);
      wr_ia(f, p->ptr_box_idx, p->cell_size);
      wr_ia(f, p->ptr_intr_grid_cell_idx_x, p->grid_intr_size);
      wr_ia(f, p->ptr_intr_grid_cell_idx_y, p->grid_intr_size);
      wr_ia(f, p->ptr_intr_grid_cell_idx_z, p->grid_intr_size);
      wr_ia(f, p->ptr_intr_box_grid_idx, p->grid_intr_size);
      wr_ia(f, p->ptr_bdry_grid_cell_idx_x, p->grid_bdry_size);
      wr_ia(f, p->ptr_bdry_grid_cell_idx_y, p->grid_bdry_size);
      wr_ia(f, p->ptr_bdry_grid_cell_idx_z, p->grid_bdry_size);
      wr_ia(f, p->ptr_bdry_box_grid_idx, p->grid_bdry_size);
      wr_ia(f, p->ptr_sim_bdry_grid_cell_idx_x, p->grid_sim_bdry_size);
      wr_ia(f, p->ptr_sim_bdry_grid_cell_idx_y, p->grid_sim_bdry_size);
      wr_ia(f, p->ptr_sim_bdry_grid_cell_idx_z, p->grid_sim_bdry_size);
      wr_ia(f, p->ptr_sim_bdry_box_grid_idx, p->grid_sim_bdry_size);
      int gcap = hdr[23];
      wr_da(f, p->ptr_d, gcap);
      wr_da(f, p->ptr_pot, gcap);
      wr_da(f, p->ptr_ax, gcap);
      wr_da(f, p->ptr_ay, gcap);
      wr_da(f, p->ptr_az, gcap);
    }
  }
  fclose(f);
}

static int force_pass(void)
{
  /* main.c:296-338 / :430-465 */
  grid_density();
  if (potential() == _FAILURE_)
    return _FAILURE_;
  if (grid_acceleration() == _FAILURE_)
    return _FAILURE_;
  if (particle_acceleration() == _FAILURE_)
    return _FAILURE_;
  return _SUCCESS_;
}

static int one_step(vtype dt)
{
  /* main.c:391-478 with dt fixed instead of timestep() */
  if (particle_updating_A(dt) == _FAILURE_)
    return _FAILURE_;
  if (reset() == _FAILURE_)
    return _FAILURE_;
  if (lmin < lmax)
  {
    if (tree_adaptation() == _FAILURE_)
      return _FAILURE_;
  }
  if (force_pass() == _FAILURE_)
    return _FAILURE_;
  particle_updating_B(dt);
  return _SUCCESS_;
}

/* main.c:341-573 from the initial force pass on, for nsteps steps */
static int adaptive_run(const char *outdir, int nsteps, int fr, int energy)
{
  vtype actualtime = 0, dt = 0;
  int Number_timesteps = 0, Number_outputs = 0;
  compute_observables_FLAG = energy ? true : false;
  FILE *fd = open_out(outdir, "dts.txt");
  if (compute_observables_FLAG && observables() == _FAILURE_)
    return _FAILURE_;
  if (output_snapshots(actualtime, Number_outputs) == _FAILURE_)
    return _FAILURE_;
  while (Number_timesteps < nsteps)
  {
    if (timestep(&dt) == _FAILURE_)
      return _FAILURE_;
    fprintf(fd, "%.17e\n", (double)dt);
    actualtime += dt;
    if (one_step(dt) == _FAILURE_)
      return _FAILURE_;
    if (Number_timesteps % fr == 0 || Number_timesteps == nsteps - 1)
    {
      ++Number_outputs;
      if (compute_observables_FLAG && observables() == _FAILURE_)
        return _FAILURE_;
      if (output_snapshots(actualtime, Number_outputs) == _FAILURE_)
        return _FAILURE_;
    }
    if ((Number_timesteps + 1) % Garbage_Collector_iter == 0 && lmin < lmax && garbage_collector() == _FAILURE_)
      return _FAILURE_;
    Number_timesteps++;
  }
  fclose(fd);
  GL_total_mass_final = GL_total_mass_initial;
  output_main_parameters(actualtime, Number_timesteps, Number_outputs);
  return _SUCCESS_;
}

int main(int argc, char **argv)
{
  if (argc < 8)
  {
    fprintf(stderr, "usage: %s particles.bin outdir lmin lmax ref_crit nsteps dt [--time] [--threads T] [--noenergy]\n", argv[0]);
    return 2;
  }
  const char *inpath = argv[1];
  const char *outdir = argv[2];
  int lmin_ = atoi(argv[3]);
  int lmax_ = atoi(argv[4]);
  int ref_crit = atoi(argv[5]);
  int nsteps = atoi(argv[6]);
  double dt = atof(argv[7]);
  int do_time = 0, threads = 1, energy = 1, snap = 0, adaptive_fr = 0, cfl_steps = 0;
  double snap_u[9] = {1, 1, 1, 0, 0, 0, 0, 0, 0}; /* dist, velocity, time, cm x y z, cm v x y z */
  for (int a = 8; a < argc; a++)
  {
    if (!strcmp(argv[a], "--time"))
      do_time = 1;
    else if (!strcmp(argv[a], "--threads") && a + 1 < argc)
      threads = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--noenergy"))
      energy = 0;
    else if (!strcmp(argv[a], "--cfl-steps"))
      cfl_steps = 1;
    else if (!strcmp(argv[a], "--adaptive") && a + 1 < argc)
      adaptive_fr = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--snapshot") && a + 9 < argc)
    {
      snap = 1;
      for (int q = 0; q < 9; q++)
        snap_u[q] = atof(argv[++a]);
    }
  }

  FILE *fin = fopen(inpath, "rb");
  if (!fin)
  {
    fprintf(stderr, "cannot open %s\n", inpath);
    return 2;
  }
  int32_t n;
  if (fread(&n, 4, 1, fin) != 1)
    return 2;
  fclose(fin);

  /* the reference prints progress on stdout; keep it off the harness output */
  fflush(stdout);
  FILE *devnull = freopen("/dev/null", "w", stdout);
  (void)devnull;

  global_variables();
  set_parameters(lmin_, lmax_, ref_crit, n);
  NUMBER_OF_THEADS = threads;
  omp_set_num_threads(threads);
  if (read_particles(inpath) != n)
  {
    fprintf(stderr, "bad particle file\n");
    return 2;
  }
  if (initialization() == _FAILURE_)
  {
    fprintf(stderr, "initialization failed\n");
    return 1;
  }
  if (tree_construction() == _FAILURE_)
  {
    fprintf(stderr, "tree_construction failed\n");
    return 1;
  }
  double t0 = now_s();
  if (force_pass() == _FAILURE_)
  {
    fprintf(stderr, "force pass failed\n");
    return 1;
  }
  double t_first = now_s() - t0;

  if (!do_time)
  {
    dump_head(outdir, "s0_head.bin");
    dump_particles(outdir, "s0_acc.bin", 0);
    dump_nodes(outdir, "s0_nodes.bin");
  }
  FILE *fe = NULL;
  if (!do_time)
  {
    fe = open_out(outdir, "energies.txt");
    if (energy)
    {
      observables();
      fprintf(fe, "%.17e %.17e %.17e %.17e %.17e %.17e\n", (double)GL_energies[0], (double)GL_energies[1],
              (double)GL_energies[2], (double)GL_momentum[0], (double)GL_momentum[1], (double)GL_momentum[2]);
    }
  }

  if (adaptive_fr > 0)
  {
    _MAX_dt_ = (vtype)dt;
    time_step_method = 1;
    if (adaptive_run(outdir, nsteps, adaptive_fr, energy) == _FAILURE_)
    {
      fprintf(stderr, "adaptive run failed\n");
      return 1;
    }
    dump_particles(outdir, "sK_ptcl.bin", 1);
    fprintf(stderr, "REF_DONE N=%d steps=%d levels_used=%d\n", n, nsteps, GL_tentacles_level_max);
    return 0;
  }

  double t_steps = 0.0, step_dt = dt;
  int timed = 0;
  if (cfl_steps)
  {
    _MAX_dt_ = (vtype)dt;
    time_step_method = 1;
  }
  for (int s = 0; s < nsteps; s++)
  {
    double ts = now_s();
    if (cfl_steps)
    {
      vtype v;
      if (timestep(&v) == _FAILURE_)
      {
        fprintf(stderr, "timestep at step %d failed\n", s);
        return 1;
      }
      step_dt = (double)v;
    }
    if (one_step((vtype)step_dt) == _FAILURE_)
    {
      fprintf(stderr, "step %d failed\n", s);
      return 1;
    }
    double te = now_s();
    if (s > 0 || nsteps == 1)
    {
      t_steps += te - ts;
      timed++;
    }
  }
  if (do_time)
  {
    fprintf(stderr, "REF_TIMING first_force_s=%.6f steps=%d timed=%d s_per_step=%.6f threads=%d N=%d lmin=%d lmax=%d levels_used=%d last_dt=%.17e\n",
            t_first, nsteps, timed, timed ? t_steps / timed : 0.0, threads, n, lmin, lmax, GL_tentacles_level_max,
            step_dt);
    return 0;
  }
  dump_particles(outdir, "sK_ptcl.bin", 1);
  dump_head(outdir, "sK_head.bin");
  dump_nodes(outdir, "sK_nodes.bin");
  if (energy)
  {
    observables();
    fprintf(fe, "%.17e %.17e %.17e %.17e %.17e %.17e\n", (double)GL_energies[0], (double)GL_energies[1],
            (double)GL_energies[2], (double)GL_momentum[0], (double)GL_momentum[1], (double)GL_momentum[2]);
  }
  if (fe)
    fclose(fe);
  if (snap)
  {
    /* output_snapshots.c:141-226 with the given unit conversions, into
     * ../output/<simulation>/Data/Data_0.bin relative to the working directory */
    _conversion_dist_ = snap_u[0];
    _conversion_velocity_ = snap_u[1];
    _conversion_time_ = snap_u[2];
    GL_cm_x = snap_u[3];
    GL_cm_y = snap_u[4];
    GL_cm_z = snap_u[5];
    GL_cm_vx = snap_u[6];
    GL_cm_vy = snap_u[7];
    GL_cm_vz = snap_u[8];
    compute_observables_FLAG = energy ? true : false;
    if (energy)
      observables();
    if (output_snapshots((vtype)(nsteps * dt), 0) == _FAILURE_)
      return 1;
  }
  fprintf(stderr, "REF_DONE N=%d steps=%d levels_used=%d\n", n, nsteps, GL_tentacles_level_max);
  return 0;
}
