/*
 * Drop-in replacement of QUANTIZER's per-step force cycle.
 *
 * This file is compiled inside a QUANTIZER tree, against that tree's own
 * headers (-I <quantizer>/headers): it takes `vtype`, `struct node` and the
 * host globals (GL_ptcl_*, lmin, lmax, _G_, solver parameters) from there and
 * exports the reference's entry points with their exact names and signatures.
 * The work runs in libqzhip.so through its C ABI (include/qz.h).
 *
 * Two builds:
 *  - default ("force cycle"): reset, grid_density, potential,
 *    potential_head_node, potential_branches, grid_acceleration,
 *    particle_acceleration and ptcl_idx_to_box_idx.  The host keeps
 *    particle_updating_A/B and tree_adaptation, which maintain its cell lists;
 *    positions go to the device at every grid_density() and accelerations come
 *    back at particle_acceleration().
 *  - QZ_DROPIN_RESIDENT: additionally particle_updating_A, particle_updating_B,
 *    tree_adaptation, timestep, observables, output_snapshots and
 *    garbage_collector.  Particles stay on the device between steps: every
 *    per-step function that reads them (the time step, the diagnostics, the
 *    snapshot writer) runs on the device, and the host arrays GL_ptcl_* are
 *    refreshed only by qz_dropin_sync_host().  The host tree is built once by
 *    tree_construction() and not adapted (garbage_collector() has nothing to
 *    trim in it).
 *
 * In both builds the device rebuilds the refinement tree from the particles at
 * every force pass with the rules of tree_construction.c (marks, zones, child
 * boxes), so the node geometry is the one tree_construction() would produce.
 * Errors follow the reference convention: _FAILURE_ with a printed reason.
 */
#include "common.h"
#include "ptcl_idx_to_box_idx.h"

#include <sys/stat.h>

#include "../../include/qz.h"

/* The globals are defined by the reference (global_variables.c); weak
 * references let the library load on its own (symbol checks) and report a
 * clear error when used outside a QUANTIZER program. */
#pragma weak GL_ptcl_x
#pragma weak GL_ptcl_y
#pragma weak GL_ptcl_z
#pragma weak GL_ptcl_vx
#pragma weak GL_ptcl_vy
#pragma weak GL_ptcl_vz
#pragma weak GL_ptcl_ax
#pragma weak GL_ptcl_ay
#pragma weak GL_ptcl_az
#pragma weak GL_ptcl_mass
#pragma weak GL_no_ptcl_final
#pragma weak lmin
#pragma weak lmax
#pragma weak _G_
#pragma weak ref_criterion_ptcl
#pragma weak n_exp
#pragma weak bdry_cond_type
#pragma weak _MAX_NUMBER_OF_ITERATIONS_IN_POISSON_EQUATION_
#pragma weak _ERROR_THRESHOLD_IN_THE_POISSON_EQUATION_
#ifdef QZ_DROPIN_RESIDENT
#pragma weak time_step_method
#pragma weak _CFL_
#pragma weak _MAX_dt_
#pragma weak potential_energy_type
#pragma weak compute_observables_FLAG
#pragma weak GL_energies
#pragma weak GL_momentum
#pragma weak _conversion_dist_
#pragma weak _conversion_time_
#pragma weak _conversion_velocity_
#pragma weak GL_cm_x
#pragma weak GL_cm_y
#pragma weak GL_cm_z
#pragma weak GL_cm_vx
#pragma weak GL_cm_vy
#pragma weak GL_cm_vz
#pragma weak Maxdt
#pragma weak _Myear_
#pragma weak GL_no_ptcl_initial
#pragma weak folder_created
#endif

static qz_ctx *g_ctx = NULL;
static int g_uploaded = 0; /* particles resident on the device (resident build) */
#if _VTYPE_ != 2
static double *g_stage = NULL; /* FP64 staging when vtype is not double */
#endif

static int qz_fail(const char *where)
{
  printf("\n\n Error running %s(): %s\n\n", where, qz_last_error());
  return _FAILURE_;
}

/* Context from the host configuration (global_variables.c:248-348). */
static int qz_ensure(const char *where)
{
  if (g_ctx)
    return _SUCCESS_;
  if (&GL_ptcl_x == NULL || &lmin == NULL)
  {
    printf("\n\n Error running %s(): QUANTIZER globals are not linked\n\n", where);
    return _FAILURE_;
  }
  if (bdry_cond_type != 1)
  {
    printf("\n\n Error running %s(): the device path implements reflexive boundaries only "
           "(bdry_cond_type = 1, global_variables.c:277)\n\n", where);
    return _FAILURE_;
  }
  qz_config cfg;
  qz_default_config(&cfg);
  cfg.lmin = lmin;
  cfg.lmax = lmax;
  cfg.n_particles = GL_no_ptcl_final;
  cfg.ref_criterion_ptcl = ref_criterion_ptcl;
  cfg.n_exp = n_exp;
  cfg.G = (double)_G_;
  cfg.tol = (double)_ERROR_THRESHOLD_IN_THE_POISSON_EQUATION_;
  cfg.tol_change = (double)_ERROR_THRESHOLD_IN_THE_POISSON_EQUATION_;
  cfg.max_iterations = _MAX_NUMBER_OF_ITERATIONS_IN_POISSON_EQUATION_;
  cfg.precision = QZ_PREC_FP64;
  const char *prec = getenv("QZ_PRECISION");
  if (prec && !strcmp(prec, "mixed"))
    cfg.precision = QZ_PREC_MIXED;
  g_ctx = qz_create(&cfg);
  if (!g_ctx)
    return qz_fail(where);
  g_uploaded = 0;
  return _SUCCESS_;
}

/* host vtype array -> FP64 view */
static const double *as_f64(const vtype *a, int which)
{
#if _VTYPE_ == 2
  (void)which;
  return (const double *)a;
#else
  int n = GL_no_ptcl_final;
  if (!g_stage)
    g_stage = (double *)malloc(sizeof(double) * 7 * (size_t)n);
  double *s = g_stage + (size_t)which * n;
  for (int i = 0; i < n; i++)
    s[i] = (double)a[i];
  return s;
#endif
}

static int qz_upload(const char *where)
{
  const double *x = as_f64(GL_ptcl_x, 0), *y = as_f64(GL_ptcl_y, 1), *z = as_f64(GL_ptcl_z, 2);
  const double *vx = as_f64(GL_ptcl_vx, 3), *vy = as_f64(GL_ptcl_vy, 4), *vz = as_f64(GL_ptcl_vz, 5);
  const double *m = as_f64(GL_ptcl_mass, 6);
  /* the first upload also sets the head's monopole start (initialization.c:258) */
  int rc = g_uploaded ? qz_load_particles(g_ctx, x, y, z, vx, vy, vz, m) : qz_set_particles(g_ctx, x, y, z, vx, vy, vz, m);
  if (rc)
    return qz_fail(where);
  g_uploaded = 1;
  return _SUCCESS_;
}

/* device -> host: the arrays selected by the mask (1 x, 2 v, 4 a) */
static int qz_download(int mask, const char *where)
{
  int n = GL_no_ptcl_final;
  double *buf = (double *)malloc(sizeof(double) * 9 * (size_t)n);
  if (!buf)
  {
    printf("\n\n Error running %s(): out of host memory\n\n", where);
    return _FAILURE_;
  }
  double *p[9];
  for (int a = 0; a < 9; a++)
    p[a] = (mask >> (a / 3)) & 1 ? buf + (size_t)a * n : NULL;
  if (qz_get_particles(g_ctx, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8]))
  {
    free(buf);
    return qz_fail(where);
  }
  vtype *dst[9] = {GL_ptcl_x, GL_ptcl_y, GL_ptcl_z, GL_ptcl_vx, GL_ptcl_vy, GL_ptcl_vz, GL_ptcl_ax, GL_ptcl_ay, GL_ptcl_az};
  for (int a = 0; a < 9; a++)
    if (p[a])
      for (int i = 0; i < n; i++)
        dst[a][i] = (vtype)p[a][i];
  free(buf);
  return _SUCCESS_;
}

/* reset.c:29-64: zero the particle accelerations and the head density (the
 * resident build keeps no host copy of the accelerations between steps) */
int reset(void)
{
  if (qz_ensure("reset") == _FAILURE_)
    return _FAILURE_;
#ifndef QZ_DROPIN_RESIDENT
  for (int i = 0; i < GL_no_ptcl_final; i++)
  {
    GL_ptcl_ax[i] = 0;
    GL_ptcl_ay[i] = 0;
    GL_ptcl_az[i] = 0;
  }
#endif
  if (g_uploaded && qz_reset(g_ctx))
    return qz_fail("reset");
  return _SUCCESS_;
}

/* grid_density.c:269-306, preceded by the device tree build */
void grid_density(void)
{
  if (qz_ensure("grid_density") == _FAILURE_)
    return;
#ifndef QZ_DROPIN_RESIDENT
  if (qz_upload("grid_density") == _FAILURE_)
    return;
#else
  if (!g_uploaded && qz_upload("grid_density") == _FAILURE_)
    return;
#endif
  if (qz_reset(g_ctx) || qz_tree_adaptation(g_ctx) || qz_grid_density(g_ctx))
    qz_fail("grid_density");
}

/* potential.c:29-51 */
int potential(void)
{
  if (qz_ensure("potential") == _FAILURE_ || qz_potential(g_ctx))
    return qz_fail("potential");
  return _SUCCESS_;
}

/* potential_head_node.c:99-179 */
int potential_head_node(void)
{
  if (qz_ensure("potential_head_node") == _FAILURE_ || qz_potential_head_node(g_ctx))
    return qz_fail("potential_head_node");
  return _SUCCESS_;
}

/* potential_branches.c:401-532 */
int potential_branches(void)
{
  if (qz_ensure("potential_branches") == _FAILURE_ || qz_potential_branches(g_ctx))
    return qz_fail("potential_branches");
  return _SUCCESS_;
}

/* grid_acceleration.c:694-738 */
int grid_acceleration(void)
{
  if (qz_ensure("grid_acceleration") == _FAILURE_ || qz_grid_acceleration(g_ctx))
    return qz_fail("grid_acceleration");
  return _SUCCESS_;
}

/* particle_acceleration.c:522-573; in the force-cycle build the accelerations
 * return to GL_ptcl_a* for the host's particle updates */
int particle_acceleration(void)
{
  if (qz_ensure("particle_acceleration") == _FAILURE_ || qz_particle_acceleration(g_ctx))
    return qz_fail("particle_acceleration");
#ifdef QZ_DROPIN_RESIDENT
  return _SUCCESS_;
#else
  return qz_download(4, "particle_acceleration");
#endif
}

/* ptcl_idx_to_box_idx.c:30-139: box index of particle ptcl_idx in node ptr_node.
 * Host code (the tree routines call it); integer arithmetic on the exact
 * product x * 2^lv, periodic wraps as the reference applies them. */
int ptcl_idx_to_box_idx(struct node *ptr_node, int ptcl_idx)
{
  const int lv = ptr_node->lv;
  const vtype scale = (vtype)(1 << lv);
  const vtype p[3] = {GL_ptcl_x[ptcl_idx] * scale, GL_ptcl_y[ptcl_idx] * scale, GL_ptcl_z[ptcl_idx] * scale};
  const vtype raw[3] = {GL_ptcl_x[ptcl_idx], GL_ptcl_y[ptcl_idx], GL_ptcl_z[ptcl_idx]};
  const int ts[3] = {ptr_node->box_ts_x, ptr_node->box_ts_y, ptr_node->box_ts_z};
  const int mx[3] = {ptr_node->box_max_x, ptr_node->box_max_y, ptr_node->box_max_z};
  const bool cross[3] = {ptr_node->pbc_crosses_sim_box_bdry_x, ptr_node->pbc_crosses_sim_box_bdry_y,
                         ptr_node->pbc_crosses_sim_box_bdry_z};
  const bool whole[3] = {ptr_node->pbc_crosses_whole_sim_box_x, ptr_node->pbc_crosses_whole_sim_box_y,
                         ptr_node->pbc_crosses_whole_sim_box_z};
  int b[3];
  for (int a = 0; a < 3; a++)
  {
    const int f = p[a] < 0 ? -1 : (int)p[a];
    b[a] = f - ts[a];
    if (ptr_node->pbc_crosses_sim_box_bdry)
    {
      if (!whole[a])
      {
        if (cross[a] && f > mx[a] + 1)
          b[a] -= 1 << lv;
      }
      else if (raw[a] < 0.)
        b[a] += 1 << lv;
      else if (raw[a] >= 1.)
        b[a] -= 1 << lv;
    }
  }
  return b[0] + b[1] * ptr_node->box_real_dim_x + b[2] * ptr_node->box_real_dim_x * ptr_node->box_real_dim_y;
}

#ifdef QZ_DROPIN_RESIDENT
/* particle_updating.c:2584-2675: kick-drift with reflexive boundaries */
int particle_updating_A(vtype dt)
{
  if (qz_ensure("particle_updating_A") == _FAILURE_)
    return _FAILURE_;
  if (!g_uploaded && qz_upload("particle_updating_A") == _FAILURE_)
    return _FAILURE_;
  if (qz_particle_updating_A(g_ctx, (double)dt))
    return qz_fail("particle_updating_A");
  return _SUCCESS_;
}

/* particle_updating.c:2787-2836: closing half kick */
void particle_updating_B(vtype dt)
{
  if (qz_ensure("particle_updating_B") == _FAILURE_)
    return;
  if (qz_particle_updating_B(g_ctx, (double)dt))
    qz_fail("particle_updating_B");
}

/* tree_adaptation.c:8250: the device rebuilds its tree at every grid_density() */
int tree_adaptation(void) { return qz_ensure("tree_adaptation"); }

/* garbage_collector.c:377: trims the capacities of the host tree, which this
 * build never changes after tree_construction(); the device pools are grow-only */
int garbage_collector(void) { return qz_ensure("garbage_collector"); }

/* timestep.c:589-699 with time_step_method = 1 (velocity and acceleration), on
 * the device: each particle uses the cell size of the finest node holding it */
int timestep(vtype *ptr_dt)
{
  if (qz_ensure("timestep") == _FAILURE_)
    return _FAILURE_;
  if (time_step_method != 1)
  {
    printf("\n\n Error running timestep(): the device path implements time_step_method = 1 only "
           "(global_variables.c)\n\n");
    return _FAILURE_;
  }
  if (!g_uploaded && qz_upload("timestep") == _FAILURE_)
    return _FAILURE_;
  double dt = 0.0;
  if (qz_timestep(g_ctx, (double)_CFL_, (double)_MAX_dt_, &dt))
    return qz_fail("timestep");
  *ptr_dt = (vtype)dt;
  return _SUCCESS_;
}

/* observables.c:661-696: GL_energies (kinetic, potential, total) and
 * GL_momentum in the user's units from the device's FP64 sums in code units.
 * With v_user = v / c_v + v_cm (observables.c:38-75, 622-659):
 *   E_kin = (1/2 sum m v^2) / c_v^2 + v_cm . (sum m v) / c_v + M |v_cm|^2 / 2,
 *   p     = (sum m v) / c_v + M v_cm,
 * and E_pot scaled by c_t^2 / c_d^2 (observables.c:110, 617). */
int observables(void)
{
  if (qz_ensure("observables") == _FAILURE_)
    return _FAILURE_;
  if (potential_energy_type != 0 && potential_energy_type != 1)
  {
    printf("Error, the potential_energy_type value is different of 0 or 1\n");
    return _FAILURE_;
  }
  if (!g_uploaded && qz_upload("observables") == _FAILURE_)
    return _FAILURE_;
  double o6[6];
  if (qz_observables(g_ctx, potential_energy_type == 0 ? 1 : 0, o6))
    return qz_fail("observables");
  double M = 0.0;
  for (int i = 0; i < GL_no_ptcl_final; i++)
    M += (double)GL_ptcl_mass[i];
  const double cv = (double)_conversion_velocity_, ct = (double)_conversion_time_, cd = (double)_conversion_dist_;
  const double vcm[3] = {(double)GL_cm_vx, (double)GL_cm_vy, (double)GL_cm_vz};
  double ek = o6[0] / (cv * cv) + 0.5 * M * (vcm[0] * vcm[0] + vcm[1] * vcm[1] + vcm[2] * vcm[2]);
  for (int a = 0; a < 3; a++)
  {
    ek += vcm[a] * o6[3 + a] / cv;
    GL_momentum[a] = (vtype)(o6[3 + a] / cv + M * vcm[a]);
  }
  GL_energies[0] = (vtype)ek;
  GL_energies[1] = (vtype)(o6[1] * (ct * ct) / (cd * cd));
  GL_energies[2] = GL_energies[0] + GL_energies[1];
  return _SUCCESS_;
}

/* output_snapshots.c: the output folders follow the reference's naming rules
 * (main folder ../output, the simulation folder named by Maxdt, the levels and
 * the particle count, with "_Release(k)" appended when it exists, snapshots in
 * its Data/ folder as Data_<n>.bin); output_main_parameters() reads folder_name. */
char file_data_name[1000];
char folder_name[1000];

static int qz_output_folders(void)
{
  const char *main_dir = "../output";
  char sim[300], tried[700];
  mkdir(main_dir, 0777);
  snprintf(sim, sizeof(sim), "%s/Simulation_Maxdt[My](%1.1f)_lv(%d+%d)_Np(%d)", main_dir, (double)(Maxdt / _Myear_),
           lmin, lmax - lmin, GL_no_ptcl_initial);
  snprintf(tried, sizeof(tried), "%s", sim);
  for (int k = 1; mkdir(tried, 0777) == -1; k++)
    snprintf(tried, sizeof(tried), "%s_Release(%d)", sim, k);
  snprintf(folder_name, sizeof(folder_name), "%s", tried);
  char data[1100];
  snprintf(data, sizeof(data), "%s/Data", folder_name);
  mkdir(data, 0777);
  snprintf(file_data_name, sizeof(file_data_name), "%s/Data_", data);
  folder_created = true;
  return _SUCCESS_;
}

/* pending snapshots are finished before the program exits */
static void qz_dropin_finish(void)
{
  if (g_ctx && qz_snapshot_wait(g_ctx))
    printf("\n\n Error writing a snapshot: %s\n\n", qz_last_error());
}

/* output_snapshots.c:141-239: the snapshot file (masses, positions and
 * velocities in the user's units, GL_energies and GL_momentum when
 * compute_observables_FLAG, then actualtime / _conversion_time_), written by
 * the device library's host thread while the simulation continues */
int output_snapshots(vtype actualtime, int snapshot)
{
  static int finish_registered = 0;
  if (qz_ensure("output_snapshots") == _FAILURE_)
    return _FAILURE_;
  if (!g_uploaded && qz_upload("output_snapshots") == _FAILURE_)
    return _FAILURE_;
  if (!folder_created && qz_output_folders() == _FAILURE_)
    return _FAILURE_;
  if (!finish_registered)
  {
    atexit(qz_dropin_finish);
    finish_registered = 1;
  }
  char name[1500];
  snprintf(name, sizeof(name), "%s%d.bin", file_data_name, snapshot);
  qz_units u;
  u.dist = (double)_conversion_dist_;
  u.velocity = (double)_conversion_velocity_;
  u.time = (double)_conversion_time_;
  u.cm[0] = (double)GL_cm_x;
  u.cm[1] = (double)GL_cm_y;
  u.cm[2] = (double)GL_cm_z;
  u.cm_v[0] = (double)GL_cm_vx;
  u.cm_v[1] = (double)GL_cm_vy;
  u.cm_v[2] = (double)GL_cm_vz;
  double obs[6];
  for (int a = 0; a < 3; a++)
  {
    obs[a] = compute_observables_FLAG ? (double)GL_energies[a] : 0.0;
    obs[3 + a] = compute_observables_FLAG ? (double)GL_momentum[a] : 0.0;
  }
  if (qz_write_snapshot_async(g_ctx, name, &u, (double)actualtime, compute_observables_FLAG ? obs : NULL))
    return qz_fail("output_snapshots");
  return _SUCCESS_;
}

/* Build-only addition: copy the device state (positions, velocities and
 * accelerations) into GL_ptcl_*, for a host that reads the particle arrays
 * itself; any snapshot still being written is finished first. */
int qz_dropin_sync_host(void)
{
  if (!g_ctx || !g_uploaded)
    return _SUCCESS_;
  if (qz_snapshot_wait(g_ctx))
    return qz_fail("qz_dropin_sync_host");
  return qz_download(1 | 2 | 4, "qz_dropin_sync_host");
}
#endif
