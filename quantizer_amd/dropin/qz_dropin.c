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
 *  - QZ_DROPIN_RESIDENT: additionally particle_updating_A, particle_updating_B
 *    and tree_adaptation; particles stay on the device between steps and the
 *    host arrays are refreshed after each update for output and diagnostics.
 *    The host tree is then built once by tree_construction() and not adapted.
 *
 * In both builds the device rebuilds the refinement tree from the particles at
 * every force pass with the rules of tree_construction.c (marks, zones, child
 * boxes), so the node geometry is the one tree_construction() would produce.
 * Errors follow the reference convention: _FAILURE_ with a printed reason.
 */
#include "common.h"
#include "ptcl_idx_to_box_idx.h"

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

/* reset.c:29-64: zero the particle accelerations and the head density */
int reset(void)
{
  if (qz_ensure("reset") == _FAILURE_)
    return _FAILURE_;
  for (int i = 0; i < GL_no_ptcl_final; i++)
  {
    GL_ptcl_ax[i] = 0;
    GL_ptcl_ay[i] = 0;
    GL_ptcl_az[i] = 0;
  }
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

/* particle_acceleration.c:522-573; the accelerations return to GL_ptcl_a* */
int particle_acceleration(void)
{
  if (qz_ensure("particle_acceleration") == _FAILURE_ || qz_particle_acceleration(g_ctx))
    return qz_fail("particle_acceleration");
  return qz_download(4, "particle_acceleration");
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
  return qz_download(1 | 2 | 4, "particle_updating_A");
}

/* particle_updating.c:2787-2836: closing half kick */
void particle_updating_B(vtype dt)
{
  if (qz_ensure("particle_updating_B") == _FAILURE_)
    return;
  if (qz_particle_updating_B(g_ctx, (double)dt))
  {
    qz_fail("particle_updating_B");
    return;
  }
  qz_download(2, "particle_updating_B");
}

/* tree_adaptation.c:8250: the device rebuilds its tree at every grid_density() */
int tree_adaptation(void) { return qz_ensure("tree_adaptation"); }
#endif
