/*
 * qz_oracle: CPU restatement (FP64, sequential) of the QUANTIZER per-step
 * particle-mesh gravity cycle.  TEST INFRASTRUCTURE ONLY: it is the checker
 * for the HIP path and the "port" CPU baseline; the product never links it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the QUANTIZER source tree).  Non-periodic boundaries only (the reference
 * default bdry_cond_type = 1, reflexive; global_variables.c:277).
 */
#ifndef QZ_ORACLE_H
#define QZ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oq_sim oq_sim;

/* Head boundary potential: 0 = monopole -GM/r about the box centre
 * (initial_potential_and_acceleration_head.c:297-335), 1 = zero Dirichlet. */
oq_sim *oq_create(int lmin, int lmax, int ref_crit, int n, const double *x, const double *y, const double *z,
                  const double *vx, const double *vy, const double *vz, const double *m, int head_bc);
void oq_destroy(oq_sim *s);

/* Tree (re)construction, tree_construction.c:2521-2577 (non-periodic). */
int oq_build_tree(oq_sim *s);
/* One force pass: reset (reset.c:29), grid_density, potential, grid_acceleration,
 * particle_acceleration (main.c:430-465).  Rebuilds the tree first when lmin<lmax. */
int oq_force(oq_sim *s);
/* particle_updating_A / B (particle_updating.c:39-207, 2677-2836). */
int oq_update_A(oq_sim *s, double dt);
void oq_update_B(oq_sim *s, double dt);
/* A, force, B (main.c:391-478 with fixed dt). */
int oq_step(oq_sim *s, double dt);

/* Individual stages (each assumes the previous ones ran). */
void oq_reset(oq_sim *s);
void oq_grid_density(oq_sim *s);
int oq_potential(oq_sim *s);
void oq_grid_acceleration(oq_sim *s);
void oq_particle_acceleration(oq_sim *s);

/* Accessors. out9: x,y,z,vx,vy,vz,ax,ay,az (each n doubles). */
int oq_num_particles(const oq_sim *s);
void oq_get_particles(const oq_sim *s, double *out9);
void oq_set_particles(oq_sim *s, const double *in9);
/* head fields rho(=4 pi G rho), pot, ax, ay, az on the (2^lmin+1)^3 solve domain */
void oq_get_head(const oq_sim *s, double *out5);
void oq_set_head_pot(oq_sim *s, const double *pot);
int oq_num_nodes(const oq_sim *s);
/* header layout matches the reference harness dump (24 ints) */
void oq_node_header(const oq_sim *s, int i, int *hdr24);
/* which: 0 cell_xyz(3*ncell) 1 cell_box(ncell) 2 intr_xyz 3 intr_idx 4 bdry_xyz 5 bdry_idx
 *        6 sim_xyz 7 sim_idx 8 box status (box cells) */
void oq_node_ints(const oq_sim *s, int i, int which, int *out);
/* fields rho, pot, ax, ay, az over the node grid box (gcap each) */
void oq_node_fields(const oq_sim *s, int i, double *out5);
/* solver statistics of the last potential(): head V-cycles, total branch CG iterations */
void oq_stats(const oq_sim *s, int *out4);
/* observables.c:661-696: E_kin, E_pot (exact pairwise), E_tot, px, py, pz */
void oq_observables(const oq_sim *s, double *out6);

/* A1 ptcl_idx_to_box_idx.c:30-139, non-periodic branch: box index of a point
 * in a node of level lv with translation ts and real box dims rd. */
int oq_key(int lv, const int *ts, const int *rd, double x, double y, double z);

/* Standalone head Poisson solve of the reference algorithm on a packed cube
 * of side 2^lmin+1 (potential_head_node.c:99-179, multigrid.c, gauss_saidel.c,
 * poisson_error.c).  Returns V-cycles used, or -1 on failure. */
int oq_head_solve(int lmin, double *phi, const double *rho, double node_mass, double tol, int max_iter);

/* FP64 reverse mode of the head-grid step (lmin == lmax).  abar/xbar and the
 * cotangent arrays are 3N doubles (x, then y, then z components).
 * oq_force_vjp: xbar += J_A(x)^T abar for the force pass at the current positions.
 * oq_adjoint: from the current state (force pass done) run nsteps fixed-dt
 * steps and return the cotangents of the initial x, v from those of the final. */
/*
 * oq_cfl_dt: min over particles i and axes k of the CFL time step of velocity
 * v3[k*n+i] and acceleration a3[k*n+i] with cell size H[i], capped at dt_max.
 * oq_timestep: the same over the sim's particles, H of each particle's finest
 * node (needs a force pass).
 */
double oq_cfl_dt(int n, const double *v3, const double *a3, const double *H, double cfl, double dt_max);
double oq_timestep(const oq_sim *s, double cfl, double dt_max);
/* 1/2 G sum m phi at each particle's finest node (needs a force pass) */
double oq_grid_pe(const oq_sim *s);
int oq_force_vjp(oq_sim *s, const double *abar, double *xbar);
int oq_adjoint(oq_sim *s, int nsteps, double dt, const double *xbN, const double *vbN, double *xb0, double *vb0);

/* Configuration knobs (defaults = reference defaults, global_variables.c:313-334). */
void oq_set_tolerance(oq_sim *s, double tol);

#ifdef __cplusplus
}
#endif
#endif
