/*
 * qz.h -- native C ABI of the MI355X particle-mesh gravity library (libqzhip.so).
 *
 * One context holds the device-resident state of one simulation: particles
 * (structure of arrays, kept sorted by head-cell key), the head grid with its
 * multigrid hierarchy, and the refinement nodes flattened into one pool per
 * level.  Every entry point returns 0 on success and 1 on failure (the
 * reference convention, common.h:81-82) and returns only after the device
 * work it queued has finished; qz_last_error() describes the last failure.
 *
 * Entry points mirror the reference's per-step functions (main.c:369-535);
 * the reference-named drop-in wrappers are declared in qz_dropin.h.
 */
#ifndef QZ_H
#define QZ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qz_ctx qz_ctx;

enum
{
  QZ_PREC_FP64 = 0,  /* particles, grids and V-cycles in FP64 (the "plain" path) */
  QZ_PREC_MIXED = 1  /* FP32 particles/grids/V-cycles, FP64 outer residual + correction */
};

enum
{
  QZ_BC_MONOPOLE = 0, /* head faces at -GM/r about the centre (initial_potential_and_acceleration_head.c:297-335) */
  QZ_BC_ZERO = 1      /* zero Dirichlet faces */
};

enum
{
  QZ_BRANCH_CG = 0, /* unpreconditioned CG per node (conjugate_gradient.c:327-460) */
  QZ_BRANCH_MG = 1  /* batched masked multigrid per tree level, FP64 outer refinement */
};

typedef struct
{
  int lmin;               /* head level: head grid has 2^lmin cells per side (global_variables.c:252) */
  int lmax;               /* finest level allowed (global_variables.c:253) */
  int n_particles;
  int precision;          /* QZ_PREC_* */
  int head_bc;            /* QZ_BC_* */
  int ref_criterion_ptcl; /* refine a cell holding >= this many particles (global_variables.c:287) */
  int n_exp;              /* refinement dilation (global_variables.c:289); only 1 is supported */
  double G;               /* gravitational constant in code units */
  double tol;             /* RMS residual / (4 pi G rho_mean) threshold (global_variables.c:318) */
  double tol_change;      /* pointwise |phi/phi_old - 1| threshold (global_variables.c:319) */
  int max_iterations;     /* V-cycle / CG iteration cap (global_variables.c:313) */
  int inner_cycles;       /* head: V-cycles per outer FP64 refinement iteration (0 = chosen per iteration
                             from its residual, 1 to 4) */
  int pre_sweeps;         /* red-black GS sweeps before restriction */
  int post_sweeps;        /* red-black GS sweeps after prolongation */
  int branch_solver;      /* QZ_BRANCH_* solver of the refinement-node boxes */
} qz_config;

typedef struct
{
  int head_outer;      /* outer FP64 residual/correction iterations of the last head solve */
  int head_vcycles;    /* inner V-cycles of the last head solve */
  double head_residual; /* final RMS residual / (4 pi G rho_mean) of the head */
  double head_change;   /* final max |phi/phi_old - 1| of the head */
  int n_nodes;         /* refinement nodes in the tree */
  int n_levels;        /* levels in use including the head */
  long long branch_points;    /* grid points over all refinement nodes */
  long long branch_intr;      /* interior (unknown) points over all refinement nodes */
  int branch_iterations;      /* max CG iterations over nodes in the last branch solve */
  long long branch_iter_points; /* sum over nodes of iterations x interior points */
  int branch_cg_levels;       /* tree levels of the last branch solve finished by CG */
} qz_stats;

void qz_default_config(qz_config *cfg);
qz_ctx *qz_create(const qz_config *cfg);
void qz_destroy(qz_ctx *ctx);
const char *qz_last_error(void);

/* host <-> device particle transfer (original particle order, FP64 host arrays) */
int qz_set_particles(qz_ctx *ctx, const double *x, const double *y, const double *z, const double *vx,
                     const double *vy, const double *vz, const double *m);
int qz_get_particles(qz_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                     double *ax, double *ay, double *az);
/* the same upload as qz_set_particles, keeping the head potential as the next
 * solve's starting guess (the reference never resets it, reset.c:54-57); for
 * hosts that move the particles themselves between force passes */
int qz_load_particles(qz_ctx *ctx, const double *x, const double *y, const double *z, const double *vx,
                      const double *vy, const double *vz, const double *m);

/* per-step stages (main.c:391-478) */
int qz_particle_updating_A(qz_ctx *ctx, double dt); /* particle_updating.c:2584 */
int qz_reset(qz_ctx *ctx);                          /* reset.c:29 */
int qz_tree_adaptation(qz_ctx *ctx);                /* tree_adaptation.c:8250 (rebuild) */
int qz_grid_density(qz_ctx *ctx);                   /* grid_density.c:269 */
int qz_potential(qz_ctx *ctx);                      /* potential.c:29 */
int qz_potential_head_node(qz_ctx *ctx);            /* potential_head_node.c:99 */
int qz_potential_branches(qz_ctx *ctx);             /* potential_branches.c:401 */
int qz_grid_acceleration(qz_ctx *ctx);              /* grid_acceleration.c:694 */
int qz_particle_acceleration(qz_ctx *ctx);          /* particle_acceleration.c:522 */
int qz_particle_updating_B(qz_ctx *ctx, double dt); /* particle_updating.c:2787 */

/* composites: initial force pass (main.c:280-338) and one fixed-dt step (main.c:391-478) */
int qz_force(qz_ctx *ctx);
int qz_step(qz_ctx *ctx, double dt);

/* field access: head rho (= 4 pi G rho), pot, ax, ay, az on the (2^lmin+1)^3
 * solve domain, x fastest; out5 holds 5 such arrays */
int qz_get_head_fields(qz_ctx *ctx, double *out5);
int qz_set_head_potential(qz_ctx *ctx, const double *pot);

/* refinement nodes: count, 24-int headers (same layout as the oracle dumps),
 * and per-node fields on the node's (rd+1)^3 box grid */
int qz_num_nodes(qz_ctx *ctx);
int qz_node_header(qz_ctx *ctx, int i, int *hdr24);
int qz_node_grid_lists(qz_ctx *ctx, int i, int *intr_xyz, int *bdry_xyz, int *sim_xyz);
int qz_node_fields(qz_ctx *ctx, int i, double *out5);

int qz_get_stats(qz_ctx *ctx, qz_stats *st);

/* FP64 diagnostics (observables.c:661-696): E_kin, E_pot, E_tot, px, py, pz.
 * exact_pe = 1 uses the pairwise sum (observables.c:77-112), 0 the grid
 * interpolation (observables.c:594-619). */
int qz_observables(qz_ctx *ctx, int exact_pe, double *out6);

/* CFL time step with velocity and acceleration (timestep.c:589-699 with
 * time_step_method = 1): min over particles and axes, each particle using the
 * cell size of the finest node holding it, capped at dt_max */
int qz_timestep(qz_ctx *ctx, double cfl, double dt_max, double *dt);

/* Adjoint of the step, head grid and refinement levels.  Cotangent and
 * gradient arrays hold 3N doubles, x then y then z components, in the
 * original particle order.  With levels the tree is held fixed (cell
 * membership is piecewise constant in the positions).
 *  qz_force_vjp: xbar = J_A(x)^T abar for the force pass a = A(x) at the
 *    current positions (CIC deposits, head and node solves with the parent
 *    boundary averages, gradients, gather from the finest node).
 *  qz_adjoint_forward: from the current state (force pass done) run nsteps
 *    fixed-dt steps keeping checkpoints every checkpoint_every steps.
 *  qz_adjoint_backward: given the cotangents of the final positions and
 *    velocities, return those of the initial ones; forward states are
 *    recomputed from the checkpoints.  The context is left at the final state.
 *  The (3, N) arrays of both calls (x, y, z rows, particle-ID order) may be
 *  host or device memory: the copies infer the direction. */
int qz_force_vjp(qz_ctx *ctx, const double *abar, double *xbar);
int qz_adjoint_forward(qz_ctx *ctx, int nsteps, double dt, int checkpoint_every);
int qz_adjoint_backward(qz_ctx *ctx, const double *xbar_n, const double *vbar_n, double *xbar_0, double *vbar_0);

/* A1 (ptcl_idx_to_box_idx.c:30-139) on the device for n host positions:
 * key = (f(x 2^lv) - ts) linearised with the real box dims rd, f(p) = p<0 ? -1 : (int)p */
int qz_keys(const double *x, const double *y, const double *z, int n, int lv, const int *ts, const int *rd,
            int *out);

/* device timing (HIP events on the context's stream) of every kernel launched
 * since qz_timing_enable(ctx, 1): per kernel name the total ms, the launch
 * count and the algorithmic bytes (each input read once, each output written
 * once), up to max entries; returns the number of entries.  Timing launches
 * kernels one by one; qz_timing_enable(ctx, 2) keeps the captured graphs and
 * records only host wall-time spans ("host:" names).  0 turns timing off. */
int qz_timing_enable(qz_ctx *ctx, int on);
int qz_timing_report(qz_ctx *ctx, char *names, int name_len, double *ms, int *launches, double *bytes, int max);
int qz_sync(qz_ctx *ctx);

/* Snapshot in the reference's binary layout (output_snapshots.c:141-226): mass,
 * x, y, z, vx, vy, vz (N values each, original particle order, user units:
 * x_user = (x - 1/2) / dist + cm + cm_v t, v_user = v / velocity + cm_v), then,
 * with observables, E_kin, E_pot (exact pairwise), E_tot and the momentum
 * (observables.c:38-112, 622-696), then t = actualtime / time; all doubles.
 * The unit factors are the reference's _conversion_dist_, _conversion_velocity_,
 * _conversion_time_ and GL_cm_* (input.c:252-295). */
typedef struct
{
  double dist, velocity, time;
  double cm[3], cm_v[3];
} qz_units;
int qz_write_snapshot(qz_ctx *ctx, const char *path, const qz_units *units, double actualtime, int with_observables);

/* The same file without stalling the caller (output_snapshots.c:141-226): the
 * state is copied on the device, ordered after the work queued so far, and the
 * call returns; a host thread of the context downloads the copy and writes
 * `path` while later steps run.  obs6 (E_kin, E_pot, E_tot, px, py, pz, already
 * in the caller's units: the reference's GL_energies and GL_momentum) is
 * written after the arrays, or nothing when obs6 is NULL.  One snapshot is in
 * flight per context: a second call first waits for the one before.
 * qz_snapshot_wait returns once the last snapshot is written: 0, or 1 with the
 * writer's error. */
int qz_write_snapshot_async(qz_ctx *ctx, const char *path, const qz_units *units, double actualtime,
                            const double *obs6);
int qz_snapshot_wait(qz_ctx *ctx);

/* Parameters.dat (output_main_parameters.c:167-230): the run's parameters as
 * 38 "name value" lines in the reference's order (example_plot.py reads them),
 * then the wall times in the reference's categories (times: 20 values in the
 * order of the reference's GL_times[0..19], in seconds, or NULL for zeros).
 * The configuration lines come from the context, the rest from info. */
typedef struct
{
  double cfl;                /* _CFL_ */
  double max_dt;             /* _MAX_dt_ (code units) */
  double box_size;           /* _User_BoxSize_ */
  double final_time;         /* actualtime at the end (code units; written divided by units->time) */
  double total_mass_initial; /* GL_total_mass_initial */
  double total_mass_final;   /* GL_total_mass_final */
  int potential_energy_type; /* 0 exact pairwise, 1 grid interpolation */
  int gc_iter;               /* Garbage_Collector_iter */
  int n_particles_initial;   /* GL_no_ptcl_initial (0: the context's particle count) */
  int number_timesteps;
  int number_outputs;
  int fr_output;
  int max_iterations;        /* MaxIterations */
  const double *times;
} qz_run_info;
int qz_write_parameters(qz_ctx *ctx, const char *path, const qz_units *units, const qz_run_info *info);

/* FP64 diagnostics without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, hosts that forbid hipMalloc).  The quantities of
 * qz_observables with exact_pe = 1 (E_kin, E_pot from the exact pairwise sum
 * observables.c:77-112, E_tot, px, py, pz; observables.c:38-75, 661-696) for n
 * particles given as FP64 device arrays in any order; with_pe = 0 skips the
 * O(n^2) pair sum and reports E_pot = 0.  out6 may be host or device memory.
 * Partial sums live in a fixed device array of the library (2 MiB per GPU,
 * held until the process ends), so n is limited to QZ_NOALLOC_MAX_PARTICLES;
 * larger n is refused with an error.  Runs on the current device's null
 * stream and returns when done.  Calls on one device share that array:
 * concurrent calls on the same device race. */
#define QZ_NOALLOC_MAX_PARTICLES (1 << 24)
int qz_observables_noalloc(const double *x, const double *y, const double *z, const double *vx, const double *vy,
                           const double *vz, const double *m, int n, double G, int with_pe, double *out6);

/* measurement helpers (no counterpart in the reference): the head's finest red
 * half sweep of the multigrid smoother replayed `reps` times back to back,
 * returning ms and algorithmic bytes per launch; and one plain device copy of
 * `bytes` bytes (kernel k_copy_probe) whose hardware byte counts calibrate them */
int qz_bench_sweep(qz_ctx *ctx, int reps, double *ms_per_launch, double *bytes_per_launch);
int qz_copy_probe(qz_ctx *ctx, size_t bytes, double *ms);

#ifdef __cplusplus
}
#endif
#endif
