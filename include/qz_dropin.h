/*
 * Entry points of libqzdropin.so / libqzdropin_resident.so: QUANTIZER's own
 * per-step functions, same names and signatures, backed by the MI355X library
 * (qz.h).  They read and write the reference's host globals (GL_ptcl_*, lmin,
 * lmax, _G_, ...), so the libraries are built inside a QUANTIZER tree against
 * its headers (quantizer_amd/dropin/qz_dropin.c) and linked in place of the
 * objects named below.  Return values follow the reference: _SUCCESS_ (0) /
 * _FAILURE_ (1) with a printed reason.  `vtype` is the reference's scalar type
 * (common.h:48-70, double by default); `struct node` is its node
 * (initialize_node.h:63-216).
 */
#ifndef QZ_DROPIN_H
#define QZ_DROPIN_H

#ifdef __cplusplus
extern "C" {
#endif

struct node;
#ifndef __COMMON__
typedef double vtype;
#endif

/* both builds (replace grid_density.o, potential*.o, grid_acceleration.o,
 * particle_acceleration.o, reset.o, ptcl_idx_to_box_idx.o) */
void grid_density(void);                                /* grid_density.h:32, grid_density.c:269 */
int potential(void);                                    /* potential.h:36, potential.c:29 */
int potential_head_node(void);                          /* potential_head_node.h:36, potential_head_node.c:99 */
int potential_branches(void);                           /* potential_branches.h:35, potential_branches.c:401 */
int grid_acceleration(void);                            /* grid_acceleration.h:32, grid_acceleration.c:694 */
int particle_acceleration(void);                        /* particle_acceleration.h:32, particle_acceleration.c:522 */
int reset(void);                                        /* reset.h:33, reset.c:29 */
int ptcl_idx_to_box_idx(struct node *ptr_node, int ptcl_idx); /* ptcl_idx_to_box_idx.h:34, ptcl_idx_to_box_idx.c:30 */

/* resident build only (also replaces particle_updating.o, tree_adaptation.o,
 * timestep.o, observables.o, output_snapshots.o and garbage_collector.o):
 * particles stay on the device, and GL_ptcl_* are refreshed only by
 * qz_dropin_sync_host() */
int particle_updating_A(vtype dt); /* particle_updating.h:33, particle_updating.c:2584 */
void particle_updating_B(vtype dt); /* particle_updating.h:34, particle_updating.c:2787 */
int tree_adaptation(void);          /* tree_adaptation.h:40, tree_adaptation.c:8250: the device rebuilds its tree */
int timestep(vtype *ptr_dt);        /* timestep.h:32, timestep.c:589 (time_step_method = 1) */
int observables(void);              /* observables.h:66, observables.c:661 */
int output_snapshots(vtype actualtime, int snapshot); /* output_snapshots.h:100, output_snapshots.c:141 (written asynchronously) */
int garbage_collector(void);        /* garbage_collector.h:34, garbage_collector.c:377: nothing to trim */
/* build-only addition (no reference counterpart): copy the device state into
 * GL_ptcl_x..az for a host that reads them; waits for pending snapshots */
int qz_dropin_sync_host(void);

#ifdef __cplusplus
}
#endif
#endif
