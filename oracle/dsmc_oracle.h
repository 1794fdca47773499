/*
 * CPU oracle for the DSMC engine — TEST INFRASTRUCTURE ONLY.
 *
 * A strict restatement of the reference's per-step DSMC path
 * (src/ProcessDSMC.cpp, src/ProcessDSMCGeometry.cpp, src/MeshparticalInitial.*)
 * for one rank that owns every cell, kept in the reference's own structure:
 * per-cell particle buckets, a recv_cache of movers, cache2chain regrouping,
 * per-cell NTC collisions and sampling.  Random draws come from the same
 * Philox streams the engine uses (the reference's std::mt19937 stream is
 * neither reproducible nor parallel), and elementary functions from the
 * shared deterministic implementations, so the engine can be compared with it
 * bit for bit.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
 * this library.  The engine never links or calls it.
 */
#ifndef DSMC_ORACLE_H
#define DSMC_ORACLE_H

#include <stdint.h>
#include "dsmc_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dsmc_oracle dsmc_oracle;

int oracle_create(const dsmc_mesh_view *mesh, const dsmc_params *p, uint64_t seed, dsmc_oracle **out);
void oracle_destroy(dsmc_oracle *o);
int oracle_set_threads(int32_t n);
/* Plants one known mistake in every oracle created afterwards (0 removes
 * it), so that a test can show that its closed-form check fails under that
 * mistake.  Never set outside those tests.
 *   1 VHS prefactor with the molecular mass where the reduced mass belongs
 *     (src/ProcessDSMC.cpp:2626-2630)
 *   2 VHS collision rate with the hard-sphere speed power (c_r^1 instead of
 *     c_r^(1-2 eta)) while the prefactor keeps its eta
 *   3 Larsen-Borgnakke exchange offered to the first partner only
 *     (src/ProcessDSMC.cpp:2700-2712 offers it to both)
 *   4 Larsen-Borgnakke energy split with the exponent 1/(1.5 - omega)
 *     instead of 1/(2.5 - omega)
 *   5 inflow normal speed drawn without the flux weight v_n
 *     (src/ProcessDSMC.cpp:2435-2441)
 *   6 inflow tangential speeds scaled by T instead of sqrt(T)
 *     (src/ProcessDSMC.cpp:2446-2449) */
#define ORACLE_PLANT_NONE 0
#define ORACLE_PLANT_VHS_MASS 1
#define ORACLE_PLANT_VHS_HS_POWER 2
#define ORACLE_PLANT_LB_ONE_PARTNER 3
#define ORACLE_PLANT_LB_EXPONENT 4
#define ORACLE_PLANT_INFLOW_NO_FLUX_WEIGHT 5
#define ORACLE_PLANT_INFLOW_TANGENT_SCALE 6
int oracle_plant(int32_t which);
int oracle_set_particles(dsmc_oracle *o, const dsmc_particles_soa *soa);
int64_t oracle_count(dsmc_oracle *o);
int oracle_get_particles(dsmc_oracle *o, dsmc_particles_soa *soa);
int oracle_step(dsmc_oracle *o, int32_t istep, uint32_t stage_mask);
int oracle_log_enable(dsmc_oracle *o, int32_t on);
int64_t oracle_log_get(dsmc_oracle *o, dsmc_cross_event *out, int64_t cap);
int oracle_macro_get(dsmc_oracle *o, int32_t which, double *out);
int oracle_ns_apply(dsmc_oracle *o, int32_t call_step, const double *ns, const uint8_t *accepted,
                    int32_t *nan_cell);
int oracle_statistic_pre(dsmc_oracle *o);
int oracle_accum_get(dsmc_oracle *o, int32_t family, double *out);
int oracle_cell_scalars(dsmc_oracle *o, int32_t *count, double *crmax, double *rem_coll, double *rem_emit,
                        uint8_t *sparse_state, int32_t *sparse_steps);
int oracle_surface_get(dsmc_oracle *o, double *out);
/* outBoundaryStressHeat (src/ProcessDSMC.cpp:3494-3557): nface x 12 wall
 * fields in dsmc_surface_fields' layout (stress(3), q_trans, q_rot,
 * normal(3), pressure, shear(3); zero off the wall), from `tally` (nface x 6)
 * or, when it is null, the oracle's own steady tallies. */
int oracle_surface_fields(dsmc_oracle *o, int32_t istep, const double *tally, double *fields);
int oracle_stats(dsmc_oracle *o, dsmc_stats *s);
/* Stage timings of the last oracle_step calls (seconds, accumulated):
 * preprocess (emission + sampling), advection, cache2chain, collision —
 * the reference driver's split (src/meshImportTest.cpp:59-64). */
int oracle_times(dsmc_oracle *o, double *t4);
/* Restated quad split choice (src/meshImportDsmcTopology.cpp:106-168). */
int oracle_split_tags(const dsmc_mesh_view *mesh, uint8_t *out);
/* Restated Moller-Trumbore test on one cached triangle
 * (src/ProcessDSMC.cpp:1174-1197): returns 1 on a hit and writes t. */
int oracle_tri_hit(const double *a, const double *e1, const double *e2, const double *x, const double *v,
                   double *t);

/* The shared deterministic functions evaluated on the host, same layout as
 * dsmc_device_math_check. */
int oracle_math(const double *u, int64_t n, double *out);
/* One Philox4x32-10 block from the shared generator (known-answer tests). */
int oracle_philox(const uint32_t *ctr4, const uint32_t *key2, uint32_t *out4);
/* First n uniforms of a Stream (key, a, b, purpose, sub). */
int oracle_stream_uniforms(uint64_t key, uint32_t a, uint32_t b, uint32_t purpose, uint32_t sub, int64_t n,
                           double *out);

#ifdef __cplusplus
}
#endif
#endif
