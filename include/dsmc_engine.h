/*
 * dsmc_engine.h — C ABI of the MI355X DSMC particle engine.
 *
 * The engine replaces the per-step particle work of the reference's
 * ProcessDSMC (src/ProcessDSMC.h:23-493): inflow emission, face-walk
 * tracking with wall/inlet/outlet handling, regrouping particles by cell,
 * NTC collisions with VHS scattering (and Larsen-Borgnakke rotational
 * exchange), and the exponential/steady/interval cell sampling.  Everything
 * stays resident in the GPU's HBM; the caller drives it once per time step in
 * the stage order of the reference driver (src/meshImportTest.cpp:207-234).
 *
 * Conventions (same as the reference, src/meshImport.h:168-208):
 *   - positions in mesh units; velocities in units of v_rms = sqrt(2 kB T_ref/m);
 *     the move uses x += v * dtime (dtime in length units, 0.2 x min edge);
 *   - cell "volume" is the reference's DsmcCell::area, face "area" its
 *     DsmcEdge::length; face normals are unit vectors oriented from the
 *     face's first cell c0 to its second cell c1 (outward on boundaries);
 *   - macro outputs are ncell x 18 doubles in MacroIndex order
 *     (src/ProcessDSMC.h:28-49): rho, ux, uy, uz, T, sxx, sxy, syy, syz, szz,
 *     szx, qx, qy, qz, Tr, qrx, qry, qrz.
 *
 * Every function returns DSMC_OK (0) or a negative DSMC_ERR_* code; the text
 * of the last error on the calling thread is returned by dsmc_last_error().
 * The engine copies what it needs: no caller pointer is kept after a call
 * returns.  One engine is not thread-safe; distinct engines may be used from
 * distinct host threads.
 */
#ifndef DSMC_ENGINE_H
#define DSMC_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DSMC_OK 0
#define DSMC_ERR_ARG (-1)      /* bad argument / size                          */
#define DSMC_ERR_HIP (-2)      /* HIP runtime failure (no GPU, launch error)   */
#define DSMC_ERR_MESH (-3)     /* inconsistent mesh                            */
#define DSMC_ERR_CAPACITY (-4) /* particle or log capacity exceeded            */
#define DSMC_ERR_STATE (-5)    /* call out of order (e.g. no particles yet)    */

/* ---- boundary roles (reference BoundaryRole, src/BoundaryCondition.h:10-24;
 *      Fluent zone tags 3/4/5/6 map to interface/inlet/outlet/wall as in
 *      BoundaryConditionTable::loadDefaults, src/BoundaryCondition.h:79-117) */
#define DSMC_TAG_INTERFACE 3
#define DSMC_TAG_INLET 4
#define DSMC_TAG_OUTLET 5
#define DSMC_TAG_WALL 6

#define DSMC_WALL_DIFFUSE 0  /* reference model (src/ProcessDSMC.cpp:1793-1825) */
#define DSMC_WALL_SPECULAR 1 /* declared but unimplemented in the reference
                                (src/BoundaryCondition.h:44); new here        */

#define DSMC_PREC_FP64 0
#define DSMC_PREC_MIXED 1 /* FP32 velocities/Ir and collision arithmetic;
                             FP64 positions, face tests, hit times, sums */

/* Quad split tags (reference meshImport::FACE_SPLIT_*, src/meshImport.h:297-324) */
#define DSMC_SPLIT_02 0
#define DSMC_SPLIT_13 1
#define DSMC_SPLIT_INVALID 255

#define DSMC_MACRO_WIDTH 18
#define DSMC_MAX_CELL_FACES 6

/* ---- mesh ---------------------------------------------------------------- */
/* Read-only view of an unstructured tet/hex/prism/pyramid mesh; the fields
 * mirror DsmcCell / DsmcEdge / localPointXY / faceSplitTag
 * (src/meshImport.h:168-189, src/ProcessDSMC.h:347-350) for a single rank
 * that owns every cell. */
typedef struct dsmc_mesh_view {
  int32_t ncell, nface, nnode;
  const double *node_xyz;      /* nnode*3                                         */
  const int32_t *face_nodes;   /* nface*4: faceMap[0..3]; [3] = -1 for triangles  */
  const int32_t *face_cells;   /* nface*2: faceMap[4..5] = c0, c1; c1 < 0 or >= ncell
                                  marks a boundary face                            */
  const int32_t *face_type;    /* nface: 3 (triangle) or 4 (quad)                  */
  const int32_t *face_tag;     /* nface: zone tag 3/4/5/6                          */
  const uint8_t *face_split;   /* nface: DSMC_SPLIT_* for quads                    */
  const double *face_normal;   /* nface*3: unit, c0 -> c1                          */
  const double *face_area;     /* nface                                            */
  const int32_t *cell_nface;   /* ncell: DsmcCell::num                              */
  const int32_t *cell_faces;   /* ncell*6: DsmcCell::cell2face, -1 padded          */
  const double *cell_volume;   /* ncell: DsmcCell::area                             */
  const double *cell_centroid; /* ncell*3: DsmcCell::cellXY                          */
} dsmc_mesh_view;

/* Host-side mesh object produced by the builders below (owns its arrays). */
typedef struct dsmc_mesh dsmc_mesh;

/* Build a mesh from raw connectivity: face nodes, face cells and tags.
 * Computes face normals/areas/centroids and cell volumes/centroids with the
 * reference's finite-volume formulas (setFaceCellNew + preprocess,
 * src/meshImport.cpp:590-728,742-797), cell->face lists in face order, and
 * the quad split tags (build_face_split_tags_from_faces,
 * src/meshImportDsmcTopology.cpp:106-168). */
int dsmc_mesh_build(int32_t ncell, int32_t nface, int32_t nnode, const double *node_xyz,
                    const int32_t *face_nodes, const int32_t *face_cells,
                    const int32_t *face_type, const int32_t *face_tag, dsmc_mesh **out);
/* Read a Fluent ASCII case mesh (readCAS, src/meshImport.cpp:177-576): nodes
 * (times `scale`, the reference's ScaleFactor), cells and faces; each face
 * zone's id becomes its faces' tag (3 interface, 4 inlet, 5 outlet, 6 wall);
 * then the geometry as dsmc_mesh_build.  Triangle and quad faces only. */
int dsmc_mesh_read_cas(const char *path, double scale, dsmc_mesh **out);
/* Write a mesh in the same format: one node zone, one cell zone (2) and one
 * face zone per run of consecutive faces with the same tag (zone id = tag). */
int dsmc_mesh_write_cas(const dsmc_mesh_view *mesh, const char *path);
/* Unit cube [0,1]^3 of n^3 hexes, each split into 6 tets (Freudenthal), every
 * boundary face tagged `wall_tag`. */
int dsmc_mesh_box_tets(int32_t n, int32_t wall_tag, dsmc_mesh **out);
/* Unit cube [0,1]^3 of n^3 hexes (quad faces), boundary tagged `wall_tag`. */
int dsmc_mesh_box_hexes(int32_t n, int32_t wall_tag, dsmc_mesh **out);
/* Capsule hex O-grid: a closed quad surface of box topology a x a x c on a
 * spherical-cap heat shield + 33 degree cone after-body of diameter
 * `diameter`, extruded in `layers` geometric layers to an outer ellipsoid.
 * Outer faces whose centroid lies on the upstream half are inlet (tag 4), the
 * rest outlet (tag 5); the body is wall (tag 6).  a=25, c=62, layers=50 gives
 * the decoded Apollo topology (372,500 cells, 380,052 nodes, 1,124,950 faces).
 * `tets` != 0 splits every hex into 12 tets around its centroid.  `jitter`
 * (0..0.3) perturbs interior nodes, seeded by `seed`. */
int dsmc_mesh_capsule(int32_t a, int32_t c, int32_t layers, double diameter, int32_t tets,
                      double jitter, uint64_t seed, dsmc_mesh **out);
/* The same mesh with its cells renumbered along a space-filling curve
 * through their centroids (DSMC_ORDER_MORTON: Morton order of the centroids
 * quantised to 2^21 steps over the bounding cube), so that neighbouring
 * cells sit near each other in memory.  The reference renumbers its mesh
 * too, after partitioning (reMesh, src/meshImport.cpp:1760).  Nodes and
 * faces keep their order; every per-cell quantity (volume, centroid, face
 * slots) is that of the original cell, bit for bit.  new_to_old (ncell, may
 * be NULL) receives the original index of each new cell: per-cell outputs of
 * a renumbered run map back to mesh order as out_mesh[new_to_old[c]] =
 * out[c]. */
#define DSMC_ORDER_MORTON 1
int dsmc_mesh_renumber(const dsmc_mesh *m, int32_t order, dsmc_mesh **out, int32_t *new_to_old);
int dsmc_mesh_get_view(const dsmc_mesh *m, dsmc_mesh_view *view);
void dsmc_mesh_destroy(dsmc_mesh *m);

/* ---- case parameters ----------------------------------------------------- */
/* Inputs are the reference's compile-time constants lifted to run time
 * (src/meshImport.h:45-67, src/meshImport.cpp:1275-1332, AoA from
 * src/ProcessDSMC.cpp:2330-2343).  dsmc_params_derive fills the derived
 * block exactly as setMeshMessage/setParticleMessage
 * (src/meshImport.cpp:1273-1425) do. */
typedef struct dsmc_params {
  /* inputs */
  double Ma, Kn, aoa_deg; /* free-stream Mach, Knudsen, angle of attack (x-y plane) */
  double T_in, T_wall;    /* free-stream and wall temperatures [K]              */
  double T_ref;           /* reference temperature [K] (273.15)                 */
  double L0;              /* reference length [m]                               */
  double gamma;           /* inflow speed uses Ma*sqrt(gamma/2)                 */
  double omega, alpha;    /* VHS viscosity index, VSS alpha (1 = VHS)           */
  double d_ref, mass;     /* reference diameter [m], molecular mass [kg]        */
  double P_relax;         /* Larsen-Borgnakke probability (0 = monatomic)       */
  double dr;              /* rotational DOF in outputs (<= 0: monatomic)        */
  int32_t Npinitial;      /* particles per cell at set-up (Neff)                */
  int32_t NSCHEME, NSS;   /* exponential window length, steady start step       */
  int32_t Nevery, Nrepeat;/* interval sampling (NGSIS = Nevery*Nrepeat)         */
  int32_t wall_model;     /* DSMC_WALL_*                                        */
  int32_t precision;      /* DSMC_PREC_*                                        */
  int32_t surface_stats;  /* 1: wall tallies after NSS (reference flag
                             EnableBoundaryStressHeatStatistic)                 */
  int32_t inflow;         /* 1: inject at inlet faces                           */
  /* derived (dsmc_params_derive) */
  double kB, n_ref, v_rms, v_in, mu_ref, p_ref, dt_ref, Twall_ref, eta, p_mass_r;
  double Neff, dtime, volume_total, min_edge;
  double u_inf[3];        /* inflow velocity in v_rms units                     */
  double T_inf;           /* T_in / T_ref                                       */
  double coll_prefactor;  /* pi d^2 (2 kB T_ref / m_r)^eta / Gamma(5/2-omega)  */
} dsmc_params;

/* Apollo case (src/meshImportTest.cpp:124, src/meshImport.cpp:1315-1332):
 * Ma 5, Kn 0.01, AoA 30, T_in 142.2 K, T_w 300 K, L0 3.912 m, N2-like gas.
 * monatomic != 0 selects the build's monatomic VHS mode (P_relax 0,
 * gamma 5/3, dr 0). */
int dsmc_params_apollo(dsmc_params *p, int32_t monatomic);
/* Equilibrium box: hard spheres (omega 0.5), T = 273.15 K, L0 = 1 m,
 * specular walls, no inflow. */
int dsmc_params_box(dsmc_params *p, double Kn);
int dsmc_params_derive(const dsmc_mesh_view *mesh, dsmc_params *p);

/* ---- particles ----------------------------------------------------------- */
/* Structure-of-arrays particle exchange record.  The reference's per-cell
 * buckets (ParticleBucketSoA, src/MeshparticalInitial.h:40-186) flattened:
 * `cell` replaces p_mesh_serial; `id` is a stable identity (the reference's
 * p_serial is renumbered every step) used to key wall random draws. */
typedef struct dsmc_particles_soa {
  int64_t n;
  double *x, *y, *z;
  double *vx, *vy, *vz;
  double *ir;
  int32_t *cell;
  uint32_t *id;
} dsmc_particles_soa;

/* Initial fill: Npc per cell proportional to volume with random rounding
 * (src/meshImport.cpp:1409-1421), positions uniform over the cell's tet
 * decomposition (src/MeshparticalInitial.cpp:347-361,481-721), a drifting
 * Maxwellian at the free-stream state (u_inf, T_inf; T = `temperature` if
 * > 0) and Ir ~ Exp(T/2) (0 when monatomic).  Call with soa->n = 0 and null
 * arrays to get the count in *count; then with arrays of that size. */
int dsmc_fill_freestream(const dsmc_mesh_view *mesh, const dsmc_params *p, double ppc,
                         double temperature, uint64_t seed, dsmc_particles_soa *soa,
                         int64_t *count);

/* ---- engine --------------------------------------------------------------- */
typedef struct dsmc_engine dsmc_engine;

/* Stage bits for dsmc_step, in the reference driver's order. */
#define DSMC_STAGE_EMIT 1u    /* preprocesseffquad  (src/ProcessDSMC.cpp:2556)  */
#define DSMC_STAGE_MOVE 2u    /* advection          (src/ProcessDSMC.cpp:1946)  */
#define DSMC_STAGE_SORT 4u    /* cache2chain        (src/ProcessDSMC.cpp:2030)  */
#define DSMC_STAGE_COLLIDE 8u /* collisionDSMC      (src/ProcessDSMC.cpp:2621)  */
#define DSMC_STAGE_SAMPLE 16u /* statistic_macro    (src/ProcessDSMC.cpp:2757)  */
#define DSMC_STAGE_ALL 31u

/* Create an engine on the current HIP device: copies the mesh, builds the
 * per-cell triangle cache (rebuildFaceTraversalCache,
 * src/ProcessDSMCGeometry.cpp:145-214), the face crossing table
 * (src/ProcessDSMCGeometry.cpp:114-136), wall normals (processFacesQuadNormals,
 * src/ProcessDSMC.cpp:1054-1128) and inflow patches (rebuildBoundaryEmitCache,
 * src/ProcessDSMC.cpp:583-768).  `capacity` bounds the particle count
 * (0: 1.5 x Npinitial x ncell + inflow headroom).  `p` must be derived. */
int dsmc_engine_create(const dsmc_mesh_view *mesh, const dsmc_params *p, uint64_t seed,
                       int64_t capacity, dsmc_engine **out);
void dsmc_engine_destroy(dsmc_engine *e);

/* Copy particles in (any order; they are regrouped by cell, keeping their
 * order inside a cell) / out (cell order).  Upload resets per-cell collision
 * state (crmax = 3.0, remainders 0 — Collision_constant_initial,
 * src/MeshparticalInitial.cpp:798-810) and all accumulators.  Download with
 * soa->n smaller than the particle count returns DSMC_ERR_CAPACITY and the
 * needed size in *count. */
int dsmc_particles_upload(dsmc_engine *e, const dsmc_particles_soa *soa);
/* Initial fill generated on the device (no reference counterpart; replaces
 * dsmc_fill_freestream + dsmc_particles_upload for cases too large to fill
 * on the host): for every case k, bit-identically the particles that
 * dsmc_fill_freestream(case k's mesh, its params, ppc[k], temperature[k],
 * seeds[k]) would produce, placed and state reset as an upload does.
 * temperature may be null (free-stream temperature). */
int dsmc_fill_device(dsmc_engine *e, const double *ppc, const double *temperature, const uint64_t *seeds,
                     int64_t *count);
int dsmc_particles_download(dsmc_engine *e, dsmc_particles_soa *soa, int64_t *count);
int dsmc_particle_count(dsmc_engine *e, int64_t *count);

/* One time step: the stages in `stage_mask`, in the order emit, move, sort,
 * collide, sample.  A stage that needs the particles grouped by cell (emit
 * appends, collide/sample read cells) requires the previous stages to have
 * run in order; DSMC_STAGE_ALL is one full reference step. */
int dsmc_step(dsmc_engine *e, int32_t istep, uint32_t stage_mask);
/* Run steps first..last (inclusive) with DSMC_STAGE_ALL, asynchronously on
 * the engine's stream; synchronises at the end. */
int dsmc_run(dsmc_engine *e, int32_t first, int32_t last);
/* Asynchronous dsmc_run (no reference counterpart; for callers that drive
 * several engines at once): enqueues the steps and returns; dsmc_run_query
 * sets *done once they have completed (then reporting their errors). */
int dsmc_run_async(dsmc_engine *e, int32_t first, int32_t last);
int dsmc_run_query(dsmc_engine *e, int32_t *done);
/* dsmc_run_async ordered on a caller's HIP stream (`stream` is a hipStream_t;
 * no reference counterpart): the steps start after the work already enqueued
 * on `stream`, run on the engine's own streams (and step graphs), and work
 * enqueued on `stream` afterwards starts after them.  Returns without
 * synchronising; dsmc_run_query reports completion and errors.  A stream
 * under graph capture is refused: the engine replays its own step graphs. */
int dsmc_run_on_stream(dsmc_engine *e, int32_t first, int32_t last, void *stream);
/* Move only (the minimum slice): advection of every particle for one dtime
 * (src/ProcessDSMC.cpp:1946-2017) followed by the regroup. */
int dsmc_track(dsmc_engine *e, int32_t istep);

/* Crossing log for tracking parity: when enabled, every face crossing or wall
 * hit appends (id, step, sequence#, face, triangle tag, kind).  kind: 1
 * interface, 2 inlet/outlet exit, 3 wall, 4 stop inside cell, 5 dropped. */
typedef struct dsmc_cross_event {
  uint32_t id;
  int32_t step;
  int32_t seq;
  int32_t face;
  int32_t tri;
  int32_t kind;
} dsmc_cross_event;
int dsmc_track_log_enable(dsmc_engine *e, int64_t max_events);
int dsmc_track_log_get(dsmc_engine *e, dsmc_cross_event *out, int64_t cap, int64_t *count);

/* Macro fields, ncell*18 doubles (MacroIndex order).  which:
 *   DSMC_MACRO_LOCAL  exponential window  ("local",        src/ProcessDSMC.cpp:2792)
 *   DSMC_MACRO_RECORD interval record     ("record",       :2849-2856)
 *   DSMC_MACRO_FINAL  steady average      ("final_record", :2818-2819)
 *   DSMC_MACRO_OUTPUT what out2dat writes (record before NSS or for sparse
 *                     cells, else final_record, :3357-3368) */
#define DSMC_MACRO_LOCAL 0
#define DSMC_MACRO_RECORD 1
#define DSMC_MACRO_FINAL 2
#define DSMC_MACRO_OUTPUT 3
int dsmc_macro_get(dsmc_engine *e, int32_t which, double *out);
/* statistic_macroPre (src/ProcessDSMC.cpp:3312-3326): local and record from
 * the current particles. */
int dsmc_statistic_pre(dsmc_engine *e);
/* Raw per-cell accumulators: family 0 stepsum, 1 steady, 2 stepinter;
 * ncell*18 doubles in the order rho(N), Sum v (3), Sum|v|^2, Sum v_i v_j
 * (xx,xy,yy,yz,zz,zx), Sum v_i|v|^2 (3), Sum Ir, Sum v_i Ir (3). */
int dsmc_accum_get(dsmc_engine *e, int32_t family, double *out);
/* Per-cell scalars, ncell each (null pointers skipped): particle count,
 * crmax, collision remainder, inflow remainder, sparse-cell state,
 * sparse accumulation steps. */
int dsmc_cell_scalars_get(dsmc_engine *e, int32_t *count, double *crmax, double *rem_coll,
                          double *rem_emit, uint8_t *sparse_state, int32_t *sparse_steps);

/* Wall tallies, nface*6 doubles: hits, momentum change (3), translational and
 * rotational energy change (src/ProcessDSMC.cpp:1659-1689); zero for
 * non-wall faces. */
int dsmc_surface_get(dsmc_engine *e, double *out);
/* Wall output (outBoundaryStressHeat, src/ProcessDSMC.cpp:3494-3557) from
 * the tallies at step `istep`: nface*12 doubles, per wall face stress(3),
 * q_trans, q_rot, normal(3), pressure, shear(3) (zero for other faces);
 * plus the drag coefficient (force on the wall along the free-stream
 * direction over q_inf * pi (D/2)^2) and the peak of q_trans + q_rot. */
#define DSMC_SURF_WIDTH 12
int dsmc_surface_fields(dsmc_engine *e, int32_t istep, double ref_diameter, double *fields,
                        double *drag_coefficient, double *peak_heat_flux);

typedef struct dsmc_stats {
  int64_t particles;     /* current count                                   */
  int64_t emitted;       /* cumulative inflow particles                     */
  int64_t removed;       /* cumulative inlet/outlet exits                   */
  int64_t dropped;       /* cumulative tracking failures (no exit face)     */
  int64_t wall_hits;     /* cumulative wall reflections                     */
  int64_t crossings;     /* cumulative interface crossings                  */
  int64_t coll_pairs;    /* cumulative NTC candidate pairs                  */
  int64_t collisions;    /* cumulative accepted collisions                  */
  int64_t particle_steps;/* sum over steps of the count at step start       */
  int32_t last_step;
  int32_t pad;
} dsmc_stats;
int dsmc_stats_get(dsmc_engine *e, dsmc_stats *s);

/* Per-kernel timing with HIP events on the engine's stream. */
#define DSMC_NKERNEL 8
typedef struct dsmc_kernel_times {
  double ms[DSMC_NKERNEL];    /* accumulated milliseconds per kernel group */
  int64_t calls[DSMC_NKERNEL];
} dsmc_kernel_times; /* groups: 0 emit, 1 move, 2 sort-count, 3 sort-scan,
                        4 sort-place, 5 sort-gather, 6 collide+sample, 7 other */
int dsmc_profile_enable(dsmc_engine *e, int32_t on);
int dsmc_profile_get(dsmc_engine *e, dsmc_kernel_times *t);
/* Wall-clock (HIP events around the whole step sequence) of the last
 * dsmc_run call, milliseconds. */
int dsmc_last_run_ms(dsmc_engine *e, double *ms);

/* ---- batches (no reference counterpart: the reference runs one case per
 *      job).  Several cases — each its own mesh, parameters and seed — run in
 *      one engine: every kernel launch of a step covers all of them.  A case's
 *      results are bit-identical to running it alone with the same seed (its
 *      random streams are keyed by its own cell, patch and particle numbers).
 *      The batch's cells are the cases' cells in order, each case padded with
 *      empty cells to a multiple of 64; dsmc_case_info gives the ranges to
 *      slice cell and face outputs, and particles uploaded to a batch use
 *      these cell numbers (ids stay per case). */
int dsmc_batch_create(int32_t ncase, const dsmc_mesh_view *meshes, const dsmc_params *params,
                      const uint64_t *seeds, int64_t capacity, dsmc_engine **out);
typedef struct dsmc_case_info_t {
  int32_t ncase;
  int32_t cell_begin, cell_end; /* the case's own cells in the batch */
  int32_t face_begin, face_end; /* the case's faces in the batch */
  int64_t particles;            /* particles of the case now */
  int64_t particle_steps;       /* sum over steps of the case's count at step start */
} dsmc_case_info_t;
int dsmc_case_info(dsmc_engine *e, int32_t kase, dsmc_case_info_t *info);
/* dsmc_surface_fields for one case of a batch (istep: the case's step). */
int dsmc_case_surface_fields(dsmc_engine *e, int32_t kase, int32_t istep, double ref_diameter, double *fields,
                             double *drag_coefficient, double *peak_heat_flux);

/* Upper bound of the device bytes dsmc_engine_create(mesh, p, seed,
 * capacity) allocates (capacity > 0). */
int dsmc_engine_estimate(const dsmc_mesh_view *mesh, const dsmc_params *p, int64_t capacity, int64_t *bytes);

/* ---- serving (no reference counterpart) ------------------------------------
 * A queue of case requests served with continuous batching: requests are
 * admitted in arrival order while their estimated footprint fits the HBM
 * budget (<= 0: 85% of the free device memory); admitted requests run in
 * engines on their own streams, advanced `chunk_steps` at a time with
 * asynchronous runs, so they overlap on the GPU.  By default each request
 * gets its own engine; with dsmc_server_set_batching(k) the queued requests
 * admitted together (up to k, consecutive in arrival order, same step count)
 * share one batch engine.  A finished request frees its memory (auto_release;
 * a batch engine goes with its last request) and queued ones are admitted.
 * Results are bit-identical to the same case run alone.  dsmc_server_pump
 * never blocks: call it repeatedly until every request has finished.
 * dsmc_server_submit takes ownership of `mesh`; capacity <= 0 picks
 * 1.25 x ppc x cells + 2M.  Times are seconds since dsmc_server_create. */
typedef struct dsmc_server dsmc_server;
#define DSMC_REQ_QUEUED 0
#define DSMC_REQ_RUNNING 1
#define DSMC_REQ_DONE 2
#define DSMC_REQ_FAILED 3
#define DSMC_REQ_RELEASED 4
typedef struct dsmc_request_info {
  int32_t state, steps, steps_done;
  int32_t kase; /* the request's case in its engine (dsmc_case_info) */
  double t_submit, t_admit, t_finish;
  int64_t estimate, footprint, particles, particle_steps;
  char error[256];
} dsmc_request_info;
int dsmc_server_create(int64_t hbm_budget_bytes, int32_t chunk_steps, int32_t auto_release, dsmc_server **out);
void dsmc_server_destroy(dsmc_server *s);
int dsmc_server_submit(dsmc_server *s, dsmc_mesh *mesh, const dsmc_params *p, uint64_t seed, double ppc,
                       int32_t steps, int64_t capacity, int64_t *request_id);
int dsmc_server_pump(dsmc_server *s, int32_t *running, int32_t *queued, int32_t *finished);
/* At most max_running requests run at once (0: no limit, the default); the
 * rest wait in the queue even when their footprint fits.  Engines that run
 * at the same time share the device's hardware queues, so a cap near their
 * number keeps each engine's kernels from waiting behind the others'. */
int dsmc_server_set_max_running(dsmc_server *s, int32_t max_running);
/* Up to max_cases requests per engine (1, the default: one engine each).
 * Requests queued at the same admission, consecutive in arrival order and
 * with the same step count, then share one batch engine (dsmc_batch_create):
 * fewer, fuller launches than one engine per request.  The running cap still
 * counts requests. */
int dsmc_server_set_batching(dsmc_server *s, int32_t max_cases);
int dsmc_server_request(dsmc_server *s, int64_t id, dsmc_request_info *info);
/* The engine of a finished (not released) request, to read its results; a
 * batched request's results are its case (dsmc_request_info.kase) of that
 * engine. */
int dsmc_server_engine(dsmc_server *s, int64_t id, dsmc_engine **engine);
/* Frees a finished request's engine share (state RELEASED).  A request still
 * queued is cancelled: it leaves the queue, its mesh is freed and its state
 * becomes FAILED with the error "cancelled before admission"; requests
 * admitted later are not affected.  A running request is refused
 * (DSMC_ERR_STATE). */
int dsmc_server_release(dsmc_server *s, int64_t id);

/* ---- Fluent-format output files -----------------------------------------------
 * The cell fields file (meshImport::out2Fluent_dsmc, src/meshImport.cpp:
 * 1107-1182; the reference calls it from out2dat, src/ProcessDSMC.cpp:3421-
 * 3424, with the gathered ncell x 18 MacroIndex array `w`): records 700
 * density, 701 velocity, 702 T and T_r, 710 stress, 711 and 712 heat fluxes,
 * scaled to free-stream units, for cell zone `zone_id` with Fluent ids
 * first_id .. first_id + ncell - 1 (the Apollo case: zone 2, first 1). */
int dsmc_write_fluent_cells(const char *path, const dsmc_params *p, int32_t ncell, const double *w,
                            int32_t zone_id, int32_t first_id);
/* The wall fields file (outBoundaryStressHeat's writer, src/ProcessDSMC.cpp:
 * 3590-3652) from dsmc_surface_fields' output: one block per run of
 * consecutive wall faces (zones zone_id, zone_id + 1, ...; Fluent face ids =
 * face index + first_id). */
int dsmc_write_fluent_walls(const char *path, const dsmc_mesh_view *mesh, const double *fields, int32_t zone_id,
                            int32_t first_id);

/* ---- continuum coupling (GSIS) ---------------------------------------------
 * Device-side molecular_velocity_change (src/ProcessGSIS.cpp:2239-2345) with
 * replication (:2368-2410), MaxwellianSample (:2424-2446), randomlocation
 * (:2459-2474) and deletion (:2487-2528): replaces the reference's update of
 * the particle buckets from the NS solution.  ns: ncell x 6 per cell = density
 * / n_ref, mean velocity (3, v_rms units), T and T_r (T_ref units), as
 * nsresult.  accepted: ncell flags (the DSMC2NS acceptance gate), or NULL for
 * every cell.  For each coupled cell, in cell order:
 *   - the count becomes Iround(rho V n_ref / Neff) by appending copies of
 *     random particles at new random points of the cell (Maxwellian samples
 *     of the NS state if the cell was empty and T >= 0), or by deleting random
 *     particles (swap with the last);
 *   - velocities become k v + b and Ir becomes kr Ir with k = sqrt(T/T_cell),
 *     b = u - k u_cell, kr = T_r/T_r,cell when the cell's T, T_r and the NS T
 *     are positive (else unchanged).
 * A NaN in the first four NS fields of a coupled cell stops the update there
 * (that cell and later cells stay as they are, as in the reference);
 * *nan_cell receives that cell or -1.  istep (> 0) keys the random streams.
 * New particles take the next ids of their case.  Call between steps. */
int dsmc_ns_apply(dsmc_engine *e, int32_t istep, const double *ns, const uint8_t *accepted, int32_t *nan_cell);
/* dsmc_ns_apply that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  ns_device (ncell x 6 doubles)
 * and accepted_device (ncell bytes, or NULL) are device memory of the
 * engine's device; results are those of dsmc_ns_apply on the same inputs,
 * bit for bit.  Its scratch (the call's counters and the replication
 * offsets) is a fixed 16 MiB array per device, held until the process ends,
 * so engines of at most 4,194,304 cells are accepted (DSMC_ERR_CAPACITY
 * beyond).  Every call on one device shares that array: calls that overlap
 * on one device race (finish one, by its return, before starting another).
 * The call synchronises the engine's stream, so it cannot be graph-captured. */
int dsmc_ns_apply_noalloc(dsmc_engine *e, int32_t istep, const double *ns_device, const uint8_t *accepted_device,
                          int32_t *nan_cell);

/* Bytes of device memory the engine holds. */
int dsmc_engine_footprint(dsmc_engine *e, int64_t *bytes);

/* Launch tuning (no reference counterpart: the reference runs one CPU thread
 * per rank).  None of these change results, which are bit-identical for every
 * setting; they only move work between kernels.
 *   wide_min   cells with more particles than this are collided and sampled
 *              by one wavefront per cell instead of one thread; < 0: never
 *              (default 128)
 *   wide_grid  workgroups of that one-wave-per-cell kernel (default 1024)
 *   move_chunk particles per workgroup of the move kernel, whose lanes take
 *              the chunk's particles one after another (64..1048576,
 *              default 4096)
 *   graph      1: dsmc_run replays pairs of steps captured as a HIP graph
 *              (steps numbered on the device; not while profiling or
 *              logging); 0: every step is launched kernel by kernel
 *              (default 1)
 *   sort_map   1: the regroup's tiles are cut into one contiguous range per
 *              XCD, so that tiles of neighbouring cells share an L2; 0:
 *              tile = workgroup index (default 0)
 *   coll_block threads per workgroup of the thread-per-cell collide+sample
 *              kernel (one cell per thread): 64, 128 or 256 (default 64)
 * Fields set to DSMC_TUNE_KEEP are left unchanged. */
#define DSMC_TUNE_KEEP (-2147483647 - 1)
typedef struct dsmc_tuning {
  int32_t wide_min;
  int32_t wide_grid;
  int32_t move_chunk;
  int32_t graph;
  int32_t sort_map;
  int32_t coll_block;
  int32_t reserved[2];
} dsmc_tuning;
int dsmc_engine_set_tuning(dsmc_engine *e, const dsmc_tuning *t);
int dsmc_engine_get_tuning(dsmc_engine *e, dsmc_tuning *t);
/* Serving (above): launch settings for every engine the server builds from
 * now on (dsmc_engine_set_tuning's fields; DSMC_TUNE_KEEP keeps a default),
 * e.g. the autotuner's pick for a representative case of the stream; null
 * restores the engine defaults.  Engines already running keep theirs.  No
 * reference counterpart. */
int dsmc_server_set_tuning(dsmc_server *s, const dsmc_tuning *t);
/* Autotuner (no reference counterpart): picks move_chunk, wide_min,
 * wide_grid, coll_block and sort_map for this case by timing the kernels on live steps first_step,
 * first_step + 1, ... (the tuning never changes results, so these are
 * ordinary steps of the run), sets the best and returns it.  *steps_used
 * gets the number of steps run; continue the run from first_step +
 * *steps_used.  steps_per_trial: 1..100 (2 is enough for cases of a few
 * million particles or more). */
int dsmc_autotune(dsmc_engine *e, int32_t first_step, int32_t steps_per_trial, dsmc_tuning *best,
                  int32_t *steps_used);

/* Diagnostic: evaluates on the device, for each input u in (0, 1], the
 * shared deterministic functions the kernels use — out[6i..6i+5] =
 * log(u), exp(-20u), sin(2 pi u), cos(2 pi u), sqrt(u), u^0.26.  Lets a test
 * check that device and host arithmetic agree bit for bit.  No counterpart
 * in the reference. */
int dsmc_device_math_check(const double *u, int64_t n, double *out);
/* Diagnostic: the collision kernels' FP32 estimate of the NTC acceptance
 * ratio coll_diam(x, pref, eta) / crm (0 where it is not used) and the
 * deterministic ratio itself, for n speeds x (out: n x 2).  The kernels
 * decide u < ratio from the estimate only when u is outside a relative band
 * of 2^-11 around it; this checks the estimate's error against that band.
 * No counterpart in the reference. */
int dsmc_device_accept_check(const double *x, int64_t n, double pref, double eta, double crm, double *out);
/* Diagnostic: device-to-device copy of `bytes` with a plain streaming kernel
 * (k_copy_probe), `reps` times; returns the achieved GB/s (read + write).
 * Used to calibrate HBM traffic counters and the achievable bandwidth.  No
 * counterpart in the reference. */
int dsmc_copy_probe(int64_t bytes, int32_t reps, double *gbs);

const char *dsmc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* DSMC_ENGINE_H */
