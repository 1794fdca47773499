// Device-side data layout and kernels of the DSMC engine (gfx950).
//
// Particle state is structure-of-arrays in HBM, grouped by cell after every
// step: positions in FP64, velocities and Ir in VT (double, or float for the
// mixed-precision engine), cell (int32) and a stable id (uint32).  Two copies
// of every array exist.  The regroup gathers the velocity records from one
// into the other; the position records stay where the move wrote them, and
// the next move reads each through the regroup's permutation and writes it
// back in cell order.
//
// All kernels read the live particle count from device memory, so a whole
// step (and a whole run) is enqueued without host round trips.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#include "dsmc_detmath.h"
#include "dsmc_rng.h"

#pragma clang fp contract(off)

namespace dsmc {

// One triangle of a cell's face cache (FaceTriCache, src/ProcessDSMC.h:109-116),
// 80 bytes: vertex A, edges e1 = B - A, e2 = C - A, and two packed words so
// that the walk decides what a crossing does without reading the face table:
//   face_tag = face | tagcode << 29   (tagcode 0: triangle face, tag -1;
//                                      1, 2: half 1 or 2 of a split quad)
//   link     = the cell across the face, if the face is an interface with a
//              cell on the other side; otherwise -1 - role (role != interface)
//              or LINK_OPEN_INTERFACE (interface tag on a boundary face)
struct TriRec {
  double ax, ay, az, e1x, e1y, e1z, e2x, e2y, e2z;
  int32_t face_tag, link;
};

enum FaceRole : int32_t { ROLE_NONE = 0, ROLE_INTERFACE = 1, ROLE_INLET = 2, ROLE_OUTLET = 3, ROLE_WALL = 4 };

constexpr int32_t TRI_FACE_BITS = 29;
constexpr int32_t TRI_FACE_MASK = (1 << TRI_FACE_BITS) - 1;
constexpr int32_t LINK_OPEN_INTERFACE = -1 - ROLE_INTERFACE;

inline __host__ __device__ int32_t tri_pack(int32_t face, int32_t tag) {
  return face | ((tag < 0 ? 0 : tag) << TRI_FACE_BITS);
}
inline __host__ __device__ int32_t tri_face(const TriRec& t) { return t.face_tag & TRI_FACE_MASK; }
inline __host__ __device__ int32_t tri_tag(const TriRec& t) {
  const int32_t c = (int32_t)((uint32_t)t.face_tag >> TRI_FACE_BITS);
  return c == 0 ? -1 : c;
}

// c0/c1 = the two cells (c1 = -1 on boundaries), role, wall slot (-1 if not a wall).
struct FaceRec {
  int32_t c0, c1, role, wall;
};

// Inflow triangle patch (BoundaryEmitPatch + BoundaryGasState,
// src/ProcessDSMC.h:157-205), gas state frozen at set-up.
struct Patch {
  double p0[3], e10[3], e20[3];
  double R[9];  // rotation built from -n, row major
  double stp1, stp2, snp, fs1, fs2, tN, sqrtT, QA;
  double flux;  // inserted particles per step before the remainder (add_num / Neff)
  int32_t cell, face, tag, kase;  // kase: the case the patch belongs to
};

// Live counters and statistics (device memory).
// Per-step event tallies.  Workgroup b adds into slot b % TALLY_SLOTS, each
// slot on its own 128-byte line: device atomics on one line serialise
// (about 12 ns apiece on MI355X), so tens of thousands of workgroups adding
// into one counter would queue behind each other.  Readers sum the slots.
constexpr int TALLY_SLOTS = 64;
struct alignas(128) TallySlot {
  unsigned long long removed, dropped, wall_hits, crossings, coll_pairs, collisions;
};

struct Counters {
  int32_t n;         // particles grouped by cell at step start
  int32_t n_total;   // n + emitted this step
  int32_t n_emit;    // emitted this step
  int32_t overflow;  // capacity exceeded (sticky)
  int32_t last_step;   // engine step being run (set by the step-begin kernel)
  int32_t park_count;  // particles parked at a wall by the move kernel this step
  int32_t wide_n;      // cells queued for the one-wave-per-cell kernel from the front (the largest)
  int32_t wide_back;   // cells queued for it from the back of the queue (the others)
  int32_t arr_spill;   // a cell got more arrivals this regroup than its arrival list holds
  unsigned long long emitted, particle_steps;
  TallySlot tally[TALLY_SLOTS];
};

// Physics and bookkeeping of one case (one entry per case of a batch; a
// single-case engine has one).  `ncell` (the accumulator stride) and
// `max_iter` are the engine's and repeated in every entry so that device code
// needs only this record.
struct CaseParams {
  uint64_t key;
  double dtime, Tw, sqrtTw, Neff, dt_ref, coll_prefactor, v_rms, omega, alpha, P_relax, eta;
  double inv_lb;  // 1 / (2.5 - omega)
  double inv_alpha;
  double dr, n_ref;
  int32_t wall_model, rot, surface, NSS, NSCHEME, Nevery, Nrepeat, ncell, max_iter;
  int32_t step0;           // engine steps before the case's step 1 (its step = engine step - step0)
  int32_t cell0, cell1;    // the case's cells [cell0, cell1)
  int32_t patch0, patch1;  // the case's inflow patches [patch0, patch1)
};

// Launch-wide parameters.
struct KParams {
  int32_t ncell, capacity, npatch, ninlet_cells, max_iter, ncase;
  int32_t park_cap;  // entries of the wall queue
  int32_t wide_min;    // cells with more particles go one wave per cell (< 0: never)
  int32_t move_chunk;  // particles per move workgroup
  const CaseParams* __restrict__ cases;
  const uint16_t* __restrict__ cell_case;  // case of every cell
};

struct CrossLog {
  int32_t* ev;  // 6 ints per event
  int32_t cap;
  int32_t* count;
};

// Particle records: position + cell + id (32 B) and velocity + Ir (32 B FP64,
// 16 B mixed).  A record is read or written with two (one) 16-byte loads, so
// a particle moves as two contiguous pieces instead of nine scattered
// scalars, and collide/sample reads only the velocity records.
struct alignas(32) PosRec {
  double x, y, z;
  int32_t cell;
  uint32_t id;
};

template <class VT>
struct alignas(4 * sizeof(VT)) VelRec {
  VT vx, vy, vz, ir;
};

// A walk stopped at a wall by the move kernel, finished by the wall kernel.
struct alignas(16) Park {
  double dtleft;
  int32_t i, cell, crossfid, tri, seq, iter, hits, ncross;
};

template <class VT>
struct PBuf {
  PosRec* pos;
  VelRec<VT>* vel;
};

// Accumulator families (stepsum, steady, stepinter), structure of arrays
// [family][field][cell], 18 fields each.
constexpr int NMOM = 18;

}  // namespace dsmc
