// Serving loop with continuous batching (no reference counterpart: the
// reference solves one case per MPI job).  Requests — a case mesh, its
// parameters, seed, particles per cell and step count — are queued in arrival
// order and admitted while their estimated device footprint fits the HBM
// budget; each admitted request gets its own engine (own HIP stream, state
// resident on the device) and advances in chunks of steps enqueued
// asynchronously, so the admitted requests' kernels overlap on the GPU.  A
// finished request frees its memory (when auto-release is on) and the next
// queued requests are admitted.  Engines share nothing, so every result is
// bit-identical to the same case run alone.
//
// dsmc_server_pump() is non-blocking: the caller calls it in its own loop
// (between submissions) until every request has finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "dsmc_engine.h"
#include "dsmc_internal.h"

using dsmc::set_error;

namespace {

enum ReqState { REQ_QUEUED = 0, REQ_RUNNING = 1, REQ_DONE = 2, REQ_FAILED = 3, REQ_RELEASED = 4 };

struct Request {
  int64_t id = 0;
  dsmc_mesh* mesh = nullptr;  // owned until admission
  dsmc_params p{};
  uint64_t seed = 0;
  double ppc = 0.0;
  int32_t steps = 0;
  int64_t capacity = 0;
  int64_t estimate = 0;   // bytes reserved at admission
  bool reserved = false;  // estimate counted in the server's in_use
  int64_t footprint = 0;  // bytes the engine holds
  dsmc_engine* eng = nullptr;
  int state = REQ_QUEUED;
  int32_t steps_done = 0, steps_enqueued = 0;
  double t_submit = 0.0, t_admit = 0.0, t_finish = 0.0;
  int64_t particles = 0, particle_steps = 0;
  std::string error;
};

}  // namespace

struct dsmc_server {
  int64_t budget = 0;
  int64_t in_use = 0;
  int32_t chunk = 50;
  int32_t auto_release = 1;
  int32_t max_running = 0;  // 0: no limit
  std::vector<Request> reqs;
  std::deque<size_t> queue;  // indices of queued requests, arrival order
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

Request* find(dsmc_server* s, int64_t id) {
  if (!s || id < 0 || id >= (int64_t)s->reqs.size()) return nullptr;
  return &s->reqs[(size_t)id];
}

void release(dsmc_server* s, Request& r) {
  if (r.eng) {
    dsmc_engine_destroy(r.eng);
    r.eng = nullptr;
  }
  if (r.reserved) {
    s->in_use -= r.estimate;
    r.reserved = false;
  }
  if (r.mesh) {
    dsmc_mesh_destroy(r.mesh);
    r.mesh = nullptr;
  }
}

void fail(dsmc_server* s, Request& r) {
  r.error = dsmc_last_error();
  r.state = REQ_FAILED;
  r.t_finish = s->now();
  release(s, r);
}

int enqueue_chunk(Request& r, int32_t chunk) {
  const int32_t first = r.steps_done + 1;
  const int32_t last = std::min(r.steps, r.steps_done + chunk);
  r.steps_enqueued = last - r.steps_done;
  return dsmc_run_async(r.eng, first, last);
}

}  // namespace

extern "C" {

int dsmc_server_create(int64_t hbm_budget_bytes, int32_t chunk_steps, int32_t auto_release, dsmc_server** out) {
  if (!out || chunk_steps < 1) return set_error(DSMC_ERR_ARG, "dsmc_server_create: bad arguments");
  dsmc_server* s = new dsmc_server();
  if (hbm_budget_bytes <= 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      delete s;
      return set_error(DSMC_ERR_HIP, "dsmc_server_create: cannot read device memory");
    }
    hbm_budget_bytes = (int64_t)(0.85 * (double)free_b);
  }
  s->budget = hbm_budget_bytes;
  s->chunk = chunk_steps;
  s->auto_release = auto_release;
  *out = s;
  return DSMC_OK;
}

void dsmc_server_destroy(dsmc_server* s) {
  if (!s) return;
  for (Request& r : s->reqs) release(s, r);
  delete s;
}

int dsmc_server_submit(dsmc_server* s, dsmc_mesh* mesh, const dsmc_params* p, uint64_t seed, double ppc,
                       int32_t steps, int64_t capacity, int64_t* request_id) {
  if (!s || !mesh || !p || !request_id || steps < 1 || !(ppc > 0.0))
    return set_error(DSMC_ERR_ARG, "dsmc_server_submit: bad arguments");
  Request r;
  r.id = (int64_t)s->reqs.size();
  r.mesh = mesh;
  r.p = *p;
  r.seed = seed;
  r.ppc = ppc;
  r.steps = steps;
  dsmc_mesh_view v;
  int rc = dsmc_mesh_get_view(mesh, &v);
  if (rc) return rc;
  r.capacity = capacity > 0 ? capacity : (int64_t)(1.25 * ppc * v.ncell) + 2000000;
  rc = dsmc_engine_estimate(&v, p, r.capacity, &r.estimate);
  if (rc) return rc;
  if (r.estimate > s->budget) return set_error(DSMC_ERR_CAPACITY, "dsmc_server_submit: request larger than the budget");
  r.t_submit = s->now();
  s->reqs.push_back(r);
  s->queue.push_back(s->reqs.size() - 1);
  *request_id = r.id;
  return DSMC_OK;
}

int dsmc_server_pump(dsmc_server* s, int32_t* running, int32_t* queued, int32_t* finished) {
  if (!s) return set_error(DSMC_ERR_ARG, "dsmc_server_pump: null");
  // progress of the running requests
  for (Request& r : s->reqs) {
    if (r.state != REQ_RUNNING) continue;
    int32_t done = 0;
    if (dsmc_run_query(r.eng, &done)) {
      fail(s, r);
      continue;
    }
    if (!done) continue;
    r.steps_done += r.steps_enqueued;
    r.steps_enqueued = 0;
    if (r.steps_done >= r.steps) {
      dsmc_stats st;
      if (dsmc_stats_get(r.eng, &st)) {
        fail(s, r);
        continue;
      }
      r.particle_steps = st.particle_steps;
      r.particles = st.particles;
      r.state = REQ_DONE;
      r.t_finish = s->now();
      if (s->auto_release) {
        release(s, r);
        r.state = REQ_RELEASED;
      }
    } else if (enqueue_chunk(r, s->chunk)) {
      fail(s, r);
    }
  }
  // admission in arrival order while the next request fits (and the
  // running-request cap allows)
  int32_t nrun = 0;
  for (const Request& r : s->reqs) nrun += (r.state == REQ_RUNNING);
  while (!s->queue.empty()) {
    Request& r = s->reqs[s->queue.front()];
    if (s->in_use + r.estimate > s->budget) break;
    if (s->max_running > 0 && nrun >= s->max_running) break;
    ++nrun;
    s->queue.pop_front();
    s->in_use += r.estimate;
    r.reserved = true;
    r.t_admit = s->now();
    r.state = REQ_RUNNING;
    dsmc_mesh_view v;
    int rc = dsmc_mesh_get_view(r.mesh, &v);
    if (!rc) rc = dsmc_engine_create(&v, &r.p, r.seed, r.capacity, &r.eng);
    if (!rc) {
      dsmc_mesh_destroy(r.mesh);
      r.mesh = nullptr;
      rc = dsmc_fill_device(r.eng, &r.ppc, nullptr, &r.seed, &r.particles);
    }
    if (!rc) rc = dsmc_engine_footprint(r.eng, &r.footprint);
    if (!rc) rc = enqueue_chunk(r, s->chunk);
    if (rc) fail(s, r);
  }
  int32_t nr = 0, nq = 0, nf = 0;
  for (const Request& r : s->reqs) {
    if (r.state == REQ_RUNNING) ++nr;
    else if (r.state == REQ_QUEUED) ++nq;
    else ++nf;
  }
  if (running) *running = nr;
  if (queued) *queued = nq;
  if (finished) *finished = nf;
  return DSMC_OK;
}

int dsmc_server_set_max_running(dsmc_server* s, int32_t max_running) {
  if (!s || max_running < 0) return set_error(DSMC_ERR_ARG, "dsmc_server_set_max_running: bad arguments");
  s->max_running = max_running;
  return DSMC_OK;
}

int dsmc_server_request(dsmc_server* s, int64_t id, dsmc_request_info* info) {
  Request* r = find(s, id);
  if (!r || !info) return set_error(DSMC_ERR_ARG, "dsmc_server_request: no such request");
  info->state = r->state;
  info->steps = r->steps;
  info->steps_done = r->steps_done;
  info->t_submit = r->t_submit;
  info->t_admit = r->t_admit;
  info->t_finish = r->t_finish;
  info->estimate = r->estimate;
  info->footprint = r->footprint;
  info->particles = r->particles;
  info->particle_steps = r->particle_steps;
  std::strncpy(info->error, r->error.c_str(), sizeof(info->error) - 1);
  info->error[sizeof(info->error) - 1] = 0;
  return DSMC_OK;
}

int dsmc_server_engine(dsmc_server* s, int64_t id, dsmc_engine** eng) {
  Request* r = find(s, id);
  if (!r || !eng) return set_error(DSMC_ERR_ARG, "dsmc_server_engine: no such request");
  if (!r->eng) return set_error(DSMC_ERR_STATE, "dsmc_server_engine: the request holds no engine");
  *eng = r->eng;
  return DSMC_OK;
}

int dsmc_server_release(dsmc_server* s, int64_t id) {
  Request* r = find(s, id);
  if (!r) return set_error(DSMC_ERR_ARG, "dsmc_server_release: no such request");
  if (r->state == REQ_RUNNING) return set_error(DSMC_ERR_STATE, "dsmc_server_release: request still running");
  release(s, *r);
  if (r->state == REQ_DONE) r->state = REQ_RELEASED;
  return DSMC_OK;
}

}  // extern "C"
