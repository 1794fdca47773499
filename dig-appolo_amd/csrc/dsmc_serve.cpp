// Serving loop with continuous batching (no reference counterpart: the
// reference solves one case per MPI job).  Requests — a case mesh, its
// parameters, seed, particles per cell and step count — are queued in arrival
// order and admitted while their estimated device footprint fits the HBM
// budget.  Admitted requests run in engines resident on the device, each
// advanced in chunks of steps enqueued asynchronously on its own streams, so
// the engines' kernels overlap on the GPU.  With batching (up to k cases per
// engine) the queued requests admitted together — consecutive in arrival
// order, with the same step count — share one batch engine, whose every
// kernel launch covers all of them.  A finished request frees its memory (when
// auto-release is on; a batch engine goes when its last request does) and
// the next queued requests are admitted.  A case's random streams are keyed
// by its own cell, patch and particle numbers, so every result is
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

using dsmc::check_tuning;
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
  int64_t footprint = 0;  // bytes its engine holds (shared by a batch's requests)
  int gen = -1;           // the engine group it runs in
  int32_t kase = 0;       // its case in that engine
  int state = REQ_QUEUED;
  int32_t steps_done = 0;
  double t_submit = 0.0, t_admit = 0.0, t_finish = 0.0;
  int64_t particles = 0, particle_steps = 0;
  std::string error;
};

// One engine and the requests (its cases, in case order) it serves.
struct Group {
  dsmc_engine* eng = nullptr;
  std::vector<size_t> members;
  int32_t steps = 0, steps_done = 0, steps_enqueued = 0;
  int live = 0;  // members still holding the engine
  bool running = false;
};

}  // namespace

struct dsmc_server {
  int64_t budget = 0;
  int64_t in_use = 0;
  int32_t chunk = 50;
  int32_t auto_release = 1;
  int32_t max_running = 0;  // 0: no limit
  int32_t max_cases = 1;    // requests per engine
  bool tuned = false;       // `tuning` goes to every engine built (else the engine defaults)
  dsmc_tuning tuning{};
  std::vector<Request> reqs;
  std::vector<Group> groups;
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

// Drops the request's hold on its engine (destroyed with the last hold), its
// reservation and its mesh.
void release(dsmc_server* s, Request& r) {
  if (r.gen >= 0) {
    Group& g = s->groups[(size_t)r.gen];
    r.gen = -1;
    if (--g.live == 0 && g.eng) {
      dsmc_engine_destroy(g.eng);
      g.eng = nullptr;
    }
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

void fail(dsmc_server* s, Request& r, const std::string& why) {
  r.error = why;
  r.state = REQ_FAILED;
  r.t_finish = s->now();
  release(s, r);
}

// A failed engine fails every request it serves.
void fail_group(dsmc_server* s, Group& g) {
  const std::string why = dsmc_last_error();
  g.running = false;
  for (size_t m : g.members)
    if (s->reqs[m].state == REQ_RUNNING) fail(s, s->reqs[m], why);
}

int enqueue_chunk(Group& g, int32_t chunk) {
  const int32_t first = g.steps_done + 1;
  const int32_t last = std::min(g.steps, g.steps_done + chunk);
  g.steps_enqueued = last - g.steps_done;
  return dsmc_run_async(g.eng, first, last);
}

// Builds the engine of an admitted group (one request: its own engine; more:
// one batch engine), fills it on the device and starts its first chunk.
int start_group(dsmc_server* s, Group& g) {
  const size_t k = g.members.size();
  std::vector<dsmc_mesh_view> views(k);
  std::vector<dsmc_params> ps(k);
  std::vector<uint64_t> seeds(k);
  std::vector<double> ppc(k);
  int64_t cap = 0;
  for (size_t c = 0; c < k; ++c) {
    Request& r = s->reqs[g.members[c]];
    int rc = dsmc_mesh_get_view(r.mesh, &views[c]);
    if (rc) return rc;
    ps[c] = r.p;
    seeds[c] = r.seed;
    ppc[c] = r.ppc;
    cap += r.capacity;
  }
  int rc = k == 1 ? dsmc_engine_create(&views[0], &ps[0], seeds[0], cap, &g.eng)
                  : dsmc_batch_create((int32_t)k, views.data(), ps.data(), seeds.data(), cap, &g.eng);
  if (rc) return rc;
  if (s->tuned && (rc = dsmc_engine_set_tuning(g.eng, &s->tuning))) return rc;
  for (size_t m : g.members) {
    Request& r = s->reqs[m];
    dsmc_mesh_destroy(r.mesh);
    r.mesh = nullptr;
  }
  int64_t n = 0, fp = 0;
  rc = dsmc_fill_device(g.eng, ppc.data(), nullptr, seeds.data(), &n);
  if (!rc) rc = dsmc_engine_footprint(g.eng, &fp);
  if (rc) return rc;
  for (size_t c = 0; c < k; ++c) {
    Request& r = s->reqs[g.members[c]];
    r.footprint = fp;
    dsmc_case_info_t ci;
    if ((rc = dsmc_case_info(g.eng, (int32_t)c, &ci))) return rc;
    r.particles = ci.particles;
  }
  g.running = true;
  return enqueue_chunk(g, s->chunk);
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
  // progress of the running engines
  for (Group& g : s->groups) {
    if (!g.running) continue;
    int32_t done = 0;
    if (dsmc_run_query(g.eng, &done)) {
      fail_group(s, g);
      continue;
    }
    if (!done) continue;
    g.steps_done += g.steps_enqueued;
    g.steps_enqueued = 0;
    for (size_t m : g.members) s->reqs[m].steps_done = g.steps_done;
    if (g.steps_done < g.steps) {
      if (enqueue_chunk(g, s->chunk)) fail_group(s, g);
      continue;
    }
    g.running = false;
    const double t = s->now();
    for (size_t c = 0; c < g.members.size(); ++c) {
      Request& r = s->reqs[g.members[c]];
      dsmc_case_info_t ci;
      if (dsmc_case_info(g.eng, (int32_t)c, &ci)) {
        fail(s, r, dsmc_last_error());
        continue;
      }
      r.particle_steps = ci.particle_steps;
      r.particles = ci.particles;
      r.state = REQ_DONE;
      r.t_finish = t;
    }
    if (s->auto_release)
      for (size_t m : g.members) {
        Request& r = s->reqs[m];
        if (r.state != REQ_DONE) continue;
        release(s, r);
        r.state = REQ_RELEASED;
      }
  }
  // admission in arrival order while the next request fits (and the
  // running-request cap allows); with batching, the queued requests that
  // follow it with the same step count join its engine while they fit too
  int32_t nrun = 0;
  for (const Request& r : s->reqs) nrun += (r.state == REQ_RUNNING);
  while (!s->queue.empty()) {
    const Request& r0 = s->reqs[s->queue.front()];
    if (s->in_use + r0.estimate > s->budget) break;
    if (s->max_running > 0 && nrun >= s->max_running) break;
    Group g;
    int64_t est = 0;
    for (size_t q : s->queue) {
      const Request& r = s->reqs[q];
      if ((int32_t)g.members.size() >= s->max_cases || r.steps != r0.steps) break;
      if (s->in_use + est + r.estimate > s->budget) break;
      if (s->max_running > 0 && nrun + (int32_t)g.members.size() >= s->max_running) break;
      g.members.push_back(q);
      est += r.estimate;
    }
    const int gi = (int)s->groups.size();
    const double t = s->now();
    for (size_t c = 0; c < g.members.size(); ++c) {
      s->queue.pop_front();
      Request& r = s->reqs[g.members[c]];
      s->in_use += r.estimate;
      r.reserved = true;
      r.t_admit = t;
      r.state = REQ_RUNNING;
      r.gen = gi;
      r.kase = (int32_t)c;
    }
    nrun += (int32_t)g.members.size();
    g.steps = r0.steps;
    g.live = (int)g.members.size();
    s->groups.push_back(std::move(g));
    Group& gg = s->groups.back();
    if (start_group(s, gg)) fail_group(s, gg);
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

int dsmc_server_set_batching(dsmc_server* s, int32_t max_cases) {
  if (!s || max_cases < 1) return set_error(DSMC_ERR_ARG, "dsmc_server_set_batching: bad arguments");
  s->max_cases = max_cases;
  return DSMC_OK;
}

int dsmc_server_set_tuning(dsmc_server* s, const dsmc_tuning* t) {
  if (!s) return set_error(DSMC_ERR_ARG, "dsmc_server_set_tuning: null server");
  if (!t) {
    s->tuned = false;
    return DSMC_OK;
  }
  if (int rc = check_tuning(t, "dsmc_server_set_tuning")) return rc;
  s->tuning = *t;
  s->tuned = true;
  return DSMC_OK;
}

int dsmc_server_request(dsmc_server* s, int64_t id, dsmc_request_info* info) {
  Request* r = find(s, id);
  if (!r || !info) return set_error(DSMC_ERR_ARG, "dsmc_server_request: no such request");
  info->state = r->state;
  info->steps = r->steps;
  info->steps_done = r->steps_done;
  info->kase = r->kase;
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
  if (r->gen < 0 || !s->groups[(size_t)r->gen].eng)
    return set_error(DSMC_ERR_STATE, "dsmc_server_engine: the request holds no engine");
  *eng = s->groups[(size_t)r->gen].eng;
  return DSMC_OK;
}

int dsmc_server_release(dsmc_server* s, int64_t id) {
  Request* r = find(s, id);
  if (!r) return set_error(DSMC_ERR_ARG, "dsmc_server_release: no such request");
  if (r->state == REQ_RUNNING) return set_error(DSMC_ERR_STATE, "dsmc_server_release: request still running");
  if (r->state == REQ_QUEUED) {
    // a cancel: the request leaves the queue before admission, so no group
    // (and no batch-mate) ever sees its mesh
    s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), (size_t)id), s->queue.end());
    r->error = "cancelled before admission";
    r->state = REQ_FAILED;
    r->t_finish = s->now();
  }
  release(s, *r);
  if (r->state == REQ_DONE) r->state = REQ_RELEASED;
  return DSMC_OK;
}

}  // extern "C"
