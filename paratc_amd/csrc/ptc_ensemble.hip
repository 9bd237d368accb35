// ptc_ensemble.hip -- serving many independent cases on one GPU.
//
// The reference advances one case per job (src/m_ChannelSystem.f90:80-145).
// An ensemble here holds any number of admitted cases, each a full context
// (its own fields, Poisson plan and HIP stream), and advances them together:
//
//  * admission by HBM footprint: a case is admitted when the device bytes its
//    context holds fit the ensemble's budget next to the cases already in it;
//    otherwise PTC_E_LIMIT tells the caller to queue it and try again later;
//  * one launch per case per step: with a fixed time step (iCFL != 1) the
//    step's kernel sequence is captured once into a HIP graph and replayed;
//    cases with CFL-adaptive dt (host-side dt from a device reduction) are
//    stepped with ordinary launches;
//  * cases run on their own streams, so small grids that cannot fill the
//    chip on their own overlap each other, and no host barrier separates one
//    round of steps from the next: ptc_ensemble_step enqueues and returns,
//    each case running ahead on its stream by at most kMaxAhead steps (the
//    host waits only on that case's own older step);
//  * every case-step records a completion event on the case's stream; its
//    latency is the time from the case's previous step completion (or its
//    admission) to its own, read once both events have completed (p50 / p99
//    on request).
#include <algorithm>
#include <deque>
#include <vector>

#include "ptc_internal.hpp"

struct ptc_ensemble {
  struct Case {
    ptc_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long gen = 0;  // the context's buffer generation at capture
    size_t bytes = 0;
    long long steps = 0;          // steps enqueued
    bool live = false;
    hipEvent_t prev = nullptr;    // completion of the case's previous step (or its admission)
    std::deque<hipEvent_t> inflight;  // step completions not yet harvested, oldest first
  };
  size_t budget = 0, used = 0;
  std::vector<Case> cases;
  std::vector<hipEvent_t> pool;  // reusable events
  std::vector<double> lat_ms;    // per case-step latencies since the last reset
};

namespace ptc {

constexpr size_t kMaxAhead = 4;  // steps in flight per case

// device bytes a context holds (fields, reduction partials, Poisson plan)
static size_t ctx_bytes(const ptc_ctx *c) {
  size_t b = 0;
  for (int f = 0; f < PTC_NFIELDS; ++f)
    if (c->d_field[f]) b += sizeof(double) * ((f == PTC_F_UY_YM || f == PTC_F_DUY_YM) ? (size_t)c->nx * c->nz : c->n3);
  b += sizeof(double) * (size_t)c->red_cap;
  if (c->pp.d_spec) b += c->pp.spec_bytes;
  if (c->pp.d_tri_scratch) b += sizeof(double) * c->pp.spec_elems;
  return b;
}

static int take_event(ptc_ensemble *e, hipEvent_t *ev) {
  if (!e->pool.empty()) {
    *ev = e->pool.back();
    e->pool.pop_back();
    return PTC_OK;
  }
  PTC_HIP(hipEventCreate(ev));
  return PTC_OK;
}

// Harvest the case's completed step events in order: each one's latency from
// the previous completion.  block: wait for the oldest (then harvest what else
// has completed); all: wait for every one.
static int harvest(ptc_ensemble *e, ptc_ensemble::Case &k, bool block, bool all) {
  bool first = true;
  while (!k.inflight.empty()) {
    hipEvent_t ev = k.inflight.front();
    if ((block && first) || all) {
      PTC_HIP(hipEventSynchronize(ev));
    } else {
      const hipError_t q = hipEventQuery(ev);
      if (q == hipErrorNotReady) {
        (void)hipGetLastError();
        break;
      }
      PTC_HIP(q);
    }
    first = false;
    float ms = 0.f;
    if (k.prev) {
      PTC_HIP(hipEventElapsedTime(&ms, k.prev, ev));
      e->lat_ms.push_back(ms);
      e->pool.push_back(k.prev);
    }
    k.prev = ev;
    k.inflight.pop_front();
  }
  return PTC_OK;
}

// every admitted case's steps done, every latency read
static int sync_all(ptc_ensemble *e) {
  for (auto &k : e->cases)
    if (k.live) PTC_TRY(harvest(e, k, true, true));
  return PTC_OK;
}

static void drop_case_graph(ptc_ensemble::Case &k) {
  if (k.exec) (void)hipGraphExecDestroy(k.exec);
  if (k.graph) (void)hipGraphDestroy(k.graph);
  k.exec = nullptr;
  k.graph = nullptr;
}

static int step_case(ptc_ensemble::Case &k) {
  // a plan change, a freed plan or grown reduction partials (clcStat, a
  // residual) since the capture: the graph holds stale addresses
  if (k.exec && k.gen != k.ctx->buf_gen) drop_case_graph(k);
  if (!k.exec && k.ctx->prm.icfl != 1 && k.ctx->pp.ready) {
    PTC_TRY(capture_step(k.ctx, &k.graph, &k.exec));
    k.gen = k.ctx->buf_gen;
  }
  if (k.exec) {
    PTC_HIP(hipGraphLaunch(k.exec, k.ctx->stream));
    k.ctx->simtime += k.ctx->prm.dtmax;  // the RK3 alphas sum to one
  } else {
    PTC_TRY(channel_step(k.ctx, nullptr));
  }
  ++k.steps;
  return PTC_OK;
}

}  // namespace ptc

using namespace ptc;

extern "C" {

int ptc_ensemble_create(size_t hbm_budget, ptc_ensemble **out) {
  PTC_CHECK(out != nullptr, PTC_E_ARG, "null output");
  *out = nullptr;
  auto *e = new ptc_ensemble();
  e->budget = hbm_budget;
  *out = e;
  return PTC_OK;
}

int ptc_ensemble_destroy(ptc_ensemble *e) {
  if (!e) return PTC_OK;
  for (auto &k : e->cases) {
    if (k.ctx) (void)hipStreamSynchronize(k.ctx->stream);
    drop_case_graph(k);
    for (hipEvent_t ev : k.inflight) (void)hipEventDestroy(ev);
    if (k.prev) (void)hipEventDestroy(k.prev);
    if (k.ctx) ptc_destroy(k.ctx);
  }
  for (hipEvent_t ev : e->pool) (void)hipEventDestroy(ev);
  delete e;
  return PTC_OK;
}

int ptc_case_footprint(ptc_ctx *ctx, size_t *bytes) {
  PTC_CHECK(ctx && bytes, PTC_E_ARG, "null argument");
  PTC_TRY(prepare_step(ctx));
  *bytes = ctx_bytes(ctx);
  return PTC_OK;
}

int ptc_ensemble_admit(ptc_ensemble *e, ptc_ctx *ctx, int *case_id) {
  PTC_CHECK(e && ctx && case_id, PTC_E_ARG, "null argument");
  for (const auto &k : e->cases)
    PTC_CHECK(!(k.live && k.ctx == ctx), PTC_E_ARG, "context already admitted");
  PTC_TRY(prepare_step(ctx));
  const size_t b = ctx_bytes(ctx);
  if (e->used + b > e->budget) {
    set_error("HBM budget: case needs " + std::to_string(b) + " B, " + std::to_string(e->budget - e->used) +
              " B free");
    return PTC_E_LIMIT;
  }
  ptc_ensemble::Case k;
  k.ctx = ctx;
  k.bytes = b;
  k.live = true;
  if (ctx->prm.icfl != 1) {
    PTC_TRY(capture_step(ctx, &k.graph, &k.exec));
    k.gen = ctx->buf_gen;
  }
  // the first step's latency runs from here
  PTC_TRY(take_event(e, &k.prev));
  PTC_HIP(hipEventRecord(k.prev, ctx->stream));
  e->used += b;
  int id = -1;
  for (size_t q = 0; q < e->cases.size(); ++q)
    if (!e->cases[q].live && !e->cases[q].ctx) {
      id = (int)q;
      break;
    }
  if (id < 0) {
    id = (int)e->cases.size();
    e->cases.push_back(k);
  } else {
    e->cases[id] = k;
  }
  *case_id = id;
  return PTC_OK;
}

int ptc_ensemble_step(ptc_ensemble *e, int nsteps) {
  PTC_CHECK(e, PTC_E_ARG, "null ensemble");
  PTC_CHECK(nsteps >= 1, PTC_E_ARG, "nsteps must be >= 1");
  // steps go out case by case, round-robin, so the streams advance together
  for (int s = 0; s < nsteps; ++s) {
    for (auto &k : e->cases) {
      if (!k.live) continue;
      if (k.inflight.size() >= kMaxAhead) PTC_TRY(harvest(e, k, true, false));
      PTC_TRY(step_case(k));
      hipEvent_t ev;
      PTC_TRY(take_event(e, &ev));
      if (hipEventRecord(ev, k.ctx->stream) != hipSuccess) {
        e->pool.push_back(ev);
        set_error("hipEventRecord failed");
        return PTC_E_HIP;
      }
      k.inflight.push_back(ev);
    }
  }
  for (auto &k : e->cases)
    if (k.live) PTC_TRY(harvest(e, k, false, false));
  return PTC_OK;
}

int ptc_ensemble_sync(ptc_ensemble *e) {
  PTC_CHECK(e, PTC_E_ARG, "null ensemble");
  return sync_all(e);
}

int ptc_ensemble_case(ptc_ensemble *e, int case_id, ptc_ctx **ctx, long long *steps) {
  PTC_CHECK(e && case_id >= 0 && case_id < (int)e->cases.size() && e->cases[case_id].live, PTC_E_ARG,
            "unknown case");
  if (ctx) *ctx = e->cases[case_id].ctx;
  if (steps) *steps = e->cases[case_id].steps;
  return PTC_OK;
}

int ptc_ensemble_release(ptc_ensemble *e, int case_id, ptc_ctx **ctx) {
  PTC_CHECK(e && case_id >= 0 && case_id < (int)e->cases.size() && e->cases[case_id].live, PTC_E_ARG,
            "unknown case");
  auto &k = e->cases[case_id];
  PTC_TRY(harvest(e, k, true, true));
  PTC_HIP(hipStreamSynchronize(k.ctx->stream));
  if (k.prev) e->pool.push_back(k.prev);
  k.prev = nullptr;
  drop_case_graph(k);
  e->used -= k.bytes;
  if (ctx) {
    *ctx = k.ctx;
  } else {
    ptc_destroy(k.ctx);
  }
  k = ptc_ensemble::Case();
  return PTC_OK;
}

int ptc_ensemble_usage(ptc_ensemble *e, size_t *used, size_t *budget, int *live) {
  PTC_CHECK(e, PTC_E_ARG, "null ensemble");
  if (used) *used = e->used;
  if (budget) *budget = e->budget;
  if (live) *live = (int)std::count_if(e->cases.begin(), e->cases.end(), [](const ptc_ensemble::Case &k) {
    return k.live;
  });
  return PTC_OK;
}

int ptc_ensemble_latency(ptc_ensemble *e, int reset, double *p50, double *p99, long long *nsamples) {
  PTC_CHECK(e, PTC_E_ARG, "null ensemble");
  PTC_TRY(sync_all(e));
  std::vector<double> v = e->lat_ms;
  if (nsamples) *nsamples = (long long)v.size();
  auto pct = [&](double q) -> double {
    if (v.empty()) return 0.0;
    const size_t i = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + (long)i, v.end());
    return v[i];
  };
  if (p50) *p50 = pct(0.50);
  if (p99) *p99 = pct(0.99);
  if (reset) e->lat_ms.clear();
  return PTC_OK;
}

}  // extern "C"
