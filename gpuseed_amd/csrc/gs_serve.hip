// Seeding server: a resident index, requests submitted from any thread, and
// continuous batching of their reads into device rounds.
//
// A request's reads are cut at its reference batches (a batch closes after
// the read that brings it to batch_bases, seed_gen.cu:2955-2991), so a chunk
// of a request seeds exactly as the request would alone: SMEM seeds depend on
// the read only, MEM row skips never leave a batch (plan_add starts a new
// batch id for every chunk).  Two threads:
//  * the planner builds the next round on the CPU while the device works:
//    every request with unscheduled reads first gets a fair share of the
//    round (round_bases / active requests, at least one reference batch),
//    then the round is filled in arrival order up to round_bases and to half
//    the device budget (two rounds can hold device memory at once: the one
//    running and the one being loaded);
//  * the device thread uploads a round into a reused batch, runs the whole
//    pipeline (one persistent-kernel launch for the reads of every request in
//    the round), downloads the results and appends each chunk's share to its
//    request.  A request is done when its last chunk is.
// Requests arriving while a round runs join the next one; a large request is
// spread over several rounds, interleaved with the requests behind it.
#include <math.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gs_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id = 0;
  const char* bases = nullptr;
  const uint64_t* offs = nullptr;
  const uint32_t* lens = nullptr;
  uint64_t n = 0, n_bases = 0;
  std::vector<uint64_t> cuts;  // read index after each reference batch; last = n
  size_t next_cut = 0;         // first cut not yet scheduled
  uint64_t next_read = 0;      // first read not yet scheduled
  uint64_t done_reads = 0;
  std::vector<uint32_t> nseeds, npos;
  std::vector<gs_seed_rec> seeds;
  std::vector<uint32_t> pos;
  double t_submit = 0, t_start = -1, t_done = -1;
  uint32_t rounds = 0;
  uint32_t inflight = 0;   // chunks planned and not yet processed
  int rc = GS_OK;
  bool done = false;       // results final (or failed)
  bool collected = false;  // gs_server_wait has returned: free once inflight is 0
};

struct Chunk {
  Request* q;
  uint64_t r0, r1;  // reads of the request
  uint64_t at;      // first read in the round
};

struct Round {
  gs::BatchPlan plan;
  std::vector<Chunk> chunks;
  uint64_t device_bytes = 0;
};

}  // namespace

struct gs_server {
  gs_index* X = nullptr;
  gs_server_opts o{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_plan, cv_dev, cv_done;
  std::deque<Request*> active;          // requests with unscheduled reads, in arrival order
  std::map<uint64_t, Request*> reqs;    // every request not yet collected by gs_server_wait
  std::unique_ptr<Round> ready;         // planned round waiting for the device
  uint64_t next_id = 1;
  bool stop = false;
  int fatal = GS_OK;
  std::thread planner, device;
  gs_batch* B = nullptr;
  gs_server_stats st{};
  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Reads of q from its first unscheduled cut up to whole reference batches
// holding at least `want` bases (one batch at least); returns the end read.
uint64_t take(Request* q, uint64_t want, uint64_t* bases_out) {
  uint64_t b = 0, r = q->next_read;
  size_t c = q->next_cut;
  while (c < q->cuts.size()) {
    const uint64_t end = q->cuts[c];
    for (uint64_t i = r; i < end; ++i) b += q->lens[i];
    r = end;
    ++c;
    if (b >= want) break;
  }
  *bases_out = b;
  return r;
}

void advance(Request* q, uint64_t end) {
  while (q->next_cut < q->cuts.size() && q->cuts[q->next_cut] <= end) ++q->next_cut;
  q->next_read = end;
}

// Builds the next round from the active requests (the server lock is held).
std::unique_ptr<Round> plan_round(gs_server* S) {
  auto R = std::make_unique<Round>();
  gs::plan_reset(R->plan, S->o.opts);
  const uint64_t cap = S->o.round_bases;
  const uint64_t mem_cap = S->o.budget_bytes / 2;
  uint64_t bases = 0, est = 0;
  // device bytes of a chunk, an upper bound of what plan_device_bytes adds:
  // bytes, codes, two candidate slots per task, work items and per-read arrays
  auto chunk_bytes = [](uint64_t b, uint64_t reads) { return b * (1 + 1 + 32 + 24 + 4) + reads * 160; };
  auto add = [&](Request* q, uint64_t end) -> bool {
    if (end <= q->next_read) return true;
    uint64_t b = 0;
    for (uint64_t i = q->next_read; i < end; ++i) b += q->lens[i];
    const uint64_t e = chunk_bytes(b, end - q->next_read);
    if (!R->chunks.empty() && est + e > mem_cap) return false;  // over the budget: a later round
    if (R->plan.n_reads + (end - q->next_read) > 0x7FFFFFFFull) return false;
    const uint64_t at = R->plan.n_reads;
    if (gs::plan_add(R->plan, q->bases, q->offs, q->lens, q->next_read, end) != GS_OK) return false;
    bases += b;
    est += e;
    R->chunks.push_back({q, q->next_read, end, at});
    ++q->inflight;
    advance(q, end);
    return true;
  };
  if (S->o.exclusive) {  // one whole request per round
    Request* q = S->active.front();
    add(q, q->n);
  } else {
    const uint64_t n_act = S->active.size();
    const uint64_t share = n_act ? (cap + n_act - 1) / n_act : cap;
    for (Request* q : S->active) {  // fair share first
      if (bases >= cap) break;
      uint64_t b = 0;
      const uint64_t end = take(q, share < cap - bases ? share : cap - bases, &b);
      if (!add(q, end)) break;
    }
    for (Request* q : S->active) {  // then arrival order
      while (bases < cap && q->next_read < q->n) {
        uint64_t b = 0;
        const uint64_t end = take(q, cap - bases, &b);
        if (!add(q, end)) break;
      }
      if (bases >= cap) break;
    }
  }
  if (R->chunks.empty()) {  // the first chunk of a round is always taken unless it cannot be planned
    Request* q = S->active.front();
    q->next_read = q->n;
    q->next_cut = q->cuts.size();
    q->rc = GS_ERANGE;
  }
  for (auto it = S->active.begin(); it != S->active.end();)
    it = (*it)->next_read >= (*it)->n ? S->active.erase(it) : it + 1;
  gs::plan_finish(R->plan);
  R->device_bytes = gs::plan_device_bytes(R->plan);
  return R;
}

void finish_request(gs_server* S, Request* q, int rc) {
  q->rc = rc;
  q->done = true;
  q->t_done = S->now();
  S->cv_done.notify_all();
}

void planner_main(gs_server* S) {
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    S->cv_plan.wait(lk, [&] { return S->stop || (!S->ready && !S->active.empty()); });
    if (S->active.empty() && S->stop) break;
    if (S->ready || S->active.empty()) continue;
    Request* front = S->active.front();
    S->ready = plan_round(S);
    if (S->ready->chunks.empty()) {  // nothing could be planned for the front request
      S->ready.reset();
      if (!front->done) finish_request(S, front, GS_ERANGE);
      continue;
    }
    S->cv_dev.notify_all();
  }
  S->cv_dev.notify_all();
}

void device_main(gs_server* S) {
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    S->cv_dev.wait(lk, [&] { return S->ready || (S->stop && S->active.empty()); });
    if (!S->ready) break;
    std::unique_ptr<Round> R = std::move(S->ready);
    S->cv_plan.notify_all();
    const double t_round = S->now();
    for (const Chunk& c : R->chunks) {
      if (c.q->t_start < 0) c.q->t_start = t_round;
      ++c.q->rounds;
    }
    lk.unlock();
    // device work without the lock: the planner builds the next round meanwhile
    int rc = gs::batch_load(S->B, R->plan);
    gs_batch_stats bs;
    memset(&bs, 0, sizeof bs);
    if (!rc) rc = gs_batch_run(S->B, -1, &bs);
    gs_result res;
    memset(&res, 0, sizeof res);
    if (!rc) rc = gs_batch_fetch(S->B, &res);
    if (!rc) {
      for (const Chunk& c : R->chunks) {  // append each chunk's results to its request
        Request* q = c.q;
        if (q->done) continue;  // failed earlier: nothing more is kept
        const uint64_t a = c.at, z = c.at + (c.r1 - c.r0);
        q->nseeds.insert(q->nseeds.end(), res.read_nseeds + a, res.read_nseeds + z);
        q->npos.insert(q->npos.end(), res.read_npos + a, res.read_npos + z);
        q->seeds.insert(q->seeds.end(), res.seeds + res.read_seed_off[a], res.seeds + res.read_seed_off[z]);
        q->pos.insert(q->pos.end(), res.pos + res.read_pos_off[a], res.pos + res.read_pos_off[z]);
      }
    }
    gs_result_free(&res);
    const double t_end = S->now();
    lk.lock();
    S->st.n_rounds++;
    S->st.busy_seconds += t_end - t_round;
    S->st.device_seconds += bs.t_total;
    S->st.n_reads += R->plan.n_reads;
    S->st.n_bases += R->plan.span;
    if (R->device_bytes > S->st.peak_round_bytes) S->st.peak_round_bytes = R->device_bytes;
    for (const Chunk& c : R->chunks) {
      Request* q = c.q;
      --q->inflight;
      if (!q->done) {
        if (rc) {
          finish_request(S, q, rc);
        } else {
          q->done_reads += c.r1 - c.r0;
          if (q->done_reads == q->n) finish_request(S, q, GS_OK);
        }
      }
      if (q->collected && q->inflight == 0) delete q;
    }
    if (rc) S->fatal = rc;
  }
}

}  // namespace

extern "C" GS_API int gs_server_create(gs_index* X, const gs_server_opts* o, gs_server** out) {
  if (!X || !o || !out) return GS_EINVAL;
  *out = nullptr;
  int rc = gs::plan_check(&o->opts);
  if (rc) return rc;
  gs_server* S = new gs_server();
  S->X = X;
  S->o = *o;
  if (!S->o.budget_bytes) S->o.budget_bytes = 64ull << 30;
  if (!S->o.round_bases) S->o.round_bases = 16000000ull;
  if (S->o.opts.kernel == 0 && !S->o.exclusive) S->o.opts.kernel = 1;
  S->t0 = Clock::now();
  S->B = gs::batch_new(X);
  S->planner = std::thread(planner_main, S);
  S->device = std::thread(device_main, S);
  *out = S;
  return GS_OK;
}

extern "C" GS_API int gs_server_submit(gs_server* S, const char* bases, const uint64_t* offsets,
                                       const uint32_t* lens, uint64_t n_reads, uint64_t* ticket) {
  if (!S || !ticket || (n_reads && (!bases || !offsets || !lens))) return GS_EINVAL;
  if (n_reads > 0xFFFFFFFFull / 2) return GS_ERANGE;
  Request* q = new Request();
  q->bases = bases;
  q->offs = offsets;
  q->lens = lens;
  q->n = n_reads;
  const uint32_t bb = S->o.opts.batch_bases ? S->o.opts.batch_bases : 1000000u;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n_reads; ++i) {
    acc += lens[i];
    q->n_bases += lens[i];
    if (acc >= bb) {
      q->cuts.push_back(i + 1);
      acc = 0;
    }
  }
  if (q->cuts.empty() || q->cuts.back() != n_reads) q->cuts.push_back(n_reads);
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) {
    delete q;
    return GS_EINVAL;
  }
  q->id = S->next_id++;
  q->t_submit = S->now();
  S->reqs[q->id] = q;
  S->st.n_requests++;
  *ticket = q->id;
  if (n_reads == 0) {
    q->t_start = q->t_submit;
    finish_request(S, q, GS_OK);
    return GS_OK;
  }
  S->active.push_back(q);
  S->cv_plan.notify_all();
  return GS_OK;
}

// Frees a collected request now, or leaves it to the device thread when a
// failed request still has chunks in a round (the server lock is held).
static void release_request(Request* q) {
  if (q->inflight == 0)
    delete q;
  else
    q->collected = true;
}

extern "C" GS_API int gs_server_wait(gs_server* S, uint64_t ticket, gs_result* out, gs_request_info* info) {
  if (!S || !out) return GS_EINVAL;
  memset(out, 0, sizeof(*out));
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return GS_EINVAL;
  Request* q = it->second;
  S->cv_done.wait(lk, [&] { return q->done; });
  S->reqs.erase(it);
  if (info) {
    memset(info, 0, sizeof(*info));
    info->t_submit = q->t_submit;
    info->t_start = q->t_start;
    info->t_done = q->t_done;
    info->n_reads = q->n;
    info->n_seeds = q->seeds.size();
    info->n_pos = q->pos.size();
    info->n_rounds = q->rounds;
  }
  if (q->rc) {
    const int rc = q->rc;
    release_request(q);
    return rc;
  }
  // a successful request has no chunk left in any round: its vectors are final
  lk.unlock();
  const uint64_t n = q->n, nr1 = n + 1;
  int rc = GS_OK;
  out->n_reads = n;
  out->n_seeds = q->seeds.size();
  out->n_pos = q->pos.size();
  out->read_nseeds = (uint32_t*)malloc(4 * nr1);
  out->read_npos = (uint32_t*)malloc(4 * nr1);
  out->read_seed_off = (uint64_t*)malloc(8 * nr1);
  out->read_pos_off = (uint64_t*)malloc(8 * nr1);
  out->seeds = (gs_seed_rec*)malloc(sizeof(gs_seed_rec) * (out->n_seeds + 1));
  out->pos = (uint32_t*)malloc(4 * (out->n_pos + 1));
  if (!out->read_nseeds || !out->read_npos || !out->read_seed_off || !out->read_pos_off || !out->seeds ||
      !out->pos) {
    gs_result_free(out);
    rc = GS_ENOMEM;
  } else {
    uint64_t so = 0, po = 0;
    for (uint64_t i = 0; i < n; ++i) {
      out->read_nseeds[i] = q->nseeds[i];
      out->read_npos[i] = q->npos[i];
      out->read_seed_off[i] = so;
      out->read_pos_off[i] = po;
      so += q->nseeds[i];
      po += q->npos[i];
    }
    out->read_seed_off[n] = so;
    out->read_pos_off[n] = po;
    memcpy(out->seeds, q->seeds.data(), sizeof(gs_seed_rec) * out->n_seeds);
    memcpy(out->pos, q->pos.data(), 4 * out->n_pos);
  }
  lk.lock();
  release_request(q);
  return rc;
}

extern "C" GS_API int gs_server_get_stats(gs_server* S, gs_server_stats* st) {
  if (!S || !st) return GS_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  *st = S->st;
  st->elapsed = S->now();
  return S->fatal;
}

extern "C" GS_API double gs_server_clock(gs_server* S) { return S ? S->now() : 0.0; }

extern "C" GS_API int gs_server_destroy(gs_server* S) {
  if (!S) return GS_EINVAL;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv_plan.notify_all();
    S->cv_dev.notify_all();
  }
  S->planner.join();
  S->device.join();
  for (auto& kv : S->reqs) delete kv.second;  // submitted but never collected
  gs::batch_free(S->B);
  const int rc = S->fatal;
  delete S;
  return rc;
}
