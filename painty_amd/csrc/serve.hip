// A request server for stroke + composite work on resident canvases.
//
// Requests (a stamp list and a composite on a canvas identified by an id)
// are submitted from any thread and return a ticket.  Worker slots -- one
// host thread, one HIP stream and one brush each -- take admitted requests
// and run them concurrently on the device: the stamps as PNT_EXEC_FUSED
// launches (FootprintBrush::imprint, renderer/FootprintBrush.hxx:73-143) and
// the composite (Renderer::compose, renderer/Renderer.hxx:48-53) into the
// canvas' resident output planes, optionally copied to a host buffer.  A
// request starts as the reference's fresh FootprintBrush would: pickup map
// clean, snapshot copied whole from the canvas on its first imprint.
//
// Admission: requests on one canvas run in submission order, one at a time;
// a request whose canvas is not resident is admitted when its canvas fits
// the HBM budget, after evicting canvases that no queued or running request
// needs, least recently used first.  An evicted canvas' state is dropped (a
// cache, not a store): later requests on its id fail with PNT_ESTATE.  Requests that arrive while others run
// are admitted as soon as a slot and memory are free, so the device always
// holds the work of up to n_slots requests.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "persistent.hpp"

namespace pnt {
int guard_call(const std::function<void()>& f);  // api.hip: status + pnt_last_error
}

using namespace pnt;
using Clock = std::chrono::steady_clock;

namespace {

struct Resident {
  pnt_canvas cv = nullptr;
  float* block = nullptr;  // canvas planes (11 x stride) then composite planes (3 x stride)
  int64_t cap = 0;         // floats in block (allocated or to allocate)
  int busy = 0;            // a request on it is admitted or running
  Clock::time_point last_use;
};

struct Req {
  int64_t ticket;
  pnt_request r;
  std::vector<pnt_stamp> stamps;
  std::vector<float> footprint;
  Clock::time_point t_submit;
  double latency_ms = 0.0;
  int status = PNT_OK;
  std::string error;
  bool done = false;
};

int64_t canvas_floats(int rows, int cols) {
  return (11 + 3) * round_up((int64_t)rows * cols, 64);  // canvas planes + composite planes
}
int64_t canvas_bytes(int rows, int cols) { return canvas_floats(rows, cols) * 4; }

}  // namespace

struct pnt_server_s {
  int device = 0;
  int64_t budget = 0;  // bytes
  int64_t held = 0;    // bytes of canvas blocks held: resident + pooled
  std::multimap<int64_t, float*> pool;  // idle blocks by capacity (floats)
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> queue;
  std::unordered_map<int64_t, Req*> tickets;
  std::unordered_map<int64_t, Resident> canvases;
  std::unordered_map<int64_t, int> pending;  // queued requests per canvas
  std::unordered_map<int64_t, char> dropped;  // evicted canvases (state gone)
  int64_t next_ticket = 1;
  bool stopping = false;
  std::vector<std::thread> workers;
  std::vector<pnt_ctx> ctxs;
  std::vector<pnt_brush> brushes;
  int64_t completed = 0;

  // Next admissible request (caller holds mu), or nullptr.  A request is
  // admissible when no earlier queued request is on its canvas, its canvas
  // is idle, and a new canvas fits the budget (evicting idle ones).  The
  // canvas is marked busy and a new one's memory reserved before the lock
  // is released.
  Req* take() {
    std::vector<int64_t> seen;
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      Req* q = *it;
      const int64_t id = q->r.canvas_id;
      if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
      seen.push_back(id);
      auto c = canvases.find(id);
      if (c != canvases.end()) {
        if (c->second.busy) continue;
      } else if (dropped.count(id)) {
        queue.erase(it);  // fails in run(): its canvas was evicted
        if (--pending[id] == 0) pending.erase(id);
        q->status = PNT_ESTATE;
        return q;
      } else {
        Resident nr;
        if (!reserve(canvas_floats(q->r.rows, q->r.cols), &nr)) continue;
        c = canvases.emplace(id, nr).first;
      }
      c->second.busy = 1;
      if (--pending[id] == 0) pending.erase(id);
      queue.erase(it);
      return q;
    }
    return nullptr;
  }

  // A block of at least `need` floats for a new canvas (caller holds mu): an
  // idle pooled block of close size, or room for a new one within the budget
  // -- evicting idle canvases into the pool, least recently used first, and
  // releasing pooled blocks that fit nothing.  False when nothing can give
  // way now.
  bool reserve(int64_t need, Resident* r) {
    for (;;) {
      auto it = pool.lower_bound(need);
      if (it != pool.end() && it->first <= 2 * need) {
        r->cap = it->first;
        r->block = it->second;
        pool.erase(it);
        return true;
      }
      if (held + need * 4 <= budget) {
        r->cap = need;
        r->block = nullptr;  // allocated by the worker
        held += need * 4;
        return true;
      }
      if (evict_one()) continue;
      if (pool.empty()) return false;
      auto big = std::prev(pool.end());  // release the largest idle block
      dev_free(big->second);
      held -= big->first * 4;
      pool.erase(big);
    }
  }

  // Moves the least recently used idle canvas with no queued requests into
  // the pool; its state is dropped.  Caller holds mu.
  bool evict_one() {
    int64_t victim = -1;
    Clock::time_point oldest = Clock::time_point::max();
    for (auto& kv : canvases) {
      const Resident& R = kv.second;
      if (R.busy || pending.count(kv.first)) continue;
      if (R.last_use < oldest) {
        oldest = R.last_use;
        victim = kv.first;
      }
    }
    if (victim < 0) return false;
    Resident& R = canvases[victim];
    if (R.cv) pnt_canvas_destroy(R.cv);  // leaves the block alone
    if (R.block) {
      pool.emplace(R.cap, R.block);
    } else {
      held -= R.cap * 4;
    }
    canvases.erase(victim);
    dropped[victim] = 1;
    return true;
  }

  void run(int slot, Req* q);
  void worker(int slot);
};

void pnt_server_s::run(int slot, Req* q) {
  pnt_ctx ctx = ctxs[slot];
  pnt_brush b = brushes[slot];
  hipStream_t s = ctx->stream;
  Resident* R = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);  // entries are stable while busy
    if (q->status == PNT_ESTATE)
      throw Error(PNT_ESTATE, "canvas " + std::to_string(q->r.canvas_id) +
                                  " was evicted from the HBM budget; its state is gone");
    R = &canvases.find(q->r.canvas_id)->second;
  }
  if (!R->cv) {  // a new canvas: substrate `background`, no paint
    if (!R->block) R->block = dev_alloc(R->cap);
    R->cv = canvas_over(ctx, q->r.rows, q->r.cols, R->block);
    const Rgb3 r0 = R->cv->R0();
    for (int c = 0; c < 3; ++c) launch_fill(r0.p[c], R->cv->n, q->r.background[c], s);
    R->cv->bg[0] = q->r.background[0];
    R->cv->bg[1] = q->r.background[1];
    R->cv->bg[2] = q->r.background[2];
  }
  PNT_REQUIRE(R->cv->rows == q->r.rows && R->cv->cols == q->r.cols,
              "request size differs from its resident canvas");
  pnt_canvas cv = R->cv;
  // a fresh brush for this request: FootprintBrush(radius), pickup map clean
  if (b->width == q->r.width) {  // same geometry: new footprint into the same buffers
    const int w = q->r.width, sm = b->size_map, pad = (sm - w) / 2;
    std::fill(b->fp_host.begin(), b->fp_host.end(), 0.0f);
    for (int i = 0; i < w; ++i)
      for (int j = 0; j < w; ++j)
        b->fp_host[(size_t)(i + pad) * sm + (j + pad)] = q->footprint[(size_t)i * w + j];
    footprint_box(b->fp_host, sm, b->fbox);
    PNT_HIP_CHECK(hipMemcpyAsync(b->fp, b->fp_host.data(), b->m * sizeof(float),
                                 hipMemcpyHostToDevice, s));
    launch_fill(b->pick, 7 * b->mstride, 0.0f, s);
    PNT_HIP_CHECK(hipStreamSynchronize(s));  // fp_host is rewritten by the next request
    b->radius = q->r.radius;
  } else {
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    brush_set_geometry(b, q->r.radius, q->footprint.data(), q->r.width);
  }
  b->pickup_rate = q->r.pickup_rate;
  b->dep_rate = q->r.deposition_rate;
  b->use_snap = q->r.use_snapshot != 0;
  for (int c = 0; c < 3; ++c) b->paintK[c] = b->paintS[c] = 0.0f;
  if (b->use_snap) brush_full_snapshot(b, cv);
  // one chain: the worker slots already keep the device busy with other canvases
  run_fused(b, cv, q->stamps.data(), (int64_t)q->stamps.size(), 1);
  float* comp = R->block + 11 * cv->stride;
  const Rgb3 out{{comp, comp + cv->stride, comp + 2 * cv->stride}};
  launch_compose(cv->n, cv->layer(), cv->R0(), out, s);
  if (q->r.out)
    for (int c = 0; c < 3; ++c)
      PNT_HIP_CHECK(hipMemcpyAsync(q->r.out + c * cv->n, out.p[c], cv->n * sizeof(float),
                                   hipMemcpyDeviceToHost, s));
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  if (*(volatile unsigned*)ctx->kernel_error) throw Error(PNT_EKERNEL, "a stamp kernel aborted");
}

void pnt_server_s::worker(int slot) {
  (void)hipSetDevice(device);
  for (;;) {
    Req* q = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stopping || (q = take()) != nullptr; });
      if (!q) return;  // stopping and nothing admissible
    }
    const int st = guard_call([&] { run(slot, q); });
    std::lock_guard<std::mutex> lk(mu);
    q->status = st;
    if (st != PNT_OK) q->error = pnt_last_error();
    auto c = canvases.find(q->r.canvas_id);
    if (c != canvases.end()) {
      c->second.busy = 0;
      c->second.last_use = Clock::now();
    }
    q->latency_ms =
        std::chrono::duration<double, std::milli>(Clock::now() - q->t_submit).count();
    q->done = true;
    ++completed;
    cv_done.notify_all();
    cv_work.notify_all();  // the canvas is free again; memory may be too
  }
}

extern "C" {

int pnt_server_create(int device, int n_slots, int64_t hbm_budget_bytes, pnt_server* out) {
  return guard_call([&] {
    PNT_REQUIRE(out, "NULL argument");
    PNT_REQUIRE(n_slots >= 1 && n_slots <= 64, "n_slots must be in [1, 64]");
    PNT_REQUIRE(hbm_budget_bytes > 0, "HBM budget must be positive");
    auto* S = new pnt_server_s;
    S->device = device;
    S->budget = hbm_budget_bytes;
    try {
      for (int i = 0; i < n_slots; ++i) {
        pnt_ctx c = nullptr;
        int st = pnt_ctx_create(device, &c);
        if (st != PNT_OK) throw Error(st, pnt_last_error());
        S->ctxs.push_back(c);
        const float ones[9] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
        pnt_brush b = nullptr;
        st = pnt_brush_create(c, 1.0, ones, 3, &b);  // placeholder geometry, set per request
        if (st != PNT_OK) throw Error(st, pnt_last_error());
        S->brushes.push_back(b);
      }
    } catch (...) {
      for (auto b : S->brushes) pnt_brush_destroy(b);
      for (auto c : S->ctxs) pnt_ctx_destroy(c);
      delete S;
      throw;
    }
    for (int i = 0; i < n_slots; ++i) S->workers.emplace_back([S, i] { S->worker(i); });
    *out = S;
  });
}

int pnt_server_submit(pnt_server S, const pnt_request* r, int64_t* ticket) {
  return guard_call([&] {
    PNT_REQUIRE(S && r && ticket, "NULL argument");
    PNT_REQUIRE(r->rows > 0 && r->cols > 0 && (int64_t)r->rows * r->cols < ((int64_t)1 << 31),
                "rows/cols out of range");
    PNT_REQUIRE(r->n_stamps >= 0 && (r->stamps || r->n_stamps == 0), "invalid stamp list");
    PNT_REQUIRE(r->footprint && r->width == (int)(2.0 * std::ceil(r->radius) + 1.0),
                "footprint width must be 2*ceil(radius)+1");
    PNT_REQUIRE(canvas_bytes(r->rows, r->cols) <= S->budget, "canvas exceeds the HBM budget");
    PNT_REQUIRE(r->radius > 0.0 && r->radius < 4096.0, "radius must be in (0, 4096)");
    auto* q = new Req;
    q->r = *r;
    q->stamps.assign(r->stamps, r->stamps + r->n_stamps);
    q->footprint.assign(r->footprint, r->footprint + (int64_t)r->width * r->width);
    q->r.stamps = nullptr;
    q->r.footprint = nullptr;
    q->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    PNT_REQUIRE(!S->stopping, "server is shutting down");
    q->ticket = S->next_ticket++;
    S->tickets[q->ticket] = q;
    S->pending[r->canvas_id]++;
    S->queue.push_back(q);
    *ticket = q->ticket;
    S->cv_work.notify_one();
  });
}

int pnt_server_wait(pnt_server S, int64_t ticket, double* latency_ms) {
  return guard_call([&] {
    PNT_REQUIRE(S, "NULL server");
    Req* q = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto it = S->tickets.find(ticket);
      PNT_REQUIRE(it != S->tickets.end(), "unknown ticket");
      q = it->second;
      S->cv_done.wait(lk, [&] { return q->done; });
      S->tickets.erase(it);
    }
    if (latency_ms) *latency_ms = q->latency_ms;
    const int st = q->status;
    const std::string msg = q->error;
    delete q;
    if (st != PNT_OK) throw Error(st, msg);
  });
}

int pnt_server_canvas(pnt_server S, int64_t canvas_id, pnt_canvas* canvas, pnt_dev_rgb* composite) {
  return guard_call([&] {
    PNT_REQUIRE(S && canvas, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto c = S->canvases.find(canvas_id);
    PNT_REQUIRE(c != S->canvases.end() && c->second.cv, "canvas is not resident");
    PNT_REQUIRE(!c->second.busy, "canvas has a running request");
    *canvas = c->second.cv;
    if (composite) {
      const int64_t st = c->second.cv->stride;
      for (int k = 0; k < 3; ++k) composite->p[k] = c->second.block + (11 + k) * st;
    }
  });
}

int pnt_server_stats(pnt_server S, int64_t* resident_bytes, int64_t* resident_canvases,
                     int64_t* completed) {
  return guard_call([&] {
    PNT_REQUIRE(S, "NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (resident_bytes) {
      int64_t b = 0;
      for (auto& kv : S->canvases) b += kv.second.cap * 4;
      *resident_bytes = b;
    }
    if (resident_canvases) *resident_canvases = (int64_t)S->canvases.size();
    if (completed) *completed = S->completed;
  });
}

int pnt_server_destroy(pnt_server S) {
  return guard_call([&] {
    if (!S) return;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_done.wait(lk, [&] {  // let queued work finish
        for (auto& kv : S->tickets)
          if (!kv.second->done) return false;
        return S->queue.empty();
      });
      S->stopping = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& kv : S->tickets) delete kv.second;
    for (auto& kv : S->canvases) {
      if (kv.second.cv) pnt_canvas_destroy(kv.second.cv);
      dev_free(kv.second.block);
    }
    for (auto& kv : S->pool) dev_free(kv.second);
    for (auto b : S->brushes) pnt_brush_destroy(b);
    for (auto c : S->ctxs) pnt_ctx_destroy(c);
    delete S;
  });
}

}  // extern "C"
