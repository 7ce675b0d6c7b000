// Request server: continuous batching of argsort / group-by / unique requests over HBM-resident
// columns.
//
// The reference serves queries one call at a time from the caller's thread (Sources/ArrowMetalC, every
// entry point synchronous; its resident mode keeps columns on the device between calls,
// Sources/ArrowMetal/Resident/ResidentMode.swift). Here a fixed set of worker threads, each with its
// own HIP stream, takes requests from one FIFO queue. A request is admitted when a worker is free
// and its estimated device footprint fits in what the budget has left (a request larger than the
// whole budget runs alone), so requests that arrive while others run join them on the device and
// their kernels overlap across streams. Admission is strictly in arrival order: a request that does
// not fit yet holds back the ones behind it, so none starves.
//
// Each request runs the same entry-point code as a plain call (am_argsort_ex2, am_group_by_keys +
// keys + am_group_agg_ex sum / count_all, am_unique_rows + am_take), so its result is bit-identical
// to the plain call's whether it runs alone or in the mix.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "am_internal.hpp"

namespace am {

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum ServeOp { OP_GROUP_SUM_COUNT = 0, OP_ARGSORT = 1, OP_UNIQUE = 2 };

struct Request {
  int64_t ticket = 0;
  int op = 0;
  std::vector<Array*> cols;   // one reference each
  Array* values = nullptr;
  int desc = 0, placement = 0, forder = 0;
  size_t footprint = 0;
  int64_t rows = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
  int state = 0;              // 0 queued, 1 running, 2 done
  bool failed = false;
  std::string error;
  std::vector<Array*> results;
  ~Request() {
    for (Array* a : cols) release_array(a);
    if (values) release_array(values);
    for (Array* a : results) release_array(a);
  }
};

// Device bytes a request may hold at its peak (upper estimates of the library's own allocations).
size_t footprint_of(int op, int64_t n, int64_t ncols) {
  const size_t un = size_t(std::max<int64_t>(n, 1));
  const size_t table = size_t(64) << 20;   // hash table, compacted keys and slot maps (<= 2^21 slots)
  switch (op) {
    case OP_ARGSORT: return 29 * un + (size_t(16) << 20);   // keys + rows ping-pong, out, status
    case OP_GROUP_SUM_COUNT: return 16 * un + 2 * table;   // ids, aggregate tables, outputs
    default: return size_t(ncols) * (12 * un + table) + 16 * un;
  }
}

}  // namespace

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Request>> tickets;
  std::vector<std::thread> workers;
  size_t budget = 0, admitted = 0, peak_admitted = 0;
  int running = 0, peak_running = 0;
  int64_t next_ticket = 1, submitted = 0, completed = 0, failed = 0;
  bool stop = false;

  void run_one(Request& r) {
    switch (r.op) {
      case OP_ARGSORT:
        r.results.push_back(argsort(*r.cols[0], r.desc != 0, r.placement, r.forder));
        break;
      case OP_GROUP_SUM_COUNT: {
        GroupBy* g = group_by_keys(r.cols.data(), int64_t(r.cols.size()));
        try {
          for (size_t c = 0; c < r.cols.size(); c++) r.results.push_back(group_by_keys_result(g, int64_t(c)));
          r.results.push_back(group_agg(g, r.values, 0, 0.0));
          r.results.push_back(group_agg(g, nullptr, 1, 0.0));
        } catch (...) {
          group_by_release(g);
          throw;
        }
        group_by_release(g);
        break;
      }
      default: {
        Array* rows = unique_rows(r.cols.data(), int64_t(r.cols.size()));
        r.results.push_back(rows);
        for (Array* c : r.cols) r.results.push_back(take(*c, *rows));
        break;
      }
    }
  }

  void worker() {
    while (true) {
      std::shared_ptr<Request> r;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] {
          if (stop && queue.empty()) return true;
          if (queue.empty()) return false;
          const size_t fp = queue.front()->footprint;
          return running == 0 || admitted + fp <= budget;
        });
        if (queue.empty()) return;   // stopping
        r = queue.front();
        queue.pop_front();
        admitted += r->footprint;
        running++;
        peak_admitted = std::max(peak_admitted, admitted);
        peak_running = std::max(peak_running, running);
        r->state = 1;
        r->t_start = now_s();
      }
      // A worker may be able to admit the next request too.
      cv_work.notify_one();
      try {
        run_one(*r);
      } catch (const std::exception& e) {
        r->failed = true;
        r->error = e.what();
      } catch (...) {
        r->failed = true;
        r->error = "unknown error";
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        r->t_end = now_s();
        r->state = 2;
        admitted -= r->footprint;
        running--;
        completed++;
        if (r->failed) failed++;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

Server* server_create(int slots, int64_t budget_bytes) {
  if (slots < 1 || slots > 16) fail("server: slots must be in 1..16");
  if (budget_bytes <= 0) fail("server: the memory budget must be positive");
  device();
  auto* s = new Server();
  s->budget = size_t(budget_bytes);
  try {
    for (int i = 0; i < slots; i++) s->workers.emplace_back([s] { s->worker(); });
  } catch (...) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    delete s;
    throw;
  }
  return s;
}

void server_release(Server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  delete s;
}

int64_t server_submit(Server* s, int op, Array* const* cols, int64_t ncols, Array* values, int desc, int placement,
                      int forder) {
  if (op < 0 || op > 2) fail("server: unknown request op " + std::to_string(op));
  if (ncols < 1) fail("server: a request needs at least one key column");
  if (op == OP_ARGSORT && ncols != 1) fail("server: an argsort request takes one column");
  if (op == OP_GROUP_SUM_COUNT && !values) fail("server: a group-by request needs a value column");
  if (op == OP_ARGSORT && (forder < 0 || forder > 2)) fail_invalid("unknown float_order " + std::to_string(forder));
  auto r = std::make_shared<Request>();
  r->op = op;
  for (int64_t c = 0; c < ncols; c++) {
    retain_array(cols[c]);
    r->cols.push_back(cols[c]);
  }
  if (values && op == OP_GROUP_SUM_COUNT) {
    retain_array(values);
    r->values = values;
  }
  r->desc = desc;
  r->placement = placement;
  r->forder = forder;
  r->rows = cols[0]->length;
  r->footprint = footprint_of(op, r->rows, ncols);
  r->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) fail("server: stopped");
    r->ticket = s->next_ticket++;
    s->submitted++;
    s->queue.push_back(r);
    s->tickets[r->ticket] = r;
  }
  s->cv_work.notify_one();
  return r->ticket;
}

// Waits for a request; hands over its results (up to `cap`, new references) and its times in ms
// (queued, running, submit to done). A failed request raises its error.
int64_t server_wait(Server* s, int64_t ticket, Array** out, int64_t cap, double* times) {
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // The waiter takes the ticket out of the map before it sleeps: submits that insert (and rehash)
    // meanwhile touch no iterator of ours, and a second wait on the same ticket finds it gone.
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) fail("server: unknown ticket " + std::to_string(ticket));
    r = std::move(it->second);
    s->tickets.erase(it);
    s->cv_done.wait(lk, [&] { return r->state == 2; });
  }
  if (times) {
    times[0] = 1e3 * (r->t_start - r->t_submit);
    times[1] = 1e3 * (r->t_end - r->t_start);
    times[2] = 1e3 * (r->t_end - r->t_submit);
  }
  if (r->failed) fail(r->error);
  const int64_t k = int64_t(r->results.size());
  for (int64_t i = 0; i < k && i < cap; i++) {
    out[i] = r->results[size_t(i)];
    r->results[size_t(i)] = nullptr;
  }
  for (Array*& a : r->results)
    if (a) { release_array(a); a = nullptr; }
  r->results.clear();
  return k;
}

void server_stats(Server* s, int64_t* out, int count) {
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t v[8] = {s->submitted, s->completed, s->failed, int64_t(s->peak_admitted), s->peak_running,
                        int64_t(s->budget), int64_t(s->queue.size()), int64_t(s->workers.size())};
  for (int i = 0; i < count && i < 8; i++) out[i] = v[i];
}

}  // namespace am
