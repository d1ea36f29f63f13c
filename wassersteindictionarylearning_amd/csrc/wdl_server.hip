// Served projection loop: continuous batching of weight-fitting requests.
//
// A request projects one histogram onto a registered dictionary: it runs
// `rounds` projection rounds, each a weights-only Sinkhorn barycenter +
// gradient (compute_gradient_both's weight chain, inverseWasserstein.h:635-673)
// with `iterations` Sinkhorn iterations, warm-started from the previous
// round's b (the warm restart of :1057-1061, :1081-1085), followed by one
// mirror-descent step on the weights (k_project_update).
//
// Requests are served from pools: a pool is one batched solver (P = its slot
// capacity) on its own context stream, for one dictionary and one iteration
// class (L rounded up to a power of two from 32, capped at the dictionary's
// maximum), so a round runs no more Sinkhorn launches than its requests need
// and short requests do not hold the histories of long ones.  A scheduler
// thread
//   * collects pools whose round has finished (stream query, no blocking),
//     counts the round for each request in the pool and retires requests that
//     have run all their rounds (their slot becomes idle: 0 iterations);
//   * admits waiting requests, oldest first, into free slots of idle pools of
//     their dictionary, creating a pool when the device-memory budget admits
//     its footprint (wdl_solver_footprint) and releasing idle empty pools of
//     other dictionaries when it does not;
//   * launches the next round of every idle pool that has active slots.
// Pools of different dictionaries (grid sizes) run concurrently on their own
// streams; requests join a pool between its rounds, so work that arrives while
// others run is admitted as soon as a slot and memory are free.  Every slot's
// arithmetic is independent of the others', so a request's result does not
// depend on what it was batched with.
//
// Mixed batching (wdl_server_set_batching): the pools of one iteration class
// share one stream, and each round of the class runs all of them -- every
// dictionary and grid size of that class -- as one group launch per stage
// kernel (wdl_group_projection_round), with a class's pools joined and left
// between its rounds.  Classes run concurrently on their own streams.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "wdl_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
  return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

struct Dict {
  int W, H, K, loss, maxL, flags;
  double gamma;
  std::vector<double> atoms;
};

// Iteration class of a request: the smallest power of two >= L (at least 32),
// at most the dictionary's maximum.
int iteration_class(int L, int maxL) {
  int c = 32;
  while (c < L) c *= 2;
  return std::min(c, maxL);
}

const int kPoolFlags = WDL_FLAG_WEIGHTS_ONLY | WDL_FLAG_WARM_RESTART;

struct Request {
  long long id;
  int dict, K, lclass;
  std::vector<double> hist;
  int iters, rounds;
  double step;
  double tSubmit, tAdmit = -1, tDone = -1;
  int pool = -1, slot = -1, doneRounds = 0;
  bool waitedMem = false;
  std::vector<double> weights;
  double loss = 0;
};

struct Pool {
  int dict, lclass;
  wdl_ctx* ctx = nullptr;  // the pool's own, or its class's when grouped
  bool grouped = false;    // mixed batching of an exact-shift dictionary: rounds run
                           // with the other pools of its class as one group
  wdl_solver* s = nullptr;
  std::vector<long long> slots;  // request id or -1
  int active = 0;
  bool busy = false;
  double idleSince = 0;
  long long bytes = 0;
};

// Mixed batching: the pools of one iteration class, their shared stream and
// the group their rounds run as.
struct ClassRun {
  wdl_ctx* ctx = nullptr;
  wdl_group* group = nullptr;
  std::vector<Pool*> members;  // the pools `group` was made of, in pool order
  bool busy = false;
};

}  // namespace

struct wdl_server {
  int device = 0;
  long long budget = 0, used = 0;
  int maxSlots = 16;
  std::vector<std::shared_ptr<const Dict>> dicts;  // under mu (appended by add_dictionary)
  std::vector<std::unique_ptr<Pool>> pools;          // worker-owned
  // worker-owned
  std::unordered_map<long long, Request> running;
  std::deque<long long> pending;
  // shared, under mu
  std::mutex mu;
  std::condition_variable cvDone;
  std::vector<Request> incoming;
  std::unordered_map<long long, Request> done;
  long long nextId = 1;
  long long stats[6] = {0, 0, 0, 0, 0, 0};  // completed, pools created, pools released,
                                             // memory waits, bytes in use, rounds launched
  std::string error;
  std::thread worker;
  std::atomic<bool> stop{false};
  std::atomic<int> mixed{0};                  // batching mode (set before the first submit)
  std::unordered_map<int, ClassRun> classes;  // worker-owned (mixed batching)

  // Pools of a dictionary run as groups in mixed batching when the dictionary
  // is exact-shift log domain (the forms a group takes); others keep streams
  // of their own.
  bool grouped(const Dict& D) const {
    return mixed.load() && (D.flags & WDL_FLAG_LOG_DOMAIN) && (D.flags & WDL_FLAG_EXACT_LSE);
  }

  ClassRun* class_of(int lclass) {
    auto it = classes.find(lclass);
    return it == classes.end() ? nullptr : &it->second;
  }

  void drop_group(ClassRun& c) {
    if (c.group) wdl_group_destroy(c.group);
    c.group = nullptr;
    c.members.clear();
  }

  int fail(const std::string& where) {
    std::lock_guard<std::mutex> g(mu);
    error = where + ": " + wdl_last_error();
    cvDone.notify_all();
    return -1;
  }

  void release_pool(size_t k) {
    Pool& p = *pools[k];
    if (ClassRun* c = p.grouped ? class_of(p.lclass) : nullptr) {
      // the class's group holds the pool's solver: it goes first (the class is
      // idle -- a pool of a running class is busy and never released)
      if (std::find(c->members.begin(), c->members.end(), &p) != c->members.end()) drop_group(*c);
    }
    if (p.s) wdl_solver_destroy(p.s);
    if (p.ctx && !p.grouped) wdl_ctx_destroy(p.ctx);
    used -= p.bytes;
    pools.erase(pools.begin() + (long)k);
    std::lock_guard<std::mutex> g(mu);
    ++stats[2];
    stats[4] = used;
  }

  // A new pool for dictionary d; false when its footprint does not fit.
  std::shared_ptr<const Dict> dict(int d) {
    std::lock_guard<std::mutex> g(mu);
    return dicts[(size_t)d];
  }

  // Slots per pool: at most maxSlots, and a pool takes at most a quarter of the
  // budget, so that several grid sizes are served at once.
  void pool_shape(const Dict& D, int lclass, int& cap, long long& bytes) const {
    const long long slot = wdl_solver_footprint(D.W, D.H, D.K, 1, lclass, D.flags | kPoolFlags);
    cap = (int)std::min<long long>(maxSlots, std::max<long long>(1, budget / 4 / slot));
    bytes = wdl_solver_footprint(D.W, D.H, D.K, cap, lclass, D.flags | kPoolFlags);
  }

  bool create_pool(int d, int lclass) {
    const std::shared_ptr<const Dict> Dp = dict(d);
    const Dict& D = *Dp;
    int cap;
    long long poolBytes;
    pool_shape(D, lclass, cap, poolBytes);
    if (used + poolBytes > budget) {
      // release idle, empty pools of other dictionaries, least recently used first
      std::vector<size_t> idle;
      for (size_t k = 0; k < pools.size(); ++k)
        if (!pools[k]->busy && pools[k]->active == 0) idle.push_back(k);
      std::sort(idle.begin(), idle.end(),
                [&](size_t a, size_t b) { return pools[a]->idleSince < pools[b]->idleSince; });
      long long freeable = 0;
      for (size_t k : idle) freeable += pools[k]->bytes;
      if (used - freeable + poolBytes > budget) return false;
      std::vector<size_t> drop;
      long long need = used + poolBytes - budget;
      for (size_t k : idle) {
        if (need <= 0) break;
        need -= pools[k]->bytes;
        drop.push_back(k);
      }
      std::sort(drop.rbegin(), drop.rend());
      for (size_t k : drop) release_pool(k);
    }
    std::unique_ptr<Pool> p(new Pool());
    p->dict = d;
    p->lclass = lclass;
    p->slots.assign((size_t)cap, -1);
    const int flags = D.flags | kPoolFlags;
    const bool mix = grouped(D);
    p->grouped = mix;
    if (mix) {  // the class's shared stream, created with its first pool
      ClassRun& c = classes[lclass];
      if (!c.ctx && wdl_ctx_create(device, &c.ctx)) {
        fail("wdl_ctx_create");
        return false;
      }
      p->ctx = c.ctx;
    } else if (wdl_ctx_create(device, &p->ctx)) {
      fail("wdl_ctx_create");
      return false;
    }
    if (wdl_solver_create(p->ctx, D.W, D.H, D.K, cap, lclass, D.gamma, D.loss, flags, &p->s)) {
      if (!mix) wdl_ctx_destroy(p->ctx);
      fail("wdl_solver_create");
      return false;
    }
    std::vector<double> lam((size_t)cap * D.K, 1.0 / D.K);
    std::vector<int> zero((size_t)cap, 0);
    if (wdl_solver_set_problem(p->s, D.atoms.data(), lam.data()) ||
        wdl_solver_set_iterations(p->s, zero.data())) {
      wdl_solver_destroy(p->s);
      if (!mix) wdl_ctx_destroy(p->ctx);
      fail("pool setup");
      return false;
    }
    p->bytes = poolBytes;
    p->idleSince = now_s();
    used += p->bytes;
    pools.push_back(std::move(p));
    std::lock_guard<std::mutex> g(mu);
    ++stats[1];
    stats[4] = used;
    return true;
  }

  // A pool's round has ended: count it for each request and retire the
  // requests that have run all their rounds.  False on failure (reported).
  bool round_done(Pool& p) {
    p.busy = false;
    p.idleSince = now_s();
    for (size_t k = 0; k < p.slots.size(); ++k) {
      if (p.slots[k] < 0) continue;
      Request& r = running.at(p.slots[k]);
      if (++r.doneRounds < r.rounds) continue;
      r.weights.resize((size_t)r.K);
      if (wdl_solver_get_slot(p.s, (int)k, r.weights.data(), &r.loss) ||
          wdl_solver_set_slot(p.s, (int)k, nullptr, nullptr, 0, 0.0)) {
        fail("retire");
        return false;
      }
      p.slots[k] = -1;
      --p.active;
      Request fin = std::move(r);
      running.erase(fin.id);
      finish(fin);
    }
    return true;
  }

  void finish(Request& r) {
    r.tDone = now_s();
    std::lock_guard<std::mutex> g(mu);
    ++stats[0];
    done.emplace(r.id, std::move(r));
    cvDone.notify_all();
  }

  void loop() {
    while (!stop.load()) {
      bool progress = false;
      {
        std::lock_guard<std::mutex> g(mu);
        for (Request& r : incoming) {
          pending.push_back(r.id);
          running.emplace(r.id, std::move(r));
        }
        incoming.clear();
      }
      // 1. finished rounds
      {
        for (auto& pp : pools) {
          Pool& p = *pp;
          if (!p.busy || p.grouped) continue;
          const int q = wdl_ctx_query(p.ctx);
          if (q < 0) {
            fail("wdl_ctx_query");
            return;
          }
          if (q == 1) continue;
          progress = true;
          if (!round_done(p)) return;
        }
        for (auto& kv : classes) {
          ClassRun& c = kv.second;
          if (!c.busy) continue;
          const int q = wdl_ctx_query(c.ctx);
          if (q < 0) {
            fail("wdl_ctx_query");
            return;
          }
          if (q == 1) continue;
          c.busy = false;
          progress = true;
          for (Pool* p : c.members)
            if (!round_done(*p)) return;
        }
      }
      // 2. admission, oldest first (mixed batching: into pools of a class
      // between its rounds, since a class's pools share its stream)
      for (auto it = pending.begin(); it != pending.end();) {
        Request& r = running.at(*it);
        const ClassRun* cr = grouped(*dict(r.dict)) ? class_of(r.lclass) : nullptr;
        if (cr && cr->busy) {
          ++it;
          continue;
        }
        Pool* target = nullptr;
        for (auto& pp : pools)
          if (pp->dict == r.dict && pp->lclass == r.lclass && !pp->busy &&
              pp->active < (int)pp->slots.size()) {
            target = pp.get();
            break;
          }
        if (!target) {
          bool busyWithRoom = false;
          for (auto& pp : pools)
            if (pp->dict == r.dict && pp->lclass == r.lclass && pp->active < (int)pp->slots.size())
              busyWithRoom = true;
          if (busyWithRoom) {  // admitted when that pool's round ends
            ++it;
            continue;
          }
          if (!create_pool(r.dict, r.lclass)) {
            if (!error.empty()) return;
            if (!r.waitedMem) {
              r.waitedMem = true;
              std::lock_guard<std::mutex> g(mu);
              ++stats[3];
            }
            ++it;
            continue;
          }
          target = pools.back().get();
        }
        int k = 0;
        while (target->slots[(size_t)k] >= 0) ++k;
        std::vector<double> lam((size_t)r.K, 1.0 / r.K);
        if (wdl_solver_set_slot(target->s, k, r.hist.data(), lam.data(), r.iters, r.step)) {
          fail("admit");
          return;
        }
        target->slots[(size_t)k] = r.id;
        ++target->active;
        r.slot = k;
        r.tAdmit = now_s();
        std::vector<double>().swap(r.hist);
        it = pending.erase(it);
        progress = true;
      }
      // 3. next rounds
      {
        for (auto& pp : pools) {
          Pool& p = *pp;
          if (p.busy || p.active == 0 || p.grouped) continue;
          if (wdl_solver_projection_round(p.s)) {
            fail("wdl_solver_projection_round");
            return;
          }
          p.busy = true;
          progress = true;
          std::lock_guard<std::mutex> g(mu);
          ++stats[5];
        }
        // every pool of an idle class with an active slot runs a round: the
        // class's pools as one group (made again when its pools changed;
        // pools with no active slot sit out its launches at no cost)
        for (auto& kv : classes) {
          ClassRun& c = kv.second;
          if (c.busy) continue;
          std::vector<Pool*> ms;
          bool any = false;
          for (auto& pp : pools)
            if (pp->grouped && pp->lclass == kv.first) {
              ms.push_back(pp.get());
              any = any || pp->active > 0;
            }
          if (!any) continue;
          if (ms != c.members) {
            drop_group(c);
            std::vector<wdl_solver*> ss;
            for (Pool* p : ms) ss.push_back(p->s);
            if (wdl_group_create(c.ctx, ss.data(), (int)ss.size(), &c.group)) {
              fail("wdl_group_create");
              return;
            }
            c.members = ms;
          }
          if (wdl_group_projection_round(c.group)) {
            fail("wdl_group_projection_round");
            return;
          }
          c.busy = true;
          for (Pool* p : ms) p->busy = true;
          progress = true;
          std::lock_guard<std::mutex> g(mu);
          ++stats[5];
        }
      }
      if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
};

static long long hbm_budget_of(const wdl_server* sv) { return sv->budget; }

extern "C" {

int wdl_server_create(int device, long long hbm_budget_bytes, int max_slots_per_pool,
                      wdl_server** out) {
  WDL_CHECK(out, WDL_ERR_ARG, "wdl_server_create: NULL argument");
  WDL_CHECK(hbm_budget_bytes > 0, WDL_ERR_ARG, "wdl_server_create: budget must be positive");
  WDL_CHECK(max_slots_per_pool >= 1 && max_slots_per_pool <= 4096, WDL_ERR_ARG,
            "wdl_server_create: max_slots_per_pool must be in 1..4096");
  int n = 0;
  WDL_HIP(hipGetDeviceCount(&n));
  WDL_CHECK(device >= 0 && device < n, WDL_ERR_ARG, "wdl_server_create: no such device");
  wdl_server* sv = new wdl_server();
  sv->device = device;
  sv->budget = hbm_budget_bytes;
  sv->maxSlots = max_slots_per_pool;
  sv->worker = std::thread([sv] {
    (void)hipSetDevice(sv->device);
    sv->loop();
  });
  *out = sv;
  return WDL_OK;
}

int wdl_server_destroy(wdl_server* sv) {
  if (!sv) return WDL_OK;
  sv->stop.store(true);
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipSetDevice(sv->device);
  while (!sv->pools.empty()) sv->release_pool(sv->pools.size() - 1);
  for (auto& kv : sv->classes) {
    sv->drop_group(kv.second);
    if (kv.second.ctx) wdl_ctx_destroy(kv.second.ctx);
  }
  delete sv;
  return WDL_OK;
}

int wdl_server_set_batching(wdl_server* sv, int mixed) {
  WDL_CHECK(sv, WDL_ERR_ARG, "wdl_server_set_batching: NULL argument");
  WDL_CHECK(mixed == WDL_SERVER_POOLS || mixed == WDL_SERVER_MIXED, WDL_ERR_ARG,
            "wdl_server_set_batching: mode must be WDL_SERVER_POOLS or WDL_SERVER_MIXED");
  std::lock_guard<std::mutex> g(sv->mu);
  WDL_CHECK(sv->nextId == 1, WDL_ERR_STATE,
            "wdl_server_set_batching: must be called before the first request");
  sv->mixed.store(mixed);
  return WDL_OK;
}

int wdl_server_add_dictionary(wdl_server* sv, int W, int H, int K, double gamma, int loss,
                              int max_iterations, int flags, const double* atoms, int* dict_id) {
  WDL_CHECK(sv && atoms && dict_id, WDL_ERR_ARG, "wdl_server_add_dictionary: NULL argument");
  WDL_CHECK(W > 0 && H > 0 && K > 0 && max_iterations > 0, WDL_ERR_ARG,
            "wdl_server_add_dictionary: W, H, K, max_iterations must be positive");
  WDL_CHECK((flags & ~(WDL_FLAG_LOG_DOMAIN | WDL_FLAG_EXACT_LSE)) == 0, WDL_ERR_ARG,
            "wdl_server_add_dictionary: flags may be WDL_FLAG_LOG_DOMAIN, WDL_FLAG_EXACT_LSE");
  const long long slot = wdl_solver_footprint(W, H, K, 1, max_iterations, flags | kPoolFlags);
  WDL_CHECK(slot > 0, WDL_ERR_ARG, "wdl_server_add_dictionary: bad sizes");
  WDL_CHECK(slot <= hbm_budget_of(sv), WDL_ERR_NOMEM,
            "wdl_server_add_dictionary: one request (" + std::to_string(slot) +
                " bytes of histories) exceeds the budget");
  std::lock_guard<std::mutex> g(sv->mu);
  std::shared_ptr<Dict> d(new Dict());
  Dict& dd = *d;
  dd.W = W;
  dd.H = H;
  dd.K = K;
  dd.loss = loss;
  dd.maxL = max_iterations;
  dd.flags = flags;
  dd.gamma = gamma;
  dd.atoms.assign(atoms, atoms + (size_t)K * W * H);
  sv->dicts.push_back(std::move(d));  // pools are created by the scheduler when requests come
  *dict_id = (int)sv->dicts.size() - 1;
  return WDL_OK;
}

int wdl_server_submit(wdl_server* sv, int dict_id, const double* histogram, int iterations,
                      int rounds, double step, long long* request_id) {
  WDL_CHECK(sv && histogram && request_id, WDL_ERR_ARG, "wdl_server_submit: NULL argument");
  std::lock_guard<std::mutex> g(sv->mu);
  WDL_CHECK(sv->error.empty(), WDL_ERR_HIP, "wdl_server_submit: server failed: " + sv->error);
  WDL_CHECK(dict_id >= 0 && dict_id < (int)sv->dicts.size(), WDL_ERR_ARG,
            "wdl_server_submit: unknown dictionary");
  const Dict& D = *sv->dicts[(size_t)dict_id];
  WDL_CHECK(iterations >= 1 && iterations <= D.maxL, WDL_ERR_ARG,
            "wdl_server_submit: iterations must be in 1..max_iterations of the dictionary");
  WDL_CHECK(rounds >= 1, WDL_ERR_ARG, "wdl_server_submit: rounds must be >= 1");
  WDL_CHECK(step >= 0 && std::isfinite(step), WDL_ERR_ARG, "wdl_server_submit: bad step");
  Request r;
  r.id = sv->nextId++;
  r.dict = dict_id;
  r.K = D.K;
  r.lclass = iteration_class(iterations, D.maxL);
  r.hist.assign(histogram, histogram + (size_t)D.W * D.H);
  r.iters = iterations;
  r.rounds = rounds;
  r.step = step;
  r.tSubmit = now_s();
  *request_id = r.id;
  sv->incoming.push_back(std::move(r));
  return WDL_OK;
}

int wdl_server_wait(wdl_server* sv, long long request_id, double timeout_s, double* weights,
                    double* loss, double* latency_ms, double* queue_ms) {
  WDL_CHECK(sv, WDL_ERR_ARG, "wdl_server_wait: NULL argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  const auto deadline =
      Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(
                         timeout_s > 0 ? timeout_s : 0.0));
  while (sv->done.find(request_id) == sv->done.end()) {
    WDL_CHECK(sv->error.empty(), WDL_ERR_HIP, "wdl_server_wait: server failed: " + sv->error);
    if (sv->cvDone.wait_until(lk, deadline) == std::cv_status::timeout &&
        sv->done.find(request_id) == sv->done.end())
      return 1;
  }
  Request& r = sv->done.at(request_id);
  if (weights) std::copy(r.weights.begin(), r.weights.end(), weights);
  if (loss) *loss = r.loss;
  if (latency_ms) *latency_ms = (r.tDone - r.tSubmit) * 1e3;
  if (queue_ms) *queue_ms = (r.tAdmit - r.tSubmit) * 1e3;
  sv->done.erase(request_id);
  return WDL_OK;
}

int wdl_server_stats(wdl_server* sv, long long* out6) {
  WDL_CHECK(sv && out6, WDL_ERR_ARG, "wdl_server_stats: NULL argument");
  std::lock_guard<std::mutex> g(sv->mu);
  std::copy(sv->stats, sv->stats + 6, out6);
  return WDL_OK;
}

}  // extern "C"
