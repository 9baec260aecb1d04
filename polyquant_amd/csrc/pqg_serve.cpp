// Serving loop: continuous batching of CI solves on one GPU.
//
// Requests (a CI problem + Davidson parameters) are queued with a submit time.  A scheduler
// admits them into `max_slots` resident slots as slots and HBM free up: admission reserves
// the request's estimated footprint (integral tables, CI vectors, Davidson subspace) against
// the context's budget, and may skip over a queued request that does not fit yet in favour
// of a later one that does.  Each slot owns a HIP stream (default priority) and a worker
// thread that runs the request's solve from setup to the final FP64 residual check; the solves
// of different slots overlap on the device, and a new request enters a slot as soon as the
// previous one retires.  A request whose setup or solve fails (out of memory, no convergence)
// is retired with that status and the loop keeps serving.  Every solve's arithmetic depends only
// on its own request (no shared reductions, no atomics), so a result is bit-identical alone
// and in the mix.
#include <cstdio>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>

#include "pqg_davidson.hpp"

namespace {

using namespace pqg;

struct OwnedDesc {
  pqg_ci_desc d{};
  std::vector<std::vector<double>> arrays;
  void copy_from(const pqg_ci_desc& s) {
    d = s;
    for (int p = 0; p < s.nspecies; p++) {
      const int n = s.n_orb[p];
      for (int sp = 0; sp < s.nspin_blocks[p]; sp++) {
        if (!s.h[p][sp]) continue;
        arrays.emplace_back(s.h[p][sp], s.h[p][sp] + (size_t)n * n);
        d.h[p][sp] = arrays.back().data();
      }
    }
    for (int a = 0; a < s.nspecies; a++)
      for (int sa = 0; sa < 2; sa++)
        for (int b = a; b < s.nspecies; b++)
          for (int sb = 0; sb < 2; sb++) {
            const double* src = s.eri[a][sa][b][sb];
            if (!src) continue;
            const size_t na = (size_t)s.n_orb[a] * (s.n_orb[a] + 1) / 2, nb = (size_t)s.n_orb[b] * (s.n_orb[b] + 1) / 2;
            arrays.emplace_back(src, src + na * nb);
            d.eri[a][sa][b][sb] = arrays.back().data();
          }
  }
};

struct Request {
  int64_t id;
  OwnedDesc desc;
  pqg_dav_params params;
  double t_submit;
  double t_admit = 0.0;
  uint64_t reserve = 0;
  double work = 0.0;  // determinants x roots: the admission order's size measure
};

// Requests waiting longer than this are admitted oldest first; below it, smallest work first.
constexpr double kMaxWaitS = 30.0;
constexpr int64_t kGraphMaxDet = 131072;  // determinants up to which a solve replays captured sigmas

}  // namespace

struct pqg_server {
  pqg_ctx* ctx = nullptr;
  int max_slots = 1;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::unique_ptr<Request>> queue;
  std::vector<std::unique_ptr<Request>> assigned;  // per slot, nullptr when free
  std::vector<pqg_result> done;
  std::vector<std::thread> workers;
  std::vector<hipStream_t> streams;
  uint64_t reserved = 0;
  int64_t next_id = 0;
  bool stop = false;
  int resident = 0;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

  // Reserve budget for a queued request that fits and hand it to a free slot.  Order: requests
  // that have waited longer than kMaxWaitS, oldest first; otherwise the smallest work first,
  // so short solves are not held behind long ones (ties: arrival order).
  void admit_locked() {
    for (int s = 0; s < max_slots; s++) {
      if (assigned[s]) continue;
      const uint64_t budget = ctx->arena->budget();
      const double t = now();
      auto best = queue.end();
      bool best_old = false;
      for (auto it = queue.begin(); it != queue.end(); ++it) {
        if (reserved + (*it)->reserve > budget) continue;
        const bool old = t - (*it)->t_submit > kMaxWaitS;
        if (best == queue.end() || (old && !best_old) ||
            (old == best_old && !old && (*it)->work < (*best)->work)) {
          best = it;
          best_old = old;
        }
      }
      auto it = best;
      if (it == queue.end()) {
        // nothing fits; a request larger than the whole budget can never run
        if (!queue.empty() && resident == 0) {
          auto big = std::find_if(queue.begin(), queue.end(), [&](const auto& r) { return r->reserve > budget; });
          if (big != queue.end()) {
            pqg_result res{};
            res.request_id = (*big)->id;
            res.status = PQG_ERR_OOM;
            std::snprintf(res.message, sizeof(res.message), "request needs more than the whole HBM budget");
            res.t_submit = (*big)->t_submit;
            res.t_admit = res.t_done = t;
            done.push_back(res);
            queue.erase(big);
            s--;  // try this slot again
            continue;
          }
        }
        return;
      }
      (*it)->t_admit = t;
      reserved += (*it)->reserve;
      resident++;
      assigned[s] = std::move(*it);
      queue.erase(it);
      cv.notify_all();
    }
  }

  void worker(int s) {
    (void)hipSetDevice(ctx->device);
    hipStream_t st = streams[s];
    while (true) {
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || (assigned[s] && !busy_[s]); });
        if (stop) return;
        r = assigned[s].get();
        busy_[s] = true;
      }
      pqg_result res{};
      res.request_id = r->id;
      res.t_submit = r->t_submit;
      res.t_admit = r->t_admit;
      res.nev = r->params.nev;
      pqg_status status = guard([&] {
        std::unique_ptr<pqg_ci> ci(create_ci(ctx, &r->desc.d, st));
        Davidson dav(ci.get(), r->params);
        // small spaces: the sigma is a few tens of short kernels, replayed from a captured graph
        if (ci->n_int <= kGraphMaxDet && !ci->xh) dav.use_graphs(true);
        while (!dav.done()) {
          dav.run_sigma(st);
          dav.post_sigma(st);
        }
        PQG_HIP(hipStreamSynchronize(st));
        const auto& e = dav.energies();
        for (int k = 0; k < dav.nev() && k < 64; k++) res.energies[k] = e[k];
        res.iterations = dav.report().iterations;
        if (dav.status() != PQG_OK) throw Error{dav.status(), "CI Calculation did not converge!"};
      });
      res.status = status;
      if (status != PQG_OK) std::snprintf(res.message, sizeof(res.message), "%s", pqg_last_error());
      res.t_done = now();
      std::lock_guard<std::mutex> lk(mu);
      done.push_back(res);
      reserved -= r->reserve;
      resident--;
      assigned[s].reset();
      busy_[s] = false;
      admit_locked();
    }
  }

  std::vector<char> busy_;
};

extern "C" {

pqg_status pqg_serve_create(pqg_ctx* ctx, int max_slots, pqg_server** out) {
  return guard([&] {
    PQG_REQUIRE(ctx && out && max_slots >= 1 && max_slots <= 64, PQG_ERR_ARG, "invalid argument");
    auto* s = new pqg_server();
    s->ctx = ctx;
    s->max_slots = max_slots;
    s->t0 = std::chrono::steady_clock::now();
    s->assigned.resize(max_slots);
    s->busy_.assign(max_slots, 0);
    s->streams.resize(max_slots);
    for (int i = 0; i < max_slots; i++) PQG_HIP(hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking));
    for (int i = 0; i < max_slots; i++) s->workers.emplace_back([s, i] { s->worker(i); });
    *out = s;
  });
}

pqg_status pqg_serve_destroy(pqg_server* srv) {
  return guard([&] {
    if (!srv) return;
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->stop = true;
    }
    srv->cv.notify_all();
    for (auto& t : srv->workers) t.join();
    for (auto st : srv->streams) (void)hipStreamDestroy(st);
    delete srv;
  });
}

pqg_status pqg_serve_submit(pqg_server* srv, const pqg_ci_desc* desc, const pqg_dav_params* params, int64_t* id) {
  return guard([&] {
    PQG_REQUIRE(srv && desc && params, PQG_ERR_ARG, "NULL argument");
    auto r = std::make_unique<Request>();
    r->desc.copy_from(*desc);
    r->params = *params;
    r->reserve = estimate_ci_bytes(desc, params);
    double ndet = 1.0;
    for (int p = 0; p < desc->nspecies; p++)
      for (int sp = 0; sp < 2; sp++) {
        const int n = desc->n_orb[p], k = sp == 0 ? desc->n_alpha[p] : desc->n_beta[p];
        double c = 1.0;
        for (int i = 1; i <= k; i++) c = c * (n - k + i) / i;
        ndet *= c;
      }
    r->work = ndet * std::max(1, params->nev);
    std::lock_guard<std::mutex> lk(srv->mu);
    r->id = srv->next_id++;
    r->t_submit = srv->now();
    if (id) *id = r->id;
    srv->queue.push_back(std::move(r));
    srv->admit_locked();
  });
}

pqg_status pqg_serve_step(pqg_server* srv, pqg_result* results, int cap, int* nresults) {
  return guard([&] {
    PQG_REQUIRE(srv && nresults, PQG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->admit_locked();
    int n = 0;
    while (n < cap && !srv->done.empty()) {
      results[n++] = srv->done.front();
      srv->done.erase(srv->done.begin());
    }
    *nresults = n;
  });
}

pqg_status pqg_serve_stats(pqg_server* srv, int* queued, int* resident, uint64_t* bytes_resident) {
  return guard([&] {
    PQG_REQUIRE(srv, PQG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (queued) *queued = (int)srv->queue.size();
    if (resident) *resident = srv->resident;
    if (bytes_resident) *bytes_resident = srv->reserved;
  });
}

double pqg_serve_clock(pqg_server* srv) { return srv ? srv->now() : 0.0; }

}  // extern "C"
