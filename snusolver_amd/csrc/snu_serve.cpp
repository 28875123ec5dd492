// Serving loop over resident factorizations (no reference counterpart).
//
// A server keeps factorized matrices resident in HBM under a byte budget and
// answers solve requests (matrix id, b, nrhs) submitted from any thread.
// Worker threads ("slots") each own a batch over the resident set and its
// stream: a free slot takes every pending request (up to a column cap),
// packs them -- several requests for one matrix become columns of one
// multi-rhs solve, different matrices with uneven column counts share one
// launch per sweep step and tree height -- and runs the packed solve.
// Requests that arrive while slots are busy wait for the next free slot;
// slots run concurrently, so their packed solves overlap on the device.
// New factorizations (snuserve_add) are analysed on the host, admitted
// against the budget by evicting least-recently-used idle matrices, then
// uploaded and factorized on the caller's thread while slots keep serving.
// An evicted matrix is parked, not forgotten: its device memory is released
// but its symbolic plan and a host copy of its values stay, and the next
// request for it admits it again (upload and factorization, no analysis) on
// the submitting thread before it is queued.
//
// Built from the public C-ABI only (snumat_*, snubatch_*).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "snu_internal.h"

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  int id = 0;
  const double *b = nullptr;
  double *x = nullptr;
  int nrhs = 0, ld = 0;
  double t_submit = 0, t_done = 0;
  int status = 0;
  bool done = false;
};

struct Resident {
  snumat_t *h = nullptr;
  int n = 0;
  size_t bytes = 0;
  double last_use = 0;
  int busy = 0;               // pending or in-flight requests
  bool parked = false;        // evicted: host-only handle, no device memory
  std::vector<double> vals;   // A's values (for admitting a parked matrix again)
};

}  // namespace

struct snuserve {
  int device = 0, prec = -1, max_cols = 256;
  size_t budget = 0, used = 0;
  std::mutex mu;                 // requests, resident bookkeeping, stats
  std::condition_variable cv_work, cv_done;
  std::shared_mutex set_mu;      // shared: a packed solve; exclusive: eviction
  std::mutex add_mu;             // one admission at a time
  std::map<int, Resident> res;  // resident and parked matrices
  int next_id = 1;
  unsigned long long generation = 0;
  std::deque<long long> pending;
  std::unordered_map<long long, Request> reqs;
  long long next_ticket = 1;
  std::vector<std::thread> workers;
  bool stop = false;
  // statistics
  std::vector<double> lat;
  double cols = 0, packs = 0, evictions = 0, admissions = 0, readmissions = 0, busy_ms = 0;
  double t_first = 0, t_last = 0;
};

namespace {

// One slot: a batch over the resident set (rebuilt when the set changes).
void worker(snuserve *s) {
  snubatch_t *batch = nullptr;
  unsigned long long gen = ~0ull;
  std::vector<int> ids;          // batch member order
  std::vector<snumat_t *> mats;  // batch members
  void *dbuf = nullptr;          // device staging of packed b and x (grow-only)
  size_t dcap = 0;
  for (;;) {
    std::vector<long long> take;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
      if (s->pending.empty()) break;  // stop requested and nothing left
      int cols = 0;
      while (!s->pending.empty()) {
        const Request &r = s->reqs[s->pending.front()];
        if (!take.empty() && cols + r.nrhs > s->max_cols) break;
        cols += r.nrhs;
        take.push_back(s->pending.front());
        s->pending.pop_front();
      }
    }
    const double t0 = now_ms();
    int rc = SNU_OK;
    {
      std::shared_lock<std::shared_mutex> sl(s->set_mu);
      {
        std::lock_guard<std::mutex> lk(s->mu);
        if (gen != s->generation) {
          ids.clear();
          mats.clear();
          for (auto &kv : s->res)
            if (!kv.second.parked) ids.push_back(kv.first), mats.push_back(kv.second.h);
          gen = s->generation;
          if (batch) snubatch_destroy(batch);
          batch = nullptr;
        }
      }
      if (!batch && !ids.empty()) rc = snubatch_create(mats.data(), (int)mats.size(), &batch);
      if (!rc && batch) {
        // pack: per member, the columns of its requests side by side in a
        // device staging buffer (b, then x), copied straight from and to
        // the requests' host arrays
        const int nm = (int)ids.size();
        std::unordered_map<int, int> slot;
        for (int m = 0; m < nm; m++) slot[ids[m]] = m;
        struct Item { int m, c0, nrhs, ld, n; const double *b; double *x; };
        std::vector<Item> items(take.size(), Item{-1, 0, 0, 0, 0, nullptr, nullptr});
        std::vector<int> nrhs(nm, 0), ld(nm, 1), nn(nm, 0);
        {
          std::lock_guard<std::mutex> lk(s->mu);
          for (int m = 0; m < nm; m++) nn[m] = s->res[ids[m]].n, ld[m] = nn[m];
          for (size_t q = 0; q < take.size(); q++) {
            const Request &r = s->reqs[take[q]];
            auto it = slot.find(r.id);
            if (it == slot.end()) continue;
            const int m = it->second;
            items[q] = Item{m, nrhs[m], r.nrhs, r.ld, nn[m], r.b, r.x};
            nrhs[m] += r.nrhs;
          }
        }
        std::vector<size_t> off(nm, 0);
        size_t tot = 0;
        for (int m = 0; m < nm; m++) off[m] = tot, tot += (size_t)nn[m] * nrhs[m];
        if (2 * tot * sizeof(double) > dcap) {
          if (dbuf) snu_dev_free(dbuf);
          dbuf = nullptr;
          dcap = 0;
          rc = snu_dev_alloc(2 * tot * sizeof(double) + 1024, &dbuf);
          if (!rc) dcap = 2 * tot * sizeof(double) + 1024;
        }
        double *db = (double *)dbuf, *dx = db + tot;
        auto copy_cols = [&](const Item &it, bool up) -> int {
          double *dev = (up ? db : dx) + off[it.m] + (size_t)it.c0 * it.n;
          if (it.ld == it.n)
            return up ? snu_dev_copy(dev, it.b, sizeof(double) * it.n * it.nrhs, 1)
                      : snu_dev_copy(it.x, dev, sizeof(double) * it.n * it.nrhs, 2);
          for (int j = 0; j < it.nrhs; j++) {
            int e = up ? snu_dev_copy(dev + (size_t)j * it.n, it.b + (size_t)j * it.ld, sizeof(double) * it.n, 1)
                       : snu_dev_copy(it.x + (size_t)j * it.ld, dev + (size_t)j * it.n, sizeof(double) * it.n, 2);
            if (e) return e;
          }
          return SNU_OK;
        };
        for (size_t q = 0; q < items.size() && !rc; q++)
          if (items[q].m >= 0) rc = copy_cols(items[q], true);
        if (!rc) {
          std::vector<const double *> pb(nm, nullptr);
          std::vector<double *> px(nm, nullptr);
          for (int m = 0; m < nm; m++)
            if (nrhs[m]) pb[m] = db + off[m], px[m] = dx + off[m];
          rc = snubatch_solve_device(batch, pb.data(), px.data(), nrhs.data(), ld.data());
        }
        if (!rc || rc == SNU_ERR_NOT_CONVERGED)
          for (size_t q = 0; q < items.size(); q++)
            if (items[q].m >= 0) {
              int e = copy_cols(items[q], false);
              if (e) rc = e;
            }
      }
    }
    const double t1 = now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (long long t : take) {
        Request &r = s->reqs[t];
        auto it = s->res.find(r.id);
        if (it == s->res.end()) {
          r.status = SNU_ERR_BAD_INPUT;
        } else {
          r.status = rc;
          it->second.busy--;
          it->second.last_use = t1;
        }
        r.t_done = t1;
        r.done = true;
        s->lat.push_back(r.t_done - r.t_submit);
        s->cols += r.nrhs;
      }
      s->packs += 1;
      s->busy_ms += t1 - t0;
      s->t_last = std::max(s->t_last, t1);
    }
    s->cv_done.notify_all();
  }
  if (batch) snubatch_destroy(batch);
  if (dbuf) snu_dev_free(dbuf);
}

}  // namespace

extern "C" int snuserve_create(int device, size_t budget_bytes, int slots, int max_cols,
                               snuserve_t **out) {
  if (!out || slots < 1 || slots > 8 || max_cols < 1) return snu::set_error(SNU_ERR_BAD_INPUT, "bad server arguments");
  int rc = snu_initialize(device);
  if (rc) return rc;
  snuserve *s = new snuserve();
  s->device = device;
  s->budget = budget_bytes;
  s->max_cols = max_cols;
  for (int i = 0; i < slots; i++) s->workers.emplace_back(worker, s);
  *out = s;
  return SNU_OK;
}

namespace {

// Room for `need` more bytes: park idle least-recently-used resident
// matrices until they fit, waiting for busy ones to finish when every
// resident matrix is busy; the bytes are then reserved.  Called with add_mu
// held (one admission at a time).
void make_room(snuserve *s, size_t need) {
  for (;;) {
    {
      std::unique_lock<std::shared_mutex> ex(s->set_mu);  // no packed solve in flight
      std::lock_guard<std::mutex> lk(s->mu);
      while (s->used + need > s->budget) {
        int best = -1;
        double t = 0;
        for (auto &kv : s->res)
          if (!kv.second.parked && kv.second.busy == 0 && (best < 0 || kv.second.last_use < t))
            best = kv.first, t = kv.second.last_use;
        if (best < 0) break;
        Resident &v = s->res[best];
        (void)snumat_detach(v.h);
        v.parked = true;
        s->used -= v.bytes;
        s->evictions += 1;
        s->generation++;
      }
      if (s->used + need <= s->budget) {
        s->used += need;
        return;
      }
    }
    // every remaining matrix is busy: wait for a request to finish
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
  }
}

}  // namespace

// Admission: host analysis, room made under the budget, upload,
// factorization.
extern "C" int snuserve_add(snuserve_t *s, const snu_csr_t *A, const snu_opts_t *opts, int *id) {
  if (!s || !A || !opts || !id) return snu::set_error(SNU_ERR_BAD_INPUT, "null");
  std::lock_guard<std::mutex> admit(s->add_mu);
  if (s->prec >= 0 && opts->precision != s->prec)
    return snu::set_error(SNU_ERR_BAD_INPUT, "a server holds one precision");
  snu_opts_t o = *opts;
  o.device = -1;
  snumat_t *h = nullptr;
  int rc = snumat_construct(A, &o, &h);
  if (rc) return rc;
  size_t need = 0;
  snumat_bytes_needed(h, &need);
  if (need > s->budget) {
    snumat_destroy(h);
    return snu::set_error(SNU_ERR_OOM, "matrix needs " + std::to_string(need) + " bytes, budget " +
                                           std::to_string(s->budget));
  }
  make_room(s, need);
  rc = snumat_attach(h, s->device, A->data);
  if (!rc) rc = snumat_factorize(h);
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc) {
    s->used -= need;
    snumat_destroy(h);
    return rc;
  }
  Resident r;
  r.h = h;
  r.n = A->n;
  r.bytes = need;
  r.last_use = now_ms();
  r.vals.assign(A->data, A->data + A->nnz);
  *id = s->next_id++;
  s->res[*id] = std::move(r);
  s->prec = opts->precision;
  s->generation++;
  s->admissions += 1;
  return SNU_OK;
}

namespace {
// A request for a parked matrix: admit it again (room, upload of the kept
// values, factorization).  Returns with the matrix resident or an error.
int readmit(snuserve *s, int id) {
  std::lock_guard<std::mutex> admit(s->add_mu);
  snumat_t *h = nullptr;
  const double *vals = nullptr;
  size_t need = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->res.find(id);
    if (it == s->res.end()) return snu::set_error(SNU_ERR_BAD_INPUT, "unknown matrix id");
    if (!it->second.parked) return SNU_OK;  // admitted again by another request
    h = it->second.h;
    vals = it->second.vals.data();
    need = it->second.bytes;
  }
  make_room(s, need);
  int rc = snumat_attach(h, s->device, vals);
  if (!rc) rc = snumat_factorize(h);
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc) {
    (void)snumat_detach(h);
    s->used -= need;
    return rc;
  }
  Resident &r = s->res[id];
  r.parked = false;
  r.last_use = now_ms();
  s->generation++;
  s->readmissions += 1;
  return SNU_OK;
}
}  // namespace

extern "C" int snuserve_remove(snuserve_t *s, int id) {
  if (!s) return snu::set_error(SNU_ERR_BAD_INPUT, "null");
  std::unique_lock<std::shared_mutex> ex(s->set_mu);
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->res.find(id);
  if (it == s->res.end()) return snu::set_error(SNU_ERR_BAD_INPUT, "unknown matrix id");
  if (it->second.busy) return snu::set_error(SNU_ERR_BAD_INPUT, "matrix has requests in flight");
  snumat_destroy(it->second.h);
  if (!it->second.parked) s->used -= it->second.bytes;
  s->res.erase(it);
  s->generation++;
  return SNU_OK;
}

extern "C" int snuserve_submit(snuserve_t *s, int id, const double *b, double *x, int nrhs, int ld,
                               long long *ticket) {
  if (!s || !b || !x || nrhs < 1 || !ticket) return snu::set_error(SNU_ERR_BAD_INPUT, "bad request");
  for (;;) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->res.find(id);
      if (it == s->res.end()) return snu::set_error(SNU_ERR_BAD_INPUT, "unknown matrix id");
      if (ld < it->second.n) return snu::set_error(SNU_ERR_BAD_INPUT, "ld < n");
      if (nrhs > s->max_cols) return snu::set_error(SNU_ERR_BAD_INPUT, "nrhs above the server's column cap");
      if (!it->second.parked) {
        it->second.busy++;  // from here on it cannot be parked
        break;
      }
    }
    int rc = readmit(s, id);
    if (rc) return rc;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    Request r;
    r.id = id;
    r.b = b;
    r.x = x;
    r.nrhs = nrhs;
    r.ld = ld;
    r.t_submit = now_ms();
    if (s->t_first == 0) s->t_first = r.t_submit;
    *ticket = s->next_ticket++;
    s->reqs[*ticket] = r;
    s->pending.push_back(*ticket);
  }
  s->cv_work.notify_one();
  return SNU_OK;
}

extern "C" int snuserve_wait(snuserve_t *s, long long ticket, double *latency_ms) {
  if (!s) return snu::set_error(SNU_ERR_BAD_INPUT, "null");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return snu::set_error(SNU_ERR_BAD_INPUT, "unknown ticket");
  s->cv_done.wait(lk, [&] { return it->second.done; });
  const int st = it->second.status;
  if (latency_ms) *latency_ms = it->second.t_done - it->second.t_submit;
  s->reqs.erase(it);
  if (st) return snu::set_error(st, "request failed with status " + std::to_string(st));
  return SNU_OK;
}

// st[0] requests answered, st[1] columns solved, st[2] packed solves,
// st[3] p50 latency ms, st[4] p99 ms, st[5] mean ms, st[6] max ms,
// st[7] evictions, st[8] admissions, st[9] resident matrices, st[10] bytes
// reserved, st[11] wall ms from the first submit to the last answer,
// st[12] summed slot busy ms, st[13] admissions of parked matrices,
// st[14] parked matrices.  reset != 0 clears the latency record.
extern "C" int snuserve_stats(snuserve_t *s, double *st16, int reset) {
  if (!s || !st16) return snu::set_error(SNU_ERR_BAD_INPUT, "null");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st16, 0, 16 * sizeof(double));
  std::vector<double> l = s->lat;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) {
    if (l.empty()) return 0.0;
    size_t k = (size_t)std::min<double>(l.size() - 1, std::max(0.0, p * l.size() - 1e-9));
    return l[k];
  };
  double sum = 0;
  for (double v : l) sum += v;
  st16[0] = (double)l.size();
  st16[1] = s->cols;
  st16[2] = s->packs;
  st16[3] = pct(0.50);
  st16[4] = pct(0.99);
  st16[5] = l.empty() ? 0 : sum / l.size();
  st16[6] = l.empty() ? 0 : l.back();
  st16[7] = s->evictions;
  st16[8] = s->admissions;
  int resident = 0;
  for (auto &kv : s->res) resident += !kv.second.parked;
  st16[9] = resident;
  st16[10] = (double)s->used;
  st16[11] = s->t_last > s->t_first ? s->t_last - s->t_first : 0;
  st16[12] = s->busy_ms;
  st16[13] = s->readmissions;
  st16[14] = (double)s->res.size() - resident;
  if (reset) {
    s->lat.clear();
    s->cols = s->packs = s->busy_ms = 0;
    s->t_first = s->t_last = 0;
  }
  return SNU_OK;
}

extern "C" int snuserve_destroy(snuserve_t *s) {
  if (!s) return SNU_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &kv : s->res) snumat_destroy(kv.second.h);
  delete s;
  return SNU_OK;
}
