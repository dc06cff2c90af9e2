// Serving engine of libt3x: continuous admission of topology optimisations
// by device footprint, one worker thread + one t3x_problem per running
// request.  Host code only; the device work is the same as a t3x_run call's.
//
// Device memory is one pool allocated when the server starts (the budget):
// an admitted request gets a contiguous range of its footprint (first fit,
// ranges coalesced when released) and its problem carves its arrays out of
// it, so serving makes no hipMalloc or hipFree per request (a hipFree can
// wait for the whole device, including the other requests' work).  Each
// worker keeps one HIP stream and one pinned read-back buffer for all the
// requests it runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "../../include/t3x.h"
#include "../../include/t3x_serve.h"
#include "t3x_pool.h"

namespace {

struct Pending {
  t3x_request req;
  long long bytes;
  double tSubmit;
  size_t offset = 0;  // the pool range, once admitted
};

// Worker-owned resources reused by every request the worker runs.
struct WorkerRes {
  hipStream_t stream = nullptr;
  double* hostBuf = nullptr;
};

struct Finished {
  t3x_completion c{};
  std::vector<double> xPhys, cHist;
  bool returned = false;
};

}  // namespace

struct t3x_server {
  t3x_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  std::mutex m;
  std::condition_variable cvWork, cvDone;
  std::deque<Pending> queue;
  std::map<long long, Finished> done;  // by tag
  std::deque<long long> readyOrder;    // finished, not yet polled
  std::set<long long> liveTags;        // queued, running or finished and not yet released
  std::vector<std::thread> workers;
  long long memInUse = 0;
  int running = 0;
  bool stopping = false;
  char* pool = nullptr;                 // memBudget bytes of device memory
  std::map<size_t, size_t> freeRanges;  // offset -> size, disjoint, coalesced

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  void worker();
  void run_one(const Pending& p, const WorkerRes& w, Finished& f);
  static size_t rounded(long long bytes) { return ((size_t)bytes + 255) & ~(size_t)255; }
  // first free range holding `bytes`, or freeRanges.end()
  std::map<size_t, size_t>::iterator find_fit(long long bytes) {
    for (auto it = freeRanges.begin(); it != freeRanges.end(); ++it)
      if (it->second >= rounded(bytes)) return it;
    return freeRanges.end();
  }
  size_t take(std::map<size_t, size_t>::iterator it, long long bytes) {
    const size_t off = it->first, sz = it->second, need = rounded(bytes);
    freeRanges.erase(it);
    if (sz > need) freeRanges[off + need] = sz - need;
    return off;
  }
  void give(size_t off, long long bytes) {
    size_t sz = rounded(bytes);
    auto next = freeRanges.lower_bound(off);
    if (next != freeRanges.end() && off + sz == next->first) {
      sz += next->second;
      next = freeRanges.erase(next);
    }
    if (next != freeRanges.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += sz;
        return;
      }
    }
    freeRanges[off] = sz;
  }
};

void t3x_server::run_one(const Pending& p, const WorkerRes& w, Finished& f) {
  const t3x_request& r = p.req;
  const int ny = r.nely, nx = r.nelx, nz = r.nelz;
  std::vector<uint8_t> vox((size_t)ny * nx * nz, 1);
  // built-in supports (536-541) and load patch (552-556) on the unpadded node grid
  std::vector<int> fixN, loadN;
  for (int jz = 0; jz <= nz; ++jz)
    for (int jy = 0; jy <= ny; ++jy) fixN.push_back(jy + (ny + 1) * (0 + (nx + 1) * jz));
  const int zTop = (int)std::llround(nz / 6.0 + 1.0);
  for (int jz = 0; jz < zTop; ++jz)
    for (int jy = 0; jy <= ny; ++jy) loadN.push_back(jy + (ny + 1) * (nx + (nx + 1) * jz));
  std::vector<uint8_t> fixS(3 * fixN.size(), 1);
  double d[3] = {r.loadDir[0], r.loadDir[1], r.loadDir[2]};
  const double nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (nrm > 0)
    for (double& v : d) v /= nrm;
  std::vector<double> loadV(3 * loadN.size());
  for (size_t i = 0; i < loadN.size(); ++i)
    for (int c = 0; c < 3; ++c) loadV[3 * i + c] = d[c] * r.loadMag / (double)loadN.size();
  t3x_problem* h = nullptr;
  t3x::PoolSlot slot;
  slot.base = pool + p.offset;
  slot.size = rounded(p.bytes);
  slot.stream = w.stream;
  slot.hostBuf = w.hostBuf;
  slot.oneLaunchCoarse = true;  // bounded process-wide by device occupancy (t3x_host)
  int rc = t3x::create_pooled(vox.data(), ny, nx, nz, (int)fixN.size(), fixN.data(), fixS.data(), (int)loadN.size(),
                              loadN.data(), loadV.data(), &cfg.params, slot, &h);
  f.c.tag = r.tag;
  f.c.footprint = p.bytes;
  if (rc == T3X_OK) {
    int info[1 + 8 * 16];
    t3x_info(h, info);
    const int nE = info[1 + 4];
    f.xPhys.assign(nE, 0.0);
    f.cHist.assign(std::max(r.nLoop, 1), 0.0);
    std::vector<int> its(std::max(r.nLoop, 1), 0);
    t3x_result R{};
    R.xPhys = f.xPhys.data();
    R.cHist = f.cHist.data();
    R.lssIts = its.data();
    rc = t3x_run(h, r.mode, r.V0, r.nLoop, r.rMin, r.rHat, &R);
    f.c.nE = nE;
    f.c.loops = R.loops;
    f.c.cSolid = R.cSolid;
    f.cHist.resize(R.loops);
    f.c.cFinal = R.loops ? f.cHist[R.loops - 1] : 0.0;
    for (int i = 0; i < R.loops; ++i) f.c.pcgIts += its[i];
  }
  if (h) t3x_destroy(h);
  f.c.status = rc;
  f.c.xPhys = f.xPhys.data();
  f.c.cHist = f.cHist.data();
}

void t3x_server::worker() {
  WorkerRes w;
  bool ok = hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking) == hipSuccess &&
            hipHostMalloc((void**)&w.hostBuf, 256 * sizeof(double), hipHostMallocDefault) == hipSuccess;
  for (;;) {
    Pending p;
    {
      std::unique_lock<std::mutex> lk(m);
      // arrival order: the head request waits until a free range of the pool
      // holds its footprint
      cvWork.wait(lk, [&] {
        return (stopping && queue.empty()) || (!queue.empty() && find_fit(queue.front().bytes) != freeRanges.end());
      });
      if (queue.empty()) break;
      p = queue.front();
      queue.pop_front();
      p.offset = take(find_fit(p.bytes), p.bytes);
      memInUse += p.bytes;
      ++running;
    }
    Finished f;
    f.c.tSubmit = p.tSubmit;
    f.c.tStart = now();
    if (ok) {
      run_one(p, w, f);
    } else {
      f.c.tag = p.req.tag;
      f.c.status = T3X_ERR_HIP;
    }
    f.c.tEnd = now();
    {
      std::lock_guard<std::mutex> lk(m);
      give(p.offset, p.bytes);
      memInUse -= p.bytes;
      --running;
      const long long tag = f.c.tag;
      done[tag] = std::move(f);
      done[tag].c.xPhys = done[tag].xPhys.data();
      done[tag].c.cHist = done[tag].cHist.data();
      readyOrder.push_back(tag);
    }
    cvWork.notify_all();
    cvDone.notify_all();
  }
  if (w.stream) (void)hipStreamDestroy(w.stream);
  if (w.hostBuf) (void)hipHostFree(w.hostBuf);
}

extern "C" {

int t3x_server_create(const t3x_server_config* cfg, t3x_server** out) {
  if (!cfg || !out || cfg->maxActive < 1) return T3X_ERR_INPUT;
  if (cfg->params.precision < 0 || cfg->params.precision > 2) return T3X_ERR_INPUT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return T3X_ERR_NO_DEVICE;
  t3x_server* s = new t3x_server();
  s->cfg = *cfg;
  if (s->cfg.memBudget <= 0) {
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) {
      delete s;
      return T3X_ERR_HIP;
    }
    s->cfg.memBudget = (long long)(0.9 * (double)freeB);
  }
  s->cfg.memBudget &= ~255LL;
  if (hipMalloc((void**)&s->pool, (size_t)s->cfg.memBudget) != hipSuccess) {
    delete s;
    return T3X_ERR_HIP;
  }
  s->freeRanges[0] = (size_t)s->cfg.memBudget;
  s->t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < cfg->maxActive; ++i) s->workers.emplace_back([s] { s->worker(); });
  *out = s;
  return T3X_OK;
}

int t3x_server_submit(t3x_server* s, const t3x_request* req) {
  if (!s || !req || req->nely < 3 || req->nelx < 3 || req->nelz < 3 || (req->mode != 0 && req->mode != 1) ||
      req->nLoop < 0)
    return T3X_ERR_INPUT;
  long long bytes = 0;
  int rc = t3x_footprint(nullptr, req->nely, req->nelx, req->nelz, &s->cfg.params, req->mode, &bytes);
  if (rc != T3X_OK) return rc;
  if ((long long)t3x_server::rounded(bytes) > s->cfg.memBudget) return T3X_ERR_INPUT;
  {
    std::lock_guard<std::mutex> lk(s->m);
    if (s->stopping) return T3X_ERR_INPUT;
    // a tag names one request until its completion is released
    if (!s->liveTags.insert(req->tag).second) return T3X_ERR_INPUT;
    s->queue.push_back(Pending{*req, bytes, s->now()});
  }
  s->cvWork.notify_all();
  return T3X_OK;
}

int t3x_server_poll(t3x_server* s, t3x_completion* out, int max, int waitMs, int* n) {
  if (!s || !n || (max > 0 && !out)) return T3X_ERR_INPUT;
  std::unique_lock<std::mutex> lk(s->m);
  if (s->readyOrder.empty() && waitMs > 0)
    s->cvDone.wait_for(lk, std::chrono::milliseconds(waitMs), [&] { return !s->readyOrder.empty(); });
  int k = 0;
  while (k < max && !s->readyOrder.empty()) {
    const long long tag = s->readyOrder.front();
    s->readyOrder.pop_front();
    Finished& f = s->done[tag];
    f.returned = true;
    out[k++] = f.c;
  }
  *n = k;
  return T3X_OK;
}

int t3x_server_release(t3x_server* s, long long tag) {
  if (!s) return T3X_ERR_INPUT;
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->done.find(tag);
  if (it == s->done.end() || !it->second.returned) return T3X_ERR_INPUT;
  s->done.erase(it);
  s->liveTags.erase(tag);
  return T3X_OK;
}

int t3x_server_stats(t3x_server* s, int* queued, int* running, int* ready, long long* memInUse) {
  if (!s) return T3X_ERR_INPUT;
  std::lock_guard<std::mutex> lk(s->m);
  if (queued) *queued = (int)s->queue.size();
  if (running) *running = s->running;
  if (ready) *ready = (int)s->readyOrder.size();
  if (memInUse) *memInUse = s->memInUse;
  return T3X_OK;
}

void t3x_server_destroy(t3x_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stopping = true;
  }
  s->cvWork.notify_all();
  for (auto& t : s->workers) t.join();
  if (s->pool) (void)hipFree(s->pool);
  delete s;
}

}  // extern "C"
