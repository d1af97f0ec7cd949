// Batch of independent timesteps: many convection instances (each its own
// mesh size, aspect ratio, Rayleigh number and viscosity contrast) stepped
// together on one GPU.  Every instance is a device-resident Sim with its own
// stream (default priority); host worker threads take instances from a shared
// counter and run their steps, so the small, latency-bound kernels of
// different instances overlap on the device instead of running one instance
// after another.  The instances share nothing, and each one's arithmetic is
// that of a lone Sim, so its results do not depend on the batching.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mw_common.h"
#include "mw_sim.h"

namespace mw {

struct Batch {
  std::vector<Sim*> sims;
  std::vector<hipStream_t> streams;
  std::vector<std::string> errors;
  int threads = 1;
  int device = 0;
  ~Batch() {
    for (size_t i = 0; i < sims.size(); i++) {
      if (sims[i]) {
        (void)hipStreamSynchronize(sims[i]->st);
        delete sims[i];
      }
    }
    for (auto s : streams)
      if (s) (void)hipStreamDestroy(s);
  }
};

}  // namespace mw

using namespace mw;

extern "C" {

int mw_batch_create(int n, const mw_sim_config* cfgs, int threads, void** batch) {
  if (n < 1 || !cfgs || !batch || threads < 1) return fail(MW_EINVAL, "batch_create: bad argument");
  *batch = nullptr;
  Batch* b = new Batch();
  b->threads = threads;
  if (hipGetDevice(&b->device) != hipSuccess) {
    delete b;
    return fail(MW_EHIP, "batch_create: no device");
  }
  b->sims.assign(n, nullptr);
  b->streams.assign(n, nullptr);
  b->errors.assign(n, std::string());
  for (int i = 0; i < n; i++) {
    if (hipStreamCreateWithFlags(&b->streams[i], hipStreamNonBlocking) != hipSuccess) {
      delete b;
      return fail(MW_EHIP, "batch_create: stream creation failed");
    }
    Sim* s = new Sim();
    const int rc = s->create(cfgs[i], b->streams[i]);
    if (rc) {
      const std::string msg = mw_last_error();
      delete s;
      delete b;
      return fail(rc, "batch_create: instance " + std::to_string(i) + ": " + msg);
    }
    b->sims[i] = s;
  }
  *batch = b;
  return MW_OK;
}

int mw_batch_sim(void* batch, int i, void** sim) {
  Batch* b = (Batch*)batch;
  if (!b || !sim || i < 0 || i >= (int)b->sims.size())
    return fail(MW_EINVAL, "batch_sim: bad argument");
  *sim = b->sims[i];
  return MW_OK;
}

// steps timesteps of every instance; stats_host (n * MW_NSTATS, may be NULL)
// receives each instance's statistics of its last step and status_host (n,
// may be NULL) its status; MW_OK only if every instance succeeded
int mw_batch_step(void* batch, int steps, double* stats_host, int* status_host) {
  Batch* b = (Batch*)batch;
  if (!b || steps < 0) return fail(MW_EINVAL, "batch_step: bad argument");
  if (prof().on) return fail(MW_EINVAL, "batch_step: kernel profiling is per process, not per batch");
  const int n = (int)b->sims.size();
  std::vector<int> status(n, MW_OK);
  std::atomic<int> next(0);
  auto work = [&]() {
    (void)hipSetDevice(b->device);
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= n) return;
      Sim* s = b->sims[i];
      int rc = MW_OK;
      for (int k = 0; k < steps && rc == MW_OK; k++) rc = s->step();
      status[i] = rc;
      b->errors[i] = rc ? std::string(mw_last_error()) : std::string();
      if (stats_host) memcpy(stats_host + (size_t)i * MW_NSTATS, s->stats, sizeof(s->stats));
    }
  };
  const int nt = b->threads < n ? b->threads : n;
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  int first_bad = -1;
  for (int i = 0; i < n; i++) {
    if (status_host) status_host[i] = status[i];
    if (status[i] && first_bad < 0) first_bad = i;
  }
  if (first_bad >= 0)
    return fail(status[first_bad], "batch_step: instance " + std::to_string(first_bad) + ": " +
                                       b->errors[first_bad]);
  return MW_OK;
}

int mw_batch_destroy(void* batch) {
  delete (Batch*)batch;
  return MW_OK;
}

}  // extern "C"
