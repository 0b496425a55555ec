// Process-wide libitx context of the host layer, device buffers and the
// contraction-precision setting.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "itensor.h"

namespace itensor {

namespace {
std::mutex g_mu;
itx_ctx *g_ctx = nullptr;
bool g_shut = false;
int g_device = 0;
int g_prec = ITX_PREC_F64;
}  // namespace

namespace {
struct TimerTable {
  double secs[HT_COUNT] = {};
  long calls[HT_COUNT] = {};
  bool on = std::getenv("ITX_HOST_TIMERS") != nullptr;
  ~TimerTable() {
    if (!on) return;
    static const char *names[HT_COUNT] = {"qnTaskList",    "itx_plan_qn",     "itx_plan_dense", "contract QN (eager)",
                                          "contract dense (eager)", "H*psi program build", "LocalOp::product",
                                          "norm / dot (with wait)", "diag_hermitian", "denmatDecomp",
                                          "LocalMPO::position", "itx_chain_create", "itx_syev_batched"};
    std::fprintf(stderr, "[itx host timers]\n");
    for (int i = 0; i < HT_COUNT; ++i)
      if (calls[i]) std::fprintf(stderr, "  %-26s %9ld calls %10.3f ms\n", names[i], calls[i], secs[i] * 1e3);
  }
};
TimerTable g_timers;
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

HostTimer::HostTimer(HostTimerId id) : id_(id), on_(g_timers.on), t0_(on_ ? now_s() : 0.0) {}
HostTimer::~HostTimer() {
  if (!on_) return;
  g_timers.secs[id_] += now_s() - t0_;
  ++g_timers.calls[id_];
}

void check(int rc, const char *what) {
  if (rc != ITX_OK) {
    const char *m = itx_last_error(g_ctx);
    throw ITError(std::string(what) + " failed (" + std::to_string(rc) + "): " + (m ? m : ""));
  }
}

itx_ctx *ctx() {
  if (g_ctx) return g_ctx;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_shut) throw ITError("itensor: the device context has been shut down");
  if (!g_ctx) {
    itx_ctx *c = nullptr;
    int rc = itx_create(g_device, &c);
    if (rc != ITX_OK) throw ITError(std::string("itx_create failed: ") + itx_last_error(nullptr));
    g_ctx = c;
  }
  return g_ctx;
}

itx_ctx *liveCtx() { return g_ctx; }

void shutdown() {
  std::lock_guard<std::mutex> lk(g_mu);
  g_shut = true;
  if (!g_ctx) return;
  itx_ctx *c = g_ctx;
  g_ctx = nullptr;
  itx_destroy(c);
}

void setDevice(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ctx) throw ITError("setDevice: context already created");
  g_device = device;
}

void setStream(void *stream) { check(itx_set_stream(ctx(), stream), "itx_set_stream"); }

int contractPrecision() { return g_prec; }
void setContractPrecision(int prec) { g_prec = prec; }

DevBuf::DevBuf(size_t n_) : n(n_) {
  void *q = nullptr;
  check(itx_malloc(ctx(), sizeof(double) * (n_ ? n_ : 1), &q), "itx_malloc");
  p = static_cast<double *>(q);
}

DevBuf::~DevBuf() {
  if (p && g_ctx) itx_free(g_ctx, p);
}

DevPtr devAlloc(size_t n) { return std::make_shared<DevBuf>(n); }

DevPtr devZeros(size_t n) {
  auto d = devAlloc(n);
  if (n) check(itx_fill(ctx(), d->p, n, 0.0), "itx_fill");
  return d;
}

}  // namespace itensor
