// runtime.hip -- device selection, stream, error state, level mirrors, timers.
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>

#include "bamhip_internal.h"

namespace bamhip {

static thread_local char g_err[1024] = "";
static hipStream_t g_stream = nullptr;
static int g_device = -1;

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_hip(hipError_t e, const char *what) {
  return set_error(BAMHIP_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

hipStream_t stream() { return g_stream; }

// A second stream (default priority, like the library stream) for work that
// may overlap the library stream's: fork makes it wait for everything issued
// so far on the library stream, join makes the library stream wait for
// everything issued on it since.  Both record events only, so a graph capture
// of the library stream takes the side stream's work in as a parallel branch.
static hipStream_t g_side = nullptr;
static hipEvent_t g_fork = nullptr, g_join = nullptr;

hipStream_t side_stream() { return g_side; }

int fork_side() {
  BH_HIP(hipEventRecord(g_fork, g_stream));
  BH_HIP(hipStreamWaitEvent(g_side, g_fork, 0));
  return BAMHIP_OK;
}

int join_side() {
  BH_HIP(hipEventRecord(g_join, g_side));
  BH_HIP(hipStreamWaitEvent(g_stream, g_join, 0));
  return BAMHIP_OK;
}

int gather_ptrs(bamhip_level *lev, const int *idx, int n, double **out, const char *what) {
  if (!idx) return set_error(BAMHIP_EINVAL, "%s: variable list is NULL", what);
  for (int i = 0; i < n; i++) {
    const int k = idx[i];
    if (k < 0 || k >= lev->nvariables || !lev->v[k])
      return set_error(BAMHIP_EINVAL, "%s: variable index %d (list entry %d) is not enabled", what, k, i);
    out[i] = lev->v[k];
  }
  return BAMHIP_OK;
}

// ------------------------------------------------------------------ timers
struct TimerRec {
  double ms = 0;
  long launches = 0;
};
static bool g_timer_on = false;
static std::map<std::string, TimerRec> g_timers;
static std::mutex g_timer_mu;

// Events are recorded around each launch on the library stream and kept in
// a pending list; they are resolved (synchronised) only when the timers are
// read, so timing does not serialise the stream.  Events are pooled.
struct Pending {
  std::string name;
  hipEvent_t a, b;
};
static std::vector<Pending> g_pending;
static std::vector<hipEvent_t> g_pool;

static hipEvent_t pool_get() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static void resolve_pending() {
  for (Pending &p : g_pending) {
    float ms = 0;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      TimerRec &r = g_timers[p.name];
      r.ms += ms;
      r.launches += 1;
    }
    g_pool.push_back(p.a);
    g_pool.push_back(p.b);
  }
  g_pending.clear();
}

static bool g_capturing = false;   // the library stream is being captured into a graph

TimerScope::TimerScope(const char *name) : name_(name), a_(nullptr), b_(nullptr), on_(g_timer_on && !g_capturing) {
  if (!on_) return;
  std::lock_guard<std::mutex> lk(g_timer_mu);
  a_ = pool_get();
  b_ = pool_get();
  if (!a_ || !b_) {
    on_ = false;
    return;
  }
  (void)hipEventRecord(a_, g_stream);
}

TimerScope::~TimerScope() {
  if (!on_) return;
  (void)hipEventRecord(b_, g_stream);
  std::lock_guard<std::mutex> lk(g_timer_mu);
  g_pending.push_back(Pending{name_, a_, b_});
  if (g_pending.size() > 4096) resolve_pending();
}

// Device arrays of grid functions are recycled: a variable freed with its
// level (or disabled) goes to a pool keyed by its size, and enabling a
// variable takes an array of that size from the pool before calling
// hipMalloc.  A regrid (move_box) frees a level and creates one of the same
// box sizes, so its ~100 variables per level cost no allocation.  The pool
// keeps at most kPoolMax bytes; arrays beyond that are freed.  Every array
// entering the pool is idle: level_free and level_disable synchronise the
// library stream first.
static std::multimap<size_t, double *> g_arrays;
static size_t g_array_bytes = 0;
constexpr size_t kPoolMax = size_t(32) << 30;

static double *array_get(size_t bytes) {
  auto it = g_arrays.find(bytes);
  if (it == g_arrays.end()) return nullptr;
  double *p = it->second;
  g_arrays.erase(it);
  g_array_bytes -= bytes;
  return p;
}

static void array_put(double *p, size_t bytes) {
  if (!p) return;
  if (g_array_bytes + bytes > kPoolMax) {
    (void)hipFree(p);
    return;
  }
  g_arrays.emplace(bytes, p);
  g_array_bytes += bytes;
}

static void array_pool_release() {
  for (auto &kv : g_arrays) (void)hipFree(kv.second);
  g_arrays.clear();
  g_array_bytes = 0;
}

}  // namespace bamhip

using namespace bamhip;

extern "C" {

int bamhip_init(int device) {
  int n = 0;
  BH_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_error(BAMHIP_EINVAL, "bamhip_init: device %d of %d", device, n);
  if (g_device >= 0 && g_device != device) array_pool_release();  // recycled arrays live on the old device
  BH_HIP(hipSetDevice(device));
  int e;
  if ((e = noalloc_scratch_init(device))) return e;
  if (!g_stream || g_device != device) {
    if (g_stream) (void)hipStreamDestroy(g_stream);
    if (g_side) (void)hipStreamDestroy(g_side);
    if (g_fork) (void)hipEventDestroy(g_fork);
    if (g_join) (void)hipEventDestroy(g_join);
    BH_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    BH_HIP(hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking));
    BH_HIP(hipEventCreateWithFlags(&g_fork, hipEventDisableTiming));
    BH_HIP(hipEventCreateWithFlags(&g_join, hipEventDisableTiming));
    g_device = device;
  }
  return BAMHIP_OK;
}

int bamhip_finalize(void) {
  if (g_side) {
    BH_HIP(hipStreamSynchronize(g_side));
    BH_HIP(hipStreamDestroy(g_side));
    g_side = nullptr;
  }
  if (g_fork) (void)hipEventDestroy(g_fork);
  if (g_join) (void)hipEventDestroy(g_join);
  g_fork = g_join = nullptr;
  if (g_stream) {
    BH_HIP(hipStreamSynchronize(g_stream));
    BH_HIP(hipStreamDestroy(g_stream));
    g_stream = nullptr;
  }
  array_pool_release();
  g_device = -1;
  return BAMHIP_OK;
}

const char *bamhip_last_error(void) { return g_err; }

void *bamhip_stream(void) { return (void *)g_stream; }

int bamhip_synchronize(void) {
  BH_HIP(hipStreamSynchronize(g_stream));
  return BAMHIP_OK;
}

int bamhip_malloc(void **ptr, size_t bytes) {
  if (!ptr) return set_error(BAMHIP_EINVAL, "bamhip_malloc: NULL");
  BH_HIP(hipMalloc(ptr, bytes ? bytes : 8));
  return BAMHIP_OK;
}

int bamhip_free(void *ptr) {
  if (ptr) BH_HIP(hipFree(ptr));
  return BAMHIP_OK;
}

int bamhip_graph_begin(void) {
  if (!g_stream) return set_error(BAMHIP_EINVAL, "bamhip_graph_begin: call bamhip_init first");
  if (g_capturing) return set_error(BAMHIP_EINVAL, "bamhip_graph_begin: already capturing");
  BH_HIP(hipStreamBeginCapture(g_stream, hipStreamCaptureModeRelaxed));
  g_capturing = true;
  return BAMHIP_OK;
}

int bamhip_graph_end(void **graph) {
  if (!g_capturing || !graph) return set_error(BAMHIP_EINVAL, "bamhip_graph_end: not capturing");
  g_capturing = false;
  hipGraph_t g = nullptr;
  BH_HIP(hipStreamEndCapture(g_stream, &g));
  hipGraphExec_t x = nullptr;
  const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  BH_HIP(e);
  *graph = (void *)x;
  return BAMHIP_OK;
}

int bamhip_graph_launch(void *graph) {
  if (!graph) return set_error(BAMHIP_EINVAL, "bamhip_graph_launch: NULL graph");
  BH_HIP(hipGraphLaunch((hipGraphExec_t)graph, g_stream));
  return BAMHIP_OK;
}

int bamhip_graph_free(void *graph) {
  if (graph) BH_HIP(hipGraphExecDestroy((hipGraphExec_t)graph));
  return BAMHIP_OK;
}

int bamhip_memcpy_htod(void *dst, const void *src, size_t bytes) {
  BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
  BH_HIP(hipStreamSynchronize(g_stream));
  return BAMHIP_OK;
}

int bamhip_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
  BH_HIP(hipStreamSynchronize(g_stream));
  return BAMHIP_OK;
}

int bamhip_level_create(bamhip_level **lev, int l, int nboxes, const bamhip_box *boxes, int nvariables) {
  if (!lev || nboxes <= 0 || !boxes || nvariables <= 0)
    return set_error(BAMHIP_EINVAL, "bamhip_level_create: bad arguments");
  if (!g_stream) return set_error(BAMHIP_EINVAL, "bamhip_level_create: call bamhip_init first");
  bamhip_level *L = new bamhip_level();
  L->l = l;
  L->nboxes = nboxes;
  L->nvariables = nvariables;
  L->v.assign(nvariables, nullptr);
  long off = 0;
  for (int b = 0; b < nboxes; b++) {
    const bamhip_box &B = boxes[b];
    if (B.m < 1 || B.n < 1 || B.o < 1 || !(B.dx > 0) || !(B.dy > 0) || !(B.dz > 0)) {
      delete L;
      return set_error(BAMHIP_EINVAL, "bamhip_level_create: box %d has bad size or spacing", b);
    }
    for (int f = 0; f < 6; f++)
      if (B.skip[f] < 1 || B.skip[f] > 16) {
        delete L;
        return set_error(BAMHIP_EINVAL,
                         "bamhip_level_create: box %d face %d skip=%d (the outermost layer is always "
                         "flagged in BAM, and at most 16 layers are supported)",
                         b, f, B.skip[f]);
      }
    DevBox D;
    D.m = B.m;
    D.n = B.n;
    D.o = B.o;
    D.noffset = off;
    D.x0 = B.x0;
    D.y0 = B.y0;
    D.z0 = B.z0;
    D.dx = B.dx;
    D.dy = B.dy;
    D.dz = B.dz;
    for (int f = 0; f < 6; f++) {
      D.skip[f] = B.skip[f];
      D.bflag[f] = B.bflag[f];
    }
    L->boxes.push_back(D);
    off += (long)B.m * B.n * B.o;
  }
  L->npoints = off;
  if (hipMalloc(&L->d_flag, sizeof(int)) != hipSuccess) {
    delete L;
    return set_error(BAMHIP_ENOMEM, "bamhip_level_create: flag allocation failed");
  }
  if (hipMemsetAsync(L->d_flag, 0, sizeof(int), g_stream) != hipSuccess ||
      hipStreamSynchronize(g_stream) != hipSuccess) {
    (void)hipFree(L->d_flag);
    delete L;
    return set_error(BAMHIP_EDEVICE, "bamhip_level_create: flag initialisation failed");
  }
  *lev = L;
  return BAMHIP_OK;
}

int bamhip_level_free(bamhip_level *lev) {
  if (!lev) return BAMHIP_OK;
  if (g_stream) (void)hipStreamSynchronize(g_stream);
  for (double *p : lev->v) array_put(p, sizeof(double) * lev->npoints);
  if (lev->d_flag) (void)hipFree(lev->d_flag);
  delete lev;
  return BAMHIP_OK;
}

int bamhip_level_npoints(const bamhip_level *lev, long *np) {
  if (!lev || !np) return set_error(BAMHIP_EINVAL, "bamhip_level_npoints: NULL");
  *np = lev->npoints;
  return BAMHIP_OK;
}

int bamhip_level_enable(bamhip_level *lev, int var) {
  if (!lev || var < 0 || var >= lev->nvariables)
    return set_error(BAMHIP_EINVAL, "bamhip_level_enable: bad variable index %d", var);
  if (lev->v[var]) return BAMHIP_OK;
  double *p = array_get(sizeof(double) * lev->npoints);
  if (!p && hipMalloc(&p, sizeof(double) * lev->npoints) != hipSuccess) {
    // return the pooled arrays to the device and try once more
    array_pool_release();
    if (hipMalloc(&p, sizeof(double) * lev->npoints) != hipSuccess)
      return set_error(BAMHIP_ENOMEM, "bamhip_level_enable: out of device memory (%ld points)", lev->npoints);
  }
  lev->v[var] = p;
  return BAMHIP_OK;
}

int bamhip_level_disable(bamhip_level *lev, int var) {
  if (!lev || var < 0 || var >= lev->nvariables)
    return set_error(BAMHIP_EINVAL, "bamhip_level_disable: bad variable index %d", var);
  if (lev->v[var]) {
    BH_HIP(hipStreamSynchronize(g_stream));
    array_put(lev->v[var], sizeof(double) * lev->npoints);
    lev->v[var] = nullptr;
  }
  return BAMHIP_OK;
}

double *bamhip_level_ptr(bamhip_level *lev, int var) {
  if (!lev || var < 0 || var >= lev->nvariables) return nullptr;
  return lev->v[var];
}

int bamhip_level_put(bamhip_level *lev, int var, const double *host) {
  double *p = bamhip_level_ptr(lev, var);
  if (!p || !host) return set_error(BAMHIP_EINVAL, "bamhip_level_put: variable %d not enabled", var);
  BH_HIP(hipMemcpyAsync(p, host, sizeof(double) * lev->npoints, hipMemcpyHostToDevice, g_stream));
  BH_HIP(hipStreamSynchronize(g_stream));
  return BAMHIP_OK;
}

int bamhip_level_get(bamhip_level *lev, int var, double *host) {
  double *p = bamhip_level_ptr(lev, var);
  if (!p || !host) return set_error(BAMHIP_EINVAL, "bamhip_level_get: variable %d not enabled", var);
  BH_HIP(hipMemcpyAsync(host, p, sizeof(double) * lev->npoints, hipMemcpyDeviceToHost, g_stream));
  BH_HIP(hipStreamSynchronize(g_stream));
  return BAMHIP_OK;
}

__global__ void set_flag_kernel(int *p, int v) { *p = v; }

int bamhip_level_set_check(bamhip_level *lev, int mode) {
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_level_set_check: NULL level");
  lev->check_mode = mode ? 1 : 0;
  return BAMHIP_OK;
}

int bamhip_level_get_check(const bamhip_level *lev, int *mode) {
  if (!lev || !mode) return set_error(BAMHIP_EINVAL, "bamhip_level_get_check: NULL argument");
  *mode = lev->check_mode;
  return BAMHIP_OK;
}

int bamhip_level_nonfinite(bamhip_level *lev, int reset, int *flag) {
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_level_nonfinite: NULL level");
  if (flag) {
    int h = 0;
    BH_HIP(hipMemcpyAsync(&h, lev->d_flag, sizeof(int), hipMemcpyDeviceToHost, g_stream));
    BH_HIP(hipStreamSynchronize(g_stream));
    *flag = h;
  }
  if (reset) {
    set_flag_kernel<<<1, 1, 0, g_stream>>>(lev->d_flag, 0);
    BH_HIP(hipGetLastError());
  }
  return BAMHIP_OK;
}

int bamhip_timer_enable(int on) {
  g_timer_on = on != 0;
  return BAMHIP_OK;
}

int bamhip_timer_reset(void) {
  std::lock_guard<std::mutex> lk(g_timer_mu);
  resolve_pending();
  g_timers.clear();
  return BAMHIP_OK;
}

int bamhip_timer_get(const char *name, double *ms_total, long *launches) {
  if (!name) return set_error(BAMHIP_EINVAL, "bamhip_timer_get: NULL name");
  std::lock_guard<std::mutex> lk(g_timer_mu);
  resolve_pending();
  auto it = g_timers.find(name);
  if (ms_total) *ms_total = (it == g_timers.end()) ? 0.0 : it->second.ms;
  if (launches) *launches = (it == g_timers.end()) ? 0 : it->second.launches;
  return BAMHIP_OK;
}

}  // extern "C"
