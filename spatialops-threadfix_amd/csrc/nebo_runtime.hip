// nebo_runtime.hip -- errors, device selection, the device memory pool and copies.
//
// The pool replaces Pool<T>::get/put (spatialops/structured/MemoryPool.cpp:102-186):
// freed blocks are kept in per-size free lists and handed out again for requests
// of the same size, so a time loop that creates temporaries allocates only once.

#include "nebo_internal.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <unordered_map>
#include <vector>

namespace nebo {

namespace {
thread_local std::string t_error;
}

int set_error(int code, const char* ctx, const std::string& msg)
{
  std::ostringstream s;
  s << "Nebo error in " << ctx << ":\n" << msg;
  t_error = s.str();
  return code;
}

void clear_error() { t_error.clear(); }

// ---- backend reporting (the reference's NEBO_REPORT_BACKEND, NeboLhs.h:167-170,
// 321-331, 407, 475: "Starting Nebo <backend>" / "Finished Nebo <backend>").  The
// reference compiles the messages in with a macro; this library is prebuilt, so
// they are switched on at run time, by the NEBO_REPORT_BACKEND environment
// variable (any value but "0") or nebo_set_report_backend.  The launches are
// asynchronous: "Finished" means the kernel is queued on its stream.
namespace {
thread_local std::string t_backend;
std::atomic<int> g_report{-1};   // -1: environment not read yet
}

bool report_backend()
{
  int r = g_report.load();
  if (r < 0) {
    const char* v = std::getenv("NEBO_REPORT_BACKEND");
    r = (v && *v && std::string(v) != "0") ? 1 : 0;
    int expected = -1;
    g_report.compare_exchange_strong(expected, r);
    r = g_report.load();
  }
  return r == 1;
}

namespace {
thread_local unsigned t_scope_serial = 0;   // BackendScopes this thread has opened
}

BackendScope::BackendScope(const std::string& name)
{
  ++t_scope_serial;
  t_backend = name;
  if (report_backend()) {
    std::printf("Starting Nebo HIP %s\n", name.c_str());
    std::fflush(stdout);
  }
}

BackendScope::~BackendScope()
{
  if (report_backend()) {
    std::printf("Finished Nebo HIP %s\n", t_backend.c_str());
    std::fflush(stdout);
  }
}

// ---- call memo (see nebo_internal.h) ----
struct MemoRecord {
  struct Step {
    bool scope;            // opened a backend scope when recorded (replayed inside one)
    std::string backend;   // the backend name in effect at the launch
    Launch launch;
    unsigned serial;       // scope serial at the launch, while recording
  };
  uint64_t epoch = 0;
  unsigned serial0 = 0;    // scope serial when recording started
  bool ok = true;
  std::vector<Step> steps;
};

namespace {
std::atomic<uint64_t> g_plan_epoch{1};
thread_local MemoRecord* t_record = nullptr;
std::mutex g_memo_mu;
std::unordered_map<std::string, std::shared_ptr<const MemoRecord>> g_memo;
constexpr size_t kMemoMaxCalls = 1024;   // a full memo is cleared and refilled

template <class T>
void put_bytes(std::string& k, const T& v)
{
  k.append(reinterpret_cast<const char*>(&v), sizeof v);
}
}  // namespace

void bump_plan_epoch() { g_plan_epoch.fetch_add(1); }

int emit(void* stream, const CallPtrs& p, Launch launch)
{
  const int rc = launch(stream, p);
  MemoRecord* r = t_record;
  if (r) {
    if (rc) {
      r->ok = false;
    } else {
      const unsigned prev = r->steps.empty() ? r->serial0 : r->steps.back().serial;
      r->steps.push_back(MemoRecord::Step{t_scope_serial != prev, t_backend, std::move(launch), t_scope_serial});
    }
  }
  return rc;
}

void CallKey::begin(char kind, int device, int flag_a, int flag_b)
{
  bytes.clear();
  bytes.push_back(kind);
  put_bytes(bytes, device);
  put_bytes(bytes, flag_a);
  put_bytes(bytes, flag_b);
}

void CallKey::field(const nebo_field_t& f) { put_bytes(bytes, f); }

bool CallKey::expr(const nebo_expr_t* e)
{
  if (!e || e->n_nodes < 0 || e->n_fields < 0 || e->n_stencils < 0 || e->n_masks < 0) return false;
  if ((e->n_nodes && !e->nodes) || (e->n_fields && !e->fields) || (e->n_stencils && !e->stencils) ||
      (e->n_masks && !e->masks))
    return false;
  put_bytes(bytes, e->n_nodes);
  put_bytes(bytes, e->n_fields);
  put_bytes(bytes, e->n_stencils);
  put_bytes(bytes, e->n_masks);
  bytes.append(reinterpret_cast<const char*>(e->nodes), sizeof(nebo_node_t) * (size_t)e->n_nodes);
  bytes.append(reinterpret_cast<const char*>(e->fields), sizeof(nebo_field_t) * (size_t)e->n_fields);
  for (int i = 0; i < e->n_stencils; ++i) {
    const nebo_stencil_t& st = e->stencils[i];
    const int n = st.npts < 0 ? 0 : st.npts > NEBO_MAX_STENCIL_POINTS ? NEBO_MAX_STENCIL_POINTS : st.npts;
    put_bytes(bytes, st.kind);
    put_bytes(bytes, st.npts);
    bytes.append(reinterpret_cast<const char*>(st.pts), sizeof st.pts[0] * (size_t)n);
    bytes.append(reinterpret_cast<const char*>(st.coefs), sizeof st.coefs[0] * (size_t)n);
  }
  bytes.append(reinterpret_cast<const char*>(e->masks), sizeof(nebo_mask_t) * (size_t)e->n_masks);
  return true;
}

std::shared_ptr<const MemoRecord> memo_find(const CallKey& key)
{
  std::lock_guard<std::mutex> lk(g_memo_mu);
  auto it = g_memo.find(key.bytes);
  if (it == g_memo.end()) return nullptr;
  if (it->second->epoch != g_plan_epoch.load()) {
    g_memo.erase(it);
    return nullptr;
  }
  return it->second;
}

int memo_run(const MemoRecord& r, void* stream, const CallPtrs& p)
{
  for (const MemoRecord::Step& st : r.steps) {
    int rc;
    if (st.scope) {
      BackendScope bs(st.backend);
      rc = st.launch(stream, p);
    } else {
      rc = st.launch(stream, p);
    }
    if (rc) return rc;
  }
  return NEBO_OK;
}

MemoRecorder::MemoRecorder(const CallKey& key) : key_(key.bytes), rec_(new MemoRecord), prev_(t_record)
{
  rec_->epoch = g_plan_epoch.load();
  rec_->serial0 = t_scope_serial;
  t_record = rec_;
}

void MemoRecorder::commit(int rc)
{
  t_record = prev_;
  if (rc != NEBO_OK || !rec_->ok) return;
  std::shared_ptr<const MemoRecord> r(rec_);
  rec_ = nullptr;
  std::lock_guard<std::mutex> lk(g_memo_mu);
  if (g_memo.size() >= kMemoMaxCalls) g_memo.clear();
  g_memo[key_] = std::move(r);
}

MemoRecorder::~MemoRecorder()
{
  t_record = prev_;
  delete rec_;
}

int check_hip(int st, const char* ctx)
{
  if (st == hipSuccess) return NEBO_OK;
  return set_error(NEBO_ERR_HIP, ctx, std::string("\t - ") + hipGetErrorString((hipError_t)st));
}

namespace {

struct DevicePool {
  std::unordered_map<size_t, std::vector<void*>> free_lists;
  std::unordered_map<void*, size_t> live;
  size_t in_use = 0;
  size_t high_water = 0;
};

std::mutex g_pool_mutex;
std::map<int, DevicePool> g_pools;

// Run f with `device` current, restoring the caller's device afterwards.
template <class F>
int with_device(int device, const char* ctx, F f)
{
  int cur = 0;
  int rc = check_hip(hipGetDevice(&cur), ctx);
  if (rc) return rc;
  if (cur != device && (rc = check_hip(hipSetDevice(device), ctx))) return rc;
  rc = f();
  if (cur != device) (void)hipSetDevice(cur);
  return rc;
}

}  // namespace
}  // namespace nebo

using namespace nebo;

extern "C" const char* nebo_last_error(void) { return t_error.c_str(); }

extern "C" int nebo_set_report_backend(int on)
{
  clear_error();
  g_report.store(on ? 1 : 0);
  return NEBO_OK;
}

extern "C" const char* nebo_last_backend(void) { return t_backend.c_str(); }

extern "C" int nebo_device_count(int* count)
{
  clear_error();
  if (!count) return set_error(NEBO_ERR_INVALID, "Nebo Device", "null argument");
  *count = 0;
  int n = 0;
  hipError_t st = hipGetDeviceCount(&n);
  if (st == hipErrorNoDevice) { (void)hipGetLastError(); return NEBO_OK; }
  int rc = check_hip(st, "Nebo Device");
  if (rc) return rc;
  *count = n;
  return NEBO_OK;
}

extern "C" int nebo_init(int device)
{
  clear_error();
  int n = 0;
  int rc = nebo_device_count(&n);
  if (rc) return rc;
  if (device < 0 || device >= n) {
    std::ostringstream m;
    m << "device index " << device << " is not a GPU of this host (" << n << " visible)";
    return set_error(NEBO_ERR_DEVICE, "Nebo Device", m.str());
  }
  if ((rc = check_hip(hipSetDevice(device), "Nebo Device"))) return rc;
  return check_hip(hipFree(nullptr), "Nebo Device");
}

extern "C" int nebo_alloc(int device, size_t bytes, void** out)
{
  clear_error();
  if (!out) return set_error(NEBO_ERR_INVALID, "Nebo Memory", "null argument");
  *out = nullptr;
  if (bytes == 0) bytes = 8;
  std::lock_guard<std::mutex> lock(g_pool_mutex);
  DevicePool& pool = g_pools[device];
  auto it = pool.free_lists.find(bytes);
  void* p = nullptr;
  if (it != pool.free_lists.end() && !it->second.empty()) {
    p = it->second.back();
    it->second.pop_back();
  } else {
    int rc = with_device(device, "Nebo Memory", [&] {
      return check_hip(hipMalloc(&p, bytes), "Nebo Memory");
    });
    if (rc) return rc;
  }
  pool.live[p] = bytes;
  pool.in_use += bytes;
  if (pool.in_use > pool.high_water) pool.high_water = pool.in_use;
  *out = p;
  return NEBO_OK;
}

extern "C" int nebo_free(int device, void* ptr)
{
  clear_error();
  if (!ptr) return NEBO_OK;
  std::lock_guard<std::mutex> lock(g_pool_mutex);
  DevicePool& pool = g_pools[device];
  auto it = pool.live.find(ptr);
  if (it == pool.live.end())
    return set_error(NEBO_ERR_INVALID, "Nebo Memory", "pointer was not allocated by nebo_alloc on this device");
  pool.in_use -= it->second;
  pool.free_lists[it->second].push_back(ptr);
  pool.live.erase(it);
  return NEBO_OK;
}

extern "C" int nebo_pool_trim(int device)
{
  clear_error();
  std::lock_guard<std::mutex> lock(g_pool_mutex);
  DevicePool& pool = g_pools[device];
  return with_device(device, "Nebo Memory", [&] {
    int rc = check_hip(hipDeviceSynchronize(), "Nebo Memory");
    if (rc) return rc;
    for (auto& kv : pool.free_lists)
      for (void* p : kv.second)
        if ((rc = check_hip(hipFree(p), "Nebo Memory"))) return rc;
    pool.free_lists.clear();
    return NEBO_OK;
  });
}

extern "C" int nebo_pool_stats(int device, size_t* in_use, size_t* high_water)
{
  clear_error();
  std::lock_guard<std::mutex> lock(g_pool_mutex);
  DevicePool& pool = g_pools[device];
  if (in_use) *in_use = pool.in_use;
  if (high_water) *high_water = pool.high_water;
  return NEBO_OK;
}

extern "C" int nebo_host_alloc(size_t bytes, void** out)
{
  clear_error();
  if (!out) return set_error(NEBO_ERR_INVALID, "Nebo Memory", "null argument");
  *out = nullptr;
  if (bytes == 0) return NEBO_OK;
  return check_hip(hipHostMalloc(out, bytes, hipHostMallocDefault), "Nebo Memory");
}

extern "C" int nebo_host_free(void* ptr)
{
  clear_error();
  if (!ptr) return NEBO_OK;
  return check_hip(hipHostFree(ptr), "Nebo Memory");
}

extern "C" int nebo_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream)
{
  clear_error();
  if (bytes == 0) return NEBO_OK;
  if (!dst || !src) return set_error(NEBO_ERR_INVALID, "Nebo Memory", "null pointer");
  hipMemcpyKind k;
  switch (kind) {
  case 0: k = hipMemcpyHostToDevice; break;
  case 1: k = hipMemcpyDeviceToHost; break;
  case 2: k = hipMemcpyDeviceToDevice; break;
  default: return set_error(NEBO_ERR_INVALID, "Nebo Memory", "unknown copy kind");
  }
  if (!stream) return check_hip(hipMemcpy(dst, src, bytes, k), "Nebo Memory");
  return check_hip(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream), "Nebo Memory");
}

extern "C" int nebo_memcpy_peer(void* dst, int dst_dev, const void* src, int src_dev,
                                size_t bytes, void* stream)
{
  clear_error();
  if (bytes == 0) return NEBO_OK;
  if (!stream)
    return check_hip(hipMemcpyPeer(dst, dst_dev, src, src_dev, bytes), "Nebo Memory");
  return check_hip(hipMemcpyPeerAsync(dst, dst_dev, src, src_dev, bytes, (hipStream_t)stream),
                   "Nebo Memory");
}

extern "C" int nebo_memset(void* dst, int value, size_t bytes, void* stream)
{
  clear_error();
  if (bytes == 0) return NEBO_OK;
  if (!stream) return check_hip(hipMemset(dst, value, bytes), "Nebo Memory");
  return check_hip(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream), "Nebo Memory");
}

extern "C" int nebo_stream_sync(void* stream)
{
  clear_error();
  return check_hip(hipStreamSynchronize((hipStream_t)stream), "Nebo Stream");
}

extern "C" int nebo_device_sync(void)
{
  clear_error();
  return check_hip(hipDeviceSynchronize(), "Nebo Stream");
}
