// Device, memory, copy, graph and timer layer of libdgk: the part of the C
// ABI that replaces gocca's device/memory API (see include/dgk.h).
#include <cstdarg>
#include <cstring>
#include <string>

#include "dgk_internal.h"

namespace dgk {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

void clear_error() { g_err.clear(); }

TileList make_tiles(int npart, const int64_t* K, int per_tile) {
  TileList t;
  for (int p = 0; p < npart; ++p)
    for (int64_t f = 0; f < K[p]; f += per_tile) {
      t.part.push_back(p);
      t.first.push_back((int32_t)f);
      t.count.push_back((int32_t)std::min<int64_t>(per_tile, K[p] - f));
    }
  return t;
}

// Minimal JSON field lookup for the device properties string: finds "key"
// and returns the following scalar token (string contents or number text).
static bool json_field(const char* js, const char* key, std::string& out) {
  std::string s(js ? js : ""), k = std::string("\"") + key + "\"";
  size_t p = s.find(k);
  if (p == std::string::npos) return false;
  p = s.find(':', p + k.size());
  if (p == std::string::npos) return false;
  ++p;
  while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n')) ++p;
  if (p < s.size() && s[p] == '"') {
    size_t e = s.find('"', p + 1);
    if (e == std::string::npos) return false;
    out = s.substr(p + 1, e - p - 1);
    return true;
  }
  size_t e = p;
  while (e < s.size() && s[e] != ',' && s[e] != '}' && s[e] != ' ') ++e;
  out = s.substr(p, e - p);
  return !out.empty();
}

}  // namespace dgk

__global__ void dgk_fill_kernel(unsigned char* __restrict__ p, unsigned char v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

extern "C" {

const char* dgk_last_error(void) { return dgk::g_err.c_str(); }

const char* dgk_version(void) { return "libdgk 0.1 gfx950"; }

int dgk_device_create(const char* props, dgk_device* out) {
  dgk::clear_error();
  DGK_REQUIRE(out != nullptr, "dgk_device_create: out is NULL");
  *out = nullptr;
  std::string mode = "HIP", idtxt;
  dgk::json_field(props, "mode", mode);
  int id = 0;
  if (dgk::json_field(props, "device_id", idtxt)) id = atoi(idtxt.c_str());
  if (mode != "HIP" && mode != "CUDA") {
    dgk::set_error("dgk_device_create: mode \"%s\" is not available (this backend runs HIP on gfx950)",
                   mode.c_str());
    return DGK_ERR_UNSUPPORTED;
  }
  int n = 0;
  DGK_HIP(hipGetDeviceCount(&n));
  DGK_REQUIRE(id >= 0 && id < n, "dgk_device_create: device_id %d out of range (%d devices)", id, n);
  DGK_HIP(hipSetDevice(id));
  hipDeviceProp_t prop;
  DGK_HIP(hipGetDeviceProperties(&prop, id));
  dgk_device d = new dgk_device_s();
  d->id = id;
  d->cu_count = prop.multiProcessorCount;
  d->lds_bytes = (int)prop.sharedMemPerMultiprocessor;
  d->arch_major = prop.major;
  d->arch_minor = prop.minor;
  snprintf(d->name, sizeof(d->name), "%s", prop.gcnArchName);
  hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete d;
    dgk::set_error("hipStreamCreateWithFlags: %s", hipGetErrorString(e));
    return DGK_ERR_HIP;
  }
  int rc = dgk::noalloc_init(d);
  if (rc != DGK_OK) {
    (void)hipStreamDestroy(d->stream);
    delete d;
    return rc;
  }
  *out = d;
  return DGK_OK;
}

int dgk_device_destroy(dgk_device dev) {
  if (!dev) return DGK_OK;
  (void)hipSetDevice(dev->id);
  (void)hipStreamSynchronize(dev->stream);
  (void)hipStreamDestroy(dev->stream);
  if (dev->kcache) (void)hipFree(dev->kcache);
  delete dev;
  return DGK_OK;
}

int dgk_device_sync(dgk_device dev) {
  DGK_REQUIRE(dev, "dgk_device_sync: NULL device");
  DGK_HIP(hipStreamSynchronize(dev->stream));
  return DGK_OK;
}

const char* dgk_device_mode(dgk_device dev) { return dev ? "HIP" : ""; }

int dgk_device_info(dgk_device dev, char* name, int* cu_count, int* lds_bytes, int* arch_major,
                    int* arch_minor) {
  DGK_REQUIRE(dev, "dgk_device_info: NULL device");
  if (name) snprintf(name, 256, "%s", dev->name);
  if (cu_count) *cu_count = dev->cu_count;
  if (lds_bytes) *lds_bytes = dev->lds_bytes;
  if (arch_major) *arch_major = dev->arch_major;
  if (arch_minor) *arch_minor = dev->arch_minor;
  return DGK_OK;
}

int dgk_malloc(dgk_device dev, size_t bytes, const void* init_host, void** out_dev) {
  dgk::clear_error();
  DGK_REQUIRE(dev && out_dev, "dgk_malloc: NULL argument");
  *out_dev = nullptr;
  DGK_HIP(hipSetDevice(dev->id));
  void* p = nullptr;
  DGK_HIP(hipMalloc(&p, bytes ? bytes : 1));
  if (init_host && bytes) {
    hipError_t e = hipMemcpy(p, init_host, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p);
      dgk::set_error("dgk_malloc: initial copy failed: %s", hipGetErrorString(e));
      return DGK_ERR_HIP;
    }
  }
  *out_dev = p;
  return DGK_OK;
}

int dgk_free(dgk_device dev, void* ptr_dev) {
  DGK_REQUIRE(dev, "dgk_free: NULL device");
  if (!ptr_dev) return DGK_OK;
  DGK_HIP(hipStreamSynchronize(dev->stream));
  DGK_HIP(hipFree(ptr_dev));
  return DGK_OK;
}

int dgk_copy_h2d(dgk_device dev, void* dst_dev, const void* src_host, size_t bytes,
                 size_t dst_offset_bytes) {
  DGK_REQUIRE(dev && dst_dev && (src_host || !bytes), "dgk_copy_h2d: NULL argument");
  if (!bytes) return DGK_OK;
  DGK_HIP(hipMemcpyAsync((char*)dst_dev + dst_offset_bytes, src_host, bytes,
                         hipMemcpyHostToDevice, dev->stream));
  DGK_HIP(hipStreamSynchronize(dev->stream));
  return DGK_OK;
}

int dgk_copy_d2h(dgk_device dev, void* dst_host, const void* src_dev, size_t bytes,
                 size_t src_offset_bytes) {
  DGK_REQUIRE(dev && src_dev && (dst_host || !bytes), "dgk_copy_d2h: NULL argument");
  if (!bytes) return DGK_OK;
  DGK_HIP(hipMemcpyAsync(dst_host, (const char*)src_dev + src_offset_bytes, bytes,
                         hipMemcpyDeviceToHost, dev->stream));
  DGK_HIP(hipStreamSynchronize(dev->stream));
  return DGK_OK;
}

int dgk_copy_d2d(dgk_device dev, void* dst_dev, const void* src_dev, size_t bytes) {
  DGK_REQUIRE(dev && dst_dev && src_dev, "dgk_copy_d2d: NULL argument");
  if (!bytes) return DGK_OK;
  DGK_HIP(hipMemcpyAsync(dst_dev, src_dev, bytes, hipMemcpyDeviceToDevice, dev->stream));
  return DGK_OK;
}

int dgk_fill_bytes(dgk_device dev, void* dst_dev, int value, size_t bytes) {
  DGK_REQUIRE(dev && dst_dev, "dgk_fill_bytes: NULL argument");
  if (!bytes) return DGK_OK;
  size_t blocks = (bytes + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(dgk_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, dev->stream,
                     (unsigned char*)dst_dev, (unsigned char)value, bytes);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

struct dgk_graph_s {
  dgk_device dev;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

int dgk_graph_begin(dgk_device dev) {
  DGK_REQUIRE(dev, "dgk_graph_begin: NULL device");
  DGK_REQUIRE(!dev->capturing, "dgk_graph_begin: capture already active");
  DGK_HIP(hipStreamBeginCapture(dev->stream, hipStreamCaptureModeThreadLocal));
  dev->capturing = true;
  return DGK_OK;
}

int dgk_graph_end(dgk_device dev, dgk_graph* out) {
  DGK_REQUIRE(dev && out, "dgk_graph_end: NULL argument");
  DGK_REQUIRE(dev->capturing, "dgk_graph_end: no capture active");
  dev->capturing = false;
  hipGraph_t g;
  DGK_HIP(hipStreamEndCapture(dev->stream, &g));
  hipGraphExec_t ex;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    dgk::set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return DGK_ERR_HIP;
  }
  *out = new dgk_graph_s{dev, g, ex};
  return DGK_OK;
}

int dgk_graph_launch(dgk_graph g) {
  DGK_REQUIRE(g, "dgk_graph_launch: NULL graph");
  DGK_HIP(hipGraphLaunch(g->exec, g->dev->stream));
  return DGK_OK;
}

int dgk_graph_destroy(dgk_graph g) {
  if (!g) return DGK_OK;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return DGK_OK;
}

struct dgk_timer_s {
  dgk_device dev;
  hipEvent_t a, b;
};

int dgk_timer_create(dgk_device dev, dgk_timer* out) {
  DGK_REQUIRE(dev && out, "dgk_timer_create: NULL argument");
  dgk_timer t = new dgk_timer_s{dev, nullptr, nullptr};
  if (hipEventCreate(&t->a) != hipSuccess || hipEventCreate(&t->b) != hipSuccess) {
    delete t;
    dgk::set_error("hipEventCreate failed");
    return DGK_ERR_HIP;
  }
  *out = t;
  return DGK_OK;
}

int dgk_timer_start(dgk_timer t) {
  DGK_REQUIRE(t, "dgk_timer_start: NULL timer");
  DGK_HIP(hipEventRecord(t->a, t->dev->stream));
  return DGK_OK;
}

int dgk_timer_stop(dgk_timer t) {
  DGK_REQUIRE(t, "dgk_timer_stop: NULL timer");
  DGK_HIP(hipEventRecord(t->b, t->dev->stream));
  return DGK_OK;
}

int dgk_timer_elapsed_ms(dgk_timer t, double* ms) {
  DGK_REQUIRE(t && ms, "dgk_timer_elapsed_ms: NULL argument");
  DGK_HIP(hipEventSynchronize(t->b));
  float f = 0.f;
  DGK_HIP(hipEventElapsedTime(&f, t->a, t->b));
  *ms = f;
  return DGK_OK;
}

int dgk_timer_destroy(dgk_timer t) {
  if (!t) return DGK_OK;
  (void)hipEventDestroy(t->a);
  (void)hipEventDestroy(t->b);
  delete t;
  return DGK_OK;
}

}  // extern "C"
