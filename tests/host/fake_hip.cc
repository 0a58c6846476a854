// A stand-in for the HIP runtime, for host-side checks of the library's own
// code on a machine without a GPU (tests/test_host_api.py builds the library
// host-only with AddressSanitizer and links it against this file). Device
// memory is host memory, so every copy and fill the library issues is a real
// memcpy / memset that AddressSanitizer checks against the allocation; kernel
// launches and graph replays do nothing. Stream capture is modelled closely
// enough to reject what the GPU runtime rejects or mis-runs:
//   * a synchronisation, blocking copy or fill on a capturing stream;
//   * hipMemsetAsync, hipMalloc or hipFree while a capture is open (fills are
//     not recorded into graphs by this library, the eager pass before a
//     capture allocates every buffer, and a free would leave a recorded node
//     pointing at released memory);
//   * a capture ended while a forked stream still has work that was never
//     joined back to the origin stream;
//   * frees of unknown pointers, launches of destroyed graph executables.
// Any violation prints "fake_hip: <what>" and aborts.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>

namespace {

struct Stream
{
  bool capturing = false;
  bool origin = false;
  long long ops = 0;    // work enqueued on the stream
  long long joined = 0; // ops another capturing stream has waited for
};
struct Event
{
  hipStream_t stream = nullptr;
  bool from_capture = false;
  long long ops = 0; // the recording stream's ops at the record
};

std::recursive_mutex mu;
std::map<void*, size_t> allocs;
std::map<hipStream_t, Stream> streams;
std::map<hipEvent_t, Event> events;
std::set<void*> graphs, execs;
int open_captures = 0;
thread_local dim3 cfg_grid, cfg_block;
thread_local size_t cfg_shmem = 0;
thread_local hipStream_t cfg_stream = nullptr;
long long launches = 0;

[[noreturn]] void fail(const char* what)
{
  std::fprintf(stderr, "fake_hip: %s\n", what);
  std::fflush(stderr);
  std::abort();
}

Stream* stream_of(hipStream_t s)
{
  auto it = streams.find(s);
  return it == streams.end() ? nullptr : &it->second;
}

void enqueue_on(hipStream_t s)
{
  Stream* st = stream_of(s);
  if (st)
    st->ops++;
}

void no_capture(hipStream_t s, const char* what)
{
  Stream* st = stream_of(s);
  if (st && st->capturing)
    fail(what);
}

} // namespace

extern "C" {

void** __hipRegisterFatBinary(const void*)
{
  static void* handle = nullptr;
  return &handle;
}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, uint3*, uint3*,
                           dim3*, dim3*, int*)
{}
void __hipRegisterVar(void**, void*, char*, char*, int, size_t, int, int) {}
void __hipUnregisterFatBinary(void**) {}

hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shmem, hipStream_t stream)
{
  cfg_grid = grid;
  cfg_block = block;
  cfg_shmem = shmem;
  cfg_stream = stream;
  return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* grid, dim3* block, size_t* shmem, hipStream_t* stream)
{
  *grid = cfg_grid;
  *block = cfg_block;
  *shmem = cfg_shmem;
  *stream = cfg_stream;
  return hipSuccess;
}
hipError_t hipLaunchKernel(const void*, dim3 grid, dim3 block, void**, size_t shmem,
                           hipStream_t stream)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (grid.x == 0 || grid.y == 0 || grid.z == 0 || block.x == 0 || block.x > 1024)
    fail("launch with an empty grid or more than 1024 threads");
  if (shmem > 64 * 1024)
    fail("launch with more than 64 KB of LDS (the library's cap)");
  if (grid.y > 65535 || grid.z > 65535)
    fail("launch with grid.y or grid.z above 65535");
  enqueue_on(stream);
  launches++;
  return hipSuccess;
}

hipError_t hipGetDevice(int* d)
{
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t attr, int)
{
  *v = attr == hipDeviceAttributeMaxSharedMemoryPerBlock ? 160 * 1024 : 1;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t* prop, int)
{
  std::memset(prop, 0, sizeof *prop);
  std::strcpy(prop->gcnArchName, "fake");
  prop->multiProcessorCount = 1;
  return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "fake_hip error"; }

hipError_t hipMalloc(void** p, size_t n)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (open_captures)
    fail("hipMalloc while a stream capture is open (the eager pass allocates everything)");
  *p = std::calloc(1, n ? n : 1);
  if (!*p)
    return hipErrorOutOfMemory;
  allocs[*p] = n;
  return hipSuccess;
}
hipError_t hipFree(void* p)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (!p)
    return hipSuccess;
  if (open_captures)
    fail("hipFree while a stream capture is open");
  auto it = allocs.find(p);
  if (it == allocs.end())
    fail("hipFree of a pointer that hipMalloc did not return (or a second free)");
  allocs.erase(it);
  std::free(p);
  return hipSuccess;
}
hipError_t hipPointerGetAttributes(hipPointerAttribute_t* a, const void* p)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  std::memset(a, 0, sizeof *a);
  auto it = allocs.upper_bound(const_cast<void*>(p));
  bool dev = false;
  if (it != allocs.begin()) {
    --it;
    dev = (const char*)p >= (const char*)it->first &&
          (const char*)p < (const char*)it->first + it->second;
  }
  a->type = dev ? hipMemoryTypeDevice : hipMemoryTypeHost;
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** p, const void* symbol)
{
  *p = const_cast<void*>(symbol);
  return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (open_captures)
    fail("blocking hipMemcpy while a stream capture is open");
  std::memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  enqueue_on(s);
  std::memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemset(void* dst, int v, size_t n)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (open_captures)
    fail("blocking hipMemset while a stream capture is open");
  std::memset(dst, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int v, size_t n, hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  no_capture(s, "hipMemsetAsync on a capturing stream (fills are not recorded into graphs)");
  std::memset(dst, v, n);
  return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  *s = reinterpret_cast<hipStream_t>(new char);
  streams[*s] = Stream{};
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  auto it = streams.find(s);
  if (it == streams.end())
    fail("hipStreamDestroy of an unknown stream");
  if (it->second.capturing)
    fail("hipStreamDestroy of a capturing stream");
  streams.erase(it);
  delete reinterpret_cast<char*>(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  no_capture(s, "hipStreamSynchronize on a capturing stream");
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  auto ev = events.find(e);
  if (ev == events.end())
    fail("hipStreamWaitEvent on an unknown event");
  Stream* st = stream_of(s);
  if (!st)
    return hipSuccess;
  if (ev->second.from_capture) {
    Stream* src = stream_of(ev->second.stream);
    if (!st->capturing) {
      // fork: the stream joins the capture with nothing to join yet
      st->capturing = true;
      st->origin = false;
      st->joined = st->ops;
    } else if (src && src != st && src->capturing && !src->origin) {
      // join: the work of the forked stream up to the record is waited for
      src->joined = std::max(src->joined, ev->second.ops);
    }
  }
  return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t* e)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  *e = reinterpret_cast<hipEvent_t>(new char);
  events[*e] = Event{};
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  auto it = events.find(e);
  if (it == events.end())
    fail("hipEventDestroy of an unknown event");
  events.erase(it);
  delete reinterpret_cast<char*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  auto it = events.find(e);
  if (it == events.end())
    fail("hipEventRecord of an unknown event");
  Stream* st = stream_of(s);
  it->second.stream = s;
  it->second.from_capture = st && st->capturing;
  it->second.ops = st ? st->ops : 0;
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  auto it = events.find(e);
  if (it == events.end())
    fail("hipEventSynchronize of an unknown event");
  if (it->second.from_capture)
    fail("hipEventSynchronize on an event recorded in a capture");
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (!events.count(a) || !events.count(b))
    fail("hipEventElapsedTime of an unknown event");
  *ms = 0.1f;
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  Stream* st = stream_of(s);
  if (!st)
    fail("capture on an unknown stream");
  if (st->capturing)
    fail("hipStreamBeginCapture on a capturing stream");
  st->capturing = true;
  st->origin = true;
  open_captures++;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  Stream* st = stream_of(s);
  if (!st || !st->capturing || !st->origin)
    fail("hipStreamEndCapture on a stream that did not begin a capture");
  for (auto& kv : streams)
    if (kv.second.capturing && !kv.second.origin) {
      if (kv.second.ops > kv.second.joined)
        fail("capture ended with unjoined work on a forked stream");
      kv.second.capturing = false;
    }
  st->capturing = false;
  st->origin = false;
  open_captures--;
  *g = reinterpret_cast<hipGraph_t>(new char);
  graphs.insert(*g);
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (!graphs.count(g))
    fail("hipGraphInstantiate of an unknown graph");
  *x = reinterpret_cast<hipGraphExec_t>(new char);
  execs.insert(*x);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (!execs.count(x))
    fail("hipGraphLaunch of an unknown or destroyed graph executable");
  no_capture(s, "hipGraphLaunch on a capturing stream");
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (!execs.erase(x))
    fail("hipGraphExecDestroy of an unknown graph executable");
  delete reinterpret_cast<char*>(x);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
  std::lock_guard<std::recursive_mutex> lk(mu);
  if (!graphs.erase(g))
    fail("hipGraphDestroy of an unknown graph");
  delete reinterpret_cast<char*>(g);
  return hipSuccess;
}

// counters for the driver's report
long long fake_hip_launches() { return launches; }
size_t fake_hip_live_allocations() { return allocs.size(); }

} // extern "C"
