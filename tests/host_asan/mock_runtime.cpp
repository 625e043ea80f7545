// Host-only stand-in for the HIP runtime, for running the library's host code under
// AddressSanitizer on a machine without a GPU: device memory is host memory, kernels are
// not run, streams / events / graphs / graph executables are heap objects (a use after
// destroy is an AddressSanitizer report), and every kernel launch is checked: grid and
// block within the hardware limits, dynamic LDS within a CU's 160 KB and, above 64 KB,
// preceded by hipFuncSetAttribute on that kernel.  Results are meaningless; the host
// logic (work lists, tuners, graph capture and cache, plan tables) runs as on a GPU.
// The __hip_fatbin_* symbols the objects reference are defined in fatbins.inc, written
// by the test from `nm` of the host objects.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

typedef int hipError_t;
struct dim3 { unsigned x, y, z; };

namespace {
struct Obj { int magic; int kind; int capturing; };
constexpr int kMagic = 0x5eed;
std::mutex mu;
std::map<const void *, int> lds_attr;   // kernel -> hipFuncAttributeMaxDynamicSharedMemorySize
dim3 cfg_grid, cfg_block;
size_t cfg_shm;
void *cfg_stream;
long n_launch = 0;
char *symbol_buffer = nullptr;

Obj *make(int kind) { return new Obj{kMagic, kind, 0}; }
void check(const void *p, const char *what)
{
    if (!p) return;   // the null stream
    if (static_cast<const Obj *>(p)->magic != kMagic) {
        fprintf(stderr, "mock runtime: bad handle in %s\n", what);
        abort();
    }
}
void destroy(void *p)
{
    if (p) delete static_cast<Obj *>(p);
}
[[noreturn]] void fail(const char *msg)
{
    fprintf(stderr, "mock runtime: %s\n", msg);
    abort();
}
}  // namespace

extern "C" {
#include "fatbins.inc"

hipError_t hipMalloc(void **p, size_t n) { *p = malloc(n ? n : 1); return 0; }
hipError_t hipFree(void *p) { free(p); return 0; }
hipError_t hipMallocAsync(void **p, size_t n, void *s) { check(s, "hipMallocAsync"); *p = malloc(n ? n : 1); return 0; }
hipError_t hipFreeAsync(void *p, void *s) { check(s, "hipFreeAsync"); free(p); return 0; }
hipError_t hipMemcpy(void *d, const void *s, size_t n, int) { if (n) memcpy(d, s, n); return 0; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, void *st)
{
    check(st, "hipMemcpyAsync");
    if (n) memcpy(d, s, n);
    return 0;
}
hipError_t hipMemset(void *d, int v, size_t n) { memset(d, v, n); return 0; }
hipError_t hipMemsetAsync(void *d, int v, size_t n, void *st) { check(st, "hipMemsetAsync"); memset(d, v, n); return 0; }
hipError_t hipGetLastError() { return 0; }
const char *hipGetErrorString(hipError_t) { return "mock runtime"; }
hipError_t hipGetDevice(int *d) { *d = 0; return 0; }
hipError_t hipSetDevice(int) { return 0; }
hipError_t hipDeviceSynchronize() { return 0; }
hipError_t hipDeviceGetAttribute(int *v, int, int) { *v = 256; return 0; }
hipError_t hipFuncSetAttribute(const void *f, int, int v)
{
    if (v > 160 * 1024) fail("LDS attribute above 160 KB");
    std::lock_guard<std::mutex> l(mu);
    lds_attr[f] = v;
    return 0;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int bs, size_t shm)
{
    if (bs <= 0 || bs > 1024 || shm > 160 * 1024) fail("occupancy query out of range");
    *n = 2;
    return 0;
}
hipError_t hipGetSymbolAddress(void **p, const void *)
{
    if (!symbol_buffer) symbol_buffer = static_cast<char *>(malloc(size_t(256) << 20));
    *p = symbol_buffer;
    return 0;
}
hipError_t hipStreamCreateWithFlags(void **s, unsigned) { *s = make(1); return 0; }
hipError_t hipStreamDestroy(void *s) { check(s, "hipStreamDestroy"); destroy(s); return 0; }
hipError_t hipStreamSynchronize(void *s) { check(s, "hipStreamSynchronize"); return 0; }
hipError_t hipStreamWaitEvent(void *s, void *e, unsigned)
{
    check(s, "hipStreamWaitEvent");
    check(e, "hipStreamWaitEvent");
    return 0;
}
hipError_t hipStreamIsCapturing(void *s, int *st)
{
    check(s, "hipStreamIsCapturing");
    *st = s ? static_cast<Obj *>(s)->capturing : 0;
    return 0;
}
hipError_t hipStreamBeginCapture(void *s, int)
{
    check(s, "hipStreamBeginCapture");
    if (!s) fail("capture on the null stream");
    static_cast<Obj *>(s)->capturing = 1;
    return 0;
}
hipError_t hipStreamEndCapture(void *s, void **g)
{
    check(s, "hipStreamEndCapture");
    static_cast<Obj *>(s)->capturing = 0;
    *g = make(3);
    return 0;
}
hipError_t hipEventCreate(void **e) { *e = make(2); return 0; }
hipError_t hipEventCreateWithFlags(void **e, unsigned) { *e = make(2); return 0; }
hipError_t hipEventDestroy(void *e) { check(e, "hipEventDestroy"); destroy(e); return 0; }
hipError_t hipEventRecord(void *e, void *s) { check(e, "hipEventRecord"); check(s, "hipEventRecord"); return 0; }
hipError_t hipEventSynchronize(void *e) { check(e, "hipEventSynchronize"); return 0; }
hipError_t hipEventElapsedTime(float *ms, void *a, void *b)
{
    check(a, "hipEventElapsedTime");
    check(b, "hipEventElapsedTime");
    *ms = 1.0f + (float)(rand() % 100) * 0.01f;   // varied, so that every tuner branch is taken
    return 0;
}
hipError_t hipGraphInstantiate(void **ex, void *g, void *, char *, size_t) { check(g, "hipGraphInstantiate"); *ex = make(4); return 0; }
hipError_t hipGraphExecUpdate(void *ex, void *g, void **, int *res)
{
    check(ex, "hipGraphExecUpdate");
    check(g, "hipGraphExecUpdate");
    *res = 0;
    return 0;
}
hipError_t hipGraphLaunch(void *ex, void *s) { check(ex, "hipGraphLaunch"); check(s, "hipGraphLaunch"); return 0; }
hipError_t hipGraphDestroy(void *g) { check(g, "hipGraphDestroy"); destroy(g); return 0; }
hipError_t hipGraphExecDestroy(void *ex) { check(ex, "hipGraphExecDestroy"); destroy(ex); return 0; }

hipError_t __hipPushCallConfiguration(dim3 g, dim3 b, size_t shm, void *s)
{
    cfg_grid = g; cfg_block = b; cfg_shm = shm; cfg_stream = s;
    return 0;
}
hipError_t __hipPopCallConfiguration(dim3 *g, dim3 *b, size_t *shm, void **s)
{
    *g = cfg_grid; *b = cfg_block; *shm = cfg_shm; *s = cfg_stream;
    return 0;
}
hipError_t hipLaunchKernel(const void *f, dim3 g, dim3 b, void **, size_t shm, void *s)
{
    check(s, "hipLaunchKernel");
    const unsigned long long lim = 1ull << 32;
    const unsigned threads = b.x * b.y * b.z;
    if (g.x == 0 || g.y == 0 || g.z == 0 || threads == 0 || threads > 1024) fail("empty grid or block above 1024 threads");
    if ((unsigned long long)g.x * b.x >= lim || (unsigned long long)g.y * b.y >= lim ||
        (unsigned long long)g.z * b.z >= lim)
        fail("grid of 2^32 work-items or more in one dimension");
    if (shm > 160 * 1024) fail("dynamic LDS above 160 KB");
    std::lock_guard<std::mutex> l(mu);
    if (shm > 65536 && lds_attr[f] < (int)shm) fail("dynamic LDS above 64 KB without hipFuncSetAttribute");
    ++n_launch;
    return 0;
}
void **__hipRegisterFatBinary(const void *) { static void *handle = nullptr; return &handle; }
void __hipRegisterFunction(void **, const void *, char *, const char *, int, void *, void *, void *, void *, int *) {}
void __hipRegisterVar(void **, void *, char *, const char *, int, size_t, int, int) {}
void __hipUnregisterFatBinary(void **) {}

long mock_launches() { return n_launch; }
}
