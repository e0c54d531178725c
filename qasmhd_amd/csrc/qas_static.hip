// qas_static.hip -- the allocation-free run (qas_run_static, include/qasmhd.h).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the
// exec_ loop of qas_run (graphs off) on a state the caller keeps in device
// memory, with the RK stage buffers and every scratch array (primitives,
// sweep accumulators, face parts of the B fluxes, edge EMFs, dt, CFL
// partial maxima) carved from one file-scope device array, `scratch`, of
// QAS_STATIC_SCRATCH_BYTES.  The launches are the ones qas_run makes, so the
// results are those of qas_run bit for bit.
//
// Each GPU holds its own copy of `scratch`; its address on a device comes
// from hipGetSymbolAddress, taken once per device with that device current
// (the array's name in host code is a host-side stand-in, never passed to a
// kernel).  Every call on one GPU shares that GPU's array and leaves its
// contents behind: a call clears the part it uses before its first kernel,
// and calls that overlap in time on one GPU race (order them, e.g. on one
// stream).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "qas_internal.hpp"

static __device__ __align__(256) unsigned char scratch[QAS_STATIC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};  // device address of `scratch` per device

int say(char* err, int errlen, int code, const char* fmt, ...)
{
  if (err && errlen > 0) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
  }
  return code;
}

__global__ __launch_bounds__(256) void k_zero_words(unsigned long long* p, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

struct Arrays8 {
  const double* a[8];
};

// the state from the RK buffer it ended in back into the caller's array
__global__ __launch_bounds__(256) void k_copy_state(Arrays8 src, double* dst, long long nd)
{
  const long long n = 8 * nd;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(i / nd);
    dst[i] = src.a[m][i - (long long)m * nd];
  }
}

__global__ void k_copy_dt(const double* src, double* dst) { *dst = *src; }

}  // namespace

using namespace qas;

extern "C" int qas_run_static(const qas_grid* g, const qas_scheme* s, double* d_state, int nsteps, long long nstep0,
                              int flg, double dt, double* d_dt_out, void* stream, char* err, int errlen)
{
  if (err && errlen > 0) err[0] = '\0';
  if (!d_state || nsteps < 0 || nstep0 < 0)
    return say(err, errlen, QAS_E_ARG, "qas_run_static: null state or negative step count");
  const std::string v = validate_config(g, s);
  if (!v.empty()) return say(err, errlen, QAS_E_ARG, "qas_run_static: %s", v.c_str());

  qas_ctx c;  // on the stack: no member allocates while it is empty
  ctx_init(&c, g, s);
  c.use_graph = false;
  const size_t need = ctx_layout(&c, nullptr, d_state);
  if (need > (size_t)QAS_STATIC_SCRATCH_BYTES)
    return say(err, errlen, QAS_E_ARG,
               "qas_run_static: the domain needs %zu bytes of scratch, the static array holds %zu "
               "(use qas_create / qas_run for larger domains)",
               need, (size_t)QAS_STATIC_SCRATCH_BYTES);

  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return say(err, errlen, QAS_E_HIP, "qas_run_static: hipGetDevice: %s", hipGetErrorName(e));
  if (dev < 0 || dev >= MAX_DEVICES) return say(err, errlen, QAS_E_ARG, "qas_run_static: device %d out of range", dev);
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
      // once per device, with that device current
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return say(err, errlen, QAS_E_HIP, "qas_run_static: hipGetSymbolAddress(scratch): %s", hipGetErrorName(e));
      g_ptr[dev] = p;
    }
    base = g_ptr[dev];
  }
  c.device = dev;
  c.stream = (hipStream_t)stream;
  ctx_layout(&c, (char*)base, d_state);
  c.slab = base;
  c.slab_bytes = need;
  c.uploaded = true;

#define QAS_SCHECK(expr)                                                                                  \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return say(err, errlen, QAS_E_HIP, "qas_run_static: %s: %s", #expr, hipGetErrorName(e_)); \
  } while (0)

  // scratch starts at zero, as a context's slab does after qas_create
  {
    const long long words = (long long)(need / sizeof(unsigned long long));
    long long nb = (words + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)nb), dim3(256), 0, c.stream, (unsigned long long*)base, words);
    QAS_SCHECK(hipGetLastError());
  }
  // dt of the first step, then the loop of qas_run (the bound before the
  // first step, setdt after even steps of the running count when flg)
  if (dt > 0.0) {
    QAS_SCHECK(launch_set_dt(&c, dt));
  } else {
    QAS_SCHECK(launch_cfl(&c, c.U[c.cur]));
    QAS_SCHECK(launch_cfl_finalize(&c, nullptr));
  }
  QAS_SCHECK(launch_bound(&c, c.U[c.cur], 8));
  for (int k = 1; k <= nsteps; k++) {
    const long long n = nstep0 + k;
    const int r = ctx_step(&c);
    if (r) return say(err, errlen, r, "qas_run_static: %s", c.err.c_str());
    if (flg && n % 2 == 0) {
      QAS_SCHECK(launch_cfl(&c, c.U[c.cur]));
      QAS_SCHECK(launch_cfl_finalize(&c, nullptr));
    }
  }
  // R_K = 1 rotates through the three buffers: bring the state home
  if (c.cur != 0) {
    Arrays8 src;
    for (int m = 0; m < 8; m++) src.a[m] = c.U[c.cur][m];
    long long nb = (8 * c.geo.nd + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_copy_state, dim3((unsigned)nb), dim3(256), 0, c.stream, src, d_state, c.geo.nd);
    QAS_SCHECK(hipGetLastError());
  }
  if (d_dt_out) {
    hipLaunchKernelGGL(k_copy_dt, dim3(1), dim3(1), 0, c.stream, (const double*)c.d_dt, d_dt_out);
    QAS_SCHECK(hipGetLastError());
  }
#undef QAS_SCHECK
  return QAS_OK;
}
