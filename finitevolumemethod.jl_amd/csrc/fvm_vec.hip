// Device vector operations used by the host-side integrators and solvers: fused
// linear combinations and deterministic reductions (fixed grid, fixed order).
#include <mutex>

#include "fvm_internal.hpp"

using namespace fvm;

namespace {

constexpr int kRedBlocks = 512;  // fixed: the summation order never depends on timing
constexpr int kRedThreads = 256;
constexpr int kMaxTerms = 8;

struct Terms {
  int n;
  double c[kMaxTerms];
  const double* x[kMaxTerms];
};

// out[i] = c_0 x_0[i] + c_1 x_1[i] + ..., the terms added left to right.  out may
// alias an input: each element is read and written by the same thread.
template <int NT>
__device__ __forceinline__ double comb(const Terms& t, const double* xv) {
  double v = t.c[0] * xv[0];
#pragma unroll
  for (int k = 1; k < NT; ++k) v = v + t.c[k] * xv[k];
  return v;
}

template <int NT>
__global__ __launch_bounds__(256) void lincomb_kernel(int64_t n, double* out, Terms t) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double xv[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) xv[k] = t.x[k][i];
    out[i] = comb<NT>(t, xv);
  }
}

// The same combination over pairs of elements with 16-byte loads and stores (every
// vector 16-byte aligned), two pairs per thread and iteration with all loads issued
// before the first store; the last element of an odd n is done by thread 0.
template <int NT>
__global__ __launch_bounds__(256) void lincomb2_kernel(int64_t n, double* out, Terms t) {
  const int64_t np = n >> 1;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += 2 * stride) {
    const int64_t p2 = p + stride;
    const bool two = p2 < np;
    double2 a[NT], b[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      a[k] = reinterpret_cast<const double2*>(t.x[k])[p];
      if (two) b[k] = reinterpret_cast<const double2*>(t.x[k])[p2];
    }
    double xa[NT], ya[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      xa[k] = a[k].x;
      ya[k] = a[k].y;
    }
    reinterpret_cast<double2*>(out)[p] = make_double2(comb<NT>(t, xa), comb<NT>(t, ya));
    if (two) {
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        xa[k] = b[k].x;
        ya[k] = b[k].y;
      }
      reinterpret_cast<double2*>(out)[p2] = make_double2(comb<NT>(t, xa), comb<NT>(t, ya));
    }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    double xv[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) xv[k] = t.x[k][n - 1];
    out[n - 1] = comb<NT>(t, xv);
  }
}

template <int NT>
void launch_lincomb(int64_t n, double* out, const Terms& t, bool aligned, hipStream_t s) {
  if (aligned) {
    int64_t blocks = ((n >> 1) + 511) / 512;  // two pairs per thread
    blocks = blocks < 1 ? 1 : (blocks > 8192 ? 8192 : blocks);
    hipLaunchKernelGGL(lincomb2_kernel<NT>, dim3((unsigned)blocks), dim3(256), 0, s, n, out, t);
  } else {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(lincomb_kernel<NT>, dim3((unsigned)blocks), dim3(256), 0, s, n, out, t);
  }
}

__global__ __launch_bounds__(256) void mul_kernel(int64_t n, double* __restrict__ out,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = x[i] * y[i];
}

// kind 0: sum x*y;  kind 1: sum (e / (abstol + reltol*max(|a|,|b|)))^2
__device__ __forceinline__ double red_term(int kind, const double* x, const double* y,
                                           const double* a, const double* b, double abstol,
                                           double reltol, int64_t i) {
  if (kind == 0) return x[i] * y[i];
  const double sc = abstol + reltol * fmax(fabs(a[i]), fabs(b[i]));
  const double r = x[i] / sc;
  return r * r;
}

__global__ __launch_bounds__(kRedThreads) void reduce_stage1(int kind, int64_t n, const double* x,
                                                             const double* y, const double* a,
                                                             const double* b, double abstol,
                                                             double reltol, double* part) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += red_term(kind, x, y, a, b, abstol, reltol, i);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// The error estimate of a Runge-Kutta step without storing the error vector: the
// kind-1 reduction of err = sum_k c_k x_k, each err_i formed as lincomb forms it and
// summed in reduce_stage1's order, so the result equals lincomb followed by the
// reduction bit for bit.
template <int NT>
__global__ __launch_bounds__(kRedThreads) void lc_err_stage1(int64_t n, Terms t, const double* a,
                                                             const double* b, double abstol, double reltol,
                                                             double* part) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double xv[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) xv[k] = t.x[k][i];
    const double e = comb<NT>(t, xv);
    const double sc = abstol + reltol * fmax(fabs(a[i]), fabs(b[i]));
    const double r = e / sc;
    acc += r * r;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kRedThreads) void reduce_stage2(const double* part, double* res) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) acc += part[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *res = sh[0];
}

// Several dot products x_j . y in one pass over y (the Gram-Schmidt step of GMRES):
// each workgroup leaves one partial per vector, reduced in a fixed order (waves by
// shuffles, then the block's waves in order), then one workgroup per vector sums the
// kRedBlocks partials in order: bitwise reproducible.
constexpr int kMaxDots = 64;
constexpr int kDotChunk = 16;  // vectors per pass (their accumulators live in registers)

struct DotChunk {
  int k;                        // vectors in this chunk
  int first;                    // index of the chunk's first vector
  const double* x[kDotChunk];
};

__global__ __launch_bounds__(kRedThreads) void mdot_stage1(int64_t n, DotChunk c, const double* __restrict__ y,
                                                           double* __restrict__ part) {
  __shared__ double sh[kRedThreads / 64][kDotChunk];
  double acc[kDotChunk];
#pragma unroll
  for (int j = 0; j < kDotChunk; ++j) acc[j] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double yi = y[i];
#pragma unroll
    for (int j = 0; j < kDotChunk; ++j)
      if (j < c.k) acc[j] += c.x[j][i] * yi;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kDotChunk; ++j) {
    double v = acc[j];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) sh[wave][j] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < c.k) {
    double v = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) v += sh[w][threadIdx.x];
    part[(int64_t)(c.first + threadIdx.x) * kRedBlocks + blockIdx.x] = v;
  }
}

__global__ __launch_bounds__(kRedThreads) void mdot_stage2(const double* __restrict__ part, double* __restrict__ res) {
  __shared__ double sh[kRedThreads];
  const double* p = part + (int64_t)blockIdx.x * kRedBlocks;
  double acc = 0.0;
  for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) acc += p[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) res[blockIdx.x] = sh[0];
}

struct AxpyTerms {
  int k;
  double c[kMaxDots];
  const double* x[kMaxDots];
};

// y += sum_j c_j x_j
__global__ __launch_bounds__(256) void maxpy_kernel(int64_t n, double* __restrict__ y, AxpyTerms t) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = y[i];
    for (int j = 0; j < t.k; ++j) v = v + t.c[j] * t.x[j][i];
    y[i] = v;
  }
}

// y -= sum_j c_j x_j with the coefficients c in device memory (the Gram-Schmidt
// step of an Arnoldi iteration whose dot products never left the device); the same
// arithmetic, in the same order, as maxpy_kernel given -c
__global__ __launch_bounds__(256) void maxpy_neg_dev_kernel(int64_t n, double* __restrict__ y, AxpyTerms t,
                                                            const double* __restrict__ c) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = y[i];
    for (int j = 0; j < t.k; ++j) v = v + (-c[j]) * t.x[j][i];
    y[i] = v;
  }
}

std::mutex g_mu;   // allocation of the scratch buffers
// The reductions share these buffers: each call holds g_use from its first launch
// to its final synchronisation, so calls from several host threads (on different
// streams) are serialised instead of mixing their partial sums.
std::mutex g_use;
double* g_ascratch = nullptr;  // Arnoldi step results: 2 kMaxDots coefficients + |w|^2
double* g_scratch = nullptr;   // kRedBlocks partials + 1 result
double* g_mscratch = nullptr;  // kMaxDots * kRedBlocks partials + kMaxDots results

int red_scratch(double** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch) FVM_HIP(hipMalloc((void**)&g_scratch, sizeof(double) * (kRedBlocks + 8)));
  *out = g_scratch;
  return FVM_OK;
}

int mscratch(double** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_mscratch) FVM_HIP(hipMalloc((void**)&g_mscratch, sizeof(double) * (size_t)kMaxDots * (kRedBlocks + 1)));
  *out = g_mscratch;
  return FVM_OK;
}

int reduce(int kind, int64_t n, const double* x, const double* y, const double* a, const double* b,
           double abstol, double reltol, double* result, hipStream_t s) {
  double* buf;
  int rc = red_scratch(&buf);
  if (rc) return rc;
  std::lock_guard<std::mutex> use(g_use);
  hipLaunchKernelGGL(reduce_stage1, dim3(kRedBlocks), dim3(kRedThreads), 0, s, kind, n, x, y, a, b,
                     abstol, reltol, buf);
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(kRedThreads), 0, s, buf, buf + kRedBlocks);
  FVM_HIP(hipGetLastError());
  FVM_HIP(hipMemcpyAsync(result, buf + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s));
  FVM_HIP(hipStreamSynchronize(s));
  return FVM_OK;
}

}  // namespace

// Scratch of fvm_vec_mdot_noalloc: one partial per (vector, workgroup), kMaxDots
// vectors (256.5 KiB per device, held until the process ends).  Host code reaches it
// only through the address hipGetSymbolAddress returns on each device.
static __device__ __align__(256) double scratch[(size_t)kMaxDots * kRedBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
double* g_na_ptr[kMaxDevices] = {};  // scratch's device address, per device

// The current device's copy of scratch, looked up once per device.
int noalloc_scratch(double** out) {
  int dev = 0;
  FVM_HIP(hipGetDevice(&dev));
  FVM_CHECK(dev >= 0 && dev < kMaxDevices, "fvm_vec_mdot_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    FVM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FVM_CHECK(p, "fvm_vec_mdot_noalloc: no device address for the scratch array");
    g_na_ptr[dev] = (double*)p;
  }
  *out = g_na_ptr[dev];
  return FVM_OK;
}
}  // namespace

namespace fvm {

// Classical Gram-Schmidt applied twice to w against xs[0..k) and |w|^2, with every
// intermediate on the device and one host synchronisation at the end:
// h_host[0..k) and h_host[k..2k) the two passes' coefficients, h_host[2k] = |w|^2.
int cgs2_norm(int64_t n, int k, const double* const* xs, double* w, double* h_host, hipStream_t s) {
  if (k < 1 || k > kMaxDots) return FVM_ERR_ARG;
  double *buf, *red;
  int rc = mscratch(&buf);
  if (rc || (rc = red_scratch(&red))) return rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ascratch) FVM_HIP(hipMalloc((void**)&g_ascratch, sizeof(double) * (2 * kMaxDots + 8)));
  }
  std::lock_guard<std::mutex> use(g_use);
  AxpyTerms t;
  t.k = k;
  for (int j = 0; j < k; ++j) {
    t.c[j] = 0.0;
    t.x[j] = xs[j];
  }
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  // results packed as [pass 0 coefficients | pass 1 coefficients | |w|^2]: one copy back
  for (int pass = 0; pass < 2; ++pass) {
    double* res = g_ascratch + pass * k;
    for (int first = 0; first < k; first += kDotChunk) {
      DotChunk c;
      c.first = first;
      c.k = k - first < kDotChunk ? k - first : kDotChunk;
      for (int j = 0; j < kDotChunk; ++j) c.x[j] = j < c.k ? xs[first + j] : xs[first];
      hipLaunchKernelGGL(mdot_stage1, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, c, w, buf);
    }
    hipLaunchKernelGGL(mdot_stage2, dim3((unsigned)k), dim3(kRedThreads), 0, s, buf, res);
    hipLaunchKernelGGL(maxpy_neg_dev_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n, w, t, res);
  }
  hipLaunchKernelGGL(reduce_stage1, dim3(kRedBlocks), dim3(kRedThreads), 0, s, 0, n, w, w, nullptr, nullptr,
                     0.0, 0.0, red);
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(kRedThreads), 0, s, red, g_ascratch + 2 * k);
  FVM_HIP(hipGetLastError());
  FVM_HIP(hipMemcpyAsync(h_host, g_ascratch, sizeof(double) * (2 * k + 1), hipMemcpyDeviceToHost, s));
  FVM_HIP(hipStreamSynchronize(s));
  return FVM_OK;
}

}  // namespace fvm

extern "C" {

int fvm_vec_lincomb(int64_t n, double* d_out, int nterms, const double* coefs,
                    const double* const* d_xs, fvm_stream s) {
  FVM_CHECK(d_out && coefs && d_xs, "fvm_vec_lincomb: null argument");
  FVM_CHECK(nterms >= 1 && nterms <= kMaxTerms, "fvm_vec_lincomb: 1..8 terms");
  if (n <= 0) return FVM_OK;
  Terms t;
  t.n = nterms;
  for (int k = 0; k < nterms; ++k) {
    FVM_CHECK(d_xs[k], "fvm_vec_lincomb: null input vector");
    t.c[k] = coefs[k];
    t.x[k] = d_xs[k];
  }
  bool aligned = ((uintptr_t)d_out & 15) == 0;
  for (int k = 0; k < nterms; ++k) aligned = aligned && ((uintptr_t)d_xs[k] & 15) == 0;
  const hipStream_t st = (hipStream_t)s;
  switch (nterms) {
    case 1: launch_lincomb<1>(n, d_out, t, aligned, st); break;
    case 2: launch_lincomb<2>(n, d_out, t, aligned, st); break;
    case 3: launch_lincomb<3>(n, d_out, t, aligned, st); break;
    case 4: launch_lincomb<4>(n, d_out, t, aligned, st); break;
    case 5: launch_lincomb<5>(n, d_out, t, aligned, st); break;
    case 6: launch_lincomb<6>(n, d_out, t, aligned, st); break;
    case 7: launch_lincomb<7>(n, d_out, t, aligned, st); break;
    default: launch_lincomb<8>(n, d_out, t, aligned, st); break;
  }
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

int fvm_vec_lincomb_err_norm(int64_t n, int nterms, const double* coefs, const double* const* d_xs,
                             const double* d_u0, const double* d_u1, double abstol, double reltol,
                             double* result, fvm_stream s) {
  FVM_CHECK(coefs && d_xs && d_u0 && d_u1 && result, "fvm_vec_lincomb_err_norm: null argument");
  FVM_CHECK(nterms >= 1 && nterms <= kMaxTerms, "fvm_vec_lincomb_err_norm: 1..8 terms");
  if (n <= 0) {
    *result = 0.0;
    return FVM_OK;
  }
  Terms t;
  t.n = nterms;
  for (int k = 0; k < nterms; ++k) {
    FVM_CHECK(d_xs[k], "fvm_vec_lincomb_err_norm: null input vector");
    t.c[k] = coefs[k];
    t.x[k] = d_xs[k];
  }
  double* buf;
  int rc = red_scratch(&buf);
  if (rc) return rc;
  std::lock_guard<std::mutex> use(g_use);
  const hipStream_t st = (hipStream_t)s;
  const dim3 g(kRedBlocks), b(kRedThreads);
  switch (nterms) {
    case 1: hipLaunchKernelGGL(lc_err_stage1<1>, g, b, 0, st, n, t, d_u0, d_u1, abstol, reltol, buf); break;
    case 2: hipLaunchKernelGGL(lc_err_stage1<2>, g, b, 0, st, n, t, d_u0, d_u1, abstol, reltol, buf); break;
    case 3: hipLaunchKernelGGL(lc_err_stage1<3>, g, b, 0, st, n, t, d_u0, d_u1, abstol, reltol, buf); break;
    case 4: hipLaunchKernelGGL(lc_err_stage1<4>, g, b, 0, st, n, t, d_u0, d_u1, abstol, reltol, buf); break;
    case 5: hipLaunchKernelGGL(lc_err_stage1<5>, g, b, 0, st, n, t, d_u0, d_u1, abstol, reltol, buf); break;
    case 6: hipLaunchKernelGGL(lc_err_stage1<6>, g, b, 0, st, n, t, d_u0, d_u1, abstol, reltol, buf); break;
    case 7: hipLaunchKernelGGL(lc_err_stage1<7>, g, b, 0, st, n, t, d_u0, d_u1, abstol, reltol, buf); break;
    default: hipLaunchKernelGGL(lc_err_stage1<8>, g, b, 0, st, n, t, d_u0, d_u1, abstol, reltol, buf); break;
  }
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(kRedThreads), 0, st, buf, buf + kRedBlocks);
  FVM_HIP(hipGetLastError());
  double sum = 0.0;
  FVM_HIP(hipMemcpyAsync(&sum, buf + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
  FVM_HIP(hipStreamSynchronize(st));
  *result = sqrt(sum / (double)n);
  return FVM_OK;
}

int fvm_vec_mdot(int64_t n, int32_t k, const double* const* d_xs, const double* d_y, double* results,
                 fvm_stream s) {
  FVM_CHECK(d_xs && d_y && results, "fvm_vec_mdot: null argument");
  FVM_CHECK(k >= 1 && k <= kMaxDots, "fvm_vec_mdot: 1..64 vectors");
  for (int j = 0; j < k; ++j) FVM_CHECK(d_xs[j], "fvm_vec_mdot: null input vector");
  double* buf;
  int rc = mscratch(&buf);
  if (rc) return rc;
  std::lock_guard<std::mutex> use(g_use);
  hipStream_t st = (hipStream_t)s;
  for (int first = 0; first < k; first += kDotChunk) {
    DotChunk c;
    c.first = first;
    c.k = k - first < kDotChunk ? k - first : kDotChunk;
    for (int j = 0; j < kDotChunk; ++j) c.x[j] = j < c.k ? d_xs[first + j] : d_xs[first];
    hipLaunchKernelGGL(mdot_stage1, dim3(kRedBlocks), dim3(kRedThreads), 0, st, n, c, d_y, buf);
  }
  double* res = buf + (size_t)kMaxDots * kRedBlocks;
  hipLaunchKernelGGL(mdot_stage2, dim3((unsigned)k), dim3(kRedThreads), 0, st, buf, res);
  FVM_HIP(hipGetLastError());
  FVM_HIP(hipMemcpyAsync(results, res, sizeof(double) * (size_t)k, hipMemcpyDeviceToHost, st));
  FVM_HIP(hipStreamSynchronize(st));
  return FVM_OK;
}

int fvm_vec_mdot_noalloc(int64_t n, int32_t k, const double* const* d_xs, const double* d_y, double* d_results,
                         fvm_stream s) {
  FVM_CHECK(d_xs && d_y && d_results, "fvm_vec_mdot_noalloc: null argument");
  FVM_CHECK(k >= 1 && k <= kMaxDots, "fvm_vec_mdot_noalloc: k must be 1..64 (the scratch holds 64 vectors' partials)");
  FVM_CHECK(n >= 0, "fvm_vec_mdot_noalloc: negative length");
  for (int j = 0; j < k; ++j) FVM_CHECK(d_xs[j], "fvm_vec_mdot_noalloc: null input vector");
  double* buf;
  int rc = noalloc_scratch(&buf);
  if (rc) return rc;
  // the same kernels, grid and summation order as fvm_vec_mdot: every partial is
  // rewritten by this call, so nothing needs clearing
  hipStream_t st = (hipStream_t)s;
  for (int first = 0; first < k; first += kDotChunk) {
    DotChunk c;
    c.first = first;
    c.k = k - first < kDotChunk ? k - first : kDotChunk;
    for (int j = 0; j < kDotChunk; ++j) c.x[j] = j < c.k ? d_xs[first + j] : d_xs[first];
    hipLaunchKernelGGL(mdot_stage1, dim3(kRedBlocks), dim3(kRedThreads), 0, st, n, c, d_y, buf);
  }
  hipLaunchKernelGGL(mdot_stage2, dim3((unsigned)k), dim3(kRedThreads), 0, st, buf, d_results);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

int fvm_vec_maxpy(int64_t n, double* d_y, int32_t k, const double* coefs, const double* const* d_xs,
                  fvm_stream s) {
  FVM_CHECK(d_y && coefs && d_xs, "fvm_vec_maxpy: null argument");
  FVM_CHECK(k >= 1 && k <= kMaxDots, "fvm_vec_maxpy: 1..64 vectors");
  if (n <= 0) return FVM_OK;
  AxpyTerms t;
  t.k = k;
  for (int j = 0; j < k; ++j) {
    FVM_CHECK(d_xs[j] && d_xs[j] != d_y, "fvm_vec_maxpy: null or aliased input vector");
    t.c[j] = coefs[j];
    t.x[j] = d_xs[j];
  }
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(maxpy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s, n, d_y, t);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

int fvm_vec_mul(int64_t n, double* d_out, const double* d_x, const double* d_y, fvm_stream s) {
  FVM_CHECK(d_out && d_x && d_y, "fvm_vec_mul: null argument");
  if (n <= 0) return FVM_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(mul_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s, n, d_out, d_x, d_y);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

int fvm_vec_dot(int64_t n, const double* d_x, const double* d_y, double* result, fvm_stream s) {
  FVM_CHECK(d_x && d_y && result, "fvm_vec_dot: null argument");
  return reduce(0, n, d_x, d_y, nullptr, nullptr, 0, 0, result, (hipStream_t)s);
}

int fvm_vec_err_norm(int64_t n, const double* d_err, const double* d_u0, const double* d_u1,
                     double abstol, double reltol, double* result, fvm_stream s) {
  FVM_CHECK(d_err && d_u0 && d_u1 && result, "fvm_vec_err_norm: null argument");
  double sum = 0;
  int rc = reduce(1, n, d_err, nullptr, d_u0, d_u1, abstol, reltol, &sum, (hipStream_t)s);
  if (rc) return rc;
  *result = n > 0 ? sqrt(sum / (double)n) : 0.0;
  return FVM_OK;
}

}  // extern "C"
