// Device-resident steepest-descent and L-BFGS steps with FP64 dot products (include/fwi.h).
//
// Replaces src/main.c:539-800: the reference copies every vector it needs a dot product of to the host
// (about 81 nx*nz transfers per iteration at memory 5) and sums in float; here the optimizer state lives
// on the device in FP64 and each dot product is a deterministic two-pass reduction (fixed block
// partition, fixed tree order), so a result is bit-reproducible from run to run.
//
// Blocks are kept in the order of the gradient produced by fwi_gradient: 0 = v_hor (the reference's
// gcx), 1 = v_po (gcz), 2 = v_nmo (gdx).  The history is a ring of m (s, y) pairs per block ordered oldest
// to newest, which is what the reference's slot shift (main.c:791-800) intends.

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fwi.h"
#include "fwi_common.h"
#include "fwi_internal.h"

namespace fwi {
namespace {

constexpr int kDotBlocks = 512;

// partial[b*3 + k] = sum over the block's cells of a_k * b_k, for the three blocks k.
__global__ void __launch_bounds__(256) k_dot3(const double* __restrict__ a, const double* __restrict__ b,
                                              long long n, double* __restrict__ partial) {
  __shared__ double red[3][4];
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] += a[i + k * n] * b[i + k * n];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    partial[blockIdx.x * 3 + k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
  }
}

__global__ void __launch_bounds__(256) k_reduce3(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  __shared__ double red[3][4];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += 256) {
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] += partial[i * 3 + k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) out[threadIdx.x] = (red[threadIdx.x][0] + red[threadIdx.x][1]) +
                                          (red[threadIdx.x][2] + red[threadIdx.x][3]);
}

struct Coef3 {
  double c[3];
};

// y_k = a_k * x_k + b_k * y_k, per block (three blocks of n).
__global__ void __launch_bounds__(256) k_axpby3(double* __restrict__ y, const double* __restrict__ x, Coef3 a,
                                                Coef3 b, long long n) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= 3 * n) return;
  const int k = (int)(i / n);
  y[i] = a.c[k] * x[i] + b.c[k] * y[i];
}

// model (3 float blocks) -> double vector, and difference with a backup: out = m - bm
__global__ void __launch_bounds__(256) k_model_diff(const float* __restrict__ m0, const float* __restrict__ m1,
                                                    const float* __restrict__ m2, const float* __restrict__ b,
                                                    double* __restrict__ out, long long n) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  out[i] = (double)m0[i] - (double)b[i];
  out[i + n] = (double)m1[i] - (double)b[i + n];
  out[i + 2 * n] = (double)m2[i] - (double)b[i + 2 * n];
}

__global__ void __launch_bounds__(256) k_backup_model(const float* __restrict__ m0, const float* __restrict__ m1,
                                                      const float* __restrict__ m2, float* __restrict__ b, long long n) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  b[i] = m0[i];
  b[i + n] = m1[i];
  b[i + 2 * n] = m2[i];
}

// Model update on rows iz > layer (parallelcalls.cu:273-286 / :343-404 with flags 1,1,1,0,0):
// m_k = base_k - alpha_k * dir_k * scale_k, then v_hor >= v_nmo.  base holds three contiguous blocks of n.
__global__ void __launch_bounds__(256) k_update3(float* __restrict__ vh, float* __restrict__ vp, float* __restrict__ vn,
                                                 const float* __restrict__ base, const double* __restrict__ dir,
                                                 Coef3 alpha, Coef3 scale, int nx, int nz) {
  const long long n = (long long)nx * nz;
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  const int iz = (int)(i / nx);
  if (iz <= kLayer) return;
  const double h = (double)base[i] - alpha.c[0] * (dir[i] * scale.c[0]);
  const double p = (double)base[i + n] - alpha.c[1] * (dir[i + n] * scale.c[1]);
  const double m = (double)base[i + 2 * n] - alpha.c[2] * (dir[i + 2 * n] * scale.c[2]);
  float hf = (float)h, nf = (float)m;
  if (hf < nf) hf = nf;
  vh[i] = hf;
  vp[i] = (float)p;
  vn[i] = nf;
}

// Physical bounds: each squared velocity clamped to [lo, hi], then v_hor >= v_nmo (the pseudo-acoustic
// system is stable for v_hor >= v_nmo >= 0 and v_po >= 0 within the time-step limit).
__global__ void __launch_bounds__(256) k_clamp3(float* __restrict__ vh, float* __restrict__ vp, float* __restrict__ vn,
                                                long long n, float lo, float hi) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  const float h = fminf(fmaxf(vh[i], lo), hi), p = fminf(fmaxf(vp[i], lo), hi), m = fminf(fmaxf(vn[i], lo), hi);
  vh[i] = h < m ? m : h;
  vp[i] = p;
  vn[i] = m;
}

// Scratch of fwi_dot3_noalloc: the 512 x 3 FP64 partial sums of k_dot3 (12 KiB), in a file-scope device array that
// every GPU the library runs on holds until the process ends.  Its host-side name is not a device address: the
// address on each device comes from hipGetSymbolAddress, taken once per device with that device current.
constexpr int kScratchDoubles = kDotBlocks * 3;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr long long kNoAllocMaxN = 1LL << 40;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, resolved on the first call on that device.
int noalloc_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess && (dev < 0 || dev >= kMaxDevices)) {
    set_message("fwi_dot3_noalloc: device index beyond the library's table of 64 devices");
    return 1;
  }
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e == hipSuccess) g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
  }
  if (e != hipSuccess) {
    set_message(std::string("fwi_dot3_noalloc: scratch address: ") + hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

}  // namespace
}  // namespace fwi

using namespace fwi;

struct fwi_lbfgs {
  int nx = 0, nz = 0, m = 5;
  long long n = 0;
  int count = 0;  // pairs stored (<= m)
  int head = 0;   // ring index of the oldest pair
  bool have_back = false;
  double* s = nullptr;      // m * 3n
  double* y = nullptr;      // m * 3n
  double* q = nullptr;      // 3n (also r)
  double* gback = nullptr;  // 3n
  float* mback = nullptr;   // 3n
  double* partial = nullptr;
  double* dots = nullptr;   // 3 device results
  std::vector<double> rho, alph;  // per pair and block
  std::string err;

  ~fwi_lbfgs() {
    (void)hipFree(s);
    (void)hipFree(y);
    (void)hipFree(q);
    (void)hipFree(gback);
    (void)hipFree(mback);
    (void)hipFree(partial);
    (void)hipFree(dots);
  }
  int fail(hipError_t e, const char* w) {
    err = std::string(w) + ": " + hipGetErrorString(e);
    set_message(err);
    return (int)e;
  }
  double* slot(double* base, int logical) const {  // logical 0 = oldest
    return base + (size_t)((head + logical) % m) * 3 * n;
  }
  unsigned blocks(long long cnt) const { return (unsigned)((cnt + 255) / 256); }

  int dot3(const double* a, const double* b, double out[3], hipStream_t st) {
    hipLaunchKernelGGL(k_dot3, dim3(kDotBlocks), dim3(256), 0, st, a, b, n, partial);
    hipLaunchKernelGGL(k_reduce3, dim3(1), dim3(256), 0, st, partial, kDotBlocks, dots);
    hipError_t e = hipMemcpyAsync(out, dots, 3 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : fail(e, "dot3");
  }
  int axpby(double* yv, const double* xv, const double a[3], const double b[3], hipStream_t st) {
    Coef3 ca{{a[0], a[1], a[2]}}, cb{{b[0], b[1], b[2]}};
    hipLaunchKernelGGL(k_axpby3, dim3(blocks(3 * n)), dim3(256), 0, st, yv, xv, ca, cb, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(e, "axpby");
  }
};

extern "C" {

int fwi_lbfgs_create(int nx, int nz, int m, fwi_lbfgs** out) {
  if (check_host("fwi_lbfgs_create", {{out, "out"}})) return 1;
  *out = nullptr;
  if (nx <= 0 || nz <= 0 || m <= 0) {
    set_message("fwi_lbfgs_create: bad sizes");
    return 1;
  }
  auto* h = new fwi_lbfgs();
  h->nx = nx;
  h->nz = nz;
  h->m = m;
  h->n = (long long)nx * nz;
  const size_t v = 3 * (size_t)h->n;
  hipError_t e = hipMalloc(&h->s, m * v * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->y, m * v * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->q, v * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->gback, v * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->mback, v * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&h->partial, kDotBlocks * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->dots, 3 * sizeof(double));
  if (e != hipSuccess) {
    int rc = h->fail(e, "fwi_lbfgs_create");
    delete h;
    return rc;
  }
  h->rho.assign((size_t)m * 3, 0.0);
  h->alph.assign((size_t)m * 3, 0.0);
  *out = h;
  return 0;
}

void fwi_lbfgs_destroy(fwi_lbfgs* h) { delete h; }

int fwi_lbfgs_pairs(const fwi_lbfgs* h) { return h->count; }

static int backup(fwi_lbfgs* h, const float* vh, const float* vp, const float* vn, const double* grad,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_backup_model, dim3(h->blocks(h->n)), dim3(256), 0, st, vh, vp, vn, h->mback, h->n);
  hipError_t e = hipMemcpyAsync(h->gback, grad, 3 * (size_t)h->n * sizeof(double), hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return h->fail(e, "backup");
  h->have_back = true;
  return 0;
}

int fwi_lbfgs_sd_step(fwi_lbfgs* h, float* vh, float* vp, float* vn, const double* grad, const double* alpha,
                      void* stream) {
  if (check_host("fwi_lbfgs_sd_step", {{h, "optimizer"}, {alpha, "alpha"}}) ||
      check_arrays("fwi_lbfgs_sd_step", {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}, {grad, "grad"}}))
    return 1;
  hipStream_t st = (hipStream_t)stream;
  if (int r = backup(h, vh, vp, vn, grad, st)) return r;
  double nn[3];
  if (int r = h->dot3(grad, grad, nn, st)) return r;
  Coef3 a{{alpha[0], alpha[1], alpha[2]}};
  Coef3 sc{{0, 0, 0}};
  for (int k = 0; k < 3; ++k) sc.c[k] = nn[k] > 0 ? 1.0 / std::sqrt(nn[k]) : 0.0;
  // base = the backup just taken: three contiguous blocks (the model fields themselves are separate arrays)
  hipLaunchKernelGGL(k_update3, dim3(h->blocks(h->n)), dim3(256), 0, st, vh, vp, vn, h->mback, grad, a, sc, h->nx,
                     h->nz);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : h->fail(e, "sd_step");
}

int fwi_lbfgs_direction(fwi_lbfgs* h, const float* vh, const float* vp, const float* vn, const double* grad,
                        void* stream) {
  if (check_host("fwi_lbfgs_direction", {{h, "optimizer"}}) ||
      check_arrays("fwi_lbfgs_direction", {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}, {grad, "grad"}}))
    return 1;
  hipStream_t st = (hipStream_t)stream;
  if (!h->have_back) {
    set_message("fwi_lbfgs_direction: no previous model/gradient (run fwi_lbfgs_sd_step first)");
    return 1;
  }
  const size_t v = 3 * (size_t)h->n;
  // push the newest pair: s = m - m_back, y = g - g_back (compute_sy, parallelcalls.cu:288-302)
  if (h->count == h->m) h->head = (h->head + 1) % h->m;
  else h->count++;
  double* sk = h->slot(h->s, h->count - 1);
  double* yk = h->slot(h->y, h->count - 1);
  hipLaunchKernelGGL(k_model_diff, dim3(h->blocks(h->n)), dim3(256), 0, st, vh, vp, vn, h->mback, sk, h->n);
  hipError_t e = hipMemcpyAsync(yk, grad, v * sizeof(double), hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return h->fail(e, "copy grad");
  {
    const double one[3] = {1, 1, 1}, mone[3] = {-1, -1, -1};
    if (int r = h->axpby(yk, h->gback, mone, one, st)) return r;
  }
  // q = g; first loop, newest to oldest (main.c:574-622)
  e = hipMemcpyAsync(h->q, grad, v * sizeof(double), hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return h->fail(e, "copy q");
  for (int l = h->count - 1; l >= 0; --l) {
    double ys[3], sq[3];
    if (int r = h->dot3(h->slot(h->y, l), h->slot(h->s, l), ys, st)) return r;
    if (int r = h->dot3(h->slot(h->s, l), h->q, sq, st)) return r;
    double ma[3], one[3] = {1, 1, 1};
    for (int k = 0; k < 3; ++k) {
      h->rho[l * 3 + k] = 1.0 / ys[k];
      h->alph[l * 3 + k] = h->rho[l * 3 + k] * sq[k];
      ma[k] = -h->alph[l * 3 + k];
    }
    if (int r = h->axpby(h->q, h->slot(h->y, l), ma, one, st)) return r;
  }
  // r = gamma q with gamma = s_k.y_k / y_k.y_k of the newest pair (main.c:624-655)
  {
    double sy[3], yy[3], g[3], zero[3] = {0, 0, 0};
    if (int r = h->dot3(sk, yk, sy, st)) return r;
    if (int r = h->dot3(yk, yk, yy, st)) return r;
    for (int k = 0; k < 3; ++k) g[k] = sy[k] / yy[k];
    if (int r = h->axpby(h->q, h->q, zero, g, st)) return r;
  }
  // second loop, oldest to newest: beta = rho y_l.r, r += (alpha_l - beta) s_l (main.c:657-682)
  for (int l = 0; l < h->count; ++l) {
    double yr[3], c[3], one[3] = {1, 1, 1};
    if (int r = h->dot3(h->slot(h->y, l), h->q, yr, st)) return r;
    for (int k = 0; k < 3; ++k) c[k] = h->alph[l * 3 + k] - h->rho[l * 3 + k] * yr[k];
    if (int r = h->axpby(h->q, h->slot(h->s, l), c, one, st)) return r;
  }
  // back up m and g (main.c:683-688)
  return backup(h, vh, vp, vn, grad, st);
}

int fwi_lbfgs_trial(fwi_lbfgs* h, float* vh, float* vp, float* vn, const double* alpha, void* stream) {
  if (check_host("fwi_lbfgs_trial", {{h, "optimizer"}, {alpha, "alpha"}}) ||
      check_arrays("fwi_lbfgs_trial", {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}}))
    return 1;
  hipStream_t st = (hipStream_t)stream;
  Coef3 a{{alpha[0], alpha[1], alpha[2]}}, one{{1, 1, 1}};
  hipLaunchKernelGGL(k_update3, dim3(h->blocks(h->n)), dim3(256), 0, st, vh, vp, vn, h->mback, h->q, a, one, h->nx,
                     h->nz);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : h->fail(e, "trial");
}

int fwi_lbfgs_restore(fwi_lbfgs* h, float* vh, float* vp, float* vn, void* stream) {
  if (check_host("fwi_lbfgs_restore", {{h, "optimizer"}}) ||
      check_arrays("fwi_lbfgs_restore", {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}}))
    return 1;
  hipStream_t st = (hipStream_t)stream;
  const size_t b = (size_t)h->n * sizeof(float);
  hipError_t e = hipMemcpyAsync(vh, h->mback, b, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(vp, h->mback + h->n, b, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(vn, h->mback + 2 * h->n, b, hipMemcpyDeviceToDevice, st);
  return e == hipSuccess ? 0 : h->fail(e, "restore");
}

int fwi_lbfgs_get_direction(fwi_lbfgs* h, double* out, void* stream) {
  if (check_host("fwi_lbfgs_get_direction", {{h, "optimizer"}}) ||
      check_arrays("fwi_lbfgs_get_direction", {{out, "out"}}))
    return 1;
  hipError_t e = hipMemcpyAsync(out, h->q, 3 * (size_t)h->n * sizeof(double), hipMemcpyDeviceToDevice,
                                (hipStream_t)stream);
  return e == hipSuccess ? 0 : h->fail(e, "get_direction");
}

// FP64 dot products of three blocks each: out[k] = sum a_k b_k (deterministic).
int fwi_dot3(const double* a, const double* b, long long n, double* out, void* stream) {
  if (check_host("fwi_dot3", {{out, "out"}}) || check_arrays("fwi_dot3", {{a, "a"}, {b, "b"}})) return 1;
  hipStream_t st = (hipStream_t)stream;
  double* partial = nullptr;
  double* d = nullptr;
  hipError_t e = hipMallocAsync((void**)&partial, kDotBlocks * 3 * sizeof(double), st);
  if (e == hipSuccess) e = hipMallocAsync((void**)&d, 3 * sizeof(double), st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_dot3, dim3(kDotBlocks), dim3(256), 0, st, a, b, n, partial);
    hipLaunchKernelGGL(k_reduce3, dim3(1), dim3(256), 0, st, partial, kDotBlocks, d);
    e = hipMemcpyAsync(out, d, 3 * sizeof(double), hipMemcpyDeviceToHost, st);
  }
  if (partial) (void)hipFreeAsync(partial, st);
  if (d) (void)hipFreeAsync(d, st);
  const hipError_t es = hipStreamSynchronize(st);
  if (e == hipSuccess) e = es;
  if (e != hipSuccess) {
    set_message(std::string("fwi_dot3: ") + hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

int fwi_dot3_noalloc(const double* a, const double* b, long long n, double* out, void* stream) {
  if (check_host("fwi_dot3_noalloc", {{a, "a"}, {b, "b"}, {out, "out"}})) return 1;
  if (n < 0 || n > kNoAllocMaxN) {
    set_message("fwi_dot3_noalloc: n must be in [0, 2^40] (the blocks are 3 x n doubles)");
    return 1;
  }
  if (check_arrays("fwi_dot3_noalloc", {{a, "a"}, {b, "b"}, {out, "out"}})) return 1;
  double* scr = nullptr;
  if (int r = noalloc_scratch(&scr)) return r;
  hipStream_t st = (hipStream_t)stream;
  // every block writes its three partial sums and the reduction writes out: nothing needs clearing
  hipLaunchKernelGGL(k_dot3, dim3(kDotBlocks), dim3(256), 0, st, a, b, n, scr);
  hipLaunchKernelGGL(k_reduce3, dim3(1), dim3(256), 0, st, scr, kDotBlocks, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_message(std::string("fwi_dot3_noalloc: ") + hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

int fwi_model_clamp(float* vh, float* vp, float* vn, long long n, float lo, float hi, void* stream) {
  if (!(lo >= 0.f) || !(hi >= lo)) {
    set_message("fwi_model_clamp: need 0 <= lo <= hi");
    return 1;
  }
  if (n <= 0) return 0;
  if (check_arrays("fwi_model_clamp", {{vh, "v_hor"}, {vp, "v_po"}, {vn, "v_nmo"}})) return 1;
  hipLaunchKernelGGL(k_clamp3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, vh, vp, vn, n, lo,
                     hi);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_message(std::string("fwi_model_clamp: ") + hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

}  // extern "C"
