// Legacy C-ABI launchers (include/fwi_serialcalls.h): the reference's 17 serialcalls functions on
// gfx950 kernels of this library.
//
// The kernels keep the reference's per-cell semantics and arithmetic (stencil sums in double, pointwise
// updates in float) so that the drop-in path reproduces the reference bit for bit on the same inputs.
// Launch shape: 64 x 4 workgroups, one wave per 64-cell row segment, so every global access is a
// coalesced row load (the reference uses 32 x 32 blocks, which splits each wave over two rows).

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fwi_common.h"
#include "fwi_internal.h"
#include "../../include/fwi_serialcalls.h"

namespace fwi {
namespace {

using S = double;  // reference arithmetic for stencil sums

constexpr int BX = 64, BY = 4;

inline dim3 grid2(int nx, int ny) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY); }

__device__ __forceinline__ bool in_omega(int ix, int iz, int nx, int nz) {
  return ix >= kRim && ix < nx - kRim && iz >= kRim && iz < nz - kRim;
}
__device__ __forceinline__ bool in_gamma(int ix, int iz, int nx, int nz, int cpml) {
  return in_omega(ix, iz, nx, nz) && (ix < cpml || ix > nx - cpml - 1 || iz > nz - cpml - 1);
}

// parallelcalls.cu:3-16
__global__ void __launch_bounds__(256) k_pre_variables(const float* vel, const float* eps, const float* del,
                                                       float* vh, float* vp, float* vn, int nx, int nz) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  long i = ix + (long)nx * iz;
  float v = vel[i] * vel[i];
  vp[i] = v;
  vh[i] = v * (1.f + 2.f * eps[i]);
  vn[i] = v * (1.f + 2.f * del[i]);
}

// parallelcalls.cu:120-130
__global__ void __launch_bounds__(256) k_second_derivate_p(const float* P, const float* F, float* p2, float* r2,
                                                           float dh2, int nx, int nz) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (!in_omega(ix, iz, nx, nz)) return;
  long i = ix + (long)nx * iz;
  p2[i] = d2_at<S>(P, i, 1, dh2);
  r2[i] = d2_at<S>(F, i, nx, dh2);
}

// parallelcalls.cu:48-63
__global__ void __launch_bounds__(256) k_psi_f(const float* P, const float* F, const float* ax, const float* bx,
                                               const float* az, const float* bz, float* psx, float* psz, int nx,
                                               int nz, float dh, int cpml) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (!in_gamma(ix, iz, nx, nz, cpml)) return;
  long i = ix + (long)nx * iz;
  float tpx = d1_at<S>(P, i, 1, dh);
  float tfz = d1_at<S>(F, i, nx, dh);
  psx[i] = bx[ix] * psx[i] + ax[ix] * tpx;
  psz[i] = bz[iz] * psz[i] + az[iz] * tfz;
}

// parallelcalls.cu:65-86
__global__ void __launch_bounds__(256) k_zeta_f(const float* p2, const float* r2, const float* ax, const float* bx,
                                                const float* az, const float* bz, const float* psx, const float* psz,
                                                float* zx, float* zz, float* atp, float* atf, int nx, int nz,
                                                float dh, int cpml) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (!in_gamma(ix, iz, nx, nz, cpml)) return;
  long i = ix + (long)nx * iz;
  float tx = d1_at<S>(psx, i, 1, dh);
  float tz = d1_at<S>(psz, i, nx, dh);
  float tpx = p2[i] + tx;
  float tfz = r2[i] + tz;
  float zxn = bx[ix] * zx[i] + ax[ix] * tpx;
  float zzn = bz[iz] * zz[i] + az[iz] * tfz;
  zx[i] = zxn;
  zz[i] = zzn;
  atp[i] = tx + zxn;
  atf[i] = tz + zzn;
}

// parallelcalls.cu:88-117.  The thread of the source cell adds the wavelet after its own update, and a
// receiver thread samples its own cell afterwards, so the per-thread order is the reference's.
__global__ void __launch_bounds__(256) k_propagator_f(const float* p2, const float* r2, float* PF, const float* Pp,
                                                      const float* Fp, float* Pm, float* Fm, const float* vh,
                                                      const float* vp, const float* vn, const float* atp,
                                                      const float* atf, const float* src, float* shot, float C,
                                                      float dh2, int nx, int nz, int sx, int sz, int it, int cpml) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  long i = ix + (long)nx * iz;
  if (in_omega(ix, iz, nx, nz)) {
    float tp = dh2 * (p2[i] + atp[i]);
    float tf = dh2 * (r2[i] + atf[i]);
    float pn = 2.f * Pp[i] - Pm[i] + vh[i] * C * tp + vp[i] * C * tf;
    float fn = 2.f * Fp[i] - Fm[i] + vn[i] * C * tp + vp[i] * C * tf;
    Pm[i] = pn;
    Fm[i] = fn;
    if (PF) PF[i] = pn + fn;
  }
  if (ix == sx - 1 && iz == sz - 1) {
    Pm[i] += src[it];
    Fm[i] += src[it];
  }
  if (ix > cpml - 1 && ix < nx - cpml && iz == sz - 1) shot[ix + (long)nx * it] = Pm[i] + Fm[i];
}

// parallelcalls.cu:133-142
__global__ void __launch_bounds__(256) k_res(float* R, const float* s1, const float* s2, int nx, int nt) {
  int ix = blockIdx.x * BX + threadIdx.x, it = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || it >= nt) return;
  long i = ix + (long)nx * it;
  R[i] = s2[i] - s1[i];
}

// parallelcalls.cu:166-177
__global__ void __launch_bounds__(256) k_entry_point(const float* Pp, const float* Fp, float* a1, float* a2,
                                                     float* a3, float* a4, const float* vh, const float* vp,
                                                     const float* vn, int nx, int nz) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  long i = ix + (long)nx * iz;
  a1[i] = vh[i] * Pp[i];
  a2[i] = vp[i] * Pp[i];
  a3[i] = vn[i] * Fp[i];
  a4[i] = vp[i] * Fp[i];
}

// parallelcalls.cu:179-198
__global__ void __launch_bounds__(256) k_psi_r(const float* a1, const float* a2, const float* a3, const float* a4,
                                               const float* ax, const float* bx, const float* az, const float* bz,
                                               float* s1, float* s2, float* s3, float* s4, int nx, int nz, float dh,
                                               int cpml) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (!in_gamma(ix, iz, nx, nz, cpml)) return;
  long i = ix + (long)nx * iz;
  float p1cx = d1_at<S>(a1, i, 1, dh);
  float p1cz = d1_at<S>(a2, i, nx, dh);
  float r1dx = d1_at<S>(a3, i, 1, dh);
  float r1cz = d1_at<S>(a4, i, nx, dh);
  s1[i] = bx[ix] * s1[i] + ax[ix] * p1cx;
  s2[i] = bz[iz] * s2[i] + az[iz] * p1cz;
  s3[i] = bx[ix] * s3[i] + ax[ix] * r1dx;
  s4[i] = bz[iz] * s4[i] + az[iz] * r1cz;
}

// parallelcalls.cu:200-214
__global__ void __launch_bounds__(256) k_second_derivate(const float* a1, const float* a2, const float* a3,
                                                         const float* a4, float* p2cx, float* p2cz, float* r2dx,
                                                         float* r2cz, float dh2, int nx, int nz) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (!in_omega(ix, iz, nx, nz)) return;
  long i = ix + (long)nx * iz;
  p2cx[i] = d2_at<S>(a1, i, 1, dh2);
  p2cz[i] = d2_at<S>(a2, i, nx, dh2);
  r2dx[i] = d2_at<S>(a3, i, 1, dh2);
  r2cz[i] = d2_at<S>(a4, i, nx, dh2);
}

// parallelcalls.cu:216-243
__global__ void __launch_bounds__(256) k_zeta_r(const float* p2cx, const float* p2cz, const float* r2dx,
                                                const float* r2cz, const float* ax, const float* bx, const float* az,
                                                const float* bz, const float* s1, const float* s2, const float* s3,
                                                const float* s4, float* z1, float* z2, float* z3, float* z4,
                                                float* t1, float* t2, float* t3, float* t4, int nx, int nz, float dh,
                                                int cpml) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (!in_gamma(ix, iz, nx, nz, cpml)) return;
  long i = ix + (long)nx * iz;
  float q1 = d1_at<S>(s1, i, 1, dh);
  float q2 = d1_at<S>(s2, i, nx, dh);
  float q3 = d1_at<S>(s3, i, 1, dh);
  float q4 = d1_at<S>(s4, i, nx, dh);
  float n1 = bx[ix] * z1[i] + ax[ix] * (q1 + p2cx[i]);
  float n2 = bz[iz] * z2[i] + az[iz] * (q2 + p2cz[i]);
  float n3 = bx[ix] * z3[i] + ax[ix] * (q3 + r2dx[i]);
  float n4 = bz[iz] * z4[i] + az[iz] * (q4 + r2cz[i]);
  z1[i] = n1;
  z2[i] = n2;
  z3[i] = n3;
  z4[i] = n4;
  t1[i] = n1 + q1;
  t2[i] = n2 + q2;
  t3[i] = n3 + q3;
  t4[i] = n4 + q4;
}

// parallelcalls.cu:144-164
__global__ void __launch_bounds__(256) k_retro(const float* Pp, const float* Fp, float* Pm, float* Fm,
                                               const float* res, float C, float dh2, int nx, int nz, int nt, int sz,
                                               int it, int cpml, const float* p2cx, const float* p2cz,
                                               const float* r2dx, const float* r2cz, const float* t1, const float* t2,
                                               const float* t3, const float* t4) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  long i = ix + (long)nx * iz;
  if (in_omega(ix, iz, nx, nz)) {
    Pm[i] = 2.f * Pp[i] - Pm[i] + dh2 * C * (p2cx[i] + t1[i] + r2dx[i] + t3[i]);
    Fm[i] = 2.f * Fp[i] - Fm[i] + dh2 * C * (p2cz[i] + r2cz[i] + t2[i] + t4[i]);
  }
  if (ix > cpml - 1 && ix < nx - cpml && iz == sz - 1) {
    float r = res[ix + (long)nx * (nt - it - 1)];
    Pm[i] += r;
    Fm[i] += r;
  }
}

// parallelcalls.cu:245-259
__global__ void __launch_bounds__(256) k_gradient(const float* Pf, const float* Ff, const float* lP, const float* lF,
                                                  float* gcx, float* gcz, float* gdx, float dh2, float dt, int nx,
                                                  int nz) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (!in_omega(ix, iz, nx, nz)) return;
  long i = ix + (long)nx * iz;
  float p2 = d2_at<S>(Pf, i, 1, dh2);
  float r2 = d2_at<S>(Ff, i, nx, dh2);
  gcx[i] += dt * p2 * lP[i];
  gcz[i] += dt * r2 * (lP[i] + lF[i]);
  gdx[i] += dt * p2 * lF[i];
}

// parallelcalls.cu:261-271
__global__ void __launch_bounds__(256) k_normal(float* gcx, float* gcz, float* gdx, int nx, int nz, float n1, float n2,
                                                float n3) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz || iz <= kLayer) return;
  long i = ix + (long)nx * iz;
  gcx[i] = gcx[i] / n1;
  gcz[i] = gcz[i] / n2;
  gdx[i] = gdx[i] / n3;
}

// parallelcalls.cu:273-286
__global__ void __launch_bounds__(256) k_update_model(const float* gcx, const float* gcz, const float* gdx, float* vh,
                                                      float* vp, float* vn, float a1, float a2, float a3, int nx,
                                                      int nz) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz || iz <= kLayer) return;
  long i = ix + (long)nx * iz;
  float p = vp[i] - (a2 * gcz[i]);
  float n = vn[i] - (a3 * gdx[i]);
  float h = vh[i] - (a1 * gcx[i]);
  if (h < n) h = n;
  vp[i] = p;
  vn[i] = n;
  vh[i] = h;
}

// parallelcalls.cu:288-302
__global__ void __launch_bounds__(256) k_compute_sy(float* s1, float* s2, float* s3, float* y1, float* y2, float* y3,
                                                    const float* m1, const float* m2, const float* m3,
                                                    const float* bm1, const float* bm2, const float* bm3,
                                                    const float* g1, const float* g2, const float* g3,
                                                    const float* bg1, const float* bg2, const float* bg3, int nx,
                                                    int nz, int k) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  long i = ix + (long)nx * iz, o = i + (long)k * nx * nz;
  s1[o] = m1[i] - bm1[i];
  s2[o] = m2[i] - bm2[i];
  s3[o] = m3[i] - bm3[i];
  y1[o] = g1[i] - bg1[i];
  y2[o] = g2[i] - bg2[i];
  y3[o] = g3[i] - bg3[i];
}

// parallelcalls.cu:304-315
__global__ void __launch_bounds__(256) k_compute_qi(float* q1, float* q2, float* q3, const float* y1, const float* y2,
                                                    const float* y3, float e1, float e2, float e3, int nx, int nz,
                                                    int k) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  long i = ix + (long)nx * iz, o = i + (long)k * nx * nz;
  q1[i] -= e1 * y1[o];
  q2[i] -= e2 * y2[o];
  q3[i] -= e3 * y3[o];
}

// parallelcalls.cu:317-328
__global__ void __launch_bounds__(256) k_compute_ri(float* r1, float* r2, float* r3, const float* q1, const float* q2,
                                                    const float* q3, float g1, float g2, float g3, int nx, int nz) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  long i = ix + (long)nx * iz;
  r1[i] = g1 * q1[i];
  r2[i] = g2 * q2[i];
  r3[i] = g3 * q3[i];
}

// parallelcalls.cu:330-341
__global__ void __launch_bounds__(256) k_update_r(float* r1, float* r2, float* r3, const float* s1, const float* s2,
                                                  const float* s3, float e1, float e2, float e3, float b1, float b2,
                                                  float b3, int nx, int nz, int k) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  long i = ix + (long)nx * iz, o = i + (long)k * nx * nz;
  r1[i] += (e1 - b1) * s1[o];
  r2[i] += (e2 - b2) * s2[o];
  r3[i] += (e3 - b3) * s3[o];
}

// parallelcalls.cu:343-404 with the clamps of include/parallelcalls.h:7-12
__global__ void __launch_bounds__(256) k_update_lbfgs(const float* r1, const float* r2, const float* r3, float* vp,
                                                      float* vn, float* vh, const float* bvp, const float* bvn,
                                                      const float* bvh, int nx, int nz, float av, float ad, float ae,
                                                      int fv, int fd, int fe, int fmin, int fmax) {
  int ix = blockIdx.x * BX + threadIdx.x, iz = blockIdx.y * BY + threadIdx.y;
  if (ix >= nx || iz >= nz || iz <= kLayer) return;
  const float min_v = 1.8e6f, max_v = 3.5e6f, min_d = 1.8e6f, max_d = 4e6f, min_e = 1.8e6f, max_e = 4.5e6f;
  long i = ix + (long)nx * iz;
  if (fv == 1) {
    float v = bvp[i] - av * r1[i];
    if (fmax == 1 && v > max_v) v = max_v;
    if (fmin == 1 && v < min_v) v = min_v;
    vp[i] = v;
  }
  if (fd == 1) {
    float v = bvn[i] - ad * r2[i];
    if (fmax == 1 && v > max_d) v = max_d;
    if (fmin == 1 && v < min_d) v = min_d;
    vn[i] = v;
  }
  if (fe == 1) {
    float v = bvh[i] - ae * r3[i];
    if (fmax == 1 && v > max_e) v = max_e;
    if (fmin == 1 && v < min_e) v = min_e;
    float n = vn[i];
    if (v < n) v = n;
    vh[i] = v;
  }
}

inline hipStream_t st(fwi_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace
}  // namespace fwi

using namespace fwi;

#define LAUNCH2(kernel, nx, ny, stream, ...)                                              \
  do {                                                                                      \
    hipLaunchKernelGGL(kernel, grid2((nx), (ny)), dim3(BX, BY), 0, (stream), __VA_ARGS__); \
    record_error(hipGetLastError());                                                        \
  } while (0)

extern "C" {

void pre_variables_h(float* vel, float* epsilon, float* delta, float* v_hor, float* v_po, float* v_nmo, int nx,
                     int nz) {
  LAUNCH2(k_pre_variables, nx, nz, 0, vel, epsilon, delta, v_hor, v_po, v_nmo, nx, nz);
}

// The C-PML profiles are evaluated once on the host in FP64 from the float parameters (the reference
// evaluates them with device exp/pow) and written into the left/right (x) or bottom (z) entries only.
static void write_cpml(float* a, float* b, int cpml, float Vmax, int n, float dt, float dh, float f, float d0,
                       float L, bool both_ends) {
  if (cpml <= 0 || n <= 0) return;
  int m = cpml < n ? cpml : n;
  std::vector<float> ta, tb;
  cpml_table_host(cpml, Vmax, dt, dh, f, d0, L, ta, tb);
  std::vector<float> ha((size_t)n), hb((size_t)n);
  // read back the current contents so cells outside the strips keep their values
  record_error(hipMemcpy(ha.data(), a, n * sizeof(float), hipMemcpyDeviceToHost));
  record_error(hipMemcpy(hb.data(), b, n * sizeof(float), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    int e = -1;
    if (both_ends && i < cpml) e = i;
    if (i > n - cpml - 1) e = n - 1 - i;
    if (e >= 0 && e < m) {
      ha[i] = ta[e];
      hb[i] = tb[e];
    }
  }
  record_error(hipMemcpy(a, ha.data(), n * sizeof(float), hipMemcpyHostToDevice));
  record_error(hipMemcpy(b, hb.data(), n * sizeof(float), hipMemcpyHostToDevice));
}

void CPML_x_h(float* a_x, float* b_x, int CPML, float Vmax, int nx, float dt, float dh, float f, float d0, float L) {
  write_cpml(a_x, b_x, CPML, Vmax, nx, dt, dh, f, d0, L, true);
}
void CPML_z_h(float* a_z, float* b_z, int CPML, float Vmax, int nz, float dt, float dh, float f, float d0, float L) {
  write_cpml(a_z, b_z, CPML, Vmax, nz, dt, dh, f, d0, L, false);
}

void PSI_F_h(float* P, float* F, float* a_x, float* b_x, float* a_z, float* b_z, float* psi_px, float* psi_fz, int nx,
             int nz, float dh, int CPML) {
  LAUNCH2(k_psi_f, nx, nz, 0, P, F, a_x, b_x, a_z, b_z, psi_px, psi_fz, nx, nz, dh, CPML);
}

void ZETA_F_h(float* p2, float* r2, float* P, float* F, float* a_x, float* b_x, float* a_z, float* b_z, float* psi_px,
              float* psi_fz, float* z_px, float* z_fz, float* aten_p, float* aten_f, int nx, int nz, float dh,
              float dh2, int CPML) {
  (void)P;
  (void)F;
  (void)dh2;
  LAUNCH2(k_zeta_f, nx, nz, 0, p2, r2, a_x, b_x, a_z, b_z, psi_px, psi_fz, z_px, z_fz, aten_p, aten_f, nx, nz, dh,
          CPML);
}

void propagator_F_h(float* p2, float* r2, float* PF, float* P_pre, float* F_pre, float* P_past, float* F_past,
                    float* v_hor, float* v_po, float* v_nmo, float* aten_p, float* aten_f, float* source, float* shot,
                    float C, float dh2, int nx, int nz, int sx, int sz, int it, int CPML) {
  LAUNCH2(k_propagator_f, nx, nz, 0, p2, r2, PF, P_pre, F_pre, P_past, F_past, v_hor, v_po, v_nmo, aten_p, aten_f,
          source, shot, C, dh2, nx, nz, sx, sz, it, CPML);
}

void res_h(float* Res, float* shot1, float* shot2, int nx, int nt) {
  LAUNCH2(k_res, nx, nt, 0, Res, shot1, shot2, nx, nt);
}

void SecondDerivate_P_h(float* P_pre, float* F_pre, float* p2, float* r2, float dh2, int nx, int nz) {
  LAUNCH2(k_second_derivate_p, nx, nz, 0, P_pre, F_pre, p2, r2, dh2, nx, nz);
}

void launchPropagation(fwi_stream_t stream1, fwi_stream_t stream2, float* PF_h, float* PF, float* P_past,
                       float* P_pre, float* F_past, float* F_pre, float* p2, float* r2, float dh2, int nx, int nz,
                       float* a_x, float* b_x, float* a_z, float* b_z, float* psi_px, float* psi_fz, float dh,
                       int CPML, float* z_px, float* z_fz, float* aten_p, float* aten_f, float* v_hor, float* v_po,
                       float* v_nmo, float* source, float* shot2, float C, int sx, int sz, int it) {
  // serialcalls.cu:60-64: snapshot of (P_past, F_past) -- contiguous, F = P + nx*nz -- on stream1, the
  // three stencil launches on stream2, the update on the null stream (which orders after both).
  record_error(hipMemcpyAsync(PF_h, P_past, 2 * (size_t)nx * nz * sizeof(float), hipMemcpyDeviceToHost, st(stream1)));
  LAUNCH2(k_second_derivate_p, nx, nz, st(stream2), P_pre, F_pre, p2, r2, dh2, nx, nz);
  LAUNCH2(k_psi_f, nx, nz, st(stream2), P_pre, F_pre, a_x, b_x, a_z, b_z, psi_px, psi_fz, nx, nz, dh, CPML);
  LAUNCH2(k_zeta_f, nx, nz, st(stream2), p2, r2, a_x, b_x, a_z, b_z, psi_px, psi_fz, z_px, z_fz, aten_p, aten_f, nx,
          nz, dh, CPML);
  LAUNCH2(k_propagator_f, nx, nz, 0, p2, r2, PF, P_pre, F_pre, P_past, F_past, v_hor, v_po, v_nmo, aten_p, aten_f,
          source, shot2, C, dh2, nx, nz, sx, sz, it, CPML);
}

void launchBackPropagation(fwi_stream_t stream1, fwi_stream_t stream2, int nx, int nz, int nt, float* PF_h,
                           float* P_pre, float* P_past, float* F_pre, float* F_past, float* aux1, float* aux2,
                           float* aux3, float* aux4, float* v_hor, float* v_po, float* v_nmo, float* p2cx,
                           float* p2cz, float* r2dx, float* r2cz, float dh2, float* P_pre2, float* F_pre2,
                           float* a_x, float* b_x, float* a_z, float* b_z, float* psixcx, float* psixcz,
                           float* psizdx, float* psizcz, float dh, int CPML, float* zitaxcx, float* zitaxcz,
                           float* zitazdx, float* zitazcz, float* atenxcx, float* atenxcz, float* atenzdx,
                           float* atenzcz, float* Res, float C, float* PF, int it, int sx, int sz, float dt,
                           float* gcx, float* gcz, float* gdx) {
  (void)PF;
  (void)sx;
  // serialcalls.cu:76-88
  record_error(
      hipMemcpyAsync(P_pre2, PF_h, 2 * (size_t)nx * nz * sizeof(float), hipMemcpyHostToDevice, st(stream1)));
  LAUNCH2(k_entry_point, nx, nz, st(stream2), P_pre, F_pre, aux1, aux2, aux3, aux4, v_hor, v_po, v_nmo, nx, nz);
  LAUNCH2(k_second_derivate, nx, nz, st(stream2), aux1, aux2, aux3, aux4, p2cx, p2cz, r2dx, r2cz, dh2, nx, nz);
  LAUNCH2(k_psi_r, nx, nz, st(stream2), aux1, aux2, aux3, aux4, a_x, b_x, a_z, b_z, psixcx, psixcz, psizdx, psizcz,
          nx, nz, dh, CPML);
  LAUNCH2(k_zeta_r, nx, nz, st(stream2), p2cx, p2cz, r2dx, r2cz, a_x, b_x, a_z, b_z, psixcx, psixcz, psizdx, psizcz,
          zitaxcx, zitaxcz, zitazdx, zitazcz, atenxcx, atenxcz, atenzdx, atenzcz, nx, nz, dh, CPML);
  LAUNCH2(k_retro, nx, nz, st(stream2), P_pre, F_pre, P_past, F_past, Res, C, dh2, nx, nz, nt, sz, it, CPML, p2cx,
          p2cz, r2dx, r2cz, atenxcx, atenxcz, atenzdx, atenzcz);
  LAUNCH2(k_gradient, nx, nz, 0, P_pre2, F_pre2, P_past, F_past, gcx, gcz, gdx, dh2, dt, nx, nz);
}

void NORMAL_h(float* gcx, float* gcz, float* gdx, int nx, int nz, float ngcx, float ngcz, float ngdx) {
  LAUNCH2(k_normal, nx, nz, 0, gcx, gcz, gdx, nx, nz, ngcx, ngcz, ngdx);
}

void UpdateModel_h(float* gcx, float* gcz, float* gdx, float* v_hor, float* v_po, float* v_nmo, float alpha1,
                   float alpha2, float alpha3, int nx, int nz) {
  LAUNCH2(k_update_model, nx, nz, 0, gcx, gcz, gdx, v_hor, v_po, v_nmo, alpha1, alpha2, alpha3, nx, nz);
}

void compute_sy_h(float* s_vp, float* s_vn, float* s_vh, float* y_b1, float* y_b2, float* y_b3, float* vp, float* vn,
                  float* vh, float* back_vp, float* back_vn, float* back_vh, float* b1, float* b2, float* b3,
                  float* back_b1, float* back_b2, float* back_b3, int nx, int nz, int k) {
  LAUNCH2(k_compute_sy, nx, nz, 0, s_vp, s_vn, s_vh, y_b1, y_b2, y_b3, vp, vn, vh, back_vp, back_vn, back_vh, b1, b2,
          b3, back_b1, back_b2, back_b3, nx, nz, k);
}

void compute_qi_h(float* q_b1, float* q_b2, float* q_b3, float* y_b1, float* y_b2, float* y_b3, float epsilon_v,
                  float epsilon_d, float epsilon_e, int nx, int nz, int k) {
  LAUNCH2(k_compute_qi, nx, nz, 0, q_b1, q_b2, q_b3, y_b1, y_b2, y_b3, epsilon_v, epsilon_d, epsilon_e, nx, nz, k);
}

void compute_ri_h(float* r_b1, float* r_b2, float* r_b3, float* q_b1, float* q_b2, float* q_b3, float gamma_v,
                  float gamma_d, float gamma_e, int nx, int nz) {
  LAUNCH2(k_compute_ri, nx, nz, 0, r_b1, r_b2, r_b3, q_b1, q_b2, q_b3, gamma_v, gamma_d, gamma_e, nx, nz);
}

void update_r_h(float* r_b1, float* r_b2, float* r_b3, float* s_vp, float* s_vn, float* s_vh, float epsilon_v,
                float epsilon_d, float epsilon_e, float beta_v, float beta_d, float beta_e, int nx, int nz, int k) {
  LAUNCH2(k_update_r, nx, nz, 0, r_b1, r_b2, r_b3, s_vp, s_vn, s_vh, epsilon_v, epsilon_d, epsilon_e, beta_v, beta_d,
          beta_e, nx, nz, k);
}

void update_LBFGS_h(float* r_b1, float* r_b2, float* r_b3, float* vp, float* vn, float* vh, float* back_vp,
                    float* back_vn, float* back_vh, int nx, int nz, float alpha_v, float alpha_d, float alpha_e,
                    int flag_v, int flag_d, int flag_e, int flag_min, int flag_max) {
  LAUNCH2(k_update_lbfgs, nx, nz, 0, r_b1, r_b2, r_b3, vp, vn, vh, back_vp, back_vn, back_vh, nx, nz, alpha_v,
          alpha_d, alpha_e, flag_v, flag_d, flag_e, flag_min, flag_max);
}

}  // extern "C"
