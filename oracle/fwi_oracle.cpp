// fwi_oracle.cpp -- CPU restatement of the reference FWI gradient path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product in fwi_amd/: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  Nothing in the product links it.
//
// The reference (danielvn/Fwi) has no CPU path: src/serialcalls.cu only launches CUDA kernels.  This
// file restates, in plain C++, the kernel bodies of src/parallelcalls.cu in the launch order of
// src/serialcalls.cu and the control flow of src/main.c.  Each function cites the lines it follows.
//
// Precision instantiations (T = storage type, S = type the finite-difference sums are evaluated in):
//   mix : T=float,  S=double -- the reference's arithmetic: its stencil coefficients are double
//         literals (parallelcalls.cu:57,127,...) so every stencil is summed in double and rounded to
//         float on store, while pointwise updates are float.
//   f32 : T=float,  S=float  -- everything in float (what the GPU FP32 kernels compute).
//   f64 : T=double, S=double -- the FP64 oracle for gradient / adjoint checks.
//   f32d: T=float,  S=F32D   -- float, second derivatives in difference form
//         sum_k c_k((u_{+k}-u_0)+(u_{-k}-u_0)) (this project's FP32 arithmetic; the weights sum to zero,
//         so it is the same operator without the cancellation of the O(|u|) terms).
//   f32i: T=float,  S=F32I   -- float, forward leapfrog in increment form: the state is (P^n, Q^n =
//         P^n - P^{n-1}) and Q^{n+1} = Q^n + d, P^{n+1} = P^n + Q^{n+1} (this project's default FP32
//         arithmetic); the same recursion as f32 in exact arithmetic.
// Build with -ffp-contract=off so the rounding of every expression is the one written here.
//
// Indexing: idx = ix + nx*iz, x fastest (include/parallelcalls.h:5).  Omega = {4<=ix<=nx-5, 4<=iz<=nz-5}
// is the updated interior; Gamma = Omega & (ix<cpml | ix>nx-cpml-1 | iz>nz-cpml-1) is the C-PML
// subset (no C-PML on the top edge, which is a free surface).
//
// Pinning (what checks each function; the tests named are in tests/):
//   d1/d2 stencils      : polynomial exactness + symbol at theta=pi      (test_oracle_pins.py)
//   cpml_table          : closed form of Komatitsch & Martin C-PML, kappa=1 (test_oracle_pins.py)
//   ricker              : closed form                                     (test_oracle_pins.py)
//   forward (all modes) : isotropic reduction (eps=delta=0 => P==F), analytic 2-D Green's function
//                         before the first reflection, linearity, time-shift invariance (bit for bit),
//                         source-receiver reciprocity (test_oracle_pins.py);
//                         f32i against f64 at 1e-5 and ahead of f32 on a long, finely sampled run
//   exact adjoint       : dot-product test of the one-step transpose, FD directional derivatives,
//                         brute-force central differences for every cell of a 24x24 grid
//   ref-compat gradient : with the C-PML off its adjoint fields equal the exact adjoint's (2e-14, FP64); its
//                         gradient equals dt sum_k lam^k D2(u^(k-2)) assembled in numpy from the recorded fields
//                         (a lag of 0 or 2 or a dt^2 scale miss by > 5e-2); with the C-PML on, a step-by-step
//                         composition of the pinned forward routines on aux1..4 reproduces it bit for bit
//   L-BFGS two-loop     : dense BFGS inverse update with H0 = gamma*I (Nocedal & Wright Alg. 7.4)
// The reference's own rounding (nvcc FMA contraction, device exp/pow in CPML_x/CPML_z) cannot be
// reproduced without running it and is "parity unpinned" at the last-ulp level.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <type_traits>

namespace {

constexpr int kRim = 4;      // stencil radius: the 4-cell rim is never updated
constexpr int kLayer = 15;   // include/parallelcalls.h:6 -- rows iz <= layer are never updated
constexpr double kPi = 3.141592653589793;   // include/parallelcalls.h:3

struct Geo {
  int nx, nz, cpml;
  inline bool omega(int ix, int iz) const {
    return ix >= kRim && ix < nx - kRim && iz >= kRim && iz < nz - kRim;
  }
  inline bool gamma(int ix, int iz) const {
    return omega(ix, iz) && (ix < cpml || ix > nx - cpml - 1 || iz > nz - cpml - 1);
  }
};

struct F32D {};  // tag: float arithmetic, difference-form second derivative
struct F32I {};  // tag: float arithmetic, forward leapfrog in increment form (state P^n, Q^n = P^n - P^{n-1})
template <class S>
struct SumType {
  using type = S;
};
template <>
struct SumType<F32D> {
  using type = float;
};
template <>
struct SumType<F32I> {
  using type = float;
};

// 8th-order central first/second derivative.  Reference: the D1 expression of parallelcalls.cu:57-58
// and the D2 expression of :127-128.  Pair sums/differences are formed in T (float operands), the
// weighted sum in S, the result is rounded to T -- the order of operations is kept.
template <class T, class S0>
inline T d2(const T* u, long i, long st, T h2) {
  if constexpr (std::is_same<S0, F32D>::value) {
    const float c1 = (float)(8.0 / 5.0), c2 = (float)(1.0 / 5.0), c3 = (float)(8.0 / 315.0), c4 = (float)(1.0 / 560.0);
    const float c = u[i];
    const float r = c1 * ((u[i + st] - c) + (u[i - st] - c)) - c2 * ((u[i + 2 * st] - c) + (u[i - 2 * st] - c)) +
                    c3 * ((u[i + 3 * st] - c) + (u[i - 3 * st] - c)) - c4 * ((u[i + 4 * st] - c) + (u[i - 4 * st] - c));
    return r / h2;
  }
  using S = typename SumType<S0>::type;
  const S c0 = (S)(205.0 / 72.0), c1 = (S)(8.0 / 5.0), c2 = (S)(1.0 / 5.0), c3 = (S)(8.0 / 315.0),
          c4 = (S)(1.0 / 560.0);
  S r = -c0 * (S)u[i] + c1 * (S)(T)(u[i + st] + u[i - st]) - c2 * (S)(T)(u[i + 2 * st] + u[i - 2 * st]) +
        c3 * (S)(T)(u[i + 3 * st] + u[i - 3 * st]) - c4 * (S)(T)(u[i + 4 * st] + u[i - 4 * st]);
  return (T)(r / (S)h2);
}
template <class T, class S0>
inline T d1(const T* u, long i, long st, T h) {
  using S = typename SumType<S0>::type;
  const S c1 = (S)(4.0 / 5.0), c2 = (S)(1.0 / 5.0), c3 = (S)(4.0 / 105.0), c4 = (S)(1.0 / 280.0);
  S r = c1 * (S)(T)(u[i + st] - u[i - st]) - c2 * (S)(T)(u[i + 2 * st] - u[i - 2 * st]) +
        c3 * (S)(T)(u[i + 3 * st] - u[i - 3 * st]) - c4 * (S)(T)(u[i + 4 * st] - u[i - 4 * st]);
  return (T)(r / (S)h);
}

// ---------------------------------------------------------------------------------------------
// Restated kernels.  One loop nest per reference __global__ kernel; same predicates, same writes.
// ---------------------------------------------------------------------------------------------

// parallelcalls.cu:3-16
template <class T>
void k_pre_variables(const T* vel, const T* eps, const T* del, T* vh, T* vp, T* vn, int nx, int nz) {
  long n = (long)nx * nz;
#pragma omp parallel for
  for (long i = 0; i < n; ++i) {
    vp[i] = vel[i] * vel[i];
    vh[i] = vp[i] * ((T)1 + (T)2 * eps[i]);
    vn[i] = vp[i] * ((T)1 + (T)2 * del[i]);
  }
}

// parallelcalls.cu:120-130
template <class T, class S>
void k_second_derivate_p(const T* P, const T* F, T* p2, T* r2, T dh2, const Geo& g) {
#pragma omp parallel for
  for (int iz = kRim; iz < g.nz - kRim; ++iz)
    for (int ix = kRim; ix < g.nx - kRim; ++ix) {
      long i = ix + (long)g.nx * iz;
      p2[i] = d2<T, S>(P, i, 1, dh2);
      r2[i] = d2<T, S>(F, i, g.nx, dh2);
    }
}

// parallelcalls.cu:48-63
template <class T, class S>
void k_psi_f(const T* P, const T* F, const T* ax, const T* bx, const T* az, const T* bz, T* psx, T* psz,
             T dh, const Geo& g) {
#pragma omp parallel for
  for (int iz = 0; iz < g.nz; ++iz)
    for (int ix = 0; ix < g.nx; ++ix) {
      if (!g.gamma(ix, iz)) continue;
      long i = ix + (long)g.nx * iz;
      T tpx = d1<T, S>(P, i, 1, dh);
      T tfz = d1<T, S>(F, i, g.nx, dh);
      psx[i] = bx[ix] * psx[i] + ax[ix] * tpx;
      psz[i] = bz[iz] * psz[i] + az[iz] * tfz;
    }
}

// parallelcalls.cu:65-86
template <class T, class S>
void k_zeta_f(const T* p2, const T* r2, const T* ax, const T* bx, const T* az, const T* bz, const T* psx,
              const T* psz, T* zx, T* zz, T* atp, T* atf, T dh, const Geo& g) {
#pragma omp parallel for
  for (int iz = 0; iz < g.nz; ++iz)
    for (int ix = 0; ix < g.nx; ++ix) {
      if (!g.gamma(ix, iz)) continue;
      long i = ix + (long)g.nx * iz;
      T tx = d1<T, S>(psx, i, 1, dh);
      T tz = d1<T, S>(psz, i, g.nx, dh);
      T tpx = p2[i] + tx;
      T tfz = r2[i] + tz;
      zx[i] = bx[ix] * zx[i] + ax[ix] * tpx;
      zz[i] = bz[iz] * zz[i] + az[iz] * tfz;
      atp[i] = tx + zx[i];
      atf[i] = tz + zz[i];
    }
}

// parallelcalls.cu:88-117.  Source and receivers follow the update of the same cell, as in the
// reference (one thread owns the cell).  Receivers: row sz-1, columns cpml..nx-cpml-1.
template <class T>
void k_propagator_f(const T* p2, const T* r2, T* PF, const T* Pp, const T* Fp, T* Pm, T* Fm, const T* vh,
                    const T* vp, const T* vn, const T* atp, const T* atf, const T* src, T* shot, T C, T dh2,
                    int sx, int sz, int it, const Geo& g) {
  const int nx = g.nx, nz = g.nz;
#pragma omp parallel for
  for (int iz = 0; iz < nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      long i = ix + (long)nx * iz;
      if (g.omega(ix, iz)) {
        T tp = dh2 * (p2[i] + atp[i]);
        T tf = dh2 * (r2[i] + atf[i]);
        Pm[i] = (T)2 * Pp[i] - Pm[i] + vh[i] * C * tp + vp[i] * C * tf;
        Fm[i] = (T)2 * Fp[i] - Fm[i] + vn[i] * C * tp + vp[i] * C * tf;
        if (PF) PF[i] = Pm[i] + Fm[i];
      }
      if (ix == sx - 1 && iz == sz - 1) {
        Pm[i] += src[it];
        Fm[i] += src[it];
      }
      if (ix > g.cpml - 1 && ix < nx - g.cpml && iz == sz - 1) shot[ix + (long)nx * it] = Pm[i] + Fm[i];
    }
}

// parallelcalls.cu:133-142
template <class T>
void k_res(T* R, const T* s1, const T* s2, long n) {
#pragma omp parallel for
  for (long i = 0; i < n; ++i) R[i] = s2[i] - s1[i];
}

// parallelcalls.cu:166-177
template <class T>
void k_entry_point(const T* Pp, const T* Fp, T* a1, T* a2, T* a3, T* a4, const T* vh, const T* vp,
                   const T* vn, long n) {
#pragma omp parallel for
  for (long i = 0; i < n; ++i) {
    a1[i] = vh[i] * Pp[i];
    a2[i] = vp[i] * Pp[i];
    a3[i] = vn[i] * Fp[i];
    a4[i] = vp[i] * Fp[i];
  }
}

// parallelcalls.cu:200-214
template <class T, class S>
void k_second_derivate(const T* a1, const T* a2, const T* a3, const T* a4, T* p2cx, T* p2cz, T* r2dx,
                       T* r2cz, T dh2, const Geo& g) {
#pragma omp parallel for
  for (int iz = kRim; iz < g.nz - kRim; ++iz)
    for (int ix = kRim; ix < g.nx - kRim; ++ix) {
      long i = ix + (long)g.nx * iz;
      p2cx[i] = d2<T, S>(a1, i, 1, dh2);
      p2cz[i] = d2<T, S>(a2, i, g.nx, dh2);
      r2dx[i] = d2<T, S>(a3, i, 1, dh2);
      r2cz[i] = d2<T, S>(a4, i, g.nx, dh2);
    }
}

// parallelcalls.cu:179-198
template <class T, class S>
void k_psi_r(const T* a1, const T* a2, const T* a3, const T* a4, const T* ax, const T* bx, const T* az,
             const T* bz, T* s1, T* s2, T* s3, T* s4, T dh, const Geo& g) {
#pragma omp parallel for
  for (int iz = 0; iz < g.nz; ++iz)
    for (int ix = 0; ix < g.nx; ++ix) {
      if (!g.gamma(ix, iz)) continue;
      long i = ix + (long)g.nx * iz;
      T p1cx = d1<T, S>(a1, i, 1, dh);
      T p1cz = d1<T, S>(a2, i, g.nx, dh);
      T r1dx = d1<T, S>(a3, i, 1, dh);
      T r1cz = d1<T, S>(a4, i, g.nx, dh);
      s1[i] = bx[ix] * s1[i] + ax[ix] * p1cx;
      s2[i] = bz[iz] * s2[i] + az[iz] * p1cz;
      s3[i] = bx[ix] * s3[i] + ax[ix] * r1dx;
      s4[i] = bz[iz] * s4[i] + az[iz] * r1cz;
    }
}

// parallelcalls.cu:216-243
template <class T, class S>
void k_zeta_r(const T* p2cx, const T* p2cz, const T* r2dx, const T* r2cz, const T* ax, const T* bx,
              const T* az, const T* bz, const T* s1, const T* s2, const T* s3, const T* s4, T* z1, T* z2,
              T* z3, T* z4, T* t1, T* t2, T* t3, T* t4, T dh, const Geo& g) {
#pragma omp parallel for
  for (int iz = 0; iz < g.nz; ++iz)
    for (int ix = 0; ix < g.nx; ++ix) {
      if (!g.gamma(ix, iz)) continue;
      long i = ix + (long)g.nx * iz;
      T q1 = d1<T, S>(s1, i, 1, dh);
      T q2 = d1<T, S>(s2, i, g.nx, dh);
      T q3 = d1<T, S>(s3, i, 1, dh);
      T q4 = d1<T, S>(s4, i, g.nx, dh);
      z1[i] = bx[ix] * z1[i] + ax[ix] * (q1 + p2cx[i]);
      z2[i] = bz[iz] * z2[i] + az[iz] * (q2 + p2cz[i]);
      z3[i] = bx[ix] * z3[i] + ax[ix] * (q3 + r2dx[i]);
      z4[i] = bz[iz] * z4[i] + az[iz] * (q4 + r2cz[i]);
      t1[i] = z1[i] + q1;
      t2[i] = z2[i] + q2;
      t3[i] = z3[i] + q3;
      t4[i] = z4[i] + q4;
    }
}

// parallelcalls.cu:144-164
template <class T>
void k_retro_propagation(const T* Pp, const T* Fp, T* Pm, T* Fm, const T* res, T C, T dh2, int nt, int sz,
                         int it, const T* p2cx, const T* p2cz, const T* r2dx, const T* r2cz, const T* t1,
                         const T* t2, const T* t3, const T* t4, const Geo& g) {
  const int nx = g.nx, nz = g.nz;
#pragma omp parallel for
  for (int iz = 0; iz < nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      long i = ix + (long)nx * iz;
      if (g.omega(ix, iz)) {
        Pm[i] = (T)2 * Pp[i] - Pm[i] + dh2 * C * (p2cx[i] + t1[i] + r2dx[i] + t3[i]);
        Fm[i] = (T)2 * Fp[i] - Fm[i] + dh2 * C * (p2cz[i] + r2cz[i] + t2[i] + t4[i]);
      }
      if (ix > g.cpml - 1 && ix < nx - g.cpml && iz == sz - 1) {
        Pm[i] += res[ix + (long)nx * (nt - it - 1)];
        Fm[i] += res[ix + (long)nx * (nt - it - 1)];
      }
    }
}

// parallelcalls.cu:245-259
template <class T, class S>
void k_gradient(const T* Pf, const T* Ff, const T* lP, const T* lF, T* gcx, T* gcz, T* gdx, T dh2, T dt,
                const Geo& g) {
#pragma omp parallel for
  for (int iz = kRim; iz < g.nz - kRim; ++iz)
    for (int ix = kRim; ix < g.nx - kRim; ++ix) {
      long i = ix + (long)g.nx * iz;
      T p2 = d2<T, S>(Pf, i, 1, dh2);
      T r2 = d2<T, S>(Ff, i, g.nx, dh2);
      gcx[i] += dt * p2 * lP[i];
      gcz[i] += dt * r2 * (lP[i] + lF[i]);
      gdx[i] += dt * p2 * lF[i];
    }
}

// parallelcalls.cu:261-271
template <class T>
void k_normal(T* gcx, T* gcz, T* gdx, int nx, int nz, T n1, T n2, T n3) {
  for (int iz = kLayer + 1; iz < nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      long i = ix + (long)nx * iz;
      gcx[i] = gcx[i] / n1;
      gcz[i] = gcz[i] / n2;
      gdx[i] = gdx[i] / n3;
    }
}

// parallelcalls.cu:273-286
template <class T>
void k_update_model(const T* gcx, const T* gcz, const T* gdx, T* vh, T* vp, T* vn, T a1, T a2, T a3, int nx,
                    int nz) {
  for (int iz = kLayer + 1; iz < nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      long i = ix + (long)nx * iz;
      vp[i] = vp[i] - (a2 * gcz[i]);
      vn[i] = vn[i] - (a3 * gdx[i]);
      vh[i] = vh[i] - (a1 * gcx[i]);
      if (vh[i] < vn[i]) vh[i] = vn[i];
    }
}

// ---------------------------------------------------------------------------------------------
// Shot-level compositions.
// ---------------------------------------------------------------------------------------------

template <class T>
struct Fields {
  long n;
  std::vector<T> buf;
  explicit Fields(long n_, int count) : n(n_), buf((size_t)n_ * count, (T)0) {}
  T* operator[](int k) { return buf.data() + (size_t)k * n; }
};

// Geometry of one shot in the generalised form used by the oracle: source at 0-based (isx, isz),
// receivers on row irz, columns [ir0, ir0+nrec).  The reference's case (propagator_F, :107-116) is
// isx=sx-1, isz=irz=sz-1, ir0=cpml, nrec=nx-2*cpml.  Gathers are [it][irec].
struct ShotGeo {
  int isx, isz, irz, ir0, nrec;
};

// One forward time step in the reference's order of launchPropagation (serialcalls.cu:55-66):
// SecondDerivate_P, PSI_F, ZETA_F, propagator_F.  Receivers are written to gather row `it`.
// Buffers: P (P^n), Pm (P^{n-1}, overwritten with P^{n+1}); likewise F.
// Rows of the FP64 near-surface band of the f32i forward (see band_step): 0 = the library's default (24 rows when
// nz - cpml - 4 >= 24, none otherwise), -1 = none, k > 4 = rows 4..k-1.  g_band_force >= -1 overrides the setting
// inside a composition that runs without the band (16-bit checkpoints).
int g_band_rows = 0;
thread_local int g_band_force = -2;
struct BandForce {  // within its scope the band setting is forced to v (>= -1), or left as it is (v = -2)
  int old;
  explicit BandForce(int v) : old(g_band_force) {
    if (v >= -1) g_band_force = v;
  }
  ~BandForce() { g_band_force = old; }
};
inline int band_rows(int nz, int cpml) {
  const int k = g_band_force >= -1 ? g_band_force : g_band_rows;
  if (k < 0) return 0;
  if (k == 0) return nz - cpml - 4 >= 24 ? 24 : 0;
  return k;
}

template <class T, class S>
struct FwdState {
  Geo g;
  long n;
  // 0 P_pre 1 P_past 2 F_pre 3 F_past 4 p2 5 r2 6 psx 7 psz 8 zx 9 zz 10 atp 11 atf 12 Q 13 R
  // (Q = P^n - P^{n-1}, R = F^n - F^{n-1}: increment form only)
  Fields<T> f;
  std::vector<T> shotbuf;
  int cur = 0;  // parity of the ping-pong
  // f32i near-surface band: rows [kRim, band) in FP64, two levels of P and F by parity plus psi_x, zeta_x
  int band = 0;
  double dtd = 0;  // the float time step, promoted (set by the composition that runs the steps)
  std::vector<double> bP[2], bF[2], bpsx, bzx, bFz;
  FwdState(const Geo& g_) : g(g_), n((long)g_.nx * g_.nz), f(n, 14) {
    if constexpr (std::is_same<S, F32I>::value) band = band_rows(g.nz, g.cpml);
    if (band)
      for (auto* v : {&bP[0], &bP[1], &bF[0], &bF[1], &bpsx, &bzx, &bFz}) v->assign((size_t)n, 0.0);
  }
  T* Pcur() { return cur == 0 ? f[0] : f[1]; }
  T* Pold() { return cur == 0 ? f[1] : f[0]; }
  T* Fcur() { return cur == 0 ? f[2] : f[3]; }
  T* Fold() { return cur == 0 ? f[3] : f[2]; }
};

// FP64 near-surface band of the f32i forward (this project's arithmetic, not the reference's): the rows
// [kRim, band), where the source, the receivers and the near-surface gradient lie, advance in FP64 with the
// two-level leapfrog of the f64 arithmetic (C-PML on the side columns; the band lies above the bottom C-PML, so
// its z terms vanish), from the float model, tables, wavelet, dh and dt promoted exactly.  F couples the band and
// the float rows through the z stencil: the band reads rows band..band+3 of the float F^n, and the float rows
// read rows band-4..band-1 of the float arrays, which hold the band's values rounded.  The band writes its rows
// of the float P, F (rounded), the imaging terms (rounded) and the receivers in it (P + F rounded once).
template <class T, class S>
void band_step(FwdState<T, S>& st, const T* vh, const T* vp, const T* vn, const T* ax, const T* bx, const T* src,
               T C, T dh, const ShotGeo& sg, int it, T* gather) {
  (void)C;
  const Geo& g = st.g;
  const int nx = g.nx, B = st.band, cpml = g.cpml;
  const double dhd = (double)dh, dh2d = dhd * dhd;
  const double dtd = st.dtd, Cd = (dtd * dtd) / dh2d;
  const int c = st.cur;
  const double* P = st.bP[c].data();
  double* Pm = st.bP[c ^ 1].data();
  const double* Fb = st.bF[c].data();
  double* Fm = st.bF[c ^ 1].data();
  double* Fz = st.bFz.data();  // F^n over rows 0..band+3: the band's, then the float rows
  const T* Ff = st.Fcur();
  for (int iz = 0; iz < B + 4 && iz < g.nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      long i = ix + (long)nx * iz;
      Fz[i] = iz < B ? Fb[i] : (double)Ff[i];
    }
  auto side = [&](int ix) { return ix < cpml || ix > nx - cpml - 1; };
  // psi_x on the side C-PML cells (PSI_F)
  for (int iz = kRim; iz < B; ++iz)
    for (int ix = kRim; ix < nx - kRim; ++ix) {
      if (!side(ix)) continue;
      long i = ix + (long)nx * iz;
      double tpx = d1<double, double>(P, i, 1, dhd);
      st.bpsx[i] = (double)bx[ix] * st.bpsx[i] + (double)ax[ix] * tpx;
    }
  T* Pf = st.Pold();
  T* Ffo = st.Fold();
  const long is = sg.isx + (long)nx * sg.isz;
  for (int iz = kRim; iz < B; ++iz)
    for (int ix = kRim; ix < nx - kRim; ++ix) {
      long i = ix + (long)nx * iz;
      double p2 = d2<double, double>(P, i, 1, dh2d);
      double r2 = d2<double, double>(Fz, i, nx, dh2d);
      double atp = 0.0, atf = 0.0;
      if (side(ix)) {
        double tx = d1<double, double>(st.bpsx.data(), i, 1, dhd);
        double tpx = p2 + tx;
        st.bzx[i] = (double)bx[ix] * st.bzx[i] + (double)ax[ix] * tpx;
        atp = tx + st.bzx[i];
      }
      double tp = dh2d * (p2 + atp);
      double tf = dh2d * (r2 + atf);
      double pn = 2.0 * P[i] - Pm[i] + (double)vh[i] * Cd * tp + (double)vp[i] * Cd * tf;
      double fn = 2.0 * Fb[i] - Fm[i] + (double)vn[i] * Cd * tp + (double)vp[i] * Cd * tf;
      if (i == is) {
        pn += (double)src[it];
        fn += (double)src[it];
      }
      Pm[i] = pn;
      Fm[i] = fn;
      Pf[i] = (T)pn;
      Ffo[i] = (T)fn;
      // the imaging terms the caller forms as f[4] + f[10], f[5] + f[11]: the band's, rounded once
      st.f[4][i] = (T)(p2 + atp);
      st.f[10][i] = (T)0;
      st.f[5][i] = (T)(r2 + atf);
      st.f[11][i] = (T)0;
    }
  if (gather && sg.irz < B)
    for (int r = 0; r < sg.nrec; ++r) {
      long i = sg.ir0 + r + (long)nx * sg.irz;
      gather[(long)it * sg.nrec + r] = (T)(Pm[i] + Fm[i]);
    }
}

template <class T, class S>
void fwd_step(FwdState<T, S>& st, const T* vh, const T* vp, const T* vn, const T* ax, const T* bx, const T* az,
              const T* bz, const T* src, T C, T dh, T dh2, const ShotGeo& sg, int it, T* gather, int nt) {
  const Geo& g = st.g;
  T* P = st.Pcur();
  T* Pm = st.Pold();
  T* F = st.Fcur();
  T* Fm = st.Fold();
  k_second_derivate_p<T, S>(P, F, st.f[4], st.f[5], dh2, g);
  k_psi_f<T, S>(P, F, ax, bx, az, bz, st.f[6], st.f[7], dh, g);
  k_zeta_f<T, S>(st.f[4], st.f[5], ax, bx, az, bz, st.f[6], st.f[7], st.f[8], st.f[9], st.f[10], st.f[11], dh, g);
  // propagator_F body with the generalised source/receiver geometry
  const int nx = g.nx, nz = g.nz;
  const T* p2 = st.f[4];
  const T* r2 = st.f[5];
  const T* atp = st.f[10];
  const T* atf = st.f[11];
  long is = sg.isx + (long)nx * sg.isz;
  if constexpr (std::is_same<S, F32I>::value) {
    // increment form (the batched kernel's FP32I): Q += d (+ source), P^{n+1} = P^n + Q; the source
    // cell is interior (the batched API validates it).  Rows above st.band are the FP64 band's.
    if (st.band) band_step<T, S>(st, vh, vp, vn, ax, bx, src, C, dh, sg, it, gather);
    T* Q = st.f[12];
    T* R = st.f[13];
#pragma omp parallel for
    for (int iz = std::max(kRim, st.band); iz < nz - kRim; ++iz)
      for (int ix = kRim; ix < nx - kRim; ++ix) {
        long i = ix + (long)nx * iz;
        T tp = dh2 * (p2[i] + atp[i]);
        T tf = dh2 * (r2[i] + atf[i]);
        T dq = vh[i] * C * tp + vp[i] * C * tf;
        T dr = vn[i] * C * tp + vp[i] * C * tf;
        if (i == is) {
          dq += src[it];
          dr += src[it];
        }
        Q[i] = Q[i] + dq;
        R[i] = R[i] + dr;
        Pm[i] = P[i] + Q[i];
        Fm[i] = F[i] + R[i];
      }
  } else {
#pragma omp parallel for
    for (int iz = kRim; iz < nz - kRim; ++iz)
      for (int ix = kRim; ix < nx - kRim; ++ix) {
        long i = ix + (long)nx * iz;
        T tp = dh2 * (p2[i] + atp[i]);
        T tf = dh2 * (r2[i] + atf[i]);
        Pm[i] = (T)2 * P[i] - Pm[i] + vh[i] * C * tp + vp[i] * C * tf;
        Fm[i] = (T)2 * F[i] - Fm[i] + vn[i] * C * tp + vp[i] * C * tf;
      }
    Pm[is] += src[it];
    Fm[is] += src[it];
  }
  if (gather && sg.irz >= st.band)  // receivers in the band: written by band_step
    for (int r = 0; r < sg.nrec; ++r) {
      long i = sg.ir0 + r + (long)nx * sg.irz;
      gather[(long)it * sg.nrec + r] = Pm[i] + Fm[i];
    }
  (void)nt;
  st.cur ^= 1;
}

}  // namespace

// =============================================================================================
// Exported C ABI (loaded by tests/ through ctypes)
// =============================================================================================
extern "C" {

// Ricker wavelet, main.c:328-333.  t accumulates in float exactly as the reference's `t += dt`.
void orc_ricker(float* src, int nt, float f, float dt) {
  float t = 0;
  for (int it = 0; it < nt; ++it) {
    t += dt;
    double a = kPi * f * (double)(t - (1 / f));
    src[it] = (float)(1 * (1 - 2 * (a * a)) * std::exp(-(a * a)));
  }
}
void orc_ricker_f64(double* src, int nt, double f, double dt) {
  for (int it = 0; it < nt; ++it) {
    double t = (it + 1) * dt;
    double a = kPi * f * (t - 1 / f);
    src[it] = (1 - 2 * a * a) * std::exp(-a * a);
  }
}

// C-PML coefficients (parallelcalls.cu:18-46; params main.c:48-52) evaluated in FP64 on the host.
// Entry e = distance of the cell from the grid edge (0 = outermost), e < cpml; depth l = (cpml-e)*dh.
//   d = d0*vmax*(l/L)^2, alpha = pi*f*(L-l)/L, b = exp(-(d+alpha)*dt), a = d*(b-1)/(d+alpha).
// As in the reference, a is formed from b already rounded to float when round_b != 0.
void orc_cpml_table(double* a, double* b, int cpml, double vmax, double dt, double dh, double f, double d0,
                    double L, int round_b) {
  for (int e = 0; e < cpml; ++e) {
    double l = (double)(cpml - e) * dh;
    double d = d0 * vmax * (l / L) * (l / L);
    double al = kPi * f * (L - l) / L;
    double bb = std::exp(-(d + al) * dt);
    if (round_b) bb = (double)(float)bb;
    b[e] = bb;
    a[e] = d * (bb - 1) / (d + al);
  }
}

// ---- per-kernel restatements on float storage; mode 0 = reference mixed precision, 1 = fp32 ----
#define ORC_DISPATCH(mode, call_mix, call_f32) \
  do {                                         \
    if ((mode) == 0) {                         \
      call_mix;                                \
    } else {                                   \
      call_f32;                                \
    }                                          \
  } while (0)

void orc_pre_variables(const float* vel, const float* eps, const float* del, float* vh, float* vp, float* vn,
                       int nx, int nz) {
  k_pre_variables<float>(vel, eps, del, vh, vp, vn, nx, nz);
}
void orc_pre_variables_f64(const double* vel, const double* eps, const double* del, double* vh, double* vp,
                           double* vn, int nx, int nz) {
  k_pre_variables<double>(vel, eps, del, vh, vp, vn, nx, nz);
}

void orc_SecondDerivate_P(int mode, const float* P, const float* F, float* p2, float* r2, float dh2, int nx, int nz) {
  Geo g{nx, nz, 0};
  ORC_DISPATCH(mode, (k_second_derivate_p<float, double>(P, F, p2, r2, dh2, g)),
               (k_second_derivate_p<float, float>(P, F, p2, r2, dh2, g)));
}
void orc_PSI_F(int mode, const float* P, const float* F, const float* ax, const float* bx, const float* az,
               const float* bz, float* psx, float* psz, int nx, int nz, float dh, int cpml) {
  Geo g{nx, nz, cpml};
  ORC_DISPATCH(mode, (k_psi_f<float, double>(P, F, ax, bx, az, bz, psx, psz, dh, g)),
               (k_psi_f<float, float>(P, F, ax, bx, az, bz, psx, psz, dh, g)));
}
void orc_ZETA_F(int mode, const float* p2, const float* r2, const float* ax, const float* bx, const float* az,
                const float* bz, const float* psx, const float* psz, float* zx, float* zz, float* atp, float* atf,
                int nx, int nz, float dh, int cpml) {
  Geo g{nx, nz, cpml};
  ORC_DISPATCH(mode, (k_zeta_f<float, double>(p2, r2, ax, bx, az, bz, psx, psz, zx, zz, atp, atf, dh, g)),
               (k_zeta_f<float, float>(p2, r2, ax, bx, az, bz, psx, psz, zx, zz, atp, atf, dh, g)));
}
void orc_propagator_F(const float* p2, const float* r2, float* PF, const float* Pp, const float* Fp, float* Pm,
                      float* Fm, const float* vh, const float* vp, const float* vn, const float* atp,
                      const float* atf, const float* src, float* shot, float C, float dh2, int nx, int nz, int sx,
                      int sz, int it, int cpml) {
  Geo g{nx, nz, cpml};
  k_propagator_f<float>(p2, r2, PF, Pp, Fp, Pm, Fm, vh, vp, vn, atp, atf, src, shot, C, dh2, sx, sz, it, g);
}
void orc_res(float* R, const float* s1, const float* s2, int nx, int nt) { k_res<float>(R, s1, s2, (long)nx * nt); }
void orc_entryPoint(const float* Pp, const float* Fp, float* a1, float* a2, float* a3, float* a4, const float* vh,
                    const float* vp, const float* vn, int nx, int nz) {
  k_entry_point<float>(Pp, Fp, a1, a2, a3, a4, vh, vp, vn, (long)nx * nz);
}
void orc_secondDerivate(int mode, const float* a1, const float* a2, const float* a3, const float* a4, float* p2cx,
                        float* p2cz, float* r2dx, float* r2cz, float dh2, int nx, int nz) {
  Geo g{nx, nz, 0};
  ORC_DISPATCH(mode, (k_second_derivate<float, double>(a1, a2, a3, a4, p2cx, p2cz, r2dx, r2cz, dh2, g)),
               (k_second_derivate<float, float>(a1, a2, a3, a4, p2cx, p2cz, r2dx, r2cz, dh2, g)));
}
void orc_PSI_R(int mode, const float* a1, const float* a2, const float* a3, const float* a4, const float* ax,
               const float* bx, const float* az, const float* bz, float* s1, float* s2, float* s3, float* s4, int nx,
               int nz, float dh, int cpml) {
  Geo g{nx, nz, cpml};
  ORC_DISPATCH(mode, (k_psi_r<float, double>(a1, a2, a3, a4, ax, bx, az, bz, s1, s2, s3, s4, dh, g)),
               (k_psi_r<float, float>(a1, a2, a3, a4, ax, bx, az, bz, s1, s2, s3, s4, dh, g)));
}
void orc_ZETA_R(int mode, const float* p2cx, const float* p2cz, const float* r2dx, const float* r2cz, const float* ax,
                const float* bx, const float* az, const float* bz, const float* s1, const float* s2, const float* s3,
                const float* s4, float* z1, float* z2, float* z3, float* z4, float* t1, float* t2, float* t3,
                float* t4, int nx, int nz, float dh, int cpml) {
  Geo g{nx, nz, cpml};
  ORC_DISPATCH(mode,
               (k_zeta_r<float, double>(p2cx, p2cz, r2dx, r2cz, ax, bx, az, bz, s1, s2, s3, s4, z1, z2, z3, z4, t1,
                                        t2, t3, t4, dh, g)),
               (k_zeta_r<float, float>(p2cx, p2cz, r2dx, r2cz, ax, bx, az, bz, s1, s2, s3, s4, z1, z2, z3, z4, t1,
                                       t2, t3, t4, dh, g)));
}
void orc_retroPropagation(const float* Pp, const float* Fp, float* Pm, float* Fm, const float* res, float C,
                          float dh2, int nx, int nz, int nt, int sz, int it, int cpml, const float* p2cx,
                          const float* p2cz, const float* r2dx, const float* r2cz, const float* t1, const float* t2,
                          const float* t3, const float* t4) {
  Geo g{nx, nz, cpml};
  k_retro_propagation<float>(Pp, Fp, Pm, Fm, res, C, dh2, nt, sz, it, p2cx, p2cz, r2dx, r2cz, t1, t2, t3, t4, g);
}
void orc_GRADIENT(int mode, const float* Pf, const float* Ff, const float* lP, const float* lF, float* gcx,
                  float* gcz, float* gdx, float dh2, float dt, int nx, int nz) {
  Geo g{nx, nz, 0};
  ORC_DISPATCH(mode, (k_gradient<float, double>(Pf, Ff, lP, lF, gcx, gcz, gdx, dh2, dt, g)),
               (k_gradient<float, float>(Pf, Ff, lP, lF, gcx, gcz, gdx, dh2, dt, g)));
}
void orc_NORMAL(float* gcx, float* gcz, float* gdx, int nx, int nz, float n1, float n2, float n3) {
  k_normal<float>(gcx, gcz, gdx, nx, nz, n1, n2, n3);
}
void orc_UpdateModel(const float* gcx, const float* gcz, const float* gdx, float* vh, float* vp, float* vn, float a1,
                     float a2, float a3, int nx, int nz) {
  k_update_model<float>(gcx, gcz, gdx, vh, vp, vn, a1, a2, a3, nx, nz);
}

// parallelcalls.cu:288-341 (L-BFGS vector kernels); history slot k of m slots, each nx*nz
void orc_compute_sy(float* s1, float* s2, float* s3, float* y1, float* y2, float* y3, const float* m1,
                    const float* m2, const float* m3, const float* bm1, const float* bm2, const float* bm3,
                    const float* g1, const float* g2, const float* g3, const float* bg1, const float* bg2,
                    const float* bg3, int nx, int nz, int k) {
  long n = (long)nx * nz, o = (long)k * n;
  for (long i = 0; i < n; ++i) {
    s1[o + i] = m1[i] - bm1[i];
    s2[o + i] = m2[i] - bm2[i];
    s3[o + i] = m3[i] - bm3[i];
    y1[o + i] = g1[i] - bg1[i];
    y2[o + i] = g2[i] - bg2[i];
    y3[o + i] = g3[i] - bg3[i];
  }
}
void orc_compute_qi(float* q1, float* q2, float* q3, const float* y1, const float* y2, const float* y3, float e1,
                    float e2, float e3, int nx, int nz, int k) {
  long n = (long)nx * nz, o = (long)k * n;
  for (long i = 0; i < n; ++i) {
    q1[i] -= e1 * y1[o + i];
    q2[i] -= e2 * y2[o + i];
    q3[i] -= e3 * y3[o + i];
  }
}
void orc_compute_ri(float* r1, float* r2, float* r3, const float* q1, const float* q2, const float* q3, float g1,
                    float g2, float g3, int nx, int nz) {
  long n = (long)nx * nz;
  for (long i = 0; i < n; ++i) {
    r1[i] = g1 * q1[i];
    r2[i] = g2 * q2[i];
    r3[i] = g3 * q3[i];
  }
}
void orc_update_r(float* r1, float* r2, float* r3, const float* s1, const float* s2, const float* s3, float e1,
                  float e2, float e3, float b1, float b2, float b3, int nx, int nz, int k) {
  long n = (long)nx * nz, o = (long)k * n;
  for (long i = 0; i < n; ++i) {
    r1[i] += (e1 - b1) * s1[o + i];
    r2[i] += (e2 - b2) * s2[o + i];
    r3[i] += (e3 - b3) * s3[o + i];
  }
}
// parallelcalls.cu:343-404 with the clamp constants of include/parallelcalls.h:7-12
void orc_update_LBFGS(const float* r1, const float* r2, const float* r3, float* vp, float* vn, float* vh,
                      const float* bvp, const float* bvn, const float* bvh, int nx, int nz, float av, float ad,
                      float ae, int fv, int fd, int fe, int fmin, int fmax) {
  const float min_v = 1.8e6f, max_v = 3.5e6f, min_d = 1.8e6f, max_d = 4e6f, min_e = 1.8e6f, max_e = 4.5e6f;
  for (int iz = kLayer + 1; iz < nz; ++iz)
    for (int ix = 0; ix < nx; ++ix) {
      long i = ix + (long)nx * iz;
      if (fv == 1) {
        vp[i] = bvp[i] - av * r1[i];
        if (fmax == 1 && vp[i] > max_v) vp[i] = max_v;
        if (fmin == 1 && vp[i] < min_v) vp[i] = min_v;
      }
      if (fd == 1) {
        vn[i] = bvn[i] - ad * r2[i];
        if (fmax == 1 && vn[i] > max_d) vn[i] = max_d;
        if (fmin == 1 && vn[i] < min_d) vn[i] = min_d;
      }
      if (fe == 1) {
        vh[i] = bvh[i] - ae * r3[i];
        if (fmax == 1 && vh[i] > max_e) vh[i] = max_e;
        if (fmin == 1 && vh[i] < min_e) vh[i] = min_e;
        if (vh[i] < vn[i]) vh[i] = vn[i];
      }
    }
}

}  // extern "C"

// =============================================================================================
// Shot-level paths: forward gather, reference-compatible gradient, exact discrete gradient.
// =============================================================================================
namespace {

template <class T>
struct Tables {
  std::vector<T> ax, bx, az, bz;
  // a/b over x: left strip ix<cpml uses entry ix, right strip ix>=nx-cpml uses entry nx-1-ix;
  // over z only the bottom strip iz>=nz-cpml (entry nz-1-iz); the top has none (main.c:276-279).
  Tables(const T* ta, const T* tb, int cpml, int nx, int nz)
      : ax((size_t)nx, (T)0), bx((size_t)nx, (T)0), az((size_t)nz, (T)0), bz((size_t)nz, (T)0) {
    for (int ix = 0; ix < nx; ++ix) {
      int e = -1;
      if (ix < cpml) e = ix;
      else if (ix > nx - cpml - 1) e = nx - 1 - ix;
      if (e >= 0) { ax[ix] = ta[e]; bx[ix] = tb[e]; }
    }
    for (int iz = 0; iz < nz; ++iz) {
      if (iz > nz - cpml - 1) { az[iz] = ta[nz - 1 - iz]; bz[iz] = tb[nz - 1 - iz]; }
    }
  }
};

struct Scal {
  double dh, dt;
};

// Forward over all nt steps; optionally stores per-step quantities q (2 fields per step):
//   qmode 1: a^n = p2 + aten_p, b^n = r2 + aten_f  (the terms multiplying the model; exact gradient)
//   qmode 2: PF_h snapshot of (P_past, F_past) taken before the update (serialcalls.cu:60), i.e. u^{n-1}
template <class T, class S>
void forward_shot(const T* vh, const T* vp, const T* vn, const Tables<T>& tb, const T* src, const Geo& g,
                  const ShotGeo& sg, int nt, Scal sc, T* gather, int qmode, std::vector<T>* q, T* rec = nullptr) {
  FwdState<T, S> st(g);
  const T dh = (T)sc.dh, dt = (T)sc.dt, dh2 = dh * dh, C = (dt * dt) / dh2;
  st.dtd = (double)dt;
  const long n = st.n;
  if (q) q->assign((size_t)nt * 2 * n, (T)0);
  // rec (optional): the state P^n, F^n that step n starts from, n = 0..nt (2 fields per n)
  auto record = [&](int k) {
    if (!rec) return;
    std::memcpy(rec + (size_t)k * 2 * n, st.Pcur(), n * sizeof(T));
    std::memcpy(rec + (size_t)k * 2 * n + n, st.Fcur(), n * sizeof(T));
  };
  for (int it = 0; it < nt; ++it) {
    record(it);
    if (qmode == 2) {
      std::memcpy(q->data() + (size_t)it * 2 * n, st.Pold(), n * sizeof(T));
      std::memcpy(q->data() + (size_t)it * 2 * n + n, st.Fold(), n * sizeof(T));
    }
    fwd_step<T, S>(st, vh, vp, vn, tb.ax.data(), tb.bx.data(), tb.az.data(), tb.bz.data(), src, C, dh, dh2, sg, it,
                   gather, nt);
    if (it == nt - 1) record(nt);
    if (qmode == 1) {
      T* qa = q->data() + (size_t)it * 2 * n;
      T* qb = qa + n;
      const T* p2 = st.f[4];
      const T* r2 = st.f[5];
      const T* atp = st.f[10];
      const T* atf = st.f[11];
      for (int iz = kRim; iz < g.nz - kRim; ++iz)
        for (int ix = kRim; ix < g.nx - kRim; ++ix) {
          long i = ix + (long)g.nx * iz;
          qa[i] = p2[i] + atp[i];
          qb[i] = r2[i] + atf[i];
        }
    }
  }
}

// Reference-compatible adjoint + imaging (main.c:457-492 driving serialcalls.cu:68-90): adjoint state
// zeroed, then per step entryPoint, secondDerivate, PSI_R, ZETA_R, retroPropagation (residual injected
// time-reversed at the receivers), GRADIENT with the snapshot PF_h[nt-it-1].
template <class T, class S>
void adjoint_ref_shot(const T* vh, const T* vp, const T* vn, const Tables<T>& tb, const Geo& g, const ShotGeo& sg,
                      int nt, Scal sc, const T* res_g, const std::vector<T>& snaps, T* gcx, T* gcz, T* gdx,
                      T* rec = nullptr) {
  const long n = (long)g.nx * g.nz;
  const T dh = (T)sc.dh, dt = (T)sc.dt, dh2 = dh * dh, C = (dt * dt) / dh2;
  Fields<T> f(n, 4 + 4 + 4 + 4 + 4 + 4);
  T *lam0 = f[0], *lam1 = f[1];  // P_pre/P_past buffers (F in f[2], f[3])
  T *Fl0 = f[2], *Fl1 = f[3];
  T *a1 = f[4], *a2 = f[5], *a3 = f[6], *a4 = f[7];
  T *p2cx = f[8], *p2cz = f[9], *r2dx = f[10], *r2cz = f[11];
  T *s1 = f[12], *s2 = f[13], *s3 = f[14], *s4 = f[15];
  T *z1 = f[16], *z2 = f[17], *z3 = f[18], *z4 = f[19];
  T *t1 = f[20], *t2 = f[21], *t3 = f[22], *t4 = f[23];
  // residual in the reference's nx-wide layout
  std::vector<T> resw((size_t)g.nx * nt, (T)0);
  for (int it = 0; it < nt; ++it)
    for (int r = 0; r < sg.nrec; ++r) resw[(size_t)it * g.nx + sg.ir0 + r] = res_g[(size_t)it * sg.nrec + r];
  T *Pp = lam0, *Pm = lam1, *Fp = Fl0, *Fm = Fl1;
  for (int it = 0; it < nt; ++it) {
    const T* Pf = snaps.data() + (size_t)(nt - it - 1) * 2 * n;
    const T* Ff = Pf + n;
    k_entry_point<T>(Pp, Fp, a1, a2, a3, a4, vh, vp, vn, n);
    k_second_derivate<T, S>(a1, a2, a3, a4, p2cx, p2cz, r2dx, r2cz, dh2, g);
    k_psi_r<T, S>(a1, a2, a3, a4, tb.ax.data(), tb.bx.data(), tb.az.data(), tb.bz.data(), s1, s2, s3, s4, dh, g);
    k_zeta_r<T, S>(p2cx, p2cz, r2dx, r2cz, tb.ax.data(), tb.bx.data(), tb.az.data(), tb.bz.data(), s1, s2, s3, s4,
                   z1, z2, z3, z4, t1, t2, t3, t4, dh, g);
    // retroPropagation with generalised receivers
#pragma omp parallel for
    for (int iz = kRim; iz < g.nz - kRim; ++iz)
      for (int ix = kRim; ix < g.nx - kRim; ++ix) {
        long i = ix + (long)g.nx * iz;
        Pm[i] = (T)2 * Pp[i] - Pm[i] + dh2 * C * (p2cx[i] + t1[i] + r2dx[i] + t3[i]);
        Fm[i] = (T)2 * Fp[i] - Fm[i] + dh2 * C * (p2cz[i] + r2cz[i] + t2[i] + t4[i]);
      }
    for (int r = 0; r < sg.nrec; ++r) {
      long i = sg.ir0 + r + (long)g.nx * sg.irz;
      T v = resw[(size_t)(nt - it - 1) * g.nx + sg.ir0 + r];
      Pm[i] += v;
      Fm[i] += v;
    }
    if (rec) {  // (lambda_P, lambda_F) after step it, stored at k - 1 with k = nt - it (App. A.4: mu^k)
      std::memcpy(rec + (size_t)(nt - it - 1) * 2 * n, Pm, n * sizeof(T));
      std::memcpy(rec + (size_t)(nt - it - 1) * 2 * n + n, Fm, n * sizeof(T));
    }
    k_gradient<T, S>(Pf, Ff, Pm, Fm, gcx, gcz, gdx, dh2, dt, g);
    std::swap(Pp, Pm);
    std::swap(Fp, Fm);
  }
}

// ---------------------------------------------------------------------------------------------
// Exact discrete adjoint of the forward step (this project's derivation; not in the reference).
//
// With chi_x = Omega & x-strip, chi_z = Omega & bottom strip, the forward step is
//   psx' = chi_x (bx psx + ax D1x P)        tx = Gamma D1x psx'      gx = D2x P + tx
//   zx'  = chi_x (bx zx + ax gx)            a  = gx + zx'            (z-part alike with F)
//   P' = 2P - P_old + dt^2 (vh a + vp b),   F' = 2F - F_old + dt^2 (vn a + vp b)   on Omega
// Its transpose, given the adjoints (P'bar, F'bar) of the new state and the memory adjoints
// (psxbar_in, zxbar_in) arriving from the next step:
//   abar = dt^2 (vh P'bar + vn F'bar)       bbar = dt^2 vp (P'bar + F'bar)
//   zt   = chi_x (zxbar_in + abar)          gxbar = abar + ax zt     zxbar_out = bx zt
//   pt   = chi_x (psxbar_in - D1x(Gamma gxbar))                      psxbar_out = bx pt
//   Pbar = 2 P'bar + D2x gxbar - D1x(ax pt) + Pshiftbar,  Pbar_old = -P'bar       (alike for F)
// using D2x^T = D2x and D1x^T = -D1x on zero-extended fields.
// ---------------------------------------------------------------------------------------------
template <class T, class S>
struct AdjOp {
  const Geo g;
  const Tables<T>& tb;
  const T *vh, *vp, *vn;
  T dh, dh2, dt2;
  long n;
  std::vector<T> abar, bbar, gx, gz, wx, wz, px, pz, ap, af;
  AdjOp(const Geo& g_, const Tables<T>& tb_, const T* vh_, const T* vp_, const T* vn_, T dh_, T dt_)
      : g(g_), tb(tb_), vh(vh_), vp(vp_), vn(vn_), dh(dh_), dh2(dh_ * dh_), dt2(dt_ * dt_),
        n((long)g_.nx * g_.nz) {
    for (auto* v : {&abar, &bbar, &gx, &gz, &wx, &wz, &px, &pz, &ap, &af}) v->assign((size_t)n, (T)0);
  }
  inline bool chix(int ix, int iz) const { return g.omega(ix, iz) && (ix < g.cpml || ix > g.nx - g.cpml - 1); }
  inline bool chiz(int ix, int iz) const { return g.omega(ix, iz) && iz > g.nz - g.cpml - 1; }

  // Generic transposed step.  In: adjoints of the new state (Pn, Fn = adjoint of P^{n+1},F^{n+1}; Ps, Fs =
  // adjoint of the shifted copy of P^n,F^n) and of the new memory (sx, zx, sz, zz, overwritten in place by
  // the adjoints of the old memory).  Out: PbarN, FbarN (adjoint of P^n, F^n), PbarO, FbarO (of P^{n-1}).
  void apply(const T* Pn, const T* Fn, const T* Ps, const T* Fs, T* sx, T* zx, T* sz, T* zz, T* PbarN, T* FbarN,
             T* PbarO, T* FbarO) {
    terms(Pn, Fn, sx, zx, sz, zz);
    const int nx = g.nx, nz = g.nz;
#pragma omp parallel for
    for (int iz = 0; iz < nz; ++iz)
      for (int ix = 0; ix < nx; ++ix) {
        long i = ix + (long)nx * iz;
        if (!g.omega(ix, iz)) {
          PbarN[i] = FbarN[i] = PbarO[i] = FbarO[i] = (T)0;
          continue;
        }
        PbarN[i] = (T)2 * Pn[i] + ap[i] + Ps[i];
        FbarN[i] = (T)2 * Fn[i] + af[i] + Fs[i];
        PbarO[i] = -Pn[i];
        FbarO[i] = -Fn[i];
      }
  }
  // ap = D2x gbar_x - D1x(a_x pt_x), af likewise in z, on Omega (the new-state terms of the transposed step);
  // updates the memory adjoints sx, zx, sz, zz in place.
  void terms(const T* Pn, const T* Fn, T* sx, T* zx, T* sz, T* zz) {
    const int nx = g.nx, nz = g.nz;
    const T* ax = tb.ax.data();
    const T* bx = tb.bx.data();
    const T* az = tb.az.data();
    const T* bz = tb.bz.data();
#pragma omp parallel for
    for (int iz = 0; iz < nz; ++iz)
      for (int ix = 0; ix < nx; ++ix) {
        long i = ix + (long)nx * iz;
        if (!g.omega(ix, iz)) {
          abar[i] = bbar[i] = gx[i] = gz[i] = wx[i] = wz[i] = (T)0;
          continue;
        }
        T ab = dt2 * (vh[i] * Pn[i] + vn[i] * Fn[i]);
        T bb = dt2 * (vp[i] * (Pn[i] + Fn[i]));
        abar[i] = ab;
        bbar[i] = bb;
        T ztx = chix(ix, iz) ? zx[i] + ab : (T)0;
        T ztz = chiz(ix, iz) ? zz[i] + bb : (T)0;
        gx[i] = ab + ax[ix] * ztx;
        gz[i] = bb + az[iz] * ztz;
        zx[i] = bx[ix] * ztx;
        zz[i] = bz[iz] * ztz;
        bool gm = g.gamma(ix, iz);
        wx[i] = gm ? gx[i] : (T)0;
        wz[i] = gm ? gz[i] : (T)0;
      }
#pragma omp parallel for
    for (int iz = 0; iz < nz; ++iz)
      for (int ix = 0; ix < nx; ++ix) {
        long i = ix + (long)nx * iz;
        T ptx = (T)0, ptz = (T)0;
        if (chix(ix, iz)) ptx = sx[i] - d1<T, S>(wx.data(), i, 1, dh);
        if (chiz(ix, iz)) ptz = sz[i] - d1<T, S>(wz.data(), i, nx, dh);
        sx[i] = bx[ix] * ptx;
        sz[i] = bz[iz] * ptz;
        px[i] = ax[ix] * ptx;
        pz[i] = az[iz] * ptz;
      }
#pragma omp parallel for
    for (int iz = 0; iz < nz; ++iz)
      for (int ix = 0; ix < nx; ++ix) {
        long i = ix + (long)nx * iz;
        if (!g.omega(ix, iz)) {
          ap[i] = af[i] = (T)0;
          continue;
        }
        ap[i] = d2<T, S>(gx.data(), i, 1, dh2) - d1<T, S>(px.data(), i, 1, dh);
        af[i] = d2<T, S>(gz.data(), i, nx, dh2) - d1<T, S>(pz.data(), i, nx, dh);
      }
  }
};

// One exact adjoint step: mu^k from mu^{k+1} with the residual res (one time sample of the gather) injected at
// the receivers; on return muP, muF hold mu^k.  Two-level form: shP, shF hold -mu^{k+2}, the adjoint of the
// shifted copy.  f32i: the rows of the forward's FP64 band ([kRim, band), band_rows) take the increment form
// instead, with D = mu^{k+1} - mu^{k+2} in DP, DF updated to mu^k - mu^{k+1} (the residual, injected on the
// band's receiver row, goes into D); near the receivers the two-level rounding is what limits the gradient.
template <class T, class S>
void adjoint_step_exact(AdjOp<T, S>& op, const Geo& g, const ShotGeo& sg, const T* res, T*& muP, T*& muF, T*& shP,
                        T*& shF, T* sx, T* zx, T* sz, T* zz, T*& nP, T*& nF, T*& oP, T*& oF, T* DP = nullptr,
                        T* DF = nullptr) {
  const int B = std::is_same<S, F32I>::value ? band_rows(g.nz, g.cpml) : 0;
  if (B > 0) {
    op.terms(muP, muF, sx, zx, sz, zz);
#pragma omp parallel for
    for (int iz = 0; iz < g.nz; ++iz)
      for (int ix = 0; ix < g.nx; ++ix) {
        long i = ix + (long)g.nx * iz;
        if (!g.omega(ix, iz)) {
          nP[i] = nF[i] = oP[i] = oF[i] = (T)0;
          continue;
        }
        if (iz < B) {
          DP[i] = DP[i] + op.ap[i];
          DF[i] = DF[i] + op.af[i];
        } else {
          nP[i] = (T)2 * muP[i] + op.ap[i] + shP[i];
          nF[i] = (T)2 * muF[i] + op.af[i] + shF[i];
        }
        oP[i] = -muP[i];
        oF[i] = -muF[i];
      }
    for (int r = 0; r < sg.nrec; ++r) {
      long i = sg.ir0 + r + (long)g.nx * sg.irz;
      T* tP = sg.irz < B ? DP : nP;
      T* tF = sg.irz < B ? DF : nF;
      tP[i] += res[r];
      tF[i] += res[r];
    }
#pragma omp parallel for
    for (int iz = kRim; iz < B; ++iz)
      for (int ix = kRim; ix < g.nx - kRim; ++ix) {
        long i = ix + (long)g.nx * iz;
        nP[i] = muP[i] + DP[i];
        nF[i] = muF[i] + DF[i];
      }
    std::swap(muP, nP);
    std::swap(muF, nF);
    std::swap(shP, oP);
    std::swap(shF, oF);
  } else {
    op.apply(muP, muF, shP, shF, sx, zx, sz, zz, nP, nF, oP, oF);
    for (int r = 0; r < sg.nrec; ++r) {
      long i = sg.ir0 + r + (long)g.nx * sg.irz;
      nP[i] += res[r];
      nF[i] += res[r];
    }
    std::swap(muP, nP);
    std::swap(muF, nF);
    std::swap(shP, oP);
    std::swap(shF, oF);
  }
}

// Exact gradient of J = 1/2 sum (d - obs)^2 with respect to (v_hor, v_po, v_nmo):
//   g_hor += dt^2 mu_P^{k} a^{k-1},  g_nmo += dt^2 mu_F^{k} a^{k-1},  g_po += dt^2 (mu_P^{k}+mu_F^{k}) b^{k-1}
// where mu^k is the adjoint of u^k (residual res^{k-1} injected at the receivers into P and F).
template <class T, class S>
void adjoint_exact_shot(const T* vh, const T* vp, const T* vn, const Tables<T>& tb, const Geo& g, const ShotGeo& sg,
                        int nt, Scal sc, const T* res_g, const std::vector<T>& q, T* gh, T* gp, T* gn,
                        T* rec = nullptr) {
  const long n = (long)g.nx * g.nz;
  const T dh = (T)sc.dh, dt = (T)sc.dt, dt2 = dt * dt;
  AdjOp<T, S> op(g, tb, vh, vp, vn, dh, dt);
  Fields<T> f(n, 14);
  T *muP = f[0], *muF = f[1], *shP = f[2], *shF = f[3];  // adjoint of u^{k+1}; shifted adjoint (= -mu^{k+2})
  T *sx = f[4], *zx = f[5], *sz = f[6], *zz = f[7];
  T *nP = f[8], *nF = f[9], *oP = f[10], *oF = f[11];
  T *DP = f[12], *DF = f[13];  // f32i band rows: mu^{k+1} - mu^{k+2}
  for (int k = nt; k >= 1; --k) {
    // mu^k = transpose-step(mu^{k+1}) + R^T res^{k-1}; mu^{nt+1} = 0
    adjoint_step_exact<T, S>(op, g, sg, res_g + (size_t)(k - 1) * sg.nrec, muP, muF, shP, shF, sx, zx, sz, zz, nP, nF,
                             oP, oF, DP, DF);
    if (rec) {  // (mu_P^k, mu_F^k) at k - 1
      std::memcpy(rec + (size_t)(k - 1) * 2 * n, muP, n * sizeof(T));
      std::memcpy(rec + (size_t)(k - 1) * 2 * n + n, muF, n * sizeof(T));
    }
    const T* qa = q.data() + (size_t)(k - 1) * 2 * n;
    const T* qb = qa + n;
#pragma omp parallel for
    for (int iz = kRim; iz < g.nz - kRim; ++iz)
      for (int ix = kRim; ix < g.nx - kRim; ++ix) {
        long i = ix + (long)g.nx * iz;
        gh[i] += dt2 * muP[i] * qa[i];
        gn[i] += dt2 * muF[i] * qa[i];
        gp[i] += dt2 * (muP[i] + muF[i]) * qb[i];
      }
  }
}

template <class T, class S>
double shot_gradient(int mode_exact, const T* vh, const T* vp, const T* vn, const T* tab_a, const T* tab_b, const T* src,
                     const T* obs, int nx, int nz, int nt, int cpml, double dh, double dt, int isx, int isz, int irz,
                     int ir0, int nrec, T* syn, T* g_hor, T* g_po, T* g_nmo, double* misfit_ref32,
                     T* fwd_rec = nullptr, T* adj_rec = nullptr) {
  Geo g{nx, nz, cpml};
  ShotGeo sg{isx, isz, irz, ir0, nrec};
  Tables<T> tb(tab_a, tab_b, cpml, nx, nz);
  Scal sc{dh, dt};
  std::vector<T> q;
  forward_shot<T, S>(vh, vp, vn, tb, src, g, sg, nt, sc, syn, mode_exact ? 1 : 2, &q, fwd_rec);
  std::vector<T> res((size_t)nt * nrec);
  double phi = 0;
  float phi32 = 0;
  for (size_t i = 0; i < res.size(); ++i) {
    res[i] = syn[i] - obs[i];
    phi += 0.5 * (double)res[i] * (double)res[i];
  }
  // the reference's host misfit: float sequential sum of powf(res,2) in nx-wide row order, halved
  // (main.c:448-454); the zero C-PML columns contribute nothing.
  for (int it = 0; it < nt; ++it)
    for (int r = 0; r < nrec; ++r) phi32 += powf((float)res[(size_t)it * nrec + r], 2);
  phi32 = phi32 / 2;
  if (misfit_ref32) *misfit_ref32 = phi32;
  if (mode_exact)
    adjoint_exact_shot<T, S>(vh, vp, vn, tb, g, sg, nt, sc, res.data(), q, g_hor, g_po, g_nmo, adj_rec);
  else  // reference arrays: gcx <-> v_hor, gcz <-> v_po, gdx <-> v_nmo (main.c:573)
    adjoint_ref_shot<T, S>(vh, vp, vn, tb, g, sg, nt, sc, res.data(), q, g_hor, g_po, g_nmo, adj_rec);
  return phi;
}

// BF16 checkpoints (the batched path's FWI_CKPT_BF16): each checkpoint field is stored as
// bf16(x * 2^-e), 2^e <= max|x| < 2^(e+1) over the field (e = 0 if it is all zero), rounded to nearest even
// on the top 16 bits, and restored as bf16 * 2^e.
inline int ck_exponent(float amax) {
  uint32_t b;
  std::memcpy(&b, &amax, 4);
  const int be = (int)((b >> 23) & 0xFFu);
  return b == 0u ? 0 : (be == 0 ? -126 : be - 127);
}
inline float ck_round_bf16(float x, int e) {
  float v = std::ldexp(x, -e);
  uint32_t u;
  std::memcpy(&u, &v, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  std::memcpy(&v, &u, 4);
  return std::ldexp(v, e);
}
// FP16 round to nearest even of a float with |v| < 2^15: on the FP16 grid (ulp 2^(floor(log2|v|)-10) for
// |v| >= 2^-14, 2^-24 below), computed exactly in double; nearbyint rounds ties to even.
inline float ck_round_fp16(float x, int e) {
  const double v = std::ldexp((double)x, 14 - e);
  if (v == 0.0) return std::ldexp(0.0f * x, 0);
  int ex;
  std::frexp(v, &ex);  // |v| in [2^(ex-1), 2^ex)
  const int ue = ex - 1 >= -14 ? ex - 11 : -24;
  const double r = std::nearbyint(std::ldexp(v, -ue));
  return (float)std::ldexp(std::ldexp(r, ue), e - 14);
}
inline void ck_round_field(float* f, long n, int fmt) {
  float m = 0.f;
  for (long i = 0; i < n; ++i) m = std::max(m, std::fabs(f[i]));
  const int e = ck_exponent(m);
  for (long i = 0; i < n; ++i) f[i] = fmt == 2 ? ck_round_fp16(f[i], e) : ck_round_bf16(f[i], e);
}

// Forward of one shot storing the imaging terms of every step (qmode 1) as the batched path ends up using
// them with checkpoints every K steps: the last segment's terms come from the forward itself, every other
// segment's from a recomputation that starts at its checkpoint (segment 0 at the zero state).  fmt 0: FP32
// checkpoints, which reproduce the forward exactly; fmt 1 / 2: BF16 / FP16 checkpoints of the
// increment-form state (P, F, Q, R and the four C-PML memory variables), so segments 1..J-2 start from
// rounded states.
template <class T, class S>
void forward_shot_ck(const T* vh, const T* vp, const T* vn, const Tables<T>& tb, const T* src, const Geo& g,
                     const ShotGeo& sg, int nt, Scal sc, T* gather, std::vector<T>* q, int K, int fmt) {
  // 16-bit checkpoints run without the FP64 band (the library refuses the combination)
  BandForce force(fmt ? -1 : -2);
  forward_shot<T, S>(vh, vp, vn, tb, src, g, sg, nt, sc, gather, 1, q);
  if (fmt == 0) return;
  static_assert(std::is_same<T, float>::value, "BF16 checkpoints of the float state");
  const T dh = (T)sc.dh, dt = (T)sc.dt, dh2 = dh * dh, C = (dt * dt) / dh2;
  const long n = (long)g.nx * g.nz;
  const int J = (nt + K - 1) / K;
  FwdState<T, S> st(g);
  std::vector<std::vector<T>> snap((size_t)J);
  auto fields = [](FwdState<T, S>& s) {
    return std::vector<T*>{s.Pcur(), s.Fcur(), s.f[12], s.f[13], s.f[6], s.f[7], s.f[8], s.f[9]};
  };
  for (int it = 0; it < nt; ++it) {
    if (it > 0 && it % K == 0 && it / K <= J - 2) {
      std::vector<T>& c = snap[(size_t)(it / K)];
      c.resize((size_t)8 * n);
      auto fs = fields(st);
      for (int k = 0; k < 8; ++k) {
        std::memcpy(c.data() + (size_t)k * n, fs[k], n * sizeof(T));
        ck_round_field(c.data() + (size_t)k * n, n, fmt);
      }
    }
    fwd_step<T, S>(st, vh, vp, vn, tb.ax.data(), tb.bx.data(), tb.az.data(), tb.bz.data(), src, C, dh, dh2, sg, it,
                   nullptr, nt);
  }
  for (int seg = 1; seg <= J - 2; ++seg) {
    FwdState<T, S> r(g);
    r.cur = (seg * K) & 1;
    auto fs = fields(r);
    for (int k = 0; k < 8; ++k) std::memcpy(fs[k], snap[(size_t)seg].data() + (size_t)k * n, n * sizeof(T));
    for (int nn = seg * K; nn < std::min((seg + 1) * K, nt); ++nn) {
      fwd_step<T, S>(r, vh, vp, vn, tb.ax.data(), tb.bx.data(), tb.az.data(), tb.bz.data(), src, C, dh, dh2, sg, nn,
                     nullptr, nt);
      T* qa = q->data() + (size_t)nn * 2 * n;
      T* qb = qa + n;
      for (int iz = kRim; iz < g.nz - kRim; ++iz)
        for (int ix = kRim; ix < g.nx - kRim; ++ix) {
          long i = ix + (long)g.nx * iz;
          qa[i] = r.f[4][i] + r.f[10][i];
          qb[i] = r.f[5][i] + r.f[11][i];
        }
    }
  }
}

// Exact gradient of a gradient group of the batched path: up to four shots, each with its own window
// (x0, nx) of the full model.  Every shot runs its own forward and exact adjoint (shot_gradient,
// adjoint_exact_shot); per step and cell of the union of the windows the shots' imaging terms c_s (zero where
// a window's interior does not reach) are added in member order, ((c_0 + c_1) + c_2) + c_3, before one update
// of the group's accumulators g (union width x nz).  With one shot this is adjoint_exact_shot's update.
// geo: (x0, nx, isx, isz, irz, ir0, nrec) per member; obs/syn: member-major gathers.
template <class T, class S>
void group_gradient(const T* vh_m, const T* vp_m, const T* vn_m, const T* tab_a, const T* tab_b, const T* src, int nxm,
                    int nz, int nt, int cpml, double dh_, double dt_, int ns, const int* geo, const T* obs, T* syn,
                    T* g_hor, T* g_po, T* g_nmo, double* misfits, int K = 0, int ckfmt = 0) {
  Scal sc{dh_, dt_};
  const T dh = (T)dh_, dt = (T)dt_, dt2 = dt * dt;
  struct Member {
    int x0, nx;
    ShotGeo sg;
    long n;
    std::vector<T> vh, vp, vn, q, res;
    Tables<T>* tb;
    AdjOp<T, S>* op;
    Fields<T>* f;
    T *muP, *muF, *shP, *shF, *sx, *zx, *sz, *zz, *nP, *nF, *oP, *oF, *DP, *DF;
  };
  BandForce force(ckfmt ? -1 : -2);  // 16-bit checkpoints run without the FP64 band
  std::vector<Member> M((size_t)ns);
  int ux0 = 1 << 30, ux1 = 0;
  long goff = 0;
  for (int s = 0; s < ns; ++s) {
    Member& m = M[(size_t)s];
    const int* gg = geo + 7 * s;
    m.x0 = gg[0];
    m.nx = gg[1];
    m.sg = ShotGeo{gg[2], gg[3], gg[4], gg[5], gg[6]};
    m.n = (long)m.nx * nz;
    ux0 = std::min(ux0, m.x0);
    ux1 = std::max(ux1, m.x0 + m.nx);
    for (auto* v : {&m.vh, &m.vp, &m.vn}) v->resize((size_t)m.n);
    for (int iz = 0; iz < nz; ++iz)
      for (int ix = 0; ix < m.nx; ++ix) {
        const long i = ix + (long)m.nx * iz, mi = m.x0 + ix + (long)nxm * iz;
        m.vh[(size_t)i] = vh_m[mi];
        m.vp[(size_t)i] = vp_m[mi];
        m.vn[(size_t)i] = vn_m[mi];
      }
    Geo g{m.nx, nz, cpml};
    m.tb = new Tables<T>(tab_a, tab_b, cpml, m.nx, nz);
    T* sy = syn + goff;
    if (ckfmt) {
      if constexpr (std::is_same<S, F32I>::value)
        forward_shot_ck<T, S>(m.vh.data(), m.vp.data(), m.vn.data(), *m.tb, src, g, m.sg, nt, sc, sy, &m.q, K, ckfmt);
    } else {
      forward_shot<T, S>(m.vh.data(), m.vp.data(), m.vn.data(), *m.tb, src, g, m.sg, nt, sc, sy, 1, &m.q);
    }
    m.res.resize((size_t)nt * m.sg.nrec);
    double phi = 0;
    for (size_t i = 0; i < m.res.size(); ++i) {
      m.res[i] = sy[i] - obs[goff + (long)i];
      phi += 0.5 * (double)m.res[i] * (double)m.res[i];
    }
    misfits[s] = phi;
    goff += (long)nt * m.sg.nrec;
    m.op = new AdjOp<T, S>(g, *m.tb, m.vh.data(), m.vp.data(), m.vn.data(), dh, dt);
    m.f = new Fields<T>(m.n, 14);
    Fields<T>& F = *m.f;
    m.muP = F[0]; m.muF = F[1]; m.shP = F[2]; m.shF = F[3];
    m.sx = F[4]; m.zx = F[5]; m.sz = F[6]; m.zz = F[7];
    m.nP = F[8]; m.nF = F[9]; m.oP = F[10]; m.oF = F[11];
    m.DP = F[12]; m.DF = F[13];
  }
  const int unx = ux1 - ux0;
  for (int k = nt; k >= 1; --k) {
    for (auto& m : M) {
      Geo g{m.nx, nz, cpml};
      adjoint_step_exact<T, S>(*m.op, g, m.sg, m.res.data() + (size_t)(k - 1) * m.sg.nrec, m.muP, m.muF, m.shP, m.shF,
                               m.sx, m.zx, m.sz, m.zz, m.nP, m.nF, m.oP, m.oF, m.DP, m.DF);
    }
#pragma omp parallel for
    for (int iz = kRim; iz < nz - kRim; ++iz)
      for (int u = 0; u < unx; ++u) {
        T ch[4] = {0, 0, 0, 0}, cn[4] = {0, 0, 0, 0}, cp[4] = {0, 0, 0, 0};
        for (int s = 0; s < ns && s < 4; ++s) {
          const Member& m = M[(size_t)s];
          const int ix = ux0 + u - m.x0;
          if (ix < kRim || ix >= m.nx - kRim) continue;
          const long i = ix + (long)m.nx * iz;
          const T qa = m.q[(size_t)(k - 1) * 2 * m.n + i], qb = m.q[(size_t)(k - 1) * 2 * m.n + m.n + i];
          const T lp = m.muP[i], lf = m.muF[i];
          ch[s] = dt2 * lp * qa;
          cn[s] = dt2 * lf * qa;
          cp[s] = dt2 * (lp + lf) * qb;
        }
        const long o = u + (long)unx * iz;
        g_hor[o] += ((ch[0] + ch[1]) + ch[2]) + ch[3];
        g_nmo[o] += ((cn[0] + cn[1]) + cn[2]) + cn[3];
        g_po[o] += ((cp[0] + cp[1]) + cp[2]) + cp[3];
      }
  }
  for (auto& m : M) {
    delete m.op;
    delete m.f;
    delete m.tb;
  }
}

}  // namespace

extern "C" {

// Forward gather only (observed-data modelling, main.c:334-366).  tab_a/tab_b: C-PML table (cpml entries).
#define ORC_FWD(SUFFIX, T, S)                                                                                   \
  void orc_forward_##SUFFIX(const T* vh, const T* vp, const T* vn, const T* tab_a, const T* tab_b, const T* src, \
                            int nx, int nz, int nt, int cpml, double dh, double dt, int isx, int isz, int irz,   \
                            int ir0, int nrec, T* gather) {                                                      \
    Geo g{nx, nz, cpml};                                                                                         \
    ShotGeo sg{isx, isz, irz, ir0, nrec};                                                                        \
    Tables<T> tb(tab_a, tab_b, cpml, nx, nz);                                                                    \
    forward_shot<T, S>(vh, vp, vn, tb, src, g, sg, nt, Scal{dh, dt}, gather, 0, nullptr);                        \
  }
ORC_FWD(mix, float, double)
ORC_FWD(f32, float, float)
ORC_FWD(f64, double, double)
ORC_FWD(f32d, float, F32D)
ORC_FWD(f32i, float, F32I)

// One shot's gradient.  mode_exact=0: reference-compatible (serialcalls.cu:68-90 semantics: one-step
// snapshot lag, dt scaling, forward-recursion C-PML on the adjoint); 1: exact discrete gradient.
// Gradients are ACCUMULATED into g_hor (gcx), g_po (gcz), g_nmo (gdx).  Returns 1/2 sum res^2 in FP64.
#define ORC_GRAD(SUFFIX, T, S)                                                                                    \
  double orc_gradient_##SUFFIX(int mode_exact, const T* vh, const T* vp, const T* vn, const T* tab_a,              \
                               const T* tab_b, const T* src, const T* obs, int nx, int nz, int nt, int cpml,       \
                               double dh, double dt, int isx, int isz, int irz, int ir0, int nrec, T* syn,         \
                               T* g_hor, T* g_po, T* g_nmo, double* misfit_ref32) {                                \
    return shot_gradient<T, S>(mode_exact, vh, vp, vn, tab_a, tab_b, src, obs, nx, nz, nt, cpml, dh, dt, isx, isz, \
                               irz, ir0, nrec, syn, g_hor, g_po, g_nmo, misfit_ref32);                             \
  }
ORC_GRAD(mix, float, double)
ORC_GRAD(f32, float, float)
ORC_GRAD(f64, double, double)
ORC_GRAD(f32d, float, F32D)
ORC_GRAD(f32i, float, F32I)

// orc_gradient_* that also records the fields it went through: fwd_states[n] = (P^n, F^n) for n = 0..nt (the
// state forward step n starts from; (nt+1) x 2 x nx*nz) and adj_states[k-1] = the adjoint (P, F) fields
// paired with k = nt..1 ((nt) x 2 x nx*nz): mu^k of the exact adjoint, or the reference-compatible adjoint's
// (P_pre, F_pre) after its step it = nt - k (main.c:478-492).
#define ORC_GRAD_STATES(SUFFIX, T, S)                                                                              \
  double orc_gradient_states_##SUFFIX(int mode_exact, const T* vh, const T* vp, const T* vn, const T* tab_a,        \
                                      const T* tab_b, const T* src, const T* obs, int nx, int nz, int nt, int cpml, \
                                      double dh, double dt, int isx, int isz, int irz, int ir0, int nrec, T* syn,   \
                                      T* g_hor, T* g_po, T* g_nmo, T* fwd_states, T* adj_states) {                  \
    return shot_gradient<T, S>(mode_exact, vh, vp, vn, tab_a, tab_b, src, obs, nx, nz, nt, cpml, dh, dt, isx, isz,  \
                               irz, ir0, nrec, syn, g_hor, g_po, g_nmo, nullptr, fwd_states, adj_states);           \
  }
ORC_GRAD_STATES(mix, float, double)
ORC_GRAD_STATES(f64, double, double)

// Exact gradient of a gradient group (1..4 shots with their own windows of the nxm-wide model), accumulated
// into g_* over the union of the windows (union width x nz); per-shot FP64 misfits into misfits[ns].
#define ORC_GROUP(SUFFIX, T, S)                                                                                    \
  void orc_group_gradient_##SUFFIX(const T* vh, const T* vp, const T* vn, const T* tab_a, const T* tab_b,           \
                                   const T* src, int nxm, int nz, int nt, int cpml, double dh, double dt, int ns,   \
                                   const int* geo, const T* obs, T* syn, T* g_hor, T* g_po, T* g_nmo,               \
                                   double* misfits) {                                                               \
    group_gradient<T, S>(vh, vp, vn, tab_a, tab_b, src, nxm, nz, nt, cpml, dh, dt, ns, geo, obs, syn, g_hor, g_po,  \
                         g_nmo, misfits);                                                                           \
  }
ORC_GROUP(mix, float, double)
ORC_GROUP(f32, float, float)
ORC_GROUP(f64, double, double)
ORC_GROUP(f32d, float, F32D)
ORC_GROUP(f32i, float, F32I)

// FP64 near-surface band of the f32i arithmetic: k = 0 the library's default (24 rows where nz - cpml - 4 >= 24),
// k < 0 none, k > 4 rows 4..k-1; and the rows it gives a grid.
void orc_set_surface_rows(int k) { g_band_rows = k; }
int orc_surface_rows(int nz, int cpml) { return band_rows(nz, cpml); }

// One checkpoint field rounded as a 16-bit checkpoint stores it (fmt 1 BF16, 2 FP16; scale from the
// field's max-abs).
void orc_ck_round(float* f, long n, int fmt) { ck_round_field(f, n, fmt); }

// The f32i group gradient with the batched path's checkpoints every K steps in format ckfmt (0 FP32,
// 1 BF16, 2 FP16, both with per-field power-of-two scaling; only the 16-bit formats change the result).
void orc_group_gradient_ck_f32i(const float* vh, const float* vp, const float* vn, const float* tab_a,
                                const float* tab_b, const float* src, int nxm, int nz, int nt, int cpml, double dh,
                                double dt, int ns, const int* geo, const float* obs, float* syn, float* g_hor,
                                float* g_po, float* g_nmo, double* misfits, int K, int ckfmt) {
  group_gradient<float, F32I>(vh, vp, vn, tab_a, tab_b, src, nxm, nz, nt, cpml, dh, dt, ns, geo, obs, syn, g_hor, g_po,
                              g_nmo, misfits, K, ckfmt);
}

// One linear forward state step without source (for the dot-product test), FP64.
// state = 8 fields of nx*nz: P, P_old, F, F_old, psx, zx, psz, zz  ->  out: P', P, F', F, psx', zx', psz', zz'
void orc_step_fwd_f64(const double* vh, const double* vp, const double* vn, const double* tab_a,
                      const double* tab_b, int nx, int nz, int cpml, double dh, double dt, const double* in,
                      double* out) {
  Geo g{nx, nz, cpml};
  Tables<double> tb(tab_a, tab_b, cpml, nx, nz);
  const long n = (long)nx * nz;
  FwdState<double, double> st(g);
  std::memcpy(st.f[0], in + 0 * n, n * sizeof(double));  // P_pre
  std::memcpy(st.f[1], in + 1 * n, n * sizeof(double));  // P_past
  std::memcpy(st.f[2], in + 2 * n, n * sizeof(double));  // F_pre
  std::memcpy(st.f[3], in + 3 * n, n * sizeof(double));  // F_past
  std::memcpy(st.f[6], in + 4 * n, n * sizeof(double));  // psx
  std::memcpy(st.f[8], in + 5 * n, n * sizeof(double));  // zx
  std::memcpy(st.f[7], in + 6 * n, n * sizeof(double));  // psz
  std::memcpy(st.f[9], in + 7 * n, n * sizeof(double));  // zz
  std::vector<double> zero_src(1, 0.0);
  ShotGeo sg{nx / 2, nz / 2, nz / 2, 0, 0};
  const double dh2 = dh * dh, C = dt * dt / dh2;
  fwd_step<double, double>(st, vh, vp, vn, tb.ax.data(), tb.bx.data(), tb.az.data(), tb.bz.data(), zero_src.data(), C,
                           dh, dh2, sg, 0, nullptr, 1);
  std::memcpy(out + 0 * n, st.Pcur(), n * sizeof(double));
  std::memcpy(out + 1 * n, st.Pold(), n * sizeof(double));
  std::memcpy(out + 2 * n, st.Fcur(), n * sizeof(double));
  std::memcpy(out + 3 * n, st.Fold(), n * sizeof(double));
  std::memcpy(out + 4 * n, st.f[6], n * sizeof(double));
  std::memcpy(out + 5 * n, st.f[8], n * sizeof(double));
  std::memcpy(out + 6 * n, st.f[7], n * sizeof(double));
  std::memcpy(out + 7 * n, st.f[9], n * sizeof(double));
}

// Transpose of orc_step_fwd_f64: in = adjoint of (P', P, F', F, psx', zx', psz', zz'), out = adjoint of
// (P, P_old, F, F_old, psx, zx, psz, zz).
void orc_step_adj_f64(const double* vh, const double* vp, const double* vn, const double* tab_a,
                      const double* tab_b, int nx, int nz, int cpml, double dh, double dt, const double* in,
                      double* out) {
  Geo g{nx, nz, cpml};
  Tables<double> tb(tab_a, tab_b, cpml, nx, nz);
  const long n = (long)nx * nz;
  AdjOp<double, double> op(g, tb, vh, vp, vn, dh, dt);
  std::vector<double> mem(in + 4 * n, in + 8 * n);
  op.apply(in + 0 * n, in + 2 * n, in + 1 * n, in + 3 * n, mem.data() + 0 * n, mem.data() + 1 * n,
           mem.data() + 2 * n, mem.data() + 3 * n, out + 0 * n, out + 2 * n, out + 1 * n, out + 3 * n);
  std::memcpy(out + 4 * n, mem.data(), 4 * n * sizeof(double));
}

}  // extern "C"
