// Fused, shot-batched stencil kernels for gfx950 (compiled as part of plan.hip).
//
// Work decomposition: a work item is (shot, 64-column band, run of z rows); one wavefront owns one item
// and marches down its rows.  The x-direction stencil samples of a row are neighbouring lanes' loads of
// the same 256-byte row segment (served by the L1 after the first touch); the z-direction samples live in
// a 9-deep register queue, so each field row is fetched from HBM once per step.  Four items (normally the
// four adjacent bands of one row run) form a 256-thread workgroup, and workgroup ids are remapped so
// that each XCD receives a contiguous block of items (whole shots), keeping halo rows in its own L2.
//
// The C-PML memory variables that are differentiated again (psi) are updated first by a small kernel over
// the C-PML cells only, so the main kernel needs a radius-4 neighbourhood everywhere.
//
// Each stored value is the IEEE result of the same expression, in the same order, as the reference
// kernel it replaces (src/parallelcalls.cu), computed with stencil sums in S (float, or double for the
// reference's mixed arithmetic).

#include <hip/hip_runtime.h>

#include "batch_kernels.h"
#include "fwi_common.h"

namespace fwi {

// ---------------------------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------------------------

// psi_px, psi_fz on the C-PML cells (parallelcalls.cu:48-63).  Cells of the side strips carry
// psi_px; cells of the bottom strip carry psi_fz (the other variable keeps a = b = 0 there and stays 0).
template <class S>
__global__ void __launch_bounds__(256) k_fwd_strip(StripArgs a, const float* __restrict__ Pc,
                                                   const float* __restrict__ Fc, float* __restrict__ psx,
                                                   float* __restrict__ psz) {
  const int2 blk = a.blocks[blockIdx.x];
  const ShotDev sd = a.g.shots[blk.x];
  const int k = blk.y + threadIdx.x;
  if (k >= sd.nstrip) return;
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml;
  int ix, iz;
  strip_cell(k, nx, nz, cpml, ix, iz);
  const long i = sd.off + ix + (long)nx * iz;
  if (ix < cpml || ix > nx - cpml - 1) {
    const float t = d1_at<S>(Pc, i, 1, a.g.dh);
    psx[i] = coef_x(a.g.tb, ix, nx, cpml) * psx[i] + coef_x(a.g.ta, ix, nx, cpml) * t;
  }
  if (iz > nz - cpml - 1) {
    const float t = d1_at<S>(Fc, i, nx, a.g.dh);
    psz[i] = coef_z(a.g.tb, iz, nz, cpml) * psz[i] + coef_z(a.g.ta, iz, nz, cpml) * t;
  }
}

// One forward step of every shot (parallelcalls.cu:120-130 SecondDerivate_P, :65-86 ZETA_F, :88-117
// propagator_F).  QM selects what is kept for the adjoint pass: 0 nothing, 1 the model-multiplied terms
// a = p2 + aten_p, b = r2 + aten_f (exact gradient), 2 p2, r2 (reference-compatible gradient).
template <int QM, class S>
__global__ void __launch_bounds__(256) k_fwd_main(FwdArgs a) {
  __shared__ float xbuf[4][72];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int item = xcd_logical_block(blockIdx.x, gridDim.x) * 4 + wave;
  if (item >= a.nitems) return;
  const int4 itm = a.items[item];
  const ShotDev sd = a.g.shots[itm.x];
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml;
  const int ix = itm.y + lane;
  const bool inb = ix < nx;
  const bool act = ix >= kRim && ix < nx - kRim;
  const int hx = itm.y + halo_col(lane);
  const bool hin = lane < 8 && hx >= 0 && hx < nx;
  const bool xstrip = ix < cpml || ix > nx - cpml - 1;
  const float axv = act ? coef_x(a.g.ta, ix, nx, cpml) : 0.f;
  const float bxv = act ? coef_x(a.g.tb, ix, nx, cpml) : 0.f;
  const float dh = a.g.dh, dh2 = a.g.dh2, C = a.C;
  const float* __restrict__ P = a.Pc + sd.off;
  const float* __restrict__ F = a.Fc + sd.off;
  float* __restrict__ Po = a.Po + sd.off;
  float* __restrict__ Fo = a.Fo + sd.off;
  const float* __restrict__ vh = a.m.vh + sd.x0;
  const float* __restrict__ vp = a.m.vp + sd.x0;
  const float* __restrict__ vn = a.m.vn + sd.x0;
  float* buf = xbuf[wave];
  const float srcv = a.src[a.it];
  const int z0 = itm.z, z1 = itm.w;
  float fq[9];
#pragma unroll
  for (int k = 0; k < 8; ++k) fq[k] = act ? F[ix + (long)nx * (z0 - 4 + k)] : 0.f;
  for (int z = z0; z < z1; ++z) {
    const long row = (long)nx * z;
    const long i = ix + row;
    fq[8] = act ? F[i + 4L * nx] : 0.f;
    const float c = inb ? P[i] : 0.f;
    const float h = hin ? P[hx + row] : 0.f;
    float u[9];
    xrow9(buf, lane, c, h, u);
    if (act) {
      const float p2 = d2s<S>(u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8], dh2);
      const float r2 = d2s<S>(fq[0], fq[1], fq[2], fq[3], fq[4], fq[5], fq[6], fq[7], fq[8], dh2);
      float atp = 0.f, atf = 0.f;
      if (xstrip || z > nz - cpml - 1) {
        const long g = sd.off + i;
        const float tx = d1_at<S>(a.psx, g, 1, dh);
        const float tz = d1_at<S>(a.psz, g, nx, dh);
        const float tpx = p2 + tx;
        const float tfz = r2 + tz;
        const float azv = coef_z(a.g.ta, z, nz, cpml), bzv = coef_z(a.g.tb, z, nz, cpml);
        const float zxn = bxv * a.zx[g] + axv * tpx;
        const float zzn = bzv * a.zz[g] + azv * tfz;
        a.zx[g] = zxn;
        a.zz[g] = zzn;
        atp = tx + zxn;
        atf = tz + zzn;
      }
      const float tp = dh2 * (p2 + atp);
      const float tf = dh2 * (r2 + atf);
      const long mi = ix + (long)a.m.nxm * z;
      const float vhv = vh[mi], vpv = vp[mi], vnv = vn[mi];
      float pn = 2.f * c - Po[i] + vhv * C * tp + vpv * C * tf;
      float fn = 2.f * fq[4] - Fo[i] + vnv * C * tp + vpv * C * tf;
      if (ix == sd.isx && z == sd.isz) {
        pn += srcv;
        fn += srcv;
      }
      Po[i] = pn;
      Fo[i] = fn;
      if (a.gather && z == sd.irz && ix >= sd.ir0 && ix < sd.ir0 + sd.nrec)
        a.gather[sd.goff + (long)a.it * sd.nrec + (ix - sd.ir0)] = pn + fn;
      if (QM == 1) {
        a.qa[sd.off + i] = p2 + atp;
        a.qb[sd.off + i] = r2 + atf;
      } else if (QM == 2) {
        a.qa[sd.off + i] = p2;
        a.qb[sd.off + i] = r2;
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) fq[k] = fq[k + 1];
  }
}

// ---------------------------------------------------------------------------------------------
// Reference-compatible adjoint (serialcalls.cu:77-88): aux1 = v_hor lP, aux2 = v_po lP, aux3 = v_nmo lF,
// aux4 = v_po lF are formed on the fly from the adjoint fields and the model.
// ---------------------------------------------------------------------------------------------

// PSI_R on the C-PML cells (parallelcalls.cu:179-198).
template <class S>
__global__ void __launch_bounds__(256) k_adj_ref_strip(StripArgs a, const float* __restrict__ Lc,
                                                       const float* __restrict__ Mc, float* __restrict__ s1,
                                                       float* __restrict__ s2, float* __restrict__ s3,
                                                       float* __restrict__ s4) {
  const int2 blk = a.blocks[blockIdx.x];
  const ShotDev sd = a.g.shots[blk.x];
  const int k = blk.y + threadIdx.x;
  if (k >= sd.nstrip) return;
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml, nxm = a.m.nxm;
  int ix, iz;
  strip_cell(k, nx, nz, cpml, ix, iz);
  const long i = sd.off + ix + (long)nx * iz;
  const long mi = sd.x0 + ix + (long)nxm * iz;
  const float dh = a.g.dh;
  if (ix < cpml || ix > nx - cpml - 1) {
    float u1[9], u3[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      u1[j] = a.m.vh[mi + j - 4] * Lc[i + j - 4];
      u3[j] = a.m.vn[mi + j - 4] * Mc[i + j - 4];
    }
    const float p1cx = d1s<S>(u1[0], u1[1], u1[2], u1[3], u1[5], u1[6], u1[7], u1[8], dh);
    const float r1dx = d1s<S>(u3[0], u3[1], u3[2], u3[3], u3[5], u3[6], u3[7], u3[8], dh);
    const float axv = coef_x(a.g.ta, ix, nx, cpml), bxv = coef_x(a.g.tb, ix, nx, cpml);
    s1[i] = bxv * s1[i] + axv * p1cx;
    s3[i] = bxv * s3[i] + axv * r1dx;
  }
  if (iz > nz - cpml - 1) {
    float u2[9], u4[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const float v = a.m.vp[mi + (long)(j - 4) * nxm];
      u2[j] = v * Lc[i + (long)(j - 4) * nx];
      u4[j] = v * Mc[i + (long)(j - 4) * nx];
    }
    const float p1cz = d1s<S>(u2[0], u2[1], u2[2], u2[3], u2[5], u2[6], u2[7], u2[8], dh);
    const float r1cz = d1s<S>(u4[0], u4[1], u4[2], u4[3], u4[5], u4[6], u4[7], u4[8], dh);
    const float azv = coef_z(a.g.ta, iz, nz, cpml), bzv = coef_z(a.g.tb, iz, nz, cpml);
    s2[i] = bzv * s2[i] + azv * p1cz;
    s4[i] = bzv * s4[i] + azv * r1cz;
  }
}

// secondDerivate (:200-214) + ZETA_R (:216-243) + retroPropagation (:144-164) + GRADIENT (:245-259).
// The imaging uses the stored p2 = D2x P_f, r2 = D2z F_f of the forward step two behind (the reference's
// one-step snapshot lag, serialcalls.cu:60,76).
template <class S>
__global__ void __launch_bounds__(256) k_adj_ref_main(AdjRefArgs a) {
  __shared__ float xbuf[4][72];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int item = xcd_logical_block(blockIdx.x, gridDim.x) * 4 + wave;
  if (item >= a.nitems) return;
  const int4 itm = a.items[item];
  const ShotDev sd = a.g.shots[itm.x];
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml, nxm = a.m.nxm;
  const int ix = itm.y + lane;
  const bool inb = ix < nx;
  const bool act = ix >= kRim && ix < nx - kRim;
  const int hx = itm.y + halo_col(lane);
  const bool hin = lane < 8 && hx >= 0 && hx < nx;
  const bool xstrip = ix < cpml || ix > nx - cpml - 1;
  const float axv = act ? coef_x(a.g.ta, ix, nx, cpml) : 0.f;
  const float bxv = act ? coef_x(a.g.tb, ix, nx, cpml) : 0.f;
  const float dh = a.g.dh, dh2 = a.g.dh2;
  const float* __restrict__ L = a.Lc + sd.off;
  const float* __restrict__ M = a.Mc + sd.off;
  float* __restrict__ Lo = a.Lo + sd.off;
  float* __restrict__ Mo = a.Mo + sd.off;
  const float* __restrict__ vh = a.m.vh + sd.x0;
  const float* __restrict__ vp = a.m.vp + sd.x0;
  const float* __restrict__ vn = a.m.vn + sd.x0;
  float* buf = xbuf[wave];
  const int z0 = itm.z, z1 = itm.w;
  // z queues: aux2 = v_po lP and aux4 = v_po lF at rows z-4..z+4; lP, lF themselves at rows z..z+4
  float q2[9], q4[9], lq[5], mq[5];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int r = z0 - 4 + k;
    const float l = inb ? L[ix + (long)nx * r] : 0.f;
    const float m = inb ? M[ix + (long)nx * r] : 0.f;
    const float v = act ? vp[ix + (long)nxm * r] : 0.f;
    q2[k] = act ? v * l : 0.f;
    q4[k] = act ? v * m : 0.f;
    if (k >= 4) {
      lq[k - 4] = l;
      mq[k - 4] = m;
    }
  }
  for (int z = z0; z < z1; ++z) {
    const long i = ix + (long)nx * z;
    const long mi = ix + (long)nxm * z;
    {
      const float l = inb ? L[i + 4L * nx] : 0.f;
      const float m = inb ? M[i + 4L * nx] : 0.f;
      const float v = act ? vp[mi + 4L * nxm] : 0.f;
      lq[4] = l;
      mq[4] = m;
      q2[8] = act ? v * l : 0.f;
      q4[8] = act ? v * m : 0.f;
    }
    // x-direction auxiliaries aux1 = v_hor lP and aux3 = v_nmo lF of the band row and its halo
    float u1o = 0.f, u3o = 0.f, u1h = 0.f, u3h = 0.f;
    if (inb) {
      u1o = vh[mi] * lq[0];
      u3o = vn[mi] * mq[0];
    }
    if (hin) {
      const long hi = hx + (long)nx * z, hm = hx + (long)nxm * z;
      u1h = vh[hm] * L[hi];
      u3h = vn[hm] * M[hi];
    }
    float u1[9], u3[9];
    xrow9(buf, lane, u1o, u1h, u1);
    xrow9(buf, lane, u3o, u3h, u3);
    if (act) {
      const float p2cx = d2s<S>(u1[0], u1[1], u1[2], u1[3], u1[4], u1[5], u1[6], u1[7], u1[8], dh2);
      const float r2dx = d2s<S>(u3[0], u3[1], u3[2], u3[3], u3[4], u3[5], u3[6], u3[7], u3[8], dh2);
      const float p2cz = d2s<S>(q2[0], q2[1], q2[2], q2[3], q2[4], q2[5], q2[6], q2[7], q2[8], dh2);
      const float r2cz = d2s<S>(q4[0], q4[1], q4[2], q4[3], q4[4], q4[5], q4[6], q4[7], q4[8], dh2);
      float t1 = 0.f, t2 = 0.f, t3 = 0.f, t4 = 0.f;
      if (xstrip || z > nz - cpml - 1) {
        const long g = sd.off + i;
        const float d1 = d1_at<S>(a.s1, g, 1, dh);
        const float d2 = d1_at<S>(a.s2, g, nx, dh);
        const float d3 = d1_at<S>(a.s3, g, 1, dh);
        const float d4 = d1_at<S>(a.s4, g, nx, dh);
        const float azv = coef_z(a.g.ta, z, nz, cpml), bzv = coef_z(a.g.tb, z, nz, cpml);
        const float n1 = bxv * a.z1[g] + axv * (d1 + p2cx);
        const float n2 = bzv * a.z2[g] + azv * (d2 + p2cz);
        const float n3 = bxv * a.z3[g] + axv * (d3 + r2dx);
        const float n4 = bzv * a.z4[g] + azv * (d4 + r2cz);
        a.z1[g] = n1;
        a.z2[g] = n2;
        a.z3[g] = n3;
        a.z4[g] = n4;
        t1 = n1 + d1;
        t2 = n2 + d2;
        t3 = n3 + d3;
        t4 = n4 + d4;
      }
      float ln = 2.f * lq[0] - Lo[i] + a.dh2C * (p2cx + t1 + r2dx + t3);
      float mn = 2.f * mq[0] - Mo[i] + a.dh2C * (p2cz + r2cz + t2 + t4);
      if (z == sd.irz && ix >= sd.ir0 && ix < sd.ir0 + sd.nrec) {
        const float r = a.res[sd.goff + (long)a.rit * sd.nrec + (ix - sd.ir0)];
        ln += r;
        mn += r;
      }
      Lo[i] = ln;
      Mo[i] = mn;
      if (a.qa) {
        const long g = sd.off + i;
        const float p2 = a.qa[g], r2 = a.qb[g];
        a.gcx[g] += a.dt * p2 * ln;
        a.gcz[g] += a.dt * r2 * (ln + mn);
        a.gdx[g] += a.dt * p2 * mn;
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      q2[k] = q2[k + 1];
      q4[k] = q4[k + 1];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      lq[k] = lq[k + 1];
      mq[k] = mq[k + 1];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Exact discrete adjoint.  With chi_x / chi_z the side / bottom C-PML cells and Gamma = chi_x U chi_z:
//   abar = dt^2 (v_hor mP + v_nmo mF), bbar = dt^2 v_po (mP + mF)
//   zt_x = chi_x (zbar_x + abar), gbar_x = abar + a_x zt_x, zbar_x <- b_x zt_x
//   pt_x = chi_x (sbar_x - D1x(Gamma gbar_x)), sbar_x <- b_x pt_x
//   m^k = 2 m^{k+1} + D2x gbar_x - D1x(a_x pt_x) - m^{k+2} (+ residual at the receivers), likewise z.
// ---------------------------------------------------------------------------------------------

// Phase 1 on the C-PML cells: zt (kept for this step) and the new zeta-bar.
__global__ void __launch_bounds__(256) k_adj_ex_strip1(StripArgs a, float dt2, const float* __restrict__ Lc,
                                                       const float* __restrict__ Mc, float* __restrict__ zx,
                                                       float* __restrict__ zz, float* __restrict__ ztx,
                                                       float* __restrict__ ztz) {
  const int2 blk = a.blocks[blockIdx.x];
  const ShotDev sd = a.g.shots[blk.x];
  const int k = blk.y + threadIdx.x;
  if (k >= sd.nstrip) return;
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml;
  int ix, iz;
  strip_cell(k, nx, nz, cpml, ix, iz);
  const long i = sd.off + ix + (long)nx * iz;
  const long mi = sd.x0 + ix + (long)a.m.nxm * iz;
  const float mP = Lc[i], mF = Mc[i];
  if (ix < cpml || ix > nx - cpml - 1) {
    const float ab = dt2 * (a.m.vh[mi] * mP + a.m.vn[mi] * mF);
    const float t = zx[i] + ab;
    ztx[i] = t;
    zx[i] = coef_x(a.g.tb, ix, nx, cpml) * t;
  }
  if (iz > nz - cpml - 1) {
    const float bb = dt2 * (a.m.vp[mi] * (mP + mF));
    const float t = zz[i] + bb;
    ztz[i] = t;
    zz[i] = coef_z(a.g.tb, iz, nz, cpml) * t;
  }
}

// gbar at cell (jx, jz) of one shot: Omega & Gamma masked on request.
__device__ __forceinline__ float gbar_x(const AdjExArgs& a, const ShotDev& sd, const float* __restrict__ L,
                                        const float* __restrict__ M, int jx, int jz, bool gamma_only) {
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml;
  if (jx < kRim || jx >= nx - kRim || jz < kRim || jz >= nz - kRim) return 0.f;
  const bool gam = jx < cpml || jx > nx - cpml - 1 || jz > nz - cpml - 1;
  if (gamma_only && !gam) return 0.f;
  const long i = jx + (long)nx * jz;
  const long mi = sd.x0 + jx + (long)a.m.nxm * jz;
  const float ab = a.dt2 * (a.m.vh[mi] * L[i] + a.m.vn[mi] * M[i]);
  return ab + coef_x(a.g.ta, jx, nx, cpml) * a.ztx[sd.off + i];
}
__device__ __forceinline__ float gbar_z(const AdjExArgs& a, const ShotDev& sd, const float* __restrict__ L,
                                        const float* __restrict__ M, int jx, int jz, bool gamma_only) {
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml;
  if (jx < kRim || jx >= nx - kRim || jz < kRim || jz >= nz - kRim) return 0.f;
  const bool gam = jx < cpml || jx > nx - cpml - 1 || jz > nz - cpml - 1;
  if (gamma_only && !gam) return 0.f;
  const long i = jx + (long)nx * jz;
  const long mi = sd.x0 + jx + (long)a.m.nxm * jz;
  const float bb = a.dt2 * (a.m.vp[mi] * (L[i] + M[i]));
  return bb + coef_z(a.g.ta, jz, nz, cpml) * a.ztz[sd.off + i];
}

// Phase 2 on the C-PML cells: psi-bar_tot -> a*psi-bar_tot (kept for this step) and the new psi-bar.
template <class S>
__global__ void __launch_bounds__(256) k_adj_ex_strip2(StripArgs s, AdjExArgs a) {
  const int2 blk = s.blocks[blockIdx.x];
  const ShotDev sd = a.g.shots[blk.x];
  const int k = blk.y + threadIdx.x;
  if (k >= sd.nstrip) return;
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml;
  int ix, iz;
  strip_cell(k, nx, nz, cpml, ix, iz);
  const long i = sd.off + ix + (long)nx * iz;
  const float* L = a.Lc + sd.off;
  const float* M = a.Mc + sd.off;
  const float dh = a.g.dh;
  if (ix < cpml || ix > nx - cpml - 1) {
    float w[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) w[j] = (j == 4) ? 0.f : gbar_x(a, sd, L, M, ix + j - 4, iz, true);
    const float t = a.sx[i] - d1s<S>(w[0], w[1], w[2], w[3], w[5], w[6], w[7], w[8], dh);
    a.sx[i] = coef_x(a.g.tb, ix, nx, cpml) * t;
    a.ptx[i] = coef_x(a.g.ta, ix, nx, cpml) * t;
  }
  if (iz > nz - cpml - 1) {
    float w[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) w[j] = (j == 4) ? 0.f : gbar_z(a, sd, L, M, ix, iz + j - 4, true);
    const float t = a.sz[i] - d1s<S>(w[0], w[1], w[2], w[3], w[5], w[6], w[7], w[8], dh);
    a.sz[i] = coef_z(a.g.tb, iz, nz, cpml) * t;
    a.ptz[i] = coef_z(a.g.ta, iz, nz, cpml) * t;
  }
}

// Main exact adjoint step with the exact imaging condition
//   g_hor += dt^2 mP^k a^{k-1}, g_nmo += dt^2 mF^k a^{k-1}, g_po += dt^2 (mP^k + mF^k) b^{k-1}.
template <class S>
__global__ void __launch_bounds__(256) k_adj_ex_main(AdjExArgs a) {
  __shared__ float xbuf[4][72];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int item = xcd_logical_block(blockIdx.x, gridDim.x) * 4 + wave;
  if (item >= a.nitems) return;
  const int4 itm = a.items[item];
  const ShotDev sd = a.g.shots[itm.x];
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml, nxm = a.m.nxm;
  const int ix = itm.y + lane;
  const bool inb = ix < nx;
  const bool act = ix >= kRim && ix < nx - kRim;
  const int hx = itm.y + halo_col(lane);
  const bool hact = lane < 8 && hx >= kRim && hx < nx - kRim;
  // bands within reach of a side strip need the C-PML terms in x
  const bool xnear = itm.y < cpml + 4 || itm.y + 64 > nx - cpml - 4;
  const float axo = act ? coef_x(a.g.ta, ix, nx, cpml) : 0.f;
  const float axh = hact ? coef_x(a.g.ta, hx, nx, cpml) : 0.f;
  const float dh = a.g.dh, dh2 = a.g.dh2, dt2 = a.dt2;
  const float* __restrict__ L = a.Lc + sd.off;
  const float* __restrict__ M = a.Mc + sd.off;
  float* __restrict__ Lo = a.Lo + sd.off;
  float* __restrict__ Mo = a.Mo + sd.off;
  const float* __restrict__ vh = a.m.vh + sd.x0;
  const float* __restrict__ vp = a.m.vp + sd.x0;
  const float* __restrict__ vn = a.m.vn + sd.x0;
  const float* __restrict__ ztz = a.ztz + sd.off;
  const float* __restrict__ ztx = a.ztx + sd.off;
  float* buf = xbuf[wave];
  const int z0 = itm.z, z1 = itm.w;
  const int zc = nz - cpml - 4;  // rows at or below this need the C-PML terms in z
  float gq[9];         // gbar_z along z (rows z-4..z+4)
  float lq[5], mq[5];  // mP, mF at rows z..z+4
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int r = z0 - 4 + k;
    const long ir = ix + (long)nx * r;
    const float l = act ? L[ir] : 0.f;
    const float m = act ? M[ir] : 0.f;
    float v = 0.f;
    if (act && r >= kRim) v = dt2 * (vp[ix + (long)nxm * r] * (l + m)) + coef_z(a.g.ta, r, nz, cpml) * ztz[ir];
    gq[k] = v;
    if (k >= 4) {
      lq[k - 4] = l;
      mq[k - 4] = m;
    }
  }
  for (int z = z0; z < z1; ++z) {
    const long i = ix + (long)nx * z;
    const long mi = ix + (long)nxm * z;
    {
      const int r = z + 4;
      const float l = act ? L[i + 4L * nx] : 0.f;
      const float m = act ? M[i + 4L * nx] : 0.f;
      lq[4] = l;
      mq[4] = m;
      float v = 0.f;
      if (act && r < nz - kRim)
        v = dt2 * (vp[mi + 4L * nxm] * (l + m)) + coef_z(a.g.ta, r, nz, cpml) * ztz[i + 4L * nx];
      gq[8] = v;
    }
    // gbar_x of the band row (zero outside Omega) and of the halo columns
    float go = 0.f, gh = 0.f;
    if (act) {
      go = dt2 * (vh[mi] * lq[0] + vn[mi] * mq[0]);
      if (xnear) go = go + axo * ztx[i];
    }
    if (hact) {
      const long hi = hx + (long)nx * z, hm = hx + (long)nxm * z;
      gh = dt2 * (vh[hm] * L[hi] + vn[hm] * M[hi]);
      if (xnear) gh = gh + axh * ztx[hi];
    }
    float gx[9];
    xrow9(buf, lane, go, gh, gx);
    if (act) {
      float ap = d2s<S>(gx[0], gx[1], gx[2], gx[3], gx[4], gx[5], gx[6], gx[7], gx[8], dh2);
      float af = d2s<S>(gq[0], gq[1], gq[2], gq[3], gq[4], gq[5], gq[6], gq[7], gq[8], dh2);
      if (xnear) ap = ap - d1_at<S>(a.ptx, sd.off + i, 1, dh);
      if (z >= zc) af = af - d1_at<S>(a.ptz, sd.off + i, nx, dh);
      float ln = 2.f * lq[0] + ap - Lo[i];
      float mn = 2.f * mq[0] + af - Mo[i];
      if (z == sd.irz && ix >= sd.ir0 && ix < sd.ir0 + sd.nrec) {
        const float r = a.res[sd.goff + (long)a.rit * sd.nrec + (ix - sd.ir0)];
        ln += r;
        mn += r;
      }
      Lo[i] = ln;
      Mo[i] = mn;
      if (a.qa) {
        const long g = sd.off + i;
        const float qa = a.qa[g], qb = a.qb[g];
        a.gh[g] += dt2 * ln * qa;
        a.gn[g] += dt2 * mn * qa;
        a.gp[g] += dt2 * (ln + mn) * qb;
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) gq[k] = gq[k + 1];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      lq[k] = lq[k + 1];
      mq[k] = mq[k + 1];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Checkpoint strips, residual / misfit, cross-shot sum, chain rule
// ---------------------------------------------------------------------------------------------

// Pack (dir=0) or unpack (dir=1) four full-field C-PML variables at the C-PML cells.
__global__ void __launch_bounds__(256) k_strip_copy(StripArgs a, int dir, float* f0, float* f1, float* f2, float* f3,
                                                    float* packed, long long nstrip_total) {
  const int2 blk = a.blocks[blockIdx.x];
  const ShotDev sd = a.g.shots[blk.x];
  const int k = blk.y + threadIdx.x;
  if (k >= sd.nstrip) return;
  int ix, iz;
  strip_cell(k, sd.nx, a.g.nz, a.g.cpml, ix, iz);
  const long i = sd.off + ix + (long)sd.nx * iz;
  const long j = sd.soff + k;
  if (dir == 0) {
    packed[j] = f0[i];
    packed[j + nstrip_total] = f1[i];
    packed[j + 2 * nstrip_total] = f2[i];
    packed[j + 3 * nstrip_total] = f3[i];
  } else {
    f0[i] = packed[j];
    f1[i] = packed[j + nstrip_total];
    f2[i] = packed[j + 2 * nstrip_total];
    f3[i] = packed[j + 3 * nstrip_total];
  }
}

// res = syn - obs (parallelcalls.cu:133-142) and per-block FP64 partial sums of res^2.
__global__ void __launch_bounds__(256) k_residual(const float* __restrict__ syn, const float* __restrict__ obs,
                                                  float* __restrict__ res, long long n, double* __restrict__ partial) {
  __shared__ double red[4];
  double acc = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float r = syn[i] - obs[i];
    if (res) res[i] = r;
    acc += (double)r * (double)r;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Sum of n doubles by one workgroup in a fixed order (deterministic).  out[0] (+)= scale * sum.
__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ partial, int n, double scale,
                                                      double* __restrict__ out, int accumulate) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double s = scale * ((red[0] + red[1]) + (red[2] + red[3]));
    out[0] = accumulate ? out[0] + s : s;
  }
}

// FP64 cross-shot sum of the batch's FP32 per-shot gradients into the model gradient (3 fields).
// Shots are added in batch order for a deterministic result.
__global__ void __launch_bounds__(256) k_cross_shot_sum(const ShotDev* __restrict__ shots, int nshots, int nxm, int nz,
                                                        const float* __restrict__ g0, const float* __restrict__ g1,
                                                        const float* __restrict__ g2, double* __restrict__ out,
                                                        long long nmodel) {
  const long long c = blockIdx.x * 256LL + threadIdx.x;
  if (c >= nmodel) return;
  const int x = (int)(c % nxm), z = (int)(c / nxm);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int s = 0; s < nshots; ++s) {
    const ShotDev sd = shots[s];
    const int lx = x - sd.x0;
    if (lx < 0 || lx >= sd.nx) continue;
    const long long i = sd.off + lx + (long long)sd.nx * z;
    a0 += (double)g0[i];
    a1 += (double)g1[i];
    a2 += (double)g2[i];
  }
  out[c] += a0;
  out[c + nmodel] += a1;
  out[c + 2 * nmodel] += a2;
  (void)nz;
}

// Chain rule to (vp, epsilon, delta).
__global__ void __launch_bounds__(256) k_chain_rule(const float* __restrict__ vp, const float* __restrict__ eps,
                                                    const float* __restrict__ dl, const double* __restrict__ g,
                                                    double* __restrict__ out, long long n) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  const double v = vp[i], e = eps[i], d = dl[i];
  const double gh = g[i], gp = g[i + n], gn = g[i + 2 * n];
  out[i] = 2.0 * v * (gp + (1.0 + 2.0 * e) * gh + (1.0 + 2.0 * d) * gn);
  out[i + n] = 2.0 * v * v * gh;
  out[i + 2 * n] = 2.0 * v * v * gn;
}

}  // namespace fwi
