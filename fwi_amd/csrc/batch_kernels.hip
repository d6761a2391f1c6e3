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

// The same step on an interior work item: a band within [cpml, nx-cpml) and rows above nz-cpml, where no cell
// carries a C-PML term and every lane updates its cell.  The body has one path (no lane-masked memory access)
// and loads each row two rows ahead, before the current row's stores: the vector memory counter retires in
// issue order, so a load issued after a store could only be waited for together with that store.  The
// stored values are those of k_fwd_main, expression by expression (its C-PML terms are +0 here).
// SHM: the four waves of the workgroup are four shots on the same model tile (same model columns and rows, the
// same row count): per row, wave w < 3 loads the row of model field w (v_hor, v_po, v_nmo) and the waves share
// the three rows through LDS (double-buffered by row parity, one workgroup barrier per row), so the model is read
// once per four shots.
template <int QM, class S, bool INC, bool SHM = false>
__device__ __forceinline__ void fwd_interior(const FwdArgs& a, const int4 itm, const ShotDev& sd, int lane,
                                             float* buf, float (*mshare)[4][64] = nullptr) {
  const int nx = sd.nx, ps = sd.ps, nxm = a.m.nxm;
  const int ix = itm.y + lane;
  const int hx = lane < 8 ? clampi(itm.y + halo_col(lane), 0, nx - 1) : ix;
  const float dh2 = a.g.dh2, C = a.C;
  const float* P = a.Pc + sd.off;
  const float* F = a.Fc + sd.off;
  float* Po = a.Po + sd.off;
  float* Fo = a.Fo + sd.off;
  const float* Qc = INC ? a.Qc + sd.off : nullptr;
  const float* Rc = INC ? a.Rc + sd.off : nullptr;
  float* Qo = INC ? a.Qo + sd.off : nullptr;
  float* Ro = INC ? a.Ro + sd.off : nullptr;
  const float* vh = a.m.vh + sd.x0;
  const float* vp = a.m.vp + sd.x0;
  const float* vn = a.m.vn + sd.x0;
  const int wave = SHM ? wave_id() : 0;
  const float* vmine = wave == 0 ? vh : wave == 1 ? vp : vn;  // SHM: the model field this wave loads (wave 3: a copy)
  const float srcv = a.src[a.it];
  const int z0 = itm.z, z1 = itm.w;
  float fq[9];  // F along z (rows z-4..z+4)
  struct Row {
    float f4;                // F at row z+4
    float c, h;              // P at row z: the lane's column, the halo column
    float pold, fold;        // two-level form: P^{n-1}, F^{n-1}; increment form: Q^n, R^n
    float vhv, vpv, vnv;     // model at row z (SHM: vhv holds this wave's field)
  };
  auto load_row = [&](int z, Row& r) __attribute__((always_inline)) {
    const unsigned o = (unsigned)(ix + ps * z);
    const unsigned mo = (unsigned)(ix + nxm * z);
    r.f4 = ldu(F, o + 4u * ps);
    r.c = ldu(P, o);
    r.h = ldu(P, (unsigned)(hx + ps * z));
    r.pold = INC ? ldn(Qc, o) : ldn(Po, o);
    r.fold = INC ? ldn(Rc, o) : ldn(Fo, o);
    if (SHM) {
      r.vhv = ldu(vmine, mo);
    } else {
      r.vhv = ldu(vh, mo);
      r.vpv = ldu(vp, mo);
      r.vnv = ldu(vn, mo);
    }
  };
  auto step = [&](int z, const Row& cur, Row& nxt) __attribute__((always_inline)) {
    const unsigned o = (unsigned)(ix + ps * z);
    fq[8] = cur.f4;
    float vhv = cur.vhv, vpv = cur.vpv, vnv = cur.vnv;
    if (SHM) {
      // a wave can only overwrite this buffer two rows later, after every wave has passed the next row's barrier
      float(*ms)[64] = mshare[z & 1];
      ms[wave][lane] = cur.vhv;
      __syncthreads();
      vhv = ms[0][lane];
      vpv = ms[1][lane];
      vnv = ms[2][lane];
    }
    float u[9];
    xrow9(buf, lane, cur.c, cur.h, u);
    load_row(min(z + 2, z1 - 1), nxt);  // two rows ahead; unused repeats on the last rows keep one path
    const float p2 = d2s<S>(u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8], dh2);
    const float r2 = d2s<S>(fq[0], fq[1], fq[2], fq[3], fq[4], fq[5], fq[6], fq[7], fq[8], dh2);
    const float atp = 0.f, atf = 0.f;
    const float tp = dh2 * (p2 + atp);
    const float tf = dh2 * (r2 + atf);
    float pn, fn;
    if (INC) {
      float dq = vhv * C * tp + vpv * C * tf;
      float dr = vnv * C * tp + vpv * C * tf;
      if (ix == sd.isx && z == sd.isz) {
        dq += srcv;
        dr += srcv;
      }
      const float qn = cur.pold + dq, rn = cur.fold + dr;
      pn = cur.c + qn;
      fn = fq[4] + rn;
      stn(Qo, o, qn);
      stn(Ro, o, rn);
    } else {
      pn = 2.f * cur.c - cur.pold + vhv * C * tp + vpv * C * tf;
      fn = 2.f * fq[4] - cur.fold + vnv * C * tp + vpv * C * tf;
      if (ix == sd.isx && z == sd.isz) {
        pn += srcv;
        fn += srcv;
      }
    }
    stn(Po, o, pn);
    stn(Fo, o, fn);
    if (QM == 1) {
      stn(a.qa + sd.off, o, p2 + atp);
      stn(a.qb + sd.off, o, r2 + atf);
    } else if (QM == 2) {
      stn(a.qa + sd.off, o, p2);
      stn(a.qb + sd.off, o, r2);
    }
    if (a.gather && z == sd.irz && ix >= sd.ir0 && ix < sd.ir0 + sd.nrec)
      a.gather[sd.goff + (long)a.it * sd.nrec + (ix - sd.ir0)] = pn + fn;
#pragma unroll
    for (int k = 0; k < 8; ++k) fq[k] = fq[k + 1];
  };
#pragma unroll
  for (int k = 0; k < 8; ++k) fq[k] = ldu(F, (unsigned)(ix + ps * (z0 - 4 + k)));
  // Two rows are in flight: each step loads the row two ahead into the buffer its predecessor consumed.  Three
  // buffers in rotation, unrolled by three, need no register copy (a copy of a value still being loaded would
  // wait for it); the first three rows are peeled, so the loop is entered, as it is repeated, after a row's
  // stores that follow the loads of the rows ahead, and the memory-counter waits at its head are the same on
  // both paths.  The last one or two rows follow the loop.
  Row ra, rb, rc;
  load_row(z0, ra);
  load_row(min(z0 + 1, z1 - 1), rb);
  int z = z0;
  if (z + 2 < z1) {
    step(z, ra, rc);
    step(z + 1, rb, ra);
    step(z + 2, rc, rb);
    for (z += 3; z + 2 < z1; z += 3) {
      step(z, ra, rc);
      step(z + 1, rb, ra);
      step(z + 2, rc, rb);
    }
  }
  if (z < z1) step(z, ra, rc);
  if (z + 1 < z1) step(z + 1, rb, ra);
}

// INC: increment form of the forward leapfrog.  The state is (P^n, Q^n = P^n - P^{n-1}) and the update is
// Q^{n+1} = Q^n + d, P^{n+1} = P^n + Q^{n+1}, which equals P^{n+1} = 2P^n - P^{n-1} + d in exact
// arithmetic.  In FP32 the stored Q keeps its own relative precision, whereas the difference of two
// stored levels loses it: with s samples per period the rounding of the two-level form is amplified
// about (s/2pi)^2 times (about 600 at 8 Hz, dt = 0.8 ms), that of the increment form about s/2pi.
// Work items: (shot, band x0, z0, z1); bit kShareModel of the shot field marks the four items of a workgroup as
// four interior items on one model tile (the host sets it on all four or none), which share the model rows.
constexpr int kShareModel = 1 << 30;

// The general body: any item, with the C-PML terms of every cell (parallelcalls.cu:48-117); with a.xr, only the
// item's column range is updated.  The C-PML memory variables psi of the step are computed in the item, as PSI_F
// computes them (parallelcalls.cu:48-63), from psi^n (a.psx, a.psz) and P^n, F^n, and psi^{n+1} is stored on the
// item's own C-PML cells (a.psxn, a.pszn): psi_x^{n+1} on the item's 64 columns and the 4 beyond each side, from an
// exchange of P^n over 80 columns; psi_z^{n+1} 4 rows ahead of the current row, from F^n in a 13-row queue (items
// that reach within 4 rows of the bottom C-PML).  Items that meet compute each other's border psi^{n+1} from the
// same inputs, so every cell reads the same bits, and psi^n and psi^{n+1} are separate arrays, so no item reads a
// value another one has already advanced.  A cell off the C-PML strips keeps psi = 0.
// buf: this wave's 152 LDS floats (80 for P^n, 72 for psi_x^{n+1}).
template <int QM, class S, bool INC, bool DEEPZ>
__device__ __forceinline__ void fwd_general_rows(const FwdArgs& a, int item, const int4 itm, const ShotDev& sd,
                                                 int lane, float* buf) {
  const int nx = sd.nx, ps = sd.ps, nz = a.g.nz, cpml = a.g.cpml, nxm = a.m.nxm;
  const int x0 = itm.y, ix = x0 + lane;
  bool act = ix >= kRim && ix < nx - kRim;
  if (a.xr) {  // shell items of the step-pair scheme update only their column range
    const int2 r = a.xr[item];
    act = act && ix >= r.x && ix < r.y;
  }
  // every lane loads from a valid column; values of columns outside the shot are never consumed
  const int cx = clampi(ix, 0, nx - 1);
  // P^n on columns x0-8 .. x0+71 (slot = column - x0 + 8): each lane its column, lanes 0..15 one more
  const int hx = lane < 8 ? clampi(x0 - 8 + lane, 0, nx - 1) : lane < 16 ? clampi(x0 + 56 + lane, 0, nx - 1) : cx;
  const int hslot = lane < 8 ? lane : lane < 16 ? 64 + lane : 8 + lane;
  // psi_x^{n+1} on columns x0-4 .. x0+67: each lane its column, lanes 0..7 one more (lanes 8.. their own again)
  const int pcol = lane < 4 ? x0 - 4 + lane : lane < 8 ? x0 + 60 + lane : ix;
  const int pslot = pcol - x0 + 8;
  const int pcx = clampi(pcol, 0, nx - 1);
  const int hidx9 = xrow9_halo_slot(lane);
  auto xs_col = [&](int c) { return (c >= kRim && c < cpml) || (c >= nx - cpml && c < nx - kRim); };
  const bool xs_own = xs_col(ix), xs_h = xs_col(pcol);
  const float pax = xs_own ? coef_x(a.g.ta, ix, nx, cpml) : 0.f, pbx = xs_own ? coef_x(a.g.tb, ix, nx, cpml) : 0.f;
  const float hax = xs_h ? coef_x(a.g.ta, pcol, nx, cpml) : 0.f, hbx = xs_h ? coef_x(a.g.tb, pcol, nx, cpml) : 0.f;
  // a side C-PML cell within 4 columns of the item: psi_x enters (wave-uniform); otherwise psi_x is 0 around it
  const bool need_px = x0 - 4 < cpml || x0 + 68 > nx - cpml;
  const bool zcol = ix >= kRim && ix < nx - kRim;  // psi_z is updated on the bottom strip rows of these columns
  const bool xstrip = ix < cpml || ix > nx - cpml - 1;
  const float axv = act ? coef_x(a.g.ta, ix, nx, cpml) : 0.f;
  const float bxv = act ? coef_x(a.g.tb, ix, nx, cpml) : 0.f;
  const float dh = a.g.dh, dh2 = a.g.dh2, C = a.C;
  const float* P = a.Pc + sd.off;
  const float* F = a.Fc + sd.off;
  float* Po = a.Po + sd.off;
  float* Fo = a.Fo + sd.off;
  const float* psx = a.psx + sd.off;
  const float* psz = a.psz + sd.off;
  float* psxn = a.psxn + sd.off;
  float* pszn = a.pszn + sd.off;
  float* zx = a.zx + sd.off;
  float* zz = a.zz + sd.off;
  const float* vh = a.m.vh + sd.x0;
  const float* vp = a.m.vp + sd.x0;
  const float* vn = a.m.vn + sd.x0;
  const float srcv = a.src[a.it];
  const int z0 = itm.z, z1 = itm.w;
  auto zs_row = [&](int r) { return zcol && r >= nz - cpml && r < nz - kRim; };
  // psi_z^{n+1} of row r from F^n at rows r-4..r+4 (f[0..8])
  auto psiz = [&](int r, const float* f) -> float {
    if (!zs_row(r)) return 0.f;
    return coef_z(a.g.tb, r, nz, cpml) * ldu(psz, (unsigned)(cx + ps * r)) +
           coef_z(a.g.ta, r, nz, cpml) * d1s<S>(f[0], f[1], f[2], f[3], f[5], f[6], f[7], f[8], dh);
  };
  constexpr int NF = DEEPZ ? 13 : 9;
  float fq[NF];  // F^n, rows z-4 .. z+4 (+4 more with DEEPZ)
  float zq[9];   // DEEPZ: psi_z^{n+1}, rows z-4 .. z+4
  if (DEEPZ) {
    float f16[16];  // rows z0-8 .. z0+7 (rows outside the grid feed no C-PML row)
#pragma unroll
    for (int k = 0; k < 16; ++k) f16[k] = ldu(F, (unsigned)(cx + ps * clampi(z0 - 8 + k, 0, nz - 1)));
#pragma unroll
    for (int j = 0; j < 8; ++j) zq[j] = psiz(z0 - 4 + j, f16 + j);
#pragma unroll
    for (int k = 0; k < 12; ++k) fq[k] = f16[k + 4];
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) fq[k] = ldu(F, (unsigned)(cx + ps * (z0 - 4 + k)));
  }
  for (int z = z0; z < z1; ++z) {
    const unsigned o = (unsigned)(cx + ps * z);
    const unsigned mo = (unsigned)(cx + nxm * z);
    fq[NF - 1] = ldu(F, (unsigned)(cx + ps * min(z + NF - 5, nz - 1)));
    const float c = ldu(P, o);
    const float h = ldu(P, (unsigned)(hx + ps * z));
    // two-level form: P^{n-1}, F^{n-1}; increment form: Q^n, R^n
    const float pold = INC ? ldn(a.Qc + sd.off, o) : ldn(Po, o);
    const float fold = INC ? ldn(a.Rc + sd.off, o) : ldn(Fo, o);
    const float vhv = ldu(vh, mo), vpv = ldu(vp, mo), vnv = ldu(vn, mo);
    float tz = 0.f;
    if (DEEPZ) {
      zq[8] = psiz(z + 4, fq + 4);
      tz = d1s<S>(zq[0], zq[1], zq[2], zq[3], zq[5], zq[6], zq[7], zq[8], dh);
    }
    float u[9];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    buf[8 + lane] = c;
    buf[hslot] = h;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int k = 0; k < 9; ++k) u[k] = buf[lane + 4 + k];
    float tx = 0.f, pxo = 0.f;
    if (need_px) {
      float v[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) v[k] = buf[pslot - 4 + k];
      pxo = xs_own ? pbx * ldu(psx, o) + pax * d1s<S>(u[0], u[1], u[2], u[3], u[5], u[6], u[7], u[8], dh) : 0.f;
      const float pxh = xs_h ? hbx * ldu(psx, (unsigned)(pcx + ps * z)) +
                                   hax * d1s<S>(v[0], v[1], v[2], v[3], v[5], v[6], v[7], v[8], dh)
                             : 0.f;
      float w[9];
      xrow9u(buf + 80, lane, hidx9, pxo, pxh, w);
      tx = d1s<S>(w[0], w[1], w[2], w[3], w[5], w[6], w[7], w[8], dh);
    }
    if (act) {
      const float p2 = d2s<S>(u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8], dh2);
      const float r2 = d2s<S>(fq[0], fq[1], fq[2], fq[3], fq[4], fq[5], fq[6], fq[7], fq[8], dh2);
      float atp = 0.f, atf = 0.f;
      if (xstrip || z > nz - cpml - 1) {
        const float tpx = p2 + tx;
        const float tfz = r2 + tz;
        const float azv = coef_z(a.g.ta, z, nz, cpml), bzv = coef_z(a.g.tb, z, nz, cpml);
        const float zxn = bxv * ldn(zx, o) + axv * tpx;
        const float zzn = bzv * ldn(zz, o) + azv * tfz;
        stn(zx, o, zxn);
        stn(zz, o, zzn);
        atp = tx + zxn;
        atf = tz + zzn;
      }
      if (xs_own) stu(psxn, o, pxo);
      if (DEEPZ && zs_row(z)) stu(pszn, o, zq[4]);
      const float tp = dh2 * (p2 + atp);
      const float tf = dh2 * (r2 + atf);
      float pn, fn;
      if (INC) {
        float dq = vhv * C * tp + vpv * C * tf;
        float dr = vnv * C * tp + vpv * C * tf;
        if (ix == sd.isx && z == sd.isz) {
          dq += srcv;
          dr += srcv;
        }
        const float qn = pold + dq, rn = fold + dr;
        pn = c + qn;
        fn = fq[4] + rn;
        stn(a.Qo + sd.off, o, qn);
        stn(a.Ro + sd.off, o, rn);
      } else {
        pn = 2.f * c - pold + vhv * C * tp + vpv * C * tf;
        fn = 2.f * fq[4] - fold + vnv * C * tp + vpv * C * tf;
        if (ix == sd.isx && z == sd.isz) {
          pn += srcv;
          fn += srcv;
        }
      }
      stn(Po, o, pn);
      stn(Fo, o, fn);
      if (a.gather && z == sd.irz && ix >= sd.ir0 && ix < sd.ir0 + sd.nrec)
        a.gather[sd.goff + (long)a.it * sd.nrec + (ix - sd.ir0)] = pn + fn;
      if (QM == 1) {
        stn(a.qa + sd.off, o, p2 + atp);
        stn(a.qb + sd.off, o, r2 + atf);
      } else if (QM == 2) {
        stn(a.qa + sd.off, o, p2);
        stn(a.qb + sd.off, o, r2);
      }
    }
#pragma unroll
    for (int k = 0; k < NF - 1; ++k) fq[k] = fq[k + 1];
    if (DEEPZ) {
#pragma unroll
      for (int k = 0; k < 8; ++k) zq[k] = zq[k + 1];
    }
  }
}
template <int QM, class S, bool INC>
__device__ __forceinline__ void fwd_general(const FwdArgs& a, int item, const int4 itm, const ShotDev& sd, int lane,
                                            float* buf) {
  if (itm.w > a.g.nz - a.g.cpml - kRim)  // within 4 rows of the bottom C-PML: psi_z enters
    fwd_general_rows<QM, S, INC, true>(a, item, itm, sd, lane, buf);
  else
    fwd_general_rows<QM, S, INC, false>(a, item, itm, sd, lane, buf);
}

// ---------------------------------------------------------------------------------------------
// FP64 near-surface band of the increment-form forward (this project's arithmetic): rows [4, band) of every
// shot, where the source, the receivers and the largest near-source fields lie, advance in FP64 with the
// two-level leapfrog (P^{n+1} = 2P^n - P^{n-1} + ...; in FP64 the time difference keeps its precision), from the
// float model, C-PML table, wavelet and dt promoted exactly.  The band lies above the bottom C-PML, so only the
// side C-PML cells carry terms (psi_x, zeta_x).  F couples band and float rows through the z stencil: rows
// band..band+3 of the float F^n enter the band's D2z, and the band writes its rows of the float P, F rounded, which
// the float rows' D2z reads.  Also written: the band's imaging terms (rounded once) and its receivers (P + F
// rounded once).  The expressions are those of the oracle's f32i band, in the same order.
// Grid (256-column chunks, groups of kBandRows band rows, shots): a thread updates one column of the group's rows,
// taking the x stencil of P from the rows staged in LDS and the z stencil of F from a register queue of the
// column (kBandRows + 8 loads instead of 9 per row).  psi_x is kept by step parity (read psx, write psxn): a chunk
// first computes psi_x^{n+1} of the side cells within 4 columns of it -- neighbouring chunks compute the cells
// they share from the same inputs and store the same bits -- then, after a workgroup barrier, its own cells.
// Measured per strip + band launch on survey1600: one row per workgroup 57.7 us, 2 rows 53.8, 3 rows 53.1, 4 rows
// 54.2, 5 rows 55.7, 10 rows 104 (LDS and registers limit the waves); a tile of 64 x 20 cells staged in LDS was
// slower still.
constexpr int kBandRows = 4;  // band rows per workgroup
template <int QM>
__device__ __forceinline__ void fwd_band_block(const BandArgs& a, int chunk, int rgroup, int shot) {
  constexpr int R = kBandRows;
  __shared__ double sp[R][264];  // P^n of the rows, columns x0-4 .. x0+259
  __shared__ double sx[R][264];  // psi_x^{n+1} of the side C-PML cells among them, 0 elsewhere (as in psxn)
  const ShotDev sd = a.g.shots[shot];
  const int nx = sd.nx, ps = sd.ps, cpml = a.g.cpml, nxm = a.m.nxm, B = a.band;
  const int x0 = chunk * 256;
  if (x0 >= nx - kRim) return;  // the whole workgroup: past this shot's window
  const int z0 = kRim + rgroup * R;
  const int nr = B - z0 < R ? B - z0 : R;  // rows of this group inside the band
  const double dh = (double)a.g.dh, dh2 = dh * dh, Cd = (a.dt * a.dt) / dh2;
  // per-shot bases (wave-uniform) and 32-bit element offsets
  const double* P = a.P + sd.boff;
  double* Pm = a.Pm + sd.boff;
  const double* Fb = a.F + sd.boff;
  double* Fm = a.Fm + sd.boff;
  const double* psx = a.psx + sd.boff;
  double* psxn = a.psxn + sd.boff;
  double* zx = a.zx + sd.boff;
  const float* Fc = a.Fc + sd.off;
  const int w = cpml - kRim;  // side C-PML cells per row and side
  for (int t = threadIdx.x; t < 264; t += 256) {
    const int c = x0 - 4 + t;
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      sp[rr][t] = (rr < nr && c >= 0 && c < nx) ? ldud(P, (unsigned)(ps * (z0 + rr) + c)) : 0.0;  // else never used
      sx[rr][t] = 0.0;
    }
  }
  const bool sides = x0 - 4 < cpml || x0 + 260 > nx - cpml;  // workgroup-uniform
  if (sides) {
    __syncthreads();
    // psi_x^{n+1} of the side C-PML cells in [x0-4, x0+260) (PSI_F); neighbouring chunks compute the cells they
    // share from the same inputs and store the same bits
    for (int j = threadIdx.x; j < 2 * w * nr; j += 256) {
      const int rr = j / (2 * w), jj = j - rr * 2 * w;
      const int ix = jj < w ? kRim + jj : nx - cpml + (jj - w);
      if (ix < x0 - 4 || ix >= x0 + 260) continue;
      const unsigned i = (unsigned)(ps * (z0 + rr) + ix);
      const double t = d1d(ldud(P, i - 4), ldud(P, i - 3), ldud(P, i - 2), ldud(P, i - 1), ldud(P, i + 1),
                           ldud(P, i + 2), ldud(P, i + 3), ldud(P, i + 4), dh);
      const double v = (double)coef_x(a.g.tb, ix, nx, cpml) * ldud(psx, i) + (double)coef_x(a.g.ta, ix, nx, cpml) * t;
      stud(psxn, i, v);
      sx[rr][ix - (x0 - 4)] = v;
    }
  }
  __syncthreads();
  const int ix = x0 + threadIdx.x;
  if (ix < kRim || ix >= nx - kRim) return;
  // F^n at rows z0-4 .. z0+R+3 of this column: FP64 in the band, the float rows below it
  double f[R + 8];
#pragma unroll
  for (int k = 0; k < R + 8; ++k) {
    const int r = z0 + k - 4;
    const unsigned j = (unsigned)(ps * r + ix);
    f[k] = (k < nr + 8) ? (r < B ? ldud(Fb, j) : (double)ldu(Fc, j)) : 0.0;
  }
  const bool xside = ix < cpml || ix > nx - cpml - 1;
  const float axv = coef_x(a.g.ta, ix, nx, cpml), bxv = coef_x(a.g.tb, ix, nx, cpml);
  const int t = threadIdx.x + 4;
#pragma unroll
  for (int rr = 0; rr < R; ++rr) {
    if (rr >= nr) break;
    const int z = z0 + rr;
    const unsigned i = (unsigned)(ps * z + ix);
    const double pc = sp[rr][t];
    const double p2 = d2d(sp[rr][t - 4], sp[rr][t - 3], sp[rr][t - 2], sp[rr][t - 1], pc, sp[rr][t + 1],
                          sp[rr][t + 2], sp[rr][t + 3], sp[rr][t + 4], dh2);
    const double r2 = d2d(f[rr], f[rr + 1], f[rr + 2], f[rr + 3], f[rr + 4], f[rr + 5], f[rr + 6], f[rr + 7],
                          f[rr + 8], dh2);
    double atp = 0.0;
    const double atf = 0.0;
    if (xside) {
      const double tx = d1d(sx[rr][t - 4], sx[rr][t - 3], sx[rr][t - 2], sx[rr][t - 1], sx[rr][t + 1], sx[rr][t + 2],
                            sx[rr][t + 3], sx[rr][t + 4], dh);
      const double tpx = p2 + tx;
      const double zn = (double)bxv * ldud(zx, i) + (double)axv * tpx;
      stud(zx, i, zn);
      atp = tx + zn;
    }
    const double tp = dh2 * (p2 + atp);
    const double tf = dh2 * (r2 + atf);
    const unsigned mi = (unsigned)(ix + nxm * z);
    const double vh = (double)ldu(a.m.vh + sd.x0, mi), vp = (double)ldu(a.m.vp + sd.x0, mi),
                 vn = (double)ldu(a.m.vn + sd.x0, mi);
    double pn = 2.0 * pc - ldud(Pm, i) + vh * Cd * tp + vp * Cd * tf;
    double fn = 2.0 * f[rr + 4] - ldud(Fm, i) + vn * Cd * tp + vp * Cd * tf;
    if (ix == sd.isx && z == sd.isz) {
      const double srcv = (double)a.src[a.it];
      pn += srcv;
      fn += srcv;
    }
    stud(Pm, i, pn);
    stud(Fm, i, fn);
    stu(a.Po + sd.off, i, (float)pn);
    stu(a.Fo + sd.off, i, (float)fn);
    if (QM == 1) {
      stu(a.qa + sd.off, i, (float)(p2 + atp));
      stu(a.qb + sd.off, i, (float)(r2 + atf));
    } else if (QM == 2) {  // the reference-compatible terms come from the float state, as in the float rows
      const float* Pf = a.Pcf + sd.off;
      a.qa[sd.off + i] = d2_at<float>(Pf, i, 1, a.g.dh2);
      a.qb[sd.off + i] = d2_at<float>(Fc, i, ps, a.g.dh2);
    }
    if (a.gather && z == sd.irz && ix >= sd.ir0 && ix < sd.ir0 + sd.nrec)
      a.gather[sd.goff + (long)a.it * sd.nrec + (ix - sd.ir0)] = (float)(pn + fn);
  }
}
// One band workgroup of a forward launch: workgroups [0, a.nband_wg) are (chunk, band row group, shot) in that order,
// nbc chunks per row, dealt to the XCDs in contiguous ranges of their own (the rows a z stencil reads come from
// workgroups on the same XCD); the last nband_wg - nband_real do nothing.
template <int QM>
__device__ __forceinline__ void fwd_band_wg(const FwdArgs& a) {
  int b = xcd_logical_block(blockIdx.x, a.nband_wg);
  if (b >= a.nband_real) return;
  const int groups = (a.band.band - kRim + kBandRows - 1) / kBandRows;
  const int chunk = b % a.nbc;
  b /= a.nbc;
  fwd_band_block<QM>(a.band, chunk, b % groups, b / groups);
}

// One forward step of every shot (parallelcalls.cu:120-130 SecondDerivate_P, :48-86 PSI_F / ZETA_F, :88-117
// propagator_F) in one launch: the FP64 near-surface band's workgroups first (increment form), then the items --
// interior items by the pipelined bodies, items with C-PML cells by the general body.  QM selects what is kept for
// the adjoint pass: 0 nothing, 1 the model-multiplied terms a = p2 + aten_p, b = r2 + aten_f (exact gradient), 2 p2,
// r2 (reference-compatible gradient).
template <int QM, class S, bool INC = false>
__global__ void __launch_bounds__(256) k_fwd_main(FwdArgs a) {
  __shared__ float xbuf[4][152];
  __shared__ float mshare[2][4][64];
  if ((int)blockIdx.x < a.nband_wg) {
    if constexpr (INC) fwd_band_wg<QM>(a);
    return;
  }
  const int wave = wave_id(), lane = threadIdx.x & 63;
  const int item = xcd_logical_block((int)blockIdx.x - a.nband_wg, (int)gridDim.x - a.nband_wg) * 4 + wave;
  if (item >= a.nitems) return;
  int4 itm = a.items[item];
  const bool share = (itm.x & kShareModel) != 0;
  itm.x &= ~kShareModel;
  const ShotDev sd = a.g.shots[itm.x];
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml;
  // Interior items (no C-PML cell; the items are cut at row nz-e, e = max(cpml,4)) take the pipelined body.
  // Items with C-PML cells take the general body, interleaved with them in the launch.
  if (share) {
    fwd_interior<QM, S, INC, true>(a, itm, sd, lane, xbuf[wave], mshare);
    return;
  }
  const int e = cpml > kRim ? cpml : kRim;
  if (!a.xr && itm.y >= e && itm.y + 64 <= nx - e && itm.w <= nz - e) {
    fwd_interior<QM, S, INC>(a, itm, sd, lane, xbuf[wave]);
    return;
  }
  fwd_general<QM, S, INC>(a, item, itm, sd, lane, xbuf[wave]);
}

// ---------------------------------------------------------------------------------------------
// Two increment-form forward steps n, n+1 per launch (temporal blocking).  The core of a window -- columns
// [cpml+4, nx-cpml-4) rounded inward to whole lines and rows [zfirst+4, nz-cpml-4), zfirst the first row of the
// float state (below the FP64 band) -- advances two steps reading only the state of step n: neither step, nor the
// halo of step n that step n+1 reads, touches a C-PML cell or the band.  A core workgroup owns up to 224 output columns [xt, xt+224) and a run of rows
// [z0, z1); its lanes compute step n on columns [xt-4, xt+228) and rows [z0-4, z1+4), marching down: at row r
// they finish step n at row r and step n+1 at row r-4.  D2x acts on P only and D2z on F only: step n+1 takes the x
// neighbours of P^{n+1} from a six-row LDS ring shared by the four waves (one workgroup barrier per row), the z
// neighbours of F^{n+1} from a register queue; Q^{n+1}, R^{n+1}, the model and the stored terms of step n wait four
// rows in register delay lines.  Per output cell the four state fields cross HBM once for two steps.  Storing, both
// steps' imaging terms are written.  Stores go through buffer resources with an out-of-range offset on the lanes that
// do not own the cell, so the loop body has one path and its memory-counter waits cover only the oldest prefetched
// row; four row buffers in rotation (two rows in flight) and the four-row delay lines give the unrolled body one
// register assignment.  Blocks [0, nshell_wg) of the same launch run step n of the shell and of the core's cells
// within 4 of it with the single-step general body (they read step n and write the set of step n+1 only).  Every
// cell is computed with the single-step kernels' expressions in the same order, so the results are identical to
// single steps.
// ---------------------------------------------------------------------------------------------
// Output columns of a core workgroup: 224 = 7 cache lines, so that, with the core's columns rounded inward to whole
// 128-byte lines, every line a core workgroup writes is written whole by it (a line shared by two workgroups would
// reach HBM as two partial writes); 232 of the 256 lanes compute step n.
constexpr int kPairCols = 224;
// The core's columns: [cpml+4, nx-cpml-4) rounded inward to multiples of 32.
__host__ __device__ inline int pair_cx0(int cpml) { return (cpml + 4 + 31) & ~31; }
__host__ __device__ inline int pair_cx1(int nx, int cpml) { return (nx - cpml - 4) & ~31; }

template <int QM>
__device__ __forceinline__ void fwd_pair_core(const FwdPairArgs& a, const int4 itm, float* xb, float (*ring)[264]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
  const ShotDev sd = a.g.shots[itm.x];
  const int nx = sd.nx, ps = sd.ps, nxm = a.m.nxm;
  const int cxb = pair_cx1(nx, a.g.cpml);    // end of the core columns
  const int xt = itm.y, z0 = itm.z, z1 = itm.w;
  const int col = xt - 4 + tid;              // step-n column of this lane (>= cpml)
  const int cc = col < nx ? col : nx - 1;    // its loads: a valid column (lanes past cxb+4 feed no output)
  const int hx = lane < 8 ? clampi(xt - 4 + 64 * wave + halo_col(lane), 0, nx - 1) : cc;
  const int hidx = xrow9_halo_slot(lane);
  const bool own = tid >= 4 && tid < 4 + kPairCols && col < cxb;
  const bool rec = a.gather != nullptr && col >= sd.ir0 && col < sd.ir0 + sd.nrec;
  const float dh2 = a.g.dh2, C = a.C;
  const float src0 = a.src[a.it], src1 = a.src[a.it + 1];
  const float* P0 = a.P0 + sd.off;
  const float* F0 = a.F0 + sd.off;
  const float* Q0 = a.Q0 + sd.off;
  const float* R0 = a.R0 + sd.off;
  const float* vh = a.m.vh + sd.x0;
  const float* vp = a.m.vp + sd.x0;
  const float* vn = a.m.vn + sd.x0;
  const unsigned fbytes = (unsigned)ps * (unsigned)a.g.nz * 4u;
  const brsrc sP2 = make_rsrc(a.P2 + sd.off, fbytes), sF2 = make_rsrc(a.F2 + sd.off, fbytes);
  const brsrc sQ2 = make_rsrc(a.Q2 + sd.off, fbytes), sR2 = make_rsrc(a.R2 + sd.off, fbytes);
  // the four imaging-term fields of the pair lie qstride bytes apart (a, b of step n, then of step n+1): one
  // resource over all four (QM != 0; otherwise an empty one that is never used)
  const unsigned qs = a.qstride;
  const brsrc sq = make_rsrc(QM != 0 ? a.qa0 + sd.off : a.P2, QM != 0 ? 3u * qs + fbytes : 0u);
  const int zend = z1 + 4;  // step n runs on rows [z0-4, zend)
  struct Row {
    float f4;          // F^n at row r+4
    float c, h;        // P^n at row r: the lane's column, the halo column
    float q, r;        // Q^n, R^n
    float vh, vp, vn;  // model
  };
  auto load_row = [&](int z, Row& w) __attribute__((always_inline)) {
    const unsigned o = (unsigned)(cc + ps * z);
    const unsigned mo = (unsigned)(cc + nxm * z);
    w.f4 = ldu(F0, o + 4u * ps);
    w.c = ldu(P0, o);
    w.h = ldu(P0, (unsigned)(hx + ps * z));
    w.q = ldn(Q0, o);
    w.r = ldn(R0, o);
    w.vh = ldu(vh, mo);
    w.vp = ldu(vp, mo);
    w.vn = ldu(vn, mo);
  };
  float fq0[9];  // F^n, rows r-4..r+4
  float fq1[9];  // F^{n+1}, rows r-8..r
  float dQ[4], dR[4], dvh[4], dvp[4], dvn[4], dta[4], dtb[4];  // rows r-4..r-1
  // MAIN: step n+1 at row r-4 and the stores of that row (rows r >= z0+4); otherwise step n only
  auto iter = [&](int r, const Row& cur, Row& nxt, auto mainc) __attribute__((always_inline)) {
    constexpr bool MAIN = decltype(mainc)::value;
    fq0[8] = cur.f4;
    float u[9];
    xrow9u(xb, lane, hidx, cur.c, cur.h, u);
    load_row(min(r + 2, zend - 1), nxt);  // two rows ahead; unused repeats on the last rows keep one path
    // step n at row r (fwd_interior's expressions; no C-PML term in the core and its halo)
    const float p2 = d2s<float>(u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8], dh2);
    const float r2 = d2s<float>(fq0[0], fq0[1], fq0[2], fq0[3], fq0[4], fq0[5], fq0[6], fq0[7], fq0[8], dh2);
    const float atp = 0.f, atf = 0.f;
    const float tp = dh2 * (p2 + atp);
    const float tf = dh2 * (r2 + atf);
    float dq = cur.vh * C * tp + cur.vp * C * tf;
    float dr = cur.vn * C * tp + cur.vp * C * tf;
    if (col == sd.isx && r == sd.isz) {
      dq += src0;
      dr += src0;
    }
    const float q1 = cur.q + dq, r1 = cur.r + dr;
    const float p1 = cur.c + q1, f1 = fq0[4] + r1;
    ring[r % 6][tid + 4] = p1;
    fq1[8] = f1;
    if (MAIN) {
      // the ring slot written above last held row r-6, which every wave read before the previous row's barrier
      __syncthreads();
      const int rb = r - 4;
      const float* rw = &ring[rb % 6][tid];
      float w[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) w[k] = rw[k];
      const float p2b = d2s<float>(w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8], dh2);
      const float r2b = d2s<float>(fq1[0], fq1[1], fq1[2], fq1[3], fq1[4], fq1[5], fq1[6], fq1[7], fq1[8], dh2);
      const float tpb = dh2 * (p2b + atp);
      const float tfb = dh2 * (r2b + atf);
      float dq2 = dvh[0] * C * tpb + dvp[0] * C * tfb;
      float dr2 = dvn[0] * C * tpb + dvp[0] * C * tfb;
      if (col == sd.isx && rb == sd.isz) {
        dq2 += src1;
        dr2 += src1;
      }
      const float q2 = dQ[0] + dq2, r2n = dR[0] + dr2;
      const float pn = w[4] + q2, fn = fq1[4] + r2n;
      const unsigned ob = own ? (unsigned)(col + ps * rb) * 4u : kOutOfRange;
      bst<kBufNT>(sP2, ob, pn);
      bst<kBufNT>(sF2, ob, fn);
      bst<kBufNT>(sQ2, ob, q2);
      bst<kBufNT>(sR2, ob, r2n);
      if (QM == 1) {
        bst<kBufNT>(sq, ob, dta[0]);
        bst<kBufNT>(sq, ob + qs, dtb[0]);
        bst<kBufNT>(sq, ob + 2u * qs, p2b + atp);
        bst<kBufNT>(sq, ob + 3u * qs, r2b + atf);
      } else if (QM == 2) {
        bst<kBufNT>(sq, ob, dta[0]);
        bst<kBufNT>(sq, ob + qs, dtb[0]);
        bst<kBufNT>(sq, ob + 2u * qs, p2b);
        bst<kBufNT>(sq, ob + 3u * qs, r2b);
      }
      if (rb == sd.irz && rec && own) {
        const long g = sd.goff + (long)a.it * sd.nrec + (col - sd.ir0);
        a.gather[g] = w[4] + fq1[4];
        a.gather[g + sd.nrec] = pn + fn;
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      fq0[k] = fq0[k + 1];
      fq1[k] = fq1[k + 1];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dQ[k] = dQ[k + 1];
      dR[k] = dR[k + 1];
      dvh[k] = dvh[k + 1];
      dvp[k] = dvp[k + 1];
      dvn[k] = dvn[k + 1];
      dta[k] = dta[k + 1];
      dtb[k] = dtb[k + 1];
    }
    dQ[3] = q1;
    dR[3] = r1;
    dvh[3] = cur.vh;
    dvp[3] = cur.vp;
    dvn[3] = cur.vn;
    dta[3] = QM == 1 ? p2 + atp : p2;
    dtb[3] = QM == 1 ? r2 + atf : r2;
  };
  using Pro = std::integral_constant<bool, false>;
  using Main = std::integral_constant<bool, true>;
#pragma unroll
  for (int k = 0; k < 8; ++k) fq0[k] = ldu(F0, (unsigned)(cc + ps * (z0 - 8 + k)));
  // four row buffers in rotation, two rows in flight: the 8 rows of step n above the first output row, the first
  // four main rows peeled (so the loop is entered as it repeats), the main rows unrolled by four, the last ones
  Row b0, b1, b2, b3;
  load_row(z0 - 4, b0);
  load_row(z0 - 3, b1);
  iter(z0 - 4, b0, b2, Pro{});
  iter(z0 - 3, b1, b3, Pro{});
  iter(z0 - 2, b2, b0, Pro{});
  iter(z0 - 1, b3, b1, Pro{});
  iter(z0, b0, b2, Pro{});
  iter(z0 + 1, b1, b3, Pro{});
  iter(z0 + 2, b2, b0, Pro{});
  iter(z0 + 3, b3, b1, Pro{});
  int z = z0 + 4;
  if (z + 3 < zend) {
    iter(z, b0, b2, Main{});
    iter(z + 1, b1, b3, Main{});
    iter(z + 2, b2, b0, Main{});
    iter(z + 3, b3, b1, Main{});
    for (z += 4; z + 3 < zend; z += 4) {
      iter(z, b0, b2, Main{});
      iter(z + 1, b1, b3, Main{});
      iter(z + 2, b2, b0, Main{});
      iter(z + 3, b3, b1, Main{});
    }
  }
  if (z < zend) iter(z, b0, b2, Main{});
  if (z + 1 < zend) iter(z + 1, b1, b3, Main{});
  if (z + 2 < zend) iter(z + 2, b2, b0, Main{});
}

// Grid: [0, sh.nband_wg) FP64 band workgroups of step n, [.., + nshell_wg) shell workgroups (four single-step items
// each), then one workgroup per core item.  Each part is dealt to the XCDs in contiguous logical ranges of its own
// (the first two counts are multiples of 8).
template <int QM>
__global__ void __launch_bounds__(256) k_fwd_pair(FwdPairArgs a) {
  __shared__ float xbuf[4][152];
  __shared__ float ring[6][264];
  const int wave = wave_id(), lane = threadIdx.x & 63;
  const int nb = a.sh.nband_wg;
  if ((int)blockIdx.x < nb) {
    fwd_band_wg<QM>(a.sh);
    return;
  }
  const int bid = (int)blockIdx.x - nb;
  if (bid < a.nshell_wg) {  // step n of the shell and the core's cells next to it
    const int item = xcd_logical_block(bid, a.nshell_wg) * 4 + wave;
    if (item >= a.sh.nitems) return;
    const int4 itm = a.sh.items[item];
    const ShotDev sd = a.g.shots[itm.x];
    fwd_general<QM, float, true>(a.sh, item, itm, sd, lane, xbuf[wave]);
    return;
  }
  const int b = xcd_logical_block(bid - a.nshell_wg, (int)gridDim.x - nb - a.nshell_wg);
  if (b >= a.nitems) return;
  fwd_pair_core<QM>(a, a.items[b], xbuf[wave], ring);
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
  const int2 blk = a.blocks[xcd_logical_block(blockIdx.x, gridDim.x)];
  const ShotDev sd = a.g.shots[blk.x];
  const int k = blk.y + threadIdx.x;
  if (k >= sd.nstrip) return;
  const int nx = sd.nx, nz = a.g.nz, cpml = a.g.cpml, nxm = a.m.nxm;
  int ix, iz;
  strip_cell(k, nx, nz, cpml, ix, iz);
  const long i = sd.off + ix + (long)sd.ps * iz;
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
      u2[j] = v * Lc[i + (long)(j - 4) * sd.ps];
      u4[j] = v * Mc[i + (long)(j - 4) * sd.ps];
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
  const int wave = wave_id(), lane = threadIdx.x & 63;
  const int item = xcd_logical_block(blockIdx.x, gridDim.x) * 4 + wave;
  if (item >= a.nitems) return;
  const int4 itm = a.items[item];
  const ShotDev sd = a.g.shots[itm.x];
  const int nx = sd.nx, ps = sd.ps, nz = a.g.nz, cpml = a.g.cpml, nxm = a.m.nxm;
  const int ix = itm.y + lane;
  const bool act = ix >= kRim && ix < nx - kRim;
  const int cx = clampi(ix, 0, nx - 1);
  const int hx = lane < 8 ? clampi(itm.y + halo_col(lane), 0, nx - 1) : cx;
  const bool xstrip = ix < cpml || ix > nx - cpml - 1;
  const float axv = act ? coef_x(a.g.ta, ix, nx, cpml) : 0.f;
  const float bxv = act ? coef_x(a.g.tb, ix, nx, cpml) : 0.f;
  const float dh = a.g.dh, dh2 = a.g.dh2;
  const float* L = a.Lc + sd.off;
  const float* M = a.Mc + sd.off;
  float* Lo = a.Lo + sd.off;
  float* Mo = a.Mo + sd.off;
  const float* vh = a.m.vh + sd.x0;
  const float* vp = a.m.vp + sd.x0;
  const float* vn = a.m.vn + sd.x0;
  float* buf = xbuf[wave];
  const int z0 = itm.z, z1 = itm.w;
  // z queues: aux2 = v_po lP and aux4 = v_po lF at rows z-4..z+4; lP, lF themselves at rows z..z+4
  float q2[9], q4[9], lq[5], mq[5];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int r = z0 - 4 + k;
    const float l = ldu(L, (unsigned)(cx + ps * r));
    const float m = ldu(M, (unsigned)(cx + ps * r));
    const float v = ldu(vp, (unsigned)(cx + nxm * r));
    q2[k] = v * l;
    q4[k] = v * m;
    if (k >= 4) {
      lq[k - 4] = l;
      mq[k - 4] = m;
    }
  }
  for (int z = z0; z < z1; ++z) {
    const unsigned o = (unsigned)(cx + ps * z);
    const unsigned mo = (unsigned)(cx + nxm * z);
    const unsigned ho = (unsigned)(hx + ps * z), hmo = (unsigned)(hx + nxm * z);
    {
      const float l = ldu(L, o + 4u * ps);
      const float m = ldu(M, o + 4u * ps);
      const float v = ldu(vp, mo + 4u * nxm);
      lq[4] = l;
      mq[4] = m;
      q2[8] = v * l;
      q4[8] = v * m;
    }
    // x-direction auxiliaries aux1 = v_hor lP and aux3 = v_nmo lF of the band row and its halo
    const float u1o = ldu(vh, mo) * lq[0];
    const float u3o = ldu(vn, mo) * mq[0];
    const float u1h = ldu(vh, hmo) * ldu(L, ho);
    const float u3h = ldu(vn, hmo) * ldu(M, ho);
    const float lold = ldn(Lo, o), mold = ldn(Mo, o);
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
        const long g = sd.off + o;
        const float d1 = d1_at<S>(a.s1, g, 1, dh);
        const float d2 = d1_at<S>(a.s2, g, ps, dh);
        const float d3 = d1_at<S>(a.s3, g, 1, dh);
        const float d4 = d1_at<S>(a.s4, g, ps, dh);
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
      float ln = 2.f * lq[0] - lold + a.dh2C * (p2cx + t1 + r2dx + t3);
      float mn = 2.f * mq[0] - mold + a.dh2C * (p2cz + r2cz + t2 + t4);
      if (z == sd.irz && ix >= sd.ir0 && ix < sd.ir0 + sd.nrec) {
        const float r = a.res[sd.goff + (long)a.rit * sd.nrec + (ix - sd.ir0)];
        ln += r;
        mn += r;
      }
      stn(Lo, o, ln);
      stn(Mo, o, mn);
      if (a.qa) {
        const float p2 = ldn(a.qa + sd.off, o), r2 = ldn(a.qb + sd.off, o);
        float* gcx = a.gcx + sd.gslot;
        float* gcz = a.gcz + sd.gslot;
        float* gdx = a.gdx + sd.gslot;
        stn(gcx, o, ldn(gcx, o) + a.dt * p2 * ln);
        stn(gcz, o, ldn(gcz, o) + a.dt * r2 * (ln + mn));
        stn(gdx, o, ldn(gdx, o) + a.dt * p2 * mn);
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
// Exact discrete adjoint.  With chi_x / chi_z the side / bottom C-PML cells of the updated region and
// Gamma = chi_x U chi_z:
//   abar = dt^2 (v_hor mP + v_nmo mF), bbar = dt^2 v_po (mP + mF)
//   zt_x = chi_x (zbar_x + abar), gbar_x = abar + a_x zt_x, zbar_x <- b_x zt_x
//   pt_x = chi_x (sbar_x - D1x(Gamma gbar_x)), sbar_x <- b_x pt_x
//   m^k = 2 m^{k+1} + D2x gbar_x - D1x(a_x pt_x) - m^{k+2} (+ residual at the receivers), likewise z.
// The work items compute these C-PML terms themselves from the memory variables zbar, sbar of step k+1 (a.zxo,
// a.sxo, a.zzo, a.szo) and store those of step k on their own C-PML cells (a.zxn, a.sxn, a.zzn, a.szn).  a_x pt_x
// enters D1x 4 columns beyond an item and is itself a D1x of gbar_x, so an item near a side strip forms gbar_x on
// the 80 columns around its 64 (exchanged through LDS); a bottom item keeps gbar_z of 13 rows and a_z pt_z of 9 in
// queues.  Neighbouring items form each other's border terms from the same inputs, and the memory variables of
// steps k+1 and k live in separate arrays, so every item reads the same bits whatever the launch order.
// ---------------------------------------------------------------------------------------------

// Main exact adjoint step with the exact imaging condition
//   g_hor += dt^2 mP^k a^{k-1}, g_nmo += dt^2 mF^k a^{k-1}, g_po += dt^2 (mP^k + mF^k) b^{k-1}.
// GROUP: the four waves of a workgroup are up to four shots of one gradient group on the same 64 model
// columns (each wave at its own window's local offset); a wave whose item has shot -(leader+1) is a
// placeholder that loads nothing.  Per row the waves' contributions meet in LDS and are added in wave order,
// ((c0 + c1) + c2) + c3 (zero where a shot's window does not reach), then waves 0, 1, 2 read-modify-write
// g_hor, g_nmo, g_po of the group's slot, which spans the union of its windows: 6 B of gradient traffic per
// cell and shot instead of 24 when the windows coincide.  Items are padded to whole workgroups, so every
// wave reaches every barrier.
//
// Interior items (bands out of reach of the side strips, rows above nz-cpml-4) are software-pipelined; items
// that need C-PML terms take a general body per row, interleaved with them in the launch.  Every path of a
// wave passes one barrier per row, so a workgroup may mix them.
// IMG: the step adds imaging terms (a.qa, a.qb are set), a launch constant so that the body has one path.
// INC: increment form of the adjoint leapfrog on the rows of the forward's FP64 band, [4, a.band) (the FP32
// increment-form arithmetic): the state there is m^{k+1} and D = m^{k+1} - m^{k+2} (a.DP, a.DF, updated in place)
// instead of m^{k+2}; D^k = D^{k+1} + (the step's terms) (+ the residual at the receivers), m^k = m^{k+1} + D^k --
// the same recursion in exact arithmetic, with the time difference stored at its own precision.  Near the
// receivers, where the residual enters, the two-level form's rounding is what limits the gradient; below, it is not.
// SHELL: an item of a step pair's shell (a.xr holds each item's column range): it computes every cell of its band as
// the single step does but stores, images and updates the gradient only in its range, with the general body.
// buf: the wave's 152 LDS floats; gsum: the workgroup's imaging buffers.
template <class S, bool GROUP, bool IMG, bool INC, bool SHELL>
__device__ __forceinline__ void adj_ex_item(const AdjExArgs& a, int item, float* buf, float (*gsum)[4][3][64]) {
  const int wave = wave_id(), lane = threadIdx.x & 63;
  if (item >= a.nitems) return;
  int4 itm = a.items[item];
  const bool ghost = GROUP && itm.x < 0;
  if (ghost) itm.x = -itm.x - 1;
  const ShotDev sd = a.g.shots[itm.x];
  const int nx = sd.nx, ps = sd.ps, nz = a.g.nz, cpml = a.g.cpml, nxm = a.m.nxm;
  const int ix = itm.y + lane;
  const int z0 = itm.z, z1 = itm.w;
  constexpr bool img = IMG;
  bool inr = true;  // the lane's cell is one this item updates
  if constexpr (SHELL) {
    const int2 r = a.xr[item];
    inr = ix >= r.x && ix < r.y;
  }
  // GROUP: field `wave` of the group's slot, read-modify-written by waves 0..2 on the model columns of the
  // union window; other lanes (and wave 3) address past the resource, so their accesses drop out
  const int mc = sd.x0 + ix;  // model column of the lane, the same in every wave of the workgroup
  const bool gown = GROUP && img && wave < 3 && mc >= sd.gx0 && mc < sd.gx0 + sd.gnx && inr;
  const brsrc grs = make_rsrc((wave == 0 ? a.gh : wave == 1 ? a.gn : a.gp) + sd.gslot, 4u * sd.gps * nz);
  const unsigned gcol = gown ? 4u * (unsigned)(mc - sd.gx0) : kOutOfRange;
  const unsigned grow = 4u * (unsigned)sd.gps;
  if (GROUP && ghost) {
    // placeholder wave: contributes zeros, takes part in the row barriers and in the slot's update
    if (!img) return;
    float gv = bld<kBufNT>(grs, gcol + grow * z0);
    for (int z = z0; z < z1; ++z) {
      const float g0 = gv;
      gv = bld<kBufNT>(grs, gcol + grow * min(z + 1, z1 - 1));
      float(*gs)[3][64] = gsum[GROUP ? (z & 1) : 0];
      gs[wave][0][lane] = 0.f;
      gs[wave][1][lane] = 0.f;
      gs[wave][2][lane] = 0.f;
      __syncthreads();
      const float sum = ((gs[0][wave][lane] + gs[1][wave][lane]) + gs[2][wave][lane]) + gs[3][wave][lane];
      bst<kBufNT>(grs, gcol + grow * z, g0 + sum);
    }
    return;
  }
  const bool act = ix >= kRim && ix < nx - kRim;
  const int cx = clampi(ix, 0, nx - 1);
  const int hxr = itm.y + halo_col(lane);
  const bool hact = lane < 8 && hxr >= kRim && hxr < nx - kRim;
  const int hx = lane < 8 ? clampi(hxr, 0, nx - 1) : cx;
  // bands within reach of a side strip need the C-PML terms in x
  const bool xnear = itm.y < cpml + 4 || itm.y + 64 > nx - cpml - 4;
  // There gbar_x is formed on the 80 columns x0-8 .. x0+71 (LDS slot = column - x0 + 8): each lane at its column
  // and at a second one, x0-8+lane for lanes 0..7, x0+56+lane for lanes 8..15, its own again for the others
  const int wxr = lane < 8 ? itm.y - 8 + lane : lane < 16 ? itm.y + 56 + lane : ix;
  const bool wact = wxr >= kRim && wxr < nx - kRim;
  const int wx = clampi(wxr, 0, nx - 1);
  const int wslot = wxr - itm.y + 8;
  auto chix = [&](int c) { return (c >= kRim && c < cpml) || (c >= nx - cpml && c < nx - kRim); };
  const bool xso = chix(ix), xsw = chix(wxr), xsh = lane < 8 && chix(hxr);
  // Gamma membership above the bottom strip of the 9 columns around the lane's column (bit k: column ix-4+k) and
  // around its halo column, as bits of a vector register (tested per row, not held as loop-invariant lane masks)
  unsigned gmo = 0, gmh = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    gmo |= chix(ix - 4 + k) ? 1u << k : 0u;
    gmh |= chix(hxr - 4 + k) ? 1u << k : 0u;
  }
  const float axo = act ? coef_x(a.g.ta, ix, nx, cpml) : 0.f;
  const float bxo = xso ? coef_x(a.g.tb, ix, nx, cpml) : 0.f;
  const float axw = wact ? coef_x(a.g.ta, wx, nx, cpml) : 0.f;
  const float axh = hact ? coef_x(a.g.ta, hx, nx, cpml) : 0.f;
  const float dh = a.g.dh, dh2 = a.g.dh2, dt2 = a.dt2;
  const float* L = a.Lc + sd.off;
  const float* M = a.Mc + sd.off;
  const float* Lp = a.Lp + sd.off;
  const float* Mp = a.Mp + sd.off;
  float* Lo = a.Lo + sd.off;
  float* Mo = a.Mo + sd.off;
  float* DP = a.DP + sd.off;  // increment-form rows: D
  float* DF = a.DF + sd.off;
  const float* vh = a.m.vh + sd.x0;
  const float* vp = a.m.vp + sd.x0;
  const float* vn = a.m.vn + sd.x0;
  const float* zxo = a.zxo + sd.off;
  const float* sxo = a.sxo + sd.off;
  const float* zzo = a.zzo + sd.off;
  const float* szo = a.szo + sd.off;
  float* zxn = a.zxn + sd.off;
  float* sxn = a.sxn + sd.off;
  float* zzn = a.zzn + sd.off;
  float* szn = a.szn + sd.off;
  const float* qa = img ? a.qa + sd.off : L;
  const float* qb = img ? a.qb + sd.off : L;
  const int zb = nz - cpml;      // first row of the bottom C-PML strip
  const int zc = nz - cpml - 4;  // rows at or below this need the C-PML terms in z
  float* gh_ = a.gh + sd.gslot;  // per-shot gradient (ungrouped)
  float* gn_ = a.gn + sd.gslot;
  float* gp_ = a.gp + sd.gslot;

  // gbar_z along z: rows z-4..z+4 in gq[0..8]; in the bottom items rows z-4..z+8 in gq[0..12], with a_z pt_z of
  // rows z-4..z+4 in pq[0..8]
  float gq[13], pq[9];
  // C-PML coefficients a_z, b_z of row r (zero above the strip), read without a branch so that the scalar loads
  // issue with the row's vector loads (rows past the grid read the outermost entry and are masked by the caller)
  auto coefs_z = [&](int r) __attribute__((always_inline)) {
    const int e = clampi(nz - 1 - r, 0, cpml - 1);
    const __attribute__((address_space(4))) float* ta = (const __attribute__((address_space(4))) float*)a.g.ta;
    const __attribute__((address_space(4))) float* tb = (const __attribute__((address_space(4))) float*)a.g.tb;
    const float av = ta[e], bv = tb[e];
    return make_float2(r >= zb ? av : 0.f, r >= zb ? bv : 0.f);
  };
  // gbar_z of row r of a bottom item from its mP, mF, v_po, zeta-bar_z^{k+1} and c = coefs_z(r); stores zeta-bar_z^k
  // of an own cell
  auto gbar_z = [&](int r, float l, float m, float vpv, float zo, float2 c) __attribute__((always_inline)) {
    const float v = dt2 * (vpv * (l + m));
    const bool zs = act && r >= zb && r < nz - kRim;
    const float tz = zs ? zo + v : 0.f;
    if (zs && inr && r >= z0 && r < z1) stu(zzn, (unsigned)(cx + ps * r), c.y * tz);
    return (act && r >= kRim && r < nz - kRim) ? v + c.x * tz : 0.f;
  };
  // a_z pt_z of row r from gbar_z of rows r-4..r+4 (g[0..8], Gamma-masked here: rows above the strip only in the
  // side strips) and psi-bar_z^{k+1} (loaded here if so is null); stores psi-bar_z^k of an own cell
  auto psi_z = [&](int r, const float* g, const float* so, float2 c) __attribute__((always_inline)) {
    if (!(act && r >= zb && r < nz - kRim)) return 0.f;
    const float sv = so ? *so : ldu(szo, (unsigned)(cx + ps * r));
    float d;
    if (r >= zb + 4) {
      d = d1s<S>(g[0], g[1], g[2], g[3], g[5], g[6], g[7], g[8], dh);
    } else {
      float w[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) w[k] = (xso || r - 4 + k >= zb) ? g[k] : 0.f;
      d = d1s<S>(w[0], w[1], w[2], w[3], w[5], w[6], w[7], w[8], dh);
    }
    const float t = sv - d;
    if (inr && r >= z0 && r < z1) stu(szn, (unsigned)(cx + ps * r), c.y * t);
    return c.x * t;
  };
  // The loads of a row are issued two rows ahead, before the current row's stores (the vector memory counter
  // retires in issue order, so a load issued after a store can only be waited for together with it), into
  // three row buffers used in rotation, so no register copy waits for them (interior items; the general body
  // loads each row at its start).
  struct Row {
    float l4, m4, vp4, zt4, s4;            // row z+4 (ZB: z+8): mP, mF, v_po, zeta-bar_z; ZB: psi-bar_z of row z+4
    float lc, mc, vhc, vnc;                // row z at the lane's column: mP, mF (an L2 hit), v_hor, v_nmo
    float vhh, vnh, lh, mh;                // row z at the halo column (XN: the lane's second gbar_x column)
    float xc, xh, pc, ph;                  // XN: zeta-bar_x at the lane's and second column, psi-bar_x at the
                                           // lane's and the halo column
    float lold, mold, qa, qb, g0, g1, g2;  // centre-only values: m^{k+2} (INC: D), imaging terms, gradient
  };
  // XN: band near a side strip; ZB: rows z >= zc (bottom C-PML terms); IN: increment form
  auto load_row = [&](auto xn_t, auto zb_t, auto in_t, int z, Row& r) __attribute__((always_inline)) {
    constexpr bool XN = decltype(xn_t)::value, ZB = decltype(zb_t)::value, IN = decltype(in_t)::value;
    const unsigned o = (unsigned)(cx + ps * z);
    const unsigned mo = (unsigned)(cx + nxm * z);
    const int hc = XN ? wx : hx;
    const unsigned ho = (unsigned)(hc + ps * z), hmo = (unsigned)(hc + nxm * z);
    const int r4 = ZB ? min(z + 8, nz - 1) : z + 4;
    const unsigned ir = (unsigned)(cx + ps * r4);
    r.l4 = ldu(L, ir);
    r.m4 = ldu(M, ir);
    r.vp4 = ldu(vp, (unsigned)(cx + nxm * r4));
    if (ZB) {
      r.zt4 = ldu(zzo, ir);
      r.s4 = ldu(szo, o + 4u * ps);
    }
    r.lc = ldu(L, o);
    r.mc = ldu(M, o);
    r.vhc = ldu(vh, mo);
    r.vnc = ldu(vn, mo);
    r.vhh = ldu(vh, hmo);
    r.lh = ldu(L, ho);
    r.vnh = ldu(vn, hmo);
    r.mh = ldu(M, ho);
    if (XN) {
      r.xc = ldu(zxo, o);
      r.xh = ldu(zxo, ho);
      r.pc = ldu(sxo, o);
      r.ph = ldu(sxo, (unsigned)(hx + ps * z));
    }
    r.lold = ldn(IN ? DP : Lp, o);
    r.mold = ldn(IN ? DF : Mp, o);
    if (img) {
      r.qa = ldn(qa, o);
      r.qb = ldn(qb, o);
      if (GROUP) {
        r.g0 = bld<kBufNT>(grs, gcol + grow * z);
      } else {
        r.g0 = ldn(gh_, o);
        r.g1 = ldn(gn_, o);
        r.g2 = ldn(gp_, o);
      }
    }
  };
  // gq (and pq) at the first row of the item
  auto prologue = [&](auto zb_t) __attribute__((always_inline)) {
    if constexpr (!decltype(zb_t)::value) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int r = z0 - 4 + k;
        const unsigned ir = (unsigned)(cx + ps * r);
        const float v = dt2 * (ldu(vp, (unsigned)(cx + nxm * r)) * (ldu(L, ir) + ldu(M, ir)));
        gq[k] = (act && r >= kRim) ? v : 0.f;
      }
    } else if (z0 <= zb - 4) {
      // an item starting 4 rows or more above the strip: pq of rows z0-4..z0+3 is zero, gq of rows z0-4..z0+7 is
      // loaded in one round trip
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        const int r = z0 - 4 + k;
        const unsigned ir = (unsigned)(cx + ps * r);
        const float zo = r >= zb ? ldu(zzo, ir) : 0.f;
        gq[k] = gbar_z(r, ldu(L, ir), ldu(M, ir), ldu(vp, (unsigned)(cx + nxm * r)), zo, coefs_z(r));
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) pq[k] = 0.f;
    } else {
      // the queues' advance over the rows before z0, as the step makes it, from zeros: gq needs rows from z0-4, pq
      // rows from z0-4, which take gbar_z from row z0-8 -- unless they lie above the strip, where pq is zero, as it
      // is up to row z0+3 when the item starts at or above the strip (then gq needs rows from z0-4 only)
#pragma unroll
      for (int k = 0; k < 13; ++k) gq[k] = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) pq[k] = 0.f;
#pragma unroll 4
      for (int z = z0 <= zb ? z0 - 12 : z0 - 16; z < z0; ++z) {
        const int r = clampi(z + 8, 0, nz - 1);
        const unsigned ir = (unsigned)(cx + ps * r);
        const float2 c8 = coefs_z(z + 8), c4 = coefs_z(z + 4);
        gq[12] = gbar_z(z + 8, ldu(L, ir), ldu(M, ir), ldu(vp, (unsigned)(cx + nxm * r)), ldu(zzo, ir), c8);
        pq[8] = psi_z(z + 4, gq + 4, nullptr, c4);
#pragma unroll
        for (int k = 0; k < 12; ++k) gq[k] = gq[k + 1];
#pragma unroll
        for (int k = 0; k < 8; ++k) pq[k] = pq[k + 1];
      }
    }
  };
  // Row z from cur.  PIPE (interior items): row min(z+2, ze-1) is loaded into nxt once cur's stencil inputs
  // are consumed and before the row's stores (unused repeats on the last rows keep one path through the
  // body).  The edge kernel loads each row at its start: its C-PML inputs leave no registers for a second row.
  auto step = [&](auto xn_t, auto zb_t, auto pipe_t, auto in_t, int z, int ze, const Row& cur, Row& nxt)
                  __attribute__((always_inline)) {
    constexpr bool XN = decltype(xn_t)::value, ZB = decltype(zb_t)::value, PIPE = decltype(pipe_t)::value;
    constexpr bool IN = decltype(in_t)::value;
    const unsigned o = (unsigned)(cx + ps * z);
    if (ZB) {
      const float2 c8 = coefs_z(z + 8), c4 = coefs_z(z + 4);
      gq[12] = gbar_z(z + 8, cur.l4, cur.m4, cur.vp4, cur.zt4, c8);
      pq[8] = psi_z(z + 4, gq + 4, &cur.s4, c4);
    } else {
      const int r = z + 4;
      const float v = dt2 * (cur.vp4 * (cur.l4 + cur.m4));
      gq[8] = (act && r < nz - kRim) ? v : 0.f;
    }
    // gbar_x of the band row (zero outside Omega) and of the halo (XN: second) columns
    float go = dt2 * (cur.vhc * cur.lc + cur.vnc * cur.mc);
    float gh = dt2 * (cur.vhh * cur.lh + cur.vnh * cur.mh);
    if (PIPE) load_row(xn_t, zb_t, in_t, min(z + 2, ze - 1), nxt);
    float ap;
    if (XN) {
      // zt_x = zeta-bar_x^{k+1} + abar on chi_x; zeta-bar_x^k of the lane's cell
      const float tx = xso ? cur.xc + go : 0.f;
      const float tw = xsw ? cur.xh + gh : 0.f;
      if (xso && inr) stu(zxn, o, bxo * tx);
      go = act ? go + axo * tx : 0.f;
      gh = wact ? gh + axw * tw : 0.f;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      buf[8 + lane] = go;
      buf[wslot] = gh;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      float u[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) u[k] = buf[lane + 4 + k];
      ap = d2s<S>(u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8], dh2);
      // a_x pt_x at the lane's column and (lanes 0..7) at its halo column, from the Gamma-masked gbar_x; on a
      // bottom strip row every cell of Omega is in Gamma
      const unsigned mo = z >= zb ? 0x1ffu : gmo, mh = z >= zb ? 0x1ffu : gmh;
      float pto = 0.f, pth = 0.f;
      if (xso) {
        float w[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = ((mo >> k) & 1u) ? u[k] : 0.f;
        const float t = cur.pc - d1s<S>(w[0], w[1], w[2], w[3], w[5], w[6], w[7], w[8], dh);
        if (inr) stu(sxn, o, bxo * t);
        pto = axo * t;
      }
      if (xsh) {
        const int hs = hxr - itm.y + 8;
        float w[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = ((mh >> k) & 1u) ? buf[hs - 4 + k] : 0.f;
        const float t = cur.ph - d1s<S>(w[0], w[1], w[2], w[3], w[5], w[6], w[7], w[8], dh);
        pth = axh * t;
      }
      float px[9];
      xrow9(buf + 80, lane, pto, pth, px);
      ap = ap - d1s<S>(px[0], px[1], px[2], px[3], px[5], px[6], px[7], px[8], dh);
    } else {
      go = act ? go : 0.f;
      gh = hact ? gh : 0.f;
      float gx[9];
      xrow9(buf, lane, go, gh, gx);
      ap = d2s<S>(gx[0], gx[1], gx[2], gx[3], gx[4], gx[5], gx[6], gx[7], gx[8], dh2);
    }
    float af;
    if (ZB) {
      af = d2s<S>(gq[0], gq[1], gq[2], gq[3], gq[4], gq[5], gq[6], gq[7], gq[8], dh2);
      af = af - d1s<S>(pq[0], pq[1], pq[2], pq[3], pq[5], pq[6], pq[7], pq[8], dh);
    } else {
      af = d2s<S>(gq[0], gq[1], gq[2], gq[3], gq[4], gq[5], gq[6], gq[7], gq[8], dh2);
    }
    // IN: ln, mn first hold D^k, then m^k
    float ln = IN ? cur.lold + ap : 2.f * cur.lc + ap - cur.lold;
    float mn = IN ? cur.mold + af : 2.f * cur.mc + af - cur.mold;
    if (z == sd.irz && sd.nrec > 0) {
      const bool rec = ix >= sd.ir0 && ix < sd.ir0 + sd.nrec;
      const float r = a.res[sd.goff + (long)a.rit * sd.nrec + (rec ? ix - sd.ir0 : 0)];
      if (rec) {
        ln += r;
        mn += r;
      }
    }
    // Every lane stores: a lane off the updated cells (the rim, or a column past the shot clamped onto the
    // rim) stores the rim's value it read, zero at every time level (IN: m^{k+1} there, as m^{k+2}).
    if (IN) {
      if (inr) {
        stn(DP, o, act ? ln : cur.lold);
        stn(DF, o, act ? mn : cur.mold);
      }
      ln = cur.lc + ln;
      mn = cur.mc + mn;
      if (inr) {
        stn(Lo, o, act ? ln : cur.lc);
        stn(Mo, o, act ? mn : cur.mc);
      }
    } else if (inr) {
      stn(Lo, o, act ? ln : cur.lold);
      stn(Mo, o, act ? mn : cur.mold);
    }
    float ch = 0.f, cn = 0.f, cp = 0.f;  // this shot's imaging contributions (GROUP)
    if (img) {
      const bool upd = act && inr;
      if (GROUP) {
        ch = upd ? dt2 * ln * cur.qa : 0.f;
        cn = upd ? dt2 * mn * cur.qa : 0.f;
        cp = upd ? dt2 * (ln + mn) * cur.qb : 0.f;
      } else if (inr) {
        stn(gh_, o, act ? cur.g0 + dt2 * ln * cur.qa : cur.g0);
        stn(gn_, o, act ? cur.g1 + dt2 * mn * cur.qa : cur.g1);
        stn(gp_, o, act ? cur.g2 + dt2 * (ln + mn) * cur.qb : cur.g2);
      }
    }
    if (ZB) {
#pragma unroll
      for (int k = 0; k < 12; ++k) gq[k] = gq[k + 1];
#pragma unroll
      for (int k = 0; k < 8; ++k) pq[k] = pq[k + 1];
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) gq[k] = gq[k + 1];
    }
    if (GROUP && img) {
      // double-buffered by row parity: one barrier per row suffices, since a wave can only overwrite this
      // buffer two rows later, after every wave has passed the next row's barrier
      float(*gs)[3][64] = gsum[GROUP ? (z & 1) : 0];
      gs[wave][0][lane] = ch;
      gs[wave][1][lane] = cn;
      gs[wave][2][lane] = cp;
      __syncthreads();
      const float sum = ((gs[0][wave][lane] + gs[1][wave][lane]) + gs[2][wave][lane]) + gs[3][wave][lane];
      bst<kBufNT>(grs, gcol + grow * z, cur.g0 + sum);
    }
  };
  auto rows = [&](auto xn_t, auto zb_t, auto pipe_t, auto in_t, int za, int ze) __attribute__((always_inline)) {
    if (za >= ze) return;
    prologue(zb_t);
    Row ra, rb;
    if (decltype(pipe_t)::value) {
      // two rows in flight, three buffers in rotation unrolled by three (no register copy waits for a load),
      // the first three rows peeled so the loop head sees the same memory-counter state from both paths
      Row rc;
      load_row(xn_t, zb_t, in_t, za, ra);
      load_row(xn_t, zb_t, in_t, min(za + 1, ze - 1), rb);
      int z = za;
      if (z + 2 < ze) {
        step(xn_t, zb_t, pipe_t, in_t, z, ze, ra, rc);
        step(xn_t, zb_t, pipe_t, in_t, z + 1, ze, rb, ra);
        step(xn_t, zb_t, pipe_t, in_t, z + 2, ze, rc, rb);
        for (z += 3; z + 2 < ze; z += 3) {
          step(xn_t, zb_t, pipe_t, in_t, z, ze, ra, rc);
          step(xn_t, zb_t, pipe_t, in_t, z + 1, ze, rb, ra);
          step(xn_t, zb_t, pipe_t, in_t, z + 2, ze, rc, rb);
        }
      }
      if (z < ze) step(xn_t, zb_t, pipe_t, in_t, z, ze, ra, rc);
      if (z + 1 < ze) step(xn_t, zb_t, pipe_t, in_t, z + 1, ze, rb, ra);
    } else {
      for (int z = za; z < ze; ++z) {
        load_row(xn_t, zb_t, in_t, z, ra);
        step(xn_t, zb_t, pipe_t, in_t, z, ze, ra, rb);
      }
    }
  };
  // an item lies entirely above row zc or entirely at or below it, and entirely above a.band or at or below it
  // (the items are split at both rows; the band ends above zc)
  using T = std::true_type;
  using F = std::false_type;
  const bool inc = INC && z1 <= a.band;
  using P = std::integral_constant<bool, !SHELL>;  // shell items take the general body
  if (xnear) {
    if (z0 >= zc) rows(T{}, T{}, F{}, F{}, z0, z1);
    else if (inc) rows(T{}, F{}, F{}, std::integral_constant<bool, INC>{}, z0, z1);
    else rows(T{}, F{}, F{}, F{}, z0, z1);
  } else {
    if (z0 >= zc) rows(F{}, T{}, F{}, F{}, z0, z1);
    else if (inc) rows(F{}, F{}, P{}, std::integral_constant<bool, INC>{}, z0, z1);
    else rows(F{}, F{}, P{}, F{}, z0, z1);
  }
}

template <class S, bool GROUP, bool IMG, bool INC = false>
__global__ void __launch_bounds__(256) k_adj_ex_main(AdjExArgs a) {
  __shared__ float xbuf[4][152];
  __shared__ float gsum[GROUP ? 2 : 1][4][3][64];
  const int wave = wave_id();
  adj_ex_item<S, GROUP, IMG, INC, false>(a, xcd_logical_block(blockIdx.x, gridDim.x) * 4 + wave, xbuf[wave], gsum);
}

// ---------------------------------------------------------------------------------------------
// Two exact adjoint steps k, k-1 per launch (temporal blocking) on each gradient group's core: the model columns
// where every member's window has columns [cpml+8, nx-cpml-8) (rounded inward to 16) and rows
// [max(4, band)+4, nz-cpml-8), so that neither step, nor the halo of step k that step k-1 reads, involves a C-PML
// term, a row of the increment-form band or the rim.  A workgroup is one group on one tile of up to 48 output
// columns, member j on wave j (placeholder waves as in the single step); the lanes of a wave compute step k on
// columns [x0-4, x0+52) and step k-1 on the tile, marching down: at row r step k at row r, step k-1 at row r-4.
// gbar_x^{k+1} and gbar_x^k are exchanged within the wave through LDS, gbar_z^{k+1} and gbar_z^k wait in register
// queues, mu^k, gbar_x^k and v_po in four-row delay lines.  mu^{k+1} is loaded once, at row r+4, and waits in a
// per-lane ring of eight rows in LDS for its uses at rows r (step k) and r-4 (step k-1): loading it again there
// missed the L2 (16 % more fetched bytes, 5 % of the launch's time) and holding it in registers would cost a wave
// per SIMD.  The imaging terms and the gradient slot are read on the owned lanes only (a lane off the tile would
// fetch a third cache line per row: 10 % of the launch's time).  Per row the members' imaging terms of both
// steps meet in LDS and waves 0-2 update their field of the group's gradient slot once, as (g + S_k) + S_{k-1}:
// the two single steps' additions in their order.  Every cell is computed with the single-step expressions in the
// same order, so the results are identical to two single steps.  Stores go through buffer resources with an
// out-of-range offset on the lanes that do not own a cell.
// ---------------------------------------------------------------------------------------------
constexpr int kAdjPairCols = 48;
__host__ __device__ inline int adj_pair_cx0(int cpml) { return (cpml + 8 + 15) & ~15; }
__host__ __device__ inline int adj_pair_cx1(int nx, int cpml) { return (nx - cpml - 8) & ~15; }

// Wave-private exchange of one value per lane: on return u[k] is lane (l-4+k)'s value.  Lanes 0..3 and 60..63 read
// slots no lane wrote; their results are not used.  buf: 72 floats of the wave's LDS.
__device__ __forceinline__ void xrow64(float* buf, int lane, float v, float (&u)[9]) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  buf[4 + lane] = v;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int k = 0; k < 9; ++k) u[k] = buf[lane + k];
}

// ring: the wave's 16 x 64 floats of LDS, mu_P^{k+1} and mu_F^{k+1} of row q in rows 2(q&7) and 2(q&7)+1.
__device__ __forceinline__ void adj_pair_core(const AdjPairArgs& a, int4 itm, float* xb, float (*gsum)[4][6][64],
                                              float (*ring)[64]) {
  const int wave = wave_id(), lane = threadIdx.x & 63;
  const bool ghost = itm.x < 0;
  if (ghost) itm.x = -itm.x - 1;
  const ShotDev sd = a.g.shots[itm.x];
  const int nx = sd.nx, ps = sd.ps, nz = a.g.nz, nxm = a.m.nxm;
  const int x0 = itm.y & 0xffff, ncols = itm.y >> 16;
  const int z0 = itm.z, z1 = itm.w;
  const int col = x0 - 8 + lane;  // the lane's column: step k on lanes 4..59, step k-1 on lanes 8..55
  const bool own = lane >= 8 && lane < 8 + ncols;
  const int fw = wave < 3 ? wave : 2;  // the field a wave sums (wave 3 sums one it does not store)
  // field `wave` of the group's gradient slot (waves 0..2) on the owned columns
  const int mc = sd.x0 + col;
  const bool gown = wave < 3 && own && mc >= sd.gx0 && mc < sd.gx0 + sd.gnx;
  const brsrc grs = make_rsrc((wave == 0 ? a.gh : wave == 1 ? a.gn : a.gp) + sd.gslot, 4u * sd.gps * nz);
  const unsigned gcol = gown ? 4u * (unsigned)(mc - sd.gx0) : kOutOfRange;
  const unsigned grow = 4u * (unsigned)sd.gps;
  if (ghost) {
    // placeholder wave: contributes zeros, takes part in the row barriers and in the slot's update
    float gv = bld<kBufNT>(grs, gcol + grow * z0);
    for (int z = z0; z < z1; ++z) {
      const float g0 = gv;
      gv = bld<kBufNT>(grs, gcol + grow * min(z + 1, z1 - 1));
      float(*gs)[6][64] = gsum[z & 1];
#pragma unroll
      for (int f = 0; f < 6; ++f) gs[wave][f][lane] = 0.f;
      __syncthreads();
      const float s0 = ((gs[0][fw][lane] + gs[1][fw][lane]) + gs[2][fw][lane]) + gs[3][fw][lane];
      const float s1 = ((gs[0][3 + fw][lane] + gs[1][3 + fw][lane]) + gs[2][3 + fw][lane]) + gs[3][3 + fw][lane];
      bst<kBufNT>(grs, gcol + grow * z, (g0 + s0) + s1);
    }
    return;
  }
  const int cc = clampi(col, 0, nx - 1);  // loads: a valid column (lanes outside the shot feed no owned cell)
  const float dh2 = a.g.dh2, dt2 = a.dt2;
  const float* L1 = a.L1 + sd.off;
  const float* M1 = a.M1 + sd.off;
  const float* L2 = a.L2 + sd.off;
  const float* M2 = a.M2 + sd.off;
  const float* vh = a.m.vh + sd.x0;
  const float* vp = a.m.vp + sd.x0;
  const float* vn = a.m.vn + sd.x0;
  const unsigned fbytes = (unsigned)ps * (unsigned)nz * 4u;
  const brsrc sL0 = make_rsrc(a.L0 + sd.off, fbytes), sM0 = make_rsrc(a.M0 + sd.off, fbytes);
  const brsrc sLm = make_rsrc(a.Lm + sd.off, fbytes), sMm = make_rsrc(a.Mm + sd.off, fbytes);
  const brsrc sqa0 = make_rsrc(a.qa0 + sd.off, fbytes), sqb0 = make_rsrc(a.qb0 + sd.off, fbytes);
  const brsrc sqa1 = make_rsrc(a.qa1 + sd.off, fbytes), sqb1 = make_rsrc(a.qb1 + sd.off, fbytes);
  const bool rec = sd.nrec > 0 && col >= sd.ir0 && col < sd.ir0 + sd.nrec;
  const float* res0 = a.res + sd.goff + (long)a.rit * sd.nrec + (rec ? col - sd.ir0 : 0);  // step k
  const float* res1 = res0 - sd.nrec;                                                       // step k-1
  const int zend = z1 + 4;  // step k runs on rows [z0-4, zend)
  struct Row {
    float vh, vn;                // row r: v_hor, v_nmo
    float l4, m4, vp4;           // row r+4: mu^{k+1}, v_po
    float lp, mp;                // row r: mu^{k+2}
  };
  auto load_row = [&](int r, Row& w) __attribute__((always_inline)) {
    const unsigned o = (unsigned)(cc + ps * r);
    const unsigned mo = (unsigned)(cc + nxm * r);
    w.vh = ldu(vh, mo);
    w.vn = ldu(vn, mo);
    w.l4 = ldu(L1, o + 4u * ps);
    w.m4 = ldu(M1, o + 4u * ps);
    w.vp4 = ldu(vp, mo + 4u * nxm);
    w.lp = ldn(L2, o);
    w.mp = ldn(M2, o);
  };
  float gq1[9];  // gbar_z^{k+1}, rows r-4..r+4
  float gq0[9];  // gbar_z^k, rows r-8..r
  float dl[4], dm[4], dgo[4], dvp[4];  // rows r-4..r-1: mu^k, gbar_x^k; v_po of rows r..r+3
  float* xa = xb;       // exchange of gbar_x^{k+1}
  float* xk = xb + 72;  // of gbar_x^k
  // MAIN: step k-1 at row r-4 (rows r >= z0+4); otherwise step k only
  auto iter = [&](int r, const Row& cur, Row& nxt, auto mainc) __attribute__((always_inline)) {
    constexpr bool MAIN = decltype(mainc)::value;
    // mu^{k+1} of rows r and r-4 from the ring; row r+4 then takes row r-4's slot
    const int sr = 2 * (r & 7), sb = 2 * ((r - 4) & 7);
    const float lc = ring[sr][lane], mc = ring[sr + 1][lane];
    const float lk1 = ring[sb][lane], mk1 = ring[sb + 1][lane];
    ring[sb][lane] = cur.l4;
    ring[sb + 1][lane] = cur.m4;
    // MAIN: row r-4's imaging terms and gradient slot (owned lanes), issued before this row's stores so that their
    // wait does not cover them
    float qa0 = 0.f, qb0 = 0.f, qa1 = 0.f, qb1 = 0.f, gv = 0.f;
    if (MAIN) {
      const unsigned ob = (unsigned)(cc + ps * (r - 4));
      const unsigned oq = own ? 4u * ob : kOutOfRange;
      qa0 = bld<kBufNT>(sqa0, oq);
      qb0 = bld<kBufNT>(sqb0, oq);
      qa1 = bld<kBufNT>(sqa1, oq);
      qb1 = bld<kBufNT>(sqb1, oq);
      gv = bld<kBufNT>(grs, gcol + grow * (r - 4));
    }
    gq1[8] = dt2 * (cur.vp4 * (cur.l4 + cur.m4));
    const float go1 = dt2 * (cur.vh * lc + cur.vn * mc);
    float u[9];
    xrow64(xa, lane, go1, u);
    load_row(min(r + 2, zend - 1), nxt);  // two rows ahead; unused repeats on the last rows keep one path
    // step k at row r (the single step's expressions; no C-PML term in the core and its halo)
    const float ap = d2s<float>(u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8], dh2);
    const float af = d2s<float>(gq1[0], gq1[1], gq1[2], gq1[3], gq1[4], gq1[5], gq1[6], gq1[7], gq1[8], dh2);
    float ln = 2.f * lc + ap - cur.lp;
    float mn = 2.f * mc + af - cur.mp;
    if (r == sd.irz && rec) {
      const float rv = res0[0];
      ln += rv;
      mn += rv;
    }
    if (r >= z0 && r < z1) {
      const unsigned ob = own ? 4u * (unsigned)(col + ps * r) : kOutOfRange;
      bst<kBufNT>(sL0, ob, ln);
      bst<kBufNT>(sM0, ob, mn);
    }
    const float go0 = dt2 * (cur.vh * ln + cur.vn * mn);  // gbar_x^k at row r
    gq0[8] = dt2 * (dvp[0] * (ln + mn));                   // gbar_z^k at row r
    if (MAIN) {
      const int rb = r - 4;
      float w[9];
      xrow64(xk, lane, dgo[0], w);
      const float ap1 = d2s<float>(w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8], dh2);
      const float af1 = d2s<float>(gq0[0], gq0[1], gq0[2], gq0[3], gq0[4], gq0[5], gq0[6], gq0[7], gq0[8], dh2);
      float ln1 = 2.f * dl[0] + ap1 - lk1;
      float mn1 = 2.f * dm[0] + af1 - mk1;
      if (rb == sd.irz && rec) {
        const float rv = res1[0];
        ln1 += rv;
        mn1 += rv;
      }
      const unsigned ob = own ? 4u * (unsigned)(col + ps * rb) : kOutOfRange;
      bst<kBufNT>(sLm, ob, ln1);
      bst<kBufNT>(sMm, ob, mn1);
      // double-buffered by row parity: a wave can only overwrite this buffer two rows later, after every wave has
      // passed the next row's barrier
      float(*gs)[6][64] = gsum[rb & 1];
      gs[wave][0][lane] = own ? dt2 * dl[0] * qa0 : 0.f;
      gs[wave][1][lane] = own ? dt2 * dm[0] * qa0 : 0.f;
      gs[wave][2][lane] = own ? dt2 * (dl[0] + dm[0]) * qb0 : 0.f;
      gs[wave][3][lane] = own ? dt2 * ln1 * qa1 : 0.f;
      gs[wave][4][lane] = own ? dt2 * mn1 * qa1 : 0.f;
      gs[wave][5][lane] = own ? dt2 * (ln1 + mn1) * qb1 : 0.f;
      __syncthreads();
      const float s0 = ((gs[0][fw][lane] + gs[1][fw][lane]) + gs[2][fw][lane]) + gs[3][fw][lane];
      const float s1 = ((gs[0][3 + fw][lane] + gs[1][3 + fw][lane]) + gs[2][3 + fw][lane]) + gs[3][3 + fw][lane];
      bst<kBufNT>(grs, gcol + grow * rb, (gv + s0) + s1);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      gq1[k] = gq1[k + 1];
      gq0[k] = gq0[k + 1];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dl[k] = dl[k + 1];
      dm[k] = dm[k + 1];
      dgo[k] = dgo[k + 1];
      dvp[k] = dvp[k + 1];
    }
    dl[3] = ln;
    dm[3] = mn;
    dgo[3] = go0;
    dvp[3] = cur.vp4;
  };
  using Pro = std::integral_constant<bool, false>;
  using Main = std::integral_constant<bool, true>;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int r = z0 - 8 + k;
    const unsigned o = (unsigned)(cc + ps * r);
    const float l = ldu(L1, o), m = ldu(M1, o);
    ring[2 * (r & 7)][lane] = l;
    ring[2 * (r & 7) + 1][lane] = m;
    gq1[k] = dt2 * (ldu(vp, (unsigned)(cc + nxm * r)) * (l + m));
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) dvp[k] = ldu(vp, (unsigned)(cc + nxm * (z0 - 4 + k)));
  // four row buffers in rotation, two rows in flight: the 8 rows of step k above the first output row, the first
  // four main rows peeled (so the loop is entered as it repeats), the main rows unrolled by four, the last ones
  Row b0, b1, b2, b3;
  load_row(z0 - 4, b0);
  load_row(z0 - 3, b1);
  iter(z0 - 4, b0, b2, Pro{});
  iter(z0 - 3, b1, b3, Pro{});
  iter(z0 - 2, b2, b0, Pro{});
  iter(z0 - 1, b3, b1, Pro{});
  iter(z0, b0, b2, Pro{});
  iter(z0 + 1, b1, b3, Pro{});
  iter(z0 + 2, b2, b0, Pro{});
  iter(z0 + 3, b3, b1, Pro{});
  int z = z0 + 4;
  if (z + 3 < zend) {
    iter(z, b0, b2, Main{});
    iter(z + 1, b1, b3, Main{});
    iter(z + 2, b2, b0, Main{});
    iter(z + 3, b3, b1, Main{});
    for (z += 4; z + 3 < zend; z += 4) {
      iter(z, b0, b2, Main{});
      iter(z + 1, b1, b3, Main{});
      iter(z + 2, b2, b0, Main{});
      iter(z + 3, b3, b1, Main{});
    }
  }
  if (z < zend) iter(z, b0, b2, Main{});
  if (z + 1 < zend) iter(z + 1, b1, b3, Main{});
  if (z + 2 < zend) iter(z + 2, b2, b0, Main{});
}

// Grid: [0, nshell_wg) shell workgroups (step k, four single-step items each), then one workgroup per core tile
// (four items).  Each part is dealt to the XCDs in contiguous logical ranges of its own.  FP32 increment-form
// arithmetic (the band rows are shell rows), gradient groups, imaging terms for both steps.
__global__ void __launch_bounds__(256) k_adj_pair(AdjPairArgs a) {
  __shared__ float xbuf[4][152];
  __shared__ float gsum[2][4][6][64];
  __shared__ float ring[4][16][64];
  const int wave = wave_id();
  const int bid = (int)blockIdx.x;
  if (bid < a.nshell_wg) {
    adj_ex_item<float, true, true, true, true>(a.sh, xcd_logical_block(bid, a.nshell_wg) * 4 + wave, xbuf[wave],
                                               reinterpret_cast<float(*)[4][3][64]>(&gsum[0][0][0][0]));
    return;
  }
  const int b = xcd_logical_block(bid - a.nshell_wg, (int)gridDim.x - a.nshell_wg);
  if (4 * b >= a.nitems) return;
  adj_pair_core(a, a.items[4 * b + wave], xbuf[wave], gsum, ring[wave]);
}

// Step k-1 of a step pair's shell: the single-step body on items with column ranges.
__global__ void __launch_bounds__(256) k_adj_shell(AdjExArgs a) {
  __shared__ float xbuf[4][152];
  __shared__ float gsum[2][4][3][64];
  const int wave = wave_id();
  adj_ex_item<float, true, true, true, true>(a, xcd_logical_block(blockIdx.x, gridDim.x) * 4 + wave, xbuf[wave],
                                             gsum);
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
  const long i = sd.off + ix + (long)sd.ps * iz;
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


// ---------------------------------------------------------------------------------------------
// 16-bit checkpoints with per-snapshot scaling.  Each of a shot's eight checkpoint fields (four cell
// fields, four C-PML memory variables on the strip cells) is stored in 16 bits after an exact
// power-of-two scaling by its max-abs over that shot at that step (2^e <= max|x| < 2^(e+1); e = 0 for an
// all-zero field), rounded to nearest even:
//   FMT 1 (BF16): bf16(x * 2^-e), so a restored value is x (1 + d), |d| <= 2^-9;
//   FMT 2 (FP16): fp16(x * 2^(14-e)): the field's maximum lands in [2^14, 2^15), |d| <= 2^-11 down to
//   2^-29 of the maximum (FP16 subnormals below that).
// ---------------------------------------------------------------------------------------------

__device__ inline unsigned short bf16_rne(float v) {
  unsigned u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_float(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
// scale exponent of a field from the bits of its max-abs
__device__ inline int ck_exponent(unsigned amax_bits) {
  const int be = (int)((amax_bits >> 23) & 0xFFu);
  return amax_bits == 0u ? 0 : (be == 0 ? -126 : be - 127);
}
template <int FMT>
__device__ inline unsigned short ck_pack(float v, int e) {
  if (FMT == 1) return bf16_rne(ldexpf(v, -e));
  const _Float16 h = (_Float16)ldexpf(v, 14 - e);
  return __builtin_bit_cast(unsigned short, h);
}
template <int FMT>
__device__ inline float ck_unpack(unsigned short h, int e) {
  if (FMT == 1) return ldexpf(bf16_to_float(h), e);
  return ldexpf((float)__builtin_bit_cast(_Float16, h), e - 14);
}

__device__ inline void wave_max_atomic(float m, unsigned* dst) {
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(dst, __float_as_uint(m));  // non-negative floats order as their bits
}

__global__ void __launch_bounds__(256) k_fill_u32(unsigned* __restrict__ p, unsigned v, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// max|x| of four cell fields per shot: amax[shot * 8 + k], k = 0..3.  Grid (blocks per shot, shots).
__global__ void __launch_bounds__(256) k_ck_absmax_cells(const ShotDev* __restrict__ shots, int nz, const float* f0,
                                                         const float* f1, const float* f2, const float* f3,
                                                         unsigned* __restrict__ amax) {
  const ShotDev sd = shots[blockIdx.y];
  const long long n = (long long)sd.ps * nz;  // padded columns included (they hold zeros)
  float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const long long c = sd.off + i;
    m0 = fmaxf(m0, fabsf(f0[c]));
    m1 = fmaxf(m1, fabsf(f1[c]));
    m2 = fmaxf(m2, fabsf(f2[c]));
    m3 = fmaxf(m3, fabsf(f3[c]));
  }
  unsigned* a = amax + blockIdx.y * 8;
  wave_max_atomic(m0, a + 0);
  wave_max_atomic(m1, a + 1);
  wave_max_atomic(m2, a + 2);
  wave_max_atomic(m3, a + 3);
}

// max|x| of four C-PML fields over each shot's strip cells: amax[shot * 8 + 4 + k].
__global__ void __launch_bounds__(256) k_ck_absmax_strips(StripArgs a, const float* f0, const float* f1,
                                                          const float* f2, const float* f3, unsigned* __restrict__ amax) {
  const int2 blk = a.blocks[blockIdx.x];
  const ShotDev sd = a.g.shots[blk.x];
  const int k = blk.y + threadIdx.x;
  float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
  if (k < sd.nstrip) {
    int ix, iz;
    strip_cell(k, sd.nx, a.g.nz, a.g.cpml, ix, iz);
    const long long i = sd.off + ix + (long long)sd.ps * iz;
    m0 = fabsf(f0[i]);
    m1 = fabsf(f1[i]);
    m2 = fabsf(f2[i]);
    m3 = fabsf(f3[i]);
  }
  unsigned* d = amax + blk.x * 8 + 4;
  wave_max_atomic(m0, d + 0);
  wave_max_atomic(m1, d + 1);
  wave_max_atomic(m2, d + 2);
  wave_max_atomic(m3, d + 3);
}

// Pack (dir 0) or unpack (dir 1) four cell fields: packed[k * ncell_total + off + i].
template <int FMT>
__global__ void __launch_bounds__(256) k_ck_pack_cells(const ShotDev* __restrict__ shots, int nz, int dir, float* f0,
                                                       float* f1, float* f2, float* f3,
                                                       unsigned short* __restrict__ packed, long long ncell_total,
                                                       const unsigned* __restrict__ amax) {
  const ShotDev sd = shots[blockIdx.y];
  const long long n = (long long)sd.ps * nz;  // padded columns included (they hold zeros)
  const unsigned* a = amax + blockIdx.y * 8;
  const int e0 = ck_exponent(a[0]), e1 = ck_exponent(a[1]), e2 = ck_exponent(a[2]), e3 = ck_exponent(a[3]);
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const long long c = sd.off + i;
    if (dir == 0) {
      packed[c] = ck_pack<FMT>(f0[c], e0);
      packed[c + ncell_total] = ck_pack<FMT>(f1[c], e1);
      packed[c + 2 * ncell_total] = ck_pack<FMT>(f2[c], e2);
      packed[c + 3 * ncell_total] = ck_pack<FMT>(f3[c], e3);
    } else {
      f0[c] = ck_unpack<FMT>(packed[c], e0);
      f1[c] = ck_unpack<FMT>(packed[c + ncell_total], e1);
      f2[c] = ck_unpack<FMT>(packed[c + 2 * ncell_total], e2);
      f3[c] = ck_unpack<FMT>(packed[c + 3 * ncell_total], e3);
    }
  }
}

// Pack / unpack four C-PML fields at the strip cells: packed[k * nstrip_total + soff + cell].
template <int FMT>
__global__ void __launch_bounds__(256) k_ck_pack_strips(StripArgs a, int dir, float* f0, float* f1, float* f2, float* f3,
                                                        unsigned short* __restrict__ packed, long long nstrip_total,
                                                        const unsigned* __restrict__ amax) {
  const int2 blk = a.blocks[blockIdx.x];
  const ShotDev sd = a.g.shots[blk.x];
  const int k = blk.y + threadIdx.x;
  if (k >= sd.nstrip) return;
  const unsigned* am = amax + blk.x * 8 + 4;
  const int e0 = ck_exponent(am[0]), e1 = ck_exponent(am[1]), e2 = ck_exponent(am[2]), e3 = ck_exponent(am[3]);
  int ix, iz;
  strip_cell(k, sd.nx, a.g.nz, a.g.cpml, ix, iz);
  const long long i = sd.off + ix + (long long)sd.ps * iz;
  const long long j = sd.soff + k;
  if (dir == 0) {
    packed[j] = ck_pack<FMT>(f0[i], e0);
    packed[j + nstrip_total] = ck_pack<FMT>(f1[i], e1);
    packed[j + 2 * nstrip_total] = ck_pack<FMT>(f2[i], e2);
    packed[j + 3 * nstrip_total] = ck_pack<FMT>(f3[i], e3);
  } else {
    f0[i] = ck_unpack<FMT>(packed[j], e0);
    f1[i] = ck_unpack<FMT>(packed[j + nstrip_total], e1);
    f2[i] = ck_unpack<FMT>(packed[j + 2 * nstrip_total], e2);
    f3[i] = ck_unpack<FMT>(packed[j + 3 * nstrip_total], e3);
  }
}

// Clears and copies of the batched path as kernels: they can be captured into a graph, where memset
// nodes are avoided.
__global__ void __launch_bounds__(256) k_zero_f32(float* __restrict__ p, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = 0.f;
}
__global__ void __launch_bounds__(256) k_copy_f32(float* __restrict__ d, const float* __restrict__ s, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
}

// Residual res = syn - obs of every shot of a batch and per-(shot, block) FP64 partial sums of res^2
// (parallelcalls.cu:133-142, main.c:448-454).  Block (b, s) covers a fixed slice of shot s's gather, so
// the misfit of a shot does not depend on which other shots share its batch.
__global__ void __launch_bounds__(256) k_residual(const ShotDev* __restrict__ shots, int nt,
                                                  const float* __restrict__ syn, const float* __restrict__ obs,
                                                  float* __restrict__ res, double* __restrict__ partial) {
  __shared__ double red[4];
  const ShotDev sd = shots[blockIdx.y];
  const long long n = (long long)nt * sd.nrec;
  double acc = 0.0;
  for (long long j = blockIdx.x * 256LL + threadIdx.x; j < n; j += (long long)gridDim.x * 256) {
    const long long i = sd.goff + j;
    const float r = syn[i] - obs[i];
    if (res) res[i] = r;
    acc += (double)r * (double)r;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.y * (long long)gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Per-shot misfits scale * sum(partials of the shot) in a fixed tree, then added to out[0] shot by shot
// in batch order (accumulate = 0 starts from zero).  shot_out (may be null) receives each shot's misfit.
__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ partial, int per_shot, int nshots,
                                                      double scale, double* __restrict__ out, int accumulate,
                                                      double* __restrict__ shot_out) {
  __shared__ double red[4];
  double total = accumulate ? out[0] : 0.0;
  for (int s = 0; s < nshots; ++s) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < per_shot; i += 256) acc += partial[(long long)s * per_shot + i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    const double v = scale * ((red[0] + red[1]) + (red[2] + red[3]));
    total += v;
    if (threadIdx.x == 0 && shot_out) shot_out[s] = v;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = total;
}

// FP64 cross-shot sum of the batch's FP32 gradient slots (one per gradient group, over the union of its
// windows, flagged on the group's leader) into the model gradient (3 fields).  Slots are added in the batch
// order of their leaders for a deterministic result.
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
    const int lx = x - sd.gx0;
    if (!sd.lead || lx < 0 || lx >= sd.gnx) continue;
    const long long i = sd.gslot + lx + (long long)sd.gps * z;
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
