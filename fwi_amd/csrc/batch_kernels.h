// Device-side descriptors and kernel parameter blocks of the batched FWI path.
#pragma once

#include <hip/hip_runtime.h>

namespace fwi {

// One shot of a batch as the kernels see it.
struct ShotDev {
  int nx;         // window width (cells)
  int ps;         // row pitch of the shot's fields: nx rounded up to whole 128-byte lines (32 floats)
  int x0;         // model column of window column 0
  int isx, isz;   // source cell
  int irz, ir0;   // receiver row and first receiver column
  int nrec;       // receivers
  int nstrip;     // C-PML cells of the window (chi_x U chi_z)
  long long off;  // offset of the shot's ps*nz fields in the batch field arrays (cells)
  long long soff; // offset of the shot's strip cells in packed strip arrays
  long long goff; // offset of the shot's gather [nt][nrec] in the batch gather arrays
  long long gslot;  // offset of the FP32 gradient slot of the shot's gradient group
  int gx0, gnx;     // the slot covers model columns [gx0, gx0 + gnx) (the union of the group's windows)
  int gps;          // row pitch of the slot (gnx rounded up to 32)
  int lead;         // 1 if the shot leads its gradient group (its slot is summed into the model gradient)
  long long boff;   // offset of the shot's band rows (band x ps cells) in the FP64 near-surface band arrays
};

// C-PML cells of a window: the side strips Omega & (ix < cpml | ix > nx-cpml-1) for rows 4..nz-5, then
// the bottom strip rows nz-cpml..nz-5 for the middle columns cpml..nx-cpml-1.  Enumerated row by row.
__host__ __device__ inline int strip_count(int nx, int nz, int cpml) {
  int w = cpml - 4;
  return 2 * w * (nz - 8) + (nx - 2 * cpml) * (cpml - 4);
}
__host__ __device__ inline void strip_cell(int k, int nx, int nz, int cpml, int& ix, int& iz) {
  int w = cpml - 4;
  int nside = 2 * w * (nz - 8);
  if (k < nside) {
    int row = k / (2 * w), j = k - row * 2 * w;
    iz = 4 + row;
    ix = j < w ? 4 + j : nx - cpml + (j - w);
  } else {
    int kk = k - nside, nmid = nx - 2 * cpml;
    int row = kk / nmid;
    iz = nz - cpml + row;
    ix = cpml + (kk - row * nmid);
  }
}

struct ModelArgs {
  const float* vh;
  const float* vp;
  const float* vn;
  int nxm;  // model width
};

struct GridArgs {
  const ShotDev* shots;
  int nz, cpml;
  float dh, dh2;
  const float* ta;  // C-PML a by distance from the edge
  const float* tb;  // C-PML b
};

// FP64 near-surface band of the increment-form forward: rows [4, band) of every shot advance in FP64 (two levels
// of P and F and of psi_x by step parity, zeta_x on the side C-PML cells); one workgroup per (256-column chunk,
// band row, shot).
struct BandArgs {
  GridArgs g;
  ModelArgs m;
  int band;           // first row of the float state
  double dt;          // the float time step, promoted
  const float* src;
  int it;
  const double* P;    // P^n, F^n (band arrays of this step's parity)
  double* Pm;         // P^{n-1}, F^{n-1}: receive P^{n+1}, F^{n+1}
  const double* F;
  double* Fm;
  const double* psx;  // psi_x^n; psi_x^{n+1} goes to psxn (two arrays by step parity)
  double* psxn;
  double* zx;
  const float* Fc;    // float F^n (rows band.. are the float state's)
  const float* Pcf;   // float P^n (reference-compatible imaging terms)
  float* Po;          // float P^{n+1}, F^{n+1}: the band's values rounded
  float* Fo;
  float* gather;
  float* qa;
  float* qb;
};

// Forward step.  P/F "cur" hold step n, "old" hold n-1 and receive n+1.  In the increment form "old"
// only receives n+1, and Q, R (the increments P^n - P^{n-1}, F^n - F^{n-1}) are read from Qc, Rc and
// written to Qo, Ro.
// The C-PML memory variables psi_x, psi_z are kept by step parity: step n reads psi^n (psx, psz) and writes psi^{n+1}
// (psxn, pszn) on the C-PML cells it updates.  Workgroups [0, nband_wg) of the launch advance the FP64 near-surface
// band of the same step (band; nbc 256-column chunks per band row), the rest the items.
struct FwdArgs {
  GridArgs g;
  ModelArgs m;
  const int4* items;  // (shot, band x0, z0, z1)
  int nitems;
  float C;            // dt^2/dh^2
  const float* src;
  const float* Pc;
  float* Po;
  const float* Fc;
  float* Fo;
  const float* psx;
  const float* psz;
  float* psxn;
  float* pszn;
  float* zx;
  float* zz;
  const float* Qc;  // increment form: Q, R of step n (read) and of step n+1 (written)
  const float* Rc;
  float* Qo;
  float* Ro;
  float* gather;  // batch gathers or null
  float* qa;      // per-step store or null
  float* qb;
  int it;
  const int2* xr; // per-item column range [x, y) the item updates, or null (its whole band)
  BandArgs band;  // the FP64 band's step (nband_wg = 0: none)
  int nband_wg;   // a multiple of 8, of which nband_real band workgroups
  int nband_real;
  int nbc;
};

// Two forward steps n, n+1 of the increment form.  Core items: (shot, first of up to 224 output columns, z0, z1)
// within the core (columns [cpml+4, nx-cpml-4) rounded inward to multiples of 32, rows [zfirst+4, nz-cpml-4)); the
// launch reads the set of step n and writes the set of step n+2 on the core and (sh: single-step arguments with their
// items and column ranges) the set of step n+1 on the FP64 band, the shell and the core's cells within 4 of it.
struct FwdPairArgs {
  GridArgs g;
  ModelArgs m;
  const int4* items;
  int nitems;
  float C;
  const float* src;
  const float* P0;
  const float* F0;
  const float* Q0;
  const float* R0;
  float* P2;
  float* F2;
  float* Q2;
  float* R2;
  float* gather;     // batch gathers or null
  float* qa0;        // imaging terms (QM 1: a, b; QM 2: p2, r2): a, b of step n at qa0, qa0 + qstride bytes, of
  unsigned qstride;  // step n+1 at qa0 + 2 qstride, qa0 + 3 qstride; qa0 null when QM = 0
  int it;            // n
  FwdArgs sh;        // step n of the shell and the core's cells next to it, and of the FP64 band
  int nshell_wg;     // shell workgroups (a multiple of 8), after the sh.nband_wg band workgroups
};

// Strip (C-PML) blocks: (shot, first strip cell) per 256-thread block.
struct StripArgs {
  GridArgs g;
  ModelArgs m;
  const int2* blocks;
  int nblocks;
};

// Adjoint step, reference-compatible (4 auxiliary fields, forward-recursion C-PML).
struct AdjRefArgs {
  GridArgs g;
  ModelArgs m;
  const int4* items;
  int nitems;
  float dh2C;  // dh2*C
  float dt;
  const float* Lc;  // lambda_P cur
  float* Lo;
  const float* Mc;  // lambda_F cur
  float* Mo;
  float* s1;
  float* s2;
  float* s3;
  float* s4;
  float* z1;
  float* z2;
  float* z3;
  float* z4;
  const float* res;  // batch residual gathers
  int rit;           // residual time index injected at this step
  const float* qa;   // forward D2x P / D2z F of the paired step, or null
  const float* qb;
  float* gcx;
  float* gcz;
  float* gdx;
};

// Adjoint step, exact discrete transpose.
struct AdjExArgs {
  GridArgs g;
  ModelArgs m;
  const int4* items;
  int nitems;
  float dt2;
  const float* Lc;  // mu_P^{k+1}
  const float* Lp;  // mu_P^{k+2} (read)
  float* Lo;        // mu_P^k (written; may be Lp)
  const float* Mc;  // mu_F alike
  const float* Mp;
  float* Mo;
  const float* sxo;  // C-PML memory variables of step k+1: psi-bar x, zeta-bar x, psi-bar z, zeta-bar z
  const float* zxo;
  const float* szo;
  const float* zzo;
  float* sxn;        // of step k (written on the C-PML cells only; zero elsewhere since the pass began)
  float* zxn;
  float* szn;
  float* zzn;
  float* DP;   // increment form (rows above `band`): D = m^{k+1} - m^{k+2}, updated in place to m^k - m^{k+1}
  float* DF;
  int band;    // rows [4, band) take the increment form (items are cut at band); 0: none
  const float* res;
  int rit;
  const float* qa;  // forward a^n, b^n of the paired step, or null
  const float* qb;
  float* gh;
  float* gp;
  float* gn;
  const int2* xr;  // shell items of a step pair: each item's column range (local to its shot), else null
};

// Two exact adjoint steps k, k-1 per launch on every gradient group's core (no C-PML term, no band row within
// reach), with step k of the shell (sh, single-step items with column ranges) in the first workgroups.
struct AdjPairArgs {
  GridArgs g;
  ModelArgs m;
  const int4* items;  // four per workgroup: (member or -(leader+1), local first output column | columns << 16, z0, z1)
  int nitems;
  float dt2;
  const float* L1;  // mu_P^{k+1}, mu_F^{k+1}
  const float* M1;
  const float* L2;  // mu^{k+2}
  const float* M2;
  float* L0;        // mu^k (written)
  float* M0;
  float* Lm;        // mu^{k-1} (written)
  float* Mm;
  const float* res;
  int rit;          // residual time of step k (step k-1: rit - 1)
  const float* qa0;  // imaging terms of step k, of step k-1
  const float* qb0;
  const float* qa1;
  const float* qb1;
  float* gh;
  float* gp;
  float* gn;
  AdjExArgs sh;
  int nshell_wg;
};

}  // namespace fwi
