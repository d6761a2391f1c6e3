/* fwi_main.c -- the reference driver's control flow (src/main.c) on the fwi_amd C ABI.
 *
 * Differences from src/main.c, all at the boundary the reference crossed by hand:
 *   - one process drives one GPU and every shot (main.c:21-31 distributed shots over MPI ranks and summed
 *     with MPI_Reduce; here fwi_gradient batches the shots and sums them on the device in FP64);
 *   - the per-step launchers and the host snapshot buffer PF_h (main.c:73-74, 422-492) are replaced by one
 *     fwi_gradient call per iteration (checkpointed recomputation on the device);
 *   - misfit, norms and L-BFGS dot products are FP64 device reductions (fwi_lbfgs_*), not host float sums;
 *   - every call is checked; the parameters come from the command line instead of #defines.
 * Geometry, C-PML parameters, wavelet, stage schedule (f += 3), steepest-descent step lengths and the
 * line-search rules are the reference's (main.c:38-62, 260-262, 275-333, 539-806).
 *
 * Usage: fwi_main --dir DIR [--nx 841 --nz 270 --dh 20 --dt 2e-3 --tend 5 --cpml 20 --nshots 15 --f 3
 *                  --stages 2 --iterations 100 --mode exact|refcompat --vmin V --vmax V]
 * --vmin/--vmax (m/s, optional) clamp every updated model with fwi_model_clamp; without them the update is
 * the reference's, whose only bound is v_hor >= v_nmo.
 * DIR holds vp_ori.bin vh_ori.bin vn_ori.bin vp_ini.bin vh_ini.bin vn_ini.bin (raw float32, x fastest:
 * vp, epsilon, delta); results are written to DIR/v_hor.bin, v_po.bin, v_nmo.bin.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fwi.h"
#include "fwi_serialcalls.h"

#define PI 3.141592653589793

static void die(const char* what) {
  fprintf(stderr, "fwi_main: %s: %s\n", what, fwi_error_message());
  exit(1);
}
static void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "fwi_main: %s: %s\n", what, hipGetErrorString(e));
    exit(1);
  }
}

static float* read_raw(const char* dir, const char* name, int n) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE* f = fopen(path, "rb");
  if (!f) {
    fprintf(stderr, "fwi_main: cannot open %s\n", path);
    exit(1);
  }
  float* a = (float*)malloc((size_t)n * sizeof(float));
  if (fread(a, sizeof(float), (size_t)n, f) != (size_t)n) {
    fprintf(stderr, "fwi_main: %s is shorter than nx*nz floats\n", path);
    exit(1);
  }
  fclose(f);
  return a;
}

static void write_raw(const char* dir, const char* name, const float* dev, int n) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  float* h = (float*)malloc((size_t)n * sizeof(float));
  hip_ok(hipMemcpy(h, dev, (size_t)n * sizeof(float), hipMemcpyDeviceToHost), "copy result");
  FILE* f = fopen(path, "wb");
  if (!f || fwrite(h, sizeof(float), (size_t)n, f) != (size_t)n) {
    fprintf(stderr, "fwi_main: cannot write %s\n", path);
    exit(1);
  }
  fclose(f);
  free(h);
}

/* Device copy of a host field. */
static float* to_device(const float* h, int n) {
  float* d = NULL;
  hip_ok(hipMalloc((void**)&d, (size_t)n * sizeof(float)), "hipMalloc");
  hip_ok(hipMemcpy(d, h, (size_t)n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  return d;
}

int main(int argc, char** argv) {
  /* main.c:38-62 */
  int CPML = 20, nx = 841, nz = 270, nshots = 15, stages = 2, iterations = 100, exact = 1;
  float dh = 20, dt = 2e-3f, tend = 5, f = 3;
  const char* dir = NULL;
  float vmin = -1, vmax = -1;
  for (int i = 1; i + 1 < argc; i += 2) {
    const char* k = argv[i];
    const char* v = argv[i + 1];
    if (!strcmp(k, "--dir")) dir = v;
    else if (!strcmp(k, "--nx")) nx = atoi(v);
    else if (!strcmp(k, "--nz")) nz = atoi(v);
    else if (!strcmp(k, "--dh")) dh = (float)atof(v);
    else if (!strcmp(k, "--dt")) dt = (float)atof(v);
    else if (!strcmp(k, "--tend")) tend = (float)atof(v);
    else if (!strcmp(k, "--cpml")) CPML = atoi(v);
    else if (!strcmp(k, "--nshots")) nshots = atoi(v);
    else if (!strcmp(k, "--f")) f = (float)atof(v);
    else if (!strcmp(k, "--stages")) stages = atoi(v);
    else if (!strcmp(k, "--iterations")) iterations = atoi(v);
    else if (!strcmp(k, "--mode")) exact = strcmp(v, "refcompat") != 0;
    else if (!strcmp(k, "--vmin")) vmin = (float)atof(v);
    else if (!strcmp(k, "--vmax")) vmax = (float)atof(v);
    else {
      fprintf(stderr, "fwi_main: unknown option %s\n", k);
      return 2;
    }
  }
  if (!dir) {
    fprintf(stderr, "usage: fwi_main --dir DIR [options]\n");
    return 2;
  }
  const int nt = (int)ceil(tend / dt);
  const int sz = 5;
  const float R = 1e-6f, L = CPML * dh, d0 = -3 * log(R) / (2 * L), Vmax = 6000;
  const int space = (int)floor(((float)(nx - 2 * CPML - 1)) / (float)nshots);
  const int n = nx * nz;

  /* main.c:101-123 */
  float* vel_h = read_raw(dir, "vp_ori.bin", n);
  float* eps_h = read_raw(dir, "vh_ori.bin", n);
  float* del_h = read_raw(dir, "vn_ori.bin", n);
  float* veli_h = read_raw(dir, "vp_ini.bin", n);
  float* epsi_h = read_raw(dir, "vh_ini.bin", n);
  float* deli_h = read_raw(dir, "vn_ini.bin", n);
  float *vel = to_device(vel_h, n), *eps = to_device(eps_h, n), *del = to_device(del_h, n);
  float *veli = to_device(veli_h, n), *epsi = to_device(epsi_h, n), *deli = to_device(deli_h, n);
  float *t_hor, *t_po, *t_nmo, *v_hor, *v_po, *v_nmo;
  hip_ok(hipMalloc((void**)&t_hor, n * sizeof(float)), "hipMalloc");
  hip_ok(hipMalloc((void**)&t_po, n * sizeof(float)), "hipMalloc");
  hip_ok(hipMalloc((void**)&t_nmo, n * sizeof(float)), "hipMalloc");
  hip_ok(hipMalloc((void**)&v_hor, n * sizeof(float)), "hipMalloc");
  hip_ok(hipMalloc((void**)&v_po, n * sizeof(float)), "hipMalloc");
  hip_ok(hipMalloc((void**)&v_nmo, n * sizeof(float)), "hipMalloc");
  pre_variables_h(vel, eps, del, t_hor, t_po, t_nmo, nx, nz);
  pre_variables_h(veli, epsi, deli, v_hor, v_po, v_nmo, nx, nz);
  if (fwi_last_error()) die("pre_variables_h");

  /* shots (main.c:260-262): 1-based sx = CPML + i*space, depth sz; receivers on row sz-1, live columns */
  fwi_shot* shots = (fwi_shot*)calloc((size_t)nshots + 1, sizeof(fwi_shot));
  for (int i = 0; i <= nshots; ++i) {
    shots[i].x0 = 0;
    shots[i].nx = nx;
    shots[i].isx = CPML + i * space - 1;
    shots[i].isz = sz - 1;
    shots[i].irz = sz - 1;
    shots[i].ir0 = CPML;
    shots[i].nrec = nx - 2 * CPML;
  }
  float* src = (float*)malloc((size_t)nt * sizeof(float));
  float* ta = (float*)malloc((size_t)CPML * sizeof(float));
  float* tb = (float*)malloc((size_t)CPML * sizeof(float));
  double* grad = NULL;
  float* obs = NULL;
  hip_ok(hipMalloc((void**)&grad, 3 * (size_t)n * sizeof(double)), "hipMalloc(grad)");
  fwi_lbfgs* opt = NULL;

  for (int scale = 0; scale < stages; ++scale) {
    /* main.c:306-317: the L-BFGS history restarts with every stage */
    if (opt) fwi_lbfgs_destroy(opt);
    if (fwi_lbfgs_create(nx, nz, 5, &opt)) die("fwi_lbfgs_create");
    /* main.c:318-333: C-PML and wavelet of the stage frequency */
    fwi_cpml_table(CPML, Vmax, dt, dh, f, d0, L, ta, tb);
    float t = 0;
    for (int it = 0; it < nt; ++it) {
      t += dt;
      src[it] = (float)(1 * (1 - 2 * pow(PI * f * (t - (1 / f)), 2)) * exp(-pow(PI * f * (t - (1 / f)), 2)));
    }
    fwi_survey sv = {nx, nz, nt, CPML, dh, dt, nshots + 1, shots, src, ta, tb};
    fwi_options o;
    fwi_options_default(&o);
    o.grad_mode = exact ? FWI_GRAD_EXACT : FWI_GRAD_REFCOMPAT;
    fwi_plan* plan = NULL;
    if (fwi_plan_create(&sv, &o, &plan)) die("fwi_plan_create");
    if (!obs) hip_ok(hipMalloc((void**)&obs, (size_t)fwi_plan_gather_size(plan) * sizeof(float)), "hipMalloc(obs)");
    /* observed data (main.c:334-366) */
    if (fwi_forward(plan, t_hor, t_po, t_nmo, obs, NULL)) die("fwi_forward");
    printf("stage %d: f = %g Hz, %d shots, nt = %d\n", scale, f, nshots + 1, nt);
    fwi_line_search ls; /* main.c:306-311: step lengths and failure count reset per stage */
    fwi_line_search_init(&ls);
    for (int it_n = 0; it_n < iterations; ++it_n) {
      double phi = 0;
      if (fwi_gradient(plan, v_hor, v_po, v_nmo, obs, grad, &phi, NULL, NULL)) die("fwi_gradient");
      printf("Error total : %.17g\n", phi);
      fflush(stdout);
      if (it_n < 1) { /* main.c:539-568 */
        const double sd[3] = {3e5, 2e5, 2.5e5};
        if (fwi_lbfgs_sd_step(opt, v_hor, v_po, v_nmo, grad, sd, NULL)) die("fwi_lbfgs_sd_step");
        if (vmax > 0 && fwi_model_clamp(v_hor, v_po, v_nmo, n, vmin > 0 ? vmin * vmin : 0.f, vmax * vmax, NULL))
          die("fwi_model_clamp");
        continue;
      }
      if (fwi_lbfgs_direction(opt, v_hor, v_po, v_nmo, grad, NULL)) die("fwi_lbfgs_direction");
      /* line search (main.c:690-786) */
      int decision = 1, stop = 0;
      while (decision == 1) {
        double trial = 0;
        if (fwi_lbfgs_trial(opt, v_hor, v_po, v_nmo, ls.alpha, NULL)) die("fwi_lbfgs_trial");
        if (vmax > 0 && fwi_model_clamp(v_hor, v_po, v_nmo, n, vmin > 0 ? vmin * vmin : 0.f, vmax * vmax, NULL))
          die("fwi_model_clamp");
        if (fwi_misfit(plan, v_hor, v_po, v_nmo, obs, &trial, NULL)) die("fwi_misfit");
        decision = fwi_line_search_step(&ls, phi, &trial);
        printf("trial misfit : %.17g\n", trial);
        if (decision == 1) printf("new alpha = %f ....\n", ls.alpha[0]);
        if (decision == 2) {
          if (fwi_lbfgs_restore(opt, v_hor, v_po, v_nmo, NULL)) die("fwi_lbfgs_restore");
          stop = 1;
        }
      }
      if (stop) break;
    }
    fwi_plan_destroy(plan);
    f += 3; /* main.c:802 */
  }
  hip_ok(hipDeviceSynchronize(), "synchronize");
  write_raw(dir, "v_hor.bin", v_hor, n);
  write_raw(dir, "v_po.bin", v_po, n);
  write_raw(dir, "v_nmo.bin", v_nmo, n);
  fwi_lbfgs_destroy(opt);
  printf("done\n");
  return 0;
}
