// Error bookkeeping and host-side setup helpers.
#include <cmath>
#include <string>
#include <vector>

#include "../../include/fwi.h"
#include "fwi_internal.h"

namespace fwi {

namespace {
thread_local int g_last_error = 0;
thread_local std::string g_message;
constexpr double kPi = 3.141592653589793;  // include/parallelcalls.h:3
}  // namespace

void record_error(hipError_t e) {
  if (e != hipSuccess && g_last_error == 0) g_last_error = (int)e;
}

void set_message(const std::string& msg) { g_message = msg; }

int check_host(const char* fn, std::initializer_list<ArrayArg> args) {
  for (const ArrayArg& a : args)
    if (!a.ptr) {
      set_message(std::string(fn) + ": " + a.name + " is null");
      return 1;
    }
  return 0;
}

int check_arrays(const char* fn, std::initializer_list<ArrayArg> args) {
  if (int r = check_host(fn, args)) return r;
  for (const ArrayArg& a : args) {
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, a.ptr);
    if (e != hipSuccess) {
      // fail closed: memory the runtime cannot classify never reaches a kernel
      (void)hipGetLastError();
      set_message(std::string(fn) + ": " + a.name + " is memory the runtime cannot classify (hipPointerGetAttributes: " +
                  hipGetErrorString(e) + "); pass device, managed or pinned memory");
      return 1;
    }
    if (at.type == hipMemoryTypeUnregistered || (at.type == hipMemoryTypeHost && !at.devicePointer)) {
      set_message(std::string(fn) + ": " + a.name +
                  " is host memory the GPU cannot address (pass device, managed or pinned memory)");
      return 1;
    }
  }
  return 0;
}

void cpml_table_host(int cpml, float vmax, float dt, float dh, float f, float d0, float L, std::vector<float>& a,
                     std::vector<float>& b) {
  a.assign((size_t)(cpml > 0 ? cpml : 0), 0.f);
  b.assign(a.size(), 0.f);
  for (int e = 0; e < cpml; ++e) {
    double l = (double)(cpml - e) * (double)dh;
    double LL = (double)L;
    double d = (double)d0 * (double)vmax * (l / LL) * (l / LL);
    double al = kPi * (double)f * (LL - l) / LL;
    float bf = (float)std::exp(-(d + al) * (double)dt);
    b[e] = bf;
    a[e] = (float)(d * ((double)bf - 1) / (d + al));
  }
}

}  // namespace fwi

extern "C" {

int fwi_last_error(void) {
  int e = fwi::g_last_error;
  fwi::g_last_error = 0;
  return e;
}

const char* fwi_error_message(void) { return fwi::g_message.c_str(); }

// The reference's backtracking line search (src/main.c:690-786) as a state machine, shared by the C and Python
// drivers: step lengths start at 1 per stage (main.c:308-310) and are only ever halved; failures accumulate over
// the stage; a NaN misfit counts as 1e30 and as a failure.
void fwi_line_search_init(fwi_line_search* ls) {
  if (!ls) return;
  ls->alpha[0] = ls->alpha[1] = ls->alpha[2] = 1.0;
  ls->count = 0;
  ls->max_failures = 5;
}

int fwi_line_search_step(fwi_line_search* ls, double phi_n, double* trial) {
  if (!ls || !trial) {
    fwi::set_message("fwi_line_search_step: null argument");
    return -1;
  }
  if (std::isnan(*trial)) {  // main.c:756-763
    *trial = 1e30;
    ls->count += 1;
    for (double& a : ls->alpha) a /= 2.0;
  } else if (*trial > phi_n) {  // main.c:764-771
    for (double& a : ls->alpha) a /= 2.0;
    ls->count += 1;
  }
  if (ls->count > ls->max_failures) return 2;  // main.c:773-784: restore the backup, end the stage
  return *trial <= phi_n ? 0 : 1;             // the loop runs while trial > phi_n (main.c:690)
}

// Host evaluation of the C-PML table (see cpml_table_host); writes cpml entries to a and b.
void fwi_cpml_table(int cpml, float vmax, float dt, float dh, float f, float d0, float L, float* a, float* b) {
  std::vector<float> ta, tb;
  fwi::cpml_table_host(cpml, vmax, dt, dh, f, d0, L, ta, tb);
  for (int e = 0; e < cpml; ++e) {
    a[e] = ta[e];
    b[e] = tb[e];
  }
}

}  // extern "C"
