// Host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <initializer_list>
#include <string>
#include <vector>

namespace fwi {

// Records the first error of the calling thread (cleared by fwi_last_error()).
void record_error(hipError_t e);
// Sets the thread's message for the status-returning API (fwi_error_message()).
void set_message(const std::string& msg);

// An array argument of a status-returning entry point and its name in messages.
struct ArrayArg {
  const void* ptr;
  const char* name;
};
// Every array must be non-null and memory the GPU can address (device, managed, or pinned host memory): a
// pointer the runtime classifies as unregistered host memory would fault the GPU inside a kernel, so it is
// refused, and so is a pointer the runtime cannot classify at all (the query fails, e.g. without a device).  On
// the first bad argument sets the message "<fn>: <name> ..." and returns 1.
int check_arrays(const char* fn, std::initializer_list<ArrayArg> args);
// Host-side arguments (read or written by the host): non-null only.
int check_host(const char* fn, std::initializer_list<ArrayArg> args);

// C-PML coefficient table by distance e from the edge (0 = outermost), e < cpml, for the profile of the
// reference's CPML_x/CPML_z kernels (src/parallelcalls.cu:18-46): depth l = (cpml-e)*dh,
// d = d0*Vmax*(l/L)^2, alpha = pi*f*(L-l)/L, b = exp(-(d+alpha)*dt), a = d*(b-1)/(d+alpha).
// Evaluated in FP64 from the float parameters; b is rounded to float before a is formed, as the
// reference forms a_x from the stored float b_x.
void cpml_table_host(int cpml, float vmax, float dt, float dh, float f, float d0, float L, std::vector<float>& a,
                     std::vector<float>& b);

}  // namespace fwi
