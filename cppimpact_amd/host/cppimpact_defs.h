// Host-only definitions of the HIP backend's C++ layer (cppimpact_defs.h of
// the reference, cppimpact_defs.h:1-40, without device qualifiers: nothing in
// this layer runs on the GPU; the step runs inside libcppimpact_hip.so).
#pragma once

#include <cstdio>
#include <cstdlib>

#include "../../include/cppimpact_hip.h"

#define CPPIMPACT_FUNCTION

// cpiErrchk(call, ctx): the gpuErrchk/gpuAssert convention
// (cppimpact_defs.h:13-22) for the C ABI: print the library's message with
// file and line, then exit with the status.
#define cpiErrchk(ans, ctx) \
  { cpiAssert((ans), (ctx), __FILE__, __LINE__); }
inline void cpiAssert(int code, const cpi_ctx *ctx, const char *file, int line,
                      bool abort = true) {
  if (code < 0) {
    std::fprintf(stderr, "cpiAssert: %s (status %d) %s %d\n", cpi_last_error(ctx), code, file,
                 line);
    if (abort) std::exit(code < 0 ? -code : code);
  }
}
