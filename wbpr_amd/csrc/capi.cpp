// Status strings, version and option defaults of the C ABI.
#include <cstring>

#include "mf_internal.h"

extern "C" {

const char *mf_status_string(mf_status s) {
  switch (s) {
    case MF_OK: return "ok";
    case MF_ERR_INVALID_ARGUMENT: return "invalid argument";
    case MF_ERR_IO: return "file missing or unreadable";
    case MF_ERR_PARSE: return "malformed input file";
    case MF_ERR_OVERFLOW: return "value does not fit int32";
    case MF_ERR_DEVICE: return "HIP runtime error";
    case MF_ERR_NO_DEVICE: return "no usable GPU";
    case MF_ERR_OUT_OF_MEMORY: return "out of memory";
    case MF_ERR_ITERATION_LIMIT: return "iteration limit exceeded";
    case MF_ERR_INTERNAL: return "internal solver invariant failed";
    case MF_ERR_BUDGET: return "request exceeds the memory budget";
    case MF_ERR_NOT_READY: return "request not finished";
    case MF_ERR_NO_PATH: return "no path from source to sink";
  }
  return "unknown status";
}

const char *mf_version(void) { return "mf 0.1 (gfx950)"; }

void mf_options_default(mf_options *opts) {
  if (!opts) return;
  memset(opts, 0, sizeof(*opts));
}

}  // extern "C"
