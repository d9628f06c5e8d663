// Error reporting and version for the pss C ABI.
#include <stdio.h>
#include <string.h>
#include "common.h"
#include "../../include/pss.h"

static thread_local char g_err[512] = "";

int pss_set_error(hipError_t e, const char* what) {
  snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
  return (int)e == 0 ? 1 : (int)e;
}

int pss_set_msg(int code, const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

extern "C" const char* pss_last_error(void) { return g_err; }
extern "C" const char* pss_version(void) { return "pss 0.3 (gfx950, fp64 + mixed)"; }
