// tis_internal.h -- context accessors shared by the library's translation
// units (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/tis_hip.h"

namespace tis {
struct BdState;
hipStream_t ctx_stream(tis_ctx *c);
int ctx_device(tis_ctx *c);
BdState *&ctx_bd(tis_ctx *c);
int ctx_fail(tis_ctx *c, int code, const char *msg);
void bd_free(tis_ctx *c);
}  // namespace tis
