// Library-internal entry points shared between the prover driver and the
// proof server (not part of the C ABI).
#pragma once
#include <stdint.h>
#include <string>
#include "vg_host.hpp"
#include "../../include/virgo.h"

// Sets this thread's vg_last_error() text.
void vg_set_error(const std::string &msg);
// Builds the circuit and plans one proof on `stream` (host work only;
// nothing allocated on the device).  Destroys the context on failure.
// host_only: no HIP call at all (sizes only, the layout of a proof on a
// caller's stream); such a context can only be measured and destroyed.
int ctx_create_planned(vg_ctx **out, const vg_circuit_desc *d, void *stream, bool host_only = false);
// Device and pinned host bytes the planned context takes from its source.
void ctx_needs(const vg_ctx *c, int64_t *dev, int64_t *pin);
// Takes the planned memory from `mem` (nullptr: hipMalloc / hipHostMalloc)
// and uploads circuit and tables; *out = c, or c destroyed on failure.
int ctx_materialize(vg_ctx *c, vg::mem_source *mem, vg_ctx **out);
