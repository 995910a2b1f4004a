// fft_gkr on the device: the GKR proof of the commitment's FFT/evaluation
// circuit (src/VPD/fft_circuit_GKR.cpp:22-845).  The reference runs it as a
// separate program ("./fft_gkr lg out", vpd_verifier.cpp:90-99), so it draws
// its challenges from a fresh rand() stream (seed 1) and depends on lg only.
#pragma once
#include <stdint.h>
#include <vector>
#include <hip/hip_runtime.h>

namespace vg {

struct fg_ctx;
struct tbuf;

// Allocates every device buffer of one fft_gkr proof with 2^lg coefficients;
// work runs on `stream` (nullptr: a stream owned by the context).
int fg_create(fg_ctx **out, int lg, hipStream_t stream);
// The same in two steps: fg_plan lays the proof out and reports the device
// and pinned host bytes it needs (nothing allocated); fg_alloc takes them
// from `mem` and uploads the static tables.  fg_destroy after either.
struct mem_source;
// host_only: no stream is made (the plan's sizes only; fg_destroy releases it)
int fg_plan(fg_ctx **out, int lg, hipStream_t stream, int64_t *dev_bytes, int64_t *pin_bytes, bool host_only = false);
int fg_alloc(fg_ctx *c, mem_source *mem);
void fg_destroy(fg_ctx *c);
int64_t fg_device_bytes(const fg_ctx *c);
// Draws the challenge stream and enqueues the whole proof (upload, circuit,
// sumchecks, copy of the messages to the host); returns without waiting.
int fg_enqueue(fg_ctx *c);
// Waits for the proof, appends the "VGF1" transcript section to `out` and
// replays the program's inline checks (its verdict).  p_ms: device time.
int fg_finish(fg_ctx *c, tbuf &out, int *accepted, int *proof_size, double *p_ms);
// Recorded on the proof's stream after its last device work (fg_enqueue).
hipEvent_t fg_done_event(const fg_ctx *c);
// Verifies a "VGF1" section (as fg_finish writes it) with the program's
// checks: *accepted = 1 or 0.  No proof may be in flight on c.
int fg_verify(fg_ctx *c, const uint8_t *sec, size_t len, int *accepted);

}  // namespace vg
