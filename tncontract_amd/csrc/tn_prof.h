// Per-kernel launch timing with HIP events recorded on the launch stream.
// Off by default; tn_prof_enable(1) turns it on for the tags it names.
#pragma once
#include "tn_common.h"

namespace tn {

// Tags of the instrumented kernels.
enum ProfTag {
  PROF_GEMM_F64 = 0,
  PROF_GEMM_F32,
  PROF_GEMM_BF16X3,
  PROF_JAC_GRAM,
  PROF_JAC_EIG,
  PROF_JAC_APPLY,
  PROF_QR_PANEL,
  PROF_DIRECT,
  PROF_NTAGS
};

bool prof_on();
// Record a start event before a launch; returns an opaque slot (or -1).
int prof_start(int tag, hipStream_t st, double flops, double bytes);
void prof_stop(int slot, hipStream_t st);
// Add work to a tag's totals without a launch of its own: for kernels whose
// work is only known after they ran (counted on the device).
void prof_add(int tag, double flops, double bytes);

struct ProfScope {
  int slot;
  hipStream_t st;
  ProfScope(int tag, hipStream_t s, double flops, double bytes) : st(s) {
    slot = prof_on() ? prof_start(tag, s, flops, bytes) : -1;
  }
  ~ProfScope() {
    if (slot >= 0) prof_stop(slot, st);
  }
};

}  // namespace tn
