// Plan executor (filled in by the plan-parser work).
#include "engine.hpp"

extern "C" int pj_plan_execute(pj_ctx *ctx, const char *plan_json, const char *base_dir,
                               char *out, uint64_t out_cap) {
  (void)ctx; (void)plan_json; (void)base_dir;
  if (out && out_cap) out[0] = 0;
  return pj_fail(PJ_E_UNSUPPORTED, "pj_plan_execute: plan executor not built");
}
