// contact_sphere_fp32.hip — FP32-force form of the sphere contact pass (placeholder until the
// cell-relative FP32 kernel lands).
#include "contact_law.h"

namespace xdem {

int launch_force_fp32(xdem_ctx* c, const ContactParams&, const ContactParams&, double, bool, bool, V3, bool,
                      int32_t) {
  return set_err(c, XDEM_ERR_UNSUPPORTED, "FP32-force form not built yet");
}

}  // namespace xdem
