// bf16_t instantiations of the sequential-walk kernels (sr_kernels.hpp).
#include "sr_kernels.hpp"

namespace segscan {

hipError_t launch_sr_bf16(int cf, int g, const SrParams& p, const Launch& l, bool seg, bool& ok) {
  return launch_sr_cf<bf16_t>(cf, g, p, l, seg, ok);
}

hipError_t launch_sddmm_bf16(int cf, int g, const SddmmParams& p, const Launch& l, bool& ok) {
  return launch_sddmm_cf<bf16_t>(cf, g, p, l, ok);
}

}  // namespace segscan
