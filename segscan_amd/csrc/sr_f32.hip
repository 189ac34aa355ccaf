// float instantiations of the sequential-walk kernels (sr_kernels.hpp).
#include "sr_kernels.hpp"

namespace segscan {

hipError_t launch_sr_f32(int cf, int g, const SrParams& p, const Launch& l, bool seg, bool& ok) {
  return launch_sr_cf<float>(cf, g, p, l, seg, ok);
}

hipError_t launch_sddmm_f32(int cf, int g, const SddmmParams& p, const Launch& l, bool& ok) {
  return launch_sddmm_cf<float>(cf, g, p, l, ok);
}

// VType dispatch for the three storage types.
hipError_t launch_spmm_sr(VType vt, int cf, int g, const SrParams& p, const Launch& l, bool& ok) {
  switch (vt) {
    case VType::F32: return launch_sr_f32(cf, g, p, l, false, ok);
    case VType::BF16: return launch_sr_bf16(cf, g, p, l, false, ok);
    case VType::F16: return launch_sr_f16(cf, g, p, l, false, ok);
  }
  ok = false;
  return hipSuccess;
}

hipError_t launch_seg_sr(VType vt, int cf, int g, const SrParams& p, const Launch& l, bool& ok) {
  switch (vt) {
    case VType::F32: return launch_sr_f32(cf, g, p, l, true, ok);
    case VType::BF16: return launch_sr_bf16(cf, g, p, l, true, ok);
    case VType::F16: return launch_sr_f16(cf, g, p, l, true, ok);
  }
  ok = false;
  return hipSuccess;
}

hipError_t launch_sddmm_sr(VType vt, int cf, int g, const SddmmParams& p, const Launch& l, bool& ok) {
  switch (vt) {
    case VType::F32: return launch_sddmm_f32(cf, g, p, l, ok);
    case VType::BF16: return launch_sddmm_bf16(cf, g, p, l, ok);
    case VType::F16: return launch_sddmm_f16(cf, g, p, l, ok);
  }
  ok = false;
  return hipSuccess;
}

}  // namespace segscan
