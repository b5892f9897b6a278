// Instantiations of the local-field kernel: one wave per chain, closed-form
// periodic lattices, records in VGPRs (kR = 2, 4, 8), up to four species.
// Exposes field_fn(canonical, kr, kbm), the kernel's address for
// hipLaunchKernel, or nullptr for a variant that is not built.
#include "ce_field.h"

namespace cemc {

namespace {

template <bool C, int KBM>
const void* field_by_kr(int kr) {
  if (kr == 2) return reinterpret_cast<const void*>(&field_kernel<C, 2, KBM>);
  if (kr == 4) return reinterpret_cast<const void*>(&field_kernel<C, 4, KBM>);
  if (kr == 8) return reinterpret_cast<const void*>(&field_kernel<C, 8, KBM>);
  return nullptr;
}

}  // namespace

const void* field_fn(bool canonical, int kr, int kbm) {
  if (kbm == 1) return canonical ? field_by_kr<true, 1>(kr) : field_by_kr<false, 1>(kr);
  if (kbm == 2) return canonical ? field_by_kr<true, 2>(kr) : field_by_kr<false, 2>(kr);
  if (kbm == 3) return canonical ? field_by_kr<true, 3>(kr) : field_by_kr<false, 3>(kr);
  return nullptr;
}

uint32_t field_lds_bytes(int occ_bytes, int trk_sites, int n_fields, int gtab_ints) {
  return field_layout(occ_bytes, trk_sites, n_fields, gtab_ints).total;
}

}  // namespace cemc
