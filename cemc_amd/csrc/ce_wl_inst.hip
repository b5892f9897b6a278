// Instantiations of the Wang-Landau kernel for one lattice path and move
// kind, chosen by CEMC_WL_PERIODIC (0 or 1) and CEMC_WL_SGC (0: canonical
// swaps, 1: semi-grand-canonical flips) at compile time.  Exposes
// wl_fn_p<P>s<S>(kr, glds, kbm, nm, mode), the kernel's address for
// hipLaunchKernel (mode 0: own histogram, 1: with adaptive windows, 2:
// walkers on one shared histogram),
// or nullptr for a variant that is not built: records in VGPRs (kR = 2, 4,
// 8; up to 4 species; nm = 2 or 3 gathered members) or in LDS (kR = 0; up
// to 8 species), energy tables in LDS or HBM.
#include "ce_wl.h"

#if !defined(CEMC_WL_PERIODIC) || !defined(CEMC_WL_SGC)
#error "CEMC_WL_PERIODIC and CEMC_WL_SGC must be defined"
#endif

namespace cemc {

namespace {

constexpr bool P = CEMC_WL_PERIODIC != 0;
constexpr bool S = CEMC_WL_SGC != 0;

template <int KBM, bool G, int MODE>
const void* wl_by_kr(int kr, int nm) {
  if (nm == 2) {
    if (kr == 2) return reinterpret_cast<const void*>(&wl_kernel<P, 2, G, KBM, 2, MODE, S>);
    if (kr == 4) return reinterpret_cast<const void*>(&wl_kernel<P, 4, G, KBM, 2, MODE, S>);
    if (kr == 8) return reinterpret_cast<const void*>(&wl_kernel<P, 8, G, KBM, 2, MODE, S>);
    return nullptr;
  }
  if (nm != 3) return nullptr;
  if (kr == 2) return reinterpret_cast<const void*>(&wl_kernel<P, 2, G, KBM, 3, MODE, S>);
  if (kr == 4) return reinterpret_cast<const void*>(&wl_kernel<P, 4, G, KBM, 3, MODE, S>);
  if (kr == 8) return reinterpret_cast<const void*>(&wl_kernel<P, 8, G, KBM, 3, MODE, S>);
  if (kr == 0) return reinterpret_cast<const void*>(&wl_kernel<P, 0, G, KBM, 3, MODE, S>);
  return nullptr;
}

template <int MODE>
const void* wl_pick(int kr, bool glds, int kbm, int nm) {
  if (kbm == 1) return glds ? wl_by_kr<1, true, MODE>(kr, nm) : wl_by_kr<1, false, MODE>(kr, nm);
  if (kbm == 2) return glds ? wl_by_kr<2, true, MODE>(kr, nm) : wl_by_kr<2, false, MODE>(kr, nm);
  if (kbm == 3) return glds ? wl_by_kr<3, true, MODE>(kr, nm) : wl_by_kr<3, false, MODE>(kr, nm);
  if (kbm == 7 && kr == 0 && nm == 3)
    return reinterpret_cast<const void*>(glds ? &wl_kernel<P, 0, true, 7, 3, MODE, S> : &wl_kernel<P, 0, false, 7, 3, MODE, S>);
  return nullptr;
}

}  // namespace

#define CEMC_WL_FN_NAME2(P_, S_) wl_fn_p##P_##s##S_
#define CEMC_WL_FN_NAME(P_, S_) CEMC_WL_FN_NAME2(P_, S_)

const void* CEMC_WL_FN_NAME(CEMC_WL_PERIODIC, CEMC_WL_SGC)(int kr, bool glds, int kbm, int nm, int mode) {
  if (mode == 0) return wl_pick<0>(kr, glds, kbm, nm);
  if (mode == 1) return wl_pick<1>(kr, glds, kbm, nm);
  if (mode == 2) return wl_pick<2>(kr, glds, kbm, nm);
  return nullptr;
}

}  // namespace cemc
