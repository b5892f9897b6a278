// Instantiations of the sweep kernel for one (waves per chain, lattice path)
// pair, chosen by CEMC_SWEEP_WAVES (1, 2, 4, 8) and CEMC_SWEEP_PERIODIC
// (0 or 1) at compile time; the build compiles this file once per pair, in
// parallel.  Exposes sweep_fn_w<W>_p<P>(canonical, kr, glds, kbm, spec, hooks, nm), the
// kernel's address for hipLaunchKernel, or nullptr for a variant that is
// not built:
//   one wave per chain:  records in VGPRs (kR = 2, 4, 8; energy tables in
//                        LDS or HBM; up to 4 species) or in LDS (up to 8)
//   2, 4 or 8 waves:     records in VGPRs only
//   speculative rounds of kSpecMoves moves: periodic lattices, records in
//                        VGPRs
//   wide rounds of kSpecMoves x W moves (W = 2, 4, 8: each wave evaluates
//                        its own eight): periodic lattices, every record in
//                        each wave's VGPRs
#include "ce_sweep.h"

#ifndef CEMC_SWEEP_WAVES
#error "CEMC_SWEEP_WAVES must be defined"
#endif
#ifndef CEMC_SWEEP_PERIODIC
#error "CEMC_SWEEP_PERIODIC must be defined"
#endif

namespace cemc {

namespace {

constexpr int W = CEMC_SWEEP_WAVES;
constexpr bool P = CEMC_SWEEP_PERIODIC != 0;

// spec: moves per speculative round, 1 (sequential) or kSpecMoves
// (periodic lattices, records in VGPRs)
template <bool C, int KBM, bool G, int K, bool H>
const void* by_kr(int kr, int nm) {
  if (nm == 2) {
    if (kr == 2) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 2, G, KBM, W, K, H, 2>);
    if (kr == 4) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 4, G, KBM, W, K, H, 2>);
    if (kr == 8) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 8, G, KBM, W, K, H, 2>);
    return nullptr;
  }
  if (nm != 3) return nullptr;
  if (kr == 2) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 2, G, KBM, W, K, H>);
  if (kr == 4) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 4, G, KBM, W, K, H>);
  if (kr == 8) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 8, G, KBM, W, K, H>);
  if constexpr (W == 1 && K == 1) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 0, G, KBM, W, K, H>);
  return nullptr;
}

template <bool C, int K, bool H>
const void* lookup(int kr, bool glds, int kbm, int nm) {
  if (kbm == 1) return glds ? by_kr<C, 1, true, K, H>(kr, nm) : by_kr<C, 1, false, K, H>(kr, nm);
  if (kbm == 2) return glds ? by_kr<C, 2, true, K, H>(kr, nm) : by_kr<C, 2, false, K, H>(kr, nm);
  if (kbm == 3) return glds ? by_kr<C, 3, true, K, H>(kr, nm) : by_kr<C, 3, false, K, H>(kr, nm);
  if constexpr (W == 1 && K == 1)
    if (kr == 0 && nm == 3)
      return reinterpret_cast<const void*>(glds ? &sweep_kernel<P, C, 0, true, 7, W, 1, H>
                                                : &sweep_kernel<P, C, 0, false, 7, W, 1, H>);
  return nullptr;
}

}  // namespace

#define CEMC_SWEEP_FN_NAME2(W_, P_) sweep_fn_w##W_##_p##P_
#define CEMC_SWEEP_FN_NAME(W_, P_) CEMC_SWEEP_FN_NAME2(W_, P_)

const void* CEMC_SWEEP_FN_NAME(CEMC_SWEEP_WAVES, CEMC_SWEEP_PERIODIC)(bool canonical, int kr, bool glds, int kbm,
                                                                       int spec, bool hooks, int nm) {
  if (hooks) {
    if constexpr (W == 1) {
      if (spec != 1) return nullptr;
      return canonical ? lookup<true, 1, true>(kr, glds, kbm, nm) : lookup<false, 1, true>(kr, glds, kbm, nm);
    }
    return nullptr;
  }
  if (spec == kSpecMoves) {
    if constexpr (P) {
      if (kr == 0) return nullptr;
      return canonical ? lookup<true, kSpecMoves, false>(kr, glds, kbm, nm)
                       : lookup<false, kSpecMoves, false>(kr, glds, kbm, nm);
    }
    return nullptr;
  }
  if constexpr (P && W > 1) {
    if (spec == kSpecMoves * W) {
      if (kr == 0) return nullptr;
      return canonical ? lookup<true, kSpecMoves * W, false>(kr, glds, kbm, nm)
                       : lookup<false, kSpecMoves * W, false>(kr, glds, kbm, nm);
    }
  }
  if (spec != 1) return nullptr;
  return canonical ? lookup<true, 1, false>(kr, glds, kbm, nm) : lookup<false, 1, false>(kr, glds, kbm, nm);
}

}  // namespace cemc
