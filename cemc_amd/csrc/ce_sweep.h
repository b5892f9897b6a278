// Sweep kernel of the cluster-expansion Metropolis engine (gfx950): device
// helpers and the kernel template.  The instantiations live in
// ce_sweep_inst.hip, compiled once per (waves per chain, lattice path).
#pragma once

#include "ce_kernels.h"

namespace cemc {

namespace {

__device__ inline int site_group(const ModelDev& m, int32_t site) {
  return (m.n_groups == 1 && !m.has_background) ? 0 : int(m.site_group[site]);
}

struct LdsLayout {
  uint32_t occ, trk, rec, gtab, total;
};

__host__ __device__ inline uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }

// occupation (padded box on periodic lattices) | species tracker (canonical)
// | records (large models) | energy tables
__host__ __device__ inline LdsLayout lds_layout(int occ_bytes, int trk_sites, int n_rec_lds, int gtab_doubles) {
  LdsLayout L;
  L.occ = 0;
  L.trk = align16(uint32_t(occ_bytes));
  L.rec = L.trk + align16(uint32_t(trk_sites) * 2u);
  L.gtab = L.rec + uint32_t(n_rec_lds) * uint32_t(sizeof(FatRec));
  L.total = L.gtab + uint32_t(gtab_doubles) * 8u;
  return L;
}

}  // namespace

// ---------------------------------------------------------------------------
// Sweep kernel.
//
// One 64-lane wavefront owns one chain.  Everything the move loop reads is
// staged before the loop: model scalars and packed species tables in SGPRs,
// the occupation, species tracker and energy tables in LDS, and each lane's
// sub-cluster records in VGPRs (kR per lane) or, for large models, in LDS
// (read in batches of 4 per lane).
//
// On closed-form periodic lattices the occupation is held in a padded
// (halo) box of (L1+2h) x (L2+2h) x (L3+2h) bytes, each cell holding the
// species of the site it wraps to; a member's address is then the reference
// site's padded index plus a per-record constant, and an accepted flip
// writes the site's (at most 8) images.  Other lattices read members through
// the translation table in HBM.
//
// Per trial move each lane issues all member-species gathers of its records
// at once (one LDS round trip), then all energy-table reads (a second round
// trip), and the partial energy changes are summed across the wave.

struct SweepCtx {
  int S, B, N;
  int L1, L2, L3, L23;
  int h, P2, P3;         // halo width and padded dims (periodic)
  uint32_t mag23, mag3;  // floor((2^32-1)/d) for the divisors L2*L3 and L3
  int DX, DY, DZ;         // padded-index steps of one lattice period along x, y, z
  int K;
  const int32_t* trans;
};

__device__ __attribute__((always_inline)) inline uint32_t fast_div(uint32_t n, uint32_t d, uint32_t magic) {
  // exact for n, d < 2^16: the estimate is the quotient or one less
  uint32_t q = __umulhi(n, magic);
  if (n - q * d >= d) ++q;
  return q;
}

struct Ref {
  int site;     // lattice site
  int p;        // padded index (periodic) or site (table)
  int x, y, z;  // cell coordinates (periodic)
};

template <bool kPeriodic>
__device__ __attribute__((always_inline)) inline Ref ref_of(const SweepCtx& c, int site) {
  Ref r;
  r.site = site;
  if constexpr (kPeriodic) {
    const uint32_t x = fast_div(uint32_t(site), uint32_t(c.L23), c.mag23);
    const uint32_t rem = uint32_t(site) - x * uint32_t(c.L23);
    const uint32_t y = fast_div(rem, uint32_t(c.L3), c.mag3);
    r.x = int(x);
    r.y = int(y);
    r.z = int(rem - y * uint32_t(c.L3));
    r.p = __umul24(__umul24(uint32_t(r.x + c.h), uint32_t(c.P2)) + uint32_t(r.y + c.h), uint32_t(c.P3)) +
          uint32_t(r.z + c.h);
  } else {
    r.x = r.y = r.z = 0;
    r.p = site;
  }
  return r;
}

// Packed wrapped cell offset from site j to site i (x | y<<8 | z<<16): a
// member of j with this wrapped offset is site i.
__device__ __attribute__((always_inline)) inline uint32_t wrapped_delta(const SweepCtx& c, const Ref& i,
                                                                        const Ref& j) {
  int dx = i.x - j.x, dy = i.y - j.y, dz = i.z - j.z;
  dx += dx < 0 ? c.L1 : 0;
  dy += dy < 0 ? c.L2 : 0;
  dz += dz < 0 ? c.L3 : 0;
  return uint32_t(dx) | (uint32_t(dy) << 8) | (uint32_t(dz) << 16);
}

// Slices per code of the interleaved table a kernel reads (0: the plain
// table): three-species tables always (a record's two values in one 16-byte
// read), four-species tables when read from HBM (three adjacent values in
// one or two cache lines); four-species tables in LDS keep the plain layout,
// whose two first slices a single ds_read2st64_b64 fetches.
template <int kBM, bool kGtabLds>
constexpr int il_width() {
  return kBM == 2 ? 2 : (kBM == 3 && !kGtabLds ? 3 : 0);
}

template <int kBM>
struct TableGeom {
  static constexpr int kBits = kBM == 1 ? 1 : (kBM <= 3 ? 2 : 3);  // code bits per species
  static constexpr int kCodes = 1 << (3 * kBits);                   // codes per decoration slice
};

// Phase 1 of a flip: addresses of every member of kR records, issue of all
// member-species gathers, energy-table indices.  With kPatch, members that
// are the patch site (periodic: wrapped offset == patch_w; table: site ==
// patch_w) read `patch_species`.
// kNM: members gathered per record (2 when no record has more than two
// other members, i.e. the model has no quadruplets; the third field of the
// code is then 0, which reads the same entry since entries repeat over the
// fields of unused members).  kIL: slices per code in the interleaved table
// (ModelDev::gtab_il), where the kIL slices of a code sit at base + kIL code
// (0: the plain table, slice d of a code at base + d * codes + code).
template <bool kPeriodic, int kR, int kBM, bool kPatch, int kNM = 3, int kIL = 0>
__device__ __attribute__((always_inline)) inline void table_index(const SweepCtx& c, const FatRec (&rec)[kR],
                                                                  const Ref& ref, const uint8_t* occ,
                                                                  uint32_t patch_w, uint32_t patch_species,
                                                                  int (&gi)[kR]) {
  static_assert(kNM == 2 || kNM == 3, "two or three gathered members");
  constexpr int kBits = TableGeom<kBM>::kBits;
  int nb[kR][3];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    if constexpr (kPeriodic) {
      nb[k][0] = ref.p + rec[k].m0;
      nb[k][1] = ref.p + rec[k].m1;
      nb[k][2] = ref.p + rec[k].m2;
    } else {
      const int32_t* row = c.trans + int64_t(ref.site) * c.K;
      nb[k][0] = row[rec[k].m0];
      nb[k][1] = row[rec[k].m1];
      nb[k][2] = kNM == 3 ? row[rec[k].m2] : 0;
    }
  }
  uint32_t sp[kR][3];
#pragma unroll
  for (int k = 0; k < kR; ++k)
#pragma unroll
    for (int q = 0; q < 3; ++q) sp[k][q] = q < kNM ? uint32_t(occ[nb[k][q]]) : 0u;
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    uint32_t s0 = sp[k][0], s1 = sp[k][1], s2 = sp[k][2];
    if constexpr (kPatch) {
      if constexpr (kPeriodic) {
        s0 = rec[k].w0 == patch_w ? patch_species : s0;
        s1 = rec[k].w1 == patch_w ? patch_species : s1;
        if constexpr (kNM == 3) s2 = rec[k].w2 == patch_w ? patch_species : s2;
      } else {
        s0 = uint32_t(nb[k][0]) == patch_w ? patch_species : s0;
        s1 = uint32_t(nb[k][1]) == patch_w ? patch_species : s1;
        if constexpr (kNM == 3) s2 = uint32_t(nb[k][2]) == patch_w ? patch_species : s2;
      }
    }
    // unused members need no masking: their table entries repeat over them
    const uint32_t code = s0 | (s1 << kBits) | (s2 << (2 * kBits));
    gi[k] = int(rec[k].meta & 0xFFFFFFu) + int(kIL ? uint32_t(kIL) * code : code);
  }
}

// The energy tables: in LDS, or in HBM through a global-address-space
// pointer (global loads count only against vmcnt; a flat load would also
// wait for, and hold up, the LDS gathers).
using GlobalDoublePtr = const __attribute__((address_space(1))) double*;

template <bool kLds>
__device__ __attribute__((always_inline)) inline auto table_ptr(const double* lds_tab, const double* hbm_tab) {
  if constexpr (kLds)
    return lds_tab;
  else
    return (GlobalDoublePtr)hbm_tab;
}

// Both slices of a code from the interleaved table in one 16-byte read.
typedef double dbl2_t __attribute__((ext_vector_type(2)));
__device__ __attribute__((always_inline)) inline dbl2_t load_pair(const double* p) {
  return *reinterpret_cast<const dbl2_t*>(p);
}
__device__ __attribute__((always_inline)) inline dbl2_t load_pair(GlobalDoublePtr p) {
  return *reinterpret_cast<const __attribute__((address_space(1))) dbl2_t*>(p);
}

// Phase 2: energy-table reads.  kBM == 1 (two species): the plain sum of the
// entries, to be scaled by the single basis difference after the reduction;
// otherwise the change for basis differences dphi.
template <int kR, int kBM, int kIL = 0, typename TabPtr>
__device__ __attribute__((always_inline)) inline double table_sum(const int (&gi)[kR], TabPtr gtab,
                                                                  const double (&dphi)[kBM]) {
  constexpr int kCodes = TableGeom<kBM>::kCodes;
  static_assert(kIL == 0 || kIL == kBM, "the interleaved table holds the read slices");
  double g[kR][kBM];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    if constexpr (kIL == 2) {
      const dbl2_t t = load_pair(gtab + gi[k]);
      g[k][0] = t.x;
      g[k][kBM - 1] = t.y;
    } else if constexpr (kIL == 3) {
      // three adjacent values (24 bytes): one or two cache lines where the
      // plain layout touches three
#pragma unroll
      for (int d = 0; d < kBM; ++d) g[k][d] = gtab[gi[k] + d];
    } else {
#pragma unroll
      for (int d = 0; d < kBM; ++d) g[k][d] = gtab[gi[k] + d * kCodes];
    }
  }
  double v = kBM == 1 ? g[0][0] : dphi[0] * g[0][0];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    if constexpr (kBM == 1) {
      if (k > 0) v += g[k][0];
    } else {
#pragma unroll
      for (int d = 0; d < kBM; ++d)
        if (k > 0 || d > 0) v = fma(dphi[d], g[k][d], v);
    }
  }
  return v;
}

// Records of one translation group read from LDS in batches of 4 per lane
// (thread `tid` of `nthreads` working on the chain).
template <bool kPeriodic, int kBM, bool kPatch, int kIL = 0, typename TabPtr>
__device__ __attribute__((always_inline)) inline double lds_flip_de(const SweepCtx& c, const FatRec* recs, int off,
                                                                    int cnt, int tid, int nthreads, const Ref& ref,
                                                                    const uint8_t* occ, TabPtr gtab,
                                                                    const double (&dphi)[kBM], uint32_t patch_w,
                                                                    uint32_t patch_species) {
  double v = 0.0;
  for (int base = 0; base < cnt; base += 4 * nthreads) {
    FatRec rb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int it = base + tid + nthreads * k;
      rb[k] = it < cnt ? recs[off + it] : FatRec{0, 0, 0, 0u, 0u, 0u, 0u, 0u};
    }
    int gi[4];
    table_index<kPeriodic, 4, kBM, kPatch, 3, kIL>(c, rb, ref, occ, patch_w, patch_species, gi);
    v += table_sum<4, kBM, kIL>(gi, gtab, dphi);
  }
  return v;
}

__device__ __attribute__((always_inline)) inline int unpack16(uint64_t lo, uint64_t hi, uint32_t i) {
  return int(((i < 4u ? lo : hi) >> (16u * (i & 3u))) & 0xFFFFu);
}

// Image word of a site on the padded lattice: bits 0-19 its padded index,
// bits 20+2d and 21+2d the other image along axis d as a signed 2-bit
// value (0: none, 1: at +L_d, -1: at -L_d; a site within h of a face has
// one other image along that axis, since L_d > 2h).
template <bool kPeriodic>
__device__ __attribute__((always_inline)) inline uint32_t image_word(const SweepCtx& c, const Ref& r) {
  if constexpr (!kPeriodic) {
    return uint32_t(r.p);
  } else {
    const uint32_t fx = r.x < c.h ? 1u : (r.x >= c.L1 - c.h ? 3u : 0u);
    const uint32_t fy = r.y < c.h ? 1u : (r.y >= c.L2 - c.h ? 3u : 0u);
    const uint32_t fz = r.z < c.h ? 1u : (r.z >= c.L3 - c.h ? 3u : 0u);
    return uint32_t(r.p) | (fx << 20) | (fy << 22) | (fz << 24);
  }
}

// Padded index of image `bits` (bit d: take the other image along axis d)
// of the site with image word w, or -1 when that image does not exist.
__device__ __attribute__((always_inline)) inline int image_of(const SweepCtx& c, uint32_t w, int bits) {
  const int ex = int(w << 10) >> 30, ey = int(w << 8) >> 30, ez = int(w << 6) >> 30;
  const int sx = (bits & 1) ? ex : 0, sy = (bits & 2) ? ey : 0, sz = (bits & 4) ? ez : 0;
  const bool ok = ((bits & 1) == 0 || ex != 0) && ((bits & 2) == 0 || ey != 0) && ((bits & 4) == 0 || ez != 0);
  return ok ? int(w & 0xFFFFFu) + sx * c.DX + sy * c.DY + sz * c.DZ : -1;
}

// Sum of x over the kW waves of the workgroup, added in wave order so every
// wave gets the identical value.  `slots` holds kW doubles of LDS that no
// wave may still be reading from an earlier call.
template <int kW>
__device__ __attribute__((always_inline)) inline double group_sum(double x, double* slots, int wave, int lane) {
  const double w = wave_sum_f64(x);
  if constexpr (kW == 1) {
    return w;
  } else {
    if (lane == 0) slots[wave] = w;
    __syncthreads();
    double t = slots[0];
#pragma unroll
    for (int v = 1; v < kW; ++v) t += slots[v];
    return t;
  }
}

// True when a change at packed cell a (x | y<<8 | z<<16) can alter the
// energy change of a flip at packed cell b: a - b (wrapped) is 0 or a
// sub-cluster member offset, per the model's interaction mask over the
// (2h+1)^3 box.  Per axis e = (a - b + h) mod L, the one residue in [0, L),
// is the smallest of d, d + L, d - L as unsigned values (d = a - b + h); the
// offset lies in the box when e <= 2h (L > 2h makes it unique).
__device__ __attribute__((always_inline)) inline bool cells_interact(uint32_t a, uint32_t b, const SweepCtx& c,
                                                                     uint64_t im0, uint64_t im1) {
  const uint32_t h = uint32_t(c.h);
  const uint32_t dx = (a & 255u) + h - (b & 255u);
  const uint32_t dy = ((a >> 8) & 255u) + h - ((b >> 8) & 255u);
  const uint32_t dz = (a >> 16) + h - (b >> 16);
  const uint32_t ex = min(min(dx, dx + uint32_t(c.L1)), dx - uint32_t(c.L1));
  const uint32_t ey = min(min(dy, dy + uint32_t(c.L2)), dy - uint32_t(c.L2));
  const uint32_t ez = min(min(dz, dz + uint32_t(c.L3)), dz - uint32_t(c.L3));
  const uint32_t w = 2u * h + 1u;
  const uint32_t idx = (ex * w + ey) * w + ez;
  const uint64_t word = idx < 64u ? im0 : im1;
  return max(max(ex, ey), ez) <= 2u * h && ((word >> (idx & 63u)) & 1u) != 0;
}

// A change of `site` from species o to n passes the per-species and
// per-site constraints (FixedElement, mc_constraints.py:84-105;
// ConstrainElementByTag and fixed sites, :171-190).
__device__ __attribute__((always_inline)) inline bool hook_allows(uint32_t fixed, const uint8_t* allowed, int site,
                                                                  uint32_t o, uint32_t n) {
  if (((fixed >> o) | (fixed >> n)) & 1u) return false;
  return allowed == nullptr || ((allowed[site] >> n) & 1u) != 0;
}

// Change of the number of ordered pairs (i, j), j a member of i in the
// constrained pair cluster, with species e1 at i and e2 at j, when site s1
// goes o1 -> n1 and (two) site s2 goes o2 -> n2.  Lanes 0-31 take the
// members of s1, lanes 32-63 those of s2.  A member outside the changed
// sites contributes both orientations; a pair of two changed sites is
// counted once from each end (member lists are symmetric).  The reference
// rescans the whole lattice for such a pair after every proposal
// (pair_constraint.cpp:14-46); with the current state's count kept, the
// proposal is valid when count + change is zero.
template <bool kPeriodic>
__device__ __attribute__((always_inline)) inline int64_t hook_pair_delta(
    const SweepCtx& c, const uint8_t* occ, const int32_t* nb_tab, int nm, int e1, int e2, int s1, uint32_t o1,
    uint32_t n1, int s2, uint32_t o2, uint32_t n2, bool two, int lane) {
  const int q = lane & 31;
  const bool second = lane >= 32;
  int v = 0;
  if (q < nm && (!second || two)) {
    const int s = second ? s2 : s1;
    const int os = int(second ? o2 : o1), ns = int(second ? n2 : n1);
    const int nb = nb_tab[int64_t(s) * nm + q];
    const bool is1 = nb == s1, is2 = two && nb == s2;
    const int on = occ[ref_of<kPeriodic>(c, nb).p];
    const int nn = is1 ? int(n1) : (is2 ? int(n2) : on);
    const bool inside = is1 || is2;
    const int before = int(os == e1 && on == e2) + int(!inside && on == e1 && os == e2);
    const int after = int(ns == e1 && nn == e2) + int(!inside && nn == e1 && ns == e2);
    v = after - before;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return int64_t(v);
}

// kR > 0: single-group model, records in VGPRs (kR per lane); kR == 0:
// records in LDS.  kBM: decoration slices of the energy tables (1: two
// species, 3: up to four, 7: up to eight).  kW: waves per chain.  With
// kW > 1 the waves split the records, every wave makes the same decision
// from the same ordered sum, and every wave writes the same commit; one
// barrier per move (inside group_sum) separates a move's commit from the
// next move's gathers in other waves.
//
// kHooks (one wave, one move at a time): the chain's constraints and bias
// (HookDev) apply; each move's proposal is redrawn while it violates a
// constraint.
template <bool kPeriodic, bool kCanonical, int kR, bool kGtabLds, int kBM, int kW, int kK, bool kHooks = false,
          int kNM = 3>
__global__ __launch_bounds__(64 * kW) void sweep_kernel(const ChainParams* __restrict__ params,
                                                   ChainState* __restrict__ states,
                                                   const ModelDev* __restrict__ models,
                                                   int8_t* __restrict__ occ_pool,
                                                   uint16_t* __restrict__ trk_pool,
                                                   const int32_t* __restrict__ chain_order,
                                                   const int64_t* __restrict__ n_moves,
                                                   uint32_t* __restrict__ accept_log, int64_t log_words,
                                                   HookDev* __restrict__ hooks, int64_t time_limit) {
  static_assert(!kHooks || (kW == 1 && kK == 1), "constraints and bias run one wave, one move at a time");
  static_assert(kNM == 3 || kR > 0, "records read from LDS gather three members");
  constexpr int kILW = il_width<kBM, kGtabLds>();
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ double dphi_tab[kMaxSpecies * kMaxSpecies * kMaxBasis];  // [(old*S+new)*B + d]
  __shared__ double dsingle_tab[kMaxSpecies * kMaxSpecies];          // singlet energy change per (old,new)
  __shared__ double phi_tab[kMaxBasis * kMaxSpecies];                // phi[d*S + s]
  __shared__ int grp_tab[2 * kMaxGroups];
  __shared__ double red_move[2][8];             // per-move wave sums, double-buffered by move parity
  __shared__ double red_init[kMaxBasis][8];     // initial singlets
  __shared__ double red_spec[2][kK == kSpecMoves ? kK * 8 : 1];  // speculative rounds: per-round wave sums
  __shared__ double red_wide[kK > kSpecMoves ? kK : 1];            // wide rounds: the round's move sums
  constexpr int kRecLds = kR == 0;
  constexpr int kThreads = 64 * kW;
  const int chain = chain_order[blockIdx.x];
  const int tid = int(threadIdx.x);
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const ChainParams* __restrict__ pp = params + chain;
  const ModelDev& m = models[pp->model];
  const int N = pp->n_sites;
  const int occ_bytes = kPeriodic ? m.pdims[0] * m.pdims[1] * m.pdims[2] : N;
  const LdsLayout L = lds_layout(occ_bytes, kCanonical ? N : 0, kRecLds ? m.n_rec : 0, kGtabLds ? m.gtab_size : 0);
  uint8_t* occ = reinterpret_cast<uint8_t*>(lds + L.occ);
  uint16_t* trk = reinterpret_cast<uint16_t*>(lds + L.trk);
  FatRec* recs_l = reinterpret_cast<FatRec*>(lds + L.rec);
  double* gtab_l = reinterpret_cast<double*>(lds + L.gtab);

  SweepCtx c;
  c.S = m.n_species;
  c.B = m.n_bfs;
  c.N = N;
  c.L1 = m.dims[0];
  c.L2 = m.dims[1];
  c.L3 = m.dims[2];
  c.L23 = c.L2 * c.L3;
  c.h = m.pad_h;
  c.P2 = m.pdims[1];
  c.P3 = m.pdims[2];
  c.mag23 = kPeriodic ? 0xFFFFFFFFu / uint32_t(c.L23) : 0u;
  c.mag3 = kPeriodic ? 0xFFFFFFFFu / uint32_t(c.L3) : 0u;
  c.DX = c.L1 * c.P2 * c.P3;
  c.DY = c.L2 * c.P3;
  c.DZ = c.L3;
  c.K = m.K;
  c.trans = m.trans;
  const bool multi_group = m.n_groups > 1 || m.has_background;
  const bool has_bkg = m.has_background != 0;
  const int n_pick = has_bkg ? m.n_movable : N;

  // ---- stage chain state, per-chain tables and model tables into LDS
  const uint8_t* gocc = reinterpret_cast<const uint8_t*>(occ_pool + pp->occ_off);
  if constexpr (kPeriodic) {
    const uint32_t p23 = uint32_t(c.P2 * c.P3), mag_p23 = 0xFFFFFFFFu / p23, mag_p3 = 0xFFFFFFFFu / uint32_t(c.P3);
    for (int p = tid; p < occ_bytes; p += kThreads) {
      const uint32_t px = fast_div(uint32_t(p), p23, mag_p23);
      const uint32_t rem = uint32_t(p) - px * p23;
      const uint32_t py = fast_div(rem, uint32_t(c.P3), mag_p3);
      int x = int(px) - c.h, y = int(py) - c.h, z = int(rem - py * uint32_t(c.P3)) - c.h;
      x += x < 0 ? c.L1 : 0;
      x -= x >= c.L1 ? c.L1 : 0;
      y += y < 0 ? c.L2 : 0;
      y -= y >= c.L2 ? c.L2 : 0;
      z += z < 0 ? c.L3 : 0;
      z -= z >= c.L3 ? c.L3 : 0;
      occ[p] = gocc[(x * c.L2 + y) * c.L3 + z];
    }
  } else {
    for (int i = tid; i < N; i += kThreads) occ[i] = gocc[i];
  }
  uint16_t* gtrk = trk_pool + pp->trk_off;
  if constexpr (kCanonical)
    for (int i = tid; i < N; i += kThreads) trk[i] = gtrk[i];
  if constexpr (kRecLds)
    for (int i = tid; i < m.n_rec; i += kThreads) recs_l[i] = m.frecs[i];
  if constexpr (kGtabLds)
    for (int i = tid; i < m.gtab_size; i += kThreads) gtab_l[i] = (kILW ? m.gtab_il : m.gtab)[i];
  const auto gtab = table_ptr<kGtabLds>(gtab_l, kILW ? m.gtab_il : m.gtab);
  if (tid < 2 * kMaxGroups) grp_tab[tid] = tid < kMaxGroups ? m.grp_rec_off[tid] : m.grp_rec_cnt[tid - kMaxGroups];
  for (int i = tid; i < c.B * c.S; i += kThreads) phi_tab[i] = m.basis[i];
  for (int i = tid; i < c.S * c.S; i += kThreads) {
    const int o = i / c.S, nw = i - o * c.S;
    double ds = 0.0;
    for (int d = 0; d < c.B; ++d) {
      const double dp = m.basis[d * c.S + nw] - m.basis[d * c.S + o];
      dphi_tab[i * c.B + d] = dp;
      ds += pp->j1[d] * dp;
    }
    dsingle_tab[i] = ds;
  }
  __syncthreads();

  // records of a single-group model stay in VGPRs for the whole launch
  // (wide rounds: every wave holds all of them, laid out as for one wave)
  constexpr bool kWide = kK > kSpecMoves;
  FatRec rr[kR > 0 ? kR : 1];
  if constexpr (kR > 0) {
    const int cnt = m.grp_rec_cnt[0];
#pragma unroll
    for (int k = 0; k < kR; ++k) {
      const int it = kWide ? lane + 64 * k : tid + kThreads * k;
      rr[k] = it < cnt ? m.frecs[it] : FatRec{0, 0, 0, 0u, 0u, 0u, 0u, 0u};
    }
  }
  // two species: the basis difference and singlet change of 0->1 and 1->0
  const double dphi01 = dphi_tab[1], dphi10 = dphi_tab[c.S];
  const double dsin01 = dsingle_tab[1], dsin10 = dsingle_tab[c.S];

  double energy = states[chain].energy;
  const uint64_t t0 = states[chain].moves_done;
  const double e_ref = states[chain].e_ref;
  const int64_t nmov = n_moves[chain];
  uint64_t n_acc = 0;
  double sum_de = 0.0, sum_de2 = 0.0;
  // SGC: lane d < B holds singlet d of the current state, sum over the
  // non-background sites of phi_d / N (ce_updater.cpp:366-371 keeps it
  // incrementally the same way)
  const double inv_n = 1.0 / double(N);
  double sing = 0.0, sum_s = 0.0, sum_s2 = 0.0, sum_se = 0.0;
  if constexpr (!kCanonical) {
    double part_s[kBM];
#pragma unroll
    for (int d = 0; d < kBM; ++d) part_s[d] = 0.0;
    // (wide rounds: every wave sums as one wave does, so the singlets and
    // their sums equal the one-wave path's bitwise)
    constexpr int kStride = kWide ? 64 : kThreads;
    for (int i = kWide ? lane : tid; i < N; i += kStride) {
      if (has_bkg && m.background[i]) continue;
      const int s = occ[ref_of<kPeriodic>(c, i).p];
#pragma unroll
      for (int d = 0; d < kBM; ++d)
        if (d < c.B) part_s[d] += phi_tab[d * c.S + s];
    }
#pragma unroll
    for (int d = 0; d < kBM; ++d) {
      const double tot = kWide ? wave_sum_f64(part_s[d]) : group_sum<kW>(part_s[d], red_init[d], wave, lane);
      if (lane == d) sing = tot / double(N);
    }
  }
  const double kT = pp->kT;
  const double inv_kT = 1.0 / kT;
  const uint64_t seed = pp->seed, uid = pp->uid;
  const uint32_t np = uint32_t(pp->n_present);
  const uint32_t present_pk = pp->present_pk;
  const uint64_t cnt_lo = pp->count_pk[0], cnt_hi = pp->count_pk[1];
  const uint64_t seg_lo = pp->segoff_pk[0], seg_hi = pp->segoff_pk[1];
  uint32_t* acc_log = accept_log ? accept_log + int64_t(chain) * log_words : nullptr;

  // Samples between accepted moves see the same state: the sums take each
  // run of equal samples at once (run = samples of the current state not yet
  // added).
  int64_t run = 0;
  auto flush = [&]() {
    const double dn = double(run);
    const double de_ref = energy - e_ref;
    sum_de = fma(dn, de_ref, sum_de);
    sum_de2 = fma(dn * de_ref, de_ref, sum_de2);
    if constexpr (!kCanonical) {
      sum_s = fma(dn, sing, sum_s);
      sum_s2 = fma(dn * sing, sing, sum_s2);
      sum_se = fma(dn * sing, de_ref, sum_se);
    }
    run = 0;
  };
  int64_t moves_made = nmov;
  // Time-sliced launches (time_limit > 0): the chain stops at the first
  // 64-move block that starts more than time_limit ticks of the constant-rate
  // wall clock after its own start; the moves it made are recorded, and the
  // next launch continues its stream where it stopped.  With several waves
  // wave 0's clock reading decides for all of them (through LDS, double
  // buffered by block parity, so every wave takes the same branch).
  __shared__ int stop_flag[2];
  const uint64_t t_begin = time_limit > 0 ? uint64_t(wall_clock64()) : 0u;
  auto out_of_time = [&](int64_t blk0) -> bool {
    if (time_limit <= 0 || blk0 == 0) return false;
    int stop = int(uint64_t(wall_clock64()) - t_begin > uint64_t(time_limit));
    if constexpr (kW > 1) {
      int* f = &stop_flag[(blk0 >> 6) & 1];
      if (tid == 0) *f = stop;
      __syncthreads();
      stop = *f;
    }
    return stop != 0;
  };
  // constraints and bias of the chain (kHooks)
  [[maybe_unused]] uint32_t hk_fixed = 0u;
  [[maybe_unused]] const uint8_t* hk_allow = nullptr;
  [[maybe_unused]] int hk_att = 1, hk_e1 = -1, hk_e2 = -1, hk_nm = 0, hk_bsp = -1, hk_bcnt = 0;
  [[maybe_unused]] const int32_t* hk_nb = nullptr;
  [[maybe_unused]] const double* hk_btab = nullptr;
  [[maybe_unused]] int64_t hk_pairs = 0;
  [[maybe_unused]] double hk_benergy = 0.0;
  [[maybe_unused]] bool hk_failed = false;
  [[maybe_unused]] uint64_t hk_fail_move = 0;
  [[maybe_unused]] bool hk_recycle = false;
  if constexpr (kHooks) {
    const HookDev* hk = hooks + chain;
    hk_fixed = hk->fixed_species;
    hk_allow = hk->allowed;
    hk_att = hk->max_attempts;
    hk_e1 = hk->pair_e1;
    hk_e2 = hk->pair_e2;
    hk_nm = hk->pair_nm;
    hk_nb = hk->pair_nb;
    hk_pairs = hk->pair_count;
    hk_bsp = hk->bias_species;
    hk_bcnt = hk->bias_count;
    hk_btab = hk->bias_tab;
    hk_benergy = hk->bias_energy;
    hk_recycle = hk->recycle != 0;
  }
  if constexpr (kK == 1) {
  for (int64_t blk0 = 0; blk0 < nmov; blk0 += 64) {
    if (out_of_time(blk0)) {
      moves_made = blk0;
      break;
    }
    // lane l holds the Philox block and log u of move blk0 + l, and decodes
    // the parts of that move that do not depend on the state
    const Philox4 blk = philox4x64_10(t0 + uint64_t(blk0) + uint64_t(lane), 0, 0, 0, seed, uid);
    const double log_u = log(unit_double(blk.w3));
    const int n_in = int(nmov - blk0 < 64 ? nmov - blk0 : 64);
    uint32_t v_old = 0, v_new = 0;  // canonical: species a and b; SGC: offset of the new species
    int v_ta = 0, v_tb = 0;         // canonical: tracker slots of the two sites
    int v_site = 0, v_p = 0;        // SGC: the site and its LDS index
    uint32_t v_img = 0;             // SGC, periodic: image word of the site
    if constexpr (kCanonical) {
      const uint32_t ia = bounded32(uint32_t(blk.w0 >> 32), np);
      uint32_t ib = ia + 1u + bounded32(uint32_t(blk.w0), np - 1u);
      ib -= ib >= np ? np : 0u;
      v_old = (present_pk >> (4u * ia)) & 15u;
      v_new = (present_pk >> (4u * ib)) & 15u;
      v_ta = unpack16(seg_lo, seg_hi, ia) + int(bounded64(blk.w1, uint64_t(unpack16(cnt_lo, cnt_hi, ia))));
      v_tb = unpack16(seg_lo, seg_hi, ib) + int(bounded64(blk.w2, uint64_t(unpack16(cnt_lo, cnt_hi, ib))));
    } else {
      const int pick = int(bounded64(blk.w1, uint64_t(n_pick)));
      v_site = has_bkg ? m.movable[pick] : pick;
      const Ref r = ref_of<kPeriodic>(c, v_site);
      v_p = r.p;
      v_img = image_word<kPeriodic>(c, r);
      v_new = bounded32(uint32_t(blk.w0), uint32_t(c.S - 1));
    }
    uint64_t acc_bits = 0;
    int l_end = n_in;  // moves of the block made (a failed constraint search ends the launch early)
    for (int l = 0; l < n_in; ++l) {
      double lu = readlane_f64(log_u, l);
      uint64_t u_bits = 0;  // hooks: the uniform of the proposal that passed

      int si, sj = 0, ta = 0, tb = 0;
      uint32_t old_i = 0, new_i = 0;
      Ref ri, rj;
      uint32_t patch_j = 0u;
      uint32_t img_h = 0u;  // hooks, SGC: image word of the flipped site
      int64_t dpair = 0;    // hooks: change of the constrained pair count
      int dcount = 0;       // hooks: change of the biased species count
      if constexpr (kHooks) {
        // proposal k of move t uses Philox counter (t, k, 0, 0); the first
        // that violates no constraint is made (montecarlo.py:970-991: with
        // max_attempts proposals none valid, the reference raises)
        const uint64_t tmove = t0 + uint64_t(blk0) + uint64_t(l);
        Philox4 pb{readlane_u64(blk.w0, l), readlane_u64(blk.w1, l), readlane_u64(blk.w2, l),
                   readlane_u64(blk.w3, l)};
        bool ok = false;
        for (int k = 0;;) {
          if constexpr (kCanonical) {
            const uint32_t ia = bounded32(uint32_t(pb.w0 >> 32), np);
            uint32_t ib = ia + 1u + bounded32(uint32_t(pb.w0), np - 1u);
            ib -= ib >= np ? np : 0u;
            old_i = (present_pk >> (4u * ia)) & 15u;
            new_i = (present_pk >> (4u * ib)) & 15u;
            ta = unpack16(seg_lo, seg_hi, ia) + int(bounded64(pb.w1, uint64_t(unpack16(cnt_lo, cnt_hi, ia))));
            tb = unpack16(seg_lo, seg_hi, ib) + int(bounded64(pb.w2, uint64_t(unpack16(cnt_lo, cnt_hi, ib))));
            si = __builtin_amdgcn_readfirstlane(trk[ta]);
            sj = __builtin_amdgcn_readfirstlane(trk[tb]);
            ok = hook_allows(hk_fixed, hk_allow, si, old_i, new_i) && hook_allows(hk_fixed, hk_allow, sj, new_i, old_i);
            if (ok && hk_e1 >= 0) {
              dpair = hook_pair_delta<kPeriodic>(c, occ, hk_nb, hk_nm, hk_e1, hk_e2, si, old_i, new_i, sj, new_i,
                                                 old_i, true, lane);
              ok = hk_pairs + dpair == 0;
            }
          } else {
            const int pick = int(bounded64(pb.w1, uint64_t(n_pick)));
            si = has_bkg ? m.movable[pick] : pick;
            old_i = occ[ref_of<kPeriodic>(c, si).p];
            old_i = __builtin_amdgcn_readfirstlane(old_i);
            new_i = old_i + 1u + bounded32(uint32_t(pb.w0), uint32_t(c.S - 1));
            new_i -= new_i >= uint32_t(c.S) ? uint32_t(c.S) : 0u;
            ok = hook_allows(hk_fixed, hk_allow, si, old_i, new_i);
            if (ok && hk_e1 >= 0) {
              dpair = hook_pair_delta<kPeriodic>(c, occ, hk_nb, hk_nm, hk_e1, hk_e2, si, old_i, new_i, si, new_i,
                                                 new_i, false, lane);
              ok = hk_pairs + dpair == 0;
            }
          }
          if (ok || ++k >= hk_att) break;
          pb = philox4x64_10(tmove, uint64_t(k), 0, 0, seed, uid);
        }
        if (!ok) {
          hk_failed = true;
          hk_fail_move = tmove;
          l_end = l;
          break;
        }
        lu = log(unit_double(pb.w3));
        u_bits = pb.w3;
        ri = ref_of<kPeriodic>(c, si);
        if constexpr (kCanonical) {
          rj = ref_of<kPeriodic>(c, sj);
          patch_j = kPeriodic ? wrapped_delta(c, ri, rj) : uint32_t(si);
        } else {
          rj = ri;
          img_h = image_word<kPeriodic>(c, ri);
          if (hk_bsp >= 0) dcount = int(new_i == uint32_t(hk_bsp)) - int(old_i == uint32_t(hk_bsp));
        }
      } else if constexpr (kCanonical) {
        old_i = __builtin_amdgcn_readlane(v_old, l);
        new_i = __builtin_amdgcn_readlane(v_new, l);
        ta = __builtin_amdgcn_readlane(v_ta, l);
        tb = __builtin_amdgcn_readlane(v_tb, l);
        si = __builtin_amdgcn_readfirstlane(trk[ta]);
        sj = __builtin_amdgcn_readfirstlane(trk[tb]);
        ri = ref_of<kPeriodic>(c, si);
        rj = ref_of<kPeriodic>(c, sj);
        patch_j = kPeriodic ? wrapped_delta(c, ri, rj) : uint32_t(si);
      } else {
        si = __builtin_amdgcn_readlane(v_site, l);
        ri.site = si;
        ri.p = __builtin_amdgcn_readlane(v_p, l);
        ri.x = ri.y = ri.z = 0;
        rj = ri;
      }
      int gi_i[kR > 0 ? kR : 1], gi_j[kR > 0 ? kR : 1];
      if constexpr (!kCanonical && !kHooks) old_i = occ[ri.p];
      if constexpr (kR > 0) {
        // all member gathers of the move are issued together
        table_index<kPeriodic, kR, kBM, false, kNM, kILW>(c, rr, ri, occ, 0u, 0u, gi_i);
        if constexpr (kCanonical)
          table_index<kPeriodic, kR, kBM, true, kNM, kILW>(c, rr, rj, occ, patch_j, new_i, gi_j);
      }
      if constexpr (!kCanonical && !kHooks) {
        new_i = old_i + 1u + __builtin_amdgcn_readlane(v_new, l);
        new_i -= new_i >= uint32_t(c.S) ? uint32_t(c.S) : 0u;
      }
      double dphi[kBM];
      double de_single = 0.0;  // a swap's two singlet changes cancel exactly
      if constexpr (kBM == 1) {
        dphi[0] = old_i == 0 ? dphi01 : dphi10;
        if constexpr (!kCanonical) de_single = old_i == 0 ? dsin01 : dsin10;
      } else {
        const int pair_i = int(old_i) * c.S + int(new_i);
#pragma unroll
        for (int d = 0; d < kBM; ++d) dphi[d] = d < c.B ? dphi_tab[pair_i * c.B + d] : 0.0;
        if constexpr (!kCanonical) de_single = dsingle_tab[pair_i];
      }
      // SGC: change of lane d's singlet if the flip is accepted
      double dsing = 0.0;
      if constexpr (!kCanonical) {
        if constexpr (kBM == 1)
          dsing = __dmul_rn(dphi[0], inv_n);
        else
          dsing = __dmul_rn(dphi_tab[(int(old_i) * c.S + int(new_i)) * c.B + (lane < c.B ? lane : 0)], inv_n);
      }
      double dphj[kBM];
#pragma unroll
      for (int d = 0; d < kBM; ++d) dphj[d] = -dphi[d];
      double part;
      if constexpr (kR > 0) {
        part = table_sum<kR, kBM, kILW>(gi_i, gtab, dphi);
        if constexpr (kCanonical) {
          // second flip of the swap: j goes b -> a and sees site i already at b
          // (ce_updater.cpp:845-852 runs update_cf twice in sequence)
          const double pj = table_sum<kR, kBM, kILW>(gi_j, gtab, dphj);
          part = kBM == 1 ? part - pj : part + pj;
        }
      } else {
        const int gi = multi_group ? int(m.site_group[si]) : 0;
        part = lds_flip_de<kPeriodic, kBM, false, kILW>(c, recs_l, grp_tab[gi], grp_tab[kMaxGroups + gi], tid, kThreads,
                                                  ri, occ, gtab, dphi, 0u, 0u);
        if constexpr (kCanonical) {
          const int gj = multi_group ? int(m.site_group[sj]) : 0;
          const double pj = lds_flip_de<kPeriodic, kBM, true, kILW>(c, recs_l, grp_tab[gj], grp_tab[kMaxGroups + gj], tid,
                                                              kThreads, rj, occ, gtab, dphj, patch_j, new_i);
          part = kBM == 1 ? part - pj : part + pj;
        }
      }
      const double sum = group_sum<kW>(part, red_move[l & 1], wave, lane);
      double dE = kBM == 1 ? fma(dphi[0], sum, de_single) : sum + de_single;
      double dbias = 0.0;
      if constexpr (kHooks) {
        // the bias of the proposed state joins its energy (montecarlo.py:926-930)
        if (dcount != 0) dbias = hk_btab[hk_bcnt + dcount] - hk_btab[hk_bcnt];
        dE += dbias;
      }
      // Metropolis test of the reference: accept if E_new < E_cur, else if
      // u <= exp(-(E_new - E_cur) / kT) (cemc/mcmc/montecarlo.py:950-956).  The
      // comparison is first made against log(u) with a margin far above the
      // rounding of log/exp; only moves inside the margin evaluate exp.  The
      // outcome is made uniform (scalar) so the branches below need no masks.
      const double x = -dE * inv_kT;
      const double margin = 1e-12 * (1.0 + fabs(x));
      int verdict = (dE < 0.0 || lu < x - margin) ? 1 : (lu > x + margin ? 0 : 2);
      [[maybe_unused]] double rw0 = 0.0, rw1 = 0.0, re0 = 0.0, rs0 = 0.0;
      if constexpr (kHooks) {
        if (hk_recycle) {
          // heat-bath choice between the current and the proposed state with
          // weights w = exp(-(E - min E) / kT): the proposed one when
          // u > w0 / (w0 + w1) (util.py:50-63, montecarlo.py:937-948)
          const double e0 = energy, e1 = energy + dE;
          const double emin = e0 < e1 ? e0 : e1;
          rw0 = exp(-(inv_kT * (e0 - emin)));
          rw1 = exp(-(inv_kT * (e1 - emin)));
          verdict = int(unit_double(u_bits) > rw0 / (rw0 + rw1));
          re0 = e0;
          rs0 = sing;
        }
      }
      verdict = __builtin_amdgcn_readfirstlane(verdict);
      if (verdict == 2)
        verdict = __builtin_amdgcn_readfirstlane(
            int(unit_double(kHooks ? u_bits : readlane_u64(blk.w3, l)) <= exp(-dE / kT)));
      if (verdict) {
        flush();
        if constexpr (kHooks) {
          hk_pairs += dpair;
          hk_bcnt += dcount;
          hk_benergy += dbias;
        }
        if constexpr (kPeriodic) {
          // lanes 0-7 write the images of site i, lanes 8-15 those of site j
          const bool second = kCanonical && lane >= 8;
          uint32_t img;
          if constexpr (kCanonical)
            img = second ? image_word<true>(c, rj) : image_word<true>(c, ri);
          else
            img = kHooks ? img_h : __builtin_amdgcn_readlane(v_img, l);
          const int idx = image_of(c, img, lane & 7);
          if (idx >= 0 && lane < (kCanonical ? 16 : 8)) occ[idx] = uint8_t(second ? old_i : new_i);
        } else {
          if (lane == 0) occ[si] = uint8_t(new_i);
          if (kCanonical && lane == 1) occ[sj] = uint8_t(old_i);
        }
        if constexpr (kCanonical) {
          if (lane == 0) {
            trk[ta] = uint16_t(sj);
            trk[tb] = uint16_t(si);
          }
        }
        energy += dE;
        if constexpr (!kCanonical) sing += dsing;
        acc_bits |= uint64_t(1) << l;
      }
      bool counted = false;
      if constexpr (kHooks) {
        if (hk_recycle) {
          // this move's sample weights the current and the proposed state
          // (montecarlo.py:805-811; singlets as mc_observers.py:231-252: the
          // state before the move and the state after it)
          const double inv_w = 1.0 / (rw0 + rw1);
          const double a = re0 - e_ref, b = re0 + dE - e_ref;
          sum_de += (a * rw0 + b * rw1) * inv_w;
          sum_de2 += (a * a * rw0 + b * b * rw1) * inv_w;
          if constexpr (!kCanonical) {
            sum_s += (rs0 * rw0 + sing * rw1) * inv_w;
            sum_s2 += (rs0 * rs0 * rw0 + sing * sing * rw1) * inv_w;
            sum_se += (rs0 * a * rw0 + sing * b * rw1) * inv_w;
          }
          counted = true;
        }
      }
      if (!counted) ++run;
      // this wave's own commit precedes its next gathers in program order
      // (LDS operations of one wave complete in order, so no wait is needed);
      // other waves are ordered by the barrier in the next move's group_sum
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    n_acc += uint64_t(__builtin_popcountll(acc_bits));
    if (acc_log != nullptr && tid < 2) {
      const int64_t w = blk0 >> 5;
      if (lane == 0) acc_log[w] = uint32_t(acc_bits);
      if (lane == 1 && n_in > 32) acc_log[w + 1] = uint32_t(acc_bits >> 32);
    }
    if constexpr (kHooks) {
      if (hk_failed) {
        moves_made = blk0 + l_end;
        break;
      }
    }
  }
    flush();
  } else if constexpr (kK == kSpecMoves) {
    // ---- rounds of kK = 8 moves (closed-form periodic lattice, records in
    // VGPRs, one translation group, no background).  The eight moves are
    // evaluated against the same state and their eight wave sums taken
    // together; lane L then decides move L >> 3 of the round.  An accepted
    // move invalidates the later moves of the round that read what it
    // changed (its sites or their sub-cluster members, or its tracker
    // slots): lane L checks move L & 7 against move L >> 3, and the round
    // ends before the first invalidated move, which the next round
    // evaluates again.  The committed moves of a round have disjoint
    // footprints, so their lanes write them at once; energy, singlets and
    // sample sums then advance in move order.  Every value therefore equals
    // the sequential loop's bitwise (same sums, same tree, same order).
    static_assert(kPeriodic && kR > 0 && kK == 8, "rounds need a periodic lattice, records in VGPRs, 8 moves");
    const uint64_t im0 = m.imask[0], im1 = m.imask[1];
    const int grp = lane >> 3;  // the move of the round this lane decides and commits
    const int sub = lane & 7;   // the earlier move it is checked against; its image (commit)
    uint32_t rnd = 0;
    for (int64_t blk0 = 0; blk0 < nmov; blk0 += 64) {
      if (out_of_time(blk0)) {
        moves_made = blk0;
        break;
      }
      const Philox4 blk = philox4x64_10(t0 + uint64_t(blk0) + uint64_t(lane), 0, 0, 0, seed, uid);
      const double log_u = log(unit_double(blk.w3));
      const int n_in = int(nmov - blk0 < 64 ? nmov - blk0 : 64);
      // lane l decodes move blk0 + l once per block
      uint32_t v_old = 0, v_new = 0;  // canonical: species a and b; SGC: offset of the new species
      int v_ta = 0, v_tb = 0;         // canonical: tracker slots of the two sites
      int v_p = 0;                    // SGC: LDS index of the site
      uint32_t v_cell = 0;            // SGC: packed cell of the site
      uint32_t v_img = 0;             // SGC: image word of the site
      if constexpr (kCanonical) {
        const uint32_t ia = bounded32(uint32_t(blk.w0 >> 32), np);
        uint32_t ib = ia + 1u + bounded32(uint32_t(blk.w0), np - 1u);
        ib -= ib >= np ? np : 0u;
        v_old = (present_pk >> (4u * ia)) & 15u;
        v_new = (present_pk >> (4u * ib)) & 15u;
        v_ta = unpack16(seg_lo, seg_hi, ia) + int(bounded64(blk.w1, uint64_t(unpack16(cnt_lo, cnt_hi, ia))));
        v_tb = unpack16(seg_lo, seg_hi, ib) + int(bounded64(blk.w2, uint64_t(unpack16(cnt_lo, cnt_hi, ib))));
      } else {
        const Ref r = ref_of<true>(c, int(bounded64(blk.w1, uint64_t(N))));
        v_p = r.p;
        v_cell = uint32_t(r.x) | (uint32_t(r.y) << 8) | (uint32_t(r.z) << 16);
        v_img = image_word<true>(c, r);
        v_new = bounded32(uint32_t(blk.w0), uint32_t(c.S - 1));
      }
      uint64_t acc_bits = 0;
      int l0 = 0;
      while (l0 < n_in) {
        const int n = n_in - l0 < kK ? n_in - l0 : kK;
        // every lane looks up its own move in the current state
        int p_i = v_p, p_j = 0;
        uint32_t cell_i = v_cell, cell_j = 0, patch = 0, o_v = v_old, n_v = v_new;
        uint32_t img_i = v_img, img_j = 0;
        uint32_t sites = 0, slots = 0;  // canonical: si | sj << 16, ta | tb << 16
        if constexpr (kCanonical) {
          const int si = trk[v_ta], sj = trk[v_tb];
          const Ref ri = ref_of<true>(c, si), rj = ref_of<true>(c, sj);
          p_i = ri.p;
          p_j = rj.p;
          cell_i = uint32_t(ri.x) | (uint32_t(ri.y) << 8) | (uint32_t(ri.z) << 16);
          cell_j = uint32_t(rj.x) | (uint32_t(rj.y) << 8) | (uint32_t(rj.z) << 16);
          img_i = image_word<true>(c, ri);
          img_j = image_word<true>(c, rj);
          patch = wrapped_delta(c, ri, rj);
          sites = uint32_t(si) | (uint32_t(sj) << 16);
          slots = uint32_t(v_ta) | (uint32_t(v_tb) << 16);
        } else {
          o_v = occ[v_p];
          n_v = o_v + 1u + v_new;
          n_v -= n_v >= uint32_t(c.S) ? uint32_t(c.S) : 0u;
        }
        // the lanes of move grp fetch it (lane l0 + grp) and the move they
        // check it against (lane l0 + sub); slots past the block read lane 63
        const int src_b = 4 * (l0 + grp < 63 ? l0 + grp : 63);
        const int src_a = 4 * (l0 + sub < 63 ? l0 + sub : 63);
        const double lu_b = join_f64(uint32_t(__builtin_amdgcn_ds_bpermute(src_b, int(lo_of(log_u)))),
                                     uint32_t(__builtin_amdgcn_ds_bpermute(src_b, int(hi_of(log_u)))));
        const uint32_t cib = uint32_t(__builtin_amdgcn_ds_bpermute(src_b, int(cell_i | (o_v << 24) | (n_v << 28))));
        const uint32_t cia = uint32_t(__builtin_amdgcn_ds_bpermute(src_a, int(cell_i)));
        const uint32_t wib = uint32_t(__builtin_amdgcn_ds_bpermute(src_b, int(img_i)));
        uint32_t cjb = 0, cja = 0, slb = 0, sla = 0, stb = 0, wjb = 0;
        if constexpr (kCanonical) {
          wjb = uint32_t(__builtin_amdgcn_ds_bpermute(src_b, int(img_j)));
          cjb = uint32_t(__builtin_amdgcn_ds_bpermute(src_b, int(cell_j)));
          cja = uint32_t(__builtin_amdgcn_ds_bpermute(src_a, int(cell_j)));
          slb = uint32_t(__builtin_amdgcn_ds_bpermute(src_b, int(slots)));
          sla = uint32_t(__builtin_amdgcn_ds_bpermute(src_a, int(slots)));
          stb = uint32_t(__builtin_amdgcn_ds_bpermute(src_b, int(sites)));
        }
        const uint32_t o_b = (cib >> 24) & 15u, n_b = cib >> 28;
        // each lane's records for every move of the round
        double part[kK];
#pragma unroll
        for (int k = 0; k < kK; ++k) {
          const int lk = l0 + k < 63 ? l0 + k : 63;
          Ref ri;
          ri.p = __builtin_amdgcn_readlane(p_i, lk);
          ri.site = 0;
          ri.x = ri.y = ri.z = 0;
          int pj_k = kCanonical ? __builtin_amdgcn_readlane(p_j, lk) : 0;
          // many-species models with many records per lane (or with eight
          // waves, 256 registers each) evaluate the moves one after the
          // other (the addresses of move k wait for the sum of move k - 1);
          // hoisting every move's loads to the top of the round would
          // overflow the registers
          if constexpr (kBM > 1 && (kR >= 4 || kW >= 8))
            if (k > 0) asm volatile("" : "+s"(ri.p), "+s"(pj_k) : "v"(part[k - 1]));
          double dphi[kBM];
          if constexpr (kBM == 1) {
            dphi[0] = 0.0;  // two species: the entries are summed plain
          } else {
            const int pair =
                int(__builtin_amdgcn_readlane(o_v, lk)) * c.S + int(__builtin_amdgcn_readlane(n_v, lk));
#pragma unroll
            for (int d = 0; d < kBM; ++d) dphi[d] = d < c.B ? dphi_tab[pair * c.B + d] : 0.0;
          }
          int gi_i[kR];
          table_index<true, kR, kBM, false, kNM, kILW>(c, rr, ri, occ, 0u, 0u, gi_i);
          double pk = table_sum<kR, kBM, kILW>(gi_i, gtab, dphi);
          if constexpr (kCanonical) {
            Ref rj;
            rj.p = pj_k;
            rj.site = 0;
            rj.x = rj.y = rj.z = 0;
            double dphj[kBM];
#pragma unroll
            for (int d = 0; d < kBM; ++d) dphj[d] = -dphi[d];
            int gi_j[kR];
            table_index<true, kR, kBM, true, kNM, kILW>(c, rr, rj, occ, __builtin_amdgcn_readlane(patch, lk),
                                                  __builtin_amdgcn_readlane(n_v, lk), gi_j);
            const double pj = table_sum<kR, kBM, kILW>(gi_j, gtab, dphj);
            pk = kBM == 1 ? pk - pj : pk + pj;
          }
          part[k] = pk;
        }
        // lane L: the energy sum of move L >> 3 (over the waves in order)
        double tot = wave_sum8_f64(part, lane);
        if constexpr (kW > 1) {
          double* sl = red_spec[rnd & 1u];
          if (sub == 0) sl[grp * 8 + wave] = tot;
          __syncthreads();
          tot = sl[grp * 8];
#pragma unroll
          for (int v = 1; v < kW; ++v) tot += sl[grp * 8 + v];
        }
        double dphi0 = 0.0, de_single = 0.0, dsl = 0.0;
        if constexpr (kBM == 1) {
          dphi0 = o_b == 0 ? dphi01 : dphi10;
          if constexpr (!kCanonical) {
            de_single = o_b == 0 ? dsin01 : dsin10;
            dsl = __dmul_rn(dphi0, inv_n);
          }
        } else if constexpr (!kCanonical) {
          const int pair = int(o_b) * c.S + int(n_b);
          de_single = dsingle_tab[pair];
          // lane (move, d): singlet d's change, d < B
          dsl = __dmul_rn(dphi_tab[pair * c.B + (sub < c.B ? sub : 0)], inv_n);
        }
        const double dE = kBM == 1 ? fma(dphi0, tot, de_single) : tot + de_single;
        // the Metropolis test of the sequential loop, per move
        const double x = -dE * inv_kT;
        const double margin = 1e-12 * (1.0 + fabs(x));
        const bool valid = grp < n;
        int verdict = (dE < 0.0 || lu_b < x - margin) ? 1 : (lu_b > x + margin ? 0 : 2);
        uint64_t unsure = __builtin_amdgcn_ballot_w64(valid && verdict == 2);
        while (unsure != 0) {
          const int lb = int(__builtin_ctzll(unsure));
          const int k = lb >> 3;
          const double dEk = readlane_f64(dE, lb);
          const int ok = int(unit_double(readlane_u64(blk.w3, l0 + k)) <= exp(-dEk / kT));
          verdict = grp == k ? ok : verdict;
          unsure &= ~(uint64_t(0xFF) << lb);
        }
        const uint64_t acc = __builtin_amdgcn_ballot_w64(valid && verdict == 1);  // byte k: move k accepted
        int stop = n;
        if (acc != 0) {
          bool hit = false;
          if (sub < grp && ((acc >> (8 * sub)) & 1u) != 0) {
            const uint32_t ib = cib & 0xFFFFFFu;
            hit = cells_interact(cia, ib, c, im0, im1);
            if constexpr (kCanonical) {
              const uint32_t ta_a = sla & 0xFFFFu, tb_a = sla >> 16, ta_b = slb & 0xFFFFu, tb_b = slb >> 16;
              hit = hit || cells_interact(cia, cjb, c, im0, im1) || cells_interact(cja, ib, c, im0, im1) ||
                    cells_interact(cja, cjb, c, im0, im1) || ta_a == ta_b || ta_a == tb_b || tb_a == ta_b ||
                    tb_a == tb_b;
            }
          }
          const uint64_t hits = __builtin_amdgcn_ballot_w64(hit);
          if (hits != 0) {
            const int first = int(__builtin_ctzll(hits)) >> 3;
            stop = first < stop ? first : stop;
          }
        }
        const uint64_t com = acc & (stop >= 8 ? ~uint64_t(0) : (uint64_t(1) << (8 * stop)) - 1u);
        int cursor = 0;
        if (com != 0) {
          if (((com >> lane) & 1u) != 0) {
            // the lanes of a committed move write its images (site j's too)
            const int idx = image_of(c, wib, sub);
            if (idx >= 0) occ[idx] = uint8_t(n_b);
            if constexpr (kCanonical) {
              const int jdx = image_of(c, wjb, sub);
              if (jdx >= 0) occ[jdx] = uint8_t(o_b);
              if (sub == 0) trk[slb & 0xFFFFu] = uint16_t(stb >> 16);  // trk[ta] = sj
              if (sub == 1) trk[slb >> 16] = uint16_t(stb & 0xFFFFu);  // trk[tb] = si
            }
          }
          // energy, singlets and sample sums in move order
          uint64_t rest = com;
          while (rest != 0) {
            const int lb = int(__builtin_ctzll(rest));
            const int k = lb >> 3;
            rest &= ~(uint64_t(0xFF) << lb);
            run += k - cursor;
            flush();
            energy += readlane_f64(dE, lb);
            if constexpr (!kCanonical) {
              double ds = readlane_f64(dsl, lb);
#pragma unroll
              for (int d = 1; d < kBM; ++d) {
                const double t = readlane_f64(dsl, lb + d);
                ds = lane == d ? t : ds;
              }
              sing += ds;
            }
            run += 1;
            cursor = k + 1;
            acc_bits |= uint64_t(1) << (l0 + k);
          }
        }
        run += stop - cursor;
        l0 += stop;
        ++rnd;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);  // as in the one-move loop
      }
      n_acc += uint64_t(__builtin_popcountll(acc_bits));
      if (acc_log != nullptr && tid < 2) {
        const int64_t w = blk0 >> 5;
        if (lane == 0) acc_log[w] = uint32_t(acc_bits);
        if (lane == 1 && n_in > 32) acc_log[w + 1] = uint32_t(acc_bits >> 32);
      }
    }
    flush();
  } else {
    // ---- wide rounds of kK = 8 kW moves (closed-form periodic lattice,
    // records in VGPRs, one translation group, no background).  Every wave
    // holds all the records, laid out as for one wave, and wave w evaluates
    // moves 8w .. 8w+7 of the round against the round's starting state.  The
    // waves' move sums meet in LDS (one barrier); then every lane decides
    // the move it decoded (lane l: move blk0 + l), and every wave finds the
    // same cut: the first move that an accepted earlier move of the round
    // affects (its sites or their sub-cluster members, or its tracker
    // slots), where the round ends as in rounds of eight.  Wave w writes the
    // commits of its own eight moves; a second barrier orders them before
    // the next round's reads.  All waves carry the same energy, singlets and
    // sums (wave 0's are written back).  Each move's energy change is the
    // one-wave sum (same records per lane, same tree), so every value equals
    // the sequential loop's bitwise.
    static_assert(kPeriodic && kR > 0 && kW > 1 && kK == kSpecMoves * kW,
                  "wide rounds need a periodic lattice, records in VGPRs and 8 moves per wave");
    const uint64_t im0 = m.imask[0], im1 = m.imask[1];
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int grp = lane >> 3;  // which of this wave's eight moves the lane commits
    const int sub = lane & 7;   // which image of it the lane writes
    for (int64_t blk0 = 0; blk0 < nmov; blk0 += 64) {
      if (out_of_time(blk0)) {
        moves_made = blk0;
        break;
      }
      const Philox4 blk = philox4x64_10(t0 + uint64_t(blk0) + uint64_t(lane), 0, 0, 0, seed, uid);
      const double log_u = log(unit_double(blk.w3));
      const int n_in = int(nmov - blk0 < 64 ? nmov - blk0 : 64);
      uint32_t v_old = 0, v_new = 0;
      int v_ta = 0, v_tb = 0;
      int v_p = 0;
      uint32_t v_cell = 0, v_img = 0;
      if constexpr (kCanonical) {
        const uint32_t ia = bounded32(uint32_t(blk.w0 >> 32), np);
        uint32_t ib = ia + 1u + bounded32(uint32_t(blk.w0), np - 1u);
        ib -= ib >= np ? np : 0u;
        v_old = (present_pk >> (4u * ia)) & 15u;
        v_new = (present_pk >> (4u * ib)) & 15u;
        v_ta = unpack16(seg_lo, seg_hi, ia) + int(bounded64(blk.w1, uint64_t(unpack16(cnt_lo, cnt_hi, ia))));
        v_tb = unpack16(seg_lo, seg_hi, ib) + int(bounded64(blk.w2, uint64_t(unpack16(cnt_lo, cnt_hi, ib))));
      } else {
        const Ref r = ref_of<true>(c, int(bounded64(blk.w1, uint64_t(N))));
        v_p = r.p;
        v_cell = uint32_t(r.x) | (uint32_t(r.y) << 8) | (uint32_t(r.z) << 16);
        v_img = image_word<true>(c, r);
        v_new = bounded32(uint32_t(blk.w0), uint32_t(c.S - 1));
      }
      uint64_t acc_bits = 0;
      int l0 = 0;
      while (l0 < n_in) {
        const int n = n_in - l0 < kK ? n_in - l0 : kK;
        // every lane looks up its own move in the current state
        int p_i = v_p, p_j = 0;
        uint32_t cell_i = v_cell, cell_j = 0, patch = 0, o_v = v_old, n_v = v_new;
        uint32_t img_i = v_img, img_j = 0;
        uint32_t sites = 0, slots = 0;  // canonical: si | sj << 16, ta | tb << 16
        if constexpr (kCanonical) {
          const int si = trk[v_ta], sj = trk[v_tb];
          const Ref ri = ref_of<true>(c, si), rj = ref_of<true>(c, sj);
          p_i = ri.p;
          p_j = rj.p;
          cell_i = uint32_t(ri.x) | (uint32_t(ri.y) << 8) | (uint32_t(ri.z) << 16);
          cell_j = uint32_t(rj.x) | (uint32_t(rj.y) << 8) | (uint32_t(rj.z) << 16);
          img_i = image_word<true>(c, ri);
          img_j = image_word<true>(c, rj);
          patch = wrapped_delta(c, ri, rj);
          sites = uint32_t(si) | (uint32_t(sj) << 16);
          slots = uint32_t(v_ta) | (uint32_t(v_tb) << 16);
        } else {
          o_v = occ[v_p];
          n_v = o_v + 1u + v_new;
          n_v -= n_v >= uint32_t(c.S) ? uint32_t(c.S) : 0u;
        }
        // this wave's eight moves: base + k
        const int base = l0 + kSpecMoves * wv;
        double part[kSpecMoves];
#pragma unroll
        for (int k = 0; k < kSpecMoves; ++k) {
          const int lk = base + k < 63 ? base + k : 63;
          Ref ri;
          ri.p = __builtin_amdgcn_readlane(p_i, lk);
          ri.site = 0;
          ri.x = ri.y = ri.z = 0;
          int pj_k = kCanonical ? __builtin_amdgcn_readlane(p_j, lk) : 0;
          // as in rounds of eight: one move after the other where hoisting
          // every move's loads would overflow the registers
          if constexpr (kBM > 1 && (kR >= 4 || kW >= 8))
            if (k > 0) asm volatile("" : "+s"(ri.p), "+s"(pj_k) : "v"(part[k - 1]));
          double dphi[kBM];
          if constexpr (kBM == 1) {
            dphi[0] = 0.0;
          } else {
            const int pair =
                int(__builtin_amdgcn_readlane(o_v, lk)) * c.S + int(__builtin_amdgcn_readlane(n_v, lk));
#pragma unroll
            for (int d = 0; d < kBM; ++d) dphi[d] = d < c.B ? dphi_tab[pair * c.B + d] : 0.0;
          }
          int gi_i[kR];
          table_index<true, kR, kBM, false, kNM, kILW>(c, rr, ri, occ, 0u, 0u, gi_i);
          double pk = table_sum<kR, kBM, kILW>(gi_i, gtab, dphi);
          if constexpr (kCanonical) {
            Ref rj;
            rj.p = pj_k;
            rj.site = 0;
            rj.x = rj.y = rj.z = 0;
            double dphj[kBM];
#pragma unroll
            for (int d = 0; d < kBM; ++d) dphj[d] = -dphi[d];
            int gi_j[kR];
            table_index<true, kR, kBM, true, kNM, kILW>(c, rr, rj, occ, __builtin_amdgcn_readlane(patch, lk),
                                                  __builtin_amdgcn_readlane(n_v, lk), gi_j);
            const double pj = table_sum<kR, kBM, kILW>(gi_j, gtab, dphj);
            pk = kBM == 1 ? pk - pj : pk + pj;
          }
          part[k] = pk;
        }
        // lane (grp, sub) holds the sum of move base + grp
        const double tot8 = wave_sum8_f64(part, lane);
        if (sub == 0) red_wide[kSpecMoves * wv + grp] = tot8;
        __syncthreads();
        // every lane: its own move, round position r
        const int r = lane - l0;
        const bool valid = r >= 0 && r < n;
        const double tot = red_wide[valid ? r : 0];
        double dphi0 = 0.0, de_single = 0.0;
        if constexpr (kBM == 1) {
          dphi0 = o_v == 0 ? dphi01 : dphi10;
          if constexpr (!kCanonical) de_single = o_v == 0 ? dsin01 : dsin10;
        } else if constexpr (!kCanonical) {
          de_single = dsingle_tab[int(o_v) * c.S + int(n_v)];
        }
        const double dE = kBM == 1 ? fma(dphi0, tot, de_single) : tot + de_single;
        const double x = -dE * inv_kT;
        const double margin = 1e-12 * (1.0 + fabs(x));
        int verdict = (dE < 0.0 || log_u < x - margin) ? 1 : (log_u > x + margin ? 0 : 2);
        if (valid && verdict == 2) verdict = int(unit_double(blk.w3) <= exp(-dE / kT));
        const uint64_t acc = __builtin_amdgcn_ballot_w64(valid && verdict == 1);  // bit l: move blk0 + l
        // the cut: the first move that an accepted earlier move affects
        int stop = n;
        uint64_t rest = acc;
        while (rest != 0) {
          const int i = int(__builtin_ctzll(rest));
          if (i - l0 >= stop - 1) break;  // later moves past the cut need no check
          rest &= rest - 1u;
          const uint32_t ca = uint32_t(__builtin_amdgcn_readlane(cell_i, i));
          bool hit = cells_interact(ca, cell_i, c, im0, im1);
          if constexpr (kCanonical) {
            const uint32_t cja = uint32_t(__builtin_amdgcn_readlane(cell_j, i));
            const uint32_t sla = uint32_t(__builtin_amdgcn_readlane(slots, i));
            const uint32_t ta_a = sla & 0xFFFFu, tb_a = sla >> 16, ta_b = slots & 0xFFFFu, tb_b = slots >> 16;
            hit = hit || cells_interact(ca, cell_j, c, im0, im1) || cells_interact(cja, cell_i, c, im0, im1) ||
                  cells_interact(cja, cell_j, c, im0, im1) || ta_a == ta_b || ta_a == tb_b || tb_a == ta_b ||
                  tb_a == tb_b;
          }
          const uint64_t hits = __builtin_amdgcn_ballot_w64(hit && valid && lane > i);
          if (hits != 0) {
            const int first = int(__builtin_ctzll(hits)) - l0;
            stop = first < stop ? first : stop;
          }
        }
        const uint64_t below = l0 + stop >= 64 ? ~uint64_t(0) : (uint64_t(1) << (l0 + stop)) - 1u;
        const uint64_t com = acc & below;
        int cursor = l0;
        if (com != 0) {
          // wave w writes the commits of its moves base + grp: lane (grp, sub)
          // image sub of the move's site i (and of site j), lanes 0 and 1 of
          // the group the tracker entries
          const int mb = base + grp;
          const int src = 4 * (mb < 63 ? mb : 63);
          const uint32_t wib = uint32_t(__builtin_amdgcn_ds_bpermute(src, int(img_i)));
          const uint32_t onb = uint32_t(__builtin_amdgcn_ds_bpermute(src, int(o_v | (n_v << 4))));
          uint32_t wjb = 0, slb = 0, stb = 0;
          if constexpr (kCanonical) {
            wjb = uint32_t(__builtin_amdgcn_ds_bpermute(src, int(img_j)));
            slb = uint32_t(__builtin_amdgcn_ds_bpermute(src, int(slots)));
            stb = uint32_t(__builtin_amdgcn_ds_bpermute(src, int(sites)));
          }
          if (mb < 64 && ((com >> (mb & 63)) & 1u) != 0) {
            const int idx = image_of(c, wib, sub);
            if (idx >= 0) occ[idx] = uint8_t(onb >> 4);
            if constexpr (kCanonical) {
              const int jdx = image_of(c, wjb, sub);
              if (jdx >= 0) occ[jdx] = uint8_t(onb & 15u);
              if (sub == 0) trk[slb & 0xFFFFu] = uint16_t(stb >> 16);  // trk[ta] = sj
              if (sub == 1) trk[slb >> 16] = uint16_t(stb & 0xFFFFu);  // trk[tb] = si
            }
          }
          // energy, singlets and sample sums in move order
          uint64_t left = com;
          while (left != 0) {
            const int i = int(__builtin_ctzll(left));
            left &= left - 1u;
            run += i - cursor;
            flush();
            energy += readlane_f64(dE, i);
            if constexpr (!kCanonical) {
              const uint32_t oi = uint32_t(__builtin_amdgcn_readlane(o_v, i));
              double ds;
              if constexpr (kBM == 1) {
                ds = __dmul_rn(oi == 0 ? dphi01 : dphi10, inv_n);
              } else {
                const int pair = int(oi) * c.S + int(__builtin_amdgcn_readlane(n_v, i));
                ds = __dmul_rn(dphi_tab[pair * c.B + (lane < c.B ? lane : 0)], inv_n);
              }
              sing += ds;
            }
            run += 1;
            cursor = i + 1;
            acc_bits |= uint64_t(1) << i;
          }
        }
        run += l0 + stop - cursor;
        l0 += stop;
        __syncthreads();  // the round's commits before the next round's reads
      }
      n_acc += uint64_t(__builtin_popcountll(acc_bits));
      if (acc_log != nullptr && tid < 2) {
        const int64_t w = blk0 >> 5;
        if (lane == 0) acc_log[w] = uint32_t(acc_bits);
        if (lane == 1 && n_in > 32) acc_log[w + 1] = uint32_t(acc_bits >> 32);
      }
    }
    flush();
  }

  // ---- write back
  if constexpr (kW > 1) __syncthreads();
  uint8_t* wocc = reinterpret_cast<uint8_t*>(occ_pool + pp->occ_off);
  for (int i = tid; i < N; i += kThreads) wocc[i] = occ[ref_of<kPeriodic>(c, i).p];
  if constexpr (kCanonical)
    for (int i = tid; i < N; i += kThreads) gtrk[i] = trk[i];
  ChainState* st = states + chain;
  if (tid == 0) {
    st->energy = energy;
    st->moves_done = t0 + uint64_t(moves_made);
    st->n_accept += n_acc;
    st->n_samples += uint64_t(moves_made);
    st->sum_de += sum_de;
    st->sum_de2 += sum_de2;
  }
  if (!kCanonical && tid < c.B) {
    st->sum_s[lane] += sum_s;
    st->sum_s2[lane] += sum_s2;
    st->sum_se[lane] += sum_se;
  }
  if constexpr (kHooks) {
    if (tid == 0) {
      HookDev* hk = hooks + chain;
      hk->pair_count = hk_pairs;
      hk->bias_count = hk_bcnt;
      hk->bias_energy = hk_benergy;
      if (hk_failed) {
        hk->status = 1;
        hk->failed_move = hk_fail_move;
      }
    }
  }
}


}  // namespace cemc
