// HIP kernels of the cluster-expansion Metropolis engine (gfx950).
//
//  sweep_kernel     one 64-lane wavefront per Markov chain; occupation,
//                   species tracker, sub-cluster records and energy tables
//                   staged in LDS; lanes split the changed sites' sub-cluster
//                   lists, partial energy changes are reduced with DPP; the
//                   Metropolis test runs in FP64.  Restates, per trial move,
//                   CEUpdater::update_cf + get_energy (cpp/src/ce_updater.cpp:
//                   236-242, 313-406) and the sampler step
//                   (cemc/mcmc/montecarlo.py:890-1038, sgc_montecarlo.py:62-76).
//  cf_partial/cf_finalize   correlation functions from scratch (the role
//                   CLEASE's CorrFunction plays for the reference,
//                   cemc/ce_calculator.py:169-177) and E = N * J.cf.
//  updater_kernel   sequence of single-site changes with per-correlation-
//                   function deltas written to a history slot, for the
//                   updater interface (ce_updater.cpp:313-406, 845-852).
//  restore_kernel   undo of popped changes (ce_updater.cpp:414-448).
//  slot_energy_kernel  E = N * sum_k J_k cf_k of one history slot.
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

uint32_t sweep_lds_bytes(int occ_bytes, int trk_sites, int n_rec_lds, int gtab_doubles) {
  return lds_layout(occ_bytes, trk_sites, n_rec_lds, gtab_doubles).total;
}

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

template <int kBM>
struct TableGeom {
  static constexpr int kBits = kBM == 1 ? 1 : (kBM == 3 ? 2 : 3);  // code bits per species
  static constexpr int kCodes = 1 << (3 * kBits);                   // codes per decoration slice
};

// Phase 1 of a flip: addresses of every member of kR records, issue of all
// member-species gathers, energy-table indices.  With kPatch, members that
// are the patch site (periodic: wrapped offset == patch_w; table: site ==
// patch_w) read `patch_species`.
template <bool kPeriodic, int kR, int kBM, bool kPatch>
__device__ __attribute__((always_inline)) inline void table_index(const SweepCtx& c, const FatRec (&rec)[kR],
                                                                  const Ref& ref, const uint8_t* occ,
                                                                  uint32_t patch_w, uint32_t patch_species,
                                                                  int (&gi)[kR]) {
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
      nb[k][2] = row[rec[k].m2];
    }
  }
  uint32_t sp[kR][3];
#pragma unroll
  for (int k = 0; k < kR; ++k)
#pragma unroll
    for (int q = 0; q < 3; ++q) sp[k][q] = occ[nb[k][q]];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    uint32_t s0 = sp[k][0], s1 = sp[k][1], s2 = sp[k][2];
    if constexpr (kPatch) {
      if constexpr (kPeriodic) {
        s0 = rec[k].w0 == patch_w ? patch_species : s0;
        s1 = rec[k].w1 == patch_w ? patch_species : s1;
        s2 = rec[k].w2 == patch_w ? patch_species : s2;
      } else {
        s0 = uint32_t(nb[k][0]) == patch_w ? patch_species : s0;
        s1 = uint32_t(nb[k][1]) == patch_w ? patch_species : s1;
        s2 = uint32_t(nb[k][2]) == patch_w ? patch_species : s2;
      }
    }
    // unused members need no masking: their table entries repeat over them
    const uint32_t code = s0 | (s1 << kBits) | (s2 << (2 * kBits));
    gi[k] = int(rec[k].meta & 0xFFFFFFu) + int(code);
  }
}

// Phase 2: energy-table reads.  kBM == 1 (two species): the plain sum of the
// entries, to be scaled by the single basis difference after the reduction;
// otherwise the change for basis differences dphi.
template <int kR, int kBM>
__device__ __attribute__((always_inline)) inline double table_sum(const int (&gi)[kR], const double* gtab,
                                                                  const double (&dphi)[kBM]) {
  constexpr int kCodes = TableGeom<kBM>::kCodes;
  double g[kR][kBM];
#pragma unroll
  for (int k = 0; k < kR; ++k)
#pragma unroll
    for (int d = 0; d < kBM; ++d) g[k][d] = gtab[gi[k] + d * kCodes];
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    if constexpr (kBM == 1) {
      v += g[k][0];
    } else {
#pragma unroll
      for (int d = 0; d < kBM; ++d) v += dphi[d] * g[k][d];
    }
  }
  return v;
}

// Records of one translation group read from LDS in batches of 4 per lane.
template <bool kPeriodic, int kBM, bool kPatch>
__device__ __attribute__((always_inline)) inline double lds_flip_de(const SweepCtx& c, const FatRec* recs, int off,
                                                                    int cnt, int lane, const Ref& ref,
                                                                    const uint8_t* occ, const double* gtab,
                                                                    const double (&dphi)[kBM], uint32_t patch_w,
                                                                    uint32_t patch_species) {
  double v = 0.0;
  for (int base = 0; base < cnt; base += 256) {
    FatRec rb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int it = base + lane + 64 * k;
      rb[k] = it < cnt ? recs[off + it] : FatRec{0, 0, 0, 0u, 0u, 0u, 0u, 0u};
    }
    int gi[4];
    table_index<kPeriodic, 4, kBM, kPatch>(c, rb, ref, occ, patch_w, patch_species, gi);
    v += table_sum<4, kBM>(gi, gtab, dphi);
  }
  return v;
}

__device__ __attribute__((always_inline)) inline int unpack16(uint64_t lo, uint64_t hi, uint32_t i) {
  return int(((i < 4u ? lo : hi) >> (16u * (i & 3u))) & 0xFFFFu);
}

// Padded index of image `bits` (bit 0: x, 1: y, 2: z take the other image)
// of the site at cell (x, y, z), or -1 when that image does not exist.
__device__ __attribute__((always_inline)) inline int image_index(const SweepCtx& c, int x, int y, int z, int bits) {
  int ix = x, iy = y, iz = z;
  bool ok = true;
  if (bits & 1) {
    ix = x < c.h ? x + c.L1 : (x >= c.L1 - c.h ? x - c.L1 : x);
    ok = ok && ix != x;
  }
  if (bits & 2) {
    iy = y < c.h ? y + c.L2 : (y >= c.L2 - c.h ? y - c.L2 : y);
    ok = ok && iy != y;
  }
  if (bits & 4) {
    iz = z < c.h ? z + c.L3 : (z >= c.L3 - c.h ? z - c.L3 : z);
    ok = ok && iz != z;
  }
  return ok ? int(__umul24(__umul24(uint32_t(ix + c.h), uint32_t(c.P2)) + uint32_t(iy + c.h), uint32_t(c.P3)) +
                  uint32_t(iz + c.h))
            : -1;
}

// kR > 0: single-group model, records in VGPRs (kR per lane); kR == 0:
// records in LDS.  kBM: decoration slices of the energy tables (1: two
// species, 3: up to four, 7: up to eight).
template <bool kPeriodic, bool kCanonical, int kR, bool kGtabLds, int kBM>
__global__ __launch_bounds__(64) void sweep_kernel(const ChainParams* __restrict__ params,
                                                   ChainState* __restrict__ states,
                                                   const ModelDev* __restrict__ models,
                                                   int8_t* __restrict__ occ_pool,
                                                   uint16_t* __restrict__ trk_pool,
                                                   const int32_t* __restrict__ chain_order,
                                                   const int64_t* __restrict__ n_moves,
                                                   uint32_t* __restrict__ accept_log, int64_t log_words) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ double dphi_tab[kMaxSpecies * kMaxSpecies * kMaxBasis];  // [(old*S+new)*B + d]
  __shared__ double dsingle_tab[kMaxSpecies * kMaxSpecies];          // singlet energy change per (old,new)
  __shared__ double phi_tab[kMaxBasis * kMaxSpecies];                // phi[d*S + s]
  __shared__ int grp_tab[2 * kMaxGroups];
  constexpr int kRecLds = kR == 0;
  const int chain = chain_order[blockIdx.x];
  const int lane = int(threadIdx.x);
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
  c.K = m.K;
  c.trans = m.trans;
  const bool multi_group = m.n_groups > 1 || m.has_background;
  const bool has_bkg = m.has_background != 0;
  const int n_pick = has_bkg ? m.n_movable : N;

  // ---- stage chain state, per-chain tables and model tables into LDS
  const uint8_t* gocc = reinterpret_cast<const uint8_t*>(occ_pool + pp->occ_off);
  if constexpr (kPeriodic) {
    const uint32_t p23 = uint32_t(c.P2 * c.P3), mag_p23 = 0xFFFFFFFFu / p23, mag_p3 = 0xFFFFFFFFu / uint32_t(c.P3);
    for (int p = lane; p < occ_bytes; p += 64) {
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
    for (int i = lane; i < N; i += 64) occ[i] = gocc[i];
  }
  uint16_t* gtrk = trk_pool + pp->trk_off;
  if constexpr (kCanonical)
    for (int i = lane; i < N; i += 64) trk[i] = gtrk[i];
  if constexpr (kRecLds)
    for (int i = lane; i < m.n_rec; i += 64) recs_l[i] = m.frecs[i];
  if constexpr (kGtabLds)
    for (int i = lane; i < m.gtab_size; i += 64) gtab_l[i] = m.gtab[i];
  const double* gtab = kGtabLds ? gtab_l : m.gtab;
  if (lane < 2 * kMaxGroups) grp_tab[lane] = lane < kMaxGroups ? m.grp_rec_off[lane] : m.grp_rec_cnt[lane - kMaxGroups];
  for (int i = lane; i < c.B * c.S; i += 64) phi_tab[i] = m.basis[i];
  for (int i = lane; i < c.S * c.S; i += 64) {
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
  FatRec rr[kR > 0 ? kR : 1];
  if constexpr (kR > 0) {
    const int cnt = m.grp_rec_cnt[0];
#pragma unroll
    for (int k = 0; k < kR; ++k) {
      const int it = lane + 64 * k;
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
    for (int i = lane; i < N; i += 64) {
      if (has_bkg && m.background[i]) continue;
      const int s = occ[ref_of<kPeriodic>(c, i).p];
#pragma unroll
      for (int d = 0; d < kBM; ++d)
        if (d < c.B) part_s[d] += phi_tab[d * c.S + s];
    }
#pragma unroll
    for (int d = 0; d < kBM; ++d) {
      const double tot = wave_sum_f64(part_s[d]);
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

  for (int64_t blk0 = 0; blk0 < nmov; blk0 += 64) {
    // lane l holds the Philox block and log u of move blk0 + l
    const Philox4 blk = philox4x64_10(t0 + uint64_t(blk0) + uint64_t(lane), 0, 0, 0, seed, uid);
    const double log_u = log(unit_double(blk.w3));
    const int n_in = int(nmov - blk0 < 64 ? nmov - blk0 : 64);
    uint64_t acc_bits = 0;
    for (int l = 0; l < n_in; ++l) {
      const uint64_t w0 = readlane_u64(blk.w0, l);
      const uint64_t w1 = readlane_u64(blk.w1, l);
      const double lu = readlane_f64(log_u, l);

      int si, sj = 0, sa = 0, sb = 0, off_a = 0, off_b = 0;
      uint32_t old_i = 0, new_i = 0;
      if constexpr (kCanonical) {
        const uint64_t w2 = readlane_u64(blk.w2, l);
        const uint32_t ia = bounded32(uint32_t(w0 >> 32), np);
        uint32_t ib = ia + 1u + bounded32(uint32_t(w0), np - 1u);
        ib -= ib >= np ? np : 0u;
        old_i = (present_pk >> (4u * ia)) & 15u;
        new_i = (present_pk >> (4u * ib)) & 15u;
        sa = int(bounded64(w1, uint64_t(unpack16(cnt_lo, cnt_hi, ia))));
        sb = int(bounded64(w2, uint64_t(unpack16(cnt_lo, cnt_hi, ib))));
        off_a = unpack16(seg_lo, seg_hi, ia);
        off_b = unpack16(seg_lo, seg_hi, ib);
        si = __builtin_amdgcn_readfirstlane(trk[off_a + sa]);
        sj = __builtin_amdgcn_readfirstlane(trk[off_b + sb]);
      } else {
        const int pick = int(bounded64(w1, uint64_t(n_pick)));
        si = has_bkg ? __builtin_amdgcn_readfirstlane(m.movable[pick]) : pick;
      }
      const Ref ri = ref_of<kPeriodic>(c, si);
      Ref rj = ri;
      uint32_t patch_j = 0u;
      if constexpr (kCanonical) {
        rj = ref_of<kPeriodic>(c, sj);
        patch_j = kPeriodic ? wrapped_delta(c, ri, rj) : uint32_t(si);
      }
      int gi_i[kR > 0 ? kR : 1], gi_j[kR > 0 ? kR : 1];
      if constexpr (!kCanonical) old_i = occ[ri.p];
      if constexpr (kR > 0) {
        // all member gathers of the move are issued together
        table_index<kPeriodic, kR, kBM, false>(c, rr, ri, occ, 0u, 0u, gi_i);
        if constexpr (kCanonical) table_index<kPeriodic, kR, kBM, true>(c, rr, rj, occ, patch_j, new_i, gi_j);
      }
      if constexpr (!kCanonical) {
        new_i = old_i + 1u + bounded32(uint32_t(w0), uint32_t(c.S - 1));
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
          dsing = dphi[0] * inv_n;
        else
          dsing = dphi_tab[(int(old_i) * c.S + int(new_i)) * c.B + (lane < c.B ? lane : 0)] * inv_n;
      }
      double dphj[kBM];
#pragma unroll
      for (int d = 0; d < kBM; ++d) dphj[d] = -dphi[d];
      double part;
      if constexpr (kR > 0) {
        part = table_sum<kR, kBM>(gi_i, gtab, dphi);
        if constexpr (kCanonical) {
          // second flip of the swap: j goes b -> a and sees site i already at b
          // (ce_updater.cpp:845-852 runs update_cf twice in sequence)
          const double pj = table_sum<kR, kBM>(gi_j, gtab, dphj);
          part = kBM == 1 ? part - pj : part + pj;
        }
      } else {
        const int gi = multi_group ? int(m.site_group[si]) : 0;
        part = lds_flip_de<kPeriodic, kBM, false>(c, recs_l, grp_tab[gi], grp_tab[kMaxGroups + gi], lane, ri, occ,
                                                  gtab, dphi, 0u, 0u);
        if constexpr (kCanonical) {
          const int gj = multi_group ? int(m.site_group[sj]) : 0;
          const double pj = lds_flip_de<kPeriodic, kBM, true>(c, recs_l, grp_tab[gj], grp_tab[kMaxGroups + gj], lane,
                                                              rj, occ, gtab, dphj, patch_j, new_i);
          part = kBM == 1 ? part - pj : part + pj;
        }
      }
      const double sum = wave_sum_f64(part);
      const double dE = (kBM == 1 ? dphi[0] * sum : sum) + de_single;
      // Metropolis test of the reference: accept if E_new < E_cur, else if
      // u <= exp(-(E_new - E_cur) / kT) (cemc/mcmc/montecarlo.py:950-956).  The
      // comparison is first made against log(u) with a margin far above the
      // rounding of log/exp; only moves inside the margin evaluate exp.
      bool accept;
      if (dE < 0.0) {
        accept = true;
      } else {
        const double x = -dE * inv_kT;
        const double margin = 1e-12 * (1.0 + fabs(x));
        if (lu < x - margin)
          accept = true;
        else if (lu > x + margin)
          accept = false;
        else
          accept = unit_double(readlane_u64(blk.w3, l)) <= exp(-dE / kT);
      }
      if (accept) {
        if constexpr (kPeriodic) {
          // lanes 0-7 write the images of site i, lanes 8-15 those of site j
          const bool second = kCanonical && lane >= 8 && lane < 16;
          const Ref& rs = second ? rj : ri;
          const int idx = lane < (kCanonical ? 16 : 8) ? image_index(c, rs.x, rs.y, rs.z, lane & 7) : -1;
          if (idx >= 0) occ[idx] = uint8_t(second ? old_i : new_i);
        } else {
          if (lane == 0) occ[si] = uint8_t(new_i);
          if (kCanonical && lane == 1) occ[sj] = uint8_t(old_i);
        }
        if constexpr (kCanonical) {
          if (lane == 0) {
            trk[off_a + sa] = uint16_t(sj);
            trk[off_b + sb] = uint16_t(si);
          }
        }
        energy += dE;
        if constexpr (!kCanonical) sing += dsing;
        acc_bits |= uint64_t(1) << l;
      }
      const double de_ref = energy - e_ref;
      sum_de += de_ref;
      sum_de2 = fma(de_ref, de_ref, sum_de2);
      if constexpr (!kCanonical) {
        sum_s += sing;
        sum_s2 = fma(sing, sing, sum_s2);
        sum_se = fma(sing, de_ref, sum_se);
      }
      __syncthreads();
    }
    n_acc += uint64_t(__builtin_popcountll(acc_bits));
    if (acc_log != nullptr && lane < 2) {
      const int64_t w = blk0 >> 5;
      if (lane == 0) acc_log[w] = uint32_t(acc_bits);
      if (lane == 1 && n_in > 32) acc_log[w + 1] = uint32_t(acc_bits >> 32);
    }
  }

  // ---- write back
  uint8_t* wocc = reinterpret_cast<uint8_t*>(occ_pool + pp->occ_off);
  for (int i = lane; i < N; i += 64) wocc[i] = occ[ref_of<kPeriodic>(c, i).p];
  if constexpr (kCanonical)
    for (int i = lane; i < N; i += 64) gtrk[i] = trk[i];
  ChainState* st = states + chain;
  if (lane == 0) {
    st->energy = energy;
    st->moves_done = t0 + uint64_t(nmov);
    st->n_accept += n_acc;
    st->n_samples += uint64_t(nmov);
    st->sum_de += sum_de;
    st->sum_de2 += sum_de2;
  }
  if (!kCanonical && lane < c.B) {
    st->sum_s[lane] += sum_s;
    st->sum_s2[lane] += sum_s2;
    st->sum_se[lane] += sum_se;
  }
}

// ---------------------------------------------------------------------------
// Correlation functions from scratch.  Block b of chain c sums, for each CF,
// the per-site contributions of sites [256 b, 256 b + 256) into
// partial[(c * n_cf + k) * n_blocks + b].
__global__ __launch_bounds__(256) void cf_partial_kernel(const ModelDev* __restrict__ models,
                                                         const ChainParams* __restrict__ params,
                                                         const int8_t* __restrict__ occ_pool,
                                                         const int32_t* __restrict__ chain_ids,
                                                         double* __restrict__ partial, int n_blocks) {
  __shared__ double red[4];
  const int c = int(blockIdx.y);
  const ChainParams* pp = params + chain_ids[c];
  const ModelDev& m = models[pp->model];
  const int8_t* occ = occ_pool + pp->occ_off;
  const int N = pp->n_sites;
  const int tid = int(threadIdx.x);
  const int i = int(blockIdx.x) * 256 + tid;
  const int g = (i < N) ? site_group(m, i) : -1;
  const bool movable = i < N && !(m.has_background && m.background[i]);
  const int S = m.n_species;
  const RefSite ref = make_ref(m, i < N ? i : 0);
  for (int k = 0; k < m.n_cf; ++k) {
    double v = 0.0;
    const int kind = m.cf_kind[k];
    if (kind == 1) {
      if (movable) v = m.basis[m.cf_dec[k] * S + occ[i]] / double(N);
    } else if (g >= 0) {
      if (kind == 2) {
        const int ent = m.cf_entry[g * m.n_cf + k];
        if (ent >= 0) {
          const ClusterEntry E = m.entries[ent];
          const int e0 = m.cf_eq_off[k], e1 = m.cf_eq_off[k + 1];
          double acc = 0.0;
          for (int mm = 0; mm < E.rec_count; ++mm) {
            const SubRec r = m.recs[E.rec_begin + mm];
            const int nm = rec_size(r) - 1;
            const int rp = rec_refpos(r);
            int spec[kMaxClusterSize];
            int q = 0;
            for (int pos = 0; pos <= nm; ++pos) {
              if (pos == rp) {
                spec[pos] = occ[i];
              } else {
                const int32_t nb = m.periodic ? neighbour<true>(m, ref, rec_slot(r, q), m.col_cells)
                                              : neighbour<false>(m, ref, rec_slot(r, q), m.col_cells);
                spec[pos] = occ[nb];
                ++q;
              }
            }
            for (int e = e0; e < e1; ++e) {
              const uint32_t deco = m.eq_decos[e];
              double prod = 1.0;
              for (int pos = 0; pos <= nm; ++pos) prod *= m.basis[int((deco >> (8 * pos)) & 0xFFu) * S + spec[pos]];
              acc += prod;
            }
          }
          v = acc / (double(e1 - e0) * double(E.norm) * double(m.grp_count[g]));
        }
      }
    }
    v = wave_sum_f64(v);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) partial[(int64_t(c) * m.n_cf + k) * n_blocks + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
  }
}

// Sums the block partials, sets c0 = 1 and computes E = N * sum_k J_k cf_k
// (ce_updater.cpp:236-242) with the chain's singlet ECIs.
__global__ __launch_bounds__(64) void cf_finalize_kernel(const ModelDev* __restrict__ models,
                                                         const ChainParams* __restrict__ params,
                                                         const int32_t* __restrict__ chain_ids,
                                                         const double* __restrict__ partial, int n_blocks,
                                                         double* __restrict__ cf_out,
                                                         double* __restrict__ energy_out) {
  const int c = int(blockIdx.x);
  const ChainParams* pp = params + chain_ids[c];
  const ModelDev& m = models[pp->model];
  double* cf = cf_out + int64_t(c) * m.n_cf;
  for (int k = int(threadIdx.x); k < m.n_cf; k += 64) {
    double s = 0.0;
    if (m.cf_kind[k] == 0) {
      s = 1.0;
    } else {
      const double* p = partial + (int64_t(c) * m.n_cf + k) * n_blocks;
      for (int b = 0; b < n_blocks; ++b) s += p[b];
    }
    cf[k] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double dot = 0.0;
    for (int k = 0; k < m.n_cf; ++k) {
      const double j = (m.cf_kind[k] == 1) ? pp->j1[m.cf_dec[k]] : m.eci[k];
      dot += j * cf[k];
    }
    energy_out[c] = dot * double(pp->n_sites);
  }
}

// ---------------------------------------------------------------------------
// Updater interface: apply `n` single-site changes in order.  Change c reads
// history slot src[c], writes slot dst[c] (CFHistoryTracker::get_next,
// cpp/src/cf_history_tracker.cpp:46-56) and sets the site's species.  After
// the last change the energy of the final slot is stored.
__global__ __launch_bounds__(256) void updater_kernel(const ModelDev* __restrict__ model,
                                                      int8_t* __restrict__ occ, double* __restrict__ ring,
                                                      double* __restrict__ slot_energy,
                                                      const int32_t* __restrict__ changes, int n,
                                                      const double* __restrict__ eci,
                                                      double* __restrict__ scratch) {
  const ModelDev& m = *model;
  const int tid = int(threadIdx.x);
  const int wave = tid >> 6;
  const int S = m.n_species;
  const int ncf = m.n_cf;
  int last_dst = -1;
  for (int c = 0; c < n; ++c) {
    const int site = changes[4 * c + 0];
    const int nw = changes[4 * c + 1];
    const int src = changes[4 * c + 2];
    const int dst = changes[4 * c + 3];
    last_dst = dst;
    const int old = occ[site];
    __syncthreads();
    if (tid == 0) occ[site] = int8_t(nw);
    __threadfence_block();
    __syncthreads();
    const double* cur = ring + int64_t(src) * ncf;
    double* nxt = ring + int64_t(dst) * ncf;
    const int g = site_group(m, site);
    const RefSite ref = make_ref(m, site);
    for (int k = 0; k < ncf; ++k) {
      const int kind = m.cf_kind[k];
      if (kind == 2) {
        const int ent = m.cf_entry[g * ncf + k];
        if (ent < 0) continue;
        const ClusterEntry E = m.entries[ent];
        const int e0 = m.cf_eq_off[k];
        const int ne = m.cf_eq_off[k + 1] - e0;
        const int total = ne * E.rec_count;
        double part = 0.0;
        for (int idx = tid; idx < total; idx += 256) {
          const int e = idx / E.rec_count;
          const int mm = idx - e * E.rec_count;
          const SubRec r = m.recs[E.rec_begin + mm];
          const int nm = rec_size(r) - 1;
          const int rp = rec_refpos(r);
          const uint32_t deco = m.eq_decos[e0 + e];
          double prod = 1.0;
          int q = 0;
          for (int pos = 0; pos <= nm; ++pos) {
            const int d = int((deco >> (8 * pos)) & 0xFFu);
            if (pos == rp) {
              prod *= m.basis[d * S + nw] - m.basis[d * S + old];
            } else {
              const int32_t nb = m.periodic ? neighbour<true>(m, ref, rec_slot(r, q), m.col_cells)
                                            : neighbour<false>(m, ref, rec_slot(r, q), m.col_cells);
              prod *= m.basis[d * S + occ[nb]];
              ++q;
            }
          }
          part += prod;
        }
        part = wave_sum_f64(part);
        if ((tid & 63) == 0) scratch[4 * k + wave] = part;
      }
    }
    __threadfence_block();
    __syncthreads();
    for (int k = tid; k < ncf; k += 256) {
      const int kind = m.cf_kind[k];
      double v = cur[k];
      if (kind == 1) {
        const int d = m.cf_dec[k];
        v = cur[k] + (m.basis[d * S + nw] - m.basis[d * S + old]) / double(unsigned(m.n_sites));
      } else if (kind == 2) {
        const int ent = m.cf_entry[g * ncf + k];
        if (ent >= 0) {
          const ClusterEntry E = m.entries[ent];
          const int ne = m.cf_eq_off[k + 1] - m.cf_eq_off[k];
          double delta = (scratch[4 * k] + scratch[4 * k + 1]) + (scratch[4 * k + 2] + scratch[4 * k + 3]);
          delta *= double(unsigned(E.size)) / double(ne);
          delta /= double(E.norm * m.grp_count[g]);
          v = cur[k] + delta;
        }
      }
      nxt[k] = v;
    }
    __threadfence_block();
    __syncthreads();
  }
  if (tid == 0 && last_dst >= 0) {
    const double* cf = ring + int64_t(last_dst) * ncf;
    double dot = 0.0;
    for (int k = 0; k < ncf; ++k) dot += eci[k] * cf[k];
    slot_energy[last_dst] = dot * double(unsigned(m.n_sites));
  }
}

__global__ void restore_kernel(int8_t* __restrict__ occ, const int32_t* __restrict__ sites,
                               const int32_t* __restrict__ species, int n) {
  // Applied in order by one thread: later pops restore earlier states.
  if (threadIdx.x == 0 && blockIdx.x == 0)
    for (int i = 0; i < n; ++i) occ[sites[i]] = int8_t(species[i]);
}

__global__ void slot_energy_kernel(const ModelDev* __restrict__ model, const double* __restrict__ ring,
                                   double* __restrict__ slot_energy, const double* __restrict__ eci, int slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const ModelDev& m = *model;
    const double* cf = ring + int64_t(slot) * m.n_cf;
    double dot = 0.0;
    for (int k = 0; k < m.n_cf; ++k) dot += eci[k] * cf[k];
    slot_energy[slot] = dot * double(unsigned(m.n_sites));
  }
}

// ---------------------------------------------------------------------------
// Launch wrappers.
// Instantiated variants: (periodic, canonical) x {records in VGPRs: kR = 2, 4,
// 8 with tables in LDS, up to 4 species} and {records in LDS: tables in LDS
// or HBM, up to 8 species}.
template <bool P, bool C>
struct SweepVariants {
  template <int KBM>
  static const void* pick(int kr, bool glds) {
    if (kr == 2) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 2, true, KBM>);
    if (kr == 4) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 4, true, KBM>);
    if (kr == 8) return reinterpret_cast<const void*>(&sweep_kernel<P, C, 8, true, KBM>);
    return glds ? reinterpret_cast<const void*>(&sweep_kernel<P, C, 0, true, KBM>)
                : reinterpret_cast<const void*>(&sweep_kernel<P, C, 0, false, KBM>);
  }
  static const void* fn(int kr, bool glds, int kbm) {
    if (kbm == 1) return pick<1>(kr, glds);
    if (kbm == 3) return pick<3>(kr, glds);
    return reinterpret_cast<const void*>(glds ? &sweep_kernel<P, C, 0, true, 7> : &sweep_kernel<P, C, 0, false, 7>);
  }
};

static const void* sweep_fn(bool periodic, bool canonical, int kr, bool glds, int kbm) {
  if (periodic) return canonical ? SweepVariants<true, true>::fn(kr, glds, kbm) : SweepVariants<true, false>::fn(kr, glds, kbm);
  return canonical ? SweepVariants<false, true>::fn(kr, glds, kbm) : SweepVariants<false, false>::fn(kr, glds, kbm);
}

bool sweep_variant_ok(int kr, bool glds, int kbm) {
  if (kr == 2 || kr == 4 || kr == 8) return glds && (kbm == 1 || kbm == 3);
  return kr == 0 && (kbm == 1 || kbm == 3 || kbm == 7);
}

uint32_t sweep_static_lds() {
  return uint32_t(sizeof(double) * (kMaxSpecies * kMaxSpecies * kMaxBasis + kMaxSpecies * kMaxSpecies +
                                    kMaxBasis * kMaxSpecies) +
                  sizeof(int) * 2 * kMaxGroups);
}

hipError_t sweep_set_max_lds(uint32_t bytes) {
  for (int p = 0; p < 2; ++p)
    for (int cn = 0; cn < 2; ++cn)
      for (int kr : {0, 2, 4, 8})
        for (int g = 0; g < 2; ++g)
          for (int kbm : {1, 3, 7}) {
            if (!sweep_variant_ok(kr, g != 0, kbm)) continue;
            const void* fn = sweep_fn(p != 0, cn != 0, kr, g != 0, kbm);
            hipFuncAttributes attr;
            hipError_t e = hipFuncGetAttributes(&attr, fn);
            if (e != hipSuccess) return e;
            const int dyn = int(bytes) - int(attr.sharedSizeBytes);
            if (dyn <= 0) return hipErrorInvalidValue;
            e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
            if (e != hipSuccess) return e;
          }
  return hipSuccess;
}

hipError_t launch_sweep(const LaunchSweep& a, hipStream_t stream) {
  if (!sweep_variant_ok(a.kr, a.gtab_lds, a.kbm)) return hipErrorInvalidValue;
  const void* fn = sweep_fn(a.periodic, a.canonical, a.kr, a.gtab_lds, a.kbm);
  const ChainParams* params = a.params;
  ChainState* states = a.states;
  const ModelDev* models = a.models;
  int8_t* occ_pool = a.occ_pool;
  uint16_t* trk_pool = a.trk_pool;
  const int32_t* order = a.chain_order;
  const int64_t* n_moves = a.n_moves;
  uint32_t* log = a.accept_log;
  int64_t log_words = a.log_words;
  void* args[] = {&params, &states, &models, &occ_pool, &trk_pool, &order, &n_moves, &log, &log_words};
  return hipLaunchKernel(fn, dim3(a.n_chains), dim3(64), args, a.lds_bytes, stream);
}

hipError_t launch_cf(const ModelDev* models, const ChainParams* params, const int8_t* occ_pool,
                     const int32_t* chain_ids, int n_chains, int max_sites, double* partial, double* cf_out,
                     double* energy_out, hipStream_t stream) {
  const int n_blocks = (max_sites + 255) / 256;
  hipLaunchKernelGGL(cf_partial_kernel, dim3(n_blocks, n_chains), dim3(256), 0, stream, models, params,
                     occ_pool, chain_ids, partial, n_blocks);
  hipLaunchKernelGGL(cf_finalize_kernel, dim3(n_chains), dim3(64), 0, stream, models, params, chain_ids,
                     partial, n_blocks, cf_out, energy_out);
  return hipGetLastError();
}

hipError_t launch_updater(const ModelDev* model, int8_t* occ, double* ring, double* slot_energy,
                          const int32_t* changes, int n, const double* eci, double* scratch, hipStream_t stream) {
  hipLaunchKernelGGL(updater_kernel, dim3(1), dim3(256), 0, stream, model, occ, ring, slot_energy, changes, n,
                     eci, scratch);
  return hipGetLastError();
}

hipError_t launch_restore(int8_t* occ, const int32_t* sites, const int32_t* species, int n, hipStream_t stream) {
  hipLaunchKernelGGL(restore_kernel, dim3(1), dim3(64), 0, stream, occ, sites, species, n);
  return hipGetLastError();
}

hipError_t launch_slot_energy(const ModelDev* model, const double* ring, double* slot_energy, const double* eci,
                              int slot, hipStream_t stream) {
  hipLaunchKernelGGL(slot_energy_kernel, dim3(1), dim3(64), 0, stream, model, ring, slot_energy, eci, slot);
  return hipGetLastError();
}

}  // namespace cemc
