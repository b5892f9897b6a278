// Wang-Landau walkers on canonical cluster-expansion chains (gfx950).  The
// instantiations live in ce_wl_inst.hip, one object per lattice path.
//
// One 64-lane wavefront owns one walker.  The move evaluation is the sweep
// kernel's one-move-at-a-time path (records in VGPRs or LDS, energy tables
// in LDS or HBM, padded occupation box on closed-form periodic lattices);
// the Metropolis test is replaced by the Wang-Landau rule of
// WangLandauSampler::step (wang_landau_sampler.cpp:232-317) on the walker's
// ln g and visit counts, which stay in LDS for the whole launch.  After
// every `check_every` moves the wave runs Histogram::is_flat
// (histogram.cpp:102-128); a flat histogram lowers f (halving, or the
// inverse-time law once a flat histogram took longer than the previous one,
// :344-367) and clears the counts, and the walker stops once f < f_min
// (:337-394).  With adaptive windows (adaptive_windows.cpp) the flatness
// test scans down from the top of the window and cuts converged upper parts
// off; the walker's lattice is recorded in HBM at the first visit of every
// bin and moved to a recorded state after a cut, and a flat window joins
// the pieces of ln g back together.  This bookkeeping runs once per check
// on values the whole wave holds, so it costs nothing per move.
#pragma once

#include "ce_sweep.h"

namespace cemc {

namespace {

// ln g (doubles) | visit counts (u32) | known flags (u8) | with adaptive
// windows: sub-bin counts (u32, 10 per bin) | recorded flags (u8); after the
// sweep layout of the chain.
struct WLLayout {
  uint32_t logdos, hist, known, sub, saved, total;
};

__host__ __device__ inline WLLayout wl_layout(uint32_t base, int n_bins, bool adaptive) {
  WLLayout L;
  L.logdos = align16(base);
  L.hist = L.logdos + align16(uint32_t(n_bins) * 8u);
  L.known = L.hist + align16(uint32_t(n_bins) * 4u);
  L.sub = L.known + align16(uint32_t(n_bins));
  L.saved = L.sub + (adaptive ? align16(uint32_t(n_bins) * 40u) : 0u);
  L.total = L.saved + (adaptive ? align16(uint32_t(n_bins)) : 0u);
  return L;
}

// Bin of a total energy from its truncated value t (wl_raw_bin):
// histogram.cpp:66-69 converts (e - e_min) * n_bins / (e_max - e_min) to int
// (truncation toward zero) and :137-140 keeps 0 <= bin < n_bins.  -1 below
// the range, n_bins when the value truncates to n_bins exactly, n_bins + 1
// above (the caller tells the last two apart, wang_landau_sampler.cpp:257-258).
__device__ __attribute__((always_inline)) inline int wl_bin(double t, double nbd, int n_bins) {
  return t < 0.0 ? -1 : (t > nbd ? n_bins + 1 : int(t));
}

// The truncated (e - e_min) * n_bins / (e_max - e_min) before any range
// check (the int Histogram::get_bin returns), limited to the int range.
__device__ __attribute__((always_inline)) inline double wl_raw_bin(double e, double e_min, double span, double nbd) {
  return fmin(fmax(trunc((e - e_min) * nbd / span), -2147483648.0), 2147483647.0);
}

__device__ __attribute__((always_inline)) inline uint32_t wave_add_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += uint32_t(__shfl_xor(int(v), o, 64));
  return v;
}

__device__ __attribute__((always_inline)) inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = min(v, uint32_t(__shfl_xor(int(v), o, 64)));
  return v;
}

}  // namespace

// kMode 0: own histogram, fixed f or the inverse-time law (the adaptive
// window code is compiled out); 1: own histogram with adaptive windows
// possible; 2 (shared): every walker of the launch visits the histogram of
// `shared` (the group's leader) in global memory, as the reference's OpenMP
// walkers share one (wang_landau_sampler.cpp:288-298): ln g and counts grow
// by atomic additions and ln g is read at device scope, so walkers see each
// other's visits as they land.  A shared launch makes one check's steps and
// leaves the flatness test, f and convergence to wl_shared_check_kernel.
// kSGC: semi-grand-canonical walkers (single-site flips, no tracker); the
// host launches canonical and semi-grand-canonical walkers apart.
template <bool kPeriodic, int kR, bool kGtabLds, int kBM, int kNM = 3, int kMode = 0, bool kSGC = false>
__global__ __launch_bounds__(64) void wl_kernel(const ChainParams* __restrict__ params,
                                                ChainState* __restrict__ states,
                                                const ModelDev* __restrict__ models,
                                                int8_t* __restrict__ occ_pool, uint16_t* __restrict__ trk_pool,
                                                const int32_t* __restrict__ chain_order, WLDev* __restrict__ walkers,
                                                double* __restrict__ logdos_pool, uint32_t* __restrict__ hist_pool,
                                                uint8_t* __restrict__ known_pool, int64_t n_checks,
                                                int64_t check_every, uint32_t* __restrict__ accept_log,
                                                int64_t log_words, const WLDev* __restrict__ shared,
                                                int64_t log_base) {
  constexpr int kILW = il_width<kBM, kGtabLds>();
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ double dphi_tab[kMaxSpecies * kMaxSpecies * kMaxBasis];  // [(old*S+new)*B + d]
  __shared__ double dsingle_tab[kMaxSpecies * kMaxSpecies];          // SGC: singlet energy change per (old,new)
  __shared__ int grp_tab[2 * kMaxGroups];
  constexpr bool kRecLds = kR == 0;
  constexpr bool kShared = kMode == 2;
  const int chain = chain_order[blockIdx.x];
  const int lane = int(threadIdx.x);
  const ChainParams* __restrict__ pp = params + chain;
  const ModelDev& m = models[pp->model];
  WLDev* __restrict__ wk = walkers + chain;
  const int N = pp->n_sites;
  const int occ_bytes = kPeriodic ? m.pdims[0] * m.pdims[1] * m.pdims[2] : N;
  const LdsLayout L = lds_layout(occ_bytes, N, kRecLds ? m.n_rec : 0, kGtabLds ? m.gtab_size : 0);
  const int nb = wk->n_bins;
  const bool adaptive = kMode == 1 && wk->adaptive != 0;
  const WLLayout WL = wl_layout(L.total, kShared ? 0 : nb, adaptive);
  uint8_t* occ = reinterpret_cast<uint8_t*>(lds + L.occ);
  uint16_t* trk = reinterpret_cast<uint16_t*>(lds + L.trk);
  FatRec* recs_l = reinterpret_cast<FatRec*>(lds + L.rec);
  double* gtab_l = reinterpret_cast<double*>(lds + L.gtab);
  const int64_t boff = kShared ? shared->bin_off : wk->bin_off;
  double* lg = kShared ? logdos_pool + boff : reinterpret_cast<double*>(lds + WL.logdos);
  uint32_t* hist = kShared ? hist_pool + boff : reinterpret_cast<uint32_t*>(lds + WL.hist);
  uint8_t* known = kShared ? known_pool + boff : reinterpret_cast<uint8_t*>(lds + WL.known);
  // ln g of a bin: shared bins are read at device scope, past this CU's cache
  auto lg_at = [&](int b) {
    if constexpr (kShared)
      return __hip_atomic_load(lg + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      return lg[b];
  };
  uint32_t* sub = reinterpret_cast<uint32_t*>(lds + WL.sub);
  uint8_t* saved = reinterpret_cast<uint8_t*>(lds + WL.saved);

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

  // A lattice in site order (global memory) into the LDS occupation: on
  // closed-form periodic lattices every padded position takes the site it
  // wraps to.
  auto load_lattice = [&](const uint8_t* src) {
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
        occ[p] = src[(x * c.L2 + y) * c.L3 + z];
      }
    } else {
      for (int i = lane; i < N; i += 64) occ[i] = src[i];
    }
  };

  // ---- stage the chain, the model tables and the walker's bins into LDS
  const uint8_t* gocc = reinterpret_cast<const uint8_t*>(occ_pool + pp->occ_off);
  load_lattice(gocc);
  uint16_t* gtrk = trk_pool + pp->trk_off;
  // semi-grand-canonical walkers flip one site at a time and have no tracker
  constexpr bool sgc = kSGC;
  const bool has_bkg = m.has_background != 0;
  const int n_pick = has_bkg ? m.n_movable : N;
  if (!sgc)
    for (int i = lane; i < N; i += 64) trk[i] = gtrk[i];
  if constexpr (kRecLds)
    for (int i = lane; i < m.n_rec; i += 64) recs_l[i] = m.frecs[i];
  if constexpr (kGtabLds)
    for (int i = lane; i < m.gtab_size; i += 64) gtab_l[i] = (kILW ? m.gtab_il : m.gtab)[i];
  const auto gtab = table_ptr<kGtabLds>(gtab_l, kILW ? m.gtab_il : m.gtab);
  if (lane < 2 * kMaxGroups) grp_tab[lane] = lane < kMaxGroups ? m.grp_rec_off[lane] : m.grp_rec_cnt[lane - kMaxGroups];
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
  if constexpr (!kShared) {
    for (int i = lane; i < nb; i += 64) {
      lg[i] = logdos_pool[boff + i];
      hist[i] = hist_pool[boff + i];
      known[i] = known_pool[boff + i];
    }
  }
  if (adaptive) {
    for (int i = lane; i < nb; i += 64) saved[i] = wk->saved[i];
    for (int i = lane; i < 10 * nb; i += 64) sub[i] = wk->sub[i];
  }
  __syncthreads();

  FatRec rr[kR > 0 ? kR : 1];
  if constexpr (kR > 0) {
    const int cnt = m.grp_rec_cnt[0];
#pragma unroll
    for (int k = 0; k < kR; ++k) {
      const int it = lane + 64 * k;
      rr[k] = it < cnt ? m.frecs[it] : FatRec{0, 0, 0, 0u, 0u, 0u, 0u, 0u};
    }
  }
  const double dphi01 = dphi_tab[1], dphi10 = dphi_tab[c.S];

  double energy = states[chain].energy;
  const uint64_t t0 = states[chain].moves_done;
  const uint64_t seed = pp->seed, uid = pp->uid;
  const uint32_t np = uint32_t(pp->n_present);
  const uint32_t present_pk = pp->present_pk;
  const uint64_t cnt_lo = pp->count_pk[0], cnt_hi = pp->count_pk[1];
  const uint64_t seg_lo = pp->segoff_pk[0], seg_hi = pp->segoff_pk[1];
  uint32_t* acc_log = accept_log ? accept_log + int64_t(chain) * log_words : nullptr;

  const double e_min = wk->e_min, span = wk->e_max - wk->e_min, nbd = double(nb);
  const double flat_c = wk->flatness, f_min = wk->f_min;
  const bool inverse_time = wk->inverse_time != 0;
  const int min_width = wk->min_width;
  double f = kShared ? shared->f : wk->f;
  int cur = wk->cur_bin, first = wk->is_first, status = wk->status;
  int converged = kShared ? shared->converged : wk->converged;
  int upper = adaptive ? wk->upper : nb;
  int n_overlap = wk->n_overlap, n_edges = wk->n_edges, has_first = wk->has_first, inv_active = wk->inv_active;
  double inv_factor = wk->inv_factor, bin_change = wk->bin_change, cur_e = wk->cur_energy;
  int64_t steps = wk->steps, since_flat = wk->since_flat, prev_conv = wk->prev_conv;
  uint64_t n_acc = 0, n_out = 0, n_self = 0, n_flat = 0, n_walk = 0, n_steps = 0;
  int64_t done = 0;  // Philox counters used in this launch (steps and walk moves)

  // lower edge of a bin, Histogram::get_energy (histogram.cpp:61-64)
  auto bin_energy = [&](int b) { return e_min + (span * double(b)) / nbd; };

  // The walker's state into record `slot` (adaptive_windows.cpp:240-260,
  // 196-214): lattice in site order, tracker, energy, bin.
  auto snapshot = [&](int slot) {
    uint8_t* so = wk->snap_occ + int64_t(slot) * N;
    uint16_t* st = wk->snap_trk + int64_t(slot) * N;
    for (int i = lane; i < N; i += 64) {
      so[i] = occ[ref_of<kPeriodic>(c, i).p];
      st[i] = trk[i];
    }
    if (lane == 0) {
      wk->snap_e[slot] = energy;
      wk->snap_bin[slot] = cur;
    }
    __threadfence_block();
  };
  // WangLandauSampler::set_updaters (wang_landau_sampler.cpp:585-600): the
  // recorded state replaces the walker's; the sub-bin energy stays.
  auto restore = [&](int slot) {
    __threadfence_block();
    load_lattice(wk->snap_occ + int64_t(slot) * N);
    const uint16_t* st = wk->snap_trk + int64_t(slot) * N;
    for (int i = lane; i < N; i += 64) trk[i] = st[i];
    energy = wk->snap_e[slot];
    cur = wk->snap_bin[slot];
    __syncthreads();
  };

  // Histogram::update (histogram.cpp:71-90) by lane 0, after the state is
  // recorded at the bin's first visit (adaptive); `with_sub` adds the
  // sub-bin count of the step's closing update (histogram.cpp:92-100).  A
  // wave's LDS operations complete in order, so the next move's reads see it.
  auto visit = [&](int b, bool with_sub) {
    if (adaptive && !saved[b]) {
      snapshot(b);
      if (lane == 0) saved[b] = 1;
    }
    if constexpr (kShared) {
      if (lane == 0) {
        known[b] = 1;
        __hip_atomic_fetch_add(hist + b, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(lg + b, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else if (lane == 0) {
      known[b] = 1;
      hist[b] += 1u;
      lg[b] += f;
      if (with_sub && adaptive) {
        const double lo = bin_energy(b), hi = bin_energy(b + 1);
        const double t = trunc((cur_e - lo) * 10.0 / (hi - lo));
        if (t >= 0.0 && t < 10.0) sub[10 * b + int(t)] += 1u;
      }
    }
  };

  // One canonical swap of Philox block `l` of the current 64-block: its
  // energy change is summed over the wave; commit() applies it.
  struct Move {
    uint32_t old_i, new_i;
    int ta, tb, si, sj;
    double e_new;
  };
  auto evaluate = [&](uint32_t old_i, uint32_t new_i, int ta, int tb) {
    Move mv;
    mv.old_i = old_i;
    mv.new_i = new_i;
    mv.ta = ta;
    mv.tb = tb;
    mv.si = __builtin_amdgcn_readfirstlane(trk[ta]);
    mv.sj = __builtin_amdgcn_readfirstlane(trk[tb]);
    const Ref ri = ref_of<kPeriodic>(c, mv.si);
    const Ref rj = ref_of<kPeriodic>(c, mv.sj);
    const uint32_t patch_j = kPeriodic ? wrapped_delta(c, ri, rj) : uint32_t(mv.si);
    double dphi[kBM];
    if constexpr (kBM == 1) {
      dphi[0] = old_i == 0 ? dphi01 : dphi10;
    } else {
      const int pair_i = int(old_i) * c.S + int(new_i);
#pragma unroll
      for (int d = 0; d < kBM; ++d) dphi[d] = d < c.B ? dphi_tab[pair_i * c.B + d] : 0.0;
    }
    double dphj[kBM];
#pragma unroll
    for (int d = 0; d < kBM; ++d) dphj[d] = -dphi[d];
    double part;
    if constexpr (kR > 0) {
      int gi_i[kR], gi_j[kR];
      table_index<kPeriodic, kR, kBM, false, kNM, kILW>(c, rr, ri, occ, 0u, 0u, gi_i);
      table_index<kPeriodic, kR, kBM, true, kNM, kILW>(c, rr, rj, occ, patch_j, new_i, gi_j);
      part = table_sum<kR, kBM, kILW>(gi_i, gtab, dphi);
      // the swap's second flip sees site i already changed
      // (ce_updater.cpp:845-852)
      const double pj = table_sum<kR, kBM, kILW>(gi_j, gtab, dphj);
      part = kBM == 1 ? part - pj : part + pj;
    } else {
      const int gi = multi_group ? int(m.site_group[mv.si]) : 0;
      const int gj = multi_group ? int(m.site_group[mv.sj]) : 0;
      part = lds_flip_de<kPeriodic, kBM, false, kILW>(c, recs_l, grp_tab[gi], grp_tab[kMaxGroups + gi], lane, 64, ri,
                                                occ, gtab, dphi, 0u, 0u);
      const double pj = lds_flip_de<kPeriodic, kBM, true, kILW>(c, recs_l, grp_tab[gj], grp_tab[kMaxGroups + gj],
                                                          lane, 64, rj, occ, gtab, dphj, patch_j, new_i);
      part = kBM == 1 ? part - pj : part + pj;
    }
    const double sum = wave_sum_f64(part);
    const double dE = kBM == 1 ? __dmul_rn(dphi[0], sum) : sum;
    mv.e_new = energy + dE;
    return mv;
  };
  // A semi-grand-canonical flip of `site` to the species `off` + 1 after its
  // own (the SGC move of the sweep kernel); the singlet ECIs carry the
  // chemical potentials.
  auto evaluate_flip = [&](int site, uint32_t off) {
    Move mv;
    const Ref ri = ref_of<kPeriodic>(c, site);
    mv.si = mv.sj = site;
    mv.ta = mv.tb = -1;
    mv.old_i = uint32_t(__builtin_amdgcn_readfirstlane(int(occ[ri.p])));
    mv.new_i = mv.old_i + 1u + off;
    mv.new_i -= mv.new_i >= uint32_t(c.S) ? uint32_t(c.S) : 0u;
    double dphi[kBM];
    double de_single;
    if constexpr (kBM == 1) {
      dphi[0] = mv.old_i == 0 ? dphi01 : dphi10;
      de_single = dsingle_tab[mv.old_i == 0 ? 1 : c.S];
    } else {
      const int pair_i = int(mv.old_i) * c.S + int(mv.new_i);
#pragma unroll
      for (int d = 0; d < kBM; ++d) dphi[d] = d < c.B ? dphi_tab[pair_i * c.B + d] : 0.0;
      de_single = dsingle_tab[pair_i];
    }
    double part;
    if constexpr (kR > 0) {
      int gi_i[kR];
      table_index<kPeriodic, kR, kBM, false, kNM, kILW>(c, rr, ri, occ, 0u, 0u, gi_i);
      part = table_sum<kR, kBM, kILW>(gi_i, gtab, dphi);
    } else {
      const int gi = multi_group ? int(m.site_group[site]) : 0;
      part = lds_flip_de<kPeriodic, kBM, false, kILW>(c, recs_l, grp_tab[gi], grp_tab[kMaxGroups + gi], lane, 64,
                                                          ri, occ, gtab, dphi, 0u, 0u);
    }
    const double sum = wave_sum_f64(part);
    const double dE = kBM == 1 ? fma(dphi[0], sum, de_single) : sum + de_single;
    mv.e_new = energy + dE;
    return mv;
  };
  auto commit = [&](const Move& mv) {
    const int n_sites_changed = sgc ? 1 : 2;
    if constexpr (kPeriodic) {
      // lanes 0-7 write the images of site i, lanes 8-15 those of site j
      const bool second = lane >= 8;
      const Ref r = ref_of<kPeriodic>(c, second ? mv.sj : mv.si);
      const uint32_t img = image_word<true>(c, r);
      const int idx = image_of(c, img, lane & 7);
      if (idx >= 0 && lane < 8 * n_sites_changed) occ[idx] = uint8_t(second ? mv.old_i : mv.new_i);
    } else {
      if (lane == 0) occ[mv.si] = uint8_t(mv.new_i);
      if (lane == 1 && !sgc) occ[mv.sj] = uint8_t(mv.old_i);
    }
    if (lane == 0 && !sgc) {
      trk[mv.ta] = uint16_t(mv.sj);
      trk[mv.tb] = uint16_t(mv.si);
    }
    energy = mv.e_new;
  };

  // Philox block of counter t0 + done + lane and the parts of that move
  // that do not depend on the state.
  struct Draw {
    Philox4 blk;
    uint32_t v_old, v_new;
    int v_ta, v_tb;
  };
  auto draw = [&]() {
    Draw d;
    d.blk = philox4x64_10(t0 + uint64_t(done) + uint64_t(lane), 0, 0, 0, seed, uid);
    if (sgc) {
      // the site (a movable one) and the offset of its new species
      const int pick = int(bounded64(d.blk.w1, uint64_t(n_pick)));
      d.v_ta = has_bkg ? m.movable[pick] : pick;
      d.v_tb = 0;
      d.v_old = 0u;
      d.v_new = bounded32(uint32_t(d.blk.w0), uint32_t(c.S - 1));
      return d;
    }
    const uint32_t ia = bounded32(uint32_t(d.blk.w0 >> 32), np);
    uint32_t ib = ia + 1u + bounded32(uint32_t(d.blk.w0), np - 1u);
    ib -= ib >= np ? np : 0u;
    d.v_old = (present_pk >> (4u * ia)) & 15u;
    d.v_new = (present_pk >> (4u * ib)) & 15u;
    d.v_ta = unpack16(seg_lo, seg_hi, ia) + int(bounded64(d.blk.w1, uint64_t(unpack16(cnt_lo, cnt_hi, ia))));
    d.v_tb = unpack16(seg_lo, seg_hi, ib) + int(bounded64(d.blk.w2, uint64_t(unpack16(cnt_lo, cnt_hi, ib))));
    return d;
  };
  auto eval_lane = [&](const Draw& d, int l) {
    if (sgc) return evaluate_flip(__builtin_amdgcn_readlane(d.v_ta, l), __builtin_amdgcn_readlane(d.v_new, l));
    return evaluate(__builtin_amdgcn_readlane(d.v_old, l), __builtin_amdgcn_readlane(d.v_new, l),
                    __builtin_amdgcn_readlane(d.v_ta, l), __builtin_amdgcn_readlane(d.v_tb, l));
  };

  // WangLandauSampler::run_until_valid_energy (wang_landau_sampler.cpp:
  // 434-560) for a walker outside its window: swaps toward the window
  // (below emax: any swap to a higher energy under emax; above it: any swap
  // to a lower energy over emin, measured from the current bin's lower
  // edge) until a proposal lands in the window, at most 10^7 swaps.
  auto walk_into_window = [&]() {
    if (cur >= 0 && cur < upper) return;
    const double emax = bin_energy(upper);
    int64_t tried = 0;
    bool found = false;
    while (!found && tried < 10000000) {
      const Draw d = draw();
      const int n_in = int(10000000 - tried < 64 ? 10000000 - tried : 64);
      int l = 0;
      for (; l < n_in; ++l) {
        const Move mv = eval_lane(d, l);
        const double raw = wl_raw_bin(mv.e_new, e_min, span, nbd);
        const double ce = bin_energy(cur);
        const int ok = __builtin_amdgcn_readfirstlane(
            int(ce >= emax ? (mv.e_new < ce && mv.e_new > e_min) : (mv.e_new > ce && mv.e_new < emax)));
        if (ok) {
          commit(mv);
          cur = __builtin_amdgcn_readfirstlane(int(raw));
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (__builtin_amdgcn_readfirstlane(int(raw >= 0.0 && raw < double(upper)))) {
          found = true;
          ++l;
          break;
        }
      }
      done += l;
      tried += l;
      n_walk += uint64_t(l);
    }
    if (!found) status = 1;
  };

  // AdaptiveWindowHistogram::is_flat (adaptive_windows.cpp:38-140), every
  // lane on the same LDS values: scanning down from upper - n_overlap, the
  // known bins above the scan point stay flat while their minimum exceeds
  // criteria times their mean; the scan stops at the first bin that breaks
  // it.  Reaching bin 0 flat means the window converged (returns true).
  // Otherwise a flat part of more than min_width bins (or of every known
  // bin) is cut off, its top n_overlap values of ln g saved, and the walker
  // moves into the window below.  With n_overlap >= upper the scan is empty
  // (the reference's unsigned start would wrap).
  auto window_flat = [&]() -> bool {
    const int top = upper > n_overlap ? upper - n_overlap : 0;
    uint32_t max_possible = 0u;
    for (int i = 0; i < top; ++i) max_possible += known[i] ? 1u : 0u;
    uint32_t sum = 0u, count = 0u, minimum = 100000000u;
    bool part = false, min_ok = false;
    int fnc = 0;
    for (int i = top; i > 0; --i) {
      if (!known[i - 1]) continue;
      const uint32_t h = hist[i - 1];
      sum += h;
      count += 1u;
      minimum = min(minimum, h);
      min_ok = double(minimum) > flat_c * (double(sum) / double(count));
      if ((count > uint32_t(min_width) || count >= max_possible) && min_ok) part = true;
      if (!min_ok) {
        fnc = i - 1;
        break;
      }
    }
    if (fnc == 0 && min_ok) return true;
    if (!part) return false;
    // is_valid_window (adaptive_windows.cpp:222-236)
    if (fnc <= min_width) return false;
    int n_known = 0;
    for (int i = 0; i < fnc; ++i) n_known += known[i] ? 1 : 0;
    if (n_known <= min_width / 2) return false;
    if (lane == 0) {
      wk->edges[n_edges] = fnc + 1;
      wk->edge_n[n_edges] = n_overlap;
      for (int k = 0; k < n_overlap; ++k) wk->edge_vals[int64_t(n_edges) * nb + k] = lg[fnc + 1 + k];
    }
    ++n_edges;
    if (upper == nb) {
      snapshot(nb);
      has_first = 1;
    }
    upper = fnc + n_overlap + 1;
    // distribute_random_walkers_evenly (adaptive_windows.cpp:262-318) with
    // one walker: to the recorded state nearest the window's middle
    // (the upper side on a tie) once more than one bin has one
    int n_rec = 0;
    for (int i = 0; i < upper; ++i) n_rec += saved[i] ? 1 : 0;
    if (n_rec > 1) {
      const int indx = upper / 2;
      int shift_pos = 2 * upper, shift_neg = 2 * upper;
      for (int j = indx; j < upper; ++j)
        if (saved[j]) {
          shift_pos = j - indx;
          break;
        }
      for (int j = indx; j >= 0; --j)
        if (saved[j]) {
          shift_neg = indx - j;
          break;
        }
      restore(shift_pos <= shift_neg ? indx + shift_pos : indx - shift_neg);
    }
    walk_into_window();
    return false;
  };

  // AdaptiveWindowHistogram::reset and make_dos_continous
  // (adaptive_windows.cpp:142-185): the full range returns; at every cut, in
  // the order they were made, the saved values go back above it and ln g at
  // and below the cut shifts to meet them (later cuts' saved values shift
  // along); the walker returns to the state recorded at the first cut.
  auto window_reset = [&]() {
    upper = nb;
    if (lane == 0) {
      for (int e = 0; e < n_edges; ++e) {
        const int E = wk->edges[e], n = wk->edge_n[e];
        const double* v = wk->edge_vals + int64_t(e) * nb;
        for (int j = 1; j < n; ++j) lg[E + j] = v[j];
        const double diff = lg[E] - v[0];
        for (int j = 0; j < E + 1; ++j) lg[j] -= diff;
        for (int e2 = e + 1; e2 < n_edges; ++e2)
          for (int k = 0; k < wk->edge_n[e2]; ++k) wk->edge_vals[int64_t(e2) * nb + k] -= diff;
      }
    }
    n_edges = 0;
    __syncthreads();
    if (has_first) restore(nb);
  };

  for (int64_t chk = 0; chk < n_checks && !converged && !status; ++chk) {
    for (int64_t blk0 = 0; blk0 < check_every; blk0 += 64) {
      // lane l holds the Philox block and log u of move done + l
      const Draw d = draw();
      const double log_u = log(unit_double(d.blk.w3));
      const int n_in = int(check_every - blk0 < 64 ? check_every - blk0 : 64);
      for (int l = 0; l < n_in; ++l) {
        const Move mv = eval_lane(d, l);
        const double raw = wl_raw_bin(mv.e_new, e_min, span, nbd);
        if (adaptive) bin_change += fabs(raw - double(cur));  // only set_overlap reads it
        const int bin = __builtin_amdgcn_readfirstlane(wl_bin(raw, nbd, nb));
        if (bin < 0 || bin >= upper) {
          // outside the range: the proposal is undone; a walker that has
          // been in range counts its current bin once more when the
          // proposal lies above the top bin (wang_landau_sampler.cpp:251-260)
          ++n_out;
          if (!first && bin > nb) visit(cur, false);
        } else {
          n_self += bin == cur ? 1u : 0u;
          // accept if u < g(cur) / g(bin) = exp(min(0, ln g(cur) - ln g(bin)))
          // (histogram.cpp:130-135), or on the walker's first in-range
          // proposal.  Compared against log u with a margin far above the
          // rounding of log/exp; only draws inside it evaluate exp.
          const double diff = lg_at(cur) - lg_at(bin);
          const double margin = 1e-12 * (1.0 + fabs(diff));
          const double lu = readlane_f64(log_u, l);
          int verdict = (first || diff > 0.0 || lu < diff - margin) ? 1 : (lu > diff + margin ? 0 : 2);
          verdict = __builtin_amdgcn_readfirstlane(verdict);
          if (verdict == 2)
            verdict = __builtin_amdgcn_readfirstlane(int(unit_double(readlane_u64(d.blk.w3, l)) < exp(diff)));
          if (verdict) {
            commit(mv);
            cur = bin;
            first = 0;
            if (adaptive) cur_e = mv.e_new;  // the sub-bin energy
            ++n_acc;
            const int64_t q = log_base + done + l;
            if (acc_log != nullptr && lane == 0 && q < 32 * log_words) atomicOr(acc_log + (q >> 5), 1u << (q & 31));
          }
          visit(cur, true);
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
      }
      done += n_in;
      n_steps += uint64_t(n_in);
    }
    steps += check_every;  // a check always makes all its steps
    since_flat += check_every;
    __syncthreads();
    if constexpr (kShared) continue;  // wl_shared_check_kernel tests the shared histogram
    bool flat;
    if (adaptive) {
      flat = window_flat();
      if (status) break;
    } else {
      // Histogram::is_flat over the known bins; the sum is an unsigned
      // 32-bit integer and the minimum starts at 100000000
      uint32_t sum = 0u, cnt = 0u, mn = 100000000u;
      for (int i = lane; i < nb; i += 64) {
        if (known[i]) {
          const uint32_t h = hist[i];
          sum += h;
          cnt += 1u;
          mn = min(mn, h);
        }
      }
      sum = wave_add_u32(sum);
      cnt = wave_add_u32(cnt);
      mn = wave_min_u32(mn);
      const double mean = double(sum) / double(cnt);
      flat = __builtin_amdgcn_readfirstlane(int(double(mn) > flat_c * mean)) != 0;
    }
    if (flat) {
      // wang_landau_sampler.cpp:337-394, convergence time in steps
      const int64_t took = since_flat;
      const double mc_time = double(steps) / nbd;
      if (inverse_time && prev_conv >= 0 && took > prev_conv) {
        if (!inv_active) inv_factor = mc_time * f;
        inv_active = 1;
      }
      prev_conv = took;
      f = inv_active ? inv_factor / mc_time : f / 2.0;
      ++n_flat;
      for (int i = lane; i < nb; i += 64) hist[i] = 0u;
      __syncthreads();
      if (adaptive) {
        window_reset();
        // set_overlap(2 * average bin change per step)
        const double ov = 2.0 * (bin_change / double(took));
        n_overlap = ov >= nbd ? nb : int(uint32_t(ov));
      }
      bin_change = 0.0;
      since_flat = 0;
    }
    if (f < f_min) converged = 1;
    __syncthreads();
  }

  // ---- write back
  __syncthreads();
  uint8_t* wocc = reinterpret_cast<uint8_t*>(occ_pool + pp->occ_off);
  for (int i = lane; i < N; i += 64) wocc[i] = occ[ref_of<kPeriodic>(c, i).p];
  if (!sgc)
    for (int i = lane; i < N; i += 64) gtrk[i] = trk[i];
  if constexpr (!kShared) {
    for (int i = lane; i < nb; i += 64) {
      logdos_pool[boff + i] = lg[i];
      hist_pool[boff + i] = hist[i];
      known_pool[boff + i] = known[i];
    }
  }
  if (adaptive) {
    for (int i = lane; i < nb; i += 64) wk->saved[i] = saved[i];
    for (int i = lane; i < 10 * nb; i += 64) wk->sub[i] = sub[i];
  }
  if (lane == 0) {
    ChainState* st = states + chain;
    st->energy = energy;
    st->moves_done = t0 + uint64_t(done);
    st->n_accept += n_acc;
    wk->cur_bin = cur;
    wk->is_first = first;
    wk->n_moves += n_steps;
    wk->n_accept += n_acc;
    wk->n_outside += n_out;
    wk->n_self += n_self;
    wk->cur_energy = cur_e;
    if (!kShared) {  // a shared histogram's state belongs to wl_shared_check_kernel
      wk->f = f;
      wk->converged = converged;
      wk->status = status;
      wk->n_flat += n_flat;
      wk->n_walk += n_walk;
      wk->inv_active = inv_active;
      wk->inv_factor = inv_factor;
      wk->bin_change = bin_change;
      wk->steps = steps;
      wk->since_flat = since_flat;
      wk->prev_conv = prev_conv;
    }
    if (adaptive) {
      wk->upper = upper;
      wk->n_overlap = n_overlap;
      wk->n_edges = n_edges;
      wk->has_first = has_first;
    }
  }
}

}  // namespace cemc
