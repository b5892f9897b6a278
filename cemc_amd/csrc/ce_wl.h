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
// (histogram.cpp:102-128); a flat histogram halves f and clears the counts,
// and the walker stops once f < f_min (wang_landau_sampler.cpp:337-394).
#pragma once

#include "ce_sweep.h"

namespace cemc {

namespace {

// ln g (doubles) | visit counts (u32) | known flags (u8), after the sweep
// layout of the chain.
struct WLLayout {
  uint32_t logdos, hist, known, total;
};

__host__ __device__ inline WLLayout wl_layout(uint32_t base, int n_bins) {
  WLLayout L;
  L.logdos = align16(base);
  L.hist = L.logdos + align16(uint32_t(n_bins) * 8u);
  L.known = L.hist + align16(uint32_t(n_bins) * 4u);
  L.total = L.known + align16(uint32_t(n_bins));
  return L;
}

// Bin of total energy e: histogram.cpp:66-69 converts
// (e - e_min) * n_bins / (e_max - e_min) to int (truncation toward zero) and
// :137-140 keeps 0 <= bin < n_bins.  -1 below the range, n_bins when the
// value truncates to n_bins exactly, n_bins + 1 above (the caller tells the
// last two apart, wang_landau_sampler.cpp:257-258).
__device__ __attribute__((always_inline)) inline int wl_bin(double e, double e_min, double span, double nbd,
                                                            int n_bins) {
  const double t = trunc((e - e_min) * nbd / span);
  return t < 0.0 ? -1 : (t > nbd ? n_bins + 1 : int(t));
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

template <bool kPeriodic, int kR, bool kGtabLds, int kBM, int kNM = 3>
__global__ __launch_bounds__(64) void wl_kernel(const ChainParams* __restrict__ params,
                                                ChainState* __restrict__ states,
                                                const ModelDev* __restrict__ models,
                                                int8_t* __restrict__ occ_pool, uint16_t* __restrict__ trk_pool,
                                                const int32_t* __restrict__ chain_order, WLDev* __restrict__ walkers,
                                                double* __restrict__ logdos_pool, uint32_t* __restrict__ hist_pool,
                                                uint8_t* __restrict__ known_pool, int64_t n_checks,
                                                int64_t check_every, uint32_t* __restrict__ accept_log,
                                                int64_t log_words) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ double dphi_tab[kMaxSpecies * kMaxSpecies * kMaxBasis];  // [(old*S+new)*B + d]
  __shared__ int grp_tab[2 * kMaxGroups];
  constexpr bool kRecLds = kR == 0;
  const int chain = chain_order[blockIdx.x];
  const int lane = int(threadIdx.x);
  const ChainParams* __restrict__ pp = params + chain;
  const ModelDev& m = models[pp->model];
  WLDev* __restrict__ wk = walkers + chain;
  const int N = pp->n_sites;
  const int occ_bytes = kPeriodic ? m.pdims[0] * m.pdims[1] * m.pdims[2] : N;
  const LdsLayout L = lds_layout(occ_bytes, N, kRecLds ? m.n_rec : 0, kGtabLds ? m.gtab_size : 0);
  const int nb = wk->n_bins;
  const WLLayout WL = wl_layout(L.total, nb);
  uint8_t* occ = reinterpret_cast<uint8_t*>(lds + L.occ);
  uint16_t* trk = reinterpret_cast<uint16_t*>(lds + L.trk);
  FatRec* recs_l = reinterpret_cast<FatRec*>(lds + L.rec);
  double* gtab_l = reinterpret_cast<double*>(lds + L.gtab);
  double* lg = reinterpret_cast<double*>(lds + WL.logdos);
  uint32_t* hist = reinterpret_cast<uint32_t*>(lds + WL.hist);
  uint8_t* known = reinterpret_cast<uint8_t*>(lds + WL.known);

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

  // ---- stage the chain, the model tables and the walker's bins into LDS
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
  for (int i = lane; i < N; i += 64) trk[i] = gtrk[i];
  if constexpr (kRecLds)
    for (int i = lane; i < m.n_rec; i += 64) recs_l[i] = m.frecs[i];
  if constexpr (kGtabLds)
    for (int i = lane; i < m.gtab_size; i += 64) gtab_l[i] = m.gtab[i];
  const auto gtab = table_ptr<kGtabLds>(gtab_l, m.gtab);
  if (lane < 2 * kMaxGroups) grp_tab[lane] = lane < kMaxGroups ? m.grp_rec_off[lane] : m.grp_rec_cnt[lane - kMaxGroups];
  for (int i = lane; i < c.S * c.S; i += 64) {
    const int o = i / c.S, nw = i - o * c.S;
    for (int d = 0; d < c.B; ++d) dphi_tab[i * c.B + d] = m.basis[d * c.S + nw] - m.basis[d * c.S + o];
  }
  const int64_t boff = wk->bin_off;
  for (int i = lane; i < nb; i += 64) {
    lg[i] = logdos_pool[boff + i];
    hist[i] = hist_pool[boff + i];
    known[i] = known_pool[boff + i];
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
  double f = wk->f;
  int cur = wk->cur_bin, first = wk->is_first, converged = wk->converged;
  uint64_t n_acc = 0, n_out = 0, n_self = 0, n_flat = 0;
  int64_t done = 0;  // moves made in this launch

  // Histogram::update (histogram.cpp:71-90), by lane 0; a wave's LDS
  // operations complete in order, so the next move's reads see it
  auto visit = [&](int b) {
    if (lane == 0) {
      known[b] = 1;
      hist[b] += 1u;
      lg[b] += f;
    }
  };

  for (int64_t chk = 0; chk < n_checks && !converged; ++chk) {
    for (int64_t blk0 = 0; blk0 < check_every; blk0 += 64) {
      // lane l holds the Philox block and log u of move done + l and decodes
      // the parts of that move that do not depend on the state
      const Philox4 blk = philox4x64_10(t0 + uint64_t(done) + uint64_t(lane), 0, 0, 0, seed, uid);
      const double log_u = log(unit_double(blk.w3));
      const int n_in = int(check_every - blk0 < 64 ? check_every - blk0 : 64);
      const uint32_t ia = bounded32(uint32_t(blk.w0 >> 32), np);
      uint32_t ib = ia + 1u + bounded32(uint32_t(blk.w0), np - 1u);
      ib -= ib >= np ? np : 0u;
      const uint32_t v_old = (present_pk >> (4u * ia)) & 15u;
      const uint32_t v_new = (present_pk >> (4u * ib)) & 15u;
      const int v_ta = unpack16(seg_lo, seg_hi, ia) + int(bounded64(blk.w1, uint64_t(unpack16(cnt_lo, cnt_hi, ia))));
      const int v_tb = unpack16(seg_lo, seg_hi, ib) + int(bounded64(blk.w2, uint64_t(unpack16(cnt_lo, cnt_hi, ib))));
      for (int l = 0; l < n_in; ++l) {
        const uint32_t old_i = __builtin_amdgcn_readlane(v_old, l);
        const uint32_t new_i = __builtin_amdgcn_readlane(v_new, l);
        const int ta = __builtin_amdgcn_readlane(v_ta, l);
        const int tb = __builtin_amdgcn_readlane(v_tb, l);
        const int si = __builtin_amdgcn_readfirstlane(trk[ta]);
        const int sj = __builtin_amdgcn_readfirstlane(trk[tb]);
        const Ref ri = ref_of<kPeriodic>(c, si);
        const Ref rj = ref_of<kPeriodic>(c, sj);
        const uint32_t patch_j = kPeriodic ? wrapped_delta(c, ri, rj) : uint32_t(si);
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
          table_index<kPeriodic, kR, kBM, false, kNM>(c, rr, ri, occ, 0u, 0u, gi_i);
          table_index<kPeriodic, kR, kBM, true, kNM>(c, rr, rj, occ, patch_j, new_i, gi_j);
          part = table_sum<kR, kBM>(gi_i, gtab, dphi);
          // the swap's second flip sees site i already changed
          // (ce_updater.cpp:845-852)
          const double pj = table_sum<kR, kBM>(gi_j, gtab, dphj);
          part = kBM == 1 ? part - pj : part + pj;
        } else {
          const int gi = multi_group ? int(m.site_group[si]) : 0;
          const int gj = multi_group ? int(m.site_group[sj]) : 0;
          part = lds_flip_de<kPeriodic, kBM, false>(c, recs_l, grp_tab[gi], grp_tab[kMaxGroups + gi], lane, 64, ri,
                                                    occ, gtab, dphi, 0u, 0u);
          const double pj = lds_flip_de<kPeriodic, kBM, true>(c, recs_l, grp_tab[gj], grp_tab[kMaxGroups + gj],
                                                              lane, 64, rj, occ, gtab, dphj, patch_j, new_i);
          part = kBM == 1 ? part - pj : part + pj;
        }
        const double sum = wave_sum_f64(part);
        const double dE = kBM == 1 ? __dmul_rn(dphi[0], sum) : sum;
        const double e_new = energy + dE;
        const int bin = __builtin_amdgcn_readfirstlane(wl_bin(e_new, e_min, span, nbd, nb));
        if (bin < 0 || bin >= nb) {
          // outside the range: the proposal is undone; a walker that has
          // been in range counts its current bin once more when the
          // proposal lies above the top bin (wang_landau_sampler.cpp:251-260)
          ++n_out;
          if (!first && bin > nb) visit(cur);
        } else {
          n_self += bin == cur ? 1u : 0u;
          // accept if u < g(cur) / g(bin) = exp(min(0, ln g(cur) - ln g(bin)))
          // (histogram.cpp:130-135), or on the walker's first in-range
          // proposal.  Compared against log u with a margin far above the
          // rounding of log/exp; only draws inside it evaluate exp.
          const double diff = lg[cur] - lg[bin];
          const double margin = 1e-12 * (1.0 + fabs(diff));
          const double lu = readlane_f64(log_u, l);
          int verdict = (first || diff > 0.0 || lu < diff - margin) ? 1 : (lu > diff + margin ? 0 : 2);
          verdict = __builtin_amdgcn_readfirstlane(verdict);
          if (verdict == 2)
            verdict = __builtin_amdgcn_readfirstlane(int(unit_double(readlane_u64(blk.w3, l)) < exp(diff)));
          if (verdict) {
            if constexpr (kPeriodic) {
              // lanes 0-7 write the images of site i, lanes 8-15 those of site j
              const bool second = lane >= 8;
              const uint32_t img = second ? image_word<true>(c, rj) : image_word<true>(c, ri);
              const int idx = image_of(c, img, lane & 7);
              if (idx >= 0 && lane < 16) occ[idx] = uint8_t(second ? old_i : new_i);
            } else {
              if (lane == 0) occ[si] = uint8_t(new_i);
              if (lane == 1) occ[sj] = uint8_t(old_i);
            }
            if (lane == 0) {
              trk[ta] = uint16_t(sj);
              trk[tb] = uint16_t(si);
            }
            energy = e_new;
            cur = bin;
            first = 0;
            ++n_acc;
            if (acc_log != nullptr && lane == 0) {
              const int64_t q = done + l;
              atomicOr(acc_log + (q >> 5), 1u << (q & 31));
            }
          }
          visit(cur);
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
      }
      done += n_in;
    }
    __syncthreads();
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
    const int flat = __builtin_amdgcn_readfirstlane(int(double(mn) > flat_c * mean));
    if (flat) {
      f /= 2.0;
      ++n_flat;
      for (int i = lane; i < nb; i += 64) hist[i] = 0u;
    }
    if (f < f_min) converged = 1;
    __syncthreads();
  }

  // ---- write back
  __syncthreads();
  uint8_t* wocc = reinterpret_cast<uint8_t*>(occ_pool + pp->occ_off);
  for (int i = lane; i < N; i += 64) wocc[i] = occ[ref_of<kPeriodic>(c, i).p];
  for (int i = lane; i < N; i += 64) gtrk[i] = trk[i];
  for (int i = lane; i < nb; i += 64) {
    logdos_pool[boff + i] = lg[i];
    hist_pool[boff + i] = hist[i];
    known_pool[boff + i] = known[i];
  }
  if (lane == 0) {
    ChainState* st = states + chain;
    st->energy = energy;
    st->moves_done = t0 + uint64_t(done);
    st->n_accept += n_acc;
    wk->f = f;
    wk->cur_bin = cur;
    wk->is_first = first;
    wk->converged = converged;
    wk->n_moves += uint64_t(done);
    wk->n_accept += n_acc;
    wk->n_outside += n_out;
    wk->n_self += n_self;
    wk->n_flat += n_flat;
  }
}

}  // namespace cemc
