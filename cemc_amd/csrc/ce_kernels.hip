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
  uint32_t occ, trk, rec, cells, gtab, total;
};

__host__ __device__ inline uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }

__host__ __device__ inline LdsLayout lds_layout(int n_sites, int canonical, int n_rec, int K, int periodic,
                                               int gtab_doubles) {
  LdsLayout L;
  L.occ = 0;
  L.trk = align16(uint32_t(n_sites));
  L.rec = L.trk + (canonical ? align16(uint32_t(n_sites) * 2u) : 0u);
  L.cells = L.rec + align16(uint32_t(n_rec) * 8u);
  L.gtab = L.cells + (periodic ? align16(uint32_t(K) * 4u) : 0u);
  L.total = L.gtab + uint32_t(gtab_doubles) * 8u;
  return L;
}

}  // namespace

uint32_t sweep_lds_bytes(int n_sites, int canonical, int n_rec, int K, int periodic, int gtab_doubles) {
  return lds_layout(n_sites, canonical, n_rec, K, periodic, gtab_doubles).total;
}

// ---------------------------------------------------------------------------
// Energy change of flipping site `ref` from species `old` to `nw`, restricted
// to the sub-cluster records [it_begin, it_end) handled by this lane.  When
// `patch_site` >= 0 the species of that site reads as `patch_species` (the
// second flip of a swap sees the first one applied, ce_updater.cpp:845-852).
template <bool kPeriodic>
__device__ inline double flip_part(const ModelDev& m, const RefSite& ref, const SubRec* recs, int rec_off,
                                   int rec_cnt, int lane, const int8_t* occ, const uint32_t* cells,
                                   const double* gtab, const double* dphi, int patch_site,
                                   int patch_species) {
  double part = 0.0;
  const int S = m.n_species;
  const int B = m.n_bfs;
  for (int it = lane; it < rec_cnt; it += 64) {
    const SubRec r = recs[rec_off + it];
    const int nm = rec_size(r) - 1;
    int code = 0;
    for (int q = nm - 1; q >= 0; --q) {
      const int32_t nb = neighbour<kPeriodic>(m, ref, rec_slot(r, q), cells);
      int s = occ[nb];
      s = (nb == patch_site) ? patch_species : s;
      code = code * S + s;
    }
    const int p = m.pow_s[nm];
    const double* g = gtab + rec_gbase(r) + code;
#pragma unroll
    for (int d = 0; d < kMaxBasis; ++d) {
      if (d < B) part += dphi[d] * g[d * p];
    }
  }
  return part;
}

template <bool kPeriodic, bool kGtabLds>
__global__ __launch_bounds__(64) void sweep_kernel(const ChainParams* __restrict__ params,
                                                   ChainState* __restrict__ states,
                                                   const ModelDev* __restrict__ models,
                                                   int8_t* __restrict__ occ_pool,
                                                   uint16_t* __restrict__ trk_pool,
                                                   const int32_t* __restrict__ chain_order,
                                                   const int64_t* __restrict__ n_moves,
                                                   uint32_t* __restrict__ accept_log, int64_t log_words) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int chain = chain_order[blockIdx.x];
  const int lane = int(threadIdx.x);
  const ChainParams* __restrict__ pp = params + chain;
  const int model_id = pp->model;
  const ModelDev& m = models[model_id];
  const int N = pp->n_sites;
  const int kind = pp->kind;
  const int canonical = kind == 0;
  const LdsLayout L = lds_layout(N, canonical, m.n_rec, m.K, m.periodic, kGtabLds ? m.gtab_size : 0);
  int8_t* occ = reinterpret_cast<int8_t*>(lds + L.occ);
  uint16_t* trk = reinterpret_cast<uint16_t*>(lds + L.trk);
  SubRec* recs = reinterpret_cast<SubRec*>(lds + L.rec);
  uint32_t* cells = reinterpret_cast<uint32_t*>(lds + L.cells);
  double* gtab_l = reinterpret_cast<double*>(lds + L.gtab);

  // ---- stage chain state and model tables into LDS
  const int8_t* gocc = occ_pool + pp->occ_off;
  for (int i = lane; i < N; i += 64) occ[i] = gocc[i];
  uint16_t* gtrk = trk_pool + pp->trk_off;
  if (canonical)
    for (int i = lane; i < N; i += 64) trk[i] = gtrk[i];
  for (int i = lane; i < m.n_rec; i += 64) recs[i] = m.recs[i];
  if (m.periodic)
    for (int i = lane; i < m.K; i += 64) cells[i] = m.col_cells[i];
  if constexpr (kGtabLds)
    for (int i = lane; i < m.gtab_size; i += 64) gtab_l[i] = m.gtab[i];
  const double* gtab = kGtabLds ? gtab_l : m.gtab;
  __syncthreads();

  double energy = states[chain].energy;
  const uint64_t t0 = states[chain].moves_done;
  const int64_t nmov = n_moves[chain];
  uint64_t n_acc = 0;
  double sum_e = 0.0, sum_e2 = 0.0;
  const double kT = pp->kT;
  const uint64_t seed = pp->seed, uid = pp->uid;
  const int S = m.n_species;
  const int B = m.n_bfs;
  double j1[kMaxBasis];
#pragma unroll
  for (int d = 0; d < kMaxBasis; ++d) j1[d] = d < B ? pp->j1[d] : 0.0;

  uint32_t* log = accept_log ? accept_log + int64_t(chain) * log_words : nullptr;
  uint32_t log_word = 0;
  Philox4 blk{0, 0, 0, 0};
  for (int64_t t = 0; t < nmov; ++t) {
    const int l = int(t & 63);
    // lane l holds the Philox block of move (t - l) + l, computed once per 64 moves
    if (l == 0) blk = philox4x64_10(t0 + uint64_t(t) + uint64_t(lane), 0, 0, 0, seed, uid);
    const uint64_t w0 = readlane_u64(blk.w0, l);
    const uint64_t w1 = readlane_u64(blk.w1, l);
    const uint64_t w2 = readlane_u64(blk.w2, l);
    const uint64_t w3 = readlane_u64(blk.w3, l);
    const double u = unit_double(w3);

    int si, sj = -1, old_i, new_i, sa = 0, sb = 0, a = 0, b = 0;
    if (canonical) {
      const int np = pp->n_present;
      const uint32_t ia = bounded32(uint32_t(w0 >> 32), uint32_t(np));
      const uint32_t ib = (ia + 1u + bounded32(uint32_t(w0), uint32_t(np - 1))) % uint32_t(np);
      a = pp->present[ia];
      b = pp->present[ib];
      sa = int(bounded64(w1, uint64_t(pp->count[a])));
      sb = int(bounded64(w2, uint64_t(pp->count[b])));
      si = trk[pp->seg_off[a] + sa];
      sj = trk[pp->seg_off[b] + sb];
      old_i = a;
      new_i = b;
    } else {
      const uint64_t pick = bounded64(w1, uint64_t(m.has_background ? m.n_movable : N));
      si = m.has_background ? m.movable[pick] : int(pick);
      old_i = occ[si];
      new_i = int((uint32_t(old_i) + 1u + bounded32(uint32_t(w0), uint32_t(S - 1))) % uint32_t(S));
    }

    double dphi[kMaxBasis];
    double de_single = 0.0;
#pragma unroll
    for (int d = 0; d < kMaxBasis; ++d) {
      dphi[d] = d < B ? m.basis[d * S + new_i] - m.basis[d * S + old_i] : 0.0;
      de_single += j1[d] * dphi[d];
    }
    const RefSite ri = make_ref(m, si);
    const int gi = site_group(m, si);
    double part = flip_part<kPeriodic>(m, ri, recs, m.grp_rec_off[gi], m.grp_rec_cnt[gi], lane, occ, cells,
                                       gtab, dphi, -1, 0);
    if (canonical) {
      double dphj[kMaxBasis];
#pragma unroll
      for (int d = 0; d < kMaxBasis; ++d) {
        dphj[d] = -dphi[d];
        de_single += j1[d] * dphj[d];
      }
      const RefSite rj = make_ref(m, sj);
      const int gj = site_group(m, sj);
      part += flip_part<kPeriodic>(m, rj, recs, m.grp_rec_off[gj], m.grp_rec_cnt[gj], lane, occ, cells, gtab,
                                   dphj, si, new_i);
    }
    const double dE = wave_sum_f64(part) + de_single;
    bool accept;
    if (dE < 0.0) {
      accept = true;
    } else {
      accept = u <= exp(-dE / kT);
    }
    if (accept) {
      if (lane == 0) {
        occ[si] = int8_t(new_i);
        if (canonical) {
          occ[sj] = int8_t(old_i);
          trk[pp->seg_off[a] + sa] = uint16_t(sj);
          trk[pp->seg_off[b] + sb] = uint16_t(si);
        }
      }
      energy += dE;
      ++n_acc;
      log_word |= 1u << (t & 31);
    }
    if (log != nullptr && ((t & 31) == 31 || t + 1 == nmov)) {
      if (lane == 0) log[t >> 5] = log_word;
      log_word = 0;
    }
    sum_e += energy;
    sum_e2 += energy * energy;
    __syncthreads();
  }

  // ---- write back
  int8_t* wocc = occ_pool + pp->occ_off;
  for (int i = lane; i < N; i += 64) wocc[i] = occ[i];
  if (canonical)
    for (int i = lane; i < N; i += 64) gtrk[i] = trk[i];
  if (lane == 0) {
    ChainState* st = states + chain;
    st->energy = energy;
    st->moves_done = t0 + uint64_t(nmov);
    st->n_accept += n_acc;
    st->sum_e += sum_e;
    st->sum_e2 += sum_e2;
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
hipError_t sweep_set_max_lds(uint32_t bytes) {
  const int b = int(bytes);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&sweep_kernel<true, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&sweep_kernel<true, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&sweep_kernel<false, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&sweep_kernel<false, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, b);
  return e;
}

hipError_t launch_sweep(const LaunchSweep& a, hipStream_t stream) {
  const dim3 grid(a.n_chains), block(64);
  if (a.periodic) {
    if (a.gtab_lds)
      hipLaunchKernelGGL((sweep_kernel<true, true>), grid, block, a.lds_bytes, stream, a.params, a.states,
                         a.models, a.occ_pool, a.trk_pool, a.chain_order, a.n_moves, a.accept_log, a.log_words);
    else
      hipLaunchKernelGGL((sweep_kernel<true, false>), grid, block, a.lds_bytes, stream, a.params, a.states,
                         a.models, a.occ_pool, a.trk_pool, a.chain_order, a.n_moves, a.accept_log, a.log_words);
  } else {
    if (a.gtab_lds)
      hipLaunchKernelGGL((sweep_kernel<false, true>), grid, block, a.lds_bytes, stream, a.params, a.states,
                         a.models, a.occ_pool, a.trk_pool, a.chain_order, a.n_moves, a.accept_log, a.log_words);
    else
      hipLaunchKernelGGL((sweep_kernel<false, false>), grid, block, a.lds_bytes, stream, a.params, a.states,
                         a.models, a.occ_pool, a.trk_pool, a.chain_order, a.n_moves, a.accept_log, a.log_words);
  }
  return hipGetLastError();
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
