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
//  updater_resident_kernel  the same calls, and restores, as commands taken
//                   from a ring in mapped host memory by a workgroup that
//                   stays resident between calls.
//  restore_kernel   undo of popped changes (ce_updater.cpp:414-448).
//  slot_energy_kernel  E = N * sum_k J_k cf_k of one history slot.
#include <atomic>
#include <map>
#include <mutex>

#include "ce_sweep.h"
#include "ce_wl.h"


namespace cemc {

uint32_t sweep_lds_bytes(int occ_bytes, int trk_sites, int n_rec_lds, int gtab_doubles) {
  return lds_layout(occ_bytes, trk_sites, n_rec_lds, gtab_doubles).total;
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

// Energy from scratch through the energy tables.  A sub-cluster of site i
// with i at canonical position r contributes sum_d phi_d(s_i) G[r][d][code],
// which is its cluster's full energy term times its size n (each cluster is
// listed once for each of its n members, and the table weights carry the
// factor n of the incremental update, ce_updater.cpp:400), so
//   E = N J_c0 + sum_i [ sum_d j1_d phi_d(s_i) + sum_sub (1/n) sum_d phi_d(s_i) G ]
// with the singlet term over non-background sites.  Block b of chain c writes
// the sum over sites [256 b, 256 b + 256) to partial[c * n_blocks + b].
__global__ __launch_bounds__(256) void energy_partial_kernel(const ModelDev* __restrict__ models,
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
  const int S = m.n_species, B = m.n_bfs;
  const int kbits = S <= 2 ? 1 : (S <= 4 ? 2 : 3);
  const int pc = 1 << (3 * kbits);
  double v = 0.0;
  if (i < N) {
    const int si = occ[i];
    if (!(m.has_background && m.background[i]))
      for (int d = 0; d < B; ++d) v += pp->j1[d] * m.basis[d * S + si];
    const int g = site_group(m, i);
    if (g >= 0) {
      double phi[kMaxBasis];
      for (int d = 0; d < B; ++d) phi[d] = m.basis[d * S + si];
      const int32_t* row = m.trans + int64_t(i) * m.K;
      const int r0 = m.grp_rec_off[g], r1 = r0 + m.grp_rec_cnt[g];
      for (int r = r0; r < r1; ++r) {
        const SubRec rec = m.recs[r];
        const int nm = rec_size(rec) - 1;
        uint32_t code = 0;
        for (int q = 0; q < nm; ++q) code |= uint32_t(occ[row[rec_slot(rec, q)]]) << (kbits * q);
        const double* gt = m.gtab + rec_gbase(rec) + code;
        double e = 0.0;
        for (int d = 0; d < B; ++d) e += phi[d] * gt[d * pc];
        v += e / double(nm + 1);
      }
    }
  }
  v = wave_sum_f64(v);
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) partial[int64_t(c) * n_blocks + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(64) void energy_finalize_kernel(const ModelDev* __restrict__ models,
                                                             const ChainParams* __restrict__ params,
                                                             const int32_t* __restrict__ chain_ids,
                                                             const double* __restrict__ partial, int n_blocks,
                                                             double* __restrict__ energy_out) {
  const int c = int(blockIdx.x);
  const ChainParams* pp = params + chain_ids[c];
  const ModelDev& m = models[pp->model];
  const int nb = (pp->n_sites + 255) / 256;
  double v = 0.0;
  for (int b = int(threadIdx.x); b < nb; b += 64) v += partial[int64_t(c) * n_blocks + b];
  v = wave_sum_f64(v);
  if (threadIdx.x == 0) {
    double j0 = 0.0;
    for (int k = 0; k < m.n_cf; ++k)
      if (m.cf_kind[k] == 0) j0 += m.eci[k];
    energy_out[c] = v + double(pp->n_sites) * j0;
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
// Updater interface: apply `n` single-site changes in order, one workgroup.
// Change c reads history slot src, writes slot dst (CFHistoryTracker::get_next,
// cpp/src/cf_history_tracker.cpp:46-56) and sets the site's species
// (ce_updater.cpp:313-406).  The correlation functions of the current slot
// stay in LDS between changes.  Per change every thread forms the products of
// its plan items (one dependent gather of the members' species) into LDS,
// then wave w sums the segments w, w + 4, ... in a fixed order, so results do
// not depend on timing.  After the last change E = N J.cf goes to the slot
// and, when `result` is given, to mapped host memory, so the host reads it
// without a copy or a stream wait.
uint32_t updater_lds_bytes(int n_cf, int chunk, int occ_sites) {
  const int ncf_pad = (n_cf + 31) / 32 * 32;
  return uint32_t(3 * sizeof(double) + sizeof(UpdSeg)) * uint32_t(ncf_pad) +
         uint32_t(sizeof(double)) * uint32_t(chunk) + uint32_t((occ_sites + 15) / 16 * 16);
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima).
__device__ inline void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Pointers into the updater's dynamic LDS: correlation functions of the
// current slot, segment sums over earlier passes, ECIs, segments of the
// change, item products of one pass, staged occupation.
struct UpdLds {
  double* cfl;
  double* acc;
  double* ecil;
  UpdSeg* segl;
  double* prod;
  int8_t* occ_l;
};
__device__ inline UpdLds upd_lds_layout(const UpdPlan& plan, double* upd_lds) {
  const int ncf_pad = (plan.n_cf + 31) / 32 * 32;
  UpdLds L;
  L.cfl = upd_lds;
  L.acc = L.cfl + ncf_pad;
  L.ecil = L.acc + ncf_pad;
  L.segl = reinterpret_cast<UpdSeg*>(L.ecil + ncf_pad);
  L.prod = reinterpret_cast<double*>(L.segl + ncf_pad);
  L.occ_l = reinterpret_cast<int8_t*>(L.prod + plan.chunk);
  return L;
}

// Stages what every call reads and no change alters: ECIs, basis functions
// and (kOccLds) the occupation.
template <bool kOccLds>
__device__ inline void updater_stage_fixed(const UpdPlan& plan, const int8_t* occ, const double* __restrict__ eci,
                                           double* upd_lds, double* phi) {
  const int tid = int(threadIdx.x);
  const UpdLds L = upd_lds_layout(plan, upd_lds);
  for (int k = tid; k < plan.n_cf; k += 256) L.ecil[k] = eci[k];
  for (int i = tid; i < plan.n_bfs * plan.n_species; i += 256) phi[i] = plan.phi[i];
  if constexpr (kOccLds) {
    // 4-byte words, then the last N mod 4 bytes (the allocation is aligned)
    const int N = plan.n_sites;
    const int nw4 = N >> 2;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(occ);
    uint32_t* dst = reinterpret_cast<uint32_t*>(L.occ_l);
    for (int i = tid; i < nw4; i += 256) dst[i] = src[i];
    if (tid < (N & 3)) L.occ_l[4 * nw4 + tid] = occ[4 * nw4 + tid];
  }
}

// What a resident kernel keeps between fused calls of one translation
// group: the group's first segment (-1: nothing held), the call's number of
// changes, and each thread's plan item.  The segments and their map stay
// staged in LDS alongside.
struct UpdHeldGroup {
  int sb = -1;
  int n = 0;
  uint32_t it[4] = {0u, 0u, 0u, 0u};
};

// One call of the updater interface: changes c = 0 .. n-1 with words
// word(c, w), then the energy of the last slot.  kOccLds: the occupation is
// staged in LDS; kPeriodic: closed-form periodic lattice (else members come
// from the translation table).  With the occupation in LDS the workgroup
// exchanges data through LDS only, so its barriers wait for LDS operations
// alone and global loads and stores stay in flight across them (a full
// barrier waits for every load and store of the wave, one memory round
// trip); otherwise the occupation written for a change is read back from
// global memory by the next change, which needs the full barrier.  A change
// in the translation group of the change before it reuses that change's
// items and segments (every swap on a one-group lattice).  kStage: a launch
// for this call alone stages ECIs, basis functions and occupation and loads
// the correlation functions of the first change's source slot; a resident
// kernel has staged the first three once and reloads the correlation
// functions only when the source slot is not the slot `held` in LDS.
// Returns the slot held in LDS afterwards (the last change's destination).
// `hg` (resident kernels): the group held from the call before; a fused call
// of the same group and size takes its items and segments from there.
template <bool kOccLds, bool kPeriodic, bool kStage, class Word>
__device__ __forceinline__ int updater_body(const UpdPlan& plan, int8_t* occ, double* __restrict__ ring,
                                            double* __restrict__ slot_energy, Word word, int n,
                                            const double* __restrict__ eci, UpdResult* result,
                                            uint64_t* result_cf, double* upd_lds, double* phi, double* red,
                                            int held, UpdHeldGroup* hg) {
  const int tid = int(threadIdx.x);
  const int wave = tid >> 6, lane = tid & 63;
  const int S = plan.n_species;
  const int N = plan.n_sites;
  const int ncf = plan.n_cf;
  const UpdLds L = upd_lds_layout(plan, upd_lds);
  double* cfl = L.cfl;
  double* acc = L.acc;
  double* ecil = L.ecil;
  UpdSeg* segl = L.segl;
  double* prod = L.prod;
  int8_t* occ_l = L.occ_l;
  auto barrier = [&] {
    if constexpr (kOccLds) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    } else {
      __threadfence_block();
      __syncthreads();
    }
  };
  auto species_at = [&](int32_t i) -> int {
    if constexpr (kOccLds) {
      return occ_l[i];
    } else {
      return occ[i];
    }
  };
  // items (m[0..2], aux) and segments (k, begin, end, pad; mul, div) are
  // held as vectors so that they stay in registers
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  static_assert(sizeof(UpdItem) == 16 && sizeof(UpdSeg) == 32, "updater plan layout");
  // cell coordinates of a site (closed-form periodic lattices)
  auto cell_of = [&](int32_t site, int32_t& rx, int32_t& ry, int32_t& rz) {
    rx = ry = rz = 0;
    if constexpr (kPeriodic) {
      const int32_t l23 = plan.dims[1] * plan.dims[2];
      rx = site / l23;
      const int32_t rem = site - rx * l23;
      ry = rem / plan.dims[2];
      rz = rem - ry * plan.dims[2];
    }
  };
  // product of one plan item for a change of `site` from `old` to `nw`,
  // the members' species read through spec(site index).  Branch-free: the
  // three member gathers and the basis-function reads are issued together,
  // then the factors are multiplied in position order (1.0 for positions
  // past the cluster's size, an exact factor), as a loop over positions
  // would multiply them.
  auto product_of = [&](const u32x4& it, int32_t site, int32_t rx, int32_t ry, int32_t rz, int nw, int old,
                        auto&& spec) {
    const uint32_t aux = it.w;
    const int nm = int((aux >> 14) & 3u);
    const int rp = int((aux >> 12) & 3u);
    int sp[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const uint32_t mq = q == 0 ? it.x : (q == 1 ? it.y : it.z);
      int32_t nb = site;
      if constexpr (kPeriodic) {
        int32_t x = rx + int32_t(mq & 0xFFu);
        int32_t y = ry + int32_t((mq >> 8) & 0xFFu);
        int32_t z = rz + int32_t((mq >> 16) & 0xFFu);
        x -= (x >= plan.dims[0]) ? plan.dims[0] : 0;
        y -= (y >= plan.dims[1]) ? plan.dims[1] : 0;
        z -= (z >= plan.dims[2]) ? plan.dims[2] : 0;
        if (q < nm) nb = (x * plan.dims[1] + y) * plan.dims[2] + z;
      } else {
        if (q < nm) nb = plan.trans[int64_t(site) * plan.K + mq];
      }
      sp[q] = spec(nb);
    }
    double f[kMaxClusterSize];
#pragma unroll
    for (int pos = 0; pos < kMaxClusterSize; ++pos) {
      const int d = pos > nm ? 0 : int((aux >> (3 * pos)) & 7u);  // positions past the size read phi[0..S)
      const int q = pos - (pos > rp ? 1 : 0);  // member at this position
      const int s = q == 0 ? sp[0] : (q == 1 ? sp[1] : sp[2]);
      const double member = phi[d * S + s];
      const double ref = phi[d * S + nw] - phi[d * S + old];
      f[pos] = pos > nm ? 1.0 : (pos == rp ? ref : member);
    }
    double p = 1.0;
#pragma unroll
    for (int pos = 0; pos < kMaxClusterSize; ++pos) p *= f[pos];
    return p;
  };
  const u32x4* items = reinterpret_cast<const u32x4*>(plan.items);
  u32x4 it0 = {0u, 0u, 0u, 0u};  // item ib + tid of the current group
  i32x4 si0 = {0, 0, 0, 0};      // segment sb + tid of the current group
  f64x2 sd0 = {0.0, 0.0};
  auto fetch = [&](int ib, int ie, int sb, int se) {
    if (ib + tid < min(ie, ib + plan.chunk)) it0 = items[ib + tid];
    if (sb + tid < se) {
      const i32x4* p = reinterpret_cast<const i32x4*>(plan.segs + sb + tid);
      si0 = p[0];
      sd0 = *reinterpret_cast<const f64x2*>(p + 1);
    }
  };
  // A resident call whose changes all lie in one translation group and fit
  // one pass forms the products of every change at once (the members read
  // the occupation as it is before their change: the call's earlier changes
  // override the staged occupation), sums every (change, segment) pair at
  // once, and then each thread runs the correlation functions it owns
  // through the changes in order.  The sums are formed in the same order as
  // change by change, so the results are identical.
  bool fuse = false;
  int fm = 0, fseg = 0;
  const int ib0 = word(0, 3), ie0 = word(0, 4), sb0 = word(0, 5), se0 = word(0, 6);
  if constexpr (!kStage) {
    fuse = n >= 2;
    for (int c = 1; c < n && fuse; ++c)
      fuse = word(c, 3) == ib0 && word(c, 4) == ie0 && word(c, 5) == sb0 && word(c, 6) == se0;
    fm = ie0 - ib0;
    fseg = se0 - sb0;
    fuse = fuse && fm > 0 && n * (fm + fseg) <= plan.chunk;
  }
  const bool reuse = fuse && hg != nullptr && hg->sb == sb0 && hg->n == n;
  if (reuse) {
    it0 = u32x4{hg->it[0], hg->it[1], hg->it[2], hg->it[3]};
  } else {
    fetch(ib0, ie0, sb0, se0);
    if (fuse && tid >= fm && tid < n * fm) it0 = items[ib0 + tid % fm];
  }
  {
    // ECIs go to LDS now, so that no wait at the end covers the stores made
    // meanwhile (loads and stores share one counter)
    const int src0 = word(0, 2) & 0xFFFF;
    if (kStage || src0 != held)
      for (int k = tid; k < ncf; k += 256) cfl[k] = ring[int64_t(src0) * ncf + k];
    if constexpr (kStage) updater_stage_fixed<kOccLds>(plan, occ, eci, upd_lds, phi);
  }
  // The loads above complete before the first barrier in any case.  Waiting
  // for all of them here, by name, tells the compiler that the item held in
  // registers has arrived, so the changes that reuse it wait for no memory
  // operation (a conservative wait there would cover the history-slot stores
  // of the change before).
  if (!reuse) vm_wait_all();
  int last_dst = -1;
  int cur_sb = -1;  // segments held in segl (first segment index)
  if (fuse) {
    int32_t* segof = reinterpret_cast<int32_t*>(acc);  // segment of each correlation function, or -1
    double* dlt = prod + n * fm;                        // change of each (change, segment) pair
    if (!reuse) {
      if (sb0 + tid < se0) segl[tid] = UpdSeg{si0.x, si0.y, si0.z, si0.w, sd0.x, sd0.y};
      for (int j = tid + 256; j < fseg; j += 256) segl[j] = plan.segs[sb0 + j];
      for (int k = tid; k < ncf; k += 256) segof[k] = -1;
      barrier();
      for (int j = tid; j < fseg; j += 256) segof[segl[j].k] = j;
    }
    for (int g = tid; g < n * fm; g += 256) {
      const int c = g / fm;
      const u32x4 it = g == tid ? it0 : items[ib0 + (g - c * fm)];
      const int site = word(c, 0);
      const int nw = word(c, 1) & 0xFF, old = (word(c, 1) >> 8) & 0xFF;
      int32_t rx, ry, rz;
      cell_of(site, rx, ry, rz);
      auto spec = [&](int32_t i) -> int {
        int v = species_at(i);
        for (int q = 0; q < c; ++q)
          if (word(q, 0) == i) v = word(q, 1) & 0xFF;
        return v;
      };
      prod[g] = product_of(it, site, rx, ry, rz, nw, old, spec);
    }
    barrier();
    for (int pj = wave; pj < n * fseg; pj += 4) {
      const int c = pj / fseg;
      const UpdSeg g = segl[pj - c * fseg];
      double v = 0.0;
      for (int i = g.begin + lane; i < g.end; i += 64) v += prod[c * fm + (i - ib0)];
      v = wave_sum_f64(v);
      if (lane == 0) dlt[pj] = ((0.0 + v) * g.mul) / g.div;
    }
    barrier();
    // the occupation after every change of the call (all products are formed)
    if (tid == 0)
      for (int c = 0; c < n; ++c) {
        occ[word(c, 0)] = int8_t(word(c, 1) & 0xFF);
        if constexpr (kOccLds) occ_l[word(c, 0)] = int8_t(word(c, 1) & 0xFF);
      }
    for (int k = tid; k < ncf; k += 256) {
      const int j = segof[k];
      int sd = -1;  // decoration of a singlet function (fixed indices: scalar loads of the arguments)
#pragma unroll
      for (int q = 0; q < kUpdMaxSinglets; ++q)
        if (q < plan.n_singlets && (plan.singlets[q] & 0xFFFF) == k) sd = plan.singlets[q] >> 16;
      double v = cfl[k];
      for (int c = 0; c < n; ++c) {
        if (j >= 0) v += dlt[c * fseg + j];
        if (sd >= 0) {
          const int nw = word(c, 1) & 0xFF, old = (word(c, 1) >> 8) & 0xFF;
          v += (phi[sd * S + nw] - phi[sd * S + old]) / double(unsigned(N));
        }
        ring[int64_t((word(c, 2) >> 16) & 0xFFFF) * ncf + k] = v;
      }
      cfl[k] = v;
    }
    last_dst = (word(n - 1, 2) >> 16) & 0xFFFF;
    if (hg != nullptr) {
      hg->sb = sb0;
      hg->n = n;
      hg->it[0] = it0.x;
      hg->it[1] = it0.y;
      hg->it[2] = it0.z;
      hg->it[3] = it0.w;
    }
  } else if (hg != nullptr) {
    hg->sb = -1;  // the calls below restage segl and reuse the map's LDS
  }
  for (int c = 0; c < (fuse ? 0 : n); ++c) {
    const int site = word(c, 0);
    const int nw = word(c, 1) & 0xFF, old = (word(c, 1) >> 8) & 0xFF;
    const int dst = (word(c, 2) >> 16) & 0xFFFF;
    const int ib = word(c, 3), ie = word(c, 4), sb = word(c, 5), se = word(c, 6);
    last_dst = dst;
    if (sb != cur_sb) {
      if (c > 0) {
        fetch(ib, ie, sb, se);
        vm_wait_all();
      }
      if (sb + tid < se) segl[tid] = UpdSeg{si0.x, si0.y, si0.z, si0.w, sd0.x, sd0.y};
      for (int j = tid + 256; j < se - sb; j += 256) segl[j] = plan.segs[sb + j];
      cur_sb = sb;
      // the staged data and the segments are visible after this barrier
      // (with the same group the barrier closing the change before suffices)
      barrier();
    }
    int32_t rx, ry, rz;
    cell_of(site, rx, ry, rz);
    auto product = [&](const u32x4& it) { return product_of(it, site, rx, ry, rz, nw, old, species_at); };
    for (int cb = ib; cb < ie; cb += plan.chunk) {
      const int ce = min(ie, cb + plan.chunk);
      int idx = cb + tid;
      if (cb == ib) {
        if (idx < ce) prod[idx - cb] = product(it0);
        idx += 256;
      }
      for (; idx < ce; idx += 256) prod[idx - cb] = product(items[idx]);
      barrier();
      const bool last_pass = ce == ie;
      for (int j = wave; j < se - sb; j += 4) {
        const UpdSeg g = segl[j];
        const int lo = max(g.begin, cb), hi = min(g.end, ce);
        double v = 0.0;
        for (int i = lo + lane; i < hi; i += 64) v += prod[i - cb];
        v = wave_sum_f64(v);
        if (lane == 0) {
          const double t = (cb == ib ? 0.0 : acc[j]) + v;
          if (last_pass) {
            cfl[g.k] += (t * g.mul) / g.div;
          } else {
            acc[j] = t;
          }
        }
      }
      if (!last_pass) barrier();
    }
    // a change without items (a group whose correlation functions are all
    // singlets) has had no barrier since the history-slot stores of the
    // change before read `cfl`
    if (ib == ie) barrier();
    // thread j updates singlet j (fixed indices: the list is read from the
    // kernel arguments with scalar loads)
#pragma unroll
    for (int j = 0; j < kUpdMaxSinglets; ++j) {
      if (tid == j && j < plan.n_singlets) {
        const int k = plan.singlets[j] & 0xFFFF, d = plan.singlets[j] >> 16;
        cfl[k] += (phi[d * S + nw] - phi[d * S + old]) / double(unsigned(N));
      }
    }
    // every gather of this change is done (the pass barrier above)
    if (tid == 0) {
      occ[site] = int8_t(nw);
      if constexpr (kOccLds) occ_l[site] = int8_t(nw);
    }
    barrier();
    for (int k = tid; k < ncf; k += 256) ring[int64_t(dst) * ncf + k] = cfl[k];
  }
  if (last_dst < 0) return held;
  // the final correlation functions to mapped host memory; the host marks
  // every value pending before the launch and waits for each, so these
  // stores need no ordering against the energy
  if (result_cf != nullptr)
    for (int k = tid; k < ncf; k += 256)
      __hip_atomic_store(&result_cf[k], uint64_t(__double_as_longlong(cfl[k])), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
  double part = 0.0;
  for (int k = tid; k < ncf; k += 256) part += ecil[k] * cfl[k];
  part = wave_sum_f64(part);
  if (lane == 0) red[wave] = part;
  barrier();
  if (tid == 0) {
    const double e = ((red[0] + red[1]) + (red[2] + red[3])) * double(unsigned(N));
    slot_energy[last_dst] = e;
    if (result != nullptr)
      __hip_atomic_store(&result->energy_bits, uint64_t(__double_as_longlong(e)), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
  }
  return last_dst;
}

// A launch for one call: the changes travel in the kernel arguments
// (kInArgs) or in `changes` in device memory.
template <bool kInArgs, bool kOccLds, bool kPeriodic>
__global__ __launch_bounds__(256) void updater_kernel(UpdPlan plan, int8_t* occ, double* __restrict__ ring,
                                                      double* __restrict__ slot_energy, UpdChanges inl,
                                                      const int32_t* __restrict__ changes, int n,
                                                      const double* __restrict__ eci, UpdResult* result,
                                                      uint64_t* result_cf) {
  extern __shared__ double upd_lds[];
  __shared__ double phi[kMaxBasis * kMaxSpecies];
  __shared__ double red[4];
  auto word = [&](int c, int w) { return kInArgs ? inl.v[kUpdWords * c + w] : changes[kUpdWords * c + w]; };
  (void)updater_body<kOccLds, kPeriodic, true>(plan, occ, ring, slot_energy, word, n, eci, result, result_cf,
                                                upd_lds, phi, red, -1, nullptr);
}

// A resident updater: one workgroup that stays on its compute unit and takes
// the calls of one updater as commands from a ring in mapped host memory, so
// a call costs no launch.  Wave 0 polls slot next % kUpdRing with one
// system-scope 8-byte load per lane, which reaches host memory every time,
// until all 64 words carry command `next`'s number, copies the command to
// LDS, runs it and stores `next` in `consumed`, which frees the slot.  The workgroup leaves, all
// waves together, on a quit command or when no command has come for
// `idle_ticks` ticks of the constant-rate wall clock; the host launches it
// again from the first command it has not copied.  Between commands the
// ECIs, basis functions, staged occupation and the correlation functions of
// the last destination slot stay in LDS.
template <bool kOccLds, bool kPeriodic>
__global__ __launch_bounds__(256) void updater_resident_kernel(UpdPlan plan, int8_t* occ, double* __restrict__ ring,
                                                               double* __restrict__ slot_energy,
                                                               const double* __restrict__ eci, UpdResult* result,
                                                               uint64_t* result_cf, UpdMailbox* mb, int64_t next,
                                                               int64_t idle_ticks) {
  extern __shared__ double upd_lds[];
  __shared__ double phi[kMaxBasis * kMaxSpecies];
  __shared__ double red[4];
  __shared__ int32_t cmd[4 + kUpdCmdWords];  // op, n, want, nr, words
  const int tid = int(threadIdx.x);
  updater_stage_fixed<kOccLds>(plan, occ, eci, upd_lds, phi);
  int8_t* occ_l = upd_lds_layout(plan, upd_lds).occ_l;
  int held = -1;
  UpdHeldGroup hg;
  for (;; ++next) {
    if (tid < 64) {
      // every lane takes part in every poll and reaches the same verdict
      uint64_t* w = &mb->cmd[next % kUpdRing].w[tid];
      const uint32_t tag = uint32_t(uint64_t(next));
      const uint64_t t0 = uint64_t(wall_clock64());
      uint64_t v = 0;
      int got = 0;
      for (;;) {
        v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (__all(uint32_t(v >> 32) == tag)) {
          got = 1;
          break;
        }
        if (uint64_t(wall_clock64()) - t0 > uint64_t(idle_ticks)) break;
      }
      if (got) {
        cmd[tid] = int32_t(uint32_t(v));
      } else if (tid == 0) {
        cmd[0] = 0;
      }
    }
    __syncthreads();
    const int op = cmd[0];
    if (op != kUpdOpChanges && op != kUpdOpRestore) {  // quit, or idle
      if (op == kUpdOpQuit && tid == 0)
        __hip_atomic_store(&mb->consumed, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      return;
    }
    const int n = cmd[1], nr = cmd[3];
    const int32_t* w = cmd + 4;
    if (nr > 0) {
      // restores of undone changes (sites, then species), in order
      if (tid == 0)
        for (int i = 0; i < nr; ++i) {
          const int site = w[i];
          occ[site] = int8_t(w[nr + i]);
          if constexpr (kOccLds) occ_l[site] = int8_t(w[nr + i]);
        }
      __syncthreads();
    }
    if (op == kUpdOpChanges) {
      const bool want = cmd[2] != 0;
      const int32_t* cw = w + 2 * nr;
      auto word = [&](int ci, int k) { return cw[kUpdWords * ci + k]; };
      held = updater_body<kOccLds, kPeriodic, false>(plan, occ, ring, slot_energy, word, n, eci,
                                                      want ? result : nullptr, want ? result_cf : nullptr, upd_lds,
                                                      phi, red, held, &hg);
    }
    // the slot is released once the command has run: a store to host
    // memory issued before it would hold up the command's first wait on
    // memory by a round trip over the bus
    if (tid == 0) __hip_atomic_store(&mb->consumed, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    // the command's reads of `cmd` and its global stores are done before
    // the next command is copied and runs
    __syncthreads();
  }
}

__global__ void restore_kernel(int8_t* __restrict__ occ, InlineChanges inl, const int32_t* __restrict__ buf,
                               int n) {
  // Applied in order by one thread: later pops restore earlier states.
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const bool in_args = n <= 2 * kInlineChanges;
    for (int i = 0; i < n; ++i) {
      const int site = in_args ? inl.v[i] : buf[i];
      const int sp = in_args ? inl.v[n + i] : buf[n + i];
      occ[site] = int8_t(sp);
    }
  }
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
// Replica exchange after a sweep of every replica (parallel tempering,
// cemc/mcmc/parallel_tempering.py:138-191).  One workgroup:
//  1. each chain's energy sums since the last cycle move into its slot's sums;
//  2. the direction of cycle c comes from Philox block (c, 0, 0, 0) of key
//     (seed, uid): bit 63 of w0 set = "up", pairs (0,1), (2,3), ...;
//     clear = "down", pairs (R-1,R-2), (R-3,R-4), ...  (the reference picks
//     the direction with random.choice, parallel_tempering.py:188-191);
//  3. pair k uses u = unit_double(w3) of block (c, 1 + k, 0, 0) and accepts
//     if u < exp((1/kT_a - 1/kT_b) (E_a - E_b)) (parallel_tempering.py:138-144,
//     166-168); an accepted pair swaps the chains of the two slots and each
//     chain takes its new slot's temperature.
__global__ __launch_bounds__(256) void pt_exchange_kernel(PTDev* __restrict__ pt_ptr, ChainState* __restrict__ states,
                                                          ChainParams* __restrict__ params) {
  __shared__ int up;
  PTDev& pt = *pt_ptr;
  const int R = pt.n_rep;
  const int tid = int(threadIdx.x);
  const uint64_t cyc = pt.cycle;
  for (int c = tid; c < R; c += 256) {
    ChainState& st = states[c];
    const uint64_t n = st.n_samples;
    if (n > 0) {
      const int t = pt.slot_of_chain[c];
      const double d = st.e_ref - pt.e_ref[t];
      const double dn = double(n);
      pt.sum_de[t] += st.sum_de + dn * d;
      pt.sum_de2[t] += st.sum_de2 + 2.0 * d * st.sum_de + dn * d * d;
      pt.n_samples[t] += n;
      st.n_samples = 0;
      st.sum_de = 0.0;
      st.sum_de2 = 0.0;
    }
  }
  if (tid == 0) {
    const Philox4 b = philox4x64_10(cyc, 0, 0, 0, pt.seed, pt.uid);
    up = int(b.w0 >> 63);
  }
  __syncthreads();
  const int n_pairs = R / 2;  // pairs (i, i+1) for even i <= R-2, or the mirror
  for (int k = tid; k < n_pairs; k += 256) {
    const int a = up ? 2 * k : R - 1 - 2 * k;
    const int b = up ? a + 1 : a - 1;
    const int ca = pt.chain_of_slot[a], cb = pt.chain_of_slot[b];
    const double ea = states[ca].energy, eb = states[cb].energy;
    const double p = exp((1.0 / pt.kT[a] - 1.0 / pt.kT[b]) * (ea - eb));
    const Philox4 blk = philox4x64_10(cyc, 1 + uint64_t(k), 0, 0, pt.seed, pt.uid);
    const bool acc = unit_double(blk.w3) < p;
    const int lo = a < b ? a : b;
    pt.n_try[lo] += 1;
    if (acc) {
      pt.n_accept[lo] += 1;
      pt.chain_of_slot[a] = cb;
      pt.chain_of_slot[b] = ca;
      pt.slot_of_chain[cb] = a;
      pt.slot_of_chain[ca] = b;
      params[cb].kT = pt.kT[a];
      params[ca].kT = pt.kT[b];
    }
  }
  __syncthreads();
  if (tid == 0) pt.cycle = cyc + 1;
}

// ---------------------------------------------------------------------------
// Launch wrappers.  The sweep kernel's instantiations are compiled in
// separate objects (ce_sweep_inst.hip) and reached through these lookups.
#define CEMC_DECLARE_SWEEP_FN(W, P) \
  const void* sweep_fn_w##W##_p##P(bool canonical, int kr, bool glds, int kbm, int spec, bool hooks, int nm);
CEMC_DECLARE_SWEEP_FN(1, 0)
CEMC_DECLARE_SWEEP_FN(1, 1)
CEMC_DECLARE_SWEEP_FN(2, 0)
CEMC_DECLARE_SWEEP_FN(2, 1)
CEMC_DECLARE_SWEEP_FN(4, 0)
CEMC_DECLARE_SWEEP_FN(4, 1)
CEMC_DECLARE_SWEEP_FN(8, 0)
CEMC_DECLARE_SWEEP_FN(8, 1)
#undef CEMC_DECLARE_SWEEP_FN

const void* field_fn(bool canonical, int kr, int kbm);
const void* wl_fn_p0s0(int kr, bool glds, int kbm, int nm, int mode);
const void* wl_fn_p1s0(int kr, bool glds, int kbm, int nm, int mode);
const void* wl_fn_p0s1(int kr, bool glds, int kbm, int nm, int mode);
const void* wl_fn_p1s1(int kr, bool glds, int kbm, int nm, int mode);

static const void* wl_fn(bool periodic, bool sgc, int kr, bool glds, int kbm, int nm, int mode) {
  if (sgc) return periodic ? wl_fn_p1s1(kr, glds, kbm, nm, mode) : wl_fn_p0s1(kr, glds, kbm, nm, mode);
  return periodic ? wl_fn_p1s0(kr, glds, kbm, nm, mode) : wl_fn_p0s0(kr, glds, kbm, nm, mode);
}

static const void* sweep_fn(bool periodic, bool canonical, int kr, bool glds, int kbm, int waves, int spec,
                            bool hooks = false, int nm = 3) {
  const bool h = hooks;
  switch (waves) {
    case 1: return periodic ? sweep_fn_w1_p1(canonical, kr, glds, kbm, spec, h, nm) : sweep_fn_w1_p0(canonical, kr, glds, kbm, spec, h, nm);
    case 2: return periodic ? sweep_fn_w2_p1(canonical, kr, glds, kbm, spec, h, nm) : sweep_fn_w2_p0(canonical, kr, glds, kbm, spec, h, nm);
    case 4: return periodic ? sweep_fn_w4_p1(canonical, kr, glds, kbm, spec, h, nm) : sweep_fn_w4_p0(canonical, kr, glds, kbm, spec, h, nm);
    case 8: return periodic ? sweep_fn_w8_p1(canonical, kr, glds, kbm, spec, h, nm) : sweep_fn_w8_p0(canonical, kr, glds, kbm, spec, h, nm);
    default: return nullptr;
  }
}

bool sweep_variant_ok(int kr, bool glds, int kbm, int waves, int spec, bool periodic) {
  (void)glds;
  if (waves != 1 && waves != 2 && waves != 4 && waves != 8) return false;
  // spec: 1, kSpecMoves, or kSpecMoves * waves (wide rounds, 2-8 waves)
  const bool wide = waves > 1 && spec == kSpecMoves * waves;
  if (spec != 1 && !((spec == kSpecMoves || wide) && periodic && kr > 0)) return false;
  if (kr == 2 || kr == 4 || kr == 8) return kbm == 1 || kbm == 2 || kbm == 3;
  return waves == 1 && kr == 0 && (kbm == 1 || kbm == 2 || kbm == 3 || kbm == 7);
}

uint32_t sweep_static_lds() {
  return uint32_t(sizeof(double) * (kMaxSpecies * kMaxSpecies * kMaxBasis + kMaxSpecies * kMaxSpecies +
                                    kMaxBasis * kMaxSpecies + 2 * 8 + kMaxBasis * 8 + 2 * kSpecMoves * 8 +
                                    kSpecMoves * 8) +
                  sizeof(int) * 2 * kMaxGroups + sizeof(int) * 2);
}

hipError_t sweep_set_max_lds(uint32_t bytes) {
  for (int p = 0; p < 2; ++p)
    for (int cn = 0; cn < 2; ++cn)
      for (int w : {1, 2, 4, 8})
        for (int kr : {0, 2, 4, 8})
          for (int g = 0; g < 2; ++g)
            for (int kbm : {1, 2, 3, 7})
              for (int sp : {1, kSpecMoves, kSpecMoves * w})
                for (int hk = 0; hk < 2; ++hk)
                  for (int nm : {2, 3}) {
              if (sp == kSpecMoves * w && w == 1) continue;
              if (!sweep_variant_ok(kr, g != 0, kbm, w, sp, p != 0)) continue;
              if (hk && (w != 1 || sp != 1)) continue;
              if (nm == 2 && kr == 0) continue;
              const void* fn = sweep_fn(p != 0, cn != 0, kr, g != 0, kbm, w, sp, hk != 0, nm);
              if (fn == nullptr) return hipErrorInvalidValue;
              hipFuncAttributes attr;
              hipError_t e = hipFuncGetAttributes(&attr, fn);
              if (e != hipSuccess) return e;
              const int dyn = int(bytes) - int(attr.sharedSizeBytes);
              if (dyn <= 0) return hipErrorInvalidValue;
              e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
              if (e != hipSuccess) return e;
            }
  for (int p = 0; p < 2; ++p)
    for (int kr : {0, 2, 4, 8})
      for (int g = 0; g < 2; ++g)
        for (int kbm : {1, 2, 3, 7})
          for (int nm : {2, 3})
            for (int mode = 0; mode < 6; ++mode) {
              if (kbm == 7 && kr != 0) continue;
              if (nm == 2 && kr == 0) continue;
              const void* fn = wl_fn(p != 0, mode >= 3, kr, g != 0, kbm, nm, mode % 3);
              if (fn == nullptr) return hipErrorInvalidValue;
              hipFuncAttributes attr;
              hipError_t e = hipFuncGetAttributes(&attr, fn);
              if (e != hipSuccess) return e;
              const int dyn = int(bytes) - int(attr.sharedSizeBytes);
              if (dyn <= 0) return hipErrorInvalidValue;
              e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
              if (e != hipSuccess) return e;
            }
  for (int cn = 0; cn < 2; ++cn)
    for (int kr : {2, 4, 8})
      for (int kbm : {1, 2, 3}) {
        const void* fn = field_fn(cn != 0, kr, kbm);
        if (fn == nullptr) return hipErrorInvalidValue;
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
  const void* fn = nullptr;
  if (a.field) {
    if (!a.periodic || a.waves != 1 || a.hooks != nullptr) return hipErrorInvalidValue;
    fn = field_fn(a.canonical, a.kr, a.kbm);
  } else {
    if (!sweep_variant_ok(a.kr, a.gtab_lds, a.kbm, a.waves, a.spec, a.periodic)) return hipErrorInvalidValue;
    if (a.hooks != nullptr && (a.waves != 1 || a.spec != 1)) return hipErrorInvalidValue;
    if (a.nm != 3 && !(a.nm == 2 && a.kr > 0)) return hipErrorInvalidValue;
    fn = sweep_fn(a.periodic, a.canonical, a.kr, a.gtab_lds, a.kbm, a.waves, a.spec, a.hooks != nullptr, a.nm);
  }
  if (fn == nullptr) return hipErrorInvalidValue;
  const ChainParams* params = a.params;
  ChainState* states = a.states;
  const ModelDev* models = a.models;
  int8_t* occ_pool = a.occ_pool;
  uint16_t* trk_pool = a.trk_pool;
  const int32_t* order = a.chain_order;
  const int64_t* n_moves = a.n_moves;
  uint32_t* log = a.accept_log;
  int64_t log_words = a.log_words;
  HookDev* hooks = a.hooks;
  int64_t time_limit = a.time_limit;
  if (a.field) {
    void* fargs[] = {&params, &states, &models, &occ_pool, &trk_pool, &order, &n_moves, &log, &log_words, &hooks};
    return hipLaunchKernel(fn, dim3(a.n_chains), dim3(64 * a.waves), fargs, a.lds_bytes, stream);
  }
  void* args[] = {&params, &states, &models, &occ_pool, &trk_pool, &order, &n_moves, &log, &log_words, &hooks,
                  &time_limit};
  return hipLaunchKernel(fn, dim3(a.n_chains), dim3(64 * a.waves), args, a.lds_bytes, stream);
}

uint32_t wl_lds_bytes(uint32_t base, int n_bins, bool adaptive) { return wl_layout(base, n_bins, adaptive).total; }

hipError_t launch_wl(const LaunchWL& a, hipStream_t stream) {
  if (!((a.kr == 0 && (a.kbm == 1 || a.kbm == 2 || a.kbm == 3 || a.kbm == 7)) ||
        ((a.kr == 2 || a.kr == 4 || a.kr == 8) && (a.kbm == 1 || a.kbm == 2 || a.kbm == 3))))
    return hipErrorInvalidValue;
  if (a.nm != 3 && !(a.nm == 2 && a.kr > 0)) return hipErrorInvalidValue;
  const void* fn =
      wl_fn(a.periodic, a.sgc, a.kr, a.gtab_lds, a.kbm, a.nm, a.shared != nullptr ? 2 : (a.adaptive ? 1 : 0));
  if (fn == nullptr) return hipErrorInvalidValue;
  const ChainParams* params = a.params;
  ChainState* states = a.states;
  const ModelDev* models = a.models;
  int8_t* occ_pool = a.occ_pool;
  uint16_t* trk_pool = a.trk_pool;
  const int32_t* order = a.chain_order;
  WLDev* walkers = a.walkers;
  double* logdos = a.logdos;
  uint32_t* hist = a.hist;
  uint8_t* known = a.known;
  int64_t n_checks = a.n_checks, check_every = a.check_every;
  uint32_t* log = a.accept_log;
  int64_t log_words = a.log_words;
  const WLDev* shared = a.shared;
  int64_t log_base = a.log_base;
  void* args[] = {&params, &states, &models, &occ_pool, &trk_pool, &order, &walkers, &logdos, &hist, &known,
                  &n_checks, &check_every, &log, &log_words, &shared, &log_base};
  return hipLaunchKernel(fn, dim3(a.n_chains), dim3(64), args, a.lds_bytes, stream);
}

// The flatness test and modification-factor update of a shared histogram
// after every walker of the group made its share of a check's steps
// (wang_landau_sampler.cpp:326-394 with the inverse-time law of
// :344-367): `total` steps join the group's time.  One wave.
__global__ __launch_bounds__(64) void wl_shared_check_kernel(WLDev* __restrict__ leader,
                                                             uint32_t* __restrict__ hist_pool,
                                                             const uint8_t* __restrict__ known_pool,
                                                             int64_t total) {
  const int lane = int(threadIdx.x);
  if (leader->converged) return;
  const int nb = leader->n_bins;
  uint32_t* hist = hist_pool + leader->bin_off;
  const uint8_t* known = known_pool + leader->bin_off;
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
  const bool flat = __builtin_amdgcn_readfirstlane(int(double(mn) > leader->flatness * (double(sum) / double(cnt))));
  const int64_t steps = leader->steps + total, took = leader->since_flat + total;
  double f = leader->f, inv_factor = leader->inv_factor;
  int inv_active = leader->inv_active;
  if (flat) {
    for (int i = lane; i < nb; i += 64) hist[i] = 0u;
    const double mc_time = double(steps) / double(nb);
    if (leader->inverse_time && leader->prev_conv >= 0 && took > leader->prev_conv) {
      if (!inv_active) inv_factor = mc_time * f;
      inv_active = 1;
    }
    f = inv_active ? inv_factor / mc_time : f / 2.0;
  }
  if (lane == 0) {
    leader->steps = steps;
    leader->since_flat = flat ? 0 : took;
    if (flat) {
      leader->prev_conv = took;
      leader->n_flat += 1;
    }
    leader->inv_active = inv_active;
    leader->inv_factor = inv_factor;
    leader->f = f;
    if (f < leader->f_min) leader->converged = 1;
  }
}

hipError_t launch_wl_shared_check(WLDev* leader, uint32_t* hist, const uint8_t* known, int64_t total,
                                  hipStream_t stream) {
  hipLaunchKernelGGL(wl_shared_check_kernel, dim3(1), dim3(64), 0, stream, leader, hist, known, total);
  return hipGetLastError();
}

hipError_t launch_pt_exchange(PTDev* pt, ChainState* states, ChainParams* params, hipStream_t stream) {
  hipLaunchKernelGGL(pt_exchange_kernel, dim3(1), dim3(256), 0, stream, pt, states, params);
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

hipError_t launch_energy(const ModelDev* models, const ChainParams* params, const int8_t* occ_pool,
                         const int32_t* chain_ids, int n_chains, int max_sites, double* partial, double* energy_out,
                         hipStream_t stream) {
  const int n_blocks = (max_sites + 255) / 256;
  hipLaunchKernelGGL(energy_partial_kernel, dim3(n_blocks, n_chains), dim3(256), 0, stream, models, params,
                     occ_pool, chain_ids, partial, n_blocks);
  hipLaunchKernelGGL(energy_finalize_kernel, dim3(n_chains), dim3(64), 0, stream, models, params, chain_ids,
                     partial, n_blocks, energy_out);
  return hipGetLastError();
}

template <bool kInArgs, bool kOccLds, bool kPeriodic>
static hipError_t launch_updater_t(const UpdPlan& plan, int8_t* occ, double* ring, double* slot_energy,
                                   const UpdChanges& inl, const int32_t* changes, int n, const double* eci,
                                   UpdResult* result, uint64_t* result_cf, uint32_t lds_bytes,
                                   hipStream_t stream) {
  auto* fn = updater_kernel<kInArgs, kOccLds, kPeriodic>;
  if (lds_bytes > 64u * 1024u) {
    static std::mutex mu;
    static uint32_t limit = 64u * 1024u;
    std::lock_guard<std::mutex> lock(mu);
    if (lds_bytes > limit) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes));
      if (e != hipSuccess) return e;
      limit = lds_bytes;
    }
  }
  hipLaunchKernelGGL(fn, dim3(1), dim3(256), lds_bytes, stream, plan, occ, ring, slot_energy, inl, changes, n, eci,
                     result, result_cf);
  return hipGetLastError();
}

hipError_t launch_updater(const UpdPlan& plan, int8_t* occ, double* ring,
                          double* slot_energy, const UpdChanges& inl, const int32_t* changes, int n,
                          const double* eci, UpdResult* result, uint64_t* result_cf, uint32_t lds_bytes,
                          hipStream_t stream) {
  const bool a = n <= kUpdInline, o = plan.occ_lds != 0, p = plan.periodic != 0;
  auto go = [&](auto fn) {
    return fn(plan, occ, ring, slot_energy, inl, changes, n, eci, result, result_cf, lds_bytes, stream);
  };
  if (a) {
    if (o) return p ? go(launch_updater_t<true, true, true>) : go(launch_updater_t<true, true, false>);
    return p ? go(launch_updater_t<true, false, true>) : go(launch_updater_t<true, false, false>);
  }
  if (o) return p ? go(launch_updater_t<false, true, true>) : go(launch_updater_t<false, true, false>);
  return p ? go(launch_updater_t<false, false, true>) : go(launch_updater_t<false, false, false>);
}

template <bool kOccLds, bool kPeriodic>
static hipError_t launch_resident_t(const UpdPlan& plan, int8_t* occ, double* ring, double* slot_energy,
                                    const double* eci, UpdResult* result, uint64_t* result_cf, UpdMailbox* mb,
                                    int64_t next, int64_t idle_ticks, uint32_t lds_bytes, hipStream_t stream) {
  auto* fn = updater_resident_kernel<kOccLds, kPeriodic>;
  if (lds_bytes > 64u * 1024u) {
    static std::mutex mu;
    static uint32_t limit = 64u * 1024u;
    std::lock_guard<std::mutex> lock(mu);
    if (lds_bytes > limit) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes));
      if (e != hipSuccess) return e;
      limit = lds_bytes;
    }
  }
  hipLaunchKernelGGL(fn, dim3(1), dim3(256), lds_bytes, stream, plan, occ, ring, slot_energy, eci, result, result_cf,
                     mb, next, idle_ticks);
  return hipGetLastError();
}

hipError_t launch_updater_resident(const UpdPlan& plan, int8_t* occ, double* ring, double* slot_energy,
                                   const double* eci, UpdResult* result, uint64_t* result_cf, UpdMailbox* mb,
                                   int64_t next, int64_t idle_ticks, uint32_t lds_bytes, hipStream_t stream) {
  if (next < 0 || idle_ticks <= 0) return hipErrorInvalidValue;
  const bool o = plan.occ_lds != 0, p = plan.periodic != 0;
  auto go = [&](auto fn) {
    return fn(plan, occ, ring, slot_energy, eci, result, result_cf, mb, next, idle_ticks, lds_bytes, stream);
  };
  if (o) return p ? go(launch_resident_t<true, true>) : go(launch_resident_t<true, false>);
  return p ? go(launch_resident_t<false, true>) : go(launch_resident_t<false, false>);
}

hipError_t launch_restore(int8_t* occ, const InlineChanges& inl, const int32_t* buf, int n, hipStream_t stream) {
  hipLaunchKernelGGL(restore_kernel, dim3(1), dim3(64), 0, stream, occ, inl, buf, n);
  return hipGetLastError();
}

hipError_t launch_slot_energy(const ModelDev* model, const double* ring, double* slot_energy, const double* eci,
                              int slot, hipStream_t stream) {
  hipLaunchKernelGGL(slot_energy_kernel, dim3(1), dim3(64), 0, stream, model, ring, slot_energy, eci, slot);
  return hipGetLastError();
}

}  // namespace cemc

