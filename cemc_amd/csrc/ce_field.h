// Local-field kernel of the cluster-expansion Metropolis engine (gfx950).
//
// The energy change of flipping site j from species a to b is
//   dE = sum_d (phi_d(b) - phi_d(a)) * F_j^d,   F_j^d = sum_r G_r[d][code_r(j)]
// over the sub-cluster records r of j (the same tables the sweep kernel sums
// per trial move; ce_updater.cpp:313-406 is the reference's per-CF form).
// F_j depends only on the species of j's neighbours, so this kernel keeps
// every field in LDS and updates them when a move is accepted: for each
// record r of the flipped site and each member q of r, the record of that
// member holding the same sub-cluster (its dual) changes code, and the
// member's field changes by the difference of two table entries.  A trial
// move then costs one LDS read of its site's field; only accepted moves pay
// for gathers and table reads.  Fields are integers (the tables scaled by
// 2^k and rounded), so updates are exact, in any order, and never drift.
//
// Per 64-move block every lane evaluates its own move against the current
// state.  The first move (in order) that is accepted, or whose outcome needs
// more than its field (a swap of two interacting sites, or a Metropolis test
// within the rounding margin), is handled by the whole wave; every move
// before it is rejected.  After an accepted move all later lanes evaluate
// again.  The accept sequence is therefore the sequential one.
#pragma once

#include "ce_sweep.h"

namespace cemc {

namespace {

struct FieldLayout {
  uint32_t occ, trk, map, fld, gtab, total;
};

// occupation (padded box) | species tracker (canonical) | site of each
// padded cell | fields (n_sites x B integers) | integer energy tables
__host__ __device__ inline FieldLayout field_layout(int occ_bytes, int trk_sites, int n_fields, int gtab_ints) {
  FieldLayout L;
  L.occ = 0;
  L.trk = align16(uint32_t(occ_bytes));
  L.map = L.trk + align16(uint32_t(trk_sites) * 2u);
  L.fld = L.map + align16(uint32_t(occ_bytes) * 2u);
  L.gtab = L.fld + uint32_t(n_fields) * 8u;
  L.total = L.gtab + uint32_t(gtab_ints) * 8u;
  return L;
}

}  // namespace

// Changes of the fields when the site at padded index p flips from species
// so to sn: each lane walks its kR records; member q of record r is the
// site at p + m_q, whose dual record sees p at code position pos_i and the
// other members of r at the positions packed in the dual word.  kSink:
// add the changes to the fields (accepted flip) or, for a swap, only
// accumulate the change of the field of site `probe` into `acc` (the second
// flip of the swap sees the first already made).
template <int kR, int kBM, bool kSink>
__device__ __attribute__((always_inline)) inline void field_flip(
    const FatRec (&rr)[kR], const uint32_t (&du)[kR][3], int p, uint32_t so, uint32_t sn, const uint8_t* occ,
    const uint16_t* map, unsigned long long* fld, const int64_t* gfx, int B, int probe, int64_t (&acc)[kBM]) {
  constexpr int kBits = TableGeom<kBM>::kBits;
  constexpr int kCodes = TableGeom<kBM>::kCodes;
  // branch-free: every (record, member) slot issues its loads and its
  // update; slots past a record's members (and padding records) add zero to
  // the field of the flipped site itself (their member offset is 0)
  uint32_t s[kR][3];
  int site[kR][3];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int mo[3] = {rr[k].m0, rr[k].m1, rr[k].m2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      s[k][q] = occ[p + mo[q]];
      site[k][q] = int(map[p + mo[q]]);
    }
  }
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int nm = int((rr[k].meta >> 24) & 3u);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const uint32_t w = du[k][q];
      uint32_t rest = 0;
      int t = 0;
#pragma unroll
      for (int q2 = 0; q2 < 3; ++q2) {
        if (q2 == q) continue;
        const uint32_t sh = kBits * ((w >> (26 + 2 * t)) & 3u);
        rest |= q2 < nm ? s[k][q2] << sh : 0u;
        ++t;
      }
      const uint32_t pi = kBits * ((w >> 24) & 3u);
      const int base = int(w & 0xFFFFFFu);
      const int c_old = base + int(rest | (so << pi));
      const int c_new = base + int(rest | (sn << pi));
      const bool live = q < nm && (!kSink || site[k][q] == probe);
#pragma unroll
      for (int d = 0; d < kBM; ++d) {
        if (d >= B) break;
        const int64_t g = gfx[c_new + d * kCodes] - gfx[c_old + d * kCodes];
        const int64_t delta = live ? g : 0;
        if constexpr (kSink)
          acc[d] += delta;
        else
          atomicAdd(&fld[site[k][q] * B + d], (unsigned long long)delta);
      }
    }
  }
}

template <bool kCanonical, int kR, int kBM>
__global__ __launch_bounds__(64) void field_kernel(const ChainParams* __restrict__ params,
                                                   ChainState* __restrict__ states,
                                                   const ModelDev* __restrict__ models,
                                                   int8_t* __restrict__ occ_pool, uint16_t* __restrict__ trk_pool,
                                                   const int32_t* __restrict__ chain_order,
                                                   const int64_t* __restrict__ n_moves,
                                                   uint32_t* __restrict__ accept_log, int64_t log_words,
                                                   HookDev* __restrict__ /*hooks: not on this path*/) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ double dphi_tab[kMaxSpecies * kMaxSpecies * kMaxBasis];  // [(old*S+new)*B + d]
  __shared__ double dsingle_tab[kMaxSpecies * kMaxSpecies];          // singlet energy change per (old,new)
  __shared__ double phi_tab[kMaxBasis * kMaxSpecies];                // phi[d*S + s]
  __shared__ unsigned long long probe_sum[kMaxBasis];                // swap of interacting sites
  constexpr int kCodes = TableGeom<kBM>::kCodes;
  const int chain = chain_order[blockIdx.x];
  const int lane = int(threadIdx.x);
  const ChainParams* __restrict__ pp = params + chain;
  const ModelDev& m = models[pp->model];
  const int N = pp->n_sites;
  const int B = m.n_bfs;
  const int occ_bytes = m.pdims[0] * m.pdims[1] * m.pdims[2];
  const FieldLayout L = field_layout(occ_bytes, kCanonical ? N : 0, N * B, m.gtab_size);
  uint8_t* occ = reinterpret_cast<uint8_t*>(lds + L.occ);
  uint16_t* trk = reinterpret_cast<uint16_t*>(lds + L.trk);
  uint16_t* map = reinterpret_cast<uint16_t*>(lds + L.map);
  unsigned long long* fld = reinterpret_cast<unsigned long long*>(lds + L.fld);
  int64_t* gfx = reinterpret_cast<int64_t*>(lds + L.gtab);

  SweepCtx c;
  c.S = m.n_species;
  c.B = B;
  c.N = N;
  c.L1 = m.dims[0];
  c.L2 = m.dims[1];
  c.L3 = m.dims[2];
  c.L23 = c.L2 * c.L3;
  c.h = m.pad_h;
  c.P2 = m.pdims[1];
  c.P3 = m.pdims[2];
  c.mag23 = 0xFFFFFFFFu / uint32_t(c.L23);
  c.mag3 = 0xFFFFFFFFu / uint32_t(c.L3);
  c.DX = c.L1 * c.P2 * c.P3;
  c.DY = c.L2 * c.P3;
  c.DZ = c.L3;
  c.K = m.K;
  c.trans = m.trans;

  // ---- stage the lattice (padded box and the site of every cell), the
  // tracker, the integer tables and the species tables
  const uint8_t* gocc = reinterpret_cast<const uint8_t*>(occ_pool + pp->occ_off);
  {
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
      const int site = (x * c.L2 + y) * c.L3 + z;
      occ[p] = gocc[site];
      map[p] = uint16_t(site);
    }
  }
  uint16_t* gtrk = trk_pool + pp->trk_off;
  if constexpr (kCanonical)
    for (int i = lane; i < N; i += 64) trk[i] = gtrk[i];
  for (int i = lane; i < m.gtab_size; i += 64) gfx[i] = m.gtab_fx[i];
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

  // ---- records (and their duals) of the lane, for the whole launch
  const int cnt = m.grp_rec_cnt[0];
  FatRec rr[kR];
  uint32_t du[kR][3];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int it = lane + 64 * k;
    rr[k] = it < cnt ? m.frecs[it] : FatRec{0, 0, 0, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
    for (int q = 0; q < 3; ++q) du[k][q] = it < cnt ? m.duals[3 * it + q] : 0u;
  }

  // ---- fields from scratch: the records travel through the wave (lane l
  // holds records l, l + 64, ...; each is broadcast with readlane) while
  // lane j sums them for sites j, j + 64, ...
  for (int j0 = 0; j0 < N; j0 += 64) {
    const int j = j0 + lane < N ? j0 + lane : N - 1;
    const Ref r = ref_of<true>(c, j);
    int64_t f[kBM];
#pragma unroll
    for (int d = 0; d < kBM; ++d) f[d] = 0;
#pragma unroll
    for (int k = 0; k < kR; ++k) {
      const int n_here = cnt - 64 * k < 64 ? cnt - 64 * k : 64;
      for (int l = 0; l < n_here; ++l) {
        FatRec one[1];
        one[0].m0 = __builtin_amdgcn_readlane(rr[k].m0, l);
        one[0].m1 = __builtin_amdgcn_readlane(rr[k].m1, l);
        one[0].m2 = __builtin_amdgcn_readlane(rr[k].m2, l);
        one[0].meta = __builtin_amdgcn_readlane(rr[k].meta, l);
        int gi[1];
        table_index<true, 1, kBM, false>(c, one, r, occ, 0u, 0u, gi);
#pragma unroll
        for (int d = 0; d < kBM; ++d)
          if (d < B) f[d] += gfx[gi[0] + d * kCodes];
      }
    }
    if (j0 + lane < N)
#pragma unroll
      for (int d = 0; d < kBM; ++d)
        if (d < B) fld[j * B + d] = (unsigned long long)f[d];
  }
  __syncthreads();

  const double dphi01 = dphi_tab[1], dphi10 = dphi_tab[c.S];
  const double dsin01 = dsingle_tab[1], dsin10 = dsingle_tab[c.S];
  const double scale = m.fx_scale;
  double energy = states[chain].energy;
  const uint64_t t0 = states[chain].moves_done;
  const double e_ref = states[chain].e_ref;
  const int64_t nmov = n_moves[chain];
  uint64_t n_acc = 0;
  double sum_de = 0.0, sum_de2 = 0.0;
  const double inv_n = 1.0 / double(N);
  double sing = 0.0, sum_s = 0.0, sum_s2 = 0.0, sum_se = 0.0;
  if constexpr (!kCanonical) {
    double part_s[kBM];
#pragma unroll
    for (int d = 0; d < kBM; ++d) part_s[d] = 0.0;
    for (int i = lane; i < N; i += 64) {
      const int s = occ[ref_of<true>(c, i).p];
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
  const uint64_t im0 = m.imask[0], im1 = m.imask[1];
  uint32_t* acc_log = accept_log ? accept_log + int64_t(chain) * log_words : nullptr;

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
  // energy change of a flip a -> b of a site with fields fd (scaled)
  auto flip_de = [&](uint32_t a, uint32_t b, const double (&fd)[kBM]) {
    double v;
    if constexpr (kBM == 1) {
      v = (a == 0 ? dphi01 : dphi10) * fd[0];
    } else {
      const int pair = int(a) * c.S + int(b);
      v = 0.0;
#pragma unroll
      for (int d = 0; d < kBM; ++d)
        if (d < c.B) v = fma(dphi_tab[pair * c.B + d], fd[d], v);
    }
    return v;
  };
  auto read_fields = [&](int site, double (&fd)[kBM]) {
#pragma unroll
    for (int d = 0; d < kBM; ++d)
      fd[d] = d < c.B ? double(int64_t(fld[site * B + d])) * scale : 0.0;
  };

  for (int64_t blk0 = 0; blk0 < nmov; blk0 += 64) {
    const Philox4 blk = philox4x64_10(t0 + uint64_t(blk0) + uint64_t(lane), 0, 0, 0, seed, uid);
    const double log_u = log(unit_double(blk.w3));
    const int n_in = int(nmov - blk0 < 64 ? nmov - blk0 : 64);
    uint32_t v_old = 0, v_new = 0;
    int v_ta = 0, v_tb = 0;
    int v_site = 0, v_p = 0;
    uint32_t v_img = 0;
    if constexpr (kCanonical) {
      const uint32_t ia = bounded32(uint32_t(blk.w0 >> 32), np);
      uint32_t ib = ia + 1u + bounded32(uint32_t(blk.w0), np - 1u);
      ib -= ib >= np ? np : 0u;
      v_old = (present_pk >> (4u * ia)) & 15u;
      v_new = (present_pk >> (4u * ib)) & 15u;
      v_ta = unpack16(seg_lo, seg_hi, ia) + int(bounded64(blk.w1, uint64_t(unpack16(cnt_lo, cnt_hi, ia))));
      v_tb = unpack16(seg_lo, seg_hi, ib) + int(bounded64(blk.w2, uint64_t(unpack16(cnt_lo, cnt_hi, ib))));
    } else {
      v_site = int(bounded64(blk.w1, uint64_t(N)));
      const Ref r = ref_of<true>(c, v_site);
      v_p = r.p;
      v_img = image_word<true>(c, r);
      v_new = bounded32(uint32_t(blk.w0), uint32_t(c.S - 1));
    }
    uint64_t acc_bits = 0;
    int cursor = 0;
    while (cursor < n_in) {
      // every lane: its move against the current state
      uint32_t o_v = v_old, n_v = v_new;
      int si = v_site, sj = 0;
      Ref ri{}, rj{};
      double dE;
      bool pair_hit = false;  // canonical: the two sites interact
      if constexpr (kCanonical) {
        si = trk[v_ta];
        sj = trk[v_tb];
        ri = ref_of<true>(c, si);
        rj = ref_of<true>(c, sj);
        double fi[kBM], fj[kBM];
        read_fields(si, fi);
        read_fields(sj, fj);
        dE = flip_de(o_v, n_v, fi) + flip_de(n_v, o_v, fj);
        const uint32_t ci = uint32_t(ri.x) | (uint32_t(ri.y) << 8) | (uint32_t(ri.z) << 16);
        const uint32_t cj = uint32_t(rj.x) | (uint32_t(rj.y) << 8) | (uint32_t(rj.z) << 16);
        pair_hit = cells_interact(ci, cj, c, im0, im1);
      } else {
        o_v = occ[v_p];
        n_v = o_v + 1u + v_new;
        n_v -= n_v >= uint32_t(c.S) ? uint32_t(c.S) : 0u;
        double fi[kBM];
        read_fields(v_site, fi);
        double de_single;
        if constexpr (kBM == 1)
          de_single = o_v == 0 ? dsin01 : dsin10;
        else
          de_single = dsingle_tab[o_v * c.S + n_v];
        dE = flip_de(o_v, n_v, fi) + de_single;
      }
      const double x = -dE * inv_kT;
      const double margin = 1e-12 * (1.0 + fabs(x));
      // 0 reject, 1 accept, 2 within the margin, 3 swap of interacting sites
      const int verdict =
          pair_hit ? 3 : ((dE < 0.0 || log_u < x - margin) ? 1 : (log_u > x + margin ? 0 : 2));
      uint64_t todo = __builtin_amdgcn_ballot_w64(lane >= cursor && lane < n_in && verdict != 0);
      bool changed = false;
      while (todo != 0 && !changed) {
        const int f = int(__builtin_ctzll(todo));
        todo &= todo - 1u;
        run += f - cursor;  // the moves before f are rejected
        cursor = f;
        int v = __builtin_amdgcn_readlane(verdict, f);
        double dEf = readlane_f64(dE, f);
        const uint32_t of = __builtin_amdgcn_readlane(o_v, f), nf = __builtin_amdgcn_readlane(n_v, f);
        Ref fi, fj;
        fi.site = __builtin_amdgcn_readlane(si, f);
        fi.x = fi.y = fi.z = 0;
        if constexpr (kCanonical)
          fi.p = __builtin_amdgcn_readlane(ri.p, f);
        else
          fi.p = __builtin_amdgcn_readlane(v_p, f);
        fj = fi;
        if constexpr (kCanonical) {
          fj.site = __builtin_amdgcn_readlane(sj, f);
          fj.p = __builtin_amdgcn_readlane(rj.p, f);
        }
        int64_t unused[kBM];
        if (v == 3) {
          // second flip of the swap sees the first: add the change of site
          // j's field from site i's flip (summed over the wave) to F_j
          if constexpr (kCanonical) {
            if (lane < kBM) probe_sum[lane] = 0ull;
            int64_t part[kBM];
#pragma unroll
            for (int d = 0; d < kBM; ++d) part[d] = 0;
            field_flip<kR, kBM, true>(rr, du, fi.p, of, nf, occ, map, fld, gfx, B, fj.site, part);
#pragma unroll
            for (int d = 0; d < kBM; ++d)
              if (d < B && part[d] != 0) atomicAdd(&probe_sum[d], (unsigned long long)part[d]);
            double fdi[kBM], fdj[kBM];
            read_fields(fi.site, fdi);
#pragma unroll
            for (int d = 0; d < kBM; ++d)
              fdj[d] = d < B ? double(int64_t(fld[fj.site * B + d] + probe_sum[d])) * scale : 0.0;
            dEf = flip_de(of, nf, fdi) + flip_de(nf, of, fdj);
            const double xf = -dEf * inv_kT;
            const double mf = 1e-12 * (1.0 + fabs(xf));
            const double luf = readlane_f64(log_u, f);
            v = (dEf < 0.0 || luf < xf - mf) ? 1 : (luf > xf + mf ? 0 : 2);
          }
        }
        if (v == 2) v = int(unit_double(readlane_u64(blk.w3, f)) <= exp(-dEf / kT));
        if (v == 1) {
          flush();
          // each flipped site (i, then j of a swap with i already flipped):
          // the fields of its partners, then its images; one copy of the
          // update code serves both flips
#pragma unroll 1
          for (int fl = 0; fl < (kCanonical ? 2 : 1); ++fl) {
            const int p_fl = fl == 0 ? fi.p : fj.p;
            const int site_fl = fl == 0 ? fi.site : fj.site;
            const uint32_t so = fl == 0 ? of : nf, sn = fl == 0 ? nf : of;
            field_flip<kR, kBM, false>(rr, du, p_fl, so, sn, occ, map, fld, gfx, B, 0, unused);
            uint32_t img;
            if constexpr (kCanonical)
              img = image_word<true>(c, ref_of<true>(c, site_fl));
            else
              img = __builtin_amdgcn_readlane(v_img, f);
            const int idx = image_of(c, img, lane & 7);
            if (lane < 8 && idx >= 0) occ[idx] = uint8_t(sn);
          }
          if constexpr (kCanonical) {
            if (lane == 0) {
              trk[__builtin_amdgcn_readlane(v_ta, f)] = uint16_t(fj.site);
              trk[__builtin_amdgcn_readlane(v_tb, f)] = uint16_t(fi.site);
            }
          }
          energy += dEf;
          if constexpr (!kCanonical) {
            double dsing;
            if constexpr (kBM == 1)
              dsing = __dmul_rn(of == 0 ? dphi01 : dphi10, inv_n);
            else
              dsing = __dmul_rn(dphi_tab[(int(of) * c.S + int(nf)) * c.B + (lane < c.B ? lane : 0)], inv_n);
            sing += dsing;
          }
          acc_bits |= uint64_t(1) << f;
          changed = true;
        }
        run += 1;
        cursor = f + 1;
      }
      if (!changed) {
        run += n_in - cursor;  // every remaining move is rejected
        cursor = n_in;
      }
    }
    n_acc += uint64_t(__builtin_popcountll(acc_bits));
    if (acc_log != nullptr && lane < 2) {
      const int64_t w = blk0 >> 5;
      if (lane == 0) acc_log[w] = uint32_t(acc_bits);
      if (lane == 1 && n_in > 32) acc_log[w + 1] = uint32_t(acc_bits >> 32);
    }
  }
  flush();

  // ---- write back
  uint8_t* wocc = reinterpret_cast<uint8_t*>(occ_pool + pp->occ_off);
  for (int i = lane; i < N; i += 64) wocc[i] = occ[ref_of<true>(c, i).p];
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

}  // namespace cemc
