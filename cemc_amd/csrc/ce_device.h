// Device-side building blocks of the cluster-expansion Metropolis engine for
// CDNA4 (gfx950): data layouts shared by host and kernels, the Philox4x64-10
// counter RNG, 64-lane DPP reductions and lattice-neighbour evaluation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cemc {

constexpr int kMaxSpecies = 8;           // species per model
constexpr int kMaxBasis = kMaxSpecies - 1;
constexpr int kMaxGroups = 32;           // translation-symmetry groups per model (2 x this <= 64 lanes)
constexpr int kMaxClusterSize = 4;       // reference limit (cpp/src/cluster.cpp:131-170)
constexpr double kBoltzmannEV = 8.617330337217213e-05;

// ---------------------------------------------------------------------------
// Sub-cluster record: one cluster containing the reference site, as listed in
// the reference's `indices`/`order` for that site (cpp/src/cluster.cpp:185-232).
//   w0: slot of member 0..2 (10 bits each, canonical order with the reference
//       site removed) | (size-2) << 30
//   w1: energy-table base (24 bits) | canonical position of the reference
//       site << 24 (the record's cluster entry is implied by the per-entry
//       record ranges)
// Slot = column index into the translation table (used columns only).
struct SubRec {
  uint32_t w0;
  uint32_t w1;
};

__host__ __device__ inline uint32_t rec_slot(SubRec r, int q) { return (r.w0 >> (10 * q)) & 1023u; }
__host__ __device__ inline int rec_size(SubRec r) { return int(r.w0 >> 30) + 2; }
__host__ __device__ inline uint32_t rec_gbase(SubRec r) { return r.w1 & 0xFFFFFFu; }
__host__ __device__ inline int rec_refpos(SubRec r) { return int((r.w1 >> 24) & 3u); }

// Sub-cluster record of the sweep kernel (32 bytes, held in VGPRs or LDS):
//   m[q]   members in canonical order with the reference site removed:
//          closed-form periodic lattices: signed offset of the member in the
//          padded (halo) lattice, index(member) = index(reference) + m[q];
//          otherwise: translation-table column.
//   w[q]   periodic lattices: member cell offset wrapped into the supercell,
//          packed x | y<<8 | z<<16 (identifies the member as a site).
//   meta   energy-table base (24 bits) | members << 24 (2 bits) |
//          member 1 unused << 28 | member 2 unused << 29 | valid << 31
// Unused members address the reference site (offset 0); the energy table
// repeats over their species field, so no masking is needed.  Energy table of
// a cluster entry: index base_r + d * 2^(3k) + code, with base_r in meta for
// the record's reference position, d the decoration slice and code the
// species of the other members, k bits each.
struct FatRec {
  int32_t m0, m1, m2;
  uint32_t w0, w1, w2;
  uint32_t meta;
  uint32_t pad;
};
constexpr uint32_t kRecUnused1 = 1u << 28;
constexpr uint32_t kRecUnused2 = 1u << 29;
constexpr uint32_t kRecValid = 1u << 31;

// Cluster entry = one (translation group, orbit) pair with its record range.
struct ClusterEntry {
  int32_t rec_begin;   // into the model's record array
  int32_t rec_count;   // M (sub-clusters for the group's reference site)
  int32_t size;        // n
  int32_t norm;        // C_p = sum over groups of M (ce_updater.cpp:137-144)
};

// Device view of one model (lattice + clusters + tables), in global memory.
struct ModelDev {
  int32_t n_sites, n_species, n_bfs, n_groups;
  int32_t n_cf, n_entries, K, periodic;
  int32_t dims[3];
  int32_t pad_h;          // halo width of the padded lattice (periodic only)
  int32_t pdims[3];       // padded dims L + 2*pad_h
  int32_t has_background;
  int32_t n_movable;
  // Speculative rounds (closed-form periodic, one translation group, no
  // background, halo h <= 2): bit ((dx+h)(2h+1) + (dy+h))(2h+1) + (dz+h) of
  // imask is set when a site at cell offset (dx, dy, dz) from a reference
  // site is that site or a member of one of its sub-clusters.
  int32_t spec_ok;
  int32_t pad_spec;
  uint64_t imask[2];
  int32_t gtab_size;      // doubles in the energy table
  int32_t n_rec;          // records over all groups
  int32_t pow_s[kMaxClusterSize];  // n_species^q
  int32_t grp_rec_off[kMaxGroups];  // energy-path records of group g: [off, off+cnt)
  int32_t grp_rec_cnt[kMaxGroups];
  int32_t grp_count[kMaxGroups];    // T_g
  double basis[kMaxBasis * kMaxSpecies];  // phi[d*n_species + s]
  const SubRec* recs;
  const FatRec* frecs;        // same order as recs
  const double* gtab;
  // three and four species (two and three slices read per record): the same
  // tables with the read slices of each code side by side, [block][code][d]
  // (blocks keep their offsets), so one 16-byte read fetches both values of
  // three species and the three values of four species share one or two
  // cache lines; the other models point this at gtab
  const double* gtab_il;
  const int32_t* trans;       // n_sites * K
  const uint32_t* col_cells;  // K packed cell offsets (x | y<<8 | z<<16)
  const int8_t* site_group;   // n_sites translation group (-1: none, background only)
  const uint8_t* background;  // n_sites, 1 = background site (never moved)
  const int32_t* movable;     // non-background sites (only read with background)
  // correlation-function tables
  const int32_t* cf_kind;     // n_cf: 0 c0, 1 singlet, 2 cluster
  const int32_t* cf_dec;      // n_cf
  const int32_t* cf_entry;    // n_groups * n_cf (cluster entry or -1)
  const int32_t* cf_eq_off;   // n_cf + 1
  const uint32_t* eq_decos;   // packed 4 x 8-bit decoration per equivalent decoration
  const ClusterEntry* entries;
  const double* eci;          // n_cf
  // Local-field path (closed-form periodic, one translation group, no
  // background; every sub-cluster seen from each of its members is a record
  // of that member).  gtab_fx = round(gtab * 2^k) as integers, fx_scale =
  // 2^-k with k chosen so that a site's field sum stays below 2^61.  duals[3r
  // + q]: for member q of record r, the record of that member holding the
  // same sub-cluster: bits 0-23 its table base, 24-25 the code position of
  // the reference site of r, 26-27 and 28-29 those of the other members of r
  // in order (q skipped).
  int32_t field_ok;
  int32_t pad_field;
  double fx_scale;
  const int64_t* gtab_fx;
  const uint32_t* duals;
};

// Per-chain parameters (host written).
struct ChainParams {
  int32_t model;
  int32_t kind;         // 0 canonical swaps, 1 semi-grand-canonical flips
  int32_t n_sites;
  int32_t n_present;
  int64_t occ_off;      // byte offset into the occupation pool
  int64_t trk_off;      // element offset into the tracker pool
  int32_t present[kMaxSpecies];
  int32_t count[kMaxSpecies];
  int32_t seg_off[kMaxSpecies];
  // the same species tables packed for register use: present ids (4 bits
  // each), counts and tracker segment offsets (16 bits each, entries 0-3 in
  // word 0, 4-7 in word 1)
  uint32_t present_pk;
  uint32_t pad0;
  uint64_t count_pk[2];
  uint64_t segoff_pk[2];
  double kT;
  double j1[kMaxBasis];  // singlet ECIs (chemical potential folded in)
  uint64_t seed;
  uint64_t uid;
};

// Per-chain mutable state (kernel written).
// Per-chain mutable state.  The sample sums run over the trial moves since
// the last reset; each move samples the state after its accept/reject, as the
// reference's observers do (cemc/mcmc/mc_observers.py:222-270).  Energies
// enter relative to e_ref, the energy at that reset, which keeps the squared
// sums well conditioned (the purpose of the reference's energy-bias probe,
// cemc/mcmc/montecarlo.py:178-217).  Singlet sums are kept for SGC chains.
struct ChainState {
  double energy;
  uint64_t moves_done;  // moves ever made: the Philox counter of the next move
  uint64_t n_accept;
  uint64_t n_samples;   // moves since the last reset
  double e_ref;
  double sum_de;        // sum of (E - e_ref)
  double sum_de2;       // sum of (E - e_ref)^2
  double pad0;
  double sum_s[kMaxBasis];   // sum of singlet d
  double sum_s2[kMaxBasis];  // sum of singlet d squared
  double sum_se[kMaxBasis];  // sum of singlet d * (E - e_ref)
  double pad1[3];
};
static_assert(sizeof(ChainState) == 256, "ChainState layout");

// Wang-Landau walker on a canonical chain (wang_landau_sampler.cpp:232-404,
// histogram.cpp:66-172): one histogram per walker, ln g, visit counts and
// known flags of its bins at bin_off in the batch's bin pools.  Counters
// accumulate over launches.
struct WLDev {
  double e_min, e_max;  // histogram range (total energy, eV)
  double f;             // modification factor added to ln g per visit
  double f_min;         // converged once f < f_min
  double flatness;      // flat when min(count) > flatness * mean(count) over known bins
  int32_t n_bins;
  int32_t cur_bin;      // bin of the walker's state
  int32_t is_first;     // the next in-range proposal is accepted whatever ln g says
  int32_t converged;
  int64_t bin_off;
  uint64_t n_moves, n_accept, n_outside, n_self, n_flat;
  // schemes of WangLandauSampler::run (wang_landau_sampler.cpp:319-404)
  int32_t inverse_time;  // use_inverse_time_algorithm
  int32_t adaptive;      // AdaptiveWindowHistogram (adaptive_windows.cpp)
  int32_t min_width;     // minimum_window_width
  int32_t inv_active;    // the inverse-time law is in force
  int32_t upper;         // bins [0, upper) are in range
  int32_t n_overlap;     // bins kept above a cut
  int32_t n_edges;       // cuts made since the last flat window
  int32_t has_first;     // the state at the first cut is recorded
  double inv_factor;     // f t at the activation of the inverse-time law
  double bin_change;     // adaptive: sum of |proposed bin - current bin| since the last flat histogram
  double cur_energy;     // energy of the last accepted state (the sub-bin energy)
  int64_t steps;         // steps since the scheme was set (the time t = steps / n_bins)
  int64_t since_flat;    // steps since the last flat histogram
  int64_t prev_conv;     // steps the previous flat histogram took (-1: none yet)
  uint64_t n_walk;       // moves made to bring the walker back into its window
  int32_t status;        // 1: no state inside the window was found
  int32_t pad;
  // adaptive windows: the walker's own device arrays (null otherwise)
  uint8_t* snap_occ;     // (n_bins + 1) x n_sites recorded lattices (row n_bins: the first cut)
  uint16_t* snap_trk;    // (n_bins + 1) x n_sites species trackers
  double* snap_e;        // n_bins + 1 energies
  int32_t* snap_bin;     // n_bins + 1 bins
  int32_t* edges;        // n_bins cut positions
  int32_t* edge_n;       // n_bins saved value counts
  double* edge_vals;     // n_bins x n_bins saved ln g above each cut
  uint8_t* saved;        // n_bins: a state was recorded at the bin's first visit
  uint32_t* sub;         // n_bins x 10 sub-bin visit counts
};
static_assert(sizeof(WLDev) == 272, "WLDev layout");

// Per-move predicates and bias of one chain (cemc/mcmc/mc_constraints.py,
// cpp/src/pair_constraint.cpp, cemc/mcmc/bias_potential.py): a trial move
// is redrawn while it violates a constraint, up to max_attempts proposals
// (montecarlo.py:970-991), and the bias of the proposed state is added to
// its energy (montecarlo.py:926-930).
struct HookDev {
  uint32_t fixed_species;  // bit s: a change from or to species s is invalid (FixedElement)
  int32_t max_attempts;    // proposals per move
  const uint8_t* allowed;  // per site, bit s: species s may occupy it (nullptr: any)
  int32_t pair_e1, pair_e2;  // PairConstraint species (-1: none): no e1 site with an e2 member
  int32_t pair_nm, pad0;     // members of the pair cluster per site (<= 32, symmetric lists)
  const int32_t* pair_nb;    // n_sites x pair_nm member sites
  int64_t pair_count;        // ordered (e1, e2) member pairs of the current state
  int32_t bias_species;      // species whose count the bias depends on (-1: none)
  int32_t bias_count;        // its count in the current state
  const double* bias_tab;    // n_sites + 1 entries: bias energy at each count
  double bias_energy;        // bias of the current state (part of ChainState::energy)
  int32_t status;            // 1: a move found no valid proposal within max_attempts
  int32_t recycle;           // 1: waste recycling (Frenkel, PNAS 101, 17571 (2004)): heat-bath
                             // acceptance and samples weighting current and proposed state
                             // (montecarlo.py:937-948, 805-811; util.py:41-63)
  uint64_t failed_move;      // Philox counter of that move
};

// Replica-exchange state of a parallel-tempering run over the R chains of a
// batch.  Slot t keeps temperature kT[t] and its own energy sums; the chain
// holding slot t is chain_of_slot[t].  Exchanges move chains between slots
// (their kT in ChainParams follows), lattices never move.
struct PTDev {
  uint64_t cycle;        // exchange cycles done: Philox counter word 0
  uint64_t seed, uid;    // Philox key of the exchange stream
  int32_t n_rep;
  int32_t pad;
  int32_t* chain_of_slot;  // R
  int32_t* slot_of_chain;  // R
  double* kT;              // R, by slot
  double* e_ref;           // R, by slot: reference of the slot's energy sums
  double* sum_de;          // R
  double* sum_de2;         // R
  uint64_t* n_samples;     // R
  uint64_t* n_accept;      // R-1, pair (t, t+1)
  uint64_t* n_try;         // R-1
};

// ---------------------------------------------------------------------------
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11).
constexpr uint64_t kPhiloxM0 = 0xD2E7470EE14C6C93ull;
constexpr uint64_t kPhiloxM1 = 0xCA5A826395121157ull;
constexpr uint64_t kPhiloxW0 = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kPhiloxW1 = 0xBB67AE8584CAA73Bull;

struct Philox4 {
  uint64_t w0, w1, w2, w3;
};

__device__ inline Philox4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0,
                                        uint64_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    const uint64_t lo0 = kPhiloxM0 * c0;
    const uint64_t hi0 = __umul64hi(kPhiloxM0, c0);
    const uint64_t lo1 = kPhiloxM1 * c2;
    const uint64_t hi1 = __umul64hi(kPhiloxM1, c2);
    const uint64_t n0 = hi1 ^ c1 ^ k0;
    const uint64_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return Philox4{c0, c1, c2, c3};
}

__device__ inline uint64_t bounded64(uint64_t w, uint64_t n) { return __umul64hi(w, n); }
__device__ inline uint32_t bounded32(uint32_t w, uint32_t n) { return __umulhi(w, n); }
__device__ inline double unit_double(uint64_t w) { return double(w >> 11) * 0x1.0p-53; }

// ---------------------------------------------------------------------------
// Cross-lane helpers (wave64).
__device__ inline uint64_t readlane_u64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), lane);
  return (uint64_t(hi) << 32) | lo;
}

__device__ inline double readlane_f64(double v, int lane) {
  return __longlong_as_double(int64_t(readlane_u64(uint64_t(__double_as_longlong(v)), lane)));
}

template <int kCtrl>
__device__ inline double dpp_f64(double v) {
  const int64_t bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp(int(uint32_t(bits)), kCtrl, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(int(uint32_t(uint64_t(bits) >> 32)), kCtrl, 0xF, 0xF, true);
  return __longlong_as_double(int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo)));
}

// Lane pairs of a wave sum, in the order they are added: lanes 32 apart,
// 16 apart, 8 apart, then 1, 2 and 4 apart.  Every reduction of the sweep
// kernel follows this one tree, so a sum taken alone (wave_sum_f64) and the
// same sum taken among eight (wave_sum8_f64) are bitwise equal.

__device__ inline double join_f64(uint32_t lo, uint32_t hi) {
  return __longlong_as_double(int64_t((uint64_t(hi) << 32) | lo));
}
__device__ inline uint32_t lo_of(double v) { return uint32_t(uint64_t(__double_as_longlong(v))); }
__device__ inline uint32_t hi_of(double v) { return uint32_t(uint64_t(__double_as_longlong(v)) >> 32); }

// x + (x of the lane 32 apart) in lanes 0-31 and y + (y of the lane 32
// apart) in lanes 32-63 (v_permlane32_swap exchanges the upper half of x
// with the lower half of y).
__device__ inline double pair32_f64(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane32_swap(lo_of(x), lo_of(y), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(hi_of(x), hi_of(y), false, false);
  return join_f64(lo[0], hi[0]) + join_f64(lo[1], hi[1]);
}

// x + (x of the lane 16 apart) in rows 0 and 2, y + (y 16 apart) in rows 1
// and 3 (v_permlane16_swap exchanges odd rows of x with even rows of y).
__device__ inline double pair16_f64(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane16_swap(lo_of(x), lo_of(y), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(hi_of(x), hi_of(y), false, false);
  return join_f64(lo[0], hi[0]) + join_f64(lo[1], hi[1]);
}

// Last three levels, within each group of 8 lanes: pairs 1 and 2 apart
// (quad_perm [1,0,3,2], [2,3,0,1]), then 4 apart (row_half_mirror: the quads
// are uniform by then).
__device__ inline double octet_sum_f64(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  return v;
}

// Sum over all 64 lanes; every lane ends with the same total.
__device__ inline double wave_sum_f64(double v) {
  v = pair32_f64(v, v);
  v = pair16_f64(v, v);
  v += dpp_f64<0x128>(v);  // row_ror:8, the lane 8 apart within the row
  return octet_sum_f64(v);
}

// Eight wave sums at once: lane L ends with the wave sum of x[L >> 3], added
// in the order of wave_sum_f64.  Each level halves the values a lane holds
// (8 -> 4 -> 2 -> 1), so the eight sums cost ten additions per lane.
__device__ inline double wave_sum8_f64(const double (&x)[8], int lane) {
  double q[4], r[2];
#pragma unroll
  for (int j = 0; j < 4; ++j) q[j] = pair32_f64(x[j], x[j + 4]);  // lane bit 5 picks j or j + 4
#pragma unroll
  for (int j = 0; j < 2; ++j) r[j] = pair16_f64(q[j], q[j + 2]);  // lane bit 4 picks j or j + 2
  const bool b3 = (lane & 8) != 0;                                // lane bit 3 picks r[0] or r[1]
  const double keep = b3 ? r[1] : r[0];
  const double send = b3 ? r[0] : r[1];
  return octet_sum_f64(keep + dpp_f64<0x128>(send));
}

// ---------------------------------------------------------------------------
// Lattice neighbours: site playing the role of column `slot` when `ref` is the
// reference site (the translation matrix, ce_updater.cpp:264).
struct RefSite {
  int32_t site;
  int32_t x, y, z;  // cell coordinates (closed-form periodic lattices only)
};

__device__ inline RefSite make_ref(const ModelDev& m, int32_t site) {
  RefSite r;
  r.site = site;
  if (m.periodic) {
    const int32_t l23 = m.dims[1] * m.dims[2];
    r.x = site / l23;
    const int32_t rem = site - r.x * l23;
    r.y = rem / m.dims[2];
    r.z = rem - r.y * m.dims[2];
  } else {
    r.x = r.y = r.z = 0;
  }
  return r;
}

template <bool kPeriodic>
__device__ inline int32_t neighbour(const ModelDev& m, const RefSite& r, uint32_t slot,
                                    const uint32_t* cells) {
  if constexpr (kPeriodic) {
    const uint32_t c = cells[slot];
    int32_t x = r.x + int32_t(c & 0xFFu);
    int32_t y = r.y + int32_t((c >> 8) & 0xFFu);
    int32_t z = r.z + int32_t((c >> 16) & 0xFFu);
    x -= (x >= m.dims[0]) ? m.dims[0] : 0;
    y -= (y >= m.dims[1]) ? m.dims[1] : 0;
    z -= (z >= m.dims[2]) ? m.dims[2] : 0;
    return (x * m.dims[1] + y) * m.dims[2] + z;
  } else {
    return m.trans[int64_t(r.site) * m.K + slot];
  }
}

}  // namespace cemc
