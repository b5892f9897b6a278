// Launch interface between the host library (ce_host.hip) and the kernels
// (ce_kernels.hip).
#pragma once

#include "ce_device.h"

namespace cemc {

// Moves per speculative round of the sweep kernel.
constexpr int kSpecMoves = 8;

struct LaunchSweep {
  const ChainParams* params;
  ChainState* states;
  const ModelDev* models;
  int8_t* occ_pool;
  uint16_t* trk_pool;
  const int32_t* chain_order;  // blockIdx -> chain
  const int64_t* n_moves;      // per chain
  int n_chains;                // blocks in this launch
  uint32_t lds_bytes;
  bool periodic;
  bool canonical;
  int kr;          // records per lane held in VGPRs (2, 4, 8) or 0 = records in LDS
  bool gtab_lds;
  int kbm;         // decoration slices read per record: n_bfs for up to 4 species (1, 2, 3), else 7
  int waves;       // wavefronts per chain: 1, 2, 4 or 8 (more than 1: kr > 0)
  int spec;        // moves per round: 1, or kSpecMoves (periodic, kr > 0)
  bool field;      // local-field kernel (periodic, one wave, kr > 0, kbm <= 3)
  HookDev* hooks;  // per chain constraints and bias (one wave, one move per round), or nullptr
  int nm;          // members gathered per record: 3, or 2 for records in VGPRs of models without quadruplets
  uint32_t* accept_log;  // optional: bit t of chain c at accept_log[c * log_words + t / 32]
  int64_t log_words;
  // time-sliced launch: a chain stops at the first 64-move block that starts
  // more than time_limit ticks of the constant-rate wall clock after the
  // chain's own start (0: no limit; ignored by the local-field kernel)
  int64_t time_limit = 0;
};

// Dynamic LDS of one chain: occupation bytes (padded box on periodic
// lattices), tracker entries (canonical chains), records held in LDS, table doubles.
uint32_t sweep_lds_bytes(int occ_bytes, int trk_sites, int n_rec_lds, int gtab_doubles);
bool sweep_variant_ok(int kr, bool gtab_lds, int kbm, int waves, int spec, bool periodic);
// Dynamic LDS of one chain on the local-field kernel: padded occupation and
// its site map, tracker entries, n_sites x B integer fields, integer tables.
uint32_t field_lds_bytes(int occ_bytes, int trk_sites, int n_fields, int gtab_ints);
// Raises the dynamic-LDS limit of every sweep kernel instantiation to
// `bytes` minus that kernel's static LDS.
hipError_t sweep_set_max_lds(uint32_t bytes);
// Static LDS of the sweep kernel (bound on what the dynamic part may use).
uint32_t sweep_static_lds();

hipError_t launch_sweep(const LaunchSweep& a, hipStream_t stream);

// Wang-Landau walkers (canonical chains, one wave each): n_checks rounds of
// check_every moves, each followed by the flatness test.
struct LaunchWL {
  const ChainParams* params;
  ChainState* states;
  const ModelDev* models;
  int8_t* occ_pool;
  uint16_t* trk_pool;
  const int32_t* chain_order;  // blockIdx -> chain
  WLDev* walkers;              // per chain
  double* logdos;              // bin pools
  uint32_t* hist;
  uint8_t* known;
  int n_chains;
  uint32_t lds_bytes;
  bool periodic;
  int kr;
  bool gtab_lds;
  int kbm;
  int nm;  // members gathered per record (2 or 3, as for the sweep kernel)
  int64_t n_checks, check_every;
  uint32_t* accept_log;  // optional, as for the sweep kernel (bits OR-ed in)
  int64_t log_words;
  bool sgc = false;               // semi-grand-canonical walkers (all of the launch)
  bool adaptive = false;          // some walker of the launch uses adaptive windows
  const WLDev* shared = nullptr;  // walkers on this leader's shared histogram (one check per launch)
  int64_t log_base = 0;           // accept-log bit of the launch's first counter
};
// Dynamic LDS of one walker: the sweep layout of its chain (base bytes)
// followed by its ln g, counts and known flags (and, with adaptive windows,
// its sub-bin counts and recorded-state flags).
uint32_t wl_lds_bytes(uint32_t base, int n_bins, bool adaptive);
hipError_t launch_wl(const LaunchWL& a, hipStream_t stream);
// Flatness test and f update of the shared histogram of `leader` after a
// check of `total` steps over its walkers.
hipError_t launch_wl_shared_check(WLDev* leader, uint32_t* hist, const uint8_t* known, int64_t total,
                                  hipStream_t stream);
hipError_t launch_pt_exchange(PTDev* pt, ChainState* states, ChainParams* params, hipStream_t stream);
hipError_t launch_cf(const ModelDev* models, const ChainParams* params, const int8_t* occ_pool,
                     const int32_t* chain_ids, int n_chains, int max_sites, double* partial, double* cf_out,
                     double* energy_out, hipStream_t stream);
hipError_t launch_energy(const ModelDev* models, const ChainParams* params, const int8_t* occ_pool,
                         const int32_t* chain_ids, int n_chains, int max_sites, double* partial, double* energy_out,
                         hipStream_t stream);
// Up to kInlineChanges site/species pairs of a restore travel in the kernel
// arguments, so an undo needs no copy to the device.
constexpr int kInlineChanges = 8;
struct InlineChanges {
  int32_t v[4 * kInlineChanges];
};

// Work plan of the updater kernel for one model.  The change of correlation
// function k at a site of translation group g sums one product per
// (equivalent decoration, sub-cluster record of the site) pair; those
// products are the items of segment (g, k), and a change at a site of group
// g touches the segments [grp_seg[g], grp_seg[g + 1]) (ce_updater.cpp:349-404).
struct UpdItem {
  // members other than the reference site, in canonical order: closed-form
  // periodic lattices: wrapped cell offset x | y<<8 | z<<16; otherwise
  // translation-table column
  uint32_t m[3];
  // decoration of position p in bits 3p..3p+2 (p = 0..3), the reference
  // site's position << 12, members other than the reference site << 14
  uint32_t aux;
};
struct UpdSeg {
  int32_t k;
  int32_t begin, end;  // items
  int32_t pad;
  double mul;  // cluster size / equivalent decorations
  double div;  // normalisation C_p * T_g
};
// Everything the kernel needs besides the lattice travels in its arguments,
// so no load waits on another one at the start.
constexpr int kUpdMaxSinglets = 8;
struct UpdPlan {
  const UpdItem* items;
  const UpdSeg* segs;
  const int32_t* trans;     // translation table (lattices without the closed form)
  int32_t singlets[kUpdMaxSinglets];  // singlet correlation functions: k | decoration << 16
  int32_t n_singlets;
  int32_t chunk;            // items per LDS pass
  int32_t n_sites, n_cf, n_species, n_bfs, K, periodic;
  int32_t dims[3];
  int32_t occ_lds;          // 1: the occupation is staged in LDS
  double phi[kMaxBasis * kMaxSpecies];  // basis functions phi[d * n_species + s]
};
// Result of an updater launch in mapped, coherent host memory: the bits of
// the energy, one 8-byte store at system scope (and, in a second buffer, the
// bits of the final correlation functions, one store each).  The host sets them to
// kUpdPending (a signalling NaN that no arithmetic result carries) before the
// launch and waits until they change, so no fence or cache write-back is
// needed on the device.
constexpr uint64_t kUpdPending = 0x7FF4C0DE0000D1CEull;
struct UpdResult {
  uint64_t energy_bits;
  uint64_t pad;
};
// Up to kUpdInline changes of the updater interface travel in the kernel
// arguments, kUpdWords ints each: site, new | old << 8, src | dst << 16
// (history slots), first item, end of items, first segment, end of segments,
// unused; longer sequences are read from `changes` in device memory.
constexpr int kUpdInline = 8;
constexpr int kUpdWords = 8;
struct UpdChanges {
  int32_t v[kUpdWords * kUpdInline];
};
// Command ring of a resident updater kernel, in mapped, coherent host
// memory.  Word i of command s (i = 0..63: op, n, want, nr, then
// kUpdCmdWords words) sits in the low half of w[i] of slot s % kUpdRing,
// with s's low 32 bits in the high half; the kernel's wave loads all 64
// words at every poll and takes the command once every word carries its
// number (8-byte words are written and read whole, so a taken command
// holds no word of an older one).  The host reuses a slot once `consumed`
// (stored by the kernel after running a command) has reached the command
// kUpdRing before.
constexpr int kUpdRing = 4;
enum : int32_t { kUpdOpChanges = 1, kUpdOpRestore = 2, kUpdOpQuit = 3 };
constexpr int kUpdCmdWords = 60;
// words: nr sites, their nr species (restored in order before the
// changes), then (kUpdOpChanges) kUpdWords words per change as in
// UpdChanges; n changes fit when kUpdWords * n + 2 * nr <= kUpdCmdWords
struct UpdCmd {
  uint64_t w[4 + kUpdCmdWords];
};
constexpr int kUpdRestoreMax = kUpdCmdWords / 2;
struct UpdMailbox {
  int64_t consumed;  // sequence number of the last command the kernel ran
  int64_t pad[7];
  UpdCmd cmd[kUpdRing];
};
static_assert(sizeof(UpdCmd) == 64 * sizeof(uint64_t), "a command is one 8-byte word per lane of a wave");

// Dynamic LDS of the updater kernel: correlation functions, segment sums,
// ECIs and segments (per correlation function), item products of one pass,
// and the occupation when staged (`occ_sites` bytes, 0 when not).
uint32_t updater_lds_bytes(int n_cf, int chunk, int occ_sites);
hipError_t launch_updater(const UpdPlan& plan, int8_t* occ, double* ring,
                          double* slot_energy, const UpdChanges& inl, const int32_t* changes, int n,
                          const double* eci, UpdResult* result, uint64_t* result_cf, uint32_t lds_bytes,
                          hipStream_t stream);
// Launches the resident updater kernel (same LDS as launch_updater), which
// runs the commands from sequence number `next` on and leaves on a quit
// command or after `idle_ticks` wall-clock ticks without one.
hipError_t launch_updater_resident(const UpdPlan& plan, int8_t* occ, double* ring, double* slot_energy,
                                   const double* eci, UpdResult* result, uint64_t* result_cf, UpdMailbox* mb,
                                   int64_t next, int64_t idle_ticks, uint32_t lds_bytes, hipStream_t stream);
// Restores sites[i] = species[i] in order; n <= 2 * kInlineChanges pairs
// come in `inl` (sites in v[0..n), species in v[n..2n)), more from `buf`
// (sites then species).
hipError_t launch_restore(int8_t* occ, const InlineChanges& inl, const int32_t* buf, int n, hipStream_t stream);
hipError_t launch_slot_energy(const ModelDev* model, const double* ring, double* slot_energy, const double* eci,
                              int slot, hipStream_t stream);

}  // namespace cemc
