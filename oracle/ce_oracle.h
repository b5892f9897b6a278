/*
 * Test infrastructure: CPU oracle for the cluster-expansion Metropolis path.
 *
 * A plain-C restatement of the reference updater's arithmetic
 * (cpp/src/ce_updater.cpp, cpp/src/cf_history_tracker.cpp,
 * cpp/src/named_array.cpp) and of the Metropolis step of the reference
 * samplers (cemc/mcmc/montecarlo.py:890-1038, cemc/mcmc/sgc_montecarlo.py:62-76),
 * driven by a fed Philox4x64-10 move stream.  Only tests, the smoke check and
 * the CPU-baseline leg of the benchmark use it; the product never does.
 */
#ifndef CE_ORACLE_H
#define CE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CEO_OK = 0, CEO_EINVAL = 1, CEO_ERUNTIME = 2 };

enum { CEO_CF_C0 = 0, CEO_CF_SINGLET = 1, CEO_CF_CLUSTER = 2 };

typedef struct ceo_model ceo_model;
typedef struct ceo_chain ceo_chain;

/* Model tables (shared, read only).
 *  basis[d*n_species + s]              phi_d(s)          (basis_function.cpp:34-42)
 *  site_group[i]                       translation group (-1: none; only allowed for background)
 *  background[i]                       nonzero: background site, never moved
 *  cf_kind[k], cf_dec[k]               per correlation function in sorted-name order
 *  cf_cluster[g*n_cf + k]              cluster id in group g (-1: not present)
 *  cf_eq_off[k]..cf_eq_off[k+1]        equivalent decorations of CF k (index into eq_decos,
 *                                      each entry is 4 ints, unused tail = 0)
 *  cl_size[c], cl_M[c], cl_norm[c]     size n, sub-clusters M, C_p = sum over groups of M
 *  cl_mem_off[c]                       offset into members (M*(n-1) trans slots)
 *  cl_ord_off[c]                       offset into order (M*n)
 *  trans[i*K + slot]                   translation table restricted to used columns
 */
ceo_model* ceo_model_create(int n_sites, int n_species, int n_bfs, const double* basis,
                            int n_groups, const int* site_group, const int* background, int n_cf, const int* cf_kind,
                            const int* cf_dec, const int* cf_cluster, const int* cf_eq_off,
                            const int* eq_decos, int n_clusters, const int* cl_size,
                            const int* cl_M, const int* cl_norm, const int* cl_mem_off,
                            const int* members, const int* cl_ord_off, const int* order, int K,
                            const int* trans);
void ceo_model_destroy(ceo_model* m);

ceo_chain* ceo_chain_create(const ceo_model* m, const int8_t* occupation, const double* cf_init,
                            const double* eci);
void ceo_chain_destroy(ceo_chain* c);
void ceo_chain_set_threads(ceo_chain* c, int n_threads);

/* Updater operations (ce_updater.cpp:313-406, 408-448, 471-519, 552-555, 236-242).
 * A change is (site, old species, new species); old/new ids only decide the
 * no-op test the reference makes on symbol strings (ce_updater.cpp:315-318). */
int ceo_update_cf(ceo_chain* c, int site, int old_id, int new_id);
int ceo_calculate(ceo_chain* c, int n, const int* site, const int* old_id, const int* new_id,
                  double* energy);
int ceo_undo_changes(ceo_chain* c);
void ceo_clear_history(ceo_chain* c);
double ceo_get_energy(const ceo_chain* c);
void ceo_get_cf(const ceo_chain* c, double* out);
void ceo_set_ecis(ceo_chain* c, const double* eci);
void ceo_get_occupation(const ceo_chain* c, int8_t* out);
int ceo_history_size(const ceo_chain* c);
const char* ceo_last_error(void);

/* Philox4x64-10 block (Salmon et al., SC'11), counter used as given. */
void ceo_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

/* Fed-stream Metropolis runs.  Move t of chain `chain_uid` uses the Philox block
 * with key (seed, chain_uid) and counter (t, 0, 0, 0).
 *  kind 0: canonical swaps with the species index tracker
 *  kind 1: semi-grand-canonical flips
 * accept_bits: one bit per move (LSB first), may be NULL.
 * e_trace: energy after every `trace_every` moves, may be NULL.
 * Returns CEO_OK or an error code. */
int ceo_run(ceo_chain* c, int kind, uint64_t seed, uint64_t chain_uid, uint64_t move_begin,
            int64_t n_moves, double temperature, uint8_t* accept_bits, double* e_trace,
            int64_t trace_every, int64_t* n_accept);

/* Constraints and bias of a run (cemc/mcmc/mc_constraints.py,
 * cpp/src/pair_constraint.cpp, cemc/mcmc/bias_potential.py).  Proposal k of
 * move t uses counter (t, k, 0, 0); the first of max_attempts proposals
 * that passes every constraint is made, else the run stops with CEO_ERUNTIME
 * and failed_move set.  The bias at the count of bias_species joins the
 * energy. */
typedef struct ceo_hooks {
  uint32_t fixed_species;   /* bit s: species s never changes */
  int32_t max_attempts;
  const uint8_t* allowed;   /* n_sites masks of species allowed per site, or NULL */
  int32_t pair_e1, pair_e2; /* -1: no pair constraint */
  int32_t pair_nm;
  const int32_t* pair_nb;   /* n_sites x pair_nm member sites */
  int32_t bias_species;     /* -1: no bias */
  const double* bias_table; /* n_sites + 1 entries */
  int64_t failed_move;      /* out: -1 or the Philox counter of the failed move */
  int32_t recycle;          /* waste recycling (montecarlo.py:937-948, 805-811; util.py:41-63) */
  int32_t pad;
  double* rec_sums;         /* out (recycle): n, sum E, sum E^2, then per singlet d: sum s_d,
                               sum s_d^2, sum s_d E (n_bfs entries each), accumulated */
} ceo_hooks;

int ceo_run_hooked(ceo_chain* c, int kind, uint64_t seed, uint64_t chain_uid, uint64_t move_begin,
                   int64_t n_moves, double temperature, uint8_t* accept_bits, double* e_trace,
                   int64_t trace_every, int64_t* n_accept, ceo_hooks* hooks);

/* Run many independent chains on OpenMP threads (one chain per thread at a
 * time), each over the same number of moves. */
int ceo_run_many(ceo_chain** chains, int n_chains, int kind, uint64_t seed,
                 const uint64_t* chain_uids, uint64_t move_begin, int64_t n_moves,
                 const double* temperatures, int n_threads, int64_t* n_accept);

/* Wang-Landau walker (wang_landau_sampler.cpp:232-404, histogram.cpp:66-172,
 * adaptive_windows.cpp) on a canonical chain: one histogram per walker.
 * Counters accumulate over calls.
 *
 * Schemes of WangLandauSampler::run:
 *  inverse_time  once the steps a flat histogram took exceed those of the
 *                previous one, f = f_act * t_act / t at every later flat
 *                histogram, t = steps / n_bins (:344-367, get_mc_time :605);
 *                the reference times convergence in CPU seconds, here in
 *                steps so that runs repeat
 *  adaptive      AdaptiveWindowHistogram: the window [0, upper) shrinks
 *                from the top as its upper part converges, the walker's
 *                state is recorded at the first visit of every bin, a cut
 *                moves the walker to the recorded state nearest the window
 *                middle, and a flat window restores the full range with
 *                ln g made continuous at the cuts; sub-bin visit counts
 *                (10 per bin) are kept in this mode only, as in the
 *                reference (Histogram::init_sub_bins is called only there).
 * The scheme fields are reset by the caller when a run starts (the
 * reference builds a new sampler per run_fast_sampler call).
 *
 * kind 1 proposes the semi-grand-canonical flip of ceo_run's SGC moves
 * (a movable site, one of the other species), so the walker samples the
 * density of states of the energy with the chemical-potential term, the
 * ensemble of the reference's Python WangLandau with ensemble
 * "semi-grand-canonical" (cemc/wanglandau/wang_landau_sampler.py:268-311). */
typedef struct ceo_wl {
  int32_t n_bins;
  int32_t cur_bin;   /* bin of the walker's state (ignored while is_first) */
  int32_t is_first;  /* the next in-range move is accepted whatever ln g says */
  int32_t converged; /* f fell below f_min */
  double e_min, e_max;
  double f, f_min, flatness;
  double* logdos;
  int32_t* hist;
  uint8_t* known;
  int64_t n_moves, n_accept, n_outside, n_self, n_flat;
  /* schemes */
  int32_t inverse_time, adaptive, min_width, inv_active;
  int32_t upper;      /* adaptive: bins [0, upper) are in range (n_bins otherwise) */
  int32_t n_overlap;  /* adaptive: bins kept above a cut (starts at 7) */
  int32_t n_edges, has_first;
  double inv_factor;  /* f_act * t_act of the inverse-time scheme */
  double bin_change;  /* adaptive: sum of |proposed bin - current bin| since the last flat histogram */
  double cur_energy;  /* energy of the last accepted state (the sub-bin energy) */
  int64_t steps;      /* steps since the scheme was set (the time t counts these) */
  int64_t since_flat; /* steps since the last flat histogram */
  int64_t prev_conv;  /* steps the previous flat histogram took, -1 before the first */
  int64_t n_walk;     /* moves made to bring the walker back into the window */
  int32_t* edges;     /* adaptive: n_bins cut positions, in order */
  int32_t* edge_n;    /* values saved at each cut */
  double* edge_vals;  /* n_bins x n_bins: ln g above each cut when it was made */
  uint8_t* saved;     /* n_bins: a state was recorded at the bin's first visit */
  int8_t* snap_occ;   /* (n_bins + 1) x n_sites: recorded occupations (row n_bins: first cut) */
  double* snap_cf;    /* (n_bins + 1) x n_cf */
  int32_t* snap_trk;  /* (n_bins + 1) x n_sites: species index trackers */
  int32_t* snap_bin;  /* n_bins + 1 */
  uint32_t* sub_hist; /* n_bins x 10 */
  int32_t kind;       /* 0: canonical swaps; 1: semi-grand-canonical flips (energies with
                         the chemical potentials folded into the singlet ECIs) */
  int32_t pad;
} ceo_wl;

/* n_checks rounds of check_every steps, each round followed by the
 * flatness test; stops early once converged.  Proposal t (steps and the
 * moves that bring the walker back into an adaptive window alike) uses the
 * Philox block of counter move_begin + t as ceo_run's canonical moves do;
 * accept_bits (bits_cap bits, may be NULL) gets bit t for accepted steps.
 * *n_used (may be NULL) receives the number of counters used. */
int ceo_wl_run(ceo_chain* c, ceo_wl* w, uint64_t seed, uint64_t chain_uid, uint64_t move_begin,
               int64_t n_checks, int64_t check_every, uint8_t* accept_bits, int64_t bits_cap, int64_t* n_used);

#ifdef __cplusplus
}
#endif
#endif
