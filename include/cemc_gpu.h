/*
 * cemc_gpu.h — C ABI of the MI355X cluster-expansion Metropolis engine
 * (libcemcgpu.so, HIP kernels for gfx950).
 *
 * Plain pointers and sizes only.  Every call returns CEMC_OK (0) or an error
 * code; cemc_last_error() gives the message of the calling thread's last
 * failure.  Species are integer ids in sorted-name order, as the reference
 * assigns them (cpp/src/symbols_with_numbers.cpp:11-21).  Correlation
 * functions and ECIs are arrays in sorted-name order, the order of the
 * reference's NamedArray (cpp/src/named_array.cpp:18-23,
 * cpp/src/cf_history_tracker.cpp:7-16).
 */
#ifndef CEMC_GPU_H
#define CEMC_GPU_H

#include <stdint.h>

/* Only the entry points below are exported; the library is built with hidden
 * default visibility. */
#define CEMC_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CEMC_OK = 0,
  CEMC_EINVAL = 1,   /* maps to ValueError   (std::invalid_argument in the reference) */
  CEMC_ERUNTIME = 2, /* maps to RuntimeError (std::runtime_error in the reference)    */
  CEMC_EDEVICE = 3,  /* HIP runtime failure                                         */
  CEMC_EFULL = 4     /* a pool has no free slot or memory for one more chain        */
};

enum { CEMC_CF_C0 = 0, CEMC_CF_SINGLET = 1, CEMC_CF_CLUSTER = 2 };
enum { CEMC_MOVE_CANONICAL = 0, CEMC_MOVE_SGC = 1 };
/* How a trial move's energy change is obtained (cemc_batch_set_energy_path). */
enum { CEMC_PATH_AUTO = 0, CEMC_PATH_EVALUATE = 1, CEMC_PATH_FIELDS = 2 };

typedef struct cemc_model cemc_model;
typedef struct cemc_updater cemc_updater;
typedef struct cemc_batch cemc_batch;
typedef struct cemc_pt cemc_pt;

/* Message of the calling thread's last failed call. */
CEMC_API const char* cemc_last_error(void);

/* Number of visible GPUs (0 when none). */
CEMC_API int cemc_device_count(int* count);

/* ------------------------------------------------------------------ model
 * Lattice, clusters, basis and ECIs: the data CEUpdater::init reads from the
 * settings object (cpp/src/ce_updater.cpp:32-234), flattened.
 *  entries    one per (translation group, cluster name) of size >= 2
 *             (ce_updater.cpp:115-147); sub-cluster m of entry c lists its
 *             n-1 member columns in members[mem_off[c] + m*(n-1) ...] as
 *             indices into trans columns, and its canonical order in
 *             order[ord_off[c] + m*n ...] with the semantics of
 *             CEUpdater::sort_indices (ce_updater.cpp:981-994).
 *  trans      n_sites x n_trans_cols translation table restricted to the used
 *             columns (RowSparseStructMatrix, row_sparse_struct_matrix.cpp:10-122).
 *  periodic_dims / col_cells   optional closed form of trans for a periodic
 *             supercell with one site per cell, site = (x*L2 + y)*L3 + z:
 *             trans(i, col) = site(cell(i) + col_cells[col]).  It is used
 *             only after the library has checked it against `trans` for
 *             every row.
 */
typedef struct cemc_model_desc {
  int32_t n_sites;
  int32_t n_species;
  int32_t n_groups;
  int32_t n_cf;
  int32_t n_entries;
  int32_t n_trans_cols;
  const double* basis;          /* (n_species-1) x n_species                  */
  const int32_t* site_group;    /* n_sites translation group; -1 (no group) is
                                   allowed for background sites only           */
  const int32_t* cf_kind;       /* n_cf, CEMC_CF_*                             */
  const int32_t* cf_dec;        /* n_cf, decoration of singlets                */
  const int32_t* cf_entry;      /* n_groups x n_cf, entry id or -1             */
  const int32_t* cf_eq_off;     /* n_cf + 1                                    */
  const int32_t* eq_decos;      /* 4 ints per equivalent decoration            */
  const int32_t* entry_group;   /* n_entries                                   */
  const int32_t* entry_size;    /* n_entries                                   */
  const int32_t* entry_M;       /* n_entries                                   */
  const int32_t* entry_norm;    /* n_entries, sub-clusters summed over groups  */
  const int32_t* entry_mem_off; /* n_entries + 1                               */
  const int32_t* members;
  const int32_t* entry_ord_off; /* n_entries + 1                               */
  const int32_t* order;
  const int32_t* trans;         /* n_sites x n_trans_cols                      */
  const int32_t* periodic_dims; /* 3 values or NULL                            */
  const int32_t* col_cells;     /* n_trans_cols x 3 or NULL                    */
  const double* eci;            /* n_cf                                        */
  const int32_t* background;    /* n_sites or NULL; nonzero = background site,
                                   never moved (ce_updater.cpp:330-332)         */
} cemc_model_desc;

/* Builds device tables; replaces CEUpdater::init's parsing
 * (cpp/src/ce_updater.cpp:32-234, cpp/src/cluster.cpp:41-232). */
CEMC_API int cemc_model_create(const cemc_model_desc* desc, cemc_model** out);
/* ECIs of the batched path (energy tables rebuilt); CEUpdater::set_ecis
 * (cpp/src/ce_updater.cpp:615-629). */
CEMC_API int cemc_model_set_ecis(cemc_model* model, const double* eci);
/* 1 when the closed-form periodic neighbour path is active, else 0. */
CEMC_API int cemc_model_is_periodic(const cemc_model* model, int* periodic);
CEMC_API int cemc_model_destroy(cemc_model* model);

/* --------------------------------------------------------------- updater
 * One lattice with the reference's history/undo semantics
 * (class CEUpdater, cpp/include/ce_updater.hpp:47-213; Python binding
 * PyCEUpdater, cemc/cpp_ext/pyce_updater.pyx:5-60). */

/* PyCEUpdater.__init__ / CEUpdater::init (ce_updater.cpp:32-234).
 * cf_init may be NULL: the correlation functions are then computed on the
 * GPU from the occupation. */
CEMC_API int cemc_updater_create(cemc_model* model, const int8_t* occupation, const double* cf_init,
                        const double* eci, cemc_updater** out);
CEMC_API int cemc_updater_destroy(cemc_updater* upd);
/* PyCEUpdater.update_cf / CEUpdater::update_cf (ce_updater.cpp:287-406).  A
 * change with old_id == new_id is skipped (ce_updater.cpp:315-318); otherwise
 * the update starts from the species the site holds, as the reference reads
 * it from its own symbols (ce_updater.cpp:335), and old_id is what an undo
 * restores.  One updater is not safe to call from several threads at once. */
CEMC_API int cemc_updater_update_cf(cemc_updater* upd, int32_t site, int32_t old_id, int32_t new_id);
/* PyCEUpdater.calculate / CEUpdater::calculate (ce_updater.cpp:471-550, 831-852).
 * Returns the total energy after the changes (same change semantics as
 * cemc_updater_update_cf). */
CEMC_API int cemc_updater_calculate(cemc_updater* upd, int32_t n, const int32_t* site, const int32_t* old_id,
                           const int32_t* new_id, double* energy);
/* PyCEUpdater.undo_changes / CEUpdater::undo_changes (ce_updater.cpp:408-448). */
CEMC_API int cemc_updater_undo_changes(cemc_updater* upd);
/* PyCEUpdater.clear_history / CEUpdater::clear_history (ce_updater.cpp:552-555). */
CEMC_API int cemc_updater_clear_history(cemc_updater* upd);
/* PyCEUpdater.get_energy / CEUpdater::get_energy (ce_updater.cpp:236-242). */
CEMC_API int cemc_updater_get_energy(cemc_updater* upd, double* energy);
/* PyCEUpdater.get_cf / CEUpdater::get_cf (ce_updater.cpp:570-583); n_cf values. */
CEMC_API int cemc_updater_get_cf(cemc_updater* upd, double* cf);
/* PyCEUpdater.set_ecis / CEUpdater::set_ecis (ce_updater.cpp:615-629); n_cf values. */
CEMC_API int cemc_updater_set_ecis(cemc_updater* upd, const double* eci);
/* PyCEUpdater.get_singlets / CEUpdater::get_singlets (ce_updater.cpp:655-683):
 * the singlet correlation functions of the current state in name order.
 * *n_singlets (if not NULL) receives their number; singlets may be NULL to
 * ask for it.  A capacity below that number fails with CEMC_ERUNTIME and the
 * reference's message. */
CEMC_API int cemc_updater_get_singlets(cemc_updater* upd, double* singlets, int32_t capacity, int32_t* n_singlets);
/* PyCEUpdater.add_linear_vib_correction / CEUpdater::add_linear_vib_correction
 * (ce_updater.cpp:685-689, 1002-1013): replaces the correction with
 * v[i] eV per kB T on correlation function cf_index[i] (name-order index). */
CEMC_API int cemc_updater_add_linear_vib_correction(cemc_updater* upd, int32_t n, const int32_t* cf_index,
                                                    const double* eci_per_kbT);
/* PyCEUpdater.vib_energy / CEUpdater::vib_energy (ce_updater.cpp:691-698,
 * linear_vib_correction.cpp:8-16): sum over the correction's functions in
 * name order of cf_k v_k kB T with kB = 8.6173303e-5 eV/K; 0 without one. */
CEMC_API int cemc_updater_vib_energy(cemc_updater* upd, double temperature, double* energy);
/* PyCEUpdater.set_num_threads / CEUpdater::set_num_threads (ce_updater.hpp:136):
 * accepted and stored (positive values only); the work is spread over GPU
 * lanes, so the count does not change the arithmetic or the results. */
CEMC_API int cemc_updater_set_num_threads(cemc_updater* upd, int32_t num_threads);
/* No counterpart in the reference: chooses how the updater's calls reach
 * the GPU.  on != 0 (the default): calls of up to 8 changes, and undos of
 * up to 32, are commands to a resident kernel, one workgroup that stays on
 * a compute unit and polls a command ring in mapped host memory, so a call
 * costs no kernel launch; it leaves 0.5 ms after the last command (and before
 * any other call of this updater that uses its stream), and while it runs
 * it holds that compute unit and the hardware queue of the updater's
 * stream, so it serves an updater only while that is the only live updater
 * of the process with it on (several, used in turn: one launch per call
 * each).  on == 0: one kernel launch per call.  Results are identical. */
CEMC_API int cemc_updater_set_resident(cemc_updater* upd, int32_t on);
/* PyCEUpdater.get_symbols (ce_updater.hpp:111) as species ids; n_sites values. */
CEMC_API int cemc_updater_get_occupation(cemc_updater* upd, int8_t* occupation);
/* Number of history entries (CFHistoryTracker::history_size, cf_history_tracker.hpp:48). */
CEMC_API int cemc_updater_history_size(cemc_updater* upd, int32_t* size);
/* Replace the whole occupation, recompute the correlation functions on the
 * GPU and clear the history: CE.set_symbols (cemc/ce_calculator.py:436-448),
 * which the reference runs as one update_cf per site then clear_history.
 * Changing a background site fails as update_cf does (ce_updater.cpp:330-332). */
CEMC_API int cemc_updater_set_occupation(cemc_updater* upd, const int8_t* occupation);

/* ----------------------------------------------------------------- batch
 * Many independent Markov chains resident on the GPU, advanced by one
 * persistent launch per call.  Each trial move restates
 * Montecarlo._mc_step (cemc/mcmc/montecarlo.py:970-1038) with the
 * proposals of _get_trial_move (montecarlo.py:890-908) or
 * SGCMonteCarlo._get_trial_move (cemc/mcmc/sgc_montecarlo.py:62-76) drawn
 * from a fed Philox4x64-10 stream: move t of a chain uses the block with key
 * (seed, uid) and counter (t, 0, 0, 0). */
typedef struct cemc_chain_desc {
  int32_t kind;             /* CEMC_MOVE_CANONICAL or CEMC_MOVE_SGC             */
  double temperature;       /* kelvin                                            */
  const double* chem_pot;   /* n_species-1 values subtracted from the singlet
                               ECIs (sgc_montecarlo.py:239-261), or NULL        */
  uint64_t seed;
  uint64_t uid;
  const int8_t* occupation; /* n_sites species ids                              */
  uint64_t first_move;      /* Philox counter of the chain's first move (0 for a
                               new chain; a resumed chain continues its stream) */
} cemc_chain_desc;

CEMC_API int cemc_batch_create(int32_t n_chains, cemc_model* const* chain_models, const cemc_chain_desc* chains,
                      cemc_batch** out);
CEMC_API int cemc_batch_destroy(cemc_batch* batch);
/* Advance chain c by n_moves[c] trial moves (n_moves NULL: `moves` each).
 * Montecarlo.runMC main loop (montecarlo.py:732-844) without equilibration. */
CEMC_API int cemc_batch_run(cemc_batch* batch, const int64_t* n_moves, int64_t moves);
/* Same as cemc_batch_run, also returning the launch's device time in ms
 * measured with HIP events on the batch's stream. */
CEMC_API int cemc_batch_run_timed(cemc_batch* batch, const int64_t* n_moves, int64_t moves, double* kernel_ms);
/* Time-sliced run for servers: like cemc_batch_run_timed, but each chain
 * stops at the first 64-move block that starts more than max_ms after its
 * own start (device wall clock), having made fewer than n_moves[c] moves if
 * its time ran out; cemc_batch_get_stats reports the moves each chain has
 * made.  A chain's moves are the same however its stream is cut into
 * launches.  Chains on the local-field energy path make all their moves.
 * The reference has no counterpart (its samplers run one chain). */
CEMC_API int cemc_batch_run_sliced(cemc_batch* batch, const int64_t* n_moves, int64_t moves, double max_ms,
                                   double* kernel_ms);
/* Same as cemc_batch_run, also recording accept (1) / reject (0) of every
 * move: bit t of chain c at accept_bits[c * words_per_chain + t / 32], bit
 * t % 32.  Used to compare acceptance sequences with the reference. */
CEMC_API int cemc_batch_run_logged(cemc_batch* batch, const int64_t* n_moves, int64_t moves, uint32_t* accept_bits,
                          int64_t words_per_chain);
/* Current energies (n_chains). */
CEMC_API int cemc_batch_get_energy(cemc_batch* batch, double* energy);
/* Occupation of one chain (n_sites of that chain). */
CEMC_API int cemc_batch_get_occupation(cemc_batch* batch, int32_t chain, int8_t* occupation);
/* Per chain: moves done and accepted since creation; sum of E and of E^2
 * over the moves since the last reset (each move samples the state after
 * its accept/reject). */
CEMC_API int cemc_batch_get_stats(cemc_batch* batch, int64_t* moves_done, int64_t* n_accept, double* sum_e,
                         double* sum_e2);
/* Zero the sample sums of one chain (chain >= 0) or of all chains (-1) and
 * take the current energy as their reference: Montecarlo.reset
 * (cemc/mcmc/montecarlo.py:349-364), SGCMonteCarlo.reset (sgc_montecarlo.py:212-218). */
CEMC_API int cemc_batch_reset_stats(cemc_batch* batch, int32_t chain);

/* Sample averages of one chain since its last reset: the quantities
 * Montecarlo.get_thermodynamic (montecarlo.py:861-888) and, for SGC chains,
 * the SGCObserver sums (cemc/mcmc/mc_observers.py:222-270) feed into
 * SGCMonteCarlo.get_thermodynamic (sgc_montecarlo.py:398-448). */
typedef struct cemc_thermo {
  int64_t n_samples;
  double energy;                 /* current energy                                  */
  double mean_energy;            /* <E>                                             */
  double var_energy;             /* <E^2> - <E>^2                                   */
  int32_t n_singlets;            /* n_species - 1                                   */
  double singlet[7];             /* <c1_d>; canonical chains: the constant value    */
  double var_singlet[7];         /* <c1_d^2> - <c1_d>^2 (SGC)                       */
  double cov_singlet_energy[7];  /* <c1_d E> - <c1_d><E> (SGC)                      */
} cemc_thermo;
CEMC_API int cemc_batch_get_thermo(cemc_batch* batch, int32_t chain, cemc_thermo* out);
/* Correlation functions of chain c computed from scratch on the GPU (n_cf of
 * that chain's model) and the energy they give. */
CEMC_API int cemc_batch_get_cf(cemc_batch* batch, int32_t chain, double* cf, double* energy);
/* Same results as cemc_batch_get_cf (the same kernels), without allocating:
 * no hipMalloc, hipFree or allocating library call.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * around HIP graph capture, tight serving loops, harnesses that forbid it).
 * Its partial sums live in a 1 MiB device array of the library that each
 * GPU holds until the process ends: a chain needs
 * 8 * (64 + n_cf rounded up to 32 + n_cf * ceil(n_sites / 256)) bytes, and a
 * larger request fails with CEMC_EINVAL.  All calls on one GPU share that
 * array, so calls that overlap on one GPU race: serialise them. */
CEMC_API int cemc_batch_get_cf_noalloc(cemc_batch* batch, int32_t chain, double* cf, double* energy);
/* Launch shape.  Each chain runs on `waves` 64-lane wavefronts of one
 * workgroup (1, 2, 4 or 8), chosen by default from the batch size so that
 * the device gets about one wave per SIMD.  waves = 0 restores that choice;
 * another value forces it wherever the chain's kernel variant allows (the
 * autotuner in cemc_amd/autotune.py times the candidates).  The reference
 * has no counterpart: its updater runs one chain per process. */
CEMC_API int cemc_batch_set_waves(cemc_batch* batch, int32_t waves);
/* Moves per round: 8 lets chains on closed-form periodic lattices
 * with one translation group, no background sites and sub-cluster records
 * in registers evaluate 8 consecutive moves together against the same state
 * and commit them in order up to the first move that an accepted earlier
 * move of the round affects (that move is evaluated again); 1 evaluates
 * every move on its own; 0 (default) takes 8 wherever the chain's variant
 * allows it and the moves of a round seldom conflict (the sites a move
 * interacts with, times 4 for swaps, are at most 12% of the lattice), else 1.
 * All give the same moves, energies and accept sequence, bit for bit; rounds
 * share one wave reduction, one barrier and one decision step among their
 * moves.  No counterpart in the reference (its moves are
 * sequential). */
CEMC_API int cemc_batch_set_moves_per_round(cemc_batch* batch, int32_t moves);
/* Energy path: CEMC_PATH_EVALUATE sums the energy tables over the flipped
 * site's sub-clusters for every trial move (the reference's update_cf per
 * move, ce_updater.cpp:313-406, in energy form); CEMC_PATH_FIELDS keeps every
 * site's field F_j^d = sum of its sub-clusters' table entries in LDS as an
 * exact integer sum, so a trial move reads one field and only accepted moves
 * update the fields of their sub-cluster partners, where the chain allows
 * it (closed-form periodic lattice, one translation group, no background, up
 * to four species, at most 512 sub-clusters per site, state within the
 * workgroup's LDS, one wave per chain); CEMC_PATH_AUTO (default) evaluates.
 * A chain's cost on the field path grows with its acceptance rate: it pays
 * for single chains at low acceptance, while in large batches the chain with
 * the highest acceptance sets the launch time.  Accept sequences and
 * lattices are the same on both paths; energies agree to rounding.  No
 * counterpart in the reference. */
CEMC_API int cemc_batch_set_energy_path(cemc_batch* batch, int32_t path);
/* Energy path a chain runs on (CEMC_PATH_EVALUATE or CEMC_PATH_FIELDS). */
CEMC_API int cemc_batch_get_energy_path(cemc_batch* batch, int32_t chain, int32_t* path);
/* Launch shape of one chain: waves per chain, sub-cluster records per lane
 * held in registers (0: records in LDS), energy tables in LDS (1) or HBM (0). */
CEMC_API int cemc_batch_get_launch_shape(cemc_batch* batch, int32_t chain, int32_t* waves,
                                         int32_t* records_per_lane, int32_t* tables_in_lds);
/* Temperature of one chain (parallel tempering uses it to exchange replicas). */
CEMC_API int cemc_batch_set_temperature(cemc_batch* batch, int32_t chain, double temperature);

/* ------------------------------------------------- constraints and bias
 * Per-move predicates and a bias on one chain of a fixed batch, checked on
 * the device.  A trial move is redrawn while it violates a constraint
 * (Montecarlo._mc_step, cemc/mcmc/montecarlo.py:970-991): proposal k of
 * move t uses Philox counter (t, k, 0, 0), and a move whose first
 * max_attempts proposals all violate ends the launch with CEMC_ERUNTIME and
 * the reference's message.  The bias of the proposed state joins its energy
 * (montecarlo.py:926-930, 998) and the chain's energy includes it.  Chains
 * with hooks run one wave and one move at a time. */
typedef struct cemc_hooks_desc {
  uint32_t fixed_species;  /* FixedElement (mc_constraints.py:84-105): bit s set = species s never changes */
  int32_t max_attempts;    /* max_constraint_attempts (> 0) */
  const uint8_t* allowed;  /* n_sites entries, bit s = species s may occupy the site (ConstrainElementByTag,
                              :171-190; a fixed site allows nothing, as FixEdgeLayers, :108-168); NULL: any */
  int32_t pair_species[2]; /* PairConstraint (pair_constraint.cpp:14-46): no site of species [0] with a member
                              of species [1] in the pair cluster; -1, -1: none */
  int32_t pair_members;    /* members of the pair cluster per site, 1-32 */
  const int32_t* pair_neighbours; /* n_sites x pair_members member sites; symmetric lists */
  int32_t bias_species;    /* bias on the count of this species (a SampledBiasPotential over a concentration,
                              bias_potential.py:52-140); -1: none */
  int32_t recycle_waste;   /* 1: waste recycling (montecarlo.py:937-948, 805-811; util.py:41-63):
                              heat-bath choice between the current and the proposed state, and each
                              move's sample weights both by exp(-E/kT) */
  const double* bias_table; /* n_sites + 1 entries: bias energy (eV) at each count of bias_species */
} cemc_hooks_desc;

/* Set (desc != NULL) or remove (NULL) the constraints and bias of a chain. */
CEMC_API int cemc_batch_set_hooks(cemc_batch* batch, int32_t chain, const cemc_hooks_desc* desc);
/* Constraint state of a chain: ordered constrained pairs in its lattice,
 * count of the biased species and the bias energy (any output may be NULL). */
CEMC_API int cemc_batch_get_hooks_state(cemc_batch* batch, int32_t chain, int64_t* pair_count, int32_t* bias_count,
                                        double* bias_energy);

/* ----------------------------------------------------------- Wang-Landau
 * A Wang-Landau walker on a chain of a (fixed) batch: the reference's
 * native sampler (cpp/src/wang_landau_sampler.cpp, bound by
 * cemc/cpp_ext/pywang_landau_sampler.pyx:4-30) on a canonical chain (its
 * swaps), or on a semi-grand-canonical chain (its flips, the energy with
 * the chemical potentials folded into the singlet ECIs: the ensemble of the
 * reference's Python WangLandau with ensemble "semi-grand-canonical",
 * cemc/wanglandau/wang_landau_sampler.py:268-311), with one histogram per walker
 * instead of one shared under a lock, so each chain is an independent
 * Wang-Landau run and their ln g can be averaged.  Energies are the
 * chain's total energy (eV). */
typedef struct cemc_wl_desc {
  int32_t n_bins;      /* Histogram::Nbins */
  int32_t current_bin; /* bin of the chain's state (ignored while is_first) */
  int32_t is_first;    /* 1: the next in-range proposal is accepted whatever ln g says */
  int32_t converged;   /* 1 once f < f_min; a converged walker makes no moves */
  double e_min, e_max; /* Histogram::Emin, Emax */
  double f;            /* modification factor added to ln g per visit */
  double f_min;        /* WangLandauSampler::min_f */
  double flatness;     /* flatness_criteria: flat when min(count) > flatness * mean(count) over known bins */
} cemc_wl_desc;

/* Give chain `chain` a walker, or replace its walker.  The constructor's
 * reads of the Python WangLandau object (wang_landau_sampler.cpp:17-150:
 * fmin, f, flatness_criteria, histogram Nbins/Emin/Emax/logdos/histogram/
 * known_state, current_bin).  logdos, hist and known hold n_bins entries
 * each, or are NULL for zeros.  The counters restart at zero. */
CEMC_API int cemc_batch_wl_set(cemc_batch* batch, int32_t chain, const cemc_wl_desc* desc, const double* logdos,
                               const int32_t* hist, const uint8_t* known);
/* Read a walker back (send_results_to_python, wang_landau_sampler.cpp:406-432).
 * Any output may be NULL.  counters[5]: moves, accepted, proposals outside
 * the range, proposals into the current bin, flat histograms seen. */
CEMC_API int cemc_batch_wl_get(cemc_batch* batch, int32_t chain, cemc_wl_desc* desc, double* logdos,
                               int32_t* hist, uint8_t* known, int64_t* counters);
/* Schemes of a walker's WangLandauSampler::run (wang_landau_sampler.cpp:
 * 319-404).  inverse_time: use_inverse_time_algorithm (:344-367): once a flat
 * histogram took more steps than the previous one, f = f_act t_act / t at
 * every later flat histogram, t = steps / n_bins (get_mc_time, :605; the
 * reference times convergence in CPU seconds, here in steps so that runs
 * repeat).  adaptive: use_adaptive_windows(min_width) (:575-580,
 * cpp/src/adaptive_windows.cpp), with sub-bin visit counts (10 per bin,
 * Histogram::update_sub_bin, histogram.cpp:92-100).  The other fields are
 * outputs of cemc_batch_wl_get_scheme. */
typedef struct cemc_wl_scheme {
  int32_t inverse_time; /* in: use the inverse-time law */
  int32_t adaptive;     /* in: adaptive energy windows */
  int32_t min_width;    /* in: minimum_window_width */
  int32_t inv_active;   /* the inverse-time law is in force */
  int32_t upper;        /* bins [0, upper) are in range */
  int32_t n_overlap;    /* bins kept above a cut */
  int32_t n_edges;      /* cuts made since the last flat window */
  int32_t has_first;    /* the state at the first cut is recorded */
  double inv_factor;    /* f t at activation of the inverse-time law */
  double bin_change;    /* adaptive: sum of |proposed bin - current bin| since the last flat histogram */
  int64_t steps;        /* steps since the scheme was set */
  int64_t since_flat;   /* steps since the last flat histogram */
  int64_t prev_conv;    /* steps the previous flat histogram took (-1: none yet) */
  int64_t n_walk;       /* moves made to bring the walker back into its window */
} cemc_wl_scheme;

/* Choose the schemes of walker `chain` and start them afresh (the
 * reference builds a new sampler per run_fast_sampler call,
 * wang_landau_sampler.py:450-469): counters, windows, recorded states and
 * sub-bin counts restart.  Adaptive windows keep (n_bins + 1) lattices and
 * trackers of the chain (3 bytes per site each) and n_bins^2 doubles of
 * saved ln g in device memory. */
CEMC_API int cemc_batch_wl_set_scheme(cemc_batch* batch, int32_t chain, const cemc_wl_scheme* scheme);
/* Read a walker's scheme state; edges and edge_n (n_bins entries), edge_vals
 * (n_bins x n_bins, row e: ln g saved above cut e), sub_bins (n_bins x 10)
 * and saved (n_bins recorded-state flags) may be NULL, and are zero
 * without adaptive windows. */
CEMC_API int cemc_batch_wl_get_scheme(cemc_batch* batch, int32_t chain, cemc_wl_scheme* scheme, int32_t* edges,
                                      int32_t* edge_n, double* edge_vals, uint32_t* sub_bins, uint8_t* saved);

/* WangLandauSampler::run (wang_landau_sampler.cpp:319-404) on every walker
 * of the batch at once: n_checks rounds of check_every moves (step,
 * :232-317), each followed by the flatness test; a flat histogram lowers f
 * (halves it, or follows the inverse-time law of cemc_wl_scheme) and clears
 * the counts, and a walker stops at the first check with f < f_min.
 * accept_bits (optional) as for cemc_batch_run_logged, bit t for
 * Philox counter t of the launch (the moves that bring a walker back into
 * an adaptive window use counters too and set no bit); kernel_ms (optional)
 * receives the device time.  Fails when an adaptive window's walker finds
 * no state inside its window within 10^7 moves. */
CEMC_API int cemc_batch_wl_run(cemc_batch* batch, int64_t n_checks, int64_t check_every, uint32_t* accept_bits,
                               int64_t words_per_chain, double* kernel_ms);
/* WangLandauSampler::run with its walkers on one shared histogram
 * (wang_landau_sampler.cpp:288-298, :321-335): every chain's walker visits
 * walker 0's bins, f and scheme; check_every steps per check (rounded up to
 * a multiple of the chain count) are split evenly over the walkers, and the
 * flatness test and f update follow each check.  Walkers read ln g as the
 * others' visits land, so runs with several walkers and f > 0 do not
 * repeat exactly (as the reference's threads do not); every walker needs
 * the same bins and range, and adaptive windows are not allowed.  Walker
 * 0's descriptor and bins hold the group's results; the others keep their
 * own lattice, current bin and counters.  accept_bits: bit t of chain c for
 * its t-th step. */
CEMC_API int cemc_batch_wl_run_shared(cemc_batch* batch, int64_t n_checks, int64_t check_every,
                                      uint32_t* accept_bits, int64_t words_per_chain, double* kernel_ms);

/* ------------------------------------------------------------------ pool
 * A batch whose chains join and leave while it runs: the resident state of
 * a job server.  Slots, model entries and device memory for occupations
 * (1 byte per site, rounded up to 16) and canonical species trackers (one
 * entry per site) are fixed at creation; a chain is admitted only when its
 * footprint fits (CEMC_EFULL otherwise) and its lattice then stays on the
 * GPU until it is removed.  cemc_batch_run* on a pool launches only the
 * active slots with a non-zero move count; the other cemc_batch_get_* and
 * cemc_batch_reset_stats calls take slot numbers as chain indices.  The
 * reference has no counterpart (one sampler per Python process); a chain's
 * moves are those of the same chain run alone. */
CEMC_API int cemc_pool_create(int32_t max_chains, int32_t max_models, int64_t occ_bytes, int64_t tracker_entries,
                              cemc_batch** out);
/* Device memory one chain of `model` needs (kind: CEMC_MOVE_*). */
CEMC_API int cemc_pool_footprint(const cemc_model* model, int32_t kind, int64_t* occ_bytes,
                                 int64_t* tracker_entries);
CEMC_API int cemc_pool_add(cemc_batch* pool, cemc_model* model, const cemc_chain_desc* chain, int32_t* slot);
CEMC_API int cemc_pool_remove(cemc_batch* pool, int32_t slot);

/* ------------------------------------------------------------ tempering
 * Parallel tempering over the chains of a batch
 * (ParallelTempering, cemc/mcmc/parallel_tempering.py:5-191), with the
 * exchange step on the GPU.  Replica slot t starts with chain t at that
 * chain's temperature.  One cycle = every chain makes `moves_per_cycle`
 * trial moves (the reference's runMC per replica, :186-187), then one
 * exchange step:
 *   direction of cycle c: Philox4x64-10 block (c, 0, 0, 0) of key
 *     (seed, uid), bit 63 of word 0 set = "up" (pairs (0,1), (2,3), ...),
 *     clear = "down" (pairs (R-1,R-2), (R-3,R-4), ...)   (:153-160, 188-191)
 *   pair k: u = (w3 >> 11) * 2^-53 of block (c, 1 + k, 0, 0); accepted if
 *     u < exp((1/kT_a - 1/kT_b) (E_a - E_b))              (:138-144, 166-168)
 * An accepted exchange moves the two chains between the slots (each takes
 * the other's temperature); lattices stay where they are.  Each chain's
 * energy samples of a cycle are added to the sums of the slot it held.
 * The chains' own sample sums are consumed by the run. */
CEMC_API int cemc_pt_create(cemc_batch* batch, uint64_t seed, uint64_t uid, cemc_pt** out);
CEMC_API int cemc_pt_destroy(cemc_pt* pt);
/* Run n_cycles cycles; use_graph != 0 replays them from a captured HIP graph
 * of up to 16 cycles.  kernel_ms (may be NULL) receives the device time. */
CEMC_API int cemc_pt_run(cemc_pt* pt, int64_t n_cycles, int64_t moves_per_cycle, int32_t use_graph,
                         double* kernel_ms);
/* Per slot (R values; NULL skips an output): chain held, temperature, energy
 * samples, their mean and variance; per pair (t, t+1) (R-1 values):
 * accepted and attempted exchanges. */
CEMC_API int cemc_pt_get(cemc_pt* pt, int32_t* chain_of_slot, double* temperatures, int64_t* n_accept,
                         int64_t* n_try, int64_t* n_samples, double* mean_energy, double* var_energy);

#ifdef __cplusplus
}
#endif
#endif
