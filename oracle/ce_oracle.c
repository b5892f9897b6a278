/*
 * Test infrastructure (see ce_oracle.h): CPU restatement of the reference
 * cluster-expansion updater and of the samplers' Metropolis step.  The loop
 * structure and floating-point operation order follow the reference so the
 * correlation functions and energies reproduce it bit for bit:
 *   update_cf                 cpp/src/ce_updater.cpp:313-406
 *   spin_product_one_atom     cpp/src/ce_updater.cpp:244-285
 *   sort_indices              cpp/src/ce_updater.cpp:981-994
 *   get_energy / dot          cpp/src/ce_updater.cpp:236-242, cpp/src/named_array.cpp:25-33
 *   calculate dispatch        cpp/src/ce_updater.cpp:471-550, 831-852, 996-1000
 *   undo / clear              cpp/src/ce_updater.cpp:408-448, 552-555
 *   history ring buffer       cpp/src/cf_history_tracker.cpp:46-149 (max_history 1000)
 *   Metropolis test           cemc/mcmc/montecarlo.py:910-956
 *   canonical proposal        cemc/mcmc/montecarlo.py:890-908,
 *                             cemc/mcmc/swap_move_index_tracker.py:39-74
 *   SGC proposal              cemc/mcmc/sgc_montecarlo.py:62-76
 * Deliberate difference: a flip of a background site is rejected before the
 * history slot is taken (the reference advances the ring first, ce_updater.cpp:323-332).
 */
#include "ce_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CEO_MAX_HISTORY 1000
#define CEO_MAX_CLUSTER 4
#define CEO_KB 8.617330337217213e-05

static _Thread_local char g_err[256];

static int fail(int code, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

const char* ceo_last_error(void) { return g_err; }

struct ceo_model {
  int n_sites, n_species, n_bfs, n_groups, n_cf, n_clusters, K;
  double* basis;
  int* site_group;
  int* background;
  int* group_count;
  int* cf_kind;
  int* cf_dec;
  int* cf_cluster;
  int* cf_eq_off;
  int* eq_decos;
  int* cl_size;
  int* cl_M;
  int* cl_norm;
  int* cl_mem_off;
  int* members;
  int* cl_ord_off;
  int* order;
  int* trans;
};

typedef struct {
  int site, old_id, new_id;
} ceo_change;

struct ceo_chain {
  const ceo_model* m;
  int8_t* occ;
  double* eci;
  double* ring;  /* CEO_MAX_HISTORY x n_cf */
  ceo_change* changes;
  unsigned current, buffer_size;
  int n_threads;
  /* species index tracker of the canonical sampler (swap_move_index_tracker.py:22-59):
   * per-species site lists, rebuilt in ascending site order when the
   * occupation was changed outside a run */
  int* trk;
  int* trk_off;
  int* trk_count;
  int trk_valid;
};

static void* dup_bytes(const void* src, size_t n) {
  void* p = malloc(n ? n : 1);
  if (p && n) memcpy(p, src, n);
  return p;
}

ceo_model* ceo_model_create(int n_sites, int n_species, int n_bfs, const double* basis,
                            int n_groups, const int* site_group, const int* background, int n_cf, const int* cf_kind,
                            const int* cf_dec, const int* cf_cluster, const int* cf_eq_off,
                            const int* eq_decos, int n_clusters, const int* cl_size,
                            const int* cl_M, const int* cl_norm, const int* cl_mem_off,
                            const int* members, const int* cl_ord_off, const int* order, int K,
                            const int* trans) {
  ceo_model* m = (ceo_model*)calloc(1, sizeof *m);
  if (!m) return NULL;
  m->n_sites = n_sites;
  m->n_species = n_species;
  m->n_bfs = n_bfs;
  m->n_groups = n_groups;
  m->n_cf = n_cf;
  m->n_clusters = n_clusters;
  m->K = K;
  m->basis = (double*)dup_bytes(basis, sizeof(double) * (size_t)n_bfs * n_species);
  m->site_group = (int*)dup_bytes(site_group, sizeof(int) * (size_t)n_sites);
  m->background = (int*)dup_bytes(background, sizeof(int) * (size_t)n_sites);
  /* trans_symm_group_count counts every site of the group, background sites
   * included (ce_updater.cpp:768-775) */
  m->group_count = (int*)calloc((size_t)n_groups + 1, sizeof(int));
  for (int i = 0; i < n_sites; ++i)
    if (site_group[i] >= 0) m->group_count[site_group[i]] += 1;
  m->cf_kind = (int*)dup_bytes(cf_kind, sizeof(int) * (size_t)n_cf);
  m->cf_dec = (int*)dup_bytes(cf_dec, sizeof(int) * (size_t)n_cf);
  m->cf_cluster = (int*)dup_bytes(cf_cluster, sizeof(int) * (size_t)n_cf * n_groups);
  m->cf_eq_off = (int*)dup_bytes(cf_eq_off, sizeof(int) * (size_t)(n_cf + 1));
  m->eq_decos = (int*)dup_bytes(eq_decos, sizeof(int) * CEO_MAX_CLUSTER * (size_t)cf_eq_off[n_cf]);
  m->cl_size = (int*)dup_bytes(cl_size, sizeof(int) * (size_t)n_clusters);
  m->cl_M = (int*)dup_bytes(cl_M, sizeof(int) * (size_t)n_clusters);
  m->cl_norm = (int*)dup_bytes(cl_norm, sizeof(int) * (size_t)n_clusters);
  m->cl_mem_off = (int*)dup_bytes(cl_mem_off, sizeof(int) * (size_t)(n_clusters + 1));
  m->members = (int*)dup_bytes(members, sizeof(int) * (size_t)cl_mem_off[n_clusters]);
  m->cl_ord_off = (int*)dup_bytes(cl_ord_off, sizeof(int) * (size_t)(n_clusters + 1));
  m->order = (int*)dup_bytes(order, sizeof(int) * (size_t)cl_ord_off[n_clusters]);
  m->trans = (int*)dup_bytes(trans, sizeof(int) * (size_t)n_sites * K);
  return m;
}

void ceo_model_destroy(ceo_model* m) {
  if (!m) return;
  free(m->basis);
  free(m->site_group);
  free(m->background);
  free(m->group_count);
  free(m->cf_kind);
  free(m->cf_dec);
  free(m->cf_cluster);
  free(m->cf_eq_off);
  free(m->eq_decos);
  free(m->cl_size);
  free(m->cl_M);
  free(m->cl_norm);
  free(m->cl_mem_off);
  free(m->members);
  free(m->cl_ord_off);
  free(m->order);
  free(m->trans);
  free(m);
}

/* ---------------------------------------------------------------- history */
static double* ring_slot(const ceo_chain* c, unsigned slot) {
  return c->ring + (size_t)slot * c->m->n_cf;
}

static unsigned current_slot(const ceo_chain* c) {
  return c->current == 0 ? CEO_MAX_HISTORY - 1 : c->current - 1;
}

ceo_chain* ceo_chain_create(const ceo_model* m, const int8_t* occupation, const double* cf_init,
                            const double* eci) {
  ceo_chain* c = (ceo_chain*)calloc(1, sizeof *c);
  if (!c) return NULL;
  c->m = m;
  c->occ = (int8_t*)dup_bytes(occupation, (size_t)m->n_sites);
  c->eci = (double*)dup_bytes(eci, sizeof(double) * (size_t)m->n_cf);
  c->ring = (double*)calloc((size_t)CEO_MAX_HISTORY * m->n_cf, sizeof(double));
  c->changes = (ceo_change*)calloc(CEO_MAX_HISTORY, sizeof(ceo_change));
  c->n_threads = 1;
  c->trk = (int*)malloc(sizeof(int) * (size_t)(m->n_sites ? m->n_sites : 1));
  c->trk_off = (int*)calloc((size_t)m->n_species + 1, sizeof(int));
  c->trk_count = (int*)calloc((size_t)m->n_species, sizeof(int));
  c->trk_valid = 0;
  /* CFHistoryTracker::insert then clear (cf_history_tracker.cpp:82-108, 146-149) */
  memcpy(ring_slot(c, 0), cf_init, sizeof(double) * (size_t)m->n_cf);
  c->current = 1;
  c->buffer_size = 1;
  return c;
}

void ceo_chain_destroy(ceo_chain* c) {
  if (!c) return;
  free(c->occ);
  free(c->eci);
  free(c->ring);
  free(c->changes);
  free(c->trk);
  free(c->trk_off);
  free(c->trk_count);
  free(c);
}

void ceo_chain_set_threads(ceo_chain* c, int n_threads) { c->n_threads = n_threads > 0 ? n_threads : 1; }

int ceo_history_size(const ceo_chain* c) { return (int)c->buffer_size; }

/* ------------------------------------------------------------- CF update */
static double spin_product_one_atom(const ceo_chain* c, int ref, int cl, const int* dec, int ref_id) {
  const ceo_model* m = c->m;
  const int n = m->cl_size[cl];
  const int n_memb = n - 1;
  const int M = m->cl_M[cl];
  const int* mem = m->members + m->cl_mem_off[cl];
  const int* ord = m->order + m->cl_ord_off[cl];
  const int* trow = m->trans + (size_t)ref * m->K;
  double sp = 0.0;
  for (int i = 0; i < M; ++i) {
    double sp_temp = 1.0;
    int indices[CEO_MAX_CLUSTER];
    int sorted[CEO_MAX_CLUSTER];
    indices[0] = ref;
    for (int j = 0; j < n_memb; ++j) indices[j + 1] = trow[mem[i * n_memb + j]];
    for (int j = 0; j < n; ++j) sorted[j] = indices[ord[i * n + j]];
    for (int j = 0; j < n; ++j) {
      int s = sorted[j] == ref ? ref_id : c->occ[sorted[j]];
      sp_temp *= m->basis[dec[j] * m->n_species + s];
    }
    sp += sp_temp;
  }
  return sp;
}

static void update_one_cf(const ceo_chain* c, int k, int site, int old_sid, int new_sid,
                          const double* cur, double* next) {
  const ceo_model* m = c->m;
  const int kind = m->cf_kind[k];
  if (kind == CEO_CF_C0) {
    next[k] = cur[k];
    return;
  }
  if (kind == CEO_CF_SINGLET) {
    const int d = m->cf_dec[k];
    next[k] = cur[k] + (m->basis[d * m->n_species + new_sid] - m->basis[d * m->n_species + old_sid]) /
                           (double)(unsigned)m->n_sites;
    return;
  }
  const int g = m->site_group[site];
  const int cl = m->cf_cluster[g * m->n_cf + k];
  if (cl < 0) {
    next[k] = cur[k];
    return;
  }
  const int n_e = m->cf_eq_off[k + 1] - m->cf_eq_off[k];
  double delta_sp = 0.0;
  for (int e = 0; e < n_e; ++e) {
    const int* deco = m->eq_decos + (size_t)(m->cf_eq_off[k] + e) * CEO_MAX_CLUSTER;
    double sp_ref = spin_product_one_atom(c, site, cl, deco, old_sid);
    double sp_new = spin_product_one_atom(c, site, cl, deco, new_sid);
    delta_sp += sp_new - sp_ref;
  }
  delta_sp *= ((double)(unsigned)m->cl_size[cl] / (double)(size_t)n_e);
  delta_sp /= (m->cl_norm[cl] * m->group_count[g]);
  next[k] = cur[k] + delta_sp;
}

static int update_cf_impl(ceo_chain* c, int site, int old_id, int new_id) {
  const ceo_model* m = c->m;
  if (old_id == new_id) return CEO_OK;
  if (site < 0 || site >= m->n_sites) return fail(CEO_EINVAL, "site index out of range");
  if (new_id < 0 || new_id >= m->n_species) return fail(CEO_EINVAL, "unknown species");
  if (m->background[site]) return fail(CEO_ERUNTIME, "Attempting to move a background atom!");
  const double* cur = ring_slot(c, current_slot(c));
  /* get_next (cf_history_tracker.cpp:46-56) */
  unsigned slot = c->current;
  c->current = (c->current + 1) % CEO_MAX_HISTORY;
  if (c->buffer_size < CEO_MAX_HISTORY) c->buffer_size += 1;
  double* next = ring_slot(c, slot);
  c->changes[slot].site = site;
  c->changes[slot].old_id = old_id;
  c->changes[slot].new_id = new_id;
  const int old_sid = c->occ[site];
  c->occ[site] = (int8_t)new_id;
  const int new_sid = new_id;
  if (c->n_threads > 1) {
#pragma omp parallel for num_threads(c->n_threads) schedule(dynamic)
    for (int k = 0; k < m->n_cf; ++k) update_one_cf(c, k, site, old_sid, new_sid, cur, next);
  } else {
    for (int k = 0; k < m->n_cf; ++k) update_one_cf(c, k, site, old_sid, new_sid, cur, next);
  }
  return CEO_OK;
}

double ceo_get_energy(const ceo_chain* c) {
  const double* cf = ring_slot(c, current_slot(c));
  double dot = 0.0;
  for (int k = 0; k < c->m->n_cf; ++k) dot += c->eci[k] * cf[k];
  return dot * (double)(unsigned)c->m->n_sites;
}

static int calculate_impl(ceo_chain* c, int n, const int* site, const int* old_id, const int* new_id,
                          double* energy) {
  int rc = CEO_OK;
  if (n == 1) {
    rc = update_cf_impl(c, site[0], old_id[0], new_id[0]);
  } else if (n > 1) {
    int all_swaps = (n % 2) == 0;
    for (int p = 0; all_swaps && p < n / 2; ++p) {
      /* is_swap_move (ce_updater.cpp:996-1000) */
      if (!(old_id[2 * p] == new_id[2 * p + 1] && new_id[2 * p + 1] == old_id[2 * p + 1]))
        all_swaps = 0;
    }
    if (all_swaps) {
      if (n / 2 >= CEO_MAX_HISTORY / 2)
        return fail(CEO_EINVAL,
                    "The length of sequence of swap move exceeds the buffer size for the history tracker");
      for (int p = 0; p < n / 2 && rc == CEO_OK; ++p) {
        int a = site[2 * p], b = site[2 * p + 1];
        if (c->occ[a] == c->occ[b])
          return fail(CEO_ERUNTIME,
                      "This version of the calculate function assumes that the provided update is "
                      "swapping two atoms");
        if (c->occ[a] != old_id[2 * p] || c->occ[b] != old_id[2 * p + 1])
          return fail(CEO_ERUNTIME, "The atom position tracker does not match the current state");
        rc = update_cf_impl(c, a, old_id[2 * p], new_id[2 * p]);
        if (rc == CEO_OK) rc = update_cf_impl(c, b, old_id[2 * p + 1], new_id[2 * p + 1]);
      }
    } else {
      for (int i = 0; i < n && rc == CEO_OK; ++i) rc = update_cf_impl(c, site[i], old_id[i], new_id[i]);
    }
  }
  if (rc != CEO_OK) return rc;
  if (energy) *energy = ceo_get_energy(c);
  return CEO_OK;
}

static void undo_impl(ceo_chain* c) {
  int steps = (int)c->buffer_size - 1;
  for (int i = 0; i < steps; ++i) {
    /* pop (cf_history_tracker.cpp:58-80) */
    if (c->buffer_size == 0) break;
    c->current = c->current == 0 ? CEO_MAX_HISTORY - 1 : c->current - 1;
    const ceo_change* ch = &c->changes[c->current];
    c->buffer_size -= 1;
    c->occ[ch->site] = (int8_t)ch->old_id;
  }
}

int ceo_update_cf(ceo_chain* c, int site, int old_id, int new_id) {
  c->trk_valid = 0;
  return update_cf_impl(c, site, old_id, new_id);
}

int ceo_calculate(ceo_chain* c, int n, const int* site, const int* old_id, const int* new_id,
                  double* energy) {
  c->trk_valid = 0;
  return calculate_impl(c, n, site, old_id, new_id, energy);
}

int ceo_undo_changes(ceo_chain* c) {
  c->trk_valid = 0;
  undo_impl(c);
  return CEO_OK;
}

void ceo_clear_history(ceo_chain* c) { c->buffer_size = 1; }

void ceo_get_cf(const ceo_chain* c, double* out) {
  memcpy(out, ring_slot(c, current_slot(c)), sizeof(double) * (size_t)c->m->n_cf);
}

void ceo_set_ecis(ceo_chain* c, const double* eci) {
  memcpy(c->eci, eci, sizeof(double) * (size_t)c->m->n_cf);
}

void ceo_get_occupation(const ceo_chain* c, int8_t* out) { memcpy(out, c->occ, (size_t)c->m->n_sites); }

/* ------------------------------------------------------------- Philox */
#define PHILOX_M0 0xD2E7470EE14C6C93ULL
#define PHILOX_M1 0xCA5A826395121157ULL
#define PHILOX_W0 0x9E3779B97F4A7C15ULL
#define PHILOX_W1 0xBB67AE8584CAA73BULL

void ceo_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  uint64_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    __uint128_t p0 = (__uint128_t)PHILOX_M0 * x0;
    __uint128_t p1 = (__uint128_t)PHILOX_M1 * x2;
    uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    x0 = hi1 ^ x1 ^ k0;
    x1 = lo1;
    x2 = hi0 ^ x3 ^ k1;
    x3 = lo0;
  }
  out[0] = x0;
  out[1] = x1;
  out[2] = x2;
  out[3] = x3;
}

static inline uint64_t bounded64(uint64_t w, uint64_t n) { return (uint64_t)(((__uint128_t)w * n) >> 64); }
static inline uint32_t bounded32(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }
static inline double unit_double(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }

/* --------------------------------------------------------- Metropolis */
/* Species index tracker of a canonical chain (built once, then kept by the
 * moves) and the species present, in id order. */
static int canonical_setup(ceo_chain* c, int* present) {
  const ceo_model* m = c->m;
  const int N = m->n_sites;
  int* count = c->trk_count;
  int* trk_off = c->trk_off;
  int* tracker = c->trk;
  int n_present = 0;
  if (!c->trk_valid) {
    memset(count, 0, sizeof(int) * (size_t)m->n_species);
    for (int i = 0; i < N; ++i)
      if (!m->background[i]) count[c->occ[i]] += 1;
    trk_off[0] = 0;
    for (int s = 0; s < m->n_species; ++s) trk_off[s + 1] = trk_off[s] + count[s];
    int* fill = (int*)calloc((size_t)(unsigned)m->n_species + 1, sizeof(int));
    for (int i = 0; i < N; ++i) {
      if (m->background[i]) continue;
      int s = c->occ[i];
      tracker[trk_off[s] + fill[s]++] = i;
    }
    free(fill);
    c->trk_valid = 1;
  }
  for (int s = 0; s < m->n_species; ++s)
    if (count[s] > 0 && n_present < 64) present[n_present++] = s;
  return n_present;
}

/* Constraint checks of the reference on a proposed change, restated
 * literally: FixedElement (mc_constraints.py:84-105), ConstrainElementByTag
 * (:171-190), and PairConstraint::elems_in_pair (pair_constraint.cpp:14-46),
 * which applies the change to a copy of the symbols and scans every site of
 * species e1 for a member of species e2. */
static int hooks_valid(const ceo_chain* c, const ceo_hooks* h, int n_ch, const int* sites, const int* olds,
                       const int* news, int8_t* scratch) {
  const int N = c->m->n_sites;
  for (int k = 0; k < n_ch; ++k) {
    if (((h->fixed_species >> olds[k]) & 1u) || ((h->fixed_species >> news[k]) & 1u)) return 0;
    if (h->allowed && !((h->allowed[sites[k]] >> news[k]) & 1u)) return 0;
  }
  if (h->pair_e1 >= 0) {
    memcpy(scratch, c->occ, (size_t)N);
    for (int k = 0; k < n_ch; ++k) scratch[sites[k]] = (int8_t)news[k];
    for (int i = 0; i < N; ++i) {
      if (scratch[i] != h->pair_e1) continue;
      for (int q = 0; q < h->pair_nm; ++q)
        if (scratch[h->pair_nb[(size_t)i * h->pair_nm + q]] == h->pair_e2) return 0;
    }
  }
  return 1;
}

static int run_impl(ceo_chain* c, int kind, uint64_t seed, uint64_t chain_uid, uint64_t move_begin,
                    int64_t n_moves, double temperature, uint8_t* accept_bits, double* e_trace,
                    int64_t trace_every, int64_t* n_accept, ceo_hooks* h) {
  const ceo_model* m = c->m;
  const int N = m->n_sites;
  int present[64];
  int n_present = 0;
  int* movable = NULL;
  int n_movable = 0;
  int64_t acc = 0;
  int rc = CEO_OK;
  int* count = c->trk_count;
  int* trk_off = c->trk_off;
  int* tracker = c->trk;
  if (kind == 0) {
    n_present = canonical_setup(c, present);
    if (n_present < 2) return fail(CEO_EINVAL, "canonical moves need at least two species present");
  } else {
    c->trk_valid = 0;
    movable = (int*)malloc(sizeof(int) * (size_t)(N ? N : 1));
    for (int i = 0; i < N; ++i)
      if (!m->background[i]) movable[n_movable++] = i;
    if (m->n_species < 2 || n_movable == 0) {
      free(movable);
      return fail(CEO_EINVAL, "SGC moves need two species and a movable site");
    }
  }
  const int attempts = h ? h->max_attempts : 1;
  int8_t* scratch = h ? (int8_t*)malloc((size_t)(N ? N : 1)) : NULL;
  int bias_cnt = 0;
  if (h && h->bias_species >= 0)
    for (int i = 0; i < N; ++i) bias_cnt += c->occ[i] == h->bias_species;
  if (h) h->failed_move = -1;
  const double kT = temperature * CEO_KB;
  const int recycle = h && h->recycle;
  const int B = m->n_bfs;
  int sing_idx[8] = {0};
  double s0[8] = {0.0};
  for (int k = 0; k < m->n_cf; ++k)
    if (m->cf_kind[k] == 1 && m->cf_dec[k] < 8) sing_idx[m->cf_dec[k]] = k;
  /* the current energy carries the bias of the state (montecarlo.py:279-282) */
  double e_cur = ceo_get_energy(c);
  if (h && h->bias_species >= 0) e_cur += h->bias_table[bias_cnt];
  const uint64_t key[2] = {seed, chain_uid};
  for (int64_t t = 0; t < n_moves; ++t) {
    int sites[2], olds[2], news[2], n_ch = 0;
    int sa = 0, sb = 0, a = 0, b = 0;
    uint64_t w[4];
    int ok = 0;
    /* proposal k of move t: counter (t, k, 0, 0); the first valid one is
     * made (montecarlo.py:970-991) */
    for (int k = 0; k < attempts && !ok; ++k) {
      const uint64_t ctr[4] = {move_begin + (uint64_t)t, (uint64_t)k, 0, 0};
      ceo_philox4x64(ctr, key, w);
      if (kind == 0) {
        uint32_t ia = bounded32((uint32_t)(w[0] >> 32), (uint32_t)n_present);
        uint32_t ib = (ia + 1 + bounded32((uint32_t)w[0], (uint32_t)(n_present - 1))) % (uint32_t)n_present;
        a = present[ia];
        b = present[ib];
        sa = (int)bounded64(w[1], (uint64_t)count[a]);
        sb = (int)bounded64(w[2], (uint64_t)count[b]);
        sites[0] = tracker[trk_off[a] + sa];
        sites[1] = tracker[trk_off[b] + sb];
        olds[0] = a;
        news[0] = b;
        olds[1] = b;
        news[1] = a;
        n_ch = 2;
      } else {
        int site = movable[bounded64(w[1], (uint64_t)n_movable)];
        int old = c->occ[site];
        int nw = (int)((old + 1 + (int)bounded32((uint32_t)w[0], (uint32_t)(m->n_species - 1))) % m->n_species);
        sites[0] = site;
        olds[0] = old;
        news[0] = nw;
        n_ch = 1;
      }
      ok = h ? hooks_valid(c, h, n_ch, sites, olds, news, scratch) : 1;
    }
    if (!ok) {
      h->failed_move = (int64_t)(move_begin + (uint64_t)t);
      rc = fail(CEO_ERUNTIME, "Did not manage to produce a trial move that does not violate any of the constraints");
      break;
    }
    const double u = unit_double(w[3]);
    double e_new = 0.0;
    if (recycle) {
      /* singlets before the move (sgc_montecarlo.py:69) */
      const double* cf0 = ring_slot(c, current_slot(c));
      for (int d = 0; d < B && d < 8; ++d) s0[d] = cf0[sing_idx[d]];
    }
    rc = calculate_impl(c, n_ch, sites, olds, news, &e_new);
    if (rc != CEO_OK) break;
    int dcnt = 0;
    if (h && h->bias_species >= 0) {
      for (int k = 0; k < n_ch; ++k) dcnt += (news[k] == h->bias_species) - (olds[k] == h->bias_species);
      e_new += h->bias_table[bias_cnt + dcnt];
    }
    int accept;
    double w0 = 0.0, w1 = 0.0;
    const double e_before = e_cur;
    if (recycle) {
      /* util.py:41-63: w = exp(-beta (E - min E)) over the current and the
       * proposed state; the new state is drawn from w / sum(w) by
       * searchsorted(cumsum, u), i.e. the proposed one when u > w0/(w0+w1) */
      const double beta = 1.0 / (CEO_KB * temperature);
      const double emin = e_cur < e_new ? e_cur : e_new;
      w0 = exp(-beta * (e_cur - emin));
      w1 = exp(-beta * (e_new - emin));
      accept = u > w0 / (w0 + w1);
    } else if (e_new < e_cur) {
      accept = 1;
    } else {
      const double energy_diff = e_new - e_cur;
      const double probability = exp(-energy_diff / kT);
      accept = u <= probability;
    }
    if (accept) {
      ceo_clear_history(c);
      e_cur = e_new;
      bias_cnt += dcnt;
      acc += 1;
      if (kind == 0) {
        tracker[trk_off[a] + sa] = sites[1];
        tracker[trk_off[b] + sb] = sites[0];
      }
      if (accept_bits) accept_bits[t >> 3] |= (uint8_t)(1u << (t & 7));
    } else {
      undo_impl(c);
    }
    if (recycle && h->rec_sums) {
      /* montecarlo.py:805-811 and mc_observers.py:231-252: energies of the
       * current and the proposed state; singlets before and after the move */
      double* r = h->rec_sums;
      const double W = w0 + w1;
      const double* cf1 = ring_slot(c, current_slot(c));
      r[0] += 1.0;
      r[1] += (e_before * w0 + e_new * w1) / W;
      r[2] += (e_before * e_before * w0 + e_new * e_new * w1) / W;
      for (int d = 0; d < B && d < 8; ++d) {
        const double s1 = cf1[sing_idx[d]];
        r[3 + d] += (s0[d] * w0 + s1 * w1) / W;
        r[3 + B + d] += (s0[d] * s0[d] * w0 + s1 * s1 * w1) / W;
        r[3 + 2 * B + d] += (s0[d] * e_before * w0 + s1 * e_new * w1) / W;
      }
    }
    if (e_trace && trace_every > 0 && ((t + 1) % trace_every) == 0) e_trace[(t + 1) / trace_every - 1] = e_cur;
  }
  free(movable);
  free(scratch);
  if (n_accept) *n_accept = acc;
  return rc;
}

int ceo_run(ceo_chain* c, int kind, uint64_t seed, uint64_t chain_uid, uint64_t move_begin,
            int64_t n_moves, double temperature, uint8_t* accept_bits, double* e_trace,
            int64_t trace_every, int64_t* n_accept) {
  return run_impl(c, kind, seed, chain_uid, move_begin, n_moves, temperature, accept_bits, e_trace, trace_every,
                  n_accept, NULL);
}

int ceo_run_hooked(ceo_chain* c, int kind, uint64_t seed, uint64_t chain_uid, uint64_t move_begin,
                   int64_t n_moves, double temperature, uint8_t* accept_bits, double* e_trace,
                   int64_t trace_every, int64_t* n_accept, ceo_hooks* hooks) {
  if (!hooks || hooks->max_attempts <= 0) return fail(CEO_EINVAL, "Max. constraint attempts has to be > 0!");
  return run_impl(c, kind, seed, chain_uid, move_begin, n_moves, temperature, accept_bits, e_trace, trace_every,
                  n_accept, hooks);
}

int ceo_run_many(ceo_chain** chains, int n_chains, int kind, uint64_t seed,
                 const uint64_t* chain_uids, uint64_t move_begin, int64_t n_moves,
                 const double* temperatures, int n_threads, int64_t* n_accept) {
  int rc = CEO_OK;
  if (n_threads < 1) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 1) reduction(| : rc)
  for (int i = 0; i < n_chains; ++i) {
    int64_t acc = 0;
    rc |= ceo_run(chains[i], kind, seed, chain_uids[i], move_begin, n_moves, temperatures[i], NULL, NULL,
                  0, &acc);
    if (n_accept) n_accept[i] = acc;
  }
  return rc;
}

/* ------------------------------------------------------- Wang-Landau */
/* Lower edge of a bin, Histogram::get_energy (histogram.cpp:61-64). */
static double wl_bin_energy(const ceo_wl* w, int bin) {
  return w->e_min + ((w->e_max - w->e_min) * bin) / (double)(unsigned)w->n_bins;
}

/* Truncated (e - Emin) * Nbins / (Emax - Emin) as a double, limited to the
 * int range: the int Histogram::get_bin returns before any range check. */
static double wl_raw_bin(const ceo_wl* w, double e) {
  double t = trunc((e - w->e_min) * (double)w->n_bins / (w->e_max - w->e_min));
  if (t < -2147483648.0) t = -2147483648.0;
  if (t > 2147483647.0) t = 2147483647.0;
  return t;
}

/* The walker's state: occupation, correlation functions, species tracker. */
static void wl_snapshot(const ceo_chain* c, ceo_wl* w, int slot) {
  const ceo_model* m = c->m;
  const int N = m->n_sites;
  memcpy(w->snap_occ + (size_t)slot * N, c->occ, (size_t)N);
  memcpy(w->snap_cf + (size_t)slot * m->n_cf, ring_slot(c, current_slot(c)), sizeof(double) * (size_t)m->n_cf);
  memcpy(w->snap_trk + (size_t)slot * N, c->trk, sizeof(int) * (size_t)c->trk_off[m->n_species]);
  w->snap_bin[slot] = w->cur_bin;
}

/* WangLandauSampler::set_updaters (wang_landau_sampler.cpp:585-600) for one
 * walker: the recorded state and its bin replace the walker's; the sub-bin
 * energy is left as it was, as in the reference. */
static void wl_restore(ceo_chain* c, ceo_wl* w, int slot) {
  const ceo_model* m = c->m;
  const int N = m->n_sites;
  memcpy(c->occ, w->snap_occ + (size_t)slot * N, (size_t)N);
  memcpy(ring_slot(c, current_slot(c)), w->snap_cf + (size_t)slot * m->n_cf, sizeof(double) * (size_t)m->n_cf);
  c->buffer_size = 1;
  memcpy(c->trk, w->snap_trk + (size_t)slot * N, sizeof(int) * (size_t)c->trk_off[m->n_species]);
  w->cur_bin = w->snap_bin[slot];
}

/* Histogram::update (histogram.cpp:71-90): the bin is known, visited once
 * more, and its ln g grows by f.  Adaptive windows first record the
 * walker's state at the bin's first visit (adaptive_windows.cpp:240-260);
 * `sub` adds the sub-bin count of the step's closing update
 * (Histogram::update_sub_bin, histogram.cpp:92-100). */
static void wl_visit(const ceo_chain* c, ceo_wl* w, int bin, int sub) {
  if (w->adaptive && !w->saved[bin]) {
    wl_snapshot(c, w, bin);
    w->saved[bin] = 1;
  }
  w->known[bin] = 1;
  w->hist[bin] += 1;
  w->logdos[bin] += w->f;
  if (sub && w->adaptive) {
    const double lo = wl_bin_energy(w, bin), hi = wl_bin_energy(w, bin + 1);
    const double t = trunc((w->cur_energy - lo) * 10.0 / (hi - lo));
    if (t >= 0.0 && t < 10.0) w->sub_hist[10 * bin + (int)t] += 1;
  }
}

/* Histogram::is_flat (histogram.cpp:102-128): over the known bins, the
 * smallest count (starting from 100000000) exceeds criteria times the mean;
 * the sum is an unsigned 32-bit integer. */
static int wl_is_flat(const ceo_wl* w) {
  uint32_t sum = 0, minimum = 100000000u, count = 0;
  for (int i = 0; i < w->n_bins; ++i) {
    if (!w->known[i]) continue;
    sum += (uint32_t)w->hist[i];
    count += 1;
    if ((uint32_t)w->hist[i] < minimum) minimum = (uint32_t)w->hist[i];
  }
  const double mean = (double)sum / (double)count;
  return (double)minimum > w->flatness * mean;
}

/* One canonical swap proposal of counter t: sites, species and the energy
 * after it (the change stays applied; the caller keeps or undoes it). */
typedef struct wl_prop {
  int a, b, sa, sb, sites[2];
  double u, e_new;
} wl_prop;

static int wl_propose(ceo_chain* c, const int* present, int n_present, const uint64_t key[2], uint64_t ctr0,
                      wl_prop* p) {
  const uint64_t ctr[4] = {ctr0, 0, 0, 0};
  uint64_t r[4];
  ceo_philox4x64(ctr, key, r);
  if (n_present < 0) {
    /* semi-grand-canonical flip: present[] holds the -n_present movable sites */
    const ceo_model* m = c->m;
    const int site = present[bounded64(r[1], (uint64_t)(-n_present))];
    const int old = c->occ[site];
    const int nw = (int)((old + 1 + (int)bounded32((uint32_t)r[0], (uint32_t)(m->n_species - 1))) % m->n_species);
    p->a = old;
    p->b = nw;
    p->sa = p->sb = -1;
    p->sites[0] = p->sites[1] = site;
    p->u = unit_double(r[3]);
    return calculate_impl(c, 1, &site, &old, &nw, &p->e_new);
  }
  uint32_t ia = bounded32((uint32_t)(r[0] >> 32), (uint32_t)n_present);
  uint32_t ib = (ia + 1 + bounded32((uint32_t)r[0], (uint32_t)(n_present - 1))) % (uint32_t)n_present;
  p->a = present[ia];
  p->b = present[ib];
  p->sa = (int)bounded64(r[1], (uint64_t)c->trk_count[p->a]);
  p->sb = (int)bounded64(r[2], (uint64_t)c->trk_count[p->b]);
  p->sites[0] = c->trk[c->trk_off[p->a] + p->sa];
  p->sites[1] = c->trk[c->trk_off[p->b] + p->sb];
  p->u = unit_double(r[3]);
  const int olds[2] = {p->a, p->b}, news[2] = {p->b, p->a};
  return calculate_impl(c, 2, p->sites, olds, news, &p->e_new);
}

static void wl_keep(ceo_chain* c, const wl_prop* p) {
  if (p->sa < 0) return; /* a flip: no species tracker */
  c->trk[c->trk_off[p->a] + p->sa] = p->sites[1];
  c->trk[c->trk_off[p->b] + p->sb] = p->sites[0];
}

/* AdaptiveWindowHistogram::distribute_random_walkers_evenly
 * (adaptive_windows.cpp:262-318) with one walker: when more than one bin
 * of the window has a recorded state, the walker moves to the recorded
 * state nearest the window's middle (the upper side on a tie). */
static void wl_distribute(ceo_chain* c, ceo_wl* w) {
  const unsigned upper = (unsigned)w->upper;
  unsigned n_known = 0;
  for (unsigned i = 0; i < upper; ++i) n_known += w->saved[i] ? 1u : 0u;
  if (n_known <= 1u) return;
  const double delta = upper / 2.0;
  const unsigned indx = (unsigned)delta;
  unsigned shift_pos = 2 * upper, shift_neg = 2 * upper;
  for (unsigned j = indx; j < upper; ++j)
    if (w->saved[j]) {
      shift_pos = j - indx;
      break;
    }
  for (int j = (int)indx; j >= 0; --j)
    if (w->saved[j]) {
      shift_neg = indx - (unsigned)j;
      break;
    }
  wl_restore(c, w, (int)(shift_pos <= shift_neg ? indx + shift_pos : indx - shift_neg));
}

/* WangLandauSampler::run_until_valid_energy (wang_landau_sampler.cpp:434-560)
 * with one walker outside the window: swaps toward the window (below emax:
 * any swap to a higher energy under emax; above: any swap to a lower energy
 * over emin, measured from the current bin's lower edge) until a proposal
 * lands in the window, at most 10^7 swaps. */
static int wl_walk_into_window(ceo_chain* c, ceo_wl* w, const int* present, int n_present, const uint64_t key[2],
                               uint64_t move_begin, int64_t* t) {
  if (w->cur_bin >= 0 && w->cur_bin < w->upper) return CEO_OK;
  const double emin = w->e_min, emax = wl_bin_energy(w, w->upper);
  for (int64_t i = 0; i < 10000000; ++i) {
    wl_prop p;
    const int rc = wl_propose(c, present, n_present, key, move_begin + (uint64_t)*t, &p);
    *t += 1;
    w->n_walk += 1;
    if (rc != CEO_OK) return rc;
    const double raw = wl_raw_bin(w, p.e_new);
    const double cur_e = wl_bin_energy(w, w->cur_bin);
    const int ok = cur_e >= emax ? (p.e_new < cur_e && p.e_new > emin) : (p.e_new > cur_e && p.e_new < emax);
    if (ok) {
      wl_keep(c, &p);
      w->cur_bin = (int)raw;
    } else {
      undo_impl(c);
    }
    ceo_clear_history(c);
    if (raw >= 0.0 && raw < (double)w->upper) return CEO_OK;
  }
  return fail(CEO_ERUNTIME, "Could not find any energy state inside the histogram range!");
}

/* AdaptiveWindowHistogram::is_flat (adaptive_windows.cpp:38-140): scanning
 * down from upper - n_overlap, the window of known bins above the scan
 * point is flat while its minimum exceeds criteria times its mean; the scan
 * stops at the first bin that breaks this.  A scan that reaches bin 0 flat
 * means the window has converged.  Otherwise, if some window of at least
 * min_width bins (or of every known bin) was flat, the upper part is cut
 * off: ln g of the n_overlap bins above the cut is saved, the walker's
 * state is recorded at the first cut, and the walker moves into the new
 * window [0, cut + n_overlap).  With n_overlap >= upper the scan is empty
 * (the reference's unsigned start would wrap). */
static int wl_window_flat(ceo_chain* c, ceo_wl* w, const int* present, int n_present, const uint64_t key[2],
                          uint64_t move_begin, int64_t* t, int* rc) {
  const int top = w->upper > w->n_overlap ? w->upper - w->n_overlap : 0;
  unsigned max_possible = 0;
  for (int i = 0; i < top; ++i) max_possible += w->known[i] ? 1u : 0u;
  uint32_t sum = 0, count = 0, minimum = 100000000u;
  int part = 0, min_ok = 0;
  int fnc = 0;
  for (int i = top; i > 0; --i) {
    if (!w->known[i - 1]) continue;
    const uint32_t h = (uint32_t)w->hist[i - 1];
    sum += h;
    count += 1;
    if (h < minimum) minimum = h;
    const double mean = (double)sum / count;
    min_ok = (double)minimum > w->flatness * mean;
    if ((count > (uint32_t)w->min_width || count >= max_possible) && min_ok) part = 1;
    if (!min_ok) {
      fnc = i - 1;
      break;
    }
  }
  if (fnc == 0 && min_ok) return 1;
  if (!part) return 0;
  /* is_valid_window (adaptive_windows.cpp:222-236) */
  if (fnc <= w->min_width) return 0;
  int n_known = 0;
  for (int i = 0; i < fnc; ++i) n_known += w->known[i] ? 1 : 0;
  if (n_known <= w->min_width / 2) return 0;
  const int e = w->n_edges++;
  w->edges[e] = fnc + 1;
  w->edge_n[e] = w->n_overlap;
  for (int k = 0; k < w->n_overlap; ++k) w->edge_vals[(size_t)e * w->n_bins + k] = w->logdos[fnc + 1 + k];
  if (w->upper == w->n_bins) {
    wl_snapshot(c, w, w->n_bins);
    w->has_first = 1;
  }
  w->upper = fnc + w->n_overlap + 1;
  wl_distribute(c, w);
  *rc = wl_walk_into_window(c, w, present, n_present, key, move_begin, t);
  return 0;
}

/* AdaptiveWindowHistogram::reset and make_dos_continous
 * (adaptive_windows.cpp:142-185): the full range is back in the window; at
 * every cut, in the order they were made, the saved values return to the
 * bins above it and everything at or below the cut shifts so that ln g is
 * continuous there (later cuts' saved values shift with it); the walker
 * returns to the state recorded at the first cut. */
static void wl_window_reset(ceo_chain* c, ceo_wl* w) {
  w->upper = w->n_bins;
  for (int e = 0; e < w->n_edges; ++e) {
    const int E = w->edges[e], n = w->edge_n[e];
    double* v = w->edge_vals + (size_t)e * w->n_bins;
    for (int j = 1; j < n; ++j) w->logdos[E + j] = v[j];
    const double diff = w->logdos[E] - v[0];
    for (int j = 0; j < E + 1; ++j) w->logdos[j] -= diff;
    for (int e2 = e + 1; e2 < w->n_edges; ++e2)
      for (int k = 0; k < w->edge_n[e2]; ++k) w->edge_vals[(size_t)e2 * w->n_bins + k] -= diff;
  }
  w->n_edges = 0;
  if (w->has_first) wl_restore(c, w, w->n_bins);
}

int ceo_wl_run(ceo_chain* c, ceo_wl* w, uint64_t seed, uint64_t chain_uid, uint64_t move_begin,
               int64_t n_checks, int64_t check_every, uint8_t* accept_bits, int64_t bits_cap, int64_t* n_used) {
  int present_buf[64];
  int* present = present_buf;
  int n_present;
  int* movable = NULL;
  if (n_used) *n_used = 0;
  if (w->kind == 1) {
    /* the movable sites, as ceo_run's SGC moves pick them */
    const ceo_model* m = c->m;
    movable = (int*)malloc(sizeof(int) * (size_t)(m->n_sites > 0 ? m->n_sites : 1));
    int n_movable = 0;
    for (int i = 0; i < m->n_sites; ++i)
      if (!m->background[i]) movable[n_movable++] = i;
    present = movable;
    n_present = -n_movable;
    c->trk_valid = 0; /* flips change the composition */
    if (n_movable < 1) {
      free(movable);
      return fail(CEO_EINVAL, "no movable site");
    }
  } else {
    n_present = canonical_setup(c, present);
    if (n_present < 2) return fail(CEO_EINVAL, "canonical moves need at least two species present");
  }
  const char* bad = NULL;
  if (w->n_bins < 1 || !(w->e_max > w->e_min)) bad = "invalid Wang-Landau histogram";
  else if (w->cur_bin < 0 || w->cur_bin >= w->n_bins) bad = "current bin outside the histogram";
  else if (w->adaptive && (w->upper < 1 || w->upper > w->n_bins || w->n_overlap < 0 || w->n_overlap > w->n_bins))
    bad = "invalid adaptive window";
  if (bad) {
    free(movable);
    return fail(CEO_EINVAL, bad);
  }
  if (!w->adaptive) w->upper = w->n_bins;
  const uint64_t key[2] = {seed, chain_uid};
  int64_t t = 0;
  int rc = CEO_OK;
  if (w->converged) { /* a converged walker makes no more moves */
    free(movable);
    return CEO_OK;
  }
  for (int64_t chk = 0; chk < n_checks && rc == CEO_OK; ++chk) {
    /* WangLandauSampler::step, wang_landau_sampler.cpp:232-317 */
    for (int64_t s = 0; s < check_every; ++s) {
      wl_prop p;
      const int64_t tm = t++;
      rc = wl_propose(c, present, n_present, key, move_begin + (uint64_t)tm, &p);
      if (rc != CEO_OK) break;
      w->n_moves += 1;
      w->steps += 1;
      w->since_flat += 1;
      const double raw = wl_raw_bin(w, p.e_new);
      if (w->adaptive) w->bin_change += fabs(raw - (double)w->cur_bin); /* only set_overlap reads it */
      const int bin = raw < 0.0 ? -1 : (raw > (double)w->n_bins ? w->n_bins + 1 : (int)raw);
      if (bin < 0 || bin >= w->upper) {
        w->n_outside += 1;
        undo_impl(c);
        if (w->is_first) continue;
        if (bin > w->n_bins) wl_visit(c, w, w->cur_bin, 0);
        continue;
      }
      if (bin == w->cur_bin) w->n_self += 1;
      /* Histogram::get_dos_ratio_old_divided_by_new (histogram.cpp:130-135) */
      const double diff = w->logdos[w->cur_bin] - w->logdos[bin];
      const double ratio = diff > 0.0 ? 1.0 : exp(diff);
      const int accept = (p.u < ratio) || w->is_first;
      if (accept) {
        w->cur_bin = bin;
        w->is_first = 0;
        w->cur_energy = p.e_new;
        w->n_accept += 1;
        wl_keep(c, &p);
        if (accept_bits && tm < bits_cap) accept_bits[tm >> 3] |= (uint8_t)(1u << (tm & 7));
      } else {
        undo_impl(c);
      }
      ceo_clear_history(c);
      wl_visit(c, w, w->cur_bin, 1);
    }
    if (rc != CEO_OK) break;
    /* WangLandauSampler::run, wang_landau_sampler.cpp:337-394 */
    const int flat = w->adaptive ? wl_window_flat(c, w, present, n_present, key, move_begin, &t, &rc) : wl_is_flat(w);
    if (rc != CEO_OK) break;
    if (flat) {
      const int64_t took = w->since_flat;
      const double mc_time = (double)w->steps / (double)(unsigned)w->n_bins;
      if (w->inverse_time && w->prev_conv >= 0 && took > w->prev_conv) {
        if (!w->inv_active) w->inv_factor = mc_time * w->f;
        w->inv_active = 1;
      }
      w->prev_conv = took;
      w->f = w->inv_active ? w->inv_factor / mc_time : w->f / 2.0;
      for (int i = 0; i < w->n_bins; ++i) w->hist[i] = 0;
      if (w->adaptive) {
        wl_window_reset(c, w);
        /* set_overlap(2 * average bin change per step) */
        const double ov = 2.0 * (w->bin_change / (double)took);
        w->n_overlap = ov >= (double)w->n_bins ? w->n_bins : (int)(unsigned)ov;
      }
      w->bin_change = 0.0;
      w->since_flat = 0;
      w->n_flat += 1;
    }
    if (w->f < w->f_min) {
      w->converged = 1;
      break;
    }
  }
  if (n_used) *n_used = t;
  free(movable);
  return rc;
}
