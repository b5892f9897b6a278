// Host side of libcemcgpu: the C ABI declared in include/cemc_gpu.h.
//
// cemc_model    parses the flattened settings (what CEUpdater::init reads,
//               cpp/src/ce_updater.cpp:32-234) into sub-cluster records,
//               per-cluster energy tables with the ECIs folded in, and the
//               correlation-function tables; uploads them to HBM.
// cemc_updater  one lattice with the reference's history ring buffer
//               (cpp/src/cf_history_tracker.cpp:46-149) and dispatch
//               (cpp/src/ce_updater.cpp:471-550, 831-852); the arithmetic
//               runs in updater_kernel.
// cemc_batch    device-resident Markov chains advanced by sweep_kernel.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <tuple>
#include <string>
#include <vector>

#include "cemc_gpu.h"
#include "ce_kernels.h"

using namespace cemc;

namespace {

thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_CHECK(expr)                                                                         \
  do {                                                                                          \
    hipError_t err_ = (expr);                                                                   \
    if (err_ != hipSuccess) throw DeviceError(std::string(#expr) + ": " + hipGetErrorString(err_)); \
  } while (0)

// A pool without room for a new chain (slot or device memory).
struct PoolFull : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class F>
int guarded(F&& f) {
  try {
    f();
    return CEMC_OK;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return CEMC_EINVAL;
  } catch (const DeviceError& e) {
    g_last_error = e.what();
    return CEMC_EDEVICE;
  } catch (const PoolFull& e) {
    g_last_error = e.what();
    return CEMC_EFULL;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return CEMC_ERUNTIME;
  }
}

void require(bool ok, const std::string& msg) {
  if (!ok) throw std::invalid_argument(msg);
}

template <class T>
T* upload(const std::vector<T>& v) {
  T* p = nullptr;
  HIP_CHECK(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

template <class T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

constexpr int kMaxHistory = 1000;  // CFHistoryTracker::max_history (cf_history_tracker.hpp:56)
constexpr uint32_t kGtabLdsBudget = 96u * 1024u;
constexpr uint32_t kUnitLds = 160u * 1024u;  // LDS of one compute unit (gfx950)

// Chains of pool launches in flight, per device.  The job server runs one
// pool per lane, each launching from its own thread onto the same device: a
// pool choosing its launch shapes counts the other lanes' chains as sharing
// the device with its own.
constexpr int kMaxDevices = 64;
std::atomic<int64_t> g_pool_chains[kMaxDevices];

int current_device_index() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  return dev >= 0 && dev < kMaxDevices ? dev : 0;
}

// Dynamic LDS a workgroup may use on the current device (160 KB on gfx950).
uint32_t lds_limit() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  size_t v = prop.sharedMemPerBlock;
  // dynamic part: what remains after the sweep kernel's static tables
  return uint32_t(std::min<size_t>(v, 160u * 1024u)) - sweep_static_lds() - 256u;
}

}  // namespace

// ============================================================================
struct cemc_model {
  int n_sites = 0, n_species = 0, n_bfs = 0, n_groups = 0, n_cf = 0, n_entries = 0, K = 0;
  bool periodic = false;
  int dims[3] = {0, 0, 0};
  int pad_h = 0;
  bool spec_ok = false;   // speculative rounds allowed (see ModelDev::imask)
  uint64_t imask[2] = {0, 0};
  int max_members = 3;  // most members other than the reference site over the records (1-3)
  bool field_ok = false;  // local-field path allowed (see ModelDev::duals)
  double fx_scale = 0.0;
  std::vector<int64_t> gtab_fx;
  std::vector<uint32_t> duals;
  int pdims[3] = {0, 0, 0};
  std::vector<int> signed_cells;
  std::vector<double> basis, eci, gtab;
  std::vector<int32_t> site_group, cf_kind, cf_dec, cf_entry, cf_eq_off, movable;
  std::vector<uint8_t> background;
  std::vector<uint32_t> eq_decos, col_cells;
  std::vector<int32_t> entry_group, entry_size, entry_M, entry_norm, entry_goff;
  std::vector<int32_t> trans;
  std::vector<SubRec> recs;
  std::vector<int> rec_entry;  // cluster entry of each record
  std::vector<FatRec> rec_slots, frecs;
  std::vector<ClusterEntry> entries;
  std::vector<int32_t> grp_count;
  std::vector<int32_t> grp_rec_off, grp_rec_cnt;
  // device
  SubRec* d_recs = nullptr;
  FatRec* d_frecs = nullptr;
  double* d_gtab = nullptr;
  std::vector<double> gtab_il;  // three or four species: gtab with the read slices of a code side by side
  double* d_gtab_il = nullptr;
  int64_t* d_gtab_fx = nullptr;
  uint32_t* d_duals = nullptr;
  int32_t* d_trans = nullptr;
  uint32_t* d_cells = nullptr;
  int8_t* d_site_group = nullptr;
  uint8_t* d_background = nullptr;
  int32_t* d_movable = nullptr;
  int32_t *d_cf_kind = nullptr, *d_cf_dec = nullptr, *d_cf_entry = nullptr, *d_cf_eq_off = nullptr;
  uint32_t* d_eq_decos = nullptr;
  ClusterEntry* d_entries = nullptr;
  double* d_eci = nullptr;
  ModelDev view{};
  ModelDev* d_view = nullptr;

  int pow_s(int q) const {
    int p = 1;
    for (int i = 0; i < q; ++i) p *= n_species;
    return p;
  }

  void build_gtab(const std::vector<double>& J);
  void build_gtab_il();
  void build_gtab_fx();
  void build_duals();
  int code_bits() const;    // bits per species in an energy-table code
  int deco_slices() const;  // decoration slices per table block (>= n_species - 1)
  int read_slices() const;  // slices the sweep kernel reads per record (its kBM)
  void refresh_view();
  ~cemc_model();
};

cemc_model::~cemc_model() {
  dev_free(d_recs);
  dev_free(d_frecs);
  dev_free(d_gtab);
  dev_free(d_gtab_il);
  dev_free(d_gtab_fx);
  dev_free(d_duals);
  dev_free(d_trans);
  dev_free(d_cells);
  dev_free(d_site_group);
  dev_free(d_background);
  dev_free(d_movable);
  dev_free(d_cf_kind);
  dev_free(d_cf_dec);
  dev_free(d_cf_entry);
  dev_free(d_cf_eq_off);
  dev_free(d_eq_decos);
  dev_free(d_entries);
  dev_free(d_eci);
  dev_free(d_view);
}

// Energy tables.  For cluster entry c (size n, group g) and decoration e over
// canonical positions, the weight is
//   W_c[e] = sum_{k on c} mult_k(e) * N * J_k * n / (|E_k| * C_c * T_g)
// where mult_k(e) counts e in the equivalent decorations of CF k.  This is
// the energy form of the per-CF increment (ce_updater.cpp:393-404) times
// N * J_k (ce_updater.cpp:236-242).  With the reference site at canonical
// position r and decoration digit d there, the table stores
//   G_c[r][d][code] = sum_{e: e_r = d} W_c[e] * prod_{q != r} phi[e_q](s_q)
// so a flip a->b of the reference site changes the energy by
//   sum_d (phi_d(b) - phi_d(a)) * G_c[r][d][code].
// Layout for the sweep kernel: code = s_0 | s_1 << k | s_2 << 2k over the
// (up to three) other members in canonical order, k = code bits per species;
// a block holds kBM decoration slices of 2^(3k) codes.  Entries of clusters
// with fewer than three other members repeat over the unused species fields,
// so the kernel never masks them.  Block 0 is all zeros (padding records).
int cemc_model::code_bits() const { return n_species <= 2 ? 1 : (n_species <= 4 ? 2 : 3); }
int cemc_model::deco_slices() const { return n_species <= 2 ? 1 : (n_species <= 4 ? 3 : 7); }
// up to four species every slice of a block holds a basis function in use
// except the third for three species, which the kernel need not read
int cemc_model::read_slices() const { return n_species <= 4 ? n_species - 1 : 7; }

void cemc_model::build_gtab(const std::vector<double>& J) {
  gtab.assign(gtab.size(), 0.0);
  const int S = n_species, B = n_bfs;
  const int k = code_bits(), Pc = 1 << (3 * k), Bp = deco_slices();
  for (int c = 0; c < n_entries; ++c) {
    const int n = entry_size[c];
    const int g = entry_group[c];
    int n_deco = 1;
    for (int i = 0; i < n; ++i) n_deco *= B;
    std::vector<double> W(n_deco, 0.0);
    for (int kk = 0; kk < n_cf; ++kk) {
      if (cf_kind[kk] != CEMC_CF_CLUSTER || cf_entry[g * n_cf + kk] != c) continue;
      const int e0 = cf_eq_off[kk], e1 = cf_eq_off[kk + 1];
      const double w = double(n_sites) * J[kk] * double(n) /
                       (double(e1 - e0) * double(entry_norm[c]) * double(grp_count[g]));
      for (int e = e0; e < e1; ++e) {
        int idx = 0;
        for (int pos = n - 1; pos >= 0; --pos) idx = idx * B + int((eq_decos[e] >> (8 * pos)) & 0xFFu);
        W[idx] += w;
      }
    }
    for (int r = 0; r < n; ++r) {
      int pos_of[3];  // canonical position of member q
      int q = 0;
      for (int pos = 0; pos < n; ++pos)
        if (pos != r) pos_of[q++] = pos;
      const int nm = n - 1;
      for (int idx = 0; idx < n_deco; ++idx) {
        if (W[idx] == 0.0) continue;
        int deco[kMaxClusterSize];
        int t = idx;
        for (int pos = 0; pos < n; ++pos) {
          deco[pos] = t % B;
          t /= B;
        }
        double* G = gtab.data() + entry_goff[c] + (r * Bp + deco[r]) * Pc;
        for (int code = 0; code < Pc; ++code) {
          double prod = W[idx];
          bool valid = true;
          for (int qq = 0; qq < nm; ++qq) {
            const int sp = (code >> (k * qq)) & ((1 << k) - 1);
            if (sp >= S) {
              valid = false;
              break;
            }
            prod *= basis[deco[pos_of[qq]] * S + sp];
          }
          if (valid) G[code] += prod;
        }
      }
    }
  }
}

// For three species the sweep and Wang-Landau kernels read slices 0 and 1 of
// every record's block together: the block's entry (d, code) moves from
// d * 64 + code to 2 * code + d, so the two values of a code are adjacent
// and 16-byte aligned (blocks start at multiples of 3 x 64 doubles).
// The tables with the read slices of each code side by side, for the
// kernels that read three-species tables (two slices: one 16-byte read) and
// four-species tables from HBM (three slices: adjacent values in one or two
// cache lines instead of three).  Blocks keep their size and offsets, so a
// record's table base is the same in both layouts.
void cemc_model::build_gtab_il() {
  gtab_il.clear();
  const int w = read_slices();
  if (w != 2 && w != 3) return;
  const int Pc = 1 << (3 * code_bits()), block = deco_slices() * Pc;  // deco_slices() >= w
  gtab_il.assign(gtab.size(), 0.0);
  for (size_t b = 0; b + size_t(block) <= gtab.size(); b += size_t(block))
    for (int code = 0; code < Pc; ++code)
      for (int d = 0; d < w; ++d) gtab_il[b + size_t(w) * code + d] = gtab[b + size_t(d) * Pc + code];
}

// Integer copy of the energy tables for the local-field path: a site's field
// F^d = sum over its records of G[d][code] is kept exactly as an integer sum
// of round(G * 2^k), k chosen so that any field stays below 2^61 in
// magnitude (the sum over records of each record's largest |G|).  Updates
// then add and subtract integers only: no rounding drift and no dependence
// on the order of concurrent updates.
void cemc_model::build_gtab_fx() {
  if (!field_ok) {
    gtab_fx.assign(1, 0);
    fx_scale = 0.0;
    return;
  }
  const int block = deco_slices() << (3 * code_bits());
  double bound = 0.0;
  for (int r = 0; r < grp_rec_cnt[0]; ++r) {
    const uint32_t gbase = frecs[r].meta & 0xFFFFFFu;
    double mx = 0.0;
    for (int i = 0; i < block; ++i) mx = std::max(mx, std::fabs(gtab[gbase + i]));
    bound += mx;
  }
  int e = 0;
  std::frexp(bound, &e);  // bound < 2^e
  const int k = 61 - e;
  fx_scale = std::ldexp(1.0, -k);
  gtab_fx.resize(gtab.size());
  for (size_t i = 0; i < gtab.size(); ++i) gtab_fx[i] = std::llround(std::ldexp(gtab[i], k));
}

// Duals of the records (see ModelDev::duals): the sub-cluster of record r
// seen from its member q (cell offset o_q) has the other members -o_q and
// o_p - o_q (p != q); the record of the same entry with exactly these
// wrapped offsets is its dual.  The path is allowed only when every dual
// exists and is unique.
void cemc_model::build_duals() {
  field_ok = false;
  duals.clear();
  if (!periodic || n_groups != 1 || int(movable.size()) != n_sites || code_bits() > 2 || n_sites > 65535) return;
  const int cnt = grp_rec_cnt[0];
  const int L[3] = {dims[0], dims[1], dims[2]};
  auto cell = [&](uint32_t packed, int a) { return int((packed >> (8 * a)) & 255u); };
  auto pack = [&](const int (&v)[3]) {
    uint32_t w = 0;
    for (int a = 0; a < 3; ++a) w |= uint32_t(((v[a] % L[a]) + L[a]) % L[a]) << (8 * a);
    return w;
  };
  // record r: entry, member count, wrapped member offsets in code order
  struct RecGeo {
    int entry, nm;
    uint32_t off[3];
  };
  std::vector<RecGeo> geo(cnt);
  for (int r = 0; r < cnt; ++r) {
    geo[r].entry = rec_entry[r];
    geo[r].nm = int((rec_slots[r].meta >> 24) & 3u);
    const int32_t sl[3] = {rec_slots[r].m0, rec_slots[r].m1, rec_slots[r].m2};
    for (int q = 0; q < 3; ++q) geo[r].off[q] = q < geo[r].nm ? col_cells[sl[q]] : 0u;
  }
  std::map<std::vector<uint32_t>, std::vector<int>> by_key;
  for (int r = 0; r < cnt; ++r) {
    std::vector<uint32_t> key(geo[r].off, geo[r].off + geo[r].nm);
    std::sort(key.begin(), key.end());
    key.insert(key.begin(), uint32_t(geo[r].entry));
    by_key[key].push_back(r);
  }
  std::vector<uint32_t> out(3 * size_t(cnt), 0u);
  for (int r = 0; r < cnt; ++r) {
    const int nm = geo[r].nm;
    for (int q = 0; q < nm; ++q) {
      // offsets seen from member q: the reference (-o_q), then the others
      uint32_t want[3];
      int n_want = 0;
      int oq[3], v[3];
      for (int a = 0; a < 3; ++a) {
        oq[a] = cell(geo[r].off[q], a);
        v[a] = -oq[a];
      }
      want[n_want++] = pack(v);
      for (int p = 0; p < nm; ++p) {
        if (p == q) continue;
        for (int a = 0; a < 3; ++a) v[a] = cell(geo[r].off[p], a) - oq[a];
        want[n_want++] = pack(v);
      }
      std::vector<uint32_t> key(want, want + n_want);
      std::sort(key.begin(), key.end());
      key.insert(key.begin(), uint32_t(geo[r].entry));
      auto it = by_key.find(key);
      if (it == by_key.end() || it->second.size() != 1) return;
      const int rd = it->second[0];
      auto pos_of = [&](uint32_t w) {
        for (int t = 0; t < geo[rd].nm; ++t)
          if (geo[rd].off[t] == w) return t;
        return -1;
      };
      uint32_t word = rec_slots[rd].meta & 0xFFFFFFu;
      const int pi = pos_of(want[0]);
      if (pi < 0) return;
      word |= uint32_t(pi) << 24;
      for (int t = 1; t < n_want; ++t) {
        const int pt = pos_of(want[t]);
        if (pt < 0) return;
        word |= uint32_t(pt) << (26 + 2 * (t - 1));
      }
      out[3 * size_t(r) + q] = word;
    }
  }
  duals.swap(out);
  field_ok = true;
}

void cemc_model::refresh_view() {
  ModelDev& v = view;
  std::memset(&v, 0, sizeof v);
  v.n_sites = n_sites;
  v.n_species = n_species;
  v.n_bfs = n_bfs;
  v.n_groups = n_groups;
  v.n_cf = n_cf;
  v.n_entries = n_entries;
  v.K = K;
  v.periodic = periodic ? 1 : 0;
  for (int i = 0; i < 3; ++i) v.dims[i] = dims[i];
  v.pad_h = pad_h;
  for (int i = 0; i < 3; ++i) v.pdims[i] = pdims[i];
  v.spec_ok = spec_ok ? 1 : 0;
  v.imask[0] = imask[0];
  v.imask[1] = imask[1];
  v.has_background = int(movable.size()) != n_sites;
  v.n_movable = int(movable.size());
  v.gtab_size = int(gtab.size());
  v.n_rec = int(recs.size());
  for (int q = 0; q < kMaxClusterSize; ++q) v.pow_s[q] = pow_s(q);
  for (int g = 0; g < n_groups; ++g) {
    v.grp_rec_off[g] = grp_rec_off[g];
    v.grp_rec_cnt[g] = grp_rec_cnt[g];
    v.grp_count[g] = grp_count[g];
  }
  for (size_t i = 0; i < basis.size(); ++i) v.basis[i] = basis[i];
  v.recs = d_recs;
  v.frecs = d_frecs;
  v.gtab = d_gtab;
  v.gtab_il = d_gtab_il ? d_gtab_il : d_gtab;
  v.trans = d_trans;
  v.col_cells = d_cells;
  v.site_group = d_site_group;
  v.background = d_background;
  v.movable = d_movable;
  v.cf_kind = d_cf_kind;
  v.cf_dec = d_cf_dec;
  v.cf_entry = d_cf_entry;
  v.cf_eq_off = d_cf_eq_off;
  v.eq_decos = d_eq_decos;
  v.entries = d_entries;
  v.eci = d_eci;
  v.field_ok = field_ok ? 1 : 0;
  v.fx_scale = fx_scale;
  v.gtab_fx = d_gtab_fx;
  v.duals = d_duals;
  if (!d_view) HIP_CHECK(hipMalloc(&d_view, sizeof(ModelDev)));
  HIP_CHECK(hipMemcpy(d_view, &v, sizeof(ModelDev), hipMemcpyHostToDevice));
}

static void build_model(cemc_model* m, const cemc_model_desc* d) {
  require(d != nullptr, "model description is NULL");
  require(d->n_sites > 0, "n_sites must be positive");
  require(d->n_species >= 2 && d->n_species <= kMaxSpecies, "n_species must be in [2, 8]");
  require(d->n_groups >= 1 && d->n_groups <= kMaxGroups, "n_groups must be in [1, 32]");
  require(d->n_cf >= 1, "at least one correlation function is required");
  require(d->n_trans_cols >= 0 && d->n_trans_cols <= 1024, "at most 1024 translation columns are supported");
  require(d->n_entries >= 0 && d->n_entries <= (1 << 20), "at most 2^20 cluster entries are supported");
  m->n_sites = d->n_sites;
  m->n_species = d->n_species;
  m->n_bfs = d->n_species - 1;
  m->n_groups = d->n_groups;
  m->n_cf = d->n_cf;
  m->n_entries = d->n_entries;
  m->K = d->n_trans_cols;
  const int N = m->n_sites, S = m->n_species, B = m->n_bfs, G = m->n_groups;
  m->basis.assign(d->basis, d->basis + B * S);
  m->site_group.assign(d->site_group, d->site_group + N);
  m->background.assign(N, 0);
  if (d->background)
    for (int i = 0; i < N; ++i) m->background[i] = d->background[i] != 0;
  m->grp_count.assign(G, 0);
  for (int i = 0; i < N; ++i) {
    const int g = m->site_group[i];
    require(g >= -1 && g < G, "site group out of range");
    require(g >= 0 || m->background[i], "a site has not been assigned to any translational symmetry group");
    // trans_symm_group_count counts background sites too (ce_updater.cpp:768-775)
    if (g >= 0) m->grp_count[g] += 1;
    if (!m->background[i]) m->movable.push_back(i);
  }
  require(!m->movable.empty(), "no movable site");
  m->cf_kind.assign(d->cf_kind, d->cf_kind + m->n_cf);
  m->cf_dec.assign(d->cf_dec, d->cf_dec + m->n_cf);
  m->cf_entry.assign(d->cf_entry, d->cf_entry + G * m->n_cf);
  m->cf_eq_off.assign(d->cf_eq_off, d->cf_eq_off + m->n_cf + 1);
  const int n_eq = m->cf_eq_off[m->n_cf];
  require(n_eq >= 0, "bad equivalent decoration offsets");
  m->eq_decos.resize(n_eq);
  for (int e = 0; e < n_eq; ++e) {
    uint32_t packed = 0;
    for (int p = 0; p < 4; ++p) {
      const int v = d->eq_decos[4 * e + p];
      require(v >= 0 && v < std::max(B, 1), "decoration digit out of range");
      packed |= uint32_t(v) << (8 * p);
    }
    m->eq_decos[e] = packed;
  }
  for (int k = 0; k < m->n_cf; ++k) {
    const int kind = m->cf_kind[k];
    require(kind >= 0 && kind <= 2, "unknown correlation-function kind");
    if (kind == CEMC_CF_SINGLET) require(m->cf_dec[k] >= 0 && m->cf_dec[k] < B, "singlet decoration out of range");
    require(m->cf_eq_off[k + 1] >= m->cf_eq_off[k], "bad equivalent decoration offsets");
    if (kind == CEMC_CF_CLUSTER) {
      for (int g = 0; g < G; ++g) {
        const int c = m->cf_entry[g * m->n_cf + k];
        require(c >= -1 && c < m->n_entries, "cluster entry out of range");
        if (c >= 0) {
          require(d->entry_group[c] == g, "cluster entry belongs to another group");
          require(m->cf_eq_off[k + 1] > m->cf_eq_off[k], "cluster correlation function without decoration");
        }
      }
    }
  }
  m->eci.assign(d->eci, d->eci + m->n_cf);
  // translation table
  m->trans.assign(d->trans, d->trans + int64_t(N) * m->K);
  for (int64_t i = 0; i < int64_t(N) * m->K; ++i)
    require(m->trans[i] >= 0 && m->trans[i] < N, "translation table entry out of range");
  // sub-cluster records, grouped by translation group then entry
  m->entry_group.assign(d->entry_group, d->entry_group + m->n_entries);
  m->entry_size.assign(d->entry_size, d->entry_size + m->n_entries);
  m->entry_M.assign(d->entry_M, d->entry_M + m->n_entries);
  m->entry_norm.assign(d->entry_norm, d->entry_norm + m->n_entries);
  m->entries.assign(m->n_entries, ClusterEntry{0, 0, 0, 0});
  m->entry_goff.assign(m->n_entries, 0);
  const int block = m->deco_slices() << (3 * m->code_bits());
  int goff = block;  // block 0: zeros read by padding records
  for (int c = 0; c < m->n_entries; ++c) {
    const int n = m->entry_size[c];
    require(n >= 2 && n <= kMaxClusterSize, "only cluster sizes 2, 3 and 4 are supported");
    require(m->entry_M[c] >= 1, "cluster without sub-clusters");
    require(m->entry_norm[c] >= m->entry_M[c], "bad cluster normalisation");
    m->entry_goff[c] = goff;
    goff += n * block;
  }
  require(goff < (1 << 24), "energy tables too large");
  m->gtab.assign(goff, 0.0);
  m->grp_rec_off.assign(G, 0);
  m->grp_rec_cnt.assign(G, 0);
  for (int g = 0; g < G; ++g) {
    m->grp_rec_off[g] = int(m->recs.size());
    for (int c = 0; c < m->n_entries; ++c) {
      if (m->entry_group[c] != g) continue;
      const int n = m->entry_size[c];
      const int M = m->entry_M[c];
      const int32_t* mem = d->members + d->entry_mem_off[c];
      const int32_t* ord = d->order + d->entry_ord_off[c];
      require(d->entry_mem_off[c + 1] - d->entry_mem_off[c] == M * (n - 1), "members size mismatch");
      require(d->entry_ord_off[c + 1] - d->entry_ord_off[c] == M * n, "order size mismatch");
      m->entries[c] = ClusterEntry{int32_t(m->recs.size()), M, n, m->entry_norm[c]};
      for (int s = 0; s < M; ++s) {
        int r = -1;
        uint32_t seen = 0;
        for (int pos = 0; pos < n; ++pos) {
          const int o = ord[s * n + pos];
          require(o >= 0 && o < n && !(seen & (1u << o)), "order is not a permutation");
          seen |= 1u << o;
          if (o == 0) r = pos;
        }
        uint32_t w0 = uint32_t(n - 2) << 30;
        int q = 0;
        for (int pos = 0; pos < n; ++pos) {
          if (pos == r) continue;
          const int slot = mem[s * (n - 1) + ord[s * n + pos] - 1];
          require(slot >= 0 && slot < m->K, "member column out of range");
          w0 |= uint32_t(slot) << (10 * q);
          ++q;
        }
        const uint32_t gbase = uint32_t(m->entry_goff[c] + r * block);
        const uint32_t w1 = gbase | (uint32_t(r) << 24);
        m->recs.push_back(SubRec{w0, w1});
        m->rec_entry.push_back(c);
        m->rec_slots.push_back(FatRec{int32_t(rec_slot(SubRec{w0, w1}, 0)), int32_t(rec_slot(SubRec{w0, w1}, 1)),
                                      int32_t(rec_slot(SubRec{w0, w1}, 2)), 0u, 0u, 0u,
                                      gbase | (uint32_t(n - 1) << 24) | kRecValid, 0u});
      }
    }
    m->grp_rec_cnt[g] = int(m->recs.size()) - m->grp_rec_off[g];
  }
  // closed-form periodic neighbours, accepted only if they reproduce `trans`
  m->col_cells.assign(std::max(m->K, 1), 0u);
  if (d->periodic_dims && d->col_cells) {
    const int L1 = d->periodic_dims[0], L2 = d->periodic_dims[1], L3 = d->periodic_dims[2];
    bool ok = L1 > 0 && L2 > 0 && L3 > 0 && L1 < 256 && L2 < 256 && L3 < 256 && int64_t(L1) * L2 * L3 == N;
    for (int j = 0; ok && j < m->K; ++j) {
      const int cx = d->col_cells[3 * j], cy = d->col_cells[3 * j + 1], cz = d->col_cells[3 * j + 2];
      ok = cx >= 0 && cx < L1 && cy >= 0 && cy < L2 && cz >= 0 && cz < L3;
      m->col_cells[j] = uint32_t(cx) | (uint32_t(cy) << 8) | (uint32_t(cz) << 16);
    }
    for (int i = 0; ok && i < N; ++i) {
      const int x = i / (L2 * L3), y = (i / L3) % L2, z = i % L3;
      for (int j = 0; ok && j < m->K; ++j) {
        const int xx = (x + d->col_cells[3 * j]) % L1, yy = (y + d->col_cells[3 * j + 1]) % L2,
                  zz = (z + d->col_cells[3 * j + 2]) % L3;
        ok = m->trans[int64_t(i) * m->K + j] == (xx * L2 + yy) * L3 + zz;
      }
    }
    // halo width: largest member offset, offsets recovered from the wrapped
    // ones; the padded box needs L > 2h in every dimension
    int h = 0;
    for (int j = 0; ok && j < m->K; ++j) {
      const int L[3] = {L1, L2, L3};
      for (int dim = 0; dim < 3; ++dim) {
        int o = d->col_cells[3 * j + dim];
        if (2 * o > L[dim]) o -= L[dim];
        h = std::max(h, std::abs(o));
      }
    }
    if (ok && 2 * h < L1 && 2 * h < L2 && 2 * h < L3 &&
        int64_t(L1 + 2 * h) * (L2 + 2 * h) * (L3 + 2 * h) < (1 << 20)) {
      m->periodic = true;
      m->dims[0] = L1;
      m->dims[1] = L2;
      m->dims[2] = L3;
      m->pad_h = h;
      m->pdims[0] = L1 + 2 * h;
      m->pdims[1] = L2 + 2 * h;
      m->pdims[2] = L3 + 2 * h;
      m->signed_cells.resize(3 * size_t(m->K));
      for (int j = 0; j < m->K; ++j) {
        const int L[3] = {L1, L2, L3};
        for (int dim = 0; dim < 3; ++dim) {
          int o = d->col_cells[3 * j + dim];
          if (2 * o > L[dim]) o -= L[dim];
          m->signed_cells[3 * j + dim] = o;
        }
      }
    }
  }
  m->build_gtab(m->eci);
  // sweep records: padded-box offsets and wrapped offsets of the members on
  // closed-form lattices, else translation-table columns
  m->frecs.clear();
  for (const auto& sr : m->rec_slots) {
    const int nm = int((sr.meta >> 24) & 3u);
    const uint32_t slot[3] = {uint32_t(sr.m0), uint32_t(sr.m1), uint32_t(sr.m2)};
    int32_t mm[3] = {0, 0, 0};
    uint32_t ww[3] = {0u, 0u, 0u};
    for (int q = 0; q < nm; ++q) {
      if (m->periodic) {
        const int* o = &m->signed_cells[3 * slot[q]];
        mm[q] = (o[0] * m->pdims[1] + o[1]) * m->pdims[2] + o[2];
        ww[q] = m->col_cells[slot[q]];
      } else {
        mm[q] = int32_t(slot[q]);
      }
    }
    uint32_t meta = sr.meta;
    if (nm < 2) meta |= kRecUnused1;
    if (nm < 3) meta |= kRecUnused2;
    m->frecs.push_back(FatRec{mm[0], mm[1], mm[2], ww[0], ww[1], ww[2], meta, 0u});
  }
  m->max_members = 1;
  for (const auto& fr : m->frecs) m->max_members = std::max(m->max_members, int((fr.meta >> 24) & 3u));
  // interaction mask of the speculative rounds: the reference site and every
  // member offset of its sub-clusters, in the (2h+1)^3 box
  m->spec_ok = m->periodic && m->n_groups == 1 && int(m->movable.size()) == N && m->pad_h <= 2;
  if (m->spec_ok) {
    const int h = m->pad_h, w = 2 * h + 1;
    auto set_bit = [&](int ox, int oy, int oz) {
      const int idx = ((ox + h) * w + (oy + h)) * w + (oz + h);
      m->imask[idx >> 6] |= uint64_t(1) << (idx & 63);
    };
    set_bit(0, 0, 0);
    for (const auto& sr : m->rec_slots) {
      const int nm = int((sr.meta >> 24) & 3u);
      const uint32_t slot[3] = {uint32_t(sr.m0), uint32_t(sr.m1), uint32_t(sr.m2)};
      for (int q = 0; q < nm; ++q) {
        const int* o = &m->signed_cells[3 * slot[q]];
        set_bit(o[0], o[1], o[2]);
      }
    }
  }
  // upload
  std::vector<int8_t> sg(N);
  for (int i = 0; i < N; ++i) sg[i] = int8_t(m->site_group[i]);
  m->d_recs = upload(m->recs);
  m->d_frecs = upload(m->frecs);
  m->d_gtab = upload(m->gtab);
  m->build_gtab_il();
  if (!m->gtab_il.empty()) m->d_gtab_il = upload(m->gtab_il);
  m->build_duals();
  m->build_gtab_fx();
  m->d_gtab_fx = upload(m->gtab_fx);
  if (!m->duals.empty()) m->d_duals = upload(m->duals);
  m->d_trans = upload(m->trans);
  m->d_cells = upload(m->col_cells);
  m->d_site_group = upload(sg);
  m->d_background = upload(m->background);
  m->d_movable = upload(m->movable);
  m->d_cf_kind = upload(m->cf_kind);
  m->d_cf_dec = upload(m->cf_dec);
  m->d_cf_entry = upload(m->cf_entry);
  m->d_cf_eq_off = upload(m->cf_eq_off);
  m->d_eq_decos = upload(m->eq_decos);
  m->d_entries = upload(m->entries);
  m->d_eci = upload(m->eci);
  m->refresh_view();
}

// Correlation functions of occupations already on the device (one launch for
// several chains).  `models` is a device array of model views.
static void compute_cf(const ModelDev* d_models, const std::vector<ChainParams>& params_host,
                       const ChainParams* d_params, const int8_t* d_occ_pool, const std::vector<int32_t>& chains,
                       int n_cf_max, std::vector<double>* cf_out, std::vector<double>* energy_out,
                       hipStream_t stream) {
  int max_sites = 1;
  for (int c : chains) max_sites = std::max(max_sites, params_host[c].n_sites);
  const int n_blocks = (max_sites + 255) / 256;
  const int nc = int(chains.size());
  int32_t* d_ids = upload(chains);
  double *d_partial = nullptr, *d_cf = nullptr, *d_e = nullptr;
  try {
    HIP_CHECK(hipMalloc(&d_partial, sizeof(double) * size_t(nc) * n_cf_max * n_blocks));
    HIP_CHECK(hipMalloc(&d_cf, sizeof(double) * size_t(nc) * n_cf_max));
    HIP_CHECK(hipMalloc(&d_e, sizeof(double) * size_t(nc)));
    HIP_CHECK(launch_cf(d_models, d_params, d_occ_pool, d_ids, nc, max_sites, d_partial, d_cf, d_e, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (cf_out) {
      cf_out->resize(size_t(nc) * n_cf_max);
      HIP_CHECK(hipMemcpy(cf_out->data(), d_cf, sizeof(double) * cf_out->size(), hipMemcpyDeviceToHost));
    }
    if (energy_out) {
      energy_out->resize(nc);
      HIP_CHECK(hipMemcpy(energy_out->data(), d_e, sizeof(double) * nc, hipMemcpyDeviceToHost));
    }
  } catch (...) {
    dev_free(d_ids);
    dev_free(d_partial);
    dev_free(d_cf);
    dev_free(d_e);
    throw;
  }
  dev_free(d_ids);
  dev_free(d_partial);
  dev_free(d_cf);
  dev_free(d_e);
}

// Energies of chains from scratch through the energy tables (the path the
// batched sweep uses; correlation functions are not formed).
static void compute_energy(const ModelDev* d_models, const std::vector<ChainParams>& params_host,
                           const ChainParams* d_params, const int8_t* d_occ_pool, const std::vector<int32_t>& chains,
                           std::vector<double>* energy_out, hipStream_t stream) {
  int max_sites = 1;
  for (int c : chains) max_sites = std::max(max_sites, params_host[c].n_sites);
  const int n_blocks = (max_sites + 255) / 256;
  const int nc = int(chains.size());
  int32_t* d_ids = upload(chains);
  double *d_partial = nullptr, *d_e = nullptr;
  try {
    HIP_CHECK(hipMalloc(&d_partial, sizeof(double) * size_t(nc) * n_blocks));
    HIP_CHECK(hipMalloc(&d_e, sizeof(double) * size_t(nc)));
    HIP_CHECK(launch_energy(d_models, d_params, d_occ_pool, d_ids, nc, max_sites, d_partial, d_e, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    energy_out->resize(nc);
    HIP_CHECK(hipMemcpy(energy_out->data(), d_e, sizeof(double) * nc, hipMemcpyDeviceToHost));
  } catch (...) {
    dev_free(d_ids);
    dev_free(d_partial);
    dev_free(d_e);
    throw;
  }
  dev_free(d_ids);
  dev_free(d_partial);
  dev_free(d_e);
}

// ============================================================================
// Scratch of the allocation-free entry point (cemc_batch_get_cf_noalloc):
// one file-scope device array per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel or a copy (in host code it
// is a stand-in, not a device address): it uses the address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device on the first call.  Every call on a GPU shares that GPU's array, so
// calls that overlap on one GPU race.
constexpr size_t kNoAllocDoubles = size_t(1) << 17;  // 1 MiB
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // device address of `scratch`, per device

double* scratch_on_current_device() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr.size() <= size_t(dev)) g_scratch_ptr.resize(size_t(dev) + 1, nullptr);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) throw DeviceError("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

// Offsets (in doubles, 256-byte aligned) inside the scratch array: the chain
// id (int32) at 0, the correlation functions, the energy, then the per-block
// partial sums of cf_partial_kernel.
struct NoAllocLayout {
  size_t cf, energy, partial, total;
};

NoAllocLayout noalloc_layout(int n_cf, int n_sites) {
  auto up = [](size_t v) { return (v + 31) / 32 * 32; };
  const size_t n_blocks = (size_t(std::max(n_sites, 1)) + 255) / 256;
  NoAllocLayout l;
  l.cf = 32;
  l.energy = l.cf + up(size_t(n_cf));
  l.partial = l.energy + 32;
  l.total = l.partial + size_t(n_cf) * n_blocks;
  return l;
}

}  // namespace

// ============================================================================
// Live updaters with the resident kernel switched on.  A resident kernel
// holds the hardware queue of its stream while it is fed, and a process has
// few of them, so an updater takes the resident path only while it is the
// only such updater; with several (replicas driven in turn) each launches
// per call instead of waiting for another's kernel to idle out.
std::atomic<int> g_resident_updaters{0};

struct cemc_updater {
  cemc_model* model = nullptr;
  hipStream_t stream = nullptr;
  int8_t* d_occ = nullptr;
  double* d_ring = nullptr;
  double* d_slot_energy = nullptr;
  double* d_eci = nullptr;
  int32_t* d_changes = nullptr;
  int32_t* d_restore = nullptr;
  int change_cap = 0;
  // work plan of updater_kernel: items and segments per translation group
  // (host copies of the segments give each change its item range)
  std::vector<UpdSeg> segs;
  std::vector<int32_t> grp_seg;  // n_groups + 1
  UpdItem* d_items = nullptr;
  UpdSeg* d_segs = nullptr;
  UpdPlan plan{};
  uint32_t lds_bytes = 0;
  // pinned host memory: the energy read back by get_energy, the result a
  // launch stores itself (coherent and mapped: `d_result` is its device
  // address), and staging for change lists too long for the kernel arguments
  double* h_energy = nullptr;
  UpdResult* h_result = nullptr;
  UpdResult* d_result = nullptr;
  // the final CFs of the last call that returned an energy (bits, mapped
  // like the result), valid for history slot `cf_slot` (-1: none)
  uint64_t* h_cf = nullptr;
  uint64_t* d_cf = nullptr;
  int cf_slot = -1;
  int32_t* h_stage = nullptr;
  size_t stage_cap = 0;  // int32 entries at h_stage
  // resident updater kernel: its command ring (mapped like the result), the
  // next sequence number, whether a launch may still run, the time of the
  // last command, and the idle time after which the kernel leaves
  UpdMailbox* h_mb = nullptr;
  UpdMailbox* d_mb = nullptr;
  bool resident = true;
  bool res_counted = false;  // counted in g_resident_updaters
  bool res_running = false;
  int64_t res_seq = 0;
  std::chrono::steady_clock::time_point res_last{};
  int64_t idle_ticks = 0;
  // the correlation functions of the last call with a result may still be
  // landing in h_cf (none is pending any more once all have)
  bool cf_landing = false;
  // restores of undone changes not yet sent: they travel in front of the
  // next command of the resident kernel
  std::vector<int32_t> rs_sites, rs_species;
  // energies read back from calls, per history slot, valid while
  // slot_e_epoch[slot] == eci_epoch
  std::vector<double> slot_e;
  std::vector<int64_t> slot_e_epoch;
  std::vector<int8_t> occ;  // host mirror of the device occupation
  std::vector<double> eci;
  std::vector<int64_t> slot_epoch;
  int64_t eci_epoch = 0;
  unsigned current = 0, buffer_size = 0;
  struct Change {
    int site, old_id, new_id;
  };
  std::vector<Change> changes;
  // LinearVibCorrection (cpp/src/linear_vib_correction.cpp): (CF index, v_k)
  // in name order, and whether one was added
  std::vector<std::pair<int, double>> vib;
  bool has_vib = false;
  int num_threads = 1;
  ~cemc_updater() {
    if (res_counted) g_resident_updaters.fetch_sub(1);
    // a resident kernel leaves on the quit command, or at the latest after
    // its idle time
    if (res_running && h_mb && __atomic_load_n(&h_mb->consumed, __ATOMIC_ACQUIRE) >= res_seq - kUpdRing) {
      UpdCmd& c = h_mb->cmd[res_seq % kUpdRing];
      const uint64_t tag = uint64_t(uint32_t(uint64_t(res_seq))) << 32;
      for (int i = 0; i < 4 + kUpdCmdWords; ++i)
        __atomic_store_n(&c.w[i], tag | (i == 0 ? uint64_t(kUpdOpQuit) : 0u), __ATOMIC_RELEASE);
    }
    if (stream) (void)hipStreamSynchronize(stream);
    if (h_mb) (void)hipHostFree(h_mb);
    if (h_energy) (void)hipHostFree(h_energy);
    if (h_result) (void)hipHostFree(h_result);
    if (h_cf) (void)hipHostFree(h_cf);
    if (h_stage) (void)hipHostFree(h_stage);
    dev_free(d_occ);
    dev_free(d_ring);
    dev_free(d_slot_energy);
    dev_free(d_eci);
    dev_free(d_items);
    dev_free(d_segs);
    dev_free(d_changes);
    dev_free(d_restore);
    if (stream) (void)hipStreamDestroy(stream);
  }
  unsigned current_slot() const { return current == 0 ? kMaxHistory - 1 : current - 1; }
};

namespace {

// ---- resident updater kernel (updater_resident_kernel)
// Idle time after which a resident kernel leaves, in milliseconds.
constexpr double kResidentIdleMs = 0.5;

int64_t updater_consumed(const cemc_updater* u) { return __atomic_load_n(&u->h_mb->consumed, __ATOMIC_ACQUIRE); }

// True once the updater's stream has no work left (a resident kernel has
// left); raises the stream's error, if any.
bool updater_stream_idle(cemc_updater* u) {
  const hipError_t q = hipStreamQuery(u->stream);
  if (q == hipErrorNotReady) return false;
  HIP_CHECK(q);
  return true;
}

// Launches the resident kernel from the first command no launch has copied.
void updater_resident_launch(cemc_updater* u) {
  HIP_CHECK(launch_updater_resident(u->plan, u->d_occ, u->d_ring, u->d_slot_energy, u->d_eci, u->d_result, u->d_cf,
                                    u->d_mb, updater_consumed(u) + 1, u->idle_ticks, u->lds_bytes, u->stream));
  u->res_running = true;
}

// Spins until ready(); every 4096 polls asks the stream whether the resident
// kernel has left (idle before a command arrived: launched again while
// commands are outstanding) and gives up after 60 s.
template <class Ready>
void updater_resident_wait(cemc_updater* u, Ready ready, const char* what) {
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t spin = 1; !ready(); ++spin) {
    if ((spin & 4095u) == 0) {
      if (updater_stream_idle(u)) {
        u->res_running = false;
        if (ready()) return;
        if (updater_consumed(u) + 1 >= u->res_seq) throw DeviceError(what);
        updater_resident_launch(u);
      }
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60)) throw DeviceError(what);
    }
    __builtin_ia32_pause();
  }
}

// Queues one command for the resident kernel (launching it when none may
// run), after the kernel has copied the command that last used the slot:
// the restores waiting in rs_sites / rs_species (they must fit beside the
// command's own words), then `words`.
void updater_post(cemc_updater* u, int32_t op, int32_t n, bool want, const int32_t* words, int n_words) {
  const int nr = int(u->rs_sites.size());
  require(2 * nr + n_words <= kUpdCmdWords, "updater command too long");
  const auto now = std::chrono::steady_clock::now();
  if (u->res_running && now - u->res_last > std::chrono::duration<double, std::milli>(kResidentIdleMs / 2) &&
      updater_stream_idle(u))
    u->res_running = false;
  if (!u->res_running) updater_resident_launch(u);
  const int64_t seq = u->res_seq;
  updater_resident_wait(u, [&] { return updater_consumed(u) >= seq - kUpdRing; },
                        "the resident updater kernel stopped taking commands");
  uint32_t cw[4 + kUpdCmdWords] = {};
  cw[0] = uint32_t(op);
  cw[1] = uint32_t(n);
  cw[2] = want ? 1u : 0u;
  cw[3] = uint32_t(nr);
  for (int i = 0; i < nr; ++i) {
    cw[4 + i] = uint32_t(u->rs_sites[i]);
    cw[4 + nr + i] = uint32_t(u->rs_species[i]);
  }
  u->rs_sites.clear();
  u->rs_species.clear();
  for (int i = 0; i < n_words; ++i) cw[4 + 2 * nr + i] = uint32_t(words[i]);
  // every word carries the command's number: the kernel takes the command
  // once all of them do, whatever order the stores arrive in
  UpdCmd& c = u->h_mb->cmd[seq % kUpdRing];
  const uint64_t tag = uint64_t(uint32_t(uint64_t(seq))) << 32;
  for (int i = 0; i < 4 + kUpdCmdWords; ++i) __atomic_store_n(&c.w[i], tag | cw[i], __ATOMIC_RELEASE);
  u->res_seq = seq + 1;
  u->res_last = std::chrono::steady_clock::now();
}

// True when this call may go to the resident kernel (see g_resident_updaters).
bool updater_resident_now(const cemc_updater* u) {
  return u->resident && g_resident_updaters.load(std::memory_order_relaxed) == 1;
}

// Ends the resident kernel after the commands queued so far, so that other
// work can use the updater's stream.
void updater_park(cemc_updater* u) {
  if (!u->res_running && u->rs_sites.empty()) return;
  if (!u->rs_sites.empty()) updater_post(u, kUpdOpRestore, 0, false, nullptr, 0);
  const int64_t quit = u->res_seq;
  updater_post(u, kUpdOpQuit, 0, false, nullptr, 0);
  for (;;) {
    HIP_CHECK(hipStreamSynchronize(u->stream));
    if (updater_consumed(u) >= quit) break;
    updater_resident_launch(u);  // it left idle before the commands arrived
  }
  u->res_running = false;
}

// Copies `n` int32 values into the pinned staging buffer and from there to
// `dst` on the updater's stream (after the stream's earlier work, which may
// still read the staging buffer).
void updater_stage_copy(cemc_updater* u, int32_t* dst, const int32_t* src, size_t n) {
  updater_park(u);
  HIP_CHECK(hipStreamSynchronize(u->stream));
  if (n > u->stage_cap) {
    if (u->h_stage) (void)hipHostFree(u->h_stage);
    u->h_stage = nullptr;
    u->stage_cap = 0;
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&u->h_stage), sizeof(int32_t) * n));
    u->stage_cap = n;
  }
  std::memcpy(u->h_stage, src, sizeof(int32_t) * n);
  HIP_CHECK(hipMemcpyAsync(dst, u->h_stage, sizeof(int32_t) * n, hipMemcpyHostToDevice, u->stream));
}

// Work plan of updater_kernel for the updater's model (see UpdItem): for
// each translation group g and cluster correlation function k with a cluster
// entry at g, one segment of items (equivalent decoration e, record r of the
// entry), in the order e-major; LDS passes of at most 4096 items.
void updater_build_plan(cemc_updater* u) {
  const cemc_model& m = *u->model;
  const int ncf = m.n_cf;
  std::vector<UpdItem> items;
  u->segs.clear();
  u->grp_seg.assign(size_t(m.n_groups) + 1, 0);
  for (int g = 0; g < m.n_groups; ++g) {
    u->grp_seg[g] = int32_t(u->segs.size());
    for (int k = 0; k < ncf; ++k) {
      if (m.cf_kind[k] != CEMC_CF_CLUSTER) continue;
      const int ent = m.cf_entry[size_t(g) * ncf + k];
      if (ent < 0) continue;
      const ClusterEntry& E = m.entries[ent];
      const int e0 = m.cf_eq_off[k], e1 = m.cf_eq_off[k + 1];
      UpdSeg sg{};
      sg.k = k;
      sg.begin = int32_t(items.size());
      for (int e = e0; e < e1; ++e) {
        const uint32_t deco = m.eq_decos[e];
        for (int r = 0; r < E.rec_count; ++r) {
          const SubRec rec = m.recs[size_t(E.rec_begin) + r];
          const int nm = rec_size(rec) - 1;
          UpdItem it{};
          for (int q = 0; q < nm; ++q) it.m[q] = m.periodic ? m.col_cells[rec_slot(rec, q)] : rec_slot(rec, q);
          for (int pos = 0; pos <= nm; ++pos) {
            const uint32_t d = (deco >> (8 * pos)) & 0xFFu;
            require(d < 8u, "decoration index out of range");
            it.aux |= d << (3 * pos);
          }
          it.aux |= uint32_t(rec_refpos(rec)) << 12;
          it.aux |= uint32_t(nm) << 14;
          items.push_back(it);
          require(items.size() < (size_t(1) << 31), "too many updater items");
        }
      }
      sg.end = int32_t(items.size());
      sg.mul = double(unsigned(E.size)) / double(e1 - e0);
      sg.div = double(E.norm * m.grp_count[g]);
      u->segs.push_back(sg);
    }
  }
  u->grp_seg[m.n_groups] = int32_t(u->segs.size());
  std::vector<int32_t> singlets;
  for (int k = 0; k < ncf; ++k)
    if (m.cf_kind[k] == CEMC_CF_SINGLET) singlets.push_back(int32_t(k | (m.cf_dec[k] << 16)));
  require(ncf < (1 << 16), "too many correlation functions for the updater");
  // the occupation is staged in LDS when it fits beside the rest (lattices
  // up to ~100k sites); passes of up to 4096 items
  constexpr uint32_t kLdsCap = 150u * 1024u;
  const int occ_sites = updater_lds_bytes(ncf, 256, m.n_sites) <= kLdsCap ? m.n_sites : 0;
  int chunk = 4096;
  while (chunk > 256 && updater_lds_bytes(ncf, chunk, occ_sites) > kLdsCap) chunk -= 256;
  u->lds_bytes = updater_lds_bytes(ncf, chunk, occ_sites);
  require(u->lds_bytes <= kLdsCap, "too many correlation functions for the updater's LDS");
  u->d_items = upload(items);
  u->d_segs = upload(u->segs);
  require(singlets.size() <= size_t(kUpdMaxSinglets), "too many singlet correlation functions");
  for (size_t j = 0; j < singlets.size(); ++j) u->plan.singlets[j] = singlets[j];
  u->plan.items = u->d_items;
  u->plan.segs = u->d_segs;
  u->plan.trans = m.d_trans;
  u->plan.n_singlets = int32_t(singlets.size());
  u->plan.chunk = chunk;
  u->plan.n_sites = m.n_sites;
  u->plan.n_cf = ncf;
  u->plan.n_species = m.n_species;
  u->plan.n_bfs = m.n_bfs;
  u->plan.K = m.K;
  u->plan.periodic = m.periodic ? 1 : 0;
  for (int i = 0; i < 3; ++i) u->plan.dims[i] = m.dims[i];
  u->plan.occ_lds = occ_sites > 0 ? 1 : 0;
  require(m.basis.size() <= size_t(kMaxBasis * kMaxSpecies), "basis table too large");
  for (size_t i = 0; i < m.basis.size(); ++i) u->plan.phi[i] = m.basis[i];
}

// One launch for a sequence of changes, queued on the updater's stream:
// short sequences travel in the kernel arguments.  With `want_result` the
// kernel stores the final energy in the mapped result, which is marked
// pending first; the host waits for it with updater_wait_result.
void updater_flush(cemc_updater* u, const std::vector<int32_t>& packed, bool want_result) {
  const int n = int(packed.size() / kUpdWords);
  if (n == 0) return;
  const int dst = (packed[packed.size() - kUpdWords + 2] >> 16) & 0xFFFF;
  if (want_result) {
    // each value is stored once per call: once none is pending, no store
    // of the previous call is still on its way
    if (u->cf_landing) {
      const int ncf = u->model->n_cf;
      updater_resident_wait(
          u,
          [&] {
            for (int k = 0; k < ncf; ++k)
              if (__atomic_load_n(&u->h_cf[k], __ATOMIC_ACQUIRE) == kUpdPending) return false;
            return true;
          },
          "the updater kernel stored no correlation functions");
      u->cf_landing = false;
    }
    for (int k = 0; k < u->model->n_cf; ++k) __atomic_store_n(&u->h_cf[k], kUpdPending, __ATOMIC_RELAXED);
    __atomic_store_n(&u->h_result->energy_bits, kUpdPending, __ATOMIC_SEQ_CST);
  }
  u->cf_slot = want_result ? dst : -1;
  if (want_result) u->cf_landing = true;
  u->slot_epoch[dst] = u->eci_epoch;
  if (updater_resident_now(u) && kUpdWords * n <= kUpdCmdWords) {
    // waiting restores ride in front of the changes when both fit
    if (2 * int(u->rs_sites.size()) + int(packed.size()) > kUpdCmdWords)
      updater_post(u, kUpdOpRestore, 0, false, nullptr, 0);
    updater_post(u, kUpdOpChanges, n, want_result, packed.data(), int(packed.size()));
    return;
  }
  updater_park(u);
  UpdChanges inl{};
  const int32_t* d_list = nullptr;
  if (n <= kUpdInline) {
    std::memcpy(inl.v, packed.data(), sizeof(int32_t) * packed.size());
  } else {
    if (n > u->change_cap) {
      HIP_CHECK(hipStreamSynchronize(u->stream));
      dev_free(u->d_changes);
      u->change_cap = std::max(n, 2 * u->change_cap);
      HIP_CHECK(hipMalloc(&u->d_changes, sizeof(int32_t) * kUpdWords * u->change_cap));
    }
    updater_stage_copy(u, u->d_changes, packed.data(), packed.size());
    d_list = u->d_changes;
  }
  HIP_CHECK(launch_updater(u->plan, u->d_occ, u->d_ring, u->d_slot_energy, inl, d_list, n, u->d_eci,
                           want_result ? u->d_result : nullptr, want_result ? u->d_cf : nullptr, u->lds_bytes,
                           u->stream));
}

// Waits for the energy of the last flush with a result: spins on the mapped
// result until it is no longer pending, and every 4096 polls asks the stream
// whether it failed or finished without storing it.
double updater_wait_result(cemc_updater* u) {
  const auto t0 = std::chrono::steady_clock::now();
  auto done = [&](double* e) {
    const uint64_t bits = __atomic_load_n(&u->h_result->energy_bits, __ATOMIC_ACQUIRE);
    if (bits == kUpdPending) return false;
    std::memcpy(e, &bits, sizeof(double));
    if (u->cf_slot >= 0) {
      u->slot_e[u->cf_slot] = *e;
      u->slot_e_epoch[u->cf_slot] = u->eci_epoch;
    }
    return true;
  };
  double e = 0.0;
  if (u->res_running) {
    updater_resident_wait(u, [&] { return done(&e); }, "the resident updater kernel stored no result");
    return e;
  }
  for (uint32_t spin = 1;; ++spin) {
    if (done(&e)) return e;
    if ((spin & 4095u) == 0) {
      const hipError_t q = hipStreamQuery(u->stream);
      if (q != hipSuccess && q != hipErrorNotReady) HIP_CHECK(q);
      if (q == hipSuccess || std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60)) {
        HIP_CHECK(hipStreamSynchronize(u->stream));
        if (done(&e)) return e;
        throw DeviceError("the updater kernel stored no result");
      }
    }
    __builtin_ia32_pause();
  }
}

// One change (ce_updater.cpp:313-347): no-op if the symbols are equal, else
// takes the next history slot and records the change.  Appends the kernel
// work item; the flip itself happens on the device.
int updater_stage(cemc_updater* u, int site, int old_id, int new_id, std::vector<int32_t>& packed) {
  const cemc_model& m = *u->model;
  if (old_id == new_id) return CEMC_OK;
  if (site < 0 || site >= m.n_sites) {
    g_last_error = "site index out of range";
    return CEMC_EINVAL;
  }
  if (new_id < 0 || new_id >= m.n_species || old_id < 0 || old_id >= m.n_species) {
    g_last_error = "species id out of range";
    return CEMC_EINVAL;
  }
  if (m.background[site]) {
    g_last_error = "Attempting to move a background atom!";
    return CEMC_ERUNTIME;
  }
  const unsigned src = u->current_slot();
  const unsigned dst = u->current;
  u->current = (u->current + 1) % kMaxHistory;
  if (u->buffer_size < unsigned(kMaxHistory)) u->buffer_size += 1;
  u->changes[dst] = cemc_updater::Change{site, old_id, new_id};
  u->slot_e_epoch[dst] = -1;
  // the species the site holds, as the reference reads it from its own
  // symbols (ce_updater.cpp:335), whatever the caller named as old
  const int cur = u->occ[site];
  u->occ[site] = int8_t(new_id);
  const int g = m.site_group[site];
  const int sb = u->grp_seg[g], se = u->grp_seg[g + 1];
  packed.push_back(site);
  packed.push_back(new_id | (cur << 8));
  packed.push_back(int32_t(src | (dst << 16)));
  packed.push_back(sb < se ? u->segs[sb].begin : 0);
  packed.push_back(sb < se ? u->segs[se - 1].end : 0);
  packed.push_back(sb);
  packed.push_back(se);
  packed.push_back(0);
  return CEMC_OK;
}

// Energy of the current slot: when it is the slot the last call with an
// energy ended in and the ECIs are unchanged since, that energy from the
// mapped result; else one copy into pinned memory on the stream (after the
// queued updates) and one wait.
double updater_energy(cemc_updater* u) {
  const unsigned slot = u->current_slot();
  if (u->slot_e_epoch[slot] == u->eci_epoch) return u->slot_e[slot];
  updater_park(u);
  if (u->slot_epoch[slot] != u->eci_epoch) {
    HIP_CHECK(launch_slot_energy(u->model->d_view, u->d_ring, u->d_slot_energy, u->d_eci, int(slot), u->stream));
    u->slot_epoch[slot] = u->eci_epoch;
  }
  HIP_CHECK(hipMemcpyAsync(u->h_energy, u->d_slot_energy + slot, sizeof(double), hipMemcpyDeviceToHost, u->stream));
  HIP_CHECK(hipStreamSynchronize(u->stream));
  u->slot_e[slot] = *u->h_energy;
  u->slot_e_epoch[slot] = u->eci_epoch;
  return *u->h_energy;
}

// Correlation functions of the current slot (n_cf values): from the mapped
// copy when the current slot is the one the last call with an energy ended
// in (each value waited for until the launch has stored it), else one copy
// from the device.
void updater_current_cf(cemc_updater* u, double* cf) {
  const int ncf = u->model->n_cf;
  if (u->cf_slot >= 0 && unsigned(u->cf_slot) == u->current_slot()) {
    bool ok = true;
    for (int k = 0; k < ncf && ok; ++k) {
      uint64_t bits = __atomic_load_n(&u->h_cf[k], __ATOMIC_ACQUIRE);
      for (uint32_t spin = 1; bits == kUpdPending; ++spin) {
        if ((spin & 4095u) == 0) {
          // the launch has ended (its energy was read): what is still
          // pending after the stream drains never comes
          updater_park(u);
          HIP_CHECK(hipStreamSynchronize(u->stream));
          bits = __atomic_load_n(&u->h_cf[k], __ATOMIC_ACQUIRE);
          if (bits == kUpdPending) ok = false;
          break;
        }
        __builtin_ia32_pause();
        bits = __atomic_load_n(&u->h_cf[k], __ATOMIC_ACQUIRE);
      }
      if (ok) std::memcpy(cf + k, &bits, sizeof(double));
    }
    if (ok) return;
    u->cf_slot = -1;
  }
  updater_park(u);
  HIP_CHECK(hipMemcpyAsync(cf, u->d_ring + size_t(u->current_slot()) * ncf, sizeof(double) * ncf,
                           hipMemcpyDeviceToHost, u->stream));
  HIP_CHECK(hipStreamSynchronize(u->stream));
}

}  // namespace

// ============================================================================
struct cemc_batch {
  int n_chains = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<cemc_model*> models;  // distinct models, index = ChainParams::model
  std::vector<ChainParams> params;
  std::vector<int> n_cf_of_chain;
  ModelDev* d_models = nullptr;
  ChainParams* d_params = nullptr;
  ChainState* d_states = nullptr;
  int8_t* d_occ = nullptr;
  uint16_t* d_trk = nullptr;
  int64_t* d_nmoves = nullptr;
  int32_t* d_order = nullptr;
  struct Class {
    bool periodic, canonical;
    int kr;
    bool gtab_lds;
    int kbm;
    int waves;
    int spec;  // moves per round
    bool field;  // local-field kernel
    bool hooks;  // constraints and bias
    int nm;      // members gathered per record
    int begin, count;
    uint32_t lds;
  };
  int spec_moves = 0;  // moves per round where the model allows it (0: automatic; 1, 8, or 8 x waves = wide)
  int energy_path = CEMC_PATH_AUTO;
  std::vector<Class> classes;
  int64_t classes_version = 0;  // bumped whenever the launch classes change
  int forced_waves = 0;
  size_t order_cap = 0;         // entries allocated at d_order
  // pool mode: chains join and leave; slots [0, n_chains) with active flags
  bool dynamic = false;
  std::vector<char> active;
  int model_capacity = 0;
  std::map<int64_t, int64_t> occ_free, trk_free;  // free ranges: offset -> length
  int n_cf_max = 1;
  // Wang-Landau walkers: descriptors per chain (n_bins 0: none), their bins
  // packed in chain order in the three pools
  // constraints and bias per chain (host copies; device arrays they point to)
  struct Hooks {
    bool on = false;
    HookDev dev{};
    uint8_t* d_allowed = nullptr;
    int32_t* d_nb = nullptr;
    double* d_btab = nullptr;
  };
  std::vector<Hooks> hooks;
  HookDev* d_hooks = nullptr;
  std::vector<WLDev> wl;
  WLDev* d_wl = nullptr;
  double* d_wl_logdos = nullptr;
  uint32_t* d_wl_hist = nullptr;
  uint8_t* d_wl_known = nullptr;
  int32_t* d_wl_order = nullptr;
  std::vector<void*> wl_arrays;  // per chain: the walker's adaptive-window arrays (one allocation)
  // pools: device scratch of the initial-energy launch at admission, kept
  // across admissions (a free per admission would wait for the whole device)
  int32_t* d_adm_id = nullptr;
  double* d_adm_e = nullptr;
  double* d_adm_partial = nullptr;
  double* h_adm = nullptr;  // pinned: chain id in and energy out
  int adm_blocks = 0;
  ~cemc_batch() {
    if (h_adm) (void)hipHostFree(h_adm);
    dev_free(d_adm_id);
    dev_free(d_adm_e);
    dev_free(d_adm_partial);
    for (auto& h : hooks) {
      dev_free(h.d_allowed);
      dev_free(h.d_nb);
      dev_free(h.d_btab);
    }
    dev_free(d_hooks);
    dev_free(d_wl);
    dev_free(d_wl_logdos);
    dev_free(d_wl_hist);
    dev_free(d_wl_known);
    dev_free(d_wl_order);
    for (void* p : wl_arrays) dev_free(p);
    dev_free(d_models);
    dev_free(d_params);
    dev_free(d_states);
    dev_free(d_occ);
    dev_free(d_trk);
    dev_free(d_nmoves);
    dev_free(d_order);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (auto e : side_done) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    for (auto st : side) (void)hipStreamDestroy(st);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // launches of different classes run concurrently on side streams (default
  // priority), forked from and joined back to `stream` with events
  static constexpr int kSideStreams = 4;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> side_done;
  hipEvent_t fork_ev = nullptr;
};

// Parallel-tempering run over the chains of one batch (replica slot t starts
// with chain t at that chain's temperature).
constexpr int kPTGraphCycles = 16;  // sweep + exchange cycles per captured graph

struct cemc_pt {
  cemc_batch* batch = nullptr;
  int n_rep = 0;
  PTDev* d_pt = nullptr;
  int32_t *d_cos = nullptr, *d_soc = nullptr;
  double *d_kT = nullptr, *d_eref = nullptr, *d_sde = nullptr, *d_sde2 = nullptr;
  uint64_t *d_n = nullptr, *d_acc = nullptr, *d_try = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_cycles = 0;
  int64_t graph_version = -1;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  ~cemc_pt() {
    drop_graph();
    dev_free(d_pt);
    dev_free(d_cos);
    dev_free(d_soc);
    dev_free(d_kT);
    dev_free(d_eref);
    dev_free(d_sde);
    dev_free(d_sde2);
    dev_free(d_n);
    dev_free(d_acc);
    dev_free(d_try);
  }
};

namespace {

// Host-side parameters of chain c on model m (index mid in the batch's
// model list) and, for canonical chains, its species index tracker: site
// lists per species in ascending site order (swap_move_index_tracker.py:22-33).
void setup_chain(cemc_batch* b, int c, cemc_model* m, int mid, const cemc_chain_desc& d, std::vector<uint16_t>& trk) {
  require(d.occupation != nullptr, "NULL occupation");
  require(d.kind == CEMC_MOVE_CANONICAL || d.kind == CEMC_MOVE_SGC, "unknown move kind");
  require(d.temperature > 0.0 && std::isfinite(d.temperature), "temperature must be positive");
  ChainParams& p = b->params[c];
  std::memset(&p, 0, sizeof p);
  p.model = mid;
  p.kind = d.kind;
  p.n_sites = m->n_sites;
  p.kT = d.temperature * kBoltzmannEV;
  p.seed = d.seed;
  p.uid = d.uid;
  b->n_cf_of_chain[c] = m->n_cf;
  b->n_cf_max = std::max(b->n_cf_max, m->n_cf);
  for (int k = 0; k < m->n_cf; ++k)
    if (m->cf_kind[k] == CEMC_CF_SINGLET) p.j1[m->cf_dec[k]] = m->eci[k];
  if (d.chem_pot)
    for (int dd = 0; dd < m->n_bfs; ++dd) p.j1[dd] -= d.chem_pot[dd];
  const int N = m->n_sites;
  for (int i = 0; i < N; ++i)
    require(d.occupation[i] >= 0 && d.occupation[i] < m->n_species, "occupation species id out of range");
  trk.clear();
  if (d.kind == CEMC_MOVE_CANONICAL) {
    require(N <= 65535, "canonical chains support at most 65535 sites");
    int off = 0;
    for (int sp = 0; sp < m->n_species; ++sp) {
      int cnt = 0;
      for (int i = 0; i < N; ++i)
        if (!m->background[i] && d.occupation[i] == sp) {
          trk.push_back(uint16_t(i));
          ++cnt;
        }
      p.seg_off[sp] = off;
      p.count[sp] = cnt;
      off += cnt;
      if (cnt > 0) p.present[p.n_present++] = sp;
    }
    require(p.n_present >= 2, "canonical moves need at least two species present");
    for (int k = 0; k < p.n_present; ++k) {
      const int a = p.present[k];
      p.present_pk |= uint32_t(a) << (4 * k);
      p.count_pk[k >> 2] |= uint64_t(uint32_t(p.count[a]) & 0xFFFFu) << (16 * (k & 3));
      p.segoff_pk[k >> 2] |= uint64_t(uint32_t(p.seg_off[a]) & 0xFFFFu) << (16 * (k & 3));
    }
    trk.resize(N, 0);
  }
}

// First-fit allocation of `len` units from a free-range map; -1 when no range fits.
int64_t range_alloc(std::map<int64_t, int64_t>& free_map, int64_t len) {
  if (len <= 0) return 0;
  for (auto it = free_map.begin(); it != free_map.end(); ++it) {
    if (it->second < len) continue;
    const int64_t off = it->first, rest = it->second - len;
    free_map.erase(it);
    if (rest > 0) free_map[off + len] = rest;
    return off;
  }
  return -1;
}

void range_free(std::map<int64_t, int64_t>& free_map, int64_t off, int64_t len) {
  if (len <= 0) return;
  auto it = free_map.emplace(off, len).first;
  auto next = std::next(it);
  if (next != free_map.end() && it->first + it->second == next->first) {
    it->second += next->second;
    free_map.erase(next);
  }
  if (it != free_map.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      free_map.erase(it);
    }
  }
}

// Launch classes: chains that share a sweep-kernel variant, each class
// ordered longest chain first.  forced_waves > 0 fixes the waves per chain
// (where the variant allows it); 0 chooses them from the batch size.
// Largest share of interacting sites for which the automatic choice takes
// rounds of kSpecMoves moves.
constexpr double kRoundShare = 0.12;

void derive_classes(cemc_batch* b, int forced_waves, const std::vector<int64_t>* nm = nullptr) {
  const int n_chains = b->n_chains;
  // pool batches launch only the active chains that have moves to make
  auto included = [&](int c) { return !b->dynamic || (b->active[c] && nm && (*nm)[c] > 0); };
  // launch classes: chains sharing a kernel variant, longest first
  std::vector<int32_t> order;
  b->classes.clear();
  const uint32_t limit = lds_limit();
  const uint32_t gtab_budget = std::min(kGtabLdsBudget, limit);
  struct Key {
    bool periodic, canonical;
    int kr;
    bool glds;
    int kbm;
    int waves;
    int spec;
    bool field;
    bool hooks;
    int nm;
    uint32_t lds;
  };
  // waves per chain: enough chains x waves to give every SIMD of the
  // device a wave (4 per CU), at most 8, and no more waves than keep
  // about two sub-clusters per lane busy
  int n_cu = 256;
  {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  int n_launch = 0;
  for (int c = 0; c < n_chains; ++c) n_launch += included(c) ? 1 : 0;
  // chains on the device during this launch: the pool's own and, for a
  // pool, those the other pools have in flight
  const int64_t on_device = n_launch + (b->dynamic ? g_pool_chains[current_device_index()].load() : 0);
  int batch_waves = 1;
  while (batch_waves < 8 && on_device * batch_waves * 2 <= int64_t(4) * n_cu) batch_waves *= 2;
  // pools: a chain with far more moves left than the launch's median (a
  // large job among many small ones) gets up to four waves even when the
  // device is crowded, so the longest jobs do not set the tail alone
  int64_t long_moves = -1;
  if (b->dynamic && nm != nullptr && n_launch > 0) {
    std::vector<int64_t> left;
    for (int c = 0; c < n_chains; ++c)
      if (included(c)) left.push_back((*nm)[c]);
    std::nth_element(left.begin(), left.begin() + left.size() / 2, left.end());
    long_moves = std::max<int64_t>(8 * left[left.size() / 2], int64_t(1) << 20);
  }
  std::vector<Key> keys(n_chains);
  std::vector<bool> done(n_chains, false);
  for (int c = 0; c < n_chains; ++c) {
    if (!included(c)) {
      done[c] = true;
      continue;
    }
    const cemc_model* m = b->models[b->params[c].model];
    const bool canon = b->params[c].kind == CEMC_MOVE_CANONICAL;
    Key k;
    k.periodic = m->periodic;
    k.canonical = canon;
    k.kbm = m->read_slices();
    const bool single = m->n_groups == 1 && int(m->movable.size()) == m->n_sites;
    const int cnt = m->grp_rec_cnt[0];
    const int occ_bytes = m->periodic ? m->pdims[0] * m->pdims[1] * m->pdims[2] : m->n_sites;
    const int trk_sites = canon ? m->n_sites : 0;
    const uint32_t with_g = sweep_lds_bytes(occ_bytes, trk_sites, 0, int(m->gtab.size()));
    const uint32_t no_g = sweep_lds_bytes(occ_bytes, trk_sites, 0, 0);
    int auto_waves = 1;
    const int chain_waves = long_moves > 0 && (*nm)[c] >= long_moves ? std::max(batch_waves, 4) : batch_waves;
    while (auto_waves < chain_waves && cnt > 128 * auto_waves) auto_waves *= 2;
    // at most four records per lane even on a crowded device: eight need
    // about twice the registers, one wave per SIMD (measured on 1024 ternary
    // FCC 12^3 chains, 494 records: two waves with four records each 6.3e8
    // swaps/s, one wave with eight 5.7e8)
    while (auto_waves < 8 && cnt > 256 * auto_waves) auto_waves *= 2;
    auto pick_kr = [&](int waves) {
      if (single && k.kbm <= 3 && no_g <= limit)
        for (int kr : {2, 4, 8})
          if (cnt <= kr * 64 * waves) return kr;
      return 0;
    };
    // a forced wave count that leaves the records no room in registers falls
    // back to the automatic choice for this chain
    k.waves = forced_waves > 0 ? forced_waves : auto_waves;
    k.kr = pick_kr(k.waves);
    if (k.kr == 0 && forced_waves > 0) {
      k.waves = auto_waves;
      k.kr = pick_kr(k.waves);
    }
    // automatic: rounds where the moves of a round seldom conflict.  An
    // accepted move invalidates a later one when their sites interact: a
    // share |interaction box| / N of the sites for a flip, four times that
    // for a swap (two sites each).  Above kRoundShare the rounds end early
    // too often and one move at a time is faster (measured on binary FCC 10^3:
    // flips, share 0.055, gain from rounds; swaps, share 0.22, loss).
    int want_spec = b->spec_moves;
    const int kr1 = pick_kr(1);
    if (want_spec == 0) {
      const int box = __builtin_popcountll(m->imask[0]) + __builtin_popcountll(m->imask[1]);
      const double share = double(box) * (canon ? 4.0 : 1.0) / double(m->n_sites);
      want_spec = share <= kRoundShare ? kSpecMoves : 1;
      // wide rounds are not chosen automatically: the k-th move of a round
      // is cut with probability about (acceptance x share x k), and the
      // acceptance is unknown here (measured on binary FCC 24^3 swaps: 2x
      // faster than rounds of eight at acceptance 0.07, 10% slower at 0.4);
      // the autotuner times them
    }
    if (want_spec > kSpecMoves) {
      // wide rounds: every wave holds all the records
      if (kr1 > 0 && m->spec_ok) {
        k.waves = want_spec / kSpecMoves;
        k.kr = kr1;
      } else {
        want_spec = kSpecMoves;
      }
    }
    k.spec = (k.kr > 0 && m->spec_ok) ? want_spec : 1;
    if (k.kr > 0) {
      k.glds = with_g <= gtab_budget;
      // a launch with more chains than can be resident at once with their
      // tables in LDS runs more of them side by side, hiding each one's
      // latency, when the tables stay in HBM (measured on 1024 ternary FCC
      // 12^3 chains, two per unit with tables in LDS: 5.7e8 against 3.4e8
      // swaps/s at one wave per chain, 6.3e8 against 4.1e8 at two)
      if (k.glds) {
        const uint32_t st = sweep_static_lds() + 256u;
        const int per_unit_lds = int(kUnitLds / (with_g + st)), per_unit_hbm = int(kUnitLds / (no_g + st));
        if (on_device > int64_t(n_cu) * per_unit_lds && per_unit_hbm > per_unit_lds) k.glds = false;
      }
      k.lds = k.glds ? with_g : no_g;
    } else {
      k.waves = 1;
      const uint32_t rec_g = sweep_lds_bytes(occ_bytes, trk_sites, int(m->recs.size()), int(m->gtab.size()));
      const uint32_t rec_only = sweep_lds_bytes(occ_bytes, trk_sites, int(m->recs.size()), 0);
      k.glds = rec_g <= gtab_budget;
      k.lds = k.glds ? rec_g : rec_only;
      require(k.lds <= limit, "chain state does not fit in the workgroup's LDS");
    }
    // local fields (one wave per chain, records in VGPRs) when asked for,
    // where the model allows them, the chain's fields fit in LDS and no other
    // wave count is forced.  Not automatic: a chain's cost on this path
    // grows with its acceptance rate, so in a batch the hottest chain sets
    // the launch time (config-style batches of 1024 chains over 300-1500 K
    // ran 3x slower than evaluation rounds); it pays for single chains at
    // low acceptance.
    k.field = false;
    if (b->energy_path == CEMC_PATH_FIELDS && m->field_ok && forced_waves <= 1) {
      const int kr1 = pick_kr(1);
      const uint32_t f_lds = field_lds_bytes(occ_bytes, trk_sites, m->n_sites * m->n_bfs, int(m->gtab_fx.size()));
      if (kr1 > 0 && k.kbm <= 3 && f_lds <= limit) {
        k.field = true;
        k.waves = 1;
        k.kr = kr1;
        k.spec = 1;
        k.glds = true;
        k.lds = f_lds;
      }
    }
    // records in VGPRs gather two members each when no record has three
    k.nm = (k.kr > 0 && !k.field && m->max_members <= 2) ? 2 : 3;
    // chains with constraints or a bias: one wave, one move at a time
    k.hooks = !b->hooks.empty() && b->hooks[c].on;
    if (k.hooks) {
      k.field = false;
      k.waves = 1;
      k.spec = 1;
      k.kr = pick_kr(1);
      if (k.kr > 0) {
        k.glds = with_g <= gtab_budget;
        k.lds = k.glds ? with_g : no_g;
      } else {
        const uint32_t rec_g = sweep_lds_bytes(occ_bytes, trk_sites, int(m->recs.size()), int(m->gtab.size()));
        const uint32_t rec_only = sweep_lds_bytes(occ_bytes, trk_sites, int(m->recs.size()), 0);
        k.glds = rec_g <= gtab_budget;
        k.lds = k.glds ? rec_g : rec_only;
        require(k.lds <= limit, "chain state does not fit in the workgroup's LDS");
      }
      k.nm = (k.kr > 0 && m->max_members <= 2) ? 2 : 3;
    }
    keys[c] = k;
  }
  for (int c0 = 0; c0 < n_chains; ++c0) {
    if (done[c0]) continue;
    const Key& k0 = keys[c0];
    std::vector<int32_t> ids;
    uint32_t lds = 0;
    for (int c = c0; c < n_chains; ++c) {
      const Key& k = keys[c];
      if (done[c] || k.periodic != k0.periodic || k.canonical != k0.canonical || k.kr != k0.kr ||
          k.glds != k0.glds || k.kbm != k0.kbm || k.waves != k0.waves || k.spec != k0.spec ||
          k.field != k0.field || k.hooks != k0.hooks || k.nm != k0.nm)
        continue;
      done[c] = true;
      ids.push_back(c);
      lds = std::max(lds, k.lds);
    }
    std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) {
      const auto* mx = b->models[b->params[x].model];
      const auto* my = b->models[b->params[y].model];
      return int64_t(mx->n_sites) * mx->recs.size() > int64_t(my->n_sites) * my->recs.size();
    });
    b->classes.push_back(cemc_batch::Class{k0.periodic, k0.canonical, k0.kr, k0.glds, k0.kbm, k0.waves, k0.spec,
                                           k0.field, k0.hooks, k0.nm, int(order.size()), int(ids.size()), lds});
    order.insert(order.end(), ids.begin(), ids.end());
  }
  if (order.size() > b->order_cap || b->d_order == nullptr) {
    dev_free(b->d_order);
    b->d_order = nullptr;
    b->order_cap = std::max<size_t>(order.size(), size_t(n_chains));
    HIP_CHECK(hipMalloc(&b->d_order, sizeof(int32_t) * b->order_cap));
  }
  if (!order.empty())
    HIP_CHECK(hipMemcpyAsync(b->d_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice,
                             b->stream));
  b->classes_version += 1;
  uint32_t need = 0;
  for (const auto& cl : b->classes) need = std::max(need, cl.lds);
  if (need + sweep_static_lds() > 64u * 1024u) HIP_CHECK(sweep_set_max_lds(limit + sweep_static_lds()));
}

// Reads the chains' constraint state back after a launch (the kernel keeps
// pair counts, bias counts and energies there); returns the reference's
// error message for a chain whose move found no valid proposal
// (montecarlo.py:986-989), clearing the flag.
std::string hooks_after_launch(cemc_batch* b) {
  if (b->d_hooks == nullptr) return std::string();
  std::vector<HookDev> st(b->hooks.size());
  HIP_CHECK(hipMemcpy(st.data(), b->d_hooks, sizeof(HookDev) * st.size(), hipMemcpyDeviceToHost));
  std::string err;
  for (size_t c = 0; c < st.size(); ++c) {
    if (!b->hooks[c].on) continue;
    b->hooks[c].dev = st[c];
    if (st[c].status != 0 && err.empty())
      err = "Did not manage to produce a trial move that does not violate any of the constraints (chain " +
            std::to_string(c) + ", move " + std::to_string(st[c].failed_move) + ")";
    b->hooks[c].dev.status = 0;
  }
  if (!err.empty()) {
    std::vector<HookDev> up;
    for (const auto& h : b->hooks) up.push_back(h.dev);
    HIP_CHECK(hipMemcpy(b->d_hooks, up.data(), sizeof(HookDev) * up.size(), hipMemcpyHostToDevice));
  }
  return err;
}

// One sweep-kernel launch per class on the batch's stream; each chain
// advances by its entry of d_nmoves.
void enqueue_sweeps(cemc_batch* b, uint32_t* d_log, int64_t log_words, int64_t time_limit = 0) {
  const bool fan_out = b->classes.size() > 1;
  if (fan_out) {
    if (b->side.empty()) {
      for (int i = 0; i < cemc_batch::kSideStreams; ++i) {
        hipStream_t st = nullptr;
        HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->side.push_back(st);
        hipEvent_t ev = nullptr;
        HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->side_done.push_back(ev);
      }
      HIP_CHECK(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
    }
    HIP_CHECK(hipEventRecord(b->fork_ev, b->stream));
    for (auto st : b->side) HIP_CHECK(hipStreamWaitEvent(st, b->fork_ev, 0));
  }
  int ci = 0;
  for (const auto& cl : b->classes) {
    LaunchSweep a;
    a.params = b->d_params;
    a.states = b->d_states;
    a.models = b->d_models;
    a.occ_pool = b->d_occ;
    a.trk_pool = b->d_trk;
    a.chain_order = b->d_order + cl.begin;
    a.n_moves = b->d_nmoves;
    a.n_chains = cl.count;
    a.lds_bytes = cl.lds;
    a.periodic = cl.periodic;
    a.canonical = cl.canonical;
    a.kr = cl.kr;
    a.gtab_lds = cl.gtab_lds;
    a.kbm = cl.kbm;
    a.waves = cl.waves;
    a.spec = cl.spec;
    a.field = cl.field;
    a.hooks = cl.hooks ? b->d_hooks : nullptr;
    a.nm = cl.nm;
    a.accept_log = d_log;
    a.log_words = log_words;
    a.time_limit = time_limit;
    hipStream_t st = fan_out ? b->side[ci % cemc_batch::kSideStreams] : b->stream;
    HIP_CHECK(launch_sweep(a, st));
    ++ci;
  }
  if (fan_out) {
    const int used = std::min<int>(ci, cemc_batch::kSideStreams);
    for (int i = 0; i < used; ++i) {
      HIP_CHECK(hipEventRecord(b->side_done[i], b->side[i]));
      HIP_CHECK(hipStreamWaitEvent(b->stream, b->side_done[i], 0));
    }
  }
}

// Ticks of the device's constant-rate wall clock (wall_clock64) per millisecond.
double wall_ticks_per_ms() {
  int dev = 0, khz = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) throw DeviceError("the device reports no wall-clock rate");
  return double(khz);
}

void batch_launch(cemc_batch* b, const int64_t* n_moves, int64_t moves, bool timed, double* ms,
                  uint32_t* host_log = nullptr, int64_t log_words = 0, int64_t time_limit = 0) {
  std::vector<int64_t> nm(b->n_chains);
  for (int c = 0; c < b->n_chains; ++c) {
    nm[c] = n_moves ? n_moves[c] : moves;
    require(nm[c] >= 0, "number of moves must be non-negative");
    if (host_log) require(nm[c] <= 32 * log_words, "accept log too small for the number of moves");
  }
  uint32_t* d_log = nullptr;
  if (host_log) {
    HIP_CHECK(hipMalloc(&d_log, sizeof(uint32_t) * size_t(b->n_chains) * log_words));
    HIP_CHECK(hipMemsetAsync(d_log, 0, sizeof(uint32_t) * size_t(b->n_chains) * log_words, b->stream));
  }
  if (b->dynamic) derive_classes(b, b->forced_waves, &nm);
  HIP_CHECK(hipMemcpyAsync(b->d_nmoves, nm.data(), sizeof(int64_t) * nm.size(), hipMemcpyHostToDevice, b->stream));
  // a pool's chains count as on the device from the launch until the wait
  // below returns (or throws)
  struct InFlight {
    std::atomic<int64_t>* counter = nullptr;
    int64_t n = 0;
    ~InFlight() {
      if (counter) counter->fetch_sub(n);
    }
  } in_flight;
  if (b->dynamic) {
    for (int c = 0; c < b->n_chains; ++c) in_flight.n += (b->active[c] && nm[c] > 0) ? 1 : 0;
    in_flight.counter = &g_pool_chains[current_device_index()];
    in_flight.counter->fetch_add(in_flight.n);
  }
  if (timed) HIP_CHECK(hipEventRecord(b->ev0, b->stream));
  enqueue_sweeps(b, d_log, log_words, time_limit);
  if (timed) HIP_CHECK(hipEventRecord(b->ev1, b->stream));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  std::string hook_error = hooks_after_launch(b);
  if (!hook_error.empty() && d_log) (void)hipFree(d_log);
  if (!hook_error.empty()) throw std::runtime_error(hook_error);
  if (d_log) {
    const hipError_t e = hipMemcpy(host_log, d_log, sizeof(uint32_t) * size_t(b->n_chains) * log_words,
                                   hipMemcpyDeviceToHost);
    (void)hipFree(d_log);
    HIP_CHECK(e);
  }
  if (timed) {
    float t = 0.f;
    HIP_CHECK(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = double(t);
  }
}

// ---------------------------------------------------------------------------
// Wang-Landau walkers.

// Bins of every walker, read back from the device (empty before the pools
// exist).
struct WLBins {
  std::vector<double> logdos;
  std::vector<uint32_t> hist;
  std::vector<uint8_t> known;
};

WLBins wl_download(cemc_batch* b) {
  WLBins h;
  int64_t total = 0;
  for (const auto& w : b->wl) total += w.n_bins;
  if (b->d_wl == nullptr || total == 0) return h;
  h.logdos.resize(size_t(total));
  h.hist.resize(size_t(total));
  h.known.resize(size_t(total));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  HIP_CHECK(hipMemcpy(h.logdos.data(), b->d_wl_logdos, sizeof(double) * size_t(total), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(h.hist.data(), b->d_wl_hist, sizeof(uint32_t) * size_t(total), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(h.known.data(), b->d_wl_known, size_t(total), hipMemcpyDeviceToHost));
  std::vector<WLDev> st(b->wl.size());
  HIP_CHECK(hipMemcpy(st.data(), b->d_wl, sizeof(WLDev) * st.size(), hipMemcpyDeviceToHost));
  b->wl = st;
  return h;
}

// Re-packs the pools after walker `chain` got n_bins bins (its new values
// follow), keeping every other walker's bins and state.
void wl_store(cemc_batch* b, int chain, const WLDev& desc, const double* logdos, const uint32_t* hist,
              const uint8_t* known) {
  if (b->wl.empty()) b->wl.assign(size_t(b->n_chains), WLDev{});
  const WLBins old = wl_download(b);
  std::vector<WLDev> next = b->wl;
  next[chain] = desc;
  WLBins nb;
  int64_t off = 0;
  for (int c = 0; c < b->n_chains; ++c) {
    WLDev& w = next[c];
    const int64_t n = w.n_bins;
    if (c == chain) {
      for (int64_t i = 0; i < n; ++i) {
        nb.logdos.push_back(logdos ? logdos[i] : 0.0);
        nb.hist.push_back(hist ? hist[i] : 0u);
        nb.known.push_back(known ? known[i] : uint8_t(0));
      }
    } else {
      const int64_t o = b->wl[c].bin_off;
      for (int64_t i = 0; i < n; ++i) {
        nb.logdos.push_back(old.logdos[size_t(o + i)]);
        nb.hist.push_back(old.hist[size_t(o + i)]);
        nb.known.push_back(old.known[size_t(o + i)]);
      }
    }
    w.bin_off = off;
    off += n;
  }
  dev_free(b->d_wl_logdos);
  dev_free(b->d_wl_hist);
  dev_free(b->d_wl_known);
  b->d_wl_logdos = nullptr;
  b->d_wl_hist = nullptr;
  b->d_wl_known = nullptr;
  const size_t n = size_t(std::max<int64_t>(off, 1));
  HIP_CHECK(hipMalloc(&b->d_wl_logdos, sizeof(double) * n));
  HIP_CHECK(hipMalloc(&b->d_wl_hist, sizeof(uint32_t) * n));
  HIP_CHECK(hipMalloc(&b->d_wl_known, n));
  if (off > 0) {
    HIP_CHECK(hipMemcpy(b->d_wl_logdos, nb.logdos.data(), sizeof(double) * size_t(off), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(b->d_wl_hist, nb.hist.data(), sizeof(uint32_t) * size_t(off), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(b->d_wl_known, nb.known.data(), size_t(off), hipMemcpyHostToDevice));
  }
  if (b->d_wl == nullptr) HIP_CHECK(hipMalloc(&b->d_wl, sizeof(WLDev) * size_t(b->n_chains)));
  HIP_CHECK(hipMemcpy(b->d_wl, next.data(), sizeof(WLDev) * next.size(), hipMemcpyHostToDevice));
  if (b->d_wl_order == nullptr) HIP_CHECK(hipMalloc(&b->d_wl_order, sizeof(int32_t) * size_t(b->n_chains)));
  b->wl = next;
}

// One launch per kernel variant over the chains that carry a walker.
// shared: every walker of the batch visits walker 0's histogram; each check
// of check_every steps (rounded up to a multiple of the walker count, as
// wang_landau_sampler.cpp:321-324 does) is one launch in which every walker
// makes its share, followed by the group's flatness test.
void wl_launch(cemc_batch* b, int64_t n_checks, int64_t check_every, uint32_t* host_log, int64_t log_words,
               double* ms, bool shared = false) {
  require(!b->dynamic, "Wang-Landau walkers run on fixed batches, not pools");
  require(n_checks >= 0, "number of convergence checks must be non-negative");
  require(check_every > 0, "check_every must be positive");
  int64_t per_walker = check_every;
  if (shared) {
    require(!b->wl.empty(), "no chain of the batch carries a Wang-Landau walker");
    const WLDev& w0 = b->wl[0];
    for (const WLDev& w : b->wl) {
      require(w.n_bins > 0, "a shared histogram needs a walker on every chain");
      require(w.n_bins == w0.n_bins && w.e_min == w0.e_min && w.e_max == w0.e_max,
              "walkers sharing a histogram need the same bins and range");
      require(!w.adaptive, "adaptive windows run with one walker per histogram");
    }
    const int64_t W = b->n_chains;
    per_walker = (check_every + W - 1) / W;
  }
  if (host_log) require(n_checks * per_walker <= 32 * log_words, "accept log too small for the number of moves");
  const uint32_t limit = lds_limit();
  struct WKey {
    bool periodic;
    int kr;
    bool glds;
    int kbm;
    int nm;
    bool adaptive;
    bool sgc;
    bool operator<(const WKey& o) const {
      return std::tie(periodic, kr, glds, kbm, nm, adaptive, sgc) <
             std::tie(o.periodic, o.kr, o.glds, o.kbm, o.nm, o.adaptive, o.sgc);
    }
  };
  std::map<WKey, std::pair<std::vector<int32_t>, uint32_t>> groups;
  for (int c = 0; c < b->n_chains; ++c) {
    if (b->wl.empty() || b->wl[c].n_bins <= 0) continue;
    const cemc_model* m = b->models[b->params[c].model];
    WKey k;
    k.periodic = m->periodic;
    k.sgc = b->params[c].kind == CEMC_MOVE_SGC;
    k.kbm = m->read_slices();
    const bool single = m->n_groups == 1 && int(m->movable.size()) == m->n_sites;
    const int cnt = m->grp_rec_cnt[0];
    k.kr = 0;
    if (single && k.kbm <= 3)
      for (int kr : {2, 4, 8})
        if (cnt <= kr * 64) {
          k.kr = kr;
          break;
        }
    k.nm = (k.kr > 0 && m->max_members <= 2) ? 2 : 3;
    const int occ_bytes = m->periodic ? m->pdims[0] * m->pdims[1] * m->pdims[2] : m->n_sites;
    const int n_rec = k.kr == 0 ? int(m->recs.size()) : 0;
    const int nbins = shared ? 0 : b->wl[c].n_bins;
    const bool adaptive = !shared && b->wl[c].adaptive != 0;
    k.adaptive = adaptive;
    const uint32_t with_g =
        wl_lds_bytes(sweep_lds_bytes(occ_bytes, m->n_sites, n_rec, int(m->gtab.size())), nbins, adaptive);
    const uint32_t no_g = wl_lds_bytes(sweep_lds_bytes(occ_bytes, m->n_sites, n_rec, 0), nbins, adaptive);
    k.glds = with_g <= std::min(kGtabLdsBudget, limit);
    const uint32_t lds = k.glds ? with_g : no_g;
    require(lds <= limit, "the walker's chain and histogram do not fit in the workgroup's LDS");
    auto& g = groups[k];
    g.first.push_back(c);
    g.second = std::max(g.second, lds);
  }
  require(!groups.empty(), "no chain of the batch carries a Wang-Landau walker");
  std::vector<int32_t> order;
  uint32_t need = 0;
  for (auto& kv : groups) {
    order.insert(order.end(), kv.second.first.begin(), kv.second.first.end());
    need = std::max(need, kv.second.second);
  }
  if (need + sweep_static_lds() > 64u * 1024u) HIP_CHECK(sweep_set_max_lds(limit + sweep_static_lds()));
  HIP_CHECK(hipMemcpyAsync(b->d_wl_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice,
                           b->stream));
  uint32_t* d_log = nullptr;
  if (host_log) {
    HIP_CHECK(hipMalloc(&d_log, sizeof(uint32_t) * size_t(b->n_chains) * log_words));
    HIP_CHECK(hipMemsetAsync(d_log, 0, sizeof(uint32_t) * size_t(b->n_chains) * log_words, b->stream));
  }
  if (ms) HIP_CHECK(hipEventRecord(b->ev0, b->stream));
  for (int64_t chk = 0; chk < (shared ? n_checks : 1); ++chk) {
  int begin = 0;
  for (auto& kv : groups) {
    LaunchWL a;
    a.params = b->d_params;
    a.states = b->d_states;
    a.models = b->d_models;
    a.occ_pool = b->d_occ;
    a.trk_pool = b->d_trk;
    a.chain_order = b->d_wl_order + begin;
    a.walkers = b->d_wl;
    a.logdos = b->d_wl_logdos;
    a.hist = b->d_wl_hist;
    a.known = b->d_wl_known;
    a.n_chains = int(kv.second.first.size());
    a.lds_bytes = kv.second.second;
    a.periodic = kv.first.periodic;
    a.kr = kv.first.kr;
    a.gtab_lds = kv.first.glds;
    a.kbm = kv.first.kbm;
    a.nm = kv.first.nm;
    a.n_checks = shared ? 1 : n_checks;
    a.check_every = per_walker;
    a.accept_log = d_log;
    a.log_words = log_words;
    a.shared = shared ? b->d_wl : nullptr;
    a.adaptive = kv.first.adaptive;
    a.sgc = kv.first.sgc;
    a.log_base = shared ? chk * per_walker : 0;
    HIP_CHECK(launch_wl(a, b->stream));
    begin += a.n_chains;
  }
  if (shared)
    HIP_CHECK(launch_wl_shared_check(b->d_wl, b->d_wl_hist, b->d_wl_known, per_walker * b->n_chains, b->stream));
  }
  if (ms) HIP_CHECK(hipEventRecord(b->ev1, b->stream));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  if (d_log) {
    const hipError_t e = hipMemcpy(host_log, d_log, sizeof(uint32_t) * size_t(b->n_chains) * log_words,
                                   hipMemcpyDeviceToHost);
    (void)hipFree(d_log);
    HIP_CHECK(e);
  }
  if (ms) {
    float t = 0.f;
    HIP_CHECK(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = double(t);
  }
  HIP_CHECK(hipMemcpy(b->wl.data(), b->d_wl, sizeof(WLDev) * b->wl.size(), hipMemcpyDeviceToHost));
  for (const WLDev& w : b->wl)
    if (w.n_bins > 0 && w.status != 0)
      throw std::runtime_error("Could not find any energy state inside the histogram range!");
}

// A walker's adaptive-window arrays in one device allocation (carved in
// 256-byte aligned pieces); saved flags and sub-bin counts start at zero.
void wl_alloc_arrays(cemc_batch* b, int chain, WLDev& w) {
  if (b->wl_arrays.empty()) b->wl_arrays.assign(size_t(b->n_chains), nullptr);
  dev_free(b->wl_arrays[chain]);
  b->wl_arrays[chain] = nullptr;
  const int64_t nb = w.n_bins, N = b->params[chain].n_sites, rows = nb + 1;
  const auto up = [](int64_t v) { return (v + 255) & ~int64_t(255); };
  const int64_t o_occ = 0, o_trk = o_occ + up(rows * N), o_e = o_trk + up(2 * rows * N), o_bin = o_e + up(8 * rows),
                o_edges = o_bin + up(4 * rows), o_en = o_edges + up(4 * nb), o_vals = o_en + up(4 * nb),
                o_saved = o_vals + up(8 * nb * nb), o_sub = o_saved + up(nb), total = o_sub + up(40 * nb);
  uint8_t* p = nullptr;
  HIP_CHECK(hipMalloc(&p, size_t(total)));
  b->wl_arrays[chain] = p;
  HIP_CHECK(hipMemset(p + o_saved, 0, size_t(up(nb) + up(40 * nb))));
  w.snap_occ = p + o_occ;
  w.snap_trk = reinterpret_cast<uint16_t*>(p + o_trk);
  w.snap_e = reinterpret_cast<double*>(p + o_e);
  w.snap_bin = reinterpret_cast<int32_t*>(p + o_bin);
  w.edges = reinterpret_cast<int32_t*>(p + o_edges);
  w.edge_n = reinterpret_cast<int32_t*>(p + o_en);
  w.edge_vals = reinterpret_cast<double*>(p + o_vals);
  w.saved = p + o_saved;
  w.sub = reinterpret_cast<uint32_t*>(p + o_sub);
}

}  // namespace

// ============================================================================
extern "C" {

const char* cemc_last_error(void) { return g_last_error.c_str(); }

int cemc_device_count(int* count) {
  return guarded([&] {
    require(count != nullptr, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

int cemc_model_create(const cemc_model_desc* desc, cemc_model** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    auto* m = new cemc_model();
    try {
      build_model(m, desc);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int cemc_model_set_ecis(cemc_model* m, const double* eci) {
  return guarded([&] {
    require(m && eci, "NULL argument");
    m->eci.assign(eci, eci + m->n_cf);
    m->build_gtab(m->eci);
    HIP_CHECK(hipMemcpy(m->d_gtab, m->gtab.data(), sizeof(double) * std::max<size_t>(m->gtab.size(), 1),
                        hipMemcpyHostToDevice));
    m->build_gtab_il();
    if (!m->gtab_il.empty())
      HIP_CHECK(hipMemcpy(m->d_gtab_il, m->gtab_il.data(), sizeof(double) * m->gtab_il.size(),
                          hipMemcpyHostToDevice));
    m->build_gtab_fx();
    HIP_CHECK(hipMemcpy(m->d_gtab_fx, m->gtab_fx.data(), sizeof(int64_t) * m->gtab_fx.size(),
                        hipMemcpyHostToDevice));
    m->refresh_view();
    HIP_CHECK(hipMemcpy(m->d_eci, m->eci.data(), sizeof(double) * m->eci.size(), hipMemcpyHostToDevice));
  });
}

int cemc_model_is_periodic(const cemc_model* m, int* periodic) {
  return guarded([&] {
    require(m && periodic, "NULL argument");
    *periodic = m->periodic ? 1 : 0;
  });
}

int cemc_model_destroy(cemc_model* m) {
  return guarded([&] { delete m; });
}

// ---------------------------------------------------------------- updater
int cemc_updater_create(cemc_model* model, const int8_t* occupation, const double* cf_init, const double* eci,
                        cemc_updater** out) {
  return guarded([&] {
    require(model && occupation && eci && out, "NULL argument");
    *out = nullptr;
    const int N = model->n_sites, ncf = model->n_cf;
    for (int i = 0; i < N; ++i)
      require(occupation[i] >= 0 && occupation[i] < model->n_species, "occupation species id out of range");
    auto* u = new cemc_updater();
    try {
      u->model = model;
      HIP_CHECK(hipStreamCreate(&u->stream));
      u->occ.assign(occupation, occupation + N);
      u->eci.assign(eci, eci + ncf);
      u->changes.assign(kMaxHistory, cemc_updater::Change{0, 0, 0});
      u->slot_epoch.assign(kMaxHistory, -1);
      u->slot_e.assign(kMaxHistory, 0.0);
      u->slot_e_epoch.assign(kMaxHistory, -1);
      u->d_occ = upload(u->occ);
      u->d_eci = upload(u->eci);
      HIP_CHECK(hipMalloc(&u->d_ring, sizeof(double) * size_t(kMaxHistory) * ncf));
      HIP_CHECK(hipMalloc(&u->d_slot_energy, sizeof(double) * kMaxHistory));
      HIP_CHECK(hipMalloc(&u->d_restore, sizeof(int32_t) * 2 * kMaxHistory));
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&u->h_energy), sizeof(double)));
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&u->h_result), sizeof(UpdResult),
                              hipHostMallocMapped | hipHostMallocCoherent));
      std::memset(u->h_result, 0, sizeof(UpdResult));
      HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&u->d_result), u->h_result, 0));
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&u->h_cf), sizeof(uint64_t) * size_t(ncf),
                              hipHostMallocMapped | hipHostMallocCoherent));
      HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&u->d_cf), u->h_cf, 0));
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&u->h_mb), sizeof(UpdMailbox),
                              hipHostMallocMapped | hipHostMallocCoherent));
      std::memset(u->h_mb, 0, sizeof(UpdMailbox));
      u->h_mb->consumed = -1;
      // no word carries the number of a command yet (the first is 0)
      for (int r = 0; r < kUpdRing; ++r)
        for (int i = 0; i < 4 + kUpdCmdWords; ++i) u->h_mb->cmd[r].w[i] = uint64_t(0xFFFFFFFFu) << 32;
      HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&u->d_mb), u->h_mb, 0));
      u->idle_ticks = int64_t(kResidentIdleMs * wall_ticks_per_ms());
      updater_build_plan(u);
      std::vector<double> cf0(ncf);
      if (cf_init) {
        cf0.assign(cf_init, cf_init + ncf);
      } else {
        ChainParams p;
        std::memset(&p, 0, sizeof p);
        p.model = 0;
        p.n_sites = N;
        for (int k = 0; k < ncf; ++k)
          if (model->cf_kind[k] == CEMC_CF_SINGLET) p.j1[model->cf_dec[k]] = u->eci[k];
        ChainParams* d_p = upload(std::vector<ChainParams>{p});
        try {
          std::vector<double> cf;
          compute_cf(model->d_view, {p}, d_p, u->d_occ, {0}, ncf, &cf, nullptr, u->stream);
          cf0 = cf;
        } catch (...) {
          dev_free(d_p);
          throw;
        }
        dev_free(d_p);
      }
      // CFHistoryTracker::insert + clear (cf_history_tracker.cpp:82-108, 146-149)
      HIP_CHECK(hipMemcpy(u->d_ring, cf0.data(), sizeof(double) * ncf, hipMemcpyHostToDevice));
      u->current = 1;
      u->buffer_size = 1;
    } catch (...) {
      delete u;
      throw;
    }
    u->res_counted = true;
    g_resident_updaters.fetch_add(1);
    *out = u;
  });
}

int cemc_updater_destroy(cemc_updater* u) {
  return guarded([&] { delete u; });
}

int cemc_updater_set_occupation(cemc_updater* u, const int8_t* occupation) {
  int rc = CEMC_OK;
  const int g = guarded([&] {
    require(u && occupation, "NULL argument");
    cemc_model* model = u->model;
    const int N = model->n_sites, ncf = model->n_cf;
    for (int i = 0; i < N; ++i)
      require(occupation[i] >= 0 && occupation[i] < model->n_species, "occupation species id out of range");
    for (int i = 0; i < N; ++i)
      if (model->background[i] && occupation[i] != u->occ[i]) {
        g_last_error = "Attempting to move a background atom!";
        rc = CEMC_ERUNTIME;
        return;
      }
    updater_park(u);
    HIP_CHECK(hipStreamSynchronize(u->stream));
    u->occ.assign(occupation, occupation + N);
    HIP_CHECK(hipMemcpy(u->d_occ, u->occ.data(), size_t(N), hipMemcpyHostToDevice));
    ChainParams p;
    std::memset(&p, 0, sizeof p);
    p.model = 0;
    p.n_sites = N;
    for (int k = 0; k < ncf; ++k)
      if (model->cf_kind[k] == CEMC_CF_SINGLET) p.j1[model->cf_dec[k]] = u->eci[k];
    ChainParams* d_p = upload(std::vector<ChainParams>{p});
    std::vector<double> cf;
    try {
      compute_cf(model->d_view, {p}, d_p, u->d_occ, {0}, ncf, &cf, nullptr, u->stream);
    } catch (...) {
      dev_free(d_p);
      throw;
    }
    dev_free(d_p);
    // history holds only the new state (cf_history_tracker.cpp:146-149)
    HIP_CHECK(hipMemcpy(u->d_ring, cf.data(), sizeof(double) * ncf, hipMemcpyHostToDevice));
    u->current = 1;
    u->buffer_size = 1;
    u->slot_epoch.assign(kMaxHistory, -1);
    u->slot_e_epoch.assign(kMaxHistory, -1);
    u->cf_slot = -1;
  });
  return g != CEMC_OK ? g : rc;
}

int cemc_updater_update_cf(cemc_updater* u, int32_t site, int32_t old_id, int32_t new_id) {
  int rc = CEMC_OK;
  const int g = guarded([&] {
    require(u != nullptr, "NULL updater");
    std::vector<int32_t> packed;
    rc = updater_stage(u, site, old_id, new_id, packed);
    updater_flush(u, packed, false);
  });
  return g != CEMC_OK ? g : rc;
}

int cemc_updater_calculate(cemc_updater* u, int32_t n, const int32_t* site, const int32_t* old_id,
                           const int32_t* new_id, double* energy) {
  int rc = CEMC_OK;
  const int g = guarded([&] {
    require(u != nullptr && n >= 0, "bad arguments");
    require(n == 0 || (site && old_id && new_id), "NULL change arrays");
    std::vector<int32_t> packed;
    bool swaps = n > 1 && (n % 2) == 0;
    for (int p = 0; swaps && p < n / 2; ++p)  // is_swap_move (ce_updater.cpp:996-1000)
      swaps = old_id[2 * p] == new_id[2 * p + 1] && new_id[2 * p + 1] == old_id[2 * p + 1];
    if (swaps) {
      if (n / 2 >= kMaxHistory / 2) {
        g_last_error = "The length of sequence of swap move exceeds the buffer size for the history tracker";
        rc = CEMC_EINVAL;
        return;
      }
      for (int p = 0; p < n / 2 && rc == CEMC_OK; ++p) {
        const int a = site[2 * p], b = site[2 * p + 1];
        require(a >= 0 && a < u->model->n_sites && b >= 0 && b < u->model->n_sites, "site index out of range");
        if (u->occ[a] == u->occ[b]) {
          g_last_error =
              "This version of the calculate function assumes that the provided update is swapping two atoms";
          rc = CEMC_ERUNTIME;
          break;
        }
        if (u->occ[a] != old_id[2 * p] || u->occ[b] != old_id[2 * p + 1]) {
          g_last_error = "The atom position tracker does not match the current state";
          rc = CEMC_ERUNTIME;
          break;
        }
        rc = updater_stage(u, a, old_id[2 * p], new_id[2 * p], packed);
        if (rc == CEMC_OK) rc = updater_stage(u, b, old_id[2 * p + 1], new_id[2 * p + 1], packed);
      }
    } else {
      for (int i = 0; i < n && rc == CEMC_OK; ++i) rc = updater_stage(u, site[i], old_id[i], new_id[i], packed);
    }
    if (rc == CEMC_OK && energy && !packed.empty()) {
      // the launch stores the energy in mapped host memory: no copy, no
      // stream wait
      updater_flush(u, packed, true);
      *energy = updater_wait_result(u);
    } else {
      updater_flush(u, packed, false);
      if (rc == CEMC_OK && energy) *energy = updater_energy(u);
    }
  });
  return g != CEMC_OK ? g : rc;
}

int cemc_updater_undo_changes(cemc_updater* u) {
  return guarded([&] {
    require(u != nullptr, "NULL updater");
    const int steps = int(u->buffer_size) - 1;
    std::vector<int32_t> sites, species;
    for (int i = 0; i < steps && u->buffer_size > 0; ++i) {  // pop (cf_history_tracker.cpp:58-80)
      u->current = u->current == 0 ? kMaxHistory - 1 : u->current - 1;
      const auto& ch = u->changes[u->current];
      u->buffer_size -= 1;
      u->occ[ch.site] = int8_t(ch.old_id);
      sites.push_back(ch.site);
      species.push_back(ch.old_id);
    }
    if (!sites.empty()) {
      // restores are queued behind the updates they undo: as a command of
      // the resident kernel, else on the stream
      std::vector<int32_t> buf(sites);
      buf.insert(buf.end(), species.begin(), species.end());
      const int n = int(sites.size());
      if (updater_resident_now(u) && n <= kUpdRestoreMax) {
        // sent with the next command (sites in pop order, applied in order)
        if (int(u->rs_sites.size()) + n > kUpdRestoreMax) updater_post(u, kUpdOpRestore, 0, false, nullptr, 0);
        u->rs_sites.insert(u->rs_sites.end(), sites.begin(), sites.end());
        u->rs_species.insert(u->rs_species.end(), species.begin(), species.end());
        return;
      }
      updater_park(u);
      InlineChanges inl{};
      const int32_t* d_buf = nullptr;
      if (n <= 2 * kInlineChanges) {
        std::memcpy(inl.v, buf.data(), sizeof(int32_t) * buf.size());
      } else {
        updater_stage_copy(u, u->d_restore, buf.data(), buf.size());
        d_buf = u->d_restore;
      }
      HIP_CHECK(launch_restore(u->d_occ, inl, d_buf, n, u->stream));
    }
  });
}

int cemc_updater_clear_history(cemc_updater* u) {
  return guarded([&] {
    require(u != nullptr, "NULL updater");
    u->buffer_size = 1;
  });
}

int cemc_updater_get_energy(cemc_updater* u, double* energy) {
  return guarded([&] {
    require(u && energy, "NULL argument");
    *energy = updater_energy(u);
  });
}

int cemc_updater_get_cf(cemc_updater* u, double* cf) {
  return guarded([&] {
    require(u && cf, "NULL argument");
    updater_current_cf(u, cf);
  });
}

int cemc_updater_get_singlets(cemc_updater* u, double* singlets, int32_t capacity, int32_t* n_singlets) {
  return guarded([&] {
    require(u != nullptr, "NULL updater");
    const cemc_model& m = *u->model;
    std::vector<int> idx;
    for (int k = 0; k < m.n_cf; ++k)
      if (m.cf_kind[k] == CEMC_CF_SINGLET) idx.push_back(k);
    if (n_singlets) *n_singlets = int32_t(idx.size());
    if (singlets == nullptr) return;
    if (capacity < int32_t(idx.size()))
      throw std::runtime_error("The passed Numpy array is too small to hold all the singlets terms!\nMinimum size: " +
                               std::to_string(idx.size()) + ". Given size: " + std::to_string(capacity));
    std::vector<double> cf(m.n_cf);
    updater_current_cf(u, cf.data());
    for (size_t i = 0; i < idx.size(); ++i) singlets[i] = cf[idx[i]];
  });
}

int cemc_updater_add_linear_vib_correction(cemc_updater* u, int32_t n, const int32_t* cf_index,
                                           const double* eci_per_kbT) {
  return guarded([&] {
    require(u != nullptr && n >= 0, "bad arguments");
    require(n == 0 || (cf_index && eci_per_kbT), "NULL correction arrays");
    std::map<int, double> by_index;  // name order = index order (CFs are in sorted-name order)
    for (int i = 0; i < n; ++i) {
      require(cf_index[i] >= 0 && cf_index[i] < u->model->n_cf, "correlation-function index out of range");
      by_index[cf_index[i]] = eci_per_kbT[i];
    }
    u->vib.assign(by_index.begin(), by_index.end());
    u->has_vib = true;
  });
}

int cemc_updater_vib_energy(cemc_updater* u, double temperature, double* energy) {
  return guarded([&] {
    require(u && energy, "NULL argument");
    double e = 0.0;
    if (u->has_vib) {
      std::vector<double> cf(u->model->n_cf);
      updater_current_cf(u, cf.data());
      // LinearVibCorrection::energy (linear_vib_correction.cpp:8-16), kB as there
      constexpr double kB = 8.6173303E-5;
      for (const auto& kv : u->vib) e += cf[kv.first] * kv.second * kB * temperature;
    }
    *energy = e;
  });
}

int cemc_updater_set_resident(cemc_updater* u, int32_t on) {
  return guarded([&] {
    require(u != nullptr, "NULL updater");
    if (!on) updater_park(u);
    u->resident = on != 0;
    if (u->resident != u->res_counted) {
      g_resident_updaters.fetch_add(u->resident ? 1 : -1);
      u->res_counted = u->resident;
    }
  });
}

int cemc_updater_set_num_threads(cemc_updater* u, int32_t num_threads) {
  return guarded([&] {
    require(u != nullptr, "NULL updater");
    require(num_threads > 0, "num_threads must be positive");
    u->num_threads = num_threads;
  });
}

int cemc_updater_set_ecis(cemc_updater* u, const double* eci) {
  return guarded([&] {
    require(u && eci, "NULL argument");
    u->eci.assign(eci, eci + u->model->n_cf);
    updater_park(u);  // a resident kernel holds the ECIs in LDS
    HIP_CHECK(hipStreamSynchronize(u->stream));  // queued updates still read the old ECIs
    HIP_CHECK(hipMemcpy(u->d_eci, u->eci.data(), sizeof(double) * u->eci.size(), hipMemcpyHostToDevice));
    u->eci_epoch += 1;
  });
}

int cemc_updater_get_occupation(cemc_updater* u, int8_t* occupation) {
  return guarded([&] {
    require(u && occupation, "NULL argument");
    // the host mirror follows every change, restore and replacement of the
    // device occupation in call order
    std::memcpy(occupation, u->occ.data(), u->occ.size());
  });
}

int cemc_updater_history_size(cemc_updater* u, int32_t* size) {
  return guarded([&] {
    require(u && size, "NULL argument");
    *size = int32_t(u->buffer_size);
  });
}

// ------------------------------------------------------------------ batch
int cemc_batch_create(int32_t n_chains, cemc_model* const* chain_models, const cemc_chain_desc* chains,
                      cemc_batch** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(n_chains > 0 && chain_models && chains, "bad arguments");
    auto* b = new cemc_batch();
    try {
      b->n_chains = n_chains;
      // non-blocking: a pool served by one host thread must not wait for
      // null-stream copies of another pool's thread, nor they for its kernels
      HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
      HIP_CHECK(hipEventCreate(&b->ev0));
      HIP_CHECK(hipEventCreate(&b->ev1));
      std::vector<int8_t> occ_pool;
      std::vector<uint16_t> trk_pool;
      b->params.resize(n_chains);
      b->n_cf_of_chain.resize(n_chains);
      for (int c = 0; c < n_chains; ++c) {
        cemc_model* m = chain_models[c];
        require(m != nullptr, "NULL model");
        auto it = std::find(b->models.begin(), b->models.end(), m);
        const int mid = int(it - b->models.begin());
        if (it == b->models.end()) b->models.push_back(m);
        std::vector<uint16_t> trk;
        setup_chain(b, c, m, mid, chains[c], trk);
        ChainParams& p = b->params[c];
        const int N = m->n_sites;
        p.occ_off = int64_t(occ_pool.size());
        occ_pool.insert(occ_pool.end(), chains[c].occupation, chains[c].occupation + N);
        occ_pool.resize((occ_pool.size() + 15) & ~size_t(15), 0);
        p.trk_off = int64_t(trk_pool.size());
        trk_pool.insert(trk_pool.end(), trk.begin(), trk.end());
      }
      std::vector<ModelDev> views;
      for (auto* m : b->models) views.push_back(m->view);
      b->d_models = upload(views);
      b->d_params = upload(b->params);
      b->d_occ = upload(occ_pool);
      b->d_trk = upload(trk_pool);
      HIP_CHECK(hipMalloc(&b->d_states, sizeof(ChainState) * n_chains));
      HIP_CHECK(hipMalloc(&b->d_nmoves, sizeof(int64_t) * n_chains));
      derive_classes(b, 0);
      // initial energies from scratch through the energy tables
      std::vector<int32_t> all(n_chains);
      std::iota(all.begin(), all.end(), 0);
      std::vector<double> energies;
      compute_energy(b->d_models, b->params, b->d_params, b->d_occ, all, &energies, b->stream);
      std::vector<ChainState> st(n_chains);
      std::memset(st.data(), 0, sizeof(ChainState) * st.size());
      for (int c = 0; c < n_chains; ++c) {
        st[c].energy = st[c].e_ref = energies[c];
        st[c].moves_done = chains[c].first_move;
      }
      HIP_CHECK(hipMemcpy(b->d_states, st.data(), sizeof(ChainState) * st.size(), hipMemcpyHostToDevice));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int cemc_batch_destroy(cemc_batch* b) {
  return guarded([&] { delete b; });
}

int cemc_batch_run(cemc_batch* b, const int64_t* n_moves, int64_t moves) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    batch_launch(b, n_moves, moves, false, nullptr);
  });
}

int cemc_batch_run_timed(cemc_batch* b, const int64_t* n_moves, int64_t moves, double* kernel_ms) {
  return guarded([&] {
    require(b && kernel_ms, "NULL argument");
    batch_launch(b, n_moves, moves, true, kernel_ms);
  });
}

int cemc_batch_run_sliced(cemc_batch* b, const int64_t* n_moves, int64_t moves, double max_ms, double* kernel_ms) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(max_ms > 0.0, "max_ms must be positive");
    const double ticks = max_ms * wall_ticks_per_ms();
    require(ticks < 9.0e18, "max_ms too large");
    double ms = 0.0;
    batch_launch(b, n_moves, moves, true, &ms, nullptr, 0, std::max<int64_t>(1, int64_t(ticks)));
    if (kernel_ms) *kernel_ms = ms;
  });
}

int cemc_batch_run_logged(cemc_batch* b, const int64_t* n_moves, int64_t moves, uint32_t* accept_bits,
                          int64_t words_per_chain) {
  return guarded([&] {
    require(b && accept_bits, "NULL argument");
    require(words_per_chain > 0, "words_per_chain must be positive");
    batch_launch(b, n_moves, moves, false, nullptr, accept_bits, words_per_chain);
  });
}

int cemc_batch_get_energy(cemc_batch* b, double* energy) {
  return guarded([&] {
    require(b && energy, "NULL argument");
    std::vector<ChainState> st(b->n_chains);
    HIP_CHECK(hipMemcpy(st.data(), b->d_states, sizeof(ChainState) * st.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < b->n_chains; ++c) energy[c] = st[c].energy;
  });
}

int cemc_batch_get_occupation(cemc_batch* b, int32_t chain, int8_t* occupation) {
  return guarded([&] {
    require(b && occupation, "NULL argument");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(!b->dynamic || b->active[chain], "slot is not in use");
    const ChainParams& p = b->params[chain];
    HIP_CHECK(hipMemcpy(occupation, b->d_occ + p.occ_off, size_t(p.n_sites), hipMemcpyDeviceToHost));
  });
}

int cemc_batch_get_stats(cemc_batch* b, int64_t* moves_done, int64_t* n_accept, double* sum_e, double* sum_e2) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    std::vector<ChainState> st(b->n_chains);
    HIP_CHECK(hipMemcpy(st.data(), b->d_states, sizeof(ChainState) * st.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < b->n_chains; ++c) {
      const double n = double(st[c].n_samples), r = st[c].e_ref;
      if (moves_done) moves_done[c] = int64_t(st[c].moves_done);
      if (n_accept) n_accept[c] = int64_t(st[c].n_accept);
      if (sum_e) sum_e[c] = n * r + st[c].sum_de;
      if (sum_e2) sum_e2[c] = st[c].sum_de2 + 2.0 * r * st[c].sum_de + n * r * r;
    }
  });
}

int cemc_batch_reset_stats(cemc_batch* b, int32_t chain) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(chain >= -1 && chain < b->n_chains, "chain index out of range");
    HIP_CHECK(hipStreamSynchronize(b->stream));
    std::vector<ChainState> st(b->n_chains);
    HIP_CHECK(hipMemcpy(st.data(), b->d_states, sizeof(ChainState) * st.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < b->n_chains; ++c) {
      if (chain >= 0 && c != chain) continue;
      ChainState& x = st[c];
      x.e_ref = x.energy;
      x.n_samples = 0;
      x.sum_de = x.sum_de2 = 0.0;
      for (int d = 0; d < kMaxBasis; ++d) x.sum_s[d] = x.sum_s2[d] = x.sum_se[d] = 0.0;
    }
    HIP_CHECK(hipMemcpy(b->d_states, st.data(), sizeof(ChainState) * st.size(), hipMemcpyHostToDevice));
  });
}

int cemc_batch_get_thermo(cemc_batch* b, int32_t chain, cemc_thermo* out) {
  return guarded([&] {
    require(b && out, "NULL argument");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(!b->dynamic || b->active[chain], "slot is not in use");
    ChainState x;
    HIP_CHECK(hipMemcpy(&x, b->d_states + chain, sizeof(ChainState), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(cemc_thermo));
    const ChainParams& p = b->params[chain];
    const cemc_model* m = b->models[p.model];
    const int B = m->n_bfs, S = m->n_species;
    out->n_samples = int64_t(x.n_samples);
    out->n_singlets = B;
    out->energy = x.energy;
    if (x.n_samples == 0) return;
    const double n = double(x.n_samples);
    const double mde = x.sum_de / n;
    out->mean_energy = x.e_ref + mde;
    out->var_energy = x.sum_de2 / n - mde * mde;
    if (p.kind == CEMC_MOVE_SGC) {
      for (int d = 0; d < B; ++d) {
        const double ms = x.sum_s[d] / n;
        out->singlet[d] = ms;
        out->var_singlet[d] = x.sum_s2[d] / n - ms * ms;
        out->cov_singlet_energy[d] = x.sum_se[d] / n - ms * mde;
      }
    } else {
      // swaps keep the composition: the singlets are constant
      for (int d = 0; d < B; ++d) {
        double v = 0.0;
        for (int k = 0; k < p.n_present; ++k) v += double(p.count[p.present[k]]) * m->basis[d * S + p.present[k]];
        out->singlet[d] = v / double(p.n_sites);
      }
    }
  });
}

int cemc_batch_get_cf(cemc_batch* b, int32_t chain, double* cf, double* energy) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(!b->dynamic || b->active[chain], "slot is not in use");
    std::vector<double> cfv, ev;
    compute_cf(b->d_models, b->params, b->d_params, b->d_occ, {chain}, b->n_cf_max, &cfv, &ev, b->stream);
    if (cf) std::copy(cfv.begin(), cfv.begin() + b->n_cf_of_chain[chain], cf);
    if (energy) *energy = ev[0];
  });
}

int cemc_batch_get_cf_noalloc(cemc_batch* b, int32_t chain, double* cf, double* energy) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(!b->dynamic || b->active[chain], "slot is not in use");
    const int n_cf = b->n_cf_of_chain[chain];
    const int n_sites = b->params[chain].n_sites;
    const NoAllocLayout lay = noalloc_layout(n_cf, n_sites);
    if (lay.total > kNoAllocDoubles)
      throw std::invalid_argument("cemc_batch_get_cf_noalloc: " + std::to_string(n_cf) + " correlation functions on " +
                                  std::to_string(n_sites) + " sites need " + std::to_string(lay.total * 8) +
                                  " bytes of scratch, more than its " + std::to_string(kNoAllocDoubles * 8) +
                                  " bytes; use cemc_batch_get_cf");
    double* base = scratch_on_current_device();
    int32_t* d_id = reinterpret_cast<int32_t*>(base);
    double* d_cf = base + lay.cf;
    double* d_e = base + lay.energy;
    double* d_partial = base + lay.partial;
    // every partial, CF and energy entry is written by the kernels: nothing
    // needs clearing between calls
    const int32_t id = chain;
    HIP_CHECK(hipMemcpyAsync(d_id, &id, sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
    HIP_CHECK(launch_cf(b->d_models, b->d_params, b->d_occ, d_id, 1, n_sites, d_partial, d_cf, d_e, b->stream));
    double e = 0.0;
    if (cf) HIP_CHECK(hipMemcpyAsync(cf, d_cf, sizeof(double) * size_t(n_cf), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipMemcpyAsync(&e, d_e, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    if (energy) *energy = e;
  });
}

int cemc_batch_set_waves(cemc_batch* b, int32_t waves) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8, "waves must be 0, 1, 2, 4 or 8");
    HIP_CHECK(hipStreamSynchronize(b->stream));
    if (!b->dynamic) derive_classes(b, waves);
    b->forced_waves = waves;
  });
}

int cemc_batch_set_moves_per_round(cemc_batch* b, int32_t moves) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(moves == 0 || moves == 1 || moves == kSpecMoves || moves == 2 * kSpecMoves || moves == 4 * kSpecMoves ||
                moves == 8 * kSpecMoves,
            "moves per round must be 0 (automatic), 1, 8, 16, 32 or 64");
    HIP_CHECK(hipStreamSynchronize(b->stream));
    b->spec_moves = moves;
    if (!b->dynamic) derive_classes(b, b->forced_waves);
  });
}

int cemc_batch_set_energy_path(cemc_batch* b, int32_t path) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(path == CEMC_PATH_AUTO || path == CEMC_PATH_EVALUATE || path == CEMC_PATH_FIELDS,
            "energy path must be CEMC_PATH_AUTO, CEMC_PATH_EVALUATE or CEMC_PATH_FIELDS");
    HIP_CHECK(hipStreamSynchronize(b->stream));
    b->energy_path = path;
    if (!b->dynamic) derive_classes(b, b->forced_waves);
  });
}

int cemc_batch_get_energy_path(cemc_batch* b, int32_t chain, int32_t* path) {
  return guarded([&] {
    require(b != nullptr && path != nullptr, "NULL argument");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    std::vector<int32_t> order(b->n_chains);
    HIP_CHECK(hipMemcpy(order.data(), b->d_order, sizeof(int32_t) * order.size(), hipMemcpyDeviceToHost));
    for (const auto& cl : b->classes)
      for (int i = cl.begin; i < cl.begin + cl.count; ++i)
        if (order[i] == chain) {
          *path = cl.field ? CEMC_PATH_FIELDS : CEMC_PATH_EVALUATE;
          return;
        }
    throw std::runtime_error("chain not found in any launch class");
  });
}

int cemc_batch_get_launch_shape(cemc_batch* b, int32_t chain, int32_t* waves, int32_t* records_per_lane,
                                int32_t* tables_in_lds) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    std::vector<int32_t> order(b->n_chains);
    HIP_CHECK(hipMemcpy(order.data(), b->d_order, sizeof(int32_t) * order.size(), hipMemcpyDeviceToHost));
    for (const auto& cl : b->classes)
      for (int i = cl.begin; i < cl.begin + cl.count; ++i)
        if (order[i] == chain) {
          if (waves) *waves = cl.waves;
          if (records_per_lane) *records_per_lane = cl.kr;
          if (tables_in_lds) *tables_in_lds = cl.gtab_lds ? 1 : 0;
          return;
        }
    throw std::runtime_error("chain not found in any launch class");
  });
}

int cemc_batch_set_temperature(cemc_batch* b, int32_t chain, double temperature) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(temperature > 0.0 && std::isfinite(temperature), "temperature must be positive");
    b->params[chain].kT = temperature * kBoltzmannEV;
    HIP_CHECK(hipMemcpy(&b->d_params[chain].kT, &b->params[chain].kT, sizeof(double), hipMemcpyHostToDevice));
  });
}

// ------------------------------------------------------ constraints, bias
int cemc_batch_set_hooks(cemc_batch* b, int32_t chain, const cemc_hooks_desc* d) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(!b->dynamic, "constraints and bias apply to fixed batches, not pools");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(d == nullptr || b->wl.empty() || b->wl[chain].n_bins <= 0,
            "constraints and bias do not apply to Wang-Landau walkers");
    const cemc_model* m = b->models[b->params[chain].model];
    const int N = m->n_sites, S = m->n_species;
    if (b->hooks.empty()) b->hooks.resize(size_t(b->n_chains));
    cemc_batch::Hooks& h = b->hooks[chain];
    // the chain's energy carries the bias of its state: take the old one out
    ChainState st{};
    HIP_CHECK(hipMemcpy(&st, b->d_states + chain, sizeof(ChainState), hipMemcpyDeviceToHost));
    if (h.on) st.energy -= h.dev.bias_energy;
    dev_free(h.d_allowed);
    dev_free(h.d_nb);
    dev_free(h.d_btab);
    h = cemc_batch::Hooks{};
    if (d != nullptr) {
      require(d->max_attempts > 0, "Max. constraint attempts has to be > 0!");
      require(d->fixed_species < (1u << S), "fixed species outside the model's species");
      std::vector<int8_t> occ(static_cast<size_t>(N));
      HIP_CHECK(hipMemcpy(occ.data(), b->d_occ + b->params[chain].occ_off, size_t(N), hipMemcpyDeviceToHost));
      HookDev& v = h.dev;
      v.fixed_species = d->fixed_species;
      v.max_attempts = d->max_attempts;
      v.pair_e1 = v.pair_e2 = -1;
      v.bias_species = -1;
      v.recycle = d->recycle_waste ? 1 : 0;
      if (d->allowed) {
        for (int i = 0; i < N; ++i) require((d->allowed[i] >> S) == 0, "allowed-species mask names unknown species");
        h.d_allowed = upload(std::vector<uint8_t>(d->allowed, d->allowed + N));
        v.allowed = h.d_allowed;
      }
      if (d->pair_species[0] >= 0 || d->pair_species[1] >= 0) {
        const int e1 = d->pair_species[0], e2 = d->pair_species[1], nm = d->pair_members;
        require(e1 >= 0 && e1 < S && e2 >= 0 && e2 < S, "pair constraint species outside the model's species");
        require(nm > 0 && nm <= 32 && d->pair_neighbours != nullptr, "pair constraint needs 1-32 members per site");
        std::vector<int32_t> nb(d->pair_neighbours, d->pair_neighbours + int64_t(N) * nm);
        int64_t count = 0;
        for (int i = 0; i < N; ++i)
          for (int q = 0; q < nm; ++q) {
            const int j = nb[size_t(i) * nm + q];
            require(j >= 0 && j < N && j != i, "pair constraint member outside the lattice");
            bool back = false;
            for (int r = 0; r < nm && !back; ++r) back = nb[size_t(j) * nm + r] == i;
            require(back, "pair constraint member lists must be symmetric (j a member of i => i of j)");
            count += (occ[i] == e1 && occ[j] == e2) ? 1 : 0;
          }
        h.d_nb = upload(nb);
        v.pair_e1 = e1;
        v.pair_e2 = e2;
        v.pair_nm = nm;
        v.pair_nb = h.d_nb;
        v.pair_count = count;
      }
      if (d->bias_species >= 0) {
        require(d->bias_species < S, "bias species outside the model's species");
        require(d->bias_table != nullptr, "bias needs a table of n_sites + 1 entries");
        std::vector<double> tab(d->bias_table, d->bias_table + N + 1);
        for (double x : tab) require(std::isfinite(x), "bias table entries must be finite");
        int cnt = 0;
        for (int i = 0; i < N; ++i) cnt += occ[i] == d->bias_species ? 1 : 0;
        h.d_btab = upload(tab);
        v.bias_species = d->bias_species;
        v.bias_count = cnt;
        v.bias_tab = h.d_btab;
        v.bias_energy = tab[size_t(cnt)];
        st.energy += v.bias_energy;
      }
      h.on = true;
    }
    HIP_CHECK(hipMemcpy(&b->d_states[chain].energy, &st.energy, sizeof(double), hipMemcpyHostToDevice));
    if (b->d_hooks == nullptr) HIP_CHECK(hipMalloc(&b->d_hooks, sizeof(HookDev) * size_t(b->n_chains)));
    std::vector<HookDev> up;
    for (const auto& x : b->hooks) up.push_back(x.dev);
    HIP_CHECK(hipMemcpy(b->d_hooks, up.data(), sizeof(HookDev) * up.size(), hipMemcpyHostToDevice));
    derive_classes(b, b->forced_waves);
  });
}

int cemc_batch_get_hooks_state(cemc_batch* b, int32_t chain, int64_t* pair_count, int32_t* bias_count,
                               double* bias_energy) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(!b->hooks.empty() && b->hooks[chain].on, "the chain has no constraints or bias");
    HookDev v{};
    HIP_CHECK(hipMemcpy(&v, b->d_hooks + chain, sizeof(HookDev), hipMemcpyDeviceToHost));
    if (pair_count) *pair_count = v.pair_count;
    if (bias_count) *bias_count = v.bias_count;
    if (bias_energy) *bias_energy = v.bias_energy;
  });
}

// ------------------------------------------------------------ Wang-Landau
int cemc_batch_wl_set(cemc_batch* b, int32_t chain, const cemc_wl_desc* desc, const double* logdos,
                      const int32_t* hist, const uint8_t* known) {
  return guarded([&] {
    require(b && desc, "NULL argument");
    require(!b->dynamic, "Wang-Landau walkers run on fixed batches, not pools");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(b->params[chain].kind == CEMC_MOVE_CANONICAL || b->params[chain].kind == CEMC_MOVE_SGC,
            "unknown move kind");
    require(b->hooks.empty() || !b->hooks[chain].on, "constraints and bias do not apply to Wang-Landau walkers");
    require(desc->n_bins > 0, "the histogram needs at least one bin");
    require(std::isfinite(desc->e_min) && std::isfinite(desc->e_max) && desc->e_max > desc->e_min,
            "the histogram range needs e_min < e_max");
    require(desc->current_bin >= 0 && desc->current_bin < desc->n_bins, "current bin outside the histogram");
    require(std::isfinite(desc->f) && desc->f >= 0.0, "modification factor must be finite and non-negative");
    WLDev w{};
    w.e_min = desc->e_min;
    w.e_max = desc->e_max;
    w.f = desc->f;
    w.f_min = desc->f_min;
    w.flatness = desc->flatness;
    w.n_bins = desc->n_bins;
    w.cur_bin = desc->current_bin;
    w.is_first = desc->is_first ? 1 : 0;
    w.converged = desc->converged ? 1 : 0;
    w.upper = w.n_bins;
    w.n_overlap = 7;
    w.prev_conv = -1;
    w.cur_energy = 0.0;  // WangLandauSampler's current_energy starts at 0 (wang_landau_sampler.cpp:37)
    if (!b->wl_arrays.empty()) {
      dev_free(b->wl_arrays[chain]);
      b->wl_arrays[chain] = nullptr;
    }
    wl_store(b, chain, w, logdos, reinterpret_cast<const uint32_t*>(hist), known);
  });
}

int cemc_batch_wl_get(cemc_batch* b, int32_t chain, cemc_wl_desc* desc, double* logdos, int32_t* hist,
                      uint8_t* known, int64_t* counters) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(!b->wl.empty() && b->wl[chain].n_bins > 0, "the chain carries no Wang-Landau walker");
    const WLBins h = wl_download(b);
    const WLDev& w = b->wl[chain];
    if (desc) {
      desc->n_bins = w.n_bins;
      desc->current_bin = w.cur_bin;
      desc->is_first = w.is_first;
      desc->converged = w.converged;
      desc->e_min = w.e_min;
      desc->e_max = w.e_max;
      desc->f = w.f;
      desc->f_min = w.f_min;
      desc->flatness = w.flatness;
    }
    for (int i = 0; i < w.n_bins; ++i) {
      const size_t o = size_t(w.bin_off + i);
      if (logdos) logdos[i] = h.logdos[o];
      if (hist) hist[i] = int32_t(h.hist[o]);
      if (known) known[i] = h.known[o];
    }
    if (counters) {
      counters[0] = int64_t(w.n_moves);
      counters[1] = int64_t(w.n_accept);
      counters[2] = int64_t(w.n_outside);
      counters[3] = int64_t(w.n_self);
      counters[4] = int64_t(w.n_flat);
    }
  });
}

int cemc_batch_wl_set_scheme(cemc_batch* b, int32_t chain, const cemc_wl_scheme* scheme) {
  return guarded([&] {
    require(b && scheme, "NULL argument");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(!b->wl.empty() && b->wl[chain].n_bins > 0, "the chain carries no Wang-Landau walker");
    require(scheme->min_width >= 0, "minimum window width must be non-negative");
    (void)wl_download(b);
    WLDev& w = b->wl[chain];
    w.inverse_time = scheme->inverse_time ? 1 : 0;
    w.adaptive = scheme->adaptive ? 1 : 0;
    w.min_width = scheme->min_width;
    w.inv_active = 0;
    w.inv_factor = 0.0;
    w.upper = w.n_bins;
    w.n_overlap = 7;
    w.n_edges = 0;
    w.has_first = 0;
    w.bin_change = 0.0;
    w.steps = 0;
    w.since_flat = 0;
    w.prev_conv = -1;
    w.status = 0;
    w.cur_energy = 0.0;
    if (w.adaptive) {
      wl_alloc_arrays(b, chain, w);
    } else if (!b->wl_arrays.empty() && b->wl_arrays[chain]) {
      dev_free(b->wl_arrays[chain]);
      b->wl_arrays[chain] = nullptr;
      w.snap_occ = nullptr;
      w.snap_trk = nullptr;
      w.snap_e = nullptr;
      w.snap_bin = nullptr;
      w.edges = nullptr;
      w.edge_n = nullptr;
      w.edge_vals = nullptr;
      w.saved = nullptr;
      w.sub = nullptr;
    }
    HIP_CHECK(hipMemcpy(b->d_wl, b->wl.data(), sizeof(WLDev) * b->wl.size(), hipMemcpyHostToDevice));
  });
}

int cemc_batch_wl_get_scheme(cemc_batch* b, int32_t chain, cemc_wl_scheme* scheme, int32_t* edges,
                             int32_t* edge_n, double* edge_vals, uint32_t* sub_bins, uint8_t* saved) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    require(chain >= 0 && chain < b->n_chains, "chain index out of range");
    require(!b->wl.empty() && b->wl[chain].n_bins > 0, "the chain carries no Wang-Landau walker");
    (void)wl_download(b);
    const WLDev& w = b->wl[chain];
    if (scheme) {
      scheme->inverse_time = w.inverse_time;
      scheme->adaptive = w.adaptive;
      scheme->min_width = w.min_width;
      scheme->inv_active = w.inv_active;
      scheme->upper = w.adaptive ? w.upper : w.n_bins;
      scheme->n_overlap = w.n_overlap;
      scheme->n_edges = w.n_edges;
      scheme->has_first = w.has_first;
      scheme->inv_factor = w.inv_factor;
      scheme->bin_change = w.bin_change;
      scheme->steps = w.steps;
      scheme->since_flat = w.since_flat;
      scheme->prev_conv = w.prev_conv;
      scheme->n_walk = int64_t(w.n_walk);
    }
    const size_t nb = size_t(w.n_bins);
    const bool arr = w.adaptive && w.saved != nullptr;
    auto get = [&](void* dst, const void* src, size_t bytes) {
      if (!dst) return;
      if (arr)
        HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
      else
        std::memset(dst, 0, bytes);
    };
    get(edges, w.edges, 4 * nb);
    get(edge_n, w.edge_n, 4 * nb);
    get(edge_vals, w.edge_vals, 8 * nb * nb);
    get(sub_bins, w.sub, 40 * nb);
    get(saved, w.saved, nb);
  });
}

int cemc_batch_wl_run(cemc_batch* b, int64_t n_checks, int64_t check_every, uint32_t* accept_bits,
                      int64_t words_per_chain, double* kernel_ms) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    if (accept_bits) require(words_per_chain > 0, "words_per_chain must be positive");
    wl_launch(b, n_checks, check_every, accept_bits, words_per_chain, kernel_ms);
  });
}

int cemc_batch_wl_run_shared(cemc_batch* b, int64_t n_checks, int64_t check_every, uint32_t* accept_bits,
                             int64_t words_per_chain, double* kernel_ms) {
  return guarded([&] {
    require(b != nullptr, "NULL batch");
    if (accept_bits) require(words_per_chain > 0, "words_per_chain must be positive");
    wl_launch(b, n_checks, check_every, accept_bits, words_per_chain, kernel_ms, true);
  });
}

// ----------------------------------------------------------------- pool
int cemc_pool_create(int32_t max_chains, int32_t max_models, int64_t occ_bytes, int64_t tracker_entries,
                     cemc_batch** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    *out = nullptr;
    require(max_chains > 0 && max_models > 0 && occ_bytes > 0 && tracker_entries >= 0, "bad pool sizes");
    auto* b = new cemc_batch();
    try {
      b->dynamic = true;
      b->n_chains = max_chains;
      b->model_capacity = max_models;
      b->active.assign(max_chains, 0);
      b->params.assign(max_chains, ChainParams{});
      for (auto& p : b->params) std::memset(&p, 0, sizeof p);
      b->n_cf_of_chain.assign(max_chains, 0);
      // non-blocking: a pool served by one host thread must not wait for
      // null-stream copies of another pool's thread, nor they for its kernels
      HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
      HIP_CHECK(hipEventCreate(&b->ev0));
      HIP_CHECK(hipEventCreate(&b->ev1));
      HIP_CHECK(hipMalloc(&b->d_models, sizeof(ModelDev) * size_t(max_models)));
      HIP_CHECK(hipMalloc(&b->d_params, sizeof(ChainParams) * size_t(max_chains)));
      HIP_CHECK(hipMalloc(&b->d_states, sizeof(ChainState) * size_t(max_chains)));
      HIP_CHECK(hipMalloc(&b->d_nmoves, sizeof(int64_t) * size_t(max_chains)));
      const int64_t occ_cap = (occ_bytes + 15) & ~int64_t(15);
      HIP_CHECK(hipMalloc(&b->d_occ, size_t(occ_cap)));
      HIP_CHECK(hipMalloc(&b->d_trk, sizeof(uint16_t) * size_t(std::max<int64_t>(tracker_entries, 8))));
      b->occ_free[0] = occ_cap;
      if (tracker_entries > 0) b->trk_free[0] = tracker_entries;
      std::vector<ChainState> st(max_chains);
      std::memset(st.data(), 0, sizeof(ChainState) * st.size());
      HIP_CHECK(hipMemcpy(b->d_states, st.data(), sizeof(ChainState) * st.size(), hipMemcpyHostToDevice));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int cemc_pool_footprint(const cemc_model* m, int32_t kind, int64_t* occ_bytes, int64_t* tracker_entries) {
  return guarded([&] {
    require(m != nullptr, "NULL model");
    if (occ_bytes) *occ_bytes = (int64_t(m->n_sites) + 15) & ~int64_t(15);
    if (tracker_entries) *tracker_entries = kind == CEMC_MOVE_CANONICAL ? int64_t(m->n_sites) : 0;
  });
}

int cemc_pool_add(cemc_batch* b, cemc_model* m, const cemc_chain_desc* d, int32_t* slot_out) {
  return guarded([&] {
    require(b && m && d && slot_out, "NULL argument");
    require(b->dynamic, "not a pool");
    int slot = -1;
    for (int c = 0; c < b->n_chains; ++c)
      if (!b->active[c]) {
        slot = c;
        break;
      }
    if (slot < 0) throw PoolFull("no free chain slot in the pool");
    // the chain's state must fit one workgroup's LDS without energy tables
    const int occ_lds = m->periodic ? m->pdims[0] * m->pdims[1] * m->pdims[2] : m->n_sites;
    const uint32_t need = sweep_lds_bytes(occ_lds, d->kind == CEMC_MOVE_CANONICAL ? m->n_sites : 0, 0, 0);
    require(need <= lds_limit(), "chain state does not fit in the workgroup's LDS");
    auto it = std::find(b->models.begin(), b->models.end(), m);
    int mid = int(it - b->models.begin());
    if (it == b->models.end()) {
      if (int(b->models.size()) >= b->model_capacity) throw PoolFull("no free model slot in the pool");
      b->models.push_back(m);
      HIP_CHECK(hipMemcpyAsync(b->d_models + mid, &m->view, sizeof(ModelDev), hipMemcpyHostToDevice, b->stream));
      HIP_CHECK(hipStreamSynchronize(b->stream));
    }
    std::vector<uint16_t> trk;
    setup_chain(b, slot, m, mid, *d, trk);
    ChainParams& p = b->params[slot];
    const int64_t occ_len = (int64_t(m->n_sites) + 15) & ~int64_t(15);
    const int64_t occ_off = range_alloc(b->occ_free, occ_len);
    if (occ_off < 0) throw PoolFull("no room for the occupation in the pool");
    int64_t trk_off = 0;
    if (!trk.empty()) {
      trk_off = range_alloc(b->trk_free, int64_t(trk.size()));
      if (trk_off < 0) {
        range_free(b->occ_free, occ_off, occ_len);
        throw PoolFull("no room for the species tracker in the pool");
      }
    }
    p.occ_off = occ_off;
    p.trk_off = trk_off;
    try {
      // every copy and launch on the pool's own stream, and no allocation or
      // free on the way (either could wait for other streams' work)
      HIP_CHECK(hipStreamSynchronize(b->stream));
      const int n_blocks = (m->n_sites + 255) / 256;
      if (b->h_adm == nullptr) {
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&b->h_adm), 2 * sizeof(double)));
        HIP_CHECK(hipMalloc(&b->d_adm_id, sizeof(int32_t)));
        HIP_CHECK(hipMalloc(&b->d_adm_e, sizeof(double)));
      }
      if (n_blocks > b->adm_blocks) {
        dev_free(b->d_adm_partial);
        b->d_adm_partial = nullptr;
        b->adm_blocks = 0;
        // grow ahead so later, larger lattices rarely reallocate
        const int cap = std::max(n_blocks, 64);
        HIP_CHECK(hipMalloc(&b->d_adm_partial, sizeof(double) * size_t(cap)));
        b->adm_blocks = cap;
      }
      HIP_CHECK(hipMemcpyAsync(b->d_occ + occ_off, d->occupation, size_t(m->n_sites), hipMemcpyHostToDevice,
                               b->stream));
      if (!trk.empty())
        HIP_CHECK(hipMemcpyAsync(b->d_trk + trk_off, trk.data(), sizeof(uint16_t) * trk.size(),
                                 hipMemcpyHostToDevice, b->stream));
      HIP_CHECK(hipMemcpyAsync(b->d_params + slot, &p, sizeof(ChainParams), hipMemcpyHostToDevice, b->stream));
      *reinterpret_cast<int32_t*>(b->h_adm) = slot;
      HIP_CHECK(hipMemcpyAsync(b->d_adm_id, b->h_adm, sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
      HIP_CHECK(launch_energy(b->d_models, b->d_params, b->d_occ, b->d_adm_id, 1, m->n_sites, b->d_adm_partial,
                              b->d_adm_e, b->stream));
      HIP_CHECK(hipMemcpyAsync(b->h_adm + 1, b->d_adm_e, sizeof(double), hipMemcpyDeviceToHost, b->stream));
      HIP_CHECK(hipStreamSynchronize(b->stream));
      ChainState st;
      std::memset(&st, 0, sizeof st);
      st.energy = st.e_ref = b->h_adm[1];
      st.moves_done = d->first_move;
      HIP_CHECK(hipMemcpyAsync(b->d_states + slot, &st, sizeof(ChainState), hipMemcpyHostToDevice, b->stream));
      HIP_CHECK(hipStreamSynchronize(b->stream));
    } catch (...) {
      range_free(b->occ_free, occ_off, occ_len);
      if (!trk.empty()) range_free(b->trk_free, trk_off, int64_t(trk.size()));
      throw;
    }
    b->active[slot] = 1;
    *slot_out = slot;
  });
}

int cemc_pool_remove(cemc_batch* b, int32_t slot) {
  return guarded([&] {
    require(b != nullptr, "NULL pool");
    require(b->dynamic, "not a pool");
    require(slot >= 0 && slot < b->n_chains && b->active[slot], "slot is not in use");
    HIP_CHECK(hipStreamSynchronize(b->stream));
    ChainParams& p = b->params[slot];
    range_free(b->occ_free, p.occ_off, (int64_t(p.n_sites) + 15) & ~int64_t(15));
    if (p.kind == CEMC_MOVE_CANONICAL) range_free(b->trk_free, p.trk_off, int64_t(p.n_sites));
    b->active[slot] = 0;
  });
}

// ------------------------------------------------------------ tempering
int cemc_pt_create(cemc_batch* b, uint64_t seed, uint64_t uid, cemc_pt** out) {
  return guarded([&] {
    require(b && out, "NULL argument");
    *out = nullptr;
    require(!b->dynamic, "parallel tempering runs on a fixed batch, not a pool");
    const int R = b->n_chains;
    require(R >= 2, "parallel tempering needs at least two replicas");
    auto* pt = new cemc_pt();
    try {
      pt->batch = b;
      pt->n_rep = R;
      HIP_CHECK(hipStreamSynchronize(b->stream));
      std::vector<ChainState> st(R);
      HIP_CHECK(hipMemcpy(st.data(), b->d_states, sizeof(ChainState) * R, hipMemcpyDeviceToHost));
      std::vector<int32_t> ident(R);
      std::iota(ident.begin(), ident.end(), 0);
      std::vector<double> kT(R), eref(R);
      for (int t = 0; t < R; ++t) {
        kT[t] = b->params[t].kT;
        eref[t] = st[t].energy;
        // chain sums restart with the run
        st[t].e_ref = st[t].energy;
        st[t].n_samples = 0;
        st[t].sum_de = st[t].sum_de2 = 0.0;
      }
      HIP_CHECK(hipMemcpy(b->d_states, st.data(), sizeof(ChainState) * R, hipMemcpyHostToDevice));
      PTDev h;
      std::memset(&h, 0, sizeof h);
      h.seed = seed;
      h.uid = uid;
      h.n_rep = R;
      h.chain_of_slot = upload(ident);
      pt->d_cos = h.chain_of_slot;
      h.slot_of_chain = upload(ident);
      pt->d_soc = h.slot_of_chain;
      h.kT = upload(kT);
      pt->d_kT = h.kT;
      h.e_ref = upload(eref);
      pt->d_eref = h.e_ref;
      std::vector<double> zd(R, 0.0);
      std::vector<uint64_t> zu(R, 0);
      h.sum_de = upload(zd);
      pt->d_sde = h.sum_de;
      h.sum_de2 = upload(zd);
      pt->d_sde2 = h.sum_de2;
      h.n_samples = upload(zu);
      pt->d_n = h.n_samples;
      h.n_accept = upload(zu);
      pt->d_acc = h.n_accept;
      h.n_try = upload(zu);
      pt->d_try = h.n_try;
      pt->d_pt = upload(std::vector<PTDev>{h});
    } catch (...) {
      delete pt;
      throw;
    }
    *out = pt;
  });
}

int cemc_pt_destroy(cemc_pt* pt) {
  return guarded([&] { delete pt; });
}

int cemc_pt_run(cemc_pt* pt, int64_t n_cycles, int64_t moves_per_cycle, int32_t use_graph, double* kernel_ms) {
  return guarded([&] {
    require(pt != nullptr, "NULL tempering run");
    require(n_cycles >= 0 && moves_per_cycle >= 0, "cycles and moves must be non-negative");
    cemc_batch* b = pt->batch;
    std::vector<int64_t> nm(b->n_chains, moves_per_cycle);
    HIP_CHECK(hipMemcpyAsync(b->d_nmoves, nm.data(), sizeof(int64_t) * nm.size(), hipMemcpyHostToDevice, b->stream));
    int64_t done = 0;
    if (kernel_ms) HIP_CHECK(hipEventRecord(b->ev0, b->stream));
    if (use_graph && n_cycles >= 2) {
      const int G = int(std::min<int64_t>(n_cycles, kPTGraphCycles));
      if (pt->exec == nullptr || pt->graph_cycles != G || pt->graph_version != b->classes_version) {
        pt->drop_graph();
        hipGraph_t graph = nullptr;
        HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        try {
          for (int g = 0; g < G; ++g) {
            enqueue_sweeps(b, nullptr, 0);
            HIP_CHECK(launch_pt_exchange(pt->d_pt, b->d_states, b->d_params, b->stream));
          }
        } catch (...) {
          (void)hipStreamEndCapture(b->stream, &graph);
          if (graph) (void)hipGraphDestroy(graph);
          throw;
        }
        HIP_CHECK(hipStreamEndCapture(b->stream, &graph));
        const hipError_t e = hipGraphInstantiate(&pt->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_CHECK(e);
        pt->graph_cycles = G;
        pt->graph_version = b->classes_version;
      }
      for (; done + pt->graph_cycles <= n_cycles; done += pt->graph_cycles)
        HIP_CHECK(hipGraphLaunch(pt->exec, b->stream));
    }
    for (; done < n_cycles; ++done) {
      enqueue_sweeps(b, nullptr, 0);
      HIP_CHECK(launch_pt_exchange(pt->d_pt, b->d_states, b->d_params, b->stream));
    }
    if (kernel_ms) HIP_CHECK(hipEventRecord(b->ev1, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    if (kernel_ms) {
      float t = 0.f;
      HIP_CHECK(hipEventElapsedTime(&t, b->ev0, b->ev1));
      *kernel_ms = double(t);
    }
    // the exchanges moved temperatures on the device: refresh the host copy
    HIP_CHECK(hipMemcpy(b->params.data(), b->d_params, sizeof(ChainParams) * b->params.size(), hipMemcpyDeviceToHost));
    const std::string hook_error = hooks_after_launch(b);
    if (!hook_error.empty()) throw std::runtime_error(hook_error);
  });
}

int cemc_pt_get(cemc_pt* pt, int32_t* chain_of_slot, double* temperatures, int64_t* n_accept, int64_t* n_try,
                int64_t* n_samples, double* mean_energy, double* var_energy) {
  return guarded([&] {
    require(pt != nullptr, "NULL tempering run");
    const int R = pt->n_rep;
    HIP_CHECK(hipStreamSynchronize(pt->batch->stream));
    std::vector<int32_t> cos(R);
    std::vector<double> kT(R), eref(R), sde(R), sde2(R);
    std::vector<uint64_t> n(R), acc(R), tr(R);
    HIP_CHECK(hipMemcpy(cos.data(), pt->d_cos, sizeof(int32_t) * R, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(kT.data(), pt->d_kT, sizeof(double) * R, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(eref.data(), pt->d_eref, sizeof(double) * R, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(sde.data(), pt->d_sde, sizeof(double) * R, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(sde2.data(), pt->d_sde2, sizeof(double) * R, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(n.data(), pt->d_n, sizeof(uint64_t) * R, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(acc.data(), pt->d_acc, sizeof(uint64_t) * R, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(tr.data(), pt->d_try, sizeof(uint64_t) * R, hipMemcpyDeviceToHost));
    for (int t = 0; t < R; ++t) {
      if (chain_of_slot) chain_of_slot[t] = cos[t];
      if (temperatures) temperatures[t] = kT[t] / kBoltzmannEV;
      if (n_samples) n_samples[t] = int64_t(n[t]);
      const double nn = double(n[t]);
      const double m = n[t] ? sde[t] / nn : 0.0;
      if (mean_energy) mean_energy[t] = eref[t] + m;
      if (var_energy) var_energy[t] = n[t] ? sde2[t] / nn - m * m : 0.0;
      if (t < R - 1) {
        if (n_accept) n_accept[t] = int64_t(acc[t]);
        if (n_try) n_try[t] = int64_t(tr[t]);
      }
    }
  });
}

}  // extern "C"
