// instance.hip -- error state, device initialisation and tamc_instance: the reference's instance
// file rules (src/util.rs:43-109), the Ising/QUBO build (src/ising.rs:286-324), device upload and
// the quad-level schedule that makes parallel sweeps equal to the sequential site order.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <string>

#include "pt_types.hpp"
#include "runtime_common.hpp"

namespace tamc {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
tamc_status fail(tamc_status st, const std::string &msg) {
  g_last_error = msg;
  return st;
}

tamc_status device_init() {
  static std::once_flag once;
  static tamc_status st = TAMC_OK;
  static std::string msg;
  std::call_once(once, []() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
      st = TAMC_E_DEVICE;
      msg = "no HIP device visible";
      return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
      st = TAMC_E_DEVICE;
      msg = "hipGetDeviceProperties failed";
      return;
    }
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
      st = TAMC_E_DEVICE;
      msg = std::string("kernels are built for gfx950, device is ") + prop.gcnArchName;
      return;
    }
    if (hipSetDevice(0) != hipSuccess) {
      st = TAMC_E_DEVICE;
      msg = "hipSetDevice(0) failed";
    }
  });
  if (st != TAMC_OK) return fail(st, msg);
  return TAMC_OK;
}

// ------------------------------------------------------------------ parsing ----
// One line of an instance file: space0 u32 space1 u32 space1 fixed, remainder ignored.
static bool parse_instance_line(const char *s, size_t len, uint32_t *oi, uint32_t *oj, double *ok) {
  size_t p = 0;
  uint32_t v[2];
  for (int f = 0; f < 2; ++f) {
    while (p < len && (s[p] == ' ' || s[p] == '\t')) ++p;
    if (p >= len || s[p] < '0' || s[p] > '9') return false;
    uint64_t x = 0;
    while (p < len && s[p] >= '0' && s[p] <= '9') {
      x = x * 10u + (uint64_t)(s[p] - '0');
      if (x > 0xFFFFFFFFull) return false;
      ++p;
    }
    v[f] = (uint32_t)x;
    if (p >= len || (s[p] != ' ' && s[p] != '\t')) return false;
    while (p < len && (s[p] == ' ' || s[p] == '\t')) ++p;
  }
  const size_t start = p;
  if (p < len && s[p] == '-') ++p;
  const size_t dig = p;
  while (p < len && s[p] >= '0' && s[p] <= '9') ++p;
  if (p == dig) return false;
  if (p < len && s[p] == '.') {
    ++p;
    while (p < len && s[p] >= '0' && s[p] <= '9') ++p;
  }
  std::string num(s + start, p - start);
  *ok = strtod(num.c_str(), nullptr);
  *oi = v[0];
  *oj = v[1];
  return true;
}

using Adj = std::vector<std::map<uint32_t, float>>;

static void adj_put(Adj &adj, uint32_t i, uint32_t j, double K, bool verbose) {
  const size_t m = (size_t)(i > j ? i : j) + 1;
  if (adj.size() < m) adj.resize(m);
  const float k = (float)K;
  auto it = adj[i].find(j);
  if (it != adj[i].end() && verbose) printf("The entry (%u, %u) :> %g was overwritten with %g.\n", i, j, it->second, K);
  adj[i][j] = k;
  if (i != j) {
    auto it2 = adj[j].find(i);
    if (it2 != adj[j].end() && verbose)
      printf("The entry (%u, %u) :> %g was overwritten with %g.\n", i, j, it2->second, K);
    adj[j][i] = k;
  }
}

static void group_classes(SiteSchedule &S, const std::vector<uint32_t> &cls, uint32_t ncls) {
  std::vector<std::vector<uint32_t>> by(ncls);
  for (uint32_t i = 0; i < (uint32_t)cls.size(); ++i) by[cls[i]].push_back(i);
  S.n_classes = ncls;
  S.class_groups.assign(ncls + 1, 0u);
  S.group_sites.clear();
  for (uint32_t l = 0; l < ncls; ++l) {
    for (size_t k = 0; k < by[l].size(); k += 4)
      for (size_t s = 0; s < 4; ++s) S.group_sites.push_back(k + s < by[l].size() ? by[l][k + s] : kNone);
    S.class_groups[l + 1] = (uint32_t)(S.group_sites.size() / 4);
  }
  S.n_groups = (uint32_t)(S.group_sites.size() / 4);
}

static void compute_schedules(tamc_instance *I) {
  const uint32_t n = I->n;
  std::vector<uint32_t> lv(n, 0u), col(n, 0u);
  uint32_t maxl = 0, maxc = 0;
  std::vector<uint8_t> used;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t l = 0;
    used.assign(I->row_ptr[i + 1] - I->row_ptr[i] + 2, 0u);
    for (uint32_t e = I->row_ptr[i]; e < I->row_ptr[i + 1]; ++e) {
      const uint32_t j = I->col[e];
      if (j < i) {
        l = std::max(l, lv[j] + 1u);
        if (col[j] < used.size()) used[col[j]] = 1;
      }
    }
    uint32_t c = 0;
    while (used[c]) ++c;
    lv[i] = l;
    col[i] = c;
    maxl = std::max(maxl, l);
    maxc = std::max(maxc, c);
  }
  group_classes(I->sched[kOrderSequential], lv, n ? maxl + 1u : 0u);
  group_classes(I->sched[kOrderColour], col, n ? maxc + 1u : 0u);
}

static tamc_instance *build(const Adj &adj, int qubo) {
  tamc_instance *I = new tamc_instance();
  const uint32_t n = (uint32_t)adj.size();
  I->n = n;
  I->qubo = qubo;
  I->bias.assign(n, 0.f);
  I->row_ptr.assign(n + 1, 0u);
  float offset = 0.f;
  std::vector<std::pair<uint32_t, uint32_t>> pat;
  for (uint32_t i = 0; i < n; ++i) {
    for (const auto &kv : adj[i]) {
      const uint32_t j = kv.first;
      const float K = kv.second;
      if (qubo) {
        if (i != j) {
          offset += K / 8.0f;
          pat.emplace_back(i, j);
          I->col.push_back(j);
          I->J.push_back(K / 4.0f);
          I->bias[i] += K / 4.0f;
        } else {
          offset += K / 2.0f;
          I->bias[i] += K / 2.0f;
          I->col.push_back(j);
          I->J.push_back(K / 2.0f);
        }
      } else if (i != j) {
        pat.emplace_back(i, j);
        I->col.push_back(j);
        I->J.push_back(K);
      } else {
        I->bias[i] = K;
      }
    }
    I->row_ptr[i + 1] = (uint32_t)I->col.size();
  }
  I->offset = offset;
  uint32_t gn = 0;
  for (auto &e : pat) gn = std::max(gn, std::max(e.first, e.second) + 1u);
  I->g_nodes = pat.empty() ? 0u : gn;
  I->g_row_ptr.assign((size_t)I->g_nodes + 1, 0u);
  for (auto &e : pat) I->g_row_ptr[e.first + 1]++;
  for (uint32_t r = 0; r < I->g_nodes; ++r) I->g_row_ptr[r + 1] += I->g_row_ptr[r];
  I->g_col.resize(pat.size());
  std::vector<uint32_t> fill(I->g_row_ptr.begin(), I->g_row_ptr.end());
  for (auto &e : pat) I->g_col[fill[e.first]++] = e.second;
  compute_schedules(I);
  return I;
}

DeviceInstance *device_instance(const tamc_instance *cinst, uint32_t G, tamc_status *st, uint32_t order) {
  tamc_instance *inst = const_cast<tamc_instance *>(cinst);
  std::lock_guard<std::mutex> lk(inst->mu);
  *st = TAMC_OK;
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> bool {
    if (b.p) return true;
    if (b.alloc(bytes) != hipSuccess) return false;
    if (bytes && hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return false;
    return true;
  };
  if (!inst->dev) inst->dev.reset(new DeviceInstance());
  DeviceInstance *D = inst->dev.get();
  std::vector<double> chi;
  for (auto &v : inst->chi) chi.insert(chi.end(), v.begin(), v.end());
  if (!up(D->row_ptr, inst->row_ptr.data(), inst->row_ptr.size() * 4) ||
      !up(D->col, inst->col.data(), inst->col.size() * 4) || !up(D->J, inst->J.data(), inst->J.size() * 4) ||
      !up(D->bias, inst->bias.data(), inst->bias.size() * 4) ||
      !up(D->g_row_ptr, inst->g_row_ptr.data(), inst->g_row_ptr.size() * 4) ||
      !up(D->g_col, inst->g_col.data(), inst->g_col.size() * 4) || !up(D->chi, chi.data(), chi.size() * 8)) {
    *st = fail(TAMC_E_DEVICE, "instance upload failed");
    return nullptr;
  }
  if (!D->ell_col.p) {
    // fixed-stride rows (ELL) of width 3, 4, 6 or 8 when every row fits: a group's couplings then
    // load with a few wide scalar loads; padding adds +0.0, which leaves every fp32 sum unchanged
    uint32_t dmax = 0;
    for (uint32_t i = 0; i < inst->n; ++i) dmax = std::max(dmax, inst->row_ptr[i + 1] - inst->row_ptr[i]);
    const uint32_t Dl = dmax <= 3 ? 3u : dmax <= 4 ? 4u : dmax <= 6 ? 6u : dmax <= 8 ? 8u : 0u;
    std::vector<uint32_t> ec((size_t)inst->n * std::max(Dl, 1u));
    std::vector<float> eJ(ec.size(), 0.0f);
    for (uint32_t i = 0; i < inst->n && Dl; ++i)
      for (uint32_t k = 0; k < Dl; ++k) {
        const uint32_t e = inst->row_ptr[i] + k;
        const bool real = e < inst->row_ptr[i + 1];
        ec[(size_t)i * Dl + k] = real ? inst->col[e] : i;
        eJ[(size_t)i * Dl + k] = real ? inst->J[e] : 0.0f;
      }
    // 16-bit lists with rows padded to a multiple of 4 entries (padding = the row's own site)
    const uint32_t Ds = nbr_stride(Dl);
    std::vector<uint16_t> ec16((size_t)inst->n * std::max(Ds, 1u));
    for (uint32_t i = 0; i < inst->n && Dl; ++i)
      for (uint32_t k = 0; k < Ds; ++k) ec16[(size_t)i * Ds + k] = (uint16_t)(k < Dl ? ec[(size_t)i * Dl + k] : i);
    if (!up(D->ell_col, ec.data(), ec.size() * 4) || !up(D->ell_J, eJ.data(), eJ.size() * 4) ||
        (inst->n < 65536u && !up(D->ell_col16, ec16.data(), ec16.size() * 2))) {
      *st = fail(TAMC_E_DEVICE, "instance upload failed");
      return nullptr;
    }
    D->ell_D = Dl;
    // periodic lattices with row-major sites: the L x L square lattice from L = 24 on and the
    // L x L x L cubic lattice with L a multiple of 4 from 12 on: ICM grows their clusters as 64-site
    // bitboards (on smaller lattices, and on cubic ones with partial tiles, the compact tiles are as
    // fast or faster)
    if ((Dl == 4u || Dl == 6u) && inst->n < 65536u) {
      const uint32_t dim = Dl / 2u;
      const uint32_t L = (uint32_t)std::lround(dim == 2u ? std::sqrt((double)inst->n) : std::cbrt((double)inst->n));
      bool lat = dim == 2u ? (L >= 24u && L * L == inst->n) : (L >= 12u && L % 4u == 0u && L * L * L == inst->n);
      for (uint32_t i = 0; i < inst->n && lat; ++i) {
        if (inst->row_ptr[i + 1] - inst->row_ptr[i] != Dl) {
          lat = false;
          break;
        }
        // site i = (a, b, c) with c fastest (a = 0 on the square lattice); neighbours +-1 along
        // each axis, periodic
        const uint32_t c = i % L, b = (i / L) % L, a = i / (L * L);
        const uint32_t sa = L * L, sb = L;
        uint32_t want[6] = {i - c + (c + 1u) % L, i - c + (c + L - 1u) % L, i - b * sb + ((b + 1u) % L) * sb,
                            i - b * sb + ((b + L - 1u) % L) * sb, 0u, 0u};
        if (dim == 3u) {
          want[4] = i - a * sa + ((a + 1u) % L) * sa;
          want[5] = i - a * sa + ((a + L - 1u) % L) * sa;
        }
        bool seen[6] = {false, false, false, false, false, false};
        for (uint32_t e = inst->row_ptr[i]; e < inst->row_ptr[i + 1] && lat; ++e) {
          uint32_t k = 0;
          while (k < Dl && (seen[k] || want[k] != inst->col[e])) ++k;
          if (k == Dl) lat = false;
          else seen[k] = true;
        }
      }
      D->lat_L = lat ? L : 0u;
      D->lat_dim = lat ? dim : 0u;
    }
    if (Dl && inst->n < 65536u) {
      // compact tiles of at most 64 sites for ICM cluster growth: balls of the coupling graph grown
      // breadth-first from the lowest unassigned site; a ball that runs out of unassigned
      // neighbours before 64 sites goes on from the next lowest unassigned site
      const uint32_t n = inst->n;
      std::vector<uint32_t> tiles, queue, mark(n, 0u);
      std::vector<uint8_t> taken(n, 0u);
      uint32_t next = 0, stamp = 0, fill = 0;
      while (true) {
        while (next < n && taken[next]) ++next;
        if (next >= n) break;
        ++stamp;
        queue.clear();
        queue.push_back(next);
        mark[next] = stamp;
        for (size_t h = 0; h < queue.size() && fill < 64u; ++h) {
          const uint32_t i = queue[h];
          taken[i] = 1u;
          tiles.push_back(i);
          ++fill;
          for (uint32_t k = 0; k < Dl; ++k) {
            const uint32_t j = ec[(size_t)i * Dl + k];
            if (!taken[j] && mark[j] != stamp) {
              mark[j] = stamp;
              queue.push_back(j);
            }
          }
        }
        if (fill == 64u) fill = 0;  // the tile is full: the next ball starts a new one
      }
      while (tiles.size() % 64u) tiles.push_back(0xFFFFFFFFu);
      // the 16-bit neighbour rows again in tile order, so that a wave loads its tile's rows with
      // one coalesced access beside the site ids (padding positions hold zeros and are not read)
      std::vector<uint16_t> trows(tiles.size() * Ds, 0u);
      for (size_t k = 0; k < tiles.size(); ++k)
        if (tiles[k] != 0xFFFFFFFFu)
          memcpy(&trows[k * Ds], &ec16[(size_t)tiles[k] * Ds], Ds * 2);
      if (!up(D->icm_tiles, tiles.data(), tiles.size() * 4) || !up(D->icm_trows, trows.data(), trows.size() * 2)) {
        *st = fail(TAMC_E_DEVICE, "instance upload failed");
        return nullptr;
      }
      D->icm_ntiles = (uint32_t)(tiles.size() / 64u);
    }
    std::vector<float> sj(inst->n, 0.0f);
    for (uint32_t i = 0; i < inst->n; ++i)
      for (uint32_t e = inst->row_ptr[i]; e < inst->row_ptr[i + 1]; ++e)
        if (inst->col[e] == i) {
          sj[i] = inst->J[e];
          D->has_self = true;
        }
    if (!up(D->self_J, sj.data(), sj.size() * 4)) {
      *st = fail(TAMC_E_DEVICE, "instance upload failed");
      return nullptr;
    }
  }
  const uint32_t key = 2u * G + order;
  if (!D->sched.count(key)) {
    const SiteSchedule &H = inst->sched[order];
    std::unique_ptr<DeviceSchedule> S(new DeviceSchedule());
    // tasks (group, word): class-major, word-minor, so task k is word k % G of group k / G
    std::vector<uint32_t> cnt(H.n_classes + 1, 0u), tasks;
    for (uint32_t l = 0; l < H.n_classes; ++l) {
      for (uint32_t q = H.class_groups[l]; q < H.class_groups[l + 1]; ++q)
        for (uint32_t g = 0; g < G; ++g) tasks.push_back(q * G + g);
      cnt[l + 1] = (uint32_t)tasks.size();
    }
    S->n_levels = H.n_classes;
    S->n_tasks = (uint32_t)tasks.size();
    const uint32_t Dl = D->ell_D, RS = 8u + 8u * Dl;
    std::vector<uint32_t> rec((size_t)H.n_groups * RS, 0u);
    for (uint32_t q = 0; q < H.n_groups; ++q) {
      uint32_t *r = rec.data() + (size_t)q * RS;
      for (uint32_t s = 0; s < 4; ++s) {
        const uint32_t i = H.group_sites[4 * q + s];
        r[s] = i;
        if (i == kNone) continue;
        memcpy(&r[4 + s], &inst->bias[i], 4);
        for (uint32_t k = 0; k < Dl; ++k) {
          const uint32_t e = inst->row_ptr[i] + k;
          const bool real = e < inst->row_ptr[i + 1];
          r[8 + s * Dl + k] = real ? inst->col[e] : i;
          const float Jv = real ? inst->J[e] : 0.0f;
          memcpy(&r[8 + 4 * Dl + s * Dl + k], &Jv, 4);
        }
      }
    }
    S->record_words = RS;
    if (!up(S->lvl_off, cnt.data(), cnt.size() * 4) || !up(S->tasks, tasks.data(), tasks.size() * 4) ||
        !up(S->groups, H.group_sites.data(), H.group_sites.size() * 4) ||
        !up(S->records, rec.data(), rec.size() * 4)) {
      *st = fail(TAMC_E_DEVICE, "schedule upload failed");
      return nullptr;
    }
    D->sched[key] = std::move(S);
  }
  return D;
}

const DeviceInstance *resident_instance(const tamc_instance *cinst) {
  std::lock_guard<std::mutex> lk(const_cast<tamc_instance *>(cinst)->mu);
  const DeviceInstance *D = cinst->dev.get();
  if (!D || !D->sa_info.p || !D->sa_blocks.p) return nullptr;
  return D;
}

const DeviceInstance *device_sa_rows(const tamc_instance *cinst, tamc_status *st) {
  DeviceInstance *D = device_instance(cinst, 1, st);
  if (!D) return nullptr;
  tamc_instance *inst = const_cast<tamc_instance *>(cinst);
  std::lock_guard<std::mutex> lk(inst->mu);
  if (D->sa_info.p) return D;
  const uint32_t n = inst->n;
  std::vector<uint32_t> info((size_t)n * 4), blocks;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t e0 = inst->row_ptr[i], len = inst->row_ptr[i + 1] - e0, nb = (len + 7) / 8;
    info[4 * i] = (uint32_t)(blocks.size() / 16);
    info[4 * i + 1] = nb;
    memcpy(&info[4 * i + 2], &inst->bias[i], 4);
    info[4 * i + 3] = len;
    for (uint32_t b = 0; b < nb; ++b) {
      uint32_t rec[16];
      for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t e = b * 8 + k;
        const float J = e < len ? inst->J[e0 + e] : 0.0f;
        rec[k] = 8u * (e < len ? inst->col[e0 + e] : i);
        memcpy(&rec[8 + k], &J, 4);
      }
      blocks.insert(blocks.end(), rec, rec + 16);
    }
  }
  // the kernel stages rows in chunks of 16 blocks and may read up to 16 blocks past a row's first
  blocks.resize(blocks.size() + 16u * 16u, 0u);
  if (D->sa_info.alloc(info.size() * 4) != hipSuccess || D->sa_blocks.alloc(blocks.size() * 4) != hipSuccess ||
      hipMemcpy(D->sa_info.p, info.data(), info.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(D->sa_blocks.p, blocks.data(), blocks.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
    D->sa_info.reset();
    D->sa_blocks.reset();
    *st = fail(TAMC_E_DEVICE, "SA row upload failed");
    return nullptr;
  }
  return D;
}

const DeviceSchedule *device_schedule(const tamc_instance *inst, uint32_t G, tamc_status *st, uint32_t order) {
  DeviceInstance *D = device_instance(inst, G, st, order);
  if (!D) return nullptr;
  std::lock_guard<std::mutex> lk(const_cast<tamc_instance *>(inst)->mu);
  return D->sched[2u * G + order].get();
}

}  // namespace tamc

using namespace tamc;

extern "C" {

const char *tamc_last_error(void) { return g_last_error.c_str(); }
int tamc_version(void) { return 100; }
tamc_status tamc_init(void) { return device_init(); }

tamc_status tamc_instance_from_file(const char *path, int qubo, tamc_instance **out) {
  clear_error();
  if (!path || !out) return fail(TAMC_E_INVALID_ARG, "null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(TAMC_E_IO, std::string("Unable to read adjancency from instance file ") + path);
  Adj adj;
  std::string line;
  uint64_t ln = 0;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    uint32_t i, j;
    double K;
    if (parse_instance_line(line.data(), line.size(), &i, &j, &K)) adj_put(adj, i, j, K, true);
    else printf("Ignoring line %llu: %s\n", (unsigned long long)ln, line.c_str());
    ++ln;
  }
  *out = build(adj, qubo);
  return TAMC_OK;
}

tamc_status tamc_instance_from_lines(uint64_t nlines, const uint32_t *i, const uint32_t *j, const double *K, int qubo,
                                     tamc_instance **out) {
  clear_error();
  if (!out || (nlines && (!i || !j || !K))) return fail(TAMC_E_INVALID_ARG, "null argument");
  Adj adj;
  for (uint64_t k = 0; k < nlines; ++k) adj_put(adj, i[k], j[k], K[k], false);
  *out = build(adj, qubo);
  return TAMC_OK;
}

tamc_status tamc_instance_create(const tamc_instance_desc *d, tamc_instance **out) {
  clear_error();
  if (!d || !out) return fail(TAMC_E_INVALID_ARG, "null argument");
  *out = nullptr;
  const uint32_t n = d->n;
  if (!d->row_ptr || (d->nnz && (!d->col || !d->J)) || (d->nchi && !d->suscept))
    return fail(TAMC_E_INVALID_ARG, "null array in the instance description");
  if (d->row_ptr[0] != 0u || d->row_ptr[n] != d->nnz) return fail(TAMC_E_INVALID_ARG, "row_ptr must run from 0 to nnz");
  for (uint32_t i = 0; i < n; ++i)
    if (d->row_ptr[i + 1] < d->row_ptr[i]) return fail(TAMC_E_INVALID_ARG, "row_ptr must be non-decreasing");
  for (uint32_t e = 0; e < d->nnz; ++e)
    if (d->col[e] >= n) return fail(TAMC_E_INVALID_ARG, "column index out of range");
  // the off-diagonal pattern must be symmetric: count (i, j) and (j, i) occurrences
  std::vector<std::pair<uint32_t, uint32_t>> fwd, rev;
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t e = d->row_ptr[i]; e < d->row_ptr[i + 1]; ++e)
      if (d->col[e] != i) {
        fwd.emplace_back(i, d->col[e]);
        rev.emplace_back(d->col[e], i);
      }
  std::sort(fwd.begin(), fwd.end());
  std::sort(rev.begin(), rev.end());
  if (std::adjacent_find(fwd.begin(), fwd.end()) != fwd.end())
    return fail(TAMC_E_INVALID_ARG, "a row lists the same column twice");
  if (fwd != rev) return fail(TAMC_E_INVALID_ARG, "the coupling pattern must be symmetric (j in row i <=> i in row j)");
  tamc_instance *I = new tamc_instance();
  I->n = n;
  I->qubo = d->qubo;
  I->row_ptr.assign(d->row_ptr, d->row_ptr + n + 1);
  I->col.assign(d->col, d->col + d->nnz);
  I->J.assign(d->J, d->J + d->nnz);
  I->bias.assign(n, 0.0f);
  if (d->bias) I->bias.assign(d->bias, d->bias + n);
  I->offset = d->offset;
  // ICM graph: the CsMat of the off-diagonal entries, rows ascending (sprs CSR order)
  uint32_t gn = 0;
  for (auto &e : fwd) gn = std::max(gn, std::max(e.first, e.second) + 1u);
  I->g_nodes = fwd.empty() ? 0u : gn;
  I->g_row_ptr.assign((size_t)I->g_nodes + 1, 0u);
  for (auto &e : fwd) I->g_row_ptr[e.first + 1]++;
  for (uint32_t r = 0; r < I->g_nodes; ++r) I->g_row_ptr[r + 1] += I->g_row_ptr[r];
  I->g_col.resize(fwd.size());
  for (size_t k = 0; k < fwd.size(); ++k) I->g_col[k] = fwd[k].second;
  I->chi.assign(d->nchi, std::vector<double>(n));
  for (uint32_t k = 0; k < d->nchi; ++k) memcpy(I->chi[k].data(), d->suscept + (size_t)k * n, n * sizeof(double));
  compute_schedules(I);
  *out = I;
  return TAMC_OK;
}

tamc_status tamc_instance_set_suscept(tamc_instance *inst, uint32_t nchi, const double *w) {
  clear_error();
  if (!inst || (nchi && !w)) return fail(TAMC_E_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(inst->mu);
  if (inst->server_uses.load())
    return fail(TAMC_E_INVALID_ARG, "instance has server requests queued or running; it is immutable until they end");
  inst->chi.assign(nchi, std::vector<double>(inst->n));
  for (uint32_t k = 0; k < nchi; ++k) memcpy(inst->chi[k].data(), w + (size_t)k * inst->n, inst->n * sizeof(double));
  if (inst->dev) inst->dev->chi.reset();
  return TAMC_OK;
}

tamc_status tamc_instance_add_suscept_file(tamc_instance *inst, const char *path) {
  clear_error();
  if (!inst || !path) return fail(TAMC_E_INVALID_ARG, "null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(TAMC_E_IO, std::string("Unable to open susceptibility file ") + path);
  std::vector<double> v;
  std::string line;
  uint64_t ln = 0;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    // space0 fixed multispace0 (src/util.rs:117-121)
    size_t p = 0;
    while (p < line.size() && (line[p] == ' ' || line[p] == '\t')) ++p;
    const size_t s = p;
    if (p < line.size() && line[p] == '-') ++p;
    const size_t d = p;
    while (p < line.size() && isdigit((unsigned char)line[p])) ++p;
    if (p == d) {
      printf("Ignoring line %llu: %s\n", (unsigned long long)ln, line.c_str());
    } else {
      if (p < line.size() && line[p] == '.') {
        ++p;
        while (p < line.size() && isdigit((unsigned char)line[p])) ++p;
      }
      v.push_back(strtod(line.substr(s, p - s).c_str(), nullptr));
    }
    ++ln;
  }
  std::lock_guard<std::mutex> lk(inst->mu);
  if (inst->server_uses.load())
    return fail(TAMC_E_INVALID_ARG, "instance has server requests queued or running; it is immutable until they end");
  if (v.size() != inst->n)
    printf("WARNING: Ignoring suscept file %s - Expected %u coefficients, but got %zu\n", path, inst->n, v.size());
  inst->chi.push_back(v);
  if (inst->dev) inst->dev->chi.reset();
  return TAMC_OK;
}

tamc_status tamc_instance_get_info(const tamc_instance *inst, tamc_instance_info *info) {
  clear_error();
  if (!inst || !info) return fail(TAMC_E_INVALID_ARG, "null argument");
  info->n = inst->n;
  info->nnz = (uint32_t)inst->col.size();
  info->graph_nnz = (uint32_t)inst->g_col.size();
  info->graph_nodes = inst->g_nodes;
  info->nchi = (uint32_t)inst->chi.size();
  info->offset = inst->offset;
  info->qubo = inst->qubo;
  return TAMC_OK;
}

tamc_status tamc_instance_get_csr(const tamc_instance *inst, uint32_t *row_ptr, uint32_t *col, float *J, float *bias) {
  clear_error();
  if (!inst) return fail(TAMC_E_INVALID_ARG, "null argument");
  if (row_ptr) memcpy(row_ptr, inst->row_ptr.data(), inst->row_ptr.size() * 4);
  if (col && !inst->col.empty()) memcpy(col, inst->col.data(), inst->col.size() * 4);
  if (J && !inst->J.empty()) memcpy(J, inst->J.data(), inst->J.size() * 4);
  if (bias && inst->n) memcpy(bias, inst->bias.data(), inst->n * 4);
  return TAMC_OK;
}

tamc_status tamc_instance_get_graph(const tamc_instance *inst, uint32_t *row_ptr, uint32_t *col) {
  clear_error();
  if (!inst) return fail(TAMC_E_INVALID_ARG, "null argument");
  if (row_ptr) memcpy(row_ptr, inst->g_row_ptr.data(), inst->g_row_ptr.size() * 4);
  if (col && !inst->g_col.empty()) memcpy(col, inst->g_col.data(), inst->g_col.size() * 4);
  return TAMC_OK;
}

// induced_subgraph_instance_sorted (src/ising.rs:327-355): the sites `nodes` (ascending) become
// 0..count-1; offset 0; biases, coupling rows (QUBO self entries included) and susceptibility
// coefficients restricted to them.  The ICM graph is the pattern of the restricted rows, self
// entries included, as the reference's TriMat of the same triplets.
tamc_status tamc_instance_induced_subgraph(const tamc_instance *inst, const uint32_t *nodes, uint32_t count,
                                           tamc_instance **out) {
  clear_error();
  if (!inst || !out || (count && !nodes)) return fail(TAMC_E_INVALID_ARG, "null argument");
  for (uint32_t k = 0; k < count; ++k) {
    if (nodes[k] >= inst->n) return fail(TAMC_E_INVALID_ARG, "induced subgraph: site index out of range");
    if (k && nodes[k] <= nodes[k - 1]) return fail(TAMC_E_INVALID_ARG, "induced subgraph: sites must ascend");
  }
  for (const auto &sc : inst->chi)
    if (sc.size() != inst->n) return fail(TAMC_E_INVALID_ARG, "induced subgraph: a susceptibility vector has the wrong length");
  tamc_instance *I = new tamc_instance();
  I->n = count;
  I->qubo = inst->qubo;
  I->offset = 0.f;
  I->bias.resize(count);
  I->row_ptr.assign((size_t)count + 1, 0u);
  std::vector<std::pair<uint32_t, uint32_t>> pat;
  for (uint32_t a = 0; a < count; ++a) {
    const uint32_t ni = nodes[a];
    I->bias[a] = inst->bias[ni];
    for (uint32_t e = inst->row_ptr[ni]; e < inst->row_ptr[ni + 1]; ++e) {
      const uint32_t* hit = std::lower_bound(nodes, nodes + count, inst->col[e]);
      if (hit == nodes + count || *hit != inst->col[e]) continue;
      const uint32_t b = (uint32_t)(hit - nodes);
      I->col.push_back(b);
      I->J.push_back(inst->J[e]);
      pat.emplace_back(a, b);
    }
    I->row_ptr[a + 1] = (uint32_t)I->col.size();
  }
  uint32_t gn = 0;
  for (auto &e : pat) gn = std::max(gn, std::max(e.first, e.second) + 1u);
  I->g_nodes = pat.empty() ? 0u : gn;
  I->g_row_ptr.assign((size_t)I->g_nodes + 1, 0u);
  for (auto &e : pat) I->g_row_ptr[e.first + 1]++;
  for (uint32_t r = 0; r < I->g_nodes; ++r) I->g_row_ptr[r + 1] += I->g_row_ptr[r];
  I->g_col.resize(pat.size());
  std::vector<uint32_t> fill(I->g_row_ptr.begin(), I->g_row_ptr.end());
  for (auto &e : pat) I->g_col[fill[e.first]++] = e.second;
  for (const auto &sc : inst->chi) {
    std::vector<double> v(count);
    for (uint32_t a = 0; a < count; ++a) v[a] = sc[nodes[a]];
    I->chi.push_back(std::move(v));
  }
  compute_schedules(I);
  *out = I;
  return TAMC_OK;
}

void tamc_instance_destroy(tamc_instance *inst) { delete inst; }

}  // extern "C"
