// tamc_oracle.cpp -- CPU restatement of the tamc sampling path.  TEST INFRASTRUCTURE ONLY.
//
// Each function cites the reference file:line it restates (paths are relative to the
// USCqserver/tamc tree).  Floating-point statements keep the reference's operation order and
// f32/f64 types; the file is compiled with -ffp-contract=off so no FMA contraction changes a
// rounding.  exp is glibc expf, which is what Rust's f32::exp lowers to on Linux.
//
// Parity pins (each is exercised by tests/test_oracle_*.py):
//   Philox4x32-10          Random123 known-answer vectors (3 KATs)
//   Xoshiro256++ / jump    reference-paper known-answer stream (splitmix64 seeding)
//   parser                 hand-written line cases covering every branch of src/util.rs:43-65
//   instance build         independent numpy rebuild of bias/offset/J (Ising and QUBO rules)
//   energy_ref / dE        brute-force Hamiltonian in numpy; ferromagnet ground state -2N;
//                          dE == E(flipped) - E exactly for +-J and QUBO; QUBO self-entry shift
//   sweep                  exact enumeration of <E>(beta) for N<=12 (statistical), detailed
//                          lock-step trace vs a pure-Python loop on a tiny instance
//   exchange               closed-form decisions from (E, u) for both variants
//   Houdayer               scipy connected_components of the q<0-induced subgraph; E1+E2 conserved
//   measurements / packing independent numpy bit packing, overlap and suscept
//   gen_index              parity unpinned: restates rand 0.8 sample_single (no vectors exist)
//   bincode/YAML layout    checked by the product's own writers against this result
#include "tamc_oracle.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// ---------------------------------------------------------------- RNG ----
inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// Xoshiro256++ (rand_xoshiro 0.6, used at src/pt.rs:361,370 and src/sa.rs:141,145)
struct Xoshiro {
  uint64_t s[4];
  uint64_t next_u64() {
    const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
  }
  // rand_xoshiro: next_u32 keeps the upper half of next_u64
  uint32_t next_u32() { return (uint32_t)(next_u64() >> 32); }
  void jump() {
    static const uint64_t J[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL, 0xa9582618e03fc9aaULL,
                                  0x39abdc4529b1661cULL};
    uint64_t a[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w)
      for (int b = 0; b < 64; ++b) {
        if (J[w] & (1ULL << b))
          for (int q = 0; q < 4; ++q) a[q] ^= s[q];
        next_u64();
      }
    std::memcpy(s, a, sizeof a);
  }
  void seed(uint64_t x) {  // seed_from_u64: four splitmix64 outputs
    for (int q = 0; q < 4; ++q) {
      x += 0x9e3779b97f4a7c15ULL;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s[q] = z ^ (z >> 31);
    }
  }
};

struct StreamExhausted : std::runtime_error {
  StreamExhausted() : std::runtime_error("uniform stream exhausted") {}
};
struct SeedOutsideGraph : std::runtime_error {
  SeedOutsideGraph() : std::runtime_error("ICM seed outside the coupling graph") {}
};

// A source of u32 words: either an owned Xoshiro or a caller-supplied array.
struct WordStream {
  bool is_feed = false;
  Xoshiro x{};
  const uint32_t *w = nullptr;
  uint64_t len = 0, pos = 0;
  uint32_t next_u32() {
    if (!is_feed) return x.next_u32();
    if (pos >= len) throw StreamExhausted();
    return w[pos++];
  }
  // rand 0.8 Standard f32: 24 high bits of next_u32 times 2^-24
  float next_f32() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }
  // rand 0.8 Uniform::new_inclusive(0i8, 1): top bit of next_u32 (src/ising.rs:172)
  int8_t next_spin() { return (next_u32() >> 31) ? (int8_t)1 : (int8_t)-1; }
  // SliceRandom::choose -> gen_index -> gen_range(0..len as u32): widening multiply with the
  // "leading zeros" rejection zone of rand 0.8 UniformInt::sample_single (src/pt.rs:53)
  uint32_t gen_index(uint32_t range) {
    const uint32_t zone = (uint32_t)((range << __builtin_clz(range)) - 1u);
    for (;;) {
      const uint64_t m = (uint64_t)next_u32() * (uint64_t)range;
      const uint32_t hi = (uint32_t)(m >> 32), lo = (uint32_t)m;
      if (lo <= zone) return hi;
    }
  }
};

inline void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Production counter layout of the GPU path (stream tags in ctr[3]).
enum : uint32_t { TAG_SWEEP = 0, TAG_EXCHANGE = 1, TAG_ICM = 2, TAG_INIT = 3 };
inline uint32_t philox_word(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t tag, uint32_t lane) {
  const uint32_t ctr[4] = {c0, c1, c2, tag};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t out[4];
  philox(ctr, key, out);
  return out[lane & 3];
}
inline float u24(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }

// ----------------------------------------------------------- instance ----
struct Inst {
  uint32_t n = 0;
  std::vector<uint32_t> row_ptr, col;  // coupling_vecs, flattened (src/ising.rs:261, src/csr.rs:18-34)
  std::vector<float> J, bias;
  float offset = 0.f;
  std::vector<uint32_t> g_row_ptr, g_col;  // CsMat nonzeros -> petgraph Csr (src/pt.rs:343-346)
  uint32_t g_nodes = 0;                    // Csr::from_sorted_edges node count = max index + 1
  std::vector<std::vector<double>> chi;    // suscept_coefs
};

// src/util.rs:43-65: space0 u32 space1 (twice), then fixed ('-'? digit1 ('.' digit0)?), rest ignored
bool parse_line(const std::string &line, uint32_t &i, uint32_t &j, double &K) {
  const size_t n = line.size();
  size_t p = 0;
  auto is_sp = [&](size_t q) { return q < n && (line[q] == ' ' || line[q] == '\t'); };
  auto is_dig = [&](size_t q) { return q < n && line[q] >= '0' && line[q] <= '9'; };
  uint32_t v[2];
  for (int k = 0; k < 2; ++k) {
    while (is_sp(p)) ++p;
    const size_t s = p;
    uint64_t acc = 0;
    while (is_dig(p)) {
      acc = acc * 10 + (uint64_t)(line[p] - '0');
      if (acc > 0xFFFFFFFFull) return false;  // the reference would panic in u32::from_str
      ++p;
    }
    if (p == s) return false;
    v[k] = (uint32_t)acc;
    const size_t s2 = p;
    while (is_sp(p)) ++p;
    if (p == s2) return false;
  }
  const size_t s = p;
  if (p < n && line[p] == '-') ++p;
  const size_t d0 = p;
  while (is_dig(p)) ++p;
  if (p == d0) return false;
  if (p < n && line[p] == '.') {
    ++p;
    while (is_dig(p)) ++p;
  }
  K = std::strtod(line.substr(s, p - s).c_str(), nullptr);
  i = v[0];
  j = v[1];
  return true;
}

typedef std::vector<std::map<uint32_t, float>> AdjList;

// src/util.rs:76-109 (BTreeMap insert overwrites; off-diagonal entries stored in both rows)
void adj_insert(AdjList &adj, uint32_t i, uint32_t j, double K) {
  const size_t m = (size_t)std::max(i, j) + 1;
  if (adj.size() < m) adj.resize(m);
  adj[i][j] = (float)K;
  if (i != j) adj[j][i] = (float)K;
}

// src/ising.rs:286-324
Inst *build_instance(const AdjList &adj, bool qubo) {
  Inst *I = new Inst();
  const uint32_t n = (uint32_t)adj.size();
  I->n = n;
  I->bias.assign(n, 0.f);
  I->row_ptr.assign(n + 1, 0);
  float offset = 0.f;
  std::vector<std::pair<uint32_t, uint32_t>> tri;  // CsMat pattern (off-diagonal only)
  for (uint32_t i = 0; i < n; ++i) {
    for (const auto &kv : adj[i]) {
      const uint32_t j = kv.first;
      const float K = kv.second;
      if (qubo) {
        if (i != j) {
          offset += K / 8.0f;
          tri.emplace_back(i, j);
          I->col.push_back(j);
          I->J.push_back(K / 4.0f);
          I->bias[i] += K / 4.0f;
        } else {
          offset += K / 2.0f;
          I->bias[i] += K / 2.0f;
          I->col.push_back(j);
          I->J.push_back(K / 2.0f);
        }
      } else {
        if (i != j) {
          tri.emplace_back(i, j);
          I->col.push_back(j);
          I->J.push_back(K);
        } else {
          I->bias[i] = K;
        }
      }
    }
    I->row_ptr[i + 1] = (uint32_t)I->col.size();
  }
  I->offset = offset;
  // ICM graph: edges of the CsMat (explicit zeros kept), rows ascending, cols ascending
  uint32_t gn = 0;
  for (auto &e : tri) gn = std::max(gn, std::max(e.first, e.second) + 1);
  I->g_nodes = tri.empty() ? 0 : gn;
  I->g_row_ptr.assign((size_t)I->g_nodes + 1, 0);
  for (auto &e : tri) I->g_row_ptr[e.first + 1]++;
  for (uint32_t r = 0; r < I->g_nodes; ++r) I->g_row_ptr[r + 1] += I->g_row_ptr[r];
  I->g_col.resize(tri.size());
  {
    std::vector<uint32_t> fill(I->g_row_ptr.begin(), I->g_row_ptr.end() - (I->g_nodes ? 1 : 0));
    for (auto &e : tri) I->g_col[fill[e.first]++] = e.second;
  }
  return I;
}

// ------------------------------------------------------- elementary ----
// src/ising.rs:438-454
float energy_ref(const Inst &I, const int8_t *s) {
  float total = I.offset;
  for (uint32_t i = 0; i < I.n; ++i) {
    float h = 0.0f;
    const float si = (float)s[i];
    h += I.bias[i];
    for (uint32_t e = I.row_ptr[i]; e < I.row_ptr[i + 1]; ++e) {
      const float sj = (float)s[I.col[e]];
      h += (I.J[e] * sj) / 2.0f;
    }
    total += h * si;
  }
  return total;
}

// src/ising.rs:468-480
inline float delta_energy(const Inst &I, const int8_t *s, uint32_t i) {
  float d = 0.0f;
  d += I.bias[i];
  for (uint32_t e = I.row_ptr[i]; e < I.row_ptr[i + 1]; ++e) d += I.J[e] * (float)s[I.col[e]];
  const float si = (float)s[i];
  d *= -2.0f * si;
  return d;
}

// tamc-core/src/metropolis.rs:43-57 inside src/ising.rs:225-233; U(i) is called only when dE >= 0
template <class U>
float sweep(const Inst &I, int8_t *s, float beta, U &&draw, const uint32_t *order = nullptr) {
  float de = 0.0f;
  for (uint32_t k = 0; k < I.n; ++k) {
    const uint32_t i = order ? order[k] : k;
    const float d = delta_energy(I, s, i);
    if (d < 0.0f) {
      s[i] = (int8_t)-s[i];
      de += d;
    } else {
      const float p = expf(-beta * d);
      const float x = draw(i);
      if (x < p) {
        s[i] = (int8_t)-s[i];
        de += d;
      }
    }
  }
  return de;
}

// Tracked-energy accounting of the production path (not a reference concept; selected by
// energy_fixed): energy changes are summed exactly as integers in units of 2^-32, and the sum is
// added to the f32 energy once per sweep (or per ICM move).
inline long long to_fix(float x) { return (long long)std::rint((double)x * 4294967296.0); }
inline float add_fix(float e, long long s) { return e + (float)((double)s * 2.3283064365386962890625e-10); }

// The Metropolis sweep of `sweep` with that accounting.  The accepted dE of the (up to) four sites
// of one RNG group are first summed in f32 in the group's site order, starting from 0; the group
// sums enter the exact total.  Returns that total; `drift` gains -2 J_ii for every accepted flip of
// a site with a QUBO self entry J_ii (those flips move the tracked energy by dE_true - 2 J_ii,
// src/ising.rs:308-310 and 468-480).
template <class U>
long long sweep_fixed(const Inst &I, int8_t *s, float beta, U &&draw, const uint32_t *order, const uint32_t *gid,
                      std::vector<float> &group_sum, const float *self_J, long long *drift) {
  std::fill(group_sum.begin(), group_sum.end(), 0.0f);
  for (uint32_t k = 0; k < I.n; ++k) {
    const uint32_t i = order ? order[k] : k;
    const float d = delta_energy(I, s, i);
    bool ok = d < 0.0f;
    if (!ok) ok = draw(i) < expf(-beta * d);
    if (ok) {
      s[i] = (int8_t)-s[i];
      group_sum[gid[i]] += d;
      if (self_J[i] != 0.0f) *drift += to_fix(-2.0f * self_J[i]);
    }
  }
  long long acc = 0;
  for (float x : group_sum) acc += to_fix(x);
  return acc;
}

// The exact energy change of replica `a` when the cluster's spins flip: the sum over cluster sites
// i of -2 s_i (h_i + sum of J_ij s_j over the row entries j outside the cluster, in row order).
long long cluster_delta_fixed(const Inst &I, const int8_t *a, const std::vector<uint8_t> &cl) {
  long long tot = 0;
  for (uint32_t i = 0; i < I.n; ++i) {
    if (!cl[i]) continue;
    float hh = I.bias[i];
    for (uint32_t e = I.row_ptr[i]; e < I.row_ptr[i + 1]; ++e)
      if (!cl[I.col[e]]) hh += I.J[e] * (float)a[I.col[e]];
    tot += to_fix(a[i] < 0 ? 2.0f * hh : -2.0f * hh);
  }
  return tot;
}

// Production sweep schedule of the GPU path (not a reference concept).  Sites are split into
// classes of mutually uncoupled sites and swept class by class, ascending inside a class:
//   order 0 (colour): greedy colouring in site order, colour(i) = the smallest colour not taken by
//            a lower-numbered site in row i;
//   order 1 (sequential): level(i) = 1 + max level of the lower-numbered sites in row i (0 if
//            none); sweeping level by level is the same as sweeping 0..n-1.
// Within a class, sites in ascending order form groups of four numbered class-major.  The PT
// uniform for (sweep t, site i, replica rho) is word slot(i) of Philox4x32-10(ctr = {t, group(i),
// rho, 0}).
struct RngGroups {
  std::vector<uint32_t> gid, slot, order;
  RngGroups(const Inst &I, uint32_t sweep_order) : gid(I.n), slot(I.n) {
    std::vector<uint32_t> lv(I.n, 0), col(I.n, 0);
    uint32_t maxl = 0, maxc = 0;
    for (uint32_t i = 0; i < I.n; ++i) {
      uint32_t l = 0;
      std::vector<bool> taken;
      for (uint32_t e = I.row_ptr[i]; e < I.row_ptr[i + 1]; ++e) {
        const uint32_t j = I.col[e];
        if (j < i) {
          l = std::max(l, lv[j] + 1);
          if (col[j] >= taken.size()) taken.resize(col[j] + 1, false);
          taken[col[j]] = true;
        }
      }
      uint32_t c = 0;
      while (c < taken.size() && taken[c]) ++c;
      lv[i] = l;
      col[i] = c;
      maxl = std::max(maxl, l);
      maxc = std::max(maxc, c);
    }
    const std::vector<uint32_t> &cls = sweep_order == 0 ? col : lv;
    std::vector<std::vector<uint32_t>> by(I.n ? (sweep_order == 0 ? maxc : maxl) + 1 : 0);
    for (uint32_t i = 0; i < I.n; ++i) by[cls[i]].push_back(i);
    uint32_t base = 0;
    for (auto &v : by) {
      for (size_t k = 0; k < v.size(); ++k) {
        gid[v[k]] = base + (uint32_t)(k / 4);
        slot[v[k]] = (uint32_t)(k % 4);
        order.push_back(v[k]);
      }
      base += (uint32_t)((v.size() + 3) / 4);
    }
  }
};

struct PhiloxGroupDraw {
  uint64_t seed;
  uint32_t t, rho;
  const RngGroups *G;
  float operator()(uint32_t i) const { return u24(philox_word(seed, t, G->gid[i], rho, TAG_SWEEP, G->slot[i])); }
};

// Philox uniforms keyed by site quads (SA path); the quad block is computed on first use.
struct PhiloxSweepDraw {
  uint64_t seed;
  uint32_t t, rho;
  uint32_t quad = 0xFFFFFFFFu;
  uint32_t out[4] = {0, 0, 0, 0};
  float operator()(uint32_t i) {
    if ((i >> 2) != quad) {
      quad = i >> 2;
      const uint32_t ctr[4] = {t, quad, rho, TAG_SWEEP};
      const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
      philox(ctr, key, out);
    }
    return u24(out[i & 3]);
  }
};

void pack_bytes(const int8_t *s, uint32_t n, uint8_t *out) {  // src/ising.rs:96-121
  const uint32_t nb = n / 8 + (n % 8 ? 1 : 0);
  std::memset(out, 0, nb);
  for (uint32_t i = 0; i < n; ++i)
    if (!(s[i] > 0)) out[i / 8] |= (uint8_t)(1u << (i % 8));
}
void pack_u64(const int8_t *s, uint32_t n, uint64_t *out) {  // src/ising.rs:123-137
  const uint32_t nw = n / 64 + (n % 64 ? 1 : 0);
  std::memset(out, 0, nw * sizeof(uint64_t));
  for (uint32_t i = 0; i < n; ++i)
    if (!(s[i] > 0)) out[i / 64] |= (1ULL << (i % 64));
}

// src/pt.rs:31-77 with the seed already chosen (index k into the ascending q<0 list).
// Returns cluster size; fills `in_cluster` if non-null.
int64_t houdayer_k(const Inst &I, int8_t *s1, int8_t *s2, uint32_t k, std::vector<uint8_t> *in_cluster) {
  const uint32_t n = I.n;
  std::vector<uint32_t> idxs;
  for (uint32_t i = 0; i < n; ++i)
    if ((int)s1[i] * (int)s2[i] < 0) idxs.push_back(i);
  if (idxs.empty()) return 0;
  const uint32_t seed = idxs[k];
  if (seed >= I.g_nodes) throw SeedOutsideGraph();  // FixedBitSet::insert panics there
  std::vector<uint8_t> disc(I.g_nodes, 0);
  std::deque<uint32_t> q;
  disc[seed] = 1;
  q.push_back(seed);
  int64_t size = 1;
  while (!q.empty()) {
    const uint32_t v = q.front();
    q.pop_front();
    for (uint32_t e = I.g_row_ptr[v]; e < I.g_row_ptr[v + 1]; ++e) {
      const uint32_t w = I.g_col[e];
      if ((int)s1[w] * (int)s2[w] < 0 && !disc[w]) {
        disc[w] = 1;
        ++size;
        q.push_back(w);
      }
    }
  }
  for (uint32_t i = 0; i < I.g_nodes; ++i)
    if (disc[i]) std::swap(s1[i], s2[i]);
  if (in_cluster) {
    in_cluster->assign(n, 0);
    for (uint32_t i = 0; i < I.g_nodes; ++i) (*in_cluster)[i] = disc[i];
  }
  return size;
}

// --------------------------------------------------------- PT runner ----
enum Tag : uint8_t { T_NONE = 0, T_MIN = 1, T_MAX = 2, T_MIN_REFL = 3, T_MAX_REFL = 4 };

struct St {
  std::vector<int8_t> s;
  float energy = 0.f;
  bool init = false;
  uint32_t id = 0;
  long long drift = 0;  // energy_fixed: tracked energy minus energy_ref, from QUBO self entries
};

struct Chain {  // tamc-core/src/pt.rs:18-48
  std::vector<St> st;
  std::vector<uint32_t> nacc;
  uint32_t rt = 0;
  std::vector<uint8_t> tags;
  std::vector<uint32_t> hist;  // [B][2]
  explicit Chain(uint32_t B) : st(B), nacc(B, 0), tags(B, T_NONE), hist(2 * (size_t)B, 0) {
    tags[0] = T_MIN;
    tags[B - 1] = T_MAX;
  }
  void swap_states(uint32_t j) {  // tamc-core/src/pt.rs:49-53
    std::swap(st[j], st[j + 1]);
    std::swap(tags[j], tags[j + 1]);
    nacc[j] += 1;
  }
  void update_round_trips() {  // tamc-core/src/pt.rs:54-84
    uint8_t &lo = tags[0];
    switch (lo) {
      case T_NONE: lo = T_MIN; break;
      case T_MAX_REFL: rt += 1; lo = T_MIN; break;
      case T_MAX: lo = T_MIN_REFL; break;
      default: break;
    }
    uint8_t &hi = tags.back();
    switch (hi) {
      case T_NONE: hi = T_MAX; break;
      case T_MIN_REFL: rt += 1; hi = T_MAX; break;
      case T_MIN: hi = T_MAX_REFL; break;
      default: break;
    }
    for (size_t b = 0; b < tags.size(); ++b) {
      if (tags[b] == T_MIN || tags[b] == T_MIN_REFL) hist[2 * b] += 1;
      else if (tags[b] == T_MAX || tags[b] == T_MAX_REFL) hist[2 * b + 1] += 1;
    }
  }
};

inline float energy(const Inst &I, St &x) {  // src/ising.rs:455-463
  if (x.init) return x.energy;
  x.energy = energy_ref(I, x.s.data());
  x.init = true;
  return x.energy;
}

// tamc-core/src/pt.rs:126-143 (variant 0) and tamc-core/src/parallel/pt.rs:202-215 (variant 1)
template <class U>
void exchange_pass(uint32_t B, const float *dbeta, const float *E, int variant, U &&draw,
                   std::vector<uint8_t> *acc, const std::function<void(uint32_t)> &swap) {
  if (B < 2) return;
  std::vector<float> dE(B - 1);
  for (uint32_t j = 0; j + 1 < B; ++j) dE[j] = E[j + 1] - E[j];
  for (uint32_t j = 0; j + 1 < B; ++j) {
    const float dlt = dbeta[j] * dE[j];
    bool ok = dlt >= 0.0f;
    if (!ok) ok = draw(j) < expf(dlt);
    if (acc) (*acc)[j] = ok;
    if (ok) {
      swap(j);
      if (variant == 0) {
        if (j + 2 < B) dE[j + 1] = dE[j + 1] + dE[j];
        dE[j] = -dE[j];
      }
    }
  }
}

struct PtRun {
  const Inst &I;
  const tamco_pt_params &P;
  uint32_t n, B, C, T, lo, meas_init, npairs, nchi;
  std::vector<float> beta, dbeta;
  std::vector<Chain> chains;
  std::vector<WordStream> streams;  // serial: 1; threaded: C*B (+ base stream first for init)
  // results
  std::vector<std::vector<float>> e;
  std::vector<std::vector<int32_t>> q;
  std::vector<std::vector<std::vector<float>>> chi;
  std::vector<std::vector<std::vector<uint8_t>>> samples;
  uint32_t n_meas = 0, n_samp_events = 0;
  bool have_min = false;
  float min_e = 0.f;
  std::vector<uint32_t> gs_steps;
  std::vector<float> gs_e;
  std::vector<uint64_t> gs_states;
  uint64_t icm_moves = 0, attempts = 0;
  RngGroups groups;
  bool fixed = false;          // production tracked-energy accounting (energy_fixed)
  std::vector<float> self_J;   // QUBO self entry of each site (0 where none)
  std::vector<float> group_sum;  // per RNG group f32 sums of a sweep's accepted dE

  PtRun(const Inst &I_, const tamco_pt_params &P_) : I(I_), P(P_), groups(I_, P_.sweep_order) {
    n = I.n;
    B = P.num_betas;
    C = P.num_chains;
    T = P.num_sweeps;
    lo = P.lo_beta_idx;
    npairs = C / 2;
    nchi = (uint32_t)I.chi.size();
    beta.assign(P.betas, P.betas + B);
    for (uint32_t j = 0; j + 1 < B; ++j) dbeta.push_back(beta[j + 1] - beta[j]);  // tamc-core/src/pt.rs:104-107
    double mi = P.warmup_fraction * (double)T;                                     // src/pt.rs:348
    meas_init = mi <= 0.0 ? 0u : (mi >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)mi);
    e.assign(B, {});
    q.assign(B, {});
    chi.assign(B, std::vector<std::vector<float>>(nchi));
    const uint32_t lvl = P.sample_limiting;
    samples.assign(lvl == 0 ? B : (lvl == 1 ? B / 2 : 1), {});
    fixed = P.rng_mode == TAMCO_RNG_PHILOX && P.energy_fixed;
    group_sum.assign((size_t)n + 4, 0.0f);  // at most one RNG group per site
    self_J.assign(n, 0.0f);
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t e = I.row_ptr[i]; e < I.row_ptr[i + 1]; ++e)
        if (I.col[e] == i) self_J[i] = I.J[e];
  }

  uint32_t stream_index(uint32_t c, uint32_t b) const { return P.threaded ? c * B + b : 0; }

  void setup_streams_and_init() {
    const bool feed = P.rng_mode == TAMCO_RNG_FEED;
    chains.clear();
    for (uint32_t c = 0; c < C; ++c) chains.emplace_back(B);
    WordStream base;
    if (!feed) base.x.seed(P.seed);
    // initial states: src/pt.rs:506-519 (chain-major, then beta, then site) or caller-supplied
    for (uint32_t c = 0; c < C; ++c)
      for (uint32_t b = 0; b < B; ++b) {
        St &x = chains[c].st[b];
        x.s.resize(n);
        x.id = c * B + b;
        x.init = false;
        if (P.init_states) {
          std::memcpy(x.s.data(), P.init_states + ((size_t)c * B + b) * n, n);
        } else if (P.rng_mode == TAMCO_RNG_PHILOX) {
          for (uint32_t i = 0; i < n; ++i)
            x.s[i] = (philox_word(P.seed, 0, i >> 2, x.id, TAG_INIT, i) >> 31) ? 1 : -1;
        } else if (!feed) {
          for (uint32_t i = 0; i < n; ++i) x.s[i] = base.next_spin();
        } else {
          throw std::invalid_argument("feed mode needs init_states");
        }
        energy(I, x);  // rand_ising_state computes the energy (src/ising.rs:175)
        if (P.init_energies) {  // a resumed IsingState keeps its cached energy
          const float e = P.init_energies[c * B + b];
          if (fixed) x.drift = (long long)std::rint(((double)e - (double)x.energy) * 4294967296.0);
          x.energy = e;
        }
      }
    for (uint32_t c = 0; c < C; ++c) {  // the rest of a resumed PTState (tamc-core/src/pt.rs:18-25)
      Chain &ch = chains[c];
      for (uint32_t b = 0; b < B; ++b) {
        if (P.init_nacc) ch.nacc[b] = P.init_nacc[c * B + b];
        if (P.init_tags) ch.tags[b] = P.init_tags[c * B + b];
        if (P.init_hist) {
          ch.hist[2 * b] = P.init_hist[2 * ((size_t)c * B + b)];
          ch.hist[2 * b + 1] = P.init_hist[2 * ((size_t)c * B + b) + 1];
        }
      }
      if (P.init_rt) ch.rt = P.init_rt[c];
    }
    // streams
    if (feed) {
      const uint32_t ns = P.threaded ? C * B : 1;
      if (P.num_streams < ns) throw std::invalid_argument("not enough feed streams");
      streams.resize(ns);
      for (uint32_t k = 0; k < ns; ++k) {
        streams[k].is_feed = true;
        streams[k].w = P.stream_words[k];
        streams[k].len = P.stream_lens[k];
        streams[k].pos = 0;
      }
    } else if (P.rng_mode == TAMCO_RNG_XOSHIRO) {
      if (P.threaded) {  // src/pt.rs:365-373: rng_vec[c][b] = clone, then jump
        streams.resize((size_t)C * B);
        for (uint32_t c = 0; c < C; ++c)
          for (uint32_t b = 0; b < B; ++b) {
            streams[c * B + b] = base;
            base.x.jump();
          }
      } else {
        streams.assign(1, base);
      }
    }
  }

  void apply_icm(uint32_t t) {  // src/pt.rs:521-540
    if (!P.icm) return;
    for (uint32_t p = 0; p < npairs; ++p)
      for (uint32_t b = lo; b < B; ++b) {
        St &r1 = chains[2 * p].st[b];
        St &r2 = chains[2 * p + 1].st[b];
        uint32_t cnt = 0;
        for (uint32_t i = 0; i < n; ++i) cnt += ((int)r1.s[i] * (int)r2.s[i] < 0);
        if (cnt == 0) continue;  // choose() on an empty slice draws nothing
        uint32_t k;
        if (P.rng_mode == TAMCO_RNG_PHILOX) {
          k = (uint32_t)(((uint64_t)philox_word(P.seed, t, b, p, TAG_ICM, 0) * cnt) >> 32);
        } else {
          k = streams[stream_index(0, 0)].gen_index(cnt);
        }
        if (fixed) {
          const std::vector<int8_t> before = r1.s;
          std::vector<uint8_t> cl;
          houdayer_k(I, r1.s.data(), r2.s.data(), k, &cl);
          const long long d = cluster_delta_fixed(I, before.data(), cl);
          r1.energy = add_fix(r1.energy, d - r1.drift);
          r2.energy = add_fix(r2.energy, -d - r2.drift);
          r1.drift = 0;
          r2.drift = 0;
        } else {
          houdayer_k(I, r1.s.data(), r2.s.data(), k, nullptr);
          r1.init = false;
          r2.init = false;
        }
        ++icm_moves;
      }
  }

  void advance_chain(uint32_t t, uint32_t c) {
    Chain &ch = chains[c];
    std::vector<float> E(B);
    for (uint32_t b = 0; b < B; ++b) E[b] = energy(I, ch.st[b]);
    const int variant = P.threaded ? 1 : 0;
    auto do_swap = [&](uint32_t j) { ch.swap_states(j); };
    if (P.rng_mode == TAMCO_RNG_PHILOX) {
      exchange_pass(B, dbeta.data(), E.data(), variant,
                    [&](uint32_t j) { return u24(philox_word(P.seed, t, j >> 2, c, TAG_EXCHANGE, j)); }, nullptr,
                    do_swap);
    } else {
      WordStream &ws = streams[stream_index(c, 0)];
      exchange_pass(B, dbeta.data(), E.data(), variant, [&](uint32_t) { return ws.next_f32(); }, nullptr, do_swap);
    }
    ch.update_round_trips();
    for (uint32_t b = 0; b < B; ++b) {
      St &x = ch.st[b];
      float de;
      if (fixed) {
        const long long acc = sweep_fixed(I, x.s.data(), beta[b], PhiloxGroupDraw{P.seed, t, x.id, &groups},
                                          P.sweep_order == 0 ? groups.order.data() : nullptr, groups.gid.data(),
                                          group_sum, self_J.data(), &x.drift);
        x.energy = add_fix(x.energy, acc);
        attempts += n;
        continue;
      } else if (P.rng_mode == TAMCO_RNG_PHILOX) {
        de = sweep(I, x.s.data(), beta[b], PhiloxGroupDraw{P.seed, t, x.id, &groups},
                   P.sweep_order == 0 ? groups.order.data() : nullptr);
      } else {
        WordStream &ws = streams[stream_index(c, b)];
        de = sweep(I, x.s.data(), beta[b], [&](uint32_t) { return ws.next_f32(); });
      }
      x.energy += de;
      attempts += n;
    }
  }

  void measure() {  // src/pt.rs:167-201
    std::vector<int8_t> ov(n);
    for (uint32_t b = 0; b < B; ++b) {
      for (uint32_t c = 0; c < C; ++c) e[b].push_back(energy(I, chains[c].st[b]));
      for (uint32_t p = 0; p < npairs; ++p) {
        const St &a = chains[2 * p].st[b];
        const St &bb = chains[2 * p + 1].st[b];
        if (nchi > 0) {
          for (uint32_t i = 0; i < n; ++i) ov[i] = (int8_t)(a.s[i] * bb.s[i]);
          for (uint32_t k = 0; k < nchi; ++k) {
            double x = 0.0;
            for (uint32_t i = 0; i < n; ++i) x += I.chi[k][i] * (double)ov[i];
            chi[b][k].push_back((float)x);
          }
        }
        int64_t qq = 0;
        for (uint32_t i = 0; i < n; ++i) qq += (int64_t)(a.s[i] * bb.s[i]);
        q[b].push_back((int32_t)qq);
      }
    }
    ++n_meas;
  }

  void sample_states() {  // src/pt.rs:202-226
    const uint32_t nb = n / 8 + (n % 8 ? 1 : 0);
    std::vector<uint8_t> buf(nb);
    auto push = [&](uint32_t slot, const St &x) {
      pack_bytes(x.s.data(), n, buf.data());
      samples[slot].push_back(buf);
    };
    if (P.sample_limiting == 0) {
      for (uint32_t b = 0; b < B; ++b)
        for (uint32_t c = 0; c < C; ++c) push(b, chains[c].st[b]);
    } else if (P.sample_limiting == 1) {
      for (uint32_t i = 0; i < B / 2; ++i)
        for (uint32_t c = 0; c < C; ++c) push(i, chains[c].st[B - i - 1]);
    } else {
      for (uint32_t c = 0; c < C; ++c) push(0, chains[c].st[B - 1]);
    }
    ++n_samp_events;
  }

  void apply_measurements(uint32_t i) {  // src/pt.rs:542-583
    if (i < meas_init) return;
    const uint32_t stp = i - meas_init;
    if (P.sample_enabled && (stp % P.sample == 0 || i == T - 1)) measure();
    if (P.sample_states_enabled && (stp % P.sample_states == 0 || i == T - 1)) sample_states();
    uint32_t best_c = 0, best_b = 0;
    float best = 0.f;
    for (uint32_t c = 0; c < C; ++c) {
      uint32_t i1 = 0;
      float e1 = energy(I, chains[c].st[0]);
      for (uint32_t b = 1; b < B; ++b) {
        const float x = energy(I, chains[c].st[b]);
        if (x < e1) {
          e1 = x;
          i1 = b;
        }
      }
      if (c == 0 || e1 < best) {
        best = e1;
        best_c = c;
        best_b = i1;
      }
    }
    if (!have_min || best < min_e) {
      have_min = true;
      min_e = best;
      const uint32_t nw = n / 64 + (n % 64 ? 1 : 0);
      const size_t o = gs_states.size();
      gs_states.resize(o + nw);
      pack_u64(chains[best_c].st[best_b].s.data(), n, gs_states.data() + o);
      gs_e.push_back(best);
      gs_steps.push_back(i);
    }
  }

  void run() {  // src/pt.rs:446-450 (threaded) and 493-497 (serial)
    setup_streams_and_init();
    for (uint32_t t = 0; t < T; ++t) {
      apply_icm(t);
      for (uint32_t c = 0; c < C; ++c) advance_chain(t, c);
      apply_measurements(t);
    }
  }
};

template <class T>
T *dup(const std::vector<T> &v) {
  T *p = (T *)std::malloc(std::max<size_t>(1, v.size() * sizeof(T)));
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

}  // namespace

struct tamco_instance {
  Inst I;
};

extern "C" {

void tamco_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }
void tamco_xoshiro_seed(uint64_t seed, uint64_t st[4]) {
  Xoshiro x;
  x.seed(seed);
  std::memcpy(st, x.s, sizeof x.s);
}
uint64_t tamco_xoshiro_next(uint64_t st[4]) {
  Xoshiro x;
  std::memcpy(x.s, st, sizeof x.s);
  const uint64_t r = x.next_u64();
  std::memcpy(st, x.s, sizeof x.s);
  return r;
}
void tamco_xoshiro_jump(uint64_t st[4]) {
  Xoshiro x;
  std::memcpy(x.s, st, sizeof x.s);
  x.jump();
  std::memcpy(st, x.s, sizeof x.s);
}
uint32_t tamco_gen_index(uint32_t len, const uint32_t *words, uint64_t nwords, uint64_t *cursor) {
  WordStream ws;
  ws.is_feed = true;
  ws.w = words;
  ws.len = nwords;
  ws.pos = *cursor;
  uint32_t r = 0xFFFFFFFFu;
  try {
    r = ws.gen_index(len);
  } catch (const StreamExhausted &) {
  }
  *cursor = ws.pos;
  return r;
}

tamco_instance *tamco_instance_from_file(const char *path, int qubo) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return nullptr;
  AdjList adj;
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    uint32_t i, j;
    double K;
    if (parse_line(line, i, j, K)) adj_insert(adj, i, j, K);
  }
  Inst *I = build_instance(adj, qubo != 0);
  tamco_instance *r = new tamco_instance{std::move(*I)};
  delete I;
  return r;
}

tamco_instance *tamco_instance_from_lines(uint64_t nlines, const uint32_t *i, const uint32_t *j, const double *K,
                                          int qubo) {
  AdjList adj;
  for (uint64_t k = 0; k < nlines; ++k) adj_insert(adj, i[k], j[k], K[k]);
  Inst *I = build_instance(adj, qubo != 0);
  tamco_instance *r = new tamco_instance{std::move(*I)};
  delete I;
  return r;
}

void tamco_instance_free(tamco_instance *p) { delete p; }
uint32_t tamco_instance_n(const tamco_instance *p) { return p->I.n; }
uint32_t tamco_instance_nnz(const tamco_instance *p) { return (uint32_t)p->I.col.size(); }
float tamco_instance_offset(const tamco_instance *p) { return p->I.offset; }
void tamco_instance_csr(const tamco_instance *p, uint32_t *row_ptr, uint32_t *col, float *J, float *bias) {
  const Inst &I = p->I;
  std::memcpy(row_ptr, I.row_ptr.data(), (I.n + 1) * sizeof(uint32_t));
  if (!I.col.empty()) {
    std::memcpy(col, I.col.data(), I.col.size() * sizeof(uint32_t));
    std::memcpy(J, I.J.data(), I.J.size() * sizeof(float));
  }
  if (I.n) std::memcpy(bias, I.bias.data(), I.n * sizeof(float));
}
uint32_t tamco_instance_graph_nnz(const tamco_instance *p) { return (uint32_t)p->I.g_col.size(); }
uint32_t tamco_instance_graph_nodes(const tamco_instance *p) { return p->I.g_nodes; }
void tamco_instance_graph(const tamco_instance *p, uint32_t *row_ptr, uint32_t *col) {
  std::memcpy(row_ptr, p->I.g_row_ptr.data(), p->I.g_row_ptr.size() * sizeof(uint32_t));
  if (!p->I.g_col.empty()) std::memcpy(col, p->I.g_col.data(), p->I.g_col.size() * sizeof(uint32_t));
}
void tamco_instance_set_suscept(tamco_instance *p, uint32_t nchi, const double *w) {
  p->I.chi.assign(nchi, std::vector<double>(p->I.n));
  for (uint32_t k = 0; k < nchi; ++k) std::memcpy(p->I.chi[k].data(), w + (size_t)k * p->I.n, p->I.n * sizeof(double));
}

float tamco_energy_ref(const tamco_instance *p, const int8_t *s) { return energy_ref(p->I, s); }
float tamco_delta_energy(const tamco_instance *p, const int8_t *s, uint32_t i) { return delta_energy(p->I, s, i); }
float tamco_sweep_feed(const tamco_instance *p, int8_t *s, float beta, const uint32_t *words, uint64_t nwords,
                       uint64_t *cursor) {
  WordStream ws;
  ws.is_feed = true;
  ws.w = words;
  ws.len = nwords;
  ws.pos = *cursor;
  float de = sweep(p->I, s, beta, [&](uint32_t) { return ws.next_f32(); });
  *cursor = ws.pos;
  return de;
}
float tamco_sweep_philox(const tamco_instance *p, int8_t *s, float beta, uint64_t seed, uint32_t t, uint32_t rho) {
  PhiloxSweepDraw d{seed, t, rho};
  return sweep(p->I, s, beta, d);
}
int64_t tamco_overlap(const int8_t *a, const int8_t *b, uint32_t n) {  // src/ising.rs:159-165
  int64_t q = 0;
  for (uint32_t i = 0; i < n; ++i) q += (int64_t)(a[i] * b[i]);
  return q;
}
double tamco_suscept(const tamco_instance *p, uint32_t k, const int8_t *ov) {  // src/ising.rs:427-433
  double x = 0.0;
  for (uint32_t i = 0; i < p->I.n; ++i) x += p->I.chi[k][i] * (double)ov[i];
  return x;
}
void tamco_pack_bytes(const int8_t *s, uint32_t n, uint8_t *out) { pack_bytes(s, n, out); }
void tamco_pack_u64(const int8_t *s, uint32_t n, uint64_t *out) { pack_u64(s, n, out); }
int64_t tamco_houdayer_k(const tamco_instance *p, int8_t *s1, int8_t *s2, uint32_t k, uint8_t *cluster_out) {
  std::vector<uint8_t> cl;
  int64_t r;
  try {
    r = houdayer_k(p->I, s1, s2, k, cluster_out ? &cl : nullptr);
  } catch (const SeedOutsideGraph &) {
    return -1;
  }
  if (cluster_out) {
    if (r > 0) std::memcpy(cluster_out, cl.data(), p->I.n);
    else std::memset(cluster_out, 0, p->I.n);
  }
  return r;
}
uint32_t tamco_num_negative_overlap(const int8_t *s1, const int8_t *s2, uint32_t n) {
  uint32_t c = 0;
  for (uint32_t i = 0; i < n; ++i) c += ((int)s1[i] * (int)s2[i] < 0);
  return c;
}
void tamco_exchange(uint32_t B, const float *betas, const float *energies, uint32_t *perm, int variant,
                    const uint32_t *words, uint64_t nwords, uint64_t *cursor, uint8_t *accept_out, uint32_t *nacc) {
  std::vector<float> dbeta;
  for (uint32_t j = 0; j + 1 < B; ++j) dbeta.push_back(betas[j + 1] - betas[j]);
  WordStream ws;
  ws.is_feed = true;
  ws.w = words;
  ws.len = nwords;
  ws.pos = *cursor;
  std::vector<uint8_t> acc(B > 0 ? B - 1 : 0, 0);
  try {
    exchange_pass(B, dbeta.data(), energies, variant, [&](uint32_t) { return ws.next_f32(); }, &acc,
                  [&](uint32_t j) {
                    std::swap(perm[j], perm[j + 1]);
                    nacc[j] += 1;
                  });
  } catch (const StreamExhausted &) {
  }
  *cursor = ws.pos;
  if (accept_out && B > 1) std::memcpy(accept_out, acc.data(), B - 1);
}
void tamco_chain_step(uint32_t B, const uint8_t *swaps, uint8_t *tags, uint32_t *nacc, uint32_t *hist, uint32_t *rt) {
  // one exchange pass with the given decisions (swap_states, tamc-core/src/pt.rs:49-53) and the
  // round-trip update that follows it (tamc-core/src/pt.rs:54-84)
  Chain ch(B);
  std::memcpy(ch.tags.data(), tags, B);
  std::memcpy(ch.nacc.data(), nacc, B * 4u);
  std::memcpy(ch.hist.data(), hist, 2u * B * 4u);
  ch.rt = *rt;
  for (uint32_t j = 0; j + 1 < B; ++j)
    if (swaps[j]) ch.swap_states(j);
  ch.update_round_trips();
  std::memcpy(tags, ch.tags.data(), B);
  std::memcpy(nacc, ch.nacc.data(), B * 4u);
  std::memcpy(hist, ch.hist.data(), 2u * B * 4u);
  *rt = ch.rt;
}
void tamco_beta_geometric(float beta_min, float beta_max, uint32_t num_beta, float *out) {
  // tamc-core/src/sa.rs:13-22 in f64, then `as f32` (src/pt.rs:249-250)
  const double l0 = std::log((double)beta_min), lf = std::log((double)beta_max);
  const double d = (lf - l0) / (double)(num_beta - 1);
  for (uint32_t i = 0; i < num_beta; ++i) out[i] = (float)std::exp(l0 + d * (double)i);
}
float tamco_expf(float x) { return expf(x); }

tamco_pt_result *tamco_pt_run(const tamco_instance *inst, const tamco_pt_params *P) {
  tamco_pt_result *R = (tamco_pt_result *)std::calloc(1, sizeof(tamco_pt_result));
  if (!inst || !P || P->num_betas < 1 || P->num_chains < 1 || (P->sample_enabled && P->sample == 0) ||
      (P->sample_states_enabled && P->sample_states == 0) || P->lo_beta_idx > P->num_betas) {
    R->error = 3;
    return R;
  }
  PtRun run(inst->I, *P);
  try {
    run.run();
  } catch (const StreamExhausted &) {
    R->error = 1;
  } catch (const SeedOutsideGraph &) {
    R->error = 2;
  } catch (const std::invalid_argument &) {
    R->error = 3;
    return R;
  }
  const uint32_t n = run.n, B = run.B, C = run.C;
  R->n = n;
  R->B = B;
  R->C = C;
  R->T = run.T;
  R->meas_init = run.meas_init;
  R->nchi = run.nchi;
  R->npairs = run.npairs;
  R->n_meas = run.n_meas;
  std::vector<float> e, chi;
  std::vector<int32_t> q;
  for (uint32_t b = 0; b < B; ++b) {
    e.insert(e.end(), run.e[b].begin(), run.e[b].end());
    q.insert(q.end(), run.q[b].begin(), run.q[b].end());
    for (uint32_t k = 0; k < run.nchi; ++k) chi.insert(chi.end(), run.chi[b][k].begin(), run.chi[b][k].end());
  }
  R->e = dup(e);
  R->q = dup(q);
  R->suscept = dup(chi);
  R->n_samp_slots = (uint32_t)run.samples.size();
  R->n_samp_events = run.n_samp_events;
  R->nbytes = n / 8 + (n % 8 ? 1 : 0);
  std::vector<uint8_t> smp;
  for (auto &slot : run.samples)
    for (auto &v : slot) smp.insert(smp.end(), v.begin(), v.end());
  R->samples = dup(smp);
  R->n_gs = (uint32_t)run.gs_e.size();
  R->gs_words = n / 64 + (n % 64 ? 1 : 0);
  R->gs_steps = dup(run.gs_steps);
  R->gs_energies = dup(run.gs_e);
  R->gs_states = dup(run.gs_states);
  std::vector<uint32_t> accs(B, 0), nacc, rts, hist, ids;
  std::vector<int8_t> fin;
  std::vector<float> fe;
  std::vector<uint8_t> tags;
  for (uint32_t c = 0; c < C && !run.chains.empty(); ++c) {
    Chain &ch = run.chains[c];
    for (uint32_t b = 0; b < B; ++b) {
      accs[b] += ch.nacc[b];
      fin.insert(fin.end(), ch.st[b].s.begin(), ch.st[b].s.end());
      fe.push_back(ch.st[b].energy);
      ids.push_back(ch.st[b].id);
    }
    nacc.insert(nacc.end(), ch.nacc.begin(), ch.nacc.end());
    rts.push_back(ch.rt);
    tags.insert(tags.end(), ch.tags.begin(), ch.tags.end());
    hist.insert(hist.end(), ch.hist.begin(), ch.hist.end());
  }
  R->acceptance_counts = dup(accs);
  R->final_states = dup(fin);
  R->final_energies = dup(fe);
  R->final_ids = dup(ids);
  R->num_acceptances = dup(nacc);
  R->round_trips = dup(rts);
  R->tags = dup(tags);
  R->diffusion_hist = dup(hist);
  R->icm_moves = run.icm_moves;
  R->attempts = run.attempts;
  return R;
}

void tamco_pt_result_free(tamco_pt_result *R) {
  if (!R) return;
  void *ptrs[] = {R->e, R->q, R->suscept, R->samples, R->gs_steps, R->gs_energies, R->gs_states,
                  R->acceptance_counts, R->final_states, R->final_energies, R->final_ids, R->num_acceptances,
                  R->round_trips, R->tags, R->diffusion_hist};
  for (void *p : ptrs) std::free(p);
  std::free(R);
}

// src/sa.rs:110-210 and tamc-core/src/sa.rs:27-48
tamco_sa_result *tamco_sa_run(const tamco_instance *inst, const tamco_sa_params *P) {
  tamco_sa_result *R = (tamco_sa_result *)std::calloc(1, sizeof(tamco_sa_result));
  if (!inst || !P || P->num_betas < 1 || (P->num_replicas > 1 && P->threads == 0)) {
    R->error = 3;
    return R;
  }
  const Inst &I = inst->I;
  const uint32_t n = I.n;
  for (uint32_t k = 1; k < P->num_betas; ++k)
    if (!(P->betas[k] - P->betas[k - 1] >= 0.0f)) {  // src/sa.rs:101-104
      R->error = 3;
      return R;
    }
  // partitions: run() for num_replicas <= 1, run_parallel() otherwise (src/sa.rs:219-224)
  const bool par = P->num_replicas > 1;
  const uint32_t parts = par ? P->threads : 1;
  const uint32_t per = par ? P->num_replicas / P->threads : P->num_replicas;
  const uint32_t Rn = parts * per;
  std::vector<int8_t> S((size_t)Rn * n);
  std::vector<WordStream> ws(parts);
  if (P->rng_mode == TAMCO_RNG_XOSHIRO) {
    WordStream base;
    base.x.seed(P->seed);
    for (uint32_t t = 0; t < parts; ++t) {
      ws[t] = base;
      if (par) base.x.jump();
    }
  }
  for (uint32_t r = 0; r < Rn; ++r) {
    int8_t *s = S.data() + (size_t)r * n;
    if (P->init_states) std::memcpy(s, P->init_states + (size_t)r * n, n);
    else if (P->rng_mode == TAMCO_RNG_PHILOX)
      for (uint32_t i = 0; i < n; ++i) s[i] = (philox_word(P->seed, 0, i >> 2, r, TAG_INIT, i) >> 31) ? 1 : -1;
    else
      for (uint32_t i = 0; i < n; ++i) s[i] = ws[r / std::max(per, 1u)].next_spin();
  }
  for (uint32_t t = 0; t < parts; ++t)
    for (uint32_t k = 0; k < P->num_betas; ++k)
      for (uint32_t rr = 0; rr < per; ++rr) {
        const uint32_t r = t * per + rr;
        int8_t *s = S.data() + (size_t)r * n;
        if (P->rng_mode == TAMCO_RNG_PHILOX) {
          PhiloxSweepDraw d{P->seed, k, r};
          sweep(I, s, P->betas[k], d);
        } else {
          sweep(I, s, P->betas[k], [&](uint32_t) { return ws[t].next_f32(); });
        }
        R->attempts += n;
      }
  std::vector<float> en(Rn);
  float mn = 0.f;
  for (uint32_t r = 0; r < Rn; ++r) {
    en[r] = energy_ref(I, S.data() + (size_t)r * n);
    if (r == 0 || en[r] < mn) mn = en[r];
  }
  R->n = n;
  R->R = Rn;
  R->min_energy = mn;
  R->energies = dup(en);
  R->final_states = dup(S);
  return R;
}

void tamco_sa_result_free(tamco_sa_result *R) {
  if (!R) return;
  std::free(R->energies);
  std::free(R->final_states);
  std::free(R);
}

}  // extern "C"
