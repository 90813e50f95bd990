// gla.hip -- GLA over the batched PT-ICM path (src/gla.rs): partitions, the partition graph, the
// sub-solves in one launch, and the joined ground state with its boundary energies.  Host code.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "runtime_common.hpp"

using namespace tamc;

namespace {

constexpr uint32_t kNoPart = 0xFFFFFFFFu;

template <class T>
T *host_copy(const std::vector<T> &v) {
  T *p = (T *)malloc(std::max<size_t>(1, v.size()) * sizeof(T));
  if (p && !v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

// read_u32_lines (src/util.rs:142-172): many1(delimited(space0, u32, multispace0)) per line; a
// line whose first token is not an index is skipped.
bool parse_u32_line(const std::string &line, std::vector<uint32_t> &out, std::string &err) {
  size_t p = 0;
  out.clear();
  for (;;) {
    size_t q = p;
    while (q < line.size() && (line[q] == ' ' || line[q] == '\t')) ++q;
    if (q >= line.size() || line[q] < '0' || line[q] > '9') break;
    uint64_t x = 0;
    while (q < line.size() && line[q] >= '0' && line[q] <= '9') {
      x = x * 10u + (uint64_t)(line[q] - '0');
      if (x > 0xFFFFFFFFull) {
        err = "partition index does not fit in u32: " + line;
        return false;
      }
      ++q;
    }
    out.push_back((uint32_t)x);
    while (q < line.size() && (line[q] == ' ' || line[q] == '\t' || line[q] == '\r' || line[q] == '\n')) ++q;
    p = q;
  }
  return true;
}

tamc_status make_partition(const tamc_instance *inst, const std::vector<std::vector<uint32_t>> &parts,
                           tamc_partition **out) {
  const uint32_t n = inst->n;
  std::vector<uint8_t> seen(n, 0u);
  for (size_t ip = 0; ip < parts.size(); ++ip)
    for (uint32_t i : parts[ip]) {
      if (i >= n)
        return fail(TAMC_E_INVALID_ARG, "Partitition must be specified for instance size " + std::to_string(n) +
                                            ". Found index " + std::to_string(i));
      if (seen[i])
        return fail(TAMC_E_INVALID_ARG,
                    "Index " + std::to_string(i) + " specified more than once in partition " + std::to_string(ip));
      seen[i] = 1;
    }
  std::vector<uint32_t> offs(1, 0u), sites;
  for (const auto &p : parts) {
    std::vector<uint32_t> s = p;
    std::sort(s.begin(), s.end());
    sites.insert(sites.end(), s.begin(), s.end());
    offs.push_back((uint32_t)sites.size());
  }
  tamc_partition *P = (tamc_partition *)calloc(1, sizeof(tamc_partition));
  P->num_partitions = (uint32_t)parts.size();
  P->offsets = host_copy(offs);
  P->sites = host_copy(sites);
  *out = P;
  return TAMC_OK;
}

}  // namespace

extern "C" {

tamc_status tamc_partition_from_file(const tamc_instance *inst, const char *path, tamc_partition **out) {
  clear_error();
  if (!inst || !path || !out) return fail(TAMC_E_INVALID_ARG, "null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(TAMC_E_IO, std::string("Unable to open partition file ") + path);
  std::vector<std::vector<uint32_t>> parts;
  std::string line, err;
  std::vector<uint32_t> v;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (!parse_u32_line(line, v, err)) return fail(TAMC_E_INVALID_ARG, err);
    if (!v.empty()) parts.push_back(v);
  }
  return make_partition(inst, parts, out);
}

tamc_status tamc_partition_create(const tamc_instance *inst, uint32_t num_partitions, const uint32_t *offsets,
                                  const uint32_t *sites, tamc_partition **out) {
  clear_error();
  if (!inst || !out || (num_partitions && !offsets)) return fail(TAMC_E_INVALID_ARG, "null argument");
  std::vector<std::vector<uint32_t>> parts(num_partitions);
  for (uint32_t p = 0; p < num_partitions; ++p) {
    if (offsets[p + 1] < offsets[p]) return fail(TAMC_E_INVALID_ARG, "partition offsets must not decrease");
    if (offsets[p + 1] > offsets[p] && !sites) return fail(TAMC_E_INVALID_ARG, "null argument");
    parts[p].assign(sites + offsets[p], sites + offsets[p + 1]);
  }
  return make_partition(inst, parts, out);
}

void tamc_partition_free(tamc_partition *p) {
  if (!p) return;
  free(p->offsets);
  free(p->sites);
  free(p);
}

tamc_status tamc_gla_run(const tamc_instance *inst, const tamc_partition *parts, const tamc_pt_params *params,
                         tamc_gla_result **out) {
  clear_error();
  if (!inst || !parts || !params || !out) return fail(TAMC_E_INVALID_ARG, "null argument");
  const uint32_t n = inst->n, np = parts->num_partitions;
  // site -> partition; every site must be in one (the reference maps every node of the energy
  // graph through the partition map, src/gla.rs:92-96)
  std::vector<uint32_t> part_of(n, kNoPart);
  for (uint32_t p = 0; p < np; ++p)
    for (uint32_t k = parts->offsets[p]; k < parts->offsets[p + 1]; ++k) {
      const uint32_t i = parts->sites[k];
      if (i >= n || part_of[i] != kNoPart) return fail(TAMC_E_INVALID_ARG, "malformed partition");
      part_of[i] = p;
    }
  for (uint32_t i = 0; i < n; ++i)
    if (part_of[i] == kNoPart) return fail(TAMC_E_INVALID_ARG, "site " + std::to_string(i) + " belongs to no partition");
  // partition graph (src/gla.rs:73-120): the energy graph's edges (coupling pattern, i < j, in
  // row order) between different partitions, grouped per directed (source, target) pair in order
  // of first appearance
  struct Edge {
    uint32_t i, j;
    float K;
  };
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> pair_index;
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  std::vector<std::vector<Edge>> edges;
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t e = inst->row_ptr[i]; e < inst->row_ptr[i + 1]; ++e) {
      const uint32_t j = inst->col[e];
      if (j <= i) continue;  // the coupling pattern has no self entries; both directions are present
      const uint32_t ps = part_of[i], pt = part_of[j];
      if (ps == pt) continue;
      auto it = pair_index.find({ps, pt});
      if (it == pair_index.end()) {
        it = pair_index.emplace(std::make_pair(ps, pt), (uint32_t)pairs.size()).first;
        pairs.emplace_back(ps, pt);
        edges.emplace_back();
      }
      edges[it->second].push_back(Edge{i, j, inst->J[e]});
    }
  // sub-instances and their solves, one batched launch
  std::vector<std::unique_ptr<tamc_instance, void (*)(tamc_instance *)>> subs;
  std::vector<const tamc_instance *> sub_ptrs;
  std::vector<tamc_pt_params> sub_params(np, *params);
  for (uint32_t p = 0; p < np; ++p) {
    tamc_instance *s = nullptr;
    const tamc_status st = tamc_instance_induced_subgraph(inst, parts->sites + parts->offsets[p],
                                                          parts->offsets[p + 1] - parts->offsets[p], &s);
    if (st) return st;
    subs.emplace_back(s, tamc_instance_destroy);
    sub_ptrs.push_back(s);
    sub_params[p].seed = params->seed + p;
  }
  std::vector<tamc_pt_result *> res(np, nullptr);
  tamc_batch_stats stats{};
  tamc_batch_options opt{};
  tamc_status st = np ? tamc_pt_run_batch(np, sub_ptrs.data(), sub_params.data(), &opt, res.data(), &stats) : TAMC_OK;
  if (st) return st;
  struct ResultsGuard {
    std::vector<tamc_pt_result *> &r;
    ~ResultsGuard() {
      for (auto *x : r) tamc_pt_result_free(x);
    }
  } rg{res};
  // each sub-solve's last ground state (src/gla.rs:165-169), joined
  std::vector<float> pe(np);
  std::vector<int8_t> joined(n, 1);
  for (uint32_t p = 0; p < np; ++p) {
    const tamc_pt_result *r = res[p];
    if (r->n_gs == 0) return fail(TAMC_E_INVALID_ARG, "GLA: a sub-solve recorded no ground state (no sweep after warm-up)");
    pe[p] = r->gs_energies[r->n_gs - 1];
    const uint32_t words = (r->n + 63u) / 64u;
    const uint64_t *w = r->gs_states + (size_t)(r->n_gs - 1) * words;
    for (uint32_t k = parts->offsets[p]; k < parts->offsets[p + 1]; ++k) {
      const uint32_t a = k - parts->offsets[p];
      joined[parts->sites[k]] = ((w[a / 64u] >> (a % 64u)) & 1u) ? -1 : 1;
    }
  }
  float bulk = 0.f;
  for (float e : pe) bulk += e;
  // boundary energies (src/gla.rs:175-190): bnd += K * (s_i s_j) in f32, edge by edge
  std::vector<uint32_t> bp;
  std::vector<float> be;
  float total_bnd = 0.f;
  for (size_t k = 0; k < pairs.size(); ++k) {
    float bnd = 0.f;
    for (const Edge &e : edges[k]) bnd += e.K * (float)(joined[e.i] * joined[e.j]);
    total_bnd += bnd;
    bp.push_back(pairs[k].first);
    bp.push_back(pairs[k].second);
    be.push_back(bnd);
  }
  tamc_gla_result *R = (tamc_gla_result *)calloc(1, sizeof(tamc_gla_result));
  R->instance_size = n;
  R->num_partitions = np;
  R->num_boundaries = (uint32_t)pairs.size();
  R->partition_energies = host_copy(pe);
  R->boundary_pairs = host_copy(bp);
  R->boundary_energies = host_copy(be);
  R->bulk_energy = bulk;
  R->boundary_energy = total_bnd;
  R->total_energy = bulk + total_bnd;
  R->joined_state = host_copy(joined);
  R->attempts = stats.attempts;
  R->device_ms = stats.device_ms;
  *out = R;
  return TAMC_OK;
}

void tamc_gla_result_free(tamc_gla_result *r) {
  if (!r) return;
  free(r->partition_energies);
  free(r->boundary_pairs);
  free(r->boundary_energies);
  free(r->joined_state);
  free(r);
}

}  // extern "C"
