// writers.hip -- bincode-1 serialisation of the reference's output structures (little-endian,
// fixed-width integers, u64 length prefixes, fields in declaration order).
//   PtIcmThermalSamples  src/pt.rs:116-125 (written at src/pt.rs:652)
//   AnnealState          src/sa.rs:67-89  (written at src/sa.rs:255-270)
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "runtime_common.hpp"

using namespace tamc;

namespace {

struct Out {
  std::vector<uint8_t> b;
  void u8(uint8_t v) { b.push_back(v); }
  void u32(uint32_t v) { raw(&v, 4); }
  void u64(uint64_t v) { raw(&v, 8); }
  void f32(float v) { raw(&v, 4); }
  void i32(int32_t v) { raw(&v, 4); }
  void raw(const void *p, size_t n) {
    const uint8_t *c = (const uint8_t *)p;
    b.insert(b.end(), c, c + n);
  }
};

tamc_status write_file(const char *path, const std::vector<uint8_t> &b) {
  FILE *f = fopen(path, "wb");
  if (!f) return fail(TAMC_E_IO, std::string("Failed to create sample output file ") + path);
  const size_t w = b.empty() ? 0 : fwrite(b.data(), 1, b.size(), f);
  fclose(f);
  if (w != b.size()) return fail(TAMC_E_IO, std::string("short write to ") + path);
  return TAMC_OK;
}

}  // namespace

extern "C" {

tamc_status tamc_write_thermal_samples(const tamc_pt_result *r, const float *betas, uint8_t compression_level,
                                       const char *path) {
  clear_error();
  if (!r || !betas || !path) return fail(TAMC_E_INVALID_ARG, "null argument");
  Out o;
  const uint32_t B = r->B, C = r->C, nm = r->n_meas, np = r->npairs;
  o.u8(compression_level);
  o.u64(r->n);
  o.u64(B);
  for (uint32_t b = 0; b < B; ++b) o.f32(betas[b]);
  o.u64(r->n_samp_slots);
  const uint64_t per = (uint64_t)r->n_samp_events * C;
  for (uint32_t s = 0; s < r->n_samp_slots; ++s) {
    o.u64(per);
    for (uint64_t k = 0; k < per; ++k) {
      o.u64(r->nbytes);
      o.raw(r->samples + ((uint64_t)s * per + k) * r->nbytes, r->nbytes);
    }
  }
  o.u64(B);
  for (uint32_t b = 0; b < B; ++b) {
    o.u64((uint64_t)nm * C);
    for (uint64_t k = 0; k < (uint64_t)nm * C; ++k) o.f32(r->e[(uint64_t)b * nm * C + k]);
  }
  o.u64(B);
  for (uint32_t b = 0; b < B; ++b) {
    o.u64((uint64_t)nm * np);
    for (uint64_t k = 0; k < (uint64_t)nm * np; ++k) o.i32(r->q[(uint64_t)b * nm * np + k]);
  }
  o.u64(B);
  for (uint32_t b = 0; b < B; ++b) {
    o.u64(r->nchi);
    for (uint32_t c = 0; c < r->nchi; ++c) {
      o.u64((uint64_t)nm * np);
      for (uint64_t k = 0; k < (uint64_t)nm * np; ++k) o.f32(r->suscept[((uint64_t)b * r->nchi + c) * nm * np + k]);
    }
  }
  return write_file(path, o.b);
}

tamc_status tamc_write_anneal_state(const tamc_sa_result *r, const char *path) {
  clear_error();
  if (!r || !path) return fail(TAMC_E_INVALID_ARG, "null argument");
  const uint32_t n = r->n, nb = (n + 7) / 8;
  const float thr = r->min_energy + 1.0e-8f;  // f32 arithmetic as in src/sa.rs:258
  std::vector<uint8_t> bytes;
  uint32_t cnt = 0;
  for (uint32_t k = 0; k < r->R; ++k) {
    if (!(r->energies[k] <= thr)) continue;
    ++cnt;
    const size_t o = bytes.size();
    bytes.resize(o + nb, 0);
    for (uint32_t i = 0; i < n; ++i)
      if (!(r->final_states[(size_t)k * n + i] > 0)) bytes[o + i / 8] |= (uint8_t)(1u << (i % 8));
  }
  Out o;
  o.u32(n);
  o.u32(cnt);
  o.u64(bytes.size());
  o.raw(bytes.data(), bytes.size());
  return write_file(path, o.b);
}

}  // extern "C"
