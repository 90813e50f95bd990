// writers.hip -- serialisation of the reference's output structures, fields in declaration order.
//   PtIcmThermalSamples  src/pt.rs:116-125 (written at src/pt.rs:640-654)
//   AnnealState          src/sa.rs:67-89  (written at src/sa.rs:255-270)
// bincode-1: little-endian, fixed-width integers, u64 length prefixes.
// Python pickle (the reference writes these with serde-pickle when the file name ends in .pkl):
// protocol 3, a struct is a dict keyed by field name, a Vec (Vec<u8> included) is a list, integers
// are ints and f32 values are floats.  The object a pickle loader rebuilds is what is specified;
// serde-pickle's exact opcode choice is not reproduced.
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

// Python pickle protocol 3 emitter.
struct Pickle {
  std::vector<uint8_t> b;
  Pickle() {
    b.push_back(0x80);  // PROTO
    b.push_back(3);
  }
  void op(uint8_t c) { b.push_back(c); }
  void integer(int64_t v) {
    if (v >= 0 && v < 256) {
      op('K');
      op((uint8_t)v);
    } else if (v >= 0 && v < 65536) {
      op('M');
      op((uint8_t)v);
      op((uint8_t)(v >> 8));
    } else if (v >= INT32_MIN && v <= INT32_MAX) {
      op('J');
      const int32_t w = (int32_t)v;
      b.insert(b.end(), (const uint8_t *)&w, (const uint8_t *)&w + 4);
    } else {
      op(0x8a);  // LONG1: little-endian two's complement
      op(8);
      b.insert(b.end(), (const uint8_t *)&v, (const uint8_t *)&v + 8);
    }
  }
  void real(double v) {
    op('G');  // BINFLOAT, big-endian
    uint64_t u;
    memcpy(&u, &v, 8);
    for (int k = 7; k >= 0; --k) op((uint8_t)(u >> (8 * k)));
  }
  void str(const char *s) {
    const uint32_t n = (uint32_t)strlen(s);
    op('X');
    b.insert(b.end(), (const uint8_t *)&n, (const uint8_t *)&n + 4);
    b.insert(b.end(), (const uint8_t *)s, (const uint8_t *)s + n);
  }
  void dict_begin() {
    op('}');
    op('(');
  }
  void dict_end() { op('u'); }
  void list_begin() {
    op(']');
    op('(');
  }
  void list_end() { op('e'); }
  void tuple3() { op(0x87); }  // TUPLE3: the last three items become a tuple
  void stop() { op('.'); }
};

tamc_status write_file(const char *path, const std::vector<uint8_t> &b,
                       const char *what = "Failed to create sample output file ") {
  FILE *f = fopen(path, "wb");
  if (!f) return fail(TAMC_E_IO, std::string(what) + path);
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

tamc_status tamc_write_thermal_samples_pickle(const tamc_pt_result *r, const float *betas, uint8_t compression_level,
                                              const char *path) {
  clear_error();
  if (!r || !betas || !path) return fail(TAMC_E_INVALID_ARG, "null argument");
  const uint32_t B = r->B, C = r->C, nm = r->n_meas, np = r->npairs;
  Pickle o;
  o.dict_begin();
  o.str("compression_level");
  o.integer(compression_level);
  o.str("instance_size");
  o.integer(r->n);
  o.str("beta_arr");
  o.list_begin();
  for (uint32_t b = 0; b < B; ++b) o.real(betas[b]);
  o.list_end();
  o.str("samples");
  o.list_begin();
  const uint64_t per = (uint64_t)r->n_samp_events * C;
  for (uint32_t s = 0; s < r->n_samp_slots; ++s) {
    o.list_begin();
    for (uint64_t k = 0; k < per; ++k) {
      o.list_begin();
      const uint8_t *p = r->samples + ((uint64_t)s * per + k) * r->nbytes;
      for (uint32_t x = 0; x < r->nbytes; ++x) o.integer(p[x]);
      o.list_end();
    }
    o.list_end();
  }
  o.list_end();
  o.str("e");
  o.list_begin();
  for (uint32_t b = 0; b < B; ++b) {
    o.list_begin();
    for (uint64_t k = 0; k < (uint64_t)nm * C; ++k) o.real(r->e[(uint64_t)b * nm * C + k]);
    o.list_end();
  }
  o.list_end();
  o.str("q");
  o.list_begin();
  for (uint32_t b = 0; b < B; ++b) {
    o.list_begin();
    for (uint64_t k = 0; k < (uint64_t)nm * np; ++k) o.integer(r->q[(uint64_t)b * nm * np + k]);
    o.list_end();
  }
  o.list_end();
  o.str("suscept");
  o.list_begin();
  for (uint32_t b = 0; b < B; ++b) {
    o.list_begin();
    for (uint32_t c = 0; c < r->nchi; ++c) {
      o.list_begin();
      for (uint64_t k = 0; k < (uint64_t)nm * np; ++k) o.real(r->suscept[((uint64_t)b * r->nchi + c) * nm * np + k]);
      o.list_end();
    }
    o.list_end();
  }
  o.list_end();
  o.dict_end();
  o.stop();
  return write_file(path, o.b);
}

tamc_status tamc_write_anneal_state_pickle(const tamc_sa_result *r, const char *path) {
  clear_error();
  if (!r || !path) return fail(TAMC_E_INVALID_ARG, "null argument");
  const uint32_t n = r->n, nb = (n + 7) / 8;
  const float thr = r->min_energy + 1.0e-8f;
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
  Pickle o;
  o.dict_begin();
  o.str("instance_size");
  o.integer(n);
  o.str("num_replicas");
  o.integer(cnt);
  o.str("bytes");
  o.list_begin();
  for (uint8_t v : bytes) o.integer(v);
  o.list_end();
  o.dict_end();
  o.stop();
  return write_file(path, o.b);
}


// GlaResults (src/gla.rs:62-70) through write_data (src/util.rs:26-36): pickle for .pkl paths,
// else bincode-1 (boundary entries are (u32, u32, f32) tuples, packed in order).
tamc_status tamc_write_gla_results(const tamc_gla_result *r, const char *path) {
  clear_error();
  if (!r || !path) return fail(TAMC_E_INVALID_ARG, "null argument");
  const size_t len = strlen(path);
  const bool pkl = len >= 4 && strcmp(path + len - 4, ".pkl") == 0;
  const char *what = "Failed to write GLA data to ";
  if (pkl) {
    Pickle o;
    o.dict_begin();
    o.str("instance_size");
    o.integer(r->instance_size);
    o.str("num_partitions");
    o.integer(r->num_partitions);
    o.str("partition_energies");
    o.list_begin();
    for (uint32_t k = 0; k < r->num_partitions; ++k) o.real(r->partition_energies[k]);
    o.list_end();
    o.str("boundary_energies");
    o.list_begin();
    for (uint32_t k = 0; k < r->num_boundaries; ++k) {
      o.integer(r->boundary_pairs[2 * k]);
      o.integer(r->boundary_pairs[2 * k + 1]);
      o.real(r->boundary_energies[k]);
      o.tuple3();
    }
    o.list_end();
    o.str("final_state");
    o.list_begin();
    o.list_end();
    o.str("total_energy");
    o.real(r->total_energy);
    o.dict_end();
    o.stop();
    return write_file(path, o.b, what);
  }
  Out o;
  o.u32(r->instance_size);
  o.u32(r->num_partitions);
  o.u64(r->num_partitions);
  for (uint32_t k = 0; k < r->num_partitions; ++k) o.f32(r->partition_energies[k]);
  o.u64(r->num_boundaries);
  for (uint32_t k = 0; k < r->num_boundaries; ++k) {
    o.u32(r->boundary_pairs[2 * k]);
    o.u32(r->boundary_pairs[2 * k + 1]);
    o.f32(r->boundary_energies[k]);
  }
  o.u64(0);  // final_state: left empty by the reference
  o.f32(r->total_energy);
  return write_file(path, o.b, what);
}

}  // extern "C"
