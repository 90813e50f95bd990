// elementary.hip -- the parity surface of the C ABI: energy_ref, delta_energy, overlap and bit
// packing on the device; the Houdayer move through the production ICM code; the lock-step PT run.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "pt_types.hpp"
#include "runtime_common.hpp"
#include "tamc_math.hpp"

namespace tamc {

struct LockParams;  // lockstep.hip

// Instance::energy_ref (src/ising.rs:438-454), thread per state, reference order.
__global__ void energy_kernel(const uint32_t *row_ptr, const uint32_t *col, const float *J, const float *bias,
                              float offset, uint32_t n, uint32_t ns, const int8_t *states, float *out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const int8_t *s = states + (size_t)k * n;
  float total = offset;
  for (uint32_t i = 0; i < n; ++i) {
    float h = 0.0f;
    const float si = (float)s[i];
    h += bias[i];
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) h += __fmul_rn(J[e], (float)s[col[e]]) / 2.0f;
    total = __fadd_rn(total, __fmul_rn(h, si));
  }
  out[k] = total;
}

// Instance::delta_energy (src/ising.rs:468-480), thread per (state, site).
__global__ void delta_kernel(const uint32_t *row_ptr, const uint32_t *col, const float *J, const float *bias,
                             uint32_t n, uint32_t ns, const int8_t *states, float *out) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)ns * n) return;
  const uint32_t k = (uint32_t)(idx / n), i = (uint32_t)(idx % n);
  const int8_t *s = states + (size_t)k * n;
  float d = 0.0f;
  d += bias[i];
  for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) d = __fadd_rn(d, __fmul_rn(J[e], (float)s[col[e]]));
  d = __fmul_rn(d, -2.0f * (float)s[i]);
  out[idx] = d;
}

// overlap (src/ising.rs:159-165) and write_to_bytes (src/ising.rs:96-121), thread per pair / state
__global__ void overlap_kernel(uint32_t n, uint32_t np, const int8_t *a, const int8_t *b, int64_t *q) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  int64_t s = 0;
  for (uint32_t i = 0; i < n; ++i) s += (int64_t)(a[(size_t)k * n + i] * b[(size_t)k * n + i]);
  q[k] = s;
}
__global__ void pack_kernel(uint32_t n, uint32_t ns, const int8_t *st, uint8_t *out) {
  const uint32_t nb = (n + 7) / 8;
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)ns * nb) return;
  const uint32_t k = (uint32_t)(idx / nb), by = (uint32_t)(idx % nb);
  uint32_t v = 0;
  for (uint32_t t = 0; t < 8; ++t) {
    const uint32_t i = by * 8 + t;
    if (i < n && !(st[(size_t)k * n + i] > 0)) v |= 1u << t;
  }
  out[idx] = (uint8_t)v;
}

}  // namespace tamc

using namespace tamc;

static tamc_status upload(DevBuf &b, const void *src, size_t bytes) {
  TAMC_HIP(b.alloc(bytes));
  if (bytes) TAMC_HIP(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
  return TAMC_OK;
}

extern "C" {

tamc_status tamc_energy(const tamc_instance *inst, uint32_t ns, const int8_t *states, float *energies) {
  clear_error();
  if (!inst || (ns && (!states || !energies))) return fail(TAMC_E_INVALID_ARG, "null argument");
  tamc_status st = device_init();
  if (st) return st;
  DeviceInstance *D = device_instance(inst, 1, &st);
  if (!D) return st;
  const uint32_t n = inst->n;
  DevBuf ds, de;
  if ((st = upload(ds, states, (size_t)ns * n))) return st;
  TAMC_HIP(de.alloc((size_t)ns * 4));
  hipLaunchKernelGGL(energy_kernel, dim3((ns + 63) / 64), dim3(64), 0, 0, (const uint32_t *)D->row_ptr.p,
                     (const uint32_t *)D->col.p, (const float *)D->J.p, (const float *)D->bias.p, inst->offset, n, ns,
                     (const int8_t *)ds.p, (float *)de.p);
  TAMC_HIP(hipGetLastError());
  TAMC_HIP(hipMemcpy(energies, de.p, (size_t)ns * 4, hipMemcpyDeviceToHost));
  return TAMC_OK;
}

tamc_status tamc_delta_energy(const tamc_instance *inst, uint32_t ns, const int8_t *states, float *out) {
  clear_error();
  if (!inst || (ns && (!states || !out))) return fail(TAMC_E_INVALID_ARG, "null argument");
  tamc_status st = device_init();
  if (st) return st;
  DeviceInstance *D = device_instance(inst, 1, &st);
  if (!D) return st;
  const uint32_t n = inst->n;
  const uint64_t tot = (uint64_t)ns * n;
  DevBuf ds, de;
  if ((st = upload(ds, states, tot))) return st;
  TAMC_HIP(de.alloc(tot * 4));
  hipLaunchKernelGGL(delta_kernel, dim3((uint32_t)((tot + 255) / 256)), dim3(256), 0, 0, (const uint32_t *)D->row_ptr.p,
                     (const uint32_t *)D->col.p, (const float *)D->J.p, (const float *)D->bias.p, n, ns,
                     (const int8_t *)ds.p, (float *)de.p);
  TAMC_HIP(hipGetLastError());
  TAMC_HIP(hipMemcpy(out, de.p, tot * 4, hipMemcpyDeviceToHost));
  return TAMC_OK;
}

tamc_status tamc_overlap(uint32_t n, uint32_t np, const int8_t *a, const int8_t *b, int64_t *q) {
  clear_error();
  if (np && (!a || !b || !q)) return fail(TAMC_E_INVALID_ARG, "null argument");
  tamc_status st = device_init();
  if (st) return st;
  DevBuf da, db, dq;
  if ((st = upload(da, a, (size_t)np * n)) || (st = upload(db, b, (size_t)np * n))) return st;
  TAMC_HIP(dq.alloc((size_t)np * 8));
  hipLaunchKernelGGL(overlap_kernel, dim3((np + 63) / 64), dim3(64), 0, 0, n, np, (const int8_t *)da.p,
                     (const int8_t *)db.p, (int64_t *)dq.p);
  TAMC_HIP(hipGetLastError());
  TAMC_HIP(hipMemcpy(q, dq.p, (size_t)np * 8, hipMemcpyDeviceToHost));
  return TAMC_OK;
}

tamc_status tamc_pack_bytes(uint32_t n, uint32_t ns, const int8_t *states, uint8_t *out) {
  clear_error();
  if (ns && (!states || !out)) return fail(TAMC_E_INVALID_ARG, "null argument");
  tamc_status st = device_init();
  if (st) return st;
  const uint32_t nb = (n + 7) / 8;
  DevBuf ds, dout;
  if ((st = upload(ds, states, (size_t)ns * n))) return st;
  TAMC_HIP(dout.alloc((size_t)ns * nb));
  const uint64_t tot = (uint64_t)ns * nb;
  hipLaunchKernelGGL(pack_kernel, dim3((uint32_t)((tot + 255) / 256)), dim3(256), 0, 0, n, ns,
                     (const int8_t *)ds.p, (uint8_t *)dout.p);
  TAMC_HIP(hipGetLastError());
  TAMC_HIP(hipMemcpy(out, dout.p, tot, hipMemcpyDeviceToHost));
  return TAMC_OK;
}

}  // extern "C"
