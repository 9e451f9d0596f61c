// engine.hip -- MI355X (gfx950) FDTD step engine behind the C ABI of
// include/ems_gpu.h.
//
// Replaces the reference engine's step scheduler and CPU kernels
// (src/engine/pool.rs:878-940, src/engine/kernel.rs). Every expression is the
// reference's, in the same order and precision; the library is compiled with
// -ffp-contract=off and IEEE denormals, so fields, probes and energies are
// bit-identical to the CPU engine on the same arrays.
//
// Device data layout (HBM): six f32 field arrays of nx*ny*nz nodes in the
// reference layout (i*ny + j)*nz + k; coefficient lanes as pooled rows plus
// a u32 row index per (i, j), or dense; CPML psi per slab box; a u64 step
// counter and per-operator probe sample buffers.
//
// Kernels (one stream; one step = H, CPML-H slabs, mid gap, E, CPML-E slabs,
// post-E gap), captured into HIP graphs of whole steps:
//   h_update / e_update: one thread per in-plane node p = j*nz + k (z fastest,
//     fully coalesced over the contiguous ny*nz plane), marching a chunk of
//     x-planes and carrying the neighbour plane it already loaded in
//     registers (uy, uz of plane i+1 for H; iy, iz of plane i-1 for E).
//   cpml_slab: one thread per node of a slab box.
//   gap_mid / gap_post_e: one thread runs the serial operator hooks in
//     registration order (they are few and must stay ordered).
//   energy_*: the invariant as f64 chains in exactly the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <string>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/ems_gpu.h"
#include "../../include/ems_host.h"

// Scratch of ems_gpu_energy_noalloc: one static device array per GPU the
// library runs on, held until the process ends (16 MiB). Host code never
// passes this name to a kernel or a copy (on the host it is a stand-in, not
// a device address): it uses the address hipGetSymbolAddress returns for
// the current device (noalloc_scratch below).
static __device__ __align__(256) double scratch[EMS_ENERGY_NOALLOC_DOUBLES];

namespace {

// ---------------------------------------------------------------- device --

// Global (1) and constant (4) address spaces: pointers the kernels load
// from memory are cast so that the compiler emits global / scalar loads
// instead of flat ones.
#define GAS __attribute__((address_space(1)))
#define CAS __attribute__((address_space(4)))

// kernel.rs:30-33
__device__ __forceinline__ float flush(float x) { return fabsf(x) < FLT_MIN ? 0.0f : x; }

// kernel.rs:38-50 (openEMS order, every op flushed) and :76, :270 (Paired).
template <bool OE>
__device__ __forceinline__ float h_expr(float da, float h, float db, float a, float b, float c,
                                        float d) {
  if (OE) {
    float curl = flush(flush(flush(a - b) + d) - c);
    return flush(flush(da * h) - flush(db * curl));
  }
  return flush(da * h - db * ((a - b) - (c - d)));
}
template <bool OE>
__device__ __forceinline__ float e_expr(float ca, float u, float cb, float a, float b, float c,
                                        float d) {
  if (OE) {
    float curl = flush(flush(flush(a - b) - c) + d);
    return flush(flush(ca * u) + flush(cb * curl));
  }
  return flush(ca * u + cb * ((a - b) - (c - d)));
}

// kernel.rs:114-117
__device__ __forceinline__ double weight(float a, double p, float b) {
  double w = (1.0 + (double)a) * p / (double)b;
  return b != 0.0f ? w : 0.0;
}

// src/constants.rs
constexpr double C0_D = 299792458.0;
constexpr double MU0_D = 1.25663706127e-6;
constexpr double EPS0_D = 1.0 / (MU0_D * C0_D * C0_D);

// oe_square (kernel.rs:198-210): x^2 in f64, infinite where its f32 would
// be (round to nearest: from the midpoint between FLT_MAX and 2^128).
__device__ __forceinline__ double oe_square(double x) {
  const double s = x * x;
  return s >= 0x1.ffffffp+127 ? __builtin_huge_val() : s;
}

// One coefficient lane: rows + per-(i,j) row index (pooled), rows only
// (dense, index == nullptr) or nothing (all zero).
struct Lane {
  const float *rows;
  const uint32_t *index;
};
struct Lanes {
  Lane l[12];
};

__device__ __forceinline__ float coef(const Lane &L, size_t r, int k, int nz, size_t flat) {
  if (L.rows == nullptr) return 0.0f;
  if (L.index == nullptr) return L.rows[flat];
  return L.rows[(size_t)L.index[r] * nz + k];
}

struct Fields {
  float *u[3];
  float *i[3];
};

// A lane by a wave-uniform runtime index, through selects with constant
// indices (a dynamic index into a kernel argument would go through memory).
__device__ __forceinline__ Lane pick_lane(const Lanes &L, int l) {
  Lane x = L.l[0];
#pragma unroll
  for (int q = 1; q < 12; q++)
    if (l == q) x = L.l[q];
  return x;
}

// coef through global-address-space pointers.
__device__ __forceinline__ float coefg(const Lane &L, size_t r, int k, int nz, size_t flat) {
  if (L.rows == nullptr) return 0.0f;
  const GAS float *rows = (const GAS float *)L.rows;
  if (L.index == nullptr) return rows[flat];
  return rows[(size_t)((const GAS uint32_t *)L.index)[r] * nz + k];
}

// The block of a one-dimensional launch of gx * nch blocks that takes
// block b's place, so that each XCD runs a contiguous share of the
// (row block, x chunk) pairs: a launch hands its blocks to the 8 XCDs in
// turn (block b to XCD b % 8), so block b + 1, whose rows neighbour block
// b's, would otherwise run beside another L2. A bijection of [0, nb).
__device__ __forceinline__ int xcd_contiguous(int b, int nb) {
  const int per = nb / 8, rem = nb % 8, x = b % 8, q = b / 8;
  return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + q;
}

// Faraday update (kernel.rs:443-511, 566-618): plane-wise in x, all three
// components at node (i, j, k); padding entries are never written.
template <bool OE>
__global__ void __launch_bounds__(256) h_update(Fields f, Lanes L, int nx, int ny, int nz,
                                                int chunk) {
  const int plane = ny * nz;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= plane) return;
  const int j = p / nz, k = p - j * nz;
  const int i0 = blockIdx.y * chunk;
  const int i1 = min(nx, i0 + chunk);
  const bool jv = j < ny - 1, kv = k < nz - 1;
  const GAS float *__restrict__ ux = (const GAS float *)f.u[0];
  const GAS float *__restrict__ uy = (const GAS float *)f.u[1];
  const GAS float *__restrict__ uz = (const GAS float *)f.u[2];
  GAS float *__restrict__ ix = (GAS float *)f.i[0];
  GAS float *__restrict__ iy = (GAS float *)f.i[1];
  GAS float *__restrict__ iz = (GAS float *)f.i[2];
  size_t o = (size_t)i0 * plane + p;
  float uyc = uy[o], uzc = uz[o];
  for (int i = i0; i < i1; i++, o += plane) {
    const bool iv = i + 1 < nx;
    const size_t r = (size_t)i * ny + j;
    const float uxc = ux[o];
    const float uyn = iv ? uy[o + plane] : 0.0f;
    const float uzn = iv ? uz[o + plane] : 0.0f;
    if (jv && kv) {  // curl_x = (uz[j+1] - uz[j]) - (uy[k+1] - uy[k])
      float da = coefg(L.l[6], r, k, nz, o), db = coefg(L.l[9], r, k, nz, o);
      ix[o] = h_expr<OE>(da, ix[o], db, uz[o + nz], uzc, uy[o + 1], uyc);
    }
    if (iv && kv) {  // curl_y = (ux[k+1] - ux[k]) - (uz[i+1] - uz[i])
      float da = coefg(L.l[7], r, k, nz, o), db = coefg(L.l[10], r, k, nz, o);
      iy[o] = h_expr<OE>(da, iy[o], db, ux[o + 1], uxc, uzn, uzc);
    }
    if (iv && jv) {  // curl_z = (uy[i+1] - uy[i]) - (ux[j+1] - ux[j])
      float da = coefg(L.l[8], r, k, nz, o), db = coefg(L.l[11], r, k, nz, o);
      iz[o] = h_expr<OE>(da, iz[o], db, uyn, uyc, ux[o + nz], uxc);
    }
    uyc = uyn;
    uzc = uzn;
  }
}

// Ampere update (kernel.rs:620-697); out-of-domain neighbours read 0.
template <bool OE>
__global__ void __launch_bounds__(256) e_update(Fields f, Lanes L, int nx, int ny, int nz,
                                                int chunk) {
  const int plane = ny * nz;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= plane) return;
  const int j = p / nz, k = p - j * nz;
  const int i0 = blockIdx.y * chunk;
  const int i1 = min(nx, i0 + chunk);
  const GAS float *__restrict__ ix = (const GAS float *)f.i[0];
  const GAS float *__restrict__ iy = (const GAS float *)f.i[1];
  const GAS float *__restrict__ iz = (const GAS float *)f.i[2];
  GAS float *__restrict__ ux = (GAS float *)f.u[0];
  GAS float *__restrict__ uy = (GAS float *)f.u[1];
  GAS float *__restrict__ uz = (GAS float *)f.u[2];
  size_t o = (size_t)i0 * plane + p;
  float iyp = i0 > 0 ? iy[o - plane] : 0.0f;
  float izp = i0 > 0 ? iz[o - plane] : 0.0f;
  for (int i = i0; i < i1; i++, o += plane) {
    const size_t r = (size_t)i * ny + j;
    const float ixc = ix[o], iyc = iy[o], izc = iz[o];
    if (i + 1 < nx) {  // curl_x = (iz[j] - iz[j-1]) - (iy[k] - iy[k-1])
      float ca = coefg(L.l[0], r, k, nz, o), cb = coefg(L.l[3], r, k, nz, o);
      float b = j > 0 ? iz[o - nz] : 0.0f;
      float d = k > 0 ? iy[o - 1] : 0.0f;
      ux[o] = e_expr<OE>(ca, ux[o], cb, izc, b, iyc, d);
    }
    if (j + 1 < ny) {  // curl_y = (ix[k] - ix[k-1]) - (iz[i] - iz[i-1])
      float ca = coefg(L.l[1], r, k, nz, o), cb = coefg(L.l[4], r, k, nz, o);
      float b = k > 0 ? ix[o - 1] : 0.0f;
      uy[o] = e_expr<OE>(ca, uy[o], cb, ixc, b, izc, izp);
    }
    if (k + 1 < nz) {  // curl_z = (iy[i] - iy[i-1]) - (ix[j] - ix[j-1])
      float ca = coefg(L.l[2], r, k, nz, o), cb = coefg(L.l[5], r, k, nz, o);
      float d = j > 0 ? ix[o - nz] : 0.0f;
      uz[o] = e_expr<OE>(ca, uz[o], cb, iyc, iyp, ixc, d);
    }
    iyp = iyc;
    izp = izc;
  }
}

// CPML slab correction (cpml.rs:469-543, 693-736).
struct SlabArg {
  int e_side, comp, other, normal;
  float sign;
  int lo[3], hi[3];
  const float *b, *a, *k;
  float *psi;
};

__global__ void __launch_bounds__(256) cpml_slab(Fields f, Lane cfl, SlabArg s, int nx, int ny,
                                                 int nz) {
  const int lx = s.hi[0] - s.lo[0], ly = s.hi[1] - s.lo[1], lz = s.hi[2] - s.lo[2];
  const size_t total = (size_t)lx * ly * lz;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int kk = (int)(q % lz);
  const size_t t2 = q / lz;
  const int jj = (int)(t2 % ly);
  const int ii = (int)(t2 / ly);
  const int i = s.lo[0] + ii, j = s.lo[1] + jj, k = s.lo[2] + kk;
  const size_t plane = (size_t)ny * nz;
  const size_t o = (size_t)i * plane + (size_t)j * nz + k;
  const bool es = s.e_side != 0;
  // the two arrays by selects on the uniform component numbers (a dynamic
  // index into the argument would go through memory)
  float *const *src = es ? f.i : f.u;
  float *const *dst = es ? f.u : f.i;
  const GAS float *cur =
      (const GAS float *)(s.other == 0 ? src[0] : (s.other == 1 ? src[1] : src[2]));
  GAS float *t = (GAS float *)(s.comp == 0 ? dst[0] : (s.comp == 1 ? dst[1] : dst[2]));
  float h, l;
  int z;
  if (s.normal == 0) {
    if (es) {
      h = cur[o];
      l = i == 0 ? 0.0f : cur[o - plane];
    } else {
      h = cur[o + plane];
      l = cur[o];
    }
    z = ii;
  } else if (s.normal == 1) {
    if (es) {
      h = cur[o];
      l = j == 0 ? 0.0f : cur[o - nz];
    } else {
      h = cur[o + nz];
      l = cur[o];
    }
    z = jj;
  } else {
    if (es) {
      h = cur[o];
      l = k == 0 ? 0.0f : cur[o - 1];
    } else {
      h = cur[o + 1];
      l = cur[o];
    }
    z = kk;
  }
  const float cf = coefg(cfl, (size_t)i * ny + j, k, nz, o);
  const float scale = es ? 1.0f : -1.0f;
  const float sv = s.sign * (h - l);
  GAS float *psi = (GAS float *)s.psi;
  const float ps = flush(((const GAS float *)s.b)[z] * psi[q] + ((const GAS float *)s.a)[z] * sv);
  psi[q] = ps;
  t[o] = flush(t[o] + scale * cf * (((const GAS float *)s.k)[z] * sv + ps));
  (void)nx;
}

// ---- fused step (H, CPML-H, E, CPML-E in one pass) ----
//
// Ping-pong buffers: reads step n from A, writes step n+1 to B, so tiles can
// recompute their neighbours' H values without racing. A block owns a tile
// of TJ rows (j) by tk nodes (k) and marches a chunk of x-planes; its
// threads cover the tile plus one halo row (j0-1) and one halo column
// (k0-1), whose H values (after the CPML correction) the E update of the
// tile needs. H values of the plane go through LDS; the previous plane's
// y and z currents stay in registers. Every value is the reference's
// expression on the same inputs, so results equal the unfused schedule bit
// for bit (the reference's fused schedule rests on the same fact,
// pool.rs:475-487).

// Address spaces. Pointers the kernels load from memory (argument structs,
// slab descriptors) are generic to the compiler, which then emits flat
// accesses (no scalar-base addressing, waits on the LDS counter too) and
// vector loads for uniform struct fields. Data pointers are cast to the
// global space; read-only descriptors at uniform addresses to the constant
// space, which the compiler reads through the scalar unit.
template <class T>
__device__ __forceinline__ T cload(const T *p, size_t i = 0) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const CAS T *)p)[i];
#else
  return p[i];  // (the host pass only checks the code)
#endif
}

// ---- CPML slabs outside the fused kernel ----
//
// A slab's H-side correction t = flush(t + (-db)*(k*s + psi')) with
// psi' = flush(b*psi + a*s) depends on t only through the final add: s is a
// difference of old voltages u^n, which the fused step only reads. So the
// increment (-db)*(k*s + psi') can be formed before the step (cpml_delta_h)
// and added by the fused kernel where the slab's turn comes, with the same
// bits. The E-side correction reads the new currents and is the last write
// to its component, so a slab whose corrections follow every other slab's
// at its nodes (a suffix of the registration order) can run after the fused
// kernel (cpml_post_e). The fused step moves the slabs normal to z there:
// in a k-tile only the few lanes inside such a slab would run its code, and
// the whole wave would pay for it.
struct DSlab {
  int comp, normal, other;
  float sign;
  int lo0, hi0, lo1, hi1, lo2, hi2, ly, lz;
  const float *b, *a, *k;
  float *psi0, *psi1;  // H side: read psi[par], write psi[1 - par]; E side: psi0
  float *delta;        // H side: the step's increments (box order)
  uint32_t first;      // the slab's first thread in its launch
};

__device__ __forceinline__ int dslab_of(const DSlab *ds, int nds, uint32_t q) {
  int s = 0;
  while (s + 1 < nds && cload(ds, s + 1).first <= q) s++;
  return s;
}

// The increments of the H-side slabs ds (cpml.rs:495-543 with scale -1),
// from u^n in A; psi read from buffer par, written to the other.
__global__ void __launch_bounds__(256) cpml_delta_h(const float *__restrict__ A, size_t N, Lanes L,
                                                    const DSlab *__restrict__ ds, int nds,
                                                    uint32_t total, int par, int nx, int ny,
                                                    int nz) {
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  if (q >= total) return;
  const DSlab d = cload(ds, dslab_of(ds, nds, q));
  const uint32_t p = q - d.first;
  const int kk = (int)(p % (uint32_t)d.lz);
  const uint32_t t2 = p / (uint32_t)d.lz;
  const int jj = (int)(t2 % (uint32_t)d.ly), ii = (int)(t2 / (uint32_t)d.ly);
  const int i = d.lo0 + ii, j = d.lo1 + jj, k = d.lo2 + kk;
  const size_t plane = (size_t)ny * nz;
  const size_t o = (size_t)i * plane + (size_t)j * nz + k;
  const GAS float *u = (const GAS float *)(A + (size_t)d.other * N);
  // s = u_other[node + e_n] - u_other[node], 0 beyond the grid
  float hv;
  int z;
  if (d.normal == 0) hv = i + 1 < nx ? u[o + plane] : 0.0f, z = ii;
  else if (d.normal == 1) hv = j + 1 < ny ? u[o + nz] : 0.0f, z = jj;
  else hv = k + 1 < nz ? u[o + 1] : 0.0f, z = kk;
  const float sv = d.sign * (hv - u[o]);
  const float cf = coefg(pick_lane(L, 9 + d.comp), (size_t)i * ny + j, k, nz, o);
  const GAS float *pin = (const GAS float *)(par == 0 ? d.psi0 : d.psi1);
  GAS float *pout = (GAS float *)(par == 1 ? d.psi0 : d.psi1);
  const float ps = flush(((const GAS float *)d.b)[z] * pin[p] + ((const GAS float *)d.a)[z] * sv);
  pout[p] = ps;
  const float scale = -1.0f;
  ((GAS float *)d.delta)[p] = scale * cf * (((const GAS float *)d.k)[z] * sv + ps);
}

// The E-side slabs ds after the fused kernel, on its output B (new u and i);
// no two of them correct one component at one node.
__global__ void __launch_bounds__(256) cpml_post_e(float *__restrict__ B, size_t N, Lanes L,
                                                   const DSlab *__restrict__ ds, int nds,
                                                   uint32_t total, int ny, int nz) {
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  if (q >= total) return;
  const DSlab d = cload(ds, dslab_of(ds, nds, q));
  const uint32_t p = q - d.first;
  const int kk = (int)(p % (uint32_t)d.lz);
  const uint32_t t2 = p / (uint32_t)d.lz;
  const int jj = (int)(t2 % (uint32_t)d.ly), ii = (int)(t2 / (uint32_t)d.ly);
  const int i = d.lo0 + ii, j = d.lo1 + jj, k = d.lo2 + kk;
  const size_t plane = (size_t)ny * nz;
  const size_t o = (size_t)i * plane + (size_t)j * nz + k;
  const GAS float *cur = (const GAS float *)(B + (size_t)(3 + d.other) * N);
  // s = i_other[node] - i_other[node - e_n], 0 below the grid
  float lv;
  int z;
  if (d.normal == 0) lv = i > 0 ? cur[o - plane] : 0.0f, z = ii;
  else if (d.normal == 1) lv = j > 0 ? cur[o - nz] : 0.0f, z = jj;
  else lv = k > 0 ? cur[o - 1] : 0.0f, z = kk;
  const float sv = d.sign * (cur[o] - lv);
  const float cf = coefg(pick_lane(L, 3 + d.comp), (size_t)i * ny + j, k, nz, o);
  GAS float *psi = (GAS float *)d.psi0;
  const float ps = flush(((const GAS float *)d.b)[z] * psi[p] + ((const GAS float *)d.a)[z] * sv);
  psi[p] = ps;
  GAS float *t = (GAS float *)(B + (size_t)d.comp * N);
  const float scale = 1.0f;
  t[o] = flush(t[o] + scale * cf * (((const GAS float *)d.k)[z] * sv + ps));
}

// The increments of the H-side slabs handled by cpml_delta_h at node
// (i, j, k), in registration order, over the slabs whose bits are set in
// `mask` (those reaching the tile and the plane).
__device__ __forceinline__ void h_out_apply(const DSlab *__restrict__ ds, uint32_t mask, int i,
                                            int j, int k, float h[3]) {
  while (mask) {
    const int q = __builtin_ctz(mask);
    mask &= mask - 1;
    const DSlab d = cload(ds, q);
    if (j < d.lo1 || j >= d.hi1 || k < d.lo2 || k >= d.hi2) continue;
    const size_t p = ((size_t)(i - d.lo0) * d.ly + (j - d.lo1)) * d.lz + (k - d.lo2);
    const float dv = ((const GAS float *)d.delta)[p];
    if (d.comp == 0) h[0] = flush(h[0] + dv);
    else if (d.comp == 1) h[1] = flush(h[1] + dv);
    else h[2] = flush(h[2] + dv);
  }
}

// One CPML slab as the fused kernel reads it (cpml.rs:394-420). H-side psi
// is double-buffered with the fields; E-side psi is updated in place (only
// the owner of a node computes its E update).
struct FSlab {
  int e_side, comp, normal;
  float sign;
  int lo0, hi0, lo1, hi1, lo2, hi2;
  int ly, lz;
  const float *b, *a, *k;
  float *psi0, *psi1;
};

// One slab's correction of component t at a node inside its box: the new
// psi and the corrected value (cpml.rs:495-543).
template <bool ESIDE>
__device__ __forceinline__ void cpml_apply(const FSlab &s, float pin, float bz, float az, float kz,
                                           float d01, float d02, float d10, float d12, float d20,
                                           float d21, float cf0, float cf1, float cf2,
                                           float &t0, float &t1, float &t2, float &ps) {
  const int c = s.comp, n = s.normal;
  float d = d01;
  d = (c == 0 && n == 2) ? d02 : d;
  d = (c == 1 && n == 0) ? d10 : d;
  d = (c == 1 && n == 2) ? d12 : d;
  d = (c == 2 && n == 0) ? d20 : d;
  d = (c == 2 && n == 1) ? d21 : d;
  const float cf = c == 0 ? cf0 : (c == 1 ? cf1 : cf2);
  float t = c == 0 ? t0 : (c == 1 ? t1 : t2);
  const float sv = s.sign * d;
  ps = flush(bz * pin + az * sv);
  const float scale = ESIDE ? 1.0f : -1.0f;
  t = flush(t + scale * cf * (kz * sv + ps));
  if (c == 0) t0 = t;
  else if (c == 1) t1 = t;
  else t2 = t;
}

__device__ __forceinline__ bool cpml_inside(const FSlab &s, int j, int k) {
  return !(j < s.lo1 || j >= s.hi1 || k < s.lo2 || k >= s.hi2);
}

__device__ __forceinline__ int cpml_depth(const FSlab &s, int i, int j, int k) {
  return s.normal == 0 ? i - s.lo0 : (s.normal == 1 ? j - s.lo1 : k - s.lo2);
}

__device__ __forceinline__ size_t cpml_index(const FSlab &s, int i, int j, int k) {
  return ((size_t)(i - s.lo0) * s.ly + (j - s.lo1)) * s.lz + (k - s.lo2);
}

// CPML corrections of one side at node (i, j, k), in registration order,
// over the slabs whose bits are set in `mask` (the slabs that can reach the
// block's tile and plane). d<c><n> is the difference h - l of the
// derivative along normal n used by component c (cpml.rs:495-521); cf<c>
// the cb (E) or db (H) of component c.
//
// One slab at a time: a form that issued every slab's loads first (up to 6
// slabs' psi, b, a, k in registers) held 24 more VGPRs, six waves per SIMD
// instead of seven for the whole tiled kernel, and measured slower on the
// 512^3 CPML box (2.976 against 2.849 ms per launch; with 2 slots 2.949).
//
// INPLACE (the unfused kernels, which compute each node once): the psi of
// parity `par` is read and written (the fused kernel's halo recompute reads
// one buffer and writes the other).
template <bool ESIDE, bool INPLACE = false>
__device__ __forceinline__ void cpml_node(const FSlab *__restrict__ slabs, uint32_t mask, int par,
                                          bool write_psi, int i, int j, int k, float d01,
                                          float d02, float d10, float d12, float d20, float d21,
                                          float cf0, float cf1, float cf2, float &t0, float &t1,
                                          float &t2) {
  while (mask) {
    const int q = __builtin_ctz(mask);
    mask &= mask - 1;
    const FSlab s = cload(slabs, q);
    // (The plane is inside the slab's x range: the caller's mask says so.)
    if (!cpml_inside(s, j, k)) continue;
    const int z = cpml_depth(s, i, j, k);
    const size_t qq = cpml_index(s, i, j, k);
    const GAS float *pin = (const GAS float *)((ESIDE || par == 0) ? s.psi0 : s.psi1);
    GAS float *pout = (GAS float *)((ESIDE || par == (INPLACE ? 0 : 1)) ? s.psi0 : s.psi1);
    float ps;
    cpml_apply<ESIDE>(s, pin[qq], ((const GAS float *)s.b)[z], ((const GAS float *)s.a)[z],
                      ((const GAS float *)s.k)[z], d01, d02, d10, d12, d20, d21, cf0, cf1, cf2, t0,
                      t1, t2, ps);
    if (write_psi) pout[qq] = ps;
  }
}

// ---- unfused H and E updates, software pipelined along x ----
//
// The same updates (same expressions, same values written) with a thread
// marching a chunk of x-planes and the next plane's inputs loaded before
// the present plane is computed, so the loads of one plane overlap the
// arithmetic and stores of the one before. Every load address is valid
// (neighbours past the grid's end fall back to the node itself; those
// values are not used).
//
// CP: the side's CPML slabs (the fused kernel's slab table, every slab of
// the side in it) corrected here, after the update, in registration order
// with the very expressions of cpml_slab, instead of in launches of their
// own; a plane that a slab's x range holds stores all three components.

// The slabs (bits of the first `nslab` table entries) whose (j, k) box
// meets the block's nodes p0..p1 of a plane: rows p0/nz..p1/nz and, over
// one row, its columns (else every column).
__device__ __forceinline__ uint32_t block_slabs(const FSlab *__restrict__ slabs, int nslab,
                                                int p0, int p1, int nz) {
  const int j0 = p0 / nz, j1 = p1 / nz;
  const int k0 = j0 == j1 ? p0 - j0 * nz : 0, k1 = j0 == j1 ? p1 - j1 * nz : nz - 1;
  uint32_t m = 0;
  for (int q = 0; q < nslab; q++) {
    const FSlab s = cload(slabs, q);
    if (s.lo1 <= j1 && j0 < s.hi1 && s.lo2 <= k1 && k0 < s.hi2) m |= 1u << q;
  }
  return m;
}

// The fold's tables: the slab table, per x-plane the slabs (H, E) whose x
// range holds it, the table's length and the H psi parity.
struct CpmlIn {
  const FSlab *slabs;
  const uint32_t *pmask;
  int nslab, par;
};

// What the H update of one node reads (kernel.rs:443-511).
struct HNode {
  float uxc, uyn, uzn, ix, iy, iz, uz_j1, uy_k1, ux_k1, ux_j1;
  float da[3], db[3];
};

__device__ __forceinline__ void h_node_load(HNode &q, const Fields &f, const Lanes &L, size_t o,
                                            size_t r, int k, int nz, size_t plane, bool iv,
                                            bool jv, bool kv) {
  const GAS float *ux = (const GAS float *)f.u[0], *uy = (const GAS float *)f.u[1];
  const GAS float *uz = (const GAS float *)f.u[2];
  const size_t oi = iv ? o + plane : o, oj = jv ? o + nz : o, ok = kv ? o + 1 : o;
  q.uxc = ux[o], q.uyn = uy[oi], q.uzn = uz[oi];
  q.ix = ((const GAS float *)f.i[0])[o], q.iy = ((const GAS float *)f.i[1])[o];
  q.iz = ((const GAS float *)f.i[2])[o];
  q.uz_j1 = uz[oj], q.ux_j1 = ux[oj], q.uy_k1 = uy[ok], q.ux_k1 = ux[ok];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    q.da[a] = coefg(L.l[6 + a], r, k, nz, o);
    q.db[a] = coefg(L.l[9 + a], r, k, nz, o);
  }
}

template <bool OE, bool CP>
__global__ void __launch_bounds__(256) h_update_pf(Fields f, Lanes L, int nx, int ny, int nz,
                                                   int chunk, CpmlIn cp) {
  const int plane = ny * nz, gx = (plane + 255) / 256;
  const int b = xcd_contiguous(blockIdx.x, gridDim.x);
  const int p = (b % gx) * blockDim.x + threadIdx.x;
  const uint32_t mblk =
      CP ? block_slabs(cp.slabs, cp.nslab, (b % gx) * 256, min((b % gx) * 256 + 255, plane - 1), nz)
         : 0u;
  if (p >= plane) return;
  const int j = p / nz, k = p - j * nz;
  const int i0 = (b / gx) * chunk;
  const int i1 = min(nx, i0 + chunk);
  const bool jv = j < ny - 1, kv = k < nz - 1;
  GAS float *ix = (GAS float *)f.i[0], *iy = (GAS float *)f.i[1], *iz = (GAS float *)f.i[2];
  size_t o = (size_t)i0 * plane + p;
  float uyc = ((const GAS float *)f.u[1])[o], uzc = ((const GAS float *)f.u[2])[o];
  HNode cur, nxt;
  h_node_load(cur, f, L, o, (size_t)i0 * ny + j, k, nz, plane, i0 + 1 < nx, jv, kv);
  for (int i = i0; i < i1; i++, o += plane) {
    const bool iv = i + 1 < nx;
    if (i + 1 < i1)
      h_node_load(nxt, f, L, o + plane, (size_t)(i + 1) * ny + j, k, nz, plane, i + 2 < nx, jv,
                  kv);
    const float uyn = iv ? cur.uyn : 0.0f, uzn = iv ? cur.uzn : 0.0f;
    const uint32_t m = CP ? mblk & cload(cp.pmask, 2 * (size_t)i) : 0u;
    if (CP && m) {
      float h0 = cur.ix, h1 = cur.iy, h2 = cur.iz;
      if (jv && kv) h0 = h_expr<OE>(cur.da[0], h0, cur.db[0], cur.uz_j1, uzc, cur.uy_k1, uyc);
      if (iv && kv) h1 = h_expr<OE>(cur.da[1], h1, cur.db[1], cur.ux_k1, cur.uxc, uzn, uzc);
      if (iv && jv) h2 = h_expr<OE>(cur.da[2], h2, cur.db[2], uyn, uyc, cur.ux_j1, cur.uxc);
      // T_p = u_other[node + e_p] - u_other[node] per (component, normal)
      cpml_node<false, true>(cp.slabs, m, cp.par, true, i, j, k, cur.uz_j1 - uzc,
                                    cur.uy_k1 - uyc, uzn - uzc, cur.ux_k1 - cur.uxc, uyn - uyc,
                                    cur.ux_j1 - cur.uxc, cur.db[0], cur.db[1], cur.db[2], h0, h1,
                                    h2);
      ix[o] = h0, iy[o] = h1, iz[o] = h2;
    } else {
      if (jv && kv)  // curl_x = (uz[j+1] - uz[j]) - (uy[k+1] - uy[k])
        ix[o] = h_expr<OE>(cur.da[0], cur.ix, cur.db[0], cur.uz_j1, uzc, cur.uy_k1, uyc);
      if (iv && kv)  // curl_y = (ux[k+1] - ux[k]) - (uz[i+1] - uz[i])
        iy[o] = h_expr<OE>(cur.da[1], cur.iy, cur.db[1], cur.ux_k1, cur.uxc, uzn, uzc);
      if (iv && jv)  // curl_z = (uy[i+1] - uy[i]) - (ux[j+1] - ux[j])
        iz[o] = h_expr<OE>(cur.da[2], cur.iz, cur.db[2], uyn, uyc, cur.ux_j1, cur.uxc);
    }
    uyc = uyn;
    uzc = uzn;
    cur = nxt;
  }
}

// What the E update of one node reads (kernel.rs:620-697).
struct ENode {
  float ix, iy, iz, iz_jm, iy_km, ix_km, ix_jm, ux, uy, uz;
  float ca[3], cb[3];
};

__device__ __forceinline__ void e_node_load(ENode &q, const Fields &f, const Lanes &L, size_t o,
                                            size_t r, int j, int k, int nz) {
  const GAS float *ix = (const GAS float *)f.i[0], *iy = (const GAS float *)f.i[1];
  const GAS float *iz = (const GAS float *)f.i[2];
  const size_t oj = j > 0 ? o - nz : o, ok = k > 0 ? o - 1 : o;
  q.ix = ix[o], q.iy = iy[o], q.iz = iz[o];
  q.iz_jm = iz[oj], q.ix_jm = ix[oj], q.iy_km = iy[ok], q.ix_km = ix[ok];
  q.ux = ((const GAS float *)f.u[0])[o], q.uy = ((const GAS float *)f.u[1])[o];
  q.uz = ((const GAS float *)f.u[2])[o];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    q.ca[a] = coefg(L.l[a], r, k, nz, o);
    q.cb[a] = coefg(L.l[3 + a], r, k, nz, o);
  }
}

template <bool OE, bool CP>
__global__ void __launch_bounds__(256) e_update_pf(Fields f, Lanes L, int nx, int ny, int nz,
                                                   int chunk, CpmlIn cp) {
  const int plane = ny * nz, gx = (plane + 255) / 256;
  const int b = xcd_contiguous(blockIdx.x, gridDim.x);
  const int p = (b % gx) * blockDim.x + threadIdx.x;
  const uint32_t mblk =
      CP ? block_slabs(cp.slabs, cp.nslab, (b % gx) * 256, min((b % gx) * 256 + 255, plane - 1), nz)
         : 0u;
  if (p >= plane) return;
  const int j = p / nz, k = p - j * nz;
  const int i0 = (b / gx) * chunk;
  const int i1 = min(nx, i0 + chunk);
  GAS float *ux = (GAS float *)f.u[0], *uy = (GAS float *)f.u[1], *uz = (GAS float *)f.u[2];
  size_t o = (size_t)i0 * plane + p;
  float iyp = i0 > 0 ? ((const GAS float *)f.i[1])[o - plane] : 0.0f;
  float izp = i0 > 0 ? ((const GAS float *)f.i[2])[o - plane] : 0.0f;
  ENode cur, nxt;
  e_node_load(cur, f, L, o, (size_t)i0 * ny + j, j, k, nz);
  for (int i = i0; i < i1; i++, o += plane) {
    if (i + 1 < i1) e_node_load(nxt, f, L, o + plane, (size_t)(i + 1) * ny + j, j, k, nz);
    const float hz_jm = j > 0 ? cur.iz_jm : 0.0f, hx_jm = j > 0 ? cur.ix_jm : 0.0f;
    const float hy_km = k > 0 ? cur.iy_km : 0.0f, hx_km = k > 0 ? cur.ix_km : 0.0f;
    const uint32_t m = CP ? mblk & cload(cp.pmask, 2 * (size_t)i + 1) : 0u;
    if (CP && m) {
      float e0 = cur.ux, e1 = cur.uy, e2 = cur.uz;
      if (i + 1 < nx) e0 = e_expr<OE>(cur.ca[0], e0, cur.cb[0], cur.iz, hz_jm, cur.iy, hy_km);
      if (j + 1 < ny) e1 = e_expr<OE>(cur.ca[1], e1, cur.cb[1], cur.ix, hx_km, cur.iz, izp);
      if (k + 1 < nz) e2 = e_expr<OE>(cur.ca[2], e2, cur.cb[2], cur.iy, iyp, cur.ix, hx_jm);
      // S_p = i_other[node] - i_other[node - e_p] (cpml.rs:495-521)
      cpml_node<true, true>(cp.slabs, m, 0, true, i, j, k, cur.iz - hz_jm, cur.iy - hy_km,
                                   cur.iz - izp, cur.ix - hx_km, cur.iy - iyp, cur.ix - hx_jm,
                                   cur.cb[0], cur.cb[1], cur.cb[2], e0, e1, e2);
      ux[o] = e0, uy[o] = e1, uz[o] = e2;
    } else {
      if (i + 1 < nx)  // curl_x = (iz[j] - iz[j-1]) - (iy[k] - iy[k-1])
        ux[o] = e_expr<OE>(cur.ca[0], cur.ux, cur.cb[0], cur.iz, hz_jm, cur.iy, hy_km);
      if (j + 1 < ny)  // curl_y = (ix[k] - ix[k-1]) - (iz[i] - iz[i-1])
        uy[o] = e_expr<OE>(cur.ca[1], cur.uy, cur.cb[1], cur.ix, hx_km, cur.iz, izp);
      if (k + 1 < nz)  // curl_z = (iy[i] - iy[i-1]) - (ix[j] - ix[j-1])
        uz[o] = e_expr<OE>(cur.ca[2], cur.uz, cur.cb[2], cur.iy, iyp, cur.ix, hx_jm);
    }
    iyp = cur.iy;
    izp = cur.iz;
    cur = nxt;
  }
}

// Loads/stores at a wave-uniform base plus a 32-bit per-lane byte offset,
// so the address is a scalar register pair plus one VGPR (saddr form)
// instead of a 64-bit vector add per access.
__device__ __forceinline__ float ldf(const float *base, unsigned boff) {
  return *(const GAS float *)((const GAS char *)base + boff);
}
__device__ __forceinline__ void stf(float *base, unsigned boff, float v) {
  *(GAS float *)((GAS char *)base + boff) = v;
}

// Coefficient l of a node row whose pool offsets are `ro` (croff + 16*r,
// wave-uniform). ro[12] flags the lanes whose row holds one value: that
// value is read once per wave through the scalar unit instead of a vector
// load per lane (the bits are the same).
// CT is the table's element type: uint32_t for pools under 4 GiB (one base
// for every load, the row offset added to the lane's in a VGPR), uint64_t
// for larger pools (the row's address a scalar sum, more scalar registers).
template <typename CT>
__device__ __forceinline__ float ldc(const float *__restrict__ cpool, const CT *ro, int l,
                                     unsigned k4) {
  if constexpr (sizeof(CT) == 4) {
    const uint32_t o = cload(ro, l);
    if ((cload(ro, 12) >> l) & 1u) return cload((const float *)((const char *)cpool + o));
    return ldf(cpool, o + k4);
  } else {
    const float *row = (const float *)((const char *)cpool + cload(ro, l));
    if ((cload(ro, 12) >> l) & 1u) return cload(row);
    return ldf(row, k4);
  }
}

// Coefficient of lane L at row r (i*ny + j, wave-uniform) at byte offset
// k4 of the row; `ob` is the row's first node ((i*ny + j)*nz).
__device__ __forceinline__ float coef_row(const Lane &L, size_t r, size_t ob, int nz,
                                          unsigned k4) {
  if (L.rows == nullptr) return 0.0f;
  const float *row = L.index == nullptr ? L.rows + ob : L.rows + (size_t)L.index[r] * nz;
  return ldf(row, k4);
}

// What one node of one x-plane reads: its three voltages and the
// neighbours the curls need, its three currents, and its 12 coefficients.
struct PlaneIn {
  float uxc, ux_j1, ux_k1, uy_k1, uz_j1, uyn, uzn;
  float ix, iy, iz;
  float c[12];
};

// Field buffer F holds u_x, u_y, u_z, i_x, i_y, i_z, N nodes each; the
// coefficient pool holds every lane's rows, and croff[16*r + l] is the
// offset of lane l's row for node row r (i*ny + j).
template <typename CT = uint32_t>
__device__ __forceinline__ void load_plane(PlaneIn &P, const float *F, size_t N,
                                           const float *cpool, const CT *croff, size_t ob,
                                           size_t r, size_t plane, int nz, unsigned k4, bool iv,
                                           bool jv, bool kv, bool need_e = true) {
  // Branch-free, every address valid (neighbour offsets fall back to the
  // node itself where the neighbour does not exist; the value is then
  // replaced by 0). One wave-uniform base per component array; the
  // neighbours are vector offsets from it, so the scalar unit computes six
  // bases per plane and nothing per neighbour.
  asm volatile("" : "+v"(k4));
  const unsigned kn = k4 + (kv ? 4u : 0u);
  const unsigned kj = k4 + (jv ? 4u * (unsigned)nz : 0u);
  const unsigned ki = k4 + (iv ? 4u * (unsigned)plane : 0u);
  const float *ux = F + ob, *uy = F + N + ob, *uz = F + 2 * N + ob;
  const float uxc = ldf(ux, k4), uxj = ldf(ux, kj), uzj = ldf(uz, kj);
  const float uxk = ldf(ux, kn), uyk = ldf(uy, kn);
  const float uyi = ldf(uy, ki), uzi = ldf(uz, ki);
  P.ix = ldf(F + 3 * N + ob, k4);
  P.iy = ldf(F + 4 * N + ob, k4);
  P.iz = ldf(F + 5 * N + ob, k4);
  // The coefficients at the pool base (or, with 64-bit row offsets, the
  // row's scalar address) plus a 32-bit byte offset per lane: no 64-bit
  // vector address math.
  const CT *ro = croff + 16 * r;
  // Rows that only update H (a tile's halo row, the plane before a chunk;
  // need_e is wave-uniform) skip the E coefficients.
#pragma unroll
  for (int l = 0; l < 12; l++) P.c[l] = (l >= 6 || need_e) ? ldc(cpool, ro, l, k4) : 0.0f;
  P.uxc = uxc;
  P.ux_j1 = jv ? uxj : 0.0f;
  P.uz_j1 = jv ? uzj : 0.0f;
  P.ux_k1 = kv ? uxk : 0.0f;
  P.uy_k1 = kv ? uyk : 0.0f;
  P.uyn = iv ? uyi : 0.0f;
  P.uzn = iv ? uzi : 0.0f;
}

// H update of a node from its plane inputs (kernel.rs:443-511) and its CPML
// correction; padding components keep their old value.
template <bool OE>
__device__ __forceinline__ void h_from(const PlaneIn &P, float uyc, float uzc, bool iv, bool jv,
                                       bool kv, const FSlab *slabs, uint32_t mask, int par,
                                       bool write_psi, int i, int j, int k, float h[3]) {
  h[0] = P.ix, h[1] = P.iy, h[2] = P.iz;
  if (jv && kv) h[0] = h_expr<OE>(P.c[6], h[0], P.c[9], P.uz_j1, uzc, P.uy_k1, uyc);
  if (iv && kv) h[1] = h_expr<OE>(P.c[7], h[1], P.c[10], P.ux_k1, P.uxc, P.uzn, uzc);
  if (iv && jv) h[2] = h_expr<OE>(P.c[8], h[2], P.c[11], P.uyn, uyc, P.ux_j1, P.uxc);
  if (mask)
    // T_p = u_other[node + e_p] - u_other[node] for each (component, normal).
    cpml_node<false>(slabs, mask, par, write_psi, i, j, k, P.uz_j1 - uzc, P.uy_k1 - uyc,
                     P.uzn - uzc, P.ux_k1 - P.uxc, P.uyn - uyc, P.ux_j1 - P.uxc, P.c[9],
                     P.c[10], P.c[11], h[0], h[1], h[2]);
}

// ---- fused step kernel ----

// Tile t of the fused step (k tiles fastest, then j tiles, then x chunks):
// the body shared by the single-model and the batched kernels.
template <bool OE, int TJ, bool PF, bool CP = true, typename CT = uint32_t>
__device__ __forceinline__ void fused_tile(const float *__restrict__ A, float *__restrict__ B,
                                           size_t N, const float *__restrict__ cpool,
                                           const CT *__restrict__ croff,
                                           const FSlab *__restrict__ slabs,
                                           const uint32_t *__restrict__ pmask, int nslab, int par,
                                           const DSlab *__restrict__ hout,
                                           const uint32_t *__restrict__ pmask_o, int nhout,
                                           int nx, int ny, int nz, int tk, int chunk, int ktiles,
                                           int jtiles, int t) {
  // Tiles that no slab reaches run an instantiation without the slab code
  // (CP false: 44 VGPRs instead of 73, eight waves per SIMD instead of six).
  if (!CP) nslab = 0, nhout = 0;
  const int kt = t % ktiles;
  const int t2 = t / ktiles;
  const int jt = t2 % jtiles;
  const int c = t2 / jtiles;
  const int tx = threadIdx.x;
  // One wave per tile row: the row (and so every row address) is
  // wave-uniform and lives in scalar registers.
  const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int k0 = kt * tk, j0 = jt * TJ;
  const int k = k0 - 1 + tx, j = j0 - 1 + ty;
  const bool row_ok = j >= 0 && j < ny;
  const bool inb = row_ok && tx <= tk && k >= 0 && k < nz;
  const bool own = inb && tx > 0 && ty > 0;
  // Addresses use the row and lane clamped into the grid, so every load is
  // valid; values of nodes outside it are not used.
  const int jc = min(max(j, 0), ny - 1);
  const unsigned ku = (unsigned)min(max(k, 0), nz - 1);
  const unsigned k4_ = ku * 4u;
  const int i0 = c * chunk, i1 = min(nx, i0 + chunk);
  const size_t plane = (size_t)ny * nz;
  const bool jv = jc < ny - 1, kv = (int)ku < nz - 1;
  // CPML slabs whose (j, k) box meets this tile (halo included).
  uint32_t mh = 0, me = 0;
  for (int q = 0; q < nslab; q++) {
    const FSlab s = cload(slabs, q);
    if (s.lo1 < j0 + TJ && s.hi1 > j0 - 1 && s.lo2 < k0 + tk && s.hi2 > k0 - 1) {
      if (s.e_side) me |= 1u << q;
      else mh |= 1u << q;
    }
  }
  // H-side slabs whose increments come from cpml_delta_h.
  uint32_t mo = 0;
  for (int q = 0; q < nhout; q++) {
    const DSlab s = cload(hout, q);
    if (s.lo1 < j0 + TJ && s.hi1 > j0 - 1 && s.lo2 < k0 + tk && s.hi2 > k0 - 1) mo |= 1u << q;
  }
  __shared__ float sh[2][3][TJ + 1][64];

  float uyc, uzc, hy_prev = 0.0f, hz_prev = 0.0f;
  PlaneIn P;
  {
    const size_t ob = ((size_t)i0 * ny + jc) * nz;
    const unsigned k4 = k4_;
    uyc = ldf(A + N + ob, k4);
    uzc = ldf(A + 2 * N + ob, k4);
    if (i0 > 0) {
      // H of plane i0-1 at this node (read-only: its owner writes it).
      PlaneIn Q;
      load_plane(Q, A, N, cpool, croff, ob - plane, (size_t)(i0 - 1) * ny + jc, plane, nz, k4,
                 true, jv, kv, false);
      const float uyp = ldf(A + N + (ob - plane), k4), uzp = ldf(A + 2 * N + (ob - plane), k4);
      if (inb) {
        float h[3];
        h_from<OE>(Q, uyp, uzp, true, jv, kv, slabs, mh & cload(pmask, 2 * (i0 - 1)), par, false, i0 - 1,
                   j, k, h);
        if (mo) h_out_apply(hout, mo & cload(pmask_o, i0 - 1), i0 - 1, j, k, h);
        hy_prev = h[1];
        hz_prev = h[2];
      }
    }
    load_plane(P, A, N, cpool, croff, ob, (size_t)i0 * ny + jc, plane, nz, k4, i0 + 1 < nx, jv,
               kv, ty > 0);
  }
  for (int i = i0; i < i1; i++) {
    const int buf = i & 1;
    const size_t r = (size_t)i * ny + jc;
    const size_t ob = r * nz;
    const bool iv = i + 1 < nx;
    // An opaque copy of the lane offset in each iteration, so the compiler
    // sees "scalar base + zero-extended 32-bit offset" at every access and
    // uses the saddr form (a hoisted 64-bit copy defeats that).
    unsigned k4 = k4_;
    asm volatile("" : "+v"(k4));
    PlaneIn Nx;
    if (PF && i + 1 < i1)
      load_plane(Nx, A, N, cpool, croff, ob + plane, r + ny, plane, nz, k4, i + 2 < nx, jv, kv,
                 ty > 0);
    // Slabs that reach this tile and this plane (bit masks, scalar).
    const uint32_t mhp = mh & cload(pmask, 2 * i), mep = me & cload(pmask, 2 * i + 1);
    float h[3] = {0.0f, 0.0f, 0.0f};
    if (inb) {
      h_from<OE>(P, uyc, uzc, iv, jv, kv, slabs, mhp, par, own, i, j, k, h);
      if (mo) h_out_apply(hout, mo & cload(pmask_o, i), i, j, k, h);
    }
    sh[buf][0][ty][tx] = h[0];
    sh[buf][1][ty][tx] = h[1];
    sh[buf][2][ty][tx] = h[2];
    __syncthreads();
    if (own) {
      const float hx_km = k > 0 ? sh[buf][0][ty][tx - 1] : 0.0f;
      const float hy_km = k > 0 ? sh[buf][1][ty][tx - 1] : 0.0f;
      const float hx_jm = j > 0 ? sh[buf][0][ty - 1][tx] : 0.0f;
      const float hz_jm = j > 0 ? sh[buf][2][ty - 1][tx] : 0.0f;
      float e0 = P.uxc, e1 = uyc, e2 = uzc;
      // curl_x = (iz[j] - iz[j-1]) - (iy[k] - iy[k-1])
      if (iv) e0 = e_expr<OE>(P.c[0], e0, P.c[3], h[2], hz_jm, h[1], hy_km);
      // curl_y = (ix[k] - ix[k-1]) - (iz[i] - iz[i-1])
      if (jv) e1 = e_expr<OE>(P.c[1], e1, P.c[4], h[0], hx_km, h[2], hz_prev);
      // curl_z = (iy[i] - iy[i-1]) - (ix[j] - ix[j-1])
      if (kv) e2 = e_expr<OE>(P.c[2], e2, P.c[5], h[1], hy_prev, h[0], hx_jm);
      if (mep)
        // S_p = i_other[node] - i_other[node - e_p] (cpml.rs:495-521).
        cpml_node<true>(slabs, mep, par, true, i, j, k, h[2] - hz_jm, h[1] - hy_km,
                        h[2] - hz_prev, h[0] - hx_km, h[1] - hy_prev, h[0] - hx_jm, P.c[3],
                        P.c[4], P.c[5], e0, e1, e2);
      stf(B + ob, k4, e0);
      stf(B + N + ob, k4, e1);
      stf(B + 2 * N + ob, k4, e2);
      stf(B + 3 * N + ob, k4, h[0]);
      stf(B + 4 * N + ob, k4, h[1]);
      stf(B + 5 * N + ob, k4, h[2]);
    }
    hy_prev = h[1];
    hz_prev = h[2];
    uyc = P.uyn;
    uzc = P.uzn;
    if (i + 1 < i1) {
      if (PF)
        P = Nx;
      else
        load_plane(P, A, N, cpool, croff, ob + plane, r + ny, plane, nz, k4, i + 2 < nx, jv, kv,
                   ty > 0);
    }
  }
}

// Block b -> tile. XCD-aware order: workgroups go round-robin to the 8
// XCDs, so give each XCD a contiguous range of tiles (k fastest, then j,
// then x chunk) and the halo rows its neighbours read stay in its own L2.
// Tiles come in rows of ktiles k-neighbours; rows are numbered
// R = c*jtiles + jt (j-neighbours adjacent). Groups of `xcd_group`
// consecutive rows go round-robin to the 8 XCDs, so k-neighbours and most
// j-neighbours share an L2 while every XCD gets a share of every x chunk.
// The grid is padded to whole groups of 8; padding blocks get -1.
__host__ __device__ __forceinline__ int tile_of_block(int b, int ktiles, int ntiles,
                                                      int xcd_group) {
  if (xcd_group <= 0) return b;
  const int x8 = b % 8, sl = b / 8;
  const int lr = sl / ktiles, ktl = sl % ktiles;
  const int R = ((lr / xcd_group) * 8 + x8) * xcd_group + lr % xcd_group;
  if (R >= ntiles / ktiles) return -1;
  return R * ktiles + ktl;
}

// tiles: the tile of each block (one class of them), or null for all tiles
// in tile_of_block's order. Without the prefetch the kernel is held to
// eight waves per SIMD: its 53 VGPRs allow them, and its scalar registers
// (106 otherwise: seven waves) spill into vector lanes instead (512^3 CPML
// box: 2.826 against 2.850 ms with 64 x chunks, 2.799 with 96).
template <bool OE, int TJ, bool PF, bool CP, typename CT>
__global__ void __launch_bounds__(64 * (TJ + 1), PF ? 1 : 8)
    step_fused(const float *__restrict__ A, float *__restrict__ B, size_t N,
               const float *__restrict__ cpool, const CT *__restrict__ croff,
               const FSlab *__restrict__ slabs, const uint32_t *__restrict__ pmask, int nslab,
               int par, const DSlab *__restrict__ hout, const uint32_t *__restrict__ pmask_o,
               int nhout, int nx, int ny, int nz, int tk, int chunk, int ktiles, int jtiles,
               int ntiles, int xcd_group, const int *__restrict__ tiles) {
  const int t = tiles ? cload(tiles, blockIdx.x) : tile_of_block(blockIdx.x, ktiles, ntiles, xcd_group);
  if (t < 0) return;
  fused_tile<OE, TJ, PF, CP, CT>(A, B, N, cpool, croff, slabs, pmask, nslab, par, hout, pmask_o,
                                 nhout, nx, ny, nz, tk, chunk, ktiles, jtiles, t);
}

// ---- fused step, bands of rows per wave ----
//
// A wave owns a k-window of 63 nodes (lane 0 is the halo node k0-1, whose
// H the window's first E update needs) and a band of TB consecutive rows
// j0..j0+TB-1, and marches a chunk of x-planes. Per plane it walks the
// band's rows in order: first the halo row j0-1 (H only), then each row
// (H, then E with the H of the row before it, of the lane before it and of
// the plane before it). So the j- and x-neighbours come from registers
// (carried from the previous row and the previous plane) and the
// k-neighbour from the lane below: no LDS, no barrier, waves independent.
// Only one halo row per TB rows, one halo lane per 63 and the plane before
// a chunk are computed twice (read-only; their owners write them).
// Same expressions on the same inputs as the other schedules: same bits.
template <bool OE, int TB>
__device__ __forceinline__ void band_tile(const float *__restrict__ A, float *__restrict__ B,
                                          size_t N, const float *__restrict__ cpool,
                                          const uint32_t *__restrict__ croff,
                                          const FSlab *__restrict__ slabs,
                                          const uint32_t *__restrict__ pmask, int nslab, int par,
                                          int nx, int ny, int nz, int chunk, int kw, int nbands,
                                          int t) {
  const int w = t % kw;
  const int t2 = t / kw;
  const int b = t2 % nbands;
  const int c = t2 / nbands;
  const int lane = threadIdx.x;
  const int k0 = 63 * w, k = k0 - 1 + lane;
  const bool kin = k >= 0 && k < nz;  // a node of the grid
  const bool own = kin && lane > 0;    // owned (E written)
  const unsigned ku = (unsigned)min(max(k, 0), nz - 1);
  const unsigned k4_ = ku * 4u;
  const bool kv = (int)ku < nz - 1;
  const int j0 = b * TB;
  const int i0 = c * chunk, i1 = min(nx, i0 + chunk);
  const size_t plane = (size_t)ny * nz;
  // CPML slabs whose (j, k) box meets this wave's rows (halo included).
  uint32_t mh = 0, me = 0;
  for (int q = 0; q < nslab; q++) {
    const FSlab s = cload(slabs, q);
    if (s.lo1 < j0 + TB && s.hi1 > j0 - 1 && s.lo2 < k0 + 63 && s.hi2 > k0 - 1) {
      if (s.e_side) me |= 1u << q;
      else mh |= 1u << q;
    }
  }
  float hyp[TB], hzp[TB], uyc[TB], uzc[TB];
#pragma unroll
  for (int r = 0; r < TB; r++) hyp[r] = hzp[r] = uyc[r] = uzc[r] = 0.0f;
#pragma unroll
  for (int r = 0; r < TB; r++) {
    const int j = j0 + r;
    if (j >= ny) break;
    unsigned k4 = k4_;
    asm volatile("" : "+v"(k4));
    const size_t ob = ((size_t)i0 * ny + j) * nz;
    uyc[r] = ldf(A + N + ob, k4);
    uzc[r] = ldf(A + 2 * N + ob, k4);
    if (i0 > 0) {
      // H of plane i0-1 (read-only: its owner writes it).
      const bool jv = j < ny - 1;
      PlaneIn Q;
      load_plane(Q, A, N, cpool, croff, ob - plane, (size_t)(i0 - 1) * ny + j, plane, nz, k4, true,
                 jv, kv, false);
      const float uyp = ldf(A + N + (ob - plane), k4), uzp = ldf(A + 2 * N + (ob - plane), k4);
      if (kin) {
        float h[3];
        h_from<OE>(Q, uyp, uzp, true, jv, kv, slabs, mh & cload(pmask, 2 * (i0 - 1)), par, false,
                   i0 - 1, j, k, h);
        hyp[r] = h[1];
        hzp[r] = h[2];
      }
    }
  }
  for (int i = i0; i < i1; i++) {
    const bool iv = i + 1 < nx;
    const uint32_t mhp = mh & cload(pmask, 2 * i), mep = me & cload(pmask, 2 * i + 1);
    // H of the halo row j0-1 (read-only): hx and hz feed row j0's E.
    float hxj = 0.0f, hzj = 0.0f;
    if (j0 > 0) {
      unsigned k4 = k4_;
      asm volatile("" : "+v"(k4));
      const int j = j0 - 1;
      const size_t r = (size_t)i * ny + j, ob = r * nz;
      PlaneIn Q;
      load_plane(Q, A, N, cpool, croff, ob, r, plane, nz, k4, iv, true, kv, false);
      const float uyh = ldf(A + N + ob, k4), uzh = ldf(A + 2 * N + ob, k4);
      if (kin) {
        float h[3];
        h_from<OE>(Q, uyh, uzh, iv, true, kv, slabs, mhp, par, false, i, j, k, h);
        hxj = h[0];
        hzj = h[2];
      }
    }
#pragma unroll
    for (int rr = 0; rr < TB; rr++) {
      const int j = j0 + rr;
      if (j >= ny) break;
      // An opaque copy of the lane offset per row: "scalar base + 32-bit
      // offset" addressing at every access.
      unsigned k4 = k4_;
      asm volatile("" : "+v"(k4));
      const bool jv = j < ny - 1;
      const size_t r = (size_t)i * ny + j, ob = r * nz;
      PlaneIn P;
      load_plane(P, A, N, cpool, croff, ob, r, plane, nz, k4, iv, jv, kv);
      float h[3] = {P.ix, P.iy, P.iz};
      if (kin) h_from<OE>(P, uyc[rr], uzc[rr], iv, jv, kv, slabs, mhp, par, own, i, j, k, h);
      // H of the lane below (k - 1); zero below k = 0.
      float hx_km = __shfl_up(h[0], 1), hy_km = __shfl_up(h[1], 1);
      if (k <= 0) hx_km = hy_km = 0.0f;
      if (own) {
        float e0 = P.uxc, e1 = uyc[rr], e2 = uzc[rr];
        // curl_x = (iz[j] - iz[j-1]) - (iy[k] - iy[k-1])
        if (iv) e0 = e_expr<OE>(P.c[0], e0, P.c[3], h[2], hzj, h[1], hy_km);
        // curl_y = (ix[k] - ix[k-1]) - (iz[i] - iz[i-1])
        if (jv) e1 = e_expr<OE>(P.c[1], e1, P.c[4], h[0], hx_km, h[2], hzp[rr]);
        // curl_z = (iy[i] - iy[i-1]) - (ix[j] - ix[j-1])
        if (kv) e2 = e_expr<OE>(P.c[2], e2, P.c[5], h[1], hyp[rr], h[0], hxj);
        if (mep)
          cpml_node<true>(slabs, mep, par, true, i, j, k, h[2] - hzj, h[1] - hy_km,
                          h[2] - hzp[rr], h[0] - hx_km, h[1] - hyp[rr], h[0] - hxj, P.c[3],
                          P.c[4], P.c[5], e0, e1, e2);
        stf(B + ob, k4, e0);
        stf(B + N + ob, k4, e1);
        stf(B + 2 * N + ob, k4, e2);
        stf(B + 3 * N + ob, k4, h[0]);
        stf(B + 4 * N + ob, k4, h[1]);
        stf(B + 5 * N + ob, k4, h[2]);
      }
      hxj = h[0];
      hzj = h[2];
      hyp[rr] = h[1];
      hzp[rr] = h[2];
      uyc[rr] = P.uyn;
      uzc[rr] = P.uzn;
    }
  }
}

template <bool OE, int TB>
__global__ void __launch_bounds__(64, 4)
    step_band(const float *__restrict__ A, float *__restrict__ B, size_t N,
              const float *__restrict__ cpool, const uint32_t *__restrict__ croff,
              const FSlab *__restrict__ slabs, const uint32_t *__restrict__ pmask, int nslab,
              int par, int nx, int ny, int nz, int chunk, int kw, int nbands, int ntiles,
              int xcd_group) {
  const int t = tile_of_block(blockIdx.x, kw, ntiles, xcd_group);
  if (t < 0) return;
  band_tile<OE, TB>(A, B, N, cpool, croff, slabs, pmask, nslab, par, nx, ny, nz, chunk, kw, nbands,
                    t);
}

// One model of a batched step: what step_fused takes as arguments.
struct FusedArgs {
  const float *A;
  float *B;
  size_t N;
  const float *cpool;
  const uint32_t *croff;
  const FSlab *slabs;
  const uint32_t *pmask;
  int nslab, par, nx, ny, nz, tk, chunk, ktiles, jtiles, ntiles;
};

// Many models of uneven sizes in one launch: seg[m] is the first block of
// model m (a multiple of 8, so the XCD order holds within each model),
// seg[nm] the grid size.
// (Held to eight waves per SIMD without the prefetch, as step_fused is.)
template <bool OE, int TJ, bool PF>
__global__ void __launch_bounds__(64 * (TJ + 1), PF ? 1 : 8)
    step_fused_batch(const FusedArgs *__restrict__ ms, const uint32_t *__restrict__ seg, int nm,
                     int xcd_group) {
  const int b = blockIdx.x;
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((int)cload(seg, mid) <= b) lo = mid;
    else hi = mid;
  }
  const FusedArgs a = cload(ms, lo);
  const int t = tile_of_block(b - (int)cload(seg, lo), a.ktiles, a.ntiles, xcd_group);
  if (t < 0) return;
  fused_tile<OE, TJ, PF>(a.A, a.B, a.N, a.cpool, a.croff, a.slabs, a.pmask, a.nslab, a.par,
                         nullptr, nullptr, 0, a.nx, a.ny, a.nz, a.tk, a.chunk, a.ktiles, a.jtiles,
                         t);
}

// ---- fused step, packed short rows ----
//
// For grids with nz <= 64: R = 64 / nz whole k-rows per wave (lane =
// jr*nz + k), so short rows fill the wave instead of leaving most lanes
// idle. A block of NW waves covers NW*R consecutive j rows, the first of
// them the halo row j0-1 whose H the E update of row j0 needs (recomputed
// read-only, as in the tiled kernel); rows are whole, so there is no k
// halo. Rows of one x-plane are contiguous, so a wave's fields are one
// contiguous run at a wave-uniform base. Same expressions and inputs as
// the other schedules, so the same bits.

// load_plane for a lane whose row differs from its wave's: fields at the
// wave base `ob` plus the lane's byte offset kf, coefficients through the
// lane's own row table entry (r) at byte offset kc within the row.
__device__ __forceinline__ void load_plane_lane(PlaneIn &P, const float *F, size_t N,
                                                const float *cpool, const uint32_t *croff,
                                                size_t ob, size_t r, size_t plane, int nz,
                                                unsigned kf, unsigned kc, bool iv, bool jv,
                                                bool kv) {
  asm volatile("" : "+v"(kf));
  const unsigned kn = kf + (kv ? 4u : 0u);
  const unsigned kj = kf + (jv ? 4u * (unsigned)nz : 0u);
  const unsigned ki = kf + (iv ? 4u * (unsigned)plane : 0u);
  const float *ux = F + ob, *uy = F + N + ob, *uz = F + 2 * N + ob;
  const float uxc = ldf(ux, kf), uxj = ldf(ux, kj), uzj = ldf(uz, kj);
  const float uxk = ldf(ux, kn), uyk = ldf(uy, kn);
  const float uyi = ldf(uy, ki), uzi = ldf(uz, ki);
  P.ix = ldf(F + 3 * N + ob, kf);
  P.iy = ldf(F + 4 * N + ob, kf);
  P.iz = ldf(F + 5 * N + ob, kf);
  const GAS uint32_t *ro = (const GAS uint32_t *)croff + 16 * r;
  const uint32_t cflag = ro[12];
#pragma unroll
  for (int l = 0; l < 12; l++) {
    // a constant row holds one value: read its first entry (same bits)
    const unsigned off = ro[l] + (((cflag >> l) & 1u) ? 0u : kc);
    P.c[l] = ldf(cpool, off);
  }
  P.uxc = uxc;
  P.ux_j1 = jv ? uxj : 0.0f;
  P.uz_j1 = jv ? uzj : 0.0f;
  P.ux_k1 = kv ? uxk : 0.0f;
  P.uy_k1 = kv ? uyk : 0.0f;
  P.uyn = iv ? uyi : 0.0f;
  P.uzn = iv ? uzi : 0.0f;
}

template <bool OE, int NW, bool PF>
__device__ __forceinline__ void fused_packed(const float *__restrict__ A, float *__restrict__ B,
                                             size_t N, const float *__restrict__ cpool,
                                             const uint32_t *__restrict__ croff,
                                             const FSlab *__restrict__ slabs,
                                             const uint32_t *__restrict__ pmask, int nslab,
                                             int par, int nx, int ny, int nz, int R, int chunk,
                                             int jtiles, int t) {
  const int jt = t % jtiles, c = t / jtiles;
  const int lane = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int owned = NW * R - 1;
  const int j0 = jt * owned;
  const int jr = lane / nz, k = lane - jr * nz;
  const bool lane_ok = jr < R;
  const int rb = w * R + jr;  // row within the block, 0 = the halo row
  const int j = j0 - 1 + rb;
  const bool inb = lane_ok && j >= 0 && j < ny;
  const bool own = inb && rb > 0;
  // The wave's first row (clamped) is the base of every field address; a
  // lane's row is clamped into [that row, ny - 1], so every load is valid.
  const int jw = min(max(j0 - 1 + w * R, 0), ny - 1);
  const int jc = min(max(j, jw), ny - 1);
  const int kk = lane_ok ? k : 0;
  const unsigned kf_ = (unsigned)(((jc - jw) * nz + kk) * 4);
  const unsigned kc = (unsigned)kk * 4u;
  const int i0 = c * chunk, i1 = min(nx, i0 + chunk);
  const size_t plane = (size_t)ny * nz;
  const bool jv = jc < ny - 1, kv = kk < nz - 1;
  // CPML slabs whose j range meets the block's rows (k is checked per lane).
  uint32_t mh = 0, me = 0;
  for (int q = 0; q < nslab; q++) {
    const FSlab s = cload(slabs, q);
    if (s.lo1 < j0 - 1 + NW * R && s.hi1 > j0 - 1) {
      if (s.e_side) me |= 1u << q;
      else mh |= 1u << q;
    }
  }
  __shared__ float sh[2][3][NW][64];

  float uyc, uzc, hy_prev = 0.0f, hz_prev = 0.0f;
  PlaneIn P;
  {
    const size_t obw = ((size_t)i0 * ny + jw) * nz;
    const size_t r = (size_t)i0 * ny + jc;
    const unsigned kf = kf_;
    uyc = ldf(A + N + obw, kf);
    uzc = ldf(A + 2 * N + obw, kf);
    if (i0 > 0) {
      // H of plane i0-1 at this node (read-only: its owner writes it).
      PlaneIn Q;
      load_plane_lane(Q, A, N, cpool, croff, obw - plane, r - ny, plane, nz, kf, kc, true, jv,
                      kv);
      const float uyp = ldf(A + N + (obw - plane), kf), uzp = ldf(A + 2 * N + (obw - plane), kf);
      if (inb) {
        float h[3];
        h_from<OE>(Q, uyp, uzp, true, jv, kv, slabs, mh & cload(pmask, 2 * (i0 - 1)), par, false, i0 - 1,
                   j, k, h);
        hy_prev = h[1];
        hz_prev = h[2];
      }
    }
    load_plane_lane(P, A, N, cpool, croff, obw, r, plane, nz, kf, kc, i0 + 1 < nx, jv, kv);
  }
  for (int i = i0; i < i1; i++) {
    const int buf = i & 1;
    const size_t r = (size_t)i * ny + jc;
    const size_t obw = ((size_t)i * ny + jw) * nz;
    const bool iv = i + 1 < nx;
    unsigned kf = kf_;
    asm volatile("" : "+v"(kf));
    PlaneIn Nx;
    if (PF && i + 1 < i1)
      load_plane_lane(Nx, A, N, cpool, croff, obw + plane, r + ny, plane, nz, kf, kc, i + 2 < nx,
                      jv, kv);
    const uint32_t mhp = mh & cload(pmask, 2 * i), mep = me & cload(pmask, 2 * i + 1);
    float h[3] = {0.0f, 0.0f, 0.0f};
    if (inb) h_from<OE>(P, uyc, uzc, iv, jv, kv, slabs, mhp, par, own, i, j, k, h);
    sh[buf][0][w][lane] = h[0];
    sh[buf][1][w][lane] = h[1];
    sh[buf][2][w][lane] = h[2];
    __syncthreads();
    if (own) {
      // k-1: the previous lane of the same row; j-1: the row before, in
      // this wave or at the end of the previous wave's rows.
      const int wp = jr > 0 ? w : w - 1;
      const int lp = jr > 0 ? lane - nz : (R - 1) * nz + k;
      const float hx_km = k > 0 ? sh[buf][0][w][lane - 1] : 0.0f;
      const float hy_km = k > 0 ? sh[buf][1][w][lane - 1] : 0.0f;
      const float hx_jm = j > 0 ? sh[buf][0][wp][lp] : 0.0f;
      const float hz_jm = j > 0 ? sh[buf][2][wp][lp] : 0.0f;
      float e0 = P.uxc, e1 = uyc, e2 = uzc;
      // curl_x = (iz[j] - iz[j-1]) - (iy[k] - iy[k-1])
      if (iv) e0 = e_expr<OE>(P.c[0], e0, P.c[3], h[2], hz_jm, h[1], hy_km);
      // curl_y = (ix[k] - ix[k-1]) - (iz[i] - iz[i-1])
      if (jv) e1 = e_expr<OE>(P.c[1], e1, P.c[4], h[0], hx_km, h[2], hz_prev);
      // curl_z = (iy[i] - iy[i-1]) - (ix[j] - ix[j-1])
      if (kv) e2 = e_expr<OE>(P.c[2], e2, P.c[5], h[1], hy_prev, h[0], hx_jm);
      if (mep)
        cpml_node<true>(slabs, mep, par, true, i, j, k, h[2] - hz_jm, h[1] - hy_km,
                        h[2] - hz_prev, h[0] - hx_km, h[1] - hy_prev, h[0] - hx_jm, P.c[3],
                        P.c[4], P.c[5], e0, e1, e2);
      stf(B + obw, kf, e0);
      stf(B + N + obw, kf, e1);
      stf(B + 2 * N + obw, kf, e2);
      stf(B + 3 * N + obw, kf, h[0]);
      stf(B + 4 * N + obw, kf, h[1]);
      stf(B + 5 * N + obw, kf, h[2]);
    }
    hy_prev = h[1];
    hz_prev = h[2];
    uyc = P.uyn;
    uzc = P.uzn;
    if (i + 1 < i1) {
      if (PF)
        P = Nx;
      else
        load_plane_lane(P, A, N, cpool, croff, obw + plane, r + ny, plane, nz, kf, kc, i + 2 < nx,
                        jv, kv);
    }
  }
}

// Without the prefetch held to eight waves per SIMD: 57 VGPRs allow them
// and the scalar registers spill into vector lanes (the 256-model patch
// batch, 16 x chunks: 1.201 ms per step against 1.331 at seven waves and
// 1.336 with the prefetch at six).
template <bool OE, int NW, bool PF>
__global__ void __launch_bounds__(64 * NW, PF ? 1 : 8)
    step_packed(const float *__restrict__ A, float *__restrict__ B, size_t N,
                const float *__restrict__ cpool, const uint32_t *__restrict__ croff,
                const FSlab *__restrict__ slabs, const uint32_t *__restrict__ pmask, int nslab,
                int par, int nx, int ny, int nz, int R, int chunk, int jtiles, int ntiles,
                int xcd_group) {
  const int t = tile_of_block(blockIdx.x, 1, ntiles, xcd_group);
  if (t < 0) return;
  fused_packed<OE, NW, PF>(A, B, N, cpool, croff, slabs, pmask, nslab, par, nx, ny, nz, R, chunk,
                           jtiles, t);
}

// The batched form: FusedArgs::tk holds R (rows per wave), ktiles is 1.
template <bool OE, int NW, bool PF>
__global__ void __launch_bounds__(64 * NW, PF ? 1 : 8)
    step_packed_batch(const FusedArgs *__restrict__ ms, const uint32_t *__restrict__ seg, int nm,
                      int xcd_group) {
  const int b = blockIdx.x;
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((int)cload(seg, mid) <= b) lo = mid;
    else hi = mid;
  }
  const FusedArgs a = cload(ms, lo);
  const int t = tile_of_block(b - (int)cload(seg, lo), 1, a.ntiles, xcd_group);
  if (t < 0) return;
  fused_packed<OE, NW, PF>(a.A, a.B, a.N, a.cpool, a.croff, a.slabs, a.pmask, a.nslab, a.par, a.nx,
                           a.ny, a.nz, a.tk, a.chunk, a.jtiles, t);
}

// ---- fused step, full-row variant ----
//
// One wave per j-row covering all k in NSEG segments of 64 lanes (k = 64m +
// lane), so a row is read and written contiguously by one wave and no cache
// line is shared between tiles along k. A block holds TJ rows plus the halo
// row j0-1 and marches a chunk of x-planes. Per plane: H of every segment of
// every row (halo row included, CPML-corrected) into LDS; barrier; E of the
// owned rows from LDS (k-1 and j-1 neighbours) and the previous plane's
// y/z currents kept in registers; barrier. Same expressions and inputs as
// the other schedules, so the same bits.
template <bool OE, int TJ, int NSEG>
__global__ void __launch_bounds__(64 * (TJ + 1))
    step_rows(const float *__restrict__ A, float *__restrict__ B, size_t N,
              const float *__restrict__ cpool, const uint32_t *__restrict__ croff,
              const FSlab *__restrict__ slabs, const uint32_t *__restrict__ pmask, int nslab,
              int par, int nx, int ny, int nz, int chunk, int jtiles, int ntiles, int xcd_group) {
  // Tiles are (x chunk, j tile), j fastest; groups of xcd_group consecutive
  // tiles go round-robin to the 8 XCDs (halo rows shared within a group).
  const int b = blockIdx.x;
  int t;
  if (xcd_group > 0) {
    const int x8 = b % 8, sl = b / 8;
    t = ((sl / xcd_group) * 8 + x8) * xcd_group + sl % xcd_group;
    if (t >= ntiles) return;
  } else {
    t = b;
  }
  const int jt = t % jtiles, c = t / jtiles;
  const int lane = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int j0 = jt * TJ, j = j0 - 1 + w;
  const bool row_ok = j >= 0 && j < ny;
  const bool own_row = row_ok && w > 0;
  const int jc = min(max(j, 0), ny - 1);
  const bool jv = jc < ny - 1;
  const int i0 = c * chunk, i1 = min(nx, i0 + chunk);
  const size_t plane = (size_t)ny * nz;
  constexpr int W = NSEG * 64;
  __shared__ float sh[3][TJ + 1][W];
  // CPML slabs whose j range meets the tile (k is checked per lane).
  uint32_t mh = 0, me = 0;
  for (int q = 0; q < nslab; q++) {
    const FSlab s = cload(slabs, q);
    if (s.lo1 < j0 + TJ && s.hi1 > j0 - 1) {
      if (s.e_side) me |= 1u << q;
      else mh |= 1u << q;
    }
  }
  float hy_prev[NSEG], hz_prev[NSEG];
#pragma unroll
  for (int m = 0; m < NSEG; m++) hy_prev[m] = hz_prev[m] = 0.0f;
  if (i0 > 0) {
    // H of plane i0-1 (read-only: its owner writes it).
    const int ip = i0 - 1;
    const size_t ob = ((size_t)ip * ny + jc) * nz, r = (size_t)ip * ny + jc;
    const uint32_t mhp = mh & cload(pmask, 2 * ip);
#pragma unroll
    for (int m = 0; m < NSEG; m++) {
      const int k = 64 * m + lane;
      const bool inb = row_ok && k < nz;
      const unsigned ku = (unsigned)min(k, nz - 1);
      const bool kv = (int)ku < nz - 1;
      PlaneIn Q;
      load_plane(Q, A, N, cpool, croff, ob, r, plane, nz, ku * 4u, true, jv, kv, false);
      const float uyc = ldf(A + N + ob, ku * 4u), uzc = ldf(A + 2 * N + ob, ku * 4u);
      if (inb) {
        float h[3];
        h_from<OE>(Q, uyc, uzc, true, jv, kv, slabs, mhp, par, false, ip, j, k, h);
        hy_prev[m] = h[1];
        hz_prev[m] = h[2];
      }
    }
  }
  for (int i = i0; i < i1; i++) {
    const size_t r = (size_t)i * ny + jc;
    const size_t ob = r * nz;
    const bool iv = i + 1 < nx;
    const uint32_t mhp = mh & cload(pmask, 2 * i), mep = me & cload(pmask, 2 * i + 1);
    // H of every segment of this row.
#pragma unroll
    for (int m = 0; m < NSEG; m++) {
      const int k = 64 * m + lane;
      const bool inb = row_ok && k < nz;
      const unsigned ku = (unsigned)min(k, nz - 1);
      const bool kv = (int)ku < nz - 1;
      PlaneIn P;
      load_plane(P, A, N, cpool, croff, ob, r, plane, nz, ku * 4u, iv, jv, kv);
      const float uyc = ldf(A + N + ob, ku * 4u), uzc = ldf(A + 2 * N + ob, ku * 4u);
      float h[3] = {0.0f, 0.0f, 0.0f};
      if (inb) h_from<OE>(P, uyc, uzc, iv, jv, kv, slabs, mhp, par, own_row, i, j, k, h);
      sh[0][w][64 * m + lane] = h[0];
      sh[1][w][64 * m + lane] = h[1];
      sh[2][w][64 * m + lane] = h[2];
      if (own_row && k < nz) {
        unsigned k4 = ku * 4u;
        asm volatile("" : "+v"(k4));
        stf(B + 3 * N + ob, k4, h[0]);
        stf(B + 4 * N + ob, k4, h[1]);
        stf(B + 5 * N + ob, k4, h[2]);
      }
    }
    __syncthreads();
    if (own_row) {
#pragma unroll
      for (int m = 0; m < NSEG; m++) {
        const int k = 64 * m + lane;
        const unsigned ku = (unsigned)min(k, nz - 1);
        unsigned k4 = ku * 4u;
        asm volatile("" : "+v"(k4));
        const float hx = sh[0][w][64 * m + lane], hy = sh[1][w][64 * m + lane];
        const float hz = sh[2][w][64 * m + lane];
        if (k < nz) {
          const bool kv = k < nz - 1;
          const float hx_km = k > 0 ? sh[0][w][64 * m + lane - 1] : 0.0f;
          const float hy_km = k > 0 ? sh[1][w][64 * m + lane - 1] : 0.0f;
          const float hx_jm = j > 0 ? sh[0][w - 1][64 * m + lane] : 0.0f;
          const float hz_jm = j > 0 ? sh[2][w - 1][64 * m + lane] : 0.0f;
          const uint32_t *ro = croff + 16 * r;
          float e0 = ldf(A + ob, k4), e1 = ldf(A + N + ob, k4), e2 = ldf(A + 2 * N + ob, k4);
          const float ca0 = ldc(cpool, ro, 0, k4), ca1 = ldc(cpool, ro, 1, k4);
          const float ca2 = ldc(cpool, ro, 2, k4), cb0 = ldc(cpool, ro, 3, k4);
          const float cb1 = ldc(cpool, ro, 4, k4), cb2 = ldc(cpool, ro, 5, k4);
          // curl_x = (iz[j] - iz[j-1]) - (iy[k] - iy[k-1])
          if (iv) e0 = e_expr<OE>(ca0, e0, cb0, hz, hz_jm, hy, hy_km);
          // curl_y = (ix[k] - ix[k-1]) - (iz[i] - iz[i-1])
          if (jv) e1 = e_expr<OE>(ca1, e1, cb1, hx, hx_km, hz, hz_prev[m]);
          // curl_z = (iy[i] - iy[i-1]) - (ix[j] - ix[j-1])
          if (kv) e2 = e_expr<OE>(ca2, e2, cb2, hy, hy_prev[m], hx, hx_jm);
          if (mep)
            cpml_node<true>(slabs, mep, par, true, i, j, k, hz - hz_jm, hy - hy_km,
                            hz - hz_prev[m], hx - hx_km, hy - hy_prev[m], hx - hx_jm, cb0, cb1,
                            cb2, e0, e1, e2);
          stf(B + ob, k4, e0);
          stf(B + N + ob, k4, e1);
          stf(B + 2 * N + ob, k4, e2);
        }
        hy_prev[m] = hy;
        hz_prev[m] = hz;
      }
    }
    __syncthreads();
  }
}

// flags[p] = 1 when pool row p holds one value (bit for bit) in all nz
// entries. A thread per row (set-up only).
__global__ void pool_row_const(const float *pool, size_t prows, int nz, uint8_t *flags) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= prows) return;
  const uint32_t *row = (const uint32_t *)(pool + p * (size_t)nz);
  const uint32_t v = row[0];
  uint8_t f = 1;
  for (int k = 1; k < nz; k++)
    if (row[k] != v) {
      f = 0;
      break;
    }
  flags[p] = f;
}

// croff[16*r + l]: byte offset in the pool of lane l's row for node row r;
// croff[16*r + 12]: bit l set when that row is constant.
template <typename CT>
__global__ void build_row_offsets(CT *croff, const uint32_t *const *index, const uint64_t *base,
                                  const uint8_t *flags, size_t rows, int nz) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  CT cm = 0;
  for (int l = 0; l < 16; l++) {
    CT v = 0;
    if (l < 12) {
      const uint32_t *ix = index[l];
      // Lanes without rows point at the pool's zero row (base[l] == 0).
      uint64_t prow = 0;
      if (base[l] != 0) prow = base[l] / (uint64_t)nz + (ix ? (uint64_t)ix[r] : (uint64_t)r);
      v = (CT)(4ull * prow * (uint64_t)nz);
      cm |= (CT)flags[prow] << l;
    }
    croff[16 * r + l] = l == 12 ? cm : v;
  }
  croff[16 * r + 12] = cm;
}

// Plain device copy (read n4 float4, write n4 float4): the stream-bandwidth
// reference for the roofline and the rocprof byte counters.
__global__ void __launch_bounds__(256) copy_f4(const float4 *__restrict__ a, float4 *__restrict__ b,
                                               size_t n4) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4;
       q += (size_t)gridDim.x * blockDim.x)
    b[q] = a[q];
}

// Serial operators (excite.rs, probe.rs, lumped.rs) and the gap rules of
// Engine::advance (mod.rs:873-915).
enum {
  OP_SOURCE = 1, OP_PROBE = 2, OP_LUMPED = 3, OP_POLES = 4, OP_PW = 5, OP_MERGED = 6, OP_MUR = 7,
  OP_UPML = 8
};

// An installed lumped element (lumped.rs:367-406) and, with columns > 0,
// its LumpedPort records (engine/port.rs:94-104).
struct DevLumped {
  int axis, n, series, src_in_branch, branch, nterm, ndrive, columns, record;
  double dt;
  const uint64_t *idx;
  const double *cb, *gsrc, *tw;
  const double *const *tv;  // per term: value by step
  const uint64_t *tlen;
  double *v, *vd;           // per-edge source / drive voltage of the step
  const double *two_l_dt, *inv_c, *z;
  double *j, *q, *javg, *uprev;
  const double *rg, *rc;
  double *rec, *pv, *pi;
  uint64_t *count;
};

struct DevOp {
  int type, kind, quantity, n;
  int f32sum;  // probes: the f32 sum of openEMS's current probes
  uint64_t interval;
  const int32_t *axis;
  const uint64_t *idx;
  const float *wf;
  const double *wd;
  const float *amp;
  uint64_t namp;
  double *vals;
  uint64_t *count;
  DevLumped *lp;
  int seg_end[2];  // per gap: the first writer at or after this op (nops: none)
};

// ---- Mur absorbing boundary (boundary/mur.rs) ----
//
// Face edges with one 1-D update each, then the corners (edges on two Mur
// faces), which read face edges already updated. u_b^n and u_1^n are saved
// before the E update (mur_capture; the H update does not change u), the
// updates run after the E update's slab corrections (mur_edges,
// mur_corners). No face edge's interior neighbour is a Mur edge, so the
// edges of one launch are independent.
struct DevMur {
  int ne, nc, oe_order, oe_corners;
  const int32_t *eax;
  const uint64_t *enode, *einner, *ehold;
  const float *ek;
  const int32_t *cax;
  const uint64_t *cnode, *cin0, *cin1, *chold;
  const float *ck0, *ck1;
  float *saved;  // 2 per edge (u_b^n, u_1^n), then 3 per corner
};

// mur_1d (mur.rs:313-320): ems's order or openEMS's.
__device__ __forceinline__ float mur_1d(float i, float k, float x, float b, int oe) {
  return oe ? (i - k * b) + k * x : i + k * (x - b);
}

__device__ __forceinline__ float *pick_u(const Fields &f, int a) {
  return a == 0 ? f.u[0] : (a == 1 ? f.u[1] : f.u[2]);
}

// pre_e / post_h_slab (mur.rs:358-373, 806-827): the saved values.
__global__ void __launch_bounds__(256) mur_capture(Fields f, DevMur m) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < m.ne) {
    const GAS float *u = (const GAS float *)pick_u(f, m.eax[q]);
    m.saved[2 * q] = u[m.enode[q]];
    m.saved[2 * q + 1] = u[m.einner[q]];
  } else if (q < m.ne + m.nc) {
    const int c = q - m.ne;
    const GAS float *u = (const GAS float *)pick_u(f, m.cax[c]);
    float *sv = m.saved + 2 * (size_t)m.ne + 3 * (size_t)c;
    sv[0] = u[m.cnode[c]], sv[1] = u[m.cin0[c]], sv[2] = u[m.cin1[c]];
  }
}

// update (mur.rs:322-355), face edges.
__global__ void __launch_bounds__(256) mur_edges(Fields f, DevMur m, const uint64_t *step) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= m.ne || *step < m.ehold[q]) return;
  GAS float *u = (GAS float *)pick_u(f, m.eax[q]);
  u[m.enode[q]] = mur_1d(m.saved[2 * q + 1], m.ek[q], u[m.einner[q]], m.saved[2 * q], m.oe_order);
}

// update (mur.rs:322-355), corners: the average of the two 1-D updates, or
// openEMS's rule (the lower axis's face alone, its new neighbour taken as 0).
__global__ void __launch_bounds__(256) mur_corners(Fields f, DevMur m, const uint64_t *step) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= m.nc || *step < m.chold[c]) return;
  GAS float *u = (GAS float *)pick_u(f, m.cax[c]);
  const float *sv = m.saved + 2 * (size_t)m.ne + 3 * (size_t)c;
  const float b = sv[0], i0 = sv[1], i1 = sv[2];
  float v;
  if (m.oe_corners) {
    v = i0 - m.ck0[c] * b;
  } else {
    const float v0 = mur_1d(i0, m.ck0[c], u[m.cin0[c]], b, m.oe_order);
    const float v1 = mur_1d(i1, m.ck1[c], u[m.cin1[c]], b, m.oe_order);
    v = 0.5f * (v0 + v1);
  }
  u[m.cnode[c]] = v;
}

// ---- openEMS's PML_n layers (boundary/upml.rs) ----
//
// The layer edges carry coefficients rewritten by the host (ca = 0 and
// cb = dt/((A/l)(1 + a_p)), da and db likewise), so the H and E updates
// leave the flux increment t in the field. Edges are grouped by side and
// axis (off[g] is the first edge of group g = 3 side + axis: E x, y, z,
// then H x, y, z); every edge is independent of the others.
struct DevUpml {
  int ne, nh;
  int off[7];
  const uint64_t *idx;  // node index per edge
  const float4 *co;     // (1-a_p)/(1+a_p), (1-a_q)/(1+a_q), (1+a_a)/((1+a_q)m), (1-a_a)/((1+a_q)m)
  float2 *fv;           // flux F, start value v
};

__device__ __forceinline__ float *pick_i(const Fields &f, int a) {
  return a == 0 ? f.i[0] : (a == 1 ? f.i[1] : f.i[2]);
}

// Parts::post_h_slab / post_e_slab (upml.rs:442-470). stage 0 (after the H
// update): E edges take v from u (Run::refresh, :391-400), H edges complete
// (Run::apply, :404-426); stage 1 (after the E update) the other way round.
__global__ void __launch_bounds__(256) upml_stage(Fields f, DevUpml m, int stage) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= m.ne + m.nh) return;
  int g = 0;
  while (g < 5 && q >= m.off[g + 1]) g++;
  const bool eside = g < 3;
  const int a = eside ? g : g - 3;
  GAS float *x = (GAS float *)(eside ? pick_u(f, a) : pick_i(f, a));
  const uint64_t n = m.idx[q];
  float2 s = m.fv[q];
  if (eside == (stage == 0)) {
    s.y = x[n];
  } else {
    const float4 c = m.co[q];
    const float f1 = flush(flush(c.x * s.x) + x[n]);
    const float v = flush(flush(flush(c.y * s.y) - flush(c.w * s.x)) + flush(c.z * f1));
    s.x = f1, s.y = v;
    x[n] = v;
  }
  m.fv[q] = s;
}

__device__ __forceinline__ float op_amp(const DevOp &o, uint64_t n) {
  return n < o.namp ? o.amp[n] : 0.0f;
}
__device__ __forceinline__ bool is_hsrc(const DevOp &o) {
  return o.type == OP_SOURCE && (o.kind == 2 || o.kind == 3);
}
__device__ __forceinline__ bool is_esrc(const DevOp &o) {
  return o.type == OP_SOURCE && (o.kind == 0 || o.kind == 1);
}
__device__ __forceinline__ bool is_iprobe(const DevOp &o) {
  return o.type == OP_PROBE && o.quantity == 1;
}
__device__ __forceinline__ bool is_vprobe(const DevOp &o) {
  return o.type == OP_PROBE && o.quantity == 0;
}
__device__ __forceinline__ bool is_lumped(const DevOp &o) { return o.type == OP_LUMPED; }
__device__ __forceinline__ bool needs(const DevOp &o, uint64_t n) {
  if (o.type == OP_SOURCE) {
    if (o.kind == 1 || o.kind == 3) return true;
    return op_amp(o, n) != 0.0f;
  }
  if (o.type == OP_LUMPED) return true;  // post_e every step
  if (o.type == OP_POLES) return false;  // slab hooks, not gap hooks
  if (o.type == OP_MUR) return false;    // its own kernels before the gap
  if (o.type == OP_UPML) return false;   // slab corrections, not gap hooks
  if (o.type == OP_PW || o.type == OP_MERGED) return true;  // post_h and post_e every step
  return n % o.interval == 0;
}

__device__ void source_apply(Fields f, const DevOp &o, uint64_t n) {
  const float a = op_amp(o, n);
  float *const *fld = (o.kind == 2 || o.kind == 3) ? f.i : f.u;
  const bool hard = o.kind == 1 || o.kind == 3;
  for (int q = 0; q < o.n; q++) {
    float *x = fld[o.axis[q]] + o.idx[q];
    if (hard)
      *x = o.wf[q] * a;
    else
      *x += o.wf[q] * a;
  }
}

__device__ void probe_sample(Fields f, const DevOp &o) {
  float *const *fld = o.quantity == 1 ? f.i : f.u;
  double v = -0.0;
  if (o.f32sum) {
    // Probe::f32_sum (probe.rs:398-404): (w as f32) * x summed in f32 from 0.
    float s = 0.0f;
    for (int q = 0; q < o.n; q++) s = s + (float)o.wd[q] * fld[o.axis[q]][o.idx[q]];
    v = (double)s;
  } else {
    for (int q = 0; q < o.n; q++) v += o.wd[q] * (double)fld[o.axis[q]][o.idx[q]];
  }
  uint64_t c = *o.count;
  o.vals[c] = v;
  *o.count = c + 1;
}

// LumpedElement::pre_e_impl (lumped.rs:1106-1121): u^n of each edge, read
// from `f` (the fields before the E update), a thread per edge.
__device__ void lumped_capture(Fields f, const DevOp *ops, int nops) {
  for (int p = 0; p < nops; p++) {
    if (!is_lumped(ops[p])) continue;
    const DevLumped &L = *ops[p].lp;
    if (!L.branch && !L.record) continue;
    const float *u = f.u[L.axis];
    for (int q = threadIdx.x; q < L.n; q += blockDim.x) L.uprev[q] = (double)u[L.idx[q]];
  }
}

__device__ __forceinline__ double lumped_value(const DevLumped &L, int t, uint64_t n) {
  return n < L.tlen[t] ? L.tv[t][n] : 0.0;
}

// LumpedElement::post_e_impl (lumped.rs:1123-1235), Installed::add_sources
// (:1240-1249) and LumpedPort::post_e (port.rs:281-297), one thread: the
// record sums run over the edges in order.
__device__ void lumped_post_e(Fields f, const DevLumped &L, uint64_t n) {
  float *u = f.u[L.axis];
  const int ne = L.n;
  bool active = false, drive = false;
  if (L.nterm) {
    for (int q = 0; q < ne; q++) L.v[q] = 0.0;
    for (int t = 0; t < L.nterm; t++) {
      const double val = lumped_value(L, t, n);
      if (val != 0.0) {
        active = true;
        for (int q = 0; q < ne; q++) L.v[q] += L.tw[(size_t)t * ne + q] * val;
      }
    }
  }
  if (L.ndrive) {
    for (int q = 0; q < ne; q++) L.vd[q] = 0.0;
    for (int t = L.nterm; t < L.nterm + L.ndrive; t++) {
      const double val = lumped_value(L, t, n);
      if (val != 0.0) {
        drive = true;
        for (int q = 0; q < ne; q++) L.vd[q] += L.tw[(size_t)t * ne + q] * val;
      }
    }
  }
  if (active && !L.src_in_branch)
    for (int q = 0; q < ne; q++) {
      const double add = L.cb[q] * L.gsrc[q] * L.v[q];
      if (add != 0.0) u[L.idx[q]] = (float)((double)u[L.idx[q]] + add);
    }
  if (L.branch) {
    for (int q = 0; q < ne; q++) {
      double v_br = active && L.src_in_branch ? L.v[q] : 0.0;
      if (drive) v_br += L.vd[q];
      const double u_star = (double)u[L.idx[q]];
      if (L.branch == 2) {
        L.j[q] += 2.0 * (L.uprev[q] - v_br) / L.two_l_dt[q];
        L.javg[q] = L.j[q];
        u[L.idx[q]] = (float)(u_star - L.cb[q] * L.j[q]);
        continue;
      }
      const double k = L.two_l_dt[q] * L.j[q] - L.q[q] * L.inv_c[q] - v_br + 0.5 * L.uprev[q];
      const double z = L.z[q];
      const double u_new = (z * u_star - L.cb[q] * k) / (z + 0.5 * L.cb[q]);
      const double jb = (k + 0.5 * u_new) / z;
      L.j[q] = 2.0 * jb - L.j[q];
      L.q[q] += L.dt * jb;
      L.javg[q] = jb;
      u[L.idx[q]] = (float)u_new;
    }
  }
  if (L.record) {
    double sum = 0.0;
    for (int q = 0; q < ne; q++) {
      const double u0 = L.uprev[q], u1 = (double)u[L.idx[q]];
      const double v = active && !L.src_in_branch ? L.v[q] : 0.0;
      sum += L.rg[q] * (0.5 * (u0 + u1) - v) + L.rc[q] * (u1 - u0) / L.dt;
      if (L.branch) sum += L.javg[q];
    }
    const double cur = sum / (double)L.series;
    const uint64_t c = *L.count;
    L.rec[c] = cur;
    if (L.columns) {
      double sv = -0.0;
      for (int q = 0; q < ne; q++) sv += (double)u[L.idx[q]];
      L.pi[c] = -cur;
      L.pv[c] = sv / (double)L.columns;
    }
    *L.count = c + 1;
  }
}

struct Gaps {
  bool gap_b, moved, gap_a;
};
// Evaluated by the whole block (every thread must call it): each thread
// looks at a share of the operators, and the flags are or-ed over the
// block, so the latency is one operator's, not a chain over all of them.
__device__ Gaps gaps_at(const DevOp *ops, int nops, uint64_t n) {
  bool need_post_e = false, mid = false, fixed = false;
  for (int q = threadIdx.x; q < nops; q += blockDim.x) {
    const DevOp &o = ops[q];
    if (!needs(o, n)) continue;
    const bool own = o.type == OP_PW || o.type == OP_MERGED;
    if (is_esrc(o) || is_vprobe(o) || is_lumped(o) || own) need_post_e = true;
    if (is_hsrc(o) || own) mid = true, fixed = true;
    if (is_iprobe(o) || is_lumped(o)) mid = true;  // lumped pre_e is movable
  }
  need_post_e = __syncthreads_or(need_post_e);
  mid = __syncthreads_or(mid);
  fixed = __syncthreads_or(fixed);
  Gaps g;
  g.gap_b = mid && fixed;
  g.moved = mid && !fixed;
  g.gap_a = need_post_e || g.moved;
  return g;
}

// The hooks of one gap in registration order. Probes only read the fields,
// so a run of consecutive probes samples in parallel (a thread each);
// writers (sources, lumped elements) run one after another, with a barrier
// on either side. `which` selects the ops of the gap: 0 the mid gap (H
// sources, current probes), 1 the post-E gap (E sources, lumped elements,
// voltage probes).
__device__ __forceinline__ bool gap_writer(const DevOp &o, int which) {
  return which == 0 ? is_hsrc(o) : (is_esrc(o) || is_lumped(o));
}
__device__ void run_gap(Fields f, const DevOp *ops, int q0, int nops, uint64_t n, int which) {
  int q = q0;
  while (q < nops) {
    const DevOp &o = ops[q];
    if (gap_writer(o, which)) {
      if (threadIdx.x == 0) {
        if (is_lumped(o))
          lumped_post_e(f, *o.lp, n);
        else
          source_apply(f, o, n);
      }
      __syncthreads();
      q++;
      continue;
    }
    const int q1 = min(o.seg_end[which], nops);
    for (int p = q + (int)threadIdx.x; p < q1; p += (int)blockDim.x) {
      const DevOp &pr = ops[p];
      const bool mine = which == 0 ? is_iprobe(pr) : is_vprobe(pr);
      if (mine && n % pr.interval == 0) probe_sample(f, pr);
    }
    __syncthreads();
    q = q1;
  }
}

// Unfused steps: the gap between the H and E updates (the fields are
// updated in place, so pre_e captures u^n here). The gap's ops run in
// registration order; a plane wave in the list (its own kernels) splits
// it into segments [q0, q1), the first of which also captures u^n.
__global__ void __launch_bounds__(256) gap_mid(Fields f, const DevOp *ops, int nops, int q0,
                                               int q1, const uint64_t *step) {
  const uint64_t n = *step;
  const Gaps g = gaps_at(ops, nops, n);
  if (g.gap_b) run_gap(f, ops, q0, q1, n, 0);
  if (q0 == 0) lumped_capture(f, ops, nops);
}

// The post-E gap on the step result `f`. Fused steps keep the state before
// the step in `fp` (the other buffer), whose u is u^n as pre_e reads it;
// capture = 0 when the mid gap has already captured it. Segments as in
// the mid gap; the last one advances the step counter.
__device__ __forceinline__ void gap_post_e_body(Fields f, Fields fp, int capture, const DevOp *ops,
                                                int nops, int q0, int q1, int last,
                                                uint64_t *step) {
  const uint64_t n = *step;
  const Gaps g = gaps_at(ops, nops, n);
  if (capture) {
    lumped_capture(fp, ops, nops);
    __syncthreads();
  }
  if (g.gap_a) {
    if (g.moved && q0 == 0) {
      // Moved post_h hooks (current probes) first (mod.rs:936-942).
      for (int p = threadIdx.x; p < nops; p += blockDim.x)
        if (is_iprobe(ops[p]) && n % ops[p].interval == 0) probe_sample(f, ops[p]);
      __syncthreads();
    }
    run_gap(f, ops, q0, q1, n, 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) *step = n + 1;
}

__global__ void __launch_bounds__(256) gap_post_e(Fields f, Fields fp, int capture,
                                                  const DevOp *ops, int nops, int q0, int q1,
                                                  int last, uint64_t *step) {
  gap_post_e_body(f, fp, capture, ops, nops, q0, q1, last, step);
}

// The post-E gaps of a batch of fused models, a block per model.
struct GapArgs {
  Fields f, fp;
  const DevOp *ops;
  int nops;
  uint64_t *step;
};
__global__ void __launch_bounds__(256) gap_post_e_batch(const GapArgs *__restrict__ gs) {
  const GapArgs g = cload(gs, blockIdx.x);
  gap_post_e_body(g.f, g.fp, 1, g.ops, g.nops, 0, g.nops, 1, g.step);
}

// ---- plane wave, total-field / scattered-field (planewave.rs) ----

// Corrections grouped by the edge they correct (in list order within a
// group): field[axis][idx] += f32(coef * sum_q w_q aux[m + q]).
struct DevPwCorr {
  int ntgt;
  const uint32_t *start;  // ntgt + 1
  const uint64_t *idx;
  const int32_t *axis;
  const double *coef, *w;  // w: 4 per correction
  const uint32_t *m;
};
struct DevPw {
  int naux;
  double *e, *h;  // auxiliary grid: naux E nodes, naux - 1 H nodes
  const double *ca, *cb, *da, *db;
  const double *src;  // e[0] after the E update of step n
  uint64_t nsrc;
};

// PlaneWave::post_h / post_e corrections (planewave.rs:1513-1545), a
// thread per corrected edge; interp is the 4-point Lagrange sum (:653-655).
template <bool ESIDE>
__global__ void __launch_bounds__(256) pw_correct(Fields f, DevPwCorr c, const double *aux) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.ntgt) return;
  const uint32_t q0 = c.start[t], q1 = c.start[t + 1];
  float *x = (ESIDE ? f.u : f.i)[c.axis[q0]] + c.idx[t];
  float v = *x;
  for (uint32_t q = q0; q < q1; q++) {
    const double *w = c.w + 4 * (size_t)q;
    const uint32_t m = c.m[q];
    const double it = w[0] * aux[m] + w[1] * aux[m + 1] + w[2] * aux[m + 2] + w[3] * aux[m + 3];
    v += (float)(c.coef[q] * it);
  }
  *x = v;
}

// The auxiliary H update (planewave.rs:1525-1528), a thread per node.
__global__ void __launch_bounds__(256) pw_aux_h(DevPw p) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= p.naux - 1) return;
  p.h[m] = p.da[m] * p.h[m] - p.db[m] * (p.e[m + 1] - p.e[m]);
}

// The auxiliary E update and the hard source at node 0
// (planewave.rs:1553-1559): e[0] = gain * a((n + 1)dt + lead), tabulated.
__global__ void __launch_bounds__(256) pw_aux_e(DevPw p, const uint64_t *step) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= p.naux - 1) return;
  if (m == 0) {
    const uint64_t n = *step;
    p.e[0] = n < p.nsrc ? p.src[n] : 0.0;
    return;
  }
  p.e[m] = p.ca[m] * p.e[m] - p.cb[m] * (p.h[m] - p.h[m - 1]);
}

// ---- reduced currents of small cut faces (MergedFaces, conformal.rs:2313-2700) ----

// Flat tables of the operator (ems_merged_desc): circulation terms per
// reduced current, blocks (rows in elimination order, envelope factor and
// B rows, columns block-local), driven faces; and the state: g^{n-1/2}
// (cur), the update's result (next), the circulations (circ) and the
// blocks' right-hand sides (x, one per block row).
struct DevMerged {
  int ng, ns, nb, nf;
  double dt;
  const double *da, *db;
  const uint32_t *ts;
  const int32_t *tax;
  const uint64_t *tidx;
  const double *tw;
  const uint32_t *bs, *mem, *ef;
  const uint64_t *eo;
  const double *ev;
  const uint32_t *brs, *bc;
  const double *bv;
  const uint32_t *cs, *cr;
  const double *cv;
  int np;
  const int32_t *pax;
  const uint64_t *pi0, *pi1;
  const double *pc;
  const uint32_t *pst, *pt;
  const double *pw;
  double *psi, *pcon;
  const int32_t *fax;
  const uint64_t *fidx;
  const uint32_t *fs, *fg;
  const double *fw;
  double *cur, *next, *circ, *x;
};

__host__ __device__ __forceinline__ float flush_hd(float x) {
  return fabsf(x) < FLT_MIN ? 0.0f : x;
}

// MergedFaces::advance for merge group q (conformal.rs:2570-2572).
__host__ __device__ __forceinline__ void mf_scalar(const DevMerged &m, int q, const double *cur,
                                                   const double *c, double *out) {
  out[q] = (double)flush_hd((float)(m.da[q] * cur[q] - m.db[q] * c[q]));
}

// Block::advance (conformal.rs:2037-2052), its right-hand side row i of the
// block whose rows start at r0: x_i = (B g)_i - dt c (the row's terms in
// order, folding from -0.0).
__host__ __device__ __forceinline__ double mf_rhs(const DevMerged &m, uint32_t r0, uint32_t i,
                                                  const double *cur, const double *c) {
  double s = -0.0;
  for (uint32_t e = m.brs[r0 + i]; e < m.brs[r0 + i + 1]; e++)
    s += m.bv[e] * cur[m.mem[r0 + m.bc[e]]];
  return s - m.dt * c[m.mem[r0 + i]];
}

// Envelope::solve (conformal.rs:2101-2123) of the block's factor on its
// nr rows, in place: L y = x forward, then L^T x' = y backward.
__host__ __device__ inline void mf_solve(const DevMerged &m, uint32_t r0, uint32_t nr,
                                         double *x) {
  for (uint32_t i = 0; i < nr; i++) {
    const uint32_t f = m.ef[r0 + i];
    const double *row = m.ev + m.eo[r0 + i];
    double s = x[i];
    for (uint32_t j = f; j < i; j++) s -= row[j - f] * x[j];
    x[i] = s / row[i - f];
  }
  for (uint32_t i = nr; i-- > 0;) {
    const uint32_t f = m.ef[r0 + i];
    const double *row = m.ev + m.eo[r0 + i];
    x[i] /= row[i - f];
    const double xi = x[i];
    for (uint32_t j = f; j < i; j++) x[j] -= row[j - f] * xi;
  }
}

// MergedFaces::circulations (conformal.rs:2541-2566): a thread per reduced
// current, its terms in order, the sum folding from -0.0.
__global__ void __launch_bounds__(256) mf_circ(Fields f, DevMerged m) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m.ng) return;
  double c = -0.0;
  for (uint32_t t = m.ts[g]; t < m.ts[g + 1]; t++)
    c += m.tw[t] * (double)f.u[m.tax[t]][m.tidx[t]];
  m.circ[g] = c;
}

// LDS bytes of a block's image in mf_advance: x, its envelope values and
// columns (f64), per row the envelope's first column, the row and column
// offsets, and the columns' rows (u32).
__host__ __device__ __forceinline__ size_t mf_lds_need(uint32_t nr, uint64_t nev, uint32_t ncol) {
  return 8 * ((size_t)nr + nev + ncol) + 4 * (3 * (size_t)nr + 2 + ncol);
}

// LDS bytes of a narrow block's image (envelopes at most 3 below the
// diagonal): per row 4 envelope values, 3 column values, x, a packed word.
__host__ __device__ __forceinline__ size_t mf_narrow_need(uint32_t nr) {
  return 68 * (size_t)nr;
}

// One narrow block by a wave: the lanes stage each row's envelope (len =
// i - f + 1 <= 4 values), its column (up to 3 rows k = j + dk, decreasing)
// and right-hand side; lane 0 then substitutes with the last three solved
// values in registers and the next row's data loaded one row ahead, the
// terms in the row form's order (ascending j forward, decreasing k
// backward).
__device__ __forceinline__ void mf_narrow(const DevMerged &m, uint32_t r0, uint32_t nr,
                                          double *xs) {
  double *rv = xs, *cw = rv + 4 * (size_t)nr, *x = cw + 3 * (size_t)nr;
  uint32_t *meta = (uint32_t *)(x + nr);
  for (uint32_t i = threadIdx.x; i < nr; i += 64) {
    const uint32_t f = m.ef[r0 + i], len = i - f + 1;
    const uint64_t e = m.eo[r0 + i];
    for (uint32_t t = 0; t < len; t++) rv[4 * (size_t)i + t] = m.ev[e + t];
    const uint32_t q0 = m.cs[r0 + i], nc = m.cs[r0 + i + 1] - q0;
    uint32_t w = len | (nc << 4);
    for (uint32_t t = 0; t < nc; t++) {
      cw[3 * (size_t)i + t] = m.cv[q0 + t];
      w |= (m.cr[q0 + t] - i) << (8 + 2 * t);
    }
    meta[i] = w;
    x[i] = mf_rhs(m, r0, i, m.cur, m.circ);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double w1 = 0.0, w2 = 0.0, w3 = 0.0;
    uint32_t mt = meta[0];
    double a0 = rv[0], a1 = rv[1], a2 = rv[2], a3 = rv[3], y = x[0];
    for (uint32_t i = 0; i < nr; i++) {
      const uint32_t len = mt & 7;
      const double b0 = a0, b1 = a1, b2 = a2, b3 = a3;
      double s = y;
      if (i + 1 < nr) {
        const double *r = rv + 4 * (size_t)(i + 1);
        mt = meta[i + 1], a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3], y = x[i + 1];
      }
      double d;
      if (len == 4) {
        s -= b0 * w3;
        s -= b1 * w2;
        s -= b2 * w1;
        d = b3;
      } else if (len == 3) {
        s -= b0 * w2;
        s -= b1 * w1;
        d = b2;
      } else if (len == 2) {
        s -= b0 * w1;
        d = b1;
      } else {
        d = b0;
      }
      const double xi = s / d;
      x[i] = xi;
      w3 = w2, w2 = w1, w1 = xi;
    }
    // backward: w1, w2, w3 = x[j + 1], x[j + 2], x[j + 3]
    w1 = w2 = w3 = 0.0;
    const uint32_t jl = nr - 1;
    mt = meta[jl];
    double c0 = cw[3 * (size_t)jl], c1 = cw[3 * (size_t)jl + 1], c2 = cw[3 * (size_t)jl + 2];
    double dg = rv[4 * (size_t)jl + (mt & 7) - 1];
    y = x[jl];
    for (uint32_t j = nr; j-- > 0;) {
      const uint32_t nc = (mt >> 4) & 3, dk = mt >> 8;
      const double e0 = c0, e1 = c1, e2 = c2, dj = dg;
      double s = y;
      if (j > 0) {
        const uint32_t jn = j - 1;
        mt = meta[jn];
        c0 = cw[3 * (size_t)jn], c1 = cw[3 * (size_t)jn + 1], c2 = cw[3 * (size_t)jn + 2];
        dg = rv[4 * (size_t)jn + (mt & 7) - 1];
        y = x[jn];
      }
      const double ev[3] = {e0, e1, e2};
      for (uint32_t t = 0; t < 3; t++) {
        if (t >= nc) break;
        const uint32_t k = (dk >> (2 * t)) & 3;
        s -= ev[t] * (k == 1 ? w1 : (k == 2 ? w2 : w3));
      }
      const double xj = s / dj;
      x[j] = xj;
      w3 = w2, w2 = w1, w1 = xj;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nr; i += 64)
    m.next[m.mem[r0 + i]] = (double)flush_hd((float)x[i]);
}

// MergedFaces::add_pml (conformal.rs:2582-2638): a thread per term, T and
// psi (kept when commit: the H update; the energy look-ahead leaves the
// state), the contribution k T + psi; then a thread per reduced current
// adds w times its terms' contributions in term order.
__global__ void __launch_bounds__(256) mf_pml_terms(Fields f, DevMerged m, int commit) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m.np) return;
  const double *c = m.pc + 4 * (size_t)q;
  const float *u = f.u[m.pax[q]];
  const double d = c[0] * ((double)u[m.pi0[q]] - (double)u[m.pi1[q]]);
  const double psi = c[1] * m.psi[q] + c[2] * d;
  if (commit) m.psi[q] = psi;
  m.pcon[q] = c[3] * d + psi;
}

__global__ void __launch_bounds__(256) mf_pml_add(DevMerged m) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m.ng || m.pst[g] == m.pst[g + 1]) return;
  double c = m.circ[g];
  for (uint32_t e = m.pst[g]; e < m.pst[g + 1]; e++) c += m.pw[e] * m.pcon[m.pt[e]];
  m.circ[g] = c;
}

// The update of the reduced currents: a wave per block of `list`, then (in
// the launch with the merge groups) a lane per merge group. The block's lanes stage its factor in LDS and form the
// right-hand sides; lane 0 runs the serial substitutions on LDS, the last
// solved value kept in a register. The backward substitution goes column by
// column (x_j = (y_j - sum over k > j, decreasing, of L_kj x_k) / L_jj), the
// very operations of the row-by-row form in the same order. A block whose
// image exceeds lds_bytes is solved in the global scratch row by row.
__global__ void __launch_bounds__(64) mf_advance(DevMerged m, const uint32_t *__restrict__ list,
                                                 int count, int lds_bytes) {
  extern __shared__ double xs[];
  if ((int)blockIdx.x >= count) {
    const int q = ((int)blockIdx.x - count) * 64 + (int)threadIdx.x;
    if (q < m.ns) mf_scalar(m, q, m.cur, m.circ, m.next);
    return;
  }
  const uint32_t b = list[blockIdx.x] & 0x7fffffffu;
  const uint32_t r0 = m.bs[b], nr = m.bs[b + 1] - r0;
  if (list[blockIdx.x] >> 31) {
    mf_narrow(m, r0, nr, xs);
    return;
  }
  const uint64_t e0 = m.eo[r0], nev = m.eo[r0 + nr] - e0;
  const uint32_t c0 = m.cs[r0], ncol = m.cs[r0 + nr] - c0;
  if (mf_lds_need(nr, nev, ncol) > (size_t)lds_bytes) {
    double *x = m.x + r0;
    for (uint32_t i = threadIdx.x; i < nr; i += 64) x[i] = mf_rhs(m, r0, i, m.cur, m.circ);
    __syncthreads();
    if (threadIdx.x == 0) mf_solve(m, r0, nr, x);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nr; i += 64)
      m.next[m.mem[r0 + i]] = (double)flush_hd((float)x[i]);
    return;
  }
  double *x = xs, *env = xs + nr, *cv = env + nev;
  uint32_t *ef = (uint32_t *)(cv + ncol), *eo = ef + nr, *cs = eo + nr + 1, *ck = cs + nr + 1;
  for (uint32_t i = threadIdx.x; i < nr; i += 64) ef[i] = m.ef[r0 + i];
  for (uint32_t i = threadIdx.x; i <= nr; i += 64) {
    eo[i] = (uint32_t)(m.eo[r0 + i] - e0);
    cs[i] = m.cs[r0 + i] - c0;
  }
  for (uint64_t e = threadIdx.x; e < nev; e += 64) env[e] = m.ev[e0 + e];
  for (uint32_t e = threadIdx.x; e < ncol; e += 64) {
    cv[e] = m.cv[c0 + e];
    ck[e] = m.cr[c0 + e];
  }
  for (uint32_t i = threadIdx.x; i < nr; i += 64) x[i] = mf_rhs(m, r0, i, m.cur, m.circ);
  __syncthreads();
  if (threadIdx.x == 0) {
    double xp = 0.0;
    for (uint32_t i = 0; i < nr; i++) {
      const uint32_t f = ef[i];
      const double *row = env + eo[i];
      double s = x[i];
      if (f < i) {
        for (uint32_t j = f; j + 1 < i; j++) s -= row[j - f] * x[j];
        s -= row[i - 1 - f] * xp;
      }
      xp = s / row[i - f];
      x[i] = xp;
    }
    for (uint32_t j = nr; j-- > 0;) {
      double s = x[j];
      uint32_t q = cs[j];
      const uint32_t q1 = cs[j + 1];
      if (q < q1) {
        for (; q + 1 < q1; q++) s -= cv[q] * x[ck[q]];
        s -= cv[q] * (ck[q] == j + 1 ? xp : x[ck[q]]);
      }
      xp = s / env[eo[j] + j - ef[j]];
      x[j] = xp;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nr; i += 64)
    m.next[m.mem[r0 + i]] = (double)flush_hd((float)x[i]);
}

// MergedFaces::post_h, the rest (conformal.rs:2662-2671): g = g', and every
// driven face's current sum w g (from -0.0) as f32.
__global__ void __launch_bounds__(256) mf_drive(Fields f, DevMerged m) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m.nf) {
    double s = -0.0;
    for (uint32_t e = m.fs[t]; e < m.fs[t + 1]; e++) s += m.fw[e] * m.next[m.fg[e]];
    f.i[m.fax[t]][m.fidx[t]] = (float)s;
  }
  if (t < m.ng) m.cur[t] = m.next[t];
}

// MergedFaces::energy's terms (conformal.rs:2684-2698), the host's
// expressions: per merge group q < ns, l/2 g_q g'_q; per block row r,
// g_r/2 (M g')_r with (M g')_r folded from -0.0 in column order. `ml`
// holds l, `mv` the M values, `mrs`/`mcol` the M rows, `row0` each row's
// block start; g' (next) is the look-ahead update.
__global__ void __launch_bounds__(256) mf_energy_terms(DevMerged m, const double *__restrict__ ml,
                                                       const uint32_t *__restrict__ mrs,
                                                       const uint32_t *__restrict__ mcol,
                                                       const double *__restrict__ mv,
                                                       const uint32_t *__restrict__ row0,
                                                       int nrows, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m.ns) {
    out[t] = 0.5 * ml[t] * m.cur[t] * m.next[t];
    return;
  }
  const int r = t - m.ns;
  if (r >= nrows) return;
  const uint32_t r0 = row0[r];
  double mn = -0.0;
  for (uint32_t k = mrs[r]; k < mrs[r + 1]; k++) mn += mv[k] * m.next[m.mem[r0 + mcol[k]]];
  out[t] = 0.5 * m.cur[m.mem[r]] * mn;
}

// ---- dispersive media (dispersive.rs) ----

struct DevPoleConst {
  double a, inv_c, inv_z;
  int kind;  // 0 trapezoidal, 1 leapfrog
};
// One side of a medium: edges (CSR over poles) and their state.
struct DevPoleSide {
  int n;
  const int32_t *axis;
  const uint64_t *idx;
  const double *sum_inv_z;
  const uint32_t *pole_end, *cidx;
  double *j, *q;
  float *prev;  // the field at the start of the update
};

// Side::add_energy's terms (dispersive.rs:1489-1517), the host's
// expression: per pole half_l j^2 + inv_c q^2 / 2, `hc` the constants'
// (half_l, inv_c) pairs.
__global__ void __launch_bounds__(256) pole_energy_terms(DevPoleSide s, int np,
                                                         const double *__restrict__ hc,
                                                         double *__restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const uint32_t c = s.cidx[p];
  out[p] = hc[2 * c] * s.j[p] * s.j[p] + 0.5 * hc[2 * c + 1] * s.q[p] * s.q[p];
}

// Side::correct / step_edge (dispersive.rs:1183-1343, 1392-1470) for
// trapezoidal and leapfrog poles, a thread per edge: `fld` the field the
// side corrects (u or i), lane 3 + axis (cb) or 9 + axis (db). With
// `fp_prev` the saved value is read from `fp` (the state before a fused
// step), else from `prev`.
template <bool ESIDE>
__global__ void __launch_bounds__(256) poles_correct(Fields f, Fields fp, int fp_prev, Lanes L,
                                                     int nz, DevPoleSide s,
                                                     const DevPoleConst *__restrict__ cs,
                                                     double dt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= s.n) return;
  const int ax = s.axis[e];
  const uint64_t idx = s.idx[e];
  float *fld = ESIDE ? f.u[ax] : f.i[ax];
  const float *pf = ESIDE ? fp.u[ax] : fp.i[ax];
  const size_t r = idx / (uint64_t)nz;
  const int k = (int)(idx - r * (uint64_t)nz);
  const double cb = (double)coef(L.l[(ESIDE ? 3 : 9) + ax], r, k, nz, idx);
  const double prev = (double)(fp_prev ? pf[idx] : s.prev[e]);
  const uint32_t p0 = e ? s.pole_end[e - 1] : 0u, p1 = s.pole_end[e];
  const double u_star = (double)fld[idx];
  const double half_prev = 0.5 * prev;
  double sum = 0.0;
  for (uint32_t p = p0; p < p1; p++) {
    const DevPoleConst c = cs[s.cidx[p]];
    if (c.kind == 1)
      sum += c.a * s.j[p] + c.inv_z * (prev - s.q[p] * c.inv_c);
    else
      sum += (c.a * s.j[p] - s.q[p] * c.inv_c + half_prev) * c.inv_z;
  }
  const float u_new = (float)((u_star - cb * sum) / (1.0 + 0.5 * cb * s.sum_inv_z[e]));
  const double half_new = 0.5 * (double)u_new;
  for (uint32_t p = p0; p < p1; p++) {
    const DevPoleConst c = cs[s.cidx[p]];
    const double jo = s.j[p], qo = s.q[p];
    if (c.kind == 1) {
      const double jn = c.a * jo + c.inv_z * (prev - qo * c.inv_c);
      s.j[p] = jn;
      s.q[p] = qo + dt * jn;
    } else {
      const double kk = c.a * jo - qo * c.inv_c + half_prev;
      const double jb = (kk + half_new) * c.inv_z;
      s.j[p] = 2.0 * jb - jo;
      s.q[p] = qo + dt * jb;
    }
  }
  fld[idx] = u_new;
}

// Side::capture (dispersive.rs:1173-1179): the field of each edge.
template <bool ESIDE>
__global__ void __launch_bounds__(256) poles_capture(Fields f, DevPoleSide s) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= s.n) return;
  const int ax = s.axis[e];
  s.prev[e] = (ESIDE ? f.u[ax] : f.i[ax])[s.idx[e]];
}

// ---- energy invariant (kernel.rs:119-244, 364-423; pool.rs:673-692) ----

// The chains below are the reference's sums, each in its own order; what is
// parallel is the terms. A wave computes a round of up to 256 consecutive
// terms at once (coalesced loads, one latency per round) into LDS, then
// the lanes that own the chains add them in order.

// acc + b[t] + b[t + st] + ... over indices < m, in that order; the LDS
// reads issued 16 at a time so that one latency serves 16 additions.
__device__ __forceinline__ double chain_add(double acc, const double *b, int t, int m, int st) {
  int q = t;
  for (; q + 15 * st < m; q += 16 * st) {
    double v[16];
#pragma unroll
    for (int s = 0; s < 16; s++) v[s] = b[q + st * s];
#pragma unroll
    for (int s = 0; s < 16; s++) acc += v[s];
  }
  for (; q < m; q += st) acc += b[q];
  return acc;
}

// A field array by a wave-uniform runtime index, through selects.
__device__ __forceinline__ const float *pick3(float *const *a, int c) {
  return c == 0 ? a[0] : (c == 1 ? a[1] : a[2]);
}

// RunningDft::push (post/dft.rs:125-136) over samples [n0, n1) of one
// recorded series, a thread per frequency: acc += phasor * x, phasor *=
// step, and at every 1024th sample the phasor re-anchored from the exact
// phase (the anchors are the host's, ems_dft_phasors). State per
// frequency: acc (re, im), phasor (re, im). Same operations in the same
// order as the host's ems_dft, so the same bits.
__global__ void __launch_bounds__(256) dft_run(const double *__restrict__ x, uint64_t n0,
                                               uint64_t n1, int nf, double *__restrict__ st,
                                               const double *__restrict__ stepv,
                                               const double *__restrict__ anchors) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nf) return;
  double ar = st[4 * f], ai = st[4 * f + 1], pr = st[4 * f + 2], pi = st[4 * f + 3];
  const double sr = stepv[2 * f], si = stepv[2 * f + 1];
  for (uint64_t s = n0; s < n1; s++) {
    const double xv = ((const GAS double *)x)[s];
    ar = ar + pr * xv;
    ai = ai + pi * xv;
    const double nr = pr * sr - pi * si, ni = pr * si + pi * sr;
    pr = nr, pi = ni;
    if ((s + 1) % 1024 == 0) {
      const double *a = anchors + ((s + 1) / 1024 - 1) * 2 * (uint64_t)nf;
      pr = a[2 * f], pi = a[2 * f + 1];
    }
  }
  st[4 * f] = ar, st[4 * f + 1] = ai, st[4 * f + 2] = pr, st[4 * f + 3] = pi;
}

// H rows, a wave per row (rows per plane: x rows ny-1, y rows ny, z rows
// ny-1, enumerated in that order): four accumulators over k = 0, 1, 2, 3
// mod 4 in increasing k (h_energy_row, kernel.rs:119-216: the weighted
// i^n i^(n+1), the update evaluated read-only), combined as
// (a0 + a1) + (a2 + a3). Rounds of 64 NT terms, NT per lane (NT sized to
// the rows: 1 for rows of at most 64 nodes, up to 8).
template <bool OE, int NT, bool MEAS>
__device__ __forceinline__ double energy_h_round(const float *A, const float *Cc, const float *H,
                                                 size_t sa, size_t sc, const Lane &lda,
                                                 const Lane &ldb, size_t base, size_t r, int nz,
                                                 int comp, int n, int k0, int t, double *b,
                                                 double acc) {
  constexpr int RND = 64 * NT;
  float va[NT], vb[NT], vc[NT], vd[NT], da[NT], db[NT], old[NT];
#pragma unroll
  for (int s = 0; s < NT; s++) {
    const int k = min(k0 + 64 * s + t, n - 1);
    const size_t o = base + k;
    va[s] = A[o + sa], vb[s] = A[o], vc[s] = Cc[o + sc], vd[s] = Cc[o], old[s] = H[o];
    da[s] = coef(lda, r, k, nz, o), db[s] = coef(ldb, r, k, nz, o);
  }
#pragma unroll
  for (int s = 0; s < NT; s++) {
    const float nw = h_expr<OE>(da[s], old[s], db[s], va[s], vb[s], vc[s], vd[s]);
    const int k = k0 + 64 * s + t;
    double term;
    if (MEAS)  // openEMS's measure: the ends of a z row weigh 4 (kernel.rs:153-156, 497-507)
      term = (comp == 2 && (k == 0 || k == n - 1) ? 4.0 : 1.0) * oe_square((double)nw);
    else
      term = weight(da[s], (double)old[s] * (double)nw, db[s]);
    b[64 * s + t] = k < n ? term : 0.0;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (t < 4) acc = chain_add(acc, b, t, min(RND, n - k0), 4);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return acc;
}

template <bool OE, int NT, bool MEAS>
__device__ __forceinline__ void energy_h_rows_at(const Fields &f, const Lanes &L, int nx, int ny,
                                                 int nz, double *rowsum, size_t blk,
                                                 double (*buf)[64 * NT]) {
  constexpr int RND = 64 * NT;
  const int wv = threadIdx.x >> 6, t = threadIdx.x & 63;
  const int rows_per_plane = 3 * ny - 2;
  const size_t rid = blk * 4 + wv;
  const size_t total = (size_t)nx * rows_per_plane;
  if (rid >= total) return;  // uniform per wave; no block barrier below
  const int gi = (int)(rid / rows_per_plane);
  const int rr = (int)(rid % rows_per_plane);
  int comp, j;
  if (rr < ny - 1) {
    comp = 0, j = rr;
  } else if (rr < 2 * ny - 1) {
    comp = 1, j = rr - (ny - 1);
  } else {
    comp = 2, j = rr - (2 * ny - 1);
  }
  const size_t plane = (size_t)ny * nz;
  const size_t base = (size_t)gi * plane + (size_t)j * nz;
  const size_t r = (size_t)gi * ny + j;
  const int n = (comp == 0 || gi + 1 < nx) ? (comp == 2 ? nz : nz - 1) : 0;
  // the curl's four voltages: a(o + sa) - b(o) - (c(o + sc) - d(o))
  const int ca_ = comp == 0 ? 2 : (comp == 1 ? 0 : 1), cc_ = comp == 0 ? 1 : (comp == 1 ? 2 : 0);
  const float *A = pick3(f.u, ca_), *Cc = pick3(f.u, cc_), *H = pick3(f.i, comp);
  const size_t sa = comp == 0 ? (size_t)nz : (comp == 1 ? 1 : plane);
  const size_t sc = comp == 0 ? 1 : (comp == 1 ? plane : (size_t)nz);
  const Lane lda = pick_lane(L, 6 + comp), ldb = pick_lane(L, 9 + comp);
  double acc = 0.0;  // lanes 0..3: the accumulator of k = lane mod 4
  double *b = buf[wv];
  // Rounds of RND terms while they fill more than NT - 1 of their 64-term
  // parts, the rest in rounds of 64 (a 257-node z row: 256 + 64 slots, not
  // 512). Every round starts at a multiple of 64, so each chain still gets
  // its terms in increasing k.
  int k0 = 0;
  for (; n - k0 > RND - 64; k0 += RND)
    acc = energy_h_round<OE, NT, MEAS>(A, Cc, H, sa, sc, lda, ldb, base, r, nz, comp, n, k0, t,
                                       b, acc);
  for (; k0 < n; k0 += 64)
    acc = energy_h_round<OE, 1, MEAS>(A, Cc, H, sa, sc, lda, ldb, base, r, nz, comp, n, k0, t, b,
                                      acc);
  const double a1 = __shfl(acc, 1), a2 = __shfl(acc, 2), a3 = __shfl(acc, 3);
  // openEMS's measure: x rows of the outer x planes and y rows j = 0, ny - 1
  // weigh 4 (h_plane, kernel.rs:457-493)
  const double rw = MEAS && ((comp == 0 && (gi == 0 || gi + 1 == nx)) ||
                             (comp == 1 && (j == 0 || j + 1 == ny)))
                        ? 4.0 : 1.0;
  if (t == 0) rowsum[rid] = MEAS ? rw * ((acc + a1) + (a2 + a3)) : (acc + a1) + (a2 + a3);
}

// The same row sums two rows per wave for rows of at most 32 nodes (half
// a wave per row, one round; lanes 0..3 of each half its accumulators):
// the same terms added in the same order. Rows with no terms (the last
// plane's y and z rows) load nothing.
template <bool OE, bool MEAS>
__device__ __forceinline__ void energy_h_rows2_at(const Fields &f, const Lanes &L, int nx, int ny,
                                                  int nz, double *rowsum, size_t blk,
                                                  double (*buf)[64]) {
  const int wv = threadIdx.x >> 6, t = threadIdx.x & 63;
  const int slot = t >> 5, tl = t & 31;
  const int rows_per_plane = 3 * ny - 2;
  const size_t total = (size_t)nx * rows_per_plane;
  const size_t rid = (blk * 4 + wv) * 2 + slot;
  if ((blk * 4 + wv) * 2 >= total) return;  // uniform per wave; no block barrier below
  const bool live = rid < total;
  const size_t rq = live ? rid : 0;  // (a slot past the end: row 0, nothing written)
  const int gi = (int)(rq / rows_per_plane);
  const int rr = (int)(rq % rows_per_plane);
  int comp, j;
  if (rr < ny - 1) {
    comp = 0, j = rr;
  } else if (rr < 2 * ny - 1) {
    comp = 1, j = rr - (ny - 1);
  } else {
    comp = 2, j = rr - (2 * ny - 1);
  }
  const size_t plane = (size_t)ny * nz;
  const size_t base = (size_t)gi * plane + (size_t)j * nz;
  const size_t r = (size_t)gi * ny + j;
  const int n = (comp == 0 || gi + 1 < nx) ? (comp == 2 ? nz : nz - 1) : 0;
  const int ca_ = comp == 0 ? 2 : (comp == 1 ? 0 : 1), cc_ = comp == 0 ? 1 : (comp == 1 ? 2 : 0);
  const float *A = pick3(f.u, ca_), *Cc = pick3(f.u, cc_), *H = pick3(f.i, comp);
  const size_t sa = comp == 0 ? (size_t)nz : (comp == 1 ? 1 : plane);
  const size_t sc = comp == 0 ? 1 : (comp == 1 ? plane : (size_t)nz);
  const Lane lda = pick_lane(L, 6 + comp), ldb = pick_lane(L, 9 + comp);
  double *b = buf[wv];
  double term = 0.0;
  if (tl < n) {  // (n <= nz <= 32)
    const int k = tl;
    const size_t o = base + k;
    const float va = A[o + sa], vb = A[o], vc = Cc[o + sc], vd = Cc[o], old = H[o];
    const float da = coef(lda, r, k, nz, o), db = coef(ldb, r, k, nz, o);
    const float nw = h_expr<OE>(da, old, db, va, vb, vc, vd);
    if (MEAS)
      term = (comp == 2 && (k == 0 || k == n - 1) ? 4.0 : 1.0) * oe_square((double)nw);
    else
      term = weight(da, (double)old * (double)nw, db);
  }
  b[t] = term;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double acc = 0.0;
  if (tl < 4) acc = chain_add(acc, b + 32 * slot, tl, n, 4);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int l0 = 32 * slot;
  const double a1 = __shfl(acc, l0 + 1), a2 = __shfl(acc, l0 + 2), a3 = __shfl(acc, l0 + 3);
  const double rw = MEAS && ((comp == 0 && (gi == 0 || gi + 1 == nx)) ||
                             (comp == 1 && (j == 0 || j + 1 == ny)))
                        ? 4.0 : 1.0;
  if (tl == 0 && live)
    rowsum[rid] = MEAS ? rw * ((acc + a1) + (a2 + a3)) : (acc + a1) + (a2 + a3);
}

template <bool OE, bool MEAS>
__global__ void __launch_bounds__(256) energy_h_rows2(Fields f, Lanes L, int nx, int ny, int nz,
                                                      double *rowsum) {
  __shared__ double buf[4][64];
  energy_h_rows2_at<OE, MEAS>(f, L, nx, ny, nz, rowsum, blockIdx.x, buf);
}

template <bool OE, int NT, bool MEAS>
__global__ void __launch_bounds__(256) energy_h_rows(Fields f, Lanes L, int nx, int ny, int nz,
                                                     double *rowsum) {
  __shared__ double buf[4][64 * NT];
  energy_h_rows_at<OE, NT, MEAS>(f, L, nx, ny, nz, rowsum, blockIdx.x, buf);
}

// E part: per (plane, component) four chains over the in-plane index
// q = 0, 1, 2, 3 mod 4 in increasing q (e_energy_row, kernel.rs:218-244),
// a wave per (plane, component), lanes 0..3 the chains. Rounds of 1024
// terms, 16 per lane.
constexpr int ECH_RND = 1024;  // terms per round of energy_e_chains (16 per lane)

template <bool MEAS>
__device__ __forceinline__ void energy_e_chains_at(const Fields &f, const Lanes &L, int nx, int ny,
                                                   int nz, double *chains, size_t blk,
                                                   double (*buf)[ECH_RND]) {
  constexpr int NT = 16, RND = 64 * NT;
  const int wv = threadIdx.x >> 6, t = threadIdx.x & 63;
  const size_t w = blk * 4 + wv;
  if (w >= (size_t)nx * 3) return;  // uniform per wave
  const int a = (int)(w % 3), gi = (int)(w / 3);
  const size_t plane = (size_t)ny * nz;
  const float *u = pick3(f.u, a) + (size_t)gi * plane;
  const Lane la = pick_lane(L, a), lb = pick_lane(L, 3 + a);
  double acc = 0.0;
  double *b = buf[wv];
  for (size_t q0 = 0; q0 < plane; q0 += RND) {
    float x[NT], ca[NT], cb[NT];
    bool in[NT];
#pragma unroll
    for (int s = 0; s < NT; s++) {
      const size_t q = min(q0 + 64 * s + t, plane - 1);
      const int j = (int)(q / nz), k = (int)(q - (size_t)j * nz);
      const size_t r = (size_t)gi * ny + j;
      const size_t o = (size_t)gi * plane + q;
      x[s] = u[q];
      // openEMS's measure leaves out the E on the three upper faces
      // (kernel.rs:410-418); the zero terms leave the chains unchanged.
      in[s] = gi + 1 < nx && j + 1 < ny && k + 1 < nz;
      if (!MEAS) ca[s] = coef(la, r, k, nz, o), cb[s] = coef(lb, r, k, nz, o);
    }
#pragma unroll
    for (int s = 0; s < NT; s++) {
      const double xd = (double)x[s];
      double term;
      if (MEAS)
        term = in[s] ? oe_square(xd) : 0.0;
      else
        term = weight(ca[s], xd * xd, cb[s]);
      b[64 * s + t] = q0 + 64 * s + t < plane ? term : 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (t < 4) acc = chain_add(acc, b, t, (int)min((size_t)RND, plane - q0), 4);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  if (t < 4) chains[w * 4 + t] = acc;
}

template <bool MEAS>
__global__ void __launch_bounds__(256) energy_e_chains(Fields f, Lanes L, int nx, int ny, int nz,
                                                       double *chains) {
  __shared__ double buf[4][ECH_RND];
  energy_e_chains_at<MEAS>(f, L, nx, ny, nz, chains, blockIdx.x, buf);
}

// The same chains with a block per (plane, component): the four waves
// compute a round of 4096 consecutive terms (16 per lane) into LDS, then
// lanes 0..3 of the first wave add them to the four chains in increasing q,
// so each chain gets the same terms in the same order. The in-plane index
// is 32-bit (planes of fewer than 2^31 nodes; the host checks), which
// turns the per-term 64-bit division into a 32-bit one by a fixed divisor.
constexpr int ECB_RND = 4096;

template <bool MEAS>
__global__ void __launch_bounds__(256) energy_e_chains_blk(Fields f, Lanes L, int nx, int ny,
                                                           int nz, double *chains) {
  constexpr int NT = 16;
  __shared__ double b[ECB_RND];
  const int t = threadIdx.x;
  const int w = blockIdx.x;  // 3 * plane + component
  const int a = w % 3, gi = w / 3;
  const uint32_t unz = (uint32_t)nz, uny = (uint32_t)ny, plane = uny * unz;
  const float *u = pick3(f.u, a) + (size_t)gi * plane;
  const Lane la = pick_lane(L, a), lb = pick_lane(L, 3 + a);
  double acc = 0.0;
  for (uint32_t q0 = 0; q0 < plane; q0 += ECB_RND) {
    float x[NT], ca[NT], cb[NT];
    bool in[NT];
#pragma unroll
    for (int s = 0; s < NT; s++) {
      const uint32_t q = min(q0 + 256u * s + t, plane - 1);
      const uint32_t j = q / unz, k = q - j * unz;
      const size_t r = (size_t)gi * ny + j;
      const size_t o = (size_t)gi * plane + q;
      x[s] = u[q];
      // openEMS's measure leaves out the E on the three upper faces
      // (kernel.rs:410-418); the zero terms leave the chains unchanged.
      in[s] = gi + 1 < nx && j + 1 < uny && k + 1 < unz;
      if (!MEAS) ca[s] = coef(la, r, (int)k, nz, o), cb[s] = coef(lb, r, (int)k, nz, o);
    }
#pragma unroll
    for (int s = 0; s < NT; s++) {
      const double xd = (double)x[s];
      double term;
      if (MEAS)
        term = in[s] ? oe_square(xd) : 0.0;
      else
        term = weight(ca[s], xd * xd, cb[s]);
      b[256 * s + t] = q0 + 256u * s + t < plane ? term : 0.0;
    }
    __syncthreads();
    if (t < 4) acc = chain_add(acc, b, t, (int)min((uint32_t)ECB_RND, plane - q0), 4);
    __syncthreads();
  }
  if (t < 4) chains[(size_t)w * 4 + t] = acc;
}

// Plane values we + wh (h_slab, kernel.rs:404-420): a wave per plane, the
// plane's row sums added in row order by lane 0 from LDS.
template <bool MEAS>
__device__ __forceinline__ void energy_planes_at(int nx, int ny, const double *rowsum,
                                                 const double *chains, double *planes, int blk,
                                                 double (*buf)[256]) {
  const int wv = threadIdx.x >> 6, t = threadIdx.x & 63;
  const int gi = blk * 4 + wv;
  if (gi >= nx) return;  // uniform per wave
  const int rows_per_plane = 3 * ny - 2;
  const double *rs = rowsum + (size_t)gi * rows_per_plane;
  const int nrows = gi + 1 < nx ? rows_per_plane : ny - 1;
  double wh = 0.0;
  double *b = buf[wv];
  for (int q0 = 0; q0 < nrows; q0 += 256) {
    double v[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int q = q0 + 64 * s + t;
      v[s] = q < nrows ? rs[q] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 4; s++) b[64 * s + t] = v[s];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (t == 0) wh = chain_add(wh, b, 0, min(256, nrows - q0), 1);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  if (t == 0) {
    double we = 0.0;
    for (int a = 0; a < 3; a++) {
      const double *c = chains + ((size_t)gi * 3 + a) * 4;
      we += (c[0] + c[1]) + (c[2] + c[3]);
    }
    planes[gi] = MEAS ? EPS0_D * we + MU0_D * wh : we + wh;
  }
}

template <bool MEAS>
__global__ void __launch_bounds__(256) energy_planes(int nx, int ny, const double *rowsum,
                                                     const double *chains, double *planes) {
  __shared__ double buf[4][256];
  energy_planes_at<MEAS>(nx, ny, rowsum, chains, planes, (int)blockIdx.x, buf);
}

// Parts::energy_openems_slab (upml.rs:505-537): per H run of the layers
// (component, plane) mu0 sum (v^2 - inc^2) over its edges in index order,
// inc the kernel's flux increment evaluated in the Paired order as written
// there, v the completed layer value; added to the plane in component
// order. A wave per plane, the terms a round at a time into LDS, lane 0
// adding them in order. hplane[c (nx + 1) + gi]: the plane's first H edge
// of component c (device order: by component, then node index).
__global__ void __launch_bounds__(256) upml_energy_planes(Fields f, Lanes L, DevUpml m,
                                                          const uint32_t *hplane, int nx, int ny,
                                                          int nz, double *planes) {
  __shared__ double buf[4][256];
  const int wv = threadIdx.x >> 6, t = threadIdx.x & 63;
  const int gi = (int)blockIdx.x * 4 + wv;
  if (gi >= nx) return;  // uniform per wave
  const size_t plane = (size_t)ny * nz;
  double *b = buf[wv];
  double pv = t == 0 ? planes[gi] : 0.0;
  for (int comp = 0; comp < 3; comp++) {
    const uint32_t r0 = hplane[comp * (nx + 1) + gi], r1 = hplane[comp * (nx + 1) + gi + 1];
    if (r0 == r1) continue;  // no run
    const Lane ldb = pick_lane(L, 9 + comp);
    double sum = 0.0;
    for (uint32_t q0 = r0; q0 < r1; q0 += 256) {
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const uint32_t q = q0 + 64 * s + t;
        double term = 0.0;
        if (q < r1) {
          const size_t idx = m.idx[q], o = idx - (size_t)gi * plane;
          const int j = (int)(o / nz), k = (int)(o - (size_t)j * nz);
          const float *ux = f.u[0], *uy = f.u[1], *uz = f.u[2];
          float c;
          if (comp == 0)
            c = (uz[idx + nz] - uz[idx]) - (uy[idx + 1] - uy[idx]);
          else if (comp == 1)
            c = (ux[idx + 1] - ux[idx]) - (uz[idx + plane] - uz[idx]);
          else
            c = (uy[idx + plane] - uy[idx]) - (ux[idx + nz] - ux[idx]);
          const float db = coef(ldb, (size_t)gi * ny + j, k, nz, idx);
          const float2 st = m.fv[q];
          const float4 co = m.co[q];
          const float inc = flush(0.0f * st.y - db * c);
          const float f1 = flush(flush(co.x * st.x) + inc);
          const float v = flush(flush(flush(co.y * st.y) - flush(co.w * st.x)) + flush(co.z * f1));
          term = (double)v * (double)v - (double)inc * (double)inc;
        }
        b[64 * s + t] = term;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (t == 0) sum = chain_add(sum, b, 0, (int)min(256u, r1 - q0), 1);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (t == 0) pv += MU0_D * sum;
  }
  if (t == 0) planes[gi] = pv;
}

// The kernel's look-ahead i^{n+1/2} of listed dual edges (dump.rs:1830-1845),
// completed by the layers where the edge is one of theirs (binary search
// in the layers' H edges of its component, sorted by node index).
template <bool OE>
__global__ void __launch_bounds__(64) lookahead_edges(Fields f, Lanes L, int n, const int32_t *axis,
                                                      const uint64_t *idx, int nx, int ny, int nz,
                                                      DevUpml m, int has_upml, float *out) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  if (q >= n) return;
  const int c = axis[q], a = (c + 1) % 3, bb = (c + 2) % 3;
  const uint64_t o = idx[q];
  const int p[3] = {(int)(o / ((uint64_t)ny * nz)), (int)((o / nz) % ny), (int)(o % nz)};
  const int nn[3] = {nx, ny, nz};
  const size_t st[3] = {(size_t)ny * nz, (size_t)nz, 1};
  const float *ua = pick3(f.u, a), *ub = pick3(f.u, bb);
  const float t0 = p[a] + 1 < nn[a] ? ub[o + st[a]] : 0.0f, t1 = ub[o];
  const float t2 = p[bb] + 1 < nn[bb] ? ua[o + st[bb]] : 0.0f, t3 = ua[o];
  const size_t r = (size_t)p[0] * ny + p[1];
  const float da = coef(pick_lane(L, 6 + c), r, p[2], nz, o);
  const float db = coef(pick_lane(L, 9 + c), r, p[2], nz, o);
  float k = h_expr<OE>(da, pick3(f.i, c)[o], db, t0, t1, t2, t3);
  if (has_upml) {
    int lo = m.off[3 + c], hi = m.off[4 + c];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (m.idx[mid] < o) lo = mid + 1; else hi = mid;
    }
    if (lo < m.off[4 + c] && m.idx[lo] == o) {
      const float2 s = m.fv[lo];
      const float4 co = m.co[lo];
      const float f1 = flush(flush(co.x * s.x) + k);
      k = flush(flush(flush(co.y * s.y) - flush(co.w * s.x)) + flush(co.z * f1));
    }
  }
  out[q] = k;
}

// The plane values summed in plane order (pool.rs:675-692), folding from
// -0.0, scaled by dt/4; one wave.
template <bool MEAS>
__device__ __forceinline__ void energy_combine_at(int nx, const double *planes, double dt,
                                                  double *out, double *buf) {
  const int t = threadIdx.x;
  double s = -0.0;
  for (int g0 = 0; g0 < nx; g0 += 256) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int g = g0 + 64 * k + t;
      buf[64 * k + t] = g < nx ? planes[g] : 0.0;
    }
    __syncthreads();
    if (t == 0) s = chain_add(s, buf, 0, min(256, nx - g0), 1);
    __syncthreads();
  }
  if (t == 0) *out = MEAS ? s : 0.25 * dt * s;  // scale_energy (pool.rs:684-689)
}

template <bool MEAS>
__global__ void __launch_bounds__(64) energy_combine(int nx, const double *planes, double dt,
                                                     double *out) {
  __shared__ double buf[256];
  energy_combine_at<MEAS>(nx, planes, dt, out, buf);
}

// ---- the invariant of many models at once (ems_gpu_energy_many) ----
//
// One launch per kernel above over the models of a table: each model's
// blocks follow the previous model's (first block hb, eb, pb of the H-row,
// E-chain and plane launches; the table's last entry holds the totals), a
// block finds its model by binary search and runs that model's block
// blk - first. The same terms in the same order as the single-model
// launches, so the same bits.
struct EnergyModel {
  Fields f;
  Lanes L;
  int nx, ny, nz;
  double dt;
  double *rowsum, *chains, *planes;
  uint32_t hb, eb, pb;
};

// The model of block b (the last whose first block is <= b), wave-uniform.
// The kernels then read the model's entry in place (a copy of it would go
// to scratch memory, its fields being passed by reference).
template <int W>
__device__ __forceinline__ int energy_model_of(const EnergyModel *__restrict__ ms, int nm,
                                               uint32_t b) {
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    const EnergyModel *m = ms + mid;
    const uint32_t first = cload(W == 0 ? &m->hb : (W == 1 ? &m->eb : &m->pb));
    if (first <= b) lo = mid; else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

template <bool OE, int NT>
__global__ void __launch_bounds__(256) energy_h_rows_many(const EnergyModel *__restrict__ ms,
                                                          int nm) {
  __shared__ double buf[4][64 * NT];
  const EnergyModel *m = ms + energy_model_of<0>(ms, nm, blockIdx.x);
  energy_h_rows_at<OE, NT, false>(m->f, m->L, m->nx, m->ny, m->nz, m->rowsum,
                                  blockIdx.x - cload(&m->hb), buf);
}

// (rows of at most 32 nodes: two per wave, eight per group, EH2_GROUPS
// groups per block, so that one search of the model table serves 32 rows)
constexpr int EH2_GROUPS = 4;

template <bool OE>
__global__ void __launch_bounds__(256) energy_h_rows2_many(const EnergyModel *__restrict__ ms,
                                                           int nm) {
  __shared__ double buf[4][64];
  const EnergyModel *m = ms + energy_model_of<0>(ms, nm, blockIdx.x);
  const size_t g0 = (size_t)(blockIdx.x - cload(&m->hb)) * EH2_GROUPS;
  for (int g = 0; g < EH2_GROUPS; g++)
    energy_h_rows2_at<OE, false>(m->f, m->L, m->nx, m->ny, m->nz, m->rowsum, g0 + g, buf);
}

__global__ void __launch_bounds__(256) energy_e_chains_many(const EnergyModel *__restrict__ ms,
                                                            int nm) {
  __shared__ double buf[4][ECH_RND];
  const EnergyModel *m = ms + energy_model_of<1>(ms, nm, blockIdx.x);
  energy_e_chains_at<false>(m->f, m->L, m->nx, m->ny, m->nz, m->chains,
                            blockIdx.x - cload(&m->eb), buf);
}

__global__ void __launch_bounds__(256) energy_planes_many(const EnergyModel *__restrict__ ms,
                                                          int nm) {
  __shared__ double buf[4][256];
  const EnergyModel *m = ms + energy_model_of<2>(ms, nm, blockIdx.x);
  energy_planes_at<false>(m->nx, m->ny, m->rowsum, m->chains, m->planes,
                          (int)(blockIdx.x - cload(&m->pb)), buf);
}

// a block per model; W^n of model q into out[q]
__global__ void __launch_bounds__(64) energy_combine_many(const EnergyModel *__restrict__ ms,
                                                          double *out) {
  __shared__ double buf[256];
  const EnergyModel *m = ms + blockIdx.x;
  energy_combine_at<false>(m->nx, m->planes, m->dt, out + blockIdx.x, buf);
}

// ------------------------------------------------------------------ host --

struct HostSlab {
  SlabArg arg;  // arg.psi is set per launch (parity)
  float *d_b = nullptr, *d_a = nullptr, *d_k = nullptr;
  float *d_psi[2] = {nullptr, nullptr};  // H side: double-buffered; E side: [0] only
  float *d_delta = nullptr;  // H side, run outside the fused kernel: its increments
  size_t len = 0;
  bool out = false;  // corrected outside the fused kernel (cpml_delta_h / cpml_post_e)
};

struct HostOp {
  int type = 0, kind = 0, quantity = 0, n = 0, f32sum = 0;
  uint64_t interval = 1;
  int32_t *d_axis = nullptr;
  uint64_t *d_idx = nullptr;
  float *d_wf = nullptr;
  double *d_wd = nullptr;
  float *d_amp = nullptr;
  uint64_t namp = 0;
  double *d_vals = nullptr;
  uint64_t cap = 0;
  uint64_t count = 0;  // samples recorded (host mirror of *d_count)
  uint64_t first = 0;  // step of sample 0
  uint64_t *d_count = nullptr;
  // lumped element: its device record (d_lp) and what it points to
  DevLumped lp{};
  DevLumped *d_lp = nullptr;
  std::vector<double *> tables;      // per term
  std::vector<uint64_t> table_len;
  double **d_tv = nullptr;
  uint64_t *d_tlen = nullptr;
  std::vector<double> two_l_dt, inv_c;  // host copies for the branch energy
  std::vector<void *> mem;              // every allocation of this op
  // dispersive medium: sides (0 electric, 1 magnetic), constants
  DevPoleSide side[2]{};
  DevPoleConst *d_consts = nullptr;
  std::vector<double> c_inv_c, c_half_l;  // host copies of the constants
  std::vector<uint32_t> side_cidx[2];
  // plane wave: auxiliary grid, corrections (0 E, 1 H), source table
  DevPw pw{};
  DevPwCorr pwc[2]{};
  double *pw_src = nullptr;
  // reduced currents of cut faces: device tables and state, host copies
  DevMerged mf{};
  // blocks by the size of their LDS image: up to 8 KiB (small), above;
  // the LDS of each launch (the largest image of its blocks, at most 64 KiB)
  uint32_t *mf_small = nullptr, *mf_big = nullptr;
  int mf_nsmall = 0, mf_nbig = 0, mf_lds_small = 0, mf_lds_big = 0;
  std::shared_ptr<struct MergedHost> mfh;
  bool circ_stale = false;  // the post-E circulations not yet formed (merged_circ_now)
  // the operator's energy terms: device array, pinned host copy, count;
  // the device tables the terms need (pole constants; reduced currents' l,
  // M rows, block starts)
  double *d_et = nullptr, *h_et = nullptr;
  size_t n_et = 0;
  double *d_ec = nullptr;
  uint32_t *d_erow0 = nullptr;
  // Mur boundary
  DevMur mur{};
  // openEMS's PML layers; upml_pos: device slot of each edge in the caller's order
  DevUpml upml{};
  std::vector<uint32_t> upml_pos;
  uint32_t *d_hplane = nullptr;  // per H component the first edge of each x plane
  // sources: host copies of the edges, weights and amplitude table
  std::vector<int32_t> h_axis;
  std::vector<uint64_t> h_idx;
  std::vector<float> h_wf, h_amp;
};

// Host copies of a reduced-current operator's tables (its energy).
// A running DFT of a lumped element's record series on the device.
struct HostDft {
  int op = -1, which = 0, nf = 0;
  double t0 = 0.0, dt = 0.0;
  uint64_t done = 0;     // samples accumulated
  uint64_t anchors = 0;  // anchors uploaded (sample 1024 m for m = 1..anchors)
  uint64_t anchor_cap = 0;
  std::vector<double> freqs;
  double *d_state = nullptr, *d_step = nullptr, *d_anchor = nullptr;
};

struct MergedHost {
  std::vector<double> da, db, l, ev, bv, mval;
  std::vector<uint32_t> bs, mem, ef, brs, bc, mrs, mcol;
  std::vector<uint64_t> eo;
};

}  // namespace

struct ems_gpu {
  int n[3] = {0, 0, 0};
  size_t N = 0, plane = 0;
  double dt = 0;
  int oe = 0;
  hipStream_t stream = nullptr;
  // the reduced currents' update beside the H half-step (created with the
  // first such operator; default priority)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_energy = nullptr;  // the energy's copies have landed
  // ems_gpu_energy_many's table and results when this engine is the first
  // active model (device and pinned host copies, capacity in models)
  EnergyModel *em_d = nullptr, *em_h = nullptr;
  double *em_dout = nullptr, *em_hout = nullptr;
  int em_cap = 0;
  Fields fb[2]{};  // ping-pong field buffers (views into fbuf)
  float *fbuf[2] = {nullptr, nullptr};  // u_x..u_z, i_x..i_z, N nodes each
  // coefficient pool: a zero row, then every lane's rows; croff[16*r + l]
  float *cpool = nullptr;
  void *croff = nullptr;   // uint32_t entries, or uint64_t when croff_wide
  bool croff_wide = false;  // the pool is 4 GiB or more (only the tiled kernel addresses it)
  bool lanes_dirty = true;
  int par = 0;     // buffer holding the current state
  Lane lanes[12]{};
  uint64_t lane_rows[12] = {0};
  std::vector<void *> lane_mem;
  std::vector<HostSlab> slabs;
  FSlab *d_fslabs = nullptr;
  uint32_t *d_pmask = nullptr;  // per x-plane: slabs (H, E) whose x range holds it
  // Slabs normal to z corrected outside the tiled fused kernel: their
  // records, the per-x-plane mask of the H-side ones, and thread counts.
  DSlab *d_hout = nullptr, *d_eout = nullptr;
  uint32_t *d_pmask_o = nullptr;
  int n_hout = 0, n_eout = 0;
  uint32_t hout_total = 0, eout_total = 0;
  bool cpml_split = false;  // ems_gpu_set_cpml_split (measured no faster: off)
  int split_state = -1;    // what the uploaded tables hold (-1: nothing yet)
  std::vector<HostOp> ops;
  std::vector<HostDft> dfts;
  DevOp *d_ops = nullptr;
  bool ops_dirty = true, slabs_dirty = true;
  uint64_t step = 0;
  uint64_t *d_step = nullptr;
  // fused schedule and its tiling
  bool fused = true, prefetch = true, rows = false, packed = false, band = false;
  int unfused_pf = 11;  // ems_gpu_set_unfused_prefetch: bits 0, 1 load ahead; 2, 3 CPML in
  int xcd_group = 2;  // tile rows per XCD visit (2 measured best of 0, 1, 2, 4, 8)
  int tj = 7, nchunk = 0;
  // tile classes of the tiled step (ems_gpu_set_tile_classes): the tiles a
  // CPML slab reaches and the others, each list padded so that a tile keeps
  // the XCD tile_of_block gives it; d_tiles holds both lists
  bool tile_class = false;
  int *d_tiles = nullptr;
  int ntl[2] = {0, 0};
  uint64_t tiles_key = 0, slab_gen = 0;
  // energy scratch
  double *d_rowsum = nullptr, *d_chains = nullptr, *d_planes = nullptr, *d_w = nullptr;
  double *h_w = nullptr;  // pinned host copy of W^n
  // graphs of whole steps, per starting buffer
  hipGraphExec_t g1[2] = {nullptr, nullptr}, gK[2] = {nullptr, nullptr};
  int K = 32;  // even: a K-step graph ends on the buffer it started from
  std::string err;
  bool dead = false;
};

namespace {

int fail(ems_gpu *h, int code, const std::string &m) {
  if (h) h->err = m;
  return code;
}

#define CK(h, x)                                                                        \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      (h)->dead = true;                                                                 \
      return fail((h), EMS_E_DEVICE, std::string(#x ": ") + hipGetErrorString(e_));     \
    }                                                                                   \
  } while (0)

int x_chunk(const ems_gpu *h, int *nchunks, bool pf) {
  // Two planes per thread by default for the plain kernels: they are
  // latency-bound, and more threads beat the reuse of a carried neighbour
  // plane (measured on 385 x 385 x 257 nodes: 11 chunks 1.73 ms per step,
  // 2-plane chunks 1.56 ms, 1-plane 1.57 ms). The kernels that load the
  // next plane ahead march 8 planes (H on the same grid: 0.548 ms with
  // 8-plane chunks, 0.565 ms with 32, 0.569 ms with 2; 0.647 ms plain). The
  // schedule's x chunks, where set, take precedence.
  int nch = h->nchunk > 0 ? h->nchunk : (pf ? (h->n[0] + 7) / 8 : (h->n[0] + 1) / 2);
  if (nch < 1) nch = 1;
  if (nch > h->n[0]) nch = h->n[0];
  int chunk = (h->n[0] + nch - 1) / nch;
  *nchunks = (h->n[0] + chunk - 1) / chunk;
  return chunk;
}

Lanes lanes_arg(const ems_gpu *h) {
  Lanes L;
  for (int l = 0; l < 12; l++) L.l[l] = h->lanes[l];
  return L;
}

bool has_hsrc(const ems_gpu *h) {
  for (auto &o : h->ops)
    if (o.type == OP_SOURCE && (o.kind == 2 || o.kind == 3)) return true;
  return false;
}

bool has_mid_ops(const ems_gpu *h) {
  for (auto &o : h->ops)
    if ((o.type == OP_SOURCE && (o.kind == 2 || o.kind == 3)) ||
        (o.type == OP_PROBE && o.quantity == 1) ||
        (o.type == OP_LUMPED && (o.lp.branch || o.lp.record)) || o.type == OP_PW ||
        o.type == OP_MERGED)
      return true;
  return false;
}

// Whether steps run fused: not while an H source needs the serial gap
// between the H and E updates (GAP_B, mod.rs:899-913; pool.rs:870-875).
// The fused kernel keeps the slabs that reach a tile in a 32-bit mask.
bool has_magnetic_poles(const ems_gpu *h) {
  for (auto &o : h->ops)
    if (o.type == OP_POLES && o.side[1].n > 0) return true;
  return false;
}

// Operators with kernels of their own in the gaps (plane waves, reduced
// currents of cut faces): they split the gaps and need the H-E gap.
bool own_gap_kernels(const HostOp &o) { return o.type == OP_PW || o.type == OP_MERGED; }
bool has_gap_kernels(const ems_gpu *h) {
  for (auto &o : h->ops)
    if (own_gap_kernels(o)) return true;
  return false;
}

bool has_merged(const ems_gpu *h) {
  for (auto &o : h->ops)
    if (o.type == OP_MERGED) return true;
  return false;
}

bool has_upml(const ems_gpu *h) {
  for (auto &o : h->ops)
    if (o.type == OP_UPML) return true;
  return false;
}

bool use_fused(const ems_gpu *h) {
  size_t ns = 0;
  for (auto &s : h->slabs) ns += s.len != 0;
  return h->fused && !has_hsrc(h) && !has_magnetic_poles(h) && !has_gap_kernels(h) &&
         !has_upml(h) && ns <= 32;
}

// The dispersive media as slab hooks (dispersive.rs:1700-1750), in
// registration order. stage 0 (after the H update and CPML-H): correct the
// magnetic poles, save u of the electric edges. stage 1 (after the E update
// and CPML-E): correct the electric poles, save i of the magnetic edges;
// with fp_prev the electric edges' u^n is read from `fp` instead.
int launch_poles(ems_gpu *h, int stage, const Fields &f, const Fields &fp, bool fp_prev) {
  for (auto &o : h->ops) {
    if (o.type != OP_POLES) continue;
    const DevPoleSide &se = o.side[0], &sh = o.side[1];
    const int nz = h->n[2];
    if (stage == 0) {
      if (sh.n)
        hipLaunchKernelGGL(poles_correct<false>, dim3((sh.n + 255) / 256), dim3(256), 0, h->stream,
                           f, f, 0, lanes_arg(h), nz, sh, (const DevPoleConst *)o.d_consts, h->dt);
      if (se.n)
        hipLaunchKernelGGL(poles_capture<true>, dim3((se.n + 255) / 256), dim3(256), 0, h->stream,
                           f, se);
    } else {
      if (se.n)
        hipLaunchKernelGGL(poles_correct<true>, dim3((se.n + 255) / 256), dim3(256), 0, h->stream,
                           f, fp, fp_prev ? 1 : 0, lanes_arg(h), nz, se,
                           (const DevPoleConst *)o.d_consts, h->dt);
      if (sh.n)
        hipLaunchKernelGGL(poles_capture<false>, dim3((sh.n + 255) / 256), dim3(256), 0, h->stream,
                           f, sh);
    }
    CK(h, hipGetLastError());
  }
  return 0;
}

// Whether side `side` (0 H, 1 E) corrects its CPML slabs inside its
// unfused update: bit 2 + side of the unfused switch with the side's
// loading-ahead kernel (bit side), and the slab table built for it (every slab but the z-normal
// suffix, whose columns would run the slab code in a few lanes of most
// waves; those keep their launches, after the kernel, so the corrections
// of a node keep the registration order; at most 32 for the planes'
// masks).
bool cpml_folded(const ems_gpu *h, int side) {
  if (use_fused(h) || !(h->unfused_pf & (4 << side)) || !(h->unfused_pf & (1 << side)) ||
      !h->d_fslabs || !h->d_pmask || h->split_state != 2)
    return false;
  int n = 0;
  for (auto &s : h->slabs) n += s.len != 0 && !s.out;
  return n > 0 && n <= 32;
}

int n_fslabs(const ems_gpu *h);

#define UNFUSED_LAUNCH(K, grid_, cp_on)                                                   \
  do {                                                                                    \
    if (h->oe) {                                                                          \
      if (cp_on)                                                                          \
        hipLaunchKernelGGL((K<true, true>), grid_, dim3(256), 0, h->stream, f, lanes_arg(h), \
                           h->n[0], h->n[1], h->n[2], chunk, cp);                         \
      else                                                                                \
        hipLaunchKernelGGL((K<true, false>), grid_, dim3(256), 0, h->stream, f, lanes_arg(h), \
                           h->n[0], h->n[1], h->n[2], chunk, cp);                         \
    } else {                                                                              \
      if (cp_on)                                                                          \
        hipLaunchKernelGGL((K<false, true>), grid_, dim3(256), 0, h->stream, f, lanes_arg(h), \
                           h->n[0], h->n[1], h->n[2], chunk, cp);                         \
      else                                                                                \
        hipLaunchKernelGGL((K<false, false>), grid_, dim3(256), 0, h->stream, f,           \
                           lanes_arg(h), h->n[0], h->n[1], h->n[2], chunk, cp);           \
    }                                                                                     \
  } while (0)

int launch_h(ems_gpu *h, const Fields &f) {
  int nch, chunk = x_chunk(h, &nch, (h->unfused_pf & 1) != 0);
  dim3 grid((unsigned)((h->plane + 255) / 256), (unsigned)nch);
  // the loading-ahead kernel runs one-dimensional (see xcd_contiguous; the
  // two-plane kernel measured slower so: E 0.526 against 0.503 ms)
  const dim3 line(grid.x * grid.y);
  if (h->unfused_pf & 1) {
    const bool cpf = cpml_folded(h, 0);
    const CpmlIn cp{h->d_fslabs, h->d_pmask, cpf ? n_fslabs(h) : 0, h->par};
    UNFUSED_LAUNCH(h_update_pf, line, cpf);
  } else if (h->oe)
    hipLaunchKernelGGL(h_update<true>, grid, dim3(256), 0, h->stream, f, lanes_arg(h), h->n[0],
                       h->n[1], h->n[2], chunk);
  else
    hipLaunchKernelGGL(h_update<false>, grid, dim3(256), 0, h->stream, f, lanes_arg(h), h->n[0],
                       h->n[1], h->n[2], chunk);
  CK(h, hipGetLastError());
  return 0;
}

int launch_e(ems_gpu *h, const Fields &f) {
  int nch, chunk = x_chunk(h, &nch, (h->unfused_pf & 2) != 0);
  dim3 grid((unsigned)((h->plane + 255) / 256), (unsigned)nch);
  const dim3 line(grid.x * grid.y);
  if (h->unfused_pf & 2) {
    const bool cpf = cpml_folded(h, 1);
    const CpmlIn cp{h->d_fslabs, h->d_pmask, cpf ? n_fslabs(h) : 0, 0};
    UNFUSED_LAUNCH(e_update_pf, line, cpf);
  } else if (h->oe)
    hipLaunchKernelGGL(e_update<true>, grid, dim3(256), 0, h->stream, f, lanes_arg(h), h->n[0],
                       h->n[1], h->n[2], chunk);
  else
    hipLaunchKernelGGL(e_update<false>, grid, dim3(256), 0, h->stream, f, lanes_arg(h), h->n[0],
                       h->n[1], h->n[2], chunk);
  CK(h, hipGetLastError());
  return 0;
}
#undef UNFUSED_LAUNCH

int launch_cpml(ems_gpu *h, const Fields &f, int e_side, int par) {
  const bool folded = cpml_folded(h, e_side);  // then only the suffix runs here
  for (auto &s : h->slabs) {
    if (s.arg.e_side != e_side || s.len == 0 || (folded && !s.out)) continue;
    const Lane cf = h->lanes[e_side ? 3 + s.arg.comp : 9 + s.arg.comp];
    SlabArg a = s.arg;
    a.psi = e_side ? s.d_psi[0] : s.d_psi[par];
    unsigned blocks = (unsigned)((s.len + 255) / 256);
    hipLaunchKernelGGL(cpml_slab, dim3(blocks), dim3(256), 0, h->stream, f, cf, a, h->n[0],
                       h->n[1], h->n[2]);
    CK(h, hipGetLastError());
  }
  return 0;
}

struct FusedGeom {
  int tk, ktiles, jtiles, chunk, nchunks, ntiles, nblocks;
};

FusedGeom fused_geom(const ems_gpu *h) {
  FusedGeom g;
  const int nz = h->n[2], ny = h->n[1], nx = h->n[0];
  // Owned k per tile <= 63 (one wave with a halo lane), spread evenly.
  g.ktiles = (nz + 62) / 63;
  g.tk = (nz + g.ktiles - 1) / g.ktiles;
  g.jtiles = (ny + h->tj - 1) / h->tj;
  int want = h->nchunk;
  if (want <= 0) {
    // ~2048 blocks (eight per CU) unless that makes chunks very short.
    const int per = g.ktiles * g.jtiles;
    want = (2048 + per - 1) / per;
    want = std::max(1, std::min(want, std::max(1, nx / 16)));
  }
  want = std::min(want, nx);
  g.chunk = (nx + want - 1) / want;
  g.nchunks = (nx + g.chunk - 1) / g.chunk;
  g.ntiles = g.ktiles * g.jtiles * g.nchunks;
  const int rows = g.jtiles * g.nchunks, G = h->xcd_group;
  const int prow = G > 0 ? (rows + 8 * G - 1) / (8 * G) * (8 * G) : rows;
  g.nblocks = prow * g.ktiles;
  return g;
}

// Marks the slabs the tiled fused step corrects outside its kernel: per
// side the longest suffix of that side's slabs (registration order) that
// are normal to z, so that every correction of a node's component outside
// the kernel follows those inside it. E-side ones must not correct one
// component at one node twice (they run in one launch); otherwise that
// side keeps all its slabs inside.
void mark_out_slabs(ems_gpu *h, bool split) {
  for (auto &s : h->slabs) s.out = false;
  if (!split) return;
  for (int side = 0; side < 2; side++) {
    std::vector<HostSlab *> suf;
    for (auto it = h->slabs.rbegin(); it != h->slabs.rend(); ++it) {
      if (it->len == 0 || it->arg.e_side != side) continue;
      if (it->arg.normal != 2) break;
      suf.push_back(&*it);
    }
    bool ok = !suf.empty() && suf.size() <= 32;
    for (size_t a = 0; ok && side == 1 && a < suf.size(); a++)
      for (size_t b = a + 1; ok && b < suf.size(); b++) {
        const SlabArg &x = suf[a]->arg, &y = suf[b]->arg;
        bool meet = x.comp == y.comp;
        for (int d = 0; d < 3 && meet; d++) meet = x.lo[d] < y.hi[d] && y.lo[d] < x.hi[d];
        ok = !meet;
      }
    if (ok)
      for (HostSlab *q : suf) q->out = true;
  }
}

int upload_out_slabs(ems_gpu *h) {
  if (h->d_hout) CK(h, hipFree(h->d_hout));
  if (h->d_eout) CK(h, hipFree(h->d_eout));
  if (h->d_pmask_o) CK(h, hipFree(h->d_pmask_o));
  h->d_hout = h->d_eout = nullptr;
  h->d_pmask_o = nullptr;
  h->n_hout = h->n_eout = 0;
  h->hout_total = h->eout_total = 0;
  std::vector<DSlab> hv, ev;
  for (auto &s : h->slabs) {
    if (!s.out) continue;
    const SlabArg &a = s.arg;
    if (!a.e_side && !s.d_delta) CK(h, hipMalloc(&s.d_delta, s.len * sizeof(float)));
    if ((a.e_side ? h->eout_total : h->hout_total) + s.len >= (1ull << 31))
      return fail(h, EMS_E_UNSUPPORTED, "CPML slabs over 2^31 nodes");
    DSlab d;
    d.comp = a.comp, d.normal = a.normal, d.other = a.other, d.sign = a.sign;
    d.lo0 = a.lo[0], d.hi0 = a.hi[0], d.lo1 = a.lo[1], d.hi1 = a.hi[1];
    d.lo2 = a.lo[2], d.hi2 = a.hi[2];
    d.ly = a.hi[1] - a.lo[1], d.lz = a.hi[2] - a.lo[2];
    d.b = s.d_b, d.a = s.d_a, d.k = s.d_k;
    d.psi0 = s.d_psi[0];
    d.psi1 = a.e_side ? s.d_psi[0] : s.d_psi[1];
    d.delta = s.d_delta;
    uint32_t &tot = a.e_side ? h->eout_total : h->hout_total;
    d.first = tot;
    tot += (uint32_t)s.len;
    (a.e_side ? ev : hv).push_back(d);
  }
  h->n_hout = (int)hv.size(), h->n_eout = (int)ev.size();
  if (!hv.empty()) {
    CK(h, hipMalloc(&h->d_hout, hv.size() * sizeof(DSlab)));
    CK(h, hipMemcpy(h->d_hout, hv.data(), hv.size() * sizeof(DSlab), hipMemcpyHostToDevice));
    std::vector<uint32_t> pm((size_t)h->n[0], 0u);
    for (size_t q = 0; q < hv.size(); q++)
      for (int i = hv[q].lo0; i < hv[q].hi0; i++) pm[i] |= 1u << q;
    CK(h, hipMalloc(&h->d_pmask_o, pm.size() * sizeof(uint32_t)));
    CK(h, hipMemcpy(h->d_pmask_o, pm.data(), pm.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  if (!ev.empty()) {
    CK(h, hipMalloc(&h->d_eout, ev.size() * sizeof(DSlab)));
    CK(h, hipMemcpy(h->d_eout, ev.data(), ev.size() * sizeof(DSlab), hipMemcpyHostToDevice));
  }
  return 0;
}

// state: 0 every slab in the table; 1 the fused kernel's split (the
// z-normal suffix corrected by its passes outside it); 2 the same suffix
// left to the unfused schedule's own slab launches (no outside tables).
int upload_fslabs(ems_gpu *h, int state) {
  if (!h->slabs_dirty && h->split_state == state) return 0;
  mark_out_slabs(h, state != 0);
  int rc = state == 2 ? 0 : upload_out_slabs(h);
  if (rc) return rc;
  if (state == 2) {
    if (h->d_hout) CK(h, hipFree(h->d_hout));
    if (h->d_eout) CK(h, hipFree(h->d_eout));
    if (h->d_pmask_o) CK(h, hipFree(h->d_pmask_o));
    h->d_hout = h->d_eout = nullptr;
    h->d_pmask_o = nullptr;
    h->n_hout = h->n_eout = 0;
    h->hout_total = h->eout_total = 0;
  }
  h->split_state = state;
  h->slab_gen++;
  if (h->d_fslabs) CK(h, hipFree(h->d_fslabs));
  h->d_fslabs = nullptr;
  std::vector<FSlab> v;
  for (auto &s : h->slabs) {
    if (s.len == 0 || s.out) continue;
    FSlab f;
    f.e_side = s.arg.e_side, f.comp = s.arg.comp, f.normal = s.arg.normal, f.sign = s.arg.sign;
    f.lo0 = s.arg.lo[0], f.hi0 = s.arg.hi[0], f.lo1 = s.arg.lo[1], f.hi1 = s.arg.hi[1];
    f.lo2 = s.arg.lo[2], f.hi2 = s.arg.hi[2];
    f.ly = s.arg.hi[1] - s.arg.lo[1], f.lz = s.arg.hi[2] - s.arg.lo[2];
    f.b = s.d_b, f.a = s.d_a, f.k = s.d_k;
    f.psi0 = s.d_psi[0];
    f.psi1 = s.arg.e_side ? s.d_psi[0] : s.d_psi[1];
    v.push_back(f);
  }
  if (!v.empty()) {
    CK(h, hipMalloc(&h->d_fslabs, v.size() * sizeof(FSlab)));
    CK(h, hipMemcpy(h->d_fslabs, v.data(), v.size() * sizeof(FSlab), hipMemcpyHostToDevice));
  }
  std::vector<uint32_t> pm(2 * (size_t)h->n[0], 0u);
  for (size_t q = 0; q < v.size() && q < 32; q++)
    for (int i = v[q].lo0; i < v[q].hi0; i++) pm[2 * i + (v[q].e_side ? 1 : 0)] |= 1u << q;
  if (h->d_pmask) CK(h, hipFree(h->d_pmask));
  CK(h, hipMalloc(&h->d_pmask, pm.size() * sizeof(uint32_t)));
  CK(h, hipMemcpy(h->d_pmask, pm.data(), pm.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  h->slabs_dirty = false;
  return 0;
}

int n_fslabs(const ems_gpu *h) {
  int n = 0;
  for (auto &s : h->slabs) n += s.len != 0 && !s.out;
  return n;
}

// The passes of the slabs corrected outside the fused kernel: before it
// (increments from u^n in buffer par) and after it (on buffer 1 - par).
int launch_cpml_out(ems_gpu *h, int par, bool after) {
  if (!after && h->n_hout)
    hipLaunchKernelGGL(cpml_delta_h, dim3((h->hout_total + 255) / 256), dim3(256), 0, h->stream,
                       (const float *)h->fbuf[par], h->N, lanes_arg(h), (const DSlab *)h->d_hout,
                       h->n_hout, h->hout_total, par, h->n[0], h->n[1], h->n[2]);
  if (after && h->n_eout)
    hipLaunchKernelGGL(cpml_post_e, dim3((h->eout_total + 255) / 256), dim3(256), 0, h->stream,
                       h->fbuf[1 - par], h->N, lanes_arg(h), (const DSlab *)h->d_eout, h->n_eout,
                       h->eout_total, h->n[1], h->n[2]);
  CK(h, hipGetLastError());
  return 0;
}

template <bool OE, int TJ, bool PF, bool CP, typename CT>
void launch_fused_ct(ems_gpu *h, int par, const FusedGeom &g, int nblocks, const int *tiles) {
  const float *A = h->fbuf[par];
  float *B = h->fbuf[1 - par];
  hipLaunchKernelGGL((step_fused<OE, TJ, PF, CP, CT>), dim3(nblocks), dim3(64, TJ + 1), 0,
                     h->stream, A, B, h->N, (const float *)h->cpool, (const CT *)h->croff,
                     h->d_fslabs, (const uint32_t *)h->d_pmask, n_fslabs(h), par,
                     (const DSlab *)h->d_hout, (const uint32_t *)h->d_pmask_o, h->n_hout, h->n[0],
                     h->n[1], h->n[2], g.tk, g.chunk, g.ktiles, g.jtiles, g.ntiles, h->xcd_group,
                     tiles);
}

template <bool OE, int TJ, bool PF, bool CP>
void launch_fused_one(ems_gpu *h, int par, const FusedGeom &g, int nblocks, const int *tiles) {
  if (nblocks <= 0) return;
  if (h->croff_wide)
    launch_fused_ct<OE, TJ, PF, CP, uint64_t>(h, par, g, nblocks, tiles);
  else
    launch_fused_ct<OE, TJ, PF, CP, uint32_t>(h, par, g, nblocks, tiles);
}

// The tiled step: one launch over every tile, or with tile classes the
// tiles a slab reaches first (the slower ones), then the others with the
// instantiation without slab code. The launches read A and write B at
// disjoint tiles, so their order does not matter for the result.
template <bool OE, int TJ, bool PF>
void launch_fused_oe(ems_gpu *h, int par, const FusedGeom &g) {
  if (h->tile_class && h->d_tiles) {
    launch_fused_one<OE, TJ, PF, true>(h, par, g, h->ntl[1], h->d_tiles + h->ntl[0]);
    launch_fused_one<OE, TJ, PF, false>(h, par, g, h->ntl[0], h->d_tiles);
  } else {
    launch_fused_one<OE, TJ, PF, true>(h, par, g, g.nblocks, nullptr);
  }
}

template <int TJ, bool PF>
void launch_fused_tj(ems_gpu *h, int par, const FusedGeom &g) {
  if (h->oe)
    launch_fused_oe<true, TJ, PF>(h, par, g);
  else
    launch_fused_oe<false, TJ, PF>(h, par, g);
}

// Full-row kernel: one wave per row, rows of up to 9 segments of 64.
template <bool OE, int TJ, int NSEG>
void launch_rows_n(ems_gpu *h, int par, int chunk, int jtiles, int ntiles, int nblocks) {
  const float *A = h->fbuf[par];
  float *B = h->fbuf[1 - par];
  hipLaunchKernelGGL((step_rows<OE, TJ, NSEG>), dim3(nblocks), dim3(64, TJ + 1), 0, h->stream, A,
                     B, h->N, (const float *)h->cpool, (const uint32_t *)h->croff, h->d_fslabs,
                     (const uint32_t *)h->d_pmask, n_fslabs(h), par, h->n[0], h->n[1], h->n[2],
                     chunk, jtiles, ntiles, h->xcd_group);
}

template <bool OE, int TJ>
int launch_rows_tj(ems_gpu *h, int par, int chunk, int jtiles, int ntiles, int nblocks) {
  switch ((h->n[2] + 63) / 64) {
    case 1: launch_rows_n<OE, TJ, 1>(h, par, chunk, jtiles, ntiles, nblocks); break;
    case 2: launch_rows_n<OE, TJ, 2>(h, par, chunk, jtiles, ntiles, nblocks); break;
    case 3: launch_rows_n<OE, TJ, 3>(h, par, chunk, jtiles, ntiles, nblocks); break;
    case 4: launch_rows_n<OE, TJ, 4>(h, par, chunk, jtiles, ntiles, nblocks); break;
    case 5: launch_rows_n<OE, TJ, 5>(h, par, chunk, jtiles, ntiles, nblocks); break;
    case 6: launch_rows_n<OE, TJ, 6>(h, par, chunk, jtiles, ntiles, nblocks); break;
    case 7: launch_rows_n<OE, TJ, 7>(h, par, chunk, jtiles, ntiles, nblocks); break;
    case 8: launch_rows_n<OE, TJ, 8>(h, par, chunk, jtiles, ntiles, nblocks); break;
    case 9: launch_rows_n<OE, TJ, 9>(h, par, chunk, jtiles, ntiles, nblocks); break;
    default: return fail(h, EMS_E_UNSUPPORTED, "row kernel: nz above 576");
  }
  return 0;
}

// Rows mode needs nz <= 576 (nine segments) and the tile height 3 or 7.
bool rows_ok(const ems_gpu *h) { return h->n[2] <= 576 && (h->tj == 3 || h->tj == 7); }

int launch_rows(ems_gpu *h, int par) {
  const int nx = h->n[0], ny = h->n[1];
  const int jtiles = (ny + h->tj - 1) / h->tj;
  int want = h->nchunk;
  if (want <= 0) want = std::max(1, std::min(nx / 16, (1024 + jtiles - 1) / jtiles));
  want = std::min(want, nx);
  const int chunk = (nx + want - 1) / want;
  const int nchunks = (nx + chunk - 1) / chunk;
  const int ntiles = jtiles * nchunks, G = h->xcd_group;
  const int nblocks = G > 0 ? (ntiles + 8 * G - 1) / (8 * G) * (8 * G) : ntiles;
  int rc;
  if (h->tj == 3)
    rc = h->oe ? launch_rows_tj<true, 3>(h, par, chunk, jtiles, ntiles, nblocks)
               : launch_rows_tj<false, 3>(h, par, chunk, jtiles, ntiles, nblocks);
  else
    rc = h->oe ? launch_rows_tj<true, 7>(h, par, chunk, jtiles, ntiles, nblocks)
               : launch_rows_tj<false, 7>(h, par, chunk, jtiles, ntiles, nblocks);
  if (rc) return rc;
  CK(h, hipGetLastError());
  return 0;
}

// Packed short rows (nz <= 64): R = 64 / nz rows per wave, PACK_NW waves
// per block, the first row of a block its halo.
constexpr int PACK_NW = 4;

bool packed_ok(const ems_gpu *h) { return h->n[2] <= 64; }

FusedGeom packed_geom(const ems_gpu *h, int nchunk) {
  FusedGeom g;
  const int nz = h->n[2], ny = h->n[1], nx = h->n[0];
  g.tk = 64 / nz;  // rows per wave
  g.ktiles = 1;
  const int owned = PACK_NW * g.tk - 1;
  g.jtiles = (ny + owned - 1) / owned;
  int want = nchunk;
  if (want <= 0) {
    want = (2048 + g.jtiles - 1) / g.jtiles;
    want = std::max(1, std::min(want, std::max(1, nx / 16)));
  }
  want = std::min(want, nx);
  g.chunk = (nx + want - 1) / want;
  g.nchunks = (nx + g.chunk - 1) / g.chunk;
  g.ntiles = g.jtiles * g.nchunks;
  const int G = h->xcd_group;
  g.nblocks = G > 0 ? (g.ntiles + 8 * G - 1) / (8 * G) * (8 * G) : g.ntiles;
  return g;
}

template <bool OE, bool PF>
void launch_packed_t(ems_gpu *h, int par, const FusedGeom &g) {
  hipLaunchKernelGGL((step_packed<OE, PACK_NW, PF>), dim3(g.nblocks), dim3(64, PACK_NW), 0,
                     h->stream, h->fbuf[par], h->fbuf[1 - par], h->N, (const float *)h->cpool,
                     (const uint32_t *)h->croff, h->d_fslabs, (const uint32_t *)h->d_pmask,
                     n_fslabs(h), par, h->n[0], h->n[1], h->n[2], g.tk, g.chunk, g.jtiles,
                     g.ntiles, h->xcd_group);
}

int launch_packed(ems_gpu *h, int par) {
  const FusedGeom g = packed_geom(h, h->nchunk);
  if (h->oe)
    h->prefetch ? launch_packed_t<true, true>(h, par, g) : launch_packed_t<true, false>(h, par, g);
  else
    h->prefetch ? launch_packed_t<false, true>(h, par, g) : launch_packed_t<false, false>(h, par, g);
  CK(h, hipGetLastError());
  return 0;
}

// Band kernel: windows of 63 k-nodes per row, bands of TB rows, x chunks
// sized for about 8 waves per SIMD unless chunks get very short.
FusedGeom band_geom(const ems_gpu *h) {
  FusedGeom g;
  const int nz = h->n[2], ny = h->n[1], nx = h->n[0];
  g.ktiles = (nz + 62) / 63;
  g.tk = 63;
  g.jtiles = (ny + h->tj - 1) / h->tj;
  int want = h->nchunk;
  if (want <= 0) {
    const int per = g.ktiles * g.jtiles;
    want = (8192 + per - 1) / per;
    want = std::max(1, std::min(want, std::max(1, nx / 8)));
  }
  want = std::min(want, nx);
  g.chunk = (nx + want - 1) / want;
  g.nchunks = (nx + g.chunk - 1) / g.chunk;
  g.ntiles = g.ktiles * g.jtiles * g.nchunks;
  const int rows = g.jtiles * g.nchunks, G = h->xcd_group;
  const int prow = G > 0 ? (rows + 8 * G - 1) / (8 * G) * (8 * G) : rows;
  g.nblocks = prow * g.ktiles;
  return g;
}

template <bool OE, int TB>
void launch_band_t(ems_gpu *h, int par, const FusedGeom &g) {
  hipLaunchKernelGGL((step_band<OE, TB>), dim3(g.nblocks), dim3(64), 0, h->stream, h->fbuf[par],
                     h->fbuf[1 - par], h->N, (const float *)h->cpool, (const uint32_t *)h->croff,
                     h->d_fslabs, (const uint32_t *)h->d_pmask, n_fslabs(h), par, h->n[0],
                     h->n[1], h->n[2], g.chunk, g.ktiles, g.jtiles, g.ntiles, h->xcd_group);
}

int launch_band(ems_gpu *h, int par) {
  const FusedGeom g = band_geom(h);
  switch (h->tj) {
    case 4: h->oe ? launch_band_t<true, 4>(h, par, g) : launch_band_t<false, 4>(h, par, g); break;
    case 8: h->oe ? launch_band_t<true, 8>(h, par, g) : launch_band_t<false, 8>(h, par, g); break;
    case 16:
      h->oe ? launch_band_t<true, 16>(h, par, g) : launch_band_t<false, 16>(h, par, g);
      break;
    default: return fail(h, EMS_E_INVALID, "band kernel: tile_rows must be 4, 8 or 16");
  }
  CK(h, hipGetLastError());
  return 0;
}

int launch_fused(ems_gpu *h, int par) {
  // (the band, packed and full-row kernels take 32-bit row offsets: with a
  // pool of 4 GiB or more the tiled kernel runs)
  if (h->band && !h->croff_wide) return launch_band(h, par);
  if (h->packed && packed_ok(h) && !h->croff_wide) return launch_packed(h, par);
  if (h->rows && rows_ok(h) && !h->croff_wide) return launch_rows(h, par);
  const FusedGeom g = fused_geom(h);
  const int key = h->tj * 2 + (h->prefetch ? 1 : 0);
  int rc = launch_cpml_out(h, par, false);
  if (rc) return rc;
  switch (key) {
    case 6: launch_fused_tj<3, false>(h, par, g); break;
    case 7: launch_fused_tj<3, true>(h, par, g); break;
    case 14: launch_fused_tj<7, false>(h, par, g); break;
    case 15: launch_fused_tj<7, true>(h, par, g); break;
    case 30: launch_fused_tj<15, false>(h, par, g); break;
    case 31: launch_fused_tj<15, true>(h, par, g); break;
    default: return fail(h, EMS_E_INVALID, "unsupported tile height");
  }
  CK(h, hipGetLastError());
  return launch_cpml_out(h, par, true);
}

// A plane wave's hooks in a gap: the corrections, then its auxiliary grid
// (stage 0 after the H update, 1 after the E update).
int launch_pw(ems_gpu *h, const HostOp &o, const Fields &f, int stage) {
  const DevPwCorr &c = o.pwc[stage == 0 ? 1 : 0];
  if (c.ntgt) {
    if (stage == 0)
      hipLaunchKernelGGL(pw_correct<false>, dim3((c.ntgt + 255) / 256), dim3(256), 0, h->stream,
                         f, c, (const double *)o.pw.e);
    else
      hipLaunchKernelGGL(pw_correct<true>, dim3((c.ntgt + 255) / 256), dim3(256), 0, h->stream,
                         f, c, (const double *)o.pw.h);
  }
  const unsigned nb = (unsigned)((o.pw.naux + 254) / 256);
  if (stage == 0)
    hipLaunchKernelGGL(pw_aux_h, dim3(nb), dim3(256), 0, h->stream, o.pw);
  else
    hipLaunchKernelGGL(pw_aux_e, dim3(nb), dim3(256), 0, h->stream, o.pw,
                       (const uint64_t *)h->d_step);
  CK(h, hipGetLastError());
  return 0;
}

// The reduced currents' hooks (MergedFaces::post_h / post_e,
// conformal.rs:2655-2682), split by what they read: stage 2 the
// circulations of u^n and the update (they read u, which the H half-step
// leaves alone, and the operator's own state, so they run beside the H
// update on the side stream); stage 0, at the operator's place in the mid
// gap, writes the driven faces' currents; stage 1 the circulations of the
// energy look-ahead in the post-E gap.
int launch_merged(ems_gpu *h, const HostOp &o, const Fields &f, int stage, hipStream_t st) {
  const DevMerged &m = o.mf;
  if (stage != 0 && m.ng)
    hipLaunchKernelGGL(mf_circ, dim3((m.ng + 255) / 256), dim3(256), 0, st, f, m);
  if (stage != 0 && m.np) {
    hipLaunchKernelGGL(mf_pml_terms, dim3((m.np + 255) / 256), dim3(256), 0, st, f, m,
                       stage == 2 ? 1 : 0);
    hipLaunchKernelGGL(mf_pml_add, dim3((m.ng + 255) / 256), dim3(256), 0, st, m);
  }
  if (stage == 2) {
    // small blocks (and the merge groups) with a small LDS image, so that
    // many waves share a CU; then the large blocks
    const int na = o.mf_nsmall + (m.ns + 63) / 64;
    if (na)
      hipLaunchKernelGGL(mf_advance, dim3(na), dim3(64), (size_t)o.mf_lds_small, st, m,
                         (const uint32_t *)o.mf_small, o.mf_nsmall, o.mf_lds_small);
    if (o.mf_nbig)
      hipLaunchKernelGGL(mf_advance, dim3(o.mf_nbig), dim3(64), (size_t)o.mf_lds_big, st, m,
                         (const uint32_t *)o.mf_big, o.mf_nbig, o.mf_lds_big);
  }
  if (stage == 0) {
    const int nd = std::max(m.nf, m.ng);
    if (nd) hipLaunchKernelGGL(mf_drive, dim3((nd + 255) / 256), dim3(256), 0, st, f, m);
  }
  CK(h, hipGetLastError());
  return 0;
}

// An own-kernel operator's hooks at its place in a gap: stage 0 the mid
// gap, 1 the post-E gap. The reduced currents' update itself (stage 2 of
// launch_merged) ran before, beside the H update (enqueue_step).
// The reduced currents' post-E stage (the look-ahead circulations of u^n+1)
// is read only by the energy's look-ahead and the state read-back: the next
// step computes the same circulations again beside its H update before it
// uses them. So the post-E gap leaves it, steps mark it stale, and
// merged_circ_now runs it, as the gap would have, when it is read.
int launch_gap_op(ems_gpu *h, const HostOp &o, const Fields &f, int stage) {
  if (o.type == OP_MERGED && stage == 1) return 0;
  return o.type == OP_PW ? launch_pw(h, o, f, stage)
                         : launch_merged(h, o, f, stage, h->stream);
}

void merged_mark_stale(ems_gpu *h) {
  for (auto &o : h->ops)
    if (o.type == OP_MERGED) o.circ_stale = true;
}

int merged_circ_now(ems_gpu *h, HostOp &o) {
  if (!o.circ_stale) return 0;
  int rc = launch_merged(h, o, h->fb[h->par], 1, h->stream);
  if (rc) return rc;
  o.circ_stale = false;
  return 0;
}

// Whether a gap segment [q0, q1) between own-kernel operators has work, so
// that empty ones are not launched (what run_gap, gap_mid and gap_post_e do
// per operator kind): mid gap, the H sources and current probes in it, and
// the first segment the lumped elements' capture of u^n; post-E gap, the E
// sources, lumped elements and voltage probes in it, and the first segment
// the current probes moved there and the capture when asked. The last
// post-E segment always runs: it advances the step counter.
bool lumped_captures(const ems_gpu *h) {
  for (auto &o : h->ops)
    if (o.type == OP_LUMPED && (o.lp.branch || o.lp.record)) return true;
  return false;
}

bool mid_segment_work(const ems_gpu *h, int q0, int q1) {
  for (int q = q0; q < q1; q++) {
    const HostOp &o = h->ops[q];
    if ((o.type == OP_SOURCE && (o.kind == 2 || o.kind == 3)) ||
        (o.type == OP_PROBE && o.quantity == 1))
      return true;
  }
  return q0 == 0 && lumped_captures(h);
}

bool post_segment_work(const ems_gpu *h, int q0, int q1, bool capture) {
  for (int q = q0; q < q1; q++) {
    const HostOp &o = h->ops[q];
    if ((o.type == OP_SOURCE && (o.kind == 0 || o.kind == 1)) || o.type == OP_LUMPED ||
        (o.type == OP_PROBE && o.quantity == 0))
      return true;
  }
  if (q0 != 0) return false;
  if (capture && lumped_captures(h)) return true;
  for (auto &o : h->ops)
    if (o.type == OP_PROBE && o.quantity == 1) return true;
  return false;
}

// Post-E gap on the result `f`; `fp` holds the state before a fused step
// (prev = true), whose u the lumped pre_e hooks capture. Plane waves run
// at their place among the gap's ops.
int launch_gap_post_e(ems_gpu *h, const Fields &f, const Fields &fp, bool prev) {
  const int nops = (int)h->ops.size();
  int q0 = 0;
  for (int q = 0; q <= nops; q++) {
    if (q < nops && !own_gap_kernels(h->ops[q])) continue;
    const bool last = q == nops;
    if (last || post_segment_work(h, q0, q, prev && q0 == 0))
      hipLaunchKernelGGL(gap_post_e, dim3(1), dim3(256), 0, h->stream, f, fp,
                         prev && q0 == 0 ? 1 : 0, h->d_ops, nops, q0, q, last ? 1 : 0,
                         h->d_step);
    CK(h, hipGetLastError());
    if (!last) {
      int rc = launch_gap_op(h, h->ops[q], f, 1);
      if (rc) return rc;
    }
    q0 = q + 1;
  }
  return 0;
}

int launch_gap_mid(ems_gpu *h, const Fields &f) {
  if (!has_mid_ops(h)) return 0;
  const int nops = (int)h->ops.size();
  int q0 = 0;
  for (int q = 0; q <= nops; q++) {
    if (q < nops && !own_gap_kernels(h->ops[q])) continue;
    if (mid_segment_work(h, q0, q))
      hipLaunchKernelGGL(gap_mid, dim3(1), dim3(256), 0, h->stream, f, h->d_ops, nops, q0, q,
                         (const uint64_t *)h->d_step);
    CK(h, hipGetLastError());
    if (q < nops) {
      int rc = launch_gap_op(h, h->ops[q], f, 0);
      if (rc) return rc;
    }
    q0 = q + 1;
  }
  return 0;
}

// The Mur boundaries: stage 0 saves u_b^n, u_1^n from `f` (before the E
// update), stage 1 updates the face edges and then the corners of `f`.
int launch_mur(ems_gpu *h, const Fields &f, int stage) {
  for (auto &o : h->ops) {
    if (o.type != OP_MUR) continue;
    const DevMur &m = o.mur;
    if (stage == 0) {
      if (m.ne + m.nc)
        hipLaunchKernelGGL(mur_capture, dim3((m.ne + m.nc + 255) / 256), dim3(256), 0, h->stream, f,
                           m);
    } else {
      if (m.ne)
        hipLaunchKernelGGL(mur_edges, dim3((m.ne + 255) / 256), dim3(256), 0, h->stream, f, m,
                           (const uint64_t *)h->d_step);
      if (m.nc)
        hipLaunchKernelGGL(mur_corners, dim3((m.nc + 255) / 256), dim3(256), 0, h->stream, f, m,
                           (const uint64_t *)h->d_step);
    }
    CK(h, hipGetLastError());
  }
  return 0;
}

// openEMS's PML layers: stage 0 right after the H update, stage 1 right
// after the E update (first among the slab corrections).
int launch_upml(ems_gpu *h, const Fields &f, int stage) {
  for (auto &o : h->ops) {
    if (o.type != OP_UPML) continue;
    const int n = o.upml.ne + o.upml.nh;
    if (n)
      hipLaunchKernelGGL(upml_stage, dim3((n + 255) / 256), dim3(256), 0, h->stream, f, o.upml,
                         stage);
    CK(h, hipGetLastError());
  }
  return 0;
}

// Unfused steps with reduced currents: their update forks onto the side
// stream before the H update and joins before the mid gap.
int merged_fork(ems_gpu *h, const Fields &f) {
  if (!h->side || !has_merged(h)) return 0;
  CK(h, hipEventRecord(h->ev_fork, h->stream));
  CK(h, hipStreamWaitEvent(h->side, h->ev_fork, 0));
  for (auto &o : h->ops) {
    if (o.type != OP_MERGED) continue;
    int rc = launch_merged(h, o, f, 2, h->side);
    if (rc) return rc;
  }
  CK(h, hipEventRecord(h->ev_join, h->side));
  return 0;
}

int merged_join(ems_gpu *h) {
  if (h->side && has_merged(h)) CK(h, hipStreamWaitEvent(h->stream, h->ev_join, 0));
  return 0;
}

// One step starting from buffer `par`; returns the buffer of the result.
int enqueue_step(ems_gpu *h, int &par) {
  int rc;
  if (use_fused(h)) {
    if ((rc = launch_fused(h, par))) return rc;
    par ^= 1;
    if ((rc = launch_poles(h, 1, h->fb[par], h->fb[par ^ 1], true))) return rc;
    // (u^n is still in the previous buffer)
    if ((rc = launch_mur(h, h->fb[par ^ 1], 0))) return rc;
    if ((rc = launch_mur(h, h->fb[par], 1))) return rc;
    return launch_gap_post_e(h, h->fb[par], h->fb[par ^ 1], true);
  }
  const Fields &f = h->fb[par];
  if ((rc = merged_fork(h, f))) return rc;
  if ((rc = launch_h(h, f))) return rc;
  if ((rc = launch_upml(h, f, 0))) return rc;
  if ((rc = launch_cpml(h, f, 0, par))) return rc;
  if ((rc = launch_poles(h, 0, f, f, false))) return rc;
  if ((rc = launch_mur(h, f, 0))) return rc;
  if ((rc = merged_join(h))) return rc;
  if ((rc = launch_gap_mid(h, f))) return rc;
  if ((rc = launch_e(h, f))) return rc;
  if ((rc = launch_upml(h, f, 1))) return rc;
  if ((rc = launch_cpml(h, f, 1, par))) return rc;
  if ((rc = launch_poles(h, 1, f, f, false))) return rc;
  if ((rc = launch_mur(h, f, 1))) return rc;
  return launch_gap_post_e(h, f, f, false);
}

void drop_graphs(ems_gpu *h) {
  for (int p = 0; p < 2; p++) {
    if (h->g1[p]) (void)hipGraphExecDestroy(h->g1[p]);
    if (h->gK[p]) (void)hipGraphExecDestroy(h->gK[p]);
    h->g1[p] = h->gK[p] = nullptr;
  }
}

int capture(ems_gpu *h, int steps, int par, hipGraphExec_t *out) {
  hipGraph_t g;
  CK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < steps && !rc; s++) rc = enqueue_step(h, par);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) return rc;
  CK(h, e);
  CK(h, hipGraphInstantiate(out, g, nullptr, nullptr, 0));
  CK(h, hipGraphDestroy(g));
  return 0;
}

// Uploads the operator and slab tables, and grows probe buffers to hold
// `steps` more steps of samples.
// Moves every lane's rows into one pool (a zero row first) and builds the
// per-row offset table the fused kernel reads; the lanes then point into
// the pool, so each coefficient is stored once.
int build_pool(ems_gpu *h) {
  const int nz = h->n[2];
  const size_t rows = (size_t)h->n[0] * h->n[1];
  uint64_t base[12] = {0};
  uint64_t total = (uint64_t)nz;  // the zero row
  for (int l = 0; l < 12; l++)
    if (h->lanes[l].rows) {
      base[l] = total;
      total += h->lane_rows[l] * (uint64_t)nz;
    }
  float *pool = nullptr;
  CK(h, hipMalloc(&pool, total * sizeof(float)));
  CK(h, hipMemset(pool, 0, (size_t)nz * sizeof(float)));
  for (int l = 0; l < 12; l++)
    if (h->lanes[l].rows)
      CK(h, hipMemcpy(pool + base[l], h->lanes[l].rows, h->lane_rows[l] * nz * sizeof(float),
                      hipMemcpyDeviceToDevice));
  // 32-bit row offsets while the pool stays under 4 GiB, else 64-bit ones
  const bool wide = total * sizeof(float) >= (1ull << 32);
  if (h->croff && wide != h->croff_wide) {
    CK(h, hipFree(h->croff));
    h->croff = nullptr;
  }
  h->croff_wide = wide;
  if (!h->croff)
    CK(h, hipMalloc(&h->croff, rows * 16 * (wide ? sizeof(uint64_t) : sizeof(uint32_t))));
  const uint32_t *ix[12];
  uint64_t b64[12];
  for (int l = 0; l < 12; l++) ix[l] = h->lanes[l].index, b64[l] = base[l];
  void *d_ix = nullptr, *d_b = nullptr;
  CK(h, hipMalloc(&d_ix, sizeof ix));
  CK(h, hipMalloc(&d_b, sizeof b64));
  CK(h, hipMemcpy(d_ix, ix, sizeof ix, hipMemcpyHostToDevice));
  CK(h, hipMemcpy(d_b, b64, sizeof b64, hipMemcpyHostToDevice));
  const size_t prows = (size_t)(total / (uint64_t)nz);
  uint8_t *d_flags = nullptr;
  CK(h, hipMalloc(&d_flags, prows));
  hipLaunchKernelGGL(pool_row_const, dim3((unsigned)((prows + 255) / 256)), dim3(256), 0,
                     h->stream, (const float *)pool, prows, nz, d_flags);
  CK(h, hipGetLastError());
  if (wide)
    hipLaunchKernelGGL(build_row_offsets<uint64_t>, dim3((unsigned)((rows + 255) / 256)), dim3(256),
                       0, h->stream, (uint64_t *)h->croff, (const uint32_t *const *)d_ix,
                       (const uint64_t *)d_b, (const uint8_t *)d_flags, rows, nz);
  else
    hipLaunchKernelGGL(build_row_offsets<uint32_t>, dim3((unsigned)((rows + 255) / 256)), dim3(256),
                       0, h->stream, (uint32_t *)h->croff, (const uint32_t *const *)d_ix,
                       (const uint64_t *)d_b, (const uint8_t *)d_flags, rows, nz);
  CK(h, hipGetLastError());
  CK(h, hipStreamSynchronize(h->stream));
  CK(h, hipFree(d_ix));
  CK(h, hipFree(d_b));
  CK(h, hipFree(d_flags));
  // Point the lanes into the pool and free their own copies.
  for (int l = 0; l < 12; l++)
    if (h->lanes[l].rows) {
      void *old = (void *)h->lanes[l].rows;
      h->lanes[l].rows = pool + base[l];
      for (auto it = h->lane_mem.begin(); it != h->lane_mem.end(); ++it)
        if (*it == old) {
          CK(h, hipFree(old));
          h->lane_mem.erase(it);
          break;
        }
    }
  if (h->cpool) CK(h, hipFree(h->cpool));
  h->cpool = pool;
  h->lanes_dirty = false;
  return 0;
}

// Whether steps run the tiled fused kernel (the one that corrects the
// z-normal CPML slabs outside itself).
bool tiled_fused(const ems_gpu *h) {
  return use_fused(h) && ((!h->band && !(h->packed && packed_ok(h)) && !(h->rows && rows_ok(h))) ||
                          h->croff_wide);
}

// The tile classes of the tiled step for the present geometry and slabs
// (host side of launch_fused_oe): a tile needs the slab code when a slab
// the kernel corrects (E or H side) or whose H increments it adds meets its
// (j, k) box with the halo row and lane and its planes i0 - 1 .. i1 - 1.
int build_tile_classes(ems_gpu *h) {
  const FusedGeom g = fused_geom(h);
  uint64_t key = 1469598103934665603ull;
  for (uint64_t v : {(uint64_t)g.tk, (uint64_t)g.ktiles, (uint64_t)g.jtiles, (uint64_t)g.chunk,
                     (uint64_t)g.nblocks, (uint64_t)h->tj, (uint64_t)h->xcd_group, h->slab_gen,
                     (uint64_t)(h->split_state + 1)})
    key = (key ^ v) * 1099511628211ull;
  if (key == h->tiles_key && h->d_tiles) return 0;
  drop_graphs(h);
  struct Box {
    int lo[3], hi[3];
  };
  std::vector<Box> boxes;
  for (auto &sl : h->slabs)
    if (sl.len != 0 && (!sl.out || !sl.arg.e_side)) {
      Box b;
      for (int d = 0; d < 3; d++) b.lo[d] = sl.arg.lo[d], b.hi[d] = sl.arg.hi[d];
      boxes.push_back(b);
    }
  std::vector<int> q[2][8];
  for (int b = 0; b < g.nblocks; b++) {
    const int t = tile_of_block(b, g.ktiles, g.ntiles, h->xcd_group);
    if (t < 0) continue;
    const int kt = t % g.ktiles, t2 = t / g.ktiles, jt = t2 % g.jtiles, c = t2 / g.jtiles;
    const int k0 = kt * g.tk, j0 = jt * h->tj, i0 = c * g.chunk;
    const int i1 = std::min(h->n[0], i0 + g.chunk), ia = std::max(i0 - 1, 0);
    bool cp = false;
    for (const Box &x : boxes)
      cp = cp || (x.lo[1] < j0 + h->tj && x.hi[1] > j0 - 1 && x.lo[2] < k0 + g.tk &&
                  x.hi[2] > k0 - 1 && x.lo[0] < i1 && x.hi[0] > ia);
    q[cp ? 1 : 0][b % 8].push_back(t);
  }
  std::vector<int> all;
  for (int cl = 0; cl < 2; cl++) {
    size_t m = 0;
    for (int x = 0; x < 8; x++) m = std::max(m, q[cl][x].size());
    for (size_t r = 0; r < m; r++)
      for (int x = 0; x < 8; x++) all.push_back(r < q[cl][x].size() ? q[cl][x][r] : -1);
    h->ntl[cl] = (int)(8 * m);
  }
  if (h->d_tiles) CK(h, hipFree(h->d_tiles));
  h->d_tiles = nullptr;
  if (!all.empty()) {
    CK(h, hipMalloc(&h->d_tiles, all.size() * sizeof(int)));
    CK(h, hipMemcpy(h->d_tiles, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  h->tiles_key = key;
  return 0;
}

int prepare(ems_gpu *h, uint64_t steps, bool allow_split = true) {
  if (h->lanes_dirty) {
    drop_graphs(h);
    int rc = build_pool(h);
    if (rc) return rc;
  }
  bool grown = false;
  for (auto &o : h->ops) {
    if (o.type == OP_LUMPED) {
      // a record per step
      if (!o.lp.record || o.count + steps <= o.cap) continue;
      uint64_t cap = o.cap ? o.cap : 1024;
      while (cap < o.count + steps) cap *= 2;
      double *nv[3] = {nullptr, nullptr, nullptr};
      double *ov[3] = {o.lp.rec, o.lp.pv, o.lp.pi};
      for (int w = 0; w < 3; w++) {
        CK(h, hipMalloc(&nv[w], cap * sizeof(double)));
        if (ov[w] && o.count)
          CK(h, hipMemcpyAsync(nv[w], ov[w], o.count * sizeof(double), hipMemcpyDeviceToDevice,
                               h->stream));
      }
      CK(h, hipStreamSynchronize(h->stream));
      for (int w = 0; w < 3; w++)
        if (ov[w]) CK(h, hipFree(ov[w]));
      o.lp.rec = nv[0], o.lp.pv = nv[1], o.lp.pi = nv[2];
      o.cap = cap;
      grown = true;
      continue;
    }
    if (o.type != OP_PROBE) continue;
    // samples taken in steps [step, step + steps): multiples of interval
    uint64_t a = h->step, b = h->step + steps;
    uint64_t more = (b + o.interval - 1) / o.interval - (a + o.interval - 1) / o.interval;
    if (o.count + more > o.cap) {
      uint64_t cap = o.cap ? o.cap : 1024;
      while (cap < o.count + more) cap *= 2;
      double *nv = nullptr;
      CK(h, hipMalloc(&nv, cap * sizeof(double)));
      if (o.d_vals) {
        CK(h, hipMemcpyAsync(nv, o.d_vals, o.count * sizeof(double), hipMemcpyDeviceToDevice,
                             h->stream));
        CK(h, hipStreamSynchronize(h->stream));
        CK(h, hipFree(o.d_vals));
      }
      o.d_vals = nv;
      o.cap = cap;
      grown = true;
    }
  }
  if (grown) h->ops_dirty = true;
  // the fused kernel's split, or the unfused kernels' CPML fold (bits 2, 3
  // of the unfused switch), which takes the slabs but the z-normal suffix
  const int split = (allow_split && h->cpml_split && tiled_fused(h))
                        ? 1
                        : ((!use_fused(h) && ((h->unfused_pf >> 2) & h->unfused_pf & 3)) ? 2 : 0);
  if (h->slabs_dirty || h->split_state != split) {
    drop_graphs(h);
    int rc = upload_fslabs(h, split);
    if (rc) return rc;
  }
  if (h->tile_class && tiled_fused(h)) {
    int rc = build_tile_classes(h);
    if (rc) return rc;
  }
  if (h->ops_dirty) {
    drop_graphs(h);
    std::vector<DevOp> dv(h->ops.size());
    for (size_t q = 0; q < h->ops.size(); q++) {
      HostOp &o = h->ops[q];
      DevOp &d = dv[q];
      d.type = o.type, d.kind = o.kind, d.quantity = o.quantity, d.n = o.n;
      d.f32sum = o.f32sum;
      d.interval = o.interval;
      d.axis = o.d_axis, d.idx = o.d_idx, d.wf = o.d_wf, d.wd = o.d_wd;
      d.amp = o.d_amp, d.namp = o.namp, d.vals = o.d_vals, d.count = o.d_count;
      d.lp = nullptr;
      if (o.type == OP_LUMPED) {
        const size_t nt = o.tables.size();
        if (nt) {
          CK(h, hipMemcpy(o.d_tv, o.tables.data(), nt * sizeof(double *), hipMemcpyHostToDevice));
          CK(h, hipMemcpy(o.d_tlen, o.table_len.data(), nt * sizeof(uint64_t),
                          hipMemcpyHostToDevice));
        }
        o.lp.count = o.d_count;
        CK(h, hipMemcpy(o.d_lp, &o.lp, sizeof(DevLumped), hipMemcpyHostToDevice));
        d.lp = o.d_lp;
      }
    }
    // Where each run of readers ends, per gap (run_gap jumps there).
    const int nop = (int)h->ops.size();
    for (int w = 0; w < 2; w++) {
      int next = nop;
      for (int q = nop - 1; q >= 0; q--) {
        const HostOp &o = h->ops[q];
        const bool writer = w == 0 ? (o.type == OP_SOURCE && (o.kind == 2 || o.kind == 3))
                                   : ((o.type == OP_SOURCE && (o.kind == 0 || o.kind == 1)) ||
                                      o.type == OP_LUMPED);
        if (writer) next = q;
        dv[q].seg_end[w] = next;
      }
    }
    if (h->d_ops) CK(h, hipFree(h->d_ops));
    h->d_ops = nullptr;
    if (!dv.empty()) {
      CK(h, hipMalloc(&h->d_ops, dv.size() * sizeof(DevOp)));
      CK(h, hipMemcpy(h->d_ops, dv.data(), dv.size() * sizeof(DevOp), hipMemcpyHostToDevice));
    }
    h->ops_dirty = false;
  }
  return 0;
}

// Host mirror of the probe sample counts (the device increments its own).
void advance_counts(ems_gpu *h, uint64_t steps) {
  for (auto &o : h->ops) {
    if (o.type == OP_LUMPED && o.lp.record) o.count += steps;
    if (o.type != OP_PROBE) continue;
    uint64_t a = h->step, b = h->step + steps;
    o.count += (b + o.interval - 1) / o.interval - (a + o.interval - 1) / o.interval;
  }
  h->step += steps;
}

// Scratch of one evaluation, in doubles: a sum per H row (nx * (3 ny - 2)),
// four chains per plane and E component (12 nx), a value per plane (nx)
// and W^n itself (1).
size_t energy_scratch_doubles(const ems_gpu *h) {
  const size_t nx = (size_t)h->n[0], ny = (size_t)h->n[1];
  return nx * (3 * ny - 2) + 12 * nx + nx + 1;
}

// The invariant's kernels on the engine's stream, with the scratch arrays
// given; W^n lands in *d_w (device memory).
template <bool MEAS>
int energy_kernels_t(ems_gpu *h, double *rowsum, double *chains, double *planes, double *d_w) {
  const int nx = h->n[0], ny = h->n[1], nz = h->n[2];
  const size_t rows = (size_t)nx * (3 * ny - 2);
  const Fields &f = h->fb[h->par];
  // a wave per H row (two per wave when rows have at most 32 nodes), per
  // (plane, component) E chains and per plane
  dim3 g1((unsigned)((rows + 3) / 4));
  // (rows longer than 256 nodes: rounds of 512 only where they fill them,
  // the rest of the row in rounds of 64)
  const int nt =
      nz <= 32 ? 0 : (nz <= 64 ? 1 : (nz <= 128 ? 2 : (nz <= 256 || nz - 1 < 512 ? 4 : 8)));
#define EH(OE, NT)                                                                     \
  hipLaunchKernelGGL((energy_h_rows<OE, NT, MEAS>), g1, dim3(256), 0, h->stream, f,       \
                     lanes_arg(h), nx, ny, nz, rowsum)
  if (nt == 0) {
    const dim3 g8((unsigned)((rows + 7) / 8));
    if (h->oe)
      hipLaunchKernelGGL((energy_h_rows2<true, MEAS>), g8, dim3(256), 0, h->stream, f,
                         lanes_arg(h), nx, ny, nz, rowsum);
    else
      hipLaunchKernelGGL((energy_h_rows2<false, MEAS>), g8, dim3(256), 0, h->stream, f,
                         lanes_arg(h), nx, ny, nz, rowsum);
  } else if (h->oe) {
    if (nt == 1) EH(true, 1); else if (nt == 2) EH(true, 2); else if (nt == 4) EH(true, 4);
    else EH(true, 8);
  } else {
    if (nt == 1) EH(false, 1); else if (nt == 2) EH(false, 2); else if (nt == 4) EH(false, 4);
    else EH(false, 8);
  }
#undef EH
  CK(h, hipGetLastError());
  if ((size_t)ny * nz < ((size_t)1 << 31)) {
    hipLaunchKernelGGL(energy_e_chains_blk<MEAS>, dim3((unsigned)nx * 3), dim3(256), 0, h->stream,
                       f, lanes_arg(h), nx, ny, nz, chains);
  } else {
    dim3 g2((unsigned)(((size_t)nx * 3 + 3) / 4));
    hipLaunchKernelGGL(energy_e_chains<MEAS>, g2, dim3(256), 0, h->stream, f, lanes_arg(h), nx,
                       ny, nz, chains);
  }
  CK(h, hipGetLastError());
  const dim3 gp((unsigned)((nx + 3) / 4));
  hipLaunchKernelGGL(energy_planes<MEAS>, gp, dim3(256), 0, h->stream, nx, ny, rowsum, chains,
                     planes);
  CK(h, hipGetLastError());
  if (MEAS)
    for (auto &o : h->ops)
      if (o.type == OP_UPML && o.upml.nh) {
        hipLaunchKernelGGL(upml_energy_planes, gp, dim3(256), 0, h->stream, f, lanes_arg(h), o.upml,
                           (const uint32_t *)o.d_hplane, nx, ny, nz, planes);
        CK(h, hipGetLastError());
      }
  hipLaunchKernelGGL(energy_combine<MEAS>, dim3(1), dim3(64), 0, h->stream, nx, planes, h->dt,
                     d_w);
  CK(h, hipGetLastError());
  return 0;
}

// The energy's kernels on the engine's stream, with the scratch arrays
// given: the invariant (measure 0) or the field part of openEMS's measure
// (1, with the layers' corrections); the value lands in *d_w (device).
int energy_kernels(ems_gpu *h, double *rowsum, double *chains, double *planes, double *d_w,
                   int measure = 0) {
  return measure ? energy_kernels_t<true>(h, rowsum, chains, planes, d_w)
                 : energy_kernels_t<false>(h, rowsum, chains, planes, d_w);
}

// The invariant's kernels and the copy of W^n into pinned host memory
// (h->h_w), queued on the engine's stream; energy_now waits for it. The
// scratch is the engine's own, allocated at the first evaluation.
int energy_enqueue(ems_gpu *h, int measure = 0) {
  const int nx = h->n[0], ny = h->n[1];
  const size_t rows = (size_t)nx * (3 * ny - 2);
  if (!h->d_rowsum) {
    CK(h, hipMalloc(&h->d_rowsum, rows * sizeof(double)));
    CK(h, hipMalloc(&h->d_chains, (size_t)nx * 12 * sizeof(double)));
    CK(h, hipMalloc(&h->d_planes, (size_t)nx * sizeof(double)));
    CK(h, hipMalloc(&h->d_w, sizeof(double)));
    CK(h, hipHostMalloc((void **)&h->h_w, sizeof(double), hipHostMallocDefault));
  }
  int rc = energy_kernels(h, h->d_rowsum, h->d_chains, h->d_planes, h->d_w, measure);
  if (rc) return rc;
  CK(h, hipMemcpyAsync(h->h_w, h->d_w, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return 0;
}

int energy_now(ems_gpu *h, double *w, int measure = 0) {
  int rc = energy_enqueue(h, measure);
  if (rc) return rc;
  CK(h, hipStreamSynchronize(h->stream));
  *w = *h->h_w;
  return 0;
}

int run_steps(ems_gpu *h, uint64_t n) {
  if (n == 0) return 0;
  int rc = prepare(h, n);
  if (rc) return rc;
  const int p0 = h->par;
  if (!h->g1[p0]) {
    for (int p = 0; p < 2; p++) {
      if ((rc = capture(h, 1, p, &h->g1[p]))) return rc;
      if ((rc = capture(h, h->K, p, &h->gK[p]))) return rc;
    }
  }
  const int flip = use_fused(h) ? 1 : 0;
  uint64_t left = n;
  while (left >= (uint64_t)h->K) {
    CK(h, hipGraphLaunch(h->gK[h->par], h->stream));
    left -= h->K;  // K is even: the buffer does not change
  }
  while (left > 0) {
    CK(h, hipGraphLaunch(h->g1[h->par], h->stream));
    h->par ^= flip;
    left--;
  }
  merged_mark_stale(h);
  advance_counts(h, n);
  return 0;
}

int upload(ems_gpu *h, void **dst, const void *src, size_t bytes) {
  CK(h, hipMalloc(dst, bytes ? bytes : 1));
  if (bytes) CK(h, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

bool node_ok(const ems_gpu *h, const int32_t *nd) {
  for (int d = 0; d < 3; d++)
    if (nd[d] < 0 || nd[d] >= h->n[d]) return false;
  return true;
}

#define ALIVE(h)                                                                          \
  do {                                                                                    \
    if (!(h)) return EMS_E_INVALID;                                                       \
    if ((h)->dead) return fail((h), EMS_E_DEVICE, "engine has a device fault: " + (h)->err); \
  } while (0)

}  // namespace

// ------------------------------------------------------------------ C ABI --

extern "C" {

int ems_gpu_create(const ems_grid_desc *d, ems_gpu **out) {
  if (!d || !out) return EMS_E_INVALID;
  *out = nullptr;
  for (int a = 0; a < 3; a++)
    if (d->nodes[a] < 2) return EMS_E_INVALID;
  if (!(d->dt > 0.0) || !std::isfinite(d->dt)) return EMS_E_INVALID;
  if (d->curl_order != 0 && d->curl_order != 1) return EMS_E_INVALID;
  const size_t N = (size_t)d->nodes[0] * d->nodes[1] * d->nodes[2];
  if ((size_t)d->nodes[1] * d->nodes[2] > (size_t)INT32_MAX) return EMS_E_UNSUPPORTED;
  ems_gpu *h = new ems_gpu();
  for (int a = 0; a < 3; a++) h->n[a] = d->nodes[a];
  h->N = N;
  h->plane = (size_t)d->nodes[1] * d->nodes[2];
  h->dt = d->dt;
  h->oe = d->curl_order;
  auto init = [&]() -> int {
    CK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (int b = 0; b < 2; b++) {
      CK(h, hipMalloc(&h->fbuf[b], 6 * N * sizeof(float)));
      CK(h, hipMemset(h->fbuf[b], 0, 6 * N * sizeof(float)));
      for (int a = 0; a < 3; a++) {
        h->fb[b].u[a] = h->fbuf[b] + a * N;
        h->fb[b].i[a] = h->fbuf[b] + (3 + a) * N;
      }
    }
    CK(h, hipMalloc(&h->d_step, sizeof(uint64_t)));
    CK(h, hipMemset(h->d_step, 0, sizeof(uint64_t)));
    return 0;
  };
  int rc = init();
  if (rc) {
    std::fprintf(stderr, "ems_gpu_create: %s\n", h->err.c_str());
    ems_gpu_destroy(h);
    return rc;
  }
  *out = h;
  return EMS_OK;
}

void ems_gpu_destroy(ems_gpu *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graphs(h);
  for (int b = 0; b < 2; b++)
    if (h->fbuf[b]) (void)hipFree(h->fbuf[b]);
  for (void *p : h->lane_mem) (void)hipFree(p);
  if (h->cpool) (void)hipFree(h->cpool);
  if (h->croff) (void)hipFree(h->croff);
  for (auto &s : h->slabs) {
    (void)hipFree(s.d_b), (void)hipFree(s.d_a), (void)hipFree(s.d_k);
    (void)hipFree(s.d_psi[0]);
    if (s.d_psi[1]) (void)hipFree(s.d_psi[1]);
    if (s.d_delta) (void)hipFree(s.d_delta);
  }
  if (h->d_fslabs) (void)hipFree(h->d_fslabs);
  if (h->d_pmask) (void)hipFree(h->d_pmask);
  if (h->d_tiles) (void)hipFree(h->d_tiles);
  for (auto &d : h->dfts) {
    if (d.d_state) (void)hipFree(d.d_state);
    if (d.d_step) (void)hipFree(d.d_step);
    if (d.d_anchor) (void)hipFree(d.d_anchor);
  }
  if (h->d_hout) (void)hipFree(h->d_hout);
  if (h->d_eout) (void)hipFree(h->d_eout);
  if (h->d_pmask_o) (void)hipFree(h->d_pmask_o);
  for (auto &o : h->ops) {
    (void)hipFree(o.d_axis), (void)hipFree(o.d_idx), (void)hipFree(o.d_wf);
    (void)hipFree(o.d_wd), (void)hipFree(o.d_amp), (void)hipFree(o.d_vals);
    (void)hipFree(o.d_count);
    for (void *p : o.mem) (void)hipFree(p);
    if (o.h_et) (void)hipHostFree(o.h_et);
    for (double *p : o.tables) (void)hipFree(p);
    if (o.pw_src) (void)hipFree(o.pw_src);
    if (o.lp.rec) (void)hipFree(o.lp.rec), (void)hipFree(o.lp.pv), (void)hipFree(o.lp.pi);
  }
  if (h->d_ops) (void)hipFree(h->d_ops);
  if (h->d_step) (void)hipFree(h->d_step);
  if (h->d_rowsum) {
    (void)hipFree(h->d_rowsum), (void)hipFree(h->d_chains);
    (void)hipFree(h->d_planes), (void)hipFree(h->d_w);
  }
  if (h->stream) (void)hipStreamDestroy(h->stream);
  if (h->side) (void)hipStreamDestroy(h->side);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_energy) (void)hipEventDestroy(h->ev_energy);
  if (h->em_d) (void)hipFree(h->em_d), (void)hipFree(h->em_dout);
  if (h->em_h) (void)hipHostFree(h->em_h), (void)hipHostFree(h->em_hout);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->h_w) (void)hipHostFree(h->h_w);
  delete h;
}

const char *ems_gpu_last_error(const ems_gpu *h) { return h ? h->err.c_str() : "null handle"; }

int ems_gpu_set_lane(ems_gpu *h, int lane, const float *rows, uint64_t nrows,
                     const uint32_t *row_index) {
  ALIVE(h);
  if (lane < 0 || lane >= 12 || !rows) return fail(h, EMS_E_INVALID, "lane out of range");
  const size_t nr = (size_t)h->n[0] * h->n[1];
  if (!row_index && nrows != nr) return fail(h, EMS_E_INVALID, "dense lane needs nx*ny rows");
  if (row_index)
    for (size_t q = 0; q < nr; q++)
      if (row_index[q] >= nrows) return fail(h, EMS_E_INVALID, "row index out of range");
  void *dr = nullptr, *di = nullptr;
  int rc = upload(h, &dr, rows, nrows * h->n[2] * sizeof(float));
  if (rc) return rc;
  h->lane_mem.push_back(dr);
  if (row_index) {
    if ((rc = upload(h, &di, row_index, nr * sizeof(uint32_t)))) return rc;
    h->lane_mem.push_back(di);
  }
  h->lanes[lane].rows = (const float *)dr;
  h->lanes[lane].index = (const uint32_t *)di;
  h->lane_rows[lane] = nrows;
  h->lanes_dirty = true;
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_add_cpml(ems_gpu *h, const ems_cpml_slab *s, int n) {
  ALIVE(h);
  for (auto &o : h->ops) {
    if (o.type == OP_POLES)
      return fail(h, EMS_E_UNSUPPORTED,
                  "CPML slabs must be registered before the dispersive media");
    if (o.type == OP_UPML)
      return fail(h, EMS_E_UNSUPPORTED, "openEMS PML layers and CPML slabs on one engine");
  }
  ALIVE(h);
  if (n > 0 && !s) return EMS_E_INVALID;
  for (int q = 0; q < n; q++) {
    const ems_cpml_slab &in = s[q];
    if (in.comp < 0 || in.comp > 2 || in.normal < 0 || in.normal > 2 || in.comp == in.normal)
      return fail(h, EMS_E_INVALID, "CPML slab: bad component/normal");
    for (int d = 0; d < 3; d++)
      if (in.lo[d] < 0 || in.hi[d] > h->n[d] || in.lo[d] > in.hi[d])
        return fail(h, EMS_E_INVALID, "CPML slab: box out of range");
    // The H-side derivative reads one node past the box along the normal;
    // an E-side box covers only existing E edges of its component.
    if (!in.e_side && in.hi[in.normal] >= h->n[in.normal] && in.lo[in.normal] < in.hi[in.normal])
      return fail(h, EMS_E_INVALID, "CPML slab: H box touches the last line");
    HostSlab hs;
    SlabArg &a = hs.arg;
    a.e_side = in.e_side ? 1 : 0;
    a.comp = in.comp;
    a.normal = in.normal;
    // Slab::new (cpml.rs:415-437)
    if (in.normal == (in.comp + 1) % 3) {
      a.sign = 1.0f;
      a.other = (in.comp + 2) % 3;
    } else {
      a.sign = -1.0f;
      a.other = (in.comp + 1) % 3;
    }
    for (int d = 0; d < 3; d++) a.lo[d] = in.lo[d], a.hi[d] = in.hi[d];
    const int depth = in.hi[in.normal] - in.lo[in.normal];
    int rc;
    if ((rc = upload(h, (void **)&hs.d_b, in.b, depth * sizeof(float)))) return rc;
    if ((rc = upload(h, (void **)&hs.d_a, in.a, depth * sizeof(float)))) return rc;
    if ((rc = upload(h, (void **)&hs.d_k, in.k, depth * sizeof(float)))) return rc;
    hs.len = (size_t)(in.hi[0] - in.lo[0]) * (in.hi[1] - in.lo[1]) * (in.hi[2] - in.lo[2]);
    const size_t bytes = (hs.len ? hs.len : 1) * sizeof(float);
    for (int b = 0; b < (a.e_side ? 1 : 2); b++) {
      CK(h, hipMalloc(&hs.d_psi[b], bytes));
      CK(h, hipMemset(hs.d_psi[b], 0, bytes));
    }
    a.b = hs.d_b, a.a = hs.d_a, a.k = hs.d_k, a.psi = nullptr;
    h->slabs.push_back(hs);
  }
  h->slabs_dirty = true;
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_add_mur(ems_gpu *h, const ems_mur_desc *d, int *id) {
  ALIVE(h);
  if (!d || d->n_edges < 0 || d->n_corners < 0) return EMS_E_INVALID;
  // Slab mode of the reference (mur.rs:553-562): no operator registered
  // before it acts on the voltages serially around the E update.
  for (auto &o : h->ops)
    if ((o.type == OP_SOURCE && (o.kind == 0 || o.kind == 1)) ||
        (o.type == OP_PROBE && o.quantity == 0) || o.type == OP_LUMPED || o.type == OP_PW ||
        o.type == OP_MERGED)
      return fail(h, EMS_E_UNSUPPORTED,
                  "register the Mur boundary before E sources, voltage probes, lumped "
                  "elements, plane waves and cut-face currents");
  const int ne = d->n_edges, nc = d->n_corners;
  if ((ne && (!d->edge_axis || !d->edge_node || !d->edge_inner || !d->edge_k)) ||
      (nc && (!d->corner_axis || !d->corner_node || !d->corner_inner0 || !d->corner_inner1 ||
              !d->corner_k0 || !d->corner_k1)))
    return fail(h, EMS_E_INVALID, "Mur arrays");
  auto flat = [&](const int32_t *nd, uint64_t *out) {
    if (!node_ok(h, nd)) return false;
    *out = ((uint64_t)nd[0] * h->n[1] + nd[1]) * h->n[2] + nd[2];
    return true;
  };
  std::vector<int32_t> eax(ne), cax(nc);
  std::vector<uint64_t> en(ne), ei(ne), eh(ne, 0), cn(nc), c0(nc), c1(nc), ch(nc, 0);
  for (int q = 0; q < ne; q++) {
    eax[q] = d->edge_axis[q];
    if (eax[q] < 0 || eax[q] > 2 || !flat(d->edge_node + 3 * q, &en[q]) ||
        !flat(d->edge_inner + 3 * q, &ei[q]))
      return fail(h, EMS_E_INVALID, "Mur edge out of range");
    if (d->edge_hold) eh[q] = d->edge_hold[q];
  }
  for (int q = 0; q < nc; q++) {
    cax[q] = d->corner_axis[q];
    if (cax[q] < 0 || cax[q] > 2 || !flat(d->corner_node + 3 * q, &cn[q]) ||
        !flat(d->corner_inner0 + 3 * q, &c0[q]) || !flat(d->corner_inner1 + 3 * q, &c1[q]))
      return fail(h, EMS_E_INVALID, "Mur corner out of range");
    if (d->corner_hold) ch[q] = d->corner_hold[q];
  }
  HostOp o;
  o.type = OP_MUR;
  DevMur &m = o.mur;
  m.ne = ne, m.nc = nc, m.oe_order = d->openems_order ? 1 : 0;
  m.oe_corners = d->openems_corners ? 1 : 0;
  int rc;
  void *p = nullptr;
#define UP(dst, src, bytes)                                       \
  if ((rc = upload(h, &p, (src), (bytes)))) return rc;           \
  o.mem.push_back(p);                                             \
  dst = (decltype(dst))p;
  UP(m.eax, eax.data(), ne * sizeof(int32_t));
  UP(m.enode, en.data(), ne * sizeof(uint64_t));
  UP(m.einner, ei.data(), ne * sizeof(uint64_t));
  UP(m.ehold, eh.data(), ne * sizeof(uint64_t));
  UP(m.ek, d->edge_k, ne * sizeof(float));
  UP(m.cax, cax.data(), nc * sizeof(int32_t));
  UP(m.cnode, cn.data(), nc * sizeof(uint64_t));
  UP(m.cin0, c0.data(), nc * sizeof(uint64_t));
  UP(m.cin1, c1.data(), nc * sizeof(uint64_t));
  UP(m.chold, ch.data(), nc * sizeof(uint64_t));
  UP(m.ck0, d->corner_k0, nc * sizeof(float));
  UP(m.ck1, d->corner_k1, nc * sizeof(float));
  std::vector<float> zero(2 * (size_t)ne + 3 * (size_t)nc, 0.0f);
  UP(m.saved, zero.data(), zero.size() * sizeof(float));
#undef UP
  h->ops.push_back(std::move(o));
  h->ops_dirty = true;
  drop_graphs(h);
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

int ems_gpu_add_upml(ems_gpu *h, const ems_upml_desc *d, int *id) {
  ALIVE(h);
  if (!d || d->n_e < 0 || d->n_h < 0) return EMS_E_INVALID;
  if (!h->slabs.empty())
    return fail(h, EMS_E_UNSUPPORTED, "openEMS PML layers and CPML slabs on one engine");
  for (auto &o : h->ops) {
    if (o.type == OP_UPML) return fail(h, EMS_E_UNSUPPORTED, "a second PML layer operator");
    if (o.type == OP_POLES || o.type == OP_MUR || o.type == OP_PW || o.type == OP_MERGED)
      return fail(h, EMS_E_UNSUPPORTED,
                  "register the PML layers before dispersive media, Mur boundaries, plane "
                  "waves and cut-face currents (they run first among the slab corrections)");
  }
  const int cnt[2] = {d->n_e, d->n_h};
  const int32_t *ax[2] = {d->e_axis, d->h_axis}, *nd[2] = {d->e_node, d->h_node};
  const float *cf[2] = {d->e_coef, d->h_coef};
  for (int s = 0; s < 2; s++)
    if (cnt[s] && (!ax[s] || !nd[s] || !cf[s])) return fail(h, EMS_E_INVALID, "PML arrays");
  const uint64_t N = (uint64_t)h->n[0] * h->n[1] * h->n[2];
  // Grouped by (side, axis), each group in the given order.
  std::vector<int> order[6];
  std::vector<uint64_t> flat[2], keys;
  for (int s = 0; s < 2; s++) {
    flat[s].resize(cnt[s]);
    for (int q = 0; q < cnt[s]; q++) {
      const int a = ax[s][q];
      const int32_t *p = nd[s] + 3 * (size_t)q;
      if (a < 0 || a > 2 || !node_ok(h, p)) return fail(h, EMS_E_INVALID, "PML edge out of range");
      for (int c = 0; c < 4; c++)
        if (!std::isfinite(cf[s][4 * (size_t)q + c]))
          return fail(h, EMS_E_INVALID, "PML coefficient not finite");
      flat[s][q] = ((uint64_t)p[0] * h->n[1] + p[1]) * h->n[2] + p[2];
      order[3 * s + a].push_back(q);
      keys.push_back((uint64_t)(3 * s + a) * N + flat[s][q]);
    }
  }
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
    return fail(h, EMS_E_INVALID, "PML edge listed twice");
  // each group by node index (the layer runs' order, upml.rs:290-330)
  for (int g = 0; g < 6; g++)
    std::stable_sort(order[g].begin(), order[g].end(),
                     [&](int x, int y) { return flat[g / 3][x] < flat[g / 3][y]; });
  HostOp o;
  o.type = OP_UPML;
  DevUpml &m = o.upml;
  m.ne = cnt[0], m.nh = cnt[1];
  const size_t n = (size_t)m.ne + m.nh;
  std::vector<uint64_t> idx;
  std::vector<float4> co;
  idx.reserve(n), co.reserve(n);
  m.off[0] = 0;
  for (int g = 0; g < 6; g++) {
    const int s = g / 3;
    for (int q : order[g]) {
      idx.push_back(flat[s][q]);
      const float *c = cf[s] + 4 * (size_t)q;
      co.push_back(make_float4(c[0], c[1], c[2], c[3]));
    }
    m.off[g + 1] = (int)idx.size();
  }
  int rc;
  void *p = nullptr;
  if ((rc = upload(h, &p, idx.data(), n * sizeof(uint64_t)))) return rc;
  o.mem.push_back(p), m.idx = (const uint64_t *)p;
  if ((rc = upload(h, &p, co.data(), n * sizeof(float4)))) return rc;
  o.mem.push_back(p), m.co = (const float4 *)p;
  std::vector<float2> zero(n, make_float2(0.0f, 0.0f));
  if ((rc = upload(h, &p, zero.data(), n * sizeof(float2)))) return rc;
  o.mem.push_back(p), m.fv = (float2 *)p;
  // The caller's order of each side, for ems_gpu_upml_state.
  std::vector<uint32_t> pos(n);
  for (int g = 0; g < 6; g++)
    for (size_t r = 0; r < order[g].size(); r++)
      pos[(g < 3 ? 0 : m.ne) + order[g][r]] = (uint32_t)(m.off[g] + r);
  o.upml_pos = std::move(pos);
  const int nx = h->n[0];
  const uint64_t plane = (uint64_t)h->n[1] * h->n[2];
  std::vector<uint32_t> hp(3 * (size_t)(nx + 1));
  for (int c = 0; c < 3; c++)
    for (int gi = 0; gi <= nx; gi++)
      hp[c * (size_t)(nx + 1) + gi] =
          (uint32_t)(std::lower_bound(idx.begin() + m.off[3 + c], idx.begin() + m.off[4 + c],
                                      (uint64_t)gi * plane) -
                     idx.begin());
  if ((rc = upload(h, &p, hp.data(), hp.size() * sizeof(uint32_t)))) return rc;
  o.mem.push_back(p), o.d_hplane = (uint32_t *)p;
  h->ops.push_back(std::move(o));
  h->ops_dirty = true;
  drop_graphs(h);
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

int ems_gpu_upml_state(ems_gpu *h, int id, int side, float *flux, float *value) {
  ALIVE(h);
  if (id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_UPML || side < 0 || side > 1)
    return fail(h, EMS_E_INVALID, "not a PML layer operator");
  const HostOp &o = h->ops[id];
  const int cnt = side ? o.upml.nh : o.upml.ne;
  if (cnt && (!flux || !value)) return EMS_E_INVALID;
  const size_t n = (size_t)o.upml.ne + o.upml.nh;
  std::vector<float2> fv(n);
  CK(h, hipStreamSynchronize(h->stream));
  if (n) CK(h, hipMemcpy(fv.data(), o.upml.fv, n * sizeof(float2), hipMemcpyDeviceToHost));
  const size_t base = side ? o.upml.ne : 0;
  for (int q = 0; q < cnt; q++) {
    const float2 s = fv[o.upml_pos[base + q]];
    flux[q] = s.x, value[q] = s.y;
  }
  return EMS_OK;
}

int ems_gpu_add_source(ems_gpu *h, int kind, int n_edges, const int32_t *axis,
                       const int32_t *nodes, const float *w, const float *amp, uint64_t n_amp,
                       int *id) {
  ALIVE(h);
  if (kind < 0 || kind > 3) return fail(h, EMS_E_INVALID, "source kind");
  if (n_edges < 0 || (n_edges && (!axis || !nodes || !w)) || (n_amp && !amp))
    return fail(h, EMS_E_INVALID, "source arrays");
  std::vector<uint64_t> idx(n_edges);
  for (int q = 0; q < n_edges; q++) {
    if (axis[q] < 0 || axis[q] > 2 || !node_ok(h, nodes + 3 * q))
      return fail(h, EMS_E_INVALID, "source edge out of range");
    idx[q] = ((uint64_t)nodes[3 * q] * h->n[1] + nodes[3 * q + 1]) * h->n[2] + nodes[3 * q + 2];
  }
  HostOp o;
  o.type = OP_SOURCE;
  o.kind = kind;
  o.n = n_edges;
  int rc;
  if ((rc = upload(h, (void **)&o.d_axis, axis, n_edges * sizeof(int32_t)))) return rc;
  if ((rc = upload(h, (void **)&o.d_idx, idx.data(), n_edges * sizeof(uint64_t)))) return rc;
  if ((rc = upload(h, (void **)&o.d_wf, w, n_edges * sizeof(float)))) return rc;
  if ((rc = upload(h, (void **)&o.d_amp, amp, n_amp * sizeof(float)))) return rc;
  o.namp = n_amp;
  if (kind == 2 || kind == 3) {  // for the H sources' share of openEMS's energy
    o.h_axis.assign(axis, axis + n_edges);
    o.h_idx = idx;
    o.h_wf.assign(w, w + n_edges);
    o.h_amp.assign(amp, amp + n_amp);
  }
  h->ops.push_back(o);
  h->ops_dirty = true;
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

int ems_gpu_add_probe(ems_gpu *h, int quantity, int n_entries, const int32_t *axis,
                      const int32_t *nodes, const double *w, uint64_t interval, int *id) {
  ALIVE(h);
  const int f32sum = (quantity & EMS_PROBE_F32_SUM) != 0;
  quantity &= ~EMS_PROBE_F32_SUM;
  if (quantity < 0 || quantity > 1) return fail(h, EMS_E_INVALID, "probe quantity");
  if (n_entries < 0 || (n_entries && (!axis || !nodes || !w)))
    return fail(h, EMS_E_INVALID, "probe arrays");
  std::vector<uint64_t> idx(n_entries);
  for (int q = 0; q < n_entries; q++) {
    if (axis[q] < 0 || axis[q] > 2 || !node_ok(h, nodes + 3 * q))
      return fail(h, EMS_E_INVALID, "probe entry out of range");
    idx[q] = ((uint64_t)nodes[3 * q] * h->n[1] + nodes[3 * q + 1]) * h->n[2] + nodes[3 * q + 2];
  }
  HostOp o;
  o.type = OP_PROBE;
  o.quantity = quantity;
  o.f32sum = f32sum;
  o.n = n_entries;
  o.interval = interval ? interval : 1;
  o.first = (h->step + o.interval - 1) / o.interval * o.interval;
  int rc;
  if ((rc = upload(h, (void **)&o.d_axis, axis, n_entries * sizeof(int32_t)))) return rc;
  if ((rc = upload(h, (void **)&o.d_idx, idx.data(), n_entries * sizeof(uint64_t)))) return rc;
  if ((rc = upload(h, (void **)&o.d_wd, w, n_entries * sizeof(double)))) return rc;
  CK(h, hipMalloc(&o.d_count, sizeof(uint64_t)));
  CK(h, hipMemset(o.d_count, 0, sizeof(uint64_t)));
  h->ops.push_back(o);
  h->ops_dirty = true;
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

int ems_gpu_add_lumped(ems_gpu *h, const ems_lumped_desc *d, int *id) {
  ALIVE(h);
  if (!d) return fail(h, EMS_E_INVALID, "null lumped description");
  const int ne = d->n_edges;
  if (d->axis < 0 || d->axis > 2 || ne < 1 || !d->nodes || !d->cb || !d->g_src)
    return fail(h, EMS_E_INVALID, "lumped element: axis, edges, cb and g_src are required");
  if (d->series < 1) return fail(h, EMS_E_INVALID, "lumped element: series must be >= 1");
  if (d->n_terms < 0 || d->n_drive < 0 || (d->n_terms && !d->term_weight) ||
      (d->n_drive && !d->drive_weight))
    return fail(h, EMS_E_INVALID, "lumped element: source weights");
  if (d->branch < 0 || d->branch > 2 || (d->branch && (!d->two_l_dt || !d->inv_c || !d->z)))
    return fail(h, EMS_E_INVALID, "lumped element: branch arrays");
  if ((d->rec_g == nullptr) != (d->rec_c == nullptr))
    return fail(h, EMS_E_INVALID, "lumped element: record needs both g and c");
  if (d->port_columns < 0 || (d->port_columns && !d->rec_g))
    return fail(h, EMS_E_INVALID, "lumped port needs the current record");
  std::vector<uint64_t> idx(ne);
  for (int q = 0; q < ne; q++) {
    const int32_t *nd = d->nodes + 3 * q;
    if (!node_ok(h, nd) || nd[d->axis] + 1 >= h->n[d->axis])
      return fail(h, EMS_E_INVALID, "lumped edge out of range");
    if (!(d->cb[q] != 0.0)) return fail(h, EMS_E_INVALID, "lumped edge is PEC (cb = 0)");
    idx[q] = ((uint64_t)nd[0] * h->n[1] + nd[1]) * h->n[2] + nd[2];
  }
  HostOp o;
  o.type = OP_LUMPED;
  o.n = ne;
  DevLumped &L = o.lp;
  L.axis = d->axis, L.n = ne, L.series = d->series, L.src_in_branch = d->src_in_branch ? 1 : 0;
  L.branch = d->branch, L.nterm = d->n_terms, L.ndrive = d->n_drive;
  L.columns = d->port_columns, L.record = d->rec_g != nullptr, L.dt = h->dt;
  auto put = [&](const void *src, size_t bytes, void **dst) -> int {
    int rc = upload(h, dst, src, bytes);
    if (!rc) o.mem.push_back(*dst);
    return rc;
  };
  auto zeros = [&](size_t n, double **dst) -> int {
    std::vector<double> z(n ? n : 1, 0.0);
    return put(z.data(), z.size() * sizeof(double), (void **)dst);
  };
  const size_t nb = (size_t)ne * sizeof(double);
  const int nt = d->n_terms + d->n_drive;
  std::vector<double> tw((size_t)(nt ? nt : 1) * ne, 0.0);
  if (d->n_terms) memcpy(tw.data(), d->term_weight, (size_t)d->n_terms * nb);
  if (d->n_drive) memcpy(tw.data() + (size_t)d->n_terms * ne, d->drive_weight, (size_t)d->n_drive * nb);
  int rc;
  if ((rc = put(idx.data(), ne * sizeof(uint64_t), (void **)&L.idx))) return rc;
  if ((rc = put(d->cb, nb, (void **)&L.cb))) return rc;
  if ((rc = put(d->g_src, nb, (void **)&L.gsrc))) return rc;
  if ((rc = put(tw.data(), tw.size() * sizeof(double), (void **)&L.tw))) return rc;
  if ((rc = zeros(ne, &L.v)) || (rc = zeros(ne, &L.vd))) return rc;
  if ((rc = zeros(ne, &L.j)) || (rc = zeros(ne, &L.q)) || (rc = zeros(ne, &L.javg)) ||
      (rc = zeros(ne, &L.uprev)))
    return rc;
  if (d->branch) {
    if ((rc = put(d->two_l_dt, nb, (void **)&L.two_l_dt))) return rc;
    if ((rc = put(d->inv_c, nb, (void **)&L.inv_c))) return rc;
    if ((rc = put(d->z, nb, (void **)&L.z))) return rc;
    o.two_l_dt.assign(d->two_l_dt, d->two_l_dt + ne);
    o.inv_c.assign(d->inv_c, d->inv_c + ne);
  }
  if (L.record) {
    if ((rc = put(d->rec_g, nb, (void **)&L.rg))) return rc;
    if ((rc = put(d->rec_c, nb, (void **)&L.rc))) return rc;
  }
  o.tables.assign(nt, nullptr);
  o.table_len.assign(nt, 0);
  if ((rc = zeros(nt, (double **)&o.d_tv))) return rc;
  if ((rc = zeros(nt, (double **)&o.d_tlen))) return rc;
  L.tv = o.d_tv, L.tlen = o.d_tlen;
  CK(h, hipMalloc(&o.d_count, sizeof(uint64_t)));
  CK(h, hipMemset(o.d_count, 0, sizeof(uint64_t)));
  if ((rc = put(&L, sizeof(DevLumped), (void **)&o.d_lp))) return rc;
  o.first = h->step;
  h->ops.push_back(o);
  h->ops_dirty = true;
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

static HostOp *lumped_op(ems_gpu *h, int id) {
  if (id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_LUMPED) return nullptr;
  return &h->ops[id];
}

int ems_gpu_lumped_table(ems_gpu *h, int id, int term, const double *values, uint64_t count) {
  ALIVE(h);
  HostOp *o = lumped_op(h, id);
  if (!o) return fail(h, EMS_E_INVALID, "not a lumped element id");
  if (term < 0 || term >= (int)o->tables.size() || (count && !values))
    return fail(h, EMS_E_INVALID, "lumped source term");
  CK(h, hipStreamSynchronize(h->stream));
  double *nt = nullptr;
  int rc = upload(h, (void **)&nt, values, count * sizeof(double));
  if (rc) return rc;
  if (o->tables[term]) CK(h, hipFree(o->tables[term]));
  o->tables[term] = nt;
  o->table_len[term] = count;
  h->ops_dirty = true;
  return EMS_OK;
}

int ems_gpu_lumped_values(ems_gpu *h, int id, int which, double *out, uint64_t cap,
                          uint64_t *count, uint64_t *first_step) {
  ALIVE(h);
  HostOp *o = lumped_op(h, id);
  if (!o) return fail(h, EMS_E_INVALID, "not a lumped element id");
  if (which < 0 || which > 2 || (which && !o->lp.columns))
    return fail(h, EMS_E_INVALID, "lumped series: 0 current, 1/2 port V/I (ports only)");
  if (!o->lp.record) {
    if (count) *count = 0;
    if (first_step) *first_step = o->first;
    return EMS_OK;
  }
  CK(h, hipStreamSynchronize(h->stream));
  const double *src = which == 0 ? o->lp.rec : which == 1 ? o->lp.pv : o->lp.pi;
  const uint64_t n = o->count < cap ? o->count : cap;
  if (n && out) CK(h, hipMemcpy(out, src, n * sizeof(double), hipMemcpyDeviceToHost));
  if (count) *count = o->count;
  if (first_step) *first_step = o->first;
  return EMS_OK;
}

int ems_gpu_add_dft(ems_gpu *h, int lumped_id, int which, const double *freqs, int nf, double t0,
                    double dt, int *dft_id) {
  ALIVE(h);
  HostOp *o = lumped_op(h, lumped_id);
  if (!o) return fail(h, EMS_E_INVALID, "not a lumped element id");
  if (which < 0 || which > 2 || (which && !o->lp.columns))
    return fail(h, EMS_E_INVALID, "dft series: 0 current, 1/2 port V/I (ports only)");
  if (!o->lp.record) return fail(h, EMS_E_INVALID, "dft: the element records no series");
  if (nf < 1 || !freqs || !(dt > 0.0) || !std::isfinite(t0)) return fail(h, EMS_E_INVALID, "dft args");
  HostDft d;
  d.op = lumped_id, d.which = which, d.nf = nf, d.t0 = t0, d.dt = dt;
  d.freqs.assign(freqs, freqs + nf);
  std::vector<double> st(4 * (size_t)nf, 0.0), step(2 * (size_t)nf), ph(2 * (size_t)nf);
  // RunningDft::new: acc 0, phasor anchored at n = 0 (dft.rs:90-108)
  ems_dft_phasors(freqs, nf, t0, dt, 0, ph.data());
  ems_dft_steps(freqs, nf, dt, step.data());
  for (int k = 0; k < nf; k++) st[4 * k + 2] = ph[2 * k], st[4 * k + 3] = ph[2 * k + 1];
  int rc;
  if ((rc = upload(h, (void **)&d.d_state, st.data(), st.size() * sizeof(double)))) return rc;
  if ((rc = upload(h, (void **)&d.d_step, step.data(), step.size() * sizeof(double)))) return rc;
  h->dfts.push_back(std::move(d));
  if (dft_id) *dft_id = (int)h->dfts.size() - 1;
  return EMS_OK;
}

int ems_gpu_dft_update(ems_gpu *h, int dft_id) {
  ALIVE(h);
  if (dft_id < 0 || dft_id >= (int)h->dfts.size()) return fail(h, EMS_E_INVALID, "not a dft id");
  HostDft &d = h->dfts[dft_id];
  HostOp *o = lumped_op(h, d.op);
  const uint64_t n = o->count;
  if (n <= d.done) return EMS_OK;
  // the anchors of samples 1024 m <= n, from the host's routine
  const uint64_t need = n / 1024;
  if (need > d.anchors) {
    const size_t per = 2 * (size_t)d.nf;
    if (need > d.anchor_cap) {
      uint64_t cap = d.anchor_cap ? 2 * d.anchor_cap : 16;
      while (cap < need) cap *= 2;
      double *na = nullptr;
      CK(h, hipMalloc(&na, cap * per * sizeof(double)));
      if (d.d_anchor) {
        CK(h, hipMemcpyAsync(na, d.d_anchor, d.anchors * per * sizeof(double),
                             hipMemcpyDeviceToDevice, h->stream));
        CK(h, hipStreamSynchronize(h->stream));
        CK(h, hipFree(d.d_anchor));
      }
      d.d_anchor = na, d.anchor_cap = cap;
    }
    std::vector<double> a((need - d.anchors) * per);
    for (uint64_t m = d.anchors + 1; m <= need; m++)
      ems_dft_phasors(d.freqs.data(), d.nf, d.t0, d.dt, 1024 * m,
                      a.data() + (m - d.anchors - 1) * per);
    CK(h, hipMemcpyAsync(d.d_anchor + d.anchors * per, a.data(), a.size() * sizeof(double),
                         hipMemcpyHostToDevice, h->stream));
    CK(h, hipStreamSynchronize(h->stream));  // (a is a host temporary)
    d.anchors = need;
  }
  const double *src = d.which == 0 ? o->lp.rec : d.which == 1 ? o->lp.pv : o->lp.pi;
  hipLaunchKernelGGL(dft_run, dim3((d.nf + 255) / 256), dim3(256), 0, h->stream, src, d.done, n,
                     d.nf, d.d_state, (const double *)d.d_step, (const double *)d.d_anchor);
  CK(h, hipGetLastError());
  d.done = n;
  return EMS_OK;
}

int ems_gpu_dft_result(ems_gpu *h, int dft_id, double *out, uint64_t *samples) {
  ALIVE(h);
  if (dft_id < 0 || dft_id >= (int)h->dfts.size() || !out)
    return fail(h, EMS_E_INVALID, "not a dft id");
  const HostDft &d = h->dfts[dft_id];
  std::vector<double> st(4 * (size_t)d.nf);
  CK(h, hipStreamSynchronize(h->stream));
  CK(h, hipMemcpy(st.data(), d.d_state, st.size() * sizeof(double), hipMemcpyDeviceToHost));
  // RunningDft::result (dft.rs:170-178): acc * dt
  for (int k = 0; k < d.nf; k++) out[2 * k] = st[4 * k] * d.dt, out[2 * k + 1] = st[4 * k + 1] * d.dt;
  if (samples) *samples = d.done;
  return EMS_OK;
}

int ems_gpu_lumped_branch(ems_gpu *h, int id, double *j, double *q, double *j_avg) {
  ALIVE(h);
  HostOp *o = lumped_op(h, id);
  if (!o) return fail(h, EMS_E_INVALID, "not a lumped element id");
  CK(h, hipStreamSynchronize(h->stream));
  const size_t nb = (size_t)o->lp.n * sizeof(double);
  if (j) CK(h, hipMemcpy(j, o->lp.j, nb, hipMemcpyDeviceToHost));
  if (q) CK(h, hipMemcpy(q, o->lp.q, nb, hipMemcpyDeviceToHost));
  if (j_avg) CK(h, hipMemcpy(j_avg, o->lp.javg, nb, hipMemcpyDeviceToHost));
  return EMS_OK;
}

int ems_gpu_add_dispersive(ems_gpu *h, const ems_pole_side *electric,
                           const ems_pole_side *magnetic, const ems_pole_consts *consts,
                           int *id) {
  ALIVE(h);
  if (!electric || !magnetic || !consts || consts->n < 0 ||
      (consts->n && (!consts->a || !consts->inv_c || !consts->inv_z || !consts->half_l ||
                     !consts->kind)))
    return fail(h, EMS_E_INVALID, "dispersive medium: null description");
  for (int c = 0; c < consts->n; c++)
    if (consts->kind[c] != 0 && consts->kind[c] != 1)
      return fail(h, EMS_E_INVALID, "dispersive medium: pole kind must be 0 or 1");
  HostOp o;
  o.type = OP_POLES;
  auto put = [&](const void *src, size_t bytes, void **dst) -> int {
    int rc = upload(h, dst, src, bytes);
    if (!rc) o.mem.push_back(*dst);
    return rc;
  };
  const ems_pole_side *sd[2] = {electric, magnetic};
  int rc;
  for (int f = 0; f < 2; f++) {
    const ems_pole_side *d = sd[f];
    const int ne = d->n_edges;
    if (ne < 0 || (ne && (!d->axis || !d->nodes || !d->sum_inv_z || !d->pole_end)) ||
        (d->n_poles && !d->cidx))
      return fail(h, EMS_E_INVALID, "dispersive side arrays");
    std::vector<uint64_t> idx(ne);
    uint32_t last = 0;
    for (int e = 0; e < ne; e++) {
      const int32_t *nd = d->nodes + 3 * e;
      const int a = d->axis[e];
      if (a < 0 || a > 2 || !node_ok(h, nd))
        return fail(h, EMS_E_INVALID, "dispersive edge out of range");
      if (d->pole_end[e] <= last || d->pole_end[e] > d->n_poles)
        return fail(h, EMS_E_INVALID, "dispersive pole_end must increase up to n_poles");
      last = d->pole_end[e];
      idx[e] = ((uint64_t)nd[0] * h->n[1] + nd[1]) * h->n[2] + nd[2];
    }
    if (ne && last != d->n_poles) return fail(h, EMS_E_INVALID, "dispersive pole count");
    for (uint64_t p = 0; p < d->n_poles; p++)
      if ((int)d->cidx[p] >= consts->n) return fail(h, EMS_E_INVALID, "dispersive pole constant");
    DevPoleSide &s = o.side[f];
    s.n = ne;
    const std::vector<double> zeros(d->n_poles ? d->n_poles : 1, 0.0);
    const std::vector<float> zf(ne ? ne : 1, 0.0f);
    if ((rc = put(d->axis, ne * sizeof(int32_t), (void **)&s.axis))) return rc;
    if ((rc = put(idx.data(), ne * sizeof(uint64_t), (void **)&s.idx))) return rc;
    if ((rc = put(d->sum_inv_z, ne * sizeof(double), (void **)&s.sum_inv_z))) return rc;
    if ((rc = put(d->pole_end, ne * sizeof(uint32_t), (void **)&s.pole_end))) return rc;
    if ((rc = put(d->cidx, d->n_poles * sizeof(uint32_t), (void **)&s.cidx))) return rc;
    if ((rc = put(zeros.data(), zeros.size() * sizeof(double), (void **)&s.j))) return rc;
    if ((rc = put(zeros.data(), zeros.size() * sizeof(double), (void **)&s.q))) return rc;
    if ((rc = put(zf.data(), zf.size() * sizeof(float), (void **)&s.prev))) return rc;
    o.side_cidx[f].assign(d->cidx, d->cidx + d->n_poles);
  }
  std::vector<DevPoleConst> cs(consts->n ? consts->n : 1);
  for (int c = 0; c < consts->n; c++)
    cs[c] = {consts->a[c], consts->inv_c[c], consts->inv_z[c], consts->kind[c]};
  if ((rc = put(cs.data(), cs.size() * sizeof(DevPoleConst), (void **)&o.d_consts))) return rc;
  o.c_inv_c.assign(consts->inv_c, consts->inv_c + consts->n);
  o.c_half_l.assign(consts->half_l, consts->half_l + consts->n);
  h->ops.push_back(o);
  h->ops_dirty = true;
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

int ems_gpu_dispersive_state(ems_gpu *h, int id, int field, double *j, double *q) {
  ALIVE(h);
  if (id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_POLES || field < 0 ||
      field > 1)
    return fail(h, EMS_E_INVALID, "not a dispersive medium id / field");
  HostOp &o = h->ops[id];
  CK(h, hipStreamSynchronize(h->stream));
  const size_t nb = o.side_cidx[field].size() * sizeof(double);
  if (nb && j) CK(h, hipMemcpy(j, o.side[field].j, nb, hipMemcpyDeviceToHost));
  if (nb && q) CK(h, hipMemcpy(q, o.side[field].q, nb, hipMemcpyDeviceToHost));
  return EMS_OK;
}

int ems_gpu_add_planewave(ems_gpu *h, const ems_pw_desc *d, int *id) {
  ALIVE(h);
  if (!d || d->n_aux < 3 || !d->ca || !d->cb || !d->da || !d->db)
    return fail(h, EMS_E_INVALID, "plane wave: auxiliary grid arrays");
  HostOp o;
  o.type = OP_PW;
  auto put = [&](const void *src, size_t bytes, void **dst) -> int {
    int rc = upload(h, dst, src, bytes);
    if (!rc) o.mem.push_back(*dst);
    return rc;
  };
  const int na = d->n_aux;
  int rc;
  const std::vector<double> z(na, 0.0);
  DevPw &p = o.pw;
  p.naux = na;
  if ((rc = put(z.data(), na * sizeof(double), (void **)&p.e))) return rc;
  if ((rc = put(z.data(), (na - 1) * sizeof(double), (void **)&p.h))) return rc;
  if ((rc = put(d->ca, na * sizeof(double), (void **)&p.ca))) return rc;
  if ((rc = put(d->cb, na * sizeof(double), (void **)&p.cb))) return rc;
  if ((rc = put(d->da, (na - 1) * sizeof(double), (void **)&p.da))) return rc;
  if ((rc = put(d->db, (na - 1) * sizeof(double), (void **)&p.db))) return rc;
  const ems_pw_corrections *cs[2] = {&d->e_corr, &d->h_corr};
  for (int side = 0; side < 2; side++) {
    const ems_pw_corrections *c = cs[side];
    const int nc = c->n;
    if (nc < 0 || (nc && (!c->axis || !c->nodes || !c->coef || !c->m || !c->w)))
      return fail(h, EMS_E_INVALID, "plane wave: correction arrays");
    // Stable order by corrected edge: each edge's corrections stay in list
    // order and one thread applies them.
    std::vector<uint64_t> key(nc);
    for (int q = 0; q < nc; q++) {
      const int32_t *nd = c->nodes + 3 * q;
      if (c->axis[q] < 0 || c->axis[q] > 2 || !node_ok(h, nd))
        return fail(h, EMS_E_INVALID, "plane wave: corrected edge out of range");
      if ((int64_t)c->m[q] + 3 >= (int64_t)(side == 0 ? na - 1 : na))
        return fail(h, EMS_E_INVALID, "plane wave: stencil beyond the auxiliary grid");
      key[q] = ((((uint64_t)nd[0] * h->n[1] + nd[1]) * h->n[2] + nd[2]) << 2) | (uint64_t)c->axis[q];
    }
    std::vector<int> ord(nc);
    for (int q = 0; q < nc; q++) ord[q] = q;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<uint32_t> start, m(nc);
    std::vector<uint64_t> idx;
    std::vector<int32_t> axis(nc);
    std::vector<double> coef(nc), w(4 * (size_t)(nc ? nc : 1));
    for (int r = 0; r < nc; r++) {
      const int q = ord[r];
      if (r == 0 || key[q] != key[ord[r - 1]]) {
        start.push_back((uint32_t)r);
        idx.push_back(key[q] >> 2);
      }
      axis[r] = c->axis[q], coef[r] = c->coef[q], m[r] = c->m[q];
      for (int t = 0; t < 4; t++) w[4 * (size_t)r + t] = c->w[4 * (size_t)q + t];
    }
    start.push_back((uint32_t)nc);
    DevPwCorr &dc = o.pwc[side];
    dc.ntgt = (int)idx.size();
    if ((rc = put(start.data(), start.size() * sizeof(uint32_t), (void **)&dc.start))) return rc;
    if ((rc = put(idx.data(), idx.size() * sizeof(uint64_t), (void **)&dc.idx))) return rc;
    if ((rc = put(axis.data(), nc * sizeof(int32_t), (void **)&dc.axis))) return rc;
    if ((rc = put(coef.data(), nc * sizeof(double), (void **)&dc.coef))) return rc;
    if ((rc = put(m.data(), nc * sizeof(uint32_t), (void **)&dc.m))) return rc;
    if ((rc = put(w.data(), w.size() * sizeof(double), (void **)&dc.w))) return rc;
  }
  h->ops.push_back(o);
  h->ops_dirty = true;
  drop_graphs(h);
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

int ems_gpu_planewave_table(ems_gpu *h, int id, const double *values, uint64_t count) {
  ALIVE(h);
  if (id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_PW || (count && !values))
    return fail(h, EMS_E_INVALID, "not a plane wave id / values");
  HostOp &o = h->ops[id];
  CK(h, hipStreamSynchronize(h->stream));
  double *nt = nullptr;
  int rc = upload(h, (void **)&nt, values, count * sizeof(double));
  if (rc) return rc;
  if (o.pw_src) CK(h, hipFree(o.pw_src));
  o.pw_src = nt;
  o.pw.src = nt;
  o.pw.nsrc = count;
  drop_graphs(h);  // the kernels take the table by value
  return EMS_OK;
}

int ems_gpu_planewave_aux(ems_gpu *h, int id, double *e, double *hh) {
  ALIVE(h);
  if (id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_PW)
    return fail(h, EMS_E_INVALID, "not a plane wave id");
  const HostOp &o = h->ops[id];
  CK(h, hipStreamSynchronize(h->stream));
  if (e) CK(h, hipMemcpy(e, o.pw.e, o.pw.naux * sizeof(double), hipMemcpyDeviceToHost));
  if (hh) CK(h, hipMemcpy(hh, o.pw.h, (o.pw.naux - 1) * sizeof(double), hipMemcpyDeviceToHost));
  return EMS_OK;
}

int ems_gpu_add_merged(ems_gpu *h, const ems_merged_desc *d, int *id) {
  ALIVE(h);
  if (!d || d->n_groups < 0 || d->n_scalar < 0 || d->n_scalar > d->n_groups ||
      d->n_blocks < 0 || d->n_faces < 0 || !(d->dt > 0.0))
    return fail(h, EMS_E_INVALID, "reduced currents: counts");
  const uint32_t ng = (uint32_t)d->n_groups, nb = (uint32_t)d->n_blocks,
                 nf = (uint32_t)d->n_faces;
  if (ng && (!d->da || !d->db || !d->l))
    return fail(h, EMS_E_INVALID, "reduced currents: groups");
  if (!d->term_start || !d->blk_start || !d->env_off || !d->brow_start || !d->mrow_start ||
      !d->face_start)
    return fail(h, EMS_E_INVALID, "reduced currents: offset tables");
  // Every index the kernels follow is checked here.
  const uint32_t nt = ng ? d->term_start[ng] : 0;
  for (uint32_t g = 0; g < ng; g++)
    if (d->term_start[g] > d->term_start[g + 1])
      return fail(h, EMS_E_INVALID, "reduced currents: term ranges");
  if (nt && (!d->term_axis || !d->term_idx || !d->term_w))
    return fail(h, EMS_E_INVALID, "reduced currents: terms");
  for (uint32_t t = 0; t < nt; t++)
    if (d->term_axis[t] < 0 || d->term_axis[t] > 2 || d->term_idx[t] >= h->N)
      return fail(h, EMS_E_INVALID, "reduced currents: term edge out of range");
  if (d->blk_start[0] != 0) return fail(h, EMS_E_INVALID, "reduced currents: block rows");
  for (uint32_t b = 0; b < nb; b++)
    if (d->blk_start[b] > d->blk_start[b + 1])
      return fail(h, EMS_E_INVALID, "reduced currents: block rows");
  const uint32_t nr = d->blk_start[nb];
  if ((uint64_t)d->n_scalar + nr != ng)
    return fail(h, EMS_E_INVALID, "reduced currents: every non-scalar current is one block row");
  if (nr && (!d->members || !d->env_first || !d->env_val))
    return fail(h, EMS_E_INVALID, "reduced currents: block tables");
  std::vector<uint8_t> seen(ng, 0);
  for (uint32_t b = 0; b < nb; b++) {
    const uint32_t r0 = d->blk_start[b], r1 = d->blk_start[b + 1];
    for (uint32_t r = r0; r < r1; r++) {
      const uint32_t i = r - r0;
      if (d->members[r] < (uint32_t)d->n_scalar || d->members[r] >= ng || seen[d->members[r]]++)
        return fail(h, EMS_E_INVALID, "reduced currents: block members");
      if (d->env_first[r] > i || d->env_off[r + 1] < d->env_off[r] ||
          d->env_off[r + 1] - d->env_off[r] != (uint64_t)(i - d->env_first[r] + 1))
        return fail(h, EMS_E_INVALID, "reduced currents: envelope rows");
      for (int w = 0; w < 2; w++) {
        const uint32_t *st = w ? d->mrow_start : d->brow_start;
        const uint32_t *col = w ? d->mcol : d->bcol;
        if (st[r + 1] < st[r]) return fail(h, EMS_E_INVALID, "reduced currents: B/M rows");
        for (uint32_t e = st[r]; e < st[r + 1]; e++)
          if (!col || col[e] >= r1 - r0)
            return fail(h, EMS_E_INVALID, "reduced currents: B/M column out of the block");
      }
    }
  }
  if (!d->col_start || d->col_start[0] != 0)
    return fail(h, EMS_E_INVALID, "reduced currents: column offsets");
  for (uint32_t b = 0; b < nb; b++) {
    const uint32_t r0 = d->blk_start[b], r1 = d->blk_start[b + 1];
    for (uint32_t r = r0; r < r1; r++) {
      if (d->col_start[r + 1] < d->col_start[r])
        return fail(h, EMS_E_INVALID, "reduced currents: column ranges");
      for (uint32_t q = d->col_start[r]; q < d->col_start[r + 1]; q++)
        if (!d->col_row || !d->col_val || d->col_row[q] <= r - r0 || d->col_row[q] >= r1 - r0)
          return fail(h, EMS_E_INVALID, "reduced currents: column row out of the block");
    }
  }
  if (d->n_pml < 0 || !d->pml_start || d->pml_start[0] != 0)
    return fail(h, EMS_E_INVALID, "reduced currents: CPML term tables");
  const uint32_t npml = (uint32_t)d->n_pml;
  for (uint32_t g = 0; g < ng; g++)
    if (d->pml_start[g + 1] < d->pml_start[g])
      return fail(h, EMS_E_INVALID, "reduced currents: CPML term ranges");
  const uint32_t npw = d->pml_start[ng];
  if ((npml && (!d->pml_axis || !d->pml_idx0 || !d->pml_idx1 || !d->pml_coef)) ||
      (npw && (!d->pml_term || !d->pml_w)))
    return fail(h, EMS_E_INVALID, "reduced currents: CPML terms");
  for (uint32_t q = 0; q < npml; q++)
    if (d->pml_axis[q] < 0 || d->pml_axis[q] > 2 || d->pml_idx0[q] >= h->N ||
        d->pml_idx1[q] >= h->N)
      return fail(h, EMS_E_INVALID, "reduced currents: CPML term voltage out of range");
  for (uint32_t e = 0; e < npw; e++)
    if (d->pml_term[e] >= npml) return fail(h, EMS_E_INVALID, "reduced currents: CPML target");
  if (d->env_off[0] != 0 || d->brow_start[0] != 0 || d->mrow_start[0] != 0)
    return fail(h, EMS_E_INVALID, "reduced currents: row offsets");
  if (d->face_start[0] != 0) return fail(h, EMS_E_INVALID, "reduced currents: face ranges");
  for (uint32_t q = 0; q < nf; q++)
    if (d->face_start[q] > d->face_start[q + 1])
      return fail(h, EMS_E_INVALID, "reduced currents: face ranges");
  const uint32_t nw = d->face_start[nf];
  if (nf && (!d->face_axis || !d->face_idx || (nw && (!d->face_group || !d->face_w))))
    return fail(h, EMS_E_INVALID, "reduced currents: faces");
  for (uint32_t q = 0; q < nf; q++)
    if (d->face_axis[q] < 0 || d->face_axis[q] > 2 || d->face_idx[q] >= h->N)
      return fail(h, EMS_E_INVALID, "reduced currents: driven face out of range");
  for (uint32_t e = 0; e < nw; e++)
    if (d->face_group[e] >= ng) return fail(h, EMS_E_INVALID, "reduced currents: face weights");
  const uint64_t nev = nr ? d->env_off[nr] : 0;
  const uint32_t nbv = nr ? d->brow_start[nr] : 0, nmv = nr ? d->mrow_start[nr] : 0;

  HostOp o;
  o.type = OP_MERGED;
  auto put = [&](const void *src, size_t bytes, void **dst) -> int {
    int rc = upload(h, dst, src, bytes);
    if (!rc) o.mem.push_back(*dst);
    return rc;
  };
  DevMerged &m = o.mf;
  m.ng = (int)ng, m.ns = d->n_scalar, m.nb = (int)nb, m.nf = (int)nf, m.dt = d->dt;
  int rc;
  const std::vector<double> z(std::max<size_t>(std::max({ng, nr, npml}), 1), 0.0);
#define PUT(dst, src, count, T)                                                           \
  if ((rc = put((src), (size_t)(count) * sizeof(T), (void **)&(dst)))) return rc;
  PUT(m.da, d->da, ng, double);
  PUT(m.db, d->db, ng, double);
  PUT(m.ts, d->term_start, ng + 1, uint32_t);
  PUT(m.tax, d->term_axis, nt, int32_t);
  PUT(m.tidx, d->term_idx, nt, uint64_t);
  PUT(m.tw, d->term_w, nt, double);
  PUT(m.bs, d->blk_start, nb + 1, uint32_t);
  PUT(m.mem, d->members, nr, uint32_t);
  PUT(m.ef, d->env_first, nr, uint32_t);
  PUT(m.eo, d->env_off, nr + 1, uint64_t);
  PUT(m.ev, d->env_val, nev, double);
  PUT(m.brs, d->brow_start, nr + 1, uint32_t);
  PUT(m.bc, d->bcol, nbv, uint32_t);
  PUT(m.bv, d->bval, nbv, double);
  const uint32_t ncv = nr ? d->col_start[nr] : 0;
  PUT(m.cs, d->col_start, nr + 1, uint32_t);
  PUT(m.cr, d->col_row, ncv, uint32_t);
  PUT(m.cv, d->col_val, ncv, double);
  m.np = (int)npml;
  PUT(m.pax, d->pml_axis, npml, int32_t);
  PUT(m.pi0, d->pml_idx0, npml, uint64_t);
  PUT(m.pi1, d->pml_idx1, npml, uint64_t);
  PUT(m.pc, d->pml_coef, 4 * (size_t)npml, double);
  PUT(m.pst, d->pml_start, ng + 1, uint32_t);
  PUT(m.pt, d->pml_term, npw, uint32_t);
  PUT(m.pw, d->pml_w, npw, double);
  PUT(m.psi, z.data(), npml, double);
  PUT(m.pcon, z.data(), npml, double);
  PUT(m.fax, d->face_axis, nf, int32_t);
  PUT(m.fidx, d->face_idx, nf, uint64_t);
  PUT(m.fs, d->face_start, nf + 1, uint32_t);
  PUT(m.fg, d->face_group, nw, uint32_t);
  PUT(m.fw, d->face_w, nw, double);
  PUT(m.cur, z.data(), ng, double);
  PUT(m.next, z.data(), ng, double);
  PUT(m.circ, z.data(), ng, double);
  PUT(m.x, z.data(), nr, double);
#undef PUT
  auto mh = std::make_shared<MergedHost>();
  auto cp = [](auto &v, const auto *src, size_t n) { v.assign(src, src + n); };
  cp(mh->da, d->da, ng);
  cp(mh->db, d->db, ng);
  cp(mh->l, d->l, ng);
  cp(mh->bs, d->blk_start, nb + 1);
  cp(mh->mem, d->members, nr);
  cp(mh->ef, d->env_first, nr);
  cp(mh->eo, d->env_off, nr + 1);
  cp(mh->ev, d->env_val, nev);
  cp(mh->brs, d->brow_start, nr + 1);
  cp(mh->bc, d->bcol, nbv);
  cp(mh->bv, d->bval, nbv);
  cp(mh->mrs, d->mrow_start, nr + 1);
  cp(mh->mcol, d->mcol, nmv);
  cp(mh->mval, d->mval, nmv);
  o.mfh = mh;
  std::vector<uint32_t> small, big;
  size_t lds_small = 0, lds_big = 0;
  for (uint32_t b = 0; b < nb; b++) {
    const uint32_t r0 = d->blk_start[b], r1 = d->blk_start[b + 1];
    bool narrow = true;
    for (uint32_t r = r0; r < r1 && narrow; r++) narrow = r - r0 - d->env_first[r] <= 3;
    size_t need = mf_lds_need(r1 - r0, d->env_off[r1] - d->env_off[r0],
                              d->col_start[r1] - d->col_start[r0]);
    uint32_t tag = b;
    if (narrow && mf_narrow_need(r1 - r0) <= 65536) {
      need = mf_narrow_need(r1 - r0);
      tag |= 0x80000000u;
    }
    if (need <= 8192) {
      small.push_back(tag);
      lds_small = std::max(lds_small, need);
    } else {
      big.push_back(tag);
      if (need <= 65536) lds_big = std::max(lds_big, need);
    }
  }
  if ((rc = put(small.data(), small.size() * sizeof(uint32_t), (void **)&o.mf_small))) return rc;
  if ((rc = put(big.data(), big.size() * sizeof(uint32_t), (void **)&o.mf_big))) return rc;
  if (!h->side) {
    CK(h, hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    CK(h, hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    CK(h, hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
  }
  o.mf_nsmall = (int)small.size(), o.mf_nbig = (int)big.size();
  o.mf_lds_small = (int)lds_small, o.mf_lds_big = (int)lds_big;
  h->ops.push_back(o);
  h->ops_dirty = true;
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

int ems_gpu_merged_state(ems_gpu *h, int id, double *cur, double *circ) {
  ALIVE(h);
  if (id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_MERGED)
    return fail(h, EMS_E_INVALID, "not a reduced-current operator id");
  HostOp &o = h->ops[id];
  int rc = merged_circ_now(h, o);
  if (rc) return rc;
  CK(h, hipStreamSynchronize(h->stream));
  const size_t bytes = (size_t)o.mf.ng * sizeof(double);
  if (cur && bytes) CK(h, hipMemcpy(cur, o.mf.cur, bytes, hipMemcpyDeviceToHost));
  if (circ && bytes) CK(h, hipMemcpy(circ, o.mf.circ, bytes, hipMemcpyDeviceToHost));
  return EMS_OK;
}

int ems_gpu_set_schedule(ems_gpu *h, int fused, int tile_rows, int x_chunks) {
  ALIVE(h);
  // fused 2: the fused kernel without the next-plane prefetch; 3: the
  // full-row fused kernel (nz <= 576, tile rows 3 or 7); 4: packed short
  // rows (nz <= 64; otherwise the tiled kernel), 5: packed without the
  // prefetch.
  // 6: bands of rows per wave (tile rows 4, 8 or 16: the band height).
  if (fused < 0 || fused > 6) return fail(h, EMS_E_INVALID, "schedule mode must be 0-6");
  if (fused == 6) {
    if (tile_rows != 4 && tile_rows != 8 && tile_rows != 16)
      return fail(h, EMS_E_INVALID, "band kernel: tile_rows must be 4, 8 or 16");
  } else if (tile_rows != 0 && tile_rows != 3 && tile_rows != 7 && tile_rows != 15) {
    return fail(h, EMS_E_INVALID, "tile_rows must be 3, 7 or 15");
  } else if (tile_rows == 0 && h->band) {
    h->tj = 7;  // leaving the band kernel: back to a tile height
  }
  h->prefetch = fused != 2 && fused != 5;
  h->rows = fused == 3;
  h->packed = fused == 4 || fused == 5;
  h->band = fused == 6;
  if (x_chunks < 0) return fail(h, EMS_E_INVALID, "x_chunks must be >= 0");
  h->fused = fused != 0;
  if (tile_rows) h->tj = tile_rows;
  h->nchunk = x_chunks;
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_set_cpml_split(ems_gpu *h, int on) {
  ALIVE(h);
  if (on != 0 && on != 1) return fail(h, EMS_E_INVALID, "cpml split must be 0 or 1");
  h->cpml_split = on != 0;
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_set_unfused_prefetch(ems_gpu *h, int which) {
  ALIVE(h);
  if (which < 0 || which > 15) return fail(h, EMS_E_INVALID, "unfused kernels: 0 to 15");
  h->unfused_pf = which;
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_set_tile_classes(ems_gpu *h, int on) {
  ALIVE(h);
  h->tile_class = on != 0;
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_set_xcd_group(ems_gpu *h, int group) {
  ALIVE(h);
  if (group < 0 || group > 64) return fail(h, EMS_E_INVALID, "xcd group must be 0..64");
  h->xcd_group = group;
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_run(ems_gpu *h, uint64_t n) {
  ALIVE(h);
  return run_steps(h, n);
}

// Engine::operator_energy (mod.rs:753-755): Operator::energy of every
// registered operator in order, Sum folding from -0.0. Every kind held here
// has 0 but a lumped element with a branch (branch_energy,
// lumped.rs:1075-1087, from the device state); the CPML slabs are one
// operator.
// The operators' energies in two parts. ops_energy_enqueue queues on the
// engine's stream what each needs from the device and its copy into pinned
// host memory: the poles' terms, the reduced currents' terms after their
// look-ahead update (written to the update's own result buffer, which the
// next step's update overwrites before anything reads it), the lumped
// branches' state. ops_energy_finish, once that copy has landed (the
// engine's ev_energy), adds them in the order of the sums it restates.
static int ops_energy_enqueue(ems_gpu *h) {
  for (auto &o : h->ops) {
    size_t n = 0;
    if (o.type == OP_POLES)
      n = o.side_cidx[0].size() + o.side_cidx[1].size();
    else if (o.type == OP_MERGED)
      n = (size_t)o.mf.ns + o.mfh->mem.size();
    else if (o.type == OP_LUMPED && o.lp.branch)
      n = 2 * (size_t)o.lp.n;
    if (!n) continue;
    if (!o.h_et) {
      CK(h, hipHostMalloc((void **)&o.h_et, n * sizeof(double), hipHostMallocDefault));
      o.n_et = n;
      if (o.type != OP_LUMPED) {
        CK(h, hipMalloc(&o.d_et, n * sizeof(double)));
        o.mem.push_back(o.d_et);
      }
      if (o.type == OP_POLES) {
        std::vector<double> hc(2 * o.c_half_l.size());
        for (size_t c = 0; c < o.c_half_l.size(); c++)
          hc[2 * c] = o.c_half_l[c], hc[2 * c + 1] = o.c_inv_c[c];
        int rc = upload(h, (void **)&o.d_ec, hc.data(), hc.size() * sizeof(double));
        if (rc) return rc;
        o.mem.push_back(o.d_ec);
      }
      if (o.type == OP_MERGED) {
        const MergedHost &mh = *o.mfh;
        const size_t nr = mh.mem.size();
        std::vector<uint32_t> row0(nr);
        for (size_t b = 0; b + 1 < mh.bs.size(); b++)
          for (uint32_t r = mh.bs[b]; r < mh.bs[b + 1]; r++) row0[r] = mh.bs[b];
        // d_ec: l (ng), then the M values; d_erow0: block starts, M row
        // starts (nr + 1), M columns
        std::vector<double> dv(mh.l);
        dv.insert(dv.end(), mh.mval.begin(), mh.mval.end());
        std::vector<uint32_t> uv(row0);
        uv.insert(uv.end(), mh.mrs.begin(), mh.mrs.end());
        uv.insert(uv.end(), mh.mcol.begin(), mh.mcol.end());
        int rc = upload(h, (void **)&o.d_ec, dv.data(), dv.size() * sizeof(double));
        if (rc) return rc;
        o.mem.push_back(o.d_ec);
        if ((rc = upload(h, (void **)&o.d_erow0, uv.data(), uv.size() * sizeof(uint32_t))))
          return rc;
        o.mem.push_back(o.d_erow0);
      }
    }
    if (o.type == OP_POLES) {
      size_t off = 0;
      for (int sd = 0; sd < 2; sd++) {
        const int np = (int)o.side_cidx[sd].size();
        if (np)
          hipLaunchKernelGGL(pole_energy_terms, dim3((np + 255) / 256), dim3(256), 0, h->stream,
                             o.side[sd], np, (const double *)o.d_ec, o.d_et + off);
        off += (size_t)np;
      }
    } else if (o.type == OP_MERGED) {
      int rc = merged_circ_now(h, o);
      if (rc) return rc;
      const DevMerged &m = o.mf;
      const int na = o.mf_nsmall + (m.ns + 63) / 64;
      if (na)
        hipLaunchKernelGGL(mf_advance, dim3(na), dim3(64), (size_t)o.mf_lds_small, h->stream, m,
                           (const uint32_t *)o.mf_small, o.mf_nsmall, o.mf_lds_small);
      if (o.mf_nbig)
        hipLaunchKernelGGL(mf_advance, dim3(o.mf_nbig), dim3(64), (size_t)o.mf_lds_big, h->stream,
                           m, (const uint32_t *)o.mf_big, o.mf_nbig, o.mf_lds_big);
      const int nr = (int)o.mfh->mem.size(), ng = (int)o.mfh->l.size();
      const uint32_t *row0 = o.d_erow0, *mrs = row0 + nr, *mcol = mrs + nr + 1;
      hipLaunchKernelGGL(mf_energy_terms, dim3((unsigned)((o.n_et + 255) / 256)), dim3(256), 0,
                         h->stream, m, (const double *)o.d_ec, mrs, mcol,
                         (const double *)o.d_ec + ng, row0, nr, o.d_et);
    }
    CK(h, hipGetLastError());
    if (o.type == OP_LUMPED) {
      const size_t ne = (size_t)o.lp.n;
      CK(h, hipMemcpyAsync(o.h_et, o.lp.j, ne * sizeof(double), hipMemcpyDeviceToHost, h->stream));
      CK(h, hipMemcpyAsync(o.h_et + ne, o.lp.q, ne * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
    } else {
      CK(h, hipMemcpyAsync(o.h_et, o.d_et, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
  }
  return 0;
}

static double ops_energy_finish(const ems_gpu *h) {
  double w = -0.0;
  if (!h->slabs.empty()) w += 0.0;
  for (auto &o : h->ops) {
    if (o.type == OP_POLES) {
      // Dispersive::energy (dispersive.rs:2789-2794): electric + magnetic,
      // each summed from 0.0 in pole order (Side::add_energy :1489-1517).
      double side_w[2] = {0.0, 0.0};
      size_t off = 0;
      for (int sd = 0; sd < 2; sd++) {
        const size_t np = o.side_cidx[sd].size();
        double acc = 0.0;
        for (size_t p = 0; p < np; p++) acc += o.h_et[off + p];
        side_w[sd] = acc;
        off += np;
      }
      w += side_w[0] + side_w[1];
      continue;
    }
    if (o.type == OP_MERGED) {
      // MergedFaces::energy (conformal.rs:2684-2698): 1/2 g^T M g' from
      // 0.0, the merge groups' terms first, then the block rows in order.
      double e = 0.0;
      for (size_t t = 0; t < o.n_et; t++) e += o.h_et[t];
      w += e;
      continue;
    }
    if (o.type != OP_LUMPED || !o.lp.branch) {
      w += 0.0;
      continue;
    }
    const int ne = o.lp.n;
    const double *j = o.h_et, *q = o.h_et + ne;
    double s = -0.0;
    for (int e = 0; e < ne; e++) {
      const double half_l = 0.25 * o.two_l_dt[e] * h->dt;
      s += half_l * j[e] * j[e] + 0.5 * o.inv_c[e] * q[e] * q[e];
    }
    w += s;
  }
  return w;
}

// The field energy's and the operators' device parts and copies, queued,
// and the event that marks their landing.
static int energy_all_enqueue(ems_gpu *h) {
  int rc = energy_enqueue(h);
  if (rc) return rc;
  if ((rc = ops_energy_enqueue(h))) return rc;
  if (!h->ev_energy) CK(h, hipEventCreateWithFlags(&h->ev_energy, hipEventDisableTiming));
  CK(h, hipEventRecord(h->ev_energy, h->stream));
  return 0;
}

// The step is queued before the wait, so the device runs it while the host
// adds the operators' terms.
int ems_gpu_step_with_energy(ems_gpu *h, double *w) {
  ALIVE(h);
  if (!w) return EMS_E_INVALID;
  int rc = energy_all_enqueue(h);
  if (rc) return rc;
  if ((rc = run_steps(h, 1))) return rc;
  CK(h, hipEventSynchronize(h->ev_energy));
  const double w_ops = ops_energy_finish(h);
  *w = *h->h_w + w_ops;
  return EMS_OK;
}

int ems_gpu_energy(ems_gpu *h, double *w) {
  ALIVE(h);
  if (!w) return EMS_E_INVALID;
  int rc = energy_all_enqueue(h);
  if (rc) return rc;
  CK(h, hipEventSynchronize(h->ev_energy));
  const double w_ops = ops_energy_finish(h);
  *w = *h->h_w + w_ops;
  return EMS_OK;
}

// lookahead_source_energy (dump.rs:1735-1782): mu0 sum w ((k + s)^2 - k^2)
// over the H sources' dual edges (sorted by axis and node, without
// repeats; w = 4 on the outer planes normal to the edge), k the kernel's
// look-ahead (completed by the layers on their edges, on the device), k + s
// with the next step's H sources applied in registration order.
static int hsource_energy(ems_gpu *h, double *out) {
  *out = 0.0;
  std::vector<std::pair<int32_t, uint64_t>> ed;
  for (auto &o : h->ops)
    if (o.type == OP_SOURCE && (o.kind == 2 || o.kind == 3))
      for (size_t q = 0; q < o.h_idx.size(); q++) ed.push_back({o.h_axis[q], o.h_idx[q]});
  if (ed.empty()) return 0;
  std::sort(ed.begin(), ed.end());
  ed.erase(std::unique(ed.begin(), ed.end()), ed.end());
  const int n = (int)ed.size();
  std::vector<int32_t> ax(n);
  std::vector<uint64_t> ix(n);
  for (int q = 0; q < n; q++) ax[q] = ed[q].first, ix[q] = ed[q].second;
  const DevUpml *pm = nullptr;
  for (auto &o : h->ops)
    if (o.type == OP_UPML) pm = &o.upml;
  int32_t *d_ax = nullptr;
  uint64_t *d_ix = nullptr;
  float *d_k = nullptr;
  int rc = 0;
  auto cleanup = [&]() {
    if (d_ax) (void)hipFree(d_ax);
    if (d_ix) (void)hipFree(d_ix);
    if (d_k) (void)hipFree(d_k);
  };
  if ((rc = upload(h, (void **)&d_ax, ax.data(), n * sizeof(int32_t))) ||
      (rc = upload(h, (void **)&d_ix, ix.data(), n * sizeof(uint64_t)))) {
    cleanup();
    return rc;
  }
  if (hipMalloc(&d_k, n * sizeof(float)) != hipSuccess) {
    cleanup();
    return fail(h, EMS_E_DEVICE, "hipMalloc (H source look-ahead)");
  }
  const Fields &f = h->fb[h->par];
  const DevUpml m = pm ? *pm : DevUpml{};
  if (h->oe)
    hipLaunchKernelGGL(lookahead_edges<true>, dim3((n + 63) / 64), dim3(64), 0, h->stream, f,
                       lanes_arg(h), n, d_ax, d_ix, h->n[0], h->n[1], h->n[2], m, pm ? 1 : 0, d_k);
  else
    hipLaunchKernelGGL(lookahead_edges<false>, dim3((n + 63) / 64), dim3(64), 0, h->stream, f,
                       lanes_arg(h), n, d_ax, d_ix, h->n[0], h->n[1], h->n[2], m, pm ? 1 : 0, d_k);
  std::vector<float> k(n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(k.data(), d_k, n * sizeof(float), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return fail(h, EMS_E_DEVICE, hipGetErrorString(e));
  double sum = -0.0;
  for (int q = 0; q < n; q++) {
    const int c = ax[q];
    float v = k[q];
    for (auto &o : h->ops) {
      if (!(o.type == OP_SOURCE && (o.kind == 2 || o.kind == 3))) continue;
      const float a = h->step < o.h_amp.size() ? o.h_amp[h->step] : 0.0f;
      for (size_t r = 0; r < o.h_idx.size(); r++)
        if (o.h_axis[r] == c && o.h_idx[r] == ix[q]) {
          const float x = o.h_wf[r] * a;
          v = o.kind == 3 ? x : v + x;
        }
    }
    const uint64_t o_ = ix[q];
    const int p[3] = {(int)(o_ / ((uint64_t)h->n[1] * h->n[2])), (int)((o_ / h->n[2]) % h->n[1]),
                      (int)(o_ % h->n[2])};
    const double wt = (p[c] == 0 || p[c] + 1 == h->n[c]) ? 4.0 : 1.0;
    sum += (double)v * (double)v * wt - (double)k[q] * (double)k[q] * wt;
  }
  *out = MU0_D * sum;
  return 0;
}

int ems_gpu_energy_in(ems_gpu *h, int measure, double *w) {
  ALIVE(h);
  if (!w || measure < 0 || measure > 1) return EMS_E_INVALID;
  if (measure == 0) return ems_gpu_energy(h, w);
  // Engine::energy_in(EnergyMeasure::OpenEms) (mod.rs:747-752)
  double wf = 0.0, ws = 0.0;
  int rc = energy_now(h, &wf, 1);
  if (rc) return rc;
  if ((rc = hsource_energy(h, &ws))) return rc;
  *w = wf + ws;
  return EMS_OK;
}

// The batched form of ems_gpu_energy_many: two or more active models, one
// curl order and one H-row round width (nz class), none with operator
// energies read from the device. Every model's earlier work is waited for,
// then one launch per energy kernel over all of them on the first model's
// stream and one copy of the results; each value then gets its operators'
// (zero) terms as ems_gpu_energy adds them.
static bool energy_many_batchable(ems_gpu *const *models, int n, const uint8_t *active) {
  int cnt = 0, oe = -1, nt = -1;
  uint64_t hb = 0, eb = 0;
  for (int q = 0; q < n; q++) {
    if (active && !active[q]) continue;
    const ems_gpu *h = models[q];
    if (!h || h->dead) return false;
    const int nz = h->n[2];
    const int t = nz <= 32 ? 0 : (nz <= 64 ? 1 : (nz <= 128 ? 2 : (nz <= 256 ? 4 : 8)));
    if ((oe >= 0 && oe != (h->oe ? 1 : 0)) || (nt >= 0 && nt != t)) return false;
    oe = h->oe ? 1 : 0, nt = t;
    for (auto &o : h->ops)
      if (o.type == OP_POLES || o.type == OP_MERGED || (o.type == OP_LUMPED && o.lp.branch))
        return false;
    const uint64_t rpb = t ? 4 : 8 * EH2_GROUPS;  // H rows per block
    hb += ((uint64_t)h->n[0] * (3 * (uint64_t)h->n[1] - 2) + rpb - 1) / rpb;
    eb += ((uint64_t)h->n[0] * 3 + 3) / 4;
    cnt++;
  }
  return cnt >= 2 && hb < (1ull << 31) && eb < (1ull << 31);
}

static int energy_many_batched(ems_gpu *const *models, int n, const uint8_t *active,
                               double *out) {
  std::vector<int> act;
  for (int q = 0; q < n; q++)
    if (!active || active[q]) act.push_back(q);
  ems_gpu *h0 = models[act[0]];
  const int nm = (int)act.size();
  for (int q : act) {
    ems_gpu *h = models[q];
    const int nx = h->n[0], ny = h->n[1];
    if (!h->d_rowsum) {
      CK(h, hipMalloc(&h->d_rowsum, (size_t)nx * (3 * ny - 2) * sizeof(double)));
      CK(h, hipMalloc(&h->d_chains, (size_t)nx * 12 * sizeof(double)));
      CK(h, hipMalloc(&h->d_planes, (size_t)nx * sizeof(double)));
      CK(h, hipMalloc(&h->d_w, sizeof(double)));
      CK(h, hipHostMalloc((void **)&h->h_w, sizeof(double), hipHostMallocDefault));
    }
    CK(h, hipStreamSynchronize(h->stream));
  }
  if (h0->em_cap < nm) {
    if (h0->em_d) {
      CK(h0, hipFree(h0->em_d));
      CK(h0, hipFree(h0->em_dout));
    }
    if (h0->em_h) {
      CK(h0, hipHostFree(h0->em_h));
      CK(h0, hipHostFree(h0->em_hout));
    }
    h0->em_d = h0->em_h = nullptr, h0->em_dout = h0->em_hout = nullptr, h0->em_cap = 0;
    CK(h0, hipMalloc(&h0->em_d, nm * sizeof(EnergyModel)));
    CK(h0, hipMalloc(&h0->em_dout, nm * sizeof(double)));
    CK(h0, hipHostMalloc((void **)&h0->em_h, nm * sizeof(EnergyModel), hipHostMallocDefault));
    CK(h0, hipHostMalloc((void **)&h0->em_hout, nm * sizeof(double), hipHostMallocDefault));
    h0->em_cap = nm;
  }
  const int nz0 = h0->n[2];
  const int nt = nz0 <= 32 ? 0 : (nz0 <= 64 ? 1 : (nz0 <= 128 ? 2 : (nz0 <= 256 ? 4 : 8)));
  const size_t rows_per_block = nt ? 4 : 8 * EH2_GROUPS;  // H rows of one block
  uint32_t hb = 0, eb = 0, pb = 0;
  for (int i = 0; i < nm; i++) {
    ems_gpu *h = models[act[i]];
    EnergyModel &m = h0->em_h[i];
    m.f = h->fb[h->par];
    m.L = lanes_arg(h);
    m.nx = h->n[0], m.ny = h->n[1], m.nz = h->n[2];
    m.dt = h->dt;
    m.rowsum = h->d_rowsum, m.chains = h->d_chains, m.planes = h->d_planes;
    m.hb = hb, m.eb = eb, m.pb = pb;
    hb += (uint32_t)(((size_t)m.nx * (3 * m.ny - 2) + rows_per_block - 1) / rows_per_block);
    eb += (uint32_t)(((size_t)m.nx * 3 + 3) / 4);
    pb += (uint32_t)((m.nx + 3) / 4);
  }
  hipStream_t st = h0->stream;
  CK(h0, hipMemcpyAsync(h0->em_d, h0->em_h, nm * sizeof(EnergyModel), hipMemcpyHostToDevice, st));
  const EnergyModel *tab = h0->em_d;
#define EHM(OE, NT) \
  hipLaunchKernelGGL((energy_h_rows_many<OE, NT>), dim3(hb), dim3(256), 0, st, tab, nm)
  if (nt == 0) {
    if (h0->oe)
      hipLaunchKernelGGL((energy_h_rows2_many<true>), dim3(hb), dim3(256), 0, st, tab, nm);
    else
      hipLaunchKernelGGL((energy_h_rows2_many<false>), dim3(hb), dim3(256), 0, st, tab, nm);
  } else if (h0->oe) {
    if (nt == 1) EHM(true, 1); else if (nt == 2) EHM(true, 2); else if (nt == 4) EHM(true, 4);
    else EHM(true, 8);
  } else {
    if (nt == 1) EHM(false, 1); else if (nt == 2) EHM(false, 2); else if (nt == 4) EHM(false, 4);
    else EHM(false, 8);
  }
#undef EHM
  CK(h0, hipGetLastError());
  hipLaunchKernelGGL(energy_e_chains_many, dim3(eb), dim3(256), 0, st, tab, nm);
  CK(h0, hipGetLastError());
  hipLaunchKernelGGL(energy_planes_many, dim3(pb), dim3(256), 0, st, tab, nm);
  CK(h0, hipGetLastError());
  hipLaunchKernelGGL(energy_combine_many, dim3(nm), dim3(64), 0, st, tab, h0->em_dout);
  CK(h0, hipGetLastError());
  CK(h0, hipMemcpyAsync(h0->em_hout, h0->em_dout, nm * sizeof(double), hipMemcpyDeviceToHost, st));
  CK(h0, hipStreamSynchronize(st));
  for (int i = 0; i < nm; i++) out[act[i]] = h0->em_hout[i] + ops_energy_finish(models[act[i]]);
  return EMS_OK;
}

int ems_gpu_energy_many(ems_gpu *const *models, int n, const uint8_t *active, double *out) {
  if (!models || n < 0 || !out) return EMS_E_INVALID;
  if (energy_many_batchable(models, n, active)) return energy_many_batched(models, n, active, out);
  for (int q = 0; q < n; q++) {
    if (active && !active[q]) continue;
    ems_gpu *h = models[q];
    ALIVE(h);
    int rc = energy_all_enqueue(h);
    if (rc) return rc;
  }
  for (int q = 0; q < n; q++) {
    if (active && !active[q]) continue;
    ems_gpu *h = models[q];
    CK(h, hipEventSynchronize(h->ev_energy));
    out[q] = *h->h_w + ops_energy_finish(h);
  }
  return EMS_OK;
}

}  // extern "C"

namespace {

// Device address of `scratch` per device, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
std::mutex g_scratch_mu;
double *g_scratch_ptr[64] = {};

int noalloc_scratch(ems_gpu *h, double **out) {
  int dev = -1;
  CK(h, hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(h, EMS_E_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(h, EMS_E_DEVICE,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int ems_gpu_energy_noalloc(ems_gpu *h, double *w) {
  ALIVE(h);
  if (!w) return EMS_E_INVALID;
  const size_t need = energy_scratch_doubles(h);
  if (need > (size_t)EMS_ENERGY_NOALLOC_DOUBLES)
    return fail(h, EMS_E_INVALID,
                "energy_noalloc: the grid needs " + std::to_string(need) +
                    " doubles of scratch, the static array holds " +
                    std::to_string((size_t)EMS_ENERGY_NOALLOC_DOUBLES) +
                    " (nx*(3*ny-2) + 13*nx + 1 at most)");
  // Operator energies read back from device state need host buffers.
  for (auto &o : h->ops)
    if (o.type == OP_POLES || o.type == OP_MERGED || (o.type == OP_LUMPED && o.lp.branch))
      return fail(h, EMS_E_UNSUPPORTED,
                  "energy_noalloc: dispersive media, reduced currents and lumped branches "
                  "carry energies read back into host buffers; use ems_gpu_energy");
  double *s = nullptr;
  int rc = noalloc_scratch(h, &s);
  if (rc) return rc;
  const size_t nx = (size_t)h->n[0], rows = nx * (3 * (size_t)h->n[1] - 2);
  double *rowsum = s, *chains = s + rows, *planes = chains + 12 * nx, *d_w = planes + nx;
  if ((rc = energy_kernels(h, rowsum, chains, planes, d_w))) return rc;
  double wf = 0.0;
  CK(h, hipMemcpyAsync(&wf, d_w, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  CK(h, hipStreamSynchronize(h->stream));
  // Engine::operator_energy of the remaining kinds (every one 0, folded
  // from -0.0 as ops_energy_finish does).
  double w_ops = -0.0;
  if (!h->slabs.empty()) w_ops += 0.0;
  for (size_t q = 0; q < h->ops.size(); q++) w_ops += 0.0;
  *w = wf + w_ops;
  return EMS_OK;
}

uint64_t ems_gpu_timestep(const ems_gpu *h) { return h ? h->step : 0; }

int ems_gpu_read_component(ems_gpu *h, int field, int axis, float *out) {
  ALIVE(h);
  if (!out || axis < 0 || axis > 2 || field < 0 || field > 1) return EMS_E_INVALID;
  const Fields &f = h->fb[h->par];
  float *src = field ? f.i[axis] : f.u[axis];
  CK(h, hipMemcpyAsync(out, src, h->N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  CK(h, hipStreamSynchronize(h->stream));
  return EMS_OK;
}

int ems_gpu_write_component(ems_gpu *h, int field, int axis, const float *in) {
  ALIVE(h);
  if (!in || axis < 0 || axis > 2 || field < 0 || field > 1) return EMS_E_INVALID;
  const Fields &f = h->fb[h->par];
  float *dst = field ? f.i[axis] : f.u[axis];
  CK(h, hipMemcpyAsync(dst, in, h->N * sizeof(float), hipMemcpyHostToDevice, h->stream));
  CK(h, hipStreamSynchronize(h->stream));
  return EMS_OK;
}

int ems_gpu_probe_values(ems_gpu *h, int id, double *out, uint64_t cap, uint64_t *count) {
  ALIVE(h);
  if (id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_PROBE)
    return fail(h, EMS_E_INVALID, "not a probe id");
  HostOp &o = h->ops[id];
  CK(h, hipStreamSynchronize(h->stream));
  uint64_t n = o.count < cap ? o.count : cap;
  if (n && out) CK(h, hipMemcpy(out, o.d_vals, n * sizeof(double), hipMemcpyDeviceToHost));
  if (count) *count = o.count;
  return EMS_OK;
}

int ems_gpu_probe_times(ems_gpu *h, int id, double *out, uint64_t cap, uint64_t *count) {
  ALIVE(h);
  if (id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_PROBE)
    return fail(h, EMS_E_INVALID, "not a probe id");
  const HostOp &o = h->ops[id];
  const double offset = o.quantity == 1 ? 0.5 : 1.0;
  const uint64_t n = o.count < cap ? o.count : cap;
  for (uint64_t k = 0; k < n && out; k++)
    out[k] = ((double)(o.first + k * o.interval) + offset) * h->dt;
  if (count) *count = o.count;
  return EMS_OK;
}

double ems_gpu_dt(const ems_gpu *h) { return h ? h->dt : 0.0; }

int ems_gpu_synchronize(ems_gpu *h) {
  ALIVE(h);
  CK(h, hipStreamSynchronize(h->stream));
  return EMS_OK;
}

int ems_gpu_time_kernels(ems_gpu *h, uint64_t steps, double out[3]) {
  ALIVE(h);
  if (!out || steps == 0) return EMS_E_INVALID;
  int rc = prepare(h, steps);
  if (rc) return rc;
  hipEvent_t ev[5];
  for (auto &e : ev) CK(h, hipEventCreate(&e));
  double t0 = 0, t1 = 0, ts = 0;
  for (uint64_t s = 0; s < steps; s++) {
    float a = 0, b = 0, c = 0;
    if (use_fused(h)) {
      CK(h, hipEventRecord(ev[0], h->stream));
      if ((rc = launch_fused(h, h->par))) return rc;
      CK(h, hipEventRecord(ev[1], h->stream));
      h->par ^= 1;
      if ((rc = launch_poles(h, 1, h->fb[h->par], h->fb[h->par ^ 1], true))) return rc;
      if ((rc = launch_mur(h, h->fb[h->par ^ 1], 0))) return rc;
      if ((rc = launch_mur(h, h->fb[h->par], 1))) return rc;
      if ((rc = launch_gap_post_e(h, h->fb[h->par], h->fb[h->par ^ 1], true))) return rc;
      CK(h, hipEventRecord(ev[4], h->stream));
      CK(h, hipEventSynchronize(ev[4]));
      CK(h, hipEventElapsedTime(&a, ev[0], ev[1]));
      CK(h, hipEventElapsedTime(&c, ev[0], ev[4]));
    } else {
      const Fields &f = h->fb[h->par];
      // The bracket starts after the side stream's launches are enqueued,
      // so that it times the H kernel rather than the host's enqueueing.
      if ((rc = merged_fork(h, f))) return rc;
      CK(h, hipEventRecord(ev[0], h->stream));
      if ((rc = launch_h(h, f))) return rc;
      CK(h, hipEventRecord(ev[1], h->stream));
      if ((rc = launch_upml(h, f, 0))) return rc;
      if ((rc = launch_cpml(h, f, 0, h->par))) return rc;
      if ((rc = launch_poles(h, 0, f, f, false))) return rc;
      if ((rc = launch_mur(h, f, 0))) return rc;
      if ((rc = merged_join(h))) return rc;
      if ((rc = launch_gap_mid(h, f))) return rc;
      CK(h, hipEventRecord(ev[2], h->stream));
      if ((rc = launch_e(h, f))) return rc;
      CK(h, hipEventRecord(ev[3], h->stream));
      if ((rc = launch_upml(h, f, 1))) return rc;
      if ((rc = launch_cpml(h, f, 1, h->par))) return rc;
      if ((rc = launch_poles(h, 1, f, f, false))) return rc;
      if ((rc = launch_mur(h, f, 1))) return rc;
      if ((rc = launch_gap_post_e(h, f, f, false))) return rc;
      CK(h, hipEventRecord(ev[4], h->stream));
      CK(h, hipEventSynchronize(ev[4]));
      CK(h, hipEventElapsedTime(&a, ev[0], ev[1]));
      CK(h, hipEventElapsedTime(&b, ev[2], ev[3]));
      CK(h, hipEventElapsedTime(&c, ev[0], ev[4]));
    }
    t0 += a, t1 += b, ts += c;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  merged_mark_stale(h);
  advance_counts(h, steps);
  out[0] = t0 / steps, out[1] = t1 / steps, out[2] = ts / steps;
  return EMS_OK;
}

int ems_gpu_traffic_model(ems_gpu *h, double out[3]) {
  ALIVE(h);
  if (!out) return EMS_E_INVALID;
  // Algorithmic bytes of one launch. Unfused: H reads u (3) and i (3) and
  // writes i (3); E reads i (3) and u (3) and writes u (3). Fused: reads u
  // and i (6) and writes u and i (6) once per node, plus psi (read + write)
  // of every CPML node. Each dense coefficient lane adds 4 B per node
  // (pooled rows stay cache-resident).
  double dense_e = 0, dense_h = 0;
  for (int l = 0; l < 12; l++)
    if (h->lanes[l].rows && !h->lanes[l].index) (l < 6 ? dense_e : dense_h) += 4.0;
  double psi = 0;
  for (auto &s : h->slabs) psi += 8.0 * (double)s.len;
  if (use_fused(h)) {
    out[0] = (48.0 + dense_e + dense_h) * (double)h->N + psi;
    out[1] = 0.0;
  } else {
    out[0] = (36.0 + dense_h) * (double)h->N;
    out[1] = (36.0 + dense_e) * (double)h->N;
  }
  // CPML slab launches of an unfused step: target r+w, psi r+w, two source
  // values and the coefficient, 4 B each; slabs corrected inside a side's
  // update add their psi (read + write) to it instead.
  double cp = 0;
  const bool fold[2] = {cpml_folded(h, 0), cpml_folded(h, 1)};
  for (auto &s : h->slabs) {
    if (!use_fused(h) && fold[s.arg.e_side ? 1 : 0] && !s.out)
      out[s.arg.e_side ? 1 : 0] += 8.0 * (double)s.len;
    else
      cp += 28.0 * (double)s.len;
  }
  out[2] = use_fused(h) ? 0.0 : cp;
  return EMS_OK;
}

int ems_gpu_copy_bandwidth(ems_gpu *h, uint64_t bytes, int reps, double *gbs) {
  ALIVE(h);
  if (!gbs || reps < 1 || bytes < 4096) return EMS_E_INVALID;
  const size_t n4 = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  CK(h, hipMalloc(&a, n4 * 16));
  CK(h, hipMalloc(&b, n4 * 16));
  CK(h, hipMemsetAsync(a, 0, n4 * 16, h->stream));
  hipEvent_t e0, e1;
  CK(h, hipEventCreate(&e0));
  CK(h, hipEventCreate(&e1));
  const unsigned blocks = (unsigned)std::min<size_t>((n4 + 255) / 256, 1u << 20);
  hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, h->stream, a, b, n4);
  CK(h, hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, h->stream, a, b, n4);
  CK(h, hipEventRecord(e1, h->stream));
  CK(h, hipEventSynchronize(e1));
  float ms = 0;
  CK(h, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
  CK(h, hipFree(a));
  CK(h, hipFree(b));
  *gbs = 2.0 * (double)(n4 * 16) * reps / (ms * 1e-3) / 1e9;
  return EMS_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batches --
//
// Many engines stepped together: one fused launch over the tiles of every
// active model and one launch for all their post-E gaps per step, captured
// in graphs. Each model computes exactly what its own engine computes (the
// same tile kernel on its own arrays), so results are bit-identical to
// stepping the models one by one. No reference counterpart: the reference
// runs one model per engine.

struct ems_gpu_batch {
  std::vector<ems_gpu *> m;
  hipStream_t stream = nullptr;
  int tj = 7, nchunk = 0, oe = 0;
  bool prefetch = true, packed = false;
  std::vector<int> active;        // the set the graphs were captured for
  FusedArgs *d_args[2] = {nullptr, nullptr};
  GapArgs *d_gap[2] = {nullptr, nullptr};
  uint32_t *d_seg = nullptr;
  size_t cap = 0;                 // models the device arrays hold
  int nblocks = 0;
  hipGraphExec_t g1[2] = {nullptr, nullptr}, gK[2] = {nullptr, nullptr};
  int K = 16;                     // even
  bool graphs = true;             // ems_gpu_batch_set_graphs
  std::string err;
};

namespace {

int bfail(ems_gpu_batch *b, int code, const std::string &msg) {
  b->err = msg;
  return code;
}

#define BCK(b, x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return bfail((b), EMS_E_DEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

void batch_drop_graphs(ems_gpu_batch *b) {
  for (int p = 0; p < 2; p++) {
    if (b->g1[p]) (void)hipGraphExecDestroy(b->g1[p]);
    if (b->gK[p]) (void)hipGraphExecDestroy(b->gK[p]);
    b->g1[p] = b->gK[p] = nullptr;
  }
  b->active.clear();
}

// Writes the per-model launch arguments of `act` for both parities.
int batch_args(ems_gpu_batch *b, const std::vector<int> &act) {
  const size_t n = act.size();
  if (n > b->cap) {
    for (int p = 0; p < 2; p++) {
      if (b->d_args[p]) BCK(b, hipFree(b->d_args[p]));
      if (b->d_gap[p]) BCK(b, hipFree(b->d_gap[p]));
      BCK(b, hipMalloc(&b->d_args[p], n * sizeof(FusedArgs)));
      BCK(b, hipMalloc(&b->d_gap[p], n * sizeof(GapArgs)));
    }
    if (b->d_seg) BCK(b, hipFree(b->d_seg));
    BCK(b, hipMalloc(&b->d_seg, (n + 1) * sizeof(uint32_t)));
    b->cap = n;
  }
  std::vector<uint32_t> seg(n + 1, 0);
  for (int p = 0; p < 2; p++) {
    std::vector<FusedArgs> fa(n);
    std::vector<GapArgs> ga(n);
    for (size_t q = 0; q < n; q++) {
      ems_gpu *h = b->m[act[q]];
      const FusedGeom g = b->packed ? packed_geom(h, b->nchunk) : fused_geom(h);
      FusedArgs &a = fa[q];
      a.A = h->fbuf[p], a.B = h->fbuf[1 - p], a.N = h->N;
      a.cpool = h->cpool, a.croff = (const uint32_t *)h->croff, a.slabs = h->d_fslabs,
      a.pmask = h->d_pmask;
      a.nslab = n_fslabs(h), a.par = p, a.nx = h->n[0], a.ny = h->n[1], a.nz = h->n[2];
      a.tk = g.tk, a.chunk = g.chunk, a.ktiles = g.ktiles, a.jtiles = g.jtiles, a.ntiles = g.ntiles;
      if (p == 0) seg[q + 1] = seg[q] + (uint32_t)g.nblocks;
      GapArgs &gg = ga[q];
      gg.f = h->fb[1 - p], gg.fp = h->fb[p], gg.ops = h->d_ops, gg.nops = (int)h->ops.size();
      gg.step = h->d_step;
    }
    BCK(b, hipMemcpy(b->d_args[p], fa.data(), n * sizeof(FusedArgs), hipMemcpyHostToDevice));
    BCK(b, hipMemcpy(b->d_gap[p], ga.data(), n * sizeof(GapArgs), hipMemcpyHostToDevice));
  }
  BCK(b, hipMemcpy(b->d_seg, seg.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
  b->nblocks = (int)seg[n];
  return 0;
}

template <bool OE, bool PF>
void batch_launch_packed(ems_gpu_batch *b, int par, int nm) {
  hipLaunchKernelGGL((step_packed_batch<OE, PACK_NW, PF>), dim3(b->nblocks), dim3(64, PACK_NW), 0,
                     b->stream, (const FusedArgs *)b->d_args[par], (const uint32_t *)b->d_seg, nm,
                     2);
}

template <bool OE, int TJ, bool PF>
void batch_launch_tj(ems_gpu_batch *b, int par, int nm) {
  hipLaunchKernelGGL((step_fused_batch<OE, TJ, PF>), dim3(b->nblocks), dim3(64, TJ + 1), 0,
                     b->stream, (const FusedArgs *)b->d_args[par], (const uint32_t *)b->d_seg, nm,
                     2);
}

// One batched step from parity `par` (every active model holds its state
// in fbuf[par]).
int batch_enqueue_step(ems_gpu_batch *b, const std::vector<int> &act, int par) {
  const int nm = (int)act.size();
  const int key = b->packed ? (b->prefetch ? 1 : 0) : b->tj * 2 + (b->prefetch ? 1 : 0);
#define BL(OE)                                                               \
  switch (key) {                                                             \
    case 0: batch_launch_packed<OE, false>(b, par, nm); break;              \
    case 1: batch_launch_packed<OE, true>(b, par, nm); break;               \
    case 6: batch_launch_tj<OE, 3, false>(b, par, nm); break;               \
    case 7: batch_launch_tj<OE, 3, true>(b, par, nm); break;                \
    case 14: batch_launch_tj<OE, 7, false>(b, par, nm); break;              \
    case 15: batch_launch_tj<OE, 7, true>(b, par, nm); break;               \
    case 30: batch_launch_tj<OE, 15, false>(b, par, nm); break;             \
    case 31: batch_launch_tj<OE, 15, true>(b, par, nm); break;              \
    default: return bfail(b, EMS_E_INVALID, "unsupported tile height");      \
  }
  if (b->oe) {
    BL(true)
  } else {
    BL(false)
  }
#undef BL
  BCK(b, hipGetLastError());
  // Dispersive media of each model, on the batch stream.
  for (int q : act) {
    ems_gpu *h = b->m[q];
    const hipStream_t own = h->stream;
    h->stream = b->stream;
    const int rc = launch_poles(h, 1, h->fb[par ^ 1], h->fb[par], true);
    h->stream = own;
    if (rc) return bfail(b, rc, h->err);
  }
  hipLaunchKernelGGL(gap_post_e_batch, dim3(nm), dim3(256), 0, b->stream,
                     (const GapArgs *)b->d_gap[par]);
  BCK(b, hipGetLastError());
  return 0;
}

int batch_capture(ems_gpu_batch *b, const std::vector<int> &act, int steps, int par,
                  hipGraphExec_t *out) {
  hipGraph_t g;
  BCK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < steps && !rc; s++, par ^= 1) rc = batch_enqueue_step(b, act, par);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) return rc;
  BCK(b, e);
  BCK(b, hipGraphInstantiate(out, g, nullptr, nullptr, 0));
  BCK(b, hipGraphDestroy(g));
  return 0;
}

}  // namespace

extern "C" {

int ems_gpu_batch_create(ems_gpu *const *models, int n, ems_gpu_batch **out) {
  if (!models || n < 1 || !out) return EMS_E_INVALID;
  *out = nullptr;
  ems_gpu_batch *b = new ems_gpu_batch();
  for (int q = 0; q < n; q++) {
    // (a handle listed twice would be stepped twice per launch, its CPML
    // psi updated in place by two blocks)
    bool dup = false;
    for (int r = 0; r < q && !dup; r++) dup = models[r] == models[q];
    if (!models[q] || models[q]->dead || dup) {
      delete b;
      return EMS_E_INVALID;
    }
    b->m.push_back(models[q]);
  }
  b->oe = models[0]->oe;
  for (auto *h : b->m)
    if (h->oe != b->oe) {
      delete b;
      return EMS_E_UNSUPPORTED;
    }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return EMS_E_DEVICE;
  }
  *out = b;
  return EMS_OK;
}

void ems_gpu_batch_destroy(ems_gpu_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graphs(b);
  for (int p = 0; p < 2; p++) {
    if (b->d_args[p]) (void)hipFree(b->d_args[p]);
    if (b->d_gap[p]) (void)hipFree(b->d_gap[p]);
  }
  if (b->d_seg) (void)hipFree(b->d_seg);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char *ems_gpu_batch_last_error(const ems_gpu_batch *b) {
  return b ? b->err.c_str() : "null batch";
}

int ems_gpu_batch_set_graphs(ems_gpu_batch *b, int on) {
  if (!b) return EMS_E_INVALID;
  b->graphs = on != 0;
  return EMS_OK;
}

int ems_gpu_batch_set_schedule(ems_gpu_batch *b, int mode, int tile_rows, int x_chunks) {
  if (!b) return EMS_E_INVALID;
  if (mode != 1 && mode != 2 && mode != 4 && mode != 5)
    return bfail(b, EMS_E_INVALID, "batch schedule mode must be 1, 2, 4 or 5");
  if (mode <= 2 && tile_rows != 3 && tile_rows != 7 && tile_rows != 15)
    return bfail(b, EMS_E_INVALID, "tile_rows must be 3, 7 or 15");
  if (x_chunks < 0) return bfail(b, EMS_E_INVALID, "x_chunks must be >= 0");
  if (mode >= 4)
    for (auto *h : b->m)
      if (!packed_ok(h)) return bfail(b, EMS_E_INVALID, "packed rows need nz <= 64 in every model");
  b->packed = mode >= 4, b->prefetch = mode == 1 || mode == 4;
  if (mode <= 2) b->tj = tile_rows;
  b->nchunk = x_chunks;
  batch_drop_graphs(b);
  return EMS_OK;
}

int ems_gpu_batch_run(ems_gpu_batch *b, const uint8_t *active, uint64_t n) {
  if (!b) return EMS_E_INVALID;
  std::vector<int> act;
  for (size_t q = 0; q < b->m.size(); q++)
    if (!active || active[q]) act.push_back((int)q);
  if (act.empty() || n == 0) return EMS_OK;
  const int par = b->m[act[0]]->par;
  for (int q : act) {
    ems_gpu *h = b->m[q];
    if (h->dead) return bfail(b, EMS_E_DEVICE, "model " + std::to_string(q) + ": " + h->err);
    h->fused = true, h->rows = false, h->band = false, h->prefetch = b->prefetch, h->tj = b->tj;
    h->packed = b->packed;
    h->nchunk = b->nchunk;
    for (auto &o : h->ops)
      if (o.type == OP_MUR)
        return bfail(b, EMS_E_UNSUPPORTED,
                     "model " + std::to_string(q) + " has a Mur boundary (not batched)");
    if (!use_fused(h))
      return bfail(b, EMS_E_UNSUPPORTED,
                   "model " + std::to_string(q) +
                       " needs the unfused schedule (H source, magnetic poles, plane wave, "
                       "cut-face currents, openEMS PML layers or more than 32 CPML slabs)");
    if (h->par != par) return bfail(b, EMS_E_INVALID, "models of a batch step in lockstep");
    const int rc = prepare(h, n, false);
    if (rc) return bfail(b, rc, h->err);
    if (h->croff_wide)
      return bfail(b, EMS_E_UNSUPPORTED,
                   "model " + std::to_string(q) + " has a coefficient pool of 4 GiB or more");
    BCK(b, hipStreamSynchronize(h->stream));
  }
  int rc = batch_args(b, act);
  if (rc) return rc;
  if (!b->graphs) {  // the same launches, queued one by one
    int p = par;
    for (uint64_t s = 0; s < n; s++, p ^= 1)
      if ((rc = batch_enqueue_step(b, act, p))) return rc;
    BCK(b, hipStreamSynchronize(b->stream));
    for (int q : act) {
      ems_gpu *h = b->m[q];
      h->par = p;
      advance_counts(h, n);
    }
    return EMS_OK;
  }
  if (act != b->active) {
    batch_drop_graphs(b);
    for (int p = 0; p < 2; p++) {
      if ((rc = batch_capture(b, act, 1, p, &b->g1[p]))) return rc;
      if ((rc = batch_capture(b, act, b->K, p, &b->gK[p]))) return rc;
    }
    b->active = act;
  }
  int p = par;
  uint64_t left = n;
  while (left >= (uint64_t)b->K) {
    BCK(b, hipGraphLaunch(b->gK[p], b->stream));
    left -= b->K;
  }
  while (left > 0) {
    BCK(b, hipGraphLaunch(b->g1[p], b->stream));
    p ^= 1;
    left--;
  }
  BCK(b, hipStreamSynchronize(b->stream));
  for (int q : act) {
    ems_gpu *h = b->m[q];
    h->par = p;
    advance_counts(h, n);
  }
  return EMS_OK;
}

}  // extern "C"
