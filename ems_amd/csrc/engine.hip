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

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ems_gpu.h"

namespace {

// ---------------------------------------------------------------- device --

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
  const float *__restrict__ ux = f.u[0];
  const float *__restrict__ uy = f.u[1];
  const float *__restrict__ uz = f.u[2];
  float *__restrict__ ix = f.i[0];
  float *__restrict__ iy = f.i[1];
  float *__restrict__ iz = f.i[2];
  size_t o = (size_t)i0 * plane + p;
  float uyc = uy[o], uzc = uz[o];
  for (int i = i0; i < i1; i++, o += plane) {
    const bool iv = i + 1 < nx;
    const size_t r = (size_t)i * ny + j;
    const float uxc = ux[o];
    const float uyn = iv ? uy[o + plane] : 0.0f;
    const float uzn = iv ? uz[o + plane] : 0.0f;
    if (jv && kv) {  // curl_x = (uz[j+1] - uz[j]) - (uy[k+1] - uy[k])
      float da = coef(L.l[6], r, k, nz, o), db = coef(L.l[9], r, k, nz, o);
      ix[o] = h_expr<OE>(da, ix[o], db, uz[o + nz], uzc, uy[o + 1], uyc);
    }
    if (iv && kv) {  // curl_y = (ux[k+1] - ux[k]) - (uz[i+1] - uz[i])
      float da = coef(L.l[7], r, k, nz, o), db = coef(L.l[10], r, k, nz, o);
      iy[o] = h_expr<OE>(da, iy[o], db, ux[o + 1], uxc, uzn, uzc);
    }
    if (iv && jv) {  // curl_z = (uy[i+1] - uy[i]) - (ux[j+1] - ux[j])
      float da = coef(L.l[8], r, k, nz, o), db = coef(L.l[11], r, k, nz, o);
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
  const float *__restrict__ ix = f.i[0];
  const float *__restrict__ iy = f.i[1];
  const float *__restrict__ iz = f.i[2];
  float *__restrict__ ux = f.u[0];
  float *__restrict__ uy = f.u[1];
  float *__restrict__ uz = f.u[2];
  size_t o = (size_t)i0 * plane + p;
  float iyp = i0 > 0 ? iy[o - plane] : 0.0f;
  float izp = i0 > 0 ? iz[o - plane] : 0.0f;
  for (int i = i0; i < i1; i++, o += plane) {
    const size_t r = (size_t)i * ny + j;
    const float ixc = ix[o], iyc = iy[o], izc = iz[o];
    if (i + 1 < nx) {  // curl_x = (iz[j] - iz[j-1]) - (iy[k] - iy[k-1])
      float ca = coef(L.l[0], r, k, nz, o), cb = coef(L.l[3], r, k, nz, o);
      float b = j > 0 ? iz[o - nz] : 0.0f;
      float d = k > 0 ? iy[o - 1] : 0.0f;
      ux[o] = e_expr<OE>(ca, ux[o], cb, izc, b, iyc, d);
    }
    if (j + 1 < ny) {  // curl_y = (ix[k] - ix[k-1]) - (iz[i] - iz[i-1])
      float ca = coef(L.l[1], r, k, nz, o), cb = coef(L.l[4], r, k, nz, o);
      float b = k > 0 ? ix[o - 1] : 0.0f;
      uy[o] = e_expr<OE>(ca, uy[o], cb, ixc, b, izc, izp);
    }
    if (k + 1 < nz) {  // curl_z = (iy[i] - iy[i-1]) - (ix[j] - ix[j-1])
      float ca = coef(L.l[2], r, k, nz, o), cb = coef(L.l[5], r, k, nz, o);
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
  const float *cur = es ? f.i[s.other] : f.u[s.other];
  float *t = es ? f.u[s.comp] : f.i[s.comp];
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
  const float cf = coef(cfl, (size_t)i * ny + j, k, nz, o);
  const float scale = es ? 1.0f : -1.0f;
  const float sv = s.sign * (h - l);
  const float ps = flush(s.b[z] * s.psi[q] + s.a[z] * sv);
  s.psi[q] = ps;
  t[o] = flush(t[o] + scale * cf * (s.k[z] * sv + ps));
  (void)nx;
}

// Serial operators (excite.rs, probe.rs) and the gap rules of
// Engine::advance (mod.rs:873-915).
enum { OP_SOURCE = 1, OP_PROBE = 2 };
struct DevOp {
  int type, kind, quantity, n;
  uint64_t interval;
  const int32_t *axis;
  const uint64_t *idx;
  const float *wf;
  const double *wd;
  const float *amp;
  uint64_t namp;
  double *vals;
  uint64_t *count;
};

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
__device__ __forceinline__ bool needs(const DevOp &o, uint64_t n) {
  if (o.type == OP_SOURCE) {
    if (o.kind == 1 || o.kind == 3) return true;
    return op_amp(o, n) != 0.0f;
  }
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
  for (int q = 0; q < o.n; q++) v += o.wd[q] * (double)fld[o.axis[q]][o.idx[q]];
  uint64_t c = *o.count;
  o.vals[c] = v;
  *o.count = c + 1;
}

struct Gaps {
  bool gap_b, moved, gap_a;
};
__device__ Gaps gaps_at(const DevOp *ops, int nops, uint64_t n) {
  bool need_post_e = false, mid = false, movable = true;
  for (int q = 0; q < nops; q++) {
    const DevOp &o = ops[q];
    if (!needs(o, n)) continue;
    if (is_esrc(o) || is_vprobe(o)) need_post_e = true;
    if (is_hsrc(o)) mid = true, movable = false;
    if (is_iprobe(o)) mid = true;
  }
  Gaps g;
  g.gap_b = mid && !movable;
  g.moved = mid && movable;
  g.gap_a = need_post_e || g.moved;
  return g;
}

__global__ void gap_mid(Fields f, const DevOp *ops, int nops, const uint64_t *step) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t n = *step;
  Gaps g = gaps_at(ops, nops, n);
  if (!g.gap_b) return;
  for (int q = 0; q < nops; q++) {
    const DevOp &o = ops[q];
    if (is_hsrc(o)) source_apply(f, o, n);
    if (is_iprobe(o) && n % o.interval == 0) probe_sample(f, o);
  }
}

__global__ void gap_post_e(Fields f, const DevOp *ops, int nops, uint64_t *step) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t n = *step;
  Gaps g = gaps_at(ops, nops, n);
  if (g.gap_a) {
    if (g.moved)
      for (int q = 0; q < nops; q++) {
        const DevOp &o = ops[q];
        if (is_iprobe(o) && n % o.interval == 0) probe_sample(f, o);
      }
    for (int q = 0; q < nops; q++) {
      const DevOp &o = ops[q];
      if (is_esrc(o)) source_apply(f, o, n);
      if (is_vprobe(o) && n % o.interval == 0) probe_sample(f, o);
    }
  }
  *step = n + 1;
}

// ---- energy invariant (kernel.rs:119-244, 364-423; pool.rs:673-692) ----

// One H row's share, read-only (COMMIT = false): 4 threads per row, lane l
// sums entries k = l mod 4 in order; lane 0 returns (a0 + a1) + (a2 + a3).
// Rows are enumerated per plane as x rows (ny-1), y rows (ny), z rows (ny-1).
template <bool OE>
__global__ void __launch_bounds__(256) energy_h_rows(Fields f, Lanes L, int nx, int ny, int nz,
                                                     double *rowsum) {
  const int rows_per_plane = 3 * ny - 2;
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = (int)(tid & 3);
  const size_t rid = tid >> 2;
  __shared__ double part[256];
  double acc = 0.0;
  const size_t total = (size_t)nx * rows_per_plane;
  const bool live = rid < total;
  if (live) {
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
    const bool has_next = gi + 1 < nx;
    if (comp == 0 || has_next) {
      const int n = comp == 2 ? nz : nz - 1;
      const float *ux = f.u[0], *uy = f.u[1], *uz = f.u[2];
      const float *h = f.i[comp];
      for (int k = lane; k < n; k += 4) {
        const size_t o = base + k;
        float a, b, c, d;
        if (comp == 0) {
          a = uz[o + nz], b = uz[o], c = uy[o + 1], d = uy[o];
        } else if (comp == 1) {
          a = ux[o + 1], b = ux[o], c = uz[o + plane], d = uz[o];
        } else {
          a = uy[o + plane], b = uy[o], c = ux[o + nz], d = ux[o];
        }
        const float da = coef(L.l[6 + comp], r, k, nz, o);
        const float db = coef(L.l[9 + comp], r, k, nz, o);
        const float old = h[o];
        const float nw = h_expr<OE>(da, old, db, a, b, c, d);
        acc += weight(da, (double)old * (double)nw, db);
      }
    }
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  if (live && lane == 0) {
    const double *pp = part + threadIdx.x;
    rowsum[rid] = (pp[0] + pp[1]) + (pp[2] + pp[3]);
  }
}

// E part: per (plane, component, lane) the chain over in-plane indices
// q = lane mod 4 in increasing order (e_energy_row, kernel.rs:218-244).
__global__ void __launch_bounds__(64) energy_e_chains(Fields f, Lanes L, int nx, int ny, int nz,
                                                      double *chains) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)nx * 3 * 4;
  if (tid >= total) return;
  const int lane = (int)(tid & 3);
  const int a = (int)((tid >> 2) % 3);
  const int gi = (int)((tid >> 2) / 3);
  const size_t plane = (size_t)ny * nz;
  const float *u = f.u[a] + (size_t)gi * plane;
  double acc = 0.0;
  for (size_t q = lane; q < plane; q += 4) {
    const int j = (int)(q / nz), k = (int)(q - (size_t)j * nz);
    const size_t r = (size_t)gi * ny + j;
    const size_t o = (size_t)gi * plane + q;
    const float ca = coef(L.l[a], r, k, nz, o), cb = coef(L.l[3 + a], r, k, nz, o);
    const double x = (double)u[q];
    acc += weight(ca, x * x, cb);
  }
  chains[tid] = acc;
}

// Plane values we + wh (h_slab, kernel.rs:404-420) and their sum in plane
// order (pool.rs:675-692); one block.
__global__ void __launch_bounds__(256) energy_combine(int nx, int ny, const double *rowsum,
                                                      const double *chains, double *planes,
                                                      double dt, double *out) {
  const int rows_per_plane = 3 * ny - 2;
  for (int gi = threadIdx.x; gi < nx; gi += blockDim.x) {
    const double *rs = rowsum + (size_t)gi * rows_per_plane;
    double wh = 0.0;
    const int nrows = gi + 1 < nx ? rows_per_plane : ny - 1;
    for (int q = 0; q < nrows; q++) wh += rs[q];
    double we = 0.0;
    for (int a = 0; a < 3; a++) {
      const double *c = chains + ((size_t)gi * 3 + a) * 4;
      we += (c[0] + c[1]) + (c[2] + c[3]);
    }
    planes[gi] = we + wh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = -0.0;
    for (int gi = 0; gi < nx; gi++) s += planes[gi];
    *out = 0.25 * dt * s;
  }
}

// ------------------------------------------------------------------ host --

struct HostSlab {
  SlabArg arg;
  float *d_b = nullptr, *d_a = nullptr, *d_k = nullptr, *d_psi = nullptr;
  size_t len = 0;
};

struct HostOp {
  int type = 0, kind = 0, quantity = 0, n = 0;
  uint64_t interval = 1;
  int32_t *d_axis = nullptr;
  uint64_t *d_idx = nullptr;
  float *d_wf = nullptr;
  double *d_wd = nullptr;
  float *d_amp = nullptr;
  uint64_t namp = 0;
  std::vector<float> amp_host;
  double *d_vals = nullptr;
  uint64_t cap = 0;
  uint64_t count = 0;  // samples recorded (host mirror of *d_count)
  uint64_t first = 0;  // step of sample 0
  uint64_t *d_count = nullptr;
};

}  // namespace

struct ems_gpu {
  int n[3] = {0, 0, 0};
  size_t N = 0, plane = 0;
  double dt = 0;
  int oe = 0;
  hipStream_t stream = nullptr;
  Fields f{};
  Lane lanes[12]{};
  std::vector<void *> lane_mem;
  std::vector<HostSlab> slabs;
  std::vector<HostOp> ops;
  DevOp *d_ops = nullptr;
  bool ops_dirty = true;
  uint64_t step = 0;
  uint64_t *d_step = nullptr;
  // energy scratch
  double *d_rowsum = nullptr, *d_chains = nullptr, *d_planes = nullptr, *d_w = nullptr;
  // graphs of whole steps
  hipGraphExec_t g1 = nullptr, gK = nullptr;
  int K = 32;
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

int x_chunk(const ems_gpu *h, int *nchunks) {
  // Enough blocks to fill 256 CUs several times over; long chunks otherwise
  // (every chunk re-reads one neighbour plane).
  const int bpp = (int)((h->plane + 255) / 256);
  const int want = 4096;
  int nch = (want + bpp - 1) / bpp;
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

int launch_h(ems_gpu *h) {
  int nch, chunk = x_chunk(h, &nch);
  dim3 grid((unsigned)((h->plane + 255) / 256), (unsigned)nch);
  if (h->oe)
    hipLaunchKernelGGL(h_update<true>, grid, dim3(256), 0, h->stream, h->f, lanes_arg(h), h->n[0],
                       h->n[1], h->n[2], chunk);
  else
    hipLaunchKernelGGL(h_update<false>, grid, dim3(256), 0, h->stream, h->f, lanes_arg(h),
                       h->n[0], h->n[1], h->n[2], chunk);
  CK(h, hipGetLastError());
  return 0;
}

int launch_e(ems_gpu *h) {
  int nch, chunk = x_chunk(h, &nch);
  dim3 grid((unsigned)((h->plane + 255) / 256), (unsigned)nch);
  if (h->oe)
    hipLaunchKernelGGL(e_update<true>, grid, dim3(256), 0, h->stream, h->f, lanes_arg(h), h->n[0],
                       h->n[1], h->n[2], chunk);
  else
    hipLaunchKernelGGL(e_update<false>, grid, dim3(256), 0, h->stream, h->f, lanes_arg(h),
                       h->n[0], h->n[1], h->n[2], chunk);
  CK(h, hipGetLastError());
  return 0;
}

int launch_cpml(ems_gpu *h, int e_side) {
  for (auto &s : h->slabs) {
    if (s.arg.e_side != e_side || s.len == 0) continue;
    const Lane cf = h->lanes[e_side ? 3 + s.arg.comp : 9 + s.arg.comp];
    unsigned blocks = (unsigned)((s.len + 255) / 256);
    hipLaunchKernelGGL(cpml_slab, dim3(blocks), dim3(256), 0, h->stream, h->f, cf, s.arg, h->n[0],
                       h->n[1], h->n[2]);
    CK(h, hipGetLastError());
  }
  return 0;
}

bool has_mid_ops(const ems_gpu *h) {
  for (auto &o : h->ops)
    if ((o.type == OP_SOURCE && (o.kind == 2 || o.kind == 3)) ||
        (o.type == OP_PROBE && o.quantity == 1))
      return true;
  return false;
}

int enqueue_step(ems_gpu *h) {
  int rc;
  if ((rc = launch_h(h))) return rc;
  if ((rc = launch_cpml(h, 0))) return rc;
  if (has_mid_ops(h)) {
    hipLaunchKernelGGL(gap_mid, dim3(1), dim3(64), 0, h->stream, h->f, h->d_ops,
                       (int)h->ops.size(), (const uint64_t *)h->d_step);
    CK(h, hipGetLastError());
  }
  if ((rc = launch_e(h))) return rc;
  if ((rc = launch_cpml(h, 1))) return rc;
  hipLaunchKernelGGL(gap_post_e, dim3(1), dim3(64), 0, h->stream, h->f, h->d_ops,
                     (int)h->ops.size(), h->d_step);
  CK(h, hipGetLastError());
  return 0;
}

void drop_graphs(ems_gpu *h) {
  if (h->g1) hipGraphExecDestroy(h->g1);
  if (h->gK) hipGraphExecDestroy(h->gK);
  h->g1 = h->gK = nullptr;
}

int capture(ems_gpu *h, int steps, hipGraphExec_t *out) {
  hipGraph_t g;
  CK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < steps && !rc; s++) rc = enqueue_step(h);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) return rc;
  CK(h, e);
  CK(h, hipGraphInstantiate(out, g, nullptr, nullptr, 0));
  CK(h, hipGraphDestroy(g));
  return 0;
}

// Uploads the operator table, and grows probe buffers to hold `steps` more
// steps of samples.
int prepare(ems_gpu *h, uint64_t steps) {
  bool grown = false;
  for (auto &o : h->ops) {
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
  if (h->ops_dirty) {
    drop_graphs(h);
    std::vector<DevOp> dv(h->ops.size());
    for (size_t q = 0; q < h->ops.size(); q++) {
      const HostOp &o = h->ops[q];
      DevOp &d = dv[q];
      d.type = o.type, d.kind = o.kind, d.quantity = o.quantity, d.n = o.n;
      d.interval = o.interval;
      d.axis = o.d_axis, d.idx = o.d_idx, d.wf = o.d_wf, d.wd = o.d_wd;
      d.amp = o.d_amp, d.namp = o.namp, d.vals = o.d_vals, d.count = o.d_count;
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
    if (o.type != OP_PROBE) continue;
    uint64_t a = h->step, b = h->step + steps;
    o.count += (b + o.interval - 1) / o.interval - (a + o.interval - 1) / o.interval;
  }
  h->step += steps;
}

int energy_now(ems_gpu *h, double *w) {
  const int nx = h->n[0], ny = h->n[1], nz = h->n[2];
  const size_t rows = (size_t)nx * (3 * ny - 2);
  if (!h->d_rowsum) {
    CK(h, hipMalloc(&h->d_rowsum, rows * sizeof(double)));
    CK(h, hipMalloc(&h->d_chains, (size_t)nx * 12 * sizeof(double)));
    CK(h, hipMalloc(&h->d_planes, (size_t)nx * sizeof(double)));
    CK(h, hipMalloc(&h->d_w, sizeof(double)));
  }
  const size_t threads = rows * 4;
  dim3 g1((unsigned)((threads + 255) / 256));
  if (h->oe)
    hipLaunchKernelGGL(energy_h_rows<true>, g1, dim3(256), 0, h->stream, h->f, lanes_arg(h), nx, ny,
                       nz, h->d_rowsum);
  else
    hipLaunchKernelGGL(energy_h_rows<false>, g1, dim3(256), 0, h->stream, h->f, lanes_arg(h), nx,
                       ny, nz, h->d_rowsum);
  CK(h, hipGetLastError());
  dim3 g2((unsigned)(((size_t)nx * 12 + 63) / 64));
  hipLaunchKernelGGL(energy_e_chains, g2, dim3(64), 0, h->stream, h->f, lanes_arg(h), nx, ny, nz,
                     h->d_chains);
  CK(h, hipGetLastError());
  hipLaunchKernelGGL(energy_combine, dim3(1), dim3(256), 0, h->stream, nx, ny, h->d_rowsum,
                     h->d_chains, h->d_planes, h->dt, h->d_w);
  CK(h, hipGetLastError());
  double v = 0;
  CK(h, hipMemcpyAsync(&v, h->d_w, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  CK(h, hipStreamSynchronize(h->stream));
  *w = v;
  return 0;
}

int run_steps(ems_gpu *h, uint64_t n) {
  if (n == 0) return 0;
  int rc = prepare(h, n);
  if (rc) return rc;
  if (!h->g1) {
    if ((rc = capture(h, 1, &h->g1))) return rc;
    if ((rc = capture(h, h->K, &h->gK))) return rc;
  }
  uint64_t left = n;
  while (left >= (uint64_t)h->K) {
    CK(h, hipGraphLaunch(h->gK, h->stream));
    left -= h->K;
  }
  while (left > 0) {
    CK(h, hipGraphLaunch(h->g1, h->stream));
    left--;
  }
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
  int rc = 0;
  auto init = [&]() -> int {
    CK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (int a = 0; a < 3; a++) {
      CK(h, hipMalloc(&h->f.u[a], N * sizeof(float)));
      CK(h, hipMalloc(&h->f.i[a], N * sizeof(float)));
      CK(h, hipMemset(h->f.u[a], 0, N * sizeof(float)));
      CK(h, hipMemset(h->f.i[a], 0, N * sizeof(float)));
    }
    CK(h, hipMalloc(&h->d_step, sizeof(uint64_t)));
    CK(h, hipMemset(h->d_step, 0, sizeof(uint64_t)));
    return 0;
  };
  rc = init();
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
  if (h->stream) hipStreamSynchronize(h->stream);
  drop_graphs(h);
  for (int a = 0; a < 3; a++) {
    if (h->f.u[a]) hipFree(h->f.u[a]);
    if (h->f.i[a]) hipFree(h->f.i[a]);
  }
  for (void *p : h->lane_mem) hipFree(p);
  for (auto &s : h->slabs) {
    hipFree(s.d_b), hipFree(s.d_a), hipFree(s.d_k), hipFree(s.d_psi);
  }
  for (auto &o : h->ops) {
    hipFree(o.d_axis), hipFree(o.d_idx), hipFree(o.d_wf), hipFree(o.d_wd), hipFree(o.d_amp);
    hipFree(o.d_vals), hipFree(o.d_count);
  }
  if (h->d_ops) hipFree(h->d_ops);
  if (h->d_step) hipFree(h->d_step);
  if (h->d_rowsum) hipFree(h->d_rowsum), hipFree(h->d_chains), hipFree(h->d_planes),
      hipFree(h->d_w);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

const char *ems_gpu_last_error(const ems_gpu *h) { return h ? h->err.c_str() : "null handle"; }

int ems_gpu_set_lane(ems_gpu *h, int lane, const float *rows, uint64_t nrows,
                     const uint32_t *row_index) {
  if (!h) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  if (lane < 0 || lane >= 12 || !rows) return fail(h, EMS_E_INVALID, "lane out of range");
  const size_t nr = (size_t)h->n[0] * h->n[1];
  if (!row_index && nrows != nr)
    return fail(h, EMS_E_INVALID, "dense lane needs nx*ny rows");
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
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_add_cpml(ems_gpu *h, const ems_cpml_slab *s, int n) {
  if (!h || (n > 0 && !s)) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  for (int q = 0; q < n; q++) {
    const ems_cpml_slab &in = s[q];
    if (in.comp < 0 || in.comp > 2 || in.normal < 0 || in.normal > 2 || in.comp == in.normal)
      return fail(h, EMS_E_INVALID, "CPML slab: bad component/normal");
    for (int d = 0; d < 3; d++)
      if (in.lo[d] < 0 || in.hi[d] > h->n[d] || in.lo[d] > in.hi[d])
        return fail(h, EMS_E_INVALID, "CPML slab: box out of range");
    // The derivative reads one node past the box on the H side.
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
    CK(h, hipMalloc(&hs.d_psi, (hs.len ? hs.len : 1) * sizeof(float)));
    CK(h, hipMemset(hs.d_psi, 0, (hs.len ? hs.len : 1) * sizeof(float)));
    a.b = hs.d_b, a.a = hs.d_a, a.k = hs.d_k, a.psi = hs.d_psi;
    h->slabs.push_back(hs);
  }
  drop_graphs(h);
  return EMS_OK;
}

int ems_gpu_add_source(ems_gpu *h, int kind, int n_edges, const int32_t *axis,
                       const int32_t *nodes, const float *w, const float *amp, uint64_t n_amp,
                       int *id) {
  if (!h) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
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
  h->ops.push_back(o);
  h->ops_dirty = true;
  if (id) *id = (int)h->ops.size() - 1;
  return EMS_OK;
}

int ems_gpu_add_probe(ems_gpu *h, int quantity, int n_entries, const int32_t *axis,
                      const int32_t *nodes, const double *w, uint64_t interval, int *id) {
  if (!h) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
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

int ems_gpu_run(ems_gpu *h, uint64_t n) {
  if (!h) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  return run_steps(h, n);
}

int ems_gpu_step_with_energy(ems_gpu *h, double *w) {
  if (!h || !w) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  // Operator energies (Operator::energy, operator.rs): 0 for every kind
  // held here; Sum folds from -0.0 (mod.rs:772-774).
  double w_ops = -0.0;
  for (size_t q = 0; q < h->ops.size() + (h->slabs.empty() ? 0 : 1); q++) w_ops += 0.0;
  double wf = 0;
  int rc = energy_now(h, &wf);
  if (rc) return rc;
  if ((rc = run_steps(h, 1))) return rc;
  *w = wf + w_ops;
  return EMS_OK;
}

int ems_gpu_energy(ems_gpu *h, double *w) {
  if (!h || !w) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  double w_ops = -0.0;
  for (size_t q = 0; q < h->ops.size() + (h->slabs.empty() ? 0 : 1); q++) w_ops += 0.0;
  double wf = 0;
  int rc = energy_now(h, &wf);
  if (rc) return rc;
  *w = wf + w_ops;
  return EMS_OK;
}

uint64_t ems_gpu_timestep(const ems_gpu *h) { return h ? h->step : 0; }

int ems_gpu_read_component(ems_gpu *h, int field, int axis, float *out) {
  if (!h || !out || axis < 0 || axis > 2 || field < 0 || field > 1) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  float *src = field ? h->f.i[axis] : h->f.u[axis];
  CK(h, hipMemcpyAsync(out, src, h->N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  CK(h, hipStreamSynchronize(h->stream));
  return EMS_OK;
}

int ems_gpu_write_component(ems_gpu *h, int field, int axis, const float *in) {
  if (!h || !in || axis < 0 || axis > 2 || field < 0 || field > 1) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  float *dst = field ? h->f.i[axis] : h->f.u[axis];
  CK(h, hipMemcpyAsync(dst, in, h->N * sizeof(float), hipMemcpyHostToDevice, h->stream));
  CK(h, hipStreamSynchronize(h->stream));
  return EMS_OK;
}

int ems_gpu_probe_values(ems_gpu *h, int id, double *out, uint64_t cap, uint64_t *count) {
  if (!h || id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_PROBE)
    return h ? fail(h, EMS_E_INVALID, "not a probe id") : EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  HostOp &o = h->ops[id];
  CK(h, hipStreamSynchronize(h->stream));
  uint64_t n = o.count < cap ? o.count : cap;
  if (n && out)
    CK(h, hipMemcpy(out, o.d_vals, n * sizeof(double), hipMemcpyDeviceToHost));
  if (count) *count = o.count;
  return EMS_OK;
}

int ems_gpu_probe_times(ems_gpu *h, int id, double *out, uint64_t cap, uint64_t *count) {
  if (!h || id < 0 || id >= (int)h->ops.size() || h->ops[id].type != OP_PROBE)
    return h ? fail(h, EMS_E_INVALID, "not a probe id") : EMS_E_INVALID;
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
  if (!h) return EMS_E_INVALID;
  CK(h, hipStreamSynchronize(h->stream));
  return EMS_OK;
}

int ems_gpu_time_kernels(ems_gpu *h, uint64_t steps, double out[3]) {
  if (!h || !out || steps == 0) return EMS_E_INVALID;
  if (h->dead) return fail(h, EMS_E_DEVICE, "engine has a device fault: " + h->err);
  int rc = prepare(h, steps);
  if (rc) return rc;
  hipEvent_t ev[5];
  for (auto &e : ev) CK(h, hipEventCreate(&e));
  double th = 0, te = 0, ts = 0;
  for (uint64_t s = 0; s < steps; s++) {
    CK(h, hipEventRecord(ev[0], h->stream));
    if ((rc = launch_h(h))) return rc;
    CK(h, hipEventRecord(ev[1], h->stream));
    if ((rc = launch_cpml(h, 0))) return rc;
    if (has_mid_ops(h)) {
      hipLaunchKernelGGL(gap_mid, dim3(1), dim3(64), 0, h->stream, h->f, h->d_ops,
                         (int)h->ops.size(), (const uint64_t *)h->d_step);
      CK(h, hipGetLastError());
    }
    CK(h, hipEventRecord(ev[2], h->stream));
    if ((rc = launch_e(h))) return rc;
    CK(h, hipEventRecord(ev[3], h->stream));
    if ((rc = launch_cpml(h, 1))) return rc;
    hipLaunchKernelGGL(gap_post_e, dim3(1), dim3(64), 0, h->stream, h->f, h->d_ops,
                       (int)h->ops.size(), h->d_step);
    CK(h, hipGetLastError());
    CK(h, hipEventRecord(ev[4], h->stream));
    CK(h, hipEventSynchronize(ev[4]));
    float a, b, c;
    CK(h, hipEventElapsedTime(&a, ev[0], ev[1]));
    CK(h, hipEventElapsedTime(&b, ev[2], ev[3]));
    CK(h, hipEventElapsedTime(&c, ev[0], ev[4]));
    th += a, te += b, ts += c;
  }
  for (auto &e : ev) hipEventDestroy(e);
  advance_counts(h, steps);
  out[0] = th / steps, out[1] = te / steps, out[2] = ts / steps;
  return EMS_OK;
}

int ems_gpu_traffic_model(ems_gpu *h, double out[3]) {
  if (!h || !out) return EMS_E_INVALID;
  // Algorithmic bytes of each launch: H reads u (3) and i (3) and writes
  // i (3); E reads i (3) and u (3) and writes u (3); each dense coefficient
  // lane adds 4 B per node (pooled rows stay cache-resident).
  double hb = 36.0, eb = 36.0;
  for (int l = 0; l < 12; l++)
    if (h->lanes[l].rows && !h->lanes[l].index) (l < 6 ? eb : hb) += 4.0;
  out[0] = hb * (double)h->N;
  out[1] = eb * (double)h->N;
  // CPML slabs per step: target r+w, psi r+w, the two source values and
  // the coefficient, 4 B each (cache hits for the neighbour not counted).
  double psi = 0;
  for (auto &s : h->slabs) psi += 24.0 * (double)s.len;
  out[2] = psi;
  return EMS_OK;
}

}  // extern "C"
