// Device helpers shared by the HIP translation units: generic site
// load/store for every storage precision, QLL host-layout indexing, and the
// counter-based random number generator.
#pragma once
#include "qamd_internal.h"
#include "qamd_objects.h"

namespace qamd {

// ------------------------------------------------------- site load / store
template <int P> struct IO;
template <> struct IO<QAMD_F64> {
  __device__ static double2 ld(const Field &f, int64_t i) { return ((const double2 *)f.data)[i]; }
  __device__ static void st(Field &f, int64_t i, double2 v) { ((double2 *)f.data)[i] = v; }
};
template <> struct IO<QAMD_F32> {
  __device__ static double2 ld(const Field &f, int64_t i) {
    float2 v = ((const float2 *)f.data)[i];
    return make_double2(v.x, v.y);
  }
  __device__ static void st(Field &f, int64_t i, double2 v) {
    ((float2 *)f.data)[i] = make_float2((float)v.x, (float)v.y);
  }
};

// generic (slow path) access to a whole site: NC complex numbers in double
// (NC = the field's ncomp, 3 or 12, fixed at compile time so the site stays
// in registers)
template <int NC>
__device__ inline void site_load(const Field &f, int rhs, int cb, double2 *v) {
  if (f.prec == QAMD_F64) {
#pragma unroll
    for (int j = 0; j < NC; j++) v[j] = IO<QAMD_F64>::ld(f, f.cidx(rhs, j, cb));
  } else if (f.prec == QAMD_F32) {
#pragma unroll
    for (int j = 0; j < NC; j++) v[j] = IO<QAMD_F32>::ld(f, f.cidx(rhs, j, cb));
  } else if (f.prec == QAMD_F16) {
    float nrm = f.norm[f.nidx(rhs, cb)];
#pragma unroll
    for (int j = 0; j < NC; j++) {
      float2 h = __half22float2(((const __half2 *)f.data)[f.cidx(rhs, j, cb)]);
      v[j] = make_double2((double)h.x * nrm, (double)h.y * nrm);
    }
  } else {
    double nrm = (double)f.norm[f.nidx(rhs, cb)] / Q16_ONE;
#pragma unroll
    for (int j = 0; j < NC; j++) {
      short2 h = ((const short2 *)f.data)[f.cidx(rhs, j, cb)];
      v[j] = make_double2(h.x * nrm, h.y * nrm);
    }
  }
}

template <int NC>
__device__ inline void site_store(Field &f, int rhs, int cb, const double2 *v) {
  if (f.prec == QAMD_F64) {
#pragma unroll
    for (int j = 0; j < NC; j++) IO<QAMD_F64>::st(f, f.cidx(rhs, j, cb), v[j]);
  } else if (f.prec == QAMD_F32) {
#pragma unroll
    for (int j = 0; j < NC; j++) IO<QAMD_F32>::st(f, f.cidx(rhs, j, cb), v[j]);
  } else {
    double m = 0;
#pragma unroll
    for (int j = 0; j < NC; j++) m = fmax(m, fmax(fabs(v[j].x), fabs(v[j].y)));
    float nrm = (float)m;
    float inv = nrm > 0 ? 1.0f / nrm : 0.0f;
#pragma unroll
    for (int j = 0; j < NC; j++) {
      if (f.prec == QAMD_F16)
        ((__half2 *)f.data)[f.cidx(rhs, j, cb)] = __floats2half2_rn((float)v[j].x * inv, (float)v[j].y * inv);
      else
        ((short2 *)f.data)[f.cidx(rhs, j, cb)] = q16_pack((float)v[j].x * inv, (float)v[j].y * inv);
    }
    f.norm[f.nidx(rhs, cb)] = nrm;
  }
}

// ---------------------------------------------------- QLL layout on device
struct QllLay {
  int phys[4], inner[4], outer[4];
  int nSitesOuter, NI, nEvenOuter, innerCb, innerCbDir;
};

// layoutIndex (lib/layout.c:120-142) restated for one rank: outer index oi
// (even outer sites first) and inner lane ii
__device__ __host__ inline void qll_index(const QllLay &l, const int c[4], int &oi2, int &ii) {
  int x[4] = {c[0], c[1], c[2], c[3]};
  ii = 0;
  for (int i = 3; i >= 0; i--) ii = ii * l.inner[i] + ((x[i] / l.outer[i]) % l.inner[i]);
  int ib = 0;
  for (int i = 0; i < 4; i++) ib += ((x[i] / l.outer[i]) % l.inner[i]) * l.outer[i];
  ib &= 1;
  x[l.innerCbDir] += l.innerCb * ib;
  int oi = 0;
  for (int i = 3; i >= 0; i--) oi = oi * l.outer[i] + (x[i] % l.outer[i]);
  int p = (c[0] + c[1] + c[2] + c[3]) & 1;
  oi2 = p ? (oi + l.nSitesOuter) / 2 : oi / 2;
}

inline QllLay to_lay(const qamd_qll_layout *l) {
  QllLay q;
  for (int i = 0; i < 4; i++) {
    q.phys[i] = l->physGeom[i];
    q.inner[i] = l->innerGeom[i];
    q.outer[i] = l->outerGeom[i];
  }
  q.nSitesOuter = l->nSitesOuter;
  q.NI = l->nSitesInner;
  q.nEvenOuter = l->nEvenOuter;
  q.innerCb = l->innerCb;
  q.innerCbDir = l->innerCbDir;
  return q;
}

// host element index of element 0 of (rhs, site) and the element stride
struct SrcMap {
  int kind;        // 0 lexicographic [rhs][V][nreal], 1 QLL half array, 2 site table
  int nreal;
  QllLay lay;
  const int *inv;  // site table: host site index of each lexicographic site
  __device__ void map(const Geom &g, int rhs, int par, int cb, int64_t &base, int64_t &stride) const {
    Site s = site_of(g, cb, par);
    if (kind == 0) {
      base = ((int64_t)rhs * g.V + s.lex) * nreal;
      stride = 1;
    } else if (kind == 2) {
      base = (int64_t)inv[s.lex] * nreal;
      stride = 1;
    } else {
      int oi, ii;
      qll_index(lay, s.c, oi, ii);
      oi -= par * lay.nEvenOuter;   // staging holds only this parity's half
      base = (int64_t)nreal * lay.NI * oi + ii;
      stride = lay.NI;
    }
  }
};

// ------------------------------------------------------------- RNG (device)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t rng64(uint64_t seed, uint64_t stream, uint64_t site, uint64_t idx) {
  const uint64_t g = 0x9E3779B97F4A7C15ULL;
  uint64_t z = mix64(seed + g * (stream + 1));
  z = mix64(z + g * (site + 1));
  return mix64(z + g * (idx + 1));
}
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t stream, uint64_t site, uint64_t idx) {
  return (double)((rng64(seed, stream, site, idx) >> 11) + 1) * 0x1.0p-53;
}
// Box-Muller on uniforms 2*pair, 2*pair+1
__device__ inline void gauss2(uint64_t seed, uint64_t stream, uint64_t site, uint64_t pair, double &a, double &b) {
  double u1 = uniform01(seed, stream, site, 2 * pair);
  double u2 = uniform01(seed, stream, site, 2 * pair + 1);
  double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincos(6.283185307179586476925 * u2, &s, &c);
  a = r * c;
  b = r * s;
}


}  // namespace qamd
