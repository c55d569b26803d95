// Colour-vector (3 complex) site loads/stores for the three storage
// precisions, used by the Dslash and the fused solver kernels.  FP16 fields
// store v / max|v| in half precision plus the FP32 max-norm per site.
#pragma once
#include "qamd_internal.h"

namespace qamd {

// complex multiply-accumulate helpers in the arithmetic type
template <class C, class R>
__device__ __forceinline__ void cmac(C &acc, const C &m, const C &v) {
  acc.x = fma(m.x, v.x, acc.x);
  acc.x = fma(-m.y, v.y, acc.x);
  acc.y = fma(m.x, v.y, acc.y);
  acc.y = fma(m.y, v.x, acc.y);
}

template <int P> struct VecLd;
template <> struct VecLd<QAMD_F64> {
  __device__ static void ld(const double2 *p, const float *, int64_t vp, int i, double2 (&v)[3], double) {
    v[0] = p[i];
    v[1] = p[i + vp];
    v[2] = p[i + 2 * vp];
  }
};
template <> struct VecLd<QAMD_F32> {
  __device__ static void ld(const float2 *p, const float *, int64_t vp, int i, float2 (&v)[3], float) {
    v[0] = p[i];
    v[1] = p[i + vp];
    v[2] = p[i + 2 * vp];
  }
};
template <> struct VecLd<QAMD_F16> {
  __device__ static void ld(const __half2 *p, const float *nrm, int64_t vp, int i, float2 (&v)[3], float sc) {
    float n = nrm[i] * sc;
    float2 a = __half22float2(p[i]), b = __half22float2(p[i + vp]), c = __half22float2(p[i + 2 * vp]);
    v[0] = make_float2(a.x * n, a.y * n);
    v[1] = make_float2(b.x * n, b.y * n);
    v[2] = make_float2(c.x * n, c.y * n);
  }
};
template <> struct VecLd<QAMD_Q16> {
  __device__ static void ld(const short2 *p, const float *nrm, int64_t vp, int i, float2 (&v)[3], float sc) {
    float n = nrm[i] * sc * (1.0f / Q16_ONE);
    short2 a = p[i], b = p[i + vp], c = p[i + 2 * vp];
    v[0] = make_float2(a.x * n, a.y * n);
    v[1] = make_float2(b.x * n, b.y * n);
    v[2] = make_float2(c.x * n, c.y * n);
  }
};

template <int P> struct VecSt;
template <> struct VecSt<QAMD_F64> {
  __device__ static void st(double2 *p, float *, int64_t vp, int i, const double2 (&v)[3]) {
    p[i] = v[0];
    p[i + vp] = v[1];
    p[i + 2 * vp] = v[2];
  }
};
template <> struct VecSt<QAMD_F32> {
  __device__ static void st(float2 *p, float *, int64_t vp, int i, const float2 (&v)[3]) {
    p[i] = v[0];
    p[i + vp] = v[1];
    p[i + 2 * vp] = v[2];
  }
};
template <> struct VecSt<QAMD_F16> {
  __device__ static void st(__half2 *p, float *nrm, int64_t vp, int i, const float2 (&v)[3]) {
    float m = 0.f;
#pragma unroll
    for (int j = 0; j < 3; j++) m = fmaxf(m, fmaxf(fabsf(v[j].x), fabsf(v[j].y)));
    float inv = m > 0.f ? 1.0f / m : 0.f;
    p[i] = __floats2half2_rn(v[0].x * inv, v[0].y * inv);
    p[i + vp] = __floats2half2_rn(v[1].x * inv, v[1].y * inv);
    p[i + 2 * vp] = __floats2half2_rn(v[2].x * inv, v[2].y * inv);
    nrm[i] = m;
  }
};
template <> struct VecSt<QAMD_Q16> {
  __device__ static void st(short2 *p, float *nrm, int64_t vp, int i, const float2 (&v)[3]) {
    float m = 0.f;
#pragma unroll
    for (int j = 0; j < 3; j++) m = fmaxf(m, fmaxf(fabsf(v[j].x), fabsf(v[j].y)));
    float inv = m > 0.f ? 1.0f / m : 0.f;
    p[i] = q16_pack(v[0].x * inv, v[0].y * inv);
    p[i + vp] = q16_pack(v[1].x * inv, v[1].y * inv);
    p[i + 2 * vp] = q16_pack(v[2].x * inv, v[2].y * inv);
    nrm[i] = m;
  }
};

}  // namespace qamd
