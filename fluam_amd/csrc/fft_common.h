// fft_common.h -- complex arithmetic and in-register DFT butterflies shared by the generic and
// the size-specialised Stockham FFT kernels.
#pragma once
#include "fhd_device.h"

namespace fhd {

FHD_DEV float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
FHD_DEV float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
FHD_DEV float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
FHD_DEV float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
FHD_DEV float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (s i), s = +-1
FHD_DEV float2 cmul_si(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// In-register DFTs with kernel e^{s 2 pi i / R}.
template <int R> struct Dft;
template <> struct Dft<2> {
  static FHD_DEV void run(float2* v, float) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <> struct Dft<3> {
  static FHD_DEV void run(float2* v, float s) {
    const float c3 = 0.86602540378443864676f;
    float2 t = cadd(v[1], v[2]);
    float2 u = csub(v[1], v[2]);
    float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    float2 r = cmul_si(cscale(u, c3), s);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};
template <> struct Dft<4> {
  static FHD_DEV void run(float2* v, float s) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = cmul_si(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};
template <> struct Dft<8> {
  static FHD_DEV void run(float2* v, float s) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4>::run(e, s);
    Dft<4>::run(o, s);
    const float h = 0.70710678118654752440f;
    // w^k = e^{s 2 pi i k / 8}
    float2 o1 = cmul(o[1], make_float2(h, s * h));
    float2 o2 = cmul_si(o[2], s);
    float2 o3 = cmul(o[3], make_float2(-h, s * h));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

template <> struct Dft<16> {
  static FHD_DEV void run(float2* v, float s) {
    float2 e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      e[k] = v[2 * k];
      o[k] = v[2 * k + 1];
    }
    Dft<8>::run(e, s);
    Dft<8>::run(o, s);
    // w^k = e^{s 2 pi i k / 16}: cos and sin of k pi / 8
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float h = 0.70710678118654752440f;
    const float2 w[8] = {make_float2(1.0f, 0.0f), make_float2(c1, s * s1), make_float2(h, s * h),
                         make_float2(s1, s * c1), make_float2(0.0f, s), make_float2(-s1, s * c1),
                         make_float2(-h, s * h), make_float2(-c1, s * s1)};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float2 t = k == 0 ? o[0] : cmul(o[k], w[k]);
      v[k] = cadd(e[k], t);
      v[k + 8] = csub(e[k], t);
    }
  }
};

// Radix plan of the size-specialised kernels: 16s first, then 8, 4, 2, 3 (fewest LDS stages).
constexpr int plan_radix(int n, int st) {
  for (int i = 0;; ++i) {
    int r = (n % 16 == 0) ? 16 : (n % 8 == 0) ? 8 : (n % 4 == 0) ? 4 : (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : 1;
    if (i == st || r == 1) return r;
    n /= r;
  }
}

// Explicit radix plans, 5 bits per stage from the first (0 ends the plan; plan 0 is the default
// above).  The schedule autotuner picks one per transform length among the candidates below.
constexpr uint32_t plan2(int r0, int r1) { return (uint32_t)r0 | ((uint32_t)r1 << 5); }
constexpr uint32_t plan3(int r0, int r1, int r2) { return plan2(r0, r1) | ((uint32_t)r2 << 10); }
constexpr int plan_at(uint32_t plan, int n, int st) { return plan ? (int)((plan >> (5 * st)) & 31u) : plan_radix(n, st); }
// Candidate plans: index 0 is the default for every length; 48, 64, 96 and 128 have two more
// (for 48 and 96, other placements of the radix-3 stage).
constexpr uint32_t kPlan48[3] = {0u, plan2(3, 16), plan3(4, 4, 3)};
constexpr uint32_t kPlan64[3] = {0u, plan2(4, 16), plan2(8, 8)};
constexpr uint32_t kPlan96[3] = {0u, plan3(8, 4, 3), plan3(3, 8, 4)};
constexpr uint32_t kPlan128[3] = {0u, plan2(8, 16), plan3(4, 4, 8)};
constexpr uint32_t plan_choice(int n, int i) {
  return i == 0     ? 0u
         : n == 48  ? kPlan48[i]
         : n == 64  ? kPlan64[i]
         : n == 96  ? kPlan96[i]
         : n == 128 ? kPlan128[i]
                    : 0u;
}

}  // namespace fhd
