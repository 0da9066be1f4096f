// fhd_device.h -- device helpers shared by the gfx950 kernels: the counter-based noise
// generator, the Peskin kernels, staggered-grid geometry and the launch-parameter structs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FHD_DEV __device__ __forceinline__

// Work item (tile, column block, or block of sorted particles) of workgroup b out of n.
// Workgroups are dealt round-robin to the eight XCDs, each with its own L2; giving every XCD a
// contiguous range of items keeps neighbouring items, which share cache lines, on one L2.
FHD_DEV int xcd_item(int b, int n) {
  constexpr int kXcd = 8;
  if (n % kXcd) return b;
  return (b % kXcd) * (n / kXcd) + b / kXcd;
}

// Device error flags (bits of the context's error word; cf. fluam's errorKernel,
// src/allocateErrorArray.cu:22-32): a non-finite particle displacement, and a displacement of
// 2^24 cells or more that was not applied.
enum { kErrNonFinite = 1, kErrClamped = 2, kErrBarrier = 4 };

// Noise streams (counter word 1 = stream << 4 | sub).
enum { FHD_STREAM_FLUID = 1, FHD_STREAM_FLUID_K = 2, FHD_STREAM_RFD = 3, FHD_STREAM_PART1 = 4,
       FHD_STREAM_PART2 = 5, FHD_STREAM_FLUID_P = 6, FHD_STREAM_INIT_POS = 8, FHD_STREAM_INIT_FLUID = 9 };

// ---------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  The 32x32->64 products map to
// v_mad_u64_u32, which runs at the full VALU rate on gfx950, so a round costs two wide
// multiplies and four xors (gfx9 has no three-input xor).
// ---------------------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

FHD_DEV u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    u32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Box-Muller (1958) on 24-bit uniforms u1 in (0,1], u2 in [0,1): z0 = r cos(2 pi u2),
// z1 = r sin(2 pi u2), r = sqrt(-2 ln u1).  The hardware transcendentals take the angle in
// turns (v_sin_f32 / v_cos_f32 compute sin(2 pi x)) and log in base 2 (v_log_f32), so no range
// reduction is needed: u2 already is the angle in turns.  The argument of the square root lies
// in [0, 33.3] (never denormal), so the hardware v_sqrt_f32 (1 ulp) needs no scaling.
FHD_DEV float2 box_muller(uint32_t a, uint32_t b) {
  const float inv24 = 1.0f / 16777216.0f;
  const float u1 = (float)((a >> 8) + 1u) * inv24;
  const float u2 = (float)(b >> 8) * inv24;
  const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));  // -2 ln 2 log2(u1)
  return make_float2(r * __builtin_amdgcn_cosf(u2), r * __builtin_amdgcn_sinf(u2));
}

// Noise key: seed words and step counter.  When `step` is non-null (steps replayed from a
// captured graph) the counter is read from device memory at kernel entry (resolve_key).
struct Key {
  uint32_t k0, k1;
  uint32_t step_lo, step_hi;
  const unsigned long long* step;
};

FHD_DEV Key resolve_key(Key k) {
  if (k.step) {
    const unsigned long long s = *k.step;
    k.step_lo = (uint32_t)s;
    k.step_hi = (uint32_t)(s >> 32);
    k.step = nullptr;
  }
  return k;
}

FHD_DEV u32x4 draw_words(const Key& key, uint32_t index, uint32_t stream, uint32_t sub) {
  u32x4 c = {index, (stream << 4) | sub, key.step_lo, key.step_hi};
  return philox4x32_10(c, key.k0, key.k1);
}

FHD_DEV void draw_normals4(const Key& key, uint32_t index, uint32_t stream, uint32_t sub, float n[4]) {
  u32x4 w = draw_words(key, index, stream, sub);
  float2 a = box_muller(w.x, w.y);
  float2 b = box_muller(w.z, w.w);
  n[0] = a.x; n[1] = a.y; n[2] = b.x; n[3] = b.y;
}

// The part of the Stokes step's noise source D.Sigma that the solve keeps, drawn per mode in
// Fourier space (the half spectrum, at mode (kx, ky, kz) with symbol g and shifts e).  The
// solve projects with P_e = E P E* (P = I - g g^T / |g|^2); for a white stress with
// f1^2 = 2 f4^2 (fluam's prefactors, src/createCellsStokesLimitGPU.cu:58-59) P_e D.Sigma has
// covariance N f4^2 |g|^2 P_e per mode, so it is drawn as u = f4 |g| E (Z1 t1 + Z2 t2) with an
// orthonormal basis t1, t2 of the plane normal to g and complex normals E|Z|^2 = N: four normals
// per mode and stress set instead of twelve.  Basis: with k_a != 0 on all three axes,
// t1 = (g_y, -g_x, 0)/|.| and t2 = g x t1 / |g|; on two axes a < b (c the other), t1 = e_c and
// t2 = (-g_b e_a + g_a e_b)/|g|; on one axis a, the other two unit vectors in axis order.  On the
// planes kx = 0 and kx = mx/2 the member of each (ky, kz), (-ky, -kz) pair with the smaller
// ky + my kz draws (counter index kx + P (ky + my kz), stream FHD_STREAM_FLUID_P, sub = stress
// set) and the other is its conjugate; g is even in k and the canonical mode's shift is
// -conj(e) on an axis with k != 0.  At a self-conjugate mode Z is real (sqrt(N) n) along a
// basis vector on the k = 0 axes and imaginary along one on the k = m/2 axes, so u is real.
FHD_DEV void kspace_noise(const Key& key, int kx, int ky, int kz, int mx, int my, int mz, int nset, float f4,
                          float sqrt_half_n, float gx, float gy, float gz, float2 ex, float2 ey, float2 ez,
                          float2 w[3]) {
  int cy = ky, cz = kz;
  bool conj = false, selfc = false;
  if (kx == 0 || 2 * kx == mx) {
    const int py = ky == 0 ? 0 : my - ky, pz = kz == 0 ? 0 : mz - kz;
    const int L = ky + my * kz, Lp = py + my * pz;
    if (Lp == L) {
      selfc = true;
    } else if (Lp < L) {
      cy = py;
      cz = pz;
      conj = true;
      if (ky != 0) ey = make_float2(-ey.x, ey.y);
      if (kz != 0) ez = make_float2(-ez.x, ez.y);
    }
  }
  const bool ax = kx != 0, ay = ky != 0, az = kz != 0;
  if (!ax && !ay && !az) {  // the mean: nothing to keep
    w[0] = w[1] = w[2] = make_float2(0.0f, 0.0f);
    return;
  }
  const float gn = sqrtf(gx * gx + gy * gy + gz * gz);
  float t1[3] = {0.0f, 0.0f, 0.0f}, t2[3] = {0.0f, 0.0f, 0.0f};
  bool im1 = false, im2 = false;  // at a self-conjugate mode: Z imaginary along t1 / t2
  if (ax && ay && az) {
    const float r = 1.0f / sqrtf(gx * gx + gy * gy);
    t1[0] = gy * r;
    t1[1] = -gx * r;
    const float ig = 1.0f / gn;
    t2[0] = (-gz * t1[1]) * ig;
    t2[1] = (gz * t1[0]) * ig;
    t2[2] = (gx * t1[1] - gy * t1[0]) * ig;
    im1 = im2 = true;
  } else if ((int)ax + (int)ay + (int)az == 2) {
    const int c = !ax ? 0 : !ay ? 1 : 2;
    const int a = c == 0 ? 1 : 0, b = c == 2 ? 1 : 2;
    const float g3[3] = {gx, gy, gz};
    const float ig = 1.0f / gn;
    t1[c] = 1.0f;
    t2[a] = -g3[b] * ig;
    t2[b] = g3[a] * ig;
    im2 = true;
  } else {
    const int a = ax ? 0 : ay ? 1 : 2;
    t1[a == 0 ? 1 : 0] = 1.0f;
    t2[a == 2 ? 1 : 2] = 1.0f;
  }
  const uint32_t index = (uint32_t)(kx + (mx / 2 + 1) * (cy + my * cz));
  float n[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int s = 0; s < nset; ++s) {
    float d[4];
    draw_normals4(key, index, FHD_STREAM_FLUID_P, (uint32_t)s, d);
#pragma unroll
    for (int k = 0; k < 4; ++k) n[k] += d[k];
  }
  float2 Z1, Z2;
  if (selfc) {
    const float sn = 1.41421356237309515f * sqrt_half_n;
    Z1 = im1 ? make_float2(0.0f, sn * n[0]) : make_float2(sn * n[0], 0.0f);
    Z2 = im2 ? make_float2(0.0f, sn * n[2]) : make_float2(sn * n[2], 0.0f);
  } else {
    Z1 = make_float2(sqrt_half_n * n[0], sqrt_half_n * n[1]);
    Z2 = make_float2(sqrt_half_n * n[2], sqrt_half_n * n[3]);
  }
  const float amp = f4 * gn;
  const float2 e3[3] = {ex, ey, ez};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float2 v = make_float2(amp * (Z1.x * t1[a] + Z2.x * t2[a]), amp * (Z1.y * t1[a] + Z2.y * t2[a]));
    const float2 u = make_float2(e3[a].x * v.x - e3[a].y * v.y, e3[a].x * v.y + e3[a].y * v.x);
    w[a] = conj ? make_float2(u.x, -u.y) : u;
  }
}

// ---------------------------------------------------------------------------------------
// Peskin kernels (r in grid units).  3-point: src/deltaGPU.cu:57-68; 4-point: Peskin (2002),
// the 3D tensor form of src/deltaGPU.cu:44-53.
// ---------------------------------------------------------------------------------------
FHD_DEV float phi3(float r) {
  float x = fabsf(r);
  if (x >= 1.5f) return 0.0f;
  if (x < 0.5f) return (1.0f + sqrtf(fmaxf(1.0f - 3.0f * x * x, 0.0f))) * (1.0f / 3.0f);
  float y = 1.0f - x;
  return (5.0f - 3.0f * x - sqrtf(fmaxf(1.0f - 3.0f * y * y, 0.0f))) * (1.0f / 6.0f);
}
FHD_DEV float dphi3(float r) {  // d phi3 / dr
  float x = fabsf(r);
  if (x >= 1.5f) return 0.0f;
  if (x <= 0.5f) return -r * rsqrtf(fmaxf(1.0f - 3.0f * r * r, 1e-30f));
  float s = r > 0.0f ? 1.0f : -1.0f;
  return -s * 0.5f * (1.0f + (1.0f - x) * rsqrtf(fmaxf(-2.0f + 6.0f * x - 3.0f * x * x, 1e-30f)));
}
FHD_DEV float phi4(float r) {
  float x = fabsf(r);
  if (x >= 2.0f) return 0.0f;
  if (x <= 1.0f) return (3.0f - 2.0f * x + sqrtf(fmaxf(1.0f + 4.0f * x - 4.0f * x * x, 0.0f))) * 0.125f;
  return (5.0f - 2.0f * x - sqrtf(fmaxf(-7.0f + 12.0f * x - 4.0f * x * x, 0.0f))) * 0.125f;
}
FHD_DEV float dphi4(float r) {
  float x = fabsf(r);
  if (x >= 2.0f) return 0.0f;
  float s = r > 0.0f ? 1.0f : -1.0f;
  if (x <= 1.0f)
    return s * (-2.0f + (2.0f - 4.0f * x) * rsqrtf(fmaxf(1.0f + 4.0f * x - 4.0f * x * x, 1e-30f))) * 0.125f;
  return s * (-2.0f - (6.0f - 4.0f * x) * rsqrtf(fmaxf(-7.0f + 12.0f * x - 4.0f * x * x, 1e-30f))) * 0.125f;
}

FHD_DEV int wrapi(int i, int m) {
  int r = i % m;
  return r < 0 ? r + m : r;
}

// ---------------------------------------------------------------------------------------
// Grid description passed by value to kernels.
// ---------------------------------------------------------------------------------------
struct GridDev {
  int mx, my, mz, N, P;        // P = mx/2 + 1 (half-spectrum row length)
  float dx, dy, dz;            // cell sizes
  float idx, idy, idz;         // inverses
  float lx, ly, lz;
};

// Particle position: integer cell (unwrapped) + offset from the cell centre in
// [-h/2, h/2) per axis.  The unwrapped coordinate is (c + 1/2) h - L/2 + d.
struct PartPos {
  int4* cell;    // x, y, z, w unused
  float4* off;   // x, y, z, w unused
};

// 1D weights of one axis for one velocity component: nodes base .. base+np-1 (unwrapped).
struct AxisWeights {
  int base;
  float w[4];
};

// Weights along an axis with node offset `stag` (0: cell-centred nodes, 1: face nodes at +h/2)
// for a particle in cell c at offset frac = d/h in [-1/2, 1/2).  Returns nodes for the 3-point
// (npts = 3) or 4-point (npts = 4) kernel.  Same stencil choice as the FP64 restatement: the
// 3-point stencil is centred on the nearest node, the 4-point one on the two nearest.
// The weight sets use one square root each: with f the offset from the stencil's reference
// node, the 3-point weights follow from phi3(f) and the moment conditions sum w = 1,
// sum (f - k) w_k = 0; the 4-point weights are Peskin's one-root form (src/deltaGPU.cu:44-53).
// They equal phi3 / phi4 at the respective distances.
template <int NPTS>
FHD_DEV void axis_weights(int c, float frac, int stag, AxisWeights& aw) {
  // t = position in node units relative to node index 0 of this staggering.
  // centred nodes: t = c + frac; face nodes: t = c - 1/2 + frac.
  if (NPTS == 3) {
    int j0;
    float f;  // t - j0, in [-1/2, 1/2)
    if (!stag) { j0 = c; f = frac; }
    else if (frac >= 0.0f) { j0 = c; f = frac - 0.5f; }
    else { j0 = c - 1; f = frac + 0.5f; }
    aw.base = j0 - 1;
    const float w0 = (1.0f + sqrtf(fmaxf(1.0f - 3.0f * f * f, 0.0f))) * (1.0f / 3.0f);
    aw.w[0] = 0.5f * (1.0f - w0 - f);  // node j0 - 1, distance f + 1
    aw.w[1] = w0;
    aw.w[2] = 0.5f * (1.0f - w0 + f);  // node j0 + 1, distance f - 1
    aw.w[3] = 0.0f;
  } else {
    int j0;
    float f;  // t - j0 in [0, 1)
    if (!stag) {
      if (frac >= 0.0f) { j0 = c; f = frac; } else { j0 = c - 1; f = frac + 1.0f; }
    } else {
      j0 = c - 1; f = frac + 0.5f;
    }
    aw.base = j0 - 1;
    const float s = 0.125f * (3.0f - 2.0f * f + sqrtf(fmaxf(1.0f + 4.0f * f * (1.0f - f), 0.0f)));
    aw.w[0] = 0.75f - 0.5f * f - s;  // distance f + 1
    aw.w[1] = s;                      // distance f
    aw.w[2] = 0.5f * f - 0.25f + s;  // distance f - 1
    aw.w[3] = 0.5f - s;               // distance f - 2
  }
}

// Derivative weights dphi/dq (per unit length: divide by h at the call site).
template <int NPTS>
FHD_DEV void axis_dweights(int c, float frac, int stag, AxisWeights& aw) {
  if (NPTS == 3) {
    int j0;
    float f;
    if (!stag) { j0 = c; f = frac; }
    else if (frac >= 0.0f) { j0 = c; f = frac - 0.5f; }
    else { j0 = c - 1; f = frac + 0.5f; }
    aw.base = j0 - 1;
    // derivatives of the one-root forms in axis_weights with respect to f
    const float d0 = -f * rsqrtf(fmaxf(1.0f - 3.0f * f * f, 1e-30f));
    aw.w[0] = 0.5f * (-d0 - 1.0f);
    aw.w[1] = d0;
    aw.w[2] = 0.5f * (-d0 + 1.0f);
    aw.w[3] = 0.0f;
  } else {
    int j0;
    float f;
    if (!stag) {
      if (frac >= 0.0f) { j0 = c; f = frac; } else { j0 = c - 1; f = frac + 1.0f; }
    } else {
      j0 = c - 1; f = frac + 0.5f;
    }
    aw.base = j0 - 1;
    const float ds = 0.125f * (-2.0f + (2.0f - 4.0f * f) * rsqrtf(fmaxf(1.0f + 4.0f * f * (1.0f - f), 1e-30f)));
    aw.w[0] = -0.5f - ds;
    aw.w[1] = ds;
    aw.w[2] = 0.5f + ds;
    aw.w[3] = -ds;
  }
}

// Renormalise (cell, offset) after a displacement so the offset lies in [-h/2, h/2).
FHD_DEV void renorm_axis(int& c, float& d, float h, float ih) {
  float k = floorf(d * ih + 0.5f);
  // Guard against non-finite displacements: keep the particle where it was.
  if (!(fabsf(k) < 1.0e9f)) { k = 0.0f; d = 0.0f; }
  c += (int)k;
  d -= k * h;
  // rounding may leave d marginally outside; fold once more
  if (d >= 0.5f * h) { c += 1; d -= h; }
  else if (d < -0.5f * h) { c -= 1; d += h; }
}
