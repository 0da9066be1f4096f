// fft.hip -- batched, LDS-staged mixed-radix Stockham 3D real FFT for gfx950, with the
// Fourier-space viscous / Stokes solve fused into the z-pencil pass.
//
// Replaces the cufftExecZ2Z calls + kernelShift + kernelUpdateV* chain of fluam
// (src/runSchemeIncompressible.cu:111-138, src/runSchemeStokesLimit.cu:198-259,
//  src/realToComplex.cu:263-303, src/kernelUpdateVIncompressible.cu:21-72,
//  src/stokesLimitFunctions.cu:2399-2451).
//
// Three passes per solve, each one read + one write of the field:
//   A  xy-forward : one workgroup per (z-plane, component): real plane -> R2C along x (an
//                   mx/2-point complex FFT plus the even/odd split) -> complex FFT along y.
//   B  z-solve    : one workgroup per (ky, tile of kx) for all three components: FFT along z,
//                   staggering shifts + projection + viscous / Stokes operator, inverse FFT.
//   C  xy-inverse : inverse of A, with an epilogue that blends the result with another field.
// Half-spectrum layout in HBM: spec[c][ky][z][kx], kx in [0, mx/2], row pitch P = mx/2 + 1, so
// a pass-B tile (fixed ky, a run of kx, every z) is one compact block of rows.
#include "fhd_device.h"
#include "fhd_kernels.h"
#include "fft_common.h"

namespace fhd {

// Lines in LDS: line l starts at (l / L1) * S1 + (l % L1) * S0, element stride es.
struct LineMap {
  int lines, L1, S1, S0, es;
  FHD_DEV int base(int l) const { return (l / L1) * S1 + (l % L1) * S0; }
};

// One Stockham radix-R stage over all lines (Govindaraju et al., SC'08 formulation):
// butterfly j of a line reads x[j + r n/R], twiddles by w^{(j mod Ns) r}, w = e^{s 2 pi i/(Ns R)},
// and writes to (j / Ns) Ns R + (j mod Ns) + r Ns.  In place: every thread reads its
// butterflies into registers, the workgroup synchronises, then writes.  Consecutive lanes take
// butterfly j of consecutive lines, which keeps the early stages' strided writes out of each
// other's LDS banks (consecutive butterflies of one line would collide R ways).
template <int R, int KMAX>
FHD_DEV void stockham_stage(float2* buf, const LineMap& lm, int n, int Ns, const float2* tw, int twmul,
                            float s) {
  const int nbl = n / R;
  const int total = lm.lines * nbl;
  float2 v[KMAX][R];
  int wbase[KMAX];
  const int stride = n / (Ns * R);
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    int b = threadIdx.x + k * blockDim.x;
    wbase[k] = -1;
    if (b < total) {
      int j = b / lm.lines;
      int l = b - j * lm.lines;
      int base = lm.base(l);
      int kk = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = buf[base + (j + r * nbl) * lm.es];
        if (r > 0 && kk > 0) {
          float2 w = tw[(kk * r * stride) * twmul];
          if (s > 0.0f) w.y = -w.y;
          x = cmul(x, w);
        }
        v[k][r] = x;
      }
      Dft<R>::run(v[k], s);
      wbase[k] = base + ((j / Ns) * Ns * R + kk) * lm.es;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    if (wbase[k] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[wbase[k] + r * Ns * lm.es] = v[k][r];
    }
  }
  __syncthreads();
}

// Maximum complex elements staged in LDS by one workgroup.
constexpr int kLdsElems = 8448;
constexpr int kThreads = 256;
constexpr int kmax_for(int R) { return (kLdsElems + R * kThreads - 1) / (R * kThreads); }

// Full FFT of length n on all lines described by lm (radix plan packed 4 bits per stage).
FHD_DEV void lds_fft(float2* buf, const LineMap& lm, int n, uint32_t plan, int nstages, const float2* tw,
                     int twmul, float s) {
  int Ns = 1;
  for (int st = 0; st < nstages; ++st) {
    int R = (plan >> (4 * st)) & 15;
    switch (R) {
      case 8: stockham_stage<8, kmax_for(8)>(buf, lm, n, Ns, tw, twmul, s); break;
      case 4: stockham_stage<4, kmax_for(4)>(buf, lm, n, Ns, tw, twmul, s); break;
      case 3: stockham_stage<3, kmax_for(3)>(buf, lm, n, Ns, tw, twmul, s); break;
      default: stockham_stage<2, kmax_for(2)>(buf, lm, n, Ns, tw, twmul, s); break;
    }
    Ns *= R;
  }
}

// ---------------------------------------------------------------------------------------
// Pass A: real plane -> half spectrum along x -> FFT along y.
// ---------------------------------------------------------------------------------------
FHD_BODY void body_fft_xy_fwd(const FftArgs& a, const float* __restrict__ in, float2* __restrict__ spec,
                              const Blk& b) {
  __shared__ float2 buf[kLdsElems];
  __shared__ float2 twx[kMaxFftLen];
  __shared__ float2 twy[kMaxFftLen];
  const int z = b.x, c = b.y;
  const int mx = a.g.mx, my = a.g.my, mz = a.g.mz, P = a.g.P, h = mx / 2;
  for (int i = threadIdx.x; i < mx; i += blockDim.x) twx[i] = a.twx[i];
  for (int i = threadIdx.x; i < my; i += blockDim.x) twy[i] = a.twy[i];
  float* fbuf = reinterpret_cast<float*>(buf);
  const float* src = in + ((size_t)c * mz + z) * (size_t)my * mx;
  for (int e = threadIdx.x; e < mx * my; e += blockDim.x) {
    int y = e / mx, x = e - y * mx;
    fbuf[y * 2 * P + x] = src[e];
  }
  __syncthreads();
  // x: h-point complex FFT of the even/odd-packed rows
  LineMap rows = {my, my, 0, P, 1};
  lds_fft(buf, rows, h, a.plan_h, a.nst_h, twx, 2, -1.0f);
  // even/odd split: X[k] = E + w^k O, X[h-k] = conj(E - w^k O)
  const int npair = h / 2 + 1;
  for (int e = threadIdx.x; e < my * npair; e += blockDim.x) {
    int y = e / npair, k = e - y * npair;
    float2* row = buf + y * P;
    float2 A = row[k];
    float2 B = cconj(row[(h - k) % h]);
    float2 E = cscale(cadd(A, B), 0.5f);
    float2 D = csub(A, B);
    float2 O = make_float2(0.5f * D.y, -0.5f * D.x);  // (A - B) / (2i)
    float2 wO = cmul(twx[k], O);
    if (k == 0) {
      row[0] = make_float2(E.x + O.x, 0.0f);
      row[h] = make_float2(E.x - O.x, 0.0f);
    } else {
      row[k] = cadd(E, wO);
      if (h - k != k) row[h - k] = cconj(csub(E, wO));
    }
  }
  __syncthreads();
  // y: columns kx = 0..P-1
  LineMap cols = {P, P, 0, 1, P};
  lds_fft(buf, cols, my, a.plan_y, a.nst_y, twy, 1, -1.0f);
  for (int e = threadIdx.x; e < P * my; e += blockDim.x) {
    int ky = e / P, kx = e - ky * P;
    spec[(((size_t)c * my + ky) * mz + z) * P + kx] = buf[e];
  }
}
FHD_KERNEL_PAIR((), kThreads, k_fft_xy_fwd, FftPass, body_fft_xy_fwd(a.f, a.in, a.spec, b))

// ---------------------------------------------------------------------------------------
// Pass B: FFT along z, Fourier-space operator, inverse FFT along z, for all three components.
// ---------------------------------------------------------------------------------------
FHD_BODY void body_fft_z_solve(const FftArgs& a, float2* __restrict__ spec, const Blk& b) {
  __shared__ float2 buf[kLdsElems];
  __shared__ float2 twz[kMaxFftLen];
  __shared__ float2 ezs[kMaxFftLen];
  __shared__ float gzs[kMaxFftLen];
  const int T = a.ztile;
  const int kx0 = b.x * T, ky = b.y;
  const int my = a.g.my, mz = a.g.mz, P = a.g.P;
  for (int i = threadIdx.x; i < mz; i += blockDim.x) {
    twz[i] = a.twz[i];
    ezs[i] = a.ez[i];
    gzs[i] = a.gz[i];
  }
  const int tot = a.ncomp * mz * T;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int c = e / (mz * T), r = e - c * mz * T, z = r / T, t = r - z * T;
    int kx = kx0 + t;
    float2 val = make_float2(0.0f, 0.0f);
    if (kx < P) val = spec[(((size_t)c * my + ky) * mz + z) * P + kx];
    buf[e] = val;
  }
  __syncthreads();
  LineMap zl = {a.ncomp * T, T, mz * T, 1, T};
  if (a.zmode != 2) lds_fft(buf, zl, mz, a.plan_z, a.nst_z, twz, 1, -1.0f);
  if (a.zmode != 0) {
    // transform-only modes (building blocks for the parity tests): no operator, no scaling
    if (a.zmode == 2) lds_fft(buf, zl, mz, a.plan_z, a.nst_z, twz, 1, 1.0f);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      int c = e / (mz * T), r = e - c * mz * T, z = r / T, t = r - z * T;
      int kx = kx0 + t;
      if (kx < P) spec[(((size_t)c * my + ky) * mz + z) * P + kx] = buf[e];
    }
    return;
  }
  // Fourier-space operator in the staggered (shifted) basis; 1/N normalisation folded in.
  const float gy = a.gy[ky];
  const float2 ey = a.ey[ky];
  const Key key = a.knoise ? resolve_key(a.key) : a.key;
  for (int e = threadIdx.x; e < mz * T; e += blockDim.x) {
    int z = e / T, t = e - z * T;
    int kx = kx0 + t;
    if (kx >= P) continue;
    float gx = a.gx[kx];
    float2 ex = a.ex[kx];
    float gz = gzs[z];
    float2 ez = ezs[z];
    float2 b0 = buf[e], b1 = buf[mz * T + e], b2 = buf[2 * mz * T + e];
    if (a.knoise) {
      // the kept part of the stochastic stress divergence, drawn in Fourier space (Stokes limit)
      float2 nw[3];
      kspace_noise(key, kx, ky, z, a.g.mx, my, mz, a.knoise, a.kn_f4, a.kn_sqrt_half_n, gx, gy, gz,
                        ex, ey, ez, nw);
      b0 = cadd(b0, nw[0]);
      b1 = cadd(b1, nw[1]);
      b2 = cadd(b2, nw[2]);
    }
    float2 w0 = cmul(b0, cconj(ex));
    float2 w1 = cmul(b1, cconj(ey));
    float2 w2 = cmul(b2, cconj(ez));
    float L = -(gx * gx + gy * gy + gz * gz);
    float2 o0, o1, o2;
    if (kx == 0 && ky == 0 && z == 0) {
      float keep = a.stokes ? 0.0f : a.inv_n;
      o0 = cscale(w0, keep);
      o1 = cscale(w1, keep);
      o2 = cscale(w2, keep);
    } else {
      float2 GW = make_float2(gx * w0.x + gy * w1.x + gz * w2.x, gx * w0.y + gy * w1.y + gz * w2.y);
      float den = a.stokes ? (-a.eta * L) : (1.0f - a.cn_coef * L);
      float sc = a.inv_n / den;
      float iL = 1.0f / L;
      o0 = cscale(make_float2(w0.x + gx * GW.x * iL, w0.y + gx * GW.y * iL), sc);
      o1 = cscale(make_float2(w1.x + gy * GW.x * iL, w1.y + gy * GW.y * iL), sc);
      o2 = cscale(make_float2(w2.x + gz * GW.x * iL, w2.y + gz * GW.y * iL), sc);
    }
    buf[e] = cmul(o0, ex);
    buf[mz * T + e] = cmul(o1, ey);
    buf[2 * mz * T + e] = cmul(o2, ez);
  }
  __syncthreads();
  lds_fft(buf, zl, mz, a.plan_z, a.nst_z, twz, 1, 1.0f);
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int c = e / (mz * T), r = e - c * mz * T, z = r / T, t = r - z * T;
    int kx = kx0 + t;
    if (kx < P) spec[(((size_t)c * my + ky) * mz + z) * P + kx] = buf[e];
  }
}
FHD_KERNEL_PAIR((), kThreads, k_fft_z_solve, FftPass, body_fft_z_solve(a.f, a.spec, b))

// ---------------------------------------------------------------------------------------
// Pass C: inverse FFT along y, C2R along x; out = alpha * result + beta * aux.
// ---------------------------------------------------------------------------------------
FHD_BODY void body_fft_xy_inv(const FftArgs& a, const float2* __restrict__ spec, float* __restrict__ out,
                              const float* __restrict__ aux, float alpha, float beta, const Blk& b) {
  __shared__ float2 buf[kLdsElems];
  __shared__ float2 twx[kMaxFftLen];
  __shared__ float2 twy[kMaxFftLen];
  const int z = b.x, c = b.y;
  const int mx = a.g.mx, my = a.g.my, mz = a.g.mz, P = a.g.P, h = mx / 2;
  for (int i = threadIdx.x; i < mx; i += blockDim.x) twx[i] = a.twx[i];
  for (int i = threadIdx.x; i < my; i += blockDim.x) twy[i] = a.twy[i];
  for (int e = threadIdx.x; e < P * my; e += blockDim.x) {
    int ky = e / P, kx = e - ky * P;
    buf[e] = spec[(((size_t)c * my + ky) * mz + z) * P + kx];
  }
  __syncthreads();
  LineMap cols = {P, P, 0, 1, P};
  lds_fft(buf, cols, my, a.plan_y, a.nst_y, twy, 1, 1.0f);
  // C2R packing: Z[k] = (X[k] + conj X[h-k]) + i (X[k] - conj X[h-k]) e^{+2 pi i k / mx}
  const int npair = h / 2 + 1;
  for (int e = threadIdx.x; e < my * npair; e += blockDim.x) {
    int y = e / npair, k = e - y * npair;
    float2* row = buf + y * P;
    float2 Xk = row[k], Xh = row[h - k];
    float2 wk = cconj(twx[k]);  // e^{+2 pi i k / mx}
    float2 B = cconj(Xh);
    float2 Zk = cadd(cadd(Xk, B), cmul_si(cmul(csub(Xk, B), wk), 1.0f));
    if (k != 0 && h - k != k) {
      float2 B2 = cconj(Xk);
      float2 wk2 = make_float2(-wk.x, wk.y);  // e^{+2 pi i (h-k)/mx} = -conj(e^{+2 pi i k/mx})
      float2 Zh = cadd(cadd(Xh, B2), cmul_si(cmul(csub(Xh, B2), wk2), 1.0f));
      row[h - k] = Zh;
    }
    row[k] = Zk;
  }
  __syncthreads();
  LineMap rows = {my, my, 0, P, 1};
  lds_fft(buf, rows, h, a.plan_h, a.nst_h, twx, 2, 1.0f);
  const float* fbuf = reinterpret_cast<const float*>(buf);
  const size_t off = ((size_t)c * mz + z) * (size_t)my * mx;
  for (int e = threadIdx.x; e < mx * my; e += blockDim.x) {
    int y = e / mx, x = e - y * mx;
    float val = alpha * fbuf[y * 2 * P + x];
    if (aux) val += beta * aux[off + e];
    out[off + e] = val;
  }
}
FHD_KERNEL_PAIR((), kThreads, k_fft_xy_inv, FftPass, body_fft_xy_inv(a.f, a.spec, a.out, a.aux, a.alpha, a.beta, b))

void launch_fft_xy_fwd(const FftArgs& a, const float* in, float2* spec, hipStream_t s) {
  if (!a.generic && fast_fft_xy_fwd(a, in, spec, s)) return;
  const FftPass x{a, in, spec, nullptr, nullptr, 0.0f, 0.0f, 0};
  launch(k_fft_xy_fwd, k_fft_xy_fwd_b, FAM_FFT_XY_FWD, dim3(a.g.mz, a.ncomp), kThreads, x, s);
}
void launch_fft_z_solve(const FftArgs& a, float2* spec, hipStream_t s) {
  if (!a.generic && fast_fft_z_solve(a, spec, s)) return;
  const FftPass x{a, nullptr, spec, nullptr, nullptr, 0.0f, 0.0f, 0};
  launch(k_fft_z_solve, k_fft_z_solve_b, FAM_FFT_Z, dim3((a.g.P + a.ztile - 1) / a.ztile, a.g.my), kThreads, x, s);
}
void launch_fft_xy_inv(const FftArgs& a, const float2* spec, float* out, const float* aux, float alpha, float beta,
                       hipStream_t s) {
  if (!a.generic && fast_fft_xy_inv(a, spec, out, aux, alpha, beta, s)) return;
  const FftPass x{a, nullptr, const_cast<float2*>(spec), out, aux, alpha, beta, 0};
  launch(k_fft_xy_inv, k_fft_xy_inv_b, FAM_FFT_XY_INV, dim3(a.g.mz, a.ncomp), kThreads, x, s);
}

}  // namespace fhd
