// fft128.hip -- the xy passes of 128 x 128 planes with the transforms in registers.
//
// The Stockham kernels (fft_fast.hip) keep the plane in LDS and sweep it once per radix stage.
// Here each transform runs as two in-register DFTs joined by one exchange:
//   x: a row of 128 reals packed as 64 complex z[n] = x[2n] + i x[2n+1] belongs to a quad of
//      lanes; with n = n1 + 8 n2, lane q holds n1 in {2q, 2q+1} for every n2 (two 16-byte loads
//      per n2 pair), takes the 8-point DFTs over n2, the twiddles W64^(n1 k2), then exchanges
//      through the row's LDS slots so that it holds k2 in {2q, 2q+1} for every n1, and takes
//      the 8-point DFTs over n1: X[k2 + 8 k1].  The real transform follows from X[k] and
//      X[64 - k] (the even/odd split of fft_fast.hip).
//   y: a column of 128 complex belongs to 8 lanes; with y = j + 8 n2, lane j takes the 16-point
//      DFT over n2, twiddles W128^(j k2), one exchange through the column's LDS slots, and two
//      8-point DFTs over n1 = j: Y[k2 + 16 k1] for k2 in {2j, 2j+1}.
// The plane sits in LDS once, between the x and the y transforms, as 64 columns of 128: the
// real columns kx = 0 and kx = 64 share one complex column (their y transforms are the real
// and imaginary parts of one complex transform, separated afterwards).  512 lanes per plane:
// every lane has work in both transforms.  Same layouts, tables and results (to FP32
// rounding) as the Stockham passes; the schedule autotuner times the two and keeps the faster.
#include "fhd_device.h"
#include "fhd_kernels.h"
#include "fft_common.h"

namespace fhd {

namespace {
constexpr int kM = 128;        // plane edge
constexpr int kLS = kM + 1;    // LDS column stride (complex values)
constexpr int kNT128 = 512;

// w^t for the 128-point table tw[k] = e^{-2 pi i k / 128}; s > 0 conjugates (inverse)
FHD_DEV float2 twp(const float2* tw, int t, float s) {
  const float2 w = tw[t & (kM - 1)];
  return s > 0.0f ? make_float2(w.x, -w.y) : w;
}

// x transform (64 complex points) of the row held by a quad: in v[b][n2] = z[2q + b + 8 n2],
// out v[b][k1] = X[2q + b + 8 k1].  The row's LDS slots (slot * kLS + r) carry the exchange.
FHD_DEV void x_transform(float2 v[2][8], float2* buf, const float2* tw, int r, int q, float s) {
#pragma unroll
  for (int b = 0; b < 2; ++b) Dft<8>::run(v[b], s);
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int n1 = 2 * q + b;
#pragma unroll
    for (int k2 = 1; k2 < 8; ++k2) v[b][k2] = cmul(v[b][k2], twp(tw, 2 * n1 * k2, s));
  }
  // M[n1][k2] to slot n1 * 8 + k2; read back k2 in {2q, 2q+1} for every n1
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) buf[((2 * q + b) * 8 + k2) * kLS + r] = v[b][k2];
  __syncthreads();
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) v[b][n1] = buf[(n1 * 8 + 2 * q + b) * kLS + r];
#pragma unroll
  for (int b = 0; b < 2; ++b) Dft<8>::run(v[b], s);
}

// y transform (128 points) of the column held by 8 lanes: in v[n2] = c[j + 8 n2], out
// o[b][k1] = Y[2j + b + 16 k1].  The column's LDS slots (col * kLS + slot) carry the exchange.
FHD_DEV void y_transform(float2 v[16], float2 o[2][8], float2* buf, const float2* tw, int col, int j, float s) {
  Dft<16>::run(v, s);
#pragma unroll
  for (int k2 = 1; k2 < 16; ++k2) v[k2] = cmul(v[k2], twp(tw, j * k2, s));
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) buf[col * kLS + j * 16 + k2] = v[k2];
  __syncthreads();
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) o[b][n1] = buf[col * kLS + n1 * 16 + 2 * j + b];
#pragma unroll
  for (int b = 0; b < 2; ++b) Dft<8>::run(o[b], s);
}
}  // namespace

// ---------------- forward: real plane -> half spectrum spec[c][ky][z][kx] ----------------
FHD_BODY void body_fftA128(const FftArgs& a, const float* __restrict__ in, float2* __restrict__ spec, const Blk& b) {
  constexpr int H = kM / 2, P = H + 1;
  __shared__ float2 buf[H * kLS];
  __shared__ float2 tw[kM];
  const int z = b.x, c = b.y, mz = a.g.mz;
  const int tid = threadIdx.x;
  if (tid < kM) tw[tid] = a.twx[tid];
  // x: row r, quad lane q
  const int r = tid >> 2, q = tid & 3;
  const float4* src = reinterpret_cast<const float4*>(in + ((size_t)c * mz + z) * (size_t)(kM * kM) + (size_t)r * kM);
  float2 v[2][8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const float4 t = src[q + 4 * m];  // z[2q + 8m], z[2q + 8m + 1]
    v[0][m] = make_float2(t.x, t.y);
    v[1][m] = make_float2(t.z, t.w);
  }
  __syncthreads();  // twiddles in LDS
  x_transform(v, buf, tw, r, q, -1.0f);
  __syncthreads();  // exchange slots read
#pragma unroll
  for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) buf[(2 * q + b2 + 8 * k1) * kLS + r] = v[b2][k1];
  __syncthreads();
  // the real transform F[k] = E + w_k O, E = (X[k] + X*[64-k]) / 2, O = -i (X[k] - X*[64-k]) / 2,
  // w_k = e^{-2 pi i k / 128}; F[0] and F[64] (real) share slot 0 as F[0] + i F[64]
  float2 f[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int k = q + 4 * t;
    const float2 A = buf[k * kLS + r];
    const float2 B = cconj(buf[(k == 0 ? 0 : H - k) * kLS + r]);
    const float2 E = cscale(cadd(A, B), 0.5f);
    const float2 D = csub(A, B);
    const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
    f[t] = k == 0 ? make_float2(E.x + O.x, E.x - O.x) : cadd(E, cmul(tw[k], O));
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 16; ++t) buf[(q + 4 * t) * kLS + r] = f[t];
  __syncthreads();
  // y: column kx (0 carries kx 0 and 64), lane j
  const int col = tid >> 3, j = tid & 7;
  float2 cv[16];
#pragma unroll
  for (int n2 = 0; n2 < 16; ++n2) cv[n2] = buf[col * kLS + j + 8 * n2];
  __syncthreads();  // the column is in registers before its slots carry the exchange
  float2 o[2][8];
  y_transform(cv, o, buf, tw, col, j, -1.0f);
  float2* dst = spec + (size_t)c * kM * mz * P + (size_t)z * P;
  if (col != 0) {
#pragma unroll
    for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) dst[(size_t)(2 * j + b2 + 16 * k1) * mz * P + col] = o[b2][k1];
  }
  __syncthreads();  // exchange slots read
  // column 0: C = F0 + i F64 with F0, F64 real along y, so F0^[ky] = (C^[ky] + C^*[-ky]) / 2 and
  // F64^[ky] = (C^[ky] - C^*[-ky]) / (2 i)
  if (col == 0) {
#pragma unroll
    for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) buf[2 * j + b2 + 16 * k1] = o[b2][k1];
  }
  __syncthreads();
  if (col == 0) {
#pragma unroll
    for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) {
        const int ky = 2 * j + b2 + 16 * k1;
        const float2 C0 = o[b2][k1];
        const float2 Cm = cconj(buf[(kM - ky) & (kM - 1)]);
        const float2 F0 = cscale(cadd(C0, Cm), 0.5f);
        const float2 d = csub(C0, Cm);
        const float2 F64 = make_float2(0.5f * d.y, -0.5f * d.x);
        dst[(size_t)ky * mz * P + 0] = F0;
        dst[(size_t)ky * mz * P + H] = F64;
      }
  }
}
FHD_KERNEL_PAIR((), kNT128, k_fftA128, FftPass, body_fftA128(a.f, a.in, a.spec, b))

// ---------------- inverse: half spectrum -> real plane, out = alpha r + beta aux ----------------
FHD_BODY void body_fftC128(const FftArgs& a, const float2* __restrict__ spec, float* __restrict__ out,
                           const float* __restrict__ aux, float alpha, float beta, const Blk& b) {
  constexpr int H = kM / 2, P = H + 1;
  __shared__ float2 buf[H * kLS];
  __shared__ float2 tw[kM];
  const int z = b.x, c = b.y, mz = a.g.mz;
  const int tid = threadIdx.x;
  if (tid < kM) tw[tid] = a.twx[tid];
  // y (inverse): column kx, with kx = 0 carrying F0 + i F64 (both Hermitian in ky, so the
  // real and imaginary parts of the result are the two real columns)
  const int col = tid >> 3, j = tid & 7;
  const float2* src = spec + (size_t)c * kM * mz * P + (size_t)z * P;
  float2 cv[16];
#pragma unroll
  for (int n2 = 0; n2 < 16; ++n2) {
    const size_t row = (size_t)(j + 8 * n2) * mz * P;
    if (col != 0) {
      cv[n2] = src[row + col];
    } else {
      const float2 f0 = src[row], f64 = src[row + H];
      cv[n2] = make_float2(f0.x - f64.y, f0.y + f64.x);  // F0 + i F64
    }
  }
  __syncthreads();  // twiddles in LDS
  float2 o[2][8];
  y_transform(cv, o, buf, tw, col, j, 1.0f);
  __syncthreads();  // exchange slots read
#pragma unroll
  for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) buf[col * kLS + 2 * j + b2 + 16 * k1] = o[b2][k1];
  __syncthreads();
  // x (inverse C2R): row r; X[k] = (Z[k] + Z*[64-k]) + i (Z[k] - Z*[64-k]) e^{+2 pi i k/128},
  // Z[0] and Z[64] the real and imaginary parts of slot 0; lane q builds X at n = 2q + b + 8 n2
  const int r = tid >> 2, q = tid & 3;
  float2 v[2][8];
#pragma unroll
  for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) {
      const int k = 2 * q + b2 + 8 * n2;
      float2 Zk, Zh;
      if (k == 0) {
        const float2 s0 = buf[r];
        Zk = make_float2(s0.x, 0.0f);
        Zh = make_float2(s0.y, 0.0f);
      } else {
        Zk = buf[k * kLS + r];
        Zh = buf[(H - k) * kLS + r];
      }
      const float2 B = cconj(Zh);
      v[b2][n2] = cadd(cadd(Zk, B), cmul_si(cmul(csub(Zk, B), cconj(tw[k])), 1.0f));
    }
  __syncthreads();  // slots read before the exchange reuses them
  x_transform(v, buf, tw, r, q, 1.0f);
  // v[b][k1] = z[2q + b + 8 k1] = (x[2n], x[2n+1]): one float4 per k1
  const size_t off = ((size_t)c * mz + z) * (size_t)(kM * kM) + (size_t)r * kM;
  float4* dst = reinterpret_cast<float4*>(out + off);
  const float4* ax = aux ? reinterpret_cast<const float4*>(aux + off) : nullptr;
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) {
    float4 res = make_float4(alpha * v[0][k1].x, alpha * v[0][k1].y, alpha * v[1][k1].x, alpha * v[1][k1].y);
    if (ax) {
      const float4 t = ax[q + 4 * k1];
      res.x += beta * t.x;
      res.y += beta * t.y;
      res.z += beta * t.z;
      res.w += beta * t.w;
    }
    dst[q + 4 * k1] = res;
  }
}
FHD_KERNEL_PAIR((), kNT128, k_fftC128, FftPass, body_fftC128(a.f, a.spec, a.out, a.aux, a.alpha, a.beta, b))

void launch_fft128_xy(const FftPass& x, int which, int mz, hipStream_t s) {
  if (which == 0) launch(k_fftA128, k_fftA128_b, FAM_FFT_XY_FWD, dim3(mz, x.f.ncomp), kNT128, x, s);
  else launch(k_fftC128, k_fftC128_b, FAM_FFT_XY_INV, dim3(mz, x.f.ncomp), kNT128, x, s);
}

}  // namespace fhd
