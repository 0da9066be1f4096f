// fft_fast.hip -- size-specialised versions of the three FFT passes (see fft.hip for the
// algorithm).  Transform lengths, radix plans and LDS extents are template parameters, so every
// index division in the Stockham stages is by a constant and the register arrays are sized for
// the actual plane instead of the worst case.  Instantiated for xy planes mx x my with both
// lengths in {16, 24, 32, 48, 64} or mx = my in {96, 128}, and z pencils of any of those
// lengths; other shapes use the generic kernels, which share the spectrum layout
// spec[c][ky][z][kx].
#include "fhd_device.h"
#include "fhd_kernels.h"
#include "fft_common.h"


namespace fhd {

// One radix-R Stockham stage, lines l = 0..lines-1 (lines <= MAXL) at base l * LS, element
// stride ES.  Consecutive lanes take butterfly j of consecutive lines: the lines are either
// adjacent elements (ES > 1, LS = 1) or an odd number of elements apart (ES = 1), so the lanes
// of a wave read and write different LDS banks in every stage, and share one twiddle.  (Taking
// consecutive butterflies of one line instead makes the writes of the early stages R-way bank
// conflicts.)
template <int N, int R, int NS, int ES, int NT, int MAXL, int TWM>
FHD_DEV void fstage(float2* buf, int lines, int LS, const float2* tw, float s) {
  constexpr int NBL = N / R;
  constexpr int KMAX = (MAXL * NBL + NT - 1) / NT;
  constexpr int STRIDE = N / (NS * R);
  float2 v[KMAX][R];
  int wb[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int b = threadIdx.x + k * NT;
    const int j = b / MAXL;
    const int l = b - j * MAXL;
    wb[k] = -1;
    if (j < NBL && l < lines) {
      const int base = l * LS;
      const int kk = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = buf[base + (j + r * NBL) * ES];
        if (NS > 1 && r > 0) {
          float2 w = tw[(kk * r * STRIDE) * TWM];
          w.y = s > 0.0f ? -w.y : w.y;
          x = cmul(x, w);
        }
        v[k][r] = x;
      }
      Dft<R>::run(v[k], s);
      wb[k] = base + ((j / NS) * NS * R + kk) * ES;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KMAX; ++k)
    if (wb[k] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[wb[k] + r * NS * ES] = v[k][r];
    }
  __syncthreads();
}

template <int N, int ST, int NS, int ES, int NT, int MAXL, int TWM, uint32_t PLAN>
struct FStages {
  static FHD_DEV void run(float2* buf, int lines, int LS, const float2* tw, float s) {
    constexpr int R = NS >= N ? 1 : plan_at(PLAN, N, ST);
    if constexpr (R > 1) {
      fstage<N, R, NS, ES, NT, MAXL, TWM>(buf, lines, LS, tw, s);
      FStages<N, ST + 1, NS * R, ES, NT, MAXL, TWM, PLAN>::run(buf, lines, LS, tw, s);
    }
  }
};

template <int N, int ES, int NT, int MAXL, int TWM, uint32_t PLAN = 0u>
FHD_DEV void ffft(float2* buf, int lines, int LS, const float2* tw, float s) {
  FStages<N, 0, 1, ES, NT, MAXL, TWM, PLAN>::run(buf, lines, LS, tw, s);
}

// ---------------- pass A: real plane -> half spectrum (x) -> FFT along y ----------------
template <int MX, int MY, int NT, int PX = 0, int PY = 0>
FHD_BODY void body_fftA(const FftArgs& a, const float* __restrict__ in, float2* __restrict__ spec, const Blk& b) {
  constexpr int H = MX / 2, P = H + 1;
  __shared__ float2 buf[MY * P];
  __shared__ float2 twx[MX];
  __shared__ float2 twy[MY];
  const int z = b.x, c = b.y, mz = a.g.mz;
  for (int i = threadIdx.x; i < MX; i += NT) twx[i] = a.twx[i];
  for (int i = threadIdx.x; i < MY; i += NT) twy[i] = a.twy[i];
  const float4* src = reinterpret_cast<const float4*>(in + ((size_t)c * mz + z) * (size_t)(MX * MY));
  constexpr int Q = MX / 4;
  // every load of this lane in flight at once, then into LDS (a load-store loop would wait out
  // one HBM round trip per element)
  constexpr int NL = (MY * Q + NT - 1) / NT;
  float4 v[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int e = threadIdx.x + k * NT;
    if (e < MY * Q) v[k] = src[e];
  }
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int e = threadIdx.x + k * NT;
    if (e < MY * Q) {
      const int y = e / Q, x4 = e - y * Q;
      buf[y * P + 2 * x4] = make_float2(v[k].x, v[k].y);
      buf[y * P + 2 * x4 + 1] = make_float2(v[k].z, v[k].w);
    }
  }
  __syncthreads();
  ffft<H, 1, NT, MY, 2, plan_choice(H, PX)>(buf, MY, P, twx, -1.0f);
  constexpr int NP = H / 2 + 1;
  for (int e = threadIdx.x; e < MY * NP; e += NT) {
    const int y = e / NP, k = e - y * NP;
    float2* row = buf + y * P;
    const float2 A = row[k];
    const float2 B = cconj(row[k == 0 ? 0 : H - k]);
    const float2 E = cscale(cadd(A, B), 0.5f);
    const float2 D = csub(A, B);
    const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
    const float2 wO = cmul(twx[k], O);
    if (k == 0) {
      row[0] = make_float2(E.x + O.x, 0.0f);
      row[H] = make_float2(E.x - O.x, 0.0f);
    } else {
      row[k] = cadd(E, wO);
      if (H - k != k) row[H - k] = cconj(csub(E, wO));
    }
  }
  __syncthreads();
  ffft<MY, P, NT, P, 1, plan_choice(MY, PY)>(buf, P, 1, twy, -1.0f);
  float2* dst = spec + (size_t)c * MY * mz * P + (size_t)z * P;
  for (int e = threadIdx.x; e < MY * P; e += NT) {
    const int ky = e / P, kx = e - ky * P;
    dst[(size_t)ky * mz * P + kx] = buf[e];
  }
}
FHD_KERNEL_PAIR((template <int MX, int MY, int NT, int PX, int PY>), NT, k_fftA, FftPass,
                body_fftA<MX, MY, NT, PX, PY>(a.f, a.in, a.spec, b))

// ---------------- pass C: FFT along y (inverse) -> C2R along x -> out = alpha r + beta aux ----
template <int MX, int MY, int NT, int PX = 0, int PY = 0>
FHD_BODY void body_fftC(const FftArgs& a, const float2* __restrict__ spec, float* __restrict__ out,
                        const float* __restrict__ aux, float alpha, float beta, const Blk& b) {
  constexpr int H = MX / 2, P = H + 1;
  __shared__ float2 buf[MY * P];
  __shared__ float2 twx[MX];
  __shared__ float2 twy[MY];
  const int z = b.x, c = b.y, mz = a.g.mz;
  for (int i = threadIdx.x; i < MX; i += NT) twx[i] = a.twx[i];
  for (int i = threadIdx.x; i < MY; i += NT) twy[i] = a.twy[i];
  const float2* src = spec + (size_t)c * MY * mz * P + (size_t)z * P;
  constexpr int NL = (MY * P + NT - 1) / NT;
  {
    float2 v[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < MY * P) {
        const int ky = e / P, kx = e - ky * P;
        v[k] = src[(size_t)ky * mz * P + kx];
      }
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < MY * P) buf[e] = v[k];
    }
  }
  __syncthreads();
  ffft<MY, P, NT, P, 1, plan_choice(MY, PY)>(buf, P, 1, twy, 1.0f);
  constexpr int NP = H / 2 + 1;
  for (int e = threadIdx.x; e < MY * NP; e += NT) {
    const int y = e / NP, k = e - y * NP;
    float2* row = buf + y * P;
    const float2 Xk = row[k], Xh = row[H - k];
    const float2 wk = cconj(twx[k]);
    const float2 B = cconj(Xh);
    const float2 Zk = cadd(cadd(Xk, B), cmul_si(cmul(csub(Xk, B), wk), 1.0f));
    if (k != 0 && H - k != k) {
      const float2 B2 = cconj(Xk);
      const float2 wk2 = make_float2(-wk.x, wk.y);
      row[H - k] = cadd(cadd(Xh, B2), cmul_si(cmul(csub(Xh, B2), wk2), 1.0f));
    }
    row[k] = Zk;
  }
  __syncthreads();
  ffft<H, 1, NT, MY, 2, plan_choice(H, PX)>(buf, MY, P, twx, 1.0f);
  const size_t off = ((size_t)c * mz + z) * (size_t)(MX * MY);
  float4* dst = reinterpret_cast<float4*>(out + off);
  const float4* ax = aux ? reinterpret_cast<const float4*>(aux + off) : nullptr;
  constexpr int Q = MX / 4;
  constexpr int NS = (MY * Q + NT - 1) / NT;
  float4 q[NS];
  if (ax) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int e = threadIdx.x + k * NT;
      if (e < MY * Q) q[k] = ax[e];
    }
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int e = threadIdx.x + k * NT;
    if (e < MY * Q) {
      const int y = e / Q, x4 = e - y * Q;
      const float2 p0 = buf[y * P + 2 * x4], p1 = buf[y * P + 2 * x4 + 1];
      float4 r = make_float4(alpha * p0.x, alpha * p0.y, alpha * p1.x, alpha * p1.y);
      if (ax) {
        r.x += beta * q[k].x;
        r.y += beta * q[k].y;
        r.z += beta * q[k].z;
        r.w += beta * q[k].w;
      }
      dst[e] = r;
    }
  }
}
FHD_KERNEL_PAIR((template <int MX, int MY, int NT, int PX, int PY>), NT, k_fftC, FftPass,
                body_fftC<MX, MY, NT, PX, PY>(a.f, a.spec, a.out, a.aux, a.alpha, a.beta, b))

// ---------------- pass B: FFT along z, operator, inverse FFT along z ----------------
// Workgroup (tile i, ky): kx columns i*T .. i*T+T-1, plus column ntile*T + i when P is not a
// multiple of T (the Nyquist column for P = mx/2 + 1).  LDS lines: (c, t) at (c*(T+1)+t)*(MZ+1).
template <int MZ, int NT, int TM, int PZ = 0>
FHD_BODY void body_fftB(const FftArgs& a, float2* __restrict__ spec, int T, const Blk& b) {
  // TM: room for T <= TM - 1 columns plus one extra column
  constexpr int LS = MZ + 1;
  __shared__ float2 buf[3 * TM * LS];
  __shared__ float2 twz[MZ], ezs[MZ];
  __shared__ float gzs[MZ];
  const int P = a.g.P, my = a.g.my;
  // neighbouring column tiles share the 128-byte lines of each z row: keep them on one XCD
  const int item = xcd_item(b.x + b.nx * b.y, b.nx * b.ny);
  const int ky = item / b.nx, tile = item - ky * b.nx;
  const int ntile = P / T, rem = P - ntile * T;
  const int TC = T + (tile < rem ? 1 : 0);
  for (int i = threadIdx.x; i < MZ; i += NT) {
    twz[i] = a.twz[i];
    ezs[i] = a.ez[i];
    gzs[i] = a.gz[i];
  }
  const size_t cstride = (size_t)my * MZ * P;
  float2* base = spec + (size_t)ky * MZ * P;
  // elements (c, z, t) on a TM-wide grid (t fastest, constant divisors); slots t >= TC idle.
  // Every load of this lane is in flight before the first LDS write.
  constexpr int EL = 3 * MZ * TM, NL = (EL + NT - 1) / NT;
  const int ncomp = a.ncomp;
  const int cs32 = (int)cstride;  // ncomp * cstride elements per ky row: far below 2^31 at every size
  {
    float2 v[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int e = threadIdx.x + k * NT;
      const int t = e % TM, r = e / TM, z = r % MZ, c = r / MZ;
      if (e < EL && c < ncomp && t < TC) {
        const int kx = t < T ? tile * T + t : ntile * T + tile;
        v[k] = base[c * cs32 + z * P + kx];
      }
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int e = threadIdx.x + k * NT;
      const int t = e % TM, r = e / TM, z = r % MZ, c = r / MZ;
      if (e < EL && c < ncomp && t < TC) buf[(c * TM + t) * LS + z] = v[k];
    }
  }
  __syncthreads();
  const int lines = a.ncomp * TM;
  if (a.zmode != 2) ffft<MZ, 1, NT, 3 * TM, 1, plan_choice(MZ, PZ)>(buf, lines, LS, twz, -1.0f);
  if (a.zmode == 0) {
    const float gy = a.gy[ky];
    const float2 ey = a.ey[ky];
    const Key key = a.knoise ? resolve_key(a.key) : a.key;
    for (int e = threadIdx.x; e < MZ * TC; e += NT) {
      const int t = e / MZ, z = e - t * MZ;
      const int kx = t < T ? tile * T + t : ntile * T + tile;
      const float gx = a.gx[kx];
      const float2 ex = a.ex[kx];
      const float gz = gzs[z];
      const float2 ez = ezs[z];
      const int i0 = (0 * TM + t) * LS + z, i1 = (1 * TM + t) * LS + z, i2 = (2 * TM + t) * LS + z;
      float2 b0 = buf[i0], b1 = buf[i1], b2 = buf[i2];
      if (a.knoise) {
        float2 nw[3];
        kspace_noise(key, kx, ky, z, a.g.mx, my, MZ, a.knoise, a.kn_f4, a.kn_sqrt_half_n, gx, gy,
                          gz, ex, ey, ez, nw);
        b0 = cadd(b0, nw[0]);
        b1 = cadd(b1, nw[1]);
        b2 = cadd(b2, nw[2]);
      }
      const float2 w0 = cmul(b0, cconj(ex));
      const float2 w1 = cmul(b1, cconj(ey));
      const float2 w2 = cmul(b2, cconj(ez));
      const float L = -(gx * gx + gy * gy + gz * gz);
      float2 o0, o1, o2;
      if (kx == 0 && ky == 0 && z == 0) {
        const float keep = a.stokes ? 0.0f : a.inv_n;
        o0 = cscale(w0, keep);
        o1 = cscale(w1, keep);
        o2 = cscale(w2, keep);
      } else {
        const float2 GW = make_float2(gx * w0.x + gy * w1.x + gz * w2.x, gx * w0.y + gy * w1.y + gz * w2.y);
        const float den = a.stokes ? (-a.eta * L) : (1.0f - a.cn_coef * L);
        const float sc = a.inv_n / den;
        const float iL = 1.0f / L;
        o0 = cscale(make_float2(w0.x + gx * GW.x * iL, w0.y + gx * GW.y * iL), sc);
        o1 = cscale(make_float2(w1.x + gy * GW.x * iL, w1.y + gy * GW.y * iL), sc);
        o2 = cscale(make_float2(w2.x + gz * GW.x * iL, w2.y + gz * GW.y * iL), sc);
      }
      buf[i0] = cmul(o0, ex);
      buf[i1] = cmul(o1, ey);
      buf[i2] = cmul(o2, ez);
    }
    __syncthreads();
  }
  if (a.zmode != 1) ffft<MZ, 1, NT, 3 * TM, 1, plan_choice(MZ, PZ)>(buf, lines, LS, twz, 1.0f);
  // the lane index, opaque to the compiler: the store indices are recomputed here rather than
  // kept live in registers from the loads through both transforms
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int e = tid + k * NT;
    const int t = e % TM, r = e / TM, z = r % MZ, c = r / MZ;
    if (e < EL && c < ncomp && t < TC) {
      const int kx = t < T ? tile * T + t : ntile * T + tile;
      base[c * cs32 + z * P + kx] = buf[(c * TM + t) * LS + z];
    }
  }
}
// 8 waves per SIMD (a 64-VGPR budget) where the LDS of that many workgroups fits in a CU
constexpr int fftb_waves(int MZ, int NT, int TM) {
  return (32 / (NT / 64)) * (3 * TM * (MZ + 1) * 8 + MZ * 20) <= 160 * 1024 ? 8 : 1;
}
FHD_KERNEL_PAIR_OCC((template <int MZ, int NT, int TM, int PZ>), NT, fftb_waves(MZ, NT, TM), k_fftB, FftPass,
                body_fftB<MZ, NT, TM, PZ>(a.f, a.spec, a.T, b))

// ---------------- dispatch ----------------
template <int MX, int MY, int NT, int PX, int PY>
static void launch_AC_plan(const FftPass& x, int which, int mz, hipStream_t s) {
  if (which == 0)
    launch(k_fftA<MX, MY, NT, PX, PY>, k_fftA_b<MX, MY, NT, PX, PY>, FAM_FFT_XY_FWD, dim3(mz, x.f.ncomp), NT, x, s);
  else
    launch(k_fftC<MX, MY, NT, PX, PY>, k_fftC_b<MX, MY, NT, PX, PY>, FAM_FFT_XY_INV, dim3(mz, x.f.ncomp), NT, x, s);
}
// The radix plans of the x and y transforms (plan_choice indices, px + 3 py) come from the
// schedule autotuner for 96 x 96 and 128 x 128 planes, one for the forward and one for the
// inverse pass (128 x 128 also has 9: the register-transform kernels of fft128.hip); other
// planes use the default plan.
template <int MX, int MY, int NT>
static void launch_AC_nt(const FftPass& x, int which, int mz, hipStream_t s) {
  if constexpr ((MX == 128 && MY == 128) || (MX == 96 && MY == 96)) {
    switch (which == 0 ? x.f.plan_xy : x.f.plan_xy_inv) {
      case 1: launch_AC_plan<MX, MY, NT, 1, 0>(x, which, mz, s); return;
      case 2: launch_AC_plan<MX, MY, NT, 2, 0>(x, which, mz, s); return;
      case 3: launch_AC_plan<MX, MY, NT, 0, 1>(x, which, mz, s); return;
      case 4: launch_AC_plan<MX, MY, NT, 1, 1>(x, which, mz, s); return;
      case 5: launch_AC_plan<MX, MY, NT, 2, 1>(x, which, mz, s); return;
      case 6: launch_AC_plan<MX, MY, NT, 0, 2>(x, which, mz, s); return;
      case 7: launch_AC_plan<MX, MY, NT, 1, 2>(x, which, mz, s); return;
      case 8: launch_AC_plan<MX, MY, NT, 2, 2>(x, which, mz, s); return;
      case 9:
        if constexpr (MX == 128) {
          launch_fft128_xy(x, which, mz, s);  // transforms in registers (fft128.hip)
          return;
        }
        break;
      default: break;
    }
  }
  launch_AC_plan<MX, MY, NT, 0, 0>(x, which, mz, s);
}
// Lanes per plane workgroup.  A single context's pass has only 3 mz plane workgroups, and more
// lanes per plane cover more of the plane load and store latency: 1024 from 64^2 up (128^3:
// 25.4 vs 26.1 / 27.3 us per pass with 512; 58-60 VGPRs, two workgroups per CU by LDS).  A
// batched dispatch already holds the planes of many contexts, and below 128^2 those pack
// better as 512-lane workgroups (ensemble step 2.30 vs 2.41 ms).
template <int MX, int MY>
static bool launch_AC(const FftArgs& a, int which, const float* in, float2* spec, float* out, const float* aux,
                      float alpha, float beta, hipStream_t s) {
  const FftPass x{a, in, spec, out, aux, alpha, beta, 0};
  if constexpr (MX * MY >= 64 * 64) {
    if (MX * MY >= 128 * 128 || t_recorder == nullptr) {
      launch_AC_nt<MX, MY, 1024>(x, which, a.g.mz, s);
      return true;
    }
    launch_AC_nt<MX, MY, 512>(x, which, a.g.mz, s);
  } else {
    launch_AC_nt<MX, MY, 256>(x, which, a.g.mz, s);
  }
  return true;
}

// Rectangular planes mx x my with both lengths in {16, 24, 32, 48, 64}.
template <int MX>
static bool launch_AC_rect(const FftArgs& a, int which, const float* in, float2* spec, float* out, const float* aux,
                           float alpha, float beta, hipStream_t s) {
  switch (a.g.my) {
    case 16: return launch_AC<MX, 16>(a, which, in, spec, out, aux, alpha, beta, s);
    case 24: return launch_AC<MX, 24>(a, which, in, spec, out, aux, alpha, beta, s);
    case 32: return launch_AC<MX, 32>(a, which, in, spec, out, aux, alpha, beta, s);
    case 48: return launch_AC<MX, 48>(a, which, in, spec, out, aux, alpha, beta, s);
    case 64: return launch_AC<MX, 64>(a, which, in, spec, out, aux, alpha, beta, s);
    default: return false;
  }
}

static bool fast_AC(const FftArgs& a, int which, const float* in, float2* spec, float* out, const float* aux,
                    float alpha, float beta, hipStream_t s) {
  if (a.g.mx == a.g.my) {
    if (a.g.mx == 96) return launch_AC<96, 96>(a, which, in, spec, out, aux, alpha, beta, s);
    if (a.g.mx == 128) return launch_AC<128, 128>(a, which, in, spec, out, aux, alpha, beta, s);
  }
  switch (a.g.mx) {
    case 16: return launch_AC_rect<16>(a, which, in, spec, out, aux, alpha, beta, s);
    case 24: return launch_AC_rect<24>(a, which, in, spec, out, aux, alpha, beta, s);
    case 32: return launch_AC_rect<32>(a, which, in, spec, out, aux, alpha, beta, s);
    case 48: return launch_AC_rect<48>(a, which, in, spec, out, aux, alpha, beta, s);
    case 64: return launch_AC_rect<64>(a, which, in, spec, out, aux, alpha, beta, s);
    default: return false;
  }
}

bool fast_fft_xy_fwd(const FftArgs& a, const float* in, float2* spec, hipStream_t s) {
  return fast_AC(a, 0, in, spec, nullptr, nullptr, 0.0f, 0.0f, s);
}
bool fast_fft_xy_inv(const FftArgs& a, const float2* spec, float* out, const float* aux, float alpha, float beta,
                     hipStream_t s) {
  return fast_AC(a, 1, nullptr, const_cast<float2*>(spec), out, aux, alpha, beta, s);
}

// Largest column-tile width T <= tmax whose leftover columns (P mod T) fit one per tile.
static int column_tile(int P, int tmax) {
  for (int T = tmax; T > 1; --T)
    if (T <= P && P % T <= P / T) return T;
  return 1;
}

template <int MZ, int NT, int PZ>
static void launch_B_plan(const FftPass& x, int bt, dim3 grid, hipStream_t s) {
  if (bt == 2) launch(k_fftB<MZ, NT, 3, PZ>, k_fftB_b<MZ, NT, 3, PZ>, FAM_FFT_Z, grid, NT, x, s);
  else if (bt == 4) launch(k_fftB<MZ, NT, 5, PZ>, k_fftB_b<MZ, NT, 5, PZ>, FAM_FFT_Z, grid, NT, x, s);
  else if (bt == 8) launch(k_fftB<MZ, NT, 9, PZ>, k_fftB_b<MZ, NT, 9, PZ>, FAM_FFT_Z, grid, NT, x, s);
  else launch(k_fftB<MZ, NT, 17, PZ>, k_fftB_b<MZ, NT, 17, PZ>, FAM_FFT_Z, grid, NT, x, s);
}
// The z transform's radix plan (plan_choice index) comes from the autotuner for 48-, 96- and
// 128-point columns.
template <int MZ, int NT>
static void launch_B_nt(const FftPass& x, int bt, dim3 grid, hipStream_t s) {
  if constexpr (MZ == 48 || MZ == 96 || MZ == 128) {
    if (x.f.plan_zs == 1) return launch_B_plan<MZ, NT, 1>(x, bt, grid, s);
    if (x.f.plan_zs == 2) return launch_B_plan<MZ, NT, 2>(x, bt, grid, s);
  }
  launch_B_plan<MZ, NT, 0>(x, bt, grid, s);
}

// Column-tile width of the z pass: 2, 4, 8 or 16 (chosen per grid shape by the context's
// schedule), with one extra LDS column for the leftover Nyquist column.
template <int MZ>
static bool launch_B2(const FftArgs& a, float2* spec, hipStream_t s) {
  const int P = a.g.P;
  const int bt = a.btile == 2 || a.btile == 4 || a.btile == 16 ? a.btile : 8;
  const int T = column_tile(P, bt);
  const dim3 grid(P / T, a.g.my);
  const FftPass x{a, nullptr, spec, nullptr, nullptr, 0.0f, 0.0f, T};
  // 512 lanes halve each lane's column loads for a single context's pass from 64-point columns
  // up (128^3: 29.6 vs 30.5 us per pass; 50 VGPRs); batched dispatches below 128 keep 256
  if constexpr (MZ >= 64) {
    if (MZ >= 128 || t_recorder == nullptr) {
      launch_B_nt<MZ, 512>(x, bt, grid, s);
      return true;
    }
  }
  launch_B_nt<MZ, 256>(x, bt, grid, s);
  return true;
}

bool fast_fft_z_solve(const FftArgs& a, float2* spec, hipStream_t s) {
  switch (a.g.mz) {
    case 16: return launch_B2<16>(a, spec, s);
    case 24: return launch_B2<24>(a, spec, s);
    case 32: return launch_B2<32>(a, spec, s);
    case 48: return launch_B2<48>(a, spec, s);
    case 64: return launch_B2<64>(a, spec, s);
    case 96: return launch_B2<96>(a, spec, s);
    case 128: return launch_B2<128>(a, spec, s);
    default: return false;
  }
}

}  // namespace fhd
