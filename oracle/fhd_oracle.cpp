// fhd_oracle.cpp -- FP64 CPU restatement of fluam's incompressible (schemeIncompressible),
// inertial-coupling (schemeQuasiNeutrallyBuoyant, with and without excess mass) and
// Stokes-limit (schemeStokesLimit, midpoint) time steps.
//
// TEST INFRASTRUCTURE: the checker for libfhd.so.  Loaded only by tests/, the smoke check
// and the benchmark's CPU-baseline leg.  Never linked into the product.
//
// Each function cites the fluam source it restates.  Deliberate differences from fluam
// (all shared with the GPU library, so parity compares like with like):
//   * noise: counter-based Philox4x32-10 + Box-Muller instead of cuRAND XORWOW
//     (src/randomNumbers.cu:25-40); same slot meaning (6 stress components per cell).  The
//     Stokes step draws the Fourier transform of a white stress field instead (same law).
//   * kernels: closed-form Peskin kernels in FP64 (fluam evaluates sqrtf in its 3-point
//     delta, src/deltaGPU.cu:57-68, and uses float textures for the QNB drift and the
//     WCA force, src/deltaGPU.cu:74-105, src/initForcesNonBonded.cu:35-66).
//   * WCA with sigma/epsilon as parameters (fluam: sigma = 2 dx, epsilon = kT).
//   * the RFD drift weights at the displaced position by default (rfd_mode).
#include "fhd_oracle.h"

#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <complex>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cplx;

// ------------------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10 (Salmon et al., SC'11), Box-Muller (1958).
// ------------------------------------------------------------------------------------
static inline void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

extern "C" void orc_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Draw = (seed, step, stream, sub, index): counter {index, stream<<4 | sub, step_lo, step_hi}.
static inline void box_muller(uint32_t a, uint32_t b, double* z0, double* z1) {
  const double inv24 = 1.0 / 16777216.0;
  double u1 = (double)((a >> 8) + 1u) * inv24;  // (0,1]
  double u2 = (double)(b >> 8) * inv24;         // [0,1)
  double r = sqrt(-2.0 * log(u1));
  double s, c;
  s = sin(2.0 * M_PI * u2);
  c = cos(2.0 * M_PI * u2);
  *z0 = r * c;
  *z1 = r * s;
}

extern "C" void orc_normals4(uint64_t seed, int64_t step, int stream, int sub, uint32_t index,
                             double out[4]) {
  uint32_t ctr[4] = {index, ((uint32_t)stream << 4) | (uint32_t)sub, (uint32_t)(uint64_t)step,
                     (uint32_t)((uint64_t)step >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox(ctr, key, w);
  box_muller(w[0], w[1], &out[0], &out[1]);
  box_muller(w[2], w[3], &out[2], &out[3]);
}

static void uniforms4(uint64_t seed, int64_t step, int stream, int sub, uint32_t index, double out[4]) {
  uint32_t ctr[4] = {index, ((uint32_t)stream << 4) | (uint32_t)sub, (uint32_t)(uint64_t)step,
                     (uint32_t)((uint64_t)step >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox(ctr, key, w);
  for (int k = 0; k < 4; ++k) out[k] = (double)(w[k] >> 8) * (1.0 / 16777216.0);
}

enum { STREAM_FLUID1 = 1, STREAM_FLUID_K = 2, STREAM_RFD = 3, STREAM_PART1 = 4, STREAM_PART2 = 5,
       STREAM_FLUID_P = 6, STREAM_INIT_POS = 8, STREAM_INIT_FLUID = 9 };

// Stress noise of one cell and step: 12 normals n_k, k = 0..11, where n_k is word k%4 of the
// Philox draw with sub = k/4 (stream STREAM_FLUID1).  Sigma_1 = n_0..n_5 and Sigma_2 = n_6..n_11
// in the component order of fluam's d_rand slots (src/kernelConstructW.cu:148-198):
// XX=0, XY=1, XZ=2, YY=3, YZ=4, ZZ=5.  which = 0 returns Sigma_1, which = 1 returns Sigma_2.
static void fluid_noise(uint64_t seed, int64_t step, int which, int N, std::vector<double>& Z) {
  Z.assign((size_t)6 * N, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    double n[12];
    for (int sub = 0; sub < 3; ++sub) orc_normals4(seed, step, STREAM_FLUID1, sub, (uint32_t)i, &n[4 * sub]);
    for (int c = 0; c < 6; ++c) Z[(size_t)c * N + i] = n[6 * which + c];
  }
}

// ------------------------------------------------------------------------------------
// FFT: mixed-radix Stockham autosort (radices 8,4,2,3,5, fallback direct DFT), unnormalised.
// Replaces cufftExecZ2Z on a cufftPlan3d(mz,my,mx) plan (src/runSchemeIncompressible.cu:52).
// ------------------------------------------------------------------------------------
struct Fft1 {
  int n = 0;
  std::vector<int> radix;
  std::vector<cplx> tw;  // e^{-2 pi i k / n}
  void init(int n_) {
    n = n_;
    radix.clear();
    int m = n;
    const int rs[5] = {8, 4, 2, 3, 5};
    for (int r : rs)
      while (m % r == 0) { radix.push_back(r); m /= r; }
    if (m != 1) radix.push_back(m);  // direct DFT stage for any remaining factor
    tw.resize(n);
    for (int k = 0; k < n; ++k) {
      double a = -2.0 * M_PI * (double)k / (double)n;
      tw[k] = cplx(cos(a), sin(a));
    }
  }
  // e^{sign 2 pi i k / n}
  inline cplx w(long k, int sign) const {
    long kk = ((k % n) + n) % n;
    return sign < 0 ? tw[kk] : std::conj(tw[kk]);
  }
  void run(cplx* a, cplx* tmp, int sign) const {
    cplx* in = a;
    cplx* out = tmp;
    int Ns = 1;
    std::vector<cplx> v(64), V(64);
    for (int R : radix) {
      if ((int)v.size() < R) { v.resize(R); V.resize(R); }
      int nb = n / R;
      for (int j = 0; j < nb; ++j) {
        int k = j % Ns;
        for (int r = 0; r < R; ++r) v[r] = in[j + r * nb] * w((long)k * r * (n / (Ns * R)), sign);
        for (int t = 0; t < R; ++t) {
          cplx s = 0;
          for (int r = 0; r < R; ++r) s += v[r] * w((long)r * t * (n / R), sign);
          V[t] = s;
        }
        int base = (j / Ns) * Ns * R + k;
        for (int r = 0; r < R; ++r) out[base + r * Ns] = V[r];
      }
      std::swap(in, out);
      Ns *= R;
    }
    if (in != a) memcpy(a, in, sizeof(cplx) * n);
  }
};

struct Fft3 {
  int mx = 0, my = 0, mz = 0;
  Fft1 fx, fy, fz;
  void init(int a, int b, int c) { mx = a; my = b; mz = c; fx.init(a); fy.init(b); fz.init(c); }
  void run(cplx* d, int sign) const {
    // x lines
#pragma omp parallel
    {
      int mmax = std::max(mx, std::max(my, mz));
      std::vector<cplx> line(mmax), tmp(mmax);
#pragma omp for schedule(static)
      for (int l = 0; l < my * mz; ++l) {
        cplx* p = d + (size_t)l * mx;
        fx.run(p, tmp.data(), sign);
      }
#pragma omp for schedule(static)
      for (int l = 0; l < mx * mz; ++l) {
        int x = l % mx, z = l / mx;
        for (int y = 0; y < my; ++y) line[y] = d[(size_t)x + (size_t)mx * (y + (size_t)my * z)];
        fy.run(line.data(), tmp.data(), sign);
        for (int y = 0; y < my; ++y) d[(size_t)x + (size_t)mx * (y + (size_t)my * z)] = line[y];
      }
#pragma omp for schedule(static)
      for (int l = 0; l < mx * my; ++l) {
        for (int z = 0; z < mz; ++z) line[z] = d[(size_t)l + (size_t)mx * my * z];
        fz.run(line.data(), tmp.data(), sign);
        for (int z = 0; z < mz; ++z) d[(size_t)l + (size_t)mx * my * z] = line[z];
      }
    }
  }
};

extern "C" int orc_fft3d(int mx, int my, int mz, double* data, int sign) {
  if (mx <= 0 || my <= 0 || mz <= 0) return 0;
  Fft3 f;
  f.init(mx, my, mz);
  f.run(reinterpret_cast<cplx*>(data), sign);
  return 1;
}

// ------------------------------------------------------------------------------------
// Peskin kernels.
// ------------------------------------------------------------------------------------
// 3-point Roma-Peskin kernel, src/deltaGPU.cu:57-68 (r in grid units).
extern "C" double orc_delta3(double r) {
  double x = fabs(r);
  if (x > 1.5) return 0.0;
  if (x < 0.5) return (1.0 + sqrt(1.0 - 3.0 * x * x)) / 3.0;
  return (5.0 - 3.0 * x - sqrt(1.0 - 3.0 * (1.0 - x) * (1.0 - x))) / 6.0;
}
// d(phi3)/dr; fluam's functionDeltaDerived (src/nonBondedForceExtraPressure.cu:21-33) returns
// the negative of this.
extern "C" double orc_ddelta3(double r) {
  double x = fabs(r);
  if (x >= 1.5) return 0.0;
  if (x <= 0.5) return -r / sqrt(1.0 - 3.0 * r * r);
  double s = (r > 0) - (r < 0);
  return -s * 0.5 * (1.0 + (1.0 - x) / sqrt(-2.0 + 6.0 * x - 3.0 * x * x));
}
// 4-point Peskin kernel (Peskin 2002), the 3D tensor form of src/deltaGPU.cu:44-53.
extern "C" double orc_delta4(double r) {
  double x = fabs(r);
  if (x >= 2.0) return 0.0;
  if (x <= 1.0) return (3.0 - 2.0 * x + sqrt(1.0 + 4.0 * x - 4.0 * x * x)) / 8.0;
  return (5.0 - 2.0 * x - sqrt(-7.0 + 12.0 * x - 4.0 * x * x)) / 8.0;
}
extern "C" double orc_ddelta4(double r) {
  double x = fabs(r);
  double s = (r > 0) - (r < 0);
  if (x >= 2.0) return 0.0;
  if (x <= 1.0) return s * (-2.0 + (4.0 - 8.0 * x) / (2.0 * sqrt(1.0 + 4.0 * x - 4.0 * x * x))) / 8.0;
  return s * (-2.0 - (12.0 - 8.0 * x) / (2.0 * sqrt(-7.0 + 12.0 * x - 4.0 * x * x))) / 8.0;
}

// ------------------------------------------------------------------------------------
// Context.
// ------------------------------------------------------------------------------------
struct orc_ctx {
  fhd_params p;
  int mx, my, mz, N;
  double dx, dy, dz, dV;
  Fft3 fft;
  std::vector<double> v[3];
  std::vector<double> q[3];
  std::vector<double> F[3];     // last particle forces
  std::vector<double> u[3];     // last particle velocities (QNB: J v~)
  std::vector<double> advPrev[3];  // QNB Adams-Bashforth history
  std::vector<double> dulag[3];    // QNB, excess mass: lagged 1/2 dt nu J L Delta v of the last step
  bool first_step_done = false;
  int64_t step = 0;
  bool projected = false;
  // harmonic bonds (bondedForces file), in file order
  struct PairBond { int i, j; double k, r0; };
  struct FixedBond { int i; double k, r0, x[3]; };
  std::vector<PairBond> pair_bonds;
  std::vector<FixedBond> fixed_bonds;
};

static inline int wrapi(int i, int m) { i %= m; return i < 0 ? i + m : i; }
static inline int cid(const orc_ctx* c, int x, int y, int z) {
  return wrapi(x, c->mx) + c->mx * (wrapi(y, c->my) + c->my * wrapi(z, c->mz));
}
static inline double wrapd(double r, double L) { return r - L * floor(r / L + 0.5); }

// One step's particle displacement along axis a.  A displacement of 2^24 cells or more (a
// blown-up step, e.g. two particles started on top of each other under the WCA core) leaves
// that coordinate where it was, as the GPU path does: its FP32 cell offsets cannot carry it.
static inline double step_disp(const orc_ctx* c, int a, double d) {
  const double h = a == 0 ? c->dx : a == 1 ? c->dy : c->dz;
  return fabs(d) < 16777216.0 * h ? d : 0.0;
}

// ---- fluid stencils (src/kernelConstructW.cu:32-259) ----
// Laplacian term 0.5 dt (eta/rho) L v, component a at cell i.
static void lap_term(const orc_ctx* c, const std::vector<double>* v, std::vector<double>* out) {
  const double ax = 1.0 / (c->dx * c->dx), ay = 1.0 / (c->dy * c->dy), az = 1.0 / (c->dz * c->dz);
  const double pre = 0.5 * c->p.dt * (c->p.shearviscosity / c->p.densfluid);
  for (int a = 0; a < 3; ++a) out[a].assign(c->N, 0.0);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < c->mz; ++z)
    for (int y = 0; y < c->my; ++y)
      for (int x = 0; x < c->mx; ++x) {
        int i = cid(c, x, y, z);
        for (int a = 0; a < 3; ++a) {
          const std::vector<double>& f = v[a];
          double L = ax * (f[cid(c, x + 1, y, z)] - 2 * f[i] + f[cid(c, x - 1, y, z)]) +
                     ay * (f[cid(c, x, y + 1, z)] - 2 * f[i] + f[cid(c, x, y - 1, z)]) +
                     az * (f[cid(c, x, y, z + 1)] - 2 * f[i] + f[cid(c, x, y, z - 1)]);
          out[a][i] = pre * L;
        }
      }
}

// Advection A(v) = (dt/4) sum_b D_b[(v_a + v_a')(v_b + v_b')] (src/kernelConstructW.cu:120-136).
static void adv_term(const orc_ctx* c, const std::vector<double>* v, std::vector<double>* A) {
  const double ix = 1.0 / c->dx, iy = 1.0 / c->dy, iz = 1.0 / c->dz;
  for (int a = 0; a < 3; ++a) A[a].assign(c->N, 0.0);
  const std::vector<double>&VX = v[0], &VY = v[1], &VZ = v[2];
#pragma omp parallel for schedule(static)
  for (int z = 0; z < c->mz; ++z)
    for (int y = 0; y < c->my; ++y)
      for (int x = 0; x < c->mx; ++x) {
        int i = cid(c, x, y, z);
        int n0 = cid(c, x, y, z - 1), n1 = cid(c, x, y - 1, z), n2 = cid(c, x - 1, y, z);
        int n3 = cid(c, x + 1, y, z), n4 = cid(c, x, y + 1, z), n5 = cid(c, x, y, z + 1);
        int pxmy = cid(c, x + 1, y - 1, z), pxmz = cid(c, x + 1, y, z - 1);
        int mxpy = cid(c, x - 1, y + 1, z), mxpz = cid(c, x - 1, y, z + 1);
        int pymz = cid(c, x, y + 1, z - 1), mypz = cid(c, x, y - 1, z + 1);
        double vx = VX[i], vy = VY[i], vz = VZ[i];
        double ax = ix * ((VX[n3] + vx) * (VX[n3] + vx) - (vx + VX[n2]) * (vx + VX[n2])) +
                    iy * ((VX[n4] + vx) * (VY[n3] + vy) - (vx + VX[n1]) * (VY[pxmy] + VY[n1])) +
                    iz * ((VX[n5] + vx) * (VZ[n3] + vz) - (vx + VX[n0]) * (VZ[pxmz] + VZ[n0]));
        double ay = ix * ((VY[n3] + vy) * (VX[n4] + vx) - (vy + VY[n2]) * (VX[mxpy] + VX[n2])) +
                    iy * ((VY[n4] + vy) * (VY[n4] + vy) - (vy + VY[n1]) * (vy + VY[n1])) +
                    iz * ((VY[n5] + vy) * (VZ[n4] + vz) - (vy + VY[n0]) * (VZ[pymz] + VZ[n0]));
        double az = ix * ((VZ[n3] + vz) * (VX[n5] + vx) - (vz + VZ[n2]) * (VX[mxpz] + VX[n2])) +
                    iy * ((VZ[n4] + vz) * (VY[n5] + vy) - (vz + VZ[n1]) * (VY[mypz] + VY[n1])) +
                    iz * ((VZ[n5] + vz) * (VZ[n5] + vz) - (vz + VZ[n0]) * (vz + VZ[n0]));
        A[0][i] = 0.25 * c->p.dt * ax;
        A[1][i] = 0.25 * c->p.dt * ay;
        A[2][i] = 0.25 * c->p.dt * az;
      }
}

// Divergence of the stochastic stress: out += scale * D.Sigma(Z).
// trace_free: diagonal (Z_aa - tr/3) (incompressible / QNB, src/kernelConstructW.cu:148-198);
// otherwise diagonal Z_aa (Stokes limit, src/stokesLimitFunctions.cu:2318-2366).
static void noise_div(const orc_ctx* c, const std::vector<double>& Z, double fact1, double fact4,
                      bool trace_free, std::vector<double>* out) {
  const int N = c->N;
  const double ix = 1.0 / c->dx, iy = 1.0 / c->dy, iz = 1.0 / c->dz;
  auto diag = [&](int comp, int j) {
    double d = Z[(size_t)comp * N + j];
    if (trace_free) d -= (Z[j] + Z[(size_t)3 * N + j] + Z[(size_t)5 * N + j]) / 3.0;
    return d;
  };
#pragma omp parallel for schedule(static)
  for (int z = 0; z < c->mz; ++z)
    for (int y = 0; y < c->my; ++y)
      for (int x = 0; x < c->mx; ++x) {
        int i = cid(c, x, y, z);
        int n0 = cid(c, x, y, z - 1), n1 = cid(c, x, y - 1, z), n2 = cid(c, x - 1, y, z);
        int n3 = cid(c, x + 1, y, z), n4 = cid(c, x, y + 1, z), n5 = cid(c, x, y, z + 1);
        const double* XY = &Z[(size_t)1 * N];
        const double* XZ = &Z[(size_t)2 * N];
        const double* YZ = &Z[(size_t)4 * N];
        double wx = ix * fact1 * (diag(0, n3) - diag(0, i)) + iy * fact4 * (XY[i] - XY[n1]) +
                    iz * fact4 * (XZ[i] - XZ[n0]);
        double wy = iy * fact1 * (diag(3, n4) - diag(3, i)) + ix * fact4 * (XY[i] - XY[n2]) +
                    iz * fact4 * (YZ[i] - YZ[n0]);
        double wz = iz * fact1 * (diag(5, n5) - diag(5, i)) + ix * fact4 * (XZ[i] - XZ[n2]) +
                    iy * fact4 * (YZ[i] - YZ[n1]);
        out[0][i] += wx;
        out[1][i] += wy;
        out[2][i] += wz;
      }
}

// ---- Fourier-space solves ----
// Stokes == false: kernelUpdateVIncompressible (src/kernelUpdateVIncompressible.cu:21-72),
// Crank-Nicolson viscous solve + projection; mode 0 kept.
// Stokes == true: kernelUpdateVstokesLimit (src/stokesLimitFunctions.cu:2399-2451),
// v = P W / (eta |g|^2); mode 0 set to zero.
// Shifts e^{-+ i pi k/m} per staggered axis: kernelShift (src/realToComplex.cu:263-303).
// g_a(k) = (2/dx_a) sin(pi k / m_a): initializePrefactorFourierSpace_2
// (src/initializePrefactorFourierSpace.cu:43-69).
static void solve(orc_ctx* c, const std::vector<double>* W, std::vector<double>* vout, bool stokes,
                  bool project_only) {
  const int N = c->N, mx = c->mx, my = c->my, mz = c->mz;
  std::vector<cplx> Wk[3];
  for (int a = 0; a < 3; ++a) {
    Wk[a].resize(N);
    for (int i = 0; i < N; ++i) Wk[a][i] = cplx(W[a][i], 0.0);
    c->fft.run(Wk[a].data(), -1);
  }
  std::vector<double> gx(mx), gy(my), gz(mz);
  std::vector<cplx> ex(mx), ey(my), ez(mz);
  for (int k = 0; k < mx; ++k) {
    gx[k] = 2.0 / c->dx * sin(M_PI * k / (double)mx);
    ex[k] = cplx(cos(M_PI * k / (double)mx), sin(M_PI * k / (double)mx));
  }
  for (int k = 0; k < my; ++k) {
    gy[k] = 2.0 / c->dy * sin(M_PI * k / (double)my);
    ey[k] = cplx(cos(M_PI * k / (double)my), sin(M_PI * k / (double)my));
  }
  for (int k = 0; k < mz; ++k) {
    gz[k] = 2.0 / c->dz * sin(M_PI * k / (double)mz);
    ez[k] = cplx(cos(M_PI * k / (double)mz), sin(M_PI * k / (double)mz));
  }
  const double nu_dt = c->p.dt * c->p.shearviscosity / c->p.densfluid;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    int kx = i % mx, ky = (i / mx) % my, kz = i / (mx * my);
    cplx w0 = Wk[0][i] * std::conj(ex[kx]);
    cplx w1 = Wk[1][i] * std::conj(ey[ky]);
    cplx w2 = Wk[2][i] * std::conj(ez[kz]);
    double g0 = gx[kx], g1 = gy[ky], g2 = gz[kz];
    double L = -(g0 * g0 + g1 * g1 + g2 * g2);
    cplx o0, o1, o2;
    if (i == 0) {
      if (stokes) { o0 = o1 = o2 = 0; }
      else { o0 = w0; o1 = w1; o2 = w2; }
    } else {
      cplx GW = g0 * w0 + g1 * w1 + g2 * w2;
      double den;
      if (project_only) den = 1.0;
      else if (stokes) den = -c->p.shearviscosity * L;
      else den = 1.0 - 0.5 * nu_dt * L;
      o0 = (w0 + g0 * GW / L) / den;
      o1 = (w1 + g1 * GW / L) / den;
      o2 = (w2 + g2 * GW / L) / den;
    }
    Wk[0][i] = o0 * ex[kx];
    Wk[1][i] = o1 * ey[ky];
    Wk[2][i] = o2 * ez[kz];
  }
  for (int a = 0; a < 3; ++a) {
    c->fft.run(Wk[a].data(), +1);
    vout[a].resize(N);
    for (int i = 0; i < N; ++i) vout[a][i] = Wk[a][i].real() / (double)N;
  }
}

// ---- immersed-boundary kernels ----
// Nodes of component comp along axis ax: staggered (face, at (k+1) dx - L/2) when ax == comp,
// centred ((k+1/2) dx - L/2) otherwise (src/createCells.cpp:68-70, src/stokesLimitFunctions.cu:767).
struct AxisW {
  int base;      // first node index (unwrapped)
  int n;         // 3 or 4
  double w[4];
  double dw[4];  // d w / d q (1/length)
};
static void axis_weights(const orc_ctx* c, int ax, int comp, double qpos, AxisW* aw) {
  const int m = ax == 0 ? c->mx : ax == 1 ? c->my : c->mz;
  const double L = ax == 0 ? c->p.lx : ax == 1 ? c->p.ly : c->p.lz;
  const double h = ax == 0 ? c->dx : ax == 1 ? c->dy : c->dz;
  const double off = (ax == comp) ? 1.0 : 0.5;
  double qw = wrapd(qpos, L);
  double t = (qw + 0.5 * L) / h - off;  // position in node units
  (void)m;
  if (c->p.kernel_points == 4) {
    int j0 = (int)floor(t);
    aw->base = j0 - 1;
    aw->n = 4;
    for (int k = 0; k < 4; ++k) {
      double r = t - (double)(j0 - 1 + k);
      aw->w[k] = orc_delta4(r);
      aw->dw[k] = orc_ddelta4(r) / h;
    }
  } else {
    int j0 = (int)floor(t + 0.5);
    aw->base = j0 - 1;
    aw->n = 3;
    for (int k = 0; k < 3; ++k) {
      double r = t - (double)(j0 - 1 + k);
      aw->w[k] = orc_delta3(r);
      aw->dw[k] = orc_ddelta3(r) / h;
    }
    aw->w[3] = aw->dw[3] = 0;
  }
}

// field_c(node) += scale * sum_p phi phi phi * F_p,c  (spread; src/stokesLimitFunctions.cu:659-1158
// with the gather of :1198-1681; the 1/dV is applied by the caller via scale).
static void spread_one(const orc_ctx* c, const double qp[3], const double Fp[3], double scale,
                       std::vector<double>* field) {
  for (int comp = 0; comp < 3; ++comp) {
    if (Fp[comp] == 0.0) continue;
    AxisW wx, wy, wz;
    axis_weights(c, 0, comp, qp[0], &wx);
    axis_weights(c, 1, comp, qp[1], &wy);
    axis_weights(c, 2, comp, qp[2], &wz);
    for (int kz = 0; kz < wz.n; ++kz)
      for (int ky = 0; ky < wy.n; ++ky)
        for (int kx = 0; kx < wx.n; ++kx) {
          int i = cid(c, wx.base + kx, wy.base + ky, wz.base + kz);
          field[comp][i] += scale * wx.w[kx] * wy.w[ky] * wz.w[kz] * Fp[comp];
        }
  }
}

// Thermal drift kT * d/dq_c [phi phi phi] on component c (src/quasiNeutrallyBuoyantFunctions.cu:
// 6281-6912), times scale.
static void spread_drift_one(const orc_ctx* c, const double qp[3], double scale,
                             std::vector<double>* field) {
  for (int comp = 0; comp < 3; ++comp) {
    AxisW wx, wy, wz;
    axis_weights(c, 0, comp, qp[0], &wx);
    axis_weights(c, 1, comp, qp[1], &wy);
    axis_weights(c, 2, comp, qp[2], &wz);
    for (int kz = 0; kz < wz.n; ++kz)
      for (int ky = 0; ky < wy.n; ++ky)
        for (int kx = 0; kx < wx.n; ++kx) {
          double d = comp == 0 ? wx.dw[kx] * wy.w[ky] * wz.w[kz]
                   : comp == 1 ? wx.w[kx] * wy.dw[ky] * wz.w[kz]
                               : wx.w[kx] * wy.w[ky] * wz.dw[kz];
          int i = cid(c, wx.base + kx, wy.base + ky, wz.base + kz);
          field[comp][i] += scale * d;
        }
  }
}

// Reference-quirk RFD spread (src/stokesLimitFunctions.cu:1741-1906): the stencil nodes are
// chosen from the displaced position qd, the weights from the undisplaced q.
static void spread_quirk_one(const orc_ctx* c, const double q0[3], const double qd[3],
                             const double Fp[3], double scale, std::vector<double>* field) {
  for (int comp = 0; comp < 3; ++comp) {
    AxisW aw[3];
    for (int ax = 0; ax < 3; ++ax) {
      axis_weights(c, ax, comp, qd[ax], &aw[ax]);
      // re-evaluate weights at q0 on the nodes chosen from qd
      const double L = ax == 0 ? c->p.lx : ax == 1 ? c->p.ly : c->p.lz;
      const double h = ax == 0 ? c->dx : ax == 1 ? c->dy : c->dz;
      const double off = (ax == comp) ? 1.0 : 0.5;
      double td = (wrapd(qd[ax], L) + 0.5 * L) / h - off;
      double shift = (q0[ax] - qd[ax]) / h;  // tiny displacement, in node units
      for (int k = 0; k < aw[ax].n; ++k) {
        double r = td - (double)(aw[ax].base + k) + shift;
        aw[ax].w[k] = c->p.kernel_points == 4 ? orc_delta4(r) : orc_delta3(r);
      }
    }
    for (int kz = 0; kz < aw[2].n; ++kz)
      for (int ky = 0; ky < aw[1].n; ++ky)
        for (int kx = 0; kx < aw[0].n; ++kx) {
          int i = cid(c, aw[0].base + kx, aw[1].base + ky, aw[2].base + kz);
          field[comp][i] += scale * aw[0].w[kx] * aw[1].w[ky] * aw[2].w[kz] * Fp[comp];
        }
  }
}

// u_c = sum phi phi phi v_c (interpolation J; src/stokesLimitFunctions.cu:2479-2892).
static void interp_one(const orc_ctx* c, const double qp[3], const std::vector<double>* v, double u[3]) {
  for (int comp = 0; comp < 3; ++comp) {
    AxisW wx, wy, wz;
    axis_weights(c, 0, comp, qp[0], &wx);
    axis_weights(c, 1, comp, qp[1], &wy);
    axis_weights(c, 2, comp, qp[2], &wz);
    double s = 0.0;
    for (int kz = 0; kz < wz.n; ++kz)
      for (int ky = 0; ky < wy.n; ++ky)
        for (int kx = 0; kx < wx.n; ++kx)
          s += wx.w[kx] * wy.w[ky] * wz.w[kz] * v[comp][cid(c, wx.base + kx, wy.base + ky, wz.base + kz)];
    u[comp] = s;
  }
}

// WCA force via a cell list of m_a = max(3, floor(L_a / cutoff)) cells per axis
// (src/createBoundariesRK2GPU.cu:134-141; pair loop src/stokesLimitFunctions.cu:160-653).
static void wca_forces(const orc_ctx* c, const std::vector<double>* q, std::vector<double>* F) {
  const int np = c->p.np;
  for (int a = 0; a < 3; ++a) F[a].assign(np, 0.0);
  if (!c->p.nonbonded || np == 0 || c->p.cutoff <= 0) return;
  const double rc = c->p.cutoff, rc2 = rc * rc;
  const double sig2 = c->p.wca_sigma * c->p.wca_sigma, eps = c->p.wca_epsilon;
  const double L[3] = {c->p.lx, c->p.ly, c->p.lz};
  int m[3];
  for (int a = 0; a < 3; ++a) m[a] = std::max(3, (int)(L[a] / rc));
  const int ncell = m[0] * m[1] * m[2];
  std::vector<int> head(ncell, -1), next(np, -1), cellof(np);
  for (int i = 0; i < np; ++i) {
    int j[3];
    for (int a = 0; a < 3; ++a) {
      double w = wrapd(q[a][i], L[a]);
      j[a] = (int)floor((w + 0.5 * L[a]) * m[a] / L[a]);
      j[a] = std::min(std::max(j[a], 0), m[a] - 1);
    }
    int ci = j[0] + m[0] * (j[1] + m[1] * j[2]);
    cellof[i] = ci;
  }
  for (int i = np - 1; i >= 0; --i) { next[i] = head[cellof[i]]; head[cellof[i]] = i; }
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < np; ++i) {
    int ci = cellof[i];
    int j0 = ci % m[0], j1 = (ci / m[0]) % m[1], j2 = ci / (m[0] * m[1]);
    double f[3] = {0, 0, 0};
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dxx = -1; dxx <= 1; ++dxx) {
          int nc = wrapi(j0 + dxx, m[0]) + m[0] * (wrapi(j1 + dy, m[1]) + m[1] * wrapi(j2 + dz, m[2]));
          for (int jj = head[nc]; jj >= 0; jj = next[jj]) {
            if (jj == i) continue;
            double r[3], r2 = 0;
            for (int a = 0; a < 3; ++a) { r[a] = wrapd(q[a][i] - q[a][jj], L[a]); r2 += r[a] * r[a]; }
            if (r2 >= rc2 || r2 == 0.0) continue;
            r2 = std::max(r2, 1e-4 * sig2);  // force evaluated at sigma/100 below that separation
            double s2 = sig2 / r2, s6 = s2 * s2 * s2;
            double fr = 48.0 * eps * (s6 * s6 - 0.5 * s6) / r2;
            for (int a = 0; a < 3; ++a) f[a] += fr * r[a];
          }
        }
    for (int a = 0; a < 3; ++a) F[a][i] = f[a];
  }
}

// Harmonic bond forces added to F (src/forceBondedGPU.cu:21-121): for a bond of stiffness k
// and rest length r0 between positions a and b, F_a += -k (1 - r0/|r|) r with r = a - b the
// plain difference (no periodic image), -k r when r0 = 0, and nothing when r = 0 and r0 != 0.
// Particle-particle bonds act on both ends (src/initializeBondedForces.cpp:126-137).
static void bond_term(const double r[3], double k, double r0, double* f) {
  double s = k;
  if (r0 != 0.0) {
    const double d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    s = d > 0.0 ? k * (1.0 - r0 / d) : 0.0;
  }
  for (int a = 0; a < 3; ++a) f[a] -= s * r[a];
}
static void bond_forces(const orc_ctx* c, const std::vector<double>* q, std::vector<double>* F) {
  for (const auto& b : c->pair_bonds) {
    double r[3], fi[3] = {0, 0, 0}, fj[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) r[a] = q[a][b.i] - q[a][b.j];
    bond_term(r, b.k, b.r0, fi);
    for (int a = 0; a < 3; ++a) r[a] = -r[a];
    bond_term(r, b.k, b.r0, fj);
    for (int a = 0; a < 3; ++a) { F[a][b.i] += fi[a]; F[a][b.j] += fj[a]; }
  }
  for (const auto& b : c->fixed_bonds) {
    double r[3], f[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) r[a] = q[a][b.i] - b.x[a];
    bond_term(r, b.k, b.r0, f);
    for (int a = 0; a < 3; ++a) F[a][b.i] += f[a];
  }
}
// All particle forces: WCA plus bonds.
static void particle_forces(const orc_ctx* c, const std::vector<double>* q, std::vector<double>* F) {
  wca_forces(c, q, F);
  bond_forces(c, q, F);
}

// Bonds file (bin/inputfiles/README:55-108): count, then "i j k r0" lines; count, then
// "i k r0 x y z" lines.
static bool orc_read_bonds(orc_ctx* c, const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  const int np = c->p.np;
  long long n = -1;
  bool ok = fscanf(f, "%lld", &n) == 1 && n >= 0;
  for (long long e = 0; ok && e < n; ++e) {
    orc_ctx::PairBond b;
    ok = fscanf(f, "%d %d %lf %lf", &b.i, &b.j, &b.k, &b.r0) == 4 && b.i >= 0 && b.i < np && b.j >= 0 && b.j < np;
    if (ok) c->pair_bonds.push_back(b);
  }
  ok = ok && fscanf(f, "%lld", &n) == 1 && n >= 0;
  for (long long e = 0; ok && e < n; ++e) {
    orc_ctx::FixedBond b;
    ok = fscanf(f, "%d %lf %lf %lf %lf %lf", &b.i, &b.k, &b.r0, &b.x[0], &b.x[1], &b.x[2]) == 6 && b.i >= 0 && b.i < np;
    if (ok) c->fixed_bonds.push_back(b);
  }
  fclose(f);
  return ok;
}

// ------------------------------------------------------------------------------------
// Public API.
// ------------------------------------------------------------------------------------
extern "C" int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

extern "C" int orc_initial_positions(const fhd_params* p, double* q) {
  const int np = p->np;
  if (p->initial_positions == 1) {
    // simpleCubic (src/simpleCubic.cpp:25-63)
    double sigma = pow(p->lx * p->ly * p->lz / np, 1.0 / 3.0);
    int nx = (int)(p->lx / sigma) ? (int)(p->lx / sigma) : 1;
    int ny = (int)(p->ly / sigma) ? (int)(p->ly / sigma) : 1;
    int nz = (int)(p->lz / sigma) ? (int)(p->lz / sigma) : 1;
    while (nx * ny * nz < np) {
      if (nx * ny * nz < np) nx++;
      if (nx * ny * nz < np) ny++;
      if (nx * ny * nz < np) nz++;
    }
    double ddx = p->lx / nx, ddy = p->ly / ny, ddz = p->lz / nz;
    int n = 0;
    for (int i = 0; i < nz && n < np; ++i)
      for (int j = 0; j < ny && n < np; ++j)
        for (int k = 0; k < nx && n < np; ++k) {
          q[n] = (k + 0.5) * ddx - p->lx / 2;
          q[np + n] = (j + 0.5) * ddy - p->ly / 2;
          q[2 * np + n] = (i + 0.5) * ddz - p->lz / 2;
          ++n;
        }
    return 1;
  }
  // seeded uniform positions (attempt 0 of every particle), or, for mode 3, random sequential
  // addition: a particle is redrawn (attempt 1, 2, ... in the step word of the counter) while
  // its minimum-image distance to an earlier particle is below the WCA sigma.  Earlier
  // particles are found through linked cell lists with cells no smaller than sigma, visiting
  // each distinct neighbour cell once.
  const bool rsa = p->initial_positions == 3 && p->nonbonded && p->cutoff > 0;
  const double dmin = p->wca_sigma > 0 ? p->wca_sigma : 2.0 * p->lx / p->mx;
  const double L[3] = {p->lx, p->ly, p->lz};
  int m[3];
  for (int a = 0; a < 3; ++a) m[a] = rsa ? std::max(1, (int)(L[a] / dmin)) : 1;
  std::vector<int> head(rsa ? (size_t)m[0] * m[1] * m[2] : 0, -1), next(rsa ? np : 0, -1);
  auto cell_index = [&](const double x[3], int j[3]) {
    for (int a = 0; a < 3; ++a) {
      j[a] = (int)((x[a] / L[a] + 0.5) * m[a]);
      j[a] = std::min(std::max(j[a], 0), m[a] - 1);
    }
    return j[0] + m[0] * (j[1] + m[1] * j[2]);
  };
  for (int i = 0; i < np; ++i) {
    double x[3];
    for (int attempt = 0;; ++attempt) {
      if (attempt >= 100000) return 0;
      double u[4];
      uniforms4(p->seed, attempt, STREAM_INIT_POS, 0, (uint32_t)i, u);
      for (int a = 0; a < 3; ++a) x[a] = (u[a] - 0.5) * L[a];
      if (!rsa) break;
      int j[3];
      cell_index(x, j);
      std::vector<int> seen;
      bool ok = true;
      for (int dz = -1; dz <= 1 && ok; ++dz)
        for (int dy = -1; dy <= 1 && ok; ++dy)
          for (int dxx = -1; dxx <= 1 && ok; ++dxx) {
            const int nc = wrapi(j[0] + dxx, m[0]) + m[0] * (wrapi(j[1] + dy, m[1]) + m[1] * wrapi(j[2] + dz, m[2]));
            if (std::find(seen.begin(), seen.end(), nc) != seen.end()) continue;
            seen.push_back(nc);
            for (int k = head[nc]; k >= 0 && ok; k = next[k]) {
              double r2 = 0.0;
              for (int a = 0; a < 3; ++a) {
                double d = x[a] - q[(size_t)a * np + k];
                d -= L[a] * floor(d / L[a] + 0.5);
                r2 += d * d;
              }
              ok = r2 >= dmin * dmin;
            }
          }
      if (ok) break;
    }
    for (int a = 0; a < 3; ++a) q[(size_t)a * np + i] = x[a];
    if (rsa) {
      int j[3];
      const int c = cell_index(x, j);
      next[i] = head[c];
      head[c] = i;
    }
  }
  return 1;
}

// Restart from a .fluidFinalConfiguration file (src/initializeFluid.cpp:168-192): skip the five
// grid lines, read "Density d" (it replaces densfluid) and the pressure line, skip the rest of
// the header, then per cell: density, vx, vy, vz.
static bool orc_read_fluid(orc_ctx* c, const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char line[1024];
  bool ok = true;
  for (int k = 0; k < 5 && ok; ++k) ok = fgets(line, sizeof line, f) != nullptr;
  double dens = 0, a0, a1, a2;
  ok = ok && fscanf(f, "%*s %lf %*s %*s %lf %lf %lf", &dens, &a0, &a1, &a2) == 4;
  for (int k = 0; k < 6 && ok; ++k) ok = fgets(line, sizeof line, f) != nullptr;
  for (int i = 0; i < c->N && ok; ++i) {
    double d;
    ok = fscanf(f, "%lf %lf %lf %lf", &d, &c->v[0][i], &c->v[1][i], &c->v[2][i]) == 4;
  }
  fclose(f);
  if (ok) c->p.densfluid = dens;
  return ok;
}

// Initial fluid velocity, initfluid 0-4 (src/initializeFluid.cpp:31-160): background; equilibrium
// fluctuations sqrt(kT/(rho dV)) N(0,1) per face (cell i: normals 0-2 of draw (seed, step 0,
// STREAM_INIT_FLUID, sub 0, index i)) minus their mean plus the background; a transverse wave
// vy0 + 0.04 sin(2 pi z/lz); a longitudinal wave 0.004 sin(16 pi (z + dz/2)/lz); vx = 0.425.
static void initial_fluid(orc_ctx* c) {
  const fhd_params& p = c->p;
  const int N = c->N;
  const double* b = p.background;
  double sum[3] = {0, 0, 0};
  for (int i = 0; i < N; ++i) {
    const int iz = i / (c->mx * c->my);
    const double z = (iz + 0.5) * c->dz - 0.5 * p.lz;
    double v[3] = {b[0], b[1], b[2]};
    if (p.initfluid == 1) {
      double n[4];
      orc_normals4(p.seed, 0, STREAM_INIT_FLUID, 0, (uint32_t)i, n);
      const double s = sqrt(p.temperature / (p.densfluid * c->dV));
      for (int a = 0; a < 3; ++a) { v[a] = s * n[a]; sum[a] += v[a]; }
    } else if (p.initfluid == 2) {
      v[1] = b[1] + 0.04 * sin(2.0 * M_PI / p.lz * z);
    } else if (p.initfluid == 3) {
      v[0] = 0.0; v[1] = 0.0;
      v[2] = 0.004 * sin(2.0 * M_PI / p.lz * 8.0 * (z + 0.5 * p.lz / c->mz));
    } else if (p.initfluid == 4) {
      v[0] = 0.425; v[1] = 0.0; v[2] = 0.0;
    }
    for (int a = 0; a < 3; ++a) c->v[a][i] = v[a];
  }
  if (p.initfluid == 1)
    for (int a = 0; a < 3; ++a) {
      const double m = sum[a] / (double)N;
      for (int i = 0; i < N; ++i) c->v[a][i] = c->v[a][i] - m + b[a];
    }
}

// Particle positions (and optional velocities) from the coordinates / velocities files
// (src/createParticles.cpp:81-120): "np" then np lines of three numbers each.
static bool read_triples(const char* path, int* np, std::vector<double>* out) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  int n = -1;
  bool ok = fscanf(f, "%d", &n) == 1 && n >= 0 && (*np < 0 || n == *np);
  if (ok) {
    *np = n;
    for (int a = 0; a < 3; ++a) out[a].assign(n, 0.0);
    for (int i = 0; i < n && ok; ++i) ok = fscanf(f, "%lf %lf %lf", &out[0][i], &out[1][i], &out[2][i]) == 3;
  }
  fclose(f);
  return ok;
}

extern "C" orc_ctx* orc_create(const fhd_params* p) {
  if (!p || p->mx <= 0 || p->my <= 0 || p->mz <= 0) return nullptr;
  orc_ctx* c = new orc_ctx();
  c->p = *p;
  std::vector<double> qf[3], uf[3];
  if (p->initial_positions == 2) {
    int n = -1;
    if (!read_triples(p->coordinates, &n, qf) || (p->velocities[0] && !read_triples(p->velocities, &n, uf))) {
      delete c;
      return nullptr;
    }
    c->p.np = n;
  }
  c->mx = p->mx; c->my = p->my; c->mz = p->mz;
  c->N = p->mx * p->my * p->mz;
  c->dx = p->lx / p->mx; c->dy = p->ly / p->my; c->dz = p->lz / p->mz;
  c->dV = c->dx * c->dy * c->dz;
  c->fft.init(c->mx, c->my, c->mz);
  for (int a = 0; a < 3; ++a) c->v[a].assign(c->N, 0.0);
  if (p->scheme != FHD_SCHEME_STOKES_LIMIT) initial_fluid(c);
  c->step = p->initial_step;
  if (p->fluid[0] && !orc_read_fluid(c, p->fluid)) {
    delete c;
    return nullptr;
  }
  const int np = c->p.np;
  for (int a = 0; a < 3; ++a) { c->q[a].assign(np, 0.0); c->F[a].assign(np, 0.0); c->u[a].assign(np, 0.0); }
  if (p->initial_positions == 2) {
    for (int a = 0; a < 3; ++a) {
      c->q[a] = qf[a];
      if (!uf[a].empty()) c->u[a] = uf[a];
    }
  } else if (np > 0) {
    std::vector<double> q0((size_t)3 * np);
    orc_initial_positions(p, q0.data());
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < np; ++i) c->q[a][i] = q0[(size_t)a * np + i];
  }
  if (p->bonded_forces[0] && !orc_read_bonds(c, p->bonded_forces)) {
    delete c;
    return nullptr;
  }
  return c;
}

extern "C" void orc_destroy(orc_ctx* c) { delete c; }
extern "C" long long orc_get_step(orc_ctx* c) { return c ? c->step : -1; }
extern "C" int orc_get_np(orc_ctx* c) { return c ? c->p.np : -1; }

extern "C" int orc_set_velocity(orc_ctx* c, const double* v) {
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->N; ++i) c->v[a][i] = v[(size_t)a * c->N + i];
  c->projected = false;
  return 1;
}
extern "C" int orc_get_velocity(orc_ctx* c, double* v) {
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->N; ++i) v[(size_t)a * c->N + i] = c->v[a][i];
  return 1;
}
extern "C" int orc_set_particles(orc_ctx* c, const double* q) {
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->p.np; ++i) c->q[a][i] = q[(size_t)a * c->p.np + i];
  return 1;
}
extern "C" int orc_get_particles(orc_ctx* c, double* q) {
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->p.np; ++i) q[(size_t)a * c->p.np + i] = c->q[a][i];
  return 1;
}
extern "C" int orc_get_particle_velocities(orc_ctx* c, double* u) {
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->p.np; ++i) u[(size_t)a * c->p.np + i] = c->u[a][i];
  return 1;
}
extern "C" int orc_get_forces(orc_ctx* c, double* f) {
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->p.np; ++i) f[(size_t)a * c->p.np + i] = c->F[a][i];
  return 1;
}

extern "C" int orc_wca(orc_ctx* c, const double* qin, double* f) {
  std::vector<double> q[3], F[3];
  for (int a = 0; a < 3; ++a) q[a].assign(qin + (size_t)a * c->p.np, qin + (size_t)(a + 1) * c->p.np);
  wca_forces(c, q, F);
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->p.np; ++i) f[(size_t)a * c->p.np + i] = F[a][i];
  return 1;
}

extern "C" int orc_bonded(orc_ctx* c, const double* qin, double* f) {
  std::vector<double> q[3], F[3];
  for (int a = 0; a < 3; ++a) {
    q[a].assign(qin + (size_t)a * c->p.np, qin + (size_t)(a + 1) * c->p.np);
    F[a].assign(c->p.np, 0.0);
  }
  bond_forces(c, q, F);
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->p.np; ++i) f[(size_t)a * c->p.np + i] = F[a][i];
  return 1;
}

extern "C" int orc_spread(orc_ctx* c, const double* q, const double* Fin, double* field) {
  std::vector<double> fld[3];
  for (int a = 0; a < 3; ++a) fld[a].assign(field + (size_t)a * c->N, field + (size_t)(a + 1) * c->N);
  const int np = c->p.np;
  for (int i = 0; i < np; ++i) {
    double qp[3] = {q[i], q[np + i], q[2 * np + i]};
    double Fp[3] = {Fin[i], Fin[np + i], Fin[2 * np + i]};
    spread_one(c, qp, Fp, 1.0 / c->dV, fld);
  }
  for (int a = 0; a < 3; ++a) memcpy(field + (size_t)a * c->N, fld[a].data(), sizeof(double) * c->N);
  return 1;
}

extern "C" int orc_interp(orc_ctx* c, const double* q, const double* field, double* u) {
  std::vector<double> fld[3];
  for (int a = 0; a < 3; ++a) fld[a].assign(field + (size_t)a * c->N, field + (size_t)(a + 1) * c->N);
  const int np = c->p.np;
  for (int i = 0; i < np; ++i) {
    double qp[3] = {q[i], q[np + i], q[2 * np + i]};
    double up[3];
    interp_one(c, qp, fld, up);
    for (int a = 0; a < 3; ++a) u[(size_t)a * np + i] = up[a];
  }
  return 1;
}

extern "C" int orc_spread_drift(orc_ctx* c, const double* q, double scale, double* field) {
  std::vector<double> fld[3];
  for (int a = 0; a < 3; ++a) fld[a].assign(field + (size_t)a * c->N, field + (size_t)(a + 1) * c->N);
  const int np = c->p.np;
  for (int i = 0; i < np; ++i) {
    double qp[3] = {q[i], q[np + i], q[2 * np + i]};
    spread_drift_one(c, qp, scale, fld);
  }
  for (int a = 0; a < 3; ++a) memcpy(field + (size_t)a * c->N, fld[a].data(), sizeof(double) * c->N);
  return 1;
}

extern "C" int orc_spread_quirk(orc_ctx* c, const double* q, const double* qd, const double* Fin, double scale,
                                double* field) {
  std::vector<double> fld[3];
  for (int a = 0; a < 3; ++a) fld[a].assign(field + (size_t)a * c->N, field + (size_t)(a + 1) * c->N);
  const int np = c->p.np;
  for (int i = 0; i < np; ++i) {
    double q0[3] = {q[i], q[np + i], q[2 * np + i]};
    double qq[3] = {qd[i], qd[np + i], qd[2 * np + i]};
    double Fp[3] = {Fin[i], Fin[np + i], Fin[2 * np + i]};
    spread_quirk_one(c, q0, qq, Fp, scale, fld);
  }
  for (int a = 0; a < 3; ++a) memcpy(field + (size_t)a * c->N, fld[a].data(), sizeof(double) * c->N);
  return 1;
}

extern "C" int orc_divergence(orc_ctx* c, const double* v, double* div) {
  const int N = c->N;
  for (int z = 0; z < c->mz; ++z)
    for (int y = 0; y < c->my; ++y)
      for (int x = 0; x < c->mx; ++x) {
        int i = cid(c, x, y, z);
        div[i] = (v[i] - v[cid(c, x - 1, y, z)]) / c->dx +
                 (v[(size_t)N + i] - v[(size_t)N + cid(c, x, y - 1, z)]) / c->dy +
                 (v[(size_t)2 * N + i] - v[(size_t)2 * N + cid(c, x, y, z - 1)]) / c->dz;
      }
  return 1;
}

extern "C" int orc_solve_incompressible(orc_ctx* c, const double* Win, double* vout) {
  std::vector<double> W[3], V[3];
  for (int a = 0; a < 3; ++a) W[a].assign(Win + (size_t)a * c->N, Win + (size_t)(a + 1) * c->N);
  solve(c, W, V, false, false);
  for (int a = 0; a < 3; ++a) memcpy(vout + (size_t)a * c->N, V[a].data(), sizeof(double) * c->N);
  return 1;
}
// Projection onto discretely divergence-free fields, mode 0 kept (projectionDivergenceFree,
// src/projectionDivergenceFree.cu:21-72).
extern "C" int orc_project(orc_ctx* c, const double* Win, double* vout) {
  std::vector<double> W[3], V[3];
  for (int a = 0; a < 3; ++a) W[a].assign(Win + (size_t)a * c->N, Win + (size_t)(a + 1) * c->N);
  solve(c, W, V, false, true);
  for (int a = 0; a < 3; ++a) memcpy(vout + (size_t)a * c->N, V[a].data(), sizeof(double) * c->N);
  return 1;
}
extern "C" int orc_solve_stokes(orc_ctx* c, const double* Win, double* vout) {
  std::vector<double> W[3], V[3];
  for (int a = 0; a < 3; ++a) W[a].assign(Win + (size_t)a * c->N, Win + (size_t)(a + 1) * c->N);
  solve(c, W, V, true, false);
  for (int a = 0; a < 3; ++a) memcpy(vout + (size_t)a * c->N, V[a].data(), sizeof(double) * c->N);
  return 1;
}

// Initial projection of v (src/runSchemeIncompressible.cu:79-97).
static void project_initial(orc_ctx* c) {
  std::vector<double> V[3];
  solve(c, c->v, V, false, true);
  for (int a = 0; a < 3; ++a) c->v[a] = V[a];
  c->projected = true;
}

// ---- incompressible step: src/runSchemeIncompressible.cu:105-139 ----
static void step_incompressible(orc_ctx* c) {
  const double dV = c->dV, rho = c->p.densfluid;
  const double fact1 = sqrt(4.0 * c->p.temperature * c->p.shearviscosity * c->p.dt / (dV * rho * rho));
  const double fact4 = sqrt(2.0 * c->p.temperature * c->p.shearviscosity * c->p.dt / (dV * rho * rho));
  std::vector<double> Z;
  fluid_noise(c->p.seed, c->step, 0, c->N, Z);
  std::vector<double> W[3], A[3], vt[3], vstar[3];
  // first substep: W = v + lap(v) - A(v) + D.Sigma
  lap_term(c, c->v, W);
  adv_term(c, c->v, A);
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->N; ++i) W[a][i] += c->v[a][i] - A[a][i];
  noise_div(c, Z, fact1, fact4, true, W);
  solve(c, W, vt, false, false);
  for (int a = 0; a < 3; ++a) {
    vstar[a].resize(c->N);
    for (int i = 0; i < c->N; ++i) vstar[a][i] = 0.5 * (vt[a][i] + c->v[a][i]);
  }
  // second substep: advection at v* = (v^n + v~)/2 (src/kernelConstructW.cu:266-521)
  lap_term(c, c->v, W);
  adv_term(c, vstar, A);
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < c->N; ++i) W[a][i] += c->v[a][i] - A[a][i];
  noise_div(c, Z, fact1, fact4, true, W);
  solve(c, W, c->v, false, false);
}

// ---- Stokes-limit midpoint step: src/runSchemeStokesLimit.cu:163-276 and
// src/boundaryParticlesFunctionStokesLimit.cu:21-249 ----
static void rfd_spread(orc_ctx* c, const std::vector<double>* q, const std::vector<double>* Wr,
                       std::vector<double>* f) {
  const int np = c->p.np;
  const double delta = c->p.rfd_delta;
  const double pre = c->p.temperature / delta / c->dV;
  for (int i = 0; i < np; ++i) {
    double W3[3] = {Wr[0][i], Wr[1][i], Wr[2][i]};
    double q0[3] = {q[0][i], q[1][i], q[2][i]};
    for (int sgn = 1; sgn >= -1; sgn -= 2) {
      double qd[3];
      for (int a = 0; a < 3; ++a) qd[a] = q0[a] + sgn * 0.5 * delta * W3[a];
      if (c->p.rfd_mode == FHD_RFD_REFERENCE) spread_quirk_one(c, q0, qd, W3, sgn * pre, f);
      else spread_one(c, qd, W3, sgn * pre, f);
    }
  }
}

// ---- The Stokes step's stochastic stress, drawn in Fourier space ----
// Sigma is white: the DFT of N independent N(0,1) values per component has, at a mode k with
// -k != k, real and imaginary parts N(0, N/2), independent of all modes but -k, where it is the
// conjugate; self-conjugate modes are real N(0, N).  The modes with kx in [0, mx/2] draw their
// values (on the planes kx = 0 and kx = mx/2 the member of each (ky, kz), (-ky, -kz) pair with
// the smaller ky + my kz draws, the other is its conjugate): stream STREAM_FLUID_K, counter index
// kx + P (ky + my kz), subs 3 s .. 3 s + 2 for stress set s, the 12 normals being the real and
// imaginary parts of XX, XY, XZ, YY, YZ, ZZ; modes with kx > mx/2 are conjugates of modes
// -k.  S[comp] gets the full spectrum of Sigma_1 (nset 1) or Sigma_1 + Sigma_2 (nset 2).
static void kspace_sigma(const orc_ctx* c, int64_t step, int nset, std::vector<cplx>* S) {
  const int mx = c->mx, my = c->my, mz = c->mz, N = c->N, P = mx / 2 + 1;
  const double shn = sqrt(0.5 * (double)N);
  for (int k = 0; k < 6; ++k) S[k].assign(N, cplx(0.0, 0.0));
#pragma omp parallel for schedule(static)
  for (int kz = 0; kz < mz; ++kz)
    for (int ky = 0; ky < my; ++ky)
      for (int kx = 0; kx < mx; ++kx) {
        int hx = kx, hy = ky, hz = kz;
        bool conj = false;
        if (2 * kx > mx) {
          hx = mx - kx;
          hy = (my - ky) % my;
          hz = (mz - kz) % mz;
          conj = true;
        }
        int cy = hy, cz = hz;
        bool selfc = false;
        if (hx == 0 || 2 * hx == mx) {
          const int py = (my - hy) % my, pz = (mz - hz) % mz;
          const int L = hy + my * hz, Lp = py + my * pz;
          if (Lp == L) {
            selfc = true;
          } else if (Lp < L) {
            cy = py;
            cz = pz;
            conj = !conj;
          }
        }
        const uint32_t index = (uint32_t)(hx + P * (cy + my * cz));
        double z[12] = {0};
        for (int s = 0; s < nset; ++s)
          for (int sub = 0; sub < 3; ++sub) {
            double n[4];
            orc_normals4(c->p.seed, step, STREAM_FLUID_K, 3 * s + sub, index, n);
            for (int k = 0; k < 4; ++k) z[4 * sub + k] += n[k];
          }
        const size_t i = (size_t)kx + (size_t)mx * (ky + (size_t)my * kz);
        for (int k = 0; k < 6; ++k) {
          cplx val = selfc ? cplx(sqrt(2.0) * shn * z[2 * k], 0.0) : cplx(shn * z[2 * k], shn * z[2 * k + 1]);
          S[k][i] = conj ? std::conj(val) : val;
        }
      }
}

// The divergence of the staggered stress in Fourier space (diagonal at cell centres,
// off-diagonal at edges, src/kernelConstructW.cu:148-198):
//   w_x = i [f1 g_x e_x S_xx + f4 g_y e_y* S_xy + f4 g_z e_z* S_xz]  (and cyclically),
// g_a = (2/h_a) sin(pi k_a/m_a), e_a = e^{i pi k_a/m_a}; returned in real space,
// w = IDFT(w_hat) / N (real: w_hat is Hermitian).
static void kspace_stress_div(orc_ctx* c, int64_t step, int nset, double f1, double f4, std::vector<double>* w) {
  const int mx = c->mx, my = c->my, mz = c->mz, N = c->N;
  std::vector<cplx> S[6], Wk[3];
  kspace_sigma(c, step, nset, S);
  const double h[3] = {c->dx, c->dy, c->dz};
  const int m[3] = {mx, my, mz};
  std::vector<double> g[3];
  std::vector<cplx> e[3];
  for (int a = 0; a < 3; ++a) {
    g[a].resize(m[a]);
    e[a].resize(m[a]);
    for (int k = 0; k < m[a]; ++k) {
      g[a][k] = 2.0 / h[a] * sin(M_PI * k / m[a]);
      e[a][k] = std::polar(1.0, M_PI * k / m[a]);
    }
  }
  for (int a = 0; a < 3; ++a) Wk[a].assign(N, cplx(0.0, 0.0));
  const cplx I(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int kz = 0; kz < mz; ++kz)
    for (int ky = 0; ky < my; ++ky)
      for (int kx = 0; kx < mx; ++kx) {
        const size_t i = (size_t)kx + (size_t)mx * (ky + (size_t)my * kz);
        const cplx ex = e[0][kx], ey = e[1][ky], ez = e[2][kz];
        const double gx = g[0][kx], gy = g[1][ky], gz = g[2][kz];
        // S: 0 XX, 1 XY, 2 XZ, 3 YY, 4 YZ, 5 ZZ
        Wk[0][i] = I * (f1 * gx * ex * S[0][i] + f4 * gy * std::conj(ey) * S[1][i] + f4 * gz * std::conj(ez) * S[2][i]);
        Wk[1][i] = I * (f1 * gy * ey * S[3][i] + f4 * gx * std::conj(ex) * S[1][i] + f4 * gz * std::conj(ez) * S[4][i]);
        Wk[2][i] = I * (f1 * gz * ez * S[5][i] + f4 * gx * std::conj(ex) * S[2][i] + f4 * gy * std::conj(ey) * S[4][i]);
      }
  for (int a = 0; a < 3; ++a) {
    c->fft.run(Wk[a].data(), +1);
    w[a].resize(N);
    for (int i = 0; i < N; ++i) w[a][i] = Wk[a][i].real() / N;
  }
}

// The part of the stress divergence D.Sigma that the Stokes solve keeps, drawn directly.  The
// solve applies P_e = E P E* (E = diag(e_a), P = I - g g^T / |g|^2 with the real symbol g) to
// the source; for the stress of kspace_stress_div with f1^2 = 2 f4^2, D.Sigma has per-mode
// covariance N f4^2 E (|g|^2 I + g g^T) E*, so P_e D.Sigma has covariance N f4^2 |g|^2 P_e: two
// independent complex normals along an orthonormal basis t1, t2 of the plane normal to g,
//   u = f4 |g| E (Z1 t1 + Z2 t2),  E|Z|^2 = N,
// the same law as P_e D.Sigma with a third of the normals.  The basis (no branch on magnitudes):
// with k_a != 0 on all three axes, t1 = (g_y, -g_x, 0)/|.| and t2 = g x t1 / |g|; on two axes
// a < b (c the other), t1 = e_c and t2 = (-g_b e_a + g_a e_b)/|g|; on one axis a, the other two
// unit vectors in axis order.  Modes are canonicalised as in kspace_sigma (a partner is the
// conjugate of its canonical mode, which draws); at a self-conjugate mode a component along a
// basis vector on the k_a = 0 axes is real (Z = sqrt(N) n) and one on the k_a = m_a/2 axes
// imaginary (Z = i sqrt(N) n), so the field is real.  Counter: stream STREAM_FLUID_P, sub s for
// stress set s, index kx + P (ky + my kz) of the canonical mode; generic modes take
// Z1 = sqrt(N/2)(n0 + i n1), Z2 = sqrt(N/2)(n2 + i n3), self-conjugate ones n0 and n2.
static void kspace_noise(orc_ctx* c, int64_t step, int nset, double f4, std::vector<double>* u,
                         double* max_imag = nullptr) {
  const int mx = c->mx, my = c->my, mz = c->mz, N = c->N, P = mx / 2 + 1;
  const double h[3] = {c->dx, c->dy, c->dz};
  const int m[3] = {mx, my, mz};
  const double shn = sqrt(0.5 * (double)N), sn = sqrt((double)N);
  std::vector<cplx> U[3];
  for (int a = 0; a < 3; ++a) U[a].assign(N, cplx(0.0, 0.0));
  const cplx I(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int kz = 0; kz < mz; ++kz)
    for (int ky = 0; ky < my; ++ky)
      for (int kx = 0; kx < mx; ++kx) {
        int hx = kx, hy = ky, hz = kz;
        bool conj = false;
        if (2 * kx > mx) {
          hx = mx - kx;
          hy = (my - ky) % my;
          hz = (mz - kz) % mz;
          conj = true;
        }
        int cy = hy, cz = hz;
        bool selfc = false;
        if (hx == 0 || 2 * hx == mx) {
          const int py = (my - hy) % my, pz = (mz - hz) % mz;
          const int L = hy + my * hz, Lp = py + my * pz;
          if (Lp == L) {
            selfc = true;
          } else if (Lp < L) {
            cy = py;
            cz = pz;
            conj = !conj;
          }
        }
        const size_t i = (size_t)kx + (size_t)mx * (ky + (size_t)my * kz);
        const int kc[3] = {hx, cy, cz};
        if (hx == 0 && cy == 0 && cz == 0) continue;  // the mean: nothing to keep
        double g[3];
        cplx e[3];
        bool act[3];
        for (int a = 0; a < 3; ++a) {
          g[a] = 2.0 / h[a] * sin(M_PI * kc[a] / m[a]);
          e[a] = std::polar(1.0, M_PI * kc[a] / m[a]);
          act[a] = kc[a] != 0;
        }
        const double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
        int s1 = 1, s2 = 1;  // +1: along the k_a = 0 axes, -1: along the active ones
        const int nact = act[0] + act[1] + act[2];
        if (nact == 3) {
          const double r = sqrt(g[0] * g[0] + g[1] * g[1]);
          t1[0] = g[1] / r;
          t1[1] = -g[0] / r;
          t2[0] = (g[1] * t1[2] - g[2] * t1[1]) / gn;
          t2[1] = (g[2] * t1[0] - g[0] * t1[2]) / gn;
          t2[2] = (g[0] * t1[1] - g[1] * t1[0]) / gn;
          s1 = s2 = -1;
        } else if (nact == 2) {
          const int cc = !act[0] ? 0 : !act[1] ? 1 : 2;
          const int aa = cc == 0 ? 1 : 0, bb = cc == 2 ? 1 : 2;
          t1[cc] = 1.0;
          t2[aa] = -g[bb] / gn;
          t2[bb] = g[aa] / gn;
          s2 = -1;
        } else {
          const int aa = act[0] ? 0 : act[1] ? 1 : 2;
          const int b1 = aa == 0 ? 1 : 0, b2 = aa == 2 ? 1 : 2;
          t1[b1] = 1.0;
          t2[b2] = 1.0;
        }
        const uint32_t index = (uint32_t)(hx + P * (cy + my * cz));
        double n[4] = {0, 0, 0, 0};
        for (int st = 0; st < nset; ++st) {
          double d[4];
          orc_normals4(c->p.seed, step, STREAM_FLUID_P, st, index, d);
          for (int k = 0; k < 4; ++k) n[k] += d[k];
        }
        cplx Z1, Z2;
        if (selfc) {
          Z1 = (s1 > 0 ? cplx(1.0, 0.0) : I) * (sn * n[0]);
          Z2 = (s2 > 0 ? cplx(1.0, 0.0) : I) * (sn * n[2]);
        } else {
          Z1 = cplx(shn * n[0], shn * n[1]);
          Z2 = cplx(shn * n[2], shn * n[3]);
        }
        for (int a = 0; a < 3; ++a) {
          const cplx v = f4 * gn * e[a] * (Z1 * t1[a] + Z2 * t2[a]);
          U[a][i] = conj ? std::conj(v) : v;
        }
      }
  double mi = 0.0;
  for (int a = 0; a < 3; ++a) {
    c->fft.run(U[a].data(), +1);
    u[a].resize(N);
    for (int i = 0; i < N; ++i) {
      u[a][i] = U[a][i].real() / N;
      mi = std::max(mi, fabs(U[a][i].imag()) / N);
    }
  }
  if (max_imag) *max_imag = mi;
}

static void step_stokes(orc_ctx* c) {
  const int N = c->N, np = c->p.np;
  const double dt = c->p.dt, dV = c->dV;
  const double fact1 = sqrt(4.0 * c->p.temperature * c->p.shearviscosity / (dV * dt));
  const double fact4 = sqrt(2.0 * c->p.temperature * c->p.shearviscosity / (dV * dt));
  const double fact2 = sqrt(c->p.temperature * c->p.extra_mobility * dt);
  std::vector<double> Wr[3], P1[3], P2[3];
  for (int a = 0; a < 3; ++a) { Wr[a].resize(np); P1[a].resize(np); P2[a].resize(np); }
  for (int i = 0; i < np; ++i) {
    double w[4], p1[4], p2[4];
    orc_normals4(c->p.seed, c->step, STREAM_RFD, 0, (uint32_t)i, w);
    orc_normals4(c->p.seed, c->step, STREAM_PART1, 0, (uint32_t)i, p1);
    orc_normals4(c->p.seed, c->step, STREAM_PART2, 0, (uint32_t)i, p2);
    for (int a = 0; a < 3; ++a) { Wr[a][i] = w[a]; P1[a][i] = p1[a]; P2[a][i] = p2[a]; }
  }
  // 1. F^n and S F^n
  std::vector<double> f[3], W[3], v[3];
  particle_forces(c, c->q, c->F);
  for (int a = 0; a < 3; ++a) f[a].assign(N, 0.0);
  for (int i = 0; i < np; ++i) {
    double qp[3] = {c->q[0][i], c->q[1][i], c->q[2][i]};
    double Fp[3] = {c->F[0][i], c->F[1][i], c->F[2][i]};
    spread_one(c, qp, Fp, 1.0 / dV, f);
  }
  // 2. W = S F + sqrt(2) D.Sigma1 (the stress drawn in Fourier space) ; solve
  std::vector<double> nz[3];
  kspace_noise(c, c->step, 1, sqrt(2.0) * fact4, nz);
  for (int a = 0; a < 3; ++a) {
    W[a] = f[a];
    for (int i = 0; i < N; ++i) W[a][i] += nz[a][i];
  }
  solve(c, W, v, true, false);
  // 3. q^{n+1/2} = q^n + dt/2 J(q^n) v (+ extra mobility)
  std::vector<double> qh[3], Fh[3];
  for (int a = 0; a < 3; ++a) qh[a].resize(np);
  for (int i = 0; i < np; ++i) {
    double qp[3] = {c->q[0][i], c->q[1][i], c->q[2][i]}, up[3];
    interp_one(c, qp, v, up);
    for (int a = 0; a < 3; ++a) {
      double d = 0.5 * dt * up[a];
      if (c->p.extra_mobility != 0.0) d += 0.5 * dt * c->p.extra_mobility * c->F[a][i] + fact2 * P1[a][i];
      qh[a][i] = qp[a] + step_disp(c, a, d);
    }
  }
  // 4. F^{n+1/2}; f = S(q^{n+1/2}) F^{n+1/2} + (kT/delta)[S(q+) W - S(q-) W]
  particle_forces(c, qh, Fh);
  for (int a = 0; a < 3; ++a) f[a].assign(N, 0.0);
  for (int i = 0; i < np; ++i) {
    double qp[3] = {qh[0][i], qh[1][i], qh[2][i]};
    double Fp[3] = {Fh[0][i], Fh[1][i], Fh[2][i]};
    spread_one(c, qp, Fp, 1.0 / dV, f);
  }
  rfd_spread(c, c->q, Wr, f);
  // 5. W = f + sqrt(1/2) D.(Sigma1 + Sigma2); solve
  kspace_noise(c, c->step, 2, sqrt(0.5) * fact4, nz);
  for (int a = 0; a < 3; ++a) {
    W[a] = f[a];
    for (int i = 0; i < N; ++i) W[a][i] += nz[a][i];
  }
  solve(c, W, v, true, false);
  // 6. q^{n+1} = q^n + dt J(q^{n+1/2}) v (+ extra mobility)
  for (int i = 0; i < np; ++i) {
    double qp[3] = {qh[0][i], qh[1][i], qh[2][i]}, up[3];
    interp_one(c, qp, v, up);
    for (int a = 0; a < 3; ++a) {
      double d = dt * up[a];
      if (c->p.extra_mobility != 0.0) d += dt * c->p.extra_mobility * Fh[a][i] + fact2 * (P1[a][i] + P2[a][i]);
      c->q[a][i] += step_disp(c, a, d);
      c->u[a][i] = up[a];
    }
  }
  for (int a = 0; a < 3; ++a) { c->v[a] = v[a]; c->F[a] = Fh[a]; }
}

// ---- QNB (mass = 0): src/runSchemeQuasiNeutrallyBuoyant.cu:178-677 and the bootstrap
// src/firstStepQuasiNeutrallyBuoyant.cu:35-1152 ----
static void qnb_forcing(orc_ctx* c, const std::vector<double>* qh, std::vector<double>* f) {
  const int N = c->N, np = c->p.np;
  particle_forces(c, qh, c->F);
  for (int a = 0; a < 3; ++a) f[a].assign(N, 0.0);
  // m_f / (m_e + m_f) with m_f = 1.5 dV volumeParticle rho (src/quasiNeutrallyBuoyantFunctions.cu:6394)
  const double mf = 1.5 * c->dV * c->p.volume_particle * c->p.densfluid;
  const double ratio = (c->p.mass + mf) != 0.0 ? mf / (c->p.mass + mf) : 1.0;
  for (int i = 0; i < np; ++i) {
    double qp[3] = {qh[0][i], qh[1][i], qh[2][i]};
    double Fp[3] = {c->F[0][i], c->F[1][i], c->F[2][i]};
    spread_one(c, qp, Fp, 1.0 / c->dV, f);
    spread_drift_one(c, qp, c->p.temperature * ratio / c->dV, f);
  }
}

static void qnb_half_positions(orc_ctx* c, const std::vector<double>* v, double frac,
                               const std::vector<double>* qbase, const std::vector<double>* qeval,
                               std::vector<double>* out) {
  const int np = c->p.np;
  for (int a = 0; a < 3; ++a) out[a].resize(np);
  for (int i = 0; i < np; ++i) {
    double qp[3] = {qeval[0][i], qeval[1][i], qeval[2][i]}, up[3];
    interp_one(c, qp, v, up);
    for (int a = 0; a < 3; ++a) out[a][i] = qbase[a][i] + step_disp(c, a, frac * c->p.dt * up[a]);
  }
}

// Interpolation of a field at every particle of q: out[a][i] = (J(q) field)_a(i).
static void interp_all(const orc_ctx* c, const std::vector<double>* q, const std::vector<double>* field,
                       std::vector<double>* out) {
  const int np = c->p.np;
  for (int a = 0; a < 3; ++a) out[a].assign(np, 0.0);
  for (int i = 0; i < np; ++i) {
    double qp[3] = {q[0][i], q[1][i], q[2][i]}, up[3];
    interp_one(c, qp, field, up);
    for (int a = 0; a < 3; ++a) out[a][i] = up[a];
  }
}
// field = scale * sum_p phi(q_p) val_p (S with the 1/dV of the caller's scale).
static void spread_all(const orc_ctx* c, const std::vector<double>* q, const std::vector<double>* val,
                       double scale, std::vector<double>* field) {
  for (int a = 0; a < 3; ++a) field[a].assign(c->N, 0.0);
  for (int i = 0; i < c->p.np; ++i) {
    double qp[3] = {q[0][i], q[1][i], q[2][i]}, vp[3] = {val[0][i], val[1][i], val[2][i]};
    spread_one(c, qp, vp, scale, field);
  }
}

// Excess-mass corrections of one QNB step (mass != 0), after the unperturbed prediction vt
// (src/runSchemeQuasiNeutrallyBuoyant.cu:336-656, STEPS 4-9 and 11):
//   du  = dulag + (J(qh) - J(q)) v                      kernelCalculateDeltau
//   dp  = m_e (u - J(qh) vt - du)                       kernelCalculateDeltap
//   dvt = P [sum_p phi(qh) f dp], f = 1.5 V_p / (1.5 V_p dV rho + m_e)
//                                                       kernelSpreadDeltap, projectionDivergenceFree
//   g   = sum_p phi(qh) (dp - m_e J(qh) dvt) / (rho dV) kernelSpreadDeltapMinusJTildev
//   u'  = J(qh) vt + J(qh) dvt + du                     updateParticleVelocityme
//   dv  = (I - dt nu/2 L)^-1 P g                        kernelUpdateVIncompressible
// (src/quasiNeutrallyBuoyantFunctions2.cu:5-2290,2679-3149).  Returns dv; u_new receives u'.
// With `first`, the particle velocity is left as it is (the bootstrap predictor skips STEP 8,
// src/firstStepQuasiNeutrallyBuoyant.cu:365-388).
static void qnb_mass_correction(orc_ctx* c, const std::vector<double>* qh, const std::vector<double>* vt,
                                const std::vector<double>* dulag, std::vector<double>* dv,
                                std::vector<double>* u_new) {
  const int np = c->p.np;
  const double me = c->p.mass, rho = c->p.densfluid, dV = c->dV, Vp = c->p.volume_particle;
  std::vector<double> jh[3], j0[3], du[3], dp[3], jvt[3], dvt[3], jdvt[3], s[3], g[3];
  interp_all(c, qh, c->v, jh);
  interp_all(c, c->q, c->v, j0);
  interp_all(c, qh, vt, jvt);
  for (int a = 0; a < 3; ++a) {
    du[a].resize(np);
    dp[a].resize(np);
    for (int i = 0; i < np; ++i) {
      du[a][i] = dulag[a][i] + jh[a][i] - j0[a][i];
      dp[a][i] = me * (c->u[a][i] - jvt[a][i] - du[a][i]);
    }
  }
  const double f = 1.5 * Vp / (1.5 * Vp * dV * rho + me);  // src/quasiNeutrallyBuoyantFunctions2.cu:1332-1335
  spread_all(c, qh, dp, f, s);
  solve(c, s, dvt, false, true);
  interp_all(c, qh, dvt, jdvt);
  for (int a = 0; a < 3; ++a) {
    s[a].resize(np);
    u_new[a].resize(np);
    for (int i = 0; i < np; ++i) {
      s[a][i] = dp[a][i] - me * jdvt[a][i];
      u_new[a][i] = jvt[a][i] + jdvt[a][i] + du[a][i];
    }
  }
  spread_all(c, qh, s, 1.0 / (rho * dV), g);
  solve(c, g, dv, false, false);
}

static void step_qnb(orc_ctx* c) {
  const int N = c->N, np = c->p.np;
  const double dV = c->dV, rho = c->p.densfluid, dt = c->p.dt;
  const bool mass = c->p.mass != 0.0 && np > 0;
  const double fact1 = sqrt(4.0 * c->p.temperature * c->p.shearviscosity * dt / (dV * rho * rho));
  const double fact4 = sqrt(2.0 * c->p.temperature * c->p.shearviscosity * dt / (dV * rho * rho));
  std::vector<double> Z;
  fluid_noise(c->p.seed, c->step, 0, N, Z);
  std::vector<double> qh[3], f[3], W[3], A[3], vt[3], vh[3];
  if (mass && c->dulag[0].size() != (size_t)np)
    for (int a = 0; a < 3; ++a) c->dulag[a].assign(np, 0.0);
  // q^{n+1/2} = q^n + dt/2 J(q^n) v^n (src/quasiNeutrallyBuoyantFunctions.cu:2618-2624)
  qnb_half_positions(c, c->v, 0.5, c->q, c->q, qh);
  qnb_forcing(c, qh, f);
  // common part: v + lap(v) + D.Sigma + dt f / rho
  std::vector<double> base[3];
  lap_term(c, c->v, base);
  noise_div(c, Z, fact1, fact4, true, base);
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < N; ++i) base[a][i] += c->v[a][i] + dt * f[a][i] / rho;
  adv_term(c, c->v, A);
  if (!c->first_step_done) {
    // predictor with first-order advection (src/kernelConstructWQuasiNeutrallyBuoyant.cu:670)
    for (int a = 0; a < 3; ++a) {
      W[a].resize(N);
      for (int i = 0; i < N; ++i) W[a][i] = base[a][i] - A[a][i];
    }
    solve(c, W, vt, false, false);
    std::vector<double> dvp[3];
    if (mass) {
      // the predictor's excess-mass corrections, with no lagged term yet
      // (src/firstStepQuasiNeutrallyBuoyant.cu:84,229-424)
      std::vector<double> zero[3], unused[3];
      for (int a = 0; a < 3; ++a) zero[a].assign(np, 0.0);
      qnb_mass_correction(c, qh, vt, zero, dvp, unused);
    }
    for (int a = 0; a < 3; ++a) {
      vh[a].resize(N);
      for (int i = 0; i < N; ++i) vh[a][i] = 0.5 * (c->v[a][i] + vt[a][i] + (mass ? dvp[a][i] : 0.0));
    }
    if (mass) {
      // lagged term for the corrector: J(q') 1/2 dt nu L dv_p at q' = q^n + dt/2 J(q^{n+1/2}) v^{n+1/2}
      // (findNeighborParticlesQuasiNeutrallyBuoyantTEST4_3, laplacianDeltaV_2,
      // interpolateLaplacianDeltaV_2: src/firstStepQuasiNeutrallyBuoyant.cu:473-560)
      std::vector<double> qp2[3], ldv[3];
      qnb_half_positions(c, vh, 0.5, c->q, qh, qp2);
      lap_term(c, dvp, ldv);
      interp_all(c, qp2, ldv, c->dulag);
    }
    std::vector<double> Ah[3];
    adv_term(c, vh, Ah);  // calculateAdvectionFluid at v^{n+1/2} (src/calculateAdvectionFluid.cu:21-114)
    // corrector (TEST5_2, src/kernelConstructWQuasiNeutrallyBuoyant.cu:1339)
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < N; ++i) W[a][i] = base[a][i] - Ah[a][i];
    for (int a = 0; a < 3; ++a) c->advPrev[a] = A[a];
    c->first_step_done = true;
  } else {
    // Adams-Bashforth (TEST5_3, src/kernelConstructWQuasiNeutrallyBuoyant.cu:2010-2025)
    for (int a = 0; a < 3; ++a) {
      W[a].resize(N);
      for (int i = 0; i < N; ++i) W[a][i] = base[a][i] - 1.5 * A[a][i] + 0.5 * c->advPrev[a][i];
      c->advPrev[a] = A[a];
    }
  }
  solve(c, W, vt, false, false);
  std::vector<double> dv[3];
  if (mass) {
    std::vector<double> un[3];
    qnb_mass_correction(c, qh, vt, c->dulag, dv, un);
    for (int a = 0; a < 3; ++a) c->u[a] = un[a];
  } else {
    // particle velocity u = J(q^{n+1/2}) v~ (updateParticleVelocityme0)
    for (int i = 0; i < np; ++i) {
      double qp[3] = {qh[0][i], qh[1][i], qh[2][i]}, up[3];
      interp_one(c, qp, vt, up);
      for (int a = 0; a < 3; ++a) c->u[a][i] = up[a];
    }
  }
  // v^{n+1} = v~ + dv; v^{n+1/2} = (v^n + v^{n+1}) / 2; q^{n+1} = q^n + dt J(q^{n+1/2}) v^{n+1/2}
  // (calculateVelocityAtHalfTimeStep, updateFluidQuasiNeutrallyBuoyantSemiImplicit)
  if (mass)
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < N; ++i) vt[a][i] += dv[a][i];
  for (int a = 0; a < 3; ++a) {
    vh[a].resize(N);
    for (int i = 0; i < N; ++i) vh[a][i] = 0.5 * (c->v[a][i] + vt[a][i]);
  }
  std::vector<double> qn[3];
  qnb_half_positions(c, vh, 1.0, c->q, qh, qn);
  if (mass) {
    // next step's lagged term J(q^{n+1/2}) 1/2 dt nu L dv (laplacianDeltaV, interpolateLaplacianDeltaV,
    // src/quasiNeutrallyBuoyantFunctions.cu:5294-5795)
    std::vector<double> ldv[3];
    lap_term(c, dv, ldv);
    interp_all(c, qh, ldv, c->dulag);
  }
  for (int a = 0; a < 3; ++a) { c->q[a] = qn[a]; c->v[a] = vt[a]; }
}

extern "C" int orc_step(orc_ctx* c, int nsteps) {
  if (!c) return 0;
  for (int s = 0; s < nsteps; ++s) {
    if (c->p.scheme == FHD_SCHEME_INCOMPRESSIBLE) {
      if (!c->projected) project_initial(c);
      step_incompressible(c);
    } else if (c->p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) {
      if (!c->projected) project_initial(c);
      step_qnb(c);
    } else if (c->p.scheme == FHD_SCHEME_STOKES_LIMIT) {
      step_stokes(c);
    } else {
      return 0;
    }
    c->step++;
  }
  return 1;
}

// ---- stencil building blocks exposed for pinning tests ----
static void unpack3(const orc_ctx* c, const double* in, std::vector<double>* out) {
  for (int a = 0; a < 3; ++a) out[a].assign(in + (size_t)a * c->N, in + (size_t)(a + 1) * c->N);
}
static void pack3(const orc_ctx* c, const std::vector<double>* in, double* out) {
  for (int a = 0; a < 3; ++a) memcpy(out + (size_t)a * c->N, in[a].data(), sizeof(double) * c->N);
}
extern "C" int orc_laplacian_term(orc_ctx* c, const double* v, double* out) {
  std::vector<double> V[3], O[3];
  unpack3(c, v, V);
  lap_term(c, V, O);
  pack3(c, O, out);
  return 1;
}
extern "C" int orc_advection_term(orc_ctx* c, const double* v, double* out) {
  std::vector<double> V[3], O[3];
  unpack3(c, v, V);
  adv_term(c, V, O);
  pack3(c, O, out);
  return 1;
}
extern "C" int orc_noise_divergence(orc_ctx* c, const double* Z6, double fact1, double fact4,
                                    int trace_free, double* out) {
  std::vector<double> Z(Z6, Z6 + (size_t)6 * c->N), O[3];
  for (int a = 0; a < 3; ++a) O[a].assign(c->N, 0.0);
  noise_div(c, Z, fact1, fact4, trace_free != 0, O);
  pack3(c, O, out);
  return 1;
}
// The Stokes step's stress drawn in Fourier space: sigma6 = IDFT(Sigma_hat) / N (real part; the
// largest imaginary part goes to *max_imag) and w3 = its divergence, IDFT(w_hat) / N.
extern "C" int orc_kspace_stress(orc_ctx* c, int64_t step, int nset, double f1, double f4, double* sigma6,
                                 double* w3, double* max_imag) {
  std::vector<cplx> S[6];
  kspace_sigma(c, step, nset, S);
  double mi = 0.0;
  for (int k = 0; k < 6; ++k) {
    c->fft.run(S[k].data(), +1);
    for (int i = 0; i < c->N; ++i) {
      sigma6[(size_t)k * c->N + i] = S[k][i].real() / c->N;
      mi = std::max(mi, fabs(S[k][i].imag()) / c->N);
    }
  }
  *max_imag = mi;
  std::vector<double> w[3];
  kspace_stress_div(c, step, nset, f1, f4, w);
  pack3(c, w, w3);
  return 1;
}
extern "C" int orc_kspace_noise(orc_ctx* c, int64_t step, int nset, double f4, double* u3, double* max_imag) {
  std::vector<double> u[3];
  kspace_noise(c, step, nset, f4, u, max_imag);
  pack3(c, u, u3);
  return 1;
}
extern "C" int orc_fluid_noise(uint64_t seed, int64_t step, int which, int N, double* Z6) {
  std::vector<double> Z;
  fluid_noise(seed, step, which, N, Z);
  memcpy(Z6, Z.data(), sizeof(double) * 6 * (size_t)N);
  return 1;
}
