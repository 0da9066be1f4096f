// fluid.hip -- the staggered-grid right-hand side W of each solve (fluam's kernelConstructW
// family) with the stochastic stress generated in registers.
//
// One workgroup owns a 16 x 16 column of cells and marches up a chunk of z-planes.  The 6
// stress normals of every cell of plane z+2 (plus the one-cell ring the divergence needs) are
// drawn from the counter-based generator into an LDS ring of four planes while plane z's
// stencil reads planes z-1..z+1, so each normal is drawn ~1.25 times instead of once per
// neighbour that reads it.
//
// Restates: kernelConstructW_1/_2 (src/kernelConstructW.cu:32-521), the QNB variants
// (src/kernelConstructWQuasiNeutrallyBuoyant.cu:26-705,1366-2025) and
// kernelConstructWstokesLimit/_2 (src/stokesLimitFunctions.cu:2281-2378,3922-4069).
#include "fhd_device.h"
#include "fhd_kernels.h"

namespace fhd {

constexpr int TX = 16, TY = 16;
constexpr int PX = TX + 2, PY = TY + 2;
constexpr int PLANE = PX * PY;

// stress slot components: 0 sXX, 1 XY, 2 XZ, 3 sYY, 4 YZ, 5 sZZ
FHD_DEV void gen_plane(const FluidWArgs& a, const Key& key, float* slot, int x0, int y0, int zz) {
  const GridDev& g = a.g;
  // zz in [-1, mz], x0 + lx in [-1, mx], y0 + ly in [-1, my]: one conditional wrap each
  const int z = zz < 0 ? zz + g.mz : (zz >= g.mz ? zz - g.mz : zz);
  const int nown = TX * TY, ntot = nown + 2 * TX + 2 * TY;
  for (int e = threadIdx.x; e < ntot; e += blockDim.x) {
    int lx, ly;
    if (e < nown) { lx = e % TX; ly = e / TX; }
    else {
      int r = e - nown;
      if (r < TY) { lx = -1; ly = r; }
      else if (r < 2 * TY) { lx = TX; ly = r - TY; }
      else if (r < 2 * TY + TX) { lx = r - 2 * TY; ly = -1; }
      else { lx = r - 2 * TY - TX; ly = TY; }
    }
    int x = x0 + lx, y = y0 + ly;
    x = x < 0 ? x + g.mx : (x >= g.mx ? x - g.mx : x);
    y = y < 0 ? y + g.my : (y >= g.my ? y - g.my : y);
    uint32_t cell = (uint32_t)(x + g.mx * (y + g.my * z));
    float n[12];
    draw_normals4(key, cell, FHD_STREAM_FLUID, 0, &n[0]);
    draw_normals4(key, cell, FHD_STREAM_FLUID, 1, &n[4]);
    float s[6];
    if (a.noise == 2) {
      draw_normals4(key, cell, FHD_STREAM_FLUID, 2, &n[8]);
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c] = n[c] + n[6 + c];
    } else {
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c] = n[c];
    }
    if (a.trace_free) {
      float tr = (s[0] + s[3] + s[5]) * (1.0f / 3.0f);
      s[0] -= tr;
      s[3] -= tr;
      s[5] -= tr;
    }
    int p = (lx + 1) + PX * (ly + 1);
#pragma unroll
    for (int c = 0; c < 6; ++c) slot[c * PLANE + p] = s[c];
  }
}

FHD_BODY void body_fluid_w(const FluidWArgs& a, const Blk& b) {
  // ring of four noise planes: while the stencil of plane z reads planes z - 1, z and z + 1,
  // plane z + 2 is drawn into the fourth slot, so one barrier per plane suffices
  constexpr int NR = 4;
  __shared__ float nbuf[NR][6 * PLANE];
  const GridDev& g = a.g;
  const Key key = resolve_key(a.key);
  const int x0 = b.x * TX, y0 = b.y * TY;
  const int zs = b.z * a.zchunk;
  const int ze = min(zs + a.zchunk, g.mz);
  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  const int x = x0 + lx, y = y0 + ly;
  const bool own = (x < g.mx) && (y < g.my);
  const int N = g.N;
  if (a.noise) {
    gen_plane(a, key, nbuf[(zs - 1 + NR) % NR], x0, y0, zs - 1);
    gen_plane(a, key, nbuf[zs % NR], x0, y0, zs);
    gen_plane(a, key, nbuf[(zs + 1) % NR], x0, y0, zs + 1);
  }
  __syncthreads();
  const int xm = wrapi(x - 1, g.mx), xp = wrapi(x + 1, g.mx);
  const int ym = wrapi(y - 1, g.my), yp = wrapi(y + 1, g.my);
  for (int z = zs; z < ze; ++z) {
    if (a.noise && z + 2 <= ze) gen_plane(a, key, nbuf[(z + 2) % NR], x0, y0, z + 2);
    if (own) {
      const int zm = wrapi(z - 1, g.mz), zp = wrapi(z + 1, g.mz);
      const int i = x + g.mx * (y + g.my * z);
      float wx = 0.0f, wy = 0.0f, wz = 0.0f;
      if (a.noise) {
        const float* cur = nbuf[z % NR];
        const float* prv = nbuf[(z - 1 + NR) % NR];
        const float* nxt = nbuf[(z + 1) % NR];
        const int p = (lx + 1) + PX * (ly + 1);
        const float f1 = a.f1, f4 = a.f4;
        wx = g.idx * f1 * (cur[0 * PLANE + p + 1] - cur[0 * PLANE + p]) +
             g.idy * f4 * (cur[1 * PLANE + p] - cur[1 * PLANE + p - PX]) +
             g.idz * f4 * (cur[2 * PLANE + p] - prv[2 * PLANE + p]);
        wy = g.idy * f1 * (cur[3 * PLANE + p + PX] - cur[3 * PLANE + p]) +
             g.idx * f4 * (cur[1 * PLANE + p] - cur[1 * PLANE + p - 1]) +
             g.idz * f4 * (cur[4 * PLANE + p] - prv[4 * PLANE + p]);
        wz = g.idz * f1 * (nxt[5 * PLANE + p] - cur[5 * PLANE + p]) +
             g.idx * f4 * (cur[2 * PLANE + p] - cur[2 * PLANE + p - 1]) +
             g.idy * f4 * (cur[4 * PLANE + p] - cur[4 * PLANE + p - PX]);
      }
      const int n0 = x + g.mx * (y + g.my * zm), n5 = x + g.mx * (y + g.my * zp);
      const int n1 = x + g.mx * (ym + g.my * z), n4 = x + g.mx * (yp + g.my * z);
      const int n2 = xm + g.mx * (y + g.my * z), n3 = xp + g.mx * (y + g.my * z);
      if (a.v) {
        const float* V = a.v;
        const float ax = g.idx * g.idx, ay = g.idy * g.idy, az = g.idz * g.idz;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float* f = V + (size_t)c * N;
          float fc = f[i];
          float L = ax * (f[n3] - 2.0f * fc + f[n2]) + ay * (f[n4] - 2.0f * fc + f[n1]) +
                    az * (f[n5] - 2.0f * fc + f[n0]);
          float t = fc + a.lap_coef * L;
          if (c == 0) wx += t; else if (c == 1) wy += t; else wz += t;
        }
      }
      if (a.vadv) {
        const float* VX = a.vadv;
        const float* VY = a.vadv + N;
        const float* VZ = a.vadv + 2 * (size_t)N;
        const int pxmy = xp + g.mx * (ym + g.my * z), pxmz = xp + g.mx * (y + g.my * zm);
        const int mxpy = xm + g.mx * (yp + g.my * z), mxpz = xm + g.mx * (y + g.my * zp);
        const int pymz = x + g.mx * (yp + g.my * zm), mypz = x + g.mx * (ym + g.my * zp);
        float vx = VX[i], vy = VY[i], vz = VZ[i];
        float vx0 = VX[n0], vx1 = VX[n1], vx2 = VX[n2], vx3 = VX[n3], vx4 = VX[n4], vx5 = VX[n5];
        float vy0 = VY[n0], vy1 = VY[n1], vy2 = VY[n2], vy3 = VY[n3], vy4 = VY[n4], vy5 = VY[n5];
        float vz0 = VZ[n0], vz1 = VZ[n1], vz2 = VZ[n2], vz3 = VZ[n3], vz4 = VZ[n4], vz5 = VZ[n5];
        // advection (dt/4) sum_b D_b[(v_a + v_a')(v_b + v_b')], src/kernelConstructW.cu:120-136
        float Ax = g.idx * ((vx3 + vx) * (vx3 + vx) - (vx + vx2) * (vx + vx2)) +
                   g.idy * ((vx4 + vx) * (vy3 + vy) - (vx + vx1) * (VY[pxmy] + vy1)) +
                   g.idz * ((vx5 + vx) * (vz3 + vz) - (vx + vx0) * (VZ[pxmz] + vz0));
        float Ay = g.idx * ((vy3 + vy) * (vx4 + vx) - (vy + vy2) * (VX[mxpy] + vx2)) +
                   g.idy * ((vy4 + vy) * (vy4 + vy) - (vy + vy1) * (vy + vy1)) +
                   g.idz * ((vy5 + vy) * (vz4 + vz) - (vy + vy0) * (VZ[pymz] + vz0));
        float Az = g.idx * ((vz3 + vz) * (vx5 + vx) - (vz + vz2) * (VX[mxpz] + vx2)) +
                   g.idy * ((vz4 + vz) * (vy5 + vy) - (vz + vz1) * (VY[mypz] + vy1)) +
                   g.idz * ((vz5 + vz) * (vz5 + vz) - (vz + vz0) * (vz + vz0));
        const float q = 0.25f;  // the dt factor is folded into adv_coef by the caller
        Ax *= q; Ay *= q; Az *= q;
        wx += a.adv_coef * Ax;
        wy += a.adv_coef * Ay;
        wz += a.adv_coef * Az;
        if (a.adv_store) {
          a.adv_store[i] = Ax;
          a.adv_store[i + N] = Ay;
          a.adv_store[i + 2 * (size_t)N] = Az;
        }
      }
      if (a.adv_prev) {
        wx += a.adv_prev_coef * a.adv_prev[i];
        wy += a.adv_prev_coef * a.adv_prev[i + N];
        wz += a.adv_prev_coef * a.adv_prev[i + 2 * (size_t)N];
      }
      if (a.force) {
        wx += a.force_coef * a.force[i];
        wy += a.force_coef * a.force[i + N];
        wz += a.force_coef * a.force[i + 2 * (size_t)N];
      }
      a.W[i] = wx;
      a.W[i + N] = wy;
      a.W[i + 2 * (size_t)N] = wz;
    }
    __syncthreads();
  }
}
FHD_KERNEL_PAIR((), 256, k_fluid_w, FluidWArgs, body_fluid_w(a, b))

// Discrete staggered divergence (diagnostic), src/kernelUpdateVIncompressible.cu invariant.
__global__ void k_divergence(GridDev g, const float* __restrict__ v, float* __restrict__ div) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  int x = i % g.mx, y = (i / g.mx) % g.my, z = i / (g.mx * g.my);
  int xm = wrapi(x - 1, g.mx), ym = wrapi(y - 1, g.my), zm = wrapi(z - 1, g.mz);
  const float* VX = v;
  const float* VY = v + g.N;
  const float* VZ = v + 2 * (size_t)g.N;
  div[i] = (VX[i] - VX[xm + g.mx * (y + g.my * z)]) * g.idx + (VY[i] - VY[x + g.mx * (ym + g.my * z)]) * g.idy +
           (VZ[i] - VZ[x + g.mx * (y + g.my * zm)]) * g.idz;
}

// Excess-mass fluid update: lap = coef L dv (seven-point staggered Laplacian per component,
// laplacianDeltaV, src/quasiNeutrallyBuoyantFunctions.cu:5294-5386) and v2 = al v + be v2 + ga dv
// (calculateVelocityAtHalfTimeStep / updateFluidQuasiNeutrallyBuoyantSemiImplicit).
FHD_BODY void body_mass_fluid(const MassFluidArgs& a, const Blk& b) {
  const GridDev& g = a.g;
  const size_t e = (size_t)b.x * blockDim.x + threadIdx.x;
  if (e >= 3 * (size_t)g.N) return;
  const int C = (int)(e / g.N);
  const int i = (int)(e - (size_t)C * g.N);
  const int x = i % g.mx, y = (i / g.mx) % g.my, z = i / (g.mx * g.my);
  const float* f = a.dv + (size_t)C * g.N;
  auto at = [&](int xx, int yy, int zz) { return f[wrapi(xx, g.mx) + g.mx * (wrapi(yy, g.my) + g.my * wrapi(zz, g.mz))]; };
  const float c0 = f[i];
  const float L = g.idx * g.idx * (at(x + 1, y, z) - 2.0f * c0 + at(x - 1, y, z)) +
                  g.idy * g.idy * (at(x, y + 1, z) - 2.0f * c0 + at(x, y - 1, z)) +
                  g.idz * g.idz * (at(x, y, z + 1) - 2.0f * c0 + at(x, y, z - 1));
  a.lap[e] = a.coef * L;
  if (a.v2) a.v2[e] = a.al * a.v[e] + a.be * a.v2[e] + a.ga * c0;
}
FHD_KERNEL_PAIR((), 256, k_mass_fluid, MassFluidArgs, body_mass_fluid(a, b))

void launch_mass_fluid(const MassFluidArgs& a, hipStream_t s) {
  launch(k_mass_fluid, k_mass_fluid_b, FAM_MASS_FLUID, dim3((unsigned)((3 * (size_t)a.g.N + 255) / 256)), 256, a, s);
}

// Normals of the shared generator, for parity tests: out[4 i + k] = draw(index i)[k].
__global__ void k_normals(Key key, int stream, int sub, int n, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v[4];
  draw_normals4(key, (uint32_t)i, (uint32_t)stream, (uint32_t)sub, v);
  for (int k = 0; k < 4; ++k) out[4 * i + k] = v[k];
}

FHD_BODY void body_zero_f(const FillArgs& a, const Blk& b) {
  const size_t i = (size_t)b.x * blockDim.x + threadIdx.x;
  if (i < a.n) a.p[i] = 0.0f;
}
FHD_KERNEL_PAIR((), 256, k_zero_f, FillArgs, body_zero_f(a, b))

void launch_normals(const Key& key, int stream, int sub, int n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_normals, dim3((n + 255) / 256), dim3(256), 0, s, key, stream, sub, n, out);
}

void launch_zero_f(float* p, size_t n, hipStream_t s) {
  const FillArgs a{p, n};
  launch(k_zero_f, k_zero_f_b, FAM_ZERO, dim3((unsigned)((n + 255) / 256)), 256, a, s);
}

// Advances the device-side step counter that graph-replayed steps key their noise with.
FHD_BODY void body_advance_step(const StepArgs& a) {
  if (threadIdx.x == 0) *a.step += 1ull;
}
FHD_KERNEL_PAIR((), 64, k_advance_step, StepArgs, body_advance_step(a))
void launch_advance_step(unsigned long long* step, hipStream_t s) {
  const StepArgs a{step};
  launch(k_advance_step, k_advance_step_b, FAM_STEP, dim3(1), 64, a, s);
}

void launch_fluid_w(const FluidWArgs& a, hipStream_t s) {
  FluidWArgs x = a;
  if (t_recorder && x.zchunk_b > 0) x.zchunk = x.zchunk_b;
  dim3 grid((x.g.mx + TX - 1) / TX, (x.g.my + TY - 1) / TY, (x.g.mz + x.zchunk - 1) / x.zchunk);
  launch(k_fluid_w, k_fluid_w_b, FAM_FLUID_W, grid, 256, x, s);
}

void launch_divergence(const GridDev& g, const float* v, float* div, hipStream_t s) {
  hipLaunchKernelGGL(k_divergence, dim3((g.N + 255) / 256), dim3(256), 0, s, g, v, div);
}

// Equal-time structure factor sample: acc[i] += |spec[i]|^2 in FP64 over the half spectrum of
// the three velocity components (the in-situ counterpart of the HydroGrid static structure
// factor that fluam's analysis hook feeds, src/hydroAnalysisIncompressible.cu:35-92).  The
// phase of the staggering shift drops out of |.|^2, so the plain transform of each component
// is used.  HBM-bound: 8 B read + 16 B read-modify-write per spectral entry.
__global__ __launch_bounds__(256) void k_sk_accum(const float2* __restrict__ spec, double* __restrict__ acc,
                                                 long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float2 v = spec[i];
    acc[i] += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
}
void launch_sk_accum(const float2* spec, double* acc, size_t n, hipStream_t s) {
  if (n == 0) return;
  long long blocks = ((long long)n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_sk_accum, dim3((unsigned)blocks), dim3(256), 0, s, spec, acc, (long long)n);
}

}  // namespace fhd
