// particles.hip -- immersed-boundary particle kernels for gfx950.
//
//  * binning: particles are bucketed by an 8x8x8-cell tile of the fluid grid (counting sort),
//    then each bucket is ordered by particle index, so every later pass visits particles in a
//    fixed order whatever order the atomics placed them in.
//  * WCA forces from the particles of the neighbouring tiles (fluam's cell-list pair loop,
//    src/stokesLimitFunctions.cu:160-653, with the analytic potential instead of the float
//    texture table of src/initForcesNonBonded.cu:21-66).
//  * spreading S F: one workgroup per tile accumulates every contribution to its nodes in LDS
//    as 64-bit fixed-point integers (scale chosen per tile from the largest source).  Integer
//    addition is associative, so the result is bit-identical whatever order the contributions
//    arrive in: deterministic without global float atomics.  Replaces the 27-slot scatter +
//    per-cell gather of src/stokesLimitFunctions.cu:659-1158,1198-1681 and
//    src/kernelSpreadParticles.cu:3012-4082.
//  * interpolation J v and the particle position updates (src/stokesLimitFunctions.cu:
//    2479-2892,3513-3894; src/quasiNeutrallyBuoyantFunctions.cu:2448-2853,4481-4862).
//
// Positions are (unwrapped integer cell, FP32 offset from the cell centre), so kernel weights
// and pair distances are computed from small offsets and exact integer cell differences.
#include "fhd_device.h"
#include "fhd_kernels.h"

namespace fhd {

FHD_DEV int tile_of(const GridDev& g, const Tiles& t, int4 c) {
  int x = wrapi(c.x, g.mx) / t.tx, y = wrapi(c.y, g.my) / t.ty, z = wrapi(c.z, g.mz) / t.tz;
  return x + t.ntx * (y + t.nty * z);
}

__global__ void k_bin(GridDev g, Tiles t, const int4* __restrict__ cell, int* count, int* slot, int np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int tile = tile_of(g, t, cell[i]);
  slot[i] = atomicAdd(&count[tile], 1);
}

// Exclusive scan of the tile counts (single workgroup); resets the counts for the next binning.
__global__ __launch_bounds__(1024) void k_scan(int n, int* count, int* start) {
  __shared__ int part[1024];
  const int T = blockDim.x;
  const int chunk = (n + T - 1) / T;
  const int lo = threadIdx.x * chunk, hi = min(lo + chunk, n);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += count[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - s;
  for (int i = lo; i < hi; ++i) {
    int c = count[i];
    start[i] = run;
    run += c;
    count[i] = 0;
  }
  if (threadIdx.x == T - 1) start[n] = part[T - 1];
}

__global__ void k_scatter(GridDev g, Tiles t, const int4* __restrict__ cell, const int* __restrict__ slot,
                          const int* __restrict__ start, int* __restrict__ tmp, int np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  tmp[start[tile_of(g, t, cell[i])] + slot[i]] = i;
}

// One wavefront per tile: rank each particle by its index within the tile, then gather.
__global__ __launch_bounds__(256) void k_tile_sort(int ntiles, const int* __restrict__ start,
                                                   const int* __restrict__ tmp, const int4* __restrict__ cell,
                                                   const float4* __restrict__ off, int* __restrict__ sid,
                                                   int4* __restrict__ scell, float4* __restrict__ soff) {
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (tile >= ntiles) return;
  const int s0 = start[tile], n = start[tile + 1] - s0;
  for (int e = lane; e < n; e += 64) {
    int id = tmp[s0 + e];
    int rank = 0;
    for (int f = 0; f < n; ++f) rank += (tmp[s0 + f] < id);
    sid[s0 + rank] = id;
    scell[s0 + rank] = cell[id];
    soff[s0 + rank] = off[id];
  }
}

FHD_DEV int min_image(int d, int m) {
  int r = wrapi(d, m);
  return (2 * r > m) ? r - m : r;
}

FHD_DEV void nb_range(int nt, int& lo, int& hi) {
  lo = nt >= 3 ? -1 : 0;
  hi = nt >= 2 ? 1 : 0;
}

// WCA: F_i = sum_j 48 eps [(s/r)^12 - (s/r)^6 / 2] / r^2 r_ij for r < rc.
__global__ __launch_bounds__(256) void k_wca(WcaArgs a) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.np) return;
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int id = a.set.id[s];
  const int4 ci = a.set.cell[s];
  const float4 di = a.set.off[s];
  const int tx = wrapi(ci.x, g.mx) / t.tx, ty = wrapi(ci.y, g.my) / t.ty, tz = wrapi(ci.z, g.mz) / t.tz;
  int lox, hix, loy, hiy, loz, hiz;
  nb_range(t.ntx, lox, hix);
  nb_range(t.nty, loy, hiy);
  nb_range(t.ntz, loz, hiz);
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  for (int dz = loz; dz <= hiz; ++dz)
    for (int dy = loy; dy <= hiy; ++dy)
      for (int dx = lox; dx <= hix; ++dx) {
        int nt = wrapi(tx + dx, t.ntx) + t.ntx * (wrapi(ty + dy, t.nty) + t.nty * wrapi(tz + dz, t.ntz));
        int j0 = a.set.start[nt], j1 = a.set.start[nt + 1];
        for (int j = j0; j < j1; ++j) {
          if (j == s) continue;
          int4 cj = a.set.cell[j];
          float4 dj = a.set.off[j];
          float rx = (float)min_image(ci.x - cj.x, g.mx) * g.dx + (di.x - dj.x);
          float ry = (float)min_image(ci.y - cj.y, g.my) * g.dy + (di.y - dj.y);
          float rz = (float)min_image(ci.z - cj.z, g.mz) * g.dz + (di.z - dj.z);
          float r2 = rx * rx + ry * ry + rz * rz;
          if (r2 < a.rc2 && r2 > 0.0f) {
            // separations below sigma/100 are evaluated at sigma/100, so the force stays finite
            // (fluam's tabulated force is bounded near r = 0 too, src/initForcesNonBonded.cu:59)
            r2 = fmaxf(r2, 1.0e-4f * a.sig2);
            float s2 = a.sig2 / r2;
            float s6 = s2 * s2 * s2;
            float f = a.eps48 * (s6 * s6 - 0.5f * s6) / r2;
            fx += f * rx;
            fy += f * ry;
            fz += f * rz;
          }
        }
      }
  a.F[id] = make_float4(fx, fy, fz, 0.0f);
}

// Weights of one axis at an explicit node base (used for the reference-quirk RFD variant):
// node k sits at integer position base + k in node units; the particle at c + frac - stag/2.
template <int NPTS>
FHD_DEV void axis_weights_at(int base, int c, float frac, int stag, AxisWeights& aw) {
  aw.base = base;
  float f = (float)(c - base) + frac - 0.5f * (float)stag;
#pragma unroll
  for (int k = 0; k < 4; ++k) aw.w[k] = (k < NPTS) ? (NPTS == 3 ? phi3(f - (float)k) : phi4(f - (float)k)) : 0.0f;
}

struct SrcWeights {
  AxisWeights cen[3], stg[3];   // centred and face weights per axis
};

template <int NPTS>
FHD_DEV void source_weights(int4 c, float3 frac, SrcWeights& w, bool deriv) {
  int cc[3] = {c.x, c.y, c.z};
  float ff[3] = {frac.x, frac.y, frac.z};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    axis_weights<NPTS>(cc[ax], ff[ax], 0, w.cen[ax]);
    if (deriv) axis_dweights<NPTS>(cc[ax], ff[ax], 1, w.stg[ax]);
    else axis_weights<NPTS>(cc[ax], ff[ax], 1, w.stg[ax]);
  }
}

constexpr int kTileMax = 512;  // cells per spreading tile (8 x 8 x 8)

template <int NPTS>
FHD_DEV void accumulate_source(unsigned long long* acc, const GridDev& g, const Tiles& t, int ox, int oy, int oz,
                               const SrcWeights& w, float3 val, float scale, bool deriv) {
  const float vals[3] = {val.x, val.y, val.z};
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    float vc = vals[comp] * scale;
    if (vc == 0.0f) continue;
    const AxisWeights& wx = comp == 0 ? w.stg[0] : w.cen[0];
    const AxisWeights& wy = comp == 1 ? w.stg[1] : w.cen[1];
    const AxisWeights& wz = comp == 2 ? w.stg[2] : w.cen[2];
    (void)deriv;
    for (int kz = 0; kz < NPTS; ++kz) {
      int lz = wrapi(wz.base + kz - oz, g.mz);
      if (lz >= t.tz) continue;
      float wzv = wz.w[kz] * vc;
      for (int ky = 0; ky < NPTS; ++ky) {
        int ly = wrapi(wy.base + ky - oy, g.my);
        if (ly >= t.ty) continue;
        float wyz = wy.w[ky] * wzv;
#pragma unroll
        for (int kx = 0; kx < NPTS; ++kx) {
          int lx = wrapi(wx.base + kx - ox, g.mx);
          if (lx >= t.tx) continue;
          long long q = __float2ll_rn(wx.w[kx] * wyz);
          atomicAdd(&acc[comp * kTileMax + lx + t.tx * (ly + t.ty * lz)], (unsigned long long)q);
        }
      }
    }
  }
}

FHD_DEV float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }

// Position (cell, offset) of source `sub` of a particle, plus its value.
FHD_DEV void rfd_normal(const Key& key, int id, float n[4]) {
  draw_normals4(key, (uint32_t)id, FHD_STREAM_RFD, 0, n);
}

template <int NPTS>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs a) {
  __shared__ unsigned long long acc[3 * kTileMax];
  __shared__ int nb_start[2][27], nb_count[2][27], nb_pref[2][28];
  __shared__ int n_nb[2];
  __shared__ unsigned int maxbits;
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int tile = blockIdx.x;
  const int tx = tile % t.ntx, ty = (tile / t.ntx) % t.nty, tz = tile / (t.ntx * t.nty);
  const int ox = tx * t.tx, oy = ty * t.ty, oz = tz * t.tz;
  const int ncell = t.tx * t.ty * t.tz;
  for (int e = threadIdx.x; e < 3 * kTileMax; e += blockDim.x) acc[e] = 0ull;
  if (threadIdx.x == 0) {
    maxbits = 0u;
    int lox, hix, loy, hiy, loz, hiz;
    nb_range(t.ntx, lox, hix);
    nb_range(t.nty, loy, hiy);
    nb_range(t.ntz, loz, hiz);
    for (int gi = 0; gi < a.ngroups; ++gi) {
      int k = 0;
      nb_pref[gi][0] = 0;
      for (int dz = loz; dz <= hiz; ++dz)
        for (int dy = loy; dy <= hiy; ++dy)
          for (int dx = lox; dx <= hix; ++dx) {
            int nt = wrapi(tx + dx, t.ntx) + t.ntx * (wrapi(ty + dy, t.nty) + t.nty * wrapi(tz + dz, t.ntz));
            int s0 = a.grp[gi].set.start[nt];
            int cnt = a.grp[gi].set.start[nt + 1] - s0;
            nb_start[gi][k] = s0;
            nb_count[gi][k] = cnt;
            nb_pref[gi][k + 1] = nb_pref[gi][k] + cnt;
            ++k;
          }
      n_nb[gi] = k;
    }
  }
  __syncthreads();
  // pass 1: largest source magnitude (sets the fixed-point scale)
  float lmax = 0.0f;
  int nsrc_total = 0;
  for (int gi = 0; gi < a.ngroups; ++gi) {
    const SpreadGroup& G = a.grp[gi];
    const int tot = nb_pref[gi][n_nb[gi]];
    nsrc_total += tot * (G.mode == SRC_RFD ? 2 : 1);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      int k = 0;
      while (nb_pref[gi][k + 1] <= e) ++k;
      int sidx = nb_start[gi][k] + (e - nb_pref[gi][k]);
      float m;
      if (G.mode == SRC_VALUE) {
        float4 v = G.val[G.set.id[sidx]];
        m = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))) * fabsf(G.coef);
      } else if (G.mode == SRC_RFD) {
        float n[4];
        rfd_normal(a.key, G.set.id[sidx], n);
        m = fmaxf(fabsf(n[0]), fmaxf(fabsf(n[1]), fabsf(n[2]))) * fabsf(G.coef);
      } else {
        m = 2.0f * fabsf(G.coef) * fmaxf(g.idx, fmaxf(g.idy, g.idz));
      }
      if (!(m <= 3.0e38f)) m = 3.0e38f;  // non-finite sources saturate
      lmax = fmaxf(lmax, m);
    }
  }
  atomicMax(&maxbits, __float_as_uint(lmax));
  __syncthreads();
  const float M = __uint_as_float(maxbits);
  if (M > 0.0f) {
    int e2;
    frexpf(M * (float)(8 * nsrc_total + 1), &e2);
    e2 = max(e2, -60);
    const float scale = ldexpf(1.0f, 62 - e2);
    for (int gi = 0; gi < a.ngroups; ++gi) {
      const SpreadGroup& G = a.grp[gi];
      const int tot = nb_pref[gi][n_nb[gi]];
      const int nsub = G.mode == SRC_RFD ? 2 : 1;
      for (int e = threadIdx.x; e < tot * nsub; e += blockDim.x) {
        int src = e / nsub, sub = e - src * nsub;
        int k = 0;
        while (nb_pref[gi][k + 1] <= src) ++k;
        int sidx = nb_start[gi][k] + (src - nb_pref[gi][k]);
        int id = G.set.id[sidx];
        int4 c = G.set.cell[sidx];
        float4 d = G.set.off[sidx];
        SrcWeights w;
        float3 val;
        if (G.mode == SRC_VALUE) {
          float4 v = G.val[id];
          val = make_float3(v.x * G.coef, v.y * G.coef, v.z * G.coef);
          source_weights<NPTS>(c, make_float3(d.x * g.idx, d.y * g.idy, d.z * g.idz), w, false);
        } else if (G.mode == SRC_DRIFT) {
          // kT d/dq_c [phi phi phi]: derivative weights on the component's own axis
          source_weights<NPTS>(c, make_float3(d.x * g.idx, d.y * g.idy, d.z * g.idz), w, true);
          val = make_float3(G.coef * g.idx, G.coef * g.idy, G.coef * g.idz);
        } else {
          float n[4];
          rfd_normal(a.key, id, n);
          const float sg = sub == 0 ? 1.0f : -1.0f;
          val = make_float3(sg * G.coef * n[0], sg * G.coef * n[1], sg * G.coef * n[2]);
          int4 cd = c;
          float3 dd = make_float3(d.x + sg * G.disp * n[0], d.y + sg * G.disp * n[1], d.z + sg * G.disp * n[2]);
          renorm_axis(cd.x, dd.x, g.dx, g.idx);
          renorm_axis(cd.y, dd.y, g.dy, g.idy);
          renorm_axis(cd.z, dd.z, g.dz, g.idz);
          float3 fd = make_float3(dd.x * g.idx, dd.y * g.idy, dd.z * g.idz);
          source_weights<NPTS>(cd, fd, w, false);
          if (G.rfd_quirk) {
            // nodes from the displaced position, weights at the undisplaced one
            int cc[3] = {c.x, c.y, c.z};
            float ff[3] = {d.x * g.idx, d.y * g.idy, d.z * g.idz};
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
              axis_weights_at<NPTS>(w.cen[ax].base, cc[ax], ff[ax], 0, w.cen[ax]);
              axis_weights_at<NPTS>(w.stg[ax].base, cc[ax], ff[ax], 1, w.stg[ax]);
            }
          }
        }
        accumulate_source<NPTS>(acc, g, t, ox, oy, oz, w, val, scale, G.mode == SRC_DRIFT);
      }
    }
    __syncthreads();
    const float inv = ldexpf(1.0f, e2 - 62) * a.out_coef;
    for (int e = threadIdx.x; e < 3 * ncell; e += blockDim.x) {
      int comp = e / ncell, l = e - comp * ncell;
      int lx = l % t.tx, ly = (l / t.tx) % t.ty, lz = l / (t.tx * t.ty);
      size_t gi = (size_t)comp * g.N + (ox + lx) + (size_t)g.mx * ((oy + ly) + (size_t)g.my * (oz + lz));
      a.out[gi] = (float)(long long)acc[comp * kTileMax + l] * inv;
    }
  } else {
    for (int e = threadIdx.x; e < 3 * ncell; e += blockDim.x) {
      int comp = e / ncell, l = e - comp * ncell;
      int lx = l % t.tx, ly = (l / t.tx) % t.ty, lz = l / (t.tx * t.ty);
      size_t gi = (size_t)comp * g.N + (ox + lx) + (size_t)g.mx * ((oy + ly) + (size_t)g.my * (oz + lz));
      a.out[gi] = 0.0f;
    }
  }
}

template <int NPTS>
FHD_DEV float interp_comp(const float* __restrict__ f, const GridDev& g, const AxisWeights& wx,
                          const AxisWeights& wy, const AxisWeights& wz) {
  float s = 0.0f;
  for (int kz = 0; kz < NPTS; ++kz) {
    int z = wrapi(wz.base + kz, g.mz);
    float sy = 0.0f;
    for (int ky = 0; ky < NPTS; ++ky) {
      int y = wrapi(wy.base + ky, g.my);
      const float* row = f + (size_t)g.mx * (y + (size_t)g.my * z);
      float sx = 0.0f;
#pragma unroll
      for (int kx = 0; kx < NPTS; ++kx) sx += wx.w[kx] * row[wrapi(wx.base + kx, g.mx)];
      sy += wy.w[ky] * sx;
    }
    s += wz.w[kz] * sy;
  }
  return s;
}

template <int NPTS>
__global__ __launch_bounds__(256) void k_interp(InterpArgs a) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.np) return;
  const GridDev& g = a.g;
  const int id = a.eval.id[s];
  const int4 c = a.eval.cell[s];
  const float4 d = a.eval.off[s];
  SrcWeights w;
  source_weights<NPTS>(c, make_float3(d.x * g.idx, d.y * g.idy, d.z * g.idz), w, false);
  float u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const AxisWeights& wx = comp == 0 ? w.stg[0] : w.cen[0];
    const AxisWeights& wy = comp == 1 ? w.stg[1] : w.cen[1];
    const AxisWeights& wz = comp == 2 ? w.stg[2] : w.cen[2];
    if (a.f1) u1[comp] = interp_comp<NPTS>(a.f1 + (size_t)comp * g.N, g, wx, wy, wz);
    if (a.f2) u2[comp] = interp_comp<NPTS>(a.f2 + (size_t)comp * g.N, g, wx, wy, wz);
  }
  float disp[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) disp[k] = a.a1 * u1[k] + a.a2 * u2[k];
  if (a.F) {
    float4 F = a.F[id];
    disp[0] += a.aF * F.x;
    disp[1] += a.aF * F.y;
    disp[2] += a.aF * F.z;
  }
  if (a.nnoise > 0) {
    float n[4];
    draw_normals4(a.key, (uint32_t)id, FHD_STREAM_PART1, 0, n);
    float z[3] = {n[0], n[1], n[2]};
    if (a.nnoise > 1) {
      draw_normals4(a.key, (uint32_t)id, FHD_STREAM_PART2, 0, n);
      z[0] += n[0];
      z[1] += n[1];
      z[2] += n[2];
    }
    disp[0] += a.aN * z[0];
    disp[1] += a.aN * z[1];
    disp[2] += a.aN * z[2];
  }
  if (a.u_out) {
    const float* u = a.f2 ? u2 : u1;
    a.u_out[id] = make_float4(u[0], u[1], u[2], 0.0f);
  }
  if (a.out_cell) {
    int4 bc = a.base_cell[id];
    float4 bo = a.base_off[id];
    bool ok = isfinite(disp[0]) && isfinite(disp[1]) && isfinite(disp[2]);
    if (!ok) {
      atomicOr(a.err, 1);
      disp[0] = disp[1] = disp[2] = 0.0f;
    }
    float ox = bo.x + disp[0], oy = bo.y + disp[1], oz = bo.z + disp[2];
    renorm_axis(bc.x, ox, g.dx, g.idx);
    renorm_axis(bc.y, oy, g.dy, g.idy);
    renorm_axis(bc.z, oz, g.dz, g.idz);
    a.out_cell[id] = bc;
    a.out_off[id] = make_float4(ox, oy, oz, 0.0f);
  }
}

__global__ void k_zero_f4(float4* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ------------------------------- launchers -------------------------------
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

void launch_bin(const GridDev& g, const Tiles& t, const int4* cell, int* count, int* slot, int np, hipStream_t s) {
  if (np > 0) hipLaunchKernelGGL(k_bin, dim3(nblk(np, 256)), dim3(256), 0, s, g, t, cell, count, slot, np);
}
void launch_scan(const Tiles& t, int* count, int* start, hipStream_t s) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, t.n, count, start);
}
void launch_scatter(const GridDev& g, const Tiles& t, const int4* cell, const int* slot, const int* start, int* tmp,
                    int np, hipStream_t s) {
  if (np > 0) hipLaunchKernelGGL(k_scatter, dim3(nblk(np, 256)), dim3(256), 0, s, g, t, cell, slot, start, tmp, np);
}
void launch_tile_sort(const Tiles& t, const int* start, const int* tmp, const int4* cell, const float4* off, int* sid,
                      int4* scell, float4* soff, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_sort, dim3(nblk(t.n, 4)), dim3(256), 0, s, t.n, start, tmp, cell, off, sid, scell, soff);
}
void launch_wca(const WcaArgs& a, hipStream_t s) {
  if (a.np > 0) hipLaunchKernelGGL(k_wca, dim3(nblk(a.np, 256)), dim3(256), 0, s, a);
}
void launch_spread(const SpreadArgs& a, hipStream_t s) {
  if (a.kernel_points == 4) hipLaunchKernelGGL(k_spread<4>, dim3(a.t.n), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_spread<3>, dim3(a.t.n), dim3(256), 0, s, a);
}
void launch_interp(const InterpArgs& a, hipStream_t s) {
  if (a.np <= 0) return;
  if (a.kernel_points == 4) hipLaunchKernelGGL(k_interp<4>, dim3(nblk(a.np, 256)), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_interp<3>, dim3(nblk(a.np, 256)), dim3(256), 0, s, a);
}
void launch_zero_f4(float4* p, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero_f4, dim3(nblk(n, 256)), dim3(256), 0, s, p, n);
}

}  // namespace fhd
