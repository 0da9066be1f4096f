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

// One wavefront per tile: rank each particle by its index within the tile, then gather.  The
// sorted copies carry the cell wrapped into the box (0 <= c < m), which is all the spatial
// passes need; the unwrapped cell stays in the per-particle arrays.
__global__ __launch_bounds__(256) void k_tile_sort(GridDev g, int ntiles, const int* __restrict__ start,
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
    int4 c = cell[id];
    sid[s0 + rank] = id;
    scell[s0 + rank] = make_int4(wrapi(c.x, g.mx), wrapi(c.y, g.my), wrapi(c.z, g.mz), 0);
    soff[s0 + rank] = off[id];
  }
}

// d = difference of two wrapped cell coordinates, |d| < m: nearest periodic image.
FHD_DEV int min_image(int d, int m) {
  if (2 * d > m) return d - m;
  if (2 * d < -m) return d + m;
  return d;
}

FHD_DEV void nb_range(int nt, int& lo, int& hi) {
  lo = nt >= 3 ? -1 : 0;
  hi = nt >= 2 ? 1 : 0;
}

// WCA: F_i = sum_j 48 eps [(s/r)^12 - (s/r)^6 / 2] / r^2 r_ij for r < rc.
FHD_DEV void wca_partial(const WcaArgs& a, int s, int q, float& fx, float& fy, float& fz);

// Four lanes per particle: lane q takes every fourth candidate of each neighbouring tile and
// the partial forces are combined by a fixed xor-shuffle tree, so the sum order is fixed.
__global__ __launch_bounds__(256) void k_wca(WcaArgs a) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = gid >> 2, q = gid & 3;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (s < a.np) wca_partial(a, s, q, fx, fy, fz);
  fx += __shfl_xor(fx, 1);
  fy += __shfl_xor(fy, 1);
  fz += __shfl_xor(fz, 1);
  fx += __shfl_xor(fx, 2);
  fy += __shfl_xor(fy, 2);
  fz += __shfl_xor(fz, 2);
  if (s < a.np && q == 0) a.F[a.set.id[s]] = make_float4(fx, fy, fz, 0.0f);
}

FHD_DEV void wca_partial(const WcaArgs& a, int s, int q, float& fx, float& fy, float& fz) {
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int4 ci = a.set.cell[s];
  const float4 di = a.set.off[s];
  const int tx = ci.x / t.tx, ty = ci.y / t.ty, tz = ci.z / t.tz;
  int lox, hix, loy, hiy, loz, hiz;
  nb_range(t.ntx, lox, hix);
  nb_range(t.nty, loy, hiy);
  nb_range(t.ntz, loz, hiz);
  // with three or more tiles along an axis, visit a neighbouring tile only when the particle
  // lies within the cutoff of that face
  const float rc = sqrtf(a.rc2);
  if (t.ntx >= 3) {
    const float lo = ((float)(ci.x - tx * t.tx) + 0.5f) * g.dx + di.x;
    if (lo >= rc) lox = 0;
    if ((float)t.tx * g.dx - lo >= rc) hix = 0;
  }
  if (t.nty >= 3) {
    const float lo = ((float)(ci.y - ty * t.ty) + 0.5f) * g.dy + di.y;
    if (lo >= rc) loy = 0;
    if ((float)t.ty * g.dy - lo >= rc) hiy = 0;
  }
  if (t.ntz >= 3) {
    const float lo = ((float)(ci.z - tz * t.tz) + 0.5f) * g.dz + di.z;
    if (lo >= rc) loz = 0;
    if ((float)t.tz * g.dz - lo >= rc) hiz = 0;
  }
  for (int dz = loz; dz <= hiz; ++dz)
    for (int dy = loy; dy <= hiy; ++dy)
      for (int dx = lox; dx <= hix; ++dx) {
        int nt = wrapi(tx + dx, t.ntx) + t.ntx * (wrapi(ty + dy, t.nty) + t.nty * wrapi(tz + dz, t.ntz));
        int j0 = a.set.start[nt], j1 = a.set.start[nt + 1];
        for (int j = j0 + q; j < j1; j += 4) {
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

// Fold a node index offset into [0, m) (the offsets here stay within a few boxes of the tile).
// Offsets passed here lie in (-m, 2m) whenever m >= 8 (wrapped cells, stencils within three
// cells, tiles of at least four cells); smaller boxes take the loop.
FHD_DEV int fold(int l, int m) {
  if (l < 0) l += m;
  else if (l >= m) l -= m;
  if (m < 8) {
    while (l < 0) l += m;
    while (l >= m) l -= m;
  }
  return l;
}

// Fixed-point image of a float for the LDS accumulators: the value is m 2^e with an integer
// mantissa m, shifted to an integer (truncating below 2^0); |x| must be below 2^62.  Exact and
// order-independent when summed as 64-bit integers.
FHD_DEV unsigned long long f2fixed(float x) {
  const unsigned b = __float_as_uint(x);
  const int be = (int)((b >> 23) & 255u);
  if (be == 0) return 0ull;  // zero or denormal
  const unsigned long long m = (unsigned long long)((b & 0x7FFFFFu) | 0x800000u);
  const int e = be - 150;
  unsigned long long v = e >= 0 ? (m << e) : (e > -24 ? (m >> -e) : 0ull);
  return (b >> 31) ? (0ull - v) : v;
}

// Candidate source staged in LDS: per axis the centred and face weights with their node bases,
// and the value (already multiplied by the fixed-point scale).
struct Cand {
  float w[2][3][4];   // [0: centred, 1: face][axis][k]
  int base[2][3];
  float val[3];
};

constexpr int kCandCap = 256;

FHD_DEV float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }

// Position (cell, offset) of source `sub` of a particle, plus its value.
FHD_DEV void rfd_normal(const Key& key, int id, float n[4]) {
  draw_normals4(key, (uint32_t)id, FHD_STREAM_RFD, 0, n);
}

// Candidate item of a spreading workgroup: group (1 bit) and source record index.
FHD_DEV int pack_item(int gi, int rec) { return (gi << 30) | rec; }

// Flattened source item -> record index (binary search over the neighbour-tile prefix sums
// held in LDS).
FHD_DEV int item_slot(const int* pref, const int* st, int n, int src) {
  int lo = 0, hi = n;  // find k with pref[k] <= src < pref[k+1]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pref[mid] <= src) lo = mid; else hi = mid;
  }
  return st[lo] + (src - pref[lo]);
}

// Does a stencil whose nodes span [lo, lo + NPTS] (both staggerings) reach the tile [o, o+T)?
template <int NPTS>
FHD_DEV bool reaches(int lo, int o, int T, int m, int ntiles) {
  if (ntiles == 1) return true;
  const int l0 = fold(lo - o, m);
  return l0 < T || l0 + NPTS >= m;
}

constexpr int kItemCap = 1024;   // compacted candidate records per round

// One record per source (sorted slot x sub-source): weights, bases, value and the cull data.
template <int NPTS>
__global__ __launch_bounds__(256) void k_src_prep(SpreadArgs a, int gi) {
  const SpreadGroup& G = a.grp[gi];
  const GridDev& g = a.g;
  const int nsub = G.mode == SRC_RFD ? 2 : 1;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= G.nsrc * nsub) return;
  const int sidx = r / nsub, sub = r - sidx * nsub;
  const int id = G.set.id[sidx];
  const int4 c = G.set.cell[sidx];
  const float4 d = G.set.off[sidx];
  float3 val;
  int4 cw = c;
  float3 dw = make_float3(d.x, d.y, d.z);
  if (G.mode == SRC_VALUE) {
    const float4 v = G.val[id];
    val = make_float3(v.x * G.coef, v.y * G.coef, v.z * G.coef);
  } else if (G.mode == SRC_DRIFT) {
    val = make_float3(G.coef * g.idx, G.coef * g.idy, G.coef * g.idz);
  } else {
    float nrm[4];
    rfd_normal(a.key, id, nrm);
    const float sg = sub == 0 ? 1.0f : -1.0f;
    val = make_float3(sg * G.coef * nrm[0], sg * G.coef * nrm[1], sg * G.coef * nrm[2]);
    dw = make_float3(d.x + sg * G.disp * nrm[0], d.y + sg * G.disp * nrm[1], d.z + sg * G.disp * nrm[2]);
    renorm_axis(cw.x, dw.x, g.dx, g.idx);
    renorm_axis(cw.y, dw.y, g.dy, g.idy);
    renorm_axis(cw.z, dw.z, g.dz, g.idz);
  }
  SrcWeights w;
  source_weights<NPTS>(cw, make_float3(dw.x * g.idx, dw.y * g.idy, dw.z * g.idz), w, G.mode == SRC_DRIFT);
  if (G.mode == SRC_RFD && G.rfd_quirk) {
    // nodes from the displaced position, weights at the undisplaced one
    const int cc[3] = {c.x, c.y, c.z};
    const float ff[3] = {d.x * g.idx, d.y * g.idy, d.z * g.idz};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      axis_weights_at<NPTS>(w.cen[ax].base, cc[ax], ff[ax], 0, w.cen[ax]);
      axis_weights_at<NPTS>(w.stg[ax].base, cc[ax], ff[ax], 1, w.stg[ax]);
    }
  }
  SrcRec rec;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    rec.base[0][ax] = w.cen[ax].base;
    rec.base[1][ax] = w.stg[ax].base;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      rec.w[0][ax][q] = w.cen[ax].w[q];
      rec.w[1][ax][q] = w.stg[ax].w[q];
    }
  }
  rec.val[0] = val.x;
  rec.val[1] = val.y;
  rec.val[2] = val.z;
  rec.pad0 = 0.0f;
  rec.pad1[0] = rec.pad1[1] = 0;
  const float4* src4 = reinterpret_cast<const float4*>(&rec);
  float4* dst4 = reinterpret_cast<float4*>(G.rec + r);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(SrcRec) / 16); ++k) dst4[k] = src4[k];
  float m = fmaxf(fabsf(val.x), fmaxf(fabsf(val.y), fabsf(val.z)));
  if (G.mode == SRC_DRIFT) m *= 2.0f;  // derivative weights reach ~1.6
  if (!(m <= 3.0e38f)) m = 3.0e38f;    // non-finite sources saturate
  G.cull[r] = make_int4(min(w.cen[0].base, w.stg[0].base), min(w.cen[1].base, w.stg[1].base),
                        min(w.cen[2].base, w.stg[2].base), (int)__float_as_uint(m));
}

// One workgroup per tile: cull the records of the sources in the neighbouring tiles, stage the
// candidates in LDS and accumulate every in-tile contribution as a 64-bit fixed-point integer.
template <int NPTS>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs a) {
  __shared__ unsigned long long acc[3 * kTileMax];
  __shared__ int nb_start[2][27], nb_pref[2][28];
  __shared__ int n_nb[2];
  __shared__ unsigned int maxbits;
  __shared__ Cand cand[kCandCap];
  __shared__ int items[kItemCap];
  __shared__ int nitem;
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int tile = blockIdx.x;
  const int tx = tile % t.ntx, ty = (tile / t.ntx) % t.nty, tz = tile / (t.ntx * t.nty);
  const int ox = tx * t.tx, oy = ty * t.ty, oz = tz * t.tz;
  const int ncell = t.tx * t.ty * t.tz;
  for (int e = threadIdx.x; e < 3 * kTileMax; e += blockDim.x) acc[e] = 0ull;
  // neighbour-tile table: one thread per (group, neighbour) loads its record range in parallel
  int lox, hix, loy, hiy, loz, hiz;
  nb_range(t.ntx, lox, hix);
  nb_range(t.nty, loy, hiy);
  nb_range(t.ntz, loz, hiz);
  const int nx = hix - lox + 1, ny = hiy - loy + 1, nz = hiz - loz + 1;
  const int nnb = nx * ny * nz;
  if ((int)threadIdx.x < a.ngroups * nnb) {
    const int gi = threadIdx.x / nnb, k = threadIdx.x - gi * nnb;
    const int nsub = a.grp[gi].mode == SRC_RFD ? 2 : 1;
    const int dx = lox + k % nx, dy = loy + (k / nx) % ny, dz = loz + k / (nx * ny);
    const int nt = wrapi(tx + dx, t.ntx) + t.ntx * (wrapi(ty + dy, t.nty) + t.nty * wrapi(tz + dz, t.ntz));
    const int s0 = a.grp[gi].set.start[nt];
    nb_start[gi][k] = s0 * nsub;
    nb_pref[gi][k + 1] = (a.grp[gi].set.start[nt + 1] - s0) * nsub;  // prefix-summed below
  }
  if (threadIdx.x == 0) {
    maxbits = 0u;
    nitem = 0;
  }
  __syncthreads();
  if ((int)threadIdx.x < a.ngroups) {
    const int gi = threadIdx.x;
    nb_pref[gi][0] = 0;
    for (int k = 0; k < nnb; ++k) nb_pref[gi][k + 1] += nb_pref[gi][k];
    n_nb[gi] = nnb;
  }
  __syncthreads();
  int nrec[2] = {0, 0};
  for (int gi = 0; gi < a.ngroups; ++gi) nrec[gi] = nb_pref[gi][n_nb[gi]];
  const int nall = nrec[0] + nrec[1];
  // pass 1: cull, compact the candidates into the LDS list (when they fit), largest magnitude
  float lmax = 0.0f;
  for (int e = threadIdx.x; e < nall; e += blockDim.x) {
    const int gi = e < nrec[0] ? 0 : 1;
    const int ri = item_slot(nb_pref[gi], nb_start[gi], n_nb[gi], gi == 0 ? e : e - nrec[0]);
    const int4 cu = a.grp[gi].cull[ri];
    if (!(reaches<NPTS>(cu.x, ox, t.tx, g.mx, t.ntx) && reaches<NPTS>(cu.y, oy, t.ty, g.my, t.nty) &&
          reaches<NPTS>(cu.z, oz, t.tz, g.mz, t.ntz)))
      continue;
    const int k = atomicAdd(&nitem, 1);
    if (k < kItemCap) items[k] = pack_item(gi, ri);
    lmax = fmaxf(lmax, __int_as_float(cu.w));
  }
  atomicMax(&maxbits, __float_as_uint(lmax));
  __syncthreads();
  const float M = __uint_as_float(maxbits);
  float inv = 0.0f;
  if (M > 0.0f) {
    int e2;
    frexpf(M * (float)(8 * nall + 1), &e2);
    e2 = max(e2, -60);
    const float scale = ldexpf(1.0f, 62 - e2);
    inv = ldexpf(1.0f, e2 - 62) * a.out_coef;
    // pass 2: the compacted list, or (when it overflowed) rounds of rescanned records
    const int nfirst = nitem;
    const bool fits = nfirst <= kItemCap;
    for (int r0 = 0; r0 < nall; r0 += kItemCap) {
      if (!fits) {
        __syncthreads();
        if (threadIdx.x == 0) nitem = 0;
        __syncthreads();
        const int r1 = min(nall, r0 + kItemCap);
        for (int e = r0 + threadIdx.x; e < r1; e += blockDim.x) {
          const int gi = e < nrec[0] ? 0 : 1;
          const int ri = item_slot(nb_pref[gi], nb_start[gi], n_nb[gi], gi == 0 ? e : e - nrec[0]);
          const int4 cu = a.grp[gi].cull[ri];
          if (reaches<NPTS>(cu.x, ox, t.tx, g.mx, t.ntx) && reaches<NPTS>(cu.y, oy, t.ty, g.my, t.nty) &&
              reaches<NPTS>(cu.z, oz, t.tz, g.mz, t.ntz))
            items[atomicAdd(&nitem, 1)] = pack_item(gi, ri);
        }
        __syncthreads();
      }
      const int ni = fits ? nfirst : nitem;
      for (int c0 = 0; c0 < ni; c0 += kCandCap) {
        const int nc = min(kCandCap, ni - c0);
        if ((int)threadIdx.x < nc) {
          const int it = items[c0 + threadIdx.x];
          const int gi = it >> 30, ri = it & ((1 << 30) - 1);
          const SrcRec* R = a.grp[gi].rec + ri;
          Cand& C = cand[threadIdx.x];
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            C.base[0][ax] = R->base[0][ax];
            C.base[1][ax] = R->base[1][ax];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              C.w[0][ax][q] = R->w[0][ax][q];
              C.w[1][ax][q] = R->w[1][ax][q];
            }
          }
          C.val[0] = R->val[0] * scale;
          C.val[1] = R->val[1] * scale;
          C.val[2] = R->val[2] * scale;
        }
        __syncthreads();
        // one work item per (candidate, component, z-row): NPTS x NPTS nodes
        const int nwork = nc * 3 * NPTS;
        for (int li = threadIdx.x; li < nwork; li += blockDim.x) {
          const int cidx = li / (3 * NPTS);
          const int r = li - cidx * (3 * NPTS);
          const int comp = r / NPTS, kz = r - comp * NPTS;
          const Cand& C = cand[cidx];
          const float vc = C.val[comp];
          if (vc == 0.0f) continue;
          const int sx = comp == 0, sy = comp == 1, sz = comp == 2;
          const int lz = fold(C.base[sz][2] + kz - oz, g.mz);
          if (lz >= t.tz) continue;
          const float wz = C.w[sz][2][kz] * vc;
          const int by = C.base[sy][1], bx = C.base[sx][0];
          int lxs[NPTS];
#pragma unroll
          for (int kx = 0; kx < NPTS; ++kx) lxs[kx] = fold(bx + kx - ox, g.mx);
#pragma unroll
          for (int ky = 0; ky < NPTS; ++ky) {
            const int ly = fold(by + ky - oy, g.my);
            if (ly >= t.ty) continue;
            const float wyz = C.w[sy][1][ky] * wz;
            unsigned long long* row = &acc[comp * kTileMax + t.tx * (ly + t.ty * lz)];
#pragma unroll
            for (int kx = 0; kx < NPTS; ++kx)
              if (lxs[kx] < t.tx) atomicAdd(&row[lxs[kx]], f2fixed(C.w[sx][0][kx] * wyz));
          }
        }
        __syncthreads();
      }
      if (fits) break;
    }
  }
  for (int e = threadIdx.x; e < 3 * ncell; e += blockDim.x) {
    const int comp = e / ncell, l = e - comp * ncell;
    const int lx = l % t.tx, ly = (l / t.tx) % t.ty, lz = l / (t.tx * t.ty);
    const size_t gi = (size_t)comp * g.N + (ox + lx) + (size_t)g.mx * ((oy + ly) + (size_t)g.my * (oz + lz));
    a.out[gi] = (float)(long long)acc[comp * kTileMax + l] * inv;
  }
}

// Sum over the NPTS x NPTS nodes of one z-row (kz) of a particle's stencil.
template <int NPTS>
FHD_DEV float interp_row(const float* __restrict__ f, const GridDev& g, const AxisWeights& wx,
                         const AxisWeights& wy, const AxisWeights& wz, int kz) {
  const int z = fold(wz.base + kz, g.mz);
  int xs[NPTS];
#pragma unroll
  for (int kx = 0; kx < NPTS; ++kx) xs[kx] = fold(wx.base + kx, g.mx);
  float sy = 0.0f;
#pragma unroll
  for (int ky = 0; ky < NPTS; ++ky) {
    const int y = fold(wy.base + ky, g.my);
    const float* row = f + (size_t)g.mx * (y + (size_t)g.my * z);
    float sx = 0.0f;
#pragma unroll
    for (int kx = 0; kx < NPTS; ++kx) sx += wx.w[kx] * row[xs[kx]];
    sy += wy.w[ky] * sx;
  }
  // kz is a lane index: select the weight instead of indexing (keeps wz in registers)
  const float wk = kz == 0 ? wz.w[0] : kz == 1 ? wz.w[1] : kz == 2 ? wz.w[2] : wz.w[3];
  return wk * sy;
}

// Four lanes per particle: lane q sums z-row q of the stencil; the rows are combined by a fixed
// xor-shuffle tree, then lane 0 applies the position update.
template <int NPTS>
__global__ __launch_bounds__(256) void k_interp(InterpArgs a) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = gid >> 2, q = gid & 3;
  const GridDev& g = a.g;
  float u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
  if (s < a.np && q < NPTS) {
    const int4 c = a.eval.cell[s];
    const float4 d = a.eval.off[s];
    SrcWeights w;
    source_weights<NPTS>(c, make_float3(d.x * g.idx, d.y * g.idy, d.z * g.idz), w, false);
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      const AxisWeights& wx = comp == 0 ? w.stg[0] : w.cen[0];
      const AxisWeights& wy = comp == 1 ? w.stg[1] : w.cen[1];
      const AxisWeights& wz = comp == 2 ? w.stg[2] : w.cen[2];
      if (a.f1) u1[comp] = interp_row<NPTS>(a.f1 + (size_t)comp * g.N, g, wx, wy, wz, q);
      if (a.f2) u2[comp] = interp_row<NPTS>(a.f2 + (size_t)comp * g.N, g, wx, wy, wz, q);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    u1[k] += __shfl_xor(u1[k], 1);
    u2[k] += __shfl_xor(u2[k], 1);
    u1[k] += __shfl_xor(u1[k], 2);
    u2[k] += __shfl_xor(u2[k], 2);
  }
  if (s >= a.np || q != 0) return;
  const int id = a.eval.id[s];
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
    const bool two = a.f2 != nullptr;
    a.u_out[id] = make_float4(two ? u2[0] : u1[0], two ? u2[1] : u1[1], two ? u2[2] : u1[2], 0.0f);
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
void launch_tile_sort(const GridDev& g, const Tiles& t, const int* start, const int* tmp, const int4* cell,
                      const float4* off, int* sid, int4* scell, float4* soff, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_sort, dim3(nblk(t.n, 4)), dim3(256), 0, s, g, t.n, start, tmp, cell, off, sid, scell,
                     soff);
}
void launch_wca(const WcaArgs& a, hipStream_t s) {
  if (a.np > 0) hipLaunchKernelGGL(k_wca, dim3(nblk(4 * a.np, 256)), dim3(256), 0, s, a);
}
void launch_spread(const SpreadArgs& a, hipStream_t s) {
  for (int gi = 0; gi < a.ngroups; ++gi) {
    const int nrec = a.grp[gi].nsrc * (a.grp[gi].mode == SRC_RFD ? 2 : 1);
    if (nrec <= 0) continue;
    if (a.kernel_points == 4) hipLaunchKernelGGL(k_src_prep<4>, dim3(nblk(nrec, 256)), dim3(256), 0, s, a, gi);
    else hipLaunchKernelGGL(k_src_prep<3>, dim3(nblk(nrec, 256)), dim3(256), 0, s, a, gi);
  }
  if (a.kernel_points == 4) hipLaunchKernelGGL(k_spread<4>, dim3(a.t.n), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_spread<3>, dim3(a.t.n), dim3(256), 0, s, a);
}
void launch_interp(const InterpArgs& a, hipStream_t s) {
  if (a.np <= 0) return;
  if (a.kernel_points == 4) hipLaunchKernelGGL(k_interp<4>, dim3(nblk(4 * a.np, 256)), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_interp<3>, dim3(nblk(4 * a.np, 256)), dim3(256), 0, s, a);
}
void launch_zero_f4(float4* p, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero_f4, dim3(nblk(n, 256)), dim3(256), 0, s, p, n);
}

}  // namespace fhd
