// particles.hip -- immersed-boundary particle kernels for gfx950.
//
//  * binning: particles are bucketed by an 8x8x8-cell tile of the fluid grid (counting sort,
//    two launches).  The order inside a bucket is the order the atomics handed out, so every
//    later sum over a tile's particles is formed in fixed point, where it does not depend on
//    that order.  The sort is kept between passes: a particle stays listed in its tile until
//    it leaves the tile's skin, and the sort launches do nothing until a flag says one has.
//  * WCA forces, one workgroup per tile, from the particles of the neighbouring tiles staged
//    in LDS (fluam's cell-list pair loop, src/stokesLimitFunctions.cu:160-653, with the
//    analytic potential instead of the float texture table of src/initForcesNonBonded.cu:21-66).
//  * spreading S F: one workgroup per tile adds the contributions of every source whose
//    stencil reaches the tile into a 64-bit fixed-point LDS accumulator.  Integer addition is
//    associative, so the field is bit-identical whatever order the contributions arrive in:
//    deterministic without float atomics.  Replaces the 27-slot scatter + per-cell gather of
//    src/stokesLimitFunctions.cu:659-1158,1198-1681 and src/kernelSpreadParticles.cu:3012-4082.
//  * interpolation J v and the particle position updates (src/stokesLimitFunctions.cu:
//    2479-2892,3513-3894; src/quasiNeutrallyBuoyantFunctions.cu:2448-2853,4481-4862).
//
// Positions are (unwrapped integer cell, FP32 offset from the cell centre), so kernel weights
// and pair distances are computed from small offsets and exact integer cell differences.
#include "fhd_device.h"
#include "fhd_kernels.h"

namespace fhd {

FHD_DEV int wave_incl_scan(int v);

FHD_DEV int tile_of(const GridDev& g, const Tiles& t, int4 c) {
  int x = wrapi(c.x, g.mx) / t.tx, y = wrapi(c.y, g.my) / t.ty, z = wrapi(c.z, g.mz) / t.tz;
  return x + t.ntx * (y + t.nty * z);
}

FHD_DEV bool sort_on(const SortArgs& a) { return a.gate == nullptr || *a.gate != 0; }

FHD_BODY void body_bin(const SortArgs& a, const Blk& b) {
  int i = b.x * blockDim.x + threadIdx.x;
  if (i >= a.np || !sort_on(a)) return;
  int tile = tile_of(a.g, a.t, a.cell[i]);
  a.slot[i] = atomicAdd(&a.count[tile], 1);
}
FHD_KERNEL_PAIR((), 256, k_bin, SortArgs, body_bin(a, b))

// Sorted copy of particle i at slot pos of tile `tile`.
FHD_DEV void place(const SortArgs& a, int i, int pos, int tile) {
  const int4 c = a.cell[i];
  a.sid[pos] = i;
  a.scell[pos] = make_int4(wrapi(c.x, a.g.mx), wrapi(c.y, a.g.my), wrapi(c.z, a.g.mz), tile);
  a.soff[pos] = a.off[i];
  if (a.slot_of) a.slot_of[i] = pos;
}

// Exclusive scan of the tile counts (single workgroup, large grids); resets the counts for the
// next binning.
FHD_BODY void body_scan(const SortArgs& a) {
  if (!sort_on(a)) return;
  const int n = a.t.n;
  int* __restrict__ count = a.count;
  int* __restrict__ start = a.start;
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
FHD_KERNEL_PAIR((), 1024, k_scan, SortArgs, body_scan(a))

FHD_BODY void body_scatter(const SortArgs& a, const Blk& b) {
  int i = b.x * blockDim.x + threadIdx.x;
  if (i >= a.np || !sort_on(a)) return;
  const int tile = tile_of(a.g, a.t, a.cell[i]);
  place(a, i, a.start[tile] + a.slot[i], tile);
}
FHD_KERNEL_PAIR((), 256, k_scatter, SortArgs, body_scatter(a, b))

// Scan + scatter in one launch for up to kScanTilesMax tiles: every workgroup forms the whole
// exclusive prefix of the tile counts (a few KB, read from L2) in LDS, workgroup 0 stores it as
// the tile starts, and each thread places its particle's sorted copy.  This set's counts are
// still being read, so the launch clears the other set's counts instead (their next binning
// comes after this set's): the clearing runs whether or not the sort itself is gated off.
constexpr int kScanTilesMax = 8192;
FHD_BODY void body_scan_scatter(const SortArgs& a, const Blk& b) {
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int4* __restrict__ cell = a.cell;
  const int* __restrict__ slot = a.slot;
  const int* __restrict__ count = a.count;
  int* __restrict__ start = a.start;
  const int np = a.np;
  __shared__ int pre[kScanTilesMax + 1];
  __shared__ int wsum[4];
  const int n = t.n, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int gtid = b.x * blockDim.x + tid, gsz = b.nx * blockDim.x;
  for (int k = gtid; k < n; k += gsz) a.count_other[k] = 0;
  if (!sort_on(a)) return;  // uniform over the grid
  const int ch = (n + 255) / 256;
  const int lo = min(n, tid * ch), hi = min(n, lo + ch);
  int sum = 0;
  for (int k = lo; k < hi; ++k) sum += count[k];
  const int incl = wave_incl_scan(sum);
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  int run = incl - sum;
  for (int w = 0; w < wv; ++w) run += wsum[w];
  for (int k = lo; k < hi; ++k) {
    pre[k] = run;
    run += count[k];
  }
  if (tid == 255) pre[n] = run;
  __syncthreads();
  if (b.x == 0)
    for (int k = tid; k <= n; k += 256) start[k] = pre[k];
  const int i = b.x * blockDim.x + tid;
  if (i < np) {
    const int tile = tile_of(g, t, cell[i]);
    place(a, i, pre[tile] + slot[i], tile);
  }
}
FHD_KERNEL_PAIR((), 256, k_scan_scatter, SortArgs, body_scan_scatter(a, b))

// Grid barrier for a launch whose workgroups are all resident at once.  The counter only grows
// (64 bits; never reset, so repeated launches and graph replays reuse it): each workgroup takes
// a ticket and waits until the counter reaches the end of its group of nblocks tickets.  The
// wait is bounded (2^22 polls of about 0.2 us); returns false if it ran out.
FHD_DEV bool grid_barrier(unsigned long long* ctr, unsigned nblocks) {
  __shared__ int met;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned long long ticket = atomicAdd(ctr, 1ull);
    const unsigned long long target = ticket - ticket % nblocks + nblocks;
    bool done = false;
    for (int spin = 0; spin < (1 << 22) && !done; ++spin) {
      done = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target;
      if (!done) __builtin_amdgcn_s_sleep(8);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    met = done;
  }
  __syncthreads();
  return met != 0;
}

// The sort in one launch, for a context launched on its own (a merged batch dispatch holds
// more workgroups than can be resident, so it keeps k_bin + k_scan_scatter).  At most
// kSortGroups workgroups, one per CU at most, all resident: they bin their particles
// (grid-stride), meet at a grid barrier, then every workgroup forms the whole prefix of the tile
// counts in LDS, workgroup 0 stores the tile starts, and each thread places its particles.  The
// other set's counts are cleared whether or not the sort runs, as in k_scan_scatter.
constexpr int kSortGroups = 256;
// The sort's work; returns 0 when the gate was off, 1 when it sorted, -1 when a barrier's wait
// ran out (the error flag raised).
FHD_DEV int sort_one(const SortArgs& a, const Blk& b) {
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int n = t.n, np = a.np, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int gtid = b.x * blockDim.x + tid, gsz = b.nx * blockDim.x;
  for (int k = gtid; k < n; k += gsz) a.count_other[k] = 0;
  if (!sort_on(a)) return 0;  // uniform over the grid
  for (int i = gtid; i < np; i += gsz) a.slot[i] = atomicAdd(&a.count[tile_of(g, t, a.cell[i])], 1);
  if (!grid_barrier(a.bar, b.nx)) {
    if (tid == 0) atomicOr(a.err, kErrBarrier);
    return -1;
  }
  __shared__ int pre[kScanTilesMax + 1];
  __shared__ int wsum[4];
  const int* __restrict__ count = a.count;
  const int ch = (n + 255) / 256;
  const int lo = min(n, tid * ch), hi = min(n, lo + ch);
  int sum = 0;
  for (int k = lo; k < hi; ++k) sum += count[k];
  const int incl = wave_incl_scan(sum);
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  int run = incl - sum;
  for (int w = 0; w < wv; ++w) run += wsum[w];
  for (int k = lo; k < hi; ++k) {
    pre[k] = run;
    run += count[k];
  }
  if (tid == 255) pre[n] = run;
  __syncthreads();
  if (b.x == 0)
    for (int k = tid; k <= n; k += 256) a.start[k] = pre[k];
  for (int i = gtid; i < np; i += gsz) {
    const int tile = tile_of(g, t, a.cell[i]);
    place(a, i, pre[tile] + a.slot[i], tile);
  }
  return 1;
}
FHD_BODY void body_sort_one(const SortArgs& a, const Blk& b) { (void)sort_one(a, b); }
FHD_KERNEL_PAIR((), 256, k_sort_one, SortArgs, body_sort_one(a, b))

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

// Inclusive prefix sum over the 64 lanes of a wave.
FHD_DEV int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// Largest value over the wave, in every lane (butterfly over lane-index bits).  An LDS
// atomicMax with per-lane values would do the same, but the compiler expands it into a
// 64-iteration scalar read-lane loop per wave, which saturates the CU's one scalar unit.
FHD_DEV unsigned wave_max_u32(unsigned v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned t = (unsigned)__shfl_xor((int)v, o);
    v = v > t ? v : t;
  }
  return v;
}

// A pair term scaled to |x| < 2^40 as a 64-bit integer; non-finite values saturate (NaN to the
// negative bound), so the conversion is always defined.
FHD_DEV long long fixed_term(float x) {
  const float c = fminf(fmaxf(x, -2199023255552.0f), 2199023255552.0f);  // +-2^41
  return (long long)c;
}

constexpr int kWcaThreads = 256;
constexpr int kWcaSlices = 16;  // most lanes sharing one particle's candidates
constexpr int kWcaCap = 512;    // culled candidates staged in LDS at once
constexpr int kWcaBuildGroups = 1024;  // most workgroups of a list build

// WCA pair term on particle i from particle j: F_ij = 48 eps [(s/r)^12 - (s/r)^6 / 2] / r^2 r_ij
// for 0 < r < rc (fluam's pair loop, src/stokesLimitFunctions.cu:160-653, with the analytic
// potential instead of the float texture table of src/initForcesNonBonded.cu:21-66).  Cells
// wrapped, offsets from the cell centre.  Returns false when the pair does not interact.
FHD_DEV bool wca_term(const WcaArgs& a, int4 ci, float4 di, int4 cj, float4 dj, float& tx, float& ty, float& tz) {
  const GridDev& g = a.g;
  const float rx = (float)min_image(ci.x - cj.x, g.mx) * g.dx + (di.x - dj.x);
  const float ry = (float)min_image(ci.y - cj.y, g.my) * g.dy + (di.y - dj.y);
  const float rz = (float)min_image(ci.z - cj.z, g.mz) * g.dz + (di.z - dj.z);
  float r2 = rx * rx + ry * ry + rz * rz;
  if (!(r2 < a.rc2 && r2 > 0.0f)) return false;
  // separations below sigma/100 are evaluated at sigma/100, so the force stays finite
  // (fluam's tabulated force is bounded near r = 0 too, src/initForcesNonBonded.cu:59)
  r2 = fmaxf(r2, 1.0e-4f * a.sig2);
  const float s2 = a.sig2 / r2;
  const float s6 = s2 * s2 * s2;
  const float f = a.eps48 * (s6 * s6 - 0.5f * s6) / r2;
  tx = f * rx;
  ty = f * ry;
  tz = f * rz;
  return true;
}

// Neighbour-list build, one workgroup per tile, run only when the lists are flagged stale.  The
// particles of the neighbouring tiles that lie within the list radius of the tile (grown by the
// sort's skin) are compacted into LDS (ballot compaction), once per tile when they fit kWcaCap
// entries, else in passes of kWcaCap; S lanes share each of the tile's particles (S = 16, or
// the largest power of two with S x particles <= 256 in crowded tiles) and append every
// candidate closer than the list radius to its list.  The order inside a list is that of the
// LDS atomics; the force pass sums in fixed point, where the order does not matter.
FHD_DEV void wca_build_tiles(const WcaArgs& a, const Blk& b) {
  __shared__ int4 ccell[kWcaCap];   // w: sorted slot
  __shared__ float4 coff[kWcaCap];
  __shared__ int nb_start[32], nb_pref[33];
  __shared__ int wsum[kWcaThreads / 64];
  __shared__ int ncnt[kWcaThreads];
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // a grid of at most kWcaBuildGroups workgroups (a stride loop over the tiles), so the launch
  // that finds the lists valid is short
  for (int tb = b.x; tb < t.n; tb += b.nx) {
    __syncthreads();  // the previous tile's LDS consumed
    const int tile = xcd_item(tb, t.n);
    const int s0 = a.set.start[tile], nown = a.set.start[tile + 1] - s0;
    if (nown == 0) continue;  // uniform: no barrier is skipped by part of the group
    const int tx = tile % t.ntx, ty = (tile / t.ntx) % t.nty, tz = tile / (t.ntx * t.nty);
    const int ox = tx * t.tx, oy = ty * t.ty, oz = tz * t.tz;
    int lox, hix, loy, hiy, loz, hiz;
    nb_range(t.ntx, lox, hix);
    nb_range(t.nty, loy, hiy);
    nb_range(t.ntz, loz, hiz);
    const int nx = hix - lox + 1, ny = hiy - loy + 1, nz = hiz - loz + 1;
    const int nnb = nx * ny * nz;  // <= 27
    if (tid < 64) {
      int cnt = 0;
      if (tid < nnb) {
        const int dx = lox + tid % nx, dy = loy + (tid / nx) % ny, dz = loz + tid / (nx * ny);
        const int nt = wrapi(tx + dx, t.ntx) + t.ntx * (wrapi(ty + dy, t.nty) + t.nty * wrapi(tz + dz, t.ntz));
        const int j0 = a.set.start[nt];
        nb_start[tid] = j0;
        cnt = a.set.start[nt + 1] - j0;
      }
      const int incl = wave_incl_scan(cnt);
      if (tid < 32) nb_pref[tid + 1] = incl;
      if (tid == 0) nb_pref[0] = 0;
    }
    __syncthreads();
    const int nall = nb_pref[nnb];
    // conservative cull radius (the pair test below is exact), measured from the tile grown by
    // the skin: a listed particle may lie up to skin_len outside its tile along each axis
    const float rcc = sqrtf(a.rl2) * 1.001f + 1e-6f;
    const float skin = a.skin_len;
    const float Tx = t.tx * g.dx, Ty = t.ty * g.dy, Tz = t.tz * g.dz;
    // The distance to the tile is measured from the candidate's image nearest the tile centre,
    // folded on whole cells, so it can be off by up to one cell near the half-box point.  Along
    // an axis where that point lies farther than the radius + one cell outside the tile, the
    // cull is exact; along an axis with one tile every candidate is inside.  Otherwise keep
    // everything.
    auto axis_ok = [&](int nt_a, int m, int T, float d) { return nt_a == 1 || 0.5f * (m - T) * d > rcc + skin + d; };
    const bool cull = axis_ok(t.ntx, g.mx, t.tx, g.dx) && axis_ok(t.nty, g.my, t.ty, g.dy) &&
                      axis_ok(t.ntz, g.mz, t.tz, g.dz);
    // Compacts the near candidates with rank in [p0, p0 + kWcaCap) into LDS; returns how many
    // near candidates there are in all.
    auto stage = [&](int p0) {
      int seen = 0;
      for (int c0 = 0; c0 < nall; c0 += kWcaThreads) {
        const int e = c0 + tid;
        bool near = false;
        int4 cj = make_int4(0, 0, 0, 0);
        float4 dj = make_float4(0.f, 0.f, 0.f, 0.f);
        if (e < nall) {
          // largest lo with nb_pref[lo] <= e (entries past nnb hold nall)
          int lo = 0;
  #pragma unroll
          for (int step = 16; step >= 1; step >>= 1)
            if (nb_pref[lo + step] <= e) lo += step;
          const int j = nb_start[lo] + (e - nb_pref[lo]);
          cj = a.set.cell[j];
          dj = a.set.off[j];
          cj.w = j;
          near = true;
          if (cull) {
            // position relative to the tile centre (nearest image), then the distance outside
            // the tile along each axis (zero along an axis with a single tile)
            const float px = ((float)min_image(cj.x - (ox + t.tx / 2), g.mx) + 0.5f) * g.dx + dj.x;
            const float py = ((float)min_image(cj.y - (oy + t.ty / 2), g.my) + 0.5f) * g.dy + dj.y;
            const float pz = ((float)min_image(cj.z - (oz + t.tz / 2), g.mz) + 0.5f) * g.dz + dj.z;
            const float ex = t.ntx == 1 ? 0.0f : fmaxf(0.0f, fabsf(px) - 0.5f * Tx - skin);
            const float ey = t.nty == 1 ? 0.0f : fmaxf(0.0f, fabsf(py) - 0.5f * Ty - skin);
            const float ez = t.ntz == 1 ? 0.0f : fmaxf(0.0f, fabsf(pz) - 0.5f * Tz - skin);
            near = ex * ex + ey * ey + ez * ez <= rcc * rcc;
          }
        }
        const unsigned long long bal = __ballot(near);
        const int before_w = __popcll(bal & ((1ull << lane) - 1ull));
        __syncthreads();  // wsum of the previous chunk consumed
        if (lane == 0) wsum[wv] = __popcll(bal);
        __syncthreads();
        int base = seen, nk = 0;
  #pragma unroll
        for (int w = 0; w < kWcaThreads / 64; ++w) {
          base += w < wv ? wsum[w] : 0;
          nk += wsum[w];
        }
        const int rank = base + before_w - p0;
        if (near && rank >= 0 && rank < kWcaCap) {
          ccell[rank] = cj;
          coff[rank] = dj;
        }
        seen += nk;
      }
      __syncthreads();
      return seen;
    };
    const int nnear = stage(0);
    const int npass = (nnear + kWcaCap - 1) / kWcaCap;
    // lanes per own particle: up to kWcaSlices, fewer in crowded tiles so that one group of
    // kWcaThreads / S particles covers the tile and the candidates are staged once
    int S = kWcaSlices;
    while (S > 1 && S * nown > kWcaThreads) S >>= 1;
    for (int i0 = 0; i0 < nown; i0 += kWcaThreads / S) {
      const int il = tid / S, sl = tid % S;
      const int si = s0 + i0 + il;
      const bool own = i0 + il < nown;
      int4 ci = make_int4(0, 0, 0, 0);
      float4 di = make_float4(0.f, 0.f, 0.f, 0.f);
      int id = 0;
      if (own) {
        ci = a.set.cell[si];
        di = a.set.off[si];
        id = a.set.id[si];
      }
      ncnt[tid] = 0;  // slot il of the group: entry il * S (lane sl == 0 of each particle)
      __syncthreads();
      for (int p = 0; p < npass; ++p) {
        if (npass > 1 && p > 0) {
          __syncthreads();  // the previous pass's candidates consumed
          stage(p * kWcaCap);
        }
        const int nk = min(kWcaCap, nnear - p * kWcaCap);
        if (own) {
          for (int k = sl; k < nk; k += S) {
            const int4 cc = ccell[k];
            if (cc.w == si) continue;
            const float4 dd = coff[k];
            const float rx = (float)min_image(ci.x - cc.x, g.mx) * g.dx + (di.x - dd.x);
            const float ry = (float)min_image(ci.y - cc.y, g.my) * g.dy + (di.y - dd.y);
            const float rz = (float)min_image(ci.z - cc.z, g.mz) * g.dz + (di.z - dd.z);
            if (rx * rx + ry * ry + rz * rz < a.rl2) {
              const int pos = atomicAdd(&ncnt[il * S], 1);
              if (pos < kNbCap) a.nb.idx[(size_t)pos * a.np + id] = a.set.id[cc.w];
            }
          }
        }
      }
      __syncthreads();
      if (own && sl == 0) {
        a.nb.cnt[id] = ncnt[il * S];
        a.nb.ref_cell[id] = ci;
        a.nb.ref_off[id] = di;
      }
      if (npass > 1 && i0 + kWcaThreads / S < nown) {
        __syncthreads();
        stage(0);  // the next group of own particles starts from the first pass again
      }
    }
  }
}
FHD_BODY void body_wca(const WcaArgs& a, const Blk& b) {
  if (*a.nb.flag == 0) return;  // lists still valid (uniform over the grid)
  wca_build_tiles(a, b);
}
FHD_KERNEL_PAIR((), kWcaThreads, k_wca, WcaArgs, body_wca(a, b))

// The WCA force from the neighbour lists, one lane per particle (by id).  The force is summed
// as 64-bit integers at a scale set by the largest force component of any of the particle's
// pair terms: each term keeps its FP32 value and integer sums are exact, so the force does not
// depend on the order of the list.  A particle whose list overflowed takes every particle of
// the 27 tiles around its tile (the sorted set) instead.  Also clears the stale-list flag that
// the build before it consumed.
FHD_DEV void wca_list_at(const WcaArgs& a, int i) {
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int4 c0 = a.cell[i];
  const int4 ci = make_int4(wrapi(c0.x, g.mx), wrapi(c0.y, g.my), wrapi(c0.z, g.mz), 0);
  const float4 di = a.off[i];
  const int n = a.nb.cnt[i];
  // the first kNbHead entries are loaded with the count and their positions gathered together
  // (entries past the count hold valid ids from earlier builds, or 0, and are not used)
  constexpr int kNbHead = 4;
  int jd[kNbHead];
#pragma unroll
  for (int j = 0; j < kNbHead; ++j) jd[j] = a.nb.idx[(size_t)j * a.np + i];
  int4 cr[kNbHead];
  float4 dr[kNbHead];
#pragma unroll
  for (int j = 0; j < kNbHead; ++j) {
    cr[j] = a.cell[jd[j]];
    dr[j] = a.off[jd[j]];
  }
  float m = 0.0f, scale = 0.0f;
  long long sx = 0, sy = 0, sz = 0;
  if (n <= kNbCap) {
    float hx[kNbHead], hy[kNbHead], hz[kNbHead];
    bool hon[kNbHead];
#pragma unroll
    for (int j = 0; j < kNbHead; ++j) {
      const int4 cj = make_int4(wrapi(cr[j].x, g.mx), wrapi(cr[j].y, g.my), wrapi(cr[j].z, g.mz), 0);
      hon[j] = j < n && wca_term(a, ci, di, cj, dr[j], hx[j], hy[j], hz[j]);
      if (hon[j]) m = fmaxf(m, fmaxf(fabsf(hx[j]), fmaxf(fabsf(hy[j]), fabsf(hz[j]))));
    }
    for (int sweep = 0; sweep < 2; ++sweep) {
      for (int j = kNbHead; j < n; ++j) {
        const int jj = a.nb.idx[(size_t)j * a.np + i];
        const int4 cq = a.cell[jj];
        const int4 cj = make_int4(wrapi(cq.x, g.mx), wrapi(cq.y, g.my), wrapi(cq.z, g.mz), 0);
        float tx, ty, tz;
        if (!wca_term(a, ci, di, cj, a.off[jj], tx, ty, tz)) continue;
        if (sweep == 0) {
          m = fmaxf(m, fmaxf(fabsf(tx), fmaxf(fabsf(ty), fabsf(tz))));
        } else {
          sx += fixed_term(tx * scale);
          sy += fixed_term(ty * scale);
          sz += fixed_term(tz * scale);
        }
      }
      if (sweep == 0) {
        if (!(m <= 3.0e38f)) m = 3.0e38f;  // an overflowed term saturates
        int e = 0;
        if (m > 0.0f) frexpf(m, &e);  // m < 2^e
        scale = ldexpf(1.0f, 40 - e);
        m = ldexpf(1.0f, e - 40);  // from here on: the inverse scale
#pragma unroll
        for (int j = 0; j < kNbHead; ++j)
          if (hon[j]) {
            sx += fixed_term(hx[j] * scale);
            sy += fixed_term(hy[j] * scale);
            sz += fixed_term(hz[j] * scale);
          }
        if (n <= kNbHead) break;
      }
    }
  } else {
    // every particle listed in the tiles around this particle's tile (a particle within rc
    // lies at most one tile away from it: tiles are 4 or more cells and the skin half a cell)
    const int tile = tile_of(g, t, c0);
    const int tx = tile % t.ntx, ty = (tile / t.ntx) % t.nty, tz = tile / (t.ntx * t.nty);
    int lox, hix, loy, hiy, loz, hiz;
    nb_range(t.ntx, lox, hix);
    nb_range(t.nty, loy, hiy);
    nb_range(t.ntz, loz, hiz);
    for (int sweep = 0; sweep < 2; ++sweep) {
      for (int dz = loz; dz <= hiz; ++dz)
        for (int dy = loy; dy <= hiy; ++dy)
          for (int dx = lox; dx <= hix; ++dx) {
            const int nt = wrapi(tx + dx, t.ntx) + t.ntx * (wrapi(ty + dy, t.nty) + t.nty * wrapi(tz + dz, t.ntz));
            for (int j = a.set.start[nt]; j < a.set.start[nt + 1]; ++j) {
              if (a.set.id[j] == i) continue;
              float fx, fy, fz;
              if (!wca_term(a, ci, di, a.set.cell[j], a.set.off[j], fx, fy, fz)) continue;
              if (sweep == 0) {
                m = fmaxf(m, fmaxf(fabsf(fx), fmaxf(fabsf(fy), fabsf(fz))));
              } else {
                sx += fixed_term(fx * scale);
                sy += fixed_term(fy * scale);
                sz += fixed_term(fz * scale);
              }
            }
          }
      if (sweep == 0) {
        if (!(m <= 3.0e38f)) m = 3.0e38f;
        int e = 0;
        if (m > 0.0f) frexpf(m, &e);
        scale = ldexpf(1.0f, 40 - e);
        m = ldexpf(1.0f, e - 40);
      }
    }
  }
  a.F[i] = make_float4((float)sx * m, (float)sy * m, (float)sz * m, 0.0f);
}
FHD_BODY void body_wca_list(const WcaArgs& a, const Blk& b) {
  const int i = b.x * blockDim.x + threadIdx.x;
  if (i == 0) *a.nb.flag = 0;
  if (i >= a.np) return;
  wca_list_at(a, i);
}
FHD_KERNEL_PAIR((), 256, k_wca_list, WcaArgs, body_wca_list(a, b))

// The list build and the list force in one launch, for a context launched on its own (merged
// batch dispatches keep the two launches).  At most kWcaOneGroups workgroups, all resident: when
// the lists are flagged stale they rebuild them (a stride loop over the tiles), meet at a grid
// barrier, and one of them clears the flag, which every workgroup read before the barrier; then
// each thread takes its particles' forces from the lists.
constexpr int kWcaOneGroups = 256;
FHD_BODY void body_wca_one(const WcaArgs& a, const Blk& b) {
  if (*a.nb.flag != 0) {  // uniform over the grid
    wca_build_tiles(a, b);
    if (!grid_barrier(a.bar, b.nx)) {
      if (threadIdx.x == 0) atomicOr(a.err, kErrBarrier);
      return;
    }
    if (b.x == 0 && threadIdx.x == 0) *a.nb.flag = 0;
  }
  for (int i = b.x * blockDim.x + threadIdx.x; i < a.np; i += b.nx * blockDim.x) wca_list_at(a, i);
}
FHD_KERNEL_PAIR((), 256, k_wca_one, WcaArgs, body_wca_one(a, b))

// The sort and the WCA pass of one force evaluation in one launch (same grid: at most 256
// workgroups of 256 lanes, all resident).  When the sort ran, the workgroups meet once more
// before the WCA pass reads the sorted set; the WCA pass's barriers use the sort's counter.
FHD_BODY void body_sort_wca(const SortWcaArgs& x, const Blk& b) {
  const int ran = sort_one(x.s, b);
  if (ran < 0) return;
  if (ran > 0 && !grid_barrier(x.s.bar, b.nx)) {
    if (threadIdx.x == 0) atomicOr(x.s.err, kErrBarrier);
    return;
  }
  WcaArgs a = x.w;
  a.bar = x.s.bar;
  body_wca_one(a, b);
}
FHD_KERNEL_PAIR((), 256, k_sort_wca, SortWcaArgs, body_sort_wca(a, b))

// Harmonic bonds, one lane per particle (src/forceBondedGPU.cu:21-121): for each bond partner
// (another particle, then fixed points, each in file order) F_i += -k (1 - r0/r) (r_i - r_j),
// with -k (r_i - r_j) when r0 = 0 and nothing when r = 0 and r0 != 0.  As in fluam the
// separation is the plain difference of the unwrapped positions, with no periodic image.
constexpr int kBondThreads = 256;
FHD_DEV void bond_term(float rx, float ry, float rz, float k, float r0, float& fx, float& fy, float& fz) {
  float s = k;
  if (r0 != 0.0f) {
    const float r = sqrtf(rx * rx + ry * ry + rz * rz);
    s = r > 0.0f ? k * (1.0f - r0 / r) : 0.0f;
  }
  fx -= s * rx;
  fy -= s * ry;
  fz -= s * rz;
}
FHD_BODY void body_bonded(const BondArgs& a, const Blk& b) {
  const int i = b.x * kBondThreads + (int)threadIdx.x;
  if (i >= a.np) return;
  const int4 ci = a.cell[i];
  const float4 oi = a.off[i];
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  const int pe = a.pp_start[i + 1];
  for (int e = a.pp_start[i]; e < pe; ++e) {
    const int4 bd = a.pp[e];
    const int4 cj = a.cell[bd.x];
    const float4 oj = a.off[bd.x];
    bond_term((float)(ci.x - cj.x) * a.g.dx + (oi.x - oj.x), (float)(ci.y - cj.y) * a.g.dy + (oi.y - oj.y),
              (float)(ci.z - cj.z) * a.g.dz + (oi.z - oj.z), __int_as_float(bd.y), __int_as_float(bd.z), fx, fy, fz);
  }
  const int fe = a.fp_start[i + 1];
  for (int e = a.fp_start[i]; e < fe; ++e) {
    const int4 cx = a.fpc[e];
    const float4 ox = a.fpo[e];
    bond_term((float)(ci.x - cx.x) * a.g.dx + (oi.x - ox.x), (float)(ci.y - cx.y) * a.g.dy + (oi.y - ox.y),
              (float)(ci.z - cx.z) * a.g.dz + (oi.z - ox.z), __int_as_float(cx.w), ox.w, fx, fy, fz);
  }
  float4 f = a.F[i];
  f.x += fx;
  f.y += fy;
  f.z += fz;
  a.F[i] = f;
}
FHD_KERNEL_PAIR((), kBondThreads, k_bonded, BondArgs, body_bonded(a, b))

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

constexpr int kTileEdge = 8;         // largest spreading tile edge (cells)
constexpr int kSpreadThreads = 256;

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

FHD_DEV float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }

// Position (cell, offset) of source `sub` of a particle, plus its value.
FHD_DEV void rfd_normal(const Key& key, int id, float n[4]) {
  draw_normals4(key, (uint32_t)id, FHD_STREAM_RFD, 0, n);
}

// Does a stencil whose nodes span [lo, lo + NPTS] (both staggerings) reach the tile [o, o+T)?
template <int NPTS>
FHD_DEV bool reaches(int lo, int o, int T, int m, int ntiles) {
  if (ntiles == 1) return true;
  const int l0 = fold(lo - o, m);
  return l0 < T || l0 + NPTS >= m;
}

// One record per source (sorted slot x sub-source): weights, bases, value and the cull data.
// One launch covers every group: threads [0, n0) build group 0's records, the rest group 1's.
template <int NPTS>
FHD_BODY void body_src_prep(const SpreadArgs& a, const Blk& b) {
  const int n0 = a.nrec0, ntot = a.nrec_tot;
  int r = b.x * blockDim.x + threadIdx.x;
  if (r >= ntot) return;
  const int gi = r < n0 ? 0 : 1;
  if (gi) r -= n0;
  const SpreadGroup& G = a.grp[gi];
  const GridDev& g = a.g;
  const int nsub = G.mode == SRC_RFD ? 2 : 1;
  const int sidx = nsub == 2 ? r >> 1 : r, sub = r - sidx * nsub;
  const int id = G.set.id[sidx];
  const int4 c = G.set.cell[sidx];
  const float4 d = G.set.off[sidx];
  float3 val;
  int4 cw = c;
  float3 dw = make_float3(d.x, d.y, d.z);
  if (G.mode == SRC_VALUE) {
    const float4 v = G.val[id];
    val = make_float3(v.x * G.coef, v.y * G.coef, v.z * G.coef);
    if (val.x == 0.0f && val.y == 0.0f && val.z == 0.0f) {
      // a zero source (e.g. a particle with no WCA neighbour) adds exactly zero to every node:
      // an empty reach mask makes every tile cull it, and its record is never read
      G.cull[r] = 0;
      return;
    }
  } else if (G.mode == SRC_DRIFT) {
    val = make_float3(G.coef * g.idx, G.coef * g.idy, G.coef * g.idz);
  } else {
    float nrm[4];
    rfd_normal(resolve_key(a.key), id, nrm);
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
  const float vv[3] = {val.x, val.y, val.z};
  SrcRec rec;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    rec.base[0][ax] = w.cen[ax].base;
    rec.base[1][ax] = w.stg[ax].base;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      rec.w[0][ax][q] = w.cen[ax].w[q];
      // an overflowed source (non-finite or beyond 3e38) is stored as +-3e38, NaN as -3e38:
      // its tile's bound then reaches 2^127, and the spreading pass writes that tile's field as
      // non-finite, which the interpolation pass flags
      const float fw = w.stg[ax].w[q] * vv[ax];
      rec.w[1][ax][q] = fabsf(fw) <= 3.0e38f ? fw : (fw > 0.0f ? 3.0e38f : -3.0e38f);
    }
  }
  rec.pad[0] = rec.pad[1] = 0;
  const float4* src4 = reinterpret_cast<const float4*>(&rec);
  float4* dst4 = reinterpret_cast<float4*>(G.rec + r);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(SrcRec) / 16); ++k) dst4[k] = src4[k];
  // bound on one contribution: centred weights are at most 1, so the largest face weight x value
  float m = 0.0f;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
#pragma unroll
    for (int q = 0; q < 4; ++q) m = fmaxf(m, fabsf(rec.w[1][ax][q]));  // <= 3e38
  // reach mask: per axis, bit e + 1 set when the stencils (nodes lo .. lo + NPTS, both
  // staggerings) reach the tile at offset e in {-1, 0, 1} from the source's sorting tile
  const int lo[3] = {min(w.cen[0].base, w.stg[0].base), min(w.cen[1].base, w.stg[1].base),
                     min(w.cen[2].base, w.stg[2].base)};
  const Tiles& t = a.t;
  const int T[3] = {t.tx, t.ty, t.tz}, NT[3] = {t.ntx, t.nty, t.ntz}, MM[3] = {g.mx, g.my, g.mz};
  // the tile the source is listed in (the sorted copy's w): the spreading pass looks for it in
  // the lists of that tile's neighbours, which its stencil may reach while it stays within the
  // skin of that tile
  const int lt[3] = {c.w % t.ntx, (c.w / t.ntx) % t.nty, c.w / (t.ntx * t.nty)};
  int mask = 0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int own = lt[ax];
#pragma unroll
    for (int e = -1; e <= 1; ++e) {
      int tt = own + e;  // in [-1, NT]
      tt += tt < 0 ? NT[ax] : (tt >= NT[ax] ? -NT[ax] : 0);
      if (reaches<NPTS>(lo[ax], tt * T[ax], T[ax], MM[ax], NT[ax])) mask |= 1 << (3 * ax + e + 1);
    }
  }
  // the bound enters the spreading pass only through its binary exponent: m < 2^e (e = -191
  // for m = 0, below any other record's), stored biased by 192 in bits 16-24
  int e = -191;
  if (m > 0.0f) frexpf(m, &e);
  G.cull[r] = mask | ((e + 192) << 16);
}
FHD_KERNEL_PAIR((template <int NPTS>), 256, k_src_prep, SpreadArgs, body_src_prep<NPTS>(a, b))

// Fixed-point image of one contribution, as a 64-bit two's-complement addend.  Integer sums are
// exact, so the accumulated field does not depend on the order in which the contributions
// arrive.  The records are finite with face weights of at most 3e38 < 2^128, and the tile's
// scale 2^(30 - e) puts its largest face weight below 2^30; centred weights are at most 1, so
// |x| <= 2^30 and the conversion is always defined without a clamp.
FHD_DEV unsigned long long fixed64(float x) {
  // truncation toward zero: at the 2^30 scale its bias (< 1 unit) is far below FP32 rounding
  const int q = (int)x;
  return (unsigned long long)(long long)q;
}

// Candidate-record reference: group in bit 30, record index below.
constexpr int kItemCap = 1024;
// Accumulator z-row stride (64-bit words): the 8 x 8 plane plus 2, so the rows that the four
// lanes of one candidate update start in different LDS banks.
constexpr int kRowStride = kTileEdge * kTileEdge + 2;

// Contributions of stencil row kz (nodes base_z + kz) of one source record to the three
// components of the tile accumulator acc[component][z row][y][x] (x, y: the 8 x 8 tile columns).
// Only nodes inside the tile are added: a lane whose row lies outside skips it, and out-of-tile
// columns are masked off, so no LDS atomic is spent on a node the tile does not own.  FAST: the
// first node relative to the tile comes from one periodic fold; otherwise (small boxes) every
// node is folded periodically and tested against the tile.
template <int NPTS, bool FAST>
FHD_DEV void add_source_row(unsigned long long* acc, const SrcRec* __restrict__ R, int kz, int ox, int oy, int oz,
                            const GridDev& g, const Tiles& t, float scale) {
  constexpr int H = NPTS - 1;
  constexpr int ZS = kTileEdge;
  constexpr int PL = kRowStride;
  const float4 cx = *reinterpret_cast<const float4*>(R->w[0][0]);
  const float4 cy = *reinterpret_cast<const float4*>(R->w[0][1]);
  const float4 fx = *reinterpret_cast<const float4*>(R->w[1][0]);
  const float4 fy = *reinterpret_cast<const float4*>(R->w[1][1]);
  const float cz = R->w[0][2][kz], fz = R->w[1][2][kz];
  const int4 b0 = *reinterpret_cast<const int4*>(&R->base[0][0]);  // cen x, y, z, face x
  const int2 b1 = *reinterpret_cast<const int2*>(&R->base[1][1]);  // face y, z
#pragma unroll
  for (int C = 0; C < 3; ++C) {
    const int bx = C == 0 ? b0.w : b0.x, by = C == 1 ? b1.x : b0.y, bz = C == 2 ? b1.y : b0.z;
    const float4 w4x = C == 0 ? fx : cx, w4y = C == 1 ? fy : cy;
    const float wx[4] = {w4x.x, w4x.y, w4x.z, w4x.w}, wy[4] = {w4y.x, w4y.y, w4y.z, w4y.w};
    const float wzs = (C == 2 ? fz : cz) * scale;
    if (FAST) {
      // first stencil node relative to the tile, folded to [-H, m - H)
      int dx = bx - ox, dy = by - oy, dz = bz - oz;
      dx += dx < -H ? g.mx : (dx >= g.mx - H ? -g.mx : 0);
      dy += dy < -H ? g.my : (dy >= g.my - H ? -g.my : 0);
      dz += dz < -H ? g.mz : (dz >= g.mz - H ? -g.mz : 0);
      const int lz = dz + kz;
      if ((unsigned)lz >= (unsigned)kTileEdge) continue;  // this row misses the tile
      const int rb = (C * ZS + lz) * PL;
#pragma unroll
      for (int ky = 0; ky < NPTS; ++ky) {
        const float wyz = wy[ky] * wzs;
        const int ly = dy + ky;
#pragma unroll
        for (int kx = 0; kx < NPTS; ++kx) {
          const int lx = dx + kx;
          if ((unsigned)ly < (unsigned)kTileEdge && (unsigned)lx < (unsigned)kTileEdge)
            atomicAdd(&acc[rb + ly * kTileEdge + lx], fixed64(wx[kx] * wyz));
        }
      }
    } else {
      const int lz = fold(bz + kz - oz, g.mz);
      if (lz >= t.tz) continue;
      unsigned long long* row = acc + (C * ZS + lz) * PL;
#pragma unroll
      for (int ky = 0; ky < NPTS; ++ky) {
        const int ly = fold(by + ky - oy, g.my);
        if (ly >= t.ty) continue;
        const float wyz = wy[ky] * wzs;
#pragma unroll
        for (int kx = 0; kx < NPTS; ++kx) {
          const int lx = fold(bx + kx - ox, g.mx);
          if (lx < t.tx) atomicAdd(&row[ly * kTileEdge + lx], fixed64(wx[kx] * wyz));
        }
      }
    }
  }
}

// The narrow accumulator of 8 x 8 x 8 tiles with at most kNarrowCand candidates: 32-bit
// fixed point, rows of kRow32 words covering x = -3 .. 10 (a halo of three columns each side,
// so every node of a stencil row that starts in [-3, 7] has a slot and the atomics need no
// column test; the halo is never read), planes of kPlane32 words (the 8 rows plus 1, so the
// rows of the four lanes of one candidate start in different LDS banks).  A node's sum has at
// most one contribution per candidate, so with contributions below 2^(30 - floor(log2 n)) for n
// candidates it stays below 2^31.
constexpr int kNarrowCand = 255;
constexpr int kRow32 = kTileEdge + 6;
constexpr int kPlane32 = kTileEdge * kRow32 + 1;
constexpr int kComp32 = kTileEdge * kPlane32;

template <int NPTS>
FHD_DEV void add_source_row32(unsigned* acc, const SrcRec* __restrict__ R, int kz, int ox, int oy, int oz,
                              const GridDev& g, float scale) {
  constexpr int H = NPTS - 1;
  const float4 cx = *reinterpret_cast<const float4*>(R->w[0][0]);
  const float4 cy = *reinterpret_cast<const float4*>(R->w[0][1]);
  const float4 fx = *reinterpret_cast<const float4*>(R->w[1][0]);
  const float4 fy = *reinterpret_cast<const float4*>(R->w[1][1]);
  const float cz = R->w[0][2][kz], fz = R->w[1][2][kz];
  const int4 b0 = *reinterpret_cast<const int4*>(&R->base[0][0]);  // cen x, y, z, face x
  const int2 b1 = *reinterpret_cast<const int2*>(&R->base[1][1]);  // face y, z
#pragma unroll
  for (int C = 0; C < 3; ++C) {
    const int bx = C == 0 ? b0.w : b0.x, by = C == 1 ? b1.x : b0.y, bz = C == 2 ? b1.y : b0.z;
    const float4 w4x = C == 0 ? fx : cx, w4y = C == 1 ? fy : cy;
    const float wx[4] = {w4x.x, w4x.y, w4x.z, w4x.w}, wy[4] = {w4y.x, w4y.y, w4y.z, w4y.w};
    const float wzs = (C == 2 ? fz : cz) * scale;
    // first stencil node relative to the tile, folded to [-H, m - H)
    int dx = bx - ox, dy = by - oy, dz = bz - oz;
    dx += dx < -H ? g.mx : (dx >= g.mx - H ? -g.mx : 0);
    dy += dy < -H ? g.my : (dy >= g.my - H ? -g.my : 0);
    dz += dz < -H ? g.mz : (dz >= g.mz - H ? -g.mz : 0);
    const int lz = dz + kz;
    // this row misses the tile, or its columns do (then dx lies outside [-H, 7])
    if ((unsigned)lz >= (unsigned)kTileEdge || dx > kTileEdge - 1) continue;
    // word index of the stencil's first node; the stencil's rows and columns are constant
    // offsets from it
    const int first = C * kComp32 + lz * kPlane32 + dy * kRow32 + dx + 3;
#pragma unroll
    for (int ky = 0; ky < NPTS; ++ky) {
      if ((unsigned)(dy + ky) >= (unsigned)kTileEdge) continue;
      const float wyz = wy[ky] * wzs;
#pragma unroll
      for (int kx = 0; kx < NPTS; ++kx)
        atomicAdd(&acc[first + ky * kRow32 + kx], (unsigned)(int)(wx[kx] * wyz));
    }
  }
}

FHD_DEV const SrcRec* rec_of(const SpreadArgs& a, int ref) {
  return a.grp[ref >> 30].rec + (ref & ((1 << 30) - 1));
}

// Spreading, one workgroup per tile.  The records of the sources in the neighbouring tiles are
// culled against the tile into a candidate list; four lanes take each candidate, one stencil
// z-row each, and add its contributions to the three components into a 64-bit fixed-point LDS
// accumulator (scale 2^(30 - e) per tile, from the largest face weight x value of its
// candidates).  Integer addition is associative, so the result is bit-identical whatever order
// the lanes' atomics land in.  FAST: tiles of 8 x 8 x 8 cells and at least two tiles per axis.
template <int NPTS, bool FAST>
FHD_BODY void body_spread(const SpreadArgs& a, const Blk& b) {
  constexpr int ZS = kTileEdge;
  constexpr int NACC = 3 * ZS * kRowStride;
  static_assert(3 * kComp32 <= 2 * NACC, "the narrow accumulator shares the 64-bit one's LDS");
  constexpr int LPC = 4;  // lanes per candidate (one stencil z-row each)
  __shared__ unsigned long long acc[NACC];
  __shared__ int nb_start[64], nb_pref[65], nb_sel[64];
  __shared__ int items[kItemCap];
  __shared__ int nitem;
  __shared__ unsigned int maxbits;
  const GridDev& g = a.g;
  const Tiles& t = a.t;
  const int tid = threadIdx.x;
  const int tile = xcd_item(b.x, t.n);
  const int tx = tile % t.ntx, ty = (tile / t.ntx) % t.nty, tz = tile / (t.ntx * t.nty);
  const int ox = tx * t.tx, oy = ty * t.ty, oz = tz * t.tz;
  // FAST tiles: each thread's output nodes are e = tid + k * kSpreadThreads, k < kOutPer; their
  // out_add values are loaded now, so the loads complete under the accumulation
  constexpr int kOutPer = 3 * kTileEdge * kTileEdge * kTileEdge / kSpreadThreads;
  float pre[kOutPer];
  if (FAST && a.out_add) {
#pragma unroll
    for (int k = 0; k < kOutPer; ++k) {
      const int e = tid + k * kSpreadThreads;
      const int C = e >> 9, lx = e & 7, ly = (e >> 3) & 7, lz = (e >> 6) & 7;
      pre[k] = a.out_add[(size_t)C * g.N + (ox + lx) + (size_t)g.mx * ((oy + ly) + (size_t)g.my * (oz + lz))];
    }
  }
  for (int e = tid; e < NACC; e += kSpreadThreads) acc[e] = 0ull;
  // neighbour-tile table: lane k of wave 0 loads the record range of (group, neighbour) k
  int lox, hix, loy, hiy, loz, hiz;
  nb_range(t.ntx, lox, hix);
  nb_range(t.nty, loy, hiy);
  nb_range(t.ntz, loz, hiz);
  const int nx = hix - lox + 1, ny = hiy - loy + 1, nz = hiz - loz + 1;
  const int nnb = nx * ny * nz;
  const int ntab = a.ngroups * nnb;  // <= 54
  if (tid < 64) {
    int cnt = 0;
    if (tid < ntab) {
      const int gi = tid / nnb, k = tid - gi * nnb;
      const int nsub = a.grp[gi].mode == SRC_RFD ? 2 : 1;
      const int dx = lox + k % nx, dy = loy + (k / nx) % ny, dz = loz + k / (nx * ny);
      const int nt = wrapi(tx + dx, t.ntx) + t.ntx * (wrapi(ty + dy, t.nty) + t.nty * wrapi(tz + dz, t.ntz));
      const int s0 = a.grp[gi].set.start[nt];
      nb_start[tid] = s0 * nsub;
      cnt = (a.grp[gi].set.start[nt + 1] - s0) * nsub;
      // this tile sits at offset -d from the source tile: the reach-mask bits to test
      nb_sel[tid] = (1 << (1 - dx)) | (1 << (4 - dy)) | (1 << (7 - dz));
    }
    const int incl = wave_incl_scan(cnt);
    nb_pref[tid + 1] = incl;
    if (tid == 0) {
      nb_pref[0] = 0;
      nitem = 0;
      maxbits = 0u;
    }
  }
  __syncthreads();
  const int nall = nb_pref[ntab];
  // record index e -> reference; -1 when the record's stencils miss the tile
  auto cull = [&](int e, int& mx) -> int {
    // largest lo with nb_pref[lo] <= e: fixed-step search (no divergent loop); entries past
    // ntab hold nall, which no record index reaches
    int lo = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
      if (nb_pref[lo + step] <= e) lo += step;
    const int gi = lo >= nnb;
    const int ri = nb_start[lo] + (e - nb_pref[lo]);
    const int cu = a.grp[gi].cull[ri];
    const int sel = nb_sel[lo];
    if ((cu & sel) != sel) return -1;
    mx = max(mx, cu >> 16);
    return (gi << 30) | ri;
  };
  // pass 1: cull every record, keep the first kItemCap candidates, largest magnitude; each wave
  // appends its survivors with one LDS atomic
  int lmax = 0;  // largest biased exponent of the candidates' bounds
  const int lane = tid & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int e0 = 0; e0 < nall; e0 += kSpreadThreads) {
    const int e = e0 + tid;
    const int ref = e < nall ? cull(e, lmax) : -1;
    const unsigned long long bal = __ballot(ref >= 0);
    int base = 0;
    if (lane == 0 && bal) base = atomicAdd(&nitem, __popcll(bal));
    base = __shfl(base, 0);
    if (ref >= 0) {
      const int k = base + __popcll(bal & below);
      if (k < kItemCap) items[k] = ref;
    }
  }
  const unsigned wmax = wave_max_u32((unsigned)lmax);
  if (lane == 0) atomicMax(&maxbits, wmax);
  __syncthreads();
  const int ntot = nitem;
  // every candidate's largest face weight x value is below 2^e2
  const int e2 = max((int)maxbits - 192, -90);
  // narrow tiles: contributions below 2^S with S = 30 - floor(log2 ntot); others 2^30 in 64 bits
  const bool narrow = FAST && ntot <= kNarrowCand;
  const int S = narrow ? 30 - (31 - __clz(ntot > 0 ? ntot : 1)) : 30;
  const float scale = ldexpf(1.0f, S - e2);
  // a bound of 2^127 or more means an overflowed source among the candidates (the record pass
  // stores those as +-3e38): the tile's field is written non-finite (every node +-inf or NaN),
  // so the step is flagged
  const float inv = e2 >= 128 ? __builtin_inff() : ldexpf(1.0f, e2 - S) * a.out_coef;
  const int kz = tid % LPC;
  unsigned* const acc32 = reinterpret_cast<unsigned*>(acc);
  if (narrow) {
    if (kz < NPTS)
      for (int c = tid / LPC; c < ntot; c += kSpreadThreads / LPC)
        add_source_row32<NPTS>(acc32, rec_of(a, items[c]), kz, ox, oy, oz, g, scale);
  } else if (ntot <= kItemCap) {
    if (kz < NPTS)
      for (int c = tid / LPC; c < ntot; c += kSpreadThreads / LPC)
        add_source_row<NPTS, FAST>(acc, rec_of(a, items[c]), kz, ox, oy, oz, g, t, scale);
  } else {
    // more candidates than the list holds: rounds over disjoint record ranges, re-culled
    for (int b0 = 0; b0 < nall; b0 += kItemCap) {
      __syncthreads();
      if (tid == 0) nitem = 0;
      __syncthreads();
      int dummy = 0;
      for (int e = b0 + tid; e < min(nall, b0 + kItemCap); e += kSpreadThreads) {
        const int ref = cull(e, dummy);
        if (ref >= 0) items[atomicAdd(&nitem, 1)] = ref;
      }
      __syncthreads();
      if (kz < NPTS)
        for (int c = tid / LPC; c < nitem; c += kSpreadThreads / LPC)
          add_source_row<NPTS, FAST>(acc, rec_of(a, items[c]), kz, ox, oy, oz, g, t, scale);
    }
  }
  __syncthreads();
  // interior nodes -> field (8 x 8 x 8 tiles: index arithmetic by shifts)
  const int ncell = t.tx * t.ty * t.tz;
  if (FAST) {
#pragma unroll
    for (int k = 0; k < kOutPer; ++k) {
      const int e = tid + k * kSpreadThreads;
      const int C = e >> 9, lx = e & 7, ly = (e >> 3) & 7, lz = (e >> 6) & 7;
      const size_t gi = (size_t)C * g.N + (ox + lx) + (size_t)g.mx * ((oy + ly) + (size_t)g.my * (oz + lz));
      const float q = narrow ? (float)(int)acc32[C * kComp32 + lz * kPlane32 + ly * kRow32 + lx + 3]
                             : (float)(long long)acc[(C * ZS + lz) * kRowStride + ly * kTileEdge + lx];
      const float val = q * inv;
      a.out[gi] = a.out_add ? pre[k] + val : val;
    }
    return;
  }
  for (int e = tid; e < 3 * ncell; e += kSpreadThreads) {
    int C, lx, ly, lz;
    if (FAST) {
      C = e >> 9;
      lx = e & 7;
      ly = (e >> 3) & 7;
      lz = (e >> 6) & 7;
    } else {
      C = e / ncell;
      const int l = e - C * ncell;
      lx = l % t.tx;
      ly = (l / t.tx) % t.ty;
      lz = l / (t.tx * t.ty);
    }
    const int ai = (C * ZS + lz) * kRowStride + ly * kTileEdge + lx;
    const size_t gi = (size_t)C * g.N + (ox + lx) + (size_t)g.mx * ((oy + ly) + (size_t)g.my * (oz + lz));
    const float val = (float)(long long)acc[ai] * inv;
    a.out[gi] = a.out_add ? a.out_add[gi] + val : val;
  }
}
FHD_KERNEL_PAIR((template <int NPTS, bool FAST>), kSpreadThreads, k_spread, SpreadArgs, body_spread<NPTS, FAST>(a, b))

// Sum over the NPTS x NPTS nodes of one z-row (kz) of a particle's stencil.
template <int NPTS>
FHD_DEV float interp_row(const float* __restrict__ f, const GridDev& g, const AxisWeights& wx,
                         const AxisWeights& wy, const AxisWeights& wz, int kz) {
  const int z = fold(wz.base + kz, g.mz);
  int xs[NPTS];
#pragma unroll
  for (int kx = 0; kx < NPTS; ++kx) xs[kx] = fold(wx.base + kx, g.mx);
  // a stencil row that does not wrap around the box and ends inside it is read as one 16-byte
  // load (global loads need only 4-byte alignment); otherwise node by node
  const bool vec = wx.base >= 0 && wx.base + 4 <= g.mx;
  float sy = 0.0f;
#pragma unroll
  for (int ky = 0; ky < NPTS; ++ky) {
    const int y = fold(wy.base + ky, g.my);
    const float* row = f + (size_t)g.mx * (y + (size_t)g.my * z);
    float r[4];
    if (vec) {
      typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
      const f4u v = *reinterpret_cast<const f4u*>(row + wx.base);
      r[0] = v.x;
      r[1] = v.y;
      r[2] = v.z;
      r[3] = v.w;
    } else {
#pragma unroll
      for (int kx = 0; kx < NPTS; ++kx) r[kx] = row[xs[kx]];
    }
    float sx = 0.0f;
#pragma unroll
    for (int kx = 0; kx < NPTS; ++kx) sx += wx.w[kx] * r[kx];
    sy += wy.w[ky] * sx;
  }
  return sy;  // weighted by wz.w[kz] when the rows are combined (static indices, no scratch)
}

// Combine the four z-rows held by lanes 4p .. 4p+3 (every lane gets the sum): sum_k wz[k] row_k
// in a fixed order, so the result does not depend on scheduling.  The four lanes of a particle
// are a DPP quad, so the rows are broadcast by quad_perm moves (plain VALU; no LDS round trip).
template <int K>
FHD_DEV float quad_lane(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), K * 0x55, 0xF, 0xF, false));
}
FHD_DEV float combine_rows(float row, const AxisWeights& wz) {
  const float r0 = quad_lane<0>(row), r1 = quad_lane<1>(row), r2 = quad_lane<2>(row), r3 = quad_lane<3>(row);
  return ((wz.w[0] * r0 + wz.w[1] * r1) + wz.w[2] * r2) + wz.w[3] * r3;
}

// Four lanes per particle: lane q sums z-row q of the stencil; the rows are combined by a fixed
// xor-shuffle tree, then lane 0 applies the position update.
// Position update of sorted particle s from its interpolated velocities u1 = J f1, u2 = J f2:
// out = base + a1 u1 + a2 u2 + extra mobility + particle noise (by particle index).
// Per-particle values the position update needs, loaded when the pass starts so that their
// latency hides under the gathers: the id, the base position and the neighbour-list reference.
struct InterpPre {
  int id;
  int4 bc, rc;
  float4 bo, ro;
};
FHD_DEV InterpPre interp_pre(const InterpArgs& a, int s) {
  InterpPre p;
  p.id = a.eval.id[s];
  p.bc = p.rc = make_int4(0, 0, 0, 0);
  p.bo = p.ro = make_float4(0.f, 0.f, 0.f, 0.f);
  if (a.out_cell) {
    p.bc = a.base_cell[p.id];
    p.bo = a.base_off[p.id];
    if (a.nb.flag) {
      p.rc = a.nb.ref_cell[p.id];
      p.ro = a.nb.ref_off[p.id];
    }
  }
  return p;
}

FHD_DEV void interp_finish(const InterpArgs& a, int s, const InterpPre& pre, const float u1[3], const float u2[3]) {
  const GridDev& g = a.g;
  const int id = pre.id;
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
    const Key key = resolve_key(a.key);
    float n[4];
    draw_normals4(key, (uint32_t)id, FHD_STREAM_PART1, 0, n);
    float z[3] = {n[0], n[1], n[2]};
    if (a.nnoise > 1) {
      draw_normals4(key, (uint32_t)id, FHD_STREAM_PART2, 0, n);
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
    int4 bc = pre.bc;
    const float4 bo = pre.bo;
    bool ok = isfinite(disp[0]) && isfinite(disp[1]) && isfinite(disp[2]);
    if (!ok) {
      atomicOr(a.err, kErrNonFinite);
      disp[0] = disp[1] = disp[2] = 0.0f;
    }
    // a displacement of 2^24 cells or more along an axis (a blown-up step, e.g. two particles
    // started on top of each other under the WCA core) leaves that coordinate where it was, and
    // the step is flagged
    const bool big = !(fabsf(disp[0]) < 16777216.0f * g.dx) || !(fabsf(disp[1]) < 16777216.0f * g.dy) ||
                     !(fabsf(disp[2]) < 16777216.0f * g.dz);
    if (big) {
      atomicOr(a.err, kErrClamped);
      if (!(fabsf(disp[0]) < 16777216.0f * g.dx)) disp[0] = 0.0f;
      if (!(fabsf(disp[1]) < 16777216.0f * g.dy)) disp[1] = 0.0f;
      if (!(fabsf(disp[2]) < 16777216.0f * g.dz)) disp[2] = 0.0f;
    }
    float ox = bo.x + disp[0], oy = bo.y + disp[1], oz = bo.z + disp[2];
    renorm_axis(bc.x, ox, g.dx, g.idx);
    renorm_axis(bc.y, oy, g.dy, g.idy);
    renorm_axis(bc.z, oz, g.dz, g.idz);
    a.out_cell[id] = bc;
    a.out_off[id] = make_float4(ox, oy, oz, 0.0f);
    if (a.nb.flag) {
      // the WCA neighbour lists hold every pair that was within rc + skin at the build; they
      // stay complete while no particle has moved more than half a skin from where it was then
      const int4 rc = pre.rc;
      const float4 ro = pre.ro;
      const float mx = (float)min_image(wrapi(bc.x, g.mx) - rc.x, g.mx) * g.dx + (ox - ro.x);
      const float my = (float)min_image(wrapi(bc.y, g.my) - rc.y, g.my) * g.dy + (oy - ro.y);
      const float mz = (float)min_image(wrapi(bc.z, g.mz) - rc.z, g.mz) * g.dz + (oz - ro.z);
      if (!(mx * mx + my * my + mz * mz <= a.nb_move2)) *a.nb.flag = 1;
    }
    if (a.sk.s_cell) {
      // the sorted copy of the new position, in the slot of the set it replaces; the particle
      // stays listed in that slot's tile unless it left the tile's skin
      const int slot = a.sk.s_slot ? a.sk.s_slot[id] : s;
      const int lt = (a.sk.s_slot ? a.sk.s_cell[slot] : a.eval.cell[s]).w;
      const int wc[3] = {wrapi(bc.x, g.mx), wrapi(bc.y, g.my), wrapi(bc.z, g.mz)};
      const int lto[3] = {(lt % a.t.ntx) * a.t.tx, ((lt / a.t.ntx) % a.t.nty) * a.t.ty,
                          (lt / (a.t.ntx * a.t.nty)) * a.t.tz};
      const int T[3] = {a.t.tx, a.t.ty, a.t.tz}, M[3] = {g.mx, g.my, g.mz};
      const float fo[3] = {ox * g.idx, oy * g.idy, oz * g.idz};
      bool out = false;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const float sk = a.sk.skin[ax];
        if (sk < 0.0f) continue;
        int d = wc[ax] - lto[ax];  // in (-m, m): fold to [-m/2, m/2)
        if (2 * d >= M[ax]) d -= M[ax];
        else if (2 * d < -M[ax]) d += M[ax];
        const float x = (float)d + 0.5f + fo[ax];  // cells from the tile's lower face
        out |= !(x >= -sk && x <= (float)T[ax] + sk);
      }
      if (out) *a.sk.skin_flag = 1;
      a.sk.s_cell[slot] = make_int4(wc[0], wc[1], wc[2], lt);
      a.sk.s_off[slot] = make_float4(ox, oy, oz, 0.0f);
      if (a.sk.sid_dst) a.sk.sid_dst[s] = id;
    }
  }
}

template <int NPTS>
FHD_BODY void body_interp(const InterpArgs& a, const Blk& b) {
  const int gid = xcd_item(b.x, b.nx) * blockDim.x + threadIdx.x;
  const int s = gid >> 2, q = gid & 3;
  const GridDev& g = a.g;
  float u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
  SrcWeights w;
  const bool act = s < a.np && q < NPTS;
  InterpPre pre;
  if (s < a.np) {
    const int4 c = a.eval.cell[s];
    const float4 d = a.eval.off[s];
    if (q == 0) pre = interp_pre(a, s);
    source_weights<NPTS>(c, make_float3(d.x * g.idx, d.y * g.idy, d.z * g.idz), w, false);
  }
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const AxisWeights& wx = comp == 0 ? w.stg[0] : w.cen[0];
    const AxisWeights& wy = comp == 1 ? w.stg[1] : w.cen[1];
    const AxisWeights& wz = comp == 2 ? w.stg[2] : w.cen[2];
    float r1 = 0.0f, r2 = 0.0f;
    if (act && a.f1) r1 = interp_row<NPTS>(a.f1 + (size_t)comp * g.N, g, wx, wy, wz, q);
    if (act && a.f2) r2 = interp_row<NPTS>(a.f2 + (size_t)comp * g.N, g, wx, wy, wz, q);
    u1[comp] = combine_rows(r1, wz);
    u2[comp] = a.f2 ? combine_rows(r2, wz) : 0.0f;
  }
  if (a.sk.s_cell) {
    // bookkeeping of the kept sort: the tile starts of the set this pass fills in, and the flag
    // that the sort before this pass consumed
    const int nth = b.nx * blockDim.x;
    if (a.sk.start_dst)
      for (int k = gid; k <= a.t.n; k += nth) a.sk.start_dst[k] = a.sk.start_src[k];
    if (gid == 0 && a.sk.clear_flag) *a.sk.clear_flag = 0;
  }
  if (s >= a.np || q != 0) return;
  interp_finish(a, s, pre, u1, u2);
}
FHD_KERNEL_PAIR((template <int NPTS>), 256, k_interp, InterpArgs, body_interp<NPTS>(a, b))


// ---------------- inertial coupling with excess mass ----------------
// Interpolation of the three components of f at one stencil (weights w): every lane of the
// particle's four gets the full sums (combine_rows broadcasts).
template <int NPTS>
FHD_DEV void interp3(const float* f, const GridDev& g, const SrcWeights& w, bool act, int q, float out[3]) {
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const AxisWeights& wx = comp == 0 ? w.stg[0] : w.cen[0];
    const AxisWeights& wy = comp == 1 ? w.stg[1] : w.cen[1];
    const AxisWeights& wz = comp == 2 ? w.stg[2] : w.cen[2];
    float r = 0.0f;
    if (act && f) r = interp_row<NPTS>(f + (size_t)comp * g.N, g, wx, wy, wz, q);
    out[comp] = combine_rows(r, wz);
  }
}

template <int NPTS>
FHD_DEV void weights_at(const GridDev& g, int4 c, float4 d, SrcWeights& w) {
  const int4 cw = make_int4(wrapi(c.x, g.mx), wrapi(c.y, g.my), wrapi(c.z, g.mz), 0);
  source_weights<NPTS>(cw, make_float3(d.x * g.idx, d.y * g.idy, d.z * g.idz), w, false);
}

FHD_DEV float4 f4_or_zero(const float4* p, int i) { return p ? p[i] : make_float4(0.f, 0.f, 0.f, 0.f); }

template <int NPTS>
FHD_BODY void body_pvec(const PVecArgs& a, const Blk& b) {
  const int gid = xcd_item(b.x, b.nx) * blockDim.x + threadIdx.x;
  const int s = gid >> 2, q = gid & 3;
  const GridDev& g = a.g;
  const bool act = s < a.np && q < NPTS;
  SrcWeights w, w0;
  int id = 0;
  if (s < a.np) {
    id = a.eval.id[s];
    weights_at<NPTS>(g, a.eval.cell[s], a.eval.off[s], w);
    if (a.f0) weights_at<NPTS>(g, a.cell0[id], a.off0[id], w0);
  }
  float j1[3], j0[3] = {0.f, 0.f, 0.f};
  interp3<NPTS>(a.f1, g, w, act, q, j1);
  if (a.f0) interp3<NPTS>(a.f0, g, w0, act, q, j0);
  if (s >= a.np || q != 0) return;
  const float4 A = f4_or_zero(a.A, id), B = f4_or_zero(a.B, id);
  a.out[id] = make_float4(a.a * A.x + a.b * B.x + a.c1 * j1[0] + a.c0 * j0[0],
                          a.a * A.y + a.b * B.y + a.c1 * j1[1] + a.c0 * j0[1],
                          a.a * A.z + a.b * B.z + a.c1 * j1[2] + a.c0 * j0[2], 0.0f);
  if (a.out2) {
    const float4 A2 = f4_or_zero(a.A2, id), B2 = f4_or_zero(a.B2, id);
    a.out2[id] = make_float4(a.a2 * A2.x + a.b2 * B2.x + a.d1 * j1[0], a.a2 * A2.y + a.b2 * B2.y + a.d1 * j1[1],
                             a.a2 * A2.z + a.b2 * B2.z + a.d1 * j1[2], 0.0f);
  }
}
FHD_KERNEL_PAIR((template <int NPTS>), 256, k_pvec, PVecArgs, body_pvec<NPTS>(a, b))

template <int NPTS>
FHD_BODY void body_lag2(const Lag2Args& a, const Blk& b) {
  const int gid = xcd_item(b.x, b.nx) * blockDim.x + threadIdx.x;
  const int s = gid >> 2, q = gid & 3;
  const GridDev& g = a.g;
  const bool act = s < a.np && q < NPTS;
  SrcWeights w;
  int id = 0;
  if (s < a.np) {
    id = a.eval.id[s];
    weights_at<NPTS>(g, a.eval.cell[s], a.eval.off[s], w);
  }
  float uh[3];
  interp3<NPTS>(a.fv, g, w, act, q, uh);
  // q' = q^n + hdt J(q^{n+1/2}) v_half, as (cell, offset)
  if (s < a.np) {
    int4 c = a.cell0[id];
    const float4 o = a.off0[id];
    float ox = o.x + a.hdt * uh[0], oy = o.y + a.hdt * uh[1], oz = o.z + a.hdt * uh[2];
    renorm_axis(c.x, ox, g.dx, g.idx);
    renorm_axis(c.y, oy, g.dy, g.idy);
    renorm_axis(c.z, oz, g.dz, g.idz);
    weights_at<NPTS>(g, c, make_float4(ox, oy, oz, 0.0f), w);
  }
  float l[3];
  interp3<NPTS>(a.fl, g, w, act, q, l);
  if (s >= a.np || q != 0) return;
  a.out[id] = make_float4(l[0], l[1], l[2], 0.0f);
}
FHD_KERNEL_PAIR((template <int NPTS>), 256, k_lag2, Lag2Args, body_lag2<NPTS>(a, b))

// ------------------------------- launchers -------------------------------
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

void launch_sort(const SortArgs& a, hipStream_t s) {
  if (t_recorder == nullptr && a.gate && a.bar && a.np > 0 && a.t.n <= kScanTilesMax) {
    const int groups = nblk(a.np, 256) < kSortGroups ? nblk(a.np, 256) : kSortGroups;
    launch(k_sort_one, k_sort_one_b, FAM_SORT_ONE, dim3(groups), 256, a, s);
    return;
  }
  if (a.np > 0) launch(k_bin, k_bin_b, FAM_BIN, dim3(nblk(a.np, 256)), 256, a, s);
  if (a.t.n <= kScanTilesMax) {
    launch(k_scan_scatter, k_scan_scatter_b, FAM_SCAN, dim3(nblk(a.np > 0 ? a.np : 1, 256)), 256, a, s);
  } else {
    // large grids: the single-workgroup scan clears this set's counts itself
    launch(k_scan, k_scan_b, FAM_SCAN, dim3(1), 1024, a, s);
    if (a.np > 0) launch(k_scatter, k_scatter_b, FAM_SCATTER, dim3(nblk(a.np, 256)), 256, a, s);
  }
}
bool wca_one_launch(const WcaArgs& a) { return t_recorder == nullptr && a.bar && a.np > 0; }
bool sort_wca_one_launch(const SortArgs& s, const WcaArgs& w) {
  return wca_one_launch(w) && s.gate && s.bar && s.np == w.np && s.t.n <= kScanTilesMax;
}
void launch_sort_wca(const SortArgs& s, const WcaArgs& w, hipStream_t st) {
  SortWcaArgs x;
  x.s = s;
  x.w = w;
  const int groups = nblk(w.np, 256) < kWcaOneGroups ? nblk(w.np, 256) : kWcaOneGroups;
  launch(k_sort_wca, k_sort_wca_b, FAM_WCA_LIST, dim3(groups), 256, x, st);
}
void launch_wca(const WcaArgs& a, hipStream_t s) {
  if (wca_one_launch(a)) {
    const int groups = nblk(a.np, 256) < kWcaOneGroups ? nblk(a.np, 256) : kWcaOneGroups;
    launch(k_wca_one, k_wca_one_b, FAM_WCA_LIST, dim3(groups), 256, a, s);
    return;
  }
  if (a.np > 0) launch(k_wca, k_wca_b, FAM_WCA, dim3(a.t.n < kWcaBuildGroups ? a.t.n : kWcaBuildGroups), kWcaThreads, a, s);
}
void launch_wca_list(const WcaArgs& a, hipStream_t s) {
  if (wca_one_launch(a)) return;  // done by launch_wca's one launch
  if (a.np > 0) launch(k_wca_list, k_wca_list_b, FAM_WCA_LIST, dim3(nblk(a.np, 256)), 256, a, s);
}
void launch_bonded(const BondArgs& a, hipStream_t s) {
  if (a.np > 0) launch(k_bonded, k_bonded_b, FAM_BONDED, dim3((a.np + kBondThreads - 1) / kBondThreads), kBondThreads, a, s);
}
void launch_src_prep(const SpreadArgs& a, hipStream_t s) {
  int nrec[2] = {0, 0};
  for (int gi = 0; gi < a.ngroups; ++gi) nrec[gi] = a.grp[gi].nsrc * (a.grp[gi].mode == SRC_RFD ? 2 : 1);
  SpreadArgs x = a;
  x.nrec0 = nrec[0];
  x.nrec_tot = nrec[0] + nrec[1];
  if (x.nrec_tot <= 0) return;
  const dim3 grid(nblk(x.nrec_tot, 256));
  if (a.kernel_points == 4) launch(k_src_prep<4>, k_src_prep_b<4>, FAM_SRC_PREP, grid, 256, x, s);
  else launch(k_src_prep<3>, k_src_prep_b<3>, FAM_SRC_PREP, grid, 256, x, s);
}
void launch_spread(const SpreadArgs& a, hipStream_t s) {
  const bool fast = a.t.tx == kTileEdge && a.t.ty == kTileEdge && a.t.tz == kTileEdge && a.t.ntx >= 2 &&
                    a.t.nty >= 2 && a.t.ntz >= 2;
  const dim3 grid(a.t.n);
  if (a.kernel_points == 4) {
    if (fast) launch(k_spread<4, true>, k_spread_b<4, true>, FAM_SPREAD, grid, kSpreadThreads, a, s);
    else launch(k_spread<4, false>, k_spread_b<4, false>, FAM_SPREAD, grid, kSpreadThreads, a, s);
  } else {
    if (fast) launch(k_spread<3, true>, k_spread_b<3, true>, FAM_SPREAD, grid, kSpreadThreads, a, s);
    else launch(k_spread<3, false>, k_spread_b<3, false>, FAM_SPREAD, grid, kSpreadThreads, a, s);
  }
}
void launch_pvec(const PVecArgs& a, hipStream_t s) {
  if (a.np <= 0) return;
  const dim3 grid(nblk(4 * a.np, 256));
  if (a.kernel_points == 4) launch(k_pvec<4>, k_pvec_b<4>, FAM_PVEC, grid, 256, a, s);
  else launch(k_pvec<3>, k_pvec_b<3>, FAM_PVEC, grid, 256, a, s);
}
void launch_lag2(const Lag2Args& a, hipStream_t s) {
  if (a.np <= 0) return;
  const dim3 grid(nblk(4 * a.np, 256));
  if (a.kernel_points == 4) launch(k_lag2<4>, k_lag2_b<4>, FAM_LAG2, grid, 256, a, s);
  else launch(k_lag2<3>, k_lag2_b<3>, FAM_LAG2, grid, 256, a, s);
}
void launch_interp(const InterpArgs& a, hipStream_t s) {
  if (a.np <= 0) return;
  const dim3 grid(nblk(4 * a.np, 256));
  if (a.kernel_points == 4) launch(k_interp<4>, k_interp_b<4>, FAM_INTERP, grid, 256, a, s);
  else launch(k_interp<3>, k_interp_b<3>, FAM_INTERP, grid, 256, a, s);
}

}  // namespace fhd
