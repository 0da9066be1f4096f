// fhd_kernels.h -- kernel argument structs and host-side launch entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fhd_device.h"
#include "fhd_launch.h"

namespace fhd {

constexpr int kMaxFftLen = 256;

struct FftArgs {
  GridDev g;
  const float2 *twx, *twy, *twz;  // e^{-2 pi i m / m_a}, m < m_a
  const float *gx, *gy, *gz;      // g_a(k) = (2/h_a) sin(pi k / m_a)
  const float2 *ex, *ey, *ez;     // e^{+i pi k / m_a}
  uint32_t plan_h, plan_y, plan_z;  // radices, 4 bits per stage
  int nst_h, nst_y, nst_z;
  int ztile;                      // kx columns per pass-B workgroup (generic kernels)
  int btile;                      // kx columns per pass-B workgroup (size-specialised: 2, 4, 8 or 16)
  int plan_xy;                    // radix plans of the forward xy pass (96 x 96, 128 x 128 planes): px + 3 py, 0 default
  int plan_xy_inv;                // the same for the inverse xy pass
  int plan_zs;                    // radix plan of the z pass (128-point columns), 0 default
  int zmode;                      // 0: FFT_z, operator, inverse FFT_z; 1: FFT_z only; 2: inverse only
  int ncomp;                      // velocity components transformed (3; 1 for a single field, transform-only modes)
  int knoise;                     // Stokes: the kept part of the stress divergence, drawn in Fourier space
                                  // and added before the solve (zmode 0): 0 none, 1 Sigma_1, 2 Sigma_1 + Sigma_2
  float kn_f1, kn_f4;             // the stress's diagonal / off-diagonal prefactors (f1^2 = 2 f4^2; the draw uses f4)
  float kn_sqrt_half_n;           // sqrt(N / 2)
  Key key;
  int generic;                    // 1: force the generic (runtime-size) kernels
  int stokes;                     // 1: v = P W / (eta |g|^2), mode 0 -> 0
  float eta;
  float cn_coef;                  // incompressible: 1 - cn_coef * L (0: projection only)
  float inv_n;
};

struct FluidWArgs {
  GridDev g;
  const float* v;          // W += v + lap_coef * L v (nullptr: term off)
  float lap_coef;
  const float* vadv;       // W += adv_coef * A(vadv) (nullptr: off)
  float adv_coef;
  const float* adv_prev;   // W += adv_prev_coef * adv_prev (nullptr: off)
  float adv_prev_coef;
  float* adv_store;        // stores A(vadv) when non-null
  const float* force;      // W += force_coef * force (nullptr: off)
  float force_coef;
  int noise;               // 0 none, 1 Sigma_1, 2 Sigma_1 + Sigma_2
  int trace_free;
  float f1, f4;            // diagonal / off-diagonal noise prefactors
  Key key;
  float* W;
  int zchunk;           // z planes per workgroup
  int zchunk_b;         // the same when the launch is recorded into a batched dispatch
};

struct FillArgs {
  float* p;
  size_t n;
};

struct StepArgs {
  unsigned long long* step;
};

// One pass of the 3D FFT: the transform parameters plus the pass's buffers.
struct FftPass {
  FftArgs f;
  const float* in;      // forward xy pass: real field
  float2* spec;         // half spectrum (in place for the z pass)
  float* out;           // inverse xy pass: out = alpha * result + beta * aux
  const float* aux;
  float alpha, beta;
  int T;                // z pass: kx columns per workgroup
};

// Tile decomposition used for particle binning and deterministic spreading.
struct Tiles {
  int tx, ty, tz;       // tile edge in cells
  int ntx, nty, ntz;    // tiles per axis
  int n;                // total tiles
};

// Counting sort of one position set by tile (bin, then scan + scatter).  The order inside a
// tile is the order of the binning atomics: every consumer is order-independent (integer
// accumulation in the spread and WCA passes, per-particle interpolation).
struct SortArgs {
  GridDev g;
  Tiles t;
  const int4* cell;     // unwrapped cells (by id)
  const float4* off;
  int* count;           // per-tile counts of this set (zero on entry)
  int* count_other;     // the other set's counts, cleared by the scatter pass for that set's next sort
  int* slot;            // rank of each particle within its tile, in atomic order
  int* start;           // tile_start[ntiles + 1]
  int* tmp;             // ids grouped by tile (large-grid path)
  int* sid;             // sorted outputs
  int4* scell;          // wrapped cell, w = the tile the slot is listed in
  float4* soff;
  int* slot_of;         // optional: sorted slot of each particle id
  const int* gate;      // optional: the sort runs only when *gate != 0 (the count clearing always runs)
  unsigned long long* bar;  // grid-barrier counter of the one-launch sort (only grows; zero at creation)
  int* err;             // the context's device error word
  int np;
};

struct SortedSet {
  const int4* cell;     // sorted copies
  const float4* off;
  const int* id;        // original particle index of each sorted slot
  const int* start;     // tile_start[ntiles + 1]
};

// Source groups for the spreading kernel.
enum { SRC_VALUE = 0, SRC_RFD = 1, SRC_DRIFT = 2 };
// Stencil record of one spreading source (128 bytes): per axis the centred weights and the face
// weights multiplied by the value's component along that axis (component a is the only one
// sampled on the a-faces, so each component's stencil is a product of one face row and two
// centred rows), then the first node of each stencil.
struct SrcRec {
  float w[2][3][4];     // [0: centred, 1: face x value][axis][k]
  int base[2][3];
  int pad[2];
};

struct SpreadGroup {
  SortedSet set;
  int nsrc;             // sorted slots in the set (particles)
  int mode;
  const float4* val;    // SRC_VALUE: per-particle vector value (by id)
  float coef;           // multiplies the value (SRC_RFD: kT/delta; SRC_DRIFT: kT ratio)
  float disp;           // SRC_RFD: displacement magnitude delta/2
  int rfd_quirk;        // SRC_RFD: weights at the undisplaced position
  SrcRec* rec;          // nsrc x (1 or 2) records, in sorted order
  int* cull;            // per record: bits 0-8 tile reach mask (3 bits per axis), bits 16-24 e + 192 where
                        // the record's largest face weight x value is below 2^e
};

struct SpreadArgs {
  GridDev g;
  Tiles t;
  int kernel_points;
  int ngroups;
  SpreadGroup grp[2];
  Key key;
  float out_coef;       // multiplies the accumulated field (1/dV)
  float* out;           // 3N field, every node written: out = out_add + coef * sum (out_add may be null)
  const float* out_add; // may alias out (each node is read, then written, by one thread)
  int nrec0, nrec_tot;  // records of group 0 and of both groups (set by launch_src_prep)
};

// Verlet neighbour lists of the WCA pass: every pair closer than rc + list skin when the list
// was built, kept until some particle has moved more than 0.45 list skins from where it was
// then (interpolation passes raise `flag`; the next force pass rebuilds).
constexpr int kNbCap = 32;  // neighbours stored per particle
struct NbList {
  int* idx;           // neighbour ids, column-major: entry j of particle i at idx[j * np + i]
  int* cnt;           // neighbours of each particle (> kNbCap: the list overflowed)
  int4* ref_cell;     // positions at the build (wrapped cell, offset), by id
  float4* ref_off;
  int* flag;          // 1: rebuild the lists before the next force pass
};

// The sorted copies of an updated position set are kept in place between sorts: a particle
// stays listed in its tile while it lies within `skin` cells of the tile along every axis that
// has three or more tiles; beyond that the pass raises *skin_flag and the set is sorted again
// before its next use.
struct SkinArgs {
  int4* s_cell;          // sorted copy to update (wrapped cell; w = listed tile, kept)
  float4* s_off;
  const int* s_slot;     // slot of each id in s_cell (null: the evaluated slot itself)
  int* skin_flag;        // set to 1 when a particle leaves its tile's skin
  int* clear_flag;       // set to 0 by the pass (a flag consumed by the sort before it)
  int* sid_dst;          // the new set's slot ids (its slot order is the evaluated set's), null: none
  const int* start_src;  // copied to start_dst (ntiles + 1 entries)
  int* start_dst;
  float skin[3];         // cells per axis; a negative value: the axis is never checked
};

struct InterpArgs {
  GridDev g;
  Tiles t;
  int kernel_points;
  SortedSet eval;        // positions where J is evaluated (sorted for locality)
  const int4* base_cell; // base positions (by id)
  const float4* base_off;
  int4* out_cell;        // out = base + a1 J f1 + a2 J f2 + extra (by id)
  float4* out_off;
  SkinArgs sk;           // sorted copies of out, kept current (s_cell null: none)
  const float* f1; float a1;
  const float* f2; float a2;
  float4* u_out;         // J f2 (or J f1 when f2 is null), optional
  const float4* F; float aF;           // extra mobility: + aF F
  float aN; int nnoise;                // + aN (Z_1 [+ Z_2]) particle noise
  Key key;
  int* err;
  NbList nb;             // neighbour lists to invalidate when a particle moves (flag null: none)
  float nb_move2;        // squared displacement from the list's reference that invalidates it
  int np;
};

// Per-particle combinations of interpolated fields (the inertial-coupling step with excess mass,
// src/quasiNeutrallyBuoyantFunctions2.cu:5-2290,2679-3149):
//   out[id]  = a A[id] + b B[id] + c1 J(f1) + c0 J0(f0)
//   out2[id] = a2 A2[id] + b2 B2[id] + d1 J(f1)           (when out2 is set)
// J interpolates at the sorted set `eval`, J0 at the by-id positions (cell0, off0).  Null
// vectors and fields count as zero.
struct PVecArgs {
  GridDev g;
  int kernel_points;
  SortedSet eval;
  const float* f1; float c1, d1;
  const float* f0; float c0;
  const int4* cell0; const float4* off0;
  const float4* A; float a;
  const float4* B; float b;
  float4* out;
  const float4* A2; float a2;
  const float4* B2; float b2;
  float4* out2;
  int np;
};

// Lagged term of the bootstrap step: q' = q0(id) + hdt J(fv) at `eval`, then out[id] = J(q') fl
// (findNeighborParticlesQuasiNeutrallyBuoyantTEST4_3 and interpolateLaplacianDeltaV_2,
// src/quasiNeutrallyBuoyantFunctions.cu:4872-5258,5879-6281).
struct Lag2Args {
  GridDev g;
  int kernel_points;
  SortedSet eval;
  const int4* cell0; const float4* off0;  // q^n by id (unwrapped cells)
  const float* fv; float hdt;
  const float* fl;
  float4* out;
  int np;
};

// lap = coef L dv, and v2 = al v + be v2 + ga dv (laplacianDeltaV, calculateVelocityAtHalfTimeStep,
// updateFluidQuasiNeutrallyBuoyantSemiImplicit).
struct MassFluidArgs {
  GridDev g;
  const float* dv;
  float* lap; float coef;
  const float* v; float* v2;
  float al, be, ga;
};

struct WcaArgs {
  GridDev g;
  Tiles t;
  SortedSet set;
  const int4* cell;   // the same positions by id (the list pass)
  const float4* off;
  float rc2, sig2, eps48;
  float rl2;       // list radius squared (rc + list skin)^2
  float skin_len;  // how far (length) a listed particle may lie outside its tile
  NbList nb;
  float4* F;   // by id
  unsigned long long* bar;  // grid-barrier counter of the one-launch WCA pass (only grows)
  int* err;    // the context's device error word
  int np;
};

// The particle sort and the WCA pass of one force evaluation, in one launch.
struct SortWcaArgs {
  SortArgs s;
  WcaArgs w;
};

// Harmonic bonds (V = k/2 (r - r0)^2) of each particle, in compressed rows by particle id:
// particle i's particle-particle bonds are pp[pp_start[i] .. pp_start[i+1]) and its fixed-point
// bonds fp*[fp_start[i] .. fp_start[i+1]).  A fixed point is stored like a position, as
// (cell, offset), so r_i - x is formed from exact cell differences.
struct BondArgs {
  GridDev g;
  const int4* cell;     // positions by id
  const float4* off;
  const int* pp_start;  // np + 1
  const int4* pp;       // {other particle, k (float bits), r0 (float bits), 0}
  const int* fp_start;  // np + 1
  const int4* fpc;      // {cell x, y, z, k (float bits)}
  const float4* fpo;    // {offset x, y, z, r0}
  float4* F;            // by id; the bond forces are added
  int np;
};

// Size-specialised FFT passes (fft_fast.hip); return false when the shape has no specialisation.
bool fast_fft_xy_fwd(const FftArgs& a, const float* in, float2* spec, hipStream_t s);
bool fast_fft_z_solve(const FftArgs& a, float2* spec, hipStream_t s);
bool fast_fft_xy_inv(const FftArgs& a, const float2* spec, float* out, const float* aux, float alpha, float beta,
                     hipStream_t s);

// Launch helpers (defined in the .hip files).
void launch_fft_xy_fwd(const FftArgs& a, const float* in, float2* spec, hipStream_t s);
// The 128 x 128 xy passes with the transforms in registers (which 0: forward, 1: inverse).
void launch_fft128_xy(const FftPass& x, int which, int mz, hipStream_t s);
void launch_fft_z_solve(const FftArgs& a, float2* spec, hipStream_t s);
void launch_fft_xy_inv(const FftArgs& a, const float2* spec, float* out, const float* aux, float alpha,
                       float beta, hipStream_t s);
void launch_fluid_w(const FluidWArgs& a, hipStream_t s);
void launch_divergence(const GridDev& g, const float* v, float* div, hipStream_t s);
void launch_normals(const Key& key, int stream, int sub, int n, float* out, hipStream_t s);
void launch_zero_f(float* p, size_t n, hipStream_t s);
void launch_advance_step(unsigned long long* step, hipStream_t s);
void launch_sk_accum(const float2* spec, double* acc, size_t n, hipStream_t s);  // acc += |spec|^2

void launch_sort(const SortArgs& a, hipStream_t s);  // bin, scan + scatter, per-tile ordering
// A context launched on its own takes the list build and the list force in one launch
// (launch_wca; launch_wca_list then does nothing); merged batch dispatches take two.
bool wca_one_launch(const WcaArgs& a);
// The sort (gated) and the WCA pass of one force evaluation in one launch, for a context
// launched on its own.
bool sort_wca_one_launch(const SortArgs& s, const WcaArgs& w);
void launch_sort_wca(const SortArgs& s, const WcaArgs& w, hipStream_t st);
void launch_wca(const WcaArgs& a, hipStream_t s);
void launch_wca_list(const WcaArgs& a, hipStream_t s);
void launch_bonded(const BondArgs& a, hipStream_t s);  // F += bond forces (after the WCA pass)
void launch_src_prep(const SpreadArgs& a, hipStream_t s);  // per-source records (run first)
void launch_spread(const SpreadArgs& a, hipStream_t s);    // tile accumulation
void launch_interp(const InterpArgs& a, hipStream_t s);
void launch_pvec(const PVecArgs& a, hipStream_t s);
void launch_lag2(const Lag2Args& a, hipStream_t s);
void launch_mass_fluid(const MassFluidArgs& a, hipStream_t s);
inline void launch_zero_f4(float4* p, int n, hipStream_t s) {
  if (n > 0) launch_zero_f(reinterpret_cast<float*>(p), 4 * (size_t)n, s);
}

}  // namespace fhd
