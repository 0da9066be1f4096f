// fhd_host.cpp -- host side of libfhd.so: data.main front end, the re-entrant step context
// that owns all device memory, the three scheme drivers and the C ABI of include/fhd.h.
//
// Everything runs on one HIP stream per context; a step is a fixed sequence of kernel
// launches with no host synchronisation, so it can be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <array>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fhd.h"
#include "fhd_kernels.h"

using namespace fhd;

namespace fhd {
thread_local Recorder* t_recorder = nullptr;
}

namespace {

thread_local std::string g_last_error;
fhd_params g_params;
bool g_params_loaded = false;

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e));     \
    }                                                                                  \
  } while (0)

// Radix plan for an n-point complex FFT: stages of 8, then 4, 2, 3, packed 4 bits each.
bool make_plan(int n, uint32_t* plan, int* nst) {
  uint32_t p = 0;
  int k = 0;
  int m = n;
  const int rs[4] = {8, 4, 2, 3};
  for (int r : rs)
    while (m % r == 0) {
      if (k >= 8) return false;
      p |= (uint32_t)r << (4 * k);
      ++k;
      m /= r;
    }
  if (m != 1) return false;
  *plan = p;
  *nst = k;
  return true;
}

int tile_edge(int m) {
  if (m % 8 == 0) return 8;
  if (m % 4 == 0) return 4;
  if (m <= 8) return m;
  return -1;
}

template <class T>
T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc(&p, n * sizeof(T)));
  return static_cast<T*>(p);
}

}  // namespace

// ------------------------------------------------------------------------------------------
// Device scratch of the no-allocation forward transform (fhd_fft_r2c_noalloc).
// ------------------------------------------------------------------------------------------
namespace {
// one component of a 128^3 half spectrum (8.5 MB), then the FFT tables
constexpr size_t kNoAllocSpecMax = (size_t)128 * 128 * 65;
constexpr size_t kNoAllocTab = 8 * (size_t)kMaxFftLen;  // 6 complex tables + 3 real ones (float2 units)
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) float2 scratch[kNoAllocSpecMax + kNoAllocTab];

namespace {
std::mutex g_noalloc_mu;
float2* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// Address of this GPU's copy of `scratch`, taken on the first call on each device (with that
// device current).  Host code uses this address, never the array's name.
float2* noalloc_scratch() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw std::runtime_error("hipGetSymbolAddress returned a null address");
    g_noalloc_ptr[dev] = static_cast<float2*>(ptr);
  }
  return g_noalloc_ptr[dev];
}

// The transform tables of build_tables, computed on the device from the same FP64 formulas:
// twiddles e^{-2 pi i k/m} (x, y, z), shifts e^{i pi k/m} (x, y, z), then the real symbols
// g(k) = 2 sin(pi k/m) (unit cells), kMaxFftLen entries per table.
__global__ void k_noalloc_tables(float2* tab, int mx, int my, int mz) {
  const int k = threadIdx.x;
  const int m[3] = {mx, my, mz};
  float* gt = reinterpret_cast<float*>(tab + 6 * kMaxFftLen);
  for (int ax = 0; ax < 3; ++ax) {
    if (k >= m[ax]) continue;
    const double a = -2.0 * M_PI * (double)k / (double)m[ax];
    tab[ax * kMaxFftLen + k] = make_float2((float)cos(a), (float)sin(a));
    const double e = M_PI * (double)k / (double)m[ax];
    tab[(3 + ax) * kMaxFftLen + k] = make_float2((float)cos(e), (float)sin(e));
    gt[ax * kMaxFftLen + k] = (float)(2.0 / 1.0 * sin(M_PI * (double)k / (double)m[ax]));
  }
}

// Half spectrum from the passes' layout [ky][z][kx] to the caller's [z][ky][kx].
__global__ void k_noalloc_to_host_layout(const float2* __restrict__ spec, float2* __restrict__ out, int my, int mz,
                                         int P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ns = (long long)mz * my * P;
  if (i >= ns) return;
  const int kx = (int)(i % P);
  const long long r = i / P;
  const int ky = (int)(r % my), z = (int)(r / my);
  out[i] = spec[((long long)ky * mz + z) * P + kx];
}
}  // namespace

// ------------------------------------------------------------------------------------------
// Context.
// ------------------------------------------------------------------------------------------
struct fhd_ctx {
  fhd_params p;
  std::string err;
  hipStream_t s = nullptr;
  GridDev g;
  Tiles t;
  FftArgs fa;
  double dx, dy, dz, dV;
  // fields (3N each)
  float *v = nullptr, *v2 = nullptr, *W = nullptr, *SF = nullptr, *adv[2] = {nullptr, nullptr};
  float* div = nullptr;
  float2* spec = nullptr;
  // Fourier tables
  float2 *twx = nullptr, *twy = nullptr, *twz = nullptr, *ex = nullptr, *ey = nullptr, *ez = nullptr;
  float *gx = nullptr, *gy = nullptr, *gz = nullptr;
  // particles: position sets 0 (q^n) and 1 (q^{n+1/2})
  int np = 0;
  int4* cell[2] = {nullptr, nullptr};
  float4* off[2] = {nullptr, nullptr};
  float4* F[2] = {nullptr, nullptr};
  float4* u = nullptr;
  int *count = nullptr, *slot = nullptr, *tmp = nullptr;  // count: 2 x ntiles (one per position set)
  int* slot_of = nullptr;  // sorted slot of each particle in set 0
  int* sflag = nullptr;    // [k]: position set k must be sorted before its next use
  unsigned long long* sbar = nullptr;  // grid-barrier counters of the one-launch sort and WCA pass
  // inertial coupling with excess mass: a field of work space and per-particle vectors by id
  float* M1 = nullptr;     // the projected correction dv~, then 1/2 dt nu L dv
  float4 *dulag = nullptr, *du = nullptr, *dp = nullptr, *jvt = nullptr, *ms = nullptr;
  float skin[3] = {-1, -1, -1};  // cells a listed particle may stray outside its tile (-1: any)
  float skin_len = 0.0f;
  // WCA neighbour lists (by id): kNbCap entries per particle, column-major; their reference
  // positions; the stale flag; the list skin (length) beyond the cutoff
  int *nb_idx = nullptr, *nb_cnt = nullptr, *nb_flag = nullptr;
  int4* nb_ref_cell = nullptr;
  float4* nb_ref_off = nullptr;
  float nb_skin = 0.0f;
  int* start[2] = {nullptr, nullptr};
  int* sid[2] = {nullptr, nullptr};
  int4* scell[2] = {nullptr, nullptr};
  float4* soff[2] = {nullptr, nullptr};
  SrcRec* rec[2] = {nullptr, nullptr};   // spreading source records per group (2 np each)
  int* cull[2] = {nullptr, nullptr};
  int* errd = nullptr;
  // harmonic bonds (bondedForces); null when the run has none
  int *bpp_start = nullptr, *bfp_start = nullptr;
  int4 *bpp = nullptr, *bfpc = nullptr;
  float4* bfpo = nullptr;
  long long step = 0;
  bool projected = false;
  bool qnb_started = false;
  bool qnb_steady = false;   // the current QNB step is an Adams-Bashforth one
  int adv_cur = 0;
  int last_force_set = 0;
  std::vector<std::pair<std::string, hipEvent_t>> marks;  // per-kernel profiling
  bool profiling = false;
  // graph replay of the steady-state step: one captured graph per step phase (QNB alternates
  // its velocity / advection-history buffers, so it has two phases)
  bool graph_mode = false;
  bool capturing = false;
  bool graphs_built = false;
  bool dev_step_valid = false;           // step_dev == step
  unsigned long long* step_dev = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int gperiod = 0, gphase = 0;
  bool own_stream = true;
  int zchunk = 8;                        // W kernel: z planes per workgroup (schedule)
  int zchunk_b = 8;                      // the same inside batched dispatches
  hipEvent_t done_ev = nullptr;          // fhd_record / fhd_query
  bool done_checked = false;             // error word read since the last fhd_record completed
  bool dev_error = false;                // fhd_query found a device error flag
  // Stokes: the stochastic part of W is built on a side stream while the particle kernels run
  hipStream_t s2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_noise = nullptr;
  bool noise_overlap = true;             // fhd_set_noise_overlap
  // equal-time structure factor: sum of |v_hat_a(k)|^2 over the samples, layout of spec
  double* sk_acc = nullptr;
  long long sk_samples = 0;
};

namespace {

void mark(fhd_ctx* c, const char* name) {
  if (!c->profiling) return;
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(e, c->s));
  c->marks.emplace_back(name, e);
}

Key make_key(const fhd_ctx* c) {
  Key k;
  k.k0 = (uint32_t)c->p.seed;
  k.k1 = (uint32_t)(c->p.seed >> 32);
  k.step_lo = (uint32_t)(uint64_t)c->step;
  k.step_hi = (uint32_t)((uint64_t)c->step >> 32);
  // a step being captured for replay reads the counter from the device at run time
  k.step = c->capturing ? c->step_dev : nullptr;
  return k;
}

void validate(const fhd_params& p) {
  if (p.scheme < 0 || p.scheme > 2) throw std::runtime_error("unknown scheme");
  if (p.mx < 2 || p.my < 1 || p.mz < 1) throw std::runtime_error("cells must be positive (mx >= 2)");
  if (p.mx % 2) throw std::runtime_error("mx must be even (real-to-complex FFT along x)");
  if (p.mx > kMaxFftLen || p.my > kMaxFftLen || p.mz > kMaxFftLen)
    throw std::runtime_error("cells per axis must be <= 256");
  uint32_t pl;
  int ns;
  if (!make_plan(p.mx / 2, &pl, &ns) || !make_plan(p.my, &pl, &ns) || !make_plan(p.mz, &pl, &ns))
    throw std::runtime_error("cells per axis must be of the form 2^a 3^b");
  if ((p.mx / 2 + 1) * p.my > 8448) throw std::runtime_error("xy plane too large for one workgroup");
  if (3 * p.mz > 8448) throw std::runtime_error("mz too large");
  if (!(p.lx > 0 && p.ly > 0 && p.lz > 0)) throw std::runtime_error("celldimension must be positive");
  if (!(p.dt > 0)) throw std::runtime_error("dt must be positive");
  if (p.densfluid <= 0 || p.shearviscosity <= 0) throw std::runtime_error("densfluid and shearviscosity must be positive");
  if (p.np < 0) throw std::runtime_error("numberparticles must be >= 0");
  if (p.np > 0) {
    if (p.scheme == FHD_SCHEME_INCOMPRESSIBLE)
      throw std::runtime_error("particles with incompressible: use quasiNeutrallyBuoyant or stokesLimit");
    if (p.kernel_points != 3 && p.kernel_points != 4) throw std::runtime_error("kernel must be 3 or 4");
    if (tile_edge(p.mx) < 0 || tile_edge(p.my) < 0 || tile_edge(p.mz) < 0)
      throw std::runtime_error("with particles, cells per axis must be a multiple of 4 or at most 8");
    if (p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) {
      // the particle's total mass m_e + m_f, m_f = 1.5 dV volumeParticle rho, must be positive
      const double mf = 1.5 * (p.lx / p.mx) * (p.ly / p.my) * (p.lz / p.mz) * p.volume_particle * p.densfluid;
      if (!(p.volume_particle > 0.0)) throw std::runtime_error("volumeParticle must be positive");
      if (!std::isfinite(p.mass) || !(p.mass + mf > 0.0))
        throw std::runtime_error("mass (the excess mass) must be finite and above -1.5 dV volumeParticle rho");
    }
    if (p.nonbonded && p.cutoff > 0) {
      double L[3] = {p.lx, p.ly, p.lz};
      int m[3] = {p.mx, p.my, p.mz};
      for (int a = 0; a < 3; ++a) {
        int te = tile_edge(m[a]);
        int nt = m[a] / te;
        double h = L[a] / m[a];
        if (nt >= 3 && p.cutoff > te * h) throw std::runtime_error("cutoff larger than a binning tile");
        if (p.cutoff >= 0.5 * L[a]) throw std::runtime_error("cutoff must be below half the box");
      }
    }
  }
}

void build_tables(fhd_ctx* c) {
  const int mx = c->p.mx, my = c->p.my, mz = c->p.mz;
  auto tw = [](int n) {
    std::vector<float2> t(n);
    for (int k = 0; k < n; ++k) {
      double a = -2.0 * M_PI * (double)k / (double)n;
      t[k] = make_float2((float)cos(a), (float)sin(a));
    }
    return t;
  };
  auto gk = [](int n, double h) {
    std::vector<float> t(n);
    for (int k = 0; k < n; ++k) t[k] = (float)(2.0 / h * sin(M_PI * (double)k / (double)n));
    return t;
  };
  auto ek = [](int n) {
    std::vector<float2> t(n);
    for (int k = 0; k < n; ++k) {
      double a = M_PI * (double)k / (double)n;
      t[k] = make_float2((float)cos(a), (float)sin(a));
    }
    return t;
  };
  auto up2 = [&](const std::vector<float2>& h, float2*& d) {
    d = dalloc<float2>(h.size());
    HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice));
  };
  auto up1 = [&](const std::vector<float>& h, float*& d) {
    d = dalloc<float>(h.size());
    HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  };
  up2(tw(mx), c->twx);
  up2(tw(my), c->twy);
  up2(tw(mz), c->twz);
  up2(ek(mx), c->ex);
  up2(ek(my), c->ey);
  up2(ek(mz), c->ez);
  up1(gk(mx, c->dx), c->gx);
  up1(gk(my, c->dy), c->gy);
  up1(gk(mz, c->dz), c->gz);
  FftArgs& a = c->fa;
  a.g = c->g;
  a.twx = c->twx; a.twy = c->twy; a.twz = c->twz;
  a.gx = c->gx; a.gy = c->gy; a.gz = c->gz;
  a.ex = c->ex; a.ey = c->ey; a.ez = c->ez;
  make_plan(mx / 2, &a.plan_h, &a.nst_h);
  make_plan(my, &a.plan_y, &a.nst_y);
  make_plan(mz, &a.plan_z, &a.nst_z);
  a.ztile = std::max(1, std::min(16, 8448 / (3 * mz)));
  a.btile = 8;
  a.plan_xy = 0;
  a.plan_xy_inv = 0;
  a.plan_zs = 0;
  a.zmode = 0;
  a.ncomp = 3;
  a.knoise = 0;
  a.kn_f1 = a.kn_f4 = 0.0f;
  a.kn_sqrt_half_n = (float)sqrt(0.5 * (double)c->g.N);
  memset(&a.key, 0, sizeof(a.key));
  const char* gen = getenv("FHD_FFT_GENERIC");  // 1: use the runtime-size FFT kernels everywhere
  a.generic = (gen && gen[0] == '1') ? 1 : 0;
  a.stokes = 0;
  a.eta = (float)c->p.shearviscosity;
  a.cn_coef = 0.0f;
  a.inv_n = (float)(1.0 / (double)c->g.N);
}

void free_ctx(fhd_ctx* c) {
  void* ptrs[] = {c->v, c->v2, c->W, c->SF, c->adv[0], c->adv[1], c->div, c->spec, c->twx, c->twy, c->twz,
                  c->ex, c->ey, c->ez, c->gx, c->gy, c->gz, c->cell[0], c->cell[1], c->off[0], c->off[1],
                  c->F[0], c->F[1], c->u, c->count, c->slot, c->tmp, c->start[0], c->start[1], c->sid[0],
                  c->sid[1], c->scell[0], c->scell[1], c->soff[0], c->soff[1], c->errd, c->rec[0], c->rec[1],
                  c->cull[0], c->cull[1], c->bpp_start, c->bfp_start, c->bpp, c->bfpc, c->bfpo, c->sk_acc,
                  c->slot_of, c->sflag, c->M1, c->dulag, c->du, c->dp, c->jvt, c->ms, c->nb_idx, c->nb_cnt,
                  c->nb_flag, c->nb_ref_cell, c->nb_ref_off, c->sbar};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (auto& m : c->marks) (void)hipEventDestroy(m.second);
  for (auto& e : c->gexec)
    if (e) (void)hipGraphExecDestroy(e);
  if (c->step_dev) (void)hipFree(c->step_dev);
  if (c->done_ev) (void)hipEventDestroy(c->done_ev);
  if (c->s2) {
    (void)hipStreamSynchronize(c->s2);
    (void)hipStreamDestroy(c->s2);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_noise) (void)hipEventDestroy(c->ev_noise);
  if (c->s && c->own_stream) (void)hipStreamDestroy(c->s);
}

// Host position <-> (cell, offset).
void to_cell(double q, double L, double h, int& cc, float& d) {
  double t = (q + 0.5 * L) / h;
  double f = floor(t);
  cc = (int)f;
  double off = q - ((f + 0.5) * h - 0.5 * L);
  d = (float)off;
  if (d >= 0.5f * (float)h) { cc += 1; d -= (float)h; }
  if (d < -0.5f * (float)h) { cc -= 1; d += (float)h; }
}
double from_cell(int cc, float d, double L, double h) { return ((double)cc + 0.5) * h - 0.5 * L + (double)d; }

void initial_positions(const fhd_params& p, std::vector<double>& q);

// ---------------- solves ----------------
void fft_forward(fhd_ctx* c, const float* in) {
  launch_fft_xy_fwd(c->fa, in, c->spec, c->s);
  mark(c, "fft_xy_fwd");
}
void fft_solve(fhd_ctx* c, int stokes, float cn_coef) {
  FftArgs a = c->fa;
  a.stokes = stokes;
  a.cn_coef = cn_coef;
  launch_fft_z_solve(a, c->spec, c->s);
  mark(c, "fft_z_solve");
}
void fft_ztransform(fhd_ctx* c, int zmode) {
  FftArgs a = c->fa;
  a.zmode = zmode;
  launch_fft_z_solve(a, c->spec, c->s);
}
void fft_inverse(fhd_ctx* c, float* out, const float* aux, float alpha, float beta) {
  launch_fft_xy_inv(c->fa, c->spec, out, aux, alpha, beta, c->s);
  mark(c, "fft_xy_inv");
}

void project_initial(fhd_ctx* c) {
  // src/runSchemeIncompressible.cu:79-97: project the initial velocity onto div-free fields
  fft_forward(c, c->v);
  fft_solve(c, 0, 0.0f);
  fft_inverse(c, c->v, nullptr, 1.0f, 0.0f);
  c->projected = true;
}

FluidWArgs base_w(fhd_ctx* c) {
  FluidWArgs w;
  memset(&w, 0, sizeof(w));
  w.g = c->g;
  w.key = make_key(c);
  w.W = c->W;
  w.zchunk = c->zchunk;
  w.zchunk_b = c->zchunk_b;
  return w;
}

// ---------------- schedule autotuner ----------------
// The z-pass column tile (4, 8 or 16 columns) and the W kernel's z-chunk (4, 8 or 16 planes)
// are timed once per grid shape on the first context of that shape and cached for the
// process: a replica's schedule depends on its shape only, never on what else runs beside it,
// and neither choice changes any result bit.
struct Sched {
  int btile, zchunk, zchunk_b, plan_xy, plan_xy_inv, plan_zs;
};
std::mutex g_sched_mu;
std::map<std::array<int, 4>, Sched> g_sched;

// Shortest of three timed rounds of reps launches (one warm-up launch first), so one slow
// round does not decide the schedule.
float time_launches(fhd_ctx* c, const std::function<void()>& f, int reps) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  f();  // warm-up
  float best = 1e30f;
  for (int round = 0; round < 3; ++round) {
    HIPCHK(hipEventRecord(a, c->s));
    for (int k = 0; k < reps; ++k) f();
    HIPCHK(hipEventRecord(b, c->s));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

void apply_schedule(fhd_ctx* c) {
  // keyed by shape and scheme: the W kernel is timed in the form the scheme launches
  const std::array<int, 4> key = {c->p.mx, c->p.my, c->p.mz, c->p.scheme};
  const char* env = getenv("FHD_AUTOTUNE");  // 0: fixed defaults (8 columns, 8 planes)
  const bool tune = !(env && env[0] == '0');
  std::lock_guard<std::mutex> lock(g_sched_mu);
  auto it = g_sched.find(key);
  Sched sc = {8, 8, 8, 0, 0, 0};
  if (it != g_sched.end()) {
    sc = it->second;
  } else if (tune) {
    // scratch runs on the context's own buffers before its state exists (contents irrelevant)
    float best = 1e30f;
    const bool stokes = c->p.scheme == FHD_SCHEME_STOKES_LIMIT;
    // z pass: column tile and, for 48-, 96- and 128-point columns, the radix plan (128: 16 8,
    // 8 16 or 4 4 8; 96: 16 2 3, 8 4 3 or 3 8 4; 48: 16 3, 3 16 or 4 4 3), timed in the form
    // the scheme launches (the Stokes pass draws its stochastic stress)
    FftArgs za = c->fa;
    if (stokes) {
      za.stokes = 1;
      za.knoise = 2;
      za.kn_f1 = za.kn_f4 = 1.0f;
      za.key = make_key(c);
    }
    const int mzp = c->p.mz;
    const int nzp = (mzp == 48 || mzp == 96 || mzp == 128) ? 3 : 1;
    for (int pz = 0; pz < nzp; ++pz)
      for (int bt : {2, 4, 8, 16}) {  // 2: more column workgroups on small grids
        za.btile = bt;
        za.plan_zs = pz;
        const float t = time_launches(c, [&] { launch_fft_z_solve(za, c->spec, c->s); }, 4);
        if (t < best) { best = t; sc.btile = bt; sc.plan_zs = pz; }
      }
    c->fa.plan_zs = sc.plan_zs;
    // xy passes of 96 x 96 and 128 x 128 planes: the radix plans of the x and y transforms
    // (three each) of the LDS-stage kernels, and for 128 x 128 the register-transform kernels
    // (plan 9), chosen for the forward and the inverse pass separately
    const bool sq128 = c->p.mx == 128 && c->p.my == 128, sq96 = c->p.mx == 96 && c->p.my == 96;
    if (sq128 || sq96) {
      best = 1e30f;
      float best_inv = 1e30f;
      for (int pxy = 0; pxy < (sq128 ? 10 : 9); ++pxy) {  // 9: the register-transform kernels
        c->fa.plan_xy = c->fa.plan_xy_inv = pxy;
        const float tf = time_launches(c, [&] { launch_fft_xy_fwd(c->fa, c->W, c->spec, c->s); }, 4);
        const float ti =
            time_launches(c, [&] { launch_fft_xy_inv(c->fa, c->spec, c->W, nullptr, 1.0f, 0.0f, c->s); }, 4);
        if (tf < best) { best = tf; sc.plan_xy = pxy; }
        if (ti < best_inv) { best_inv = ti; sc.plan_xy_inv = pxy; }
      }
      c->fa.plan_xy = sc.plan_xy;
      c->fa.plan_xy_inv = sc.plan_xy_inv;
    }
    // W kernel (incompressible and QNB; the Stokes step has none): 2 planes per workgroup fill
    // the chip on small grids (64^3: 512 workgroups) when the context launches alone; a
    // batched dispatch already holds many contexts' workgroups, where the redundant halo planes
    // of short chunks cost more, so it keeps the best of 4-16
    best = 1e30f;
    float best_b = 1e30f;
    for (int zc : {2, 4, 8, 16}) {
      if (stokes) break;
      c->zchunk = zc;
      FluidWArgs w = base_w(c);
      w.noise = 1;
      w.f1 = w.f4 = 1.0f;
      w.v = c->v;
      w.vadv = c->v;
      w.lap_coef = 1.0f;
      w.adv_coef = 1.0f;
      w.trace_free = 1;
      const float t = time_launches(c, [&] { launch_fluid_w(w, c->s); }, 4);
      if (t < best) { best = t; sc.zchunk = zc; }
      if (zc >= 4 && t < best_b) { best_b = t; sc.zchunk_b = zc; }
    }
    g_sched[key] = sc;
  }
  c->fa.btile = sc.btile;
  c->fa.plan_xy = sc.plan_xy;
  c->fa.plan_xy_inv = sc.plan_xy_inv;
  c->fa.plan_zs = sc.plan_zs;
  c->zchunk = sc.zchunk;
  c->zchunk_b = sc.zchunk_b;
  // explicit radix plans (tests and experiments): FHD_PLAN_XY in 0-9 (128 x 128 planes; 0-8
  // for 96 x 96), FHD_PLAN_Z in 0-2 (48-, 96- and 128-point columns)
  if (const char* e = getenv("FHD_PLAN_XY")) {
    const int v = atoi(e);
    if (v < 0 || v > 9) throw std::runtime_error("FHD_PLAN_XY must be 0-9");
    c->fa.plan_xy = c->fa.plan_xy_inv = v;
  }
  if (const char* e = getenv("FHD_PLAN_Z")) {
    const int v = atoi(e);
    if (v < 0 || v > 2) throw std::runtime_error("FHD_PLAN_Z must be 0-2");
    c->fa.plan_zs = v;
  }
  // explicit schedule (tests and experiments): FHD_BTILE in {2, 4, 8, 16}, FHD_ZCHUNK >= 1
  if (const char* e = getenv("FHD_BTILE")) {
    const int v = atoi(e);
    if (v != 2 && v != 4 && v != 8 && v != 16) throw std::runtime_error("FHD_BTILE must be 2, 4, 8 or 16");
    c->fa.btile = v;
  }
  if (const char* e = getenv("FHD_ZCHUNK")) {
    const int v = atoi(e);
    if (v < 1) throw std::runtime_error("FHD_ZCHUNK must be at least 1");
    c->zchunk = c->zchunk_b = v;
  }
}

// ---------------- particles ----------------
// Counting sort of position set k by tile (gated off on the device while every particle is
// still inside its listed tile's skin).
SortArgs sort_args(fhd_ctx* c, int k) {
  SortArgs a;
  a.g = c->g;
  a.t = c->t;
  a.cell = c->cell[k];
  a.off = c->off[k];
  a.count = c->count + (size_t)k * c->t.n;
  a.count_other = c->count + (size_t)(1 - k) * c->t.n;
  a.gate = c->sflag + k;
  a.slot_of = k == 0 ? c->slot_of : nullptr;
  a.slot = c->slot;
  a.start = c->start[k];
  a.tmp = c->tmp;
  a.sid = c->sid[k];
  a.scell = c->scell[k];
  a.soff = c->soff[k];
  a.bar = c->sbar;
  a.err = c->errd;
  a.np = c->np;
  return a;
}
void sort_set(fhd_ctx* c, int k) {
  launch_sort(sort_args(c, k), c->s);
  mark(c, "sort");
}

SortedSet sset(fhd_ctx* c, int k) {
  SortedSet s;
  s.cell = c->scell[k];
  s.off = c->soff[k];
  s.id = c->sid[k];
  s.start = c->start[k];
  return s;
}
// Harmonic bond forces added to F (position set k), after the WCA pass wrote it.
void bonded(fhd_ctx* c, int k, float4* F) {
  if (!c->bpp) return;
  BondArgs a;
  a.g = c->g;
  a.cell = c->cell[k];
  a.off = c->off[k];
  a.pp_start = c->bpp_start;
  a.pp = c->bpp;
  a.fp_start = c->bfp_start;
  a.fpc = c->bfpc;
  a.fpo = c->bfpo;
  a.F = F;
  a.np = c->np;
  launch_bonded(a, c->s);
  mark(c, "bonded");
}
NbList nb_list(fhd_ctx* c) {
  NbList n;
  n.idx = c->nb_idx;
  n.cnt = c->nb_cnt;
  n.ref_cell = c->nb_ref_cell;
  n.ref_off = c->nb_ref_off;
  n.flag = c->nb_flag;
  return n;
}

WcaArgs wca_args(fhd_ctx* c, int k, float4* F) {
  WcaArgs a;
  a.g = c->g;
  a.t = c->t;
  a.set = sset(c, k);
  a.cell = c->cell[k];
  a.off = c->off[k];
  a.rc2 = (float)(c->p.cutoff * c->p.cutoff);
  a.rl2 = (float)((c->p.cutoff + c->nb_skin) * (c->p.cutoff + c->nb_skin));
  a.sig2 = (float)(c->p.wca_sigma * c->p.wca_sigma);
  a.eps48 = (float)(48.0 * c->p.wca_epsilon);
  a.skin_len = c->skin_len;
  a.nb = nb_list(c);
  a.F = F;
  a.bar = c->sbar + 1;
  a.err = c->errd;
  a.np = c->np;
  return a;
}

void wca(fhd_ctx* c, int k, float4* F) {
  if (!c->p.nonbonded || c->p.cutoff <= 0) {
    launch_zero_f4(F, c->np, c->s);
    bonded(c, k, F);
    return;
  }
  const WcaArgs a = wca_args(c, k, F);
  if (wca_one_launch(a)) {
    launch_wca(a, c->s);  // the list build (while the lists are stale) and the force in one launch
    mark(c, "wca");
  } else {
    launch_wca(a, c->s);  // the list build, which returns at once while the lists are valid
    mark(c, "wca_build");
    launch_wca_list(a, c->s);
    mark(c, "wca");
  }
  bonded(c, k, F);
}
// Sort position set k (gated) and take the WCA forces (plus bonded ones) into F: for a context
// launched on its own with WCA, the sort and the WCA pass are one launch.
void sort_and_wca(fhd_ctx* c, int k, float4* F) {
  if (c->p.nonbonded && c->p.cutoff > 0) {
    const SortArgs s = sort_args(c, k);
    const WcaArgs w = wca_args(c, k, F);
    if (sort_wca_one_launch(s, w)) {
      launch_sort_wca(s, w, c->s);
      mark(c, "wca");
      bonded(c, k, F);
      return;
    }
  }
  sort_set(c, k);
  wca(c, k, F);
}
// Attach the per-group record buffers, then prepare the source records and spread.
void run_spread(fhd_ctx* c, SpreadArgs& sa) {
  for (int k = 0; k < sa.ngroups; ++k) {
    sa.grp[k].nsrc = c->np;
    sa.grp[k].rec = c->rec[k];
    sa.grp[k].cull = c->cull[k];
  }
  launch_src_prep(sa, c->s);
  mark(c, "src_prep");
  launch_spread(sa, c->s);
  mark(c, "spread");
}

SpreadArgs spread_base(fhd_ctx* c) {
  SpreadArgs a;
  memset(&a, 0, sizeof(a));
  a.g = c->g;
  a.t = c->t;
  a.kernel_points = c->p.kernel_points;
  a.key = make_key(c);
  a.out_coef = (float)(1.0 / c->dV);
  a.out = c->SF;
  return a;
}
InterpArgs interp_base(fhd_ctx* c, int eval_set) {
  InterpArgs a;
  memset(&a, 0, sizeof(a));
  a.g = c->g;
  a.t = c->t;
  a.kernel_points = c->p.kernel_points;
  a.eval = sset(c, eval_set);
  a.key = make_key(c);
  a.err = c->errd;
  a.nb = nb_list(c);
  const double move = 0.45 * c->nb_skin;
  a.nb_move2 = (float)(move * move);
  a.np = c->np;
  return a;
}

// The kept sort.  Interpolation passes that move particles also keep the sorted copies of the
// position set they fill: from q^n (set 0) to q^{n+1/2} (set 1) the slot order of set 0 is
// reused, and the next update writes q^{n+1} back into set 0's slots.  A particle that leaves
// its tile's skin raises the set's flag, and the set is sorted again (the sort kernels are
// gated on that flag).
void skin_to_half(fhd_ctx* c, InterpArgs& ia) {
  SkinArgs& k = ia.sk;
  k.s_cell = c->scell[1];
  k.s_off = c->soff[1];
  k.s_slot = nullptr;
  k.skin_flag = c->sflag + 1;
  k.clear_flag = c->sflag;
  k.sid_dst = c->sid[1];
  k.start_src = c->start[0];
  k.start_dst = c->start[1];
  for (int a = 0; a < 3; ++a) k.skin[a] = c->skin[a];
}
void skin_to_full(fhd_ctx* c, InterpArgs& ia) {
  SkinArgs& k = ia.sk;
  k.s_cell = c->scell[0];
  k.s_off = c->soff[0];
  k.s_slot = c->slot_of;
  k.skin_flag = c->sflag;
  k.clear_flag = c->sflag + 1;
  k.sid_dst = nullptr;
  k.start_src = nullptr;
  k.start_dst = nullptr;
  for (int a = 0; a < 3; ++a) k.skin[a] = c->skin[a];
}

// ---------------- schemes ----------------
// Incompressible: src/runSchemeIncompressible.cu:105-139.
void step_incompressible(fhd_ctx* c) {
  if (!c->projected) project_initial(c);
  const fhd_params& p = c->p;
  const double dV = c->dV, rho = p.densfluid;
  FluidWArgs w = base_w(c);
  w.v = c->v;
  w.lap_coef = (float)(0.5 * p.dt * p.shearviscosity / rho);
  w.vadv = c->v;
  w.adv_coef = (float)(-p.dt);
  w.noise = 1;
  w.trace_free = 1;
  w.f1 = (float)sqrt(4.0 * p.temperature * p.shearviscosity * p.dt / (dV * rho * rho));
  w.f4 = (float)sqrt(2.0 * p.temperature * p.shearviscosity * p.dt / (dV * rho * rho));
  const float cn = (float)(0.5 * p.dt * p.shearviscosity / rho);
  launch_fluid_w(w, c->s);
  mark(c, "fluid_w");
  fft_forward(c, c->W);
  fft_solve(c, 0, cn);
  fft_inverse(c, c->v2, c->v, 0.5f, 0.5f);  // v* = (v~ + v^n) / 2
  w.vadv = c->v2;
  launch_fluid_w(w, c->s);
  mark(c, "fluid_w");
  fft_forward(c, c->W);
  fft_solve(c, 0, cn);
  fft_inverse(c, c->v, nullptr, 1.0f, 0.0f);
}

// Stokes limit, midpoint: src/runSchemeStokesLimit.cu:163-276,
// src/boundaryParticlesFunctionStokesLimit.cu:21-249.
// W = S F (+ RFD drift) + noise D.Sigma.  The noise is drawn in Fourier space by the z pass as
// the part of D.Sigma that the solve keeps (kspace_noise: two complex normals per mode in the
// plane normal to g, the law of the projected divergence of a white stress whose prefactors
// satisfy f1^2 = 2 f4^2, as fluam's do), so W in real space holds only the particle terms.
void fft_solve_stokes(fhd_ctx* c, int knoise, float f1, float f4) {
  FftArgs a = c->fa;
  a.stokes = 1;
  a.cn_coef = 0.0f;
  a.knoise = (f1 != 0.0f || f4 != 0.0f) ? knoise : 0;
  a.kn_f1 = f1;
  a.kn_f4 = f4;
  a.key = make_key(c);
  launch_fft_z_solve(a, c->spec, c->s);
  mark(c, "fft_z_solve");
}

void step_stokes(fhd_ctx* c) {
  const fhd_params& p = c->p;
  const double dV = c->dV, dt = p.dt;
  const float f1 = (float)sqrt(4.0 * p.temperature * p.shearviscosity / (dV * dt));
  const float f4 = (float)sqrt(2.0 * p.temperature * p.shearviscosity / (dV * dt));
  const float fact2 = (float)sqrt(p.temperature * p.extra_mobility * dt);
  const bool xm = p.extra_mobility != 0.0;
  // solve 1: W = S F^n, and the kept part of sqrt(2) D.Sigma_1, drawn in Fourier space
  if (c->np > 0) {
    sort_and_wca(c, 0, c->F[0]);
    SpreadArgs sa = spread_base(c);
    sa.ngroups = 1;
    sa.grp[0].set = sset(c, 0);
    sa.grp[0].mode = SRC_VALUE;
    sa.grp[0].val = c->F[0];
    sa.grp[0].coef = 1.0f;
    sa.out = c->W;
    sa.out_add = nullptr;
    run_spread(c, sa);
  } else {
    launch_zero_f(c->W, 3 * (size_t)c->g.N, c->s);
  }
  fft_forward(c, c->W);
  fft_solve_stokes(c, 1, (float)(sqrt(2.0) * f1), (float)(sqrt(2.0) * f4));
  fft_inverse(c, c->v, nullptr, 1.0f, 0.0f);
  if (c->np > 0) {
    InterpArgs ia = interp_base(c, 0);
    ia.base_cell = c->cell[0];
    ia.base_off = c->off[0];
    ia.out_cell = c->cell[1];
    ia.out_off = c->off[1];
    skin_to_half(c, ia);
    ia.f1 = c->v;
    ia.a1 = (float)(0.5 * dt);
    if (xm) {
      ia.F = c->F[0];
      ia.aF = (float)(0.5 * dt * p.extra_mobility);
      ia.aN = fact2;
      ia.nnoise = 1;
    }
    launch_interp(ia, c->s);
    mark(c, "interp");
    sort_and_wca(c, 1, c->F[1]);
    SpreadArgs sa = spread_base(c);
    sa.ngroups = 2;
    sa.grp[0].set = sset(c, 1);
    sa.grp[0].mode = SRC_VALUE;
    sa.grp[0].val = c->F[1];
    sa.grp[0].coef = 1.0f;
    sa.grp[1].set = sset(c, 0);
    sa.grp[1].mode = SRC_RFD;
    sa.grp[1].coef = (float)(p.temperature / p.rfd_delta);
    sa.grp[1].disp = (float)(0.5 * p.rfd_delta);
    sa.grp[1].rfd_quirk = p.rfd_mode == FHD_RFD_REFERENCE;
    sa.out = c->W;
    sa.out_add = nullptr;
    run_spread(c, sa);
  } else {
    launch_zero_f(c->W, 3 * (size_t)c->g.N, c->s);
  }
  // solve 2: W = S F^{n+1/2} + RFD drift, and the kept part of sqrt(1/2) D.(Sigma_1 + Sigma_2)
  fft_forward(c, c->W);
  fft_solve_stokes(c, 2, (float)(sqrt(0.5) * f1), (float)(sqrt(0.5) * f4));
  fft_inverse(c, c->v, nullptr, 1.0f, 0.0f);
  if (c->np > 0) {
    InterpArgs ia = interp_base(c, 1);
    ia.base_cell = c->cell[0];
    ia.base_off = c->off[0];
    ia.out_cell = c->cell[0];
    ia.out_off = c->off[0];
    skin_to_full(c, ia);
    ia.f1 = c->v;
    ia.a1 = (float)dt;
    ia.u_out = c->u;
    if (xm) {
      ia.F = c->F[1];
      ia.aF = (float)(dt * p.extra_mobility);
      ia.aN = fact2;
      ia.nnoise = 2;
    }
    launch_interp(ia, c->s);
    mark(c, "interp");
    c->last_force_set = 1;
  }
}

// ---------------- inertial coupling with excess mass ----------------
// The corrections of one QNB step with m_e != 0 after the unperturbed prediction v~ (in v2),
// at q^{n+1/2} (sorted set 1) with q^n by id (src/runSchemeQuasiNeutrallyBuoyant.cu:336-656):
//   du  = lag + J(q^{n+1/2}) v^n - J(q^n) v^n                     kernelCalculateDeltau
//   dp  = m_e (u - J v~ - du)                                      kernelCalculateDeltap
//   dv~ = P [sum phi f dp], f = 1.5 V_p / (1.5 V_p dV rho + m_e)  -> M1
//   s   = dp - m_e J dv~;  u' = J v~ + J dv~ + du (update_u)       updateParticleVelocityme
//   dv  = (I - dt nu/2 L)^-1 P S(s) / rho                          -> W
// `lag` false: the bootstrap predictor, which has no lagged term yet.
PVecArgs pvec_base(fhd_ctx* c) {
  PVecArgs a;
  memset(&a, 0, sizeof(a));
  a.g = c->g;
  a.kernel_points = c->p.kernel_points;
  a.eval = sset(c, 1);
  a.np = c->np;
  return a;
}
void pvec(fhd_ctx* c, const PVecArgs& a) {
  launch_pvec(a, c->s);
  mark(c, "mass");
}
void qnb_mass_correction(fhd_ctx* c, bool lag, bool update_u) {
  const fhd_params& p = c->p;
  const double me = p.mass, rho = p.densfluid, dV = c->dV, Vp = p.volume_particle;
  const float cn = (float)(0.5 * p.dt * p.shearviscosity / rho);
  PVecArgs a = pvec_base(c);
  a.f1 = c->v;
  a.c1 = 1.0f;
  a.f0 = c->v;
  a.c0 = -1.0f;
  a.cell0 = c->cell[0];
  a.off0 = c->off[0];
  a.A = lag ? c->dulag : nullptr;
  a.a = 1.0f;
  a.out = c->du;
  pvec(c, a);
  a = pvec_base(c);
  a.f1 = c->v2;
  a.c1 = (float)-me;
  a.d1 = 1.0f;
  a.A = c->u;
  a.a = (float)me;
  a.B = c->du;
  a.b = (float)-me;
  a.out = c->dp;
  a.out2 = c->jvt;
  pvec(c, a);
  const double f = 1.5 * Vp / (1.5 * Vp * dV * rho + me);  // src/quasiNeutrallyBuoyantFunctions2.cu:1332-1335
  SpreadArgs sa = spread_base(c);
  sa.ngroups = 1;
  sa.grp[0].set = sset(c, 1);
  sa.grp[0].mode = SRC_VALUE;
  sa.grp[0].val = c->dp;
  sa.grp[0].coef = (float)(f * dV);  // the pass scales by 1/dV: sum phi f dp
  sa.out = c->W;
  sa.out_add = nullptr;
  run_spread(c, sa);
  fft_forward(c, c->W);
  fft_solve(c, 0, 0.0f);  // projectionDivergenceFree (mode 0 kept)
  fft_inverse(c, c->M1, nullptr, 1.0f, 0.0f);
  a = pvec_base(c);
  a.f1 = c->M1;
  a.c1 = (float)-me;
  a.A = c->dp;
  a.a = 1.0f;
  a.out = c->ms;
  if (update_u) {
    a.d1 = 1.0f;
    a.A2 = c->jvt;
    a.a2 = 1.0f;
    a.B2 = c->du;
    a.b2 = 1.0f;
    a.out2 = c->u;
  }
  pvec(c, a);
  sa.grp[0].val = c->ms;
  sa.grp[0].coef = (float)(1.0 / rho);  // S(s) / rho = sum phi s / (rho dV)
  run_spread(c, sa);
  fft_forward(c, c->W);
  fft_solve(c, 0, cn);
  fft_inverse(c, c->W, nullptr, 1.0f, 0.0f);
}
// M1 = 1/2 dt nu L dv (dv in W) and v2 = al v + be v2 + ga dv.
void qnb_mass_fluid(fhd_ctx* c, float al, float be, float ga) {
  MassFluidArgs m;
  m.g = c->g;
  m.dv = c->W;
  m.lap = c->M1;
  m.coef = (float)(0.5 * c->p.dt * c->p.shearviscosity / c->p.densfluid);
  m.v = c->v;
  m.v2 = c->v2;
  m.al = al;
  m.be = be;
  m.ga = ga;
  launch_mass_fluid(m, c->s);
  mark(c, "mass");
}

// Host-side state a QNB step leaves behind (also applied after each replay of a captured step):
// the new velocity becomes v, and in the Adams-Bashforth regime the history buffers alternate.
void qnb_after_step(fhd_ctx* c) {
  if (c->qnb_steady) c->adv_cur = 1 - c->adv_cur;
  c->qnb_steady = c->qnb_started;
  std::swap(c->v, c->v2);
}

// Quasi-neutrally-buoyant particles, excess mass 0: src/runSchemeQuasiNeutrallyBuoyant.cu:178-677
// with the first-step bootstrap of src/firstStepQuasiNeutrallyBuoyant.cu:35-1152.
void step_qnb(fhd_ctx* c) {
  if (!c->projected) project_initial(c);
  const fhd_params& p = c->p;
  const double dV = c->dV, rho = p.densfluid, dt = p.dt;
  FluidWArgs w = base_w(c);
  w.v = c->v;
  w.lap_coef = (float)(0.5 * dt * p.shearviscosity / rho);
  w.noise = 1;
  w.trace_free = 1;
  w.f1 = (float)sqrt(4.0 * p.temperature * p.shearviscosity * dt / (dV * rho * rho));
  w.f4 = (float)sqrt(2.0 * p.temperature * p.shearviscosity * dt / (dV * rho * rho));
  const float cn = (float)(0.5 * dt * p.shearviscosity / rho);
  if (c->np > 0) {
    sort_set(c, 0);
    InterpArgs ia = interp_base(c, 0);
    ia.base_cell = c->cell[0];
    ia.base_off = c->off[0];
    ia.out_cell = c->cell[1];
    ia.out_off = c->off[1];
    skin_to_half(c, ia);
    ia.f1 = c->v;
    ia.a1 = (float)(0.5 * dt);
    launch_interp(ia, c->s);
    mark(c, "interp");
    sort_and_wca(c, 1, c->F[1]);
    const double mf = 1.5 * dV * p.volume_particle * rho;
    const double ratio = (p.mass + mf) != 0.0 ? mf / (p.mass + mf) : 1.0;
    SpreadArgs sa = spread_base(c);
    sa.ngroups = 2;
    sa.grp[0].set = sset(c, 1);
    sa.grp[0].mode = SRC_VALUE;
    sa.grp[0].val = c->F[1];
    sa.grp[0].coef = 1.0f;
    sa.grp[1].set = sset(c, 1);
    sa.grp[1].mode = SRC_DRIFT;
    sa.grp[1].coef = (float)(p.temperature * ratio);
    run_spread(c, sa);
    w.force = c->SF;
    w.force_coef = (float)(dt / rho);
    c->last_force_set = 1;
  }
  const bool mass = c->M1 != nullptr;
  if (!c->qnb_started) {
    // predictor with first-order advection; stores A(v^n) as the Adams-Bashforth history
    w.vadv = c->v;
    w.adv_coef = (float)(-dt);
    w.adv_store = c->adv[c->adv_cur];
    launch_fluid_w(w, c->s);
    mark(c, "fluid_w");
    fft_forward(c, c->W);
    fft_solve(c, 0, cn);
    if (!mass) {
      fft_inverse(c, c->v2, c->v, 0.5f, 0.5f);  // v^{n+1/2}
    } else {
      // the predictor's excess-mass corrections (no lagged term yet, u kept), v^{n+1/2} =
      // (v^n + v~ + dv) / 2, and the corrector's lagged term J(q') 1/2 dt nu L dv at
      // q' = q^n + dt/2 J(q^{n+1/2}) v^{n+1/2} (src/firstStepQuasiNeutrallyBuoyant.cu:84,229-560)
      fft_inverse(c, c->v2, nullptr, 1.0f, 0.0f);
      qnb_mass_correction(c, false, false);
      qnb_mass_fluid(c, 0.5f, 0.5f, 0.5f);
      Lag2Args l;
      l.g = c->g;
      l.kernel_points = p.kernel_points;
      l.eval = sset(c, 1);
      l.cell0 = c->cell[0];
      l.off0 = c->off[0];
      l.fv = c->v2;
      l.hdt = (float)(0.5 * dt);
      l.fl = c->M1;
      l.out = c->dulag;
      l.np = c->np;
      launch_lag2(l, c->s);
      mark(c, "mass");
    }
    // corrector: advection at v^{n+1/2}
    w.vadv = c->v2;
    w.adv_store = nullptr;
    launch_fluid_w(w, c->s);
    mark(c, "fluid_w");
    c->qnb_started = true;
  } else {
    w.vadv = c->v;
    w.adv_coef = (float)(-1.5 * dt);
    w.adv_prev = c->adv[c->adv_cur];
    w.adv_prev_coef = (float)(0.5 * dt);
    w.adv_store = c->adv[1 - c->adv_cur];
    launch_fluid_w(w, c->s);
    mark(c, "fluid_w");
  }
  fft_forward(c, c->W);
  fft_solve(c, 0, cn);
  fft_inverse(c, c->v2, nullptr, 1.0f, 0.0f);  // v~^{n+1}
  if (mass) {
    // excess mass: u' and dv, v^{n+1} = v~ + dv, and the next step's lagged term
    // J(q^{n+1/2}) 1/2 dt nu L dv (interpolateLaplacianDeltaV)
    qnb_mass_correction(c, true, true);
    qnb_mass_fluid(c, 0.0f, 1.0f, 1.0f);
    PVecArgs a = pvec_base(c);
    a.f1 = c->M1;
    a.c1 = 1.0f;
    a.out = c->dulag;
    pvec(c, a);
  }
  if (c->np > 0) {
    // q^{n+1} = q^n + dt/2 J(q^{n+1/2}) (v^n + v^{n+1}); u = J(q^{n+1/2}) v~ (mass 0)
    InterpArgs ia = interp_base(c, 1);
    ia.base_cell = c->cell[0];
    ia.base_off = c->off[0];
    ia.out_cell = c->cell[0];
    ia.out_off = c->off[0];
    skin_to_full(c, ia);
    ia.f1 = c->v;
    ia.a1 = (float)(0.5 * dt);
    ia.f2 = c->v2;
    ia.a2 = (float)(0.5 * dt);
    ia.u_out = mass ? nullptr : c->u;
    launch_interp(ia, c->s);
    mark(c, "interp");
  }
  qnb_after_step(c);
}

void scheme_step(fhd_ctx* c) {
  switch (c->p.scheme) {
    case FHD_SCHEME_INCOMPRESSIBLE: step_incompressible(c); break;
    case FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT: step_qnb(c); break;
    case FHD_SCHEME_STOKES_LIMIT: step_stokes(c); break;
    default: throw std::runtime_error("unknown scheme");
  }
}

// Eager step: every kernel launched with the step counter as an argument.
void do_step(fhd_ctx* c) {
  scheme_step(c);
  c->step++;
  c->dev_step_valid = false;
  if (c->graphs_built) c->gphase = (c->gphase + 1) % c->gperiod;
}

// The step sequence repeats (with period 1, or 2 for QNB) once the start-up steps are done.
bool steady(const fhd_ctx* c) {
  switch (c->p.scheme) {
    case FHD_SCHEME_INCOMPRESSIBLE: return c->projected;
    case FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT: return c->projected && c->qnb_steady;
    default: return true;
  }
}

// Capture one graph per phase of the steady-state step.  Capturing runs the host logic of the
// step (which advances host-side state exactly as an eager step would) and records the launches;
// the step counter comes from device memory, advanced by the last node of each graph.
void build_graphs(fhd_ctx* c) {
  if (!c->step_dev) c->step_dev = dalloc<unsigned long long>(1);
  c->gperiod = c->p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT ? 2 : 1;
  for (int ph = 0; ph < c->gperiod; ++ph) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    try {
      scheme_step(c);
      launch_advance_step(c->step_dev, c->s);
    } catch (...) {
      c->capturing = false;
      (void)hipStreamEndCapture(c->s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    c->capturing = false;
    HIPCHK(hipStreamEndCapture(c->s, &g));
    HIPCHK(hipGraphInstantiate(&c->gexec[ph], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
  }
  // after gperiod captured steps the host state is back where it started
  c->gphase = 0;
  c->graphs_built = true;
}

void graph_step(fhd_ctx* c) {
  if (!c->dev_step_valid) {
    HIPCHK(hipStreamSynchronize(c->s));
    const unsigned long long v = (unsigned long long)c->step;
    HIPCHK(hipMemcpy(c->step_dev, &v, sizeof(v), hipMemcpyHostToDevice));
    c->dev_step_valid = true;
  }
  HIPCHK(hipGraphLaunch(c->gexec[c->gphase], c->s));
  if (c->p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) qnb_after_step(c);
  c->gphase = (c->gphase + 1) % c->gperiod;
  c->step++;
}

// The context's device error word (fluam's errorKernel, checked by its gpuToHost* calls,
// src/allocateErrorArray.cu:22-32, src/gpuToHostStokesLimit.cu:23-30).  The flags are sticky: a
// context that has flagged an error keeps reporting it.
void check_device_flags(fhd_ctx* c) {
  int e[8];
  HIPCHK(hipMemcpy(e, c->errd, sizeof(e), hipMemcpyDeviceToHost));
  if (e[0] & kErrNonFinite)
    throw std::runtime_error("device error: non-finite particle displacement (a force or velocity overflowed)");
  if (e[0] & kErrClamped)
    throw std::runtime_error("device error: a particle displacement of 2^24 cells or more was not applied "
                             "(blown-up step)");
  if (e[0] & kErrBarrier)
    throw std::runtime_error("device error: the one-launch particle sort's workgroups did not all meet "
                             "(its bounded wait ran out)");
}

void advance(fhd_ctx* c) {
  if (c->graph_mode && c->profiling == false && steady(c)) {
    if (!c->graphs_built) build_graphs(c);
    graph_step(c);
  } else {
    do_step(c);
  }
}

}  // namespace

// ------------------------------------------------------------------------------------------
// data.main front end (grammar of src/loadDataMain.cu:173-683).
// ------------------------------------------------------------------------------------------
extern "C" unsigned long fhd_params_size(void) { return (unsigned long)sizeof(fhd_params); }

extern "C" void fhd_default_params(fhd_params* p) {
  memset(p, 0, sizeof(*p));
  p->scheme = -1;
  p->densfluid = 1.0;
  p->shearviscosity = 1.0;
  p->temperature = 0.0;
  p->dt = 0.0;
  p->seed = 0;
  p->kernel_points = 3;
  p->nonbonded = 1;
  p->cutoff = 0.0;
  p->wca_sigma = -1.0;   // -1: 2 dx (src/initForcesNonBonded.cu:24)
  p->wca_epsilon = -1.0; // -1: kT  (src/initForcesNonBonded.cu:25)
  p->mass = 0.0;
  p->volume_particle = 8.0;  // src/createParticles.cpp:141-143
  p->rfd_delta = 1e-3;
  p->rfd_mode = FHD_RFD_DISPLACED;
  p->max_part_in_cell = 6;
  p->max_part_in_cell_nonbonded = 6;
  p->initial_positions = 1;  // simple cubic unless coordinates are given
  strcpy(p->outputname, "run");
}

namespace {

// The particle count a coordinates file declares (its first number): with loadparticles 1 it
// replaces numberparticles, as src/createParticles.cpp:88 reads np from the file.
bool coordinates_count(const char* path, int* np, std::string& err) {
  std::ifstream f(path);
  long long n = -1;
  if (!f || !(f >> n) || n < 0 || n > (1LL << 30)) {
    err = std::string("cannot read the particle count of coordinates file ") + path;
    return false;
  }
  *np = (int)n;
  return true;
}

bool parse_datamain(const char* path, fhd_params* out, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = std::string("cannot open ") + path;
    return false;
  }
  fhd_params p;
  fhd_default_params(&p);
  bool incompressible = false, qnb = false, stokes = false, setparticles = false, seed_set = false;
  bool vol_set = false;
  double vol = 8.0;
  std::string word, last;
  auto need = [&](bool ok) {
    if (!ok) throw std::runtime_error("ERROR INPUT FILE: bad value after " + word);
  };
  const char* unsupported[] = {"binaryMixture", "binaryMixtureWall", "giantFluctuations", "continuousGradient",
                               "thermostat", "incompressibleBoundaryRK2", "quasiNeutrallyBuoyant2D",
                               "quasiNeutrallyBuoyant4pt2D", "quasi2D", "stokesLimit2D", "particlesWall",
                               "freeEnergyCompressibleParticles", "semiImplicitCompressibleParticles",
                               "momentumCoupling", "stokesLimitFirstOrder", "incompressibleBinaryMixture",
                               "incompressibleBinaryMixtureMidPoint", "ghost", "boundary"};
  try {
    while (f >> word) {
      if (word[0] == '#') {
        std::string rest;
        std::getline(f, rest);
        continue;
      }
      bool handled = true;
      if (word == "incompressible") incompressible = true;
      else if (word == "quasiNeutrallyBuoyant") qnb = true;
      else if (word == "stokesLimit") { stokes = true; setparticles = true; }
      else if (word == "particles") { int v; need(bool(f >> v)); setparticles = v != 0; }
      else if (word == "numberparticles") need(bool(f >> p.np));
      else if (word == "mass") need(bool(f >> p.mass));
      else if (word == "volumeParticle") { need(bool(f >> vol)); vol_set = true; }
      else if (word == "cutoff") need(bool(f >> p.cutoff));
      else if (word == "maxNumberPartInCell") need(bool(f >> p.max_part_in_cell));
      else if (word == "maxNumberPartInCellNonBonded") need(bool(f >> p.max_part_in_cell_nonbonded));
      else if (word == "densfluid") need(bool(f >> p.densfluid));
      else if (word == "shearviscosity") need(bool(f >> p.shearviscosity));
      else if (word == "bulkviscosity") { double v; need(bool(f >> v)); }
      else if (word == "temperature") need(bool(f >> p.temperature));
      else if (word == "cells") need(bool(f >> p.mx >> p.my >> p.mz));
      else if (word == "celldimension") need(bool(f >> p.lx >> p.ly >> p.lz));
      else if (word == "initfluid") need(bool(f >> p.initfluid));
      else if (word == "backgroundvelocity") need(bool(f >> p.background[0] >> p.background[1] >> p.background[2]));
      else if (word == "seed") { need(bool(f >> p.seed)); seed_set = true; }
      else if (word == "numstepsRelaxation") need(bool(f >> p.numsteps_relaxation));
      else if (word == "numsteps") need(bool(f >> p.numsteps));
      else if (word == "samplefreq") need(bool(f >> p.samplefreq));
      else if (word == "savefreq") { long long v; need(bool(f >> v)); }
      else if (word == "dt") need(bool(f >> p.dt));
      else if (word == "saveVTK" || word == "saveFluid" || word == "setDevice") { int v; need(bool(f >> v)); }
      else if (word == "outputname") { std::string s; need(bool(f >> s)); snprintf(p.outputname, sizeof(p.outputname), "%s", s.c_str()); }
      else if (word == "loadparticles") { int v; need(bool(f >> v)); if (v) p.initial_positions = 2; }
      else if (word == "coordinates") { std::string s; need(bool(f >> s)); snprintf(p.coordinates, sizeof(p.coordinates), "%s", s.c_str()); }
      else if (word == "fluid") { std::string s; need(bool(f >> s)); snprintf(p.fluid, sizeof(p.fluid), "%s", s.c_str()); }
      else if (word == "initialStep") need(bool(f >> p.initial_step));
      else if (word == "velocities") { std::string s; need(bool(f >> s)); snprintf(p.velocities, sizeof(p.velocities), "%s", s.c_str()); }
      else if (word == "computeNonBondedForces") need(bool(f >> p.nonbonded));
      else if (word == "extraMobility") need(bool(f >> p.extra_mobility));
      else if (word == "bondedForces") { std::string s; need(bool(f >> s)); snprintf(p.bonded_forces, sizeof(p.bonded_forces), "%s", s.c_str()); }
      else if (word == "bondedForcesVersion") {
        int v;
        need(bool(f >> v));
        if (v) throw std::runtime_error("bondedForcesVersion 1 (the older bond-file layout) is not supported");
      }
      // extension keywords
      else if (word == "kernel") need(bool(f >> p.kernel_points));
      else if (word == "sigma") need(bool(f >> p.wca_sigma));
      else if (word == "epsilon") need(bool(f >> p.wca_epsilon));
      else if (word == "rfdMode") need(bool(f >> p.rfd_mode));
      else if (word == "rfdDelta") need(bool(f >> p.rfd_delta));
      else if (word == "initialPositions") need(bool(f >> p.initial_positions));
      else if (word == "structureFactor") need(bool(f >> p.structure_factor));
      else handled = false;
      if (!handled) {
        for (const char* u : unsupported)
          if (word == u) throw std::runtime_error("scheme or option '" + word + "' is outside this library (fluam's " +
                                                  "incompressible, quasiNeutrallyBuoyant and stokesLimit paths only)");
        throw std::runtime_error("ERROR INPUT FILE; LAST WORD READ: " + last + "; UNKNOWN: " + word);
      }
      last = word;
    }
  } catch (const std::exception& e) {
    err = e.what();
    return false;
  }
  // scheme precedence of src/main.cpp:70-99 restricted to the three schemes
  if (setparticles && incompressible) {
    err = "particles with incompressible select fluam's incompressibleBoundary scheme, which is outside this library";
    return false;
  }
  if (incompressible) p.scheme = FHD_SCHEME_INCOMPRESSIBLE;
  else if (qnb) p.scheme = FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT;
  else if (stokes) p.scheme = FHD_SCHEME_STOKES_LIMIT;
  else {
    err = "no scheme keyword (incompressible, quasiNeutrallyBuoyant or stokesLimit)";
    return false;
  }
  if (!setparticles && p.scheme != FHD_SCHEME_STOKES_LIMIT) p.np = 0;
  if (!seed_set) p.seed = (unsigned long long)(time(nullptr) % 18671);  // src/initializeRandomNumber.cpp:24-28
  double dx = p.lx / (p.mx > 0 ? p.mx : 1);
  double cvol = (p.lx / std::max(p.mx, 1)) * (p.ly / std::max(p.my, 1)) * (p.lz / std::max(p.mz, 1));
  p.volume_particle = vol_set ? vol / cvol : 8.0;  // src/createParticles.cpp:127-143
  if (p.wca_sigma < 0) p.wca_sigma = 2.0 * dx;
  if (p.wca_epsilon < 0) p.wca_epsilon = p.temperature;
  if (p.initial_positions == 2 && p.coordinates[0] == 0) {
    err = "loadparticles 1 needs a coordinates file";
    return false;
  }
  if (p.initial_positions == 2 && !coordinates_count(p.coordinates, &p.np, err)) return false;
  *out = p;
  return true;
}

// coordinates file: first number np, then np lines "x y z" (src/createParticles.cpp:81-100).
// .fluidFinalConfiguration reader: five header lines (cell count, "mx my mz", their values,
// "lx ly lz", their values), "Density d", "Pressure parameters a0 a1 a2", five more lines,
// then per cell its density and vx vy vz (src/initializeFluid.cpp:168-192).
bool read_fluid_file(const char* path, int mx, int my, int mz, std::vector<double>& v, double& dens,
                     std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = std::string("cannot open fluid file ") + path;
    return false;
  }
  std::string line, w;
  std::getline(f, line);
  std::getline(f, line);
  int fm[3] = {0, 0, 0};
  f >> fm[0] >> fm[1] >> fm[2];
  if (!f || fm[0] != mx || fm[1] != my || fm[2] != mz) {
    err = std::string("fluid file ") + path + ": grid differs from cells";
    return false;
  }
  std::getline(f, line);
  std::getline(f, line);
  std::getline(f, line);
  double a0, a1, a2;
  f >> w >> dens >> w >> w >> a0 >> a1 >> a2;
  for (int k = 0; k < 6; ++k) std::getline(f, line);
  const long long N = (long long)mx * my * mz;
  v.assign(3 * (size_t)N, 0.0);
  for (long long i = 0; i < N; ++i) {
    double d;
    if (!(f >> d >> v[i] >> v[N + i] >> v[2 * N + i])) {
      err = std::string("fluid file ") + path + " truncated";
      return false;
    }
  }
  return true;
}

bool read_coordinates(const fhd_params& p, std::vector<double>& q, std::string& err) {
  std::ifstream f(p.coordinates);
  if (!f) {
    err = std::string("cannot open coordinates file ") + p.coordinates;
    return false;
  }
  int n = 0;
  f >> n;
  if (n != p.np) {
    err = "coordinates file particle count changed while it was read";
    return false;
  }
  q.assign(3 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i)
    if (!(f >> q[i] >> q[n + i] >> q[2 * (size_t)n + i])) {
      err = "coordinates file truncated";
      return false;
    }
  return true;
}

// velocities file: "np" then np lines "vx vy vz"; its count must equal the coordinates file's
// (src/createParticles.cpp:103-120).  Returned component-major, like positions.
bool read_velocities(const fhd_params& p, std::vector<double>& u, std::string& err) {
  std::ifstream f(p.velocities);
  if (!f) {
    err = std::string("cannot open velocities file ") + p.velocities;
    return false;
  }
  int n = -1;
  if (!(f >> n) || n != p.np) {
    err = "ERROR NP DIFFERENT IN PARTICLESCOOR AND PARTICLESVEL (velocities file count differs from the coordinates)";
    return false;
  }
  u.assign(3 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i)
    if (!(f >> u[i] >> u[n + i] >> u[2 * (size_t)n + i])) {
      err = "velocities file truncated";
      return false;
    }
  return true;
}

// Bonds file (bin/inputfiles/README:55-108): the number of particle-particle bonds, one line
// "i j k r0" each, then the number of particle-fixed-point bonds, one line "i k r0 x y z" each.
// Each particle-particle bond is listed under both particles, in file order, as
// src/initializeBondedForces.cpp:110-160 builds its per-particle rows.
struct HostBonds {
  std::vector<int> pp_start, fp_start;
  std::vector<int4> pp, fpc;
  std::vector<float4> fpo;
  bool any() const { return !pp.empty() || !fpc.empty(); }
};
bool read_bonds(const fhd_params& p, double h[3], HostBonds& b, std::string& err) {
  std::ifstream f(p.bonded_forces);
  if (!f) {
    err = std::string("cannot open bonds file ") + p.bonded_forces;
    return false;
  }
  const int np = p.np;
  auto bad = [&](const char* what) {
    err = std::string("bonds file ") + p.bonded_forces + ": " + what;
    return false;
  };
  long long npp = -1, nfp = -1;
  if (!(f >> npp) || npp < 0 || npp > (1LL << 28)) return bad("bad particle-particle bond count");
  std::vector<int> pi(npp), pj(npp);
  std::vector<double> pk(npp), pr(npp);
  for (long long e = 0; e < npp; ++e) {
    if (!(f >> pi[e] >> pj[e] >> pk[e] >> pr[e])) return bad("particle-particle bonds truncated");
    if (pi[e] < 0 || pi[e] >= np || pj[e] < 0 || pj[e] >= np) return bad("bond index outside 0..numberparticles-1");
  }
  if (!(f >> nfp) || nfp < 0 || nfp > (1LL << 28)) return bad("bad fixed-point bond count");
  std::vector<int> fi(nfp);
  std::vector<double> fk(nfp), fr(nfp), fx(3 * nfp);
  for (long long e = 0; e < nfp; ++e) {
    if (!(f >> fi[e] >> fk[e] >> fr[e] >> fx[3 * e] >> fx[3 * e + 1] >> fx[3 * e + 2]))
      return bad("fixed-point bonds truncated");
    if (fi[e] < 0 || fi[e] >= np) return bad("bond index outside 0..numberparticles-1");
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(fx[3 * e + a])) return bad("non-finite fixed point");
  }
  b.pp_start.assign(np + 1, 0);
  b.fp_start.assign(np + 1, 0);
  for (long long e = 0; e < npp; ++e) { b.pp_start[pi[e] + 1]++; b.pp_start[pj[e] + 1]++; }
  for (long long e = 0; e < nfp; ++e) b.fp_start[fi[e] + 1]++;
  for (int i = 0; i < np; ++i) { b.pp_start[i + 1] += b.pp_start[i]; b.fp_start[i + 1] += b.fp_start[i]; }
  b.pp.assign(2 * npp, make_int4(0, 0, 0, 0));
  b.fpc.assign(nfp, make_int4(0, 0, 0, 0));
  b.fpo.assign(nfp, make_float4(0, 0, 0, 0));
  std::vector<int> fill(b.pp_start.begin(), b.pp_start.end() - 1);
  auto fbits = [](double v) { float x = (float)v; int r; memcpy(&r, &x, 4); return r; };
  for (long long e = 0; e < npp; ++e) {
    b.pp[fill[pi[e]]++] = make_int4(pj[e], fbits(pk[e]), fbits(pr[e]), 0);
    b.pp[fill[pj[e]]++] = make_int4(pi[e], fbits(pk[e]), fbits(pr[e]), 0);
  }
  fill.assign(b.fp_start.begin(), b.fp_start.end() - 1);
  const double L[3] = {p.lx, p.ly, p.lz};
  for (long long e = 0; e < nfp; ++e) {
    int cc[3];
    float d[3];
    for (int a = 0; a < 3; ++a) to_cell(fx[3 * e + a], L[a], h[a], cc[a], d[a]);
    const int slot = fill[fi[e]]++;
    b.fpc[slot] = make_int4(cc[0], cc[1], cc[2], fbits(fk[e]));
    b.fpo[slot] = make_float4(d[0], d[1], d[2], (float)fr[e]);
  }
  return true;
}

}  // namespace

// Philox-based uniform draw shared with the FP64 restatement (seeded uniform positions).
namespace {
void philox_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1,
             n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Four standard normals of one host draw (Box-Muller on 24-bit uniforms, FP64), the same
// construction as the device generator's.
void host_normals4(uint64_t seed, uint32_t index, uint32_t stream, double out[4]) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const uint32_t ctr[4] = {index, stream << 4, 0u, 0u};
  uint32_t w[4];
  philox_host(ctr, key, w);
  for (int k = 0; k < 2; ++k) {
    const double u1 = (double)((w[2 * k] >> 8) + 1u) * (1.0 / 16777216.0);
    const double u2 = (double)(w[2 * k + 1] >> 8) * (1.0 / 16777216.0);
    const double r = sqrt(-2.0 * log(u1));
    out[2 * k] = r * cos(2.0 * M_PI * u2);
    out[2 * k + 1] = r * sin(2.0 * M_PI * u2);
  }
}

// Initial velocity for initfluid 0-4 (src/initializeFluid.cpp:31-160): 0 the background velocity;
// 1 equilibrium fluctuations sqrt(kT / (rho dV)) N(0,1) per face (cell i draws its three normals
// from stream FHD_STREAM_INIT_FLUID at index i in place of fluam's host gauss()), minus their
// mean, plus the background; 2 a transverse wave vy = vy0 + 0.04 sin(2 pi z / lz); 3 a
// longitudinal wave vz = 0.004 sin(16 pi (z + dz/2) / lz); 4 vx = 0.425.  z is the cell centre,
// (iz + 1/2) dz - lz/2 (src/createCells.cpp:66-69).
void initial_fluid(const fhd_params& p, std::vector<double>& v) {
  const int mx = p.mx, my = p.my, mz = p.mz;
  const long long N = (long long)mx * my * mz;
  const double dx = p.lx / mx, dy = p.ly / my, dz = p.lz / mz, dV = dx * dy * dz;
  v.assign(3 * (size_t)N, 0.0);
  const double* b = p.background;
  double mean[3] = {0.0, 0.0, 0.0};
  for (long long i = 0; i < N; ++i) {
    const double z = (double)(i / ((long long)mx * my) + 0.5) * dz - 0.5 * p.lz;
    double c[3] = {b[0], b[1], b[2]};
    switch (p.initfluid) {
      case 1: {
        double n[4];
        host_normals4(p.seed, (uint32_t)i, FHD_STREAM_INIT_FLUID, n);
        const double s = sqrt(p.temperature / (p.densfluid * dV));
        for (int a = 0; a < 3; ++a) {
          c[a] = s * n[a];
          mean[a] += c[a];
        }
        break;
      }
      case 2: c[1] = b[1] + 0.04 * sin(2.0 * M_PI / p.lz * z); break;
      case 3: c[0] = 0.0; c[1] = 0.0; c[2] = 0.004 * sin(2.0 * M_PI / p.lz * 8.0 * (z + 0.5 * p.lz / mz)); break;
      case 4: c[0] = 0.425; c[1] = 0.0; c[2] = 0.0; break;
      default: break;
    }
    for (int a = 0; a < 3; ++a) v[(size_t)a * N + i] = c[a];
  }
  if (p.initfluid == 1)
    for (int a = 0; a < 3; ++a) {
      const double m = mean[a] / (double)N;
      for (long long i = 0; i < N; ++i) v[(size_t)a * N + i] = v[(size_t)a * N + i] - m + b[a];
    }
}

void initial_positions(const fhd_params& p, std::vector<double>& q) {
  const int np = p.np;
  q.assign(3 * (size_t)np, 0.0);
  if (p.initial_positions == 1) {
    // simple cubic lattice, src/simpleCubic.cpp:25-63
    double sigma = pow(p.lx * p.ly * p.lz / np, 1.0 / 3.0);
    int nx = (int)(p.lx / sigma) ? (int)(p.lx / sigma) : 1;
    int ny = (int)(p.ly / sigma) ? (int)(p.ly / sigma) : 1;
    int nz = (int)(p.lz / sigma) ? (int)(p.lz / sigma) : 1;
    while (nx * ny * nz < np) {
      if (nx * ny * nz < np) nx++;
      if (nx * ny * nz < np) ny++;
      if (nx * ny * nz < np) nz++;
    }
    double ddx = p.lx / nx, ddy = p.ly / ny, ddz = p.lz / nz;
    int n = 0;
    for (int i = 0; i < nz && n < np; ++i)
      for (int j = 0; j < ny && n < np; ++j)
        for (int k = 0; k < nx && n < np; ++k) {
          q[n] = (k + 0.5) * ddx - p.lx / 2;
          q[np + n] = (j + 0.5) * ddy - p.ly / 2;
          q[2 * (size_t)np + n] = (i + 0.5) * ddz - p.lz / 2;
          ++n;
        }
    return;
  }
  // seeded uniform positions: Philox draw (seed, attempt k, stream 8, index i), 24-bit uniforms.
  // Mode 0 takes attempt 0 of every particle.  Mode 3 (random sequential addition) redraws a
  // particle (attempt 1, 2, ...) while it lies closer than the WCA sigma to an earlier one, by
  // the minimum-image distance, so no pair starts inside the repulsive core.
  const uint32_t key[2] = {(uint32_t)p.seed, (uint32_t)(p.seed >> 32)};
  const double L[3] = {p.lx, p.ly, p.lz};
  auto draw = [&](int i, uint32_t attempt, double x[3]) {
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)FHD_STREAM_INIT_POS << 4, attempt, 0u}, w[4];
    philox_host(ctr, key, w);
    for (int a = 0; a < 3; ++a) x[a] = ((double)(w[a] >> 8) / 16777216.0 - 0.5) * L[a];
  };
  const bool rsa = p.initial_positions == 3 && p.nonbonded && p.cutoff > 0;
  const double dmin = p.wca_sigma > 0 ? p.wca_sigma : 2.0 * p.lx / p.mx;
  int m[3] = {1, 1, 1};
  for (int a = 0; a < 3; ++a) m[a] = rsa ? std::max(1, (int)(L[a] / dmin)) : 1;
  std::vector<std::vector<int>> cells(rsa ? (size_t)m[0] * m[1] * m[2] : 0);
  auto cell_of = [&](const double x[3], int j[3]) {
    for (int a = 0; a < 3; ++a) j[a] = std::min(m[a] - 1, std::max(0, (int)((x[a] / L[a] + 0.5) * m[a])));
  };
  constexpr int kMaxAttempts = 100000;
  for (int i = 0; i < np; ++i) {
    double x[3];
    int attempt = 0;
    for (;; ++attempt) {
      if (attempt >= kMaxAttempts)
        throw std::runtime_error("initialPositions 3: no overlap-free place for particle " + std::to_string(i) +
                                 " (too dense for random sequential addition; use the lattice, initialPositions 1)");
      draw(i, (uint32_t)attempt, x);
      if (!rsa) break;
      int j[3];
      cell_of(x, j);
      bool ok = true;
      const int lo[3] = {m[0] >= 3 ? -1 : 0, m[1] >= 3 ? -1 : 0, m[2] >= 3 ? -1 : 0};
      const int hi[3] = {m[0] >= 2 ? 1 : 0, m[1] >= 2 ? 1 : 0, m[2] >= 2 ? 1 : 0};
      for (int dz = lo[2]; dz <= hi[2] && ok; ++dz)
        for (int dy = lo[1]; dy <= hi[1] && ok; ++dy)
          for (int dx = lo[0]; dx <= hi[0] && ok; ++dx) {
            const int cx = (j[0] + dx + m[0]) % m[0], cy = (j[1] + dy + m[1]) % m[1], cz = (j[2] + dz + m[2]) % m[2];
            for (int k : cells[cx + (size_t)m[0] * (cy + (size_t)m[1] * cz)]) {
              double r2 = 0.0;
              for (int a = 0; a < 3; ++a) {
                double d = x[a] - q[(size_t)a * np + k];
                d -= L[a] * std::floor(d / L[a] + 0.5);
                r2 += d * d;
              }
              if (r2 < dmin * dmin) {
                ok = false;
                break;
              }
            }
          }
      if (ok) break;
    }
    for (int a = 0; a < 3; ++a) q[(size_t)a * np + i] = x[a];
    if (rsa) {
      int j[3];
      cell_of(x, j);
      cells[j[0] + (size_t)m[0] * (j[1] + (size_t)m[1] * j[2])].push_back(i);
    }
  }
}
}  // namespace

// ------------------------------------------------------------------------------------------
// C ABI.
// ------------------------------------------------------------------------------------------
extern "C" {

const char* fhd_last_error(void) { return g_last_error.c_str(); }

int fhd_read_fluid_file(const char* path, int mx, int my, int mz, double* v, double* densfluid) {
  std::vector<double> vf;
  std::string err;
  double dens = 0.0;
  if (!path || !v || !read_fluid_file(path, mx, my, mz, vf, dens, err)) {
    g_last_error = err.empty() ? "bad arguments" : err;
    return 0;
  }
  std::copy(vf.begin(), vf.end(), v);
  if (densfluid) *densfluid = dens;
  return 1;
}

int fhd_load_datamain(const char* path) {
  fhd_params p;
  std::string err;
  if (!parse_datamain(path, &p, err)) {
    g_last_error = err;
    return 0;
  }
  g_params = p;
  g_params_loaded = true;
  return 1;
}

int fhd_get_params(fhd_params* out) {
  if (!g_params_loaded) {
    g_last_error = "no data.main loaded";
    return 0;
  }
  *out = g_params;
  return 1;
}

int fhd_set_params(const fhd_params* in) {
  g_params = *in;
  g_params_loaded = true;
  return 1;
}

fhd_ctx* fhd_create(const fhd_params* pin) {
  fhd_ctx* c = new fhd_ctx();
  try {
    c->p = *pin;
    if (c->p.initial_positions == 2) {
      // loadparticles 1: the coordinates file's count is the particle count
      std::string err;
      if (!coordinates_count(c->p.coordinates, &c->p.np, err)) throw std::runtime_error(err);
    }
    if (c->p.initfluid < 0 || c->p.initfluid > 4)
      throw std::runtime_error("initfluid must be 0-4 (the options of src/initializeFluid.cpp for this path)");
    if (c->p.wca_sigma < 0) c->p.wca_sigma = 2.0 * c->p.lx / c->p.mx;
    if (c->p.wca_epsilon < 0) c->p.wca_epsilon = c->p.temperature;
    validate(c->p);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("no HIP device");
    HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev_noise, hipEventDisableTiming));
    const fhd_params& p = c->p;
    c->dx = p.lx / p.mx;
    c->dy = p.ly / p.my;
    c->dz = p.lz / p.mz;
    c->dV = c->dx * c->dy * c->dz;
    GridDev& g = c->g;
    g.mx = p.mx; g.my = p.my; g.mz = p.mz;
    g.N = p.mx * p.my * p.mz;
    g.P = p.mx / 2 + 1;
    g.dx = (float)c->dx; g.dy = (float)c->dy; g.dz = (float)c->dz;
    g.idx = (float)(1.0 / c->dx); g.idy = (float)(1.0 / c->dy); g.idz = (float)(1.0 / c->dz);
    g.lx = (float)p.lx; g.ly = (float)p.ly; g.lz = (float)p.lz;
    const size_t N3 = 3 * (size_t)g.N;
    c->v = dalloc<float>(N3);
    c->v2 = dalloc<float>(N3);
    c->W = dalloc<float>(N3);
    c->SF = dalloc<float>(N3);
    c->div = dalloc<float>(g.N);
    c->spec = dalloc<float2>(3 * (size_t)p.mz * p.my * g.P);
    if (p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) {
      c->adv[0] = dalloc<float>(N3);
      c->adv[1] = dalloc<float>(N3);
    }
    c->errd = dalloc<int>(8);
    HIPCHK(hipMemset(c->errd, 0, 8 * sizeof(int)));
    build_tables(c);
    // initial velocity (initfluid, src/initializeFluid.cpp:31-160); the Stokes limit has no
    // velocity state
    std::vector<float> v0(N3, 0.0f);
    if (p.scheme != FHD_SCHEME_STOKES_LIMIT) {
      std::vector<double> vi;
      initial_fluid(p, vi);
      for (size_t i = 0; i < N3; ++i) v0[i] = (float)vi[i];
    }
    if (p.fluid[0]) {
      // restart: velocity and density from a saved configuration (fluid keyword)
      std::vector<double> vf;
      std::string err;
      double dens = p.densfluid;
      if (!read_fluid_file(p.fluid, p.mx, p.my, p.mz, vf, dens, err)) throw std::runtime_error(err);
      for (size_t i = 0; i < N3; ++i) v0[i] = (float)vf[i];
      c->p.densfluid = dens;
    }
    c->step = p.initial_step;
    HIPCHK(hipMemcpy(c->v, v0.data(), N3 * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->v2, 0, N3 * sizeof(float)));
    HIPCHK(hipMemset(c->SF, 0, N3 * sizeof(float)));
    apply_schedule(c);
    c->np = p.np;
    if (c->np > 0) {
      Tiles& t = c->t;
      t.tx = tile_edge(p.mx); t.ty = tile_edge(p.my); t.tz = tile_edge(p.mz);
      t.ntx = p.mx / t.tx; t.nty = p.my / t.ty; t.ntz = p.mz / t.tz;
      t.n = t.ntx * t.nty * t.ntz;
      for (int k = 0; k < 2; ++k) {
        c->cell[k] = dalloc<int4>(c->np);
        c->off[k] = dalloc<float4>(c->np);
        c->F[k] = dalloc<float4>(c->np);
        c->start[k] = dalloc<int>(t.n + 1);
        c->sid[k] = dalloc<int>(c->np);
        c->scell[k] = dalloc<int4>(c->np);
        c->soff[k] = dalloc<float4>(c->np);
        c->rec[k] = dalloc<SrcRec>(2 * (size_t)c->np);
        c->cull[k] = dalloc<int>(2 * (size_t)c->np);
        HIPCHK(hipMemset(c->F[k], 0, sizeof(float4) * c->np));
        // the sorted sets start empty and in range (every tile empty, every slot particle 0),
        // so a sort that never completes (its grid barrier gave up and raised the device error)
        // leaves the later passes of that step reading valid indices
        HIPCHK(hipMemset(c->start[k], 0, sizeof(int) * (t.n + 1)));
        HIPCHK(hipMemset(c->sid[k], 0, sizeof(int) * c->np));
        HIPCHK(hipMemset(c->scell[k], 0, sizeof(int4) * c->np));
        HIPCHK(hipMemset(c->soff[k], 0, sizeof(float4) * c->np));
      }
      c->u = dalloc<float4>(c->np);
      HIPCHK(hipMemset(c->u, 0, sizeof(float4) * c->np));
      if (p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT && p.mass != 0.0) {
        c->M1 = dalloc<float>(N3);
        float4** pv[5] = {&c->dulag, &c->du, &c->dp, &c->jvt, &c->ms};
        for (float4** q : pv) {
          *q = dalloc<float4>(c->np);
          HIPCHK(hipMemset(*q, 0, sizeof(float4) * c->np));
        }
      }
      c->count = dalloc<int>(2 * (size_t)t.n);
      HIPCHK(hipMemset(c->count, 0, 2 * sizeof(int) * t.n));
      c->slot_of = dalloc<int>(c->np);
      c->sflag = dalloc<int>(2);
      const int flags0[2] = {1, 0};
      HIPCHK(hipMemcpy(c->sflag, flags0, sizeof(flags0), hipMemcpyHostToDevice));
      c->sbar = dalloc<unsigned long long>(2);  // [0] the sort's, [1] the WCA pass's
      HIPCHK(hipMemset(c->sbar, 0, 2 * sizeof(unsigned long long)));
      // skin of the WCA neighbour lists: a quarter of the cutoff, so that a list built at
      // radius rc + skin stays complete for hundreds of steps at the particles' ~1e-3 cells
      // per step; the list radius stays within a tile (axes with three or more tiles) and
      // below half the box
      const int T[3] = {t.tx, t.ty, t.tz}, NT[3] = {t.ntx, t.nty, t.ntz};
      const double h[3] = {c->dx, c->dy, c->dz};
      const bool wca_on = p.nonbonded && p.cutoff > 0;
      if (wca_on) {
        double sk = 0.25 * p.cutoff;
        const int m[3] = {p.mx, p.my, p.mz};
        for (int a = 0; a < 3; ++a) {
          if (NT[a] >= 3) sk = std::min(sk, T[a] * h[a] - p.cutoff);
          sk = std::min(sk, 0.49 * m[a] * h[a] - p.cutoff);
        }
        c->nb_skin = (float)std::max(sk, 0.0);
        c->nb_idx = dalloc<int>((size_t)kNbCap * c->np);
        c->nb_cnt = dalloc<int>(c->np);
        c->nb_ref_cell = dalloc<int4>(c->np);
        c->nb_ref_off = dalloc<float4>(c->np);
        c->nb_flag = dalloc<int>(1);
        HIPCHK(hipMemset(c->nb_idx, 0, sizeof(int) * kNbCap * c->np));  // every entry a valid id
        HIPCHK(hipMemset(c->nb_cnt, 0, sizeof(int) * c->np));
        HIPCHK(hipMemset(c->nb_ref_cell, 0, sizeof(int4) * c->np));
        HIPCHK(hipMemset(c->nb_ref_off, 0, sizeof(float4) * c->np));
        const int one = 1;
        HIPCHK(hipMemcpy(c->nb_flag, &one, sizeof(int), hipMemcpyHostToDevice));
      }
      // skin of the kept sort, per axis with three or more tiles: half a cell (the list build
      // culls from the tile grown by it, so it is kept small; particles move ~1e-3 cells per
      // step, so a sort is rarely needed).  The spreading pass finds a source in its tile's
      // neighbours while its stencil (two cells each way, one more for an RFD displacement)
      // stays within them, s <= T - 3; the list build finds every pair within its radius while
      // no tile beyond the neighbours can hold a partner, 2 s h + rc + list skin <= T h.
      {
        c->skin_len = 0.0f;
        for (int a = 0; a < 3; ++a) {
          if (NT[a] <= 2) { c->skin[a] = -1.0f; continue; }
          double sk = std::min(0.5, (double)(T[a] - 3));
          if (wca_on) sk = std::min(sk, (T[a] * h[a] - p.cutoff - c->nb_skin) / (2.0 * h[a]));
          c->skin[a] = (float)std::max(sk, 0.0);
          c->skin_len = std::max(c->skin_len, (float)(c->skin[a] * h[a]));
        }
      }
      c->slot = dalloc<int>(c->np);
      c->tmp = dalloc<int>(c->np);
      std::vector<double> q;
      if (p.initial_positions == 2) {
        std::string err;
        if (!read_coordinates(p, q, err)) throw std::runtime_error(err);
      } else {
        initial_positions(p, q);
      }
      if (!fhd_set_particles(c, q.data())) throw std::runtime_error(c->err);
      if (p.initial_positions == 2 && p.velocities[0]) {
        // initial particle velocities (reported until the first step replaces them)
        std::vector<double> uv;
        std::string err;
        if (!read_velocities(p, uv, err)) throw std::runtime_error(err);
        std::vector<float4> hu(c->np);
        for (int i = 0; i < c->np; ++i)
          hu[i] = make_float4((float)uv[i], (float)uv[c->np + i], (float)uv[2 * (size_t)c->np + i], 0.0f);
        HIPCHK(hipMemcpy(c->u, hu.data(), c->np * sizeof(float4), hipMemcpyHostToDevice));
      }
      if (p.bonded_forces[0]) {
        HostBonds hb;
        std::string err;
        double h[3] = {c->dx, c->dy, c->dz};
        if (!read_bonds(p, h, hb, err)) throw std::runtime_error(err);
        if (hb.any()) {
          c->bpp_start = dalloc<int>(c->np + 1);
          c->bfp_start = dalloc<int>(c->np + 1);
          HIPCHK(hipMemcpy(c->bpp_start, hb.pp_start.data(), (c->np + 1) * sizeof(int), hipMemcpyHostToDevice));
          HIPCHK(hipMemcpy(c->bfp_start, hb.fp_start.data(), (c->np + 1) * sizeof(int), hipMemcpyHostToDevice));
          c->bpp = dalloc<int4>(std::max<size_t>(hb.pp.size(), 1));
          c->bfpc = dalloc<int4>(std::max<size_t>(hb.fpc.size(), 1));
          c->bfpo = dalloc<float4>(std::max<size_t>(hb.fpo.size(), 1));
          if (!hb.pp.empty())
            HIPCHK(hipMemcpy(c->bpp, hb.pp.data(), hb.pp.size() * sizeof(int4), hipMemcpyHostToDevice));
          if (!hb.fpc.empty()) {
            HIPCHK(hipMemcpy(c->bfpc, hb.fpc.data(), hb.fpc.size() * sizeof(int4), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(c->bfpo, hb.fpo.data(), hb.fpo.size() * sizeof(float4), hipMemcpyHostToDevice));
          }
        }
      }
    } else if (p.bonded_forces[0]) {
      throw std::runtime_error("bondedForces needs particles");
    }
    // the uploads above are synchronous; wait for the context's own stream (the schedule timing
    // runs there), not for other contexts' queued work
    HIPCHK(hipStreamSynchronize(c->s));
  } catch (const std::exception& e) {
    g_last_error = e.what();
    free_ctx(c);
    delete c;
    return nullptr;
  }
  return c;
}

void fhd_destroy(fhd_ctx* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->s);
  free_ctx(c);
  delete c;
}

const char* fhd_ctx_error(const fhd_ctx* c) { return c ? c->err.c_str() : g_last_error.c_str(); }

#define CTX_TRY(c, ...)                  \
  do {                                   \
    if (!(c)) {                          \
      g_last_error = "null context";     \
      return 0;                          \
    }                                    \
    try {                                \
      __VA_ARGS__;                       \
    } catch (const std::exception& e) {  \
      (c)->err = e.what();               \
      g_last_error = e.what();           \
      return 0;                          \
    }                                    \
  } while (0)

int fhd_set_velocity(fhd_ctx* c, const double* v) {
  CTX_TRY(c, {
    const size_t n = 3 * (size_t)c->g.N;
    std::vector<float> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = (float)v[i];
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(c->v, h.data(), n * sizeof(float), hipMemcpyHostToDevice));
    c->projected = false;
  });
  return 1;
}

int fhd_get_velocity(fhd_ctx* c, double* v) {
  CTX_TRY(c, {
    const size_t n = 3 * (size_t)c->g.N;
    std::vector<float> h(n);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(h.data(), c->v, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) v[i] = h[i];
  });
  return 1;
}

int fhd_set_particles(fhd_ctx* c, const double* q) {
  CTX_TRY(c, {
    const int np = c->np;
    std::vector<int4> hc(np);
    std::vector<float4> ho(np);
    const double L[3] = {c->p.lx, c->p.ly, c->p.lz}, h[3] = {c->dx, c->dy, c->dz};
    for (int i = 0; i < np; ++i) {
      int cc[3];
      float d[3];
      for (int a = 0; a < 3; ++a) {
        double qa = q[(size_t)a * np + i];
        if (!std::isfinite(qa)) throw std::runtime_error("non-finite particle position");
        to_cell(qa, L[a], h[a], cc[a], d[a]);
      }
      hc[i] = make_int4(cc[0], cc[1], cc[2], 0);
      ho[i] = make_float4(d[0], d[1], d[2], 0.0f);
    }
    HIPCHK(hipStreamSynchronize(c->s));
    if (np > 0) {
      HIPCHK(hipMemcpy(c->cell[0], hc.data(), np * sizeof(int4), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->off[0], ho.data(), np * sizeof(float4), hipMemcpyHostToDevice));
      const int one = 1;  // the sorted copies are stale: sort again before the next use
      HIPCHK(hipMemcpy(c->sflag, &one, sizeof(int), hipMemcpyHostToDevice));
      if (c->nb_flag) HIPCHK(hipMemcpy(c->nb_flag, &one, sizeof(int), hipMemcpyHostToDevice));  // and rebuild the lists
    }
  });
  return 1;
}

int fhd_get_particles(fhd_ctx* c, double* q) {
  CTX_TRY(c, {
    const int np = c->np;
    std::vector<int4> hc(np);
    std::vector<float4> ho(np);
    HIPCHK(hipStreamSynchronize(c->s));
    if (np > 0) {
      HIPCHK(hipMemcpy(hc.data(), c->cell[0], np * sizeof(int4), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(ho.data(), c->off[0], np * sizeof(float4), hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < np; ++i) {
      q[i] = from_cell(hc[i].x, ho[i].x, c->p.lx, c->dx);
      q[np + i] = from_cell(hc[i].y, ho[i].y, c->p.ly, c->dy);
      q[2 * (size_t)np + i] = from_cell(hc[i].z, ho[i].z, c->p.lz, c->dz);
    }
  });
  return 1;
}

static int get_f4(fhd_ctx* c, const float4* d, double* out) {
  CTX_TRY(c, {
    const int np = c->np;
    std::vector<float4> h(np);
    HIPCHK(hipStreamSynchronize(c->s));
    if (np > 0 && d) HIPCHK(hipMemcpy(h.data(), d, np * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < np; ++i) {
      out[i] = h[i].x;
      out[np + i] = h[i].y;
      out[2 * (size_t)np + i] = h[i].z;
    }
  });
  return 1;
}

int fhd_get_particle_velocities(fhd_ctx* c, double* u) { return get_f4(c, c ? c->u : nullptr, u); }
int fhd_get_forces(fhd_ctx* c, double* f) { return get_f4(c, c ? c->F[c->last_force_set] : nullptr, f); }

int fhd_step(fhd_ctx* c, int nsteps) {
  CTX_TRY(c, {
    for (int k = 0; k < nsteps; ++k) advance(c);
    HIPCHK(hipGetLastError());
  });
  return 1;
}

int fhd_set_graph(fhd_ctx* c, int on) {
  CTX_TRY(c, { c->graph_mode = on != 0; });
  return 1;
}

// Equal-time structure factor of the velocity (fluam's HydroGrid hook,
// src/hydroAnalysisIncompressible.cu:35-92, computed in place on the GPU).
int fhd_sk_reset(fhd_ctx* c) {
  CTX_TRY(c, {
    const size_t n = 3 * (size_t)c->p.mz * c->p.my * c->g.P;
    if (!c->sk_acc) c->sk_acc = dalloc<double>(n);
    HIPCHK(hipMemsetAsync(c->sk_acc, 0, n * sizeof(double), c->s));
    c->sk_samples = 0;
  });
  return 1;
}

int fhd_sk_sample(fhd_ctx* c) {
  if (c && !c->sk_acc && !fhd_sk_reset(c)) return 0;
  CTX_TRY(c, {
    // the spectrum buffer is free between steps; the transform is the step's own forward pass
    fft_forward(c, c->v);
    fft_ztransform(c, 1);
    launch_sk_accum(c->spec, c->sk_acc, 3 * (size_t)c->p.mz * c->p.my * c->g.P, c->s);
    HIPCHK(hipGetLastError());
    c->sk_samples++;
  });
  return 1;
}

int fhd_sk_read(fhd_ctx* c, double* s, long long* nsamples) {
  CTX_TRY(c, {
    const int P = c->g.P, my = c->p.my, mz = c->p.mz;
    const size_t per = (size_t)mz * my * P, n = 3 * per;
    std::vector<double> h(n, 0.0);
    HIPCHK(hipStreamSynchronize(c->s));
    if (c->sk_acc) HIPCHK(hipMemcpy(h.data(), c->sk_acc, n * sizeof(double), hipMemcpyDeviceToHost));
    const double kT = c->p.temperature;
    const double mult = kT != 0.0 ? c->p.densfluid * c->dV / kT : 1.0;
    const double norm = c->sk_samples > 0 ? mult / ((double)c->g.N * (double)c->sk_samples) : 0.0;
    // device [a][ky][z][kx] -> caller [a][z][ky][kx]
    for (int a = 0; a < 3; ++a)
      for (int z = 0; z < mz; ++z)
        for (int ky = 0; ky < my; ++ky)
          for (int kx = 0; kx < P; ++kx)
            s[a * per + ((size_t)z * my + ky) * P + kx] = norm * h[a * per + ((size_t)ky * mz + z) * P + kx];
    if (nsamples) *nsamples = c->sk_samples;
  });
  return 1;
}

int fhd_set_noise_overlap(fhd_ctx* c, int on) {
  CTX_TRY(c, {
    const bool v = on != 0;
    if (v != c->noise_overlap) {
      // captured steps hold the old stream layout: drop them, the next graph step recaptures
      HIPCHK(hipStreamSynchronize(c->s));
      for (auto& e : c->gexec)
        if (e) {
          HIPCHK(hipGraphExecDestroy(e));
          e = nullptr;
        }
      c->graphs_built = false;
      c->noise_overlap = v;
    }
  });
  return 1;
}

int fhd_sync(fhd_ctx* c) {
  CTX_TRY(c, {
    HIPCHK(hipStreamSynchronize(c->s));
    check_device_flags(c);
  });
  return 1;
}

long long fhd_get_step(const fhd_ctx* c) { return c ? c->step : -1; }
int fhd_get_np(const fhd_ctx* c) { return c ? c->np : -1; }

int fhd_time_steps(fhd_ctx* c, int nsteps, double* ms) {
  CTX_TRY(c, {
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, c->s));
    for (int k = 0; k < nsteps; ++k) advance(c);
    HIPCHK(hipEventRecord(b, c->s));
    HIPCHK(hipEventSynchronize(b));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  });
  return 1;
}

int fhd_profile_kernels(fhd_ctx* c, int nsteps, double* kern_ms, char* kern_names, int nkern) {
  int nout = 0;
  CTX_TRY(c, {
    for (auto& m : c->marks) (void)hipEventDestroy(m.second);
    c->marks.clear();
    c->profiling = true;
    mark(c, "begin");
    for (int k = 0; k < nsteps; ++k) do_step(c);
    c->profiling = false;
    HIPCHK(hipStreamSynchronize(c->s));
    std::vector<std::string> names;
    std::vector<double> tot;
    std::vector<int> cnt;
    for (size_t i = 1; i < c->marks.size(); ++i) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, c->marks[i - 1].second, c->marks[i].second));
      const std::string& nm = c->marks[i].first;
      size_t j = 0;
      for (; j < names.size(); ++j)
        if (names[j] == nm) break;
      if (j == names.size()) { names.push_back(nm); tot.push_back(0); cnt.push_back(0); }
      tot[j] += t;
      cnt[j] += 1;
    }
    for (size_t j = 0; j < names.size() && (int)j < nkern; ++j) {
      kern_ms[2 * j] = tot[j] / std::max(cnt[j], 1);
      kern_ms[2 * j + 1] = cnt[j];
      snprintf(kern_names + 48 * j, 48, "%s", names[j].c_str());
      ++nout;
    }
    for (auto& m : c->marks) (void)hipEventDestroy(m.second);
    c->marks.clear();
  });
  return nout;
}

int fhd_get_schedule(const fhd_ctx* c, int out[6]) {
  if (!c || !out) return 0;
  out[0] = c->fa.btile;
  out[1] = c->fa.plan_xy;
  out[2] = c->fa.plan_zs;
  out[3] = c->zchunk;
  out[4] = c->zchunk_b;
  out[5] = c->fa.plan_xy_inv;
  return 1;
}

int fhd_max_divergence(fhd_ctx* c, double* maxdiv) {
  CTX_TRY(c, {
    launch_divergence(c->g, c->v, c->div, c->s);
    std::vector<float> h(c->g.N);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(h.data(), c->div, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    double m = 0;
    for (float x : h) m = std::max(m, (double)fabsf(x));
    *maxdiv = m;
  });
  return 1;
}

// ---------------- building blocks for parity tests ----------------
int fhd_fft_r2c(int mx, int my, int mz, const double* in, double* out_spec) {
  fhd_params p;
  fhd_default_params(&p);
  p.scheme = FHD_SCHEME_STOKES_LIMIT;
  p.mx = mx; p.my = my; p.mz = mz;
  p.lx = mx; p.ly = my; p.lz = mz;
  p.dt = 1.0;
  fhd_ctx* c = fhd_create(&p);
  if (!c) return 0;
  int ok = 1;
  try {
    const size_t N3 = 3 * (size_t)c->g.N;
    std::vector<float> h(N3, 0.0f);
    for (size_t i = 0; i < (size_t)c->g.N; ++i) h[i] = (float)in[i];
    HIPCHK(hipMemcpy(c->W, h.data(), N3 * sizeof(float), hipMemcpyHostToDevice));
    fft_forward(c, c->W);
    fft_ztransform(c, 1);
    const size_t ns = (size_t)mz * my * c->g.P;
    std::vector<float2> sp(ns);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(sp.data(), c->spec, ns * sizeof(float2), hipMemcpyDeviceToHost));
    // device layout [ky][z][kx] -> host layout [z][ky][kx]
    const int P = c->g.P;
    for (int z = 0; z < mz; ++z)
      for (int ky = 0; ky < my; ++ky)
        for (int kx = 0; kx < P; ++kx) {
          size_t h = ((size_t)z * my + ky) * P + kx, d = ((size_t)ky * mz + z) * P + kx;
          out_spec[2 * h] = sp[d].x;
          out_spec[2 * h + 1] = sp[d].y;
        }
  } catch (const std::exception& e) {
    g_last_error = e.what();
    ok = 0;
  }
  fhd_destroy(c);
  return ok;
}

int fhd_fft_c2r(int mx, int my, int mz, const double* in_spec, double* out) {
  fhd_params p;
  fhd_default_params(&p);
  p.scheme = FHD_SCHEME_STOKES_LIMIT;
  p.mx = mx; p.my = my; p.mz = mz;
  p.lx = mx; p.ly = my; p.lz = mz;
  p.dt = 1.0;
  fhd_ctx* c = fhd_create(&p);
  if (!c) return 0;
  int ok = 1;
  try {
    const size_t ns = (size_t)mz * my * c->g.P;
    std::vector<float2> sp(3 * ns, make_float2(0.0f, 0.0f));
    const int P = c->g.P;
    for (int z = 0; z < mz; ++z)
      for (int ky = 0; ky < my; ++ky)
        for (int kx = 0; kx < P; ++kx) {
          size_t h = ((size_t)z * my + ky) * P + kx, d = ((size_t)ky * mz + z) * P + kx;
          sp[d] = make_float2((float)in_spec[2 * h], (float)in_spec[2 * h + 1]);
        }
    HIPCHK(hipMemcpy(c->spec, sp.data(), 3 * ns * sizeof(float2), hipMemcpyHostToDevice));
    fft_ztransform(c, 2);
    fft_inverse(c, c->W, nullptr, 1.0f, 0.0f);
    std::vector<float> h(c->g.N);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(h.data(), c->W, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
  } catch (const std::exception& e) {
    g_last_error = e.what();
    ok = 0;
  }
  fhd_destroy(c);
  return ok;
}

// ---------------- the no-allocation forward transform ----------------
// fhd_fft_r2c_noalloc works in a file-scope device array instead of allocated memory: the half
// spectrum in the passes' [ky][z][kx] layout, then the transform tables.  Each GPU holds its own
// copy until the process ends; its address is taken once per device with hipGetSymbolAddress
// and only that address is handed to kernels.
int fhd_fft_r2c_noalloc(int mx, int my, int mz, const float* in, float* out_spec, void* stream) {
  try {
    if (!in || !out_spec) throw std::runtime_error("fhd_fft_r2c_noalloc: null buffer");
    fhd_params p;
    fhd_default_params(&p);
    p.scheme = FHD_SCHEME_STOKES_LIMIT;
    p.mx = mx; p.my = my; p.mz = mz;
    p.lx = mx; p.ly = my; p.lz = mz;
    p.dt = 1.0;
    validate(p);
    const int P = mx / 2 + 1;
    if ((size_t)mz * my * P > kNoAllocSpecMax)
      throw std::runtime_error("fhd_fft_r2c_noalloc: mz * my * (mx/2 + 1) exceeds " +
                               std::to_string(kNoAllocSpecMax) + " (a 128^3 grid)");
    float2* sc = noalloc_scratch();
    float2* spec = sc;
    float2* tab = sc + kNoAllocSpecMax;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_noalloc_tables, dim3(1), dim3(kMaxFftLen), 0, s, tab, mx, my, mz);
    HIPCHK(hipGetLastError());
    FftArgs a;
    memset(&a, 0, sizeof(a));
    GridDev& g = a.g;
    g.mx = mx; g.my = my; g.mz = mz;
    g.N = mx * my * mz;
    g.P = P;
    g.dx = g.dy = g.dz = 1.0f;
    g.idx = g.idy = g.idz = 1.0f;
    g.lx = (float)mx; g.ly = (float)my; g.lz = (float)mz;
    a.twx = tab; a.twy = tab + kMaxFftLen; a.twz = tab + 2 * kMaxFftLen;
    a.ex = tab + 3 * kMaxFftLen; a.ey = tab + 4 * kMaxFftLen; a.ez = tab + 5 * kMaxFftLen;
    const float* gt = reinterpret_cast<const float*>(tab + 6 * kMaxFftLen);
    a.gx = gt; a.gy = gt + kMaxFftLen; a.gz = gt + 2 * kMaxFftLen;
    make_plan(mx / 2, &a.plan_h, &a.nst_h);
    make_plan(my, &a.plan_y, &a.nst_y);
    make_plan(mz, &a.plan_z, &a.nst_z);
    a.ztile = std::max(1, std::min(16, 8448 / (3 * mz)));
    a.btile = 8;
    a.zmode = 1;
    a.ncomp = 1;
    const char* gen = getenv("FHD_FFT_GENERIC");
    a.generic = (gen && gen[0] == '1') ? 1 : 0;
    a.eta = 1.0f;
    a.inv_n = (float)(1.0 / (double)g.N);
    launch_fft_xy_fwd(a, in, spec, s);
    launch_fft_z_solve(a, spec, s);
    const long long ns = (long long)mz * my * P;
    hipLaunchKernelGGL(k_noalloc_to_host_layout, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s,
                       (const float2*)spec, reinterpret_cast<float2*>(out_spec), my, mz, P);
    HIPCHK(hipGetLastError());
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 0;
  }
  return 1;
}

int fhd_normals(unsigned long long seed, long long step, int stream, int sub, int n, double* out) {
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("no HIP device");
    Key k;
    k.k0 = (uint32_t)seed;
    k.k1 = (uint32_t)(seed >> 32);
    k.step_lo = (uint32_t)(uint64_t)step;
    k.step_hi = (uint32_t)((uint64_t)step >> 32);
    k.step = nullptr;
    float* d = dalloc<float>(4 * (size_t)n);
    launch_normals(k, stream, sub, n, d, 0);
    std::vector<float> h(4 * (size_t)n);
    HIPCHK(hipMemcpy(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipFree(d));
    for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 0;
  }
  return 1;
}

// Internal: lets the scheme drivers report through fhd_last_error().
void fhd_internal_set_error(const char* e) { g_last_error = e ? e : ""; }

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Batches of independent replicas and the completion / memory queries the serving loop uses.
// No fluam counterpart: fluam runs one simulation per process.
// ------------------------------------------------------------------------------------------
constexpr int kMaxBranches = 8;
struct fhd_batch {
  std::vector<fhd_ctx*> reps;  // each replica owns its memory and stream
  bool owns = true;            // false for a batch over contexts the caller keeps (fhd_batch_adopt)
  hipStream_t timer = nullptr;
  std::string err;
  // whole-batch graphs: one graph per phase holding every replica's step as an independent
  // branch (fork from / join to the timer stream), replayed with a single launch per step
  // 0: each replica replays its own graphs on its stream; 1: one batch graph of per-replica
  // branches; 2: merged stages, one dispatch per stage and kernel instance for all replicas
  int mode = 0;
  bool built = false;
  int period = 1, phase = 0;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  // merged mode: per phase a graph of batched dispatches reading the tables in `tab`
  bool merged_built = false;
  hipGraphExec_t mexec[2] = {nullptr, nullptr};
  void* tab[2] = {nullptr, nullptr};
  int ndispatch[2] = {0, 0};
  bool timer_ahead = false;  // merged work on the timer stream not yet ordered before replica streams
  int branches = 4;          // parallel graph branches per round of merged dispatches
  hipStream_t side[kMaxBranches] = {};
  hipEvent_t side_done[kMaxBranches] = {};
  hipEvent_t round_fork = nullptr;
};

namespace {

// Device bytes a context with these parameters allocates (mirrors fhd_create).
long long ctx_bytes(const fhd_params& p) {
  const long long N = (long long)p.mx * p.my * p.mz, N3 = 3 * N;
  const long long P = p.mx / 2 + 1;
  long long b = 6 * N3 * 4 + N * 4 + 3LL * p.mz * p.my * P * 8;  // v v2 W SF adv[2], div, spectrum
  b += 3LL * (p.mx + p.my + p.mz) * 20;                            // tables
  if (p.np > 0) {
    const long long np = p.np;
    const int te[3] = {tile_edge(p.mx), tile_edge(p.my), tile_edge(p.mz)};
    long long nt = 1;
    const int m[3] = {p.mx, p.my, p.mz};
    for (int a = 0; a < 3; ++a) nt *= te[a] > 0 ? m[a] / te[a] : 1;
    b += 2 * np * (16 + 16 + 16 + 4 + 16 + 16) + 2 * (nt + 1) * 4;  // positions, forces, sorted sets
    b += np * 16 + nt * 4 + 2 * np * 4;                               // u, counts, slots, scratch
    b += 2 * 2 * np * ((long long)sizeof(SrcRec) + 4);                // spreading records
    if (p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT && p.mass != 0.0) b += N3 * 4 + 5 * np * 16;
  }
  return b;
}

}  // namespace

extern "C" {

long long fhd_bytes_needed(const fhd_params* p) {
  if (!p) return -1;
  fhd_params q = *p;
  std::string err;
  if (q.initial_positions == 2 && !coordinates_count(q.coordinates, &q.np, err)) {
    g_last_error = err;
    return -1;
  }
  return ctx_bytes(q);
}

int fhd_mem_info(unsigned long long* free_bytes, unsigned long long* total_bytes) {
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) {
    g_last_error = "hipMemGetInfo failed";
    return 0;
  }
  *free_bytes = f;
  *total_bytes = t;
  return 1;
}

int fhd_record(fhd_ctx* c) {
  CTX_TRY(c, {
    if (!c->done_ev) HIPCHK(hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(c->done_ev, c->s));
    c->done_checked = false;
  });
  return 1;
}

// Once the marked work has completed, the device error word is read once (a few bytes): a
// context whose steps flagged an error reports -1 from then on, so a serving loop can evict it.
int fhd_query(fhd_ctx* c) {
  if (!c || !c->done_ev) return 1;
  if (c->dev_error) return -1;
  const hipError_t e = hipEventQuery(c->done_ev);
  if (e == hipErrorNotReady) return 0;
  if (e != hipSuccess) {
    c->err = hipGetErrorString(e);
    return -1;
  }
  if (!c->done_checked) {
    c->done_checked = true;
    try {
      check_device_flags(c);
    } catch (const std::exception& ex) {
      c->err = ex.what();
      c->dev_error = true;
      return -1;
    }
  }
  return 1;
}

fhd_batch* fhd_batch_create(const fhd_params* p, int n) {
  fhd_batch* b = new fhd_batch();
  try {
    if (n <= 0) throw std::runtime_error("a batch needs at least one replica");
    for (int i = 0; i < n; ++i) {
      fhd_ctx* c = fhd_create(&p[i]);
      if (!c) throw std::runtime_error("replica " + std::to_string(i) + ": " + g_last_error);
      c->graph_mode = true;
      b->reps.push_back(c);
    }
    HIPCHK(hipStreamCreate(&b->timer));
  } catch (const std::exception& e) {
    g_last_error = e.what();
    for (fhd_ctx* c : b->reps) fhd_destroy(c);
    delete b;
    return nullptr;
  }
  return b;
}

fhd_batch* fhd_batch_adopt(fhd_ctx* const* ctxs, int n) {
  fhd_batch* b = new fhd_batch();
  b->owns = false;
  try {
    if (n <= 0 || !ctxs) throw std::runtime_error("a batch needs at least one context");
    for (int i = 0; i < n; ++i) {
      fhd_ctx* c = ctxs[i];
      if (!c) throw std::runtime_error("null context " + std::to_string(i));
      for (fhd_ctx* o : b->reps)
        if (o == c) throw std::runtime_error("a context appears twice in the batch");
      c->graph_mode = true;
      b->reps.push_back(c);
    }
    HIPCHK(hipStreamCreate(&b->timer));
  } catch (const std::exception& e) {
    g_last_error = e.what();
    delete b;
    return nullptr;
  }
  return b;
}

void fhd_batch_destroy(fhd_batch* b) {
  if (!b) return;
  if (b->timer) (void)hipStreamSynchronize(b->timer);
  for (auto& e : b->exec)
    if (e) (void)hipGraphExecDestroy(e);
  for (auto& e : b->mexec)
    if (e) (void)hipGraphExecDestroy(e);
  for (auto& t : b->tab)
    if (t) (void)hipFree(t);
  for (auto& x : b->side)
    if (x) (void)hipStreamDestroy(x);
  for (auto& x : b->side_done)
    if (x) (void)hipEventDestroy(x);
  if (b->round_fork) (void)hipEventDestroy(b->round_fork);
  for (auto& e : b->join)
    if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  // (adopted contexts stay with the caller; their streams were ordered after the batch's steps
  // by fhd_batch_step, and the timer stream has drained above)
  if (b->owns)
    for (fhd_ctx* c : b->reps) fhd_destroy(c);
  if (b->timer) (void)hipStreamDestroy(b->timer);
  delete b;
}

int fhd_batch_size(const fhd_batch* b) { return b ? (int)b->reps.size() : 0; }

fhd_ctx* fhd_batch_replica(fhd_batch* b, int i) {
  return (b && i >= 0 && i < (int)b->reps.size()) ? b->reps[i] : nullptr;
}

const char* fhd_batch_error(const fhd_batch* b) { return b ? b->err.c_str() : g_last_error.c_str(); }

int fhd_batch_set_mode(fhd_batch* b, int mode) {
  if (!b) return 0;
  if (mode < 0 || mode > 2) {
    b->err = "batch mode must be 0, 1 or 2";
    return 0;
  }
  b->mode = mode;
  return 1;
}

}  // extern "C"

namespace {

// Capture one graph per phase for the whole batch: the timer stream forks to every replica
// stream, each replica records its steady-state step there, and the timer stream joins them.
void build_batch_graphs(fhd_batch* b) {
  b->period = 1;
  for (fhd_ctx* c : b->reps)
    if (c->p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) b->period = 2;
  if (!b->fork) HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  if (b->join.empty()) {
    b->join.resize(b->reps.size(), nullptr);
    for (auto& e : b->join) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  for (fhd_ctx* c : b->reps) {
    if (!c->step_dev) c->step_dev = dalloc<unsigned long long>(1);
    HIPCHK(hipStreamSynchronize(c->s));
    const unsigned long long v = (unsigned long long)c->step;
    HIPCHK(hipMemcpy(c->step_dev, &v, sizeof(v), hipMemcpyHostToDevice));
    c->dev_step_valid = true;
  }
  for (int ph = 0; ph < b->period; ++ph) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(b->timer, hipStreamCaptureModeGlobal));
    HIPCHK(hipEventRecord(b->fork, b->timer));
    for (size_t i = 0; i < b->reps.size(); ++i) {
      fhd_ctx* c = b->reps[i];
      HIPCHK(hipStreamWaitEvent(c->s, b->fork, 0));
      c->capturing = true;
      scheme_step(c);
      launch_advance_step(c->step_dev, c->s);
      c->capturing = false;
      HIPCHK(hipEventRecord(b->join[i], c->s));
      HIPCHK(hipStreamWaitEvent(b->timer, b->join[i], 0));
    }
    HIPCHK(hipStreamEndCapture(b->timer, &g));
    HIPCHK(hipGraphInstantiate(&b->exec[ph], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
  }
  // each replica's host state went through `period` steps: back where it started for QNB
  // (period 2) and unchanged for the other schemes
  b->phase = 0;
  b->built = true;
}

// ---- merged stages (mode 2) ----
// Each replica's steady-state step is recorded (not launched): the list of its launches with
// their argument structs.  The lists are merged round by round: a round takes the stage family
// at the head of most lists and issues one batched dispatch per kernel instance covering every
// replica whose next launch is of that family.  Launches of one replica keep their order, and
// the dispatches run in order on one stream, so every replica sees its own step unchanged.
struct MergedDispatch {
  const void* fn;
  int threads;
  unsigned nblocks;
  size_t blk, dims, args;  // byte offsets into the phase's device table
  int stride;
  int round;               // dispatches of one round touch disjoint replicas
};

std::vector<LaunchRecord> record_step(fhd_ctx* c) {
  Recorder rec;
  t_recorder = &rec;
  c->capturing = true;
  const int btile = c->fa.btile;
  // one column tile for every replica, so replicas of equal z-length share one kernel instance
  // (the column tile changes the work split only, never a bit of the result)
  c->fa.btile = 8;
  try {
    scheme_step(c);
    launch_advance_step(c->step_dev, c->s);
  } catch (...) {
    t_recorder = nullptr;
    c->capturing = false;
    c->fa.btile = btile;
    throw;
  }
  t_recorder = nullptr;
  c->capturing = false;
  c->fa.btile = btile;
  return std::move(rec.list);
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Merge the replicas' launch lists into dispatches and their device tables (host image).
void merge_lists(const std::vector<std::vector<LaunchRecord>>& lists, std::vector<MergedDispatch>& out,
                 std::vector<unsigned char>& image) {
  const int R = (int)lists.size();
  std::vector<size_t> pos(R, 0);
  int round = 0;
  auto put = [&](const void* src, size_t n) {
    const size_t at = align_up(image.size(), 256);
    image.resize(at + n);
    memcpy(image.data() + at, src, n);
    return at;
  };
  for (;;) {
    // family with the most lists at it; ties go to the family whose list has most left
    std::vector<std::pair<int, std::pair<int, size_t>>> fam;  // family, (count, longest rest)
    for (int r = 0; r < R; ++r) {
      if (pos[r] >= lists[r].size()) continue;
      const int f = lists[r][pos[r]].family;
      const size_t rest = lists[r].size() - pos[r];
      bool found = false;
      for (auto& e : fam)
        if (e.first == f) {
          e.second.first++;
          e.second.second = std::max(e.second.second, rest);
          found = true;
        }
      if (!found) fam.push_back({f, {1, rest}});
    }
    if (fam.empty()) break;
    int best = 0;
    for (int k = 1; k < (int)fam.size(); ++k)
      if (fam[k].second > fam[best].second) best = k;
    const int f = fam[best].first;
    // group the heads of that family by kernel instance, in replica order
    std::vector<std::pair<const void*, std::vector<int>>> groups;
    for (int r = 0; r < R; ++r) {
      if (pos[r] >= lists[r].size() || lists[r][pos[r]].family != f) continue;
      const void* fn = lists[r][pos[r]].fn;
      bool found = false;
      for (auto& g : groups)
        if (g.first == fn) {
          g.second.push_back(r);
          found = true;
        }
      if (!found) groups.push_back({fn, {r}});
    }
    for (auto& g : groups) {
      const LaunchRecord& first = lists[g.second[0]][pos[g.second[0]]];
      MergedDispatch d;
      d.round = round;
      d.fn = g.first;
      d.threads = first.threads;
      d.stride = (int)align_up(first.args.size(), 16);
      std::vector<int4> blk, dims;
      std::vector<unsigned char> args(g.second.size() * (size_t)d.stride, 0);
      for (size_t e = 0; e < g.second.size(); ++e) {
        const LaunchRecord& L = lists[g.second[e]][pos[g.second[e]]];
        if (L.threads != d.threads || L.args.size() != first.args.size())
          throw std::runtime_error("merged batch: launches of one kernel instance disagree in shape");
        dims.push_back(make_int4((int)L.grid.x, (int)L.grid.y, (int)L.grid.z, 0));
        for (unsigned z = 0; z < L.grid.z; ++z)
          for (unsigned y = 0; y < L.grid.y; ++y)
            for (unsigned x = 0; x < L.grid.x; ++x) blk.push_back(make_int4((int)e, (int)x, (int)y, (int)z));
        memcpy(args.data() + e * d.stride, L.args.data(), L.args.size());
      }
      if (blk.size() > 0x7fffffffu) throw std::runtime_error("merged batch: dispatch too large");
      d.nblocks = (unsigned)blk.size();
      d.blk = put(blk.data(), blk.size() * sizeof(int4));
      d.dims = put(dims.data(), dims.size() * sizeof(int4));
      d.args = put(args.data(), args.size());
      out.push_back(d);
      for (int r : g.second) pos[r]++;
    }
    ++round;
  }
}

// Order the timer stream after work queued on the replica streams, or the reverse.
void join_replicas_to_timer(fhd_batch* b) {
  if (b->join.empty()) {
    b->join.resize(b->reps.size(), nullptr);
    for (auto& e : b->join) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  for (size_t i = 0; i < b->reps.size(); ++i) {
    HIPCHK(hipEventRecord(b->join[i], b->reps[i]->s));
    HIPCHK(hipStreamWaitEvent(b->timer, b->join[i], 0));
  }
}
void fork_timer_to_replicas(fhd_batch* b) {
  if (!b->timer_ahead) return;
  if (!b->fork) HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  HIPCHK(hipEventRecord(b->fork, b->timer));
  for (fhd_ctx* c : b->reps) HIPCHK(hipStreamWaitEvent(c->s, b->fork, 0));
  b->timer_ahead = false;
}

void build_merged(fhd_batch* b) {
  if (const char* e = getenv("FHD_BATCH_BRANCHES")) b->branches = std::max(1, std::min(kMaxBranches, atoi(e)));
  if (!b->round_fork) HIPCHK(hipEventCreateWithFlags(&b->round_fork, hipEventDisableTiming));
  for (int k = 0; k < b->branches; ++k) {
    if (!b->side[k]) HIPCHK(hipStreamCreate(&b->side[k]));
    if (!b->side_done[k]) HIPCHK(hipEventCreateWithFlags(&b->side_done[k], hipEventDisableTiming));
  }
  b->period = 1;
  for (fhd_ctx* c : b->reps)
    if (c->p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) b->period = 2;
  for (fhd_ctx* c : b->reps) {
    if (!c->step_dev) c->step_dev = dalloc<unsigned long long>(1);
    HIPCHK(hipStreamSynchronize(c->s));
    const unsigned long long v = (unsigned long long)c->step;
    HIPCHK(hipMemcpy(c->step_dev, &v, sizeof(v), hipMemcpyHostToDevice));
    c->dev_step_valid = true;
  }
  for (int ph = 0; ph < b->period; ++ph) {
    std::vector<std::vector<LaunchRecord>> lists;
    for (fhd_ctx* c : b->reps) lists.push_back(record_step(c));
    std::vector<MergedDispatch> ds;
    std::vector<unsigned char> image;
    merge_lists(lists, ds, image);
    HIPCHK(hipMalloc(&b->tab[ph], std::max<size_t>(image.size(), 256)));
    HIPCHK(hipMemcpy(b->tab[ph], image.data(), image.size(), hipMemcpyHostToDevice));
    const unsigned char* base = static_cast<const unsigned char*>(b->tab[ph]);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(b->timer, hipStreamCaptureModeThreadLocal));
    try {
      // the dispatches of a round are independent: up to `branches` of them run as parallel
      // branches of the graph (fork from the timer stream, join back before the next round)
      for (size_t i0 = 0; i0 < ds.size();) {
        size_t i1 = i0;
        while (i1 < ds.size() && ds[i1].round == ds[i0].round) ++i1;
        const int nb = std::min<int>((int)(i1 - i0), b->branches);
        if (nb > 1) {
          HIPCHK(hipEventRecord(b->round_fork, b->timer));
          for (int k = 0; k < nb; ++k) HIPCHK(hipStreamWaitEvent(b->side[k], b->round_fork, 0));
        }
        for (size_t i = i0; i < i1; ++i) {
          const MergedDispatch& d = ds[i];
          const int4* blk = reinterpret_cast<const int4*>(base + d.blk);
          const int4* dims = reinterpret_cast<const int4*>(base + d.dims);
          const unsigned char* args = base + d.args;
          int stride = d.stride;
          void* kargs[4] = {&blk, &dims, &args, &stride};
          hipStream_t st = nb > 1 ? b->side[(i - i0) % nb] : b->timer;
          HIPCHK(hipLaunchKernel(d.fn, dim3(d.nblocks), dim3(d.threads), kargs, 0, st));
        }
        if (nb > 1)
          for (int k = 0; k < nb; ++k) {
            HIPCHK(hipEventRecord(b->side_done[k], b->side[k]));
            HIPCHK(hipStreamWaitEvent(b->timer, b->side_done[k], 0));
          }
        i0 = i1;
      }
    } catch (...) {
      (void)hipStreamEndCapture(b->timer, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(b->timer, &g));
    HIPCHK(hipGraphInstantiate(&b->mexec[ph], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    b->ndispatch[ph] = (int)ds.size();
  }
  b->phase = 0;
  b->merged_built = true;
}

void batch_advance(fhd_batch* b) {
  bool all_steady = true;
  for (fhd_ctx* c : b->reps) all_steady = all_steady && steady(c);
  if (b->mode == 0 || !all_steady) {
    fork_timer_to_replicas(b);
    for (fhd_ctx* c : b->reps) advance(c);
    if (b->built || b->merged_built) b->phase = (b->phase + 1) % b->period;  // host state moved one step
    return;
  }
  if (b->mode == 2 ? !b->merged_built : !b->built) {
    if (b->built || b->merged_built) throw std::runtime_error("batch mode changed after its step was captured");
    if (b->mode == 2) build_merged(b);
    else build_batch_graphs(b);
  }
  bool joined = false;
  for (fhd_ctx* c : b->reps)
    if (!c->dev_step_valid) {  // an eager step ran since the counter was last on the device
      HIPCHK(hipStreamSynchronize(c->s));
      const unsigned long long v = (unsigned long long)c->step;
      HIPCHK(hipMemcpy(c->step_dev, &v, sizeof(v), hipMemcpyHostToDevice));
      c->dev_step_valid = true;
      joined = true;
    }
  // the batch graph runs on the timer stream: order it after work queued on the replica streams
  // (an eager step, a structure-factor sample, a velocity upload) and those streams after it
  if (!b->timer_ahead && !joined) join_replicas_to_timer(b);
  b->timer_ahead = true;
  HIPCHK(hipGraphLaunch(b->mode == 2 ? b->mexec[b->phase] : b->exec[b->phase], b->timer));
  for (fhd_ctx* c : b->reps) {
    if (c->p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) qnb_after_step(c);
    c->step++;
    if (c->graphs_built) c->gphase = (c->gphase + 1) % c->gperiod;
  }
  b->phase = (b->phase + 1) % b->period;
}

}  // namespace

extern "C" {

// All replicas advance step by step, each on its own stream, so their kernels overlap.
int fhd_batch_step(fhd_batch* b, int nsteps) {
  if (!b) return 0;
  try {
    for (int k = 0; k < nsteps; ++k) batch_advance(b);
    fork_timer_to_replicas(b);
    HIPCHK(hipGetLastError());
  } catch (const std::exception& e) {
    b->err = e.what();
    return 0;
  }
  return 1;
}

int fhd_batch_sync(fhd_batch* b) {
  if (!b) return 0;
  for (fhd_ctx* c : b->reps)
    if (!fhd_sync(c)) {
      b->err = c->err;
      return 0;
    }
  return 1;
}

// Wall time (HIP events) of nsteps batch steps: every replica stream starts after one event
// and the timer stream waits for all of them to finish.
int fhd_batch_time_steps(fhd_batch* b, int nsteps, double* ms) {
  if (!b) return 0;
  std::vector<hipEvent_t> ends(b->reps.size(), nullptr);
  hipEvent_t a = nullptr, e = nullptr;
  try {
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&e));
    for (auto& x : ends) HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    for (fhd_ctx* c : b->reps) HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipEventRecord(a, b->timer));
    for (fhd_ctx* c : b->reps) HIPCHK(hipStreamWaitEvent(c->s, a, 0));
    for (int k = 0; k < nsteps; ++k) batch_advance(b);
    for (size_t i = 0; i < b->reps.size(); ++i) {
      HIPCHK(hipEventRecord(ends[i], b->reps[i]->s));
      HIPCHK(hipStreamWaitEvent(b->timer, ends[i], 0));
    }
    HIPCHK(hipEventRecord(e, b->timer));
    fork_timer_to_replicas(b);
    HIPCHK(hipEventSynchronize(e));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, a, e));
    *ms = t;
  } catch (const std::exception& ex) {
    b->err = ex.what();
    for (auto& x : ends)
      if (x) (void)hipEventDestroy(x);
    if (a) (void)hipEventDestroy(a);
    if (e) (void)hipEventDestroy(e);
    return 0;
  }
  for (auto& x : ends) (void)hipEventDestroy(x);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(e);
  return 1;
}

}  // extern "C"
