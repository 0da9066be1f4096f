/*
 * fhd.h -- C ABI of the MI355X (gfx950) fluctuating-hydrodynamics step library
 * (libfhd.so).  Plain C types only: pointers, sizes, doubles.  Every function that
 * returns int returns 1 on success and 0 on failure, the convention of the fluam
 * scheme entry points; the reason for a failure is in fhd_last_error() (global API)
 * or fhd_ctx_error() (context API).
 *
 * Reference interfaces replaced (fbusabiaga/fluam):
 *   fhd_load_datamain                 <- bool loadDataMain(int, char*[])      src/loadDataMain.cu:173
 *   fhd_scheme_incompressible         <- bool schemeIncompressible()           src/schemeIncompressible.cu:25, src/header.h:200
 *   fhd_scheme_quasi_neutrally_buoyant<- bool schemeQuasiNeutrallyBuoyant()    src/schemeQuasiNeutrallyBuoyant.cu:27, src/header.h:222
 *   fhd_scheme_stokes_limit           <- bool schemeStokesLimit()              src/schemeStokesLimit.cu:27, src/header.h:310
 *   fhd_main                          <- int main(int, char*[]) scheme dispatch src/main.cpp:28-112
 *   fhd_create/fhd_step/...           <- the runScheme*() time loops with their device
 *                                        state held in globals (src/runSchemeIncompressible.cu:21,
 *                                        src/runSchemeQuasiNeutrallyBuoyant.cu:21,
 *                                        src/runSchemeStokesLimit.cu:21); here the state is a
 *                                        re-entrant context instead of process globals.
 *
 * Host array layouts (double precision on the host side of the boundary):
 *   velocity: 3*N values, component-major: vx[0..N) vy[N..2N) vz[2N..3N),
 *             cell i = ix + mx*(iy + my*iz), v_x on the +x face of cell i
 *             (src/createCells.cpp:68-70, src/stokesLimitFunctions.cu:767).
 *   particles: 3*np values, component-major: rx[0..np) ry[np..2np) rz[2np..3np),
 *             unwrapped coordinates, box centred at the origin.
 */
#ifndef FHD_H
#define FHD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FHD_SCHEME_INCOMPRESSIBLE 0
#define FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT 1
#define FHD_SCHEME_STOKES_LIMIT 2

/* Random-finite-difference variant of the Stokes-limit thermal drift.
 * FHD_RFD_DISPLACED: bin and weight the particle at q +- delta/2 W (the term
 *   S(q + delta/2 W) W documented at src/boundaryParticlesFunctionStokesLimit.cu:152-153).
 * FHD_RFD_REFERENCE: bin at the displaced position but evaluate the kernel weights at
 *   the undisplaced q, as src/stokesLimitFunctions.cu:1741-1799,1872-1906 does. */
#define FHD_RFD_DISPLACED 0
#define FHD_RFD_REFERENCE 1

/* Run parameters: the data.main keywords of the three schemes (src/loadDataMain.cu:38-170)
 * plus the extension keywords listed next to each field. */
typedef struct fhd_params {
  int scheme;                 /* FHD_SCHEME_* */
  int mx, my, mz;             /* cells */
  double lx, ly, lz;          /* celldimension */
  double densfluid;           /* densfluid */
  double shearviscosity;      /* shearviscosity */
  double temperature;         /* temperature (kT) */
  double dt;                  /* dt */
  unsigned long long seed;    /* seed */
  long long numsteps;         /* numsteps */
  long long numsteps_relaxation; /* numstepsRelaxation */
  long long samplefreq;       /* samplefreq */
  int initfluid;              /* initfluid (src/initializeFluid.cpp:31-160): 0 background velocity;
                                 1 background + equilibrium fluctuations sqrt(kT/(rho dV)) N(0,1) per
                                 face with the mean removed (normals from the library's counter
                                 generator, stream 9); 2 transverse wave vy += 0.04 sin(2 pi z/lz);
                                 3 longitudinal wave vz = 0.004 sin(16 pi (z + dz/2)/lz); 4 vx = 0.425 */
  double background[3];       /* backgroundvelocity */
  int np;                     /* numberparticles */
  int kernel_points;          /* extension keyword `kernel` 3|4 (Peskin 3- or 4-point) */
  int nonbonded;              /* computeNonBondedForces */
  double cutoff;              /* cutoff */
  double wca_sigma;           /* extension keyword `sigma` (default 2*dx, src/initForcesNonBonded.cu:24) */
  double wca_epsilon;         /* extension keyword `epsilon` (default kT, src/initForcesNonBonded.cu:25) */
  double mass;                /* mass: the particles' excess mass m_e over the fluid they carry,
                                 m_f = 1.5 dV volumeParticle rho (quasiNeutrallyBuoyant; m_e + m_f > 0) */
  double volume_particle;     /* volumeParticle */
  double extra_mobility;      /* extraMobility (0 = off) */
  double rfd_delta;           /* RFD displacement (0.001, src/stokesLimitFunctions.cu:1719) */
  int rfd_mode;               /* extension keyword `rfdMode` FHD_RFD_* */
  int max_part_in_cell;       /* maxNumberPartInCell (capacity hint, per spreading tile) */
  int max_part_in_cell_nonbonded; /* maxNumberPartInCellNonBonded (capacity hint) */
  int initial_positions;      /* 0: seeded uniform-random, 1: simple cubic (src/simpleCubic.cpp:25),
                                 2: coordinates file, 3: seeded uniform-random without pairs closer
                                 than the WCA sigma (random sequential addition) */
  char outputname[512];       /* outputname */
  char coordinates[512];      /* coordinates file (loadparticles 1): its first number is the particle
                                 count, which replaces numberparticles (src/createParticles.cpp:81-100) */
  char fluid[512];            /* fluid: initial velocity and density from a .fluidFinalConfiguration
                                 file (src/loadDataMain.cu:443-445, src/initializeFluid.cpp:168-192) */
  long long initial_step;     /* initialStep (extension): step counter of the first step, so a run
                                 restarted from its saved state draws the same noise */
  char bonded_forces[512];    /* bondedForces: harmonic bonds file (format of bin/inputfiles/README:55-108,
                                 read as src/initializeBondedForces.cpp:29-160); empty = no bonds */
  int structure_factor;       /* extension keyword `structureFactor 1`: the scheme run samples the
                                 equal-time structure factor every samplefreq steps and writes
                                 <outputname>.structureFactor (fluam's HydroGrid hook,
                                 src/hydroAnalysisIncompressible.cu:35-92) */
  char velocities[512];       /* velocities: initial particle velocities, "np" then np lines "vx vy vz";
                                 np must equal the coordinates file's (src/createParticles.cpp:103-120) */
} fhd_params;

/* Fill *p with the defaults of src/loadDataMain.cu:177-299 (restated). */
void fhd_default_params(fhd_params* p);
/* sizeof(fhd_params) as compiled into the library (binding sanity check; no fluam counterpart). */
unsigned long fhd_params_size(void);

/* ---------------- global (fluam-style) API ---------------- */
int fhd_load_datamain(const char* path);          /* parse data.main into the global params */
/* Reads a .fluidFinalConfiguration file (the format of src/saveFluidFinalConfiguration.cu:30-75,
 * read as src/initializeFluid.cpp:168-192 does) for an mx x my x mz grid: v[3N] component-major,
 * x fastest; *densfluid receives the file's density.  CPU only; 1 on success. */
int fhd_read_fluid_file(const char* path, int mx, int my, int mz, double* v, double* densfluid);
int fhd_get_params(fhd_params* out);
int fhd_set_params(const fhd_params* in);
int fhd_scheme_incompressible(void);
int fhd_scheme_quasi_neutrally_buoyant(void);
int fhd_scheme_stokes_limit(void);
int fhd_main(const char* datamain_path);          /* load + dispatch in main.cpp order */
const char* fhd_last_error(void);

/* ---------------- re-entrant context API ---------------- */
typedef struct fhd_ctx fhd_ctx;
fhd_ctx* fhd_create(const fhd_params* p);         /* owns all device memory; NULL on error */
void fhd_destroy(fhd_ctx* c);
const char* fhd_ctx_error(const fhd_ctx* c);
int fhd_set_velocity(fhd_ctx* c, const double* v);       /* 3*N, layout above */
int fhd_get_velocity(fhd_ctx* c, double* v);
int fhd_set_particles(fhd_ctx* c, const double* q);      /* 3*np unwrapped positions */
int fhd_get_particles(fhd_ctx* c, double* q);
int fhd_get_particle_velocities(fhd_ctx* c, double* u);  /* J v at the last step (QNB) */
int fhd_get_forces(fhd_ctx* c, double* f);              /* last particle forces, 3*np */
int fhd_step(fhd_ctx* c, int nsteps);                    /* enqueue nsteps steps (async) */
/* Wait for the context's queued work, then check its device error flags (fluam's errorKernel):
 * 0 with the reason in fhd_ctx_error when a step produced a non-finite particle displacement or
 * one of 2^24 cells or more (not applied).  The flags are sticky. */
int fhd_sync(fhd_ctx* c);
/* Replay the steady-state step from captured HIP graphs (on != 0) instead of launching each
 * kernel.  No fluam counterpart (fluam launches every kernel each step); results are identical. */
int fhd_set_graph(fhd_ctx* c, int on);
/* Kept for callers of an earlier version of the library: the Stokes-limit step used to build
 * its noise field on a second stream.  The stress is now drawn in Fourier space inside the
 * solve, so the setting has no effect; it is accepted and returns 1.  No fluam counterpart. */
int fhd_set_noise_overlap(fhd_ctx* c, int on);

/* Equal-time structure factor of the velocity, computed on the GPU from the current field
 * (replaces the HydroGrid library fluam hands its velocities to, src/hydroAnalysisIncompressible.cu:35-92).
 * fhd_sk_sample adds |v_hat_a(k)|^2 (v_hat: unnormalised DFT) of the current velocity to an FP64
 * accumulator; fhd_sk_read writes S_a(k) = rho dV / kT * <|v_hat_a(k)|^2> / N (multiplier 1 when
 * kT = 0) over the half spectrum, s[3][mz][my][mx/2+1] with kx fastest, and the sample count.  At
 * equilibrium S_a(k) is the diagonal of the discrete projector, 1 - g_a^2/|g|^2.  fhd_sk_reset
 * clears the accumulator. */
int fhd_sk_reset(fhd_ctx* c);
int fhd_sk_sample(fhd_ctx* c);
int fhd_sk_read(fhd_ctx* c, double* s, long long* nsamples);
long long fhd_get_step(const fhd_ctx* c);
/* Particles the context holds: numberparticles, or the count its coordinates file declares
 * (loadparticles 1, src/createParticles.cpp:88); -1 for a null context. */
int fhd_get_np(const fhd_ctx* c);
/* Time nsteps steps with HIP events on the context's stream: total ms into *ms.
 * fhd_profile_kernels runs nsteps steps with HIP events around each kernel group and reports up
 * to nkern groups: kern_ms must hold 2*nkern doubles, filled with (mean ms per launch, launch
 * count) pairs, kern_ms[2j] and kern_ms[2j+1] for group j; kern_names must hold 48*nkern bytes,
 * group j's NUL-terminated name at kern_names + 48*j.  Returns the number of groups reported. */
int fhd_time_steps(fhd_ctx* c, int nsteps, double* ms);
int fhd_profile_kernels(fhd_ctx* c, int nsteps, double* kern_ms, char* kern_names, int nkern);

/* The schedule the context's autotuner chose for its grid shape (no fluam counterpart): out[0] the
 * z-pass column tile, out[1] the forward xy pass's radix plan (x plan + 3 y plan, 0-8, for 128 x 128
 * planes: x 16 4 | 4 16 | 8 8, y 16 8 | 8 16 | 4 4 8, and 9: the transforms in registers; for
 * 96 x 96 planes: x 16 3 | 3 16 | 4 4 3, y 16 2 3 | 8 4 3 | 3 8 4; 0 elsewhere), out[2] the z
 * pass's radix plan (0-2 for 48-, 96- and 128-point columns, the same candidates as y), out[3]
 * and out[4] the W kernel's z chunk alone and in batched dispatches, out[5] the inverse xy pass's
 * radix plan (the same candidates as out[1]). */
int fhd_get_schedule(const fhd_ctx* c, int out[6]);

/* Divergence diagnostics of the current velocity: max |D v| over cells, in units of
 * 1/time (discrete staggered divergence, src/kernelUpdateVIncompressible.cu:21-72 invariant). */
int fhd_max_divergence(fhd_ctx* c, double* maxdiv);

/* ---------------- batches and serving (no fluam counterpart) ----------------
 * For callers that run many independent simulations on one GPU: ensembles of replicas and
 * trajectory-serving loops.  fluam runs one simulation per process. */
/* Device bytes fhd_create would allocate for these parameters (admission against a budget). */
long long fhd_bytes_needed(const fhd_params* p);
/* Free and total device memory of the current GPU. */
int fhd_mem_info(unsigned long long* free_bytes, unsigned long long* total_bytes);
/* Mark the work enqueued so far on the context; fhd_query returns 1 once it has completed,
 * 0 while it runs, -1 on a device error (a HIP error, or a device error flag of fhd_sync's
 * kinds raised by the completed work; sticky). */
int fhd_record(fhd_ctx* c);
int fhd_query(fhd_ctx* c);
/* n replicas (own parameters, shapes and seeds), each on its own stream with graph-replayed
 * steps; fhd_batch_step advances all of them step by step so their kernels overlap.
 * Results equal those of each replica run alone. */
typedef struct fhd_batch fhd_batch;
fhd_batch* fhd_batch_create(const fhd_params* p, int n);   /* NULL on error (fhd_last_error) */
/* A batch over contexts the caller keeps (for a serving loop that regroups its running jobs as
 * they join and leave): fhd_batch_destroy leaves them alive, and each can be stepped alone or
 * adopted by another batch afterwards.  A context belongs to at most one batch at a time. */
fhd_batch* fhd_batch_adopt(fhd_ctx* const* ctxs, int n);
void fhd_batch_destroy(fhd_batch* b);
int fhd_batch_size(const fhd_batch* b);
fhd_ctx* fhd_batch_replica(fhd_batch* b, int i);           /* owned by the batch */
const char* fhd_batch_error(const fhd_batch* b);
/* mode 0: every replica replays its own graphs on its stream; mode 1: the whole batch step is
 * one captured graph (each replica an independent branch), replayed with one launch; mode 2:
 * merged stages, one batched dispatch per stage and kernel instance covering every replica,
 * captured as one graph.  Set the mode before the first steady-state step; changing it after
 * the step was captured is an error. */
int fhd_batch_set_mode(fhd_batch* b, int mode);
int fhd_batch_step(fhd_batch* b, int nsteps);
int fhd_batch_sync(fhd_batch* b);
int fhd_batch_time_steps(fhd_batch* b, int nsteps, double* ms);

/* ---------------- building blocks (exposed for parity tests) ---------------- */
/* Batched 3D real FFT round trip on device: real field (mz,my,mx) -> half spectrum
 * (mz,my,mx/2+1) complex (interleaved re,im) -> back; unnormalised both ways. */
int fhd_fft_r2c(int mx, int my, int mz, const double* in, double* out_spec);
int fhd_fft_c2r(int mx, int my, int mz, const double* in_spec, double* out);
/* No-allocation variant of fhd_fft_r2c.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  in: device pointer to a real FP32 field (mz,my,mx), x fastest;
 * out_spec: device pointer to (mz,my,mx/2+1) complex FP32 values, interleaved re,im;
 * unnormalised, the transform of fhd_fft_r2c.  Enqueued on `stream` (a hipStream_t; NULL is the
 * null stream) without waiting for it.  It allocates nothing: the work space is a file-scope
 * device array of 128*128*65 complex values plus tables (about 8.5 MB, held by each GPU it runs
 * on until the process ends), so mz*my*(mx/2+1) must not exceed 128*128*65 (a 128^3 grid) or the
 * call fails.  Every call on one GPU uses that GPU's one array: calls that overlap on a GPU (on
 * different streams or host threads) race.  1 on success, 0 with fhd_last_error() on failure. */
int fhd_fft_r2c_noalloc(int mx, int my, int mz, const float* in, float* out_spec, void* stream);
/* Counter-based normal generator used for every noise term (Philox4x32-10 + Box-Muller).
 * Writes n normals of (seed, step, stream, index=0..n/4, sub) as the kernels draw them. */
int fhd_normals(unsigned long long seed, long long step, int stream, int sub, int n, double* out);

#ifdef __cplusplus
}
#endif
#endif /* FHD_H */
