/*
 * fhd_oracle.h -- FP64 CPU restatement of fluam's incompressible, quasi-neutrally-buoyant
 * and Stokes-limit time steps.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for libfhd.so: only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.  Nothing in the
 * product links or calls it.
 *
 * It consumes the same counter-based noise stream as the GPU kernels (Philox4x32-10 keyed
 * by the seed, counter = (index, stream<<4|sub, step_lo, step_hi), Box-Muller on 24-bit
 * uniforms), so
 * GPU and oracle runs on the same seed see the same noise up to FP32 rounding.
 *
 * Parity pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_philox            Random123 known-answer vectors
 *   orc_fft3d             numpy.fft golden fixtures (tests/golden) + brute-force DFT
 *   orc_delta3/orc_delta4 closed-form values and Peskin moment conditions
 *   orc_ddelta3/4         central finite differences of the pinned kernels; moments
 *   orc_spread_drift      the QNB thermal drift equals the finite-difference gradient of the
 *                         pinned spread (3- and 4-point)
 *   orc_spread_quirk      built directly from its definition (nodes at q +- d/2, weights at q)
 *   orc_laplacian_term,   dense staggered L and D on a 4^3 grid; fluctuation-dissipation
 *   orc_noise_divergence  D D^T = -L
 *   orc_advection_term    momentum conservation, skew properties, continuum limit
 *   orc_kspace_stress     Hermitian (real after the inverse transform), white N(0,1) real
 *                         fields, per-mode second moments, divergence symbol equal to the
 *                         pinned real-space noise divergence
 *   orc_kspace_noise      real, discretely divergence-free, unchanged by the pinned projection;
 *                         per-mode second moments equal those of the pinned projection of the
 *                         pinned stress divergence (over many draws)
 *   orc_wca               analytic WCA (zero at 2^(1/6) sigma), Newton's third law
 *   orc_bonded            closed-form traps and springs, F = -grad V by finite differences of
 *                         V = sum k/2 (r - r0)^2, Newton's third law
 *   orc_spread/orc_interp direct triple loops; adjointness <J v, F> = dV <v, S F>
 *   incompressible step   dense-matrix brute force (I - dt nu L/2)^-1 P on a 4^3 grid,
 *                         divergence, momentum conservation, equal-time S(k)
 *   Stokes solve          dense saddle-point solve on a 4^3 grid, zero mean
 *   Stokes step           the midpoint rule over the pinned trap force, spread, Stokes solve,
 *                         interpolation, Fourier-space stress and RFD spreads (kT = 0 and
 *                         kT > 0), with extra mobility and with the reference's RFD variant
 *   QNB step              mass=0 reduces to the incompressible step with no particles;
 *                         particle velocities = J v; divergence; the bootstrap and
 *                         Adams-Bashforth steps over the pinned blocks at kT = 0 and kT > 0
 *                         (noise and thermal drift)
 *   QNB excess mass       m_e -> 0 reproduces the mass-0 step; total momentum
 *                         rho dV sum v + m_e sum u is conserved; at kT = 0 the bootstrap and
 *                         Adams-Bashforth steps with the excess-mass corrections over the
 *                         pinned blocks (spread, interpolation, projection, Crank-Nicolson
 *                         solve, Laplacian and advection terms)
 *   orc_project           idempotent, divergence-free output, kills discrete gradients,
 *                         keeps divergence-free fields and the mean
 */
#ifndef FHD_ORACLE_H
#define FHD_ORACLE_H
#include <stdint.h>
#include "../include/fhd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Philox4x32-10 block: ctr[4], key[2] -> out[4]. */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Four normals of one draw: (seed, step, stream, sub, index). */
void orc_normals4(uint64_t seed, int64_t step, int stream, int sub, uint32_t index, double out[4]);

/* Unnormalised in-place complex 3D FFT, layout [mz][my][mx] interleaved (re,im). sign=-1 fwd. */
int orc_fft3d(int mx, int my, int mz, double* data, int sign);

double orc_delta3(double r);   /* 3-point Roma-Peskin kernel, r in grid units */
double orc_delta4(double r);   /* 4-point Peskin kernel, r in grid units */
double orc_ddelta3(double r);  /* derivative of the 3-point kernel */
double orc_ddelta4(double r);  /* derivative of the 4-point kernel */

orc_ctx* orc_create(const fhd_params* p);
void orc_destroy(orc_ctx* c);
int orc_set_velocity(orc_ctx* c, const double* v);
int orc_get_velocity(orc_ctx* c, double* v);
int orc_set_particles(orc_ctx* c, const double* q);
int orc_get_particles(orc_ctx* c, double* q);
int orc_get_particle_velocities(orc_ctx* c, double* u);
int orc_get_forces(orc_ctx* c, double* f);
int orc_step(orc_ctx* c, int nsteps);
long long orc_get_step(orc_ctx* c);
int orc_get_np(orc_ctx* c);   /* numberparticles, or the count of the coordinates file */
int orc_set_threads(int n);

/* Building blocks on the context's grid / particles. */
int orc_wca(orc_ctx* c, const double* q, double* f);               /* q,f: 3*np */
int orc_bonded(orc_ctx* c, const double* q, double* f);            /* harmonic bonds only, f = bond forces */
int orc_spread(orc_ctx* c, const double* q, const double* F, double* field); /* field += S F (3*N) */
int orc_interp(orc_ctx* c, const double* q, const double* field, double* u);  /* u = J field */
/* field_c += scale * d/dq_c [phi phi phi] summed over the particles (the QNB thermal-drift
 * spread without its kT m_f/(m_e+m_f)/dV factor, which the caller passes in scale). */
int orc_spread_drift(orc_ctx* c, const double* q, double scale, double* field);
/* field += scale * S_quirk F: stencil nodes chosen from the displaced positions qd, kernel
 * weights evaluated at the undisplaced q (the reference's RFD spread, rfdMode 1). */
int orc_spread_quirk(orc_ctx* c, const double* q, const double* qd, const double* F, double scale,
                     double* field);
int orc_divergence(orc_ctx* c, const double* v, double* div);     /* div: N */
/* One incompressible solve: v_out = shift^-1 FFT^-1 [ M(k) shift FFT W ] / N with
 * M = (I - g g^T/|g|^2)/(1 - dt nu L/2) (mode 0 identity), as kernelUpdateVIncompressible. */
int orc_solve_incompressible(orc_ctx* c, const double* W, double* v);
/* One Stokes solve: v = P W / (eta |g|^2), mode 0 zero (kernelUpdateVstokesLimit). */
int orc_solve_stokes(orc_ctx* c, const double* W, double* v);
/* projection onto divergence-free fields, mode 0 kept (projectionDivergenceFree) */
int orc_project(orc_ctx* c, const double* W, double* v);
/* Stencil terms: 0.5 dt (eta/rho) L v; advection (dt/4) D[(v+v')(v+v')]; noise divergence. */
int orc_laplacian_term(orc_ctx* c, const double* v, double* out);
int orc_advection_term(orc_ctx* c, const double* v, double* out);
int orc_noise_divergence(orc_ctx* c, const double* Z6, double fact1, double fact4, int trace_free,
                         double* out);
int orc_fluid_noise(uint64_t seed, int64_t step, int which, int N, double* Z6); /* which: 0 Sigma_1, 1 Sigma_2 */
/* Stokes-limit stress drawn in Fourier space (nset 1: Sigma_1, 2: Sigma_1 + Sigma_2): the real
 * fields sigma6 (6N), their divergence w3 (3N) with prefactors f1, f4, and the largest imaginary
 * part of the inverse transform of Sigma_hat */
int orc_kspace_stress(orc_ctx* c, int64_t step, int nset, double f1, double f4, double* sigma6, double* w3,
                      double* max_imag);
/* The part of the Stokes noise source that the solve keeps, drawn directly in Fourier space:
 * u = f4 |g| E (Z1 t1 + Z2 t2), the projected stress divergence's law (f1^2 = 2 f4^2); nset 1:
 * set 1, nset 2: sets 1 + 2.  u3: 3 x N real-space values (IDFT / N); *max_imag: the largest
 * imaginary part of the IDFT (zero for a Hermitian spectrum). */
int orc_kspace_noise(orc_ctx* c, int64_t step, int nset, double f4, double* u3, double* max_imag);
/* Initial positions used when the caller provides none (seeded uniform / simple cubic). */
int orc_initial_positions(const fhd_params* p, double* q);

#ifdef __cplusplus
}
#endif
#endif
