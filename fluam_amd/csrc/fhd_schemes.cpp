// fhd_schemes.cpp -- fluam-style scheme entry points on top of the context API: read the global
// data.main parameters, run numstepsRelaxation + numsteps steps, and write fluam's output files
// (<out>.data.main, .seed, .particles / .velocityParticles / .velocityParticlesI,
// .fluidFinalConfiguration, .temperature, .time; .structureFactor when asked for).
//
// Restates schemeIncompressible (src/schemeIncompressible.cu:25-55), schemeQuasiNeutrallyBuoyant
// (src/schemeQuasiNeutrallyBuoyant.cu:27-96), schemeStokesLimit (src/schemeStokesLimit.cu:27-98)
// and the dispatch of src/main.cpp:28-112.
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/fhd.h"

extern "C" void fhd_internal_set_error(const char* e);

namespace {

std::string g_datamain_path;
std::string g_err;

std::string out_name(const fhd_params& p, const char* ext) { return std::string(p.outputname) + ext; }

bool fail(const std::string& e) {
  g_err = e;
  fhd_internal_set_error(e.c_str());
  return false;
}

// saveFluidFinalConfiguration (src/saveFluidFinalConfiguration.cu:30-75)
void save_fluid_final(const fhd_params& p, const std::vector<double>& v) {
  std::ofstream f(out_name(p, ".fluidFinalConfiguration"));
  f.precision(15);
  const long long N = (long long)p.mx * p.my * p.mz;
  f << "Number of cells " << N << "\n";
  f << "mx my mz\n" << p.mx << " " << p.my << " " << p.mz << "\n";
  f << "lx ly lz\n" << p.lx << " " << p.ly << " " << p.lz << "\n";
  f << "Density " << p.densfluid << "\n";
  f << "Pressure parameters 0 0 0\n";
  f << "Thermostat 0\n";
  f << "Temperature " << p.temperature << "\n";
  f << "Cells properties \nDensity \nVelocity\n";
  for (long long i = 0; i < N; ++i) {
    f << p.densfluid << "\n";
    f << v[i] << " " << v[N + i] << " " << v[2 * N + i] << "\n";
  }
}

// Fluid temperature summary written as <out>.temperature by the incompressible and inertial
// schemes: temperatureFunction (src/temperature.cu:38-212) and the fluid lines of
// temperatureBoundary (src/temperatureBoundary.cu).  Per sample, t_a = (1/N) sum_i rho dV
// (v_a,i - <v_a>)^2 (uniform density, so the face density average is rho) and t = mean of t_a.
struct TemperatureSummary {
  double t_sum = 0, dev2 = 0, tx = 0, ty = 0, tz = 0;
  long long n = 0;
  void sample(const fhd_params& p, const std::vector<double>& v) {
    const long long N = (long long)p.mx * p.my * p.mz;
    const double dV = p.lx * p.ly * p.lz / (double)N, rho = p.densfluid;
    double t[3];
    for (int a = 0; a < 3; ++a) {
      double mean = 0.0;
      for (long long i = 0; i < N; ++i) mean += v[a * N + i];
      mean /= (double)N;
      double s = 0.0;
      for (long long i = 0; i < N; ++i) {
        const double d = v[a * N + i] - mean;
        s += rho * dV * d * d;
      }
      t[a] = s / (double)N;
    }
    const double tt = (t[0] + t[1] + t[2]) / 3.0;
    dev2 += (tt - p.temperature) * (tt - p.temperature);
    t_sum += tt;
    tx += t[0];
    ty += t[1];
    tz += t[2];
    ++n;
  }
  void write(const fhd_params& p, const std::vector<double>& v) const {
    if (n == 0) return;
    const long long N = (long long)p.mx * p.my * p.mz;
    const double dV = p.lx * p.ly * p.lz / (double)N, rho = p.densfluid, T = p.temperature;
    std::ofstream f(out_name(p, ".temperature"));
    f.precision(15);
    const double tm = t_sum / (double)n;
    f << "TEMPERATURE FLUID " << tm << " " << (tm - T) / T << " " << std::sqrt(dev2 / (double)n) / T << "  " << n
      << "\n";
    f << "tx                " << tx / (double)n << "\n";
    f << "ty                " << ty / (double)n << "\n";
    f << "tz                " << tz / (double)n << "\n";
    double P[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a)
      for (long long i = 0; i < N; ++i) P[a] += rho * dV * v[a * N + i];
    f << "PX " << P[0] << "\nPY " << P[1] << "\nPZ " << P[2] << "\n\n";
    f << "MASS FLUID    " << rho * dV * (double)N << "\n\n";
  }
};

struct ParticleFiles {
  std::ofstream pos, vel, velI;
  void open(const fhd_params& p) {
    pos.open(out_name(p, ".particles"));
    vel.open(out_name(p, ".velocityParticles"));
    pos.precision(15);
    vel.precision(15);
    pos << "#NUMBER PARTICLES " << p.np << "\n";
    vel << "#NUMBER PARTICLES " << p.np << "\n";
    if (p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) {
      velI.open(out_name(p, ".velocityParticlesI"));
      velI.precision(15);
      velI << "#NUMBER PARTICLES " << p.np << "\n";
    }
  }
  // saveParticles(1, step) (src/saveParticles.cu:55-73)
  void sample(const fhd_params& p, long long step, const std::vector<double>& q, const std::vector<double>& u) {
    const int np = p.np;
    pos << step * p.dt << "\n";
    vel << step * p.dt << "\n";
    if (velI.is_open()) velI << step * p.dt << "\n";
    for (int i = 0; i < np; ++i) {
      pos << q[i] << " " << q[np + i] << " " << q[2 * np + i] << "\n";
      vel << u[i] << " " << u[np + i] << " " << u[2 * np + i] << "\n";
      if (velI.is_open()) velI << u[i] << " " << u[np + i] << " " << u[2 * np + i] << "\n";
    }
  }
};

// <out>.structureFactor: the equal-time structure factor sampled every samplefreq steps (the
// quantity fluam's HydroGrid hook accumulates, src/hydroAnalysisIncompressible.cu:35-92), over
// the half spectrum: signed mode numbers n, wavenumbers k = 2 pi n / L, then S_x, S_y, S_z.
void save_structure_factor(const fhd_params& p, fhd_ctx* c) {
  const int P = p.mx / 2 + 1;
  const size_t per = (size_t)p.mz * p.my * P;
  std::vector<double> sk(3 * per);
  long long ns = 0;
  if (!fhd_sk_read(c, sk.data(), &ns)) return;
  std::ofstream f(out_name(p, ".structureFactor"));
  f.precision(10);
  f << "# equal-time structure factor S_a(k) = rho dV / kT <|v_a(k)|^2> / N, samples " << ns << "\n";
  f << "# nx ny nz kx ky kz S_x S_y S_z\n";
  const double two_pi = 6.283185307179586;
  auto sgn = [](int n, int m) { return n <= m / 2 ? n : n - m; };
  for (int z = 0; z < p.mz; ++z)
    for (int y = 0; y < p.my; ++y)
      for (int x = 0; x < P; ++x) {
        const int nx = x, ny = sgn(y, p.my), nz = sgn(z, p.mz);
        const size_t i = ((size_t)z * p.my + y) * P + x;
        f << nx << " " << ny << " " << nz << " " << two_pi * nx / p.lx << " " << two_pi * ny / p.ly << " "
          << two_pi * nz / p.lz << " " << sk[i] << " " << sk[per + i] << " " << sk[2 * per + i] << "\n";
      }
}

bool run_scheme(int scheme, const char* label) {
  fhd_params p;
  if (!fhd_get_params(&p)) return fail("no data.main loaded");
  if (p.scheme != scheme) p.scheme = scheme;
  if (p.samplefreq <= 0) return fail("samplefreq must be positive (fluam divides the step count by it)");
  // echo the input (src/loadDataMain.cu:671-676) and the seed (src/saveSeed.cpp:23-31)
  if (!g_datamain_path.empty()) {
    std::ifstream in(g_datamain_path);
    std::ofstream out(out_name(p, ".data.main"));
    out << in.rdbuf() << "\n";
  }
  {
    std::ofstream f(out_name(p, ".seed"));
    f << p.seed << "\n";
  }
  fhd_ctx* c = fhd_create(&p);
  if (!c) return fail(fhd_last_error());
  time_t t0 = time(nullptr);
  ParticleFiles pf;
  const bool parts = p.np > 0;
  if (parts) pf.open(p);
  std::vector<double> q(3 * (size_t)p.np), u(3 * (size_t)p.np);
  const bool temp = scheme != FHD_SCHEME_STOKES_LIMIT;  // fluam's Stokes scheme writes none
  TemperatureSummary ts;
  std::vector<double> vs;
  if (temp) vs.resize(3 * (size_t)p.mx * p.my * p.mz);
  long long step = -p.numsteps_relaxation;
  bool ok = true;
  while (step < p.numsteps) {
    if (!fhd_step(c, 1)) { ok = fail(fhd_ctx_error(c)); break; }
    ++step;
    if (step > 0 && step % p.samplefreq == 0) {
      if (!fhd_sync(c)) { ok = fail(fhd_ctx_error(c)); break; }
      std::cout << label << " " << step << std::endl;
      if (parts) {
        fhd_get_particles(c, q.data());
        fhd_get_particle_velocities(c, u.data());
        pf.sample(p, step, q, u);
      }
      if (temp) {
        fhd_get_velocity(c, vs.data());
        ts.sample(p, vs);
      }
      if (p.structure_factor && !fhd_sk_sample(c)) { ok = fail(fhd_ctx_error(c)); break; }
    }
  }
  if (ok && !fhd_sync(c)) ok = fail(fhd_ctx_error(c));
  if (ok) {
    std::vector<double> v(3 * (size_t)p.mx * p.my * p.mz);
    fhd_get_velocity(c, v.data());
    save_fluid_final(p, v);
    if (temp) ts.write(p, v);
    if (p.structure_factor) save_structure_factor(p, c);
    long long t = (long long)(time(nullptr) - t0);
    std::cout << "\nTIME RUN_GPU                       " << t << "\n\n";
    std::ofstream f(out_name(p, ".time"));
    f << "TIME RUN GPU     " << t << "\n";
  }
  fhd_destroy(c);
  return ok;
}

}  // namespace

extern "C" {

int fhd_scheme_incompressible(void) { return run_scheme(FHD_SCHEME_INCOMPRESSIBLE, "INCOMPRESSIBLE") ? 1 : 0; }
int fhd_scheme_quasi_neutrally_buoyant(void) {
  return run_scheme(FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT, "INCOMPRESSIBLE BOUNDARY") ? 1 : 0;
}
int fhd_scheme_stokes_limit(void) { return run_scheme(FHD_SCHEME_STOKES_LIMIT, "Stokes Limit ") ? 1 : 0; }

const char* fhd_scheme_error(void) { return g_err.c_str(); }

int fhd_main(const char* datamain_path) {
  const char* path = datamain_path ? datamain_path : "data.main";
  if (!fhd_load_datamain(path)) {
    g_err = fhd_last_error();
    std::cerr << g_err << std::endl;
    return 0;
  }
  g_datamain_path = path;
  fhd_params p;
  fhd_get_params(&p);
  int ok = 0;
  if (p.scheme == FHD_SCHEME_INCOMPRESSIBLE) ok = fhd_scheme_incompressible();
  else if (p.scheme == FHD_SCHEME_QUASI_NEUTRALLY_BUOYANT) ok = fhd_scheme_quasi_neutrally_buoyant();
  else if (p.scheme == FHD_SCHEME_STOKES_LIMIT) ok = fhd_scheme_stokes_limit();
  if (!ok) std::cerr << g_err << std::endl;
  return ok;
}

}  // extern "C"
