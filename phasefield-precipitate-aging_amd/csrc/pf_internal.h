// pf_internal.h -- types shared by the kernels, the C-ABI and the C++ drop-in.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace pf {

// Kinetic and discretisation parameters of one step (by value into kernels),
// with dt, the mobilities and the constant prefactors folded in on the host.
// The gamma fictitious composition x_g = xe_g + A_g mu is linear in the
// potential mu, and so is the Laplacian, so the Cahn-Hilliard update
// x' = x + dt D lap(x_g) (src/discretization.cu:859-863) is
//   x_Cr' = x_Cr + lap(q_Cr),  q_Cr = qCr[0] mu0 + qCr[1] mu1 + qCr[2]
// with q_Cr = dt (D_CrCr x_g,Cr + D_CrNb x_g,Nb), and likewise for Nb.  The
// Allen-Cahn update (src/discretization.cu:888-942) is
//   phi' = phi + u (G(mu) u + w (2 phi - 1)) - c phi phi_o^2 + k lap(phi)
// with u = phi (1 - phi), p'(phi) = 30 u^2, phi (phi - 1)(2 phi - 1) = -u (2 phi - 1)
// and G = 30 dt L dOmega = g[0] mu0^2 + g[1] mu0 mu1 + g[2] mu1^2 + g[3] mu0 + g[4] mu1.
//
// A symmetric five-point mask (zero corners, m[1] = m[7] != 0, m[3] = m[5];
// every five-point mask the reference builds) is stored divided by m[1], and
// q and k carry the factor m[1] instead, so that a Laplacian is
// (S + N) + m[3] (W + E) + m[4] C: four operations instead of five.  Any
// other mask is stored as given (scale 1) and runs the general nine-term sum.
//
// Stored compositions may be offset: an array holds x - xoff (FP32 storage
// with xoff near the mean composition keeps 4-5 more bits of the small
// per-step increments; xoff = 0 otherwise).  The step is offset-invariant but
// for the residual x - xe_gamma of the KKS solve, so the kernels work on stored
// values throughout with xe_g = xe_gamma - xoff; only the nucleation test
// (which evaluates g and D at x) adds xoff back.
struct StepParams {
  double qCr[3], qNb[3];  // scale * dt * (D row) . (A_gamma mu + xe_gamma)
  double g[2][5];         // 30 dt L * grand-potential difference  (delta, Laves)
  double w[2];            // 2 omega dt L
  double c[2];            // 2 alpha dt L
  double k[2];            // scale * kappa dt L
  double m[9];            // 3x3 Laplacian mask / scale, row-major (y-1, y, y+1)
  double xe_g[2];         // xe_gamma - xoff (Cr, Nb)
  double xoff[2];         // composition offset of the stored arrays (Cr, Nb)
};

// The masks the folded five-point Laplacian takes (as given by the caller).
inline bool symmetric_five_point(const double* m) {
  return m[0] == 0.0 && m[2] == 0.0 && m[6] == 0.0 && m[8] == 0.0 && m[1] == m[7] &&
         m[3] == m[5] && m[1] != 0.0 && m[1] - m[1] == 0.0;  // last: finite
}
// ... and such a mask once folded (StepParams::m): the kernels pick the
// five-point arithmetic exactly when this holds.
inline bool folded_five_point(const double* m) {
  return m[0] == 0.0 && m[2] == 0.0 && m[6] == 0.0 && m[8] == 0.0 && m[1] == 1.0 &&
         m[7] == 1.0 && m[3] == m[5];
}

// Nucleation pass parameters, derived on the host from the reference's
// device_nucleation arguments (src/discretization.cu:1179-1209).
struct NucParams {
  double sigma_del, sigma_lav;
  double Vatom, n_gam, dV, dt;  // dt = nuc_interval * step dt
  double dx_arg;                // meshres * nx, as the reference passes it
  double R_scan;                // 1.75e-9 / dx_arg
  uint64_t seed, call;
  uint32_t instance;
  double xoff[2];  // composition offset of the stored arrays (StepParams::xoff)
};

// A nucleation event whose embedding cannot be done in place by the
// deciding thread (embed or scan radius of a cell or more).
struct Event {
  uint32_t cell;   // padded index nx*y + x
  uint32_t phase;  // 0 = delta, 1 = Laves
  double xCr, xNb, R;
};

// Device scratch owned by a context: nucleation events and reduction partials.
// Free-energy functional parameters (src/alloy625.cpp:865-919 with
// src/parameters.h:36-48): mesh spacing, cell volume dx*dy, gradient and
// well coefficients per phase, three-phase coefficient.
struct EnergyParams {
  double dx, dy, dV;
  double kappa[2], omega[2];
  double Omega;
  double xoff[2];  // composition offset of the stored arrays (StepParams::xoff)
};

struct Scratch {
  uint32_t* counters;  // [0] event count, [1] overflow flag, [2] wide-event total
  Event* events;
  uint32_t event_capacity;
  double* partials;  // reduction partials: REDUCE_BLOCKS * 10 doubles
};

constexpr int REDUCE_BLOCKS = 1024;
constexpr uint32_t EVENT_CAPACITY = 1u << 16;

// Tile configuration of the marching step kernel.
struct TileConfig {
  int rows = 0;   // output rows per block (0 = heuristic)
  int remap = 1;  // XCD-aware block order
  int depth = 3;  // rows of loads in flight ahead of the computed row (2, 3 or 6)
  int rows2 = 0;  // the same for the two-step kernel
  int depth2 = 3;
  int rows3 = 0;  // and for the three-step kernel
  int depth3 = 3;
  int rows4 = 0;  // and the four-step kernel
  int depth4 = 3;
};

hipError_t scratch_alloc(Scratch& s);
void scratch_free(Scratch& s);

// ---- launchers (pf_step.hip); T = double or float storage, FP64 compute
template <typename T>
hipError_t launch_step(int nx, int ny, int bc, const StepParams& P, const T* iCr, const T* iNb,
                       const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd, T* opl,
                       const NucParams* nuc, const Scratch& S, const TileConfig& tc,
                       hipStream_t st);
// k = 2, 3 or 4 steps per launch (folded symmetric five-point mask, no nucleation); grids
// need Nx, Ny >= 3 (k = 4: >= 4)
template <typename T>
hipError_t launch_stepk(int k, int nx, int ny, int bc, const StepParams& P, const T* iCr,
                        const T* iNb, const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd,
                        T* opl, const TileConfig& tc, hipStream_t st);
template <typename T>
hipError_t launch_nucleation(int nx, int ny, T* xCr, T* xNb, T* pd, T* pl, const NucParams& N,
                             const Scratch& S, hipStream_t st);
template <typename T>
hipError_t launch_apply_events(int nx, int ny, T* xCr, T* xNb, T* pd, T* pl, const Scratch& S,
                               const double* xoff, hipStream_t st);
template <typename T>
hipError_t launch_energy(int nx, int ny, int bc, const EnergyParams& E, const T* xCr,
                         const T* xNb, const T* pd, const T* pl, double* chem, double* grad,
                         double* out3, const Scratch& S, hipStream_t st);
template <typename T>
hipError_t launch_summarize(int nx, int ny, const T* xCr, const T* xNb, const T* pd,
                            const T* pl, const double* xoff, double* out5, const Scratch& S,
                            hipStream_t st);
hipError_t launch_dataviz(int nx, int ny, const double* xCr, const double* xNb, const double* pd,
                          const double* pl, double* conc_Ni, double* phi, hipStream_t st);
hipError_t launch_thermo_eval(int what, int n, const double* in, double* out, hipStream_t st);
hipError_t launch_philox_eval(int n, const uint32_t* in, uint32_t* out, hipStream_t st);

int default_rows(int nx, int ny);
int default_rows_k(int k, int nx, int ny, int storage_bytes);

NucParams make_nuc(double sigma_del, double sigma_lav, double unit_a, double dx, double dy,
                   double dz, double dt, uint64_t seed, uint64_t call, uint32_t instance,
                   int nx, int ny);
// xoff: composition offset of the stored arrays (null = none)
StepParams make_step(double dt, const double* DCr, const double* DNb, const double* kappa,
                     const double* omega, const double* Lmob, double alpha, const double* mask,
                     const double* xoff = nullptr);

void set_error(const std::string& msg);
const char* last_error();

}  // namespace pf
