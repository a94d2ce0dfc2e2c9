// pf_aux.hip -- the kernels around the step: standalone nucleation pass,
// deferred embedding, deterministic compensated reductions, visualisation
// fields and pointwise evaluation (used by the parity tests).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "nucleation.cuh"
#include "parabola625.cuh"
#include "pf_internal.h"
#include "pf_march.h"
#include "pf_thermo.h"

namespace pf {

// ------------------------------------------------------------ error state
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

// ------------------------------------------------------------ scratch
hipError_t scratch_alloc(Scratch& s) {
  s.event_capacity = EVENT_CAPACITY;
  hipError_t e = hipMalloc((void**)&s.counters, 4 * sizeof(uint32_t));
  if (e != hipSuccess) return e;
  e = hipMemset(s.counters, 0, 4 * sizeof(uint32_t));
  if (e != hipSuccess) return e;
  e = hipMalloc((void**)&s.events, (size_t)s.event_capacity * sizeof(Event));
  if (e != hipSuccess) return e;
  // per-block partials of up to 5 compensated sums, then 8 doubles of final sums
  e = hipMalloc((void**)&s.partials, ((size_t)REDUCE_BLOCKS * 10 + 8) * sizeof(double));
  return e;
}
void scratch_free(Scratch& s) {
  if (s.counters) (void)hipFree(s.counters);
  if (s.events) (void)hipFree(s.events);
  if (s.partials) (void)hipFree(s.partials);
  s = Scratch{};
}

NucParams make_nuc(double sigma_del, double sigma_lav, double unit_a, double dx, double dy,
                   double dz, double dt, uint64_t seed, uint64_t call, uint32_t instance,
                   int nx, int ny) {
  // src/discretization.cu:1093-1095,1105,1206 (the kernel receives dx*nx, dy*ny)
  NucParams N{};
  N.sigma_del = sigma_del;
  N.sigma_lav = sigma_lav;
  N.dx_arg = dx * nx;
  const double dy_arg = dy * ny;
  N.dV = N.dx_arg * dy_arg * dz;
  N.Vatom = 0.25 * unit_a * unit_a * unit_a;
  N.n_gam = N.dV / N.Vatom;
  N.dt = dt;
  N.R_scan = 1.75e-9 / N.dx_arg;
  N.seed = seed;
  N.call = call;
  N.instance = instance;
  return N;
}

StepParams make_step(double dt, const double* DCr, const double* DNb, const double* kappa,
                     const double* omega, const double* Lmob, double alpha, const double* mask,
                     const double* xoff) {
  StepParams P{};
  // q = dt D (xe_g + A_g mu): coefficients of mu0, mu1 and the constant
  const double* rows[2] = {DCr, DNb};
  double* q[2] = {P.qCr, P.qNb};
  for (int i = 0; i < 2; i++) {
    q[i][0] = dt * (rows[i][0] * A[0][0] + rows[i][1] * A[0][1]);
    q[i][1] = dt * (rows[i][0] * A[0][1] + rows[i][1] * A[0][2]);
    q[i][2] = dt * (rows[i][0] * XE[0][0] + rows[i][1] * XE[0][1]);
  }
  // 30 dt L (mu.B.mu - mu.e) as a quadratic form in mu0, mu1
  const double* B[2] = {B_DEL, B_LAV};
  const double* E[2] = {E_DEL, E_LAV};
  for (int i = 0; i < 2; i++) {
    const double dtL = dt * Lmob[i];
    const double a = 30.0 * dtL;
    P.g[i][0] = a * B[i][0];
    P.g[i][1] = a * (2.0 * B[i][1]);
    P.g[i][2] = a * B[i][2];
    P.g[i][3] = -a * E[i][0];
    P.g[i][4] = -a * E[i][1];
    P.w[i] = 2.0 * omega[i] * dtL;
    P.c[i] = 2.0 * alpha * dtL;
    P.k[i] = kappa[i] * dtL;
  }
  // fold the mask's common factor into what the Laplacians multiply
  const double scale = symmetric_five_point(mask) ? mask[1] : 1.0;
  for (int i = 0; i < 3; i++) {
    P.qCr[i] *= scale;
    P.qNb[i] *= scale;
  }
  for (int i = 0; i < 2; i++) P.k[i] *= scale;
  for (int i = 0; i < 9; i++) P.m[i] = mask[i] / scale;
  for (int c = 0; c < 2; c++) {
    P.xoff[c] = xoff ? xoff[c] : 0.0;
    P.xe_g[c] = XE[0][c] - P.xoff[c];
  }
  return P;
}

// ------------------------------------------------------------ nucleation
// Decision pass over interior cells (src/discretization.cu:1079-1177).
// Scan radius R < 1 (every realistic domain): each cell reads only itself,
// so centre-only embeds are applied in place.  Otherwise every event is
// deferred so that all decisions see the entry fields.
template <typename T>
__global__ __launch_bounds__(256) void nucleation_kernel(T* __restrict__ xCr, T* __restrict__ xNb,
                                                         T* __restrict__ pd, T* __restrict__ pl,
                                                         int nx, int ny, NucParams N, Scratch S) {
  const int x = 1 + blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = 1 + blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x > nx - 2 || y > ny - 2) return;
  const long n = (long)nx * ny;
  const double R = N.R_scan;
  const bool in_place = R < 1.0;
  double phi_pre = 0.;
  for (int i = (int)(-R); i < R; i++) {
    for (int j = (int)(-R); j < R; j++) {
      const long idn = (long)nx * (y + j) + (x + i);
      if (idn >= 0 && idn < n && i * i + j * j < R * R) {
        const double v = p((double)pd[idn]) + p((double)pl[idn]);
        phi_pre = fmax(phi_pre, v);
      }
    }
  }
  if (!(phi_pre < 1e-10)) return;
  const uint32_t idx = (uint32_t)(nx * y + x);
  const double cCr = (double)xCr[idx] + N.xoff[0], cNb = (double)xNb[idx] + N.xoff[1];
  double u1, u2;
  nuc_uniforms(N.seed, N.call, N.instance, idx, u1, u2);
  const double dG = g(0, cCr, cNb);
  const double DCrCr = D(0, 0, cCr, cNb);
  const double DNbNb = D(1, 1, cCr, cNb);
  for (int ph = 0; ph < 2; ph++) {
    double Rstar, Pn;
    nucleation_probability(cCr, cNb, dG, DCrCr, DNbNb, ph == 0 ? N.sigma_del : N.sigma_lav,
                           N.Vatom, N.n_gam, N.dV, N.dt, Rstar, Pn);
    if (!(Rstar > 0.)) continue;
    const double u = ph == 0 ? u1 : u2;
    if (!(Pn - u > 0)) continue;
    const double Rp = Rstar / N.dx_arg;
    if (in_place && Rp < 1.0) {
      xCr[idx] = (T)(XE[1 + ph][0] - N.xoff[0]);
      xNb[idx] = (T)(XE[1 + ph][1] - N.xoff[1]);
      if (ph == 0)
        pd[idx] = (T)1.0;
      else
        pl[idx] = (T)1.0;
    } else {
      const uint32_t slot = atomicAdd(&S.counters[0], 1u);
      atomicAdd(&S.counters[2], 1u);
      if (slot < S.event_capacity)
        S.events[slot] = Event{idx, (uint32_t)ph, cCr, cNb, Rp};
      else
        atomicOr(&S.counters[1], 1u);
    }
  }
}

// Deferred embeds, one thread, deterministic order: centre-only events first,
// then wide ones, each by (cell, phase).  embed_OPC_kernel semantics
// (src/discretization.cu:1046-1077).  Clears the event count afterwards.
// Events carry absolute compositions; the arrays hold x - (off0, off1).
template <typename T>
__global__ void apply_events_kernel(T* xCr, T* xNb, T* pd, T* pl, int nx, int ny, Scratch S,
                                    double off0, double off1) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t n = S.counters[0];
  if (n > S.event_capacity) n = S.event_capacity;
  Event* ev = S.events;
  // insertion sort by (wide, cell, phase); events are rare
  for (uint32_t i = 1; i < n; i++) {
    const Event e = ev[i];
    const uint64_t ke = ((uint64_t)(e.R >= 1.0) << 40) | ((uint64_t)e.cell << 1) | e.phase;
    uint32_t j = i;
    while (j > 0) {
      const Event f = ev[j - 1];
      const uint64_t kf = ((uint64_t)(f.R >= 1.0) << 40) | ((uint64_t)f.cell << 1) | f.phase;
      if (kf <= ke) break;
      ev[j] = f;
      j--;
    }
    ev[j] = e;
  }
  const long ncell = (long)nx * ny;
  for (uint32_t k = 0; k < n; k++) {
    const Event e = ev[k];
    const int x = (int)(e.cell % (uint32_t)nx), y = (int)(e.cell / (uint32_t)nx);
    const int ph = (int)e.phase;
    const double xeCr = XE[1 + ph][0], xeNb = XE[1 + ph][1];
    T* phi = ph == 0 ? pd : pl;
    const double R = e.R;
    const double RdCr = R * sqrt((xeCr - e.xCr) / (e.xCr - XE[0][0]));
    const double RdNb = R * sqrt((xeNb - e.xNb) / (e.xNb - XE[0][1]));
    for (int i = (int)(-R); i < R; i++) {
      for (int j = (int)(-R); j < R; j++) {
        const long idn = (long)nx * (y + j) + (x + i);
        const double r = sqrt((double)(i * i + j * j));
        if (idn >= 0 && idn < ncell) {
          if (r <= R) {
            xCr[idn] = (T)(xeCr - off0);
            xNb[idn] = (T)(xeNb - off1);
            phi[idn] = (T)1.0;
          } else {
            if (r <= RdCr) xCr[idn] = (T)(XE[0][0] - off0);
            if (r <= RdNb) xNb[idn] = (T)(XE[0][1] - off1);
          }
        }
      }
    }
  }
  S.counters[0] = 0;
}

template <typename T>
hipError_t launch_nucleation(int nx, int ny, T* xCr, T* xNb, T* pd, T* pl, const NucParams& N,
                             const Scratch& S, hipStream_t st) {
  dim3 grid((nx - 2 + 15) / 16, (ny - 2 + 15) / 16);
  hipLaunchKernelGGL(nucleation_kernel<T>, grid, dim3(256), 0, st, xCr, xNb, pd, pl, nx, ny, N,
                     S);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(apply_events_kernel<T>, dim3(1), dim3(64), 0, st, xCr, xNb, pd, pl, nx, ny,
                     S, N.xoff[0], N.xoff[1]);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_apply_events(int nx, int ny, T* xCr, T* xNb, T* pd, T* pl, const Scratch& S,
                               const double* xoff, hipStream_t st) {
  hipLaunchKernelGGL(apply_events_kernel<T>, dim3(1), dim3(64), 0, st, xCr, xNb, pd, pl, nx, ny,
                     S, xoff ? xoff[0] : 0.0, xoff ? xoff[1] : 0.0);
  return hipGetLastError();
}
template hipError_t launch_apply_events<double>(int, int, double*, double*, double*, double*,
                                                const Scratch&, const double*, hipStream_t);
template hipError_t launch_apply_events<float>(int, int, float*, float*, float*, float*,
                                               const Scratch&, const double*, hipStream_t);

template hipError_t launch_nucleation<double>(int, int, double*, double*, double*, double*,
                                              const NucParams&, const Scratch&, hipStream_t);
template hipError_t launch_nucleation<float>(int, int, float*, float*, float*, float*,
                                             const NucParams&, const Scratch&, hipStream_t);

// ------------------------------------------------------------ reductions
// Neumaier-compensated FP64 sums in a fixed order (grid-stride over cells with
// a fixed grid, then a fixed tree), replacing the nondeterministic OpenMP
// 'critical' sum of summarize_fields (src/alloy625.cpp:811-863).
struct Kahan {
  double s, c;
};
__device__ __forceinline__ void kadd(Kahan& k, double x) {
  const double t = k.s + x;
  if (fabs(k.s) >= fabs(x))
    k.c += (k.s - t) + x;
  else
    k.c += (x - t) + k.s;
  k.s = t;
}
__device__ __forceinline__ void kmerge(Kahan& a, const Kahan& b) {
  kadd(a, b.s);
  a.c += b.c;
}

template <typename T>
__global__ __launch_bounds__(256) void summarize_partial_kernel(const T* __restrict__ xCr,
                                                                const T* __restrict__ xNb,
                                                                const T* __restrict__ pd,
                                                                const T* __restrict__ pl, int nx,
                                                                int ny, double* partials) {
  const long Nx = nx - 2, Ny = ny - 2, N = Nx * Ny;
  Kahan acc[5] = {};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long)gridDim.x * 256) {
    const long y = i / Nx, x = i - y * Nx;
    const size_t o = (size_t)(y + 1) * nx + (x + 1);
    const double d = p((double)pd[o]), l = p((double)pl[o]);
    kadd(acc[0], (double)xCr[o]);
    kadd(acc[1], (double)xNb[o]);
    kadd(acc[2], 1.0 - d - l);
    kadd(acc[3], d);
    kadd(acc[4], l);
  }
  __shared__ Kahan sh[5][256];
  for (int k = 0; k < 5; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 5; k++) kmerge(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; k++) {
      partials[(size_t)blockIdx.x * 10 + 2 * k] = sh[k][0].s;
      partials[(size_t)blockIdx.x * 10 + 2 * k + 1] = sh[k][0].c;
    }
}

// Final pass over the per-block partials of `nq` quantities (stride 2*nq
// doubles per block): block k merges quantity k, 64 threads over a fixed
// partition, then a fixed tree.  out[k] = sum * scale (+ add0 / add1 for
// quantities 0 / 1).
__global__ __launch_bounds__(64) void reduce_final_kernel(const double* partials, int nblocks,
                                                          int nq, double scale, double* out,
                                                          double add0, double add1) {
  const int k = blockIdx.x, t = threadIdx.x;
  Kahan a{0.0, 0.0};
  for (int b = t; b < nblocks; b += 64) {
    Kahan v{partials[(size_t)b * 2 * nq + 2 * k], partials[(size_t)b * 2 * nq + 2 * k + 1]};
    kmerge(a, v);
  }
  __shared__ Kahan sh[64];
  sh[t] = a;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (t < w) kmerge(sh[t], sh[t + w]);
    __syncthreads();
  }
  if (t == 0) out[k] = (sh[0].s + sh[0].c) * scale + (k == 0 ? add0 : (k == 1 ? add1 : 0.0));
}

// The mean compositions are the offset plus the mean of the stored values.
template <typename T>
hipError_t launch_summarize(int nx, int ny, const T* xCr, const T* xNb, const T* pd,
                            const T* pl, const double* xoff, double* out5, const Scratch& S,
                            hipStream_t st) {
  const long N = (long)(nx - 2) * (ny - 2);
  int blocks = (int)((N + 255) / 256);
  if (blocks > REDUCE_BLOCKS) blocks = REDUCE_BLOCKS;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(summarize_partial_kernel<T>, dim3(blocks), dim3(256), 0, st, xCr, xNb, pd,
                     pl, nx, ny, S.partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reduce_final_kernel, dim3(5), dim3(64), 0, st, S.partials, blocks, 5,
                     1.0 / (double)N, out5, xoff ? xoff[0] : 0.0, xoff ? xoff[1] : 0.0);
  return hipGetLastError();
}
template hipError_t launch_summarize<double>(int, int, const double*, const double*,
                                             const double*, const double*, const double*,
                                             double*, const Scratch&, hipStream_t);
template hipError_t launch_summarize<float>(int, int, const float*, const float*, const float*,
                                            const float*, const double*, double*,
                                            const Scratch&, hipStream_t);

// ------------------------------------------------------------ free energy
// src/alloy625.cpp:752-784 (gibbs), :786-806 (maskedgradient) and :865-919
// (summarize_energy) per interior cell, times dV = dx dy:
//   chem = sum_a p_a g_a(x_a) + sum_i omega_i phi_i^2 (1 - phi_i)^2
//          + 2 Omega phi_del^2 phi_lav^2
//   grad = sum_i kappa_i |grad phi_i|^2   (central differences)
// With the KKS potential mu of the cell, x_a - xe_a = A_a mu and
// g_a(x_a) = 1/2 mu.A_a.mu, so sum_a p_a g_a(x_a) = 1/2 mu.(M mu) = 1/2 mu.r,
// r = x - sum_a p_a xe_a.  Neighbours of the gradients go through the
// boundary map (no-flux: the MMSP Neumann ghost copy; periodic: wrap).
template <typename T, int BC>
__global__ __launch_bounds__(256) void energy_partial_kernel(
    const T* __restrict__ xCr, const T* __restrict__ xNb, const T* __restrict__ pd,
    const T* __restrict__ pl, int nx, int ny, EnergyParams E, double* __restrict__ chem_out,
    double* __restrict__ grad_out, double* partials) {
  const int Nx = nx - 2, Ny = ny - 2;
  const long N = (long)Nx * Ny;
  Kahan acc[2] = {};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long)gridDim.x * 256) {
    const int y = (int)(i / Nx), x = (int)(i - (long)y * Nx);
    const size_t o = (size_t)(y + 1) * nx + (x + 1);
    const double cr = (double)xCr[o] + E.xoff[0], nb = (double)xNb[o] + E.xoff[1];
    const double fd = (double)pd[o], fl = (double)pl[o];
    const double pD = p(fd), pL = p(fl), pG = 1.0 - pD - pL;
    double mu0, mu1;
    kks_mu(cr, nb, pD, pG, pL, mu0, mu1);
    const double r0 = cr - pG * XE[0][0] - pD * XE[1][0] - pL * XE[2][0];
    const double r1 = nb - pG * XE[0][1] - pD * XE[1][1] - pL * XE[2][1];
    double g = 0.5 * (mu0 * r0 + mu1 * r1);
    g += E.omega[0] * (fd * fd) * ((1.0 - fd) * (1.0 - fd));
    g += E.omega[1] * (fl * fl) * ((1.0 - fl) * (1.0 - fl));
    g += 2.0 * E.Omega * (fd * fd) * (fl * fl);
    const size_t xl = (size_t)(y + 1) * nx + 1 + bc_map<BC>(x - 1, Nx);
    const size_t xh = (size_t)(y + 1) * nx + 1 + bc_map<BC>(x + 1, Nx);
    const size_t yl = (size_t)(1 + bc_map<BC>(y - 1, Ny)) * nx + (x + 1);
    const size_t yh = (size_t)(1 + bc_map<BC>(y + 1, Ny)) * nx + (x + 1);
    const T* phi[2] = {pd, pl};
    double ge = 0.0;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const double gx = ((double)phi[k][xh] - (double)phi[k][xl]) / (2.0 * E.dx);
      const double gy = ((double)phi[k][yh] - (double)phi[k][yl]) / (2.0 * E.dy);
      ge += E.kappa[k] * (gx * gx + gy * gy);
    }
    const double c = g * E.dV, gg = ge * E.dV;
    if (chem_out) chem_out[o] = c;
    if (grad_out) grad_out[o] = gg;
    kadd(acc[0], c);
    kadd(acc[1], gg);
  }
  __shared__ Kahan sh[2][256];
  for (int k = 0; k < 2; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 2; k++) kmerge(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 2; k++) {
      partials[(size_t)blockIdx.x * 4 + 2 * k] = sh[k][0].s;
      partials[(size_t)blockIdx.x * 4 + 2 * k + 1] = sh[k][0].c;
    }
}

__global__ void energy_total_kernel(const double* sums, double* out3) {
  if (threadIdx.x == 0) {
    out3[1] = sums[0];
    out3[2] = sums[1];
    out3[0] = sums[0] + sums[1];
  }
}

template <typename T>
hipError_t launch_energy(int nx, int ny, int bc, const EnergyParams& E, const T* xCr,
                         const T* xNb, const T* pd, const T* pl, double* chem, double* grad,
                         double* out3, const Scratch& S, hipStream_t st) {
  const long N = (long)(nx - 2) * (ny - 2);
  int blocks = (int)((N + 255) / 256);
  if (blocks > REDUCE_BLOCKS) blocks = REDUCE_BLOCKS;
  if (blocks < 1) blocks = 1;
  if (bc == 0)
    hipLaunchKernelGGL((energy_partial_kernel<T, 0>), dim3(blocks), dim3(256), 0, st, xCr, xNb,
                       pd, pl, nx, ny, E, chem, grad, S.partials);
  else
    hipLaunchKernelGGL((energy_partial_kernel<T, 1>), dim3(blocks), dim3(256), 0, st, xCr, xNb,
                       pd, pl, nx, ny, E, chem, grad, S.partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the two sums land in out3[1..2] via a scratch slot after the partials
  double* sums = S.partials + (size_t)REDUCE_BLOCKS * 10;
  hipLaunchKernelGGL(reduce_final_kernel, dim3(2), dim3(64), 0, st, S.partials, blocks, 2, 1.0,
                     sums, 0.0, 0.0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(energy_total_kernel, dim3(1), dim3(64), 0, st, sums, out3);
  return hipGetLastError();
}
template hipError_t launch_energy<double>(int, int, int, const EnergyParams&, const double*,
                                          const double*, const double*, const double*, double*,
                                          double*, double*, const Scratch&, hipStream_t);
template hipError_t launch_energy<float>(int, int, int, const EnergyParams&, const float*,
                                         const float*, const float*, const float*, double*,
                                         double*, double*, const Scratch&, hipStream_t);

// ------------------------------------------------------------ dataviz
// src/discretization.cu:1211-1229 over every cell (ghosts included)
__global__ void dataviz_kernel(const double* xCr, const double* xNb, const double* pd,
                               const double* pl, double* cNi, double* phi, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  cNi[i] = 1.0 - xCr[i] - xNb[i];
  phi[i] = p(pd[i]) + p(pl[i]);
}
hipError_t launch_dataviz(int nx, int ny, const double* xCr, const double* xNb, const double* pd,
                          const double* pl, double* conc_Ni, double* phi, hipStream_t st) {
  const long n = (long)nx * ny;
  hipLaunchKernelGGL(dataviz_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, xCr,
                     xNb, pd, pl, conc_Ni, phi, n);
  return hipGetLastError();
}

// ------------------------------------------------------------ pointwise eval
// what = 0: in x[n]                -> out {p, pPrime}[n][2]
// what = 1: in {xCr,xNb,pD,pG,pL}  -> out fict {gCr,gNb,dCr,dNb,lCr,lNb}[n][6]
// what = 2: in {xCr,xNb}           -> out {g_gam,g_del,g_lav, dg x6, D_CrCr,D_CrNb,D_NbCr,D_NbNb}[n][13]
// what = 3: in nucleation args[10] -> out {Rstar, P}[n][2]
__global__ void thermo_eval_kernel(int what, int n, const double* in, double* out) {
  // evaluates through the reference-named __device__ API (parabola625.cuh,
  // nucleation.cuh) so that API is what the parity tests exercise
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (what == 0) {
    const double x = in[i];
    out[2 * i] = d_p(x);
    out[2 * i + 1] = d_pPrime(x);
  } else if (what == 1) {
    const double* a = in + 5 * (size_t)i;
    double* o = out + 6 * (size_t)i;
    o[0] = d_fict_gam_Cr(a[0], a[1], a[2], a[3], a[4]);
    o[1] = d_fict_gam_Nb(a[0], a[1], a[2], a[3], a[4]);
    o[2] = d_fict_del_Cr(a[0], a[1], a[2], a[3], a[4]);
    o[3] = d_fict_del_Nb(a[0], a[1], a[2], a[3], a[4]);
    o[4] = d_fict_lav_Cr(a[0], a[1], a[2], a[3], a[4]);
    o[5] = d_fict_lav_Nb(a[0], a[1], a[2], a[3], a[4]);
  } else if (what == 2) {
    const double xCr = in[2 * (size_t)i], xNb = in[2 * (size_t)i + 1];
    double* o = out + 13 * (size_t)i;
    o[0] = d_g_gam(xCr, xNb);
    o[1] = d_g_del(xCr, xNb);
    o[2] = d_g_lav(xCr, xNb);
    o[3] = d_dg_gam_dxCr(xCr, xNb);
    o[4] = d_dg_gam_dxNb(xCr, xNb);
    o[5] = d_dg_del_dxCr(xCr, xNb);
    o[6] = d_dg_del_dxNb(xCr, xNb);
    o[7] = d_dg_lav_dxCr(xCr, xNb);
    o[8] = d_dg_lav_dxNb(xCr, xNb);
    o[9] = d_D_CrCr(xCr, xNb);
    o[10] = d_D_CrNb(xCr, xNb);
    o[11] = d_D_NbCr(xCr, xNb);
    o[12] = d_D_NbNb(xCr, xNb);
  } else if (what == 3) {
    const double* a = in + 10 * (size_t)i;
    double R, P;
    d_nucleation_probability_sphere(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9],
                                    &R, &P);
    out[2 * (size_t)i] = R;
    out[2 * (size_t)i + 1] = P;
  }
}
hipError_t launch_thermo_eval(int what, int n, const double* in, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(thermo_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, st, what, n, in,
                     out);
  return hipGetLastError();
}

// in: n x {ctr0..3, key0, key1}; out: n x 4 words
__global__ void philox_eval_kernel(int n, const uint32_t* in, uint32_t* out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t* a = in + 6 * (size_t)i;
  const u32x4 r = philox4x32_10(u32x4{a[0], a[1], a[2], a[3]}, a[4], a[5]);
  out[4 * (size_t)i] = r.x;
  out[4 * (size_t)i + 1] = r.y;
  out[4 * (size_t)i + 2] = r.z;
  out[4 * (size_t)i + 3] = r.w;
}
hipError_t launch_philox_eval(int n, const uint32_t* in, uint32_t* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(philox_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, in, out);
  return hipGetLastError();
}

}  // namespace pf
