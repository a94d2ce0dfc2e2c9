// pf_step.hip -- fused KKS phase-field step for gfx950 (MI355X).
//
// One launch replaces the reference's per-step sequence (src/main.cpp:289-305):
//   boundaries -> fictitious compositions -> boundaries -> 4 Laplacians ->
//   boundaries -> Cahn-Hilliard || Allen-Cahn   (29 launches, 24 host syncs)
// and, on nucleation steps, the nucleation pass (src/discretization.cu:1079-1177).
// The per-strip arithmetic is pf::march (pf_march.h).  This kernel maps one
// wave to one (strip, row chunk), four adjacent strips per block, with blocks
// renumbered so that each of the 8 XCDs (each with its own L2) works on a
// contiguous band of the domain and halo re-reads stay inside one L2.
#include <hip/hip_runtime.h>

#include "pf_march.h"

namespace pf {

template <typename T, int BC, bool NINE, bool NUC, int DEPTH>
__global__ __launch_bounds__(WAVE * WPB) void step_kernel(
    const T* __restrict__ iCr, const T* __restrict__ iNb, const T* __restrict__ ipd,
    const T* __restrict__ ipl, T* __restrict__ oCr, T* __restrict__ oNb, T* __restrict__ opd,
    T* __restrict__ opl, int Nx, int Ny, int pitch, int strips, int chunks, int rows, int remap,
    StepParams P, NucParams N, Scratch S) {
  const int nwaves = strips * chunks;
  const int nblocks = (nwaves + WPB - 1) / WPB;
  const int b = (remap ? xcd_remap(blockIdx.x, nblocks) : (int)blockIdx.x) * WPB +
                __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));  // wave-uniform
  if (b >= nwaves) return;  // whole wave exits; no barriers in this kernel
  march<T, BC, NINE, NUC, DEPTH>(iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, Nx, Ny, pitch,
                                 b % strips, b / strips, rows, P, N, S);
}

int default_rows(int nx, int ny) {
  // enough blocks to fill 256 CUs several times over, long enough chunks to
  // amortise the two halo rows
  const int Nx = nx - 2, Ny = ny - 2;
  const int strips = (Nx + STRIP - 1) / STRIP;
  int rows = 16;
  while (rows > 4 && (long)strips * ((Ny + rows - 1) / rows) < 16384) rows >>= 1;
  return rows;
}

template <typename T, int BC, bool NINE, bool NUC, int DEPTH>
static hipError_t launch_step_t(int nx, int ny, const StepParams& P, const T* iCr, const T* iNb,
                                const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd, T* opl,
                                const NucParams& N, const Scratch& S, const TileConfig& tc,
                                hipStream_t st) {
  const int Nx = nx - 2, Ny = ny - 2;
  const int rows = tc.rows > 0 ? tc.rows : default_rows(nx, ny);
  const int strips = (Nx + STRIP - 1) / STRIP;
  const int chunks = (Ny + rows - 1) / rows;
  const int blocks = (strips * chunks + WPB - 1) / WPB;
  hipLaunchKernelGGL((step_kernel<T, BC, NINE, NUC, DEPTH>), dim3(blocks), dim3(WAVE * WPB),
                     0, st, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, Nx, Ny, nx, strips, chunks,
                     rows, tc.remap, P, N, S);
  return hipGetLastError();
}

// The plain five-point step without nucleation (every step but one in ~900)
// is instantiated for prefetch depths 2, 3 and 6; the rarer variants use 3.
template <typename T, int BC>
static hipError_t launch_plain(int nx, int ny, const StepParams& P, const T* iCr, const T* iNb,
                               const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd, T* opl,
                               const NucParams& N, const Scratch& S, const TileConfig& tc,
                               hipStream_t st) {
  if (tc.depth == 2)
    return launch_step_t<T, BC, false, false, 2>(nx, ny, P, iCr, iNb, ipd, ipl, oCr, oNb, opd,
                                                 opl, N, S, tc, st);
  if (tc.depth == 6)
    return launch_step_t<T, BC, false, false, 6>(nx, ny, P, iCr, iNb, ipd, ipl, oCr, oNb, opd,
                                                 opl, N, S, tc, st);
  return launch_step_t<T, BC, false, false, 3>(nx, ny, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl,
                                               N, S, tc, st);
}

template <typename T>
hipError_t launch_step(int nx, int ny, int bc, const StepParams& P, const T* iCr, const T* iNb,
                       const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd, T* opl,
                       const NucParams* nuc, const Scratch& S, const TileConfig& tc,
                       hipStream_t st) {
  const bool nine = !folded_five_point(P.m);  // the general nine-term Laplacian
  NucParams N{};
  if (nuc) N = *nuc;
#define PF_LAUNCH(BCV, NI, NU)                                                                  \
  return launch_step_t<T, BCV, NI, NU, 3>(nx, ny, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, N, \
                                          S, tc, st)
  if (bc == 0) {
    if (nine) {
      if (nuc) PF_LAUNCH(0, true, true);
      PF_LAUNCH(0, true, false);
    }
    if (nuc) PF_LAUNCH(0, false, true);
    return launch_plain<T, 0>(nx, ny, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, N, S, tc, st);
  } else {
    if (nine) {
      if (nuc) PF_LAUNCH(1, true, true);
      PF_LAUNCH(1, true, false);
    }
    if (nuc) PF_LAUNCH(1, false, true);
    return launch_plain<T, 1>(nx, ny, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, N, S, tc, st);
  }
#undef PF_LAUNCH
}

// ---- K steps per launch (marchk, K = 2, 3 or 4) ----

template <typename T, int BC, int K, int DEPTH>
__global__ __launch_bounds__(WAVE * WPB) void stepk_kernel(
    const T* __restrict__ iCr, const T* __restrict__ iNb, const T* __restrict__ ipd,
    const T* __restrict__ ipl, T* __restrict__ oCr, T* __restrict__ oNb, T* __restrict__ opd,
    T* __restrict__ opl, int Nx, int Ny, int pitch, int strips, int chunks, int rows, int remap,
    StepParams P) {
  const int nwaves = strips * chunks;
  const int nblocks = (nwaves + WPB - 1) / WPB;
  const int b = (remap ? xcd_remap(blockIdx.x, nblocks) : (int)blockIdx.x) * WPB +
                __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));  // wave-uniform
  if (b >= nwaves) return;
  marchk<T, BC, K, DEPTH>(iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, Nx, Ny, pitch, b % strips,
                          b / strips, rows, P);
}

// Four steps: the rows behind of the four levels' windows live in LDS (16 KiB
// per wave, 64 KiB per block of WPB = 4 waves), which caps the kernel at two
// blocks, so two waves per SIMD, per CU; the register budget is set to match.
template <typename T, int BC, int DEPTH>
__global__ __launch_bounds__(WAVE * WPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
void step4_kernel(const T* __restrict__ iCr, const T* __restrict__ iNb,
                  const T* __restrict__ ipd, const T* __restrict__ ipl, T* __restrict__ oCr,
                  T* __restrict__ oNb, T* __restrict__ opd, T* __restrict__ opl, int Nx, int Ny,
                  int pitch, int strips, int chunks, int rows, int remap, StepParams P) {
  constexpr int K = 4;
  constexpr int PER_WAVE = K * 2 * 4 * WAVE;  // levels x parity x values x lanes
  __shared__ double win_lds[WPB * PER_WAVE];
  const int nwaves = strips * chunks;
  const int nblocks = (nwaves + WPB - 1) / WPB;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));  // wave-uniform
  const int b = (remap ? xcd_remap(blockIdx.x, nblocks) : (int)blockIdx.x) * WPB + w;
  if (b >= nwaves) return;
  marchk<T, BC, K, DEPTH, WAVE - 2 * K, true>(iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, Nx, Ny,
                                              pitch, b % strips, b / strips, rows, P,
                                              (LdsPtr)win_lds + w * PER_WAVE);
}

// Chunk height of the four-step kernel from a model of its launch on
// MI355X: 256 CUs x 4 SIMDs hold 2048 waves at two waves per SIMD.  A wave's
// time is proportional to its rows plus the 3 rows (K - 1) its levels
// recompute per chunk; a SIMD running two waves delivers 1.3x one wave's
// throughput, so a full round of resident waves costs 2/1.3 wave-times and a
// last partial round of at most one wave per SIMD costs one.  The height
// that minimises (rows + 3) x (rounds) is the measured best, or within 0.7%
// of it, from 1024^2 to 8192^2 (e.g. 10 rows at 1024^2, 38 at 2048^2, 75 at
// 4096^2, 119 at 8192^2).
static int rows4_model(int strips, int Ny) {
  constexpr long RESIDENT = 2048;
  constexpr double PAIR = 2.0 / 1.3, HALO = 3.0;
  int best = 1;
  double best_cost = 1e300;
  for (int rows = 1; rows <= min(Ny, 128); rows++) {
    const long waves = (long)strips * ((Ny + rows - 1) / rows);
    const long full = waves / RESIDENT, rem = waves % RESIDENT;
    const double tail = rem == 0 ? 0.0 : (rem <= RESIDENT / 2 ? 1.0 : PAIR);
    const double cost = (rows + HALO) * (PAIR * (double)full + tail);
    if (cost < best_cost) {
      best_cost = cost;
      best = rows;
    }
  }
  return best;
}

int default_rows_k(int k, int nx, int ny, int storage_bytes) {
  // The intermediate levels recompute (K-1)K rows per chunk in all, so chunks
  // are longer than the one-step kernel's while the grid still fills the chip.
  // Two steps: FP64 storage is near the HBM bound and prefers more waves in
  // flight (16 rows at 4096^2); FP32 storage is bound by arithmetic and
  // prefers less recomputation (32).  Three steps are bound by arithmetic in
  // both and take 48 rows at 4096^2.  Four steps: rows4_model.
  const int Nx = nx - 2, Ny = ny - 2;
  const int sw = WAVE - 2 * k;
  const int strips = (Nx + sw - 1) / sw;
  if (k == 4) return rows4_model(strips, Ny);
  int rows = k == 3 ? 48 : (storage_bytes == 8 ? 16 : 32);
  const long min_waves = k == 3 ? 4096 : 8192;
  while (rows > 6 && (long)strips * ((Ny + rows - 1) / rows) < min_waves) rows >>= 1;
  return rows;
}

template <typename T, int BC, int K, int DEPTH>
static hipError_t launch_stepk_t(int nx, int ny, const StepParams& P, const T* iCr,
                                 const T* iNb, const T* ipd, const T* ipl, T* oCr, T* oNb,
                                 T* opd, T* opl, int rows_cfg, int remap, hipStream_t st) {
  const int Nx = nx - 2, Ny = ny - 2;
  const int rows = rows_cfg > 0 ? rows_cfg : default_rows_k(K, nx, ny, (int)sizeof(T));
  const int sw = WAVE - 2 * K;
  const int strips = (Nx + sw - 1) / sw;
  const int chunks = (Ny + rows - 1) / rows;
  const int blocks = (strips * chunks + WPB - 1) / WPB;
  if constexpr (K == 4)
    hipLaunchKernelGGL((step4_kernel<T, BC, DEPTH>), dim3(blocks), dim3(WAVE * WPB), 0, st, iCr,
                       iNb, ipd, ipl, oCr, oNb, opd, opl, Nx, Ny, nx, strips, chunks, rows, remap,
                       P);
  else
    hipLaunchKernelGGL((stepk_kernel<T, BC, K, DEPTH>), dim3(blocks), dim3(WAVE * WPB), 0, st,
                       iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, Nx, Ny, nx, strips, chunks, rows,
                       remap, P);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_stepk(int k, int nx, int ny, int bc, const StepParams& P, const T* iCr,
                        const T* iNb, const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd,
                        T* opl, const TileConfig& tc, hipStream_t st) {
  const int rows = k == 4 ? tc.rows4 : k == 3 ? tc.rows3 : tc.rows2;
  const int depth = k == 4 ? tc.depth4 : k == 3 ? tc.depth3 : tc.depth2;
#define PF_LAUNCHK(BCV, KV, D)                                                             \
  return launch_stepk_t<T, BCV, KV, D>(nx, ny, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, \
                                       rows, tc.remap, st)
#define PF_LAUNCHK_D(BCV, KV)          \
  do {                                 \
    if (depth == 2) PF_LAUNCHK(BCV, KV, 2); \
    if (depth == 6) PF_LAUNCHK(BCV, KV, 6); \
    PF_LAUNCHK(BCV, KV, 3);            \
  } while (0)
  if (k == 4) {
    if (bc == 0) PF_LAUNCHK_D(0, 4);
    PF_LAUNCHK_D(1, 4);
  }
  if (k == 3) {
    if (bc == 0) PF_LAUNCHK_D(0, 3);
    PF_LAUNCHK_D(1, 3);
  }
  if (bc == 0) PF_LAUNCHK_D(0, 2);
  PF_LAUNCHK_D(1, 2);
#undef PF_LAUNCHK_D
#undef PF_LAUNCHK
}

template hipError_t launch_stepk<double>(int, int, int, int, const StepParams&, const double*,
                                         const double*, const double*, const double*, double*,
                                         double*, double*, double*, const TileConfig&,
                                         hipStream_t);
template hipError_t launch_stepk<float>(int, int, int, int, const StepParams&, const float*,
                                        const float*, const float*, const float*, float*, float*,
                                        float*, float*, const TileConfig&, hipStream_t);

template hipError_t launch_step<double>(int, int, int, const StepParams&, const double*,
                                        const double*, const double*, const double*, double*,
                                        double*, double*, double*, const NucParams*,
                                        const Scratch&, const TileConfig&, hipStream_t);
template hipError_t launch_step<float>(int, int, int, const StepParams&, const float*,
                                       const float*, const float*, const float*, float*, float*,
                                       float*, float*, const NucParams*, const Scratch&,
                                       const TileConfig&, hipStream_t);

}  // namespace pf
