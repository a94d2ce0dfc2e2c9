// pf_step.hip -- fused KKS phase-field step for gfx950 (MI355X).
//
// One launch replaces the reference's per-step sequence (src/main.cpp:289-305):
//   boundaries -> fictitious compositions -> boundaries -> 4 Laplacians ->
//   boundaries -> Cahn-Hilliard || Allen-Cahn   (29 launches, 24 host syncs)
// and, on nucleation steps, the nucleation pass (src/discretization.cu:1079-1177).
//
// Work decomposition ("marching strips"): one wave64 per block owns a strip of
// 64 columns (62 output columns + one halo column each side) and marches down
// a chunk of rows.  Each lane keeps a three-row window (y-1, y, y+1) of the
// quantities the 5-point stencil needs -- phi_delta, phi_Laves and the
// fictitious gamma compositions -- in registers, so every cell is loaded once
// per strip and its KKS solve is done once; x-neighbours come from adjacent
// lanes by shuffles.  Row y+2 is prefetched while row y is computed.  Blocks
// are renumbered so that each of the 8 XCDs (each with its own L2) works on a
// contiguous band of the domain, keeping halo re-reads inside one L2.
//
// Boundary conditions are index maps (no ghost reads): no-flux clamps to the
// nearest interior cell, which is exactly the reference's ghost copy
// (src/discretization.cu:360-394) for every interior output; periodic wraps.
// Ghost cells of the outputs are written with the value of the cell they
// mirror.
//
// Storage T is double (plain form) or float (precision form); all arithmetic
// is FP64.
#include <hip/hip_runtime.h>

#include "pf_internal.h"
#include "pf_thermo.h"

namespace pf {

constexpr int WAVE = 64;
constexpr int STRIP = WAVE - 2;  // output columns per strip

__device__ __forceinline__ int xcd_remap(int b, int nb) {
  // Hardware dispatches block b to XCD b % 8.  Give XCD x the contiguous
  // logical range [x*per + min(x, rem), ... + count_x).
  const int per = nb >> 3, rem = nb & 7;
  const int x = b & 7, k = b >> 3;
  return x * per + min(x, rem) + k;
}

// interior index map for i in [-1, n] (values beyond are clamped, unused)
template <int BC>
__device__ __forceinline__ int bc_map(int i, int n) {
  if (BC == 0) return i < 0 ? 0 : (i >= n ? n - 1 : i);
  if (i < 0) return i + n;
  if (i >= n) return (i - n) < n ? i - n : n - 1;
  return i;
}

struct Point {       // what the neighbours of a cell need
  double pd, pl;     // phase fields
  double gCr, gNb;   // fictitious gamma composition
};
struct Centre {      // extra state of a cell when it is the stencil centre
  double xCr, xNb;   // compositions
  double mu0, mu1;   // common diffusion potential
};

template <typename T>
__device__ __forceinline__ void derive(T rCr, T rNb, T rpd, T rpl, Point& v, Centre& c) {
  v.pd = (double)rpd;
  v.pl = (double)rpl;
  c.xCr = (double)rCr;
  c.xNb = (double)rNb;
  const double pD = p(v.pd);
  const double pL = p(v.pl);
  const double pG = 1.0 - pD - pL;
  kks_mu(c.xCr, c.xNb, pD, pG, pL, c.mu0, c.mu1);
  v.gCr = fict(0, 0, c.mu0, c.mu1);
  v.gNb = fict(0, 1, c.mu0, c.mu1);
}

__device__ __forceinline__ double shfl_up1(double v) { return __shfl_up(v, 1, WAVE); }
__device__ __forceinline__ double shfl_dn1(double v) { return __shfl_down(v, 1, WAVE); }

// Nucleation of one cell from its new values (centre-only scan; the caller
// guarantees R_scan < 1).  Mirrors src/discretization.cu:1104-1176 with the
// event order delta then Laves on the pre-embed composition.
__device__ __forceinline__ void nucleate_cell(const NucParams& N, uint32_t cell, double& xCr,
                                              double& xNb, double& pd, double& pl,
                                              const Scratch& S) {
  const double phi_pre = p(pd) + p(pl);
  if (!(phi_pre < 1e-10)) return;
  double u1, u2;
  nuc_uniforms(N.seed, N.call, N.instance, cell, u1, u2);
  const double cCr = xCr, cNb = xNb;
  const double dG = g(0, cCr, cNb);
  const double DCrCr = D(0, 0, cCr, cNb);
  const double DNbNb = D(1, 1, cCr, cNb);
#pragma unroll
  for (int ph = 0; ph < 2; ph++) {
    double Rstar, Pn;
    nucleation_probability(cCr, cNb, dG, DCrCr, DNbNb, ph == 0 ? N.sigma_del : N.sigma_lav,
                           N.Vatom, N.n_gam, N.dV, N.dt, Rstar, Pn);
    if (Rstar > 0.) {
      const double u = ph == 0 ? u1 : u2;
      if (Pn - u > 0) {
        const double R = Rstar / N.dx_arg;
        if (R < 1.0) {
          // embed loop collapses to the centre (r = 0 <= R)
          xCr = XE[1 + ph][0];
          xNb = XE[1 + ph][1];
          if (ph == 0)
            pd = 1.0;
          else
            pl = 1.0;
        } else {
          const uint32_t slot = atomicAdd(&S.counters[0], 1u);
          atomicAdd(&S.counters[2], 1u);
          if (slot < S.event_capacity) {
            S.events[slot] = Event{cell, (uint32_t)ph, cCr, cNb, R};
          } else {
            atomicOr(&S.counters[1], 1u);
          }
        }
      }
    }
  }
}

template <typename T, int BC, bool NINE, bool NUC>
__global__ __launch_bounds__(WAVE) void step_kernel(
    const T* __restrict__ iCr, const T* __restrict__ iNb, const T* __restrict__ ipd,
    const T* __restrict__ ipl, T* __restrict__ oCr, T* __restrict__ oNb, T* __restrict__ opd,
    T* __restrict__ opl, int Nx, int Ny, int pitch, int strips, int chunks, int rows, int remap,
    StepParams P, NucParams N, Scratch S) {
  const int lane = threadIdx.x;
  const int nb = strips * chunks;
  const int b = remap ? xcd_remap(blockIdx.x, nb) : (int)blockIdx.x;
  const int strip = b % strips;
  const int chunk = b / strips;

  const int gx = strip * STRIP + lane - 1;  // interior column owned by this lane
  const int px = 1 + bc_map<BC>(gx, Nx);    // padded column it reads
  const bool out_lane = (lane >= 1) && (lane <= STRIP) && (gx < Nx);

  const int y0 = chunk * rows;
  const int y1 = min(y0 + rows, Ny);
  if (y0 >= Ny) return;

  auto row_base = [&](int y) -> size_t {
    return (size_t)(1 + bc_map<BC>(y, Ny)) * (size_t)pitch + (size_t)px;
  };

  // window: rows y-1 (prev), y (cur), y+1 (next)
  Point prv, cur, nxt;
  Centre ccur, cnxt, cdump;
  {
    const size_t a = row_base(y0 - 1);
    derive<T>(iCr[a], iNb[a], ipd[a], ipl[a], prv, cdump);
    const size_t c = row_base(y0);
    derive<T>(iCr[c], iNb[c], ipd[c], ipl[c], cur, ccur);
  }
  T rCr, rNb, rpd, rpl;  // raw loads of the next row, in flight during compute
  {
    const size_t a = row_base(y0 + 1);
    rCr = iCr[a];
    rNb = iNb[a];
    rpd = ipd[a];
    rpl = ipl[a];
  }
  // diagonal neighbours for a nine-point mask
  Point prvW, prvE;
  if (NINE) {
    prvW.pd = shfl_up1(prv.pd); prvW.pl = shfl_up1(prv.pl);
    prvW.gCr = shfl_up1(prv.gCr); prvW.gNb = shfl_up1(prv.gNb);
    prvE.pd = shfl_dn1(prv.pd); prvE.pl = shfl_dn1(prv.pl);
    prvE.gCr = shfl_dn1(prv.gCr); prvE.gNb = shfl_dn1(prv.gNb);
  }

  const double* m = P.m;
  for (int y = y0; y < y1; y++) {
    derive<T>(rCr, rNb, rpd, rpl, nxt, cnxt);
    if (y + 1 < y1) {
      const size_t a = row_base(y + 2);
      rCr = iCr[a];
      rNb = iNb[a];
      rpd = ipd[a];
      rpl = ipl[a];
    }
    Point W, E;
    W.pd = shfl_up1(cur.pd); W.pl = shfl_up1(cur.pl);
    W.gCr = shfl_up1(cur.gCr); W.gNb = shfl_up1(cur.gNb);
    E.pd = shfl_dn1(cur.pd); E.pl = shfl_dn1(cur.pl);
    E.gCr = shfl_dn1(cur.gCr); E.gNb = shfl_dn1(cur.gNb);
    Point nW, nE;
    if (NINE) {
      nW.pd = shfl_up1(nxt.pd); nW.pl = shfl_up1(nxt.pl);
      nW.gCr = shfl_up1(nxt.gCr); nW.gNb = shfl_up1(nxt.gNb);
      nE.pd = shfl_dn1(nxt.pd); nE.pl = shfl_dn1(nxt.pl);
      nE.gCr = shfl_dn1(nxt.gCr); nE.gNb = shfl_dn1(nxt.gNb);
    }

    // Laplacians accumulated in the reference's mask order
    // (src/discretization.cu:589-595): rows y-1, y, y+1; columns x-1, x, x+1.
    double lap[4];
    {
      const double sv[4] = {prv.pd, prv.pl, prv.gCr, prv.gNb};
      const double wv[4] = {W.pd, W.pl, W.gCr, W.gNb};
      const double cv[4] = {cur.pd, cur.pl, cur.gCr, cur.gNb};
      const double ev[4] = {E.pd, E.pl, E.gCr, E.gNb};
      const double nv[4] = {nxt.pd, nxt.pl, nxt.gCr, nxt.gNb};
#pragma unroll
      for (int f = 0; f < 4; f++) {
        double v;
        if (NINE) {
          const double swv[4] = {prvW.pd, prvW.pl, prvW.gCr, prvW.gNb};
          const double sev[4] = {prvE.pd, prvE.pl, prvE.gCr, prvE.gNb};
          const double nwv[4] = {nW.pd, nW.pl, nW.gCr, nW.gNb};
          const double nev[4] = {nE.pd, nE.pl, nE.gCr, nE.gNb};
          v = m[0] * swv[f];
          v = fma(m[1], sv[f], v);
          v = fma(m[2], sev[f], v);
          v = fma(m[3], wv[f], v);
          v = fma(m[4], cv[f], v);
          v = fma(m[5], ev[f], v);
          v = fma(m[6], nwv[f], v);
          v = fma(m[7], nv[f], v);
          v = fma(m[8], nev[f], v);
        } else {
          v = m[1] * sv[f];
          v = fma(m[3], wv[f], v);
          v = fma(m[4], cv[f], v);
          v = fma(m[5], ev[f], v);
          v = fma(m[7], nv[f], v);
        }
        lap[f] = v;
      }
    }

    // Cahn-Hilliard (src/discretization.cu:859-863)
    double nCr = ccur.xCr + P.dt * (P.DCr0 * lap[2] + P.DCr1 * lap[3]);
    double nNb = ccur.xNb + P.dt * (P.DNb0 * lap[2] + P.DNb1 * lap[3]);

    // Allen-Cahn (src/discretization.cu:888-942) with the closed-form
    // grand-potential difference at the solved potential
    const double fd = cur.pd, fl = cur.pl;
    const double dOd = dOmega_del(ccur.mu0, ccur.mu1);
    const double dOl = dOmega_lav(ccur.mu0, ccur.mu1);
    const double dFd = -pPrime(fd) * dOd + 2.0 * P.w0 * fd * (fd - 1.0) * (2.0 * fd - 1.0) +
                       2.0 * P.alpha * fd * fl * fl - P.k0 * lap[0];
    const double dFl = -pPrime(fl) * dOl + 2.0 * P.w1 * fl * (fl - 1.0) * (2.0 * fl - 1.0) +
                       2.0 * P.alpha * fl * fd * fd - P.k1 * lap[1];
    double npd = fd - P.dt * P.L0 * dFd;
    double npl = fl - P.dt * P.L1 * dFl;

    const int py = 1 + y;
    if (NUC && out_lane) {
      nucleate_cell(N, (uint32_t)(py * pitch + 1 + gx), nCr, nNb, npd, npl, S);
    }

    if (out_lane) {
      const int qx = 1 + gx;
      const size_t o = (size_t)py * pitch + qx;
      oCr[o] = (T)nCr;
      oNb[o] = (T)nNb;
      opd[o] = (T)npd;
      opl[o] = (T)npl;
      // ghost images of edge cells
      const bool ex = (gx == 0) || (gx == Nx - 1);
      const bool ey = (y == 0) || (y == Ny - 1);
      if (ex || ey) {
        int cx[3], cy[3], ncx = 1, ncy = 1;
        cx[0] = qx;
        cy[0] = py;
        if (BC == 0) {
          if (gx == 0) cx[ncx++] = 0;
          if (gx == Nx - 1) cx[ncx++] = Nx + 1;
          if (y == 0) cy[ncy++] = 0;
          if (y == Ny - 1) cy[ncy++] = Ny + 1;
        } else {
          if (gx == Nx - 1) cx[ncx++] = 0;
          if (gx == 0) cx[ncx++] = Nx + 1;
          if (y == Ny - 1) cy[ncy++] = 0;
          if (y == 0) cy[ncy++] = Ny + 1;
        }
        for (int j = 0; j < ncy; j++)
          for (int i = 0; i < ncx; i++) {
            if (i == 0 && j == 0) continue;
            const size_t g2 = (size_t)cy[j] * pitch + cx[i];
            oCr[g2] = (T)nCr;
            oNb[g2] = (T)nNb;
            opd[g2] = (T)npd;
            opl[g2] = (T)npl;
          }
      }
    }

    if (NINE) {
      prvW = W;
      prvE = E;
    }
    prv = cur;
    cur = nxt;
    ccur = cnxt;
  }
}

int default_rows(int nx, int ny) {
  // enough blocks to fill 256 CUs several times over, long enough chunks to
  // amortise the two halo rows
  const int Nx = nx - 2, Ny = ny - 2;
  const int strips = (Nx + STRIP - 1) / STRIP;
  int rows = 64;
  while (rows > 8 && (long)strips * ((Ny + rows - 1) / rows) < 4096) rows >>= 1;
  return rows;
}

template <typename T, int BC, bool NINE, bool NUC>
static hipError_t launch_step_t(int nx, int ny, const StepParams& P, const T* iCr, const T* iNb,
                                const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd, T* opl,
                                const NucParams& N, const Scratch& S, const TileConfig& tc,
                                hipStream_t st) {
  const int Nx = nx - 2, Ny = ny - 2;
  const int rows = tc.rows > 0 ? tc.rows : default_rows(nx, ny);
  const int strips = (Nx + STRIP - 1) / STRIP;
  const int chunks = (Ny + rows - 1) / rows;
  hipLaunchKernelGGL((step_kernel<T, BC, NINE, NUC>), dim3(strips * chunks), dim3(WAVE), 0, st,
                     iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, Nx, Ny, nx, strips, chunks, rows,
                     tc.remap, P, N, S);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_step(int nx, int ny, int bc, const StepParams& P, const T* iCr, const T* iNb,
                       const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd, T* opl,
                       const NucParams* nuc, const Scratch& S, const TileConfig& tc,
                       hipStream_t st) {
  const bool nine = P.m[0] != 0.0 || P.m[2] != 0.0 || P.m[6] != 0.0 || P.m[8] != 0.0;
  NucParams N{};
  if (nuc) N = *nuc;
#define PF_LAUNCH(BCV, NI, NU)                                                               \
  return launch_step_t<T, BCV, NI, NU>(nx, ny, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, N, \
                                       S, tc, st)
  if (bc == 0) {
    if (nine) {
      if (nuc) PF_LAUNCH(0, true, true);
      PF_LAUNCH(0, true, false);
    }
    if (nuc) PF_LAUNCH(0, false, true);
    PF_LAUNCH(0, false, false);
  } else {
    if (nine) {
      if (nuc) PF_LAUNCH(1, true, true);
      PF_LAUNCH(1, true, false);
    }
    if (nuc) PF_LAUNCH(1, false, true);
    PF_LAUNCH(1, false, false);
  }
#undef PF_LAUNCH
}

template hipError_t launch_step<double>(int, int, int, const StepParams&, const double*,
                                        const double*, const double*, const double*, double*,
                                        double*, double*, double*, const NucParams*,
                                        const Scratch&, const TileConfig&, hipStream_t);
template hipError_t launch_step<float>(int, int, int, const StepParams&, const float*,
                                       const float*, const float*, const float*, float*, float*,
                                       float*, float*, const NucParams*, const Scratch&,
                                       const TileConfig&, hipStream_t);

}  // namespace pf
