// pf_march.h -- device code shared by the plain and the batched step kernels:
// the per-strip "march" that advances one wave's strip of one row chunk by one
// explicit Euler step.  Both kernels run exactly this arithmetic, so a
// simulation gives the same bits whether it runs alone or inside a batch.
//
// One wave64 owns a strip of 64 columns (62 output columns + one halo column
// each side) and marches down a chunk of rows.  Each lane keeps a three-row
// window (y-1, y, y+1) of what the stencil needs -- phi_delta, phi_Laves and
// the fictitious gamma composition -- in registers, so each cell is loaded
// once per strip and its KKS solve is done once.  x-neighbours come from
// adjacent lanes through DPP wave shifts (no LDS round trip).  Loads run DEPTH
// rows ahead of the row being computed (a register ring), so a wave keeps
// several rows of HBM requests in flight; the row loop is unrolled by 6 so the
// window and the ring rotate by register renaming instead of moves.
//
// Boundary conditions are index maps (no ghost reads): no-flux clamps to the
// nearest interior cell, which is exactly the reference's ghost copy
// (src/discretization.cu:360-394) for every interior output; periodic wraps.
// Ghost cells of the outputs are written with the value of the cell they
// mirror.  Storage T is double or float; all arithmetic is FP64.
#pragma once

#include <hip/hip_runtime.h>

#include "pf_internal.h"
#include "pf_thermo.h"

namespace pf {

constexpr int WAVE = 64;
constexpr int STRIP = WAVE - 2;  // output columns per strip
constexpr int WPB = 4;           // waves (strips) per block: adjacent strips of one row chunk
constexpr int UNROLL = 6;        // multiple of the 3-row window and of every ring depth

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

// DPP wave shifts (GFX9 wave_shr:1 / wave_shl:1): lane i receives lane i-1 /
// lane i+1.  Lanes 0 and 63 receive 0; their results are never stored.
__device__ __forceinline__ double from_left(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
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
struct Raw {
  T cr, nb, pd, pl;
};

template <typename T>
__device__ __forceinline__ void derive(const Raw<T>& r, Point& v, Centre& c) {
  v.pd = (double)r.pd;
  v.pl = (double)r.pl;
  c.xCr = (double)r.cr;
  c.xNb = (double)r.nb;
  const double pD = p(v.pd);
  const double pL = p(v.pl);
  const double pG = 1.0 - pD - pL;
  kks_mu(c.xCr, c.xNb, pD, pG, pL, c.mu0, c.mu1);
  v.gCr = fict(0, 0, c.mu0, c.mu1);
  v.gNb = fict(0, 1, c.mu0, c.mu1);
}

__device__ __forceinline__ Point shift_left(const Point& a) {
  return Point{from_left(a.pd), from_left(a.pl), from_left(a.gCr), from_left(a.gNb)};
}
__device__ __forceinline__ Point shift_right(const Point& a) {
  return Point{from_right(a.pd), from_right(a.pl), from_right(a.gCr), from_right(a.gNb)};
}

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

// Laplacians of the four stencil quantities in the reference's mask order
// (src/discretization.cu:589-595): rows y-1, y, y+1; columns x-1, x, x+1.
template <bool NINE>
__device__ __forceinline__ void laplacians(const double* m, const Point& S, const Point& W,
                                           const Point& C, const Point& E, const Point& Nn,
                                           const Point& SW, const Point& SE, const Point& NW,
                                           const Point& NE, double lap[4]) {
  const double s[4] = {S.pd, S.pl, S.gCr, S.gNb};
  const double w[4] = {W.pd, W.pl, W.gCr, W.gNb};
  const double c[4] = {C.pd, C.pl, C.gCr, C.gNb};
  const double e[4] = {E.pd, E.pl, E.gCr, E.gNb};
  const double n[4] = {Nn.pd, Nn.pl, Nn.gCr, Nn.gNb};
#pragma unroll
  for (int f = 0; f < 4; f++) {
    double v;
    if (NINE) {
      const double sw[4] = {SW.pd, SW.pl, SW.gCr, SW.gNb};
      const double se[4] = {SE.pd, SE.pl, SE.gCr, SE.gNb};
      const double nw[4] = {NW.pd, NW.pl, NW.gCr, NW.gNb};
      const double ne[4] = {NE.pd, NE.pl, NE.gCr, NE.gNb};
      v = m[0] * sw[f];
      v = fma(m[1], s[f], v);
      v = fma(m[2], se[f], v);
      v = fma(m[3], w[f], v);
      v = fma(m[4], c[f], v);
      v = fma(m[5], e[f], v);
      v = fma(m[6], nw[f], v);
      v = fma(m[7], n[f], v);
      v = fma(m[8], ne[f], v);
    } else {
      v = m[1] * s[f];
      v = fma(m[3], w[f], v);
      v = fma(m[4], c[f], v);
      v = fma(m[5], e[f], v);
      v = fma(m[7], n[f], v);
    }
    lap[f] = v;
  }
}

template <typename T>
__device__ __forceinline__ void store4(T* oCr, T* oNb, T* opd, T* opl, size_t o, double a,
                                       double b, double c, double d) {
  oCr[o] = (T)a;
  oNb[o] = (T)b;
  opd[o] = (T)c;
  opl[o] = (T)d;
}

// Write the ghost-ring images of interior cell (gx, y) (0-based interior
// coordinates): no-flux mirrors edge cells onto the adjacent ghost cells,
// periodic onto the opposite side.  Corner ghosts get the corner cell's value.
template <typename T, int BC>
__device__ __forceinline__ void store_ghost_images(T* oCr, T* oNb, T* opd, T* opl, int gx, int y,
                                                   int Nx, int Ny, int pitch, double a, double b,
                                                   double c, double d) {
  const bool ex = (gx == 0) || (gx == Nx - 1);
  const bool ey = (y == 0) || (y == Ny - 1);
  if (!(ex || ey)) return;
  int cx[3], cy[3], ncx = 1, ncy = 1;
  cx[0] = 1 + gx;
  cy[0] = 1 + y;
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
      store4(oCr, oNb, opd, opl, (size_t)cy[j] * pitch + cx[i], a, b, c, d);
    }
}

// Advance strip `strip` of row chunk `chunk` (rows [chunk*rows, ...)) by one
// step, reading the old fields and writing the new ones.  Called by a whole
// wave; uses no barriers.
template <typename T, int BC, bool NINE, bool NUC, int DEPTH>
__device__ __forceinline__ void march(
    const T* __restrict__ iCr, const T* __restrict__ iNb, const T* __restrict__ ipd,
    const T* __restrict__ ipl, T* __restrict__ oCr, T* __restrict__ oNb, T* __restrict__ opd,
    T* __restrict__ opl, int Nx, int Ny, int pitch, int strip, int chunk, int rows,
    const StepParams& P, const NucParams& N, const Scratch& S) {
  static_assert(UNROLL % DEPTH == 0, "ring depth must divide the unroll factor");
  const int lane = threadIdx.x & (WAVE - 1);

  const int gx = strip * STRIP + lane - 1;  // interior column owned by this lane
  const int px = 1 + bc_map<BC>(gx, Nx);    // padded column it reads
  const bool out_lane = (lane >= 1) && (lane <= STRIP) && (gx < Nx);

  const int y0 = chunk * rows;
  const int y1 = min(y0 + rows, Ny);
  if (y0 >= Ny) return;

  auto row_base = [&](int y) -> size_t {
    return (size_t)(1 + bc_map<BC>(y, Ny)) * (size_t)pitch + (size_t)px;
  };
  auto load = [&](Raw<T>& r, int y) {
    const size_t a = row_base(y);
    r.cr = iCr[a];
    r.nb = iNb[a];
    r.pd = ipd[a];
    r.pl = ipl[a];
  };

  // window slots (row y0 + k - 1 + 3j lives in slot k % 3 pattern below)
  Point win[3];
  Centre cen[3];
  Point winW[3], winE[3];  // DPP-shifted neighbours (nine-point masks only)
  Raw<T> ring[DEPTH];

  {
    Raw<T> r;
    load(r, y0 - 1);
    derive<T>(r, win[2], cen[2]);
    load(r, y0);
    derive<T>(r, win[0], cen[0]);
  }
#pragma unroll
  for (int j = 0; j < DEPTH; j++) load(ring[j], min(y0 + 1 + j, y1));
  if (NINE) {
    winW[2] = shift_left(win[2]);
    winE[2] = shift_right(win[2]);
    winW[0] = shift_left(win[0]);
    winE[0] = shift_right(win[0]);
  }

  const double* m = P.m;
  for (int yb = y0;; yb += UNROLL) {
#pragma unroll
    for (int k = 0; k < UNROLL; k++) {
      const int y = yb + k;
      if (y >= y1) return;
      const int ps = (k + 2) % 3, cs = k % 3, ns = (k + 1) % 3;
      const int rs = k % DEPTH;
      derive<T>(ring[rs], win[ns], cen[ns]);
      load(ring[rs], min(y + 1 + DEPTH, y1));

      Point W, E;
      if (NINE) {
        winW[ns] = shift_left(win[ns]);
        winE[ns] = shift_right(win[ns]);
        W = winW[cs];
        E = winE[cs];
      } else {
        W = shift_left(win[cs]);
        E = shift_right(win[cs]);
      }
      double lap[4];
      laplacians<NINE>(m, win[ps], W, win[cs], E, win[ns], winW[ps], winE[ps], winW[ns],
                       winE[ns], lap);

      const Point& cur = win[cs];
      const Centre& cc = cen[cs];
      // Cahn-Hilliard (src/discretization.cu:859-863)
      double nCr = cc.xCr + P.dt * (P.DCr0 * lap[2] + P.DCr1 * lap[3]);
      double nNb = cc.xNb + P.dt * (P.DNb0 * lap[2] + P.DNb1 * lap[3]);
      // Allen-Cahn (src/discretization.cu:888-942) with the closed-form
      // grand-potential difference at the solved potential
      const double fd = cur.pd, fl = cur.pl;
      const double dOd = dOmega_del(cc.mu0, cc.mu1);
      const double dOl = dOmega_lav(cc.mu0, cc.mu1);
      const double dFd = -pPrime(fd) * dOd + 2.0 * P.w0 * fd * (fd - 1.0) * (2.0 * fd - 1.0) +
                         2.0 * P.alpha * fd * fl * fl - P.k0 * lap[0];
      const double dFl = -pPrime(fl) * dOl + 2.0 * P.w1 * fl * (fl - 1.0) * (2.0 * fl - 1.0) +
                         2.0 * P.alpha * fl * fd * fd - P.k1 * lap[1];
      double npd = fd - P.dt * P.L0 * dFd;
      double npl = fl - P.dt * P.L1 * dFl;

      const int py = 1 + y;
      if (NUC && out_lane) {
        // decide on the stored (storage-rounded) values, as a separate pass would
        nCr = (double)(T)nCr;
        nNb = (double)(T)nNb;
        npd = (double)(T)npd;
        npl = (double)(T)npl;
        nucleate_cell(N, (uint32_t)(py * pitch + 1 + gx), nCr, nNb, npd, npl, S);
      }

      if (out_lane) {
        const int qx = 1 + gx;
        store4(oCr, oNb, opd, opl, (size_t)py * pitch + qx, nCr, nNb, npd, npl);
        store_ghost_images<T, BC>(oCr, oNb, opd, opl, gx, y, Nx, Ny, pitch, nCr, nNb, npd, npl);
      }
    }
  }
}

}  // namespace pf
