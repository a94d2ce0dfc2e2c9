// pf_march.h -- device code shared by the plain and the batched step kernels:
// the per-strip "march" that advances one wave's strip of one row chunk by one
// explicit Euler step.  Both kernels run exactly this arithmetic, so a
// simulation gives the same bits whether it runs alone or inside a batch.
//
// One wave64 owns a strip of 64 columns (62 output columns + one halo column
// each side) and marches down a chunk of rows.  Each lane keeps a three-row
// window (y-1, y, y+1) of what the stencil needs -- phi_delta, phi_Laves and
// the composition potentials q = dt D x_gamma -- in registers, so each cell
// is loaded once per strip and its KKS solve is done once.  x-neighbours come
// from adjacent lanes through DPP wave shifts (no LDS round trip).  Loads run DEPTH
// rows ahead of the row being computed (a register ring), so a wave keeps
// several rows of HBM requests in flight; the row loop is unrolled by 6 so the
// window and the ring rotate by register renaming instead of moves.
//
// Boundary conditions are index maps (no ghost reads): no-flux reflects about
// the domain faces, which at the one ghost layer the one-step march reads is
// exactly the reference's ghost copy (src/discretization.cu:360-394); periodic
// wraps.
// Ghost cells of the outputs are written with the value of the cell they
// mirror.  Storage T is double or float; all arithmetic is FP64.
#pragma once

#include <hip/hip_runtime.h>

#include "pf_internal.h"
#include "pf_thermo.h"

namespace pf {

constexpr int WAVE = 64;
constexpr int STRIP = WAVE - 2;   // output columns per strip
constexpr int STRIP2 = WAVE - 4;  // output columns per strip of the two-step march
constexpr int STRIP3 = WAVE - 6;  // and of the three-step march (the batched form's strips)
constexpr int WPB = 4;           // waves (strips) per block: adjacent strips of one row chunk
constexpr int UNROLL = 6;        // multiple of the 3-row window and of every ring depth

// a pointer into LDS (address space 3), for the multi-step march's row windows
typedef double __attribute__((address_space(3)))* LdsPtr;

__device__ __forceinline__ int xcd_remap(int b, int nb) {
  // Hardware dispatches block b to XCD b % 8.  Give XCD x the contiguous
  // logical range [x*per + min(x, rem), ... + count_x).
  const int per = nb >> 3, rem = nb & 7;
  const int x = b & 7, k = b >> 3;
  return x * per + min(x, rem) + k;
}

// Interior index map.  No-flux (BC 0): even reflection about the domain
// faces, period 2n (-1 -> 0, -2 -> 1, n -> n-1, n+1 -> n-2, ...).  At the one
// ghost layer (i = -1, n) this is the reference's ghost copy
// (src/discretization.cu:360-394); deeper halo cells (the multi-step march)
// are the mirror images that make every intermediate level's halo equal,
// bit for bit, to the ghost copy of that level (see marchk_impl).  Any i maps
// into [0, n).  Periodic (BC 1): wrap, for i in [-n, 2n) (values beyond are
// clamped, unused).
template <int BC>
__device__ __forceinline__ int bc_map(int i, int n) {
  if (BC == 0) {
    const int p = 2 * n;
    int m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - 1 - m;
  }
  if (i < 0) return i + n;
  if (i >= n) return (i - n) < n ? i - n : n - 1;
  return i;
}

// DPP wave shifts (GFX9 wave_shr:1 / wave_shl:1): lane i receives lane i-1 /
// lane i+1.  With bound_ctrl set, lanes 0 and 63 (no source lane) receive 0;
// their results are never stored.
__device__ __forceinline__ double from_left(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

struct Point {       // what the neighbours of a cell need
  double pd, pl;     // phase fields
  double qCr, qNb;   // dt D x_gamma: the Cahn-Hilliard update is x' = x + lap(q)
};
struct Centre {      // extra state of a cell when it is the stencil centre
  double xCr, xNb;   // compositions as stored (x - StepParams::xoff)
  double mu0, mu1;   // common diffusion potential
};
template <typename T>
struct Raw {
  T cr, nb, pd, pl;
};

// Field access as a wave-uniform row byte offset (scalar registers) plus a
// 32-bit per-lane byte offset.  The compiler forms each address with one
// 64-bit multiply-add of the row index onto a hoisted per-lane base; forcing
// a scalar row base instead costs more in scalar-register spills than the
// add it saves.
template <typename T>
__device__ __forceinline__ T ld_at(const T* base, size_t row_bytes, uint32_t lane_bytes) {
  return *(const T*)((const char*)base + row_bytes + lane_bytes);
}
template <typename T>
__device__ __forceinline__ void st_at(T* base, size_t row_bytes, uint32_t lane_bytes, T v) {
  *(T*)((char*)base + row_bytes + lane_bytes) = v;
}
template <typename T>
__device__ __forceinline__ void load_row(Raw<T>& r, const T* iCr, const T* iNb, const T* ipd,
                                         const T* ipl, size_t row_bytes, uint32_t lane_bytes) {
  r.cr = ld_at(iCr, row_bytes, lane_bytes);
  r.nb = ld_at(iNb, row_bytes, lane_bytes);
  r.pd = ld_at(ipd, row_bytes, lane_bytes);
  r.pl = ld_at(ipl, row_bytes, lane_bytes);
}
template <typename T>
__device__ __forceinline__ void store_row(T* oCr, T* oNb, T* opd, T* opl, size_t row_bytes,
                                          uint32_t lane_bytes, double a, double b, double c,
                                          double d) {
  st_at(oCr, row_bytes, lane_bytes, (T)a);
  st_at(oNb, row_bytes, lane_bytes, (T)b);
  st_at(opd, row_bytes, lane_bytes, (T)c);
  st_at(opl, row_bytes, lane_bytes, (T)d);
}

// 1/d to within an ulp or so: hardware reciprocal and two Newton steps (the
// refinement an IEEE division does before its final rounding correction).
// Measured on 4M random doubles over 200 binades: the hardware reciprocal
// alone is good to 2^-24.4, one step leaves results up to 2.2e-15 relative
// off the correctly rounded 1/d, two give it on every input.
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}
// One Newton step (relative error <= 2.2e-15): for FP32 storage, whose every
// stored value is rounded to 2^-24 relative, so the second step changes no
// stored result beyond that rounding; 2 FP64 operations less per cell
// (the FP32-storage kernels are bound by FP64 arithmetic).
__device__ __forceinline__ double recip1(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

// interpolator p(x) = x^3 (6x^2 - 15x + 10), returned divided by 6:
// x^3 (x (x - 5/2) + 5/3).  The factor 6 is folded into the constants p
// multiplies (DA_*, DXE_*), which keeps every operation to one constant
// operand.
__device__ __forceinline__ double p_interp6(double x) {
  return (x * x) * x * fma(x, x - 2.5, 5.0 / 3.0);
}

// KKS solve of a cell (src/parabola625.cu:245-297 restated): with
// p_gamma = 1 - p_delta - p_Laves eliminated, M = A_g + p_d (A_d - A_g) +
// p_l (A_l - A_g) and r = x - xe_g - p_d (xe_d - xe_g) - p_l (xe_l - xe_g);
// mu = M^-1 r, and the neighbours need q = dt D x_g, x_g = xe_g + A_g mu
// (linear in mu; coefficients in StepParams).  The constants here carry the
// factor 6 of p_interp6.  x - xe_g is formed from the stored composition and
// P.xe_g (xe_g less the storage offset).
constexpr double DA_DEL[3] = {6.0 * (A[1][0] - A[0][0]), 6.0 * (A[1][1] - A[0][1]),
                              6.0 * (A[1][2] - A[0][2])};
constexpr double DA_LAV[3] = {6.0 * (A[2][0] - A[0][0]), 6.0 * (A[2][1] - A[0][1]),
                              6.0 * (A[2][2] - A[0][2])};
constexpr double DXE_DEL[2] = {6.0 * (XE[1][0] - XE[0][0]), 6.0 * (XE[1][1] - XE[0][1])};
constexpr double DXE_LAV[2] = {6.0 * (XE[2][0] - XE[0][0]), 6.0 * (XE[2][1] - XE[0][1])};

template <typename T>
__device__ __forceinline__ void derive(const StepParams& P, const Raw<T>& r, Point& v,
                                       Centre& c) {
  v.pd = (double)r.pd;
  v.pl = (double)r.pl;
  c.xCr = (double)r.cr;
  c.xNb = (double)r.nb;
  const double pD = p_interp6(v.pd);
  const double pL = p_interp6(v.pl);
  const double M0 = fma(pD, DA_DEL[0], fma(pL, DA_LAV[0], A[0][0]));
  const double M1 = fma(pD, DA_DEL[1], fma(pL, DA_LAV[1], A[0][1]));
  const double M2 = fma(pD, DA_DEL[2], fma(pL, DA_LAV[2], A[0][2]));
  const double r0 = fma(-pD, DXE_DEL[0], fma(-pL, DXE_LAV[0], c.xCr - P.xe_g[0]));
  const double r1 = fma(-pD, DXE_DEL[1], fma(-pL, DXE_LAV[1], c.xNb - P.xe_g[1]));
  const double det = fma(M0, M2, -(M1 * M1));
  const double inv = sizeof(T) == sizeof(float) ? recip1(det) : recip(det);
  c.mu0 = fma(M2, r0, -(M1 * r1)) * inv;
  c.mu1 = fma(M0, r1, -(M1 * r0)) * inv;
  v.qCr = fma(P.qCr[0], c.mu0, fma(P.qCr[1], c.mu1, P.qCr[2]));
  v.qNb = fma(P.qNb[0], c.mu0, fma(P.qNb[1], c.mu1, P.qNb[2]));
}

__device__ __forceinline__ Point shift_left(const Point& a) {
  return Point{from_left(a.pd), from_left(a.pl), from_left(a.qCr), from_left(a.qNb)};
}
__device__ __forceinline__ Point shift_right(const Point& a) {
  return Point{from_right(a.pd), from_right(a.pl), from_right(a.qCr), from_right(a.qNb)};
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

// Laplacians of the four stencil quantities with the reference's mask
// (src/discretization.cu:589-595): rows y-1, y, y+1; columns x-1, x, x+1.
// NINE: the general nine-term sum in mask order.  Otherwise the mask is a
// folded symmetric five-point one (m[1] = m[7] = 1, m[3] = m[5]; the factor
// the caller's mask had at m[1] is in q and k), and a Laplacian is
// (S + N) + m[3] (W + E) + m[4] C.
template <bool NINE>
__device__ __forceinline__ void laplacians(const double* m, const Point& S, const Point& W,
                                           const Point& C, const Point& E, const Point& Nn,
                                           const Point& SW, const Point& SE, const Point& NW,
                                           const Point& NE, double lap[4]) {
  const double s[4] = {S.pd, S.pl, S.qCr, S.qNb};
  const double w[4] = {W.pd, W.pl, W.qCr, W.qNb};
  const double c[4] = {C.pd, C.pl, C.qCr, C.qNb};
  const double e[4] = {E.pd, E.pl, E.qCr, E.qNb};
  const double n[4] = {Nn.pd, Nn.pl, Nn.qCr, Nn.qNb};
#pragma unroll
  for (int f = 0; f < 4; f++) {
    double v;
    if (NINE) {
      const double sw[4] = {SW.pd, SW.pl, SW.qCr, SW.qNb};
      const double se[4] = {SE.pd, SE.pl, SE.qCr, SE.qNb};
      const double nw[4] = {NW.pd, NW.pl, NW.qCr, NW.qNb};
      const double ne[4] = {NE.pd, NE.pl, NE.qCr, NE.qNb};
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
      v = fma(m[3], w[f] + e[f], s[f] + n[f]);
      v = fma(m[4], c[f], v);
    }
    lap[f] = v;
  }
}

// The explicit Euler update of one cell from its stencil Laplacians:
// Cahn-Hilliard (src/discretization.cu:859-863) and Allen-Cahn
// (src/discretization.cu:888-942) with the closed-form grand-potential
// difference at the solved potential, as the quadratic form of StepParams.
// The composition increment lap(q) is formed on its own and added to x once,
// as the reference adds dt D lap(x_g): where the increment is ~0 (the
// uniform matrix) x comes back unchanged, so rounding adds no drift there.
__device__ __forceinline__ void update_cell(const StepParams& P, const double lap[4],
                                            const Point& cur, const Centre& cc, double& nCr,
                                            double& nNb, double& npd, double& npl) {
  nCr = cc.xCr + lap[2];
  nNb = cc.xNb + lap[3];
  const double fd = cur.pd, fl = cur.pl;
  const double m00 = cc.mu0 * cc.mu0, m01 = cc.mu0 * cc.mu1, m11 = cc.mu1 * cc.mu1;
  const double Gd = fma(P.g[0][0], m00, fma(P.g[0][1], m01, fma(P.g[0][2], m11,
                    fma(P.g[0][3], cc.mu0, P.g[0][4] * cc.mu1))));
  const double Gl = fma(P.g[1][0], m00, fma(P.g[1][1], m01, fma(P.g[1][2], m11,
                    fma(P.g[1][3], cc.mu0, P.g[1][4] * cc.mu1))));
  const double ud = fma(-fd, fd, fd), ul = fma(-fl, fl, fl);  // phi (1 - phi)
  const double td = ud * fma(ud, Gd, P.w[0] * fma(2.0, fd, -1.0));
  const double tl = ul * fma(ul, Gl, P.w[1] * fma(2.0, fl, -1.0));
  npd = fma(-P.c[0] * fd, fl * fl, fma(P.k[0], lap[0], fd) + td);
  npl = fma(-P.c[1] * fl, fd * fd, fma(P.k[1], lap[1], fl) + tl);
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
// wave; uses no barriers.  SW is the number of output columns per strip (62,
// or 60 where the strips are laid out for the two-step march).
template <typename T, int BC, bool NINE, bool NUC, int DEPTH, int SW = STRIP>
__device__ __forceinline__ void march(
    const T* __restrict__ iCr, const T* __restrict__ iNb, const T* __restrict__ ipd,
    const T* __restrict__ ipl, T* __restrict__ oCr, T* __restrict__ oNb, T* __restrict__ opd,
    T* __restrict__ opl, int Nx, int Ny, int pitch, int strip, int chunk, int rows,
    const StepParams& P, const NucParams& N, const Scratch& S) {
  static_assert(UNROLL % DEPTH == 0, "ring depth must divide the unroll factor");
  const int lane = threadIdx.x & (WAVE - 1);

  static_assert(SW >= 1 && SW <= STRIP, "strip width");
  const int gx = strip * SW + lane - 1;   // interior column owned by this lane
  const int px = 1 + bc_map<BC>(gx, Nx);  // padded column it reads
  const bool out_lane = (lane >= 1) && (lane <= SW) && (gx < Nx);

  const int y0 = chunk * rows;
  const int y1 = min(y0 + rows, Ny);
  if (y0 >= Ny) return;

  const uint32_t in_lane = (uint32_t)px * (uint32_t)sizeof(T);
  const uint32_t out_lane_b = (uint32_t)(1 + gx) * (uint32_t)sizeof(T);
  const size_t row_b = (size_t)pitch * sizeof(T);
  auto load = [&](Raw<T>& r, int y) {
    load_row(r, iCr, iNb, ipd, ipl, (size_t)(1 + bc_map<BC>(y, Ny)) * row_b, in_lane);
  };

  // window slots (row y0 + k - 1 + 3j lives in slot k % 3 pattern below)
  Point win[3];
  Centre cen[3];
  Point winW[3], winE[3];  // DPP-shifted neighbours (nine-point masks only)
  Raw<T> ring[DEPTH];

  {
    Raw<T> r;
    load(r, y0 - 1);
    derive<T>(P, r, win[2], cen[2]);
    load(r, y0);
    derive<T>(P, r, win[0], cen[0]);
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
      derive<T>(P, ring[rs], win[ns], cen[ns]);
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

      double nCr, nNb, npd, npl;
      update_cell(P, lap, win[cs], cen[cs], nCr, nNb, npd, npl);

      const int py = 1 + y;
      if (NUC && out_lane) {
        // decide on the stored (storage-rounded) values, as a separate pass
        // would, with the compositions in absolute units; an embed stores
        // xe - xoff, anything else the stored value unchanged
        const double sCr = (double)(T)nCr, sNb = (double)(T)nNb;
        const double aCr = sCr + P.xoff[0], aNb = sNb + P.xoff[1];
        double eCr = aCr, eNb = aNb;
        npd = (double)(T)npd;
        npl = (double)(T)npl;
        nucleate_cell(N, (uint32_t)(py * pitch + 1 + gx), eCr, eNb, npd, npl, S);
        const bool embedded = eCr != aCr || eNb != aNb;
        nCr = embedded ? eCr - P.xoff[0] : sCr;
        nNb = embedded ? eNb - P.xoff[1] : sNb;
      }

      if (out_lane) {
        store_row(oCr, oNb, opd, opl, (size_t)py * row_b, out_lane_b, nCr, nNb, npd, npl);
        store_ghost_images<T, BC>(oCr, oNb, opd, opl, gx, y, Nx, Ny, pitch, nCr, nNb, npd, npl);
      }
    }
  }
}


// ---------------------------------------------------------------------------
// K explicit steps per pass (temporal blocking, K = 2, 3 or 4).  The reference
// advances one step per sequence of launches and so reads and writes every
// field once per step; here a wave marches its strip through K time levels at
// once, so the old fields are read once and the fields K steps later written
// once: 1/K of the HBM traffic per step of the one-step march.
//
// A strip has 64 - 2K output columns and a K-column halo each side.  Level
// l (1..K) is computed on every lane and is valid on lanes l .. 63-l; level K
// is stored, on lanes K .. 63-K.  Level l reads level l-1 through the same DPP
// shifts as the one-step march.
//
// The levels form a pipeline along the march: at iteration t level 0 (the old
// fields, loaded) reaches row y0-K+t and level l computes row y0-K+t-l from
// level l-1's rows y0-K+t-l-1 .. y0-K+t-l+1, the last of which level l-1 has
// just produced.  Level l starts at iteration 2l, so the chunk reads old rows
// y0-K .. y1-1+K and level l covers rows y0-(K-l) .. y1-1+(K-l).  Each level
// keeps a three-row window in registers, row computed at iteration t in slot
// t % 3; the first 2K iterations (the prologue) are unrolled at compile time.
//
// With a window array `lw` in LDS (LW = true, the four-step kernel), each
// level's row also goes to LDS at the end of the iteration that computed it
// and comes back from there two iterations later as the row behind: the
// register windows then hold two live rows per level, not three, which is
// what lets four levels fit in the registers of two waves per SIMD.  Each
// lane reads back only what it wrote itself, so no barrier is needed.  The
// array holds 2 slots (iteration parity) x K levels x 4 values x 64 lanes.
//
// The intermediate levels are exactly what one-step launches would have
// stored: values are rounded to the storage type before the next level uses
// them, and values at positions outside the interior are the images the
// one-step kernel writes into the ghost ring.  Both come out of the index map
// of the old fields (bc_map):
//  - periodic: a halo lane at gx = -1 reads the neighbourhood of column Nx-1;
//  - no-flux: the old fields are read through the even reflection (column -1
//    is column 0, -2 is 1, ...).  The step is pointwise apart from the
//    five-point Laplacian, and the Laplacian of a cell and of its mirror
//    image add the same two values in x (W + E = u[1] + u[0] and u[0] + u[1])
//    and in y, in the same order otherwise; floating-point addition being
//    commutative, every level's value at -1 is its value at 0 bit for bit,
//    which is the ghost copy a one-step launch makes.  So no image needs a
//    fix-up.
// So K steps here give the same result as K one-step launches.  Symmetric
// five-point masks, no nucleation (the caller runs nucleation steps one at a
// time).
template <typename T, int BC, int K, int DEPTH, int SW, bool LW>
__device__ __forceinline__ void marchk_impl(
    const T* __restrict__ iCr, const T* __restrict__ iNb, const T* __restrict__ ipd,
    const T* __restrict__ ipl, T* __restrict__ oCr, T* __restrict__ oNb, T* __restrict__ opd,
    T* __restrict__ opl, int Nx, int Ny, int pitch, int strip, int chunk, int rows,
    const StepParams& P, LdsPtr lw) {
  static_assert(K >= 2 && K <= 4, "two to four steps per pass");
  static_assert(UNROLL % 2 == 0, "the LDS slot is the iteration parity");
  static_assert(UNROLL % DEPTH == 0, "ring depth must divide the unroll factor");
  // SW output columns per strip (64 - 2K, or fewer where the strips are laid
  // out for a deeper march); lanes K .. K+SW-1 store
  static_assert(SW >= 1 && SW <= WAVE - 2 * K, "strip width");
  const int lane = threadIdx.x & (WAVE - 1);

  const int gx = strip * SW + lane - K;  // interior column of this lane (may be a halo)
  const int px = 1 + bc_map<BC>(gx, Nx);
  const bool out_lane = (lane >= K) && (lane < K + SW) && (gx < Nx);

  const int y0 = chunk * rows;
  const int y1 = min(y0 + rows, Ny);
  if (y0 >= Ny) return;
  const int nrows = y1 - y0;
  const int y_last = y1 - 1 + K;  // last old row the chunk reads

  const uint32_t in_lane = (uint32_t)px * (uint32_t)sizeof(T);
  const uint32_t out_lane_b = (uint32_t)(1 + gx) * (uint32_t)sizeof(T);
  const size_t row_b = (size_t)pitch * sizeof(T);
  auto load = [&](Raw<T>& r, int y) {
    load_row(r, iCr, iNb, ipd, ipl, (size_t)(1 + bc_map<BC>(y, Ny)) * row_b, in_lane);
  };
  const double* m = P.m;

  // windows: level l (0 = old fields) row of iteration t in slot t % 3
  Point win[K][3];
  Centre cen[K][3];
  Raw<T> ring[DEPTH];

  // LDS copy of level l's row of an iteration of parity `slot`
  auto put = [&](int l, int slot, const Point& v) {
    LdsPtr q = lw + ((l * 2 + slot) * 4) * WAVE + lane;
    q[0] = v.pd;
    q[WAVE] = v.pl;
    q[2 * WAVE] = v.qCr;
    q[3 * WAVE] = v.qNb;
  };
  // level l's row behind at iteration t (slot sb of the window; in LDS, the
  // row stored at iteration t - 2, of the same parity)
  auto behind = [&](int l, int sb, int par) -> Point {
    if constexpr (LW) {
      const LdsPtr q = lw + ((l * 2 + par) * 4) * WAVE + lane;
      return Point{q[0], q[WAVE], q[2 * WAVE], q[3 * WAVE]};
    } else {
      return win[l][sb];
    }
  };

  // One step at the centre of level `lv`'s window (slots behind / centre /
  // ahead): the next level's values, rounded to storage, derived into `v`, `c`.
  auto level_up = [&](const Point& Sth, const Point& C, const Centre& cc, const Point& Nth,
                      Point& v, Centre& c) {
    const Point W = shift_left(C), E = shift_right(C);
    double lap[4];
    laplacians<false>(m, Sth, W, C, E, Nth, Sth, Sth, Nth, Nth, lap);
    double n0, n1, n2, n3;
    update_cell(P, lap, C, cc, n0, n1, n2, n3);
    Raw<T> r;
    r.cr = (T)n0;
    r.nb = (T)n1;
    r.pd = (T)n2;
    r.pl = (T)n3;
    derive<T>(P, r, v, c);
  };

  // Levels 1..K-1 at iteration t (slots are compile-time in both callers):
  // level l computes row y0-K+t-l.
  auto middle_levels = [&](int sa, int sc, int sb, int lmax_active, int par) {
#pragma unroll
    for (int l = 1; l < K; l++) {
      if (l > lmax_active) break;
      level_up(behind(l - 1, sb, par), win[l - 1][sc], cen[l - 1][sc], win[l - 1][sa],
               win[l][sa], cen[l][sa]);
    }
  };
  // end of iteration t: the rows of the active levels to LDS
  auto put_rows = [&](int sa, int lmax_active, int par) {
    if constexpr (LW) {
#pragma unroll
      for (int l = 0; l < K; l++) {
        if (l > lmax_active) break;
        put(l, par, win[l][sa]);
      }
    }
  };

  // ---- prologue: iterations t = 0 .. 2K-1 (level K not yet active).  Every
  // row the prologue reads, and the ring, is requested before the first is
  // used, so a wave waits for one memory latency at its start, not 2K.
  {
    Raw<T> pro[2 * K];
#pragma unroll
    for (int t = 0; t < 2 * K; t++) load(pro[t], y0 - K + t);
#pragma unroll
    for (int j = 0; j < DEPTH; j++) load(ring[j], min(y0 + K + j, y_last));
#pragma unroll
    for (int t = 0; t < 2 * K; t++) {
      const int sa = t % 3, sc = (t + 2) % 3, sb = (t + 1) % 3;
      derive<T>(P, pro[t], win[0][sa], cen[0][sa]);
      // level l is active from iteration 2l on
      middle_levels(sa, sc, sb, t / 2, t % 2);
      put_rows(sa, t / 2, t % 2);
    }
  }

  // ---- main loop: iteration t = 2K + i, output row y = y0 + i (t and k
  // have the same parity: 2K and UNROLL are even)
  for (int ib = 0;; ib += UNROLL) {
#pragma unroll
    for (int k = 0; k < UNROLL; k++) {
      if (ib + k >= nrows) return;
      const int y = y0 + ib + k;
      const int sa = (2 * K + k) % 3, sc = (2 * K + k + 2) % 3, sb = (2 * K + k + 1) % 3;
      const int rs = k % DEPTH;
      derive<T>(P, ring[rs], win[0][sa], cen[0][sa]);  // old row y + K
      load(ring[rs], min(y + K + DEPTH, y_last));
      middle_levels(sa, sc, sb, K - 1, k % 2);
      // level K: row y from level K-1's rows y-1 (behind), y, y+1 (ahead)
      const Point Sth = behind(K - 1, sb, k % 2);
      const Point& C = win[K - 1][sc];
      const Point& Nth = win[K - 1][sa];
      const Point W = shift_left(C), E = shift_right(C);
      double lap[4];
      laplacians<false>(m, Sth, W, C, E, Nth, Sth, Sth, Nth, Nth, lap);
      double nCr, nNb, npd, npl;
      update_cell(P, lap, C, cen[K - 1][sc], nCr, nNb, npd, npl);
      if (out_lane) {
        store_row(oCr, oNb, opd, opl, (size_t)(1 + y) * row_b, out_lane_b, nCr, nNb, npd, npl);
        store_ghost_images<T, BC>(oCr, oNb, opd, opl, gx, y, Nx, Ny, pitch, nCr, nNb, npd, npl);
      }
      put_rows(sa, K - 1, k % 2);
    }
  }
}

template <typename T, int BC, int K, int DEPTH, int SW = WAVE - 2 * K, bool LW = false>
__device__ __forceinline__ void marchk(
    const T* __restrict__ iCr, const T* __restrict__ iNb, const T* __restrict__ ipd,
    const T* __restrict__ ipl, T* __restrict__ oCr, T* __restrict__ oNb, T* __restrict__ opd,
    T* __restrict__ opl, int Nx, int Ny, int pitch, int strip, int chunk, int rows,
    const StepParams& P, LdsPtr lw = nullptr) {
  marchk_impl<T, BC, K, DEPTH, SW, LW>(iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, Nx, Ny, pitch,
                                       strip, chunk, rows, P, lw);
}

}  // namespace pf
