// pf_mobility.hip -- the variable-mobility Cahn-Hilliard divergence: the
// face-averaged sums of chemical_convolution_Cr_kernel / _Nb_kernel
// (src/discretization.cu:605-778), with the mobility arrays as inputs.  In the
// reference these kernels compile but their launch is commented out
// (src/discretization.cu:809-835), and the mobility kernels that would fill
// the arrays call d_M_* functions no file defines; so the caller supplies
// the mobilities.
//
// Per cell and component i (Cr, Nb):
//   div_i = sum_{t<6} [ F^x_t(x+1/2) - F^x_t(x-1/2) ] / (2 dx^2)
//                   + [ F^y_t(y+1/2) - F^y_t(y-1/2) ] / (2 dy^2)
//         + sum_{t=6,7} [ F^x_t(x+1/2) - F^x_t(x-1/2) ] / (2 dx^2)
// with the face flux F(a+1/2) = (M[a+1] + M[a]) (c[a+1] - c[a]) of the pair
// (mob_i[t], conc[t]) for t < 6 and (mob_i[6 + k], phi_k) for the phase
// terms, which the reference differences in x only.  Every cell of the padded
// arrays is computed, ghost ring included; the neighbour beyond the array
// edge is the cell itself, as the reference's index clamp makes it, so the
// flux through the array edge is zero.
//
// One wave64 owns a strip of 62 output columns (+1 halo lane each side) and
// marches down a chunk of rows.  Each face flux is formed once: an x flux on
// the lane left of the face, handed to the lane right of it by a DPP shift;
// a y flux once per row pair, carried to the next row in registers.  So
// every input is loaded once per strip (24 loads, 2 stores per cell: the
// kernel is HBM-bound at 208 B per cell).
#include <hip/hip_runtime.h>

#include <string>

#include "pf_internal.h"
#include "pf_march.h"
#include "pfpa.h"

namespace pf {

constexpr int MOB_TERMS = 6;   // composition terms (x and y)
constexpr int MOB_PHASE = 2;   // phase-field terms (x only)

struct MobArgs {
  const double* conc[6];
  const double* mob[2][8];
  const double* phi[2];
  double* out[2];
};

__global__ __launch_bounds__(WAVE * WPB) void mobility_divergence_kernel(MobArgs a, int nx, int ny,
                                                                         int rows, int strips,
                                                                         double kx, double ky) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int nb = gridDim.x;
  const int b = xcd_remap(blockIdx.x, nb);
  const int groups = (strips + WPB - 1) / WPB;
  const int strip = (b % groups) * WPB + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int chunk = b / groups;
  if (strip >= strips) return;
  const int y0 = chunk * rows;
  const int y1 = min(y0 + rows, ny);
  const int gx = strip * STRIP + lane - 1;
  const int cx = gx < 0 ? 0 : (gx >= nx ? nx - 1 : gx);  // the reference's clamp at the array edge
  const bool out_lane = lane >= 1 && lane <= STRIP && gx < nx;

  // a wave-uniform row address (made scalar explicitly, so the loads take it
  // from scalar registers) plus a 32-bit per-lane byte offset: no per-lane
  // 64-bit base per array
  const uint32_t lane_bytes = (uint32_t)cx * sizeof(double);
  auto at = [&](const double* p, int y) -> double {
    const uintptr_t row = (uintptr_t)p + (size_t)y * (size_t)nx * sizeof(double);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)row);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(row >> 32));
    typedef const char __attribute__((address_space(1)))* gbytes;   // global memory
    gbytes r = (gbytes)(((uintptr_t)hi << 32) | lo);
    return *(const double __attribute__((address_space(1)))*)(r + lane_bytes);
  };
  auto yc = [&](int y) { return y < 0 ? 0 : (y >= ny ? ny - 1 : y); };

  // current-row values of the y-differenced arrays, the carried y fluxes
  // F(y - 1/2), per composition term: conc, mob Cr, mob Nb
  double c[MOB_TERMS], m0[MOB_TERMS], m1[MOB_TERMS], fy0[MOB_TERMS], fy1[MOB_TERMS];
  {
    const int yp = yc(y0 - 1), yq = yc(y0);
#pragma unroll
    for (int t = 0; t < MOB_TERMS; t++) {
      const double cp = at(a.conc[t], yp), mp0 = at(a.mob[0][t], yp), mp1 = at(a.mob[1][t], yp);
      c[t] = at(a.conc[t], yq);
      m0[t] = at(a.mob[0][t], yq);
      m1[t] = at(a.mob[1][t], yq);
      const double dc = c[t] - cp;
      fy0[t] = (m0[t] + mp0) * dc;
      fy1[t] = (m1[t] + mp1) * dc;
    }
  }
  for (int y = y0; y < y1; y++) {
    const int yn = yc(y + 1);
    double cn[MOB_TERMS], mn0[MOB_TERMS], mn1[MOB_TERMS];
#pragma unroll
    for (int t = 0; t < MOB_TERMS; t++) {
      cn[t] = at(a.conc[t], yn);
      mn0[t] = at(a.mob[0][t], yn);
      mn1[t] = at(a.mob[1][t], yn);
    }
    double ph[MOB_PHASE], mp[2][MOB_PHASE];
#pragma unroll
    for (int k = 0; k < MOB_PHASE; k++) {
      ph[k] = at(a.phi[k], y);
      mp[0][k] = at(a.mob[0][6 + k], y);
      mp[1][k] = at(a.mob[1][6 + k], y);
    }
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int t = 0; t < MOB_TERMS; t++) {
      // x: the flux through the right face, and the left lane's right face
      const double dcx = from_right(c[t]) - c[t];
      const double fr0 = (from_right(m0[t]) + m0[t]) * dcx;
      const double fr1 = (from_right(m1[t]) + m1[t]) * dcx;
      const double fl0 = from_left(fr0), fl1 = from_left(fr1);
      s0 += (fr0 - fl0) * kx;
      s1 += (fr1 - fl1) * kx;
      // y: the flux through the upper face; the lower one is carried
      const double dcy = cn[t] - c[t];
      const double fu0 = (mn0[t] + m0[t]) * dcy;
      const double fu1 = (mn1[t] + m1[t]) * dcy;
      s0 += (fu0 - fy0[t]) * ky;
      s1 += (fu1 - fy1[t]) * ky;
      fy0[t] = fu0;
      fy1[t] = fu1;
      c[t] = cn[t];
      m0[t] = mn0[t];
      m1[t] = mn1[t];
    }
#pragma unroll
    for (int k = 0; k < MOB_PHASE; k++) {
      const double dcx = from_right(ph[k]) - ph[k];
      const double fr0 = (from_right(mp[0][k]) + mp[0][k]) * dcx;
      const double fr1 = (from_right(mp[1][k]) + mp[1][k]) * dcx;
      s0 += (fr0 - from_left(fr0)) * kx;
      s1 += (fr1 - from_left(fr1)) * kx;
    }
    if (out_lane) {
      const size_t o = (size_t)y * nx + gx;
      a.out[0][o] = s0;
      a.out[1][o] = s1;
    }
  }
}

}  // namespace pf

namespace {

int mfail(int code, const std::string& msg) {
  pf::set_error(msg);
  return code;
}

}  // namespace

extern "C" int pf_mobility_divergence_f64(int nx, int ny, double dx, double dy,
                                          const pf_mobility_fields* f, double* div_Cr,
                                          double* div_Nb, void* stream) {
  if (!f || !div_Cr || !div_Nb) return mfail(PF_ERR_ARG, "null argument");
  if (nx < 1 || ny < 1) return mfail(PF_ERR_ARG, "grid must have at least one cell");
  if ((long long)nx * (long long)ny >= (1LL << 31))
    return mfail(PF_ERR_ARG, "grid too large for 32-bit cell indices");
  if (!(dx > 0.0) || !(dy > 0.0)) return mfail(PF_ERR_ARG, "mesh spacing must be positive");
  pf::MobArgs a{};
  for (int t = 0; t < 6; t++) {
    if (!f->conc[t]) return mfail(PF_ERR_ARG, "null composition array");
    a.conc[t] = f->conc[t];
  }
  for (int t = 0; t < 8; t++) {
    if (!f->mob_Cr[t] || !f->mob_Nb[t]) return mfail(PF_ERR_ARG, "null mobility array");
    a.mob[0][t] = f->mob_Cr[t];
    a.mob[1][t] = f->mob_Nb[t];
  }
  if (!f->phi_del || !f->phi_lav) return mfail(PF_ERR_ARG, "null phase-field array");
  a.phi[0] = f->phi_del;
  a.phi[1] = f->phi_lav;
  // the outputs are written while every input is read: no overlap allowed
  const uintptr_t n = (uintptr_t)nx * (uintptr_t)ny * sizeof(double);
  const double* ins[24];
  for (int t = 0; t < 6; t++) ins[t] = a.conc[t];
  for (int t = 0; t < 8; t++) ins[6 + t] = a.mob[0][t], ins[14 + t] = a.mob[1][t];
  ins[22] = a.phi[0];
  ins[23] = a.phi[1];
  const uintptr_t o0 = (uintptr_t)div_Cr, o1 = (uintptr_t)div_Nb;
  if (o0 < o1 + n && o1 < o0 + n) return mfail(PF_ERR_ARG, "div_Cr and div_Nb overlap");
  for (int k = 0; k < 24; k++) {
    const uintptr_t i = (uintptr_t)ins[k];
    if ((o0 < i + n && i < o0 + n) || (o1 < i + n && i < o1 + n))
      return mfail(PF_ERR_ARG, "an output array overlaps an input array");
  }
  a.out[0] = div_Cr;
  a.out[1] = div_Nb;
  const int rows = 32;
  const int strips = (nx + pf::STRIP - 1) / pf::STRIP;
  const int groups = (strips + pf::WPB - 1) / pf::WPB;
  const int chunks = (ny + rows - 1) / rows;
  const long long blocks = (long long)groups * chunks;
  if (blocks > 0x7fffffffLL) return mfail(PF_ERR_ARG, "grid too large");
  hipLaunchKernelGGL(pf::mobility_divergence_kernel, dim3((unsigned)blocks),
                     dim3(pf::WAVE * pf::WPB), 0, (hipStream_t)stream, a, nx, ny, rows, strips,
                     1.0 / (2.0 * dx * dx), 1.0 / (2.0 * dy * dy));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return mfail(PF_ERR_HIP, std::string("mobility divergence launch: ") + hipGetErrorString(e));
  return PF_OK;
}
