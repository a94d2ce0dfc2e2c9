// pf_thermo.h -- pointwise thermodynamics of the ternary Cr-Nb-Ni KKS model,
// usable from host and device code.
//
// Model tables (values must equal the reference's, so they are kept alike):
//   kT, RT, Vm                       thermo/parabola625.c:38-63
//   equilibrium compositions xe      thermo/parabola625.c:65-117
//   paraboloid Hessians d2g          thermo/parabola625.c:1305-1411
//   interfacial energies             thermo/parabola625.c:299-316
//   mobility prefactors / exponents  thermo/parabola625.c:1860-1894
//
// Everything else is derived here.  The free energies are paraboloids
//   g_a(x) = 1/2 (x - xe_a)^T H_a (x - xe_a),  a in {gamma, delta, Laves},
// so the KKS condition (equal diffusion potentials mu, mass balance
// x = sum_a p_a x_a) is a 2x2 linear solve
//   M = sum_a p_a A_a   (A_a = H_a^-1),   mu = M^-1 (x - sum_a p_a xe_a),
//   x_a = xe_a + A_a mu,
// which replaces the sympy-expanded rational functions the reference evaluates
// (thermo/parabola625.c:245-297) and needs one reciprocal per point.  With the
// solved mu, g_gamma(x_gamma) = 1/2 mu^T A_gamma mu and dg_gamma/dx = mu, so the
// Allen-Cahn grand-potential difference (src/discretization.cu:904,932) is
//   dOmega_a = mu^T B_a mu - mu . e_a,  B_a = (A_a - A_gamma)/2,  e_a = xe_gamma - xe_a.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define PF_HD __host__ __device__ __forceinline__

namespace pf {

// ---------------------------------------------------------------- tables
constexpr double kT = 1.5782889043500002e-20;
constexpr double RT = 9504.6840941999999;
constexpr double Vm = 1.0000000000000001e-5;

// phase index: 0 = gamma (matrix), 1 = delta, 2 = Laves
constexpr double XE[3][2] = {
    {0.52421634830562147, 0.01299272922003303},
    {0.022966218927631978, 0.24984563695705883},
    {0.37392129441013022, 0.25826261799015571}};

// symmetric Hessians {CrCr, CrNb, NbNb}
constexpr double H[3][3] = {
    {4534264425.5240498, 15095482346.486225, 110386166481.37187},
    {42692985981.597771, 16906497386.287418, 6170738265863.7695},
    {17733460569.613991, 24191004361.532181, 196588620559.76782}};

constexpr double SIGMA_DEL = 0.13;
constexpr double SIGMA_LAV = 0.13;
constexpr double LATTICE = 0.352e-9;  // FCC Ni lattice constant, src/parameters.h:27

// mobility prefactors c_k (k = Cr, Nb, Ni) of D_ij = sum_k (d_ik - x_i) x_k c_k e^{Q_k} dmu_k/dx_j
constexpr double MOB_C[3] = {3.5027570743586952e-21, 1.4170990277916469e-20,
                             1.8295676400933012e-21};

// ---------------------------------------------------------------- derived
constexpr double det2(int a) { return H[a][0] * H[a][2] - H[a][1] * H[a][1]; }
constexpr double A[3][3] = {
    {H[0][2] / det2(0), -H[0][1] / det2(0), H[0][0] / det2(0)},
    {H[1][2] / det2(1), -H[1][1] / det2(1), H[1][0] / det2(1)},
    {H[2][2] / det2(2), -H[2][1] / det2(2), H[2][0] / det2(2)}};
constexpr double B_DEL[3] = {0.5 * (A[1][0] - A[0][0]), 0.5 * (A[1][1] - A[0][1]),
                             0.5 * (A[1][2] - A[0][2])};
constexpr double B_LAV[3] = {0.5 * (A[2][0] - A[0][0]), 0.5 * (A[2][1] - A[0][1]),
                             0.5 * (A[2][2] - A[0][2])};
constexpr double E_DEL[2] = {XE[0][0] - XE[1][0], XE[0][1] - XE[1][1]};
constexpr double E_LAV[2] = {XE[0][0] - XE[2][0], XE[0][1] - XE[2][1]};

// ---------------------------------------------------------------- pointwise
// interpolator p(x) = x^3 (6x^2 - 15x + 10) (src/parabola625.cu:11-18 form)
PF_HD double p(double x) { return x * x * x * (6.0 * x * x - 15.0 * x + 10.0); }
// p'(x) = 30 x^2 (1-x)^2 (src/parabola625.cu:20-27 form)
PF_HD double pPrime(double x) { return 30.0 * x * x * (1.0 - x) * (1.0 - x); }

PF_HD double g(int a, double xCr, double xNb) {
  const double u = xCr - XE[a][0], v = xNb - XE[a][1];
  return 0.5 * H[a][0] * u * u + H[a][1] * u * v + 0.5 * H[a][2] * v * v;
}
PF_HD double dg_dxCr(int a, double xCr, double xNb) {
  return H[a][0] * (xCr - XE[a][0]) + H[a][1] * (xNb - XE[a][1]);
}
PF_HD double dg_dxNb(int a, double xCr, double xNb) {
  return H[a][1] * (xCr - XE[a][0]) + H[a][2] * (xNb - XE[a][1]);
}

// Common diffusion potential of the KKS split (general pGam, as the reference's
// fict_* functions take it).
PF_HD void kks_mu(double xCr, double xNb, double pDel, double pGam, double pLav,
                  double& mu0, double& mu1) {
  const double M0 = pGam * A[0][0] + pDel * A[1][0] + pLav * A[2][0];
  const double M1 = pGam * A[0][1] + pDel * A[1][1] + pLav * A[2][1];
  const double M2 = pGam * A[0][2] + pDel * A[1][2] + pLav * A[2][2];
  const double r0 = xCr - pGam * XE[0][0] - pDel * XE[1][0] - pLav * XE[2][0];
  const double r1 = xNb - pGam * XE[0][1] - pDel * XE[1][1] - pLav * XE[2][1];
  const double inv = 1.0 / (M0 * M2 - M1 * M1);
  mu0 = (M2 * r0 - M1 * r1) * inv;
  mu1 = (M0 * r1 - M1 * r0) * inv;
}
PF_HD double fict(int a, int comp, double mu0, double mu1) {
  return comp == 0 ? XE[a][0] + A[a][0] * mu0 + A[a][1] * mu1
                   : XE[a][1] + A[a][1] * mu0 + A[a][2] * mu1;
}

// grand-potential difference gamma - a (a = 1 delta, 2 Laves) at potential mu
PF_HD double dOmega_del(double mu0, double mu1) {
  return mu0 * (B_DEL[0] * mu0 + 2.0 * B_DEL[1] * mu1) + B_DEL[2] * mu1 * mu1 -
         (mu0 * E_DEL[0] + mu1 * E_DEL[1]);
}
PF_HD double dOmega_lav(double mu0, double mu1) {
  return mu0 * (B_LAV[0] * mu0 + 2.0 * B_LAV[1] * mu1) + B_LAV[2] * mu1 * mu1 -
         (mu0 * E_LAV[0] + mu1 * E_LAV[1]);
}

// mobility exponent factor e^{Q_k(x)} (thermo/parabola625.c:1860-1894 tables)
PF_HD double mob_exp(int k, double xCr, double xNb) {
  const double xNi = 1.0 - xCr - xNb;
  if (k == 0) return exp(-7.1543718172640256 * xCr * xNi + 3.3541967644499233 * xCr);
  if (k == 1) return exp(-34.982563536715503 * xNb * xNi + 14.437947842988088 * xNb);
  return exp(-8.5221193705645018 * xCr * xNi + 4.0036674816806448 * xCr -
             442.62877981749585 * xNb * xNi - 0.82164267100913768 * xNb);
}

// diffusivity D_ij(x), i, j in {0: Cr, 1: Nb}
PF_HD double D(int i, int j, double xCr, double xNb) {
  const double h0 = H[0][0] * Vm, h1 = H[0][1] * Vm, h2 = H[0][2] * Vm;
  const double x[3] = {xCr, xNb, 1.0 - xCr - xNb};
  double dmu[3];  // dmu_k/dx_j, k = Cr, Nb, Ni
  if (j == 0) {
    dmu[0] = (1.0 - xCr) * h0 - xNb * h1;
    dmu[1] = (1.0 - xNb) * h1 - xCr * h0;
    dmu[2] = -xCr * h0 - xNb * h1;
  } else {
    dmu[0] = (1.0 - xCr) * h1 - xNb * h2;
    dmu[1] = (1.0 - xNb) * h2 - xCr * h1;
    dmu[2] = -xCr * h1 - xNb * h2;
  }
  double s = 0.0;
  for (int k = 0; k < 3; k++) {
    const double kron = (i == k) ? 1.0 : 0.0;
    s += (kron - x[i]) * x[k] * MOB_C[k] * mob_exp(k, xCr, xNb) * dmu[k];
  }
  return s;
}

// Classical nucleation probability for a sphere (thermo/nucleation.c:52-81),
// with the reference's signs of dc_Cr and dc_Nb.
PF_HD void nucleation_probability(double xCr, double xNb, double dG, double DCrCr, double DNbNb,
                                  double sigma, double Vatom, double n_gam, double dV, double dt,
                                  double& Rstar, double& P) {
  const double pi = 3.14159265358979323846;
  const double s3 = sigma * sigma * sigma;
  const double Zeldov = (Vatom * dG * dG) / (8. * pi * sqrt(kT * s3));
  const double Gstar = (16. * pi * s3) / (3. * dG * dG);
  const double a4 = LATTICE * LATTICE * LATTICE * LATTICE;
  const double BCr = (3. * Gstar * DCrCr * xCr) / (sigma * a4);
  const double BNb = (3. * Gstar * DNbNb * xNb) / (sigma * a4);
  const double k1Cr = BCr * Zeldov * n_gam;
  const double k1Nb = BNb * Zeldov * n_gam;
  const double k2 = Gstar / kT;
  const double dcCr = -xCr + XE[0][0];
  const double dcNb = xNb - XE[0][1];
  const double JCr = k1Cr * exp(-k2 / dcCr);
  const double JNb = k1Nb * exp(-k2 / dcNb);
  P = 1. - exp(-dt * dV * (JCr + JNb));
  Rstar = (2. * sigma) / dG;
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based, stateless.
struct u32x4 {
  uint32_t x, y, z, w;
};
PF_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    u32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Nucleation stream: key = seed, counter = (cell, instance, call lo, call hi);
// two 53-bit uniforms in (0, 1]: u1 (delta test), u2 (Laves test).
PF_HD void nuc_uniforms(uint64_t seed, uint64_t call, uint32_t instance, uint32_t cell, double& u1,
                        double& u2) {
  const u32x4 r = philox4x32_10(u32x4{cell, instance, (uint32_t)call, (uint32_t)(call >> 32)},
                                (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = ((((uint64_t)r.y) << 32) | r.x) >> 11;
  const uint64_t b = ((((uint64_t)r.w) << 32) | r.z) >> 11;
  u1 = (double)(a + 1) * (1.0 / 9007199254740992.0);
  u2 = (double)(b + 1) * (1.0 / 9007199254740992.0);
}

}  // namespace pf
