/**
 \file  parabola625.cuh
 \brief Device pointwise thermodynamics (drop-in for src/parabola625.cuh:13-95,
        the subset the step path uses).

 Same names and signatures as the reference's __device__ functions; defined
 inline over pf_thermo.h (paraboloid free energies, closed-form KKS solve,
 mobility-model diffusivities).  The CALPHAD family (d_GCAL*, d_dGCAL*,
 d_d2GCAL*, d_mu_*), d_xr_* and the composition-range constants are not on
 the step path and are not provided.
*/
#ifndef __D_PARABOLA625_H__
#define __D_PARABOLA625_H__

#include "pf_thermo.h"
#include "type.h"

__device__ inline double d_p(double x) { return pf::p(x); }
__device__ inline double d_pPrime(double x) { return pf::pPrime(x); }
__device__ inline double d_interface_profile(double z) { return 0.5 - 0.5 * tanh(z); }
__device__ inline double d_kT() { return pf::kT; }
__device__ inline double d_RT() { return pf::RT; }
__device__ inline double d_Vm() { return pf::Vm; }
__device__ inline double d_xe_gam_Cr() { return pf::XE[0][0]; }
__device__ inline double d_xe_gam_Nb() { return pf::XE[0][1]; }
__device__ inline double d_xe_del_Cr() { return pf::XE[1][0]; }
__device__ inline double d_xe_del_Nb() { return pf::XE[1][1]; }
__device__ inline double d_xe_lav_Cr() { return pf::XE[2][0]; }
__device__ inline double d_xe_lav_Nb() { return pf::XE[2][1]; }

__device__ inline double d_fict_phase(int a, int c, double XCR, double XNB, double pDel,
                                      double pGam, double pLav) {
	double mu0, mu1;
	pf::kks_mu(XCR, XNB, pDel, pGam, pLav, mu0, mu1);
	return pf::fict(a, c, mu0, mu1);
}
__device__ inline double d_fict_gam_Cr(double XCR, double XNB, double pDel, double pGam, double pLav) { return d_fict_phase(0, 0, XCR, XNB, pDel, pGam, pLav); }
__device__ inline double d_fict_gam_Nb(double XCR, double XNB, double pDel, double pGam, double pLav) { return d_fict_phase(0, 1, XCR, XNB, pDel, pGam, pLav); }
__device__ inline double d_fict_del_Cr(double XCR, double XNB, double pDel, double pGam, double pLav) { return d_fict_phase(1, 0, XCR, XNB, pDel, pGam, pLav); }
__device__ inline double d_fict_del_Nb(double XCR, double XNB, double pDel, double pGam, double pLav) { return d_fict_phase(1, 1, XCR, XNB, pDel, pGam, pLav); }
__device__ inline double d_fict_lav_Cr(double XCR, double XNB, double pDel, double pGam, double pLav) { return d_fict_phase(2, 0, XCR, XNB, pDel, pGam, pLav); }
__device__ inline double d_fict_lav_Nb(double XCR, double XNB, double pDel, double pGam, double pLav) { return d_fict_phase(2, 1, XCR, XNB, pDel, pGam, pLav); }

__device__ inline double d_s_delta() { return pf::SIGMA_DEL; }
__device__ inline double d_s_laves() { return pf::SIGMA_LAV; }

__device__ inline double d_g_gam(double XCR, double XNB) { return pf::g(0, XCR, XNB); }
__device__ inline double d_g_del(double XCR, double XNB) { return pf::g(1, XCR, XNB); }
__device__ inline double d_g_lav(double XCR, double XNB) { return pf::g(2, XCR, XNB); }

__device__ inline double d_dg_gam_dxCr(double XCR, double XNB) { return pf::dg_dxCr(0, XCR, XNB); }
__device__ inline double d_dg_gam_dxNb(double XCR, double XNB) { return pf::dg_dxNb(0, XCR, XNB); }
__device__ inline double d_dg_del_dxCr(double XCR, double XNB) { return pf::dg_dxCr(1, XCR, XNB); }
__device__ inline double d_dg_del_dxNb(double XCR, double XNB) { return pf::dg_dxNb(1, XCR, XNB); }
__device__ inline double d_dg_lav_dxCr(double XCR, double XNB) { return pf::dg_dxCr(2, XCR, XNB); }
__device__ inline double d_dg_lav_dxNb(double XCR, double XNB) { return pf::dg_dxNb(2, XCR, XNB); }

__device__ inline double d_d2g_gam_dxCrCr() { return pf::H[0][0]; }
__device__ inline double d_d2g_gam_dxCrNb() { return pf::H[0][1]; }
__device__ inline double d_d2g_gam_dxNbCr() { return pf::H[0][1]; }
__device__ inline double d_d2g_gam_dxNbNb() { return pf::H[0][2]; }
__device__ inline double d_d2g_del_dxCrCr() { return pf::H[1][0]; }
__device__ inline double d_d2g_del_dxCrNb() { return pf::H[1][1]; }
__device__ inline double d_d2g_del_dxNbCr() { return pf::H[1][1]; }
__device__ inline double d_d2g_del_dxNbNb() { return pf::H[1][2]; }
__device__ inline double d_d2g_lav_dxCrCr() { return pf::H[2][0]; }
__device__ inline double d_d2g_lav_dxCrNb() { return pf::H[2][1]; }
__device__ inline double d_d2g_lav_dxNbCr() { return pf::H[2][1]; }
__device__ inline double d_d2g_lav_dxNbNb() { return pf::H[2][2]; }

__device__ inline double d_D_CrCr(double XCR, double XNB) { return pf::D(0, 0, XCR, XNB); }
__device__ inline double d_D_CrNb(double XCR, double XNB) { return pf::D(0, 1, XCR, XNB); }
__device__ inline double d_D_NbCr(double XCR, double XNB) { return pf::D(1, 0, XCR, XNB); }
__device__ inline double d_D_NbNb(double XCR, double XNB) { return pf::D(1, 1, XCR, XNB); }

#endif /* __D_PARABOLA625_H__ */
