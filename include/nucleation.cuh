/**
 \file  nucleation.cuh
 \brief Device nucleation functions (drop-in for src/nucleation.cuh:8-20).
*/
#ifndef __D_NUCLEATION_H_
#define __D_NUCLEATION_H_

#include "pf_thermo.h"
#include "type.h"

/** src/nucleation.cu:12-16 */
__device__ inline fp_t d_tanh_interp(const fp_t r, const fp_t w) { return 0.5 - 0.5 * tanh(r / w); }

/** Driving force = g_gamma(x) (src/nucleation.cu:18-22). */
__device__ inline void d_nucleation_driving_force_delta(const fp_t& xCr, const fp_t& xNb, fp_t* dG)
{
	*dG = pf::g(0, xCr, xNb);
}

/** Driving force = g_gamma(x) (src/nucleation.cu:24-28). */
__device__ inline void d_nucleation_driving_force_laves(const fp_t& xCr, const fp_t& xNb, fp_t* dG)
{
	*dG = pf::g(0, xCr, xNb);
}

/** Classical nucleation probability of a sphere (src/nucleation.cu:30-70). */
__device__ inline void d_nucleation_probability_sphere(const fp_t& xCr, const fp_t& xNb,
                                                       const fp_t& dG_chem,
                                                       const fp_t& D_CrCr, const fp_t& D_NbNb,
                                                       const fp_t& sigma,
                                                       const fp_t& Vatom,
                                                       const fp_t& n_gam, const fp_t& dV,
                                                       const fp_t& dt,
                                                       fp_t* Rstar, fp_t* P_nuc)
{
	pf::nucleation_probability(xCr, xNb, dG_chem, D_CrCr, D_NbNb, sigma, Vatom, n_gam, dV, dt,
	                           *Rstar, *P_nuc);
}

#endif /* __D_NUCLEATION_H_ */
