/**
 \file  type.h
 \brief Scalar type and container structures of the KKS step (drop-in for
        src/type.h:13-107 of the reference).

 HostData is unchanged.  CudaData keeps every member the driver touches
 (the *_old / *_new field pointers it swaps, src/main.cpp:318-321) and the
 visualisation arrays; the fictitious-composition arrays are gone (the fused
 step keeps them on chip), the per-cell cuRAND state becomes a counter-based
 Philox stream, and the context of libpfpa.so rides along.
*/
#ifndef _TYPE_H_
#define _TYPE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NP 2 /* number of phases */
#define NC 2 /* number of components */

typedef double fp_t;

/** Host arrays (src/type.h:28-53). */
struct HostData {
	fp_t** mask_lap;
	fp_t** conc_Ni;
	fp_t** phi;

	fp_t** chem_nrg;
	fp_t** grad_nrg;

	fp_t** conc_Cr_old;
	fp_t** conc_Cr_new;
	fp_t** conc_Cr_lap;
	fp_t** conc_Cr_gam;

	fp_t** conc_Nb_old;
	fp_t** conc_Nb_new;
	fp_t** conc_Nb_lap;
	fp_t** conc_Nb_gam;

	fp_t** phi_del_old;
	fp_t** phi_del_new;
	fp_t** phi_del_lap;

	fp_t** phi_lav_old;
	fp_t** phi_lav_new;
	fp_t** phi_lav_lap;
};

/** Counter-based random stream replacing curandState* prng (src/type.h:59). */
struct PhiloxStream {
	uint64_t seed;   /* Philox4x32-10 key */
	uint64_t calls;  /* nucleation passes drawn so far */
};

struct pf_ctx;

/** Device containers (src/type.h:58-107). */
struct CudaData {
	struct PhiloxStream prng;
	hipStream_t str_A, str_B, str_C, str_D;
	hipEvent_t ev_A, ev_B, ev_C, ev_D;

	fp_t* conc_Ni;
	fp_t* phi;

	fp_t* conc_Cr_old;
	fp_t* conc_Cr_new;

	fp_t* conc_Nb_old;
	fp_t* conc_Nb_new;

	fp_t* phi_del_old;
	fp_t* phi_del_new;

	fp_t* phi_lav_old;
	fp_t* phi_lav_new;

	/* libpfpa state */
	struct pf_ctx* ctx;
	int bc;              /* 0 no-flux (the reference's), 1 periodic */
	fp_t DCr[NC], DNb[NC];
	fp_t kappa[NP], omega[NP], Lmob[NP];
	fp_t mask[9];
	/* grouped launches: steps issued by device_evolution and not yet run */
	int group;           /* steps per launch (1, 3 or 4) */
	int pend;            /* pending steps (0..7) */
	fp_t* pend_old[4];   /* old and new fields of the first pending step */
	fp_t* pend_new[4];
	fp_t pend_alpha, pend_dt;
	int pend_nx, pend_ny;
};

#endif /* _TYPE_H_ */
