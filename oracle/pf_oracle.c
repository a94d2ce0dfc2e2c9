/*
 * pf_oracle.c -- CPU oracle for the ternary Cr-Nb-Ni KKS phase-field step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * file; only tests/, the smoke check and bench.py's cpu_baseline leg use it, and
 * only as the checker or as the timed CPU baseline.
 *
 * It is a plain-C FP64 restatement of the reference step, written from the
 * reference semantics (each function cites the reference file:line it follows,
 * paths relative to the reference repository root):
 *
 *   pointwise thermodynamics   thermo/parabola625.c (constants and functions)
 *   nucleation probability     thermo/nucleation.c:52-81
 *   step sequence              src/main.cpp:289-321
 *   boundary ghost copy        src/discretization.cu:360-394
 *   5-point Laplacian          src/discretization.cu:550-603, src/mesh.cpp:187-196
 *   Cahn-Hilliard update       src/discretization.cu:843-886
 *   Allen-Cahn update          src/discretization.cu:888-980
 *   nucleation + embedding     src/discretization.cu:1046-1177
 *   diffusion matrix / dt      src/alloy625.cpp:80-121, src/main.cpp:259-268
 *   field summary              src/alloy625.cpp:811-863
 *   variable-mobility CH       src/discretization.cu:605-778 (dead code there:
 *   divergence                 its launch is commented out; mobilities are inputs)
 *
 * Pinning (what checks each function; see tests/):
 *   p, pPrime, g_*, dg_*, fict_*, D_*  golden vectors from the compiled reference
 *                                      thermo/parabola625.c plus closed forms
 *   nucleation probability            thermo/nucleation.c cannot be compiled here
 *                                      (it includes src/type.h -> curand_kernel.h);
 *                                      pinned by textbook classical-nucleation
 *                                      identities and the Poisson zero-event law
 *   step                              literal replay of the reference kernel
 *                                      sequence on the compiled reference thermo,
 *                                      mass conservation, fixed points, Laplacian
 *                                      on quadratics and Fourier modes
 *   Philox4x32-10                     Random123 known-answer vectors
 *   nucleation uniforms               word order, shift and (0,1] offset computed
 *                                     exactly from the known-answer words
 *   diffusion matrix, mobility L      golden D_* at (0.30, 0.02); L against the
 *                                     closed form on those values, 1/w^2 scaling
 *                                     and the hand-evaluated 4.0015e-8
 *   field summary                     uniform interiors (exact p(0), p(1/2), p(1)),
 *                                     NaN / 1e300 ghost rings, dyadic fields with
 *                                     hand-computed means
 *   mobility divergence               constant mobility = M times the five-point
 *                                     Laplacian; linear mobility on quadratic fields
 *                                     = the continuous divergence (exact for the
 *                                     face average); zero interior sum with no-flux
 *                                     and periodic ghosts; x-only phase terms
 *
 * Array layout: the reference's, (nx = Nx+2) x (ny = Ny+2) row-major FP64 with a
 * one-cell ghost ring, idx = nx*y + x.  Interior cells are 1..nx-2 by 1..ny-2.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Model constants (thermo/parabola625.c)                                     */
/* ------------------------------------------------------------------------- */

/* kT, RT, Vm: thermo/parabola625.c:38-63 */
static const double ORC_KT = 1.5782889043500002e-20;
static const double ORC_RT = 9504.6840941999999;
static const double ORC_VM = 1.0000000000000001e-5;

/* equilibrium compositions xe_{gam,del,lav}_{Cr,Nb}: thermo/parabola625.c:65-117 */
static const double ORC_XE[3][2] = {
	{0.52421634830562147, 0.01299272922003303},   /* gamma  */
	{0.022966218927631978, 0.24984563695705883},  /* delta  */
	{0.37392129441013022, 0.25826261799015571}    /* Laves  */
};

/* paraboloid Hessians d2g_{phase}_dx{CrCr,CrNb,NbNb}: thermo/parabola625.c:1305-1411 */
static const double ORC_H[3][3] = {
	{4534264425.5240498, 15095482346.486225, 110386166481.37187},
	{42692985981.597771, 16906497386.287418, 6170738265863.7695},
	{17733460569.613991, 24191004361.532181, 196588620559.76782}
};

/* interfacial energies s_delta, s_laves: thermo/parabola625.c:299-316 */
static const double ORC_SIGMA[2] = {0.13, 0.13};

/* FCC Ni lattice constant used inside the nucleation rate: src/parameters.h:27 */
static const double ORC_LATTICE = 0.352e-9;

/* Mobility model behind D_{CrCr,CrNb,NbCr,NbNb} (thermo/parabola625.c:1860-1894).
 * D_ij = sum_k (delta_ik - x_i) x_k M_k(x) dmu_k/dx_j with the gamma paraboloid
 * in molar units; M_k = c_k exp(Q_k(x)).  Prefactors c_k and the exponent
 * polynomials Q_k are the reference's tabulated values. */
static const double ORC_MOB_C[3] = {3.5027570743586952e-21,  /* Cr */
                                    1.4170990277916469e-20,  /* Nb */
                                    1.8295676400933012e-21}; /* Ni */

static double orc_mob_exp(int k, double xCr, double xNb)
{
	const double xNi = 1.0 - xCr - xNb;
	switch (k) {
	case 0:
		return exp(-7.1543718172640256 * xCr * xNi + 3.3541967644499233 * xCr);
	case 1:
		return exp(-34.982563536715503 * xNb * xNi + 14.437947842988088 * xNb);
	default:
		return exp(-8.5221193705645018 * xCr * xNi + 4.0036674816806448 * xCr
		           - 442.62877981749585 * xNb * xNi - 0.82164267100913768 * xNb);
	}
}

/* ------------------------------------------------------------------------- */
/* Pointwise thermodynamics                                                   */
/* ------------------------------------------------------------------------- */

double orc_kT(void) { return ORC_KT; }
double orc_RT(void) { return ORC_RT; }
double orc_Vm(void) { return ORC_VM; }
double orc_xe(int phase, int comp) { return ORC_XE[phase][comp]; }
double orc_d2g(int phase, int which) { return ORC_H[phase][which]; }
double orc_sigma(int which) { return ORC_SIGMA[which]; }

/* interpolator p(x) = x^3 (6x^2 - 15x + 10): thermo/parabola625.c:11-18 */
double orc_p(double x)
{
	return x * x * x * (6.0 * x * x - 15.0 * x + 10.0);
}

/* p'(x) = 30 x^2 (1-x)^2: thermo/parabola625.c:20-27 */
double orc_pPrime(double x)
{
	return 30.0 * x * x * (1.0 - x) * (1.0 - x);
}

/* paraboloid free energy g_a(x) = 1/2 (x-xe)^T H (x-xe): thermo/parabola625.c:418-443 */
double orc_g(int phase, double xCr, double xNb)
{
	const double a = xCr - ORC_XE[phase][0];
	const double b = xNb - ORC_XE[phase][1];
	const double *h = ORC_H[phase];
	return 0.5 * h[0] * a * a + h[1] * a * b + 0.5 * h[2] * b * b;
}

/* gradient dg_a/dx_comp: thermo/parabola625.c:655-707 */
double orc_dg(int phase, int comp, double xCr, double xNb)
{
	const double a = xCr - ORC_XE[phase][0];
	const double b = xNb - ORC_XE[phase][1];
	const double *h = ORC_H[phase];
	return comp == 0 ? h[0] * a + h[1] * b : h[1] * a + h[2] * b;
}

/* KKS fictitious compositions (thermo/parabola625.c:245-297).
 * The reference evaluates sympy-expanded rational functions of
 * (XCR, XNB, pDel, pGam, pLav).  They solve
 *     x = pDel x_del + pGam x_gam + pLav x_lav,
 *     H_gam (x_gam - xe_gam) = H_del (x_del - xe_del) = H_lav (x_lav - xe_lav) = mu,
 * with pGam independent.  Restated as the 2x2 solve
 *     M = sum_a p_a H_a^-1,  mu = M^-1 (x - sum_a p_a xe_a),  x_a = xe_a + H_a^-1 mu.
 * out[] = {gam_Cr, gam_Nb, del_Cr, del_Nb, lav_Cr, lav_Nb}. */
void orc_fict(double xCr, double xNb, double pDel, double pGam, double pLav, double *out)
{
	double A[3][3];
	for (int a = 0; a < 3; a++) {
		const double det = ORC_H[a][0] * ORC_H[a][2] - ORC_H[a][1] * ORC_H[a][1];
		A[a][0] = ORC_H[a][2] / det;
		A[a][1] = -ORC_H[a][1] / det;
		A[a][2] = ORC_H[a][0] / det;
	}
	const double w[3] = {pGam, pDel, pLav};
	double M0 = 0., M1 = 0., M2 = 0., r0 = xCr, r1 = xNb;
	for (int a = 0; a < 3; a++) {
		M0 += w[a] * A[a][0];
		M1 += w[a] * A[a][1];
		M2 += w[a] * A[a][2];
		r0 -= w[a] * ORC_XE[a][0];
		r1 -= w[a] * ORC_XE[a][1];
	}
	const double detM = M0 * M2 - M1 * M1;
	const double mu0 = (M2 * r0 - M1 * r1) / detM;
	const double mu1 = (M0 * r1 - M1 * r0) / detM;
	for (int a = 0; a < 3; a++) {
		out[2 * a + 0] = ORC_XE[a][0] + A[a][0] * mu0 + A[a][1] * mu1;
		out[2 * a + 1] = ORC_XE[a][1] + A[a][1] * mu0 + A[a][2] * mu1;
	}
}

/* Diffusivity matrix entry D_{ij}(x), i,j in {0: Cr, 1: Nb}:
 * thermo/parabola625.c:1860-1894 (mobility model restated, see above). */
double orc_D(int i, int j, double xCr, double xNb)
{
	const double h0 = ORC_H[0][0] * ORC_VM, h1 = ORC_H[0][1] * ORC_VM, h2 = ORC_H[0][2] * ORC_VM;
	const double x[3] = {xCr, xNb, 1.0 - xCr - xNb};
	/* dmu_k/dx_j for k = Cr, Nb, Ni (gamma paraboloid, molar) */
	double dmu[3][2];
	dmu[0][0] = (1.0 - xCr) * h0 - xNb * h1;
	dmu[0][1] = (1.0 - xCr) * h1 - xNb * h2;
	dmu[1][0] = (1.0 - xNb) * h1 - xCr * h0;
	dmu[1][1] = (1.0 - xNb) * h2 - xCr * h1;
	dmu[2][0] = -xCr * h0 - xNb * h1;
	dmu[2][1] = -xCr * h1 - xNb * h2;
	double D = 0.0;
	for (int k = 0; k < 3; k++) {
		const double kron = (i == k) ? 1.0 : 0.0;
		D += (kron - x[i]) * x[k] * ORC_MOB_C[k] * orc_mob_exp(k, xCr, xNb) * dmu[k][j];
	}
	return D;
}

/* Classical nucleation probability for a spherical nucleus:
 * thermo/nucleation.c:52-81 (device twin src/nucleation.cu:30-70).
 * Note the reference's asymmetric signs of dc_Cr and dc_Nb (kept). */
void orc_nucleation_probability(double xCr, double xNb, double dG_chem,
                                double D_CrCr, double D_NbNb, double sigma,
                                double Vatom, double n_gam, double dV, double dt,
                                double *Rstar, double *P_nuc)
{
	const double sigma_cubed = sigma * sigma * sigma;
	const double Zeldov = (Vatom * dG_chem * dG_chem) / (8. * M_PI * sqrt(ORC_KT * sigma_cubed));
	const double Gstar = (16. * M_PI * sigma_cubed) / (3. * dG_chem * dG_chem);
	const double a4 = ORC_LATTICE * ORC_LATTICE * ORC_LATTICE * ORC_LATTICE;
	const double BstarCr = (3. * Gstar * D_CrCr * xCr) / (sigma * a4);
	const double BstarNb = (3. * Gstar * D_NbNb * xNb) / (sigma * a4);
	const double k1Cr = BstarCr * Zeldov * n_gam;
	const double k1Nb = BstarNb * Zeldov * n_gam;
	const double k2 = Gstar / ORC_KT;
	const double dc_Cr = -xCr + ORC_XE[0][0];
	const double dc_Nb = xNb - ORC_XE[0][1];
	const double JCr = k1Cr * exp(-k2 / dc_Cr);
	const double JNb = k1Nb * exp(-k2 / dc_Nb);
	const double P0 = exp(-dt * dV * (JCr + JNb));
	*P_nuc = 1. - P0;
	*Rstar = (2. * sigma) / dG_chem;
}

/* ------------------------------------------------------------------------- */
/* Host-side setup (src/alloy625.cpp:80-121, src/main.cpp:259-268, mesh.cpp)  */
/* ------------------------------------------------------------------------- */

/* set_diffusion_matrix: src/alloy625.cpp:80-109 (MobStab = 1) */
void orc_set_diffusion_matrix(double xCr0, double xNb0, double ifce_width,
                              double *DCr, double *DNb, double *Lmob)
{
	DCr[0] = orc_D(0, 0, xCr0, xNb0);
	DCr[1] = orc_D(0, 1, xCr0, xNb0);
	DNb[0] = orc_D(1, 0, xCr0, xNb0);
	DNb[1] = orc_D(1, 1, xCr0, xNb0);
	const double Dnorm = sqrt(DCr[0] * DCr[0] + DNb[1] * DNb[1]);
	Lmob[0] = 1.0 * Dnorm / (ifce_width * ifce_width * ORC_RT / ORC_VM);
	Lmob[1] = Lmob[0];
}

/* five_point_Laplacian_stencil: src/mesh.cpp:187-196, mask[3][3] row-major */
void orc_five_point_mask(double dx, double dy, double *mask)
{
	for (int i = 0; i < 9; i++)
		mask[i] = 0.0;
	mask[0 * 3 + 1] = 1. / (dy * dy);
	mask[1 * 3 + 0] = 1. / (dx * dx);
	mask[1 * 3 + 1] = -2. * (dx * dx + dy * dy) / (dx * dx * dy * dy);
	mask[1 * 3 + 2] = 1. / (dx * dx);
	mask[2 * 3 + 1] = 1. / (dy * dy);
}

/* ------------------------------------------------------------------------- */
/* Time step                                                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
	double dt;
	double DCr[2], DNb[2];
	double kappa[2], omega[2], Lmob[2];
	double alpha;      /* three-phase coefficient Omega */
	double mask[9];
} orc_params;

/* boundary mode: 0 = no-flux (ghost := adjacent interior, as boundary_kernel
 * does at src/discretization.cu:371-393), 1 = periodic. */
static inline int orc_bc(int i, int n, int bc)
{
	/* i is a padded index in [0, n+1]; interior is 1..n */
	if (i >= 1 && i <= n)
		return i;
	if (bc == 0)
		return i < 1 ? 1 : n;
	return i < 1 ? i + n : i - n;
}

/*
 * One explicit Euler step of the reference sequence (src/main.cpp:289-305):
 *   boundaries(old) -> fictitious(old) -> boundaries(fict) -> Laplacians ->
 *   boundaries(lap) -> Cahn-Hilliard || Allen-Cahn.
 * Ghost cells are realised by index mapping (no-flux clamp or periodic wrap);
 * for interior outputs the clamp is exactly the reference ghost copy.
 * Interior cells of out[] are written; ghost cells of out[] are set to the
 * value of the cell they mirror (what the reference leaves there on edges).
 * in/out: {conc_Cr, conc_Nb, phi_del, phi_lav}; nx, ny include the ghost ring.
 */
void orc_step(int nx, int ny, int bc, const orc_params *P,
              const double *xCr, const double *xNb, const double *pd, const double *pl,
              double *oCr, double *oNb, double *opd, double *opl)
{
	const int Nx = nx - 2, Ny = ny - 2;
	const size_t n = (size_t)nx * ny;
	double *gCr = (double *)malloc(n * sizeof(double));
	double *gNb = (double *)malloc(n * sizeof(double));

	/* fictitious_gam_kernel over the interior: src/discretization.cu:50-74 */
	#pragma omp parallel for schedule(static)
	for (int y = 1; y <= Ny; y++) {
		for (int x = 1; x <= Nx; x++) {
			const size_t i = (size_t)nx * y + x;
			const double pDel = orc_p(pd[i]);
			const double pLav = orc_p(pl[i]);
			const double pGam = 1.0 - pDel - pLav;
			double f[6];
			orc_fict(xCr[i], xNb[i], pDel, pGam, pLav, f);
			gCr[i] = f[0];
			gNb[i] = f[1];
		}
	}

	const double *m = P->mask;
	#pragma omp parallel for schedule(static)
	for (int y = 1; y <= Ny; y++) {
		for (int x = 1; x <= Nx; x++) {
			const size_t c = (size_t)nx * y + x;
			size_t nb[9];
			for (int j = 0; j < 3; j++)
				for (int k = 0; k < 3; k++)
					nb[j * 3 + k] = (size_t)nx * orc_bc(y + j - 1, Ny, bc) + orc_bc(x + k - 1, Nx, bc);

			/* convolution_kernel accumulation order: src/discretization.cu:589-595 */
			double lap[4] = {0., 0., 0., 0.};
			const double *src[4] = {pd, pl, gCr, gNb};
			for (int f = 0; f < 4; f++) {
				double v = 0.;
				for (int j = 0; j < 9; j++)
					v += m[j] * src[f][nb[j]];
				lap[f] = v;
			}

			/* composition_kernel: src/discretization.cu:859-863 */
			const double dt = P->dt;
			oCr[c] = xCr[c] + dt * (P->DCr[0] * lap[2] + P->DCr[1] * lap[3]);
			oNb[c] = xNb[c] + dt * (P->DNb[0] * lap[2] + P->DNb[1] * lap[3]);

			/* allen_cahn_kernel + delta_kernel + laves_kernel:
			 * src/discretization.cu:888-980 */
			const double phd = pd[c], phl = pl[c];
			const double pDel = orc_p(phd);
			const double pLav = orc_p(phl);
			const double pGam = 1.0 - pDel - pLav;
			double f[6];
			orc_fict(xCr[c], xNb[c], pDel, pGam, pLav, f);
			const double gam_nrg = orc_g(0, f[0], f[1]);
			const double dgGdxCr = orc_dg(0, 0, f[0], f[1]);
			const double dgGdxNb = orc_dg(0, 1, f[0], f[1]);

			const double del_nrg = orc_g(1, f[2], f[3]);
			const double gp_del = gam_nrg - del_nrg - dgGdxCr * (f[0] - f[2]) - dgGdxNb * (f[1] - f[3]);
			const double dF_del = -orc_pPrime(phd) * gp_del
			                      + 2.0 * P->omega[0] * phd * (phd - 1.0) * (2.0 * phd - 1.0)
			                      + 2.0 * P->alpha * phd * phl * phl
			                      - P->kappa[0] * lap[0];
			opd[c] = phd - dt * P->Lmob[0] * dF_del;

			const double lav_nrg = orc_g(2, f[4], f[5]);
			const double gp_lav = gam_nrg - lav_nrg - dgGdxCr * (f[0] - f[4]) - dgGdxNb * (f[1] - f[5]);
			const double dF_lav = -orc_pPrime(phl) * gp_lav
			                      + 2.0 * P->omega[1] * phl * (phl - 1.0) * (2.0 * phl - 1.0)
			                      + 2.0 * P->alpha * phl * phd * phd
			                      - P->kappa[1] * lap[1];
			opl[c] = phl - dt * P->Lmob[1] * dF_lav;
		}
	}

	/* ghost ring of the outputs: mirror image of the interior */
	double *outs[4] = {oCr, oNb, opd, opl};
	for (int f = 0; f < 4; f++) {
		double *o = outs[f];
		for (int y = 0; y < ny; y++) {
			for (int x = 0; x < nx; x++) {
				if (x >= 1 && x <= Nx && y >= 1 && y <= Ny)
					continue;
				o[(size_t)nx * y + x] = o[(size_t)nx * orc_bc(y, Ny, bc) + orc_bc(x, Nx, bc)];
			}
		}
	}
	free(gCr);
	free(gNb);
}

/* Run nsteps steps, ping-ponging between the two buffer sets; the result ends
 * in a[] (the buffers are swapped as src/main.cpp:318-321 does). */
void orc_run(int nx, int ny, int bc, const orc_params *P, int nsteps,
             double *aCr, double *aNb, double *apd, double *apl,
             double *bCr, double *bNb, double *bpd, double *bpl)
{
	const size_t n = (size_t)nx * ny;
	double *src[4] = {aCr, aNb, apd, apl}, *dst[4] = {bCr, bNb, bpd, bpl};
	for (int s = 0; s < nsteps; s++) {
		orc_step(nx, ny, bc, P, src[0], src[1], src[2], src[3], dst[0], dst[1], dst[2], dst[3]);
		for (int k = 0; k < 4; k++) {
			double *t = src[k];
			src[k] = dst[k];
			dst[k] = t;
		}
	}
	if (src[0] != aCr)
		for (int k = 0; k < 4; k++)
			memcpy(dst[k], src[k], n * sizeof(double));
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                          */
/* ------------------------------------------------------------------------- */

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
	uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
	uint32_t k0 = key_in[0], k1 = key_in[1];
	for (int r = 0; r < 10; r++) {
		const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
		const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
		const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
		const uint32_t n1 = (uint32_t)p1;
		const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
		const uint32_t n3 = (uint32_t)p0;
		c0 = n0; c1 = n1; c2 = n2; c3 = n3;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Nucleation stream: key = seed, counter = (cell index, instance, call lo, call hi).
 * Two 53-bit uniforms in (0, 1] per cell per nucleation call: u1 (delta test)
 * from words 0,1 and u2 (Laves test) from words 2,3. */
void orc_nuc_uniforms(uint64_t seed, uint64_t call, uint32_t instance, uint32_t cell,
                      double *u1, double *u2)
{
	const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
	const uint32_t ctr[4] = {cell, instance, (uint32_t)call, (uint32_t)(call >> 32)};
	uint32_t r[4];
	orc_philox4x32_10(ctr, key, r);
	const uint64_t a = (((uint64_t)r[1] << 32) | r[0]) >> 11;
	const uint64_t b = (((uint64_t)r[3] << 32) | r[2]) >> 11;
	*u1 = (double)(a + 1) * (1.0 / 9007199254740992.0);
	*u2 = (double)(b + 1) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------- */
/* Nucleation (src/discretization.cu:1046-1177)                               */
/* ------------------------------------------------------------------------- */

typedef struct {
	int cell;      /* padded index nx*y + x */
	int phase;     /* 0 = delta, 1 = Laves */
	double xCr, xNb, R_precip;
} orc_event;

/* embed_OPC_kernel: src/discretization.cu:1046-1077.  phi points at the
 * phase's own field (the Laves call passes d_phi_lav for both phase
 * arguments, :1168-1169, so it sets phi_lav). */
static void orc_embed(double *xCr, double *xNb, double *phi, int nx, int ny,
                      int x, int y, double cxCr, double cxNb,
                      double par_xe_Cr, double par_xe_Nb, double R_precip)
{
	const double R_dep_Cr = R_precip * sqrt((par_xe_Cr - cxCr) / (cxCr - ORC_XE[0][0]));
	const double R_dep_Nb = R_precip * sqrt((par_xe_Nb - cxNb) / (cxNb - ORC_XE[0][1]));
	for (int i = (int)(-R_precip); i < R_precip; i++) {
		for (int j = (int)(-R_precip); j < R_precip; j++) {
			const long idn = (long)nx * (y + j) + (x + i);
			const double r = sqrt((double)(i * i + j * j));
			if (idn >= 0 && idn < (long)nx * ny) {
				if (r <= R_precip) {
					xCr[idn] = par_xe_Cr;
					xNb[idn] = par_xe_Nb;
					phi[idn] = 1.;
				} else {
					if (r <= R_dep_Cr)
						xCr[idn] = ORC_XE[0][0];
					if (r <= R_dep_Nb)
						xNb[idn] = ORC_XE[0][1];
				}
			}
		}
	}
}

/*
 * One nucleation pass on the *_new fields (src/main.cpp:307-316 ->
 * src/discretization.cu:1079-1177) over the interior cells.
 * dx_arg, dy_arg are what the kernel receives (the reference passes
 * meshres*nx and meshres*ny, :1206); dt_arg is nuc_interval*dt.
 * Deterministic two-phase form: all decisions are taken on the fields as
 * they are on entry, then embeds are applied: those confined to their own
 * cell first, then wider ones, each in cell order (delta before Laves).  For
 * scan and embed radii below one cell (every realistic domain) this is
 * identical to the reference kernel, which races on wider ones.
 * Returns the number of events; event_flags (optional, nx*ny bytes) gets
 * bit0 = delta event, bit1 = Laves event per cell.
 */
int orc_nucleation(int nx, int ny, double *xCr, double *xNb, double *pd, double *pl,
                   double sigma_del, double sigma_lav, double unit_a,
                   double dx_arg, double dy_arg, double dz, double dt_arg,
                   uint64_t seed, uint64_t call, uint32_t instance,
                   unsigned char *event_flags)
{
	const double dV = dx_arg * dy_arg * dz;
	const double Vatom = 0.25 * unit_a * unit_a * unit_a;
	const double n_gam = dV / Vatom;
	const double R = 1.75e-9 / dx_arg;
	const int Nx = nx - 2, Ny = ny - 2;
	const size_t n = (size_t)nx * ny;
	orc_event *ev = (orc_event *)malloc(2 * n * sizeof(orc_event));
	int nev = 0;
	if (event_flags)
		memset(event_flags, 0, n);

	for (int y = 1; y <= Ny; y++) {
		for (int x = 1; x <= Nx; x++) {
			double phi_pre = 0.;
			for (int i = (int)(-R); i < R; i++) {
				for (int j = (int)(-R); j < R; j++) {
					const long idn = (long)nx * (y + j) + (x + i);
					if (idn >= 0 && idn < (long)n && i * i + j * j < R * R) {
						const double v = orc_p(pd[idn]) + orc_p(pl[idn]);
						phi_pre = v > phi_pre ? v : phi_pre;
					}
				}
			}
			if (!(phi_pre < 1e-10))
				continue;
			const int idx = nx * y + x;
			const double cCr = xCr[idx], cNb = xNb[idx];
			double u1, u2;
			orc_nuc_uniforms(seed, call, instance, (uint32_t)idx, &u1, &u2);
			const double DCrCr = orc_D(0, 0, cCr, cNb);
			const double DNbNb = orc_D(1, 1, cCr, cNb);
			/* driving force = g_gam(x) for both phases (thermo/nucleation.c:40-50) */
			const double dG = orc_g(0, cCr, cNb);
			const double sig[2] = {sigma_del, sigma_lav};
			for (int ph = 0; ph < 2; ph++) {
				double Rstar, Pn;
				orc_nucleation_probability(cCr, cNb, dG, DCrCr, DNbNb, sig[ph], Vatom, n_gam, dV,
				                           dt_arg, &Rstar, &Pn);
				if (Rstar > 0.) {
					const double u = ph == 0 ? u1 : u2;
					if (Pn - u > 0) {
						ev[nev].cell = idx;
						ev[nev].phase = ph;
						ev[nev].xCr = cCr;
						ev[nev].xNb = cNb;
						ev[nev].R_precip = Rstar / dx_arg;
						nev++;
						if (event_flags)
							event_flags[idx] |= (unsigned char)(1u << ph);
					}
				}
			}
		}
	}
	/* apply: embeds confined to their own cell first, then wider ones; each
	 * pass in (cell, phase) order, the order the decisions were recorded in */
	for (int pass = 0; pass < 2; pass++) {
		for (int e = 0; e < nev; e++) {
			const int wide = ev[e].R_precip >= 1.0;
			if (wide != pass)
				continue;
			const int x = ev[e].cell % nx, y = ev[e].cell / nx;
			const int ph = ev[e].phase;
			orc_embed(xCr, xNb, ph == 0 ? pd : pl, nx, ny, x, y, ev[e].xCr, ev[e].xNb,
			          ORC_XE[1 + ph][0], ORC_XE[1 + ph][1], ev[e].R_precip);
		}
	}
	free(ev);
	return nev;
}

/* ------------------------------------------------------------------------- */
/* Field summary (src/alloy625.cpp:811-863)                                   */
/* ------------------------------------------------------------------------- */

/* Means over interior cells of {x_Cr, x_Nb, gamma, delta, Laves}; long double
 * accumulation (the reference's OpenMP 'critical' sum has no fixed order). */
void orc_summarize(int nx, int ny, const double *xCr, const double *xNb,
                   const double *pd, const double *pl, double *out5)
{
	long double s[5] = {0, 0, 0, 0, 0};
	for (int y = 1; y <= ny - 2; y++) {
		for (int x = 1; x <= nx - 2; x++) {
			const size_t i = (size_t)nx * y + x;
			const double d = orc_p(pd[i]), l = orc_p(pl[i]);
			s[0] += xCr[i];
			s[1] += xNb[i];
			s[2] += 1.0 - d - l;
			s[3] += d;
			s[4] += l;
		}
	}
	const long double N = (long double)(nx - 2) * (ny - 2);
	for (int k = 0; k < 5; k++)
		out5[k] = (double)(s[k] / N);
}

/* Free energy of the fields (src/alloy625.cpp:752-784 gibbs, :786-806
 * maskedgradient, :865-919 summarize_energy), restated literally: per
 * interior cell
 *   chem = (pGam g_gam(x_gam) + pDel g_del(x_del) + pLav g_lav(x_lav)
 *           + sum_i omega_i phi_i^2 (1 - phi_i)^2 + 2 Omega phi_del^2 phi_lav^2) dV
 *   grad = sum_i kappa_i |grad phi_i|^2 dV,
 * fictitious compositions from the KKS solve (orc_fict), gradients by central
 * differences (phi(x+1) - phi(x-1)) / (2 dx) with the boundary map for the
 * neighbours (no-flux: the MMSP Neumann ghost copy; periodic: wrap), dV = dx dy.
 * chem[] and grad[] (nx*ny, may be NULL) receive the per-cell values at
 * interior cells; out3 = {total, sum chem, sum grad} (long double sums). */
void orc_energy(int nx, int ny, int bc, double dx, double dy, const double kappa[2],
                const double omega[2], double Omega, const double *xCr, const double *xNb,
                const double *pd, const double *pl, double *chem, double *grad, double out3[3])
{
	const int Nx = nx - 2, Ny = ny - 2;
	const double dV = dx * dy;
	long double sc = 0.0L, sg = 0.0L;
	for (int y = 1; y <= Ny; y++) {
		for (int x = 1; x <= Nx; x++) {
			const size_t i = (size_t)nx * y + x;
			const double pDel = orc_p(pd[i]), pLav = orc_p(pl[i]);
			const double pGam = 1.0 - pDel - pLav;
			double f[6];
			orc_fict(xCr[i], xNb[i], pDel, pGam, pLav, f);
			double g = pGam * orc_g(0, f[0], f[1]);
			g += pDel * orc_g(1, f[2], f[3]);
			g += pLav * orc_g(2, f[4], f[5]);
			const double phi[2] = {pd[i], pl[i]};
			for (int k = 0; k < 2; k++)
				g += omega[k] * (phi[k] * phi[k]) * pow(1.0 - phi[k], 2);
			g += 2.0 * Omega * (phi[0] * phi[0]) * (phi[1] * phi[1]);
			const int xl = orc_bc(x - 1, Nx, bc), xh = orc_bc(x + 1, Nx, bc);
			const int yl = orc_bc(y - 1, Ny, bc), yh = orc_bc(y + 1, Ny, bc);
			const double *fl[2] = {pd, pl};
			double ge = 0.0;
			for (int k = 0; k < 2; k++) {
				const double gxv = (fl[k][(size_t)nx * y + xh] - fl[k][(size_t)nx * y + xl]) / (2.0 * dx);
				const double gyv = (fl[k][(size_t)nx * yh + x] - fl[k][(size_t)nx * yl + x]) / (2.0 * dy);
				ge += kappa[k] * (gxv * gxv + gyv * gyv);
			}
			const double c = g * dV, gg = ge * dV;
			if (chem)
				chem[i] = c;
			if (grad)
				grad[i] = gg;
			sc += c;
			sg += gg;
		}
	}
	out3[0] = (double)(sc + sg);
	out3[1] = (double)sc;
	out3[2] = (double)sg;
}

/* ------------------------------------------------------------------------- */
/* Variable-mobility Cahn-Hilliard divergence (src/discretization.cu:605-778) */
/* ------------------------------------------------------------------------- */

/* discrete_laplacian (src/discretization.cu:605-616): the face-averaged
 * 1-D second difference ((D_hi + D_mid)(c_hi - c_mid)
 * - (D_mid + D_low)(c_mid - c_low)) / (2 h^2). */
static double orc_face_lap(double Dl, double Dm, double Dh, double cl, double cm, double ch,
                           double h)
{
	return ((Dh + Dm) * (ch - cm) - (Dm + Dl) * (cm - cl)) / (2.0 * h * h);
}

/* chemical_convolution_Cr_kernel / _Nb_kernel (src/discretization.cu:618-778),
 * evaluated at every cell of the padded (nx x ny) arrays, ghost ring
 * included, with the reference's index clamp at the array edges (the left /
 * bottom neighbour of column / row 0 and the right / top neighbour of the
 * last column / row is the cell itself).  Per component i (Cr, Nb) the sum
 * runs in the reference's order: terms 1-6, each in x then in y, over
 * (mob[i][2a+j], conc[2a+j]) for phase a = gamma, delta, Laves and component
 * j = Cr, Nb; then terms 7-8, (mob[i][6], phi_del) and (mob[i][7], phi_lav),
 * in x ONLY (the reference has no y part for them, :688-695 and :769-776).
 *   conc[6]    = x_gam_Cr, x_gam_Nb, x_del_Cr, x_del_Nb, x_lav_Cr, x_lav_Nb
 *   mob[0][8]  = gam_CrCr, gam_CrNb, del_CrCr, del_CrNb, lav_CrCr, lav_CrNb,
 *                phi_del_Cr, phi_lav_Cr
 *   mob[1][8]  = gam_NbCr, gam_NbNb, del_NbCr, del_NbNb, lav_NbCr, lav_NbNb,
 *                phi_del_Nb, phi_lav_Nb
 * out[0] / out[1] receive div(M grad mu) for Cr / Nb (what the reference
 * stores in conc_Cr_new / conc_Nb_new). */
void orc_mobility_divergence(int nx, int ny, double dx, double dy, const double *const conc[6],
                             const double *const mobCr[8], const double *const mobNb[8],
                             const double *pd, const double *pl, double *outCr, double *outNb)
{
	const double *const *mob[2] = {mobCr, mobNb};
	double *out[2] = {outCr, outNb};
	#pragma omp parallel for schedule(static)
	for (int y = 0; y < ny; y++) {
		for (int x = 0; x < nx; x++) {
			const size_t mid = (size_t)nx * y + x;
			const size_t lft = (x == 0) ? mid : mid - 1;
			const size_t rgt = (x == nx - 1) ? mid : mid + 1;
			const size_t bot = (y == 0) ? mid : mid - nx;
			const size_t top = (y == ny - 1) ? mid : mid + nx;
			for (int i = 0; i < 2; i++) {
				double s = 0.0;
				for (int t = 0; t < 6; t++) {
					const double *M = mob[i][t], *c = conc[t];
					s += orc_face_lap(M[lft], M[mid], M[rgt], c[lft], c[mid], c[rgt], dx);
					s += orc_face_lap(M[bot], M[mid], M[top], c[bot], c[mid], c[top], dy);
				}
				const double *phi[2] = {pd, pl};
				for (int t = 0; t < 2; t++) {
					const double *M = mob[i][6 + t], *c = phi[t];
					s += orc_face_lap(M[lft], M[mid], M[rgt], c[lft], c[mid], c[rgt], dx);
				}
				out[i][mid] = s;
			}
		}
	}
}

int orc_num_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

/* Threads of the following parallel loops (the CPU baseline's single-core leg). */
void orc_set_threads(int n)
{
#ifdef _OPENMP
	if (n > 0) omp_set_num_threads(n);
#else
	(void)n;
#endif
}
