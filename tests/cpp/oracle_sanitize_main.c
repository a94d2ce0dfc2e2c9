/*
 * Drives every array-walking function of the CPU oracle (oracle/pf_oracle.c)
 * over edge-case grids: one interior cell, one interior row or column,
 * ragged sizes, both boundary modes, nucleation with embeds wider than a
 * cell.  Built with -fsanitize=address,undefined by tests/test_sanitize.py,
 * so any out-of-bounds access or undefined behaviour in the checker aborts.
 * Every array is allocated at exactly nx * ny doubles (no slack).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct {
	double dt;
	double DCr[2], DNb[2];
	double kappa[2], omega[2], Lmob[2];
	double alpha;
	double mask[9];
} orc_params;

void orc_set_diffusion_matrix(double xCr0, double xNb0, double ifce_width, double *DCr,
                              double *DNb, double *Lmob);
void orc_five_point_mask(double dx, double dy, double *mask);
void orc_step(int nx, int ny, int bc, const orc_params *P, const double *xCr, const double *xNb,
              const double *pd, const double *pl, double *oCr, double *oNb, double *opd,
              double *opl);
void orc_run(int nx, int ny, int bc, const orc_params *P, int nsteps, double *aCr, double *aNb,
             double *apd, double *apl, double *bCr, double *bNb, double *bpd, double *bpl);
int orc_nucleation(int nx, int ny, double *xCr, double *xNb, double *pd, double *pl,
                   double sigma_del, double sigma_lav, double unit_a, double dx_arg,
                   double dy_arg, double dz, double dt_arg, uint64_t seed, uint64_t call,
                   uint32_t instance, unsigned char *event_flags);
void orc_summarize(int nx, int ny, const double *xCr, const double *xNb, const double *pd,
                   const double *pl, double *out5);
void orc_energy(int nx, int ny, int bc, double dx, double dy, const double kappa[2],
                const double omega[2], double Omega, const double *xCr, const double *xNb,
                const double *pd, const double *pl, double *chem, double *grad, double out3[3]);
void orc_mobility_divergence(int nx, int ny, double dx, double dy, const double *const conc[6],
                             const double *const mobCr[8], const double *const mobNb[8],
                             const double *pd, const double *pl, double *outCr, double *outNb);

static double urand(uint64_t *s)
{
	*s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
	return (double)(*s >> 11) * (1.0 / 9007199254740992.0);
}

static double *field(size_t n, double lo, double hi, uint64_t *s)
{
	double *a = malloc(n * sizeof(double));
	for (size_t i = 0; i < n; i++)
		a[i] = lo + (hi - lo) * urand(s);
	return a;
}

int main(void)
{
	const int shapes[][2] = {{1, 1}, {1, 5}, {5, 1}, {2, 3}, {7, 4}, {13, 9}, {33, 20}};
	uint64_t seed = 12345;
	orc_params P;
	orc_set_diffusion_matrix(0.30, 0.02, 2.5e-9, P.DCr, P.DNb, P.Lmob);
	for (int k = 0; k < 2; k++) {
		P.kappa[k] = 3.0 * 0.13 * 2.5e-9 / 2.2;
		P.omega[k] = 18.0 * 0.13 * 0.13 / P.kappa[k];
	}
	P.alpha = 1.07e11;
	P.dt = 1e-7;
	orc_five_point_mask(0.25e-9, 0.3e-9, P.mask);
	long checks = 0;
	for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); s++) {
		const int nx = shapes[s][0] + 2, ny = shapes[s][1] + 2;
		const size_t n = (size_t)nx * ny;
		double *a[4] = {field(n, 0.28, 0.32, &seed), field(n, 0.01, 0.03, &seed),
		                field(n, 0.0, 1.0, &seed), field(n, 0.0, 1.0, &seed)};
		double *b[4];
		for (int k = 0; k < 4; k++)
			b[k] = malloc(n * sizeof(double));
		for (int bc = 0; bc < 2; bc++) {
			orc_step(nx, ny, bc, &P, a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]);
			orc_run(nx, ny, bc, &P, 3, a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]);
			double out3[3];
			double *chem = malloc(n * sizeof(double)), *grad = malloc(n * sizeof(double));
			orc_energy(nx, ny, bc, 0.25e-9, 0.3e-9, P.kappa, P.omega, P.alpha, a[0], a[1],
			           a[2], a[3], chem, grad, out3);
			free(chem);
			free(grad);
			checks += 3;
		}
		double out5[5];
		orc_summarize(nx, ny, a[0], a[1], a[2], a[3], out5);
		/* nucleation: P inside (0, 1) and embeds wider than a cell (a tiny
		 * mesh spacing and a huge interval), so the embed loop runs to the
		 * edges of the grid */
		double *c[4] = {field(n, 0.15, 0.3, &seed), field(n, 0.19, 0.24, &seed),
		                field(n, 0.0, 0.0, &seed), field(n, 0.0, 0.0, &seed)};
		unsigned char *flags = malloc(n);
		const double dx = 2e-13;
		orc_nucleation(nx, ny, c[0], c[1], c[2], c[3], 0.0105, 0.0108, 0.352e-9, dx * nx,
		               dx * ny, dx, 3e38, 77, 2, 0, flags);
		orc_nucleation(nx, ny, c[0], c[1], c[2], c[3], 0.13, 0.13, 0.352e-9, 0.25e-9 * nx,
		               0.25e-9 * ny, 0.25e-9, 1e-4, 42, 0, 0, flags);
		free(flags);
		/* variable-mobility divergence */
		double *conc[6], *mc[8], *mn[8];
		for (int k = 0; k < 6; k++)
			conc[k] = field(n, 0.0, 0.5, &seed);
		for (int k = 0; k < 8; k++) {
			mc[k] = field(n, 0.5, 2.0, &seed);
			mn[k] = field(n, 0.5, 2.0, &seed);
		}
		double *oc = malloc(n * sizeof(double)), *on = malloc(n * sizeof(double));
		orc_mobility_divergence(nx, ny, 0.5, 0.4, (const double *const *)conc,
		                        (const double *const *)mc, (const double *const *)mn, a[2], a[3],
		                        oc, on);
		free(oc);
		free(on);
		for (int k = 0; k < 6; k++)
			free(conc[k]);
		for (int k = 0; k < 8; k++) {
			free(mc[k]);
			free(mn[k]);
		}
		for (int k = 0; k < 4; k++) {
			free(a[k]);
			free(b[k]);
			free(c[k]);
		}
		checks += 4;
	}
	printf("oracle sanitizer run: %ld calls clean\n", checks);
	return 0;
}
