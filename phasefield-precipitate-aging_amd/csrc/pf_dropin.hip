// pf_dropin.hip -- the reference's C++ host API for the step path
// (src/mesh.h, src/data.cuh, src/discretization.cuh), implemented over the
// C-ABI of include/pfpa.h so both front ends run the same kernels.
#include <hip/hip_runtime.h>

#include <assert.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "data.cuh"
#include "discretization.cuh"
#include "mesh.h"
#include "pf_internal.h"  // symmetric_five_point: the masks the multi-step entry points take
#include "pfpa.h"

// ---------------------------------------------------------------- errors

hipError_t checkCuda(hipError_t result) {
  if (result != hipSuccess) {
    fprintf(stderr, "HIP Runtime Error: %s\n", hipGetErrorString(result));
    abort();
  }
  return result;
}

static void check_pf(int rc, const char* what) {
  if (rc != PF_OK) {
    fprintf(stderr, "libpfpa error in %s: %s\n", what, pf_last_error());
    abort();
  }
}

// ---------------------------------------------------------------- mesh.h

void make_arrays(struct HostData* host, const int nx, const int ny, const int nm) {
  // 2-D row pointers over 1-D storage, as src/mesh.cpp:12-97 (sized by ny rows)
  fp_t*** fields[] = {&host->conc_Cr_old, &host->conc_Cr_new, &host->conc_Cr_lap,
                      &host->conc_Cr_gam, &host->conc_Nb_old, &host->conc_Nb_new,
                      &host->conc_Nb_lap, &host->conc_Nb_gam, &host->conc_Ni,
                      &host->phi_del_old, &host->phi_del_new, &host->phi_del_lap,
                      &host->phi_lav_old, &host->phi_lav_new, &host->phi_lav_lap,
                      &host->phi, &host->chem_nrg, &host->grad_nrg};
  for (fp_t*** f : fields) {
    *f = (fp_t**)calloc(ny, sizeof(fp_t*));
    (*f)[0] = (fp_t*)calloc((size_t)nx * ny, sizeof(fp_t));
    for (int i = 1; i < ny; i++) (*f)[i] = &((*f)[0])[(size_t)nx * i];
  }
  host->mask_lap = (fp_t**)calloc(nm, sizeof(fp_t*));
  host->mask_lap[0] = (fp_t*)calloc(nm * nm, sizeof(fp_t));
  for (int i = 1; i < nm; i++) host->mask_lap[i] = &(host->mask_lap[0])[nm * i];
}

void free_arrays(struct HostData* host) {
  fp_t** fields[] = {host->conc_Cr_old, host->conc_Cr_new, host->conc_Cr_lap,
                     host->conc_Cr_gam, host->conc_Nb_old, host->conc_Nb_new,
                     host->conc_Nb_lap, host->conc_Nb_gam, host->conc_Ni,
                     host->phi_del_old, host->phi_del_new, host->phi_del_lap,
                     host->phi_lav_old, host->phi_lav_new, host->phi_lav_lap,
                     host->phi, host->chem_nrg, host->grad_nrg, host->mask_lap};
  for (fp_t** f : fields) {
    if (!f) continue;
    free(f[0]);
    free(f);
  }
}

void swap_pointers(fp_t*** conc_old, fp_t*** conc_new) {
  fp_t** t = *conc_old;
  *conc_old = *conc_new;
  *conc_new = t;
}

void swap_pointers_1D(fp_t** conc_old, fp_t** conc_new) {
  fp_t* t = *conc_old;
  *conc_old = *conc_new;
  *conc_new = t;
}

void five_point_Laplacian_stencil(const fp_t dx, const fp_t dy, fp_t** mask_lap, const int nm) {
  assert(nm == 3);
  mask_lap[0][1] = 1. / (dy * dy);
  mask_lap[1][0] = 1. / (dx * dx);
  mask_lap[1][1] = -2. * (dx * dx + dy * dy) / (dx * dx * dy * dy);
  mask_lap[1][2] = 1. / (dx * dx);
  mask_lap[2][1] = 1. / (dy * dy);
}

void nine_point_Laplacian_stencil(const fp_t dx, const fp_t dy, fp_t** mask_lap, const int nm) {
  assert(nm == 3);
  mask_lap[0][0] = 1. / (6. * dx * dy);
  mask_lap[0][1] = 4. / (6. * dy * dy);
  mask_lap[0][2] = 1. / (6. * dx * dy);
  mask_lap[1][0] = 4. / (6. * dx * dx);
  mask_lap[1][1] = -10. * (dx * dx + dy * dy) / (6. * dx * dx * dy * dy);
  mask_lap[1][2] = 4. / (6. * dx * dx);
  mask_lap[2][0] = 1. / (6. * dx * dy);
  mask_lap[2][1] = 4. / (6. * dy * dy);
  mask_lap[2][2] = 1. / (6. * dx * dy);
}

void set_mask(const fp_t dx, const fp_t dy, const int code, fp_t** mask_lap, const int nm) {
  if (code == 93)
    nine_point_Laplacian_stencil(dx, dy, mask_lap, nm);
  else
    five_point_Laplacian_stencil(dx, dy, mask_lap, nm);
  assert(nm <= MAX_MASK_W);
  assert(nm <= MAX_MASK_H);
}

fp_t grad_sq(fp_t** conc, const int x, const int y, const fp_t dx, const fp_t dy, const int nx,
             const int ny) {
  assert(x > 0 && x < nx - 1);
  assert(y > 0 && y < ny - 1);
  const fp_t gx = conc[y][x + 1] - conc[y][x - 1];
  const fp_t gy = conc[y + 1][x] - conc[y - 1][x];
  return gx * gx / (4. * dx * dx) + gy * gy / (4. * dy * dy);
}

// ---------------------------------------------------------------- data.cuh

void init_cuda(struct HostData* host, const int nx, const int ny, const int nm,
               const fp_t* kappa, const fp_t* omega, const fp_t* Lmob, const fp_t* D_Cr,
               const fp_t* D_Nb, struct CudaData* dev) {
  assert(nm == 3);
  const size_t bytes = (size_t)nx * ny * sizeof(fp_t);
  dev->prng.seed = 42;
  dev->prng.calls = 0;
  // a blocking stream, as the reference's cudaStreamCreate (src/data.cu:19-22):
  // a caller's work on the default stream (a hipMemcpy of the device arrays)
  // is ordered after the steps queued here
  (void)checkCuda(hipStreamCreate(&dev->str_A));
  dev->str_B = dev->str_C = dev->str_D = dev->str_A;
  (void)checkCuda(hipEventCreateWithFlags(&dev->ev_A, hipEventDisableTiming));
  dev->ev_B = dev->ev_C = dev->ev_D = dev->ev_A;
  fp_t** fields[] = {&dev->conc_Cr_old, &dev->conc_Cr_new, &dev->conc_Nb_old, &dev->conc_Nb_new,
                     &dev->phi_del_old, &dev->phi_del_new, &dev->phi_lav_old, &dev->phi_lav_new,
                     &dev->conc_Ni, &dev->phi};
  for (fp_t** f : fields) (void)checkCuda(hipMalloc((void**)f, bytes));
  check_pf(pf_ctx_create(&dev->ctx), "init_cuda");
  dev->bc = PF_BC_NOFLUX;
  dev->group = 4;
  dev->pend = 0;
  for (int i = 0; i < NC; i++) {
    dev->DCr[i] = D_Cr[i];
    dev->DNb[i] = D_Nb[i];
  }
  for (int i = 0; i < NP; i++) {
    dev->kappa[i] = kappa[i];
    dev->omega[i] = omega[i];
    dev->Lmob[i] = Lmob[i];
  }
  for (int i = 0; i < 9; i++) dev->mask[i] = host->mask_lap[0][i];
  (void)checkCuda(hipMemcpy(dev->conc_Cr_old, host->conc_Cr_old[0], bytes, hipMemcpyHostToDevice));
  (void)checkCuda(hipMemcpy(dev->conc_Nb_old, host->conc_Nb_old[0], bytes, hipMemcpyHostToDevice));
  (void)checkCuda(hipMemcpy(dev->phi_del_old, host->phi_del_old[0], bytes, hipMemcpyHostToDevice));
  (void)checkCuda(hipMemcpy(dev->phi_lav_old, host->phi_lav_old[0], bytes, hipMemcpyHostToDevice));
  // the new fields start as a copy, so that their ghost rings are defined
  (void)checkCuda(hipMemcpy(dev->conc_Cr_new, dev->conc_Cr_old, bytes, hipMemcpyDeviceToDevice));
  (void)checkCuda(hipMemcpy(dev->conc_Nb_new, dev->conc_Nb_old, bytes, hipMemcpyDeviceToDevice));
  (void)checkCuda(hipMemcpy(dev->phi_del_new, dev->phi_del_old, bytes, hipMemcpyDeviceToDevice));
  (void)checkCuda(hipMemcpy(dev->phi_lav_new, dev->phi_lav_old, bytes, hipMemcpyDeviceToDevice));
}

void free_cuda(struct CudaData* dev) {
  device_flush(dev);
  (void)checkCuda(hipStreamSynchronize(dev->str_A));
  fp_t* fields[] = {dev->conc_Cr_old, dev->conc_Cr_new, dev->conc_Nb_old, dev->conc_Nb_new,
                    dev->phi_del_old, dev->phi_del_new, dev->phi_lav_old, dev->phi_lav_new,
                    dev->conc_Ni, dev->phi};
  for (fp_t* f : fields) (void)checkCuda(hipFree(f));
  pf_ctx_destroy(dev->ctx);
  dev->ctx = nullptr;
  (void)checkCuda(hipEventDestroy(dev->ev_A));
  (void)checkCuda(hipStreamDestroy(dev->str_A));
}

void read_out_result(struct CudaData* dev, struct HostData* host, const int nx, const int ny) {
  device_flush(dev);
  const size_t bytes = (size_t)nx * ny * sizeof(fp_t);
  (void)checkCuda(hipStreamSynchronize(dev->str_A));
  (void)checkCuda(hipMemcpy(host->conc_Cr_new[0], dev->conc_Cr_old, bytes, hipMemcpyDeviceToHost));
  (void)checkCuda(hipMemcpy(host->conc_Nb_new[0], dev->conc_Nb_old, bytes, hipMemcpyDeviceToHost));
  (void)checkCuda(hipMemcpy(host->phi_del_new[0], dev->phi_del_old, bytes, hipMemcpyDeviceToHost));
  (void)checkCuda(hipMemcpy(host->phi_lav_new[0], dev->phi_lav_old, bytes, hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------- discretization.cuh

float nTiles(int domain_size, int tile_loc, int mask_size) {
  return ceil(float(domain_size) / float(tile_loc - mask_size + 1));
}

// The fused step does these inside device_evolution.
void device_fictitious(struct CudaData*, const int, const int, const int, const int, const int) {}
void device_boundaries(struct CudaData*, const int, const int, const int, const int, const int) {}
void fictitious_boundaries(struct CudaData*, const int, const int, const int, const int,
                           const int) {}
void device_laplacian_boundaries(struct CudaData*, const int, const int, const int, const int,
                                 const int) {}
void device_laplacian(struct CudaData*, const int, const int, const int, const int, const int,
                      const fp_t, const fp_t) {}

static pf_params params_of(const struct CudaData* dev, fp_t alpha, fp_t dt) {
  pf_params P;
  P.dt = dt;
  for (int i = 0; i < 2; i++) {
    P.DCr[i] = dev->DCr[i];
    P.DNb[i] = dev->DNb[i];
    P.kappa[i] = dev->kappa[i];
    P.omega[i] = dev->omega[i];
    P.Lmob[i] = dev->Lmob[i];
  }
  P.alpha = alpha;
  for (int i = 0; i < 9; i++) P.mask[i] = dev->mask[i];
  return P;
}

static void one_step(struct CudaData* dev, int nx, int ny, fp_t alpha, fp_t dt,
                     fp_t* const old_f[4], fp_t* const new_f[4]) {
  const pf_params P = params_of(dev, alpha, dt);
  check_pf(pf_step_f64(dev->ctx, nx, ny, dev->bc, &P, old_f[0], old_f[1], old_f[2], old_f[3],
                       new_f[0], new_f[1], new_f[2], new_f[3], nullptr, dev->str_A),
           "device_evolution");
}

// Runs the pending steps.  The driver swaps its pointers after every step, so
// after p steps the newest fields are in the new set of the first step if p
// is odd and in its old set if p is even.  A launch of k steps reads one set
// and writes the other, so the launches alternate between the sets and their
// number must have the parity of p: pairs of four-step launches (where the
// grid and the steps-per-launch setting allow), then three-step launches,
// then one-step launches.
static void run_pending(struct CudaData* dev) {
  const int nx = dev->pend_nx, ny = dev->pend_ny;
  const pf_params P = params_of(dev, dev->pend_alpha, dev->pend_dt);
  int left = dev->pend, launches = 0;
  auto launch = [&](int k) {
    fp_t* const* src = launches % 2 ? dev->pend_new : dev->pend_old;
    fp_t* const* dst = launches % 2 ? dev->pend_old : dev->pend_new;
    if (k == 1)
      one_step(dev, nx, ny, dev->pend_alpha, dev->pend_dt, src, dst);
    else
      check_pf((k == 4 ? pf_step4_f64 : pf_step3_f64)(dev->ctx, nx, ny, dev->bc, &P, src[0], src[1],
                                                      src[2], src[3], dst[0], dst[1], dst[2],
                                                      dst[3], dev->str_A),
               "device_evolution");
    launches++;
    left -= k;
  };
  const bool four = dev->group == 4 && nx >= 6 && ny >= 6;
  while (four && left >= 8) {
    launch(4);
    launch(4);
  }
  while (left >= 3) launch(3);
  while (left > 0) launch(1);
  dev->pend = 0;
}

void device_flush(struct CudaData* dev) {
  if (dev->pend > 0) run_pending(dev);
}

void device_set_steps_per_launch(struct CudaData* dev, int steps) {
  device_flush(dev);
  dev->group = steps == 4 ? 4 : steps == 3 ? 3 : 1;
}

// One step, old -> new (src/discretization.cu:1000-1020 and the kernels the
// driver calls before it).  Grouped: the call is recorded, and eight recorded
// steps run as two four-step launches (three as one three-step launch with
// three steps per launch).
void device_evolution(struct CudaData* dev, const int nx, const int ny, const int nm,
                      const int bx, const int by, const fp_t alpha, const fp_t dt) {
  (void)nm;
  (void)bx;
  (void)by;
  fp_t* const cur_old[4] = {dev->conc_Cr_old, dev->conc_Nb_old, dev->phi_del_old,
                            dev->phi_lav_old};
  fp_t* const cur_new[4] = {dev->conc_Cr_new, dev->conc_Nb_new, dev->phi_del_new,
                            dev->phi_lav_new};
  const bool five_point = pf::symmetric_five_point(dev->mask);
  if (dev->group == 1 || !five_point || nx < 5 || ny < 5) {
    device_flush(dev);
    one_step(dev, nx, ny, alpha, dt, cur_old, cur_new);
    return;
  }
  if (dev->pend > 0) {
    // the next step of the group reads what the previous one wrote
    fp_t* const* want_old = dev->pend % 2 ? dev->pend_new : dev->pend_old;
    fp_t* const* want_new = dev->pend % 2 ? dev->pend_old : dev->pend_new;
    bool chain = alpha == dev->pend_alpha && dt == dev->pend_dt && nx == dev->pend_nx &&
                 ny == dev->pend_ny;
    for (int k = 0; k < 4; k++) chain = chain && cur_old[k] == want_old[k] && cur_new[k] == want_new[k];
    if (!chain) device_flush(dev);
  }
  if (dev->pend == 0) {
    for (int k = 0; k < 4; k++) {
      dev->pend_old[k] = cur_old[k];
      dev->pend_new[k] = cur_new[k];
    }
    dev->pend_alpha = alpha;
    dev->pend_dt = dt;
    dev->pend_nx = nx;
    dev->pend_ny = ny;
  }
  if (++dev->pend == (dev->group == 4 && nx >= 6 && ny >= 6 ? 8 : 3)) run_pending(dev);
}

void device_init_prng(struct CudaData* dev, const int, const int, const int, const int,
                      const int) {
  dev->prng.calls = 0;
}

void device_nucleation(struct CudaData* dev, const int nx, const int ny, const int nm,
                       const int bx, const int by, const fp_t sigma_del, const fp_t sigma_lav,
                       const fp_t unit_a, const fp_t ifce_width, const fp_t dx, const fp_t dy,
                       const fp_t dz, const fp_t dt) {
  (void)nm;
  (void)bx;
  (void)by;
  device_flush(dev);
  pf_nuc_params N;
  N.sigma_del = sigma_del;
  N.sigma_lav = sigma_lav;
  N.unit_a = unit_a;
  N.ifce_width = ifce_width;
  N.dx = dx;
  N.dy = dy;
  N.dz = dz;
  N.dt = dt;
  N.seed = dev->prng.seed;
  N.call = dev->prng.calls;
  N.instance = 0;
  N.reserved = 0;
  check_pf(pf_nucleation_f64(dev->ctx, nx, ny, dev->conc_Cr_new, dev->conc_Nb_new,
                             dev->phi_del_new, dev->phi_lav_new, &N, dev->str_A),
           "device_nucleation");
  dev->prng.calls++;
}

void device_dataviz(struct CudaData* dev, struct HostData* host, const int nx, const int ny,
                    const int nm, const int bx, const int by) {
  (void)nm;
  (void)bx;
  (void)by;
  device_flush(dev);
  check_pf(pf_dataviz_f64(nx, ny, dev->conc_Cr_old, dev->conc_Nb_old, dev->phi_del_old,
                          dev->phi_lav_old, dev->conc_Ni, dev->phi, dev->str_A),
           "device_dataviz");
  const size_t bytes = (size_t)nx * ny * sizeof(fp_t);
  (void)checkCuda(hipStreamSynchronize(dev->str_A));
  (void)checkCuda(hipMemcpy(host->conc_Cr_new[0], dev->conc_Cr_old, bytes, hipMemcpyDeviceToHost));
  (void)checkCuda(hipMemcpy(host->conc_Nb_new[0], dev->conc_Nb_old, bytes, hipMemcpyDeviceToHost));
  (void)checkCuda(hipMemcpy(host->conc_Ni[0], dev->conc_Ni, bytes, hipMemcpyDeviceToHost));
  (void)checkCuda(hipMemcpy(host->phi_del_new[0], dev->phi_del_old, bytes, hipMemcpyDeviceToHost));
  (void)checkCuda(hipMemcpy(host->phi_lav_new[0], dev->phi_lav_old, bytes, hipMemcpyDeviceToHost));
  (void)checkCuda(hipMemcpy(host->phi[0], dev->phi, bytes, hipMemcpyDeviceToHost));
}

void device_set_boundary(struct CudaData* dev, int bc) {
  device_flush(dev);
  dev->bc = bc ? PF_BC_PERIODIC : PF_BC_NOFLUX;
}

void device_seed_prng(struct CudaData* dev, uint64_t seed) {
  device_flush(dev);
  dev->prng.seed = seed;
  dev->prng.calls = 0;
}
