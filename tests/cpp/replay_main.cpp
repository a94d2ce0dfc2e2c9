// replay_main.cpp -- drives the drop-in C++ API exactly as the reference's
// driver does (src/main.cpp:225-327): make_arrays, five_point_Laplacian_stencil,
// init_cuda, device_init_prng, then per step device_boundaries,
// device_fictitious, fictitious_boundaries, device_laplacian,
// device_laplacian_boundaries, device_evolution, [device_nucleation],
// swap_pointers_1D x4, and on I/O steps device_dataviz + read_out_result.
//
//   replay_main <in.bin> <out.bin> <steps> <dt> <dx> <nuc_interval> <seed> <io_interval>
//               [steps_per_launch [bc [mask_code]]]
//
// in.bin / out.bin: int32 nx, int32 ny, then conc_Cr, conc_Nb, phi_del,
// phi_lav as nx*ny doubles each (ghost ring included).  nuc_interval 0 = off.
// steps_per_launch (1 or 3, 0 = the library's default) is the drop-in's
// grouping of device_evolution calls; bc 0 = no-flux (the reference's), 1 =
// periodic (device_set_boundary); mask_code 53 = five-point, 93 = nine-point
// (set_mask, src/mesh.cpp:170-185).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <chrono>

#include "data.cuh"
#include "discretization.cuh"
#include "mesh.h"
#include "pfpa.h"

int main(int argc, char* argv[]) {
  if (argc < 9 || argc > 12) {
    fprintf(stderr,
            "usage: %s in.bin out.bin steps dt dx nuc_interval seed io_interval "
            "[steps_per_launch [bc [mask_code]]]\n",
            argv[0]);
    return 2;
  }
  const int steps_per_launch = argc >= 10 ? atoi(argv[9]) : 0;
  const int bc = argc >= 11 ? atoi(argv[10]) : 0;
  const int mask_code = argc >= 12 ? atoi(argv[11]) : 53;
  const long steps = atol(argv[3]);
  const double dt = atof(argv[4]);
  const double meshres = atof(argv[5]);
  const long nuc_interval = atol(argv[6]);
  const uint64_t seed = strtoull(argv[7], nullptr, 10);
  const long io_interval = atol(argv[8]);

  FILE* fi = fopen(argv[1], "rb");
  if (!fi) return 3;
  int dims[2];
  if (fread(dims, sizeof(int), 2, fi) != 2) return 3;
  const int nx = dims[0], ny = dims[1];

  // src/main.cpp:225-237
  struct HostData host;
  struct CudaData dev;
  const int bx = 13, by = 13, nm = 3;
  make_arrays(&host, nx, ny, nm);
  if (argc >= 12)
    set_mask(meshres, meshres, mask_code, host.mask_lap, nm);
  else
    five_point_Laplacian_stencil(meshres, meshres, host.mask_lap, nm);
  fp_t** fields[4] = {host.conc_Cr_old, host.conc_Nb_old, host.phi_del_old, host.phi_lav_old};
  for (int k = 0; k < 4; k++)
    if (fread(fields[k][0], sizeof(double), (size_t)nx * ny, fi) != (size_t)nx * ny) return 3;
  fclose(fi);

  // src/main.cpp:254 with src/alloy625.cpp:80-109 (MobStab = 1, ifce_width = 2.5 nm)
  fp_t D_Cr[NC], D_Nb[NC], Lmob[NP];
  D_Cr[0] = pf_host_D(0, 0, 0.30, 0.02);
  D_Cr[1] = pf_host_D(0, 1, 0.30, 0.02);
  D_Nb[0] = pf_host_D(1, 0, 0.30, 0.02);
  D_Nb[1] = pf_host_D(1, 1, 0.30, 0.02);
  const double ifce_width = 2.5e-9, RT = 9504.6840941999999, Vm = 1.0000000000000001e-5;
  Lmob[0] = sqrt(D_Cr[0] * D_Cr[0] + D_Nb[1] * D_Nb[1]) / (ifce_width * ifce_width * RT / Vm);
  Lmob[1] = Lmob[0];
  // src/parameters.h:36-48
  const double sigma[NP] = {0.13, 0.13};
  const double kappa[NP] = {3. * sigma[0] * ifce_width / 2.2, 3. * sigma[1] * ifce_width / 2.2};
  const double omega[NP] = {18. * sigma[0] * sigma[0] / kappa[0], 18. * sigma[1] * sigma[1] / kappa[1]};
  const double Omega = 1.07e11, lattice_const = 0.352e-9;

  init_cuda(&host, nx, ny, nm, kappa, omega, Lmob, D_Cr, D_Nb, &dev);
  device_init_prng(&dev, nx, ny, nm, bx, by);
  device_seed_prng(&dev, seed);
  if (steps_per_launch) device_set_steps_per_launch(&dev, steps_per_launch);
  if (bc) device_set_boundary(&dev, bc);

  // the last two thirds of the steps are timed (host wall clock from a point
  // where the device has finished the first third until it has finished all),
  // so that code loading on the first launches is excluded; the final read-out
  // is timed on its own
  long io_steps = 0;
  const long t_from = steps / 3;
  auto t0 = std::chrono::steady_clock::now();
  for (long kernel_time = 0; kernel_time < steps; kernel_time++) {
    if (kernel_time == t_from) {
      device_flush(&dev);
      (void)hipDeviceSynchronize();
      t0 = std::chrono::steady_clock::now();
    }
    device_boundaries(&dev, nx, ny, nm, bx, by);
    device_fictitious(&dev, nx, ny, nm, bx, by);
    fictitious_boundaries(&dev, nx, ny, nm, bx, by);
    device_laplacian(&dev, nx, ny, nm, bx, by, meshres, meshres);
    device_laplacian_boundaries(&dev, nx, ny, nm, bx, by);
    device_evolution(&dev, nx, ny, nm, bx, by, Omega, dt);
    if (nuc_interval > 0 && kernel_time % nuc_interval == 0)
      device_nucleation(&dev, nx, ny, nm, bx, by, sigma[0], sigma[1], lattice_const, ifce_width,
                        meshres, meshres, meshres, nuc_interval * dt);
    swap_pointers_1D(&(dev.conc_Cr_old), &(dev.conc_Cr_new));
    swap_pointers_1D(&(dev.conc_Nb_old), &(dev.conc_Nb_new));
    swap_pointers_1D(&(dev.phi_del_old), &(dev.phi_del_new));
    swap_pointers_1D(&(dev.phi_lav_old), &(dev.phi_lav_new));
    if (io_interval > 0 && (kernel_time + 1) % io_interval == 0) {
      device_dataviz(&dev, &host, nx, ny, nm, bx, by);
      read_out_result(&dev, &host, nx, ny);
      io_steps++;
    }
  }
  device_flush(&dev);
  (void)hipDeviceSynchronize();
  const auto t1 = std::chrono::steady_clock::now();
  const double secs = std::chrono::duration<double>(t1 - t0).count();
  read_out_result(&dev, &host, nx, ny);
  const double read_secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();

  FILE* fo = fopen(argv[2], "wb");
  if (!fo) return 4;
  fwrite(dims, sizeof(int), 2, fo);
  fp_t** outs[4] = {host.conc_Cr_new, host.conc_Nb_new, host.phi_del_new, host.phi_lav_new};
  for (int k = 0; k < 4; k++) fwrite(outs[k][0], sizeof(double), (size_t)nx * ny, fo);
  // visualisation fields of the last I/O step
  fwrite(host.conc_Ni[0], sizeof(double), (size_t)nx * ny, fo);
  fwrite(host.phi[0], sizeof(double), (size_t)nx * ny, fo);
  fclose(fo);
  free_cuda(&dev);
  free_arrays(&host);
  printf("replayed %ld steps, %ld I/O steps; last %ld steps in %.6f s; read-out %.6f s\n",
         steps, io_steps, steps - t_from, secs, read_secs);
  return 0;
}
