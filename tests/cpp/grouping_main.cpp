// grouping_main.cpp -- the drop-in's grouping of device_evolution calls
// (eight to two four-step launches, or three to a three-step launch) under
// call sequences the reference's driver does not
// make but a caller of the same API may: a change of dt inside a group, a
// step whose pointers were not swapped after the previous one (the same
// old -> new twice), a nucleation pass, and a direct read of the device arrays
// after device_flush, and pending tails of every length the sequence reaches.
// The scripted sequence runs three times: four and three steps per launch,
// and one launch per call (device_set_steps_per_launch(&dev, 1)); the final
// fields and the directly read intermediate must be the same bits.
//
//   grouping_main <in.bin> <dt> <dx>
//
// in.bin: int32 nx, int32 ny, then conc_Cr, conc_Nb, phi_del, phi_lav as
// nx*ny doubles each (ghost ring included).  Prints "grouping ok" and exits 0
// when both runs agree.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "data.cuh"
#include "discretization.cuh"
#include "mesh.h"
#include "pfpa.h"

struct Result {
  std::vector<double> fields;  // 4 x nx*ny after the sequence
  std::vector<double> mid;     // conc_Cr read directly from the device mid-sequence
};

static void swap_all(CudaData* dev) {
  swap_pointers_1D(&(dev->conc_Cr_old), &(dev->conc_Cr_new));
  swap_pointers_1D(&(dev->conc_Nb_old), &(dev->conc_Nb_new));
  swap_pointers_1D(&(dev->phi_del_old), &(dev->phi_del_new));
  swap_pointers_1D(&(dev->phi_lav_old), &(dev->phi_lav_new));
}

static int run(const char* in, double dt, double meshres, int steps_per_launch, Result& out) {
  FILE* fi = fopen(in, "rb");
  if (!fi) return 3;
  int dims[2];
  if (fread(dims, sizeof(int), 2, fi) != 2) return 3;
  const int nx = dims[0], ny = dims[1];
  const int bx = 13, by = 13, nm = 3;
  struct HostData host;
  struct CudaData dev;
  make_arrays(&host, nx, ny, nm);
  five_point_Laplacian_stencil(meshres, meshres, host.mask_lap, nm);
  fp_t** fields[4] = {host.conc_Cr_old, host.conc_Nb_old, host.phi_del_old, host.phi_lav_old};
  for (int k = 0; k < 4; k++)
    if (fread(fields[k][0], sizeof(double), (size_t)nx * ny, fi) != (size_t)nx * ny) return 3;
  fclose(fi);

  fp_t D_Cr[NC], D_Nb[NC], Lmob[NP];
  D_Cr[0] = pf_host_D(0, 0, 0.30, 0.02);
  D_Cr[1] = pf_host_D(0, 1, 0.30, 0.02);
  D_Nb[0] = pf_host_D(1, 0, 0.30, 0.02);
  D_Nb[1] = pf_host_D(1, 1, 0.30, 0.02);
  const double ifce_width = 2.5e-9, RT = 9504.6840941999999, Vm = 1.0000000000000001e-5;
  Lmob[0] = sqrt(D_Cr[0] * D_Cr[0] + D_Nb[1] * D_Nb[1]) / (ifce_width * ifce_width * RT / Vm);
  Lmob[1] = Lmob[0];
  const double sigma[NP] = {0.13, 0.13};
  const double kappa[NP] = {3. * sigma[0] * ifce_width / 2.2, 3. * sigma[1] * ifce_width / 2.2};
  const double omega[NP] = {18. * sigma[0] * sigma[0] / kappa[0],
                            18. * sigma[1] * sigma[1] / kappa[1]};
  const double Omega = 1.07e11;

  init_cuda(&host, nx, ny, nm, kappa, omega, Lmob, D_Cr, D_Nb, &dev);
  device_init_prng(&dev, nx, ny, nm, bx, by);
  device_seed_prng(&dev, 7);
  device_set_steps_per_launch(&dev, steps_per_launch);

  auto step = [&](double h, bool swap) {
    device_evolution(&dev, nx, ny, nm, bx, by, Omega, h);
    if (swap) swap_all(&dev);
  };
  step(dt, true);            // 0
  step(dt, true);            // 1: two pending
  step(0.5 * dt, true);      // 2: dt changes inside the group
  step(0.5 * dt, true);      // 3
  step(0.5 * dt, true);      // 4: a full group at the new dt
  step(dt, false);           // 5: no swap afterwards ...
  step(dt, true);            // 6: ... so old -> new again (the chain breaks)
  step(dt, true);            // 7
  step(dt, true);            // 8
  // 9: nucleation on the new fields before the swap (huge dt: events happen)
  device_evolution(&dev, nx, ny, nm, bx, by, Omega, dt);
  device_nucleation(&dev, nx, ny, nm, bx, by, 0.0105, 0.0108, 0.352e-9, ifce_width, meshres,
                    meshres, meshres, 3e38);
  swap_all(&dev);
  step(dt, true);            // 10
  step(dt, true);            // 11
  // a direct read of the device arrays after device_flush
  device_flush(&dev);
  out.mid.resize((size_t)nx * ny);
  if (hipMemcpy(out.mid.data(), dev.conc_Cr_old, sizeof(double) * nx * ny,
                hipMemcpyDeviceToHost) != hipSuccess)
    return 5;
  // 12..34: with four steps per launch two groups of eight and a tail of
  // seven (three-, three- and one-step launches); with three, seven groups and
  // a tail of two
  for (int s = 12; s < 35; s++) step(dt, true);
  read_out_result(&dev, &host, nx, ny);

  out.fields.resize((size_t)4 * nx * ny);
  fp_t** outs[4] = {host.conc_Cr_new, host.conc_Nb_new, host.phi_del_new, host.phi_lav_new};
  for (int k = 0; k < 4; k++)
    memcpy(out.fields.data() + (size_t)k * nx * ny, outs[k][0], sizeof(double) * nx * ny);
  free_cuda(&dev);
  free_arrays(&host);
  return 0;
}

int main(int argc, char* argv[]) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s in.bin dt dx\n", argv[0]);
    return 2;
  }
  const double dt = atof(argv[2]), dx = atof(argv[3]);
  Result single;
  if (int r = run(argv[1], dt, dx, 1, single)) return r;
  for (int spl : {4, 3}) {
    Result grouped;
    if (int r = run(argv[1], dt, dx, spl, grouped)) return r;
    if (memcmp(grouped.mid.data(), single.mid.data(), sizeof(double) * grouped.mid.size()) != 0) {
      fprintf(stderr, "%d steps per launch: direct read after device_flush differs\n", spl);
      return 1;
    }
    if (memcmp(grouped.fields.data(), single.fields.data(),
               sizeof(double) * grouped.fields.size()) != 0) {
      fprintf(stderr, "%d steps per launch: final fields differ\n", spl);
      return 1;
    }
  }
  printf("grouping ok\n");
  return 0;
}
