/**
 \file  discretization.cuh
 \brief Host entry points of the step (drop-in for src/discretization.cuh:34-118).

 device_evolution launches one fused kernel for the whole step (boundaries,
 fictitious compositions, Laplacians, Cahn-Hilliard, Allen-Cahn); the four
 calls the driver makes before it are kept for the call sequence of
 src/main.cpp:289-303 and do no device work.  All work goes to dev->str_A in
 order, so the calls are ordered without host synchronisation; results are
 complete when read_out_result / device_dataviz return.
*/
#ifndef _CUDA_KERNELS_H_
#define _CUDA_KERNELS_H_

#include "mesh.h"

/** Tiles along an axis for a block of tile_loc with a mask_size halo (src/discretization.cu:45-48). */
float nTiles(int domain_size, int tile_loc, int mask_size);

void device_fictitious(struct CudaData* dev, const int nx, const int ny, const int nm,
                       const int bx, const int by);

void device_boundaries(struct CudaData* dev, const int nx, const int ny, const int nm,
                       const int bx, const int by);

void fictitious_boundaries(struct CudaData* dev, const int nx, const int ny, const int nm,
                           const int bx, const int by);

void device_laplacian_boundaries(struct CudaData* dev, const int nx, const int ny, const int nm,
                                 const int bx, const int by);

void device_laplacian(struct CudaData* dev, const int nx, const int ny, const int nm,
                      const int bx, const int by, const fp_t dx, const fp_t dy);

/** One explicit Euler step *_old -> *_new (src/discretization.cu:982-1020). */
void device_evolution(struct CudaData* dev, const int nx, const int ny, const int nm,
                      const int bx, const int by, const fp_t alpha, const fp_t dt);

/** Key the nucleation random stream (src/discretization.cu:1033-1044). */
void device_init_prng(struct CudaData* dev, const int nx, const int ny, const int nm,
                      const int bx, const int by);

/** Nucleation pass on *_new (src/discretization.cu:1179-1209). */
void device_nucleation(struct CudaData* dev, const int nx, const int ny, const int nm,
                       const int bx, const int by,
                       const fp_t sigma_del, const fp_t sigma_lav,
                       const fp_t unit_a, const fp_t ifce_width,
                       const fp_t dx, const fp_t dy, const fp_t dz, const fp_t dt);

/** x_Ni and phi of the latest fields, copied to the host with the fields
 *  (src/discretization.cu:1231-1263). */
void device_dataviz(struct CudaData* dev, struct HostData* host, const int nx, const int ny,
                    const int nm, const int bx, const int by);

/* ---- additions with no counterpart in the reference ---- */

/** Boundary mode of subsequent steps: 0 no-flux (default, the reference's), 1 periodic. */
void device_set_boundary(struct CudaData* dev, int bc);

/** Seed of the nucleation stream (default 42); resets the stream position. */
void device_seed_prng(struct CudaData* dev, uint64_t seed);

/** Steps per launch of device_evolution: 4 (default), 3 or 1.  With 4 or 3,
 * calls to device_evolution are grouped: consecutive steps (the driver
 * swapping the old/new pointers in between, src/main.cpp:318-321) run as
 * multi-step launches, each reading one pointer set and writing the other.
 * With 4, eight steps run as two four-step launches (an even number of
 * launches, as eight swaps leave the newest fields in the first step's old
 * set); with 3, three steps as one three-step launch.  Every function of this
 * API that reads the fields (device_nucleation, device_dataviz,
 * read_out_result, free_cuda) first runs the pending steps, as launches whose
 * number has the parity of the steps (pairs of four-step launches, three-step
 * launches, one-step launches), so what the driver observes is unchanged.
 * Grids below 4x4 interior cells group three steps per launch.  Code that
 * reads the device arrays by other means calls device_flush first. */
void device_set_steps_per_launch(struct CudaData* dev, int steps);

/** Queue the steps device_evolution has deferred (see above).  The steps run
 * on the API's stream, which is a blocking stream as in the reference: work
 * the caller then issues on the default stream (a hipMemcpy of the device
 * arrays) runs after them. */
void device_flush(struct CudaData* dev);

#endif /* _CUDA_KERNELS_H_ */
