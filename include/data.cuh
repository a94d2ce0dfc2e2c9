/**
 \file  data.cuh
 \brief Device data lifecycle (drop-in for src/data.cuh:16-35 of the reference).
*/
#ifndef _CUDA_DATA_H_
#define _CUDA_DATA_H_

#include "type.h"

/** Print and abort on a HIP error (src/discretization.cu:23-30 checks cudaError_t). */
hipError_t checkCuda(hipError_t result);

/** Allocate device fields, store the kinetic parameters and upload the host
 *  *_old fields (src/data.cu:10-92). */
void init_cuda(struct HostData* host,
               const int nx, const int ny, const int nm,
               const fp_t* kappa, const fp_t* omega, const fp_t* Lmob,
               const fp_t* D_Cr, const fp_t* D_Nb,
               struct CudaData* dev);

/** Release everything init_cuda allocated (src/data.cu:94-154). */
void free_cuda(struct CudaData* dev);

/** Copy the latest fields (device *_old) into host *_new; synchronous
 *  (src/data.cu:156-172). */
void read_out_result(struct CudaData* dev, struct HostData* host, const int nx, const int ny);

#endif /* _CUDA_DATA_H_ */
