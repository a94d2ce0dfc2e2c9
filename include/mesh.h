/**
 \file  mesh.h
 \brief Host mesh helpers (drop-in for src/mesh.h:20-93 of the reference).
*/
#ifndef _MESH_H_
#define _MESH_H_

#include "type.h"

/** Allocate 2-D host arrays over 1-D storage (src/mesh.cpp:12-97). */
void make_arrays(struct HostData* host, const int nx, const int ny, const int nm);

/** Free the arrays of make_arrays (src/mesh.cpp:99-154). */
void free_arrays(struct HostData* host);

/** Swap 2-D array pointers (src/mesh.cpp:156-161). */
void swap_pointers(fp_t*** conc_old, fp_t*** conc_new);

/** Swap 1-D (device) array pointers (src/mesh.cpp:163-168). */
void swap_pointers_1D(fp_t** conc_old, fp_t** conc_new);

#define MAX_MASK_W 5
#define MAX_MASK_H 5

/** Select a Laplacian mask by code: 53 five-point, 93 nine-point (src/mesh.cpp:170-185). */
void set_mask(const fp_t dx, const fp_t dy, const int code, fp_t** mask_lap, const int nm);

/** 3x3 five-point Laplacian mask (src/mesh.cpp:187-196). */
void five_point_Laplacian_stencil(const fp_t dx, const fp_t dy, fp_t** mask_lap, const int nm);

/** 3x3 nine-point Laplacian mask (src/mesh.cpp:198-213). */
void nine_point_Laplacian_stencil(const fp_t dx, const fp_t dy, fp_t** mask_lap, const int nm);

/** Squared gradient at an interior point (src/mesh.cpp:215-244). */
fp_t grad_sq(fp_t** conc, const int x, const int y, const fp_t dx, const fp_t dy,
             const int nx, const int ny);

#endif /* _MESH_H_ */
