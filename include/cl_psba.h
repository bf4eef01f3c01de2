/*
 * cl_psba.h -- the problem state handle of the drop-in interface.
 *
 * Same type and function names as PSBA/cl_psba.h:9-126 so existing callers
 * compile unchanged; the state is HIP-backed (one psba_ctx per PSBA_struct,
 * one device, one stream).  API errors print a message and terminate the
 * process, like the reference's checkErr (PSBA/cl_psba.cpp:275-283).
 */
#pragma once

#include "psba.h"
#include "psba_capi.h"

typedef struct {
  psba_ctx* ctx; /* device state; NULL until setup_cl */
  int cnp, pnp, mnp, nCams, n3Dpts, n2Dprojs;
  int device;    /* HIP device used by setup_cl (default 0) */
} PSBA_struct, *PSBA_structPtr;

/* PSBA/cl_psba.h:93 -- allocate the device state for one problem */
void setup_cl(PSBA_structPtr psbaPtr, int cnp, int pnp, int mnp, int nCams, int n3Dpts, int n2Dprojs);
/* PSBA/cl_psba.h:94 */
void release_buffer(PSBA_structPtr psbaPtr);
/* PSBA/cl_psba.h:114-121 -- upload intrinsics, observations (point-major,
 * camera-ascending), initial rotations, camera parameters and points */
void fill_initBuffer2(PSBA_structPtr psbaPtr, int cnp, int pnp, int mnp, int nCams, int n3Dpts, int n2Dprojs,
                      dtype* Kparas, dtype* impts_data, dtype* initcams_data, dtype* camsExParas, dtype* pts3Ds);
/* PSBA/cl_psba.h:123-126 -- the device index is built from iidx/jidx;
 * comm3DIdx may be NULL (see generate_idxs) */
void fill_idxBuffer(PSBA_structPtr psbaPtr, int nCams, int n3Dpts, int n2Dprojs, int* comm3DIdx, int* comm3DIdxCnt,
                    int* iidx, int* jidx, int* blk_idx);
/* addition: read the optimised parameters back (the reference leaves them in
 * cams_buffer / pts3D_buffer) */
void read_params(PSBA_structPtr psbaPtr, dtype* cams, dtype* pts);
