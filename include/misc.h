/*
 * misc.h -- host helpers of the drop-in interface (PSBA/misc.h:8-23).
 */
#pragma once

#include "psba.h"

/* PSBA/misc.cpp:21-49: normalise a quaternion (scalar >= 0) and keep its
 * vector part; leading intrinsics/distortion are copied */
void quat2vec(dtype* inp, int nin, dtype* outp, int nout);
/* PSBA/misc.cpp:60-85: inverse of quat2vec */
void vec2quat(dtype* inp, int nin, dtype* outp, int nout);
/* PSBA/misc.cpp:151-157 */
dtype compute_L2_sq(int n, dtype* x);
/* PSBA/misc.cpp:162-169 */
dtype dotProduct(dtype* a, dtype* b, int size);
/* PSBA/misc.cpp:178-218 (bit-exact); comm3DIdx may be NULL to skip the
 * nCams*nCams*n3Dpts list */
void generate_idxs(int nCams, int n3Dpts, int n2Dprojs, dtype* impts_data, char* vmask, int* comm3DIdx,
                   int* comm3DIdxCnt, int* iidx, int* jidx, int* blk_idx);
