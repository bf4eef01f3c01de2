/*
 * misc.h -- host helpers of the drop-in interface (PSBA/misc.h:1-23).
 */
#pragma once

#include <stddef.h>

#include "psba.h"

/* PSBA/misc.h:5: checked allocation that names the caller's file and line */
#define emalloc(sz) emalloc_(__FILE__, __LINE__, sz)

/* PSBA/misc.cpp:21-49: camera-line filter of the sba reader.  The quaternion
 * q (4 values after nin-7 leading intrinsics/distortion values) becomes the
 * vector part of the unit quaternion with non-negative scalar, q_v*sign(q0)/|q|;
 * the values after the quaternion (translation) are copied. */
void quat2vec(dtype* inp, int nin, dtype* outp, int nout);
/* PSBA/misc.cpp:60-85: inverse of quat2vec (scalar part sqrt(1 - |v|^2)) */
void vec2quat(dtype* inp, int nin, dtype* outp, int nout);
/* PSBA/misc.cpp:119-133: per camera, quaternion (dim 3: vector part, dim 4:
 * s,v) + stubForTrans translation values -> row-major 3x3 rotation followed
 * by the translation; input stride dim+stubForTrans, output stride
 * 9+stubForTrans.  nin is the number of cameras. */
void camsFormatTrans(dtype* in, int nin, int dim, int stubForTrans, dtype* out);
/* PSBA/misc.cpp:91-113: row-major rotation matrix of a unit quaternion given
 * as its vector part (nin = 3) or as (s, v) (nin = 4) */
void quat2matrix(dtype* in, int nin, dtype* out);
/* PSBA/misc.cpp:135-146: malloc that prints the caller's file/line and exits
 * on failure.  The file name is taken as const char* (callers pass
 * __FILE__ through the emalloc macro, as the reference's do). */
void* emalloc_(const char* file, int line, size_t sz);
/* PSBA/misc.cpp:151-157 */
dtype compute_L2_sq(int n, dtype* x);
/* PSBA/misc.cpp:162-169 */
dtype dotProduct(dtype* a, dtype* b, int size);
/* PSBA/misc.cpp:178-218 (bit-exact); comm3DIdx may be NULL to skip the
 * nCams*nCams*n3Dpts list */
void generate_idxs(int nCams, int n3Dpts, int n2Dprojs, dtype* impts_data, char* vmask, int* comm3DIdx,
                   int* comm3DIdxCnt, int* iidx, int* jidx, int* blk_idx);
