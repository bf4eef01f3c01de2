// A host program written against the reference's own headers and call
// sequence (PSBA/main.cpp:88-208): read the sba files, split intrinsics from
// the pose, set up the device buffers, build the indices, and alternate the
// LM and dogleg drivers.  Built against libpsba.so it exercises the drop-in
// layer end to end.
//
//   main_dropin CAMS PTS OUT [origin_cnp=11]
//
// Prints the reference's report (iter_flag, "initial error", "final error",
// "total iteration", PSBA/main.cpp:210-218) plus exact itno / init_err /
// final_err lines; uses the library's emalloc and the levmar-set initErr as
// the reference program does.  Writes the final camera
// (cnp = origin_cnp - 5) and point parameters to OUT as raw doubles.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>

#include "cl_psba.h"
#include "levmar.h"
#include "misc.h"
#include "psba.h"
#include "readparams.h"
#include "trust_region.h"

dtype initErr;  // set by levmar (PSBA/main.cpp:24, PSBA/levmar.cpp:95)
int itno = 0;   // the shared iteration counter the drivers advance (PSBA/main.cpp:27)

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s CAMS PTS OUT [origin_cnp]\n", argv[0]);
    return 2;
  }
  int origin_cnp = argc > 4 ? atoi(argv[4]) : 11;
  int pnp = 3, mnp = 2;
  int n3Dpts, nCams, n2Dprojs;
  dtype *motion_data, *initrot, *impts_data, *impts_cov;
  char* vmask;
  readInitialSBAEstimate(argv[1], argv[2], origin_cnp, pnp, mnp, quat2vec, origin_cnp + 1, &nCams, &n3Dpts,
                         &n2Dprojs, &motion_data, &initrot, &impts_data, &impts_cov, &vmask);
  // the rotation now lives in initrot: zero the vector part of every camera
  for (int j = 0; j < nCams; ++j) {
    int base = (j + 1) * origin_cnp;
    motion_data[base - 4] = motion_data[base - 5] = motion_data[base - 6] = 0;
  }
  int cnp = origin_cnp - 5;
  dtype* Kparas = (dtype*)emalloc(sizeof(dtype) * nCams * 5);
  dtype* camsEx = (dtype*)emalloc(sizeof(dtype) * nCams * cnp);
  for (int j = 0; j < nCams; ++j) {
    for (int k = 0; k < 5; ++k) Kparas[j * 5 + k] = motion_data[j * origin_cnp + k];
    for (int k = 0; k < cnp; ++k) camsEx[j * cnp + k] = motion_data[j * origin_cnp + k + 5];
  }
  PSBA_struct ps;
  memset(&ps, 0, sizeof(ps));
  setup_cl(&ps, cnp, pnp, mnp, nCams, n3Dpts, n2Dprojs);
  fill_initBuffer2(&ps, cnp, pnp, mnp, nCams, n3Dpts, n2Dprojs, Kparas, impts_data, initrot, camsEx,
                   &motion_data[nCams * origin_cnp]);
  int* iidx = (int*)emalloc(n2Dprojs * sizeof(int));
  int* jidx = (int*)emalloc(n2Dprojs * sizeof(int));
  int* blk_idx = (int*)emalloc((size_t)n3Dpts * nCams * sizeof(int));
  int* comm3DIdx = (int*)emalloc((size_t)n3Dpts * nCams * nCams * sizeof(int));
  int* comm3DIdxCnt = (int*)emalloc((size_t)nCams * nCams * sizeof(int));
  generate_idxs(nCams, n3Dpts, n2Dprojs, impts_data, vmask, comm3DIdx, comm3DIdxCnt, iidx, jidx, blk_idx);
  fill_idxBuffer(&ps, nCams, n3Dpts, n2Dprojs, comm3DIdx, comm3DIdxCnt, iidx, jidx, blk_idx);

  dtype finalErr = 0;
  int iter_flag;
  clock_t start_time = clock();
  while (true) {
    iter_flag = levmar(&ps, cnp, pnp, mnp, n3Dpts, nCams, n2Dprojs, &finalErr);
    if (iter_flag != ITER_TURN_TO_TR) break;
    iter_flag = trust_region(&ps, cnp, pnp, mnp, n3Dpts, nCams, n2Dprojs, blk_idx, &finalErr);
    if (iter_flag != ITER_TURN_TO_LM) break;
  }
  clock_t end_time = clock();
  dtype* cams = (dtype*)emalloc(sizeof(dtype) * nCams * cnp);
  dtype* pts = (dtype*)emalloc(sizeof(dtype) * n3Dpts * pnp);
  read_params(&ps, cams, pts);
  release_buffer(&ps);
  // the report of PSBA/main.cpp:210-218
  printf("iter_flag=%d\n", iter_flag);
  printf("time eclipse %lf s\n", double(end_time - start_time) / CLOCKS_PER_SEC);
  printf("initial error: %.15E \n", sqrt(initErr) / n2Dprojs);
  printf("final error: %.15E \n", sqrt(finalErr) / n2Dprojs);
  printf("total iteration: %d\n", itno);
  // exact values for the tests
  printf("itno=%d\n", itno);
  printf("init_err=%.17g\n", initErr);
  printf("final_err=%.17g\n", finalErr);
  FILE* f = fopen(argv[3], "wb");
  if (!f) return 3;
  fwrite(cams, sizeof(dtype), (size_t)nCams * cnp, f);
  fwrite(pts, sizeof(dtype), (size_t)n3Dpts * pnp, f);
  fclose(f);
  return 0;
}
