// C++ drop-in layer: the reference's entry points (same names, signatures and
// argument meaning) implemented on the C-ABI.  API errors print a message and
// terminate, as the reference's checkErr does (PSBA/cl_psba.cpp:275-283);
// numerical outcomes are the ITER_* return codes.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "cl_psba.h"
#include "levmar.h"
#include "misc.h"
#include "psba_capi.h"
#include "psba_io.h"
#include "trust_region.h"

// The shared iteration counter.  The application normally defines it (as
// PSBA/main.cpp:27 does); this weak definition is used otherwise.
__attribute__((weak)) int itno = 0;

static void die_on(int rc, const char* where) {
  if (rc >= 0) return;
  fprintf(stderr, "%s failed: %s (%s)\n", where, psba_strerror(rc), psba_last_error());
  exit(1);
}

void setup_cl(PSBA_structPtr p, int cnp, int pnp, int mnp, int nCams, int n3Dpts, int n2Dprojs) {
  psba_desc d{cnp, pnp, mnp, nCams, n3Dpts, n2Dprojs, PSBA_PREC_FP64, 0};
  if (!p) die_on(PSBA_ERR_ARG, "setup_cl");
  d.device = p->device;
  p->ctx = nullptr;
  die_on(psba_create(&d, &p->ctx), "setup_cl");
  p->cnp = cnp;
  p->pnp = pnp;
  p->mnp = mnp;
  p->nCams = nCams;
  p->n3Dpts = n3Dpts;
  p->n2Dprojs = n2Dprojs;
}

void release_buffer(PSBA_structPtr p) {
  if (!p) return;
  psba_destroy(p->ctx);
  p->ctx = nullptr;
}

void fill_initBuffer2(PSBA_structPtr p, int cnp, int pnp, int mnp, int nCams, int n3Dpts, int n2Dprojs,
                      dtype* Kparas, dtype* impts_data, dtype* initcams_data, dtype* camsExParas, dtype* pts3Ds) {
  if (!p || !p->ctx || cnp != p->cnp || pnp != p->pnp || mnp != p->mnp || nCams != p->nCams ||
      n3Dpts != p->n3Dpts || n2Dprojs != p->n2Dprojs)
    die_on(PSBA_ERR_ARG, "fill_initBuffer2");
  die_on(psba_upload(p->ctx, Kparas, impts_data, initcams_data, camsExParas, pts3Ds), "fill_initBuffer2");
}

void fill_idxBuffer(PSBA_structPtr p, int nCams, int n3Dpts, int n2Dprojs, int* comm3DIdx, int* comm3DIdxCnt,
                    int* iidx, int* jidx, int* blk_idx) {
  (void)comm3DIdx;
  (void)comm3DIdxCnt;
  (void)blk_idx;
  if (!p || !p->ctx || nCams != p->nCams || n3Dpts != p->n3Dpts || n2Dprojs != p->n2Dprojs)
    die_on(PSBA_ERR_ARG, "fill_idxBuffer");
  die_on(psba_set_index(p->ctx, iidx, jidx), "fill_idxBuffer");
}

void read_params(PSBA_structPtr p, dtype* cams, dtype* pts) {
  if (!p || !p->ctx) die_on(PSBA_ERR_ARG, "read_params");
  die_on(psba_read_params(p->ctx, cams, pts), "read_params");
}

int levmar(PSBA_structPtr p, int cnp, int pnp, int mnp, int n3Dpts, int nCams, int n2Dprojs, dtype* finalErr) {
  if (!p || !p->ctx || cnp != p->cnp || pnp != p->pnp || mnp != p->mnp || nCams != p->nCams ||
      n3Dpts != p->n3Dpts || n2Dprojs != p->n2Dprojs || !finalErr)
    die_on(PSBA_ERR_ARG, "levmar");
  int rc = psba_levmar(p->ctx, nullptr, &itno, finalErr, nullptr, 0, nullptr);
  die_on(rc, "levmar");
  return rc;
}

int trust_region(PSBA_structPtr p, int cnp, int pnp, int mnp, int n3Dpts, int nCams, int n2Dprojs, int* blk_idx,
                 dtype* finalErr) {
  (void)blk_idx;
  if (!p || !p->ctx || cnp != p->cnp || pnp != p->pnp || mnp != p->mnp || nCams != p->nCams ||
      n3Dpts != p->n3Dpts || n2Dprojs != p->n2Dprojs || !finalErr)
    die_on(PSBA_ERR_ARG, "trust_region");
  int rc = psba_trust_region(p->ctx, nullptr, &itno, finalErr, nullptr, 0, nullptr);
  die_on(rc, "trust_region");
  return rc;
}

// PSBA/misc.cpp:21-49
void quat2vec(dtype* inp, int nin, dtype* outp, int nout) { psba_quat2vec(inp, nin, outp, nout); }

// PSBA/misc.cpp:60-85
void vec2quat(dtype* inp, int nin, dtype* outp, int nout) {
  int i = 0;
  if (nin > 6)
    for (; i < nin - 6; ++i) outp[i] = inp[i];
  const dtype* v = inp + i;
  outp[i + 1] = v[0];
  outp[i + 2] = v[1];
  outp[i + 3] = v[2];
  outp[i] = std::sqrt(1.0 - v[0] * v[0] - v[1] * v[1] - v[2] * v[2]);
  i += 4;
  for (; i < nout; ++i) outp[i] = inp[i - 1];
}

dtype compute_L2_sq(int n, dtype* x) {
  dtype s = 0;
  for (int i = 0; i < n; ++i) s += x[i] * x[i];
  return s;
}

dtype dotProduct(dtype* a, dtype* b, int size) {
  dtype s = 0.0;
  for (int i = 0; i < size; ++i) s += a[i] * b[i];
  return s;
}

void generate_idxs(int nCams, int n3Dpts, int n2Dprojs, dtype* impts_data, char* vmask, int* comm3DIdx,
                   int* comm3DIdxCnt, int* iidx, int* jidx, int* blk_idx) {
  (void)impts_data;
  die_on(psba_generate_idxs(nCams, n3Dpts, n2Dprojs, vmask, comm3DIdx, comm3DIdxCnt, iidx, jidx, blk_idx),
         "generate_idxs");
}
