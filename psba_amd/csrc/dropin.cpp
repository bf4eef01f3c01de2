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
// The SSE at the entry of the last levmar call (PSBA/levmar.cpp:95; defined by
// the application at PSBA/main.cpp:24 and printed at :216).  Weak, like itno.
__attribute__((weak)) dtype initErr = 0;

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
  die_on(psba_levmar_entry_cost(p->ctx, &initErr), "levmar");
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

// PSBA/misc.cpp:60-85.  Layout [lead..., v1 v2 v3, tail...] with lead = nin - 6
// leading values (none for a bare pose); the quaternion's scalar part is the
// one that makes it a unit quaternion, taken non-negative.
void vec2quat(dtype* inp, int nin, dtype* outp, int nout) {
  const int lead = nin > 6 ? nin - 6 : 0;
  const dtype* v = inp + lead;
  dtype* q = outp + lead;
  for (int k = 0; k < lead; ++k) outp[k] = inp[k];
  const dtype vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  q[1] = v[0];
  q[2] = v[1];
  q[3] = v[2];
  q[0] = std::sqrt(1.0 - vv);
  const int tail = nout - lead - 4;  // values after the quaternion (translation)
  for (int k = 0; k < tail; ++k) q[4 + k] = v[3 + k];
}

// PSBA/misc.cpp:91-113.  R(q) for the unit quaternion q = (s, x, y, z):
//   R = (s^2 - |v|^2) I + 2 v v^T + 2 s [v]_x, with s^2 = 1 - |v|^2 written as
//   1 - 2(...) on the diagonal.
void quat2matrix(dtype* in, int nin, dtype* out) {
  dtype s, x, y, z;
  if (nin == 3) {
    x = in[0];
    y = in[1];
    z = in[2];
    s = std::sqrt(1 - x * x - y * y - z * z);
  } else {
    s = in[0];
    x = in[1];
    y = in[2];
    z = in[3];
  }
  const dtype xx = 2 * x * x, yy = 2 * y * y, zz = 2 * z * z;
  const dtype xy = 2 * x * y, xz = 2 * x * z, yz = 2 * y * z;
  const dtype sx = 2 * x * s, sy = 2 * y * s, sz = 2 * z * s;
  dtype* r0 = out;
  dtype* r1 = out + 3;
  dtype* r2 = out + 6;
  r0[0] = 1 - yy - zz;
  r1[1] = 1 - xx - zz;
  r2[2] = 1 - xx - yy;
  r0[1] = xy - sz;
  r1[0] = xy + sz;
  r0[2] = xz + sy;
  r2[0] = xz - sy;
  r1[2] = yz - sx;
  r2[1] = yz + sx;
}

// PSBA/misc.cpp:119-133: nin cameras of [quaternion (dim values), translation
// (stubForTrans values)] -> [R row-major (9 values), translation].
void camsFormatTrans(dtype* in, int nin, int dim, int stubForTrans, dtype* out) {
  const int sin = dim + stubForTrans, sout = 9 + stubForTrans;
  for (int j = 0; j < nin; ++j) {
    dtype* src = in + (size_t)j * sin;
    dtype* dst = out + (size_t)j * sout;
    quat2matrix(src, dim, dst);
    for (int k = 0; k < stubForTrans; ++k) dst[9 + k] = src[dim + k];
  }
}

// PSBA/misc.cpp:135-146
void* emalloc_(const char* file, int line, size_t sz) {
  void* ptr = malloc(sz);
  if (!ptr) {
    fprintf(stderr, "SBA: memory allocation request for %zu bytes failed in file %s, line %d, exiting", sz, file,
            line);
    exit(1);
  }
  return ptr;
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
