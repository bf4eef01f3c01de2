/*
 * psba_io.h -- reader of the sba text format (C ABI, part of libpsba.so).
 *
 * Replaces the input side of PSBA/readparams.cpp (readInitialSBAEstimate,
 * PSBA/readparams.h:8-11, PSBA/readparams.cpp:444-541) with status codes in
 * place of exit().
 *
 * Camera file: one camera per line, `filecnp` numbers each (e.g. for the
 *   reference's 11-parameter model `fu u0 v0 ar s  q0 q1 q2 q3  tx ty tz`).
 *   When `infilter` is given each line is passed through it (the reference
 *   passes quat2vec to turn the quaternion into its vector part), producing
 *   `cnp` parameters; otherwise a line must hold exactly `cnp` numbers.
 * Point file: one point per line,
 *   `X_0 .. X_{pnp-1}  nframes  frame x_0 .. x_{mnp-1} [cov] ...`
 *   with an optional full (mnp*mnp) or upper-triangular (mnp(mnp+1)/2)
 *   covariance per projection; the first point line decides which.
 * Lines whose first character is '#' are comments; lines holding only white
 *   space are skipped.
 *
 * Outputs follow the reference: motstruct = cameras (n_cams*cnp) then points
 * (n_pts*pnp); initrot = unit quaternion per camera rebuilt from the vector
 * part at cams[cnp-6..cnp-4] (PSBA/readparams.cpp:237-241); imgpts in file
 * order with each point's projections in ascending frame order (the order
 * generate_idxs numbers them in, PSBA/misc.cpp:189-216; the reference keeps
 * file order, PSBA/readparams.cpp:364-415, which disagrees for unsorted frame
 * lists); covimgpts full mnp*mnp per projection or NULL; vmask[pt*n_cams+cam].
 */
#ifndef PSBA_IO_H
#define PSBA_IO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*psba_cam_filter)(double* pin, int nin, double* pout, int nout);

typedef struct {
  int n_cams, n_pts, n_obs;
  int cnp, pnp, mnp;
  double* motstruct; /* n_cams*cnp + n_pts*pnp */
  double* initrot;   /* 4*n_cams */
  double* imgpts;    /* n_obs*mnp, file order */
  double* covimgpts; /* n_obs*mnp*mnp, or NULL when the file has none */
  char* vmask;       /* n_pts*n_cams, 1 where the point is seen */
} psba_sba_data;

/* PSBA/readparams.cpp:444-541 (readInitialSBAEstimate).  filecnp is the
 * number of values per camera line when infilter is set (ignored otherwise).
 * On success the arrays are malloc'ed and owned by the caller (release with
 * psba_sba_free); on failure nothing is allocated and psba_last_error() names
 * the file and line. */
int psba_read_sba(const char* camsfname, const char* ptsfname, int cnp, int pnp, int mnp, psba_cam_filter infilter,
                  int filecnp, psba_sba_data* out);

void psba_sba_free(psba_sba_data* d);

/* Writes solved parameters in the format psba_read_sba reads (no reference
 * counterpart: PSBA/readparams.h:13-25 leaves its printSBA* writers commented
 * out).  cams: cnp solver parameters per camera [intrinsics..., v, t] with v
 * the local rotation relative to initrot (4 per camera), as the drivers leave
 * them; each camera line holds K's 5 values (when K != NULL), the leading
 * intrinsics, the whole rotation q_v (x) initrot as a unit quaternion with
 * non-negative scalar, and t.  pts: 3 per point.  The projections (and
 * covariances) come from obs, a reader result for the same problem.  Reading
 * the files back with the quat2vec filter and filecnp = cnp + 1 (+5 with K)
 * reproduces the parameters' projections.  Returns PSBA_OK or PSBA_ERR_ARG. */
int psba_write_sba(const char* camsfname, const char* ptsfname, int n_cams, int n_pts, int cnp, const double* K,
                   const double* initrot, const double* cams, const double* pts, const psba_sba_data* obs);

/* PSBA/misc.cpp:21-49 (quat2vec): the camera-line filter the reference passes
 * to its reader; copies the leading nin-7 values, replaces the unit quaternion
 * (q0 q1 q2 q3) by its vector part sign(q0)*(q1 q2 q3)/|q| and copies the
 * translation.  The C++ drop-in quat2vec (misc.h) forwards here. */
void psba_quat2vec(double* inp, int nin, double* outp, int nout);

#ifdef __cplusplus
}
#endif

#endif
