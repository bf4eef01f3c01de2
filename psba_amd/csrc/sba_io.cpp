// Reader of the sba text format: cameras one per line, points one per line
// with their projections (PSBA/readparams.cpp:444-541 is the reference's).
// The whole file is read into memory and tokenised line by line with strtod;
// errors name the file and line and return a status code.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "psba_capi.h"
#include "psba_io.h"
#include "readparams.h"

namespace {

struct Line {
  int no;            // 1-based line number in the file
  const char* text;  // NUL-terminated
};

// Splits the file into non-comment, non-blank lines (the buffer is modified
// in place: newlines become NULs).
static bool load_lines(const char* path, std::string& buf, std::vector<Line>& lines) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  char chunk[1 << 16];
  size_t n;
  while ((n = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, n);
  fclose(f);
  buf.push_back('\n');
  int no = 0;
  size_t start = 0;
  for (size_t i = 0; i < buf.size(); ++i) {
    if (buf[i] != '\n') continue;
    buf[i] = '\0';
    ++no;
    const char* t = buf.c_str() + start;
    size_t len = i - start;
    if (len > 0 && t[len - 1] == '\r') buf[i - 1] = '\0';
    start = i + 1;
    if (t[0] == '#') continue;  // comment (the '#' must open the line)
    const char* q = t;
    while (*q == ' ' || *q == '\t' || *q == '\r') ++q;
    if (*q == '\0') continue;  // blank
    lines.push_back({no, t});
  }
  return true;
}

struct Tok {
  const char* p;
  bool next_double(double* v) {
    char* e;
    errno = 0;
    double x = strtod(p, &e);
    if (e == p) return false;
    p = e;
    *v = x;
    return true;
  }
  bool next_int(int* v) {
    char* e;
    long x = strtol(p, &e, 10);
    if (e == p) return false;
    // an integer field must not continue as a real number ("3.5")
    if (*e == '.' || *e == 'e' || *e == 'E') return false;
    p = e;
    *v = (int)x;
    return true;
  }
  int count_doubles() {
    int n = 0;
    double v;
    Tok t{p};
    while (t.next_double(&v)) ++n;
    return n;
  }
  bool at_end() const {
    const char* q = p;
    while (*q == ' ' || *q == '\t' || *q == '\r') ++q;
    return *q == '\0';
  }
};

static int bad(const char* path, int line, const char* msg) {
  char buf[400];
  snprintf(buf, sizeof(buf), "%s", msg);
  psba::set_last_error(path, line, buf);
  return PSBA_ERR_ARG;
}

enum { NOCOV = 0, FULLCOV = 1, TRICOV = 2 };

}  // namespace

// Layout [lead..., q0 q1 q2 q3, tail...] with lead = nin - 7 leading values
// (intrinsics, distortion; none for a bare pose).  q and -q are the same
// rotation; the representative with q0 >= 0 is kept and normalised, and only
// its vector part goes out.  The tail (translation) is copied.
extern "C" void psba_quat2vec(double* inp, int nin, double* outp, int nout) {
  const int lead = nin > 7 ? nin - 7 : 0;
  const double* q = inp + lead;
  double* v = outp + lead;
  for (int k = 0; k < lead; ++k) outp[k] = inp[k];
  const double qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  const double scale = (q[0] < 0.0 ? -1.0 : 1.0) / std::sqrt(qq);
  for (int k = 0; k < 3; ++k) v[k] = q[1 + k] * scale;
  const int tail = nout - lead - 3;
  for (int k = 0; k < tail; ++k) v[3 + k] = q[4 + k];
}

extern "C" void psba_sba_free(psba_sba_data* d) {
  if (!d) return;
  free(d->motstruct);
  free(d->initrot);
  free(d->imgpts);
  free(d->covimgpts);
  free(d->vmask);
  memset(d, 0, sizeof(*d));
}

extern "C" int psba_read_sba(const char* camsfname, const char* ptsfname, int cnp, int pnp, int mnp,
                             psba_cam_filter infilter, int filecnp, psba_sba_data* out) {
  if (!camsfname || !ptsfname || !out || cnp < 6 || pnp < 1 || mnp < 1) {
    psba::set_last_error(__FILE__, __LINE__, "psba_read_sba: bad arguments");
    return PSBA_ERR_ARG;
  }
  memset(out, 0, sizeof(*out));
  if (!infilter || filecnp <= 0) {
    infilter = nullptr;
    filecnp = cnp;
  }
  std::string cbuf, pbuf;
  std::vector<Line> cl, pl;
  if (!load_lines(camsfname, cbuf, cl)) return bad(camsfname, 0, "cannot open file");
  if (!load_lines(ptsfname, pbuf, pl)) return bad(ptsfname, 0, "cannot open file");
  const int ncams = (int)cl.size(), npts = (int)pl.size();
  if (ncams == 0) return bad(camsfname, 0, "no cameras");

  // ---- cameras (PSBA/readparams.cpp:187-247)
  std::vector<double> cams((size_t)ncams * cnp), rot((size_t)ncams * 4), tmp(filecnp);
  for (int j = 0; j < ncams; ++j) {
    Tok t{cl[j].text};
    int n = t.count_doubles();
    if (n != filecnp) {
      char m[160];
      snprintf(m, sizeof(m), "camera line holds %d values, expected %d", n, filecnp);
      return bad(camsfname, cl[j].no, m);
    }
    for (int k = 0; k < filecnp; ++k) t.next_double(&tmp[k]);
    double* prm = &cams[(size_t)j * cnp];
    if (infilter)
      infilter(tmp.data(), filecnp, prm, cnp);
    else
      for (int k = 0; k < cnp; ++k) prm[k] = tmp[k];
    double* q = &rot[(size_t)j * 4];
    q[1] = prm[cnp - 6];
    q[2] = prm[cnp - 5];
    q[3] = prm[cnp - 4];
    q[0] = std::sqrt(1.0 - q[1] * q[1] - q[2] * q[2] - q[3] * q[3]);
  }

  // ---- covariance kind from the first point line (PSBA/readparams.cpp:263-291)
  int havecov = NOCOV;
  const int full = mnp * mnp, tri = mnp * (mnp + 1) / 2;
  if (npts > 0) {
    Tok t{pl[0].text};
    int nfirst = t.count_doubles() - (pnp + 1);
    double v;
    for (int k = 0; k < pnp; ++k)
      if (!t.next_double(&v)) return bad(ptsfname, pl[0].no, "expecting the point's coordinates");
    int nframes = 0;
    if (!t.next_int(&nframes)) return bad(ptsfname, pl[0].no, "expecting the number of frames of the point");
    if (nframes > 0 && nfirst == nframes * (mnp + 1 + full))
      havecov = FULLCOV;
    else if (nframes > 0 && nfirst == nframes * (mnp + 1 + tri))
      havecov = TRICOV;
  }
  const int ncov = havecov == FULLCOV ? full : (havecov == TRICOV ? tri : 0);

  // ---- points and projections (PSBA/readparams.cpp:332-436)
  std::vector<double> pts((size_t)npts * pnp), proj, cov;
  std::vector<char> vm((size_t)npts * ncams, 0);
  std::vector<double> cv(full > 0 ? full : 1);
  // One point's projections as read, in file order: frame, then mnp values
  // (and the full covariance) per projection in pproj / pcov.  generate_idxs
  // numbers a point's observations in ascending camera order
  // (PSBA/misc.cpp:189-216) while the reference stores them in file order
  // (PSBA/readparams.cpp:364-415); the two agree only for sorted frame lists,
  // so each point's projections are put in ascending frame order here.
  std::vector<int> pframe, perm;
  std::vector<double> pproj, pcov;
  for (int i = 0; i < npts; ++i) {
    Tok t{pl[i].text};
    for (int k = 0; k < pnp; ++k)
      if (!t.next_double(&pts[(size_t)i * pnp + k])) return bad(ptsfname, pl[i].no, "expecting the point's coordinates");
    int nframes = 0;
    if (!t.next_int(&nframes) || nframes < 0) return bad(ptsfname, pl[i].no, "expecting the number of frames of the point");
    const size_t proj0 = proj.size(), cov0 = cov.size();
    pframe.clear();
    for (int f = 0; f < nframes; ++f) {
      int frame = 0;
      if (!t.next_int(&frame)) return bad(ptsfname, pl[i].no, "line holds fewer projections than its frame count");
      if (frame < 0 || frame >= ncams) {
        char m[160];
        snprintf(m, sizeof(m), "projection for frame %d but only %d cameras are given", frame, ncams);
        return bad(ptsfname, pl[i].no, m);
      }
      if (vm[(size_t)i * ncams + frame]) return bad(ptsfname, pl[i].no, "point projected twice in one frame");
      vm[(size_t)i * ncams + frame] = 1;
      pframe.push_back(frame);
      for (int k = 0; k < mnp; ++k) {
        double x;
        if (!t.next_double(&x)) return bad(ptsfname, pl[i].no, "line holds fewer projections than its frame count");
        proj.push_back(x);
      }
      if (ncov) {
        for (int k = 0; k < ncov; ++k)
          if (!t.next_double(&cv[k])) return bad(ptsfname, pl[i].no, "projection covariance is incomplete");
        if (havecov == FULLCOV) {
          for (int k = 0; k < full; ++k) cov.push_back(cv[k]);
        } else {
          // upper triangle row by row -> full symmetric matrix
          std::vector<double> m(full);
          int w = 0;
          for (int r = 0; r < mnp; ++r)
            for (int c = r; c < mnp; ++c) m[r * mnp + c] = m[c * mnp + r] = cv[w++];
          cov.insert(cov.end(), m.begin(), m.end());
        }
      }
    }
    if (!t.at_end()) return bad(ptsfname, pl[i].no, "line holds more values than its frame count implies");
    if (nframes > 1 && !std::is_sorted(pframe.begin(), pframe.end())) {
      perm.resize(nframes);
      for (int f = 0; f < nframes; ++f) perm[f] = f;
      std::sort(perm.begin(), perm.end(), [&](int a, int b) { return pframe[a] < pframe[b]; });
      pproj.assign(proj.begin() + proj0, proj.end());
      for (int f = 0; f < nframes; ++f)
        std::copy_n(pproj.begin() + (size_t)perm[f] * mnp, mnp, proj.begin() + proj0 + (size_t)f * mnp);
      if (ncov) {
        pcov.assign(cov.begin() + cov0, cov.end());
        for (int f = 0; f < nframes; ++f)
          std::copy_n(pcov.begin() + (size_t)perm[f] * full, full, cov.begin() + cov0 + (size_t)f * full);
      }
    }
  }

  const int nobs = (int)(proj.size() / mnp);
  out->n_cams = ncams;
  out->n_pts = npts;
  out->n_obs = nobs;
  out->cnp = cnp;
  out->pnp = pnp;
  out->mnp = mnp;
  size_t nms = (size_t)ncams * cnp + (size_t)npts * pnp;
  out->motstruct = (double*)malloc(nms * sizeof(double));
  out->initrot = (double*)malloc((size_t)ncams * 4 * sizeof(double));
  out->imgpts = (double*)malloc((proj.empty() ? 1 : proj.size()) * sizeof(double));
  out->covimgpts = ncov ? (double*)malloc(cov.size() * sizeof(double)) : nullptr;
  out->vmask = (char*)malloc(vm.empty() ? 1 : vm.size());
  if (!out->motstruct || !out->initrot || !out->imgpts || (ncov && !out->covimgpts) || !out->vmask) {
    psba_sba_free(out);
    psba::set_last_error(__FILE__, __LINE__, "psba_read_sba: out of host memory");
    return PSBA_ERR_NOMEM;
  }
  memcpy(out->motstruct, cams.data(), cams.size() * sizeof(double));
  if (!pts.empty()) memcpy(out->motstruct + cams.size(), pts.data(), pts.size() * sizeof(double));
  memcpy(out->initrot, rot.data(), rot.size() * sizeof(double));
  if (!proj.empty()) memcpy(out->imgpts, proj.data(), proj.size() * sizeof(double));
  if (ncov) memcpy(out->covimgpts, cov.data(), cov.size() * sizeof(double));
  if (!vm.empty()) memcpy(out->vmask, vm.data(), vm.size());
  return PSBA_OK;
}

// PSBA/readparams.h:8-11: the reference's reader; errors terminate the program.
void readInitialSBAEstimate(char* camsfname, char* ptsfname, int cnp, int pnp, int mnp,
                            void (*infilter)(dtype* pin, int nin, dtype* pout, int nout), int cnfp, int* ncams,
                            int* n3Dpts, int* n2Dprojs, dtype** motstruct, dtype** initrot, dtype** imgpts,
                            dtype** covimgpts, char** vmask) {
  psba_sba_data d;
  int rc = psba_read_sba(camsfname, ptsfname, cnp, pnp, mnp, infilter, cnfp, &d);
  if (rc != PSBA_OK) {
    fprintf(stderr, "readInitialSBAEstimate: %s\n", psba_last_error());
    exit(1);
  }
  *ncams = d.n_cams;
  *n3Dpts = d.n_pts;
  *n2Dprojs = d.n_obs;
  *motstruct = d.motstruct;
  *initrot = d.initrot;
  *imgpts = d.imgpts;
  *covimgpts = d.covimgpts;
  *vmask = d.vmask;
}

// Writer of the same text format (the reference declares printSBA* writers but
// leaves them commented out, PSBA/readparams.h:13-25).  A camera line is
// [K (5 values, when given)] [the model's leading intrinsics] [q0' q1' q2' q3']
// [t], where q' = q_v (x) q0 is the camera's whole rotation (the optimised
// local rotation composed with its initial one, as the projection uses it),
// normalised with q0' >= 0.  A point line is X Y Z, the number of frames and,
// per frame in ascending order, the frame index and the measured projection
// (with its full covariance when obs holds one).  Numbers use %.17g, so
// reading the files back gives the same doubles.
extern "C" int psba_write_sba(const char* camsfname, const char* ptsfname, int n_cams, int n_pts, int cnp,
                              const double* K, const double* initrot, const double* cams, const double* pts,
                              const psba_sba_data* obs) {
  if (!camsfname || !ptsfname || n_cams <= 0 || n_pts < 0 || cnp < 6 || !initrot || !cams || !pts || !obs ||
      obs->n_cams != n_cams || obs->n_pts != n_pts || !obs->vmask || (obs->n_obs > 0 && !obs->imgpts)) {
    psba::set_last_error(__FILE__, __LINE__, "psba_write_sba: bad arguments");
    return PSBA_ERR_ARG;
  }
  const int mnp = obs->mnp > 0 ? obs->mnp : 2;
  FILE* fc = fopen(camsfname, "w");
  if (!fc) return bad(camsfname, 0, "cannot create file");
  for (int j = 0; j < n_cams; ++j) {
    const double* c = cams + (size_t)j * cnp;
    const double* v = c + (cnp - 6);
    const double* q0 = initrot + (size_t)j * 4;
    const double s = std::sqrt(1.0 - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    // q = (s, v) (x) q0, Hamilton product, scalar first
    double q[4] = {s * q0[0] - v[0] * q0[1] - v[1] * q0[2] - v[2] * q0[3],
                   s * q0[1] + v[0] * q0[0] + v[1] * q0[3] - v[2] * q0[2],
                   s * q0[2] - v[0] * q0[3] + v[1] * q0[0] + v[2] * q0[1],
                   s * q0[3] + v[0] * q0[2] - v[1] * q0[1] + v[2] * q0[0]};
    const double nq = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const double sc = (q[0] < 0 ? -1.0 : 1.0) / nq;
    if (K)
      for (int k = 0; k < 5; ++k) fprintf(fc, "%.17g ", K[(size_t)j * 5 + k]);
    for (int k = 0; k < cnp - 6; ++k) fprintf(fc, "%.17g ", c[k]);
    for (int k = 0; k < 4; ++k) fprintf(fc, "%.17g ", q[k] * sc);
    fprintf(fc, "%.17g %.17g %.17g\n", c[cnp - 3], c[cnp - 2], c[cnp - 1]);
  }
  if (fclose(fc) != 0) return bad(camsfname, 0, "write failed");
  FILE* fp = fopen(ptsfname, "w");
  if (!fp) return bad(ptsfname, 0, "cannot create file");
  long o = 0;
  for (int i = 0; i < n_pts; ++i) {
    const char* vm = obs->vmask + (size_t)i * n_cams;
    int nf = 0;
    for (int j = 0; j < n_cams; ++j) nf += vm[j] != 0;
    fprintf(fp, "%.17g %.17g %.17g %d", pts[(size_t)i * 3], pts[(size_t)i * 3 + 1], pts[(size_t)i * 3 + 2], nf);
    for (int j = 0; j < n_cams; ++j) {
      if (!vm[j]) continue;
      if (o >= obs->n_obs) {
        fclose(fp);
        return bad(ptsfname, i + 1, "vmask lists more projections than obs->n_obs");
      }
      fprintf(fp, "  %d", j);
      for (int k = 0; k < mnp; ++k) fprintf(fp, " %.17g", obs->imgpts[(size_t)o * mnp + k]);
      if (obs->covimgpts)
        for (int k = 0; k < mnp * mnp; ++k) fprintf(fp, " %.17g", obs->covimgpts[(size_t)o * mnp * mnp + k]);
      ++o;
    }
    fprintf(fp, "\n");
  }
  if (fclose(fp) != 0) return bad(ptsfname, 0, "write failed");
  if (o != obs->n_obs) return bad(ptsfname, 0, "vmask lists fewer projections than obs->n_obs");
  return PSBA_OK;
}
