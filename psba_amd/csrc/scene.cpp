// Seeded synthetic scenes for configurations 0-4 (see include/psba_scene.h).
#include "psba_scene.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "model.hpp"

namespace {

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {  // Box-Muller, one value per call
    double u1 = 1.0 - uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  int randint(int lo, int hi) {  // inclusive
    uint64_t span = (uint64_t)(hi - lo + 1);
    return lo + (int)(next() % span);
  }
  void unit(double* d) {
    double n;
    do {
      d[0] = normal(); d[1] = normal(); d[2] = normal();
      n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    } while (n < 1e-12);
    for (int k = 0; k < 3; ++k) d[k] /= n;
  }
};

void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
void normalize(double* v) {
  double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; ++k) v[k] /= n;
}

// rotation matrix (rows = camera axes) -> unit quaternion (s >= 0)
void rot_to_quat(const double* R, double* q) {
  double tr = R[0] + R[4] + R[8], w, x, y, z;
  if (tr > 0) {
    double s = std::sqrt(tr + 1.0) * 2;
    w = 0.25 * s; x = (R[7] - R[5]) / s; y = (R[2] - R[6]) / s; z = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2;
    w = (R[7] - R[5]) / s; x = 0.25 * s; y = (R[1] + R[3]) / s; z = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2;
    w = (R[2] - R[6]) / s; x = (R[1] + R[3]) / s; y = 0.25 * s; z = (R[5] + R[7]) / s;
  } else {
    double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2;
    w = (R[3] - R[1]) / s; x = (R[2] + R[6]) / s; y = (R[5] + R[7]) / s; z = 0.25 * s;
  }
  double n = std::sqrt(w * w + x * x + y * y + z * z);
  double sg = w < 0 ? -1.0 : 1.0;
  q[0] = sg * w / n; q[1] = sg * x / n; q[2] = sg * y / n; q[3] = sg * z / n;
}

// camera at centre c looking at target p, world up u: q and t = -R c
void look_at(const double* c, const double* p, const double* up, double* q, double* t) {
  double zc[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  normalize(zc);
  double xc[3];
  cross(zc, up, xc);
  normalize(xc);
  double yc[3];
  cross(zc, xc, yc);
  double R[9] = {xc[0], xc[1], xc[2], yc[0], yc[1], yc[2], zc[0], zc[1], zc[2]};
  rot_to_quat(R, q);
  for (int r = 0; r < 3; ++r) t[r] = -(R[3 * r] * c[0] + R[3 * r + 1] * c[1] + R[3 * r + 2] * c[2]);
}

void qmul(const double* a, const double* b, double* o) {
  psba::Quat<double> A{a[0], a[1], a[2], a[3]}, B{b[0], b[1], b[2], b[3]};
  psba::Quat<double> O = psba::qmul(A, B);
  o[0] = O.s; o[1] = O.x; o[2] = O.y; o[3] = O.z;
}

bool project(int cnp, const double* K, const double* q0, const double* cam, const double* M, double* uv,
             double* depth) {
  // depth of the point in the camera frame (positive in front)
  psba::Quat<double> q{q0[0], q0[1], q0[2], q0[3]};
  double R[9];
  psba::qmatrix(q, R);
  const double* t = cam + cnp - 3;
  *depth = R[6] * M[0] + R[7] * M[1] + R[8] * M[2] + t[2];
  switch (cnp) {
    case 6: psba::project<double, 6, false>(K, q0, cam, M, uv, nullptr, nullptr); break;
    case 9: psba::project<double, 9, false>(K, q0, cam, M, uv, nullptr, nullptr); break;
    case 11: psba::project<double, 11, false>(K, q0, cam, M, uv, nullptr, nullptr); break;
    case 16: psba::project<double, 16, false>(K, q0, cam, M, uv, nullptr, nullptr); break;
    default: return false;
  }
  return std::isfinite(uv[0]) && std::isfinite(uv[1]);
}

// truncated Zipf on [lo, hi] with exponent s
struct Zipf {
  std::vector<double> cdf;
  int lo;
  Zipf(double s, int lo_, int hi) : lo(lo_) {
    double acc = 0;
    for (int t = lo; t <= hi; ++t) {
      acc += std::pow((double)t, -s);
      cdf.push_back(acc);
    }
    for (auto& v : cdf) v /= acc;
  }
  int sample(Rng& r) const {
    double u = r.uniform();
    int k = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    if (k >= (int)cdf.size()) k = (int)cdf.size() - 1;
    return lo + k;
  }
  static double mean(double s, int lo, int hi) {
    double a = 0, b = 0;
    for (int t = lo; t <= hi; ++t) {
      double w = std::pow((double)t, -s);
      a += w * t;
      b += w;
    }
    return a / b;
  }
  static double fit(double target_mean, int lo, int hi) {
    double a = 0.5, b = 6.0;
    for (int it = 0; it < 100; ++it) {
      double m = 0.5 * (a + b);
      if (mean(m, lo, hi) > target_mean) a = m;
      else b = m;
    }
    return 0.5 * (a + b);
  }
};

struct Spec {
  int cnp = 6, nc = 0, np = 0;
  double W = 640, H = 480;
  double noise = 0.5, outlier = 0.0;
  double rot_pert = 0.02, trans_pert = 0.1, pt_pert = 0.05, intr_pert = 0.0;
  // track sampling: 0 = all visible cameras, 1 = contiguous window, 2 = random subset
  int track_mode = 0;
  int tmin = 2, tmax = 2;
  int track_dist = 0;  // 0 uniform, 1 geometric(p), 2 zipf
  double geo_p = 0.26, zipf_s = 0;
  double box_lo[3] = {-2, -2, -2}, box_hi[3] = {2, 2, 2};
};

struct Built {
  std::vector<double> K, gq, gcams, q0, cams, gpts, pts, meas;
  std::vector<int> iidx, jidx;
  std::vector<unsigned char> fix;
  double st2 = 0;
  int nout = 0;
};

void build(const Spec& sp, Rng& rng, const std::vector<double>& centres, const std::vector<double>& targets,
           const std::vector<double>& Kbase, const std::vector<double>& intr, Built& b) {
  const int nc = sp.nc, C = sp.cnp;
  b.K = Kbase;
  b.gq.resize(4 * nc);
  b.gcams.assign((size_t)C * nc, 0.0);
  const double up[3] = {0, 0, 1};
  for (int j = 0; j < nc; ++j) {
    double t[3];
    look_at(&centres[3 * j], &targets[3 * j], up, &b.gq[4 * j], t);
    for (int k = 0; k < C - 6; ++k) b.gcams[(size_t)C * j + k] = intr[(size_t)(C - 6) * j + k];
    for (int k = 0; k < 3; ++k) b.gcams[(size_t)C * j + C - 3 + k] = t[k];
  }
  Zipf* zipf = nullptr;
  if (sp.track_dist == 2) zipf = new Zipf(sp.zipf_s, sp.tmin, sp.tmax);
  // points and tracks
  std::vector<int> vis;
  std::vector<double> uvs;
  b.gpts.reserve(3 * (size_t)sp.np);
  for (int i = 0; i < sp.np; ++i) {
    for (int attempt = 0;; ++attempt) {
      double M[3];
      for (int k = 0; k < 3; ++k) M[k] = rng.uniform(sp.box_lo[k], sp.box_hi[k]);
      vis.clear();
      uvs.clear();
      for (int j = 0; j < nc; ++j) {
        double uv[2], d;
        if (!project(C, &b.K[5 * j], &b.gq[4 * j], &b.gcams[(size_t)C * j], M, uv, &d)) continue;
        if (d <= 0 || uv[0] < 0 || uv[0] >= sp.W || uv[1] < 0 || uv[1] >= sp.H) continue;
        vis.push_back(j);
        uvs.push_back(uv[0]);
        uvs.push_back(uv[1]);
      }
      std::vector<int> track;
      if (sp.track_mode == 0) {
        track = vis;
      } else {
        int t;
        if (sp.track_dist == 1) {
          t = sp.tmin;
          while (t < sp.tmax && rng.uniform() >= sp.geo_p) ++t;
        } else if (sp.track_dist == 2) {
          t = zipf->sample(rng);
        } else {
          t = rng.randint(sp.tmin, sp.tmax);
        }
        if (sp.track_mode == 1) {
          int a = rng.randint(0, nc - t);
          bool ok = true;
          for (int j = a; j < a + t; ++j) ok = ok && std::find(vis.begin(), vis.end(), j) != vis.end();
          if (ok)
            for (int j = a; j < a + t; ++j) track.push_back(j);
        } else {
          if ((int)vis.size() >= 2) {
            t = std::min(t, (int)vis.size());
            std::vector<int> pool = vis;
            for (int q = 0; q < t; ++q) {  // partial Fisher-Yates
              int r = q + (int)(rng.next() % (uint64_t)(pool.size() - q));
              std::swap(pool[q], pool[r]);
            }
            track.assign(pool.begin(), pool.begin() + t);
            std::sort(track.begin(), track.end());
          }
        }
      }
      if ((int)track.size() < 2) {
        if (attempt > 1000) abort();
        continue;
      }
      for (int k = 0; k < 3; ++k) b.gpts.push_back(M[k]);
      b.st2 += (double)track.size() * track.size();
      for (int j : track) {
        size_t vi = std::find(vis.begin(), vis.end(), j) - vis.begin();
        b.iidx.push_back(i);
        b.jidx.push_back(j);
        b.meas.push_back(uvs[2 * vi]);
        b.meas.push_back(uvs[2 * vi + 1]);
      }
      break;
    }
  }
  delete zipf;
  // measurement noise and outliers, in observation order
  const int no = (int)b.iidx.size();
  for (int o = 0; o < no; ++o) {
    if (sp.noise > 0) {
      b.meas[2 * o] += sp.noise * rng.normal();
      b.meas[2 * o + 1] += sp.noise * rng.normal();
    }
  }
  if (sp.outlier > 0) {
    for (int o = 0; o < no; ++o) {
      if (rng.uniform() < sp.outlier) {
        b.meas[2 * o] = rng.uniform(0, sp.W);
        b.meas[2 * o + 1] = rng.uniform(0, sp.H);
        b.nout++;
      }
    }
  }
  // initial estimate: camera 0 exact and fixed
  b.q0 = b.gq;
  b.cams = b.gcams;
  for (int j = 1; j < nc; ++j) {
    double axis[3];
    rng.unit(axis);
    double h = 0.5 * sp.rot_pert;
    double qp[4] = {std::cos(h), std::sin(h) * axis[0], std::sin(h) * axis[1], std::sin(h) * axis[2]};
    qmul(qp, &b.gq[4 * j], &b.q0[4 * j]);
    double d[3];
    rng.unit(d);
    for (int k = 0; k < 3; ++k) b.cams[(size_t)C * j + C - 3 + k] += sp.trans_pert * d[k];
    for (int k = 0; k < C - 6; ++k) {
      double& v = b.cams[(size_t)C * j + k];
      v = v * (1.0 + sp.intr_pert * rng.normal());
    }
  }
  b.pts = b.gpts;
  for (auto& v : b.pts) v += sp.pt_pert * rng.normal();
  b.fix.assign((size_t)C * nc, 0);
  for (int k = 0; k < C; ++k) b.fix[k] = 1;
}

template <typename T>
T* dup(const std::vector<T>& v) {
  T* p = (T*)std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

}  // namespace

extern "C" {

unsigned long long psba_problem_seed(unsigned long long seed, unsigned long long k) {
  uint64_t x = seed ^ k;
  return splitmix(x);
}

psba_scene* psba_scene_generate(int config, unsigned long long seed, const psba_scene_opts* o) {
  psba_scene_opts d;
  std::memset(&d, 0, sizeof(d));
  d.noise = -1;
  d.outlier_frac = -1;
  d.perturb = -1;
  if (o) d = *o;
  Spec sp;
  uint64_t s = seed;
  if (config == 3 || config == 4) s = psba_problem_seed(seed, d.problem);
  Rng rng(s);
  std::vector<double> centres, targets, K, intr;
  switch (config) {
    case 0: {
      sp.cnp = 6; sp.nc = d.n_cams > 0 ? d.n_cams : 7; sp.np = d.n_pts > 0 ? d.n_pts : 1000;
      sp.noise = 0.5;
      for (int j = 0; j < sp.nc; ++j) {
        double th = 2 * M_PI * j / sp.nc;
        double c[3] = {10 * std::cos(th), 10 * std::sin(th), 0};
        centres.insert(centres.end(), c, c + 3);
        targets.insert(targets.end(), {0, 0, 0});
        K.insert(K.end(), {800, 320, 240, 1, 0});
      }
      break;
    }
    case 1: {
      sp.cnp = (d.cnp == 16) ? 16 : 11;
      sp.nc = d.n_cams > 0 ? d.n_cams : 54; sp.np = d.n_pts > 0 ? d.n_pts : 5200;
      sp.noise = 1.0; sp.intr_pert = 0.02;
      sp.track_mode = 1; sp.track_dist = 1; sp.tmin = 2; sp.tmax = std::min(20, sp.nc); sp.geo_p = 0.26;
      for (int k = 0; k < 3; ++k) { sp.box_lo[k] = -3; sp.box_hi[k] = 3; }
      for (int j = 0; j < sp.nc; ++j) {
        // an arc at uneven distances, each camera aimed near the origin
        double ph = (-60.0 + 120.0 * j / std::max(1, sp.nc - 1)) * M_PI / 180.0;
        double r = 20.0 * (1.0 + 0.15 * rng.normal());
        double c[3] = {r * std::cos(ph), r * std::sin(ph), 2.0 * std::sin(2 * ph) + 0.5 * rng.normal()};
        centres.insert(centres.end(), c, c + 3);
        targets.insert(targets.end(), {0.7 * rng.normal(), 0.7 * rng.normal(), 0.7 * rng.normal()});
        K.insert(K.end(), {800, 320, 240, 1, 0});
        double fx = 800 * (1 + 0.02 * rng.normal());
        double cx = 320 + 5 * rng.normal(), cy = 240 + 5 * rng.normal(), ar = 1 + 0.005 * rng.normal();
        intr.insert(intr.end(), {fx, cx, cy, ar, 0.0});
        if (sp.cnp == 16) {
          double kc[5] = {-0.08 + 0.01 * rng.normal(), 0.02 + 0.005 * rng.normal(), 0.001 * rng.normal(),
                          0.001 * rng.normal(), 0.001 * rng.normal()};
          intr.insert(intr.end(), kc, kc + 5);
        }
      }
      break;
    }
    case 2: {
      sp.cnp = 9; sp.nc = d.n_cams > 0 ? d.n_cams : 88; sp.np = d.n_pts > 0 ? d.n_pts : 64298;
      sp.W = 1024; sp.H = 768; sp.noise = 1.0; sp.outlier = 0.01;
      sp.rot_pert = 0.005; sp.trans_pert = 0.5; sp.pt_pert = 0.2; sp.intr_pert = 0.0;
      sp.track_mode = 2; sp.track_dist = 2; sp.tmin = 2; sp.tmax = sp.nc;
      sp.zipf_s = d.zipf_s > 0 ? d.zipf_s : Zipf::fit(384000.0 / 64298.0, 2, sp.nc);
      sp.box_lo[0] = -50; sp.box_hi[0] = 50; sp.box_lo[1] = -50; sp.box_hi[1] = 50;
      sp.box_lo[2] = -10; sp.box_hi[2] = 10;
      const double V[4][3] = {{-100, -140, 70}, {-35, -160, 80}, {35, -130, 75}, {100, -150, 80}};
      double seg[3], total = 0;
      for (int q = 0; q < 3; ++q) {
        seg[q] = std::sqrt(std::pow(V[q + 1][0] - V[q][0], 2) + std::pow(V[q + 1][1] - V[q][1], 2) +
                           std::pow(V[q + 1][2] - V[q][2], 2));
        total += seg[q];
      }
      for (int j = 0; j < sp.nc; ++j) {
        double a = total * j / std::max(1, sp.nc - 1);
        int q = 0;
        while (q < 2 && a > seg[q]) { a -= seg[q]; ++q; }
        double f = std::min(1.0, a / seg[q]);
        double c[3];
        for (int k = 0; k < 3; ++k) c[k] = V[q][k] + f * (V[q + 1][k] - V[q][k]);
        centres.insert(centres.end(), c, c + 3);
        targets.insert(targets.end(), {3 * rng.normal(), 3 * rng.normal(), 0});
        K.insert(K.end(), {900, 512, 384, 1, 0});
        intr.insert(intr.end(), {900 * (1 + 0.02 * rng.normal()), -0.05 + 0.01 * rng.normal(),
                                 0.01 + 0.003 * rng.normal()});
      }
      break;
    }
    case 3: {
      sp.cnp = 6;
      sp.nc = d.n_cams > 0 ? d.n_cams : rng.randint(5, 30);
      sp.np = d.n_pts > 0 ? d.n_pts : rng.randint(200, 4000);
      sp.noise = 0.5; sp.track_mode = 2; sp.track_dist = 0; sp.tmin = 2; sp.tmax = std::min(8, sp.nc);
      for (int j = 0; j < sp.nc; ++j) {
        double ph = (-45.0 + 90.0 * j / std::max(1, sp.nc - 1)) * M_PI / 180.0;
        double c[3] = {10 * std::cos(ph), 10 * std::sin(ph), 0};
        centres.insert(centres.end(), c, c + 3);
        targets.insert(targets.end(), {0, 0, 0});
        K.insert(K.end(), {800, 320, 240, 1, 0});
      }
      break;
    }
    case 4: {
      sp.cnp = 6; sp.noise = 0.5; sp.track_mode = 2; sp.tmin = 2;
      double span = 90;
      if (d.size_class == 0) {
        sp.nc = 10; sp.np = 1000; sp.track_dist = 0; sp.tmax = 10;
      } else if (d.size_class == 1) {
        sp.nc = 50; sp.np = 20000; sp.track_dist = 0; sp.tmax = 8; span = 180;
      } else {
        sp.nc = 138; sp.np = 19878; sp.track_dist = 2; sp.tmax = 138; span = 270;
        sp.zipf_s = Zipf::fit(100000.0 / 19878.0, 2, 138);
      }
      if (d.n_cams > 0) sp.nc = d.n_cams;
      if (d.n_pts > 0) sp.np = d.n_pts;
      sp.tmax = std::min(sp.tmax, sp.nc);
      for (int j = 0; j < sp.nc; ++j) {
        double ph = (-span / 2 + span * j / std::max(1, sp.nc - 1)) * M_PI / 180.0;
        double c[3] = {10 * std::cos(ph), 10 * std::sin(ph), 0.5 * std::sin(3 * ph)};
        centres.insert(centres.end(), c, c + 3);
        targets.insert(targets.end(), {0, 0, 0});
        K.insert(K.end(), {800, 320, 240, 1, 0});
      }
      break;
    }
    default:
      return nullptr;
  }
  if (d.noise >= 0) sp.noise = d.noise;
  if (d.outlier_frac >= 0) sp.outlier = d.outlier_frac;
  if (d.perturb >= 0) {
    sp.rot_pert *= d.perturb; sp.trans_pert *= d.perturb; sp.pt_pert *= d.perturb; sp.intr_pert *= d.perturb;
  }
  Built b;
  build(sp, rng, centres, targets, K, intr, b);
  psba_scene* out = (psba_scene*)std::calloc(1, sizeof(psba_scene));
  out->cnp = sp.cnp;
  out->n_cams = sp.nc;
  out->n_pts = sp.np;
  out->n_obs = (int)b.iidx.size();
  out->K = dup(b.K);
  out->q0 = dup(b.q0);
  out->cams = dup(b.cams);
  out->pts = dup(b.pts);
  out->meas = dup(b.meas);
  out->iidx = dup(b.iidx);
  out->jidx = dup(b.jidx);
  out->fixmask = dup(b.fix);
  out->gt_q0 = dup(b.gq);
  out->gt_cams = dup(b.gcams);
  out->gt_pts = dup(b.gpts);
  out->sum_t2 = b.st2;
  out->n_outliers = b.nout;
  out->zipf_s = sp.zipf_s;
  return out;
}

void psba_scene_free(psba_scene* s) {
  if (!s) return;
  void* p[] = {s->K, s->q0, s->cams, s->pts, s->meas, s->iidx, s->jidx, s->fixmask, s->gt_q0, s->gt_cams, s->gt_pts};
  for (void* q : p) std::free(q);
  std::free(s);
}

}  // extern "C"
