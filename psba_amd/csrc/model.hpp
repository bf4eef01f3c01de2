// Camera models of the bundle adjuster: projection and analytic Jacobians.
//
// Pose: camera parameters end with [v1 v2 v3 t1 t2 t3]; v is the vector part of
// a local unit quaternion q_v = (sqrt(1-|v|^2), v) composed on the left of the
// camera's fixed initial rotation q0 (q = q_v (x) q0, Hamilton, scalar first),
// X' = R(q) M + t.  This is the pose model of PSBA/CL_files/compute_exQT.cl:
// 18-71 and compute_jacobiQT.cl:7-141.  Pixel: x = X'/Z', y = Y'/Z',
// (xd, yd) = distortion(x, y), u = fx xd + skew yd + cx, v = fx ar yd + cy.
//
//   CNP 6 : [v t]                      fx cx cy ar skew fixed (K5)
//   CNP 9 : [f k1 k2 v t]              radial 1 + k1 r^2 + k2 r^4; cx cy ar skew fixed
//   CNP 11: [fx cx cy ar skew v t]
//   CNP 16: [fx cx cy ar skew kc1..kc5 v t]   radial kc1,kc2,kc5; tangential kc3,kc4
//
// JA is the 2 x CNP row-major Jacobian of (u,v) w.r.t. the camera parameters,
// JB the 2 x 3 Jacobian w.r.t. the point.  T is double or float.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace psba {

template <typename T>
struct Quat {
  T s, x, y, z;
};

template <typename T>
__host__ __device__ inline Quat<T> qmul(const Quat<T>& a, const Quat<T>& b) {
  Quat<T> o;
  o.s = a.s * b.s - a.x * b.x - a.y * b.y - a.z * b.z;
  o.x = a.s * b.x + a.x * b.s + a.y * b.z - a.z * b.y;
  o.y = a.s * b.y - a.x * b.z + a.y * b.s + a.z * b.x;
  o.z = a.s * b.z + a.x * b.y - a.y * b.x + a.z * b.s;
  return o;
}

// R(q) for a unit quaternion in the homogeneous form
// R = (s^2 - |u|^2) I + 2 u u^T + 2 s [u]_x ; row-major
template <typename T>
__host__ __device__ inline void qmatrix(const Quat<T>& q, T* R) {
  T w = q.s * q.s - (q.x * q.x + q.y * q.y + q.z * q.z);
  T xs = 2 * q.s * q.x, ys = 2 * q.s * q.y, zs = 2 * q.s * q.z;
  R[0] = w + 2 * q.x * q.x; R[1] = 2 * q.x * q.y - zs;  R[2] = 2 * q.x * q.z + ys;
  R[3] = 2 * q.y * q.x + zs; R[4] = w + 2 * q.y * q.y;  R[5] = 2 * q.y * q.z - xs;
  R[6] = 2 * q.z * q.x - ys; R[7] = 2 * q.z * q.y + xs; R[8] = w + 2 * q.z * q.z;
}

template <int CNP>
struct ModelTraits {
  static constexpr bool supported = (CNP == 6 || CNP == 9 || CNP == 11 || CNP == 16);
  static constexpr int n_intr = CNP - 6;
};

// Projection (and optionally Jacobians) of point M through camera `cam`.
// K5 = [fx cx cy ar skew] supplies whatever the model does not optimise.
template <typename T, int CNP, bool WANT_JAC>
__host__ __device__ inline void project(const T* K5, const T* q0p, const T* cam, const T* M, T* uv,
                                        T* JA, T* JB) {
  static_assert(ModelTraits<CNP>::supported, "unsupported camera model");
  const T* v = cam + (CNP - 6);
  const T* tr = cam + (CNP - 3);
  T s = sqrt(T(1) - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
  Quat<T> q0{q0p[0], q0p[1], q0p[2], q0p[3]};
  Quat<T> qv{s, v[0], v[1], v[2]};
  Quat<T> q = qmul(qv, q0);
  T R[9];
  qmatrix(q, R);
  T X = R[0] * M[0] + R[1] * M[1] + R[2] * M[2] + tr[0];
  T Y = R[3] * M[0] + R[4] * M[1] + R[5] * M[2] + tr[1];
  T Z = R[6] * M[0] + R[7] * M[1] + R[8] * M[2] + tr[2];
  T iz = T(1) / Z;
  T x = X * iz, y = Y * iz;

  T fx, cx, cy, ar, sk;
  if constexpr (CNP == 6) {
    fx = K5[0]; cx = K5[1]; cy = K5[2]; ar = K5[3]; sk = K5[4];
  } else if constexpr (CNP == 9) {
    fx = cam[0]; cx = K5[1]; cy = K5[2]; ar = K5[3]; sk = K5[4];
  } else {
    fx = cam[0]; cx = cam[1]; cy = cam[2]; ar = cam[3]; sk = cam[4];
  }

  T r2 = x * x + y * y;
  T xd = x, yd = y;
  // Jacobian of (xd, yd) w.r.t. (x, y)
  T j00 = 1, j01 = 0, j10 = 0, j11 = 1;
  T radial = 1, dradial = 0;
  if constexpr (CNP == 9) {
    T k1 = cam[1], k2 = cam[2];
    radial = 1 + r2 * (k1 + r2 * k2);
    dradial = k1 + 2 * k2 * r2;
    xd = x * radial;
    yd = y * radial;
  } else if constexpr (CNP == 16) {
    const T* kc = cam + 5;
    radial = 1 + r2 * (kc[0] + r2 * (kc[1] + r2 * kc[4]));
    dradial = kc[0] + r2 * (2 * kc[1] + 3 * kc[4] * r2);
    T xy = x * y;
    xd = x * radial + 2 * kc[2] * xy + kc[3] * (r2 + 2 * x * x);
    yd = y * radial + kc[2] * (r2 + 2 * y * y) + 2 * kc[3] * xy;
  }
  uv[0] = fx * xd + sk * yd + cx;
  uv[1] = fx * ar * yd + cy;
  if constexpr (!WANT_JAC) {
    return;
  } else {
    if constexpr (CNP == 9) {
      j00 = radial + 2 * x * x * dradial;
      j01 = 2 * x * y * dradial;
      j10 = j01;
      j11 = radial + 2 * y * y * dradial;
    } else if constexpr (CNP == 16) {
      const T* kc = cam + 5;
      j00 = radial + 2 * x * x * dradial + 2 * kc[2] * y + 6 * kc[3] * x;
      j01 = 2 * x * y * dradial + 2 * kc[2] * x + 2 * kc[3] * y;
      j10 = j01;
      j11 = radial + 2 * y * y * dradial + 6 * kc[2] * y + 2 * kc[3] * x;
    }
    // H = d(u,v)/d(xd,yd) * d(xd,yd)/d(x,y)
    T fa = fx * ar;
    T h00 = fx * j00 + sk * j10, h01 = fx * j01 + sk * j11;
    T h10 = fa * j10, h11 = fa * j11;
    // G = d(u,v)/dX' (2x3)
    T G[6] = {h00 * iz, h01 * iz, -(h00 * x + h01 * y) * iz,
              h10 * iz, h11 * iz, -(h10 * x + h11 * y) * iz};
    // point block
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 3; ++c)
        JB[3 * r + c] = G[3 * r] * R[c] + G[3 * r + 1] * R[3 + c] + G[3 * r + 2] * R[6 + c];
    // rotation block via dq/dv_k = (-v_k/s) q0 + (0,e_k) (x) q0 and the
    // derivative of R(q)M = (s^2-|u|^2) M + 2 (u.M) u + 2 s (u x M) w.r.t. q
    T uM = q.x * M[0] + q.y * M[1] + q.z * M[2];
    T cxM[3] = {q.y * M[2] - q.z * M[1], q.z * M[0] - q.x * M[2], q.x * M[1] - q.y * M[0]};
    T dS[3], dU[3][3];  // dRM/ds, dRM/du_m (column m)
    for (int r = 0; r < 3; ++r) dS[r] = 2 * q.s * M[r] + 2 * cxM[r];
    const T u[3] = {q.x, q.y, q.z};
    for (int m = 0; m < 3; ++m) {
      // -2 u_m M + 2 M_m u + 2 (u.M) e_m + 2 s (e_m x M)
      T em_x_M[3] = {T(0), T(0), T(0)};
      if (m == 0) { em_x_M[1] = -M[2]; em_x_M[2] = M[1]; }
      if (m == 1) { em_x_M[0] = M[2]; em_x_M[2] = -M[0]; }
      if (m == 2) { em_x_M[0] = -M[1]; em_x_M[1] = M[0]; }
      for (int r = 0; r < 3; ++r)
        dU[r][m] = -2 * u[m] * M[r] + 2 * M[m] * u[r] + (r == m ? 2 * uM : T(0)) + 2 * q.s * em_x_M[r];
    }
    T is = T(1) / s;
    for (int k = 0; k < 3; ++k) {
      Quat<T> ek{T(0), T(k == 0), T(k == 1), T(k == 2)};
      Quat<T> dq = qmul(ek, q0);
      T a = -v[k] * is;
      dq.s += a * q0.s; dq.x += a * q0.x; dq.y += a * q0.y; dq.z += a * q0.z;
      T dX[3];
      for (int r = 0; r < 3; ++r) dX[r] = dS[r] * dq.s + dU[r][0] * dq.x + dU[r][1] * dq.y + dU[r][2] * dq.z;
      for (int r = 0; r < 2; ++r)
        JA[r * CNP + (CNP - 6) + k] = G[3 * r] * dX[0] + G[3 * r + 1] * dX[1] + G[3 * r + 2] * dX[2];
    }
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 2; ++r) JA[r * CNP + (CNP - 3) + k] = G[3 * r + k];
    // intrinsic / distortion columns
    if constexpr (CNP == 9) {
      JA[0] = xd; JA[CNP] = ar * yd;
      T xr = x * r2, yr = y * r2;
      JA[1] = fx * xr + sk * yr; JA[CNP + 1] = fa * yr;
      JA[2] = (fx * xr + sk * yr) * r2; JA[CNP + 2] = fa * yr * r2;
    } else if constexpr (CNP == 11 || CNP == 16) {
      JA[0] = xd;    JA[CNP + 0] = ar * yd;
      JA[1] = 1;     JA[CNP + 1] = 0;
      JA[2] = 0;     JA[CNP + 2] = 1;
      JA[3] = 0;     JA[CNP + 3] = fx * yd;
      JA[4] = yd;    JA[CNP + 4] = 0;
      if constexpr (CNP == 16) {
        T r4 = r2 * r2;
        T gx[5] = {x * r2, x * r4, 2 * x * y, r2 + 2 * x * x, x * r4 * r2};
        T gy[5] = {y * r2, y * r4, r2 + 2 * y * y, 2 * x * y, y * r4 * r2};
        for (int k = 0; k < 5; ++k) {
          JA[5 + k] = fx * gx[k] + sk * gy[k];
          JA[CNP + 5 + k] = fa * gy[k];
        }
      }
    }
  }
}

}  // namespace psba
