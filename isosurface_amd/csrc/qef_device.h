// Feature placement of dual contouring on the device: TangentPlanes
// classification in f32 and the QEF solve of MinimiseQEF through the SVD of
// the 8 x 3 normal matrix in f64.
//
// Bit-identical results with the reference require its exact arithmetic, so
// every step follows feature/mod.rs:91-148, feature/qef.rs:36-68 and
// math/svd.rs:25-84, 130-972 operation for operation (the library is built
// with -ffp-contract=off; f64 sqrt and division are correctly rounded).  The
// one liberty taken is data movement: the reference copies the active column
// window into scratch matrices before each QR step, here the step works on
// the window in place, which performs the same arithmetic on the same values.
#pragma once
#include <hip/hip_runtime.h>

namespace iso {
namespace qef {

constexpr int kRows = 8;     // dual contouring: the 8 cell corners
constexpr int kMaxRows = 12;  // extended marching cubes: up to 12 edge crossings

struct Vec3f {
  float x, y, z;
};

__device__ __forceinline__ float dot3(Vec3f a, Vec3f b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;  // vector.rs:240-242, 290-292
}

__device__ __forceinline__ Vec3f cross3(Vec3f a, Vec3f b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Vec3::normalised().unwrap_or_default() (vector.rs:135-144).
__device__ __forceinline__ Vec3f normalised_or_zero(Vec3f v) {
  const float l = __builtin_sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
  if (__builtin_fabsf(l) < 1.1920929e-07f) return {0.0f, 0.0f, 0.0f};
  return {v.x / l, v.y / l, v.z / l};
}

// The SVD of a rows x 3 matrix (rows <= R): every loop over rows runs to
// `rows`, so the arithmetic is that of the reference's rows-row matrix.
template <int R>
struct SvdT {
  double u[R][3];
  double v[3][3];
  double d[3];
  int rows;
};

// Householder reflection of the n entries x[0], x[stride], ... (svd.rs:251-293).
__device__ __forceinline__ double householder(double* x, int n, int stride) {
  double tau = 0.0;
  if (n > 1) {
    double xnorm;
    if (n == 2) {
      xnorm = __builtin_fabs(x[stride]);
    } else {
      double scl = 0.0, ssq = 1.0;
      for (int i = 1; i < n; ++i) {
        const double a = __builtin_fabs(x[i * stride]);
        if (a != 0.0) {
          if (scl < a) {
            ssq = 1.0 + ssq * (scl / a) * (scl / a);
            scl = a;
          } else {
            ssq += a / scl * (a / scl);
          }
        }
      }
      xnorm = scl * __builtin_sqrt(ssq);
    }
    if (xnorm != 0.0) {
      const double alpha = x[0];
      double beta = __builtin_sqrt(alpha * alpha + xnorm * xnorm);
      if (alpha >= 0.0) beta = -beta;
      tau = (beta - alpha) / beta;
      const double scl = 1.0 / (alpha - beta);
      x[0] = beta;
      for (int i = 1; i < n; ++i) x[i * stride] *= scl;
    }
  }
  return tau;
}

// Bidiagonal factorisation (svd.rs:149-249).
template <int R>
__device__ __forceinline__ void bidiagonalise(double m[R][3], double tau_u[3], double tau_v[2],
                                              int rows) {
  for (int i = 0; i < 3; ++i) {
    double tau = householder(&m[i][i], rows - i, 3);
    tau_u[i] = tau;
    if (i + 1 < 3) {
      if (tau != 0.0) {
        for (int x = i + 1; x < 3; ++x) {
          double wx = m[i][x];
          for (int y = i + 1; y < rows; ++y) wx += m[y][x] * m[y][i];
          const double tw = tau * wx;
          m[i][x] -= tw;
          for (int y = i + 1; y < rows; ++y) m[y][x] -= tw * m[y][i];
        }
      }
      tau = householder(&m[i][i + 1], i == 0 ? 2 : 1, 1);
      tau_v[i] = tau;
      if (tau != 0.0) {
        for (int y = i + 1; y < rows; ++y) {
          double wy = m[y][i + 1];
          if (i == 0) wy += m[y][i + 2] * m[i][i + 2];
          const double tw = tau * wy;
          m[y][i + 1] -= tw;
          if (i == 0) m[y][i + 2] -= tw * m[i][i + 2];
        }
      }
    }
  }
}

// Accumulates the reflections into U and V (svd.rs:295-410).
template <int R>
__device__ __forceinline__ void unpack(double u[R][3], double v[3][3], double tau_u[3],
                                       double tau_v[2], int rows) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i][j] = i == j ? 1.0 : 0.0;
  for (int i = 1; i >= 0; --i) {
    const double tau = tau_v[i];
    if (tau != 0.0) {
      for (int x = i + 1; x < 3; ++x) {
        double wx = v[i + 1][x];
        for (int y = i + 2; y < 3; ++y) wx += v[y][x] * u[i][y];
        const double tw = tau * wx;
        v[i + 1][x] -= tw;
        for (int y = i + 2; y < 3; ++y) v[y][x] -= tw * u[i][y];
      }
    }
  }
  for (int i = 0; i < 2; ++i) tau_v[i] = u[i][i + 1];
  for (int i = 2; i >= 0; --i) {
    const double tau = tau_u[i];
    tau_u[i] = u[i][i];
    if (tau == 0.0) {
      u[i][i] = 1.0;
      if (i < 2) {
        u[i][2] = 0.0;
        if (i < 1) u[i][1] = 0.0;
      }
      for (int y = i + 1; y < rows; ++y) u[y][i] = 0.0;
    } else {
      for (int x = i + 1; x < 3; ++x) {
        double wx = 0.0;
        for (int y = i + 1; y < rows; ++y) wx += u[y][x] * u[y][i];
        const double tw = tau * wx;
        u[i][x] = -tw;
        for (int y = i + 1; y < rows; ++y) u[y][x] -= tw * u[y][i];
      }
      for (int y = i + 1; y < rows; ++y) u[y][i] = u[y][i] * -tau;
      u[i][i] = 1.0 - tau;
    }
  }
}

__device__ __forceinline__ void givens(double a, double b, double* c, double* s) {  // svd.rs:825-842
  if (b == 0.0) {
    *c = 1.0;
    *s = 0.0;
  } else if (__builtin_fabs(b) > __builtin_fabs(a)) {
    const double t = -a / b;
    const double s1 = 1.0 / __builtin_sqrt(1.0 + t * t);
    *s = s1;
    *c = s1 * t;
  } else {
    const double t = -b / a;
    const double c1 = 1.0 / __builtin_sqrt(1.0 + t * t);
    *c = c1;
    *s = c1 * t;
  }
}

__device__ __forceinline__ void schur(double a1, double a2, double a3, double* c, double* s) {  // svd.rs:844-865
  const double apq = a1 * a2 * 2.0;
  if (apq == 0.0) {
    *c = 1.0;
    *s = 0.0;
  } else {
    const double tau = (a2 * a2 + (a3 + a1) * (a3 - a1)) / apq;
    const double t = tau >= 0.0 ? 1.0 / (tau + __builtin_sqrt(1.0 + tau * tau))
                                : -1.0 / (__builtin_sqrt(1.0 + tau * tau) - tau);
    *c = 1.0 / __builtin_sqrt(1.0 + t * t);
    *s = t * *c;
  }
}

__device__ __forceinline__ void chop(double* a, double* b, int n) {  // svd.rs:481-493
  double ai = a[0];
  for (int i = 0; i < n - 1; ++i) {
    const double bi = b[i];
    const double ai1 = a[i + 1];
    if (__builtin_fabs(bi) < 1e-5 * (__builtin_fabs(ai) + __builtin_fabs(ai1))) b[i] = 0.0;
    ai = ai1;
  }
}

// Rotates columns p and q (absolute) of the first `rows` rows by (c, s).
template <int R>
__device__ __forceinline__ void rotate(double m[R][3], int rows, int p, int q, double c,
                                       double s) {
  for (int i = 0; i < rows; ++i) {
    const double mp = m[i][p], mq = m[i][q];
    m[i][p] = mp * c - mq * s;
    m[i][q] = mp * s + mq * c;
  }
}

__device__ __forceinline__ void swap_cols01(double v[3][3], int a) {
  for (int i = 0; i < 3; ++i) {
    const double t = v[i][a];
    v[i][a] = v[i][a + 1];
    v[i][a + 1] = t;
  }
}

// QR step on the 2-column window [a, a+1] (svd.rs:689-823).
template <int R>
__device__ __forceinline__ void qr_cols2(SvdT<R>& S, double* tu, double* tv, int a) {
  double c, s;
  if (tu[0] == 0.0) {
    givens(tv[0], tu[1], &c, &s);
    tu[0] = tv[0] * c - tu[1] * s;
    tv[0] = tv[0] * s + tu[1] * c;
    tu[1] = 0.0;
    rotate<R>(S.u, S.rows, a, a + 1, c, s);
    swap_cols01(S.v, a);
  } else if (tu[1] == 0.0) {
    givens(tu[0], tv[0], &c, &s);
    tu[0] = tu[0] * c - tv[0] * s;
    tv[0] = 0.0;
    rotate<3>(S.v, 3, a, a + 1, c, s);
  } else {
    schur(tu[0], tv[0], tu[1], &c, &s);
    double a11 = tu[0] * c - tv[0] * s;
    double a21 = -tu[1] * s;
    double a12 = tu[0] * s + tv[0] * c;
    double a22 = tu[1] * c;
    rotate<3>(S.v, 3, a, a + 1, c, s);
    if (a11 * a11 + a21 * a21 < a12 * a12 + a22 * a22) {
      double t = a11;
      a11 = a12;
      a12 = t;
      t = a21;
      a21 = a22;
      a22 = t;
      swap_cols01(S.v, a);
    }
    givens(a11, a21, &c, &s);
    tu[0] = a11 * c - a21 * s;
    tv[0] = a12 * c - a22 * s;
    tu[1] = a12 * s + a22 * c;
    rotate<R>(S.u, S.rows, a, a + 1, c, s);
  }
}

// A zero on the diagonal inside the 3-column window (svd.rs:604-641).
template <int R>
__device__ __forceinline__ void qr_middle(SvdT<R>& S, double* tu, double* tv, int cols,
                                          int col) {
  double x = tv[col], y = tu[col + 1];
  for (int j = col; j < cols - 1; ++j) {
    double c, s;
    givens(y, -x, &c, &s);
    for (int i = 0; i < S.rows; ++i) {
      const double up = S.u[i][col], uq = S.u[i][j + 1];
      S.u[i][col] = up * c - uq * s;
      S.u[i][j + 1] = up * s + uq * c;
    }
    tu[j + 1] = x * s + y * c;
    if (j == col) tv[j] = x * c - y * s;
    if (j < cols - 2) {
      const double z = tv[j + 1];
      tv[j + 1] *= c;
      x = z * -s;
      y = tu[j + 2];
    }
  }
}

// A zero at the end of the diagonal (svd.rs:643-671).
template <int R>
__device__ __forceinline__ void qr_end(SvdT<R>& S, double* tu, double* tv) {
  double x = tu[1], y = tv[1];
  for (int k = 1; k >= 0; --k) {
    double c, s;
    givens(x, y, &c, &s);
    for (int i = 0; i < 3; ++i) {
      const double vp = S.v[i][k], vq = S.v[i][2];
      S.v[i][k] = vp * c - vq * s;
      S.v[i][2] = vp * s + vq * c;
    }
    tu[k] = x * c - y * s;
    if (k == 1) tv[k] = x * s + y * c;
    if (k > 0) {
      const double z = tv[k - 1];
      tv[k - 1] *= c;
      x = tu[k - 1];
      y = z * s;
    }
  }
}

__device__ __forceinline__ double qr_shift(const double* tu, const double* tv) {  // svd.rs:673-687
  const double ta = tu[1] * tu[1] + tv[0] * tv[0];
  const double tb = tu[2] * tu[2] + tv[1] * tv[1];
  const double tab = tu[1] * tv[1];
  const double dt = (ta - tb) / 2.0;
  if (dt >= 0.0) return tb - tab * tab / (dt + __builtin_sqrt(dt * dt + tab * tab));
  return tb + tab * tab / (__builtin_sqrt(dt * dt + tab * tab) - dt);
}

// One implicit-shift QR sweep on the window [a, a+cols) (svd.rs:495-602).
template <int R>
__device__ __forceinline__ void qr_step(SvdT<R>& S, double* tu, double* tv, int a, int cols) {
  if (cols == 2) {
    qr_cols2(S, tu, tv, a);
    return;
  }
  // cols == 3 here, so the window is the whole matrix (a == 0).
  for (int i = 0; i < cols - 1; ++i) {
    if (tu[i] == 0.0) {
      qr_middle(S, tu, tv, cols, i);
      return;
    }
  }
  if (tu[cols - 1] == 0.0) {
    qr_end(S, tu, tv);
    return;
  }
  const double mu = qr_shift(tu, tv);
  double y = tu[0] * tu[0] - mu;
  double z = tu[0] * tv[0];
  double ak = 0.0, bk = 0.0, zk = 0.0;
  double ap = tu[0], bp = tv[0], aq = tu[1];
  for (int k = 0; k < cols - 1; ++k) {
    double c, s;
    givens(y, z, &c, &s);
    rotate<3>(S.v, 3, k, k + 1, c, s);
    double bk1 = bk * c - z * s;
    double ap1 = ap * c - bp * s;
    double bp1 = ap * s + bp * c;
    double zp1 = aq * -s;
    double aq1 = aq * c;
    if (k > 0) tv[k - 1] = bk1;
    ak = ap1;
    bk = bp1;
    zk = zp1;
    ap = aq1;
    bp = k < cols - 2 ? tv[k + 1] : 0.0;
    y = ak;
    z = zk;
    givens(y, z, &c, &s);
    rotate<R>(S.u, S.rows, k, k + 1, c, s);
    const double ak1 = ak * c - zk * s;
    bk1 = bk * c - ap * s;
    const double zk1 = bp * -s;
    ap1 = bk * s + ap * c;
    bp1 = bp * c;
    tu[k] = ak1;
    ak = ak1;
    bk = bk1;
    zk = zk1;
    ap = ap1;
    bp = bp1;
    aq = k < cols - 2 ? tu[k + 2] : 0.0;
    y = bk;
    z = zk;
  }
  tv[cols - 2] = bk;
  tu[cols - 1] = ap;
}

// Diagonalisation (svd.rs:413-479).  The reference has no iteration cap;
// `max_sweeps` bounds the loop on the device so that every thread finishes
// (a cell that hits it is reported through *stalled).
template <int R>
__device__ __forceinline__ void diagonalise(SvdT<R>& S, double tu[3], double tv[2], int max_sweeps,
                                            bool* stalled) {
  chop(tu, tv, 3);
  int b = 2;
  int sweeps = 0;
  while (b > 0) {
    if (tv[b - 1] == 0.0) {
      --b;
      continue;
    }
    if (++sweeps > max_sweeps) {
      *stalled = true;
      return;
    }
    int a = b - 1;
    while (a > 0 && tv[a - 1] != 0.0) --a;
    const int n = b - a + 1;
    qr_step(S, &tu[a], &tv[a], a, n);
    chop(&tu[a], &tv[a], n);
  }
}

template <int R>
__device__ __forceinline__ void sort_singular(SvdT<R>& S) {  // svd.rs:867-926
  for (int j = 0; j < 3; ++j) {
    if (S.d[j] < 0.0) {
      for (int i = 0; i < 3; ++i) S.v[i][j] = -S.v[i][j];
      S.d[j] = -S.d[j];
    }
  }
  for (int i = 0; i < 3; ++i) {
    double dmax = S.d[i];
    int imax = i;
    for (int j = i + 1; j < 3; ++j) {
      if (S.d[j] > dmax) {
        dmax = S.d[j];
        imax = j;
      }
    }
    if (imax != i) {
      double t = S.d[i];
      S.d[i] = S.d[imax];
      S.d[imax] = t;
      for (int y = 0; y < S.rows; ++y) {
        t = S.u[y][i];
        S.u[y][i] = S.u[y][imax];
        S.u[y][imax] = t;
      }
      for (int y = 0; y < 3; ++y) {
        t = S.v[y][i];
        S.v[y][i] = S.v[y][imax];
        S.v[y][imax] = t;
      }
    }
  }
}

// TangentPlanes::new (feature/mod.rs:91-148) over n <= R points and normals
// + MinimiseQEF::place_feature_with_tangents (qef.rs:36-68); *feature
// receives whether the planes were classified Edge or Corner.
template <int R>
__device__ __forceinline__ Vec3f place_feature_n(const Vec3f* corners, const Vec3f* normals,
                                                 int n, bool* stalled, bool* feature) {
  const float kFeatureAngle = 0.866025388f;  // f32(0.8660254037844387), mod.rs:22
  Vec3f com{0.0f, 0.0f, 0.0f}, axis{0.0f, 0.0f, 0.0f};
  float min_angle = 3.40282347e+38f;
  float count = 0.0f;
  *feature = false;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      const float angle = dot3(normals[i], normals[j]);
      if (angle < min_angle) {
        axis = cross3(normals[i], normals[j]);
        min_angle = angle;
      }
    }
    com.x += corners[i].x;
    com.y += corners[i].y;
    com.z += corners[i].z;
    count += 1.0f;
  }
  com.x /= count;
  com.y /= count;
  com.z /= count;
  if (min_angle > kFeatureAngle) return com;  // planar
  *feature = true;
  axis = normalised_or_zero(axis);
  float min_c = 1.0f, max_c = -1.0f;
  for (int i = 0; i < n; ++i) {
    const float c = dot3(axis, normals[i]);
    min_c = fminf(min_c, c);
    max_c = fmaxf(max_c, c);
  }
  float c = fmaxf(__builtin_fabsf(min_c), __builtin_fabsf(max_c));
  c = __builtin_sqrtf(1.0f - c * c);
  const bool edge = c > kFeatureAngle;

  SvdT<R> S;
  S.rows = n;
  double b[R];
  for (int i = 0; i < n; ++i) {
    S.u[i][0] = normals[i].x;
    S.u[i][1] = normals[i].y;
    S.u[i][2] = normals[i].z;
    const Vec3f r{corners[i].x - com.x, corners[i].y - com.y, corners[i].z - com.z};
    b[i] = static_cast<double>(dot3(r, normals[i]));
  }
  double tau_v[2] = {0.0, 0.0};
  bidiagonalise<R>(S.u, S.d, tau_v, n);
  unpack<R>(S.u, S.v, S.d, tau_v, n);
  diagonalise(S, S.d, tau_v, 1000, stalled);
  sort_singular(S);
  if (edge) {  // qef.rs:51-63: drop the smallest singular value
    double smin = 1.7976931348623157e308;
    int id = 0;
    for (int i = 0; i < 3; ++i) {
      if (S.d[i] < smin) {
        smin = S.d[i];
        id = i;
      }
    }
    S.d[id] = 0.0;
  }
  // SVD::solve (svd.rs:55-84) + solveSVD (svd.rs:928-972)
  if (S.d[2] < 0.1) S.d[2] = 0.0;
  if (S.d[1] < 0.1) S.d[1] = 0.0;
  if (S.d[0] < 0.1) S.d[0] = 0.0;
  double w[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < n; ++i)
    if (b[i] != 0.0)
      for (int j = 0; j < 3; ++j) w[j] += b[i] * S.u[i][j];
  for (int i = 0; i < 3; ++i)
    if (S.d[i] != 0.0) w[i] /= S.d[i];
  float x[3];
  for (int i = 0; i < 3; ++i) {
    double t = 0.0;
    for (int j = 0; j < 3; ++j) t += w[j] * S.v[i][j];
    x[i] = static_cast<float>(t);
  }
  return {com.x + x[0], com.y + x[1], com.z + x[2]};
}

// The dual-contouring form: the 8 cell corners.
__device__ __forceinline__ Vec3f place_feature(const Vec3f corners[8], const Vec3f normals[8],
                                               bool* stalled) {
  bool feature;
  return place_feature_n<kRows>(corners, normals, kRows, stalled, &feature);
}

}  // namespace qef
}  // namespace iso
