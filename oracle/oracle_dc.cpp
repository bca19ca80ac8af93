// CPU oracle, dual contouring — TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates DualContouring::extract with MinimiseQEF and CentralDifference
// normals (or, with epsilon 0, the program's own HermiteSource normals,
// implicit/*.rs sample_normal and csg.rs:47-51): dual_contouring.rs:65-105, traversal/dual_grid.rs:52-95,
// marching_cubes_impl.rs:58-69, source.rs:82-94, feature/mod.rs:91-148,
// feature/qef.rs:36-68 and the SVD of math/svd.rs:25-84, 130-972 (f64).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../isosurface_amd/csrc/mc_tables.h"
#include "oracle.h"
#include "oracle_internal.h"

namespace {

using iso::tables::kCorner;
using iso::tables::kEdgeCorners;
using iso::tables::kTables;
using orc::F3;

// ---------------------------------------------------------------------------
// SVD of an n x 3 matrix (n <= 12), f64 (math/svd.rs).  Householder
// bidiagonalisation, implicit-shift QR with Givens rotations, sort.
// ---------------------------------------------------------------------------
struct Svd {
  int rows = 0;
  double u[12][3] = {};
  double v[3][3] = {};
  double d[3] = {};
};

// Householder vector of the entries x[0..n) (pointers into the matrix).
double householder(double* x[], int n) {  // svd.rs:251-293
  double tau = 0.0;
  if (n > 1) {
    double xnorm;
    if (n == 2) {
      xnorm = std::fabs(*x[1]);
    } else {
      double scl = 0.0, ssq = 1.0;
      for (int i = 1; i < n; ++i) {
        const double a = std::fabs(*x[i]);
        if (a != 0.0) {
          if (scl < a) {
            ssq = 1.0 + ssq * (scl / a) * (scl / a);
            scl = a;
          } else {
            ssq += a / scl * (a / scl);
          }
        }
      }
      xnorm = scl * std::sqrt(ssq);
    }
    if (xnorm != 0.0) {
      const double alpha = *x[0];
      double beta = std::sqrt(alpha * alpha + xnorm * xnorm);
      if (alpha >= 0.0) beta = -beta;
      tau = (beta - alpha) / beta;
      const double scl = 1.0 / (alpha - beta);
      *x[0] = beta;
      for (int i = 1; i < n; ++i) *x[i] *= scl;
    }
  }
  return tau;
}

void bidiagonalise(double m[12][3], double tau_u[3], double tau_v[2], int rows) {  // svd.rs:149-249
  for (int i = 0; i < 3; ++i) {
    double* col[12];
    int n = rows - i;
    for (int q = 0; q < n; ++q) col[q] = &m[q + i][i];
    double tau = householder(col, n);
    tau_u[i] = tau;
    if (i + 1 < 3) {
      if (tau != 0.0) {
        for (int x = i + 1; x < 3; ++x) {
          double wx = m[i][x];
          for (int y = i + 1; y < rows; ++y) wx += m[y][x] * *col[y - i];
          const double tw = tau * wx;
          m[i][x] -= tw;
          for (int y = i + 1; y < rows; ++y) m[y][x] -= tw * *col[y - i];
        }
      }
      double* row[2] = {&m[i][i + 1], i == 0 ? &m[i][i + 2] : nullptr};
      const int nr = i == 0 ? 2 : 1;
      tau = householder(row, nr);
      tau_v[i] = tau;
      if (tau != 0.0) {
        for (int y = i + 1; y < rows; ++y) {
          double wy = m[y][i + 1];
          if (i == 0) wy += m[y][i + 2] * *row[1];
          const double tw = tau * wy;
          m[y][i + 1] -= tw;
          if (i == 0) m[y][i + 2] -= tw * *row[1];
        }
      }
    }
  }
}

void unpack(double u[12][3], double v[3][3], double tau_u[3], double tau_v[2], int rows) {  // svd.rs:295-410
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

void givens(double a, double b, double* c, double* s) {  // svd.rs:825-842
  if (b == 0.0) {
    *c = 1.0;
    *s = 0.0;
  } else if (std::fabs(b) > std::fabs(a)) {
    const double t = -a / b;
    const double s1 = 1.0 / std::sqrt(1.0 + t * t);
    *s = s1;
    *c = s1 * t;
  } else {
    const double t = -b / a;
    const double c1 = 1.0 / std::sqrt(1.0 + t * t);
    *c = c1;
    *s = c1 * t;
  }
}

void schur(double a1, double a2, double a3, double* c, double* s) {  // svd.rs:844-865
  const double apq = a1 * a2 * 2.0;
  if (apq == 0.0) {
    *c = 1.0;
    *s = 0.0;
  } else {
    const double tau = (a2 * a2 + (a3 + a1) * (a3 - a1)) / apq;
    const double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau))
                                : -1.0 / (std::sqrt(1.0 + tau * tau) - tau);
    *c = 1.0 / std::sqrt(1.0 + t * t);
    *s = t * *c;
  }
}

void chop(double* a, double* b, int n) {  // svd.rs:481-493
  double ai = a[0];
  for (int i = 0; i < n - 1; ++i) {
    const double bi = b[i];
    const double ai1 = a[i + 1];
    if (std::fabs(bi) < 1e-5 * (std::fabs(ai) + std::fabs(ai1))) b[i] = 0.0;
    ai = ai1;
  }
}

// Rotates columns p and q of an n x 3 matrix by (c, s).
template <int N>
void rotate(double m[N][3], int rows, int p, int q, double c, double s) {
  for (int i = 0; i < rows; ++i) {
    const double mp = m[i][p], mq = m[i][q];
    m[i][p] = mp * c - mq * s;
    m[i][q] = mp * s + mq * c;
  }
}

void qr_cols2(double u[12][3], double v[3][3], double* tu, double* tv, int rows) {  // svd.rs:689-823
  double c, s;
  if (tu[0] == 0.0) {
    givens(tv[0], tu[1], &c, &s);
    tu[0] = tv[0] * c - tu[1] * s;
    tv[0] = tv[0] * s + tu[1] * c;
    tu[1] = 0.0;
    rotate<12>(u, rows, 0, 1, c, s);
    for (int i = 0; i < 3; ++i) {
      const double t = v[i][0];
      v[i][0] = v[i][1];
      v[i][1] = t;
    }
  } else if (tu[1] == 0.0) {
    givens(tu[0], tv[0], &c, &s);
    tu[0] = tu[0] * c - tv[0] * s;
    tv[0] = 0.0;
    rotate<3>(v, 3, 0, 1, c, s);
  } else {
    schur(tu[0], tv[0], tu[1], &c, &s);
    double a11 = tu[0] * c - tv[0] * s;
    double a21 = -tu[1] * s;
    double a12 = tu[0] * s + tv[0] * c;
    double a22 = tu[1] * c;
    rotate<3>(v, 3, 0, 1, c, s);
    if (a11 * a11 + a21 * a21 < a12 * a12 + a22 * a22) {
      double t = a11;
      a11 = a12;
      a12 = t;
      t = a21;
      a21 = a22;
      a22 = t;
      for (int i = 0; i < 3; ++i) {
        t = v[i][0];
        v[i][0] = v[i][1];
        v[i][1] = t;
      }
    }
    givens(a11, a21, &c, &s);
    tu[0] = a11 * c - a21 * s;
    tv[0] = a12 * c - a22 * s;
    tu[1] = a12 * s + a22 * c;
    rotate<12>(u, rows, 0, 1, c, s);
  }
}

void qr_middle(double u[12][3], double* tu, double* tv, int rows, int cols, int col) {  // svd.rs:604-641
  double x = tv[col], y = tu[col + 1];
  for (int j = col; j < cols - 1; ++j) {
    double c, s;
    givens(y, -x, &c, &s);
    for (int i = 0; i < rows; ++i) {
      const double up = u[i][col], uq = u[i][j + 1];
      u[i][col] = up * c - uq * s;
      u[i][j + 1] = up * s + uq * c;
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

void qr_end(double v[3][3], double* tu, double* tv) {  // svd.rs:643-671
  double x = tu[1], y = tv[1];
  for (int k = 1; k >= 0; --k) {
    double c, s;
    givens(x, y, &c, &s);
    for (int i = 0; i < 3; ++i) {
      const double vp = v[i][k], vq = v[i][2];
      v[i][k] = vp * c - vq * s;
      v[i][2] = vp * s + vq * c;
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

double qr_shift(const double* tu, const double* tv) {  // svd.rs:673-687
  const double ta = tu[1] * tu[1] + tv[0] * tv[0];
  const double tb = tu[2] * tu[2] + tv[1] * tv[1];
  const double tab = tu[1] * tv[1];
  const double dt = (ta - tb) / 2.0;
  if (dt >= 0.0) return tb - tab * tab / (dt + std::sqrt(dt * dt + tab * tab));
  return tb + tab * tab / (std::sqrt(dt * dt + tab * tab) - dt);
}

void qr_step(double u[12][3], double v[3][3], double* tu, double* tv, int rows, int cols) {  // svd.rs:495-602
  if (cols == 2) {
    qr_cols2(u, v, tu, tv, rows);
    return;
  }
  for (int i = 0; i < cols - 1; ++i) {
    if (tu[i] == 0.0) {
      qr_middle(u, tu, tv, rows, cols, i);
      return;
    }
  }
  if (tu[cols - 1] == 0.0) {
    qr_end(v, tu, tv);
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
    rotate<3>(v, 3, k, k + 1, c, s);
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
    rotate<12>(u, rows, k, k + 1, c, s);
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

void diagonalise(double u[12][3], double v[3][3], double tu[3], double tv[2], int rows) {  // svd.rs:413-479
  chop(tu, tv, 3);
  int b = 2;
  while (b > 0) {
    if (tv[b - 1] == 0.0) {
      --b;
      continue;
    }
    int a = b - 1;
    while (a > 0 && tv[a - 1] != 0.0) --a;
    const int n = b - a + 1;
    double u1[12][3] = {}, v1[3][3] = {};
    for (int j = a; j <= b; ++j) {
      for (int i = 0; i < rows; ++i) u1[i][j - a] = u[i][j];
      for (int i = 0; i < 3; ++i) v1[i][j - a] = v[i][j];
    }
    qr_step(u1, v1, &tu[a], &tv[a], rows, n);
    for (int j = a; j <= b; ++j) {
      for (int i = 0; i < rows; ++i) u[i][j] = u1[i][j - a];
      for (int i = 0; i < 3; ++i) v[i][j] = v1[i][j - a];
    }
    chop(&tu[a], &tv[a], n);
  }
}

void sort_singular(double u[12][3], double v[3][3], double d[3], int rows) {  // svd.rs:867-926
  for (int j = 0; j < 3; ++j) {
    if (d[j] < 0.0) {
      for (int i = 0; i < 3; ++i) v[i][j] = -v[i][j];
      d[j] = -d[j];
    }
  }
  for (int i = 0; i < 3; ++i) {
    double dmax = d[i];
    int imax = i;
    for (int j = i + 1; j < 3; ++j) {
      if (d[j] > dmax) {
        dmax = d[j];
        imax = j;
      }
    }
    if (imax != i) {
      double t = d[i];
      d[i] = d[imax];
      d[imax] = t;
      for (int y = 0; y < rows; ++y) {
        t = u[y][i];
        u[y][i] = u[y][imax];
        u[y][imax] = t;
      }
      for (int y = 0; y < 3; ++y) {
        t = v[y][i];
        v[y][i] = v[y][imax];
        v[y][imax] = t;
      }
    }
  }
}

Svd svd(const double (*a)[3], int rows) {  // svd.rs:25-52, 130-147
  Svd r;
  r.rows = rows;
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < 3; ++j) r.u[i][j] = a[i][j];
  double tau_v[2] = {0.0, 0.0};
  bidiagonalise(r.u, r.d, tau_v, rows);
  unpack(r.u, r.v, r.d, tau_v, rows);
  diagonalise(r.u, r.v, r.d, tau_v, rows);
  sort_singular(r.u, r.v, r.d, rows);
  return r;
}

// SVD::solve (svd.rs:55-84) + solveSVD (svd.rs:928-972).
void svd_solve(Svd& s, const double* b, double x[3]) {
  if (s.d[2] < 0.1) s.d[2] = 0.0;
  if (s.d[1] < 0.1) s.d[1] = 0.0;
  if (s.d[0] < 0.1) s.d[0] = 0.0;
  double w[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < s.rows; ++i)
    if (b[i] != 0.0)
      for (int j = 0; j < 3; ++j) w[j] += b[i] * s.u[i][j];
  for (int i = 0; i < 3; ++i)
    if (s.d[i] != 0.0) w[i] /= s.d[i];
  for (int i = 0; i < 3; ++i) {
    double t = 0.0;
    for (int j = 0; j < 3; ++j) t += w[j] * s.v[i][j];
    x[i] = t;
  }
}

// ---------------------------------------------------------------------------
// Feature placement (feature/mod.rs, feature/qef.rs)
// ---------------------------------------------------------------------------
const float kFeatureAngle = static_cast<float>(0.8660254037844387);  // mod.rs:22, cos 30 deg

float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }  // vector.rs:240-242,290-292
F3 cross(F3 a, F3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Vec3::normalised().unwrap_or_default() (vector.rs:135-144).
F3 normalised_or_zero(F3 v) {
  const float l = std::sqrt((v.x * v.x + v.y * v.y) + v.z * v.z);
  if (std::fabs(l) < 1.1920929e-07f) return {0.0f, 0.0f, 0.0f};
  return {v.x / l, v.y / l, v.z / l};
}

// TangentPlanes::new (feature/mod.rs:91-148) over n points and normals +
// MinimiseQEF::place_feature_with_tangents (qef.rs:36-68); `feature`, when
// not null, receives whether the planes were classified Edge or Corner.
F3 place_feature_n(const F3* corners, const F3* normals, int n, bool* feature) {
  F3 com{0.0f, 0.0f, 0.0f}, axis{0.0f, 0.0f, 0.0f};
  float min_angle = 3.40282347e+38f;
  float count = 0.0f;
  if (feature) *feature = false;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      const float angle = dot(normals[i], normals[j]);
      if (angle < min_angle) {
        axis = cross(normals[i], normals[j]);
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
  if (min_angle > kFeatureAngle) return com;  // Planar
  if (feature) *feature = true;
  axis = normalised_or_zero(axis);
  float min_c = 1.0f, max_c = -1.0f;
  for (int i = 0; i < n; ++i) {
    const float c = dot(axis, normals[i]);
    min_c = std::fmin(min_c, c);
    max_c = std::fmax(max_c, c);
  }
  float c = std::fmax(std::fabs(min_c), std::fabs(max_c));
  c = std::sqrt(1.0f - c * c);
  const bool edge = c > kFeatureAngle;
  double a[12][3], b[12];
  for (int i = 0; i < n; ++i) {
    a[i][0] = normals[i].x;
    a[i][1] = normals[i].y;
    a[i][2] = normals[i].z;
    const F3 r{corners[i].x - com.x, corners[i].y - com.y, corners[i].z - com.z};
    b[i] = static_cast<double>(dot(r, normals[i]));
  }
  Svd s = svd(a, n);
  if (edge) {  // drop the smallest singular value (qef.rs:51-63)
    double smin = 1.7976931348623157e308;
    int id = 0;
    for (int i = 0; i < 3; ++i) {
      if (s.d[i] < smin) {
        smin = s.d[i];
        id = i;
      }
    }
    s.d[id] = 0.0;
  }
  double x[3];
  svd_solve(s, b, x);
  return {com.x + static_cast<float>(x[0]), com.y + static_cast<float>(x[1]),
          com.z + static_cast<float>(x[2])};
}

F3 place_feature(const F3 corners[8], const F3 normals[8]) {
  return place_feature_n(corners, normals, 8, nullptr);
}

}  // namespace

namespace orc {
F3 place_feature_planes(const F3* points, const F3* normals, int n, bool* feature) {
  return place_feature_n(points, normals, n, feature);
}
F3 unit_or_zero(F3 v) { return normalised_or_zero(v); }
}  // namespace orc

// ---------------------------------------------------------------------------
// DualContouring::extract (dual_contouring.rs:65-105)
// ---------------------------------------------------------------------------
extern "C" int or_extract_dc(const iso_node* prog, uint32_t n, const iso_grid_desc* desc,
                             float epsilon, or_mesh* out) {
  std::memset(out, 0, sizeof(*out));
  const iso_grid_desc& d = *desc;
  if (d.nx < 2 || d.ny < 2 || d.nz < 2) return ISO_ERR_INVALID_ARGUMENT;
  orc::Program p;
  if (!orc::load_program(prog, n, &p)) return ISO_ERR_PROGRAM;
  if (!(epsilon >= 0.0f)) return ISO_ERR_INVALID_ARGUMENT;
  if (epsilon == 0.0f && !orc::program_has_hermite(p)) return ISO_ERR_PROGRAM;
  const int64_t nx = d.nx, ny = d.ny, nz = d.nz;
  const int64_t dnx = nx - 1;  // dual row width (size_minus_one, dual_grid.rs:62)

  struct Sample {
    F3 p;
    float v;
  };
  std::vector<Sample> layers[2] = {std::vector<Sample>(nx * ny), std::vector<Sample>(nx * ny)};
  std::vector<Sample> duals[2] = {std::vector<Sample>(dnx * (ny - 1)),
                                  std::vector<Sample>(dnx * (ny - 1))};
  auto sample = [&](int64_t x, int64_t y, int64_t z) -> Sample {
    F3 q{orc::coord(d, 0, x), orc::coord(d, 1, y), orc::coord(d, 2, z)};
    return {q, orc::sample(p, q)};
  };
  // CentralDifference::sample_normal (source.rs:82-94), or with epsilon 0
  // the program's HermiteSource::sample_normal, normalised or zero
  // (marching_cubes_impl.rs:58-69).
  auto normal = [&](F3 q) -> F3 {
    if (epsilon == 0.0f) return normalised_or_zero(orc::program_hermite_normal(p, q));
    const float vx = orc::sample(p, F3{q.x + epsilon, q.y + 0.0f, q.z + 0.0f}) -
                     orc::sample(p, F3{q.x - epsilon, q.y - 0.0f, q.z - 0.0f});
    const float vy = orc::sample(p, F3{q.x + 0.0f, q.y + epsilon, q.z + 0.0f}) -
                     orc::sample(p, F3{q.x - 0.0f, q.y - epsilon, q.z - 0.0f});
    const float vz = orc::sample(p, F3{q.x + 0.0f, q.y + 0.0f, q.z + epsilon}) -
                     orc::sample(p, F3{q.x - 0.0f, q.y - 0.0f, q.z - epsilon});
    const float e2 = 2.0f * epsilon;
    return normalised_or_zero(F3{vx / e2, vy / e2, vz / e2});
  };

  orc::MeshBuilder mb;
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x) layers[0][y * nx + x] = sample(x, y, 0);

  F3 corners[8], normals[8];
  float values[8];
  orc::Pt keys[8];
  for (int64_t z = 0; z + 1 < nz; ++z) {
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) layers[1][y * nx + x] = sample(x, y, z + 1);
    for (int64_t y = 0; y + 1 < ny; ++y) {
      for (int64_t x = 0; x + 1 < nx; ++x) {
        for (int i = 0; i < 8; ++i) {
          const Sample& s = layers[kCorner[i][2]][(y + kCorner[i][1]) * nx + x + kCorner[i][0]];
          corners[i] = s.p;
          values[i] = s.v;
        }
        // dual vertex (dual_grid.rs:73-81; dual_contouring.rs:78-87)
        const int cube = orc::classify(values);
        F3 vert;
        if (cube != 0 && cube != 255) {
          for (int i = 0; i < 8; ++i) normals[i] = normal(corners[i]);
          vert = place_feature(corners, normals);
        } else {
          vert = F3{0.5f * corners[0].x + 0.5f * corners[6].x, 0.5f * corners[0].y + 0.5f * corners[6].y,
                    0.5f * corners[0].z + 0.5f * corners[6].z};
        }
        const float dval = (1.0f - 0.5f) * values[0] + 0.5f * values[6];
        duals[z % 2][y * dnx + x] = {vert, dval};

        if (x > 0 && y > 0 && z > 0) {  // dual cube (x-1, y-1, z-1), dual_grid.rs:83-93
          F3 dc[8];
          float dv[8];
          for (int i = 0; i < 8; ++i) {
            const int64_t kx = x - 1 + kCorner[i][0], ky = y - 1 + kCorner[i][1],
                          kz = z - 1 + kCorner[i][2];
            keys[i] = {kx, ky, kz};
            const Sample& s = duals[kz % 2][ky * dnx + kx];
            dc[i] = s.p;
            dv[i] = s.v;
          }
          orc::march(mb, keys, dc, dv);
        }
      }
    }
    std::swap(layers[0], layers[1]);
  }
  orc::finish(mb, out);
  return ISO_OK;
}

// Test hook: the SVD solve of an n x 3 system with the reference's
// thresholds (and optionally the Edge rank drop), for direct checks.
extern "C" int or_svd_solve(const double* a, const double* b, int rows, int drop_min,
                            double* x_out, double* d_out) {
  if (rows < 1 || rows > 12) return ISO_ERR_INVALID_ARGUMENT;
  Svd s = svd(reinterpret_cast<const double(*)[3]>(a), rows);
  for (int i = 0; i < 3; ++i) d_out[i] = s.d[i];
  if (drop_min) {
    double smin = 1.7976931348623157e308;
    int id = 0;
    for (int i = 0; i < 3; ++i)
      if (s.d[i] < smin) {
        smin = s.d[i];
        id = i;
      }
    s.d[id] = 0.0;
  }
  svd_solve(s, b, x_out);
  return ISO_OK;
}

// Test hook: the dual vertex a mixed cell gets (corners + normals given).
extern "C" void or_place_feature(const float* corners, const float* normals, float* out) {
  F3 c[8], nn[8];
  for (int i = 0; i < 8; ++i) {
    c[i] = {corners[3 * i], corners[3 * i + 1], corners[3 * i + 2]};
    nn[i] = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
  }
  const F3 r = place_feature(c, nn);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
