// Device-side evaluation of lowered SDF programs (DevOp lists).
//
// Every operation follows the reference's f32 operation order exactly and
// the library is compiled with -ffp-contract=off, so no a*b+c is fused;
// sqrt and division are hipcc's correctly rounded f32 sequences.  The
// results are therefore bit-identical to the crate's sample_scalar:
//   Sphere    p.len() - r                               implicit/sphere.rs:35-39
//   Torus     len(|sqrt(x*x+y*y)| - R, z) - r           implicit/torus.rs:40-46
//   Box       |max(q,0)| + min(max_comp(q), 0)          implicit/rectangular_prism.rs:36-41
//   Cylinder  min(max(qx,qz),0) + |(max(qx,0),max(qz,0),0)|   implicit/cylinder.rs:41-48
//   len_sq = (x*x + y*y) + z*z                           math/vector.rs:45-52,125-127
#pragma once
#include <hip/hip_runtime.h>

#include "sdf_program.h"

namespace iso {

__device__ __forceinline__ float f_len3(float x, float y, float z) {
  return __builtin_sqrtf((x * x + y * y) + z * z);
}

__device__ __forceinline__ float prim_value(uint32_t code, const float* prm, float x, float y,
                                            float z) {
  if (code == kSphere) {
    return f_len3(x, y, z) - prm[0];
  } else if (code == kTorus) {
    const float qx = __builtin_fabsf(__builtin_sqrtf(x * x + y * y)) - prm[0];
    const float l = __builtin_sqrtf(qx * qx + z * z);
    return l - prm[1];
  } else if (code == kBox) {
    const float qx = __builtin_fabsf(x) - prm[0];
    const float qy = __builtin_fabsf(y) - prm[1];
    const float qz = __builtin_fabsf(z) - prm[2];
    const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
    const float maxc = fmaxf(qx, fmaxf(qy, qz));
    return f_len3(mx, my, mz) + fminf(maxc, 0.0f);
  } else {  // kCylinder
    const float qx = __builtin_fabsf(__builtin_sqrtf(x * x + y * y)) - prm[0];
    const float qz = __builtin_fabsf(z) - prm[1];
    const float dx = fmaxf(qx, 0.0f), dy = fmaxf(qz, 0.0f);
    return fminf(fmaxf(qx, qz), 0.0f) + f_len3(dx, dy, 0.0f);
  }
}

// Polynomial smooth minimum: h = max(k - |a-b|, 0) / k; min(a,b) - h*h*k*0.25.
__device__ __forceinline__ float smooth_min(float a, float b, float k) {
  const float h = fmaxf(k - __builtin_fabsf(a - b), 0.0f) / k;
  return fminf(a, b) - ((h * h) * k) * 0.25f;
}

// Evaluates the program at (px, py, pz).  `ops` must be the same for every
// lane of the wavefront (it is a kernel argument), so all control flow below
// is uniform and runs on the scalar unit.
__device__ __forceinline__ float sdf_eval(const DevOp* __restrict__ ops, int n, float px,
                                          float py, float pz) {
  float top = 0.0f;
  float st[kMaxValueStack];
  int sp = 0;  // values below `top`
  float cx = px, cy = py, cz = pz;
  float sx[kMaxPointStack], sy[kMaxPointStack], sz[kMaxPointStack];
  int psp = 0;
  bool have_top = false;
  for (int i = 0; i < n; ++i) {
    const uint32_t code = ops[i].code;
    if (code <= kCylinder) {
      float qx = cx, qy = cy, qz = cz;
      if (ops[i].translated) {
        qx = cx - ops[i].off[0];
        qy = cy - ops[i].off[1];
        qz = cz - ops[i].off[2];
      }
      const float v = prim_value(code, ops[i].prm, qx, qy, qz);
      const uint32_t mode = ops[i].combine;
      if (mode == kPush) {
        if (have_top) st[sp++] = top;
        top = v;
        have_top = true;
      } else if (mode == kMin) {
        top = fminf(top, v);
      } else if (mode == kMax) {
        top = fmaxf(top, v);
      } else if (mode == kDiff) {
        top = fmaxf(v, -top);
      } else if (mode == kDiffRev) {
        top = fmaxf(top, -v);
      } else {
        top = smooth_min(top, v, ops[i].k);
      }
    } else if (code == kPushTranslate) {
      sx[psp] = cx;
      sy[psp] = cy;
      sz[psp] = cz;
      ++psp;
      cx = cx - ops[i].off[0];
      cy = cy - ops[i].off[1];
      cz = cz - ops[i].off[2];
    } else if (code == kPopTranslate) {
      --psp;
      cx = sx[psp];
      cy = sy[psp];
      cz = sz[psp];
    } else {
      // Generic binary operator on (below, top).
      const float t = top;
      const float u = st[--sp];
      if (code == kUnion) {
        top = fminf(u, t);
      } else if (code == kIntersection) {
        top = fmaxf(u, t);
      } else if (code == kDifference) {  // a = u, b = t
        top = fmaxf(t, -u);
      } else if (code == kDifferenceRev) {  // b = u, a = t
        top = fmaxf(u, -t);
      } else {
        top = smooth_min(u, t, ops[i].k);
      }
    }
  }
  return top;
}

// ---------------------------------------------------------------------------
// Directed distances: VectorSource::sample_vector (source.rs:32-38), the
// signed distance to the surface along each cardinal axis, or f32::MAX where
// the axis line through the point misses the shape.  Each restates the
// reference's f32 operation order; a = |p| componentwise.
// ---------------------------------------------------------------------------
constexpr float kDirectedMiss = 3.40282347e+38f;  // f32::MAX

__device__ __forceinline__ void prim_vector(uint32_t code, const float* prm, float x, float y,
                                            float z, float* out) {
  const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y), az = __builtin_fabsf(z);
  if (code == kSphere) {  // implicit/sphere.rs:41-55
    const float r2 = prm[0] * prm[0];
    const float l_yz = r2 - (ay * ay + az * az);
    const float l_xz = r2 - (ax * ax + az * az);
    const float l_xy = r2 - (ax * ax + ay * ay);
    out[0] = l_yz < 0.0f ? kDirectedMiss : ax - __builtin_sqrtf(l_yz);
    out[1] = l_xz < 0.0f ? kDirectedMiss : ay - __builtin_sqrtf(l_xz);
    out[2] = l_xy < 0.0f ? kDirectedMiss : az - __builtin_sqrtf(l_xy);
  } else if (code == kTorus) {  // implicit/torus.rs:48-91
    const float R = prm[0], r = prm[1];
    const float l = __builtin_sqrtf(ax * ax + ay * ay);
    const float l_xy = l - R;
    const float tz = __builtin_sqrtf(r * r - az * az);
    const float rx = R + tz, ry = R - tz;
    if (az > r || ay > R + tz)
      out[0] = kDirectedMiss;
    else if (ax == 0.0f)
      out[0] = __builtin_fabsf(ay - R) - r;
    else
      out[0] = fmaxf(ax - __builtin_sqrtf(rx * rx - ay * ay), __builtin_sqrtf(ry * ry - ay * ay) - ax);
    if (az > r || ax > R + tz)
      out[1] = kDirectedMiss;
    else if (ay == 0.0f)
      out[1] = __builtin_fabsf(ax - R) - r;
    else
      out[1] = fmaxf(ay - __builtin_sqrtf(rx * rx - ax * ax), __builtin_sqrtf(ry * ry - ax * ax) - ay);
    out[2] = __builtin_fabsf(l_xy) > r ? kDirectedMiss : az - __builtin_sqrtf(r * r - l_xy * l_xy);
  } else if (code == kBox) {  // implicit/rectangular_prism.rs:43-73
    const float hx = prm[0], hy = prm[1], hz = prm[2];
    const float mx = ((ay - hy) > 0.0f || (az - hz) > 0.0f) ? kDirectedMiss : -kDirectedMiss;
    const float my = ((ax - hx) > 0.0f || (az - hz) > 0.0f) ? kDirectedMiss : -kDirectedMiss;
    const float mz = ((ax - hx) > 0.0f || (ay - hy) > 0.0f) ? kDirectedMiss : -kDirectedMiss;
    const bool in = ax < hx && ay < hy && az < hz;
    const float cx = in ? fmaxf(ax, hx) : fminf(ax, hx);
    const float cy = in ? fmaxf(ay, hy) : fminf(ay, hy);
    const float cz = in ? fmaxf(az, hz) : fminf(az, hz);
    out[0] = fmaxf(ax - cx, mx);
    out[1] = fmaxf(ay - cy, my);
    out[2] = fmaxf(az - cz, mz);
  } else {  // kCylinder, implicit/cylinder.rs:50-71
    const float r = prm[0], h = prm[1];
    out[0] = (az > h || ay > r) ? kDirectedMiss : ax - __builtin_sqrtf(r * r - ay * ay);
    out[1] = (az > h || ax > r) ? kDirectedMiss : ay - __builtin_sqrtf(r * r - ax * ax);
    out[2] = (ax * ax + ay * ay) > r * r ? kDirectedMiss : az - h;
  }
}

// sdf_eval's stack machine over Directed values: the CSG operators act
// componentwise (csg.rs:41-45, 74-78, 102-106: min, max, and b.max(-a)).
// Programs with a smooth union have no directed form; the host rejects them
// before launch.
__device__ __forceinline__ void sdf_eval_vector(const DevOp* __restrict__ ops, int n, float px,
                                                float py, float pz, float* out) {
  float top[3] = {0.0f, 0.0f, 0.0f};
  float st[kMaxValueStack][3];
  int sp = 0;
  float cx = px, cy = py, cz = pz;
  float sx[kMaxPointStack], sy[kMaxPointStack], sz[kMaxPointStack];
  int psp = 0;
  bool have_top = false;
  for (int i = 0; i < n; ++i) {
    const uint32_t code = ops[i].code;
    if (code <= kCylinder) {
      float qx = cx, qy = cy, qz = cz;
      if (ops[i].translated) {
        qx = cx - ops[i].off[0];
        qy = cy - ops[i].off[1];
        qz = cz - ops[i].off[2];
      }
      float v[3];
      prim_vector(code, ops[i].prm, qx, qy, qz, v);
      const uint32_t mode = ops[i].combine;
      if (mode == kPush) {
        if (have_top) {
          st[sp][0] = top[0];
          st[sp][1] = top[1];
          st[sp][2] = top[2];
          ++sp;
        }
        top[0] = v[0];
        top[1] = v[1];
        top[2] = v[2];
        have_top = true;
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (mode == kMin)
            top[k] = fminf(top[k], v[k]);
          else if (mode == kMax)
            top[k] = fmaxf(top[k], v[k]);
          else if (mode == kDiff)
            top[k] = fmaxf(v[k], -top[k]);
          else  // kDiffRev
            top[k] = fmaxf(top[k], -v[k]);
        }
      }
    } else if (code == kPushTranslate) {
      sx[psp] = cx;
      sy[psp] = cy;
      sz[psp] = cz;
      ++psp;
      cx = cx - ops[i].off[0];
      cy = cy - ops[i].off[1];
      cz = cz - ops[i].off[2];
    } else if (code == kPopTranslate) {
      --psp;
      cx = sx[psp];
      cy = sy[psp];
      cz = sz[psp];
    } else {
      --sp;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float t = top[k];
        const float u = st[sp][k];
        if (code == kUnion)
          top[k] = fminf(u, t);
        else if (code == kIntersection)
          top[k] = fmaxf(u, t);
        else if (code == kDifference)  // a = u, b = t
          top[k] = fmaxf(t, -u);
        else  // kDifferenceRev: b = u, a = t
          top[k] = fmaxf(u, -t);
      }
    }
  }
  out[0] = top[0];
  out[1] = top[1];
  out[2] = top[2];
}

// ---------------------------------------------------------------------------
// HermiteSource::sample_normal (source.rs:41-48): the shapes' analytic
// normals, not normalised, for programs of shapes, unions and translations
// (Intersection and Difference have no HermiteSource; the host rejects
// them).  Union takes the componentwise minimum of its operands' normals
// (csg.rs:47-51).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void prim_normal(uint32_t code, const float* prm, float x, float y,
                                            float z, float* out) {
  if (code == kSphere) {  // sphere.rs:58-62
    out[0] = x;
    out[1] = y;
    out[2] = z;
  } else if (code == kTorus) {  // torus.rs:98-107
    // (x, y, 0).normalised().unwrap_or((1, 0, 0)) * radius
    const float l = __builtin_sqrtf((x * x + y * y) + 0.0f * 0.0f);
    float rx = 1.0f, ry = 0.0f, rz = 0.0f;
    if (!(__builtin_fabsf(l) < 1.1920929e-07f)) {
      rx = x / l;
      ry = y / l;
      rz = 0.0f / l;
    }
    out[0] = x - rx * prm[0];
    out[1] = y - ry * prm[0];
    out[2] = z - rz * prm[0];
  } else if (code == kBox) {  // rectangular_prism.rs:76-79, vector.rs:278-287
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y), az = __builtin_fabsf(z);
    out[0] = 0.0f;
    out[1] = 0.0f;
    out[2] = 0.0f;
    if (ax > ay && ax > az)
      out[0] = x;
    else if (ay > az)
      out[1] = y;
    else
      out[2] = z;
  } else {  // kCylinder, cylinder.rs:74-84
    const float zz = __builtin_fabsf(z) / prm[1];
    const float r = __builtin_sqrtf(x * x + y * y) / prm[0];
    if (zz > r) {
      out[0] = 0.0f;
      out[1] = 0.0f;
      out[2] = z;
    } else {
      out[0] = x;
      out[1] = y;
      out[2] = 0.0f;
    }
  }
}

__device__ __forceinline__ void sdf_eval_hermite(const DevOp* __restrict__ ops, int n, float px,
                                                 float py, float pz, float* out) {
  float top[3] = {0.0f, 0.0f, 0.0f};
  float st[kMaxValueStack][3];
  int sp = 0;
  float cx = px, cy = py, cz = pz;
  float sx[kMaxPointStack], sy[kMaxPointStack], sz[kMaxPointStack];
  int psp = 0;
  bool have_top = false;
  for (int i = 0; i < n; ++i) {
    const uint32_t code = ops[i].code;
    if (code <= kCylinder) {
      float qx = cx, qy = cy, qz = cz;
      if (ops[i].translated) {
        qx = cx - ops[i].off[0];
        qy = cy - ops[i].off[1];
        qz = cz - ops[i].off[2];
      }
      float v[3];
      prim_normal(code, ops[i].prm, qx, qy, qz, v);
      if (ops[i].combine == kPush) {
        if (have_top) {
          st[sp][0] = top[0];
          st[sp][1] = top[1];
          st[sp][2] = top[2];
          ++sp;
        }
        top[0] = v[0];
        top[1] = v[1];
        top[2] = v[2];
        have_top = true;
      } else {  // kMin: Union(top, v)
        top[0] = fminf(top[0], v[0]);
        top[1] = fminf(top[1], v[1]);
        top[2] = fminf(top[2], v[2]);
      }
    } else if (code == kPushTranslate) {
      sx[psp] = cx;
      sy[psp] = cy;
      sz[psp] = cz;
      ++psp;
      cx = cx - ops[i].off[0];
      cy = cy - ops[i].off[1];
      cz = cz - ops[i].off[2];
    } else if (code == kPopTranslate) {
      --psp;
      cx = sx[psp];
      cy = sy[psp];
      cz = sz[psp];
    } else {  // kUnion of (below, top)
      --sp;
      top[0] = fminf(st[sp][0], top[0]);
      top[1] = fminf(st[sp][1], top[1]);
      top[2] = fminf(st[sp][2], top[2]);
    }
  }
  out[0] = top[0];
  out[1] = top[1];
  out[2] = top[2];
}

// CentralDifference::sample_normal (source.rs:82-94): not normalised.
__device__ __forceinline__ void sdf_central_difference(const DevOp* ops, int n_ops, float px,
                                                       float py, float pz, float eps,
                                                       float* out) {
  const float vx = sdf_eval(ops, n_ops, px + eps, py + 0.0f, pz + 0.0f) -
                   sdf_eval(ops, n_ops, px - eps, py - 0.0f, pz - 0.0f);
  const float vy = sdf_eval(ops, n_ops, px + 0.0f, py + eps, pz + 0.0f) -
                   sdf_eval(ops, n_ops, px - 0.0f, py - eps, pz - 0.0f);
  const float vz = sdf_eval(ops, n_ops, px + 0.0f, py + 0.0f, pz + eps) -
                   sdf_eval(ops, n_ops, px - 0.0f, py - 0.0f, pz - eps);
  const float e2 = 2.0f * eps;
  out[0] = vx / e2;
  out[1] = vy / e2;
  out[2] = vz / e2;
}

}  // namespace iso
