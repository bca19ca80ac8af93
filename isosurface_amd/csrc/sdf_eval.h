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

}  // namespace iso
