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
// sdf_eval for a whole wavefront of nearby points, skipping the union and
// smooth-union operands that cannot change any lane's value.  Returns
// sdf_eval's value bit for bit; every lane of the wavefront must call it
// (all 64 active) with finite coordinates.
//
// A run of primitives that join the value by min (kMin) or smooth min
// (kSmooth, k > 0) is taken 64 operands at a time.  Lane j bounds operand j
// over the box of the wavefront's current points (the f32 points after the
// enclosing translations, as the primitives see them):
// lb_j <= its f32 value <= ub_j at every point.  A sphere's and a box's
// values are nondecreasing in each |q_c|, so the least and greatest |q_c|
// over the box give the bounds (other primitives are not bounded and are
// always evaluated).  Before operand j the running value is at most
// T_j = min(max over lanes of the value at the group's start, ub_i for
// i < j), because min and smooth min never exceed either operand.  Operand j
// is skipped when
//   kMin:    lb_j > T_j           (fminf(top, v) = top exactly), or
//   kSmooth: lb_j - T_j >= k      (with room for the rounding of |top - v|:
//            the blend weight h is then exactly 0 and the result is top).
// The bounds carry 128 ulps of every magnitude involved, far above the few
// roundings of a sphere's f32 evaluation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_reduce_min(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_reduce_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
// A wave-uniform value moved to a scalar register.
__device__ __forceinline__ float wave_uniform(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ float sdf_eval_culled(const DevOp* __restrict__ ops, int n, float px,
                                                 float py, float pz) {
  constexpr float u = 5.9604645e-08f;  // 2^-24
  const int lane = __lane_id();
  // Box of the lanes' current (translated) points, found at the first run
  // of a translation scope.
  float blo[3], bhi[3], bmag = 0.0f;
  bool have_box = false;
  float top = 0.0f;
  float st[kMaxValueStack];
  int sp = 0;
  float cx = px, cy = py, cz = pz;
  float sx[kMaxPointStack], sy[kMaxPointStack], sz[kMaxPointStack];
  int psp = 0;
  bool have_top = false;
  // One primitive operand joined to the value stack.
  auto primitive = [&](const DevOp& op) {
    float qx = cx, qy = cy, qz = cz;
    if (op.translated) {
      qx = cx - op.off[0];
      qy = cy - op.off[1];
      qz = cz - op.off[2];
    }
    const float v = prim_value(op.code, op.prm, qx, qy, qz);
    const uint32_t mode = op.combine;
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
      top = smooth_min(top, v, op.k);
    }
  };
  auto cullable = [](const DevOp& op) {
    return op.code <= kCylinder &&
           (op.combine == kMin || (op.combine == kSmooth && op.k > 0.0f && op.k < 3.0e38f));
  };
  int i = 0;
  while (i < n) {
    if (have_top && cullable(ops[i])) {
      if (!have_box) {
        blo[0] = wave_uniform(wave_reduce_min(cx));
        blo[1] = wave_uniform(wave_reduce_min(cy));
        blo[2] = wave_uniform(wave_reduce_min(cz));
        bhi[0] = wave_uniform(wave_reduce_max(cx));
        bhi[1] = wave_uniform(wave_reduce_max(cy));
        bhi[2] = wave_uniform(wave_reduce_max(cz));
        bmag = fmaxf(fmaxf(fmaxf(__builtin_fabsf(blo[0]), __builtin_fabsf(bhi[0])),
                           fmaxf(__builtin_fabsf(blo[1]), __builtin_fabsf(bhi[1]))),
                     fmaxf(__builtin_fabsf(blo[2]), __builtin_fabsf(bhi[2])));
        have_box = true;
      }
      // Lane j looks at operand i + j.
      const int j = i + lane;
      bool in_run = j < n && cullable(ops[j]);
      const uint64_t run = __ballot(in_run);
      // operands i .. i+m-1 (m >= 1: lane 0's operand is ops[i])
      const int m = ~run ? static_cast<int>(__builtin_ctzll(~run)) : 64;
      in_run = lane < m;
      float lb = -3.0e38f, ub = 3.0e38f, k = 0.0f;
      bool smooth = false;
      if (in_run && (ops[j].code == kSphere || ops[j].code == kBox)) {
        const DevOp& op = ops[j];
        const float o[3] = {op.translated ? op.off[0] : 0.0f, op.translated ? op.off[1] : 0.0f,
                            op.translated ? op.off[2] : 0.0f};
        // Per axis, the least and greatest |q| over the box.
        float an[3], af[3], omag = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          an[a] = fmaxf(fmaxf(blo[a] - o[a], o[a] - bhi[a]), 0.0f);
          af[a] = fmaxf(__builtin_fabsf(blo[a] - o[a]), __builtin_fabsf(bhi[a] - o[a]));
          omag = fmaxf(omag, __builtin_fabsf(o[a]));
        }
        const float far = __builtin_sqrtf((af[0] * af[0] + af[1] * af[1]) + af[2] * af[2]);
        const float pm = __builtin_fabsf(op.prm[0]) + __builtin_fabsf(op.prm[1]) +
                         __builtin_fabsf(op.prm[2]);
        const float margin = 128.0f * u * (far + pm + bmag + omag) + 1.0e-30f;
        // Both values are nondecreasing in each |q_c|, so the box's least
        // and greatest |q| per axis bound them.
        lb = __builtin_sqrtf((an[0] * an[0] + an[1] * an[1]) + an[2] * an[2]) - op.prm[0];
        ub = far - op.prm[0];
        // Boxes only when the group holds one: a wave-uniform branch, so
        // sphere-only programs do not pay for the box bounds.
        const bool is_box = op.code == kBox;
        if (__ballot(is_box)) {
          if (is_box) {
            lb = prim_value(kBox, op.prm, an[0], an[1], an[2]);
            ub = prim_value(kBox, op.prm, af[0], af[1], af[2]);
          }
        }
        lb -= margin;
        ub += margin;
        smooth = op.combine == kSmooth;
        k = op.k;
      }
      // T_j: the running value's bound before operand j.
      float pre = ub;  // inclusive prefix minimum of ub over the run's lanes
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const float t = __shfl_up(pre, o, 64);
        if (lane >= o) pre = fminf(pre, t);
      }
      float excl = __shfl_up(pre, 1, 64);
      if (lane == 0) excl = 3.0e38f;
      const float T = fminf(wave_uniform(wave_reduce_max(top)), excl);
      const bool skip =
          in_run && (smooth ? (lb - T) * (1.0f - 4.0f * u) >= k * (1.0f + 4.0f * u) : lb > T);
      uint64_t keep = __ballot(in_run && !skip);
      while (keep) {
        const int b = static_cast<int>(__builtin_ctzll(keep));
        keep &= keep - 1;
        primitive(ops[i + b]);
      }
      i += m;
      continue;
    }
    const uint32_t code = ops[i].code;
    if (code <= kCylinder) {
      primitive(ops[i]);
    } else if (code == kPushTranslate) {
      sx[psp] = cx;
      sy[psp] = cy;
      sz[psp] = cz;
      ++psp;
      cx = cx - ops[i].off[0];
      cy = cy - ops[i].off[1];
      cz = cz - ops[i].off[2];
      have_box = false;
    } else if (code == kPopTranslate) {
      --psp;
      cx = sx[psp];
      cy = sy[psp];
      cz = sz[psp];
      have_box = false;
    } else {
      const float t = top;
      const float w = st[--sp];
      if (code == kUnion) {
        top = fminf(w, t);
      } else if (code == kIntersection) {
        top = fmaxf(w, t);
      } else if (code == kDifference) {
        top = fmaxf(t, -w);
      } else if (code == kDifferenceRev) {
        top = fmaxf(w, -t);
      } else {
        top = smooth_min(w, t, ops[i].k);
      }
    }
    ++i;
  }
  return top;
}

// ---------------------------------------------------------------------------
// Evaluation with an error bound, for proofs.  Returns sdf_eval's value at
// (px, py, pz) and sets *err to a bound E such that, for every point x within
// Euclidean distance `widen` of (px, py, pz), the f32 evaluation at x differs
// from the program's exact (real-arithmetic) value at x by at most E.  The
// bound follows the evaluation: with u = 2^-24, a subtraction or sum rounds
// by at most u times its magnitude, a primitive by at most 16u times its
// operands' magnitude (squares, sums, square roots and the final subtraction
// together, with room to spare), min, max and negation are exact, and the
// smooth minimum (whose exact form is 1-Lipschitz in the max-norm of its
// operands) adds its own roundings wherever the blend can be active in the
// ball.  Every magnitude is taken over the ball (the value at the centre plus
// `widen` plus its error), so E holds at every point of it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bound_max3(float a, float b, float c) {
  return fmaxf(a, fmaxf(b, c));
}

__device__ __forceinline__ float sdf_eval_bound(const DevOp* __restrict__ ops, int n, float px,
                                                float py, float pz, float widen, float* err) {
  constexpr float u = 5.9604645e-08f;  // 2^-24
  float top = 0.0f, etop = 0.0f;
  float st[kMaxValueStack], est[kMaxValueStack];
  int sp = 0;
  float cx = px, cy = py, cz = pz;
  float pe = 0.0f;  // bound on the rounding of the current (translated) point, per coordinate
  float sx[kMaxPointStack], sy[kMaxPointStack], sz[kMaxPointStack], spe[kMaxPointStack];
  int psp = 0;
  bool have_top = false;
  auto translate_err = [&](float qx, float qy, float qz, float e) {
    return e + u * (bound_max3(__builtin_fabsf(qx), __builtin_fabsf(qy), __builtin_fabsf(qz)) +
                    widen + e);
  };
  auto blend = [&](float a, float ea, float b, float eb, float k) {
    // exact form's sensitivity: max(ea, eb); roundings only where the blend
    // can be active somewhere in the ball
    float e = fmaxf(ea, eb);
    if (__builtin_fabsf(a - b) < k + 2.0f * widen + 2.0f * (ea + eb) + u * (__builtin_fabsf(a) +
                                                                            __builtin_fabsf(b))) {
      const float mag = fminf(__builtin_fabsf(a), __builtin_fabsf(b)) + widen + e;
      e += 8.0f * u * k + 2.0f * u * (mag + k);
    }
    return e;
  };
  for (int i = 0; i < n; ++i) {
    const uint32_t code = ops[i].code;
    if (code <= kCylinder) {
      float qx = cx, qy = cy, qz = cz, qe = pe;
      if (ops[i].translated) {
        qx = cx - ops[i].off[0];
        qy = cy - ops[i].off[1];
        qz = cz - ops[i].off[2];
        qe = translate_err(qx, qy, qz, pe);
      }
      const float v = prim_value(code, ops[i].prm, qx, qy, qz);
      const float qmax =
          bound_max3(__builtin_fabsf(qx), __builtin_fabsf(qy), __builtin_fabsf(qz)) + widen + qe;
      const float pm = __builtin_fabsf(ops[i].prm[0]) + __builtin_fabsf(ops[i].prm[1]) +
                       __builtin_fabsf(ops[i].prm[2]) + __builtin_fabsf(ops[i].prm[3]);
      // point rounding (1-Lipschitz in the Euclidean norm) + the primitive's own
      const float ev = 1.7320509f * qe + 16.0f * u * (1.7320509f * qmax + pm);
      const uint32_t mode = ops[i].combine;
      if (mode == kPush) {
        if (have_top) {
          st[sp] = top;
          est[sp] = etop;
          ++sp;
        }
        top = v;
        etop = ev;
        have_top = true;
      } else if (mode == kMin) {
        top = fminf(top, v);
        etop = fmaxf(etop, ev);
      } else if (mode == kMax) {
        top = fmaxf(top, v);
        etop = fmaxf(etop, ev);
      } else if (mode == kDiff) {
        top = fmaxf(v, -top);
        etop = fmaxf(etop, ev);
      } else if (mode == kDiffRev) {
        top = fmaxf(top, -v);
        etop = fmaxf(etop, ev);
      } else {
        etop = blend(top, etop, v, ev, ops[i].k);
        top = smooth_min(top, v, ops[i].k);
      }
    } else if (code == kPushTranslate) {
      sx[psp] = cx;
      sy[psp] = cy;
      sz[psp] = cz;
      spe[psp] = pe;
      ++psp;
      cx = cx - ops[i].off[0];
      cy = cy - ops[i].off[1];
      cz = cz - ops[i].off[2];
      pe = translate_err(cx, cy, cz, pe);
    } else if (code == kPopTranslate) {
      --psp;
      cx = sx[psp];
      cy = sy[psp];
      cz = sz[psp];
      pe = spe[psp];
    } else {
      const float t = top, et = etop;
      --sp;
      const float w = st[sp], ew = est[sp];
      if (code == kUnion) {
        top = fminf(w, t);
        etop = fmaxf(ew, et);
      } else if (code == kIntersection) {
        top = fmaxf(w, t);
        etop = fmaxf(ew, et);
      } else if (code == kDifference) {
        top = fmaxf(t, -w);
        etop = fmaxf(ew, et);
      } else if (code == kDifferenceRev) {
        top = fmaxf(w, -t);
        etop = fmaxf(ew, et);
      } else {
        etop = blend(w, ew, t, et, ops[i].k);
        top = smooth_min(w, t, ops[i].k);
      }
    }
  }
  // the bound's own f32 arithmetic rounds too
  *err = etop * 1.001f;
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
