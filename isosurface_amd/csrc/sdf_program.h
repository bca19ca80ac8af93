// Lowered SDF programs: the form the GPU evaluates.
//
// iso_program_upload turns a post-order iso_node list (the C ABI's mirror of
// the crate's source types, implicit/*.rs) into a flat list of DevOps that a
// wavefront walks in lock-step: every lane runs the same op at the same time,
// so dispatch is scalar-unit work and only the arithmetic is per lane.
//
// Lowering is exact — it never changes an f32 result bit that reaches the
// mesh:
//   * TRANSLATE around a single primitive folds into the primitive's offset
//     (q = p - offset, the same subtraction the reference's wrapper does);
//   * a primitive that is the second operand of a binary operator carries
//     the operator as its "combine" mode, so chains need one stack slot;
//   * runs of UNION (min) or INTERSECTION (max) are re-associated into left
//     folds.  min/max are exact and associative, so the value is the same;
//     the only freedom is the sign of a zero result, which neither the
//     classification (v > 0) nor the interpolation can observe.
#pragma once
#include <stdint.h>

#include "../../include/isosurface.h"

namespace iso {

enum DevCode : uint32_t {
  kSphere = 1,
  kTorus = 2,
  kBox = 3,
  kCylinder = 4,
  kUnion = 16,
  kIntersection = 17,
  kDifference = 18,
  kSmoothUnion = 19,
  kDifferenceRev = 20,  // a on top, b below: max(b, -a)
  kPushTranslate = 32,
  kPopTranslate = 33,
};

// How a primitive's value v joins the value stack.
enum Combine : uint32_t {
  kPush = 0,      // push v
  kMin = 1,       // top = min(top, v)          Union(top, v)
  kMax = 2,       // top = max(top, v)          Intersection(top, v)
  kDiff = 3,      // top = max(v, -top)         Difference(top, v): v minus top
  kSmooth = 4,    // top = smooth_min(top, v, k)
  kDiffRev = 5,   // top = max(top, -v)         Difference(v, top): top minus v
};

struct DevOp {
  uint32_t code;
  uint32_t combine;
  uint32_t translated;  // primitive: sample at (p - off)
  float k;              // smooth-min parameter for kSmooth / kSmoothUnion
  float prm[4];         // primitive parameters
  float off[4];         // translation (primitive offset or kPushTranslate)
};
static_assert(sizeof(DevOp) == 48, "DevOp layout");

constexpr int kMaxValueStack = 8;
constexpr int kMaxPointStack = 4;
constexpr int kMaxProgramOps = 8192;

}  // namespace iso
