// Host-side validation and lowering of iso_node programs into DevOps.
#include "program.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

namespace iso {

namespace {

struct Node {
  uint32_t op;
  float p[6];
  int a = -1, b = -1;
};

bool is_primitive(uint32_t op) {
  return op == ISO_OP_SPHERE || op == ISO_OP_TORUS || op == ISO_OP_BOX || op == ISO_OP_CYLINDER;
}

class Lowerer {
 public:
  explicit Lowerer(const std::vector<Node>& n) : nodes_(n), need_(n.size(), -1) {}

  void emit(int id) {
    const Node& n = nodes_[id];
    if (leaf_like(id)) {
      emit_leaf(id, kPush, 0.0f);
      return;
    }
    if (n.op == ISO_OP_TRANSLATE) {
      DevOp o{};
      o.code = kPushTranslate;
      o.off[0] = n.p[0];
      o.off[1] = n.p[1];
      o.off[2] = n.p[2];
      ops_.push_back(o);
      emit(n.a);
      DevOp c{};
      c.code = kPopTranslate;
      ops_.push_back(c);
      return;
    }
    if (n.op == ISO_OP_UNION || n.op == ISO_OP_INTERSECTION) {
      // min / max are exact and associative: evaluate the operand list as a
      // left fold, deepest operand first, primitives fused into the fold.
      std::vector<int> operands;
      flatten(id, n.op, &operands);
      std::stable_sort(operands.begin(), operands.end(),
                       [&](int x, int y) { return need(x) > need(y); });
      const Combine mode = n.op == ISO_OP_UNION ? kMin : kMax;
      const uint32_t code = n.op == ISO_OP_UNION ? kUnion : kIntersection;
      emit(operands[0]);
      for (size_t i = 1; i < operands.size(); ++i) join(operands[i], mode, code, 0.0f);
      return;
    }
    // DIFFERENCE (max(b, -a)) or SMOOTH_UNION (symmetric in a, b).
    const bool diff = n.op == ISO_OP_DIFFERENCE;
    const float k = diff ? 0.0f : n.p[0];
    if (leaf_like(n.b) || (!leaf_like(n.a) && need(n.a) >= need(n.b))) {
      emit(n.a);  // a below / on top, then b
      join(n.b, diff ? kDiff : kSmooth, diff ? kDifference : kSmoothUnion, k);
    } else {
      emit(n.b);  // b first, then a
      join(n.a, diff ? kDiffRev : kSmooth, diff ? kDifferenceRev : kSmoothUnion, k);
    }
  }

  std::vector<DevOp> take() { return std::move(ops_); }

 private:
  // A (possibly translated) primitive.
  bool leaf_like(int id) const {
    const Node& n = nodes_[id];
    return is_primitive(n.op) || (n.op == ISO_OP_TRANSLATE && is_primitive(nodes_[n.a].op));
  }

  // Value-stack slots needed to evaluate node `id` (Sethi-Ullman number with
  // primitives fused into their parent's combine step).
  int need(int id) {
    if (need_[id] >= 0) return need_[id];
    const Node& n = nodes_[id];
    int r = 1;
    if (leaf_like(id)) {
      r = 1;
    } else if (n.op == ISO_OP_TRANSLATE) {
      r = need(n.a);
    } else if (n.op == ISO_OP_UNION || n.op == ISO_OP_INTERSECTION) {
      std::vector<int> ops;
      flatten(id, n.op, &ops);
      std::vector<int> needs;
      for (int o : ops) needs.push_back(leaf_like(o) ? 0 : need(o));
      std::sort(needs.begin(), needs.end(), std::greater<int>());
      r = needs[0] > 0 ? needs[0] : 1;
      for (size_t i = 1; i < needs.size(); ++i)
        if (needs[i] > 0) r = std::max(r, needs[i] + 1);
    } else {
      const bool la = leaf_like(n.a), lb = leaf_like(n.b);
      if (la && lb) {
        r = 1;
      } else if (lb) {
        r = need(n.a);
      } else if (la) {
        r = need(n.b);
      } else {
        const int na = need(n.a), nb = need(n.b);
        r = na == nb ? na + 1 : std::max(na, nb);
      }
    }
    need_[id] = r;
    return r;
  }

  void join(int id, Combine mode, uint32_t code, float k) {
    if (leaf_like(id)) {
      emit_leaf(id, mode, k);
    } else {
      emit(id);
      DevOp o{};
      o.code = code;
      o.k = k;
      ops_.push_back(o);
    }
  }

  void emit_leaf(int id, Combine mode, float k) {
    const Node* n = &nodes_[id];
    DevOp o{};
    if (n->op == ISO_OP_TRANSLATE) {
      o.translated = 1;
      o.off[0] = n->p[0];
      o.off[1] = n->p[1];
      o.off[2] = n->p[2];
      n = &nodes_[n->a];
    }
    o.code = n->op;  // primitive opcodes coincide with DevCode values
    o.combine = mode;
    o.k = k;
    for (int i = 0; i < 4; ++i) o.prm[i] = n->p[i];
    ops_.push_back(o);
  }

  void flatten(int id, uint32_t op, std::vector<int>* out) const {
    const Node& n = nodes_[id];
    if (n.op != op) {
      out->push_back(id);
      return;
    }
    flatten(n.a, op, out);
    flatten(n.b, op, out);
  }

  const std::vector<Node>& nodes_;
  std::vector<int> need_;
  std::vector<DevOp> ops_;
};

}  // namespace

int lower_program(const iso_node* prog, uint32_t n, std::vector<DevOp>* out, std::string* err) {
  if (prog == nullptr || n == 0) {
    *err = "empty program";
    return ISO_ERR_PROGRAM;
  }
  if (n > static_cast<uint32_t>(kMaxProgramOps)) {
    *err = "program longer than " + std::to_string(kMaxProgramOps) + " nodes";
    return ISO_ERR_PROGRAM;
  }
  std::vector<Node> nodes;
  std::vector<int> st;
  std::vector<std::pair<size_t, int>> scopes;
  for (uint32_t i = 0; i < n; ++i) {
    const iso_node& in = prog[i];
    if (in.reserved != 0) {
      *err = "node " + std::to_string(i) + ": reserved field must be 0";
      return ISO_ERR_PROGRAM;
    }
    Node nd;
    nd.op = in.op;
    std::memcpy(nd.p, in.p, sizeof(nd.p));
    const size_t floor_sz = scopes.empty() ? 0 : scopes.back().first;
    switch (in.op) {
      case ISO_OP_SPHERE:
      case ISO_OP_TORUS:
      case ISO_OP_BOX:
      case ISO_OP_CYLINDER:
        nodes.push_back(nd);
        st.push_back(static_cast<int>(nodes.size()) - 1);
        break;
      case ISO_OP_UNION:
      case ISO_OP_INTERSECTION:
      case ISO_OP_DIFFERENCE:
      case ISO_OP_SMOOTH_UNION:
        if (st.size() < floor_sz + 2) {
          *err = "node " + std::to_string(i) + ": binary operator needs two operands";
          return ISO_ERR_PROGRAM;
        }
        if (in.op == ISO_OP_SMOOTH_UNION && !(in.p[0] > 0.0f)) {
          *err = "node " + std::to_string(i) + ": smooth union needs k > 0";
          return ISO_ERR_PROGRAM;
        }
        nd.b = st.back();
        st.pop_back();
        nd.a = st.back();
        st.pop_back();
        nodes.push_back(nd);
        st.push_back(static_cast<int>(nodes.size()) - 1);
        break;
      case ISO_OP_TRANSLATE:
        nodes.push_back(nd);
        scopes.push_back({st.size(), static_cast<int>(nodes.size()) - 1});
        break;
      case ISO_OP_END_TRANSLATE: {
        if (scopes.empty()) {
          *err = "node " + std::to_string(i) + ": END_TRANSLATE without TRANSLATE";
          return ISO_ERR_PROGRAM;
        }
        auto sc = scopes.back();
        scopes.pop_back();
        if (st.size() != sc.first + 1) {
          *err = "node " + std::to_string(i) + ": translate scope must hold exactly one value";
          return ISO_ERR_PROGRAM;
        }
        nodes[sc.second].a = st.back();
        st.pop_back();
        st.push_back(sc.second);
        break;
      }
      default:
        *err = "node " + std::to_string(i) + ": unknown opcode " + std::to_string(in.op);
        return ISO_ERR_PROGRAM;
    }
  }
  if (!scopes.empty()) {
    *err = "unterminated TRANSLATE";
    return ISO_ERR_PROGRAM;
  }
  if (st.size() != 1) {
    *err = "program leaves " + std::to_string(st.size()) + " values (needs exactly 1)";
    return ISO_ERR_PROGRAM;
  }
  Lowerer lw(nodes);
  lw.emit(st.back());
  std::vector<DevOp> ops = lw.take();

  // Depth check of the lowered program against the device's register stacks.
  int vdepth = 0, vmax = 0, pdepth = 0, pmax = 0;
  for (const DevOp& o : ops) {
    switch (o.code) {
      case kSphere:
      case kTorus:
      case kBox:
      case kCylinder:
        if (o.combine == kPush) vdepth++;
        break;
      case kUnion:
      case kIntersection:
      case kDifference:
      case kDifferenceRev:
      case kSmoothUnion:
        vdepth--;
        break;
      case kPushTranslate:
        pdepth++;
        break;
      case kPopTranslate:
        pdepth--;
        break;
    }
    vmax = vdepth > vmax ? vdepth : vmax;
    pmax = pdepth > pmax ? pdepth : pmax;
  }
  if (vmax > kMaxValueStack) {
    *err = "program needs a value stack of " + std::to_string(vmax) + " (max " +
           std::to_string(kMaxValueStack) + ")";
    return ISO_ERR_PROGRAM;
  }
  if (pmax > kMaxPointStack) {
    *err = "program nests " + std::to_string(pmax) + " translations (max " +
           std::to_string(kMaxPointStack) + ")";
    return ISO_ERR_PROGRAM;
  }
  *out = std::move(ops);
  return ISO_OK;
}

}  // namespace iso
