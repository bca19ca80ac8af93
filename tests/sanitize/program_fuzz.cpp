// Drives the host-side program lowering (isosurface_amd/csrc/program.cpp)
// with seeded random programs, well-formed and malformed, under
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile).
// Well-formed programs must lower; malformed ones must be refused; neither
// may touch memory out of bounds or hit undefined behaviour.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../isosurface_amd/csrc/program.h"

namespace {

uint64_t g_state = 1;
uint64_t next_u64() {  // SplitMix64
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
float uniform(float lo, float hi) {
  return lo + (hi - lo) * static_cast<float>((next_u64() >> 40) * (1.0 / (1 << 24)));
}

iso_node node(uint32_t op) {
  iso_node n{};
  n.op = op;
  for (float& p : n.p) p = uniform(0.01f, 0.5f);
  return n;
}

// A well-formed random tree of `depth` levels in post order.
void tree(int depth, std::vector<iso_node>* out) {
  const uint64_t r = next_u64() % 10;
  if (depth == 0 || r < 3) {
    static const uint32_t prims[4] = {ISO_OP_SPHERE, ISO_OP_TORUS, ISO_OP_BOX, ISO_OP_CYLINDER};
    out->push_back(node(prims[next_u64() % 4]));
    return;
  }
  if (r < 5) {  // translate around a subtree
    out->push_back(node(ISO_OP_TRANSLATE));
    tree(depth - 1, out);
    out->push_back(node(ISO_OP_END_TRANSLATE));
    return;
  }
  static const uint32_t bins[4] = {ISO_OP_UNION, ISO_OP_INTERSECTION, ISO_OP_DIFFERENCE,
                                   ISO_OP_SMOOTH_UNION};
  tree(depth - 1, out);
  tree(depth - 1, out);
  out->push_back(node(bins[next_u64() % 4]));
}

}  // namespace

int main(int argc, char** argv) {
  const int rounds = argc > 1 ? std::atoi(argv[1]) : 20000;
  int lowered = 0, refused = 0;
  for (int i = 0; i < rounds; ++i) {
    g_state = 0x1234567ull + static_cast<uint64_t>(i);
    std::vector<iso_node> prog;
    tree(1 + static_cast<int>(next_u64() % 9), &prog);
    const bool corrupt = (i % 3) == 2;
    if (corrupt && !prog.empty()) {
      // drop, duplicate or rewrite nodes: most such programs are malformed
      const uint64_t k = next_u64() % prog.size();
      switch (next_u64() % 5) {
        case 0: prog.erase(prog.begin() + static_cast<long>(k)); break;
        case 1: prog.insert(prog.begin() + static_cast<long>(k), prog[k]); break;
        case 2: prog[k].op = static_cast<uint32_t>(next_u64() % 40); break;
        case 3: prog[k].reserved = 1; break;
        default: prog.resize(k); break;
      }
    }
    std::vector<iso::DevOp> ops;
    std::string err;
    const int st = iso::lower_program(prog.data(), static_cast<uint32_t>(prog.size()), &ops, &err);
    if (st == ISO_OK) {
      ++lowered;
      if (ops.empty()) {
        std::fprintf(stderr, "round %d: a lowered program has no ops\n", i);
        return 1;
      }
    } else {
      ++refused;
      if (!corrupt && prog.size() <= 4096) {
        // a well-formed tree may still exceed the device stacks; say why
        if (err.find("stack") == std::string::npos && err.find("deep") == std::string::npos &&
            err.find("nests") == std::string::npos) {
          std::fprintf(stderr, "round %d: well-formed program refused: %s\n", i, err.c_str());
          return 1;
        }
      }
    }
  }
  std::printf("program_fuzz: %d lowered, %d refused\n", lowered, refused);
  return 0;
}
