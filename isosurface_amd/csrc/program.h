#pragma once
#include <string>
#include <vector>

#include "sdf_program.h"

namespace iso {

// Validates a post-order iso_node program and lowers it to DevOps.
// Returns ISO_OK or ISO_ERR_PROGRAM with a message in *err.
int lower_program(const iso_node* prog, uint32_t n, std::vector<DevOp>* out, std::string* err);

}  // namespace iso
