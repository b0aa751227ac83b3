// dop_build.hpp -- validation of a dop_desc and construction of the host-side plan
// tables (before upload).
#pragma once
#include <string>
#include <vector>

#include "../../include/dop.h"
#include "dop_plan.hpp"
#include "dop_rows.hpp"

namespace dop {

struct HostPlan {
    PlanDev meta{};             // everything but the device table pointers
    AxisTables axis[kMaxDim];
    int64_t in_elems = 0;       // elements of u (or M)
    int64_t out_elems = 0;      // elements of du
    int max_derivative_order = 0;
};

// Returns DOP_OK or a DOP_E* code with *err set.
int build_host_plan(const dop_desc& D, HostPlan* out, std::string* err);

// Two plans have the same structure (every field but the BC taps / constants and the table
// contents and pointers): the launches of one serve the other's values.
bool same_structure(const PlanDev& a, const PlanDev& b);

}  // namespace dop
