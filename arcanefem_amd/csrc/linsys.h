// linsys.h -- internal interface between the linear-system instance and the
// solver.
#pragma once

#include "common.h"

namespace afem {

// femutils/FemUtilsGlobal.h:51-62
constexpr uint8_t kEliminateNone = 0;
constexpr uint8_t kEliminateRow = 1;
constexpr uint8_t kEliminateRowColumn = 2;

void system_require_matrix(const afem_system* s);
void system_matrix_transformation(afem_system* s);
void system_rhs_transformation(afem_system* s);
void system_release_workspace(afem_system* s);

} // namespace afem
