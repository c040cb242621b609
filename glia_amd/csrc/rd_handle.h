// rd_handle.h -- the object behind glia_rd_t (include/glia_rd.h), shared by the C-ABI files that
// compose the engine with other solvers.
#pragma once
#include <memory>
#include <string>

#include "rd_solver.h"

struct glia_rd_s {
  std::unique_ptr<glia::RDEnsemble> e;
  std::string err;
};
