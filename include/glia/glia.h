/*
 * glia.h -- the reference's class interface for the reaction-diffusion path (Parameters, Tumor,
 * MatProp, DiffCoef, ReacCoef, Phi, Obs, Data, DiffusionSolver, PdeOperatorsRD,
 * DerivativeOperatorsRD, TILOptimizer, RDOptimizer, SparseTILOptimizer, SemiLagrangianSolver,
 * VariableLinearElasticitySolver), header-only C++17 over the C ABI of libglia_rd.so
 * (include/glia_rd.h, include/glia_opt.h, include/glia_adv.h, include/glia_el.h).
 * Link with -lglia_rd; every method returns 0 or the ABI's error code.
 */
#ifndef GLIA_GLIA_H_
#define GLIA_GLIA_H_

#include "AdvectionSolver.h"
#include "DerivativeOperators.h"
#include "ElasticitySolver.h"
#include "Optimizer.h"
#include "Parameters.h"
#include "PdeOperators.h"
#include "Tumor.h"
#include "Vec.h"

#endif
