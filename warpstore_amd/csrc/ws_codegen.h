// Code generation for per-query compiled plans (host side; see ws_codegen.hip).
#pragma once
#include <string>

#include "warpstore_amd/ws_abi.h"

namespace ws {
namespace codegen {

// The evaluator struct (`struct QEval`) of one plan: its WHERE, keys and aggregate arguments as
// typed device code, shape queries as compile-time constants. *rows gets the rows per thread per
// step the kernels should take with it. Throws ws::Error (1105) on a malformed plan.
std::string evaluator_source(const ws_expr_env& env, int ncols, const ws_agg_plan& plan, int* rows);
// Whether a program's value is an exact int64 (integer columns and literals through + - * and
// integer-valued operators), as the compiled evaluator types it (reference jit.h:51 progIsInt).
bool program_is_int(const ws_expr_env& env, ws_prog p);

}  // namespace codegen
}  // namespace ws
