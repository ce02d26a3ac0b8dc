// Thin C entry points over the reference's own expression semantics (its src/devtypes.h, compiled
// from where it lies under the reference tree by oracle/build.py into oracle/_ref/). Used only to
// pin the CPU oracle: tests compare oracle results with these on the same inputs, and
// tests/golden/make_golden.py records them as fixtures. No reference source is copied here.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "devtypes.h"

extern "C" {

double ref_eval_prog(const void* env, const void* cols, int16_t start, int16_t len, int64_t row) {
  const gdb::ExprEnv& e = *(const gdb::ExprEnv*)env;
  gdb::Prog p;
  p.start = start;
  p.len = len;
  return gdb::evalProgC(e, cols ? (const gdb::ColView*)cols : e.cols, p, row);
}
double ref_load_col(const void* cv, int64_t r) { return gdb::loadCol(*(const gdb::ColView*)cv, r); }
int64_t ref_col_int(const void* cv, int64_t r) { return gdb::colInt(*(const gdb::ColView*)cv, r); }
double ref_col_dbl(const void* cv, int64_t r) { return gdb::colDbl(*(const gdb::ColView*)cv, r); }
double ref_call1(int fn, double x) { return gdb::gdbCall1(fn, x); }
double ref_call2(int fn, double a, double b) { return gdb::gdbCall2(fn, a, b); }
int64_t ref_days_from_civil(int64_t y, unsigned m, unsigned d) { return gdb::daysFromCivil(y, m, d); }
void ref_civil_from_days(int64_t z, int64_t* y, unsigned* m, unsigned* d) {
  int64_t yy;
  gdb::civilFromDays(z, yy, *m, *d);
  *y = yy;
}
int64_t ref_checked(int op, int64_t a, int64_t b, int* f) {
  return op == 0 ? gdb::iaddc(a, b, f) : op == 1 ? gdb::isubc(a, b, f) : gdb::imulc(a, b, f);
}
// sizes of the wire structs, for the layout check
int64_t ref_sizeof(int which) {
  switch (which) {
    case 0: return sizeof(gdb::Instr);
    case 1: return sizeof(gdb::ColView);
    case 2: return sizeof(gdb::Prog);
    case 3: return sizeof(gdb::ExprEnv);
    case 4: return sizeof(gdb::AggPlan);
    case 5: return sizeof(gdb::DenseSpec);
    case 6: return sizeof(gdb::BatchSet);
  }
  return -1;
}
// field offsets of the plan structs
int64_t ref_offsetof(int which) {
  switch (which) {
    case 0: return offsetof(gdb::AggPlan, nkeys);
    case 1: return offsetof(gdb::AggPlan, naggs);
    case 2: return offsetof(gdb::AggPlan, ops);
    case 3: return offsetof(gdb::AggPlan, keyInt);
    case 4: return offsetof(gdb::AggPlan, genv);
    case 5: return offsetof(gdb::ExprEnv, cols);
    case 6: return offsetof(gdb::ExprEnv, ovf);
    case 7: return offsetof(gdb::DenseSpec, kmin);
    case 8: return offsetof(gdb::DenseSpec, slotOff);
    case 9: return offsetof(gdb::DenseSpec, kind);
    case 10: return offsetof(gdb::AggPlan, aggBits);
    case 11: return offsetof(gdb::AggPlan, densePriv);
  }
  return -1;
}

}  // extern "C"
