/* Plain-data wire format between the engine's planner and the GPU operators.
 *
 * Every struct here is byte-compatible with the engine's own plan structs so a planner that fills
 * them (bytecode compiler, aggregate slot layout, dense-group spec, batch descriptors) can hand
 * them over by pointer without conversion. The layouts follow src/devtypes.h of the reference:
 *   ws_instr      = Instr      (devtypes.h:85-89)      4 bytes
 *   ws_colview    = ColView    (devtypes.h:91-102)    16 bytes
 *   ws_prog       = Prog       (devtypes.h:179-181)    4 bytes
 *   ws_expr_env   = ExprEnv    (devtypes.h:185-194) 4488 bytes
 *   ws_agg_plan   = AggPlan    (devtypes.h:503-520)  648 bytes
 *   ws_dense_spec = DenseSpec  (devtypes.h:528-538)  536 bytes
 *   ws_batch_set  = BatchSet   (devtypes.h:541-547)   32 bytes
 * and the numeric codes (opcodes, column type codes, encodings, aggregate ops, dense kinds,
 * scalar function ids) are the same values, because the bytecode is data the engine produces.
 *
 * This header is C, C++ and hipRTC compatible (no host headers under hipRTC). */
#ifndef WARPSTORE_AMD_WS_ABI_H
#define WARPSTORE_AMD_WS_ABI_H

#if defined(__HIPCC_RTC__)
typedef signed char int8_t;
typedef unsigned char uint8_t;
typedef short int16_t;
typedef unsigned short uint16_t;
typedef int int32_t;
typedef unsigned int uint32_t;
typedef long long int64_t;
typedef unsigned long long uint64_t;
#else
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- limits (devtypes.h:168-177) */
#define WS_MAX_STACK 16
#define WS_MAX_COLS 128
#define WS_MAX_CODE 320
#define WS_MAX_CONSTS 96
#define WS_MAX_SETS 16
#define WS_MAX_KEYS 8
#define WS_MAX_AGGS 64
#define WS_INTERP_KEYS 4  /* interpreter limits; bigger plans need a compiled plan (error 1235) */
#define WS_INTERP_AGGS 16
#define WS_ENC_HEADER 16  /* encoded column: int64 base (+ double 10^s for enc 5/6), then codes */

/* ---- NULL sentinels (devtypes.h:27-28); doubles use NaN */
#define WS_NULL_I32 ((int32_t)0x80000000)
#define WS_NULL_I64 ((int64_t)0x8000000000000000ull)

/* ---- column type codes, ws_colview.t (devtypes.h:93-94) */
enum {
  WS_T_I32 = 0,
  WS_T_I64 = 1,
  WS_T_F64 = 2,      /* NULL = NaN */
  WS_T_I32N = 3,     /* int32 with NULL sentinel */
  WS_T_I64N = 4,     /* int64 with NULL sentinel */
  WS_T_VEC = 5,      /* float32 vectors */
  WS_T_DECIMAL = 6   /* int64 scaled by 10^scale, NULL sentinel */
};

/* ---- physical storage types (common.h:24) */
enum { WS_P_I32 = 0, WS_P_I64 = 1, WS_P_F64 = 2, WS_P_VEC = 5 };

/* ---- bytecode opcodes (devtypes.h:64-83) */
enum {
  WS_OP_COL = 0, WS_OP_CONST,
  WS_OP_ADD, WS_OP_SUB, WS_OP_MUL, WS_OP_DIV, WS_OP_MOD, WS_OP_NEG,
  WS_OP_EQ, WS_OP_NE, WS_OP_LT, WS_OP_LE, WS_OP_GT, WS_OP_GE,
  WS_OP_AND, WS_OP_OR, WS_OP_NOT,
  WS_OP_INSET, WS_OP_SELECT, WS_OP_YEAR, WS_OP_ISNULL, WS_OP_COALESCE, WS_OP_NULLIF,
  WS_OP_CALL1, WS_OP_CALL2, WS_OP_DICTMAP, WS_OP_INSORTED, WS_OP_VDOT, WS_OP_VEUC, WS_OP_COLRAW
};

/* ---- scalar function ids (devtypes.h:307-317): part of the bytecode, numbering is fixed */
enum {
  WS_FN_ABS = 0, WS_FN_CEIL, WS_FN_FLOOR, WS_FN_ROUND, WS_FN_SQRT, WS_FN_EXP, WS_FN_LN, WS_FN_LOG10, WS_FN_LOG2,
  WS_FN_SIN, WS_FN_COS, WS_FN_TAN, WS_FN_SIGN, WS_FN_MONTH, WS_FN_DAY, WS_FN_DAYOFWEEK, WS_FN_DAYOFYEAR,
  WS_FN_QUARTER, WS_FN_WEEKDAY, WS_FN_LAST_DAY, WS_FN_DT_DATE, WS_FN_DT_HOUR, WS_FN_DT_MINUTE, WS_FN_DT_SECOND,
  WS_FN_DATE_TO_DT, WS_FN_DT_UNIX, WS_FN_UNIX_DT, WS_FN_DT_YEAR, WS_FN_DT_MONTH, WS_FN_DT_DAY, WS_FN_ASIN,
  WS_FN_ACOS, WS_FN_ATAN, WS_FN_DEGREES, WS_FN_RADIANS, WS_FN_WEEK, WS_FN_TO_DAYS, WS_FN_FROM_DAYS,
  WS_FN_DT_MICRO, WS_FN_YEARWEEK,
  WS_FN_ROUND2 = 100, WS_FN_POW, WS_FN_ATAN2, WS_FN_LEAST, WS_FN_GREATEST, WS_FN_TRUNCATE2, WS_FN_ADD_MONTHS,
  WS_FN_DT_ADD_MONTHS, WS_FN_LOG_B, WS_FN_IDIV, WS_FN_DATE_TRUNC_MONTH, WS_FN_DT_DIFF_UNIT
};

/* ---- aggregate ops (devtypes.h:501) and dense accumulator kinds (devtypes.h:527) */
enum { WS_AGG_SUM = 0, WS_AGG_COUNT = 1, WS_AGG_MIN = 2, WS_AGG_MAX = 3 };
enum { WS_DK_COUNT = 0, WS_DK_SUM_I, WS_DK_SUM_D, WS_DK_MIN_I, WS_DK_MAX_I, WS_DK_MIN_D, WS_DK_MAX_D };

/* ---- kernel kinds of a compiled plan (jit.h:14) */
enum { WS_JF_NOKEYS = 0, WS_JF_GROUPED = 1, WS_JF_DENSE = 2, WS_JF_FLAG = 3, WS_JF_FLAGM = 4 };

/* ---- codegen modes (jit.h:36-42) */
enum { WS_JIT_OFF = 0, WS_JIT_ON = 1, WS_JIT_SYNC = 2 };

typedef struct ws_instr {
  uint8_t op;
  uint8_t is_int; /* OP_CONST: the literal is an integer (its int64 bits sit in the double slot) */
  uint16_t arg;
} ws_instr;

typedef struct ws_colview {
  const void* p;
  uint8_t t;     /* WS_T_* */
  uint8_t scale; /* DECIMAL scale */
  uint8_t enc;   /* 0 raw; 1/2/3 u8/u16/u32 frame of reference; 4 f64 dictionary; 5/6 decimal doubles.
                    An encoded buffer is header + codes + 16 bytes (the segment format's tail, as the
                    engine's encoders allocate it): kernels read codes as aligned 32-bit words. */
  int32_t off;   /* first row of an encoded column (range scans) */
} ws_colview;

typedef struct ws_prog {
  int16_t start, len;
} ws_prog;

typedef struct ws_expr_env {
  ws_instr code[WS_MAX_CODE];
  double consts[WS_MAX_CONSTS];
  const uint32_t* sets[WS_MAX_SETS];
  int32_t set_size[WS_MAX_SETS];
  const double* maps[WS_MAX_SETS];
  int32_t map_size[WS_MAX_SETS];
  ws_colview cols[WS_MAX_COLS];
  int* ovf; /* device flag set by checked int64 arithmetic that overflows */
} ws_expr_env;

typedef struct ws_agg_plan {
  uint8_t has_where; /* C++ bool in the engine's struct: one byte */
  ws_prog where;
  int32_t nkeys;
  ws_prog keys[WS_MAX_KEYS];
  int32_t naggs; /* slot 0 is always a hidden COUNT(*) */
  ws_prog aggs[WS_MAX_AGGS];
  uint8_t ops[WS_MAX_AGGS];
  uint8_t agg_int[WS_MAX_AGGS];  /* argument is integer valued */
  uint8_t agg_bits[WS_MAX_AGGS]; /* GROUP BY: slot holds exact int64 bits */
  uint8_t agg_part[WS_MAX_AGGS]; /* with agg_bits on a SUM: 1 = sum of low 32 bits, 2 = of high 32 */
  uint8_t agg_null[WS_MAX_AGGS]; /* nullable int column (t 3/4): NULL sentinel skipped */
  uint8_t dense_priv;            /* 1 per-thread, 2 copies/global; 0 unknown */
  uint8_t key_int[WS_MAX_KEYS];  /* key stored as exact int64 bits in its double slot */
  const ws_expr_env* genv;       /* interpreted kernels: device copy of the env */
} ws_agg_plan;

typedef struct ws_dense_spec {
  int64_t G; /* 0 => hash aggregation */
  int32_t copies;
  int32_t copy_bytes;
  int32_t priv;
  int64_t kmin[WS_MAX_KEYS];
  int64_t range[WS_MAX_KEYS];
  int64_t stride[WS_MAX_KEYS];
  int32_t slot_off[WS_MAX_AGGS];
  uint8_t kind[WS_MAX_AGGS];
} ws_dense_spec;

typedef struct ws_batch_set {
  int32_t nb;
  int32_t ncols;
  const int64_t* n;
  const uint32_t* const* del;
  const ws_colview* cols;
} ws_batch_set;

#ifdef __cplusplus
}
#endif

#if defined(__cplusplus) && !defined(__HIPCC_RTC__)
static_assert(sizeof(ws_instr) == 4, "Instr layout");
static_assert(sizeof(ws_colview) == 16, "ColView layout");
static_assert(sizeof(ws_prog) == 4, "Prog layout");
static_assert(sizeof(ws_expr_env) == 4488, "ExprEnv layout");
static_assert(sizeof(ws_agg_plan) == 648, "AggPlan layout");
static_assert(sizeof(ws_dense_spec) == 536, "DenseSpec layout");
static_assert(sizeof(ws_batch_set) == 32, "BatchSet layout");
static_assert(offsetof(ws_agg_plan, nkeys) == 8 && offsetof(ws_agg_plan, naggs) == 44 &&
                  offsetof(ws_agg_plan, ops) == 304 && offsetof(ws_agg_plan, key_int) == 625 &&
                  offsetof(ws_agg_plan, genv) == 640,
              "AggPlan field offsets");
static_assert(offsetof(ws_expr_env, cols) == 2432 && offsetof(ws_expr_env, ovf) == 4480, "ExprEnv field offsets");
static_assert(offsetof(ws_dense_spec, kmin) == 24 && offsetof(ws_dense_spec, slot_off) == 216 &&
                  offsetof(ws_dense_spec, kind) == 472,
              "DenseSpec field offsets");
#endif

#endif /* WARPSTORE_AMD_WS_ABI_H */
