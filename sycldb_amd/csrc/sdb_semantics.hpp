// Per-row integer semantics of the reference functors, bit for bit:
//   comparisons / logical ops  kernels/selection.hpp:7-96
//   int32 binary ops           kernels/aggregation.hpp:11-35 (wrapping)
//   direct-mapped hash         kernels/join.hpp:11-15
// Division: the reference leaves x/0 and INT_MIN/-1 undefined; here x/0 = 0
// and INT_MIN/-1 = INT_MIN (two's-complement wrap).
//
// Shared by host code, the compiled kernels and the fused kernels compiled at
// run time (which see this file as an in-memory header).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "sycldb.h"

namespace sdb {

__host__ __device__ __forceinline__ bool compare(int cmp, int32_t a, int32_t b)
{
    switch (cmp) {
    case SDB_EQ: return a == b;
    case SDB_NE: return a != b;
    case SDB_LT: return a < b;
    case SDB_LE: return a <= b;
    case SDB_GT: return a > b;
    case SDB_GE: return a >= b;
    default: return false;
    }
}

__host__ __device__ __forceinline__ bool logical(int logic, bool a, bool b)
{
    switch (logic) {
    case SDB_AND: return a && b;
    case SDB_OR: return a || b;
    case SDB_NONE: return b;
    default: return false;
    }
}

__host__ __device__ __forceinline__ int32_t element_op(int op, int32_t a, int32_t b)
{
    const uint32_t ua = (uint32_t)a, ub = (uint32_t)b;
    switch (op) {
    case SDB_MUL: return (int32_t)(ua * ub);
    case SDB_ADD: return (int32_t)(ua + ub);
    case SDB_SUB: return (int32_t)(ua - ub);
    case SDB_DIV:
        if (b == 0) return 0;
        if (a == INT32_MIN && b == -1) return INT32_MIN;
        return a / b;
    default: return 0;
    }
}

// Slot of key x in a direct-mapped table over [min, min + len - 1]; only
// meaningful for keys inside that range (callers check it first).
__host__ __device__ __forceinline__ uint32_t slot_of(int32_t x, int32_t min)
{
    return (uint32_t)x - (uint32_t)min;
}

} // namespace sdb
