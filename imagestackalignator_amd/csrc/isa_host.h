// Host-side helpers of libisa_hip.so: status codes, error text, launch checks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/isa_hip.h"

namespace isa {

void set_error(const char* fmt, ...);

inline hipStream_t to_stream(isa_stream_t s) { return reinterpret_cast<hipStream_t>(s); }
inline int div_up(int a, int b) { return (a + b - 1) / b; }

// Checks the launch that was just issued (configuration errors surface here).
inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return ISA_ERR_HIP;
    }
    return ISA_OK;
}

inline int check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return ISA_ERR_HIP;
    }
    return ISA_OK;
}

#define ISA_REQUIRE(cond, ...)              \
    do {                                    \
        if (!(cond)) {                      \
            ::isa::set_error(__VA_ARGS__);  \
            return ISA_ERR_ARG;             \
        }                                   \
    } while (0)

#define ISA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != ISA_OK) return _rc; \
    } while (0)

// Copies a host table (bytes a multiple of 4, dst 4-byte aligned) into device memory in stream
// order through kernel arguments: the host buffer may be reused as soon as this returns, and
// nothing waits for the device. For small tables of the per-call entry points.
int upload_by_args(void* dst_dev, const void* src, size_t bytes, hipStream_t s);

// Shift-collection tables (R:PEFStudioDX/ShiftCollection.cs:113-496); host integers.
// Fills pairs (reference, toTrack) in GetShiftPairs order, the n x n index
// table and the (a, b) endpoints of every measurement row (row k measures
// x_b - x_a, x_j = sum_{i<j} r_i). Any output may be null. Returns m or < 0.
int shift_tables(int n, int ref, int strategy, int block, int* pairs, int* index, int* rows);

}  // namespace isa
