// Internal declarations of libmr_aha_c (not installed).
//
// Error plumbing: every internal function returns an mr_status int and
// records a message with mr_set_error(); public entry points clear the
// message on entry (src/common/mr_config.cpp:21-50 of the reference has the
// same contract).
#pragma once

#include "mr_aha/mr_aha.h"

#include <cstdint>
#include <cstdio>

void mr_set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void mr_clear_error();
// Backend hook of mr_clear_error: drops the HIP runtime's sticky-free
// "last error" so a failure the caller already handled is not reported as a
// later launch failure (reference: src/cuda/mr_backend_cuda.cu:11-21).
void mr_backend_clear_error();

#define MR_FAIL(status, ...)                                                   \
    do {                                                                       \
        mr_set_error(__VA_ARGS__);                                             \
        return (status);                                                       \
    } while (0)

#define MR_REQUIRE(cond, ...)                                                  \
    do {                                                                       \
        if (!(cond)) MR_FAIL(MR_ERR_INVALID_ARG, __VA_ARGS__);                 \
    } while (0)

#define MR_REQUIRE_CFG(cond, ...)                                              \
    do {                                                                       \
        if (!(cond)) MR_FAIL(MR_ERR_INVALID_CONFIG, __VA_ARGS__);              \
    } while (0)

#define MR_TRY(expr)                                                           \
    do {                                                                       \
        int mr_rc_ = (expr);                                                   \
        if (mr_rc_ != MR_SUCCESS) return mr_rc_;                               \
    } while (0)

// Geometry by value for kernels (re-entrant: no constant memory).
struct mr_dims {
    int n[3];          // image extents, unused axes 1
    int m[3];          // upsampled extents, unused axes 1
    int c_n[3];        // floor(n/2)
    int c_m[3];        // floor(m/2)
    int L;
    int tri_upper;
    long long image_size;  // prod n
    long long pad_size;    // prod m
};

void mr_dims_from_config(const mr_config *cfg, mr_dims *d);

inline long long mr_ceil_div(long long a, long long b) { return (a + b - 1) / b; }

// Read an integer environment variable (def when unset or empty).
long long mr_env_ll(const char *name, long long def);
