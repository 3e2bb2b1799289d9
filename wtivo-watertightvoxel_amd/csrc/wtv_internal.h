// wtv_internal.h -- shared declarations of libwtv_gpupr (not part of the C-ABI).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "wtv_gpupr.h"

namespace wtv {

// Typed failure carried up to the C-ABI boundary, where it becomes a status
// code plus a thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

// Message prefix kept from the reference (gpu_push_relabel_fast.cu:26,354) so
// callers that match on it keep working.
constexpr const char* kPrefix = "GPUPr-FAST-v6.30 ";

// Solve a reduced graph already resident on the device (wtv_solver.hip).
void pr_solve(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term,
              uint8_t* partition, bool partition_on_device, const wtv_params& p,
              wtv_stats& st, void* stream);

// Solve G independent graphs as one disjoint union (wtv_solver.hip).
void batch_solve(uint32_t G, const uint64_t* h_off, const int32_t* d_nbr, uint32_t* d_res, int32_t* d_term,
                 uint8_t* partition, bool partition_on_device, int64_t* h_flows, const wtv_params& p,
                 wtv_stats& st, void* stream);

// Thread-local stats of the last solve (wtv_capi.cpp).
void set_last_stats(const wtv_stats& st);

}  // namespace wtv
