// kernels.h -- small device utilities shared by the engines.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "engine.h"

namespace bd {

// Fill n words with v on stream (a kernel, so it can sit inside a captured
// graph where a memset node must not).
void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t stream);

// Scratch for exclusive_scan_u32, grown on demand and kept between calls.
struct ScanTemp {
    DevBuf<uint32_t> level[4];
};

// out[i] = in[0] + ... + in[i-1] over n words (n < 2^32 total), in and out may
// alias.  Block-local scans use wavefront (64-lane) prefix scans; block sums
// are scanned recursively.
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, hipStream_t stream, ScanTemp &tmp,
                       uint64_t &launches, std::string &err);

}  // namespace bd
