// The forward NTT at d = 2048 on the matrix cores (ntt_mfma.hip): its
// constant tables and launcher.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kernels.hpp"

namespace inspire {

// Host tables for the context's moduli and roots (psi per modulus), packed
// as one byte array with the offsets MfmaNttTables needs relative to its start.
struct MfmaNttHost {
    std::vector<uint8_t> bytes;
    size_t o_tw1, o_tw2, o_w64, o_w32, o_c1;
};
MfmaNttHost ntt_mfma_tables(uint32_t K, const uint64_t* q, const uint64_t* psi);
// In-place or strided forward transform of count [K][2048] polys, output as
// launch_ntt's (canonical, scaled by sc when enabled).
hipError_t launch_ntt_fwd_mfma(const uint32_t* src, size_t src_stride, uint32_t* dst, size_t dst_stride, size_t count,
                               const MfmaNttTables& T, const CtxDev& C, const ScaleDev& sc, hipStream_t st);

}  // namespace inspire
