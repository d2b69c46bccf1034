// Hand-written matrix-core GEMMs of the Γ projector contractions (qbh_gemm.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace qbh {

// Slices the reduction of an R x C output with reduction length Q is cut into (want > 0:
// that many, else enough for ~2048 workgroups, each slice at least 512 long); the count
// actually used (every slice non-empty).
int mfma_slices(int R, int C, long Q, int want);
// Length of each slice for `slices` slices (a multiple of the 16-element stage).
long mfma_slice_len(long Q, int slices);

// O_s[r + c ldo] = sum_{q in slice s} X[q + r ldx] Y[q + c ldy], s < slices, with
// O_s = O + s ostride (the fnl product: r = projector, c = band, q = real-view row).
// FP64 or FP32 operands and outputs.  Returns 0, or -1 for invalid sizes.
int mfma_gemm_qq(hipStream_t st, bool fp32, int R, int C, long Q, const void* X, long ldx,
                 const void* Y, long ldy, void* O, long ldo, int slices, long ostride);

// O[c + r ldo] (+)= sum_q X[q + r ldx] Y[c + q ldy] (the enl product taken transposed:
// r = band, c = real-view row, q = projector; X = F, Y = the projector table).
int mfma_gemm_qc(hipStream_t st, bool fp32, int R, int C, long Q, const void* X, long ldx,
                 const void* Y, long ldy, void* O, long ldo, bool accumulate);

}  // namespace qbh
