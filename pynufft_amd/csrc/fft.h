// Line-FFT descriptor shared by the planner (api.hip) and the kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PNH_FFT_MAX_PASS 16
#define PNH_FFT_MAX_GENERIC 13

// Addressing of one side (source or destination) of a line transform.  Line
// (o0, o1, b), element t lives at o0*o_stride[0] + o1*o_stride[1] +
// b*b_stride + t*t_stride (in elements).
struct FftSide {
  int64_t o_stride[2];
  int64_t b_stride;
  int64_t t_stride;
};

struct FftDesc {
  int n;                // transform length
  int n_in;             // valid input elements per line (rest are zero)
  int n_out;            // outputs kept per line
  int sign;             // -1 forward, +1 inverse
  double scale;         // applied on store
  int64_t n_outer;      // outer lines (split as o0 * o_split + o1)
  int64_t o_split;
  int64_t n_inner;      // inner (contiguous-stride) lines
  FftSide src, dst;
  int sn_side;          // 0 none, 1 multiply on load, 2 multiply on store
  const void* sn;       // real apodization (float or double, matches dtype)
  int64_t sn_o_stride[2];
  int64_t sn_t_stride;
  int npass;
  int radix[PNH_FFT_MAX_PASS];
  int ns[PNH_FFT_MAX_PASS];
  int lines_per_block;
  int threads;
  int inner_major;      // 1: threads walk lines fastest (strided axis)
};

int pnh_fft_launch(const FftDesc& d, int fp64, const void* in, void* out,
                   const void* tw, const int* skip, hipStream_t stream);
int pnh_fft_factor(int n, int* radix, int* ns, int* npass);
// LDS line length in elements: one pad element per 16 (bank spreading).
__host__ __device__ inline int pnh_fft_lds_line(int n) { return n + (n >> 4) + 1; }
// LDS bytes of one workgroup: line buffers (none for a single pass, one for
// two passes, two for more) plus per-line offsets.
size_t pnh_fft_lds_bytes(const FftDesc& d, int fp64);
