// Box-pass kernels for line-length group 2 (fft_lengths.h).
#include "box_kernels.h"

namespace smk_box {
SMK_FFT_LENGTHS_2(SMK_LEN_INSTANTIATE)
}  // namespace smk_box
