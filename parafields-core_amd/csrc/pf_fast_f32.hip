// FP32 instantiations of the specialised line-FFT kernels.
#include "pf_fast_impl.hpp"
namespace pfamd {
PF_INSTANTIATE(float)
}
