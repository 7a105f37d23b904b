// The wave stage kernels instantiated for one metric storage type (MetBf16Delta).
#include "wave_stage_kernels.h"

namespace cgfdm {
namespace wave {
template void launch_wave_dispatch<MetBf16Delta>(const WaveArgs<MetBf16Delta>&, const Ctx&, int, int, int, hipStream_t);
}  // namespace wave
}  // namespace cgfdm
