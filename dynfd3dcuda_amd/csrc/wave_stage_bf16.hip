// The wave stage kernels instantiated for one metric storage type (uint16_t).
#include "wave_stage_kernels.h"

namespace cgfdm {
namespace wave {
template void launch_wave_dispatch<uint16_t>(const WaveArgs<uint16_t>&, const Ctx&, int, int, int, hipStream_t);
}  // namespace wave
}  // namespace cgfdm
