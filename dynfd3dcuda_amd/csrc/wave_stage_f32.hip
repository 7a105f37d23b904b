// The wave stage kernels instantiated for one metric storage type (float).
#include "wave_stage_kernels.h"

namespace cgfdm {
namespace wave {
template void launch_wave_dispatch<float>(const WaveArgs<float>&, const Ctx&, int, int, int, hipStream_t);
}  // namespace wave
}  // namespace cgfdm
