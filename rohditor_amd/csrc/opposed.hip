// Opposed highlight reconstruction (placeholder until the GPU port lands).
#include "rhd_internal.hpp"

namespace rhd {
size_t opposed_scratch_bytes(int, int) { return 256; }
hipError_t launch_opposed(const OpposedArgs&, hipStream_t) { return hipErrorNotSupported; }
}  // namespace rhd
