// Host-visible view of an MLP handle for the fused MLP-IBL resolve.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "althea/althea.h"
#include "mlp_layout.hpp"

namespace althea {

// What the MLP-IBL resolve needs from an althea_mlp handle: the packed
// weights, the monitor's per-cell acceptance (built for one cube size and
// mip count; mon_mips == 0 before the first build) and the handle's queue of
// deferred pixels: one region of kMlpQueueRegion records (three float4 per
// pixel the MLP serves, see resolve_kernels.hip) per resolve wave, and one
// record count per region.
constexpr int kMlpQueueRegion = 64;
struct MlpInferenceView {
  MlpDeviceWeights w;
  const uint8_t* cell_ok;  // [mip][face][cy][cx], ALTHEA_MLP_MONITOR_CELLS per face
  int mon_size, mon_mips;
  int mon_accepted;  // accepted cells of that build
  float4* queue;
  uint32_t* region_count;  // [queue_regions]
  size_t queue_regions;
};
althea_status mlp_inference_view(const althea_mlp* m, MlpInferenceView* v);
// Grows the handle's queue to at least `regions` regions (device allocation
// after a device synchronisation when it grows) and fills the view's queue
// fields.
althea_status mlp_resolve_queue(const althea_mlp* m, size_t regions, MlpInferenceView* v);

}  // namespace althea
