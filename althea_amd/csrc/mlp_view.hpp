// Host-visible view of an MLP handle for the fused MLP-IBL resolve.
#pragma once

#include <cstdint>

#include "althea/althea.h"
#include "mlp_layout.hpp"

namespace althea {

// What the fused resolve needs from an althea_mlp handle: the packed
// weights and the monitor's per-cell acceptance (built for one cube size and
// mip count; mon_mips == 0 before the first build).
struct MlpInferenceView {
  MlpDeviceWeights w;
  const uint8_t* cell_ok;  // [mip][face][cy][cx], ALTHEA_MLP_MONITOR_CELLS per face
  int mon_size, mon_mips;
  int mon_accepted;  // accepted cells of that build
};
althea_status mlp_inference_view(const althea_mlp* m, MlpInferenceView* v);

}  // namespace althea
