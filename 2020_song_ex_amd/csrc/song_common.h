// song_common.h -- constants shared by the host runtime, the NSW build and the kernels.
#pragma once

namespace song {

enum { METRIC_L2 = 0, METRIC_IP = 1, METRIC_HAMMING = 2 };

}  // namespace song
