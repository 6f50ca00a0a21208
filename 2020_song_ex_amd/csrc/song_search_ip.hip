// song_search_ip.hip -- instantiates the search kernels for METRIC_IP.
#include "song_search_impl.h"

namespace song {

hipError_t launch_search_ip(const SearchArgs& a, int S, int G, int qch, const LaunchShape& shape, hipStream_t st) {
    return launch_s<METRIC_IP>(a, S, G, qch, shape, st);
}

}  // namespace song
