// song_search_l2.hip -- instantiates the search kernels for METRIC_L2.
#include "song_search_impl.h"

namespace song {

hipError_t launch_search_l2(const SearchArgs& a, int S, int G, int qch, const LaunchShape& shape, hipStream_t st) {
    return launch_s<METRIC_L2>(a, S, G, qch, shape, st);
}

hipError_t launch_search(const SearchArgs& a, int metric, int S, int G, int qch, const LaunchShape& shape,
                         hipStream_t st) {
    switch (metric) {
    case METRIC_L2: return launch_search_l2(a, S, G, qch, shape, st);
    case METRIC_IP: return launch_search_ip(a, S, G, qch, shape, st);
    case METRIC_HAMMING: return launch_search_hamming(a, S, G, qch, shape, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace song
