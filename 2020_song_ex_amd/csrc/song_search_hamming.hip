// song_search_hamming.hip -- instantiates the search kernels for METRIC_HAMMING.
#include "song_search_impl.h"

namespace song {

hipError_t launch_search_hamming(const SearchArgs& a, int S, int G, int qch, const LaunchShape& shape, hipStream_t st) {
    return launch_s<METRIC_HAMMING>(a, S, G, qch, shape, st);
}

}  // namespace song
