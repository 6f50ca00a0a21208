// song_search_part.hip -- instantiates the search kernels of one metric and
// one set size (S slots per lane).  The Makefile compiles this file once per
// (SONG_PART_METRIC, SONG_PART_S) pair, so the kernels build in parallel;
// song_search.hip dispatches to the parts.
#include "song_search_impl.h"

#if !defined(SONG_PART_METRIC) || !defined(SONG_PART_S)
#error "compile with -DSONG_PART_METRIC=<metric> -DSONG_PART_S=<slots>"
#endif

#define SONG_PART_CAT2(a, b, c) launch_search_part_##a##_##b
#define SONG_PART_CAT(a, b) SONG_PART_CAT2(a, b, 0)

namespace song {

hipError_t SONG_PART_CAT(SONG_PART_METRIC, SONG_PART_S)(const SearchArgs& a, int G, int qch, const LaunchShape& shape,
                                                        hipStream_t st) {
    return launch_g<SONG_PART_METRIC, SONG_PART_S>(a, G, qch, shape, st);
}

}  // namespace song
