// song_server.h -- internal hooks the serving loop uses from the index runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/song.h"

namespace song {

void set_last_error(const std::string& msg);
void index_info(const song_index* ix, int* metric, int* device);
// Per-request budgets, launched on `stream` with the caller's queue counter
// (persistent mode) and planning workspace (learned entry / cost order,
// plan_ws_words(nq) int32); no synchronisation.  Result rows are k_stride
// long; k_limit (<= k_stride, 0 = k_stride) bounds the requests' k, which is
// what the options are checked against.
int launch_mixed_on_stream(song_index* ix, const void* d_q, int nq, const int32_t* d_ks, const int32_t* d_efs,
                           int k_stride, int ef_max, const song_search_opts* opts, int32_t* d_ids, float* d_dists,
                           int32_t* d_counter, int32_t* d_plan_ws, hipStream_t stream, int k_limit = 0);
int64_t plan_ws_words(int64_t nq);
// The checks a launch of these options makes, for requests up to k_max / ef_max
// (learned searches need k <= 64, ...): SONG_OK or the launch's error.
int check_search_opts(const song_index* ix, int32_t k_max, int32_t ef_max, const song_search_opts* opts);
// The resident serving grid over a request ring (see ResidentQueue); d_queue
// is the device copy of the ring descriptor.  No synchronisation.
int launch_resident_on_stream(song_index* ix, const void* d_q, int ring, const int32_t* d_ks, const int32_t* d_efs,
                              int k_stride, int ef_max, const song_search_opts* opts, int32_t* d_ids, float* d_dists,
                              const void* d_queue, hipStream_t stream);

}  // namespace song
