// Launchers of one evaluation tile shape: TSL_TILE_K particles per
// wavefront, TSL_TILE_P points per lane.  Compiled once per shape (the build
// passes -DTSL_TILE_K=<1|2> -DTSL_TILE_P=<1|2|4>), each object holding that
// shape's instantiations of k_eval and k_eval_batch.
#include "eval_kernels.h"

#if !defined(TSL_TILE_K) || !defined(TSL_TILE_P)
#error "compile with -DTSL_TILE_K=<1|2> -DTSL_TILE_P=<1|2|4>"
#endif

#define TSL_TILE_NAME2(base, k, p) base##_##k##_##p
#define TSL_TILE_NAME(base, k, p) TSL_TILE_NAME2(base, k, p)

namespace tsl {

int TSL_TILE_NAME(launch_eval_tile, TSL_TILE_K, TSL_TILE_P)(int mode, bool u24, bool brick, bool rank, dim3 grid,
                                                            size_t lds, hipStream_t stream, const tsl_filter* f,
                                                            const Tile& tile, size_t N, float* partial,
                                                            void* trace_out)
{
#define TSL_EVAL_FLAGS u24, brick, rank, grid, lds, stream, f, tile, N, partial, trace_out
    if (mode == 0) return launch_eval_flags<TSL_TILE_K, TSL_TILE_P, 0>(TSL_EVAL_FLAGS);
#if TSL_TILE_P == 1
    if (mode == 1) return launch_eval_flags<TSL_TILE_K, 1, 1>(TSL_EVAL_FLAGS);
    if (mode == 2) return launch_eval_flags<TSL_TILE_K, 1, 2>(TSL_EVAL_FLAGS);
#endif
#undef TSL_EVAL_FLAGS
    return set_error(TSL_ERR_INVALID_ARGUMENT, "evaluation mode %d with %d points per lane", mode, TSL_TILE_P);
}

int TSL_TILE_NAME(launch_eval_batch_tile, TSL_TILE_K, TSL_TILE_P)(const BatchLaunch& b, hipStream_t stream,
                                                                  uint32_t* grid_out)
{
    return launch_eval_batch_flags<TSL_TILE_K, TSL_TILE_P>(b, stream, grid_out);
}

}  // namespace tsl
