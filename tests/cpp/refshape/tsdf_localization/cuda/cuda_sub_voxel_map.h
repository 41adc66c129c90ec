// Shape of include/tsdf_localization/cuda/cuda_sub_voxel_map.h:13-50,141,166,271:
// the MapCoef layout and the raw accessors CudaEvaluator reads.
#pragma once
#include <cstddef>
#include <vector>
namespace tsdf_localization {
using OCC_T = int;
template <typename DataT, typename IndexT>
class CudaSubVoxelMap {
public:
    struct MapCoef {
        size_t dim_x_, dim_y_, dim_z_;
        IndexT min_x_, min_y_, min_z_;
        IndexT max_x_, max_y_, max_z_;
        IndexT resolution_;
        DataT init_value_;
        size_t up_dim_x_, up_dim_y_, up_dim_z_, up_dim_2_;
        size_t sub_dim_, sub_dim_2_;
        size_t grid_occ_size_, data_size_;
    };
    OCC_T* rawGridOcc() { return grid_occ_.data(); }
    DataT* rawData() { return data_.data(); }
    MapCoef& coef() { return coef_; }
private:
    MapCoef coef_{};
    std::vector<OCC_T> grid_occ_;
    std::vector<DataT> data_;
};
}  // namespace tsdf_localization
