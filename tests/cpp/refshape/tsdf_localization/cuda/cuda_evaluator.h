// Shape of include/tsdf_localization/cuda/cuda_evaluator.h:22-39,94-107:
// CudaPoint and the Evaluator interface (two virtual evaluate overloads, no
// virtual destructor).
#pragma once
#include <stdexcept>
#include <vector>
#include <geometry_msgs/PoseWithCovariance.h>
#include <sensor_msgs/PointCloud2.h>
#include <tsdf_localization/cuda/cuda_sub_voxel_map.h>
#include <tsdf_localization/particle.h>
namespace tsdf_localization {
struct CudaPoint {
    FLOAT_T x, y, z;
    CudaPoint(FLOAT_T x = 0, FLOAT_T y = 0, FLOAT_T z = 0) : x(x), y(y), z(z) {}
};
class Evaluator {
public:
    virtual geometry_msgs::PoseWithCovariance evaluate(std::vector<Particle>& particles,
                                                       const sensor_msgs::PointCloud2& real_cloud,
                                                       FLOAT_T tf_matrix[16])
    {
        throw std::runtime_error("CUDA acceleration is not supported. Please install CUDA!");
    }
    virtual geometry_msgs::PoseWithCovariance evaluate(std::vector<Particle>& particles,
                                                       const std::vector<CudaPoint>& points, FLOAT_T tf_matrix[16])
    {
        throw std::runtime_error("CUDA acceleration is not supported. Please install CUDA!");
    }
};
}  // namespace tsdf_localization
