// The integration snippets, compiled against declarations with the shape of
// the reference's headers (tests/cpp/refshape): TSDFEvaluator's construction
// of its evaluator (tsdf_evaluator.h:43,78), both evaluate overloads
// through the base pointer (tsdf_evaluator.cpp:80-83, cuda_evaluator.h:97-102),
// and mcl_3d's resampler (mcl_3d.cpp:105,252,425).  Compile-only.
#define TSL_WITH_REFERENCE_HEADERS
#include <tsdf_localization/hip_evaluator.h>

using namespace tsdf_localization;

geometry_msgs::PoseWithCovariance drive(const std::shared_ptr<CudaSubVoxelMap<FLOAT_T, FLOAT_T>>& map_ptr,
                                        std::vector<Particle>& particles, const std::vector<CudaPoint>& points,
                                        const sensor_msgs::PointCloud2& cloud, ParticleCloud& particle_cloud)
{
    std::shared_ptr<Evaluator> cuda_evaluator_;
    cuda_evaluator_.reset(new HipEvaluator(*map_ptr, false, 0.9f, 0.1f, 0.0f, 100.0f));
    FLOAT_T tf_matrix[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    geometry_msgs::PoseWithCovariance pose = cuda_evaluator_->evaluate(particles, points, tf_matrix);
    pose = cuda_evaluator_->evaluate(particles, cloud, tf_matrix);
    std::unique_ptr<Resampler> resampler_ptr_;
    resampler_ptr_.reset(new HipSystematicResampler());
    resampler_ptr_->resample(particle_cloud);
    return pose;
}
