// C++ drop-in for the reference's GPU sensor update and systematic
// resampler, built on the C ABI in tsl.h.
//
//   HipEvaluator            replaces CudaEvaluator
//                           (include/tsdf_localization/cuda/cuda_evaluator.h:109-150,
//                           src/cuda/cuda_evaluator.cu:18-428): same
//                           constructor arguments, same evaluate() contract
//                           (weights overwritten with normalised weights,
//                           weighted-mean pose returned, std::runtime_error
//                           "No particle is valid!" on a zero weight sum).
//   HipSystematicResampler  replaces SystematicResampler
//                           (include/tsdf_localization/resampling/novel_resampling.h:38-73):
//                           same uniform draw from the base class's mt19937,
//                           same emitted particles (count may differ from M
//                           exactly as the reference's float recurrence does).
//
// Inside the reference tree define TSL_WITH_REFERENCE_HEADERS before
// including this file: the classes then derive from the reference's own
// Evaluator and Resampler and use its Particle, CudaPoint, ParticleCloud,
// CudaSubVoxelMap, geometry_msgs and sensor_msgs types.  Without it (no ROS
// available), minimal stand-ins with the same names, member names, layouts
// and virtual interfaces are declared here: Evaluator and Resampler have no
// virtual destructor, exactly as in the reference, so owners keep deleting
// through the concrete type (TSDFEvaluator's shared_ptr, constructed from
// the derived pointer, does).
#ifndef TSL_HIP_EVALUATOR_H
#define TSL_HIP_EVALUATOR_H

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../tsl.h"

#ifdef TSL_WITH_REFERENCE_HEADERS
#include <tsdf_localization/cuda/cuda_evaluator.h>
#include <tsdf_localization/cuda/cuda_sub_voxel_map.h>
#include <tsdf_localization/particle_cloud.h>
#include <tsdf_localization/resampling/resampler.h>
#else
#include <array>
#include <cstdint>

namespace geometry_msgs
{
struct Point { double x = 0, y = 0, z = 0; };
struct Quaternion { double x = 0, y = 0, z = 0, w = 1; };
struct Pose { Point position; Quaternion orientation; };
struct PoseWithCovariance { Pose pose; std::array<double, 36> covariance{}; };
}  // namespace geometry_msgs

// sensor_msgs/PointCloud2 and PointField: the fields a cloud reader uses
namespace sensor_msgs
{
struct PointField {
    enum : uint8_t { INT8 = 1, UINT8 = 2, INT16 = 3, UINT16 = 4, INT32 = 5, UINT32 = 6, FLOAT32 = 7, FLOAT64 = 8 };
    std::string name;
    uint32_t offset = 0;
    uint8_t datatype = 0;
    uint32_t count = 1;
};
struct PointCloud2 {
    uint32_t height = 1, width = 0;
    std::vector<PointField> fields;
    bool is_bigendian = false;
    uint32_t point_step = 0, row_step = 0;
    std::vector<uint8_t> data;
    bool is_dense = false;
};
}  // namespace sensor_msgs

namespace tsdf_localization
{
// particle.h:26-55
struct Particle {
    float first[6];
    float second;
};

// cuda_evaluator.h:22-39
struct CudaPoint {
    float x, y, z;
    CudaPoint(float x = 0, float y = 0, float z = 0) : x(x), y(y), z(z) {}
};

// cuda_evaluator.h:94-107
class Evaluator
{
public:
    virtual geometry_msgs::PoseWithCovariance evaluate(std::vector<Particle>&, const sensor_msgs::PointCloud2&,
                                                       float[16])
    {
        throw std::runtime_error("CUDA acceleration is not supported. Please install CUDA!");
    }

    virtual geometry_msgs::PoseWithCovariance evaluate(std::vector<Particle>&, const std::vector<CudaPoint>&,
                                                       float[16])
    {
        throw std::runtime_error("CUDA acceleration is not supported. Please install CUDA!");
    }
};

// particle_cloud.h:186-196 (the members a resampler uses)
class ParticleCloud
{
    std::vector<Particle> m_particles;

public:
    ParticleCloud() = default;
    explicit ParticleCloud(std::vector<Particle> p) : m_particles(std::move(p)) {}
    Particle& operator[](unsigned int index) { return m_particles[index]; }
    std::size_t size() const { return m_particles.size(); }
    std::vector<Particle>& particles() { return m_particles; }
};

// resampler.h:16-31
class Resampler
{
public:
    Resampler()
    {
        std::random_device device{};
        m_generator_ptr.reset(new std::mt19937(device()));
    }

    virtual void resample(ParticleCloud& particle_cloud) = 0;

protected:
    std::unique_ptr<std::mt19937> m_generator_ptr;
};

// cuda_sub_voxel_map.h:17-335, the host side only: constructor arithmetic,
// setData and the raw arrays the evaluator uploads.
template <typename DataT, typename IndexT>
class CudaSubVoxelMap;

template <>
class CudaSubVoxelMap<float, float>
{
public:
    using MapCoef = tsl_map_coef;

    CudaSubVoxelMap(float min_x, float min_y, float min_z, float max_x, float max_y, float max_z, float resolution,
                    float init_value = 0)
    {
        if (tsl_map_coef_init(min_x, min_y, min_z, max_x, max_y, max_z, resolution, init_value, &coef_) != TSL_OK)
            throw std::runtime_error(tsl_last_error());
    }

    // cuda_sub_voxel_map.tcc:171-231
    void setData(const std::vector<std::tuple<float, float, float, float>>& data)
    {
        if (data.empty()) return;
        std::vector<float> xyz(3 * data.size()), val(data.size());
        for (size_t i = 0; i < data.size(); ++i) {
            xyz[3 * i] = std::get<0>(data[i]);
            xyz[3 * i + 1] = std::get<1>(data[i]);
            xyz[3 * i + 2] = std::get<2>(data[i]);
            val[i] = std::get<3>(data[i]);
        }
        grid_occ_.assign(coef_.grid_occ_size, -1);
        int rc = tsl_map_set_data(&coef_, xyz.data(), val.data(), data.size(), grid_occ_.data(), nullptr, 0);
        if (rc != TSL_OK && rc != TSL_ERR_CAPACITY) throw std::runtime_error(tsl_last_error());
        data_.assign(coef_.data_size, coef_.init_value);
        rc = tsl_map_set_data(&coef_, xyz.data(), val.data(), data.size(), grid_occ_.data(), data_.data(),
                              data_.size());
        if (rc != TSL_OK) throw std::runtime_error(tsl_last_error());
    }

    MapCoef& coef() { return coef_; }
    int* rawGridOcc() { return grid_occ_.data(); }
    float* rawData() { return data_.data(); }
    size_t gridOccSize() const { return coef_.grid_occ_size; }
    size_t dataSize() const { return coef_.data_size; }

private:
    MapCoef coef_{};
    std::vector<int> grid_occ_;
    std::vector<float> data_;
};
}  // namespace tsdf_localization
#endif  // TSL_WITH_REFERENCE_HEADERS

namespace tsdf_localization
{

static_assert(sizeof(Particle) == sizeof(tsl_particle), "Particle layout");
static_assert(sizeof(CudaPoint) == sizeof(tsl_point), "CudaPoint layout");

namespace tsl_detail
{
inline void throw_on(int rc)
{
    if (rc == TSL_OK) return;
    // the reference's message for a zero weight sum (cuda_evaluator.cu:368)
    if (rc == TSL_ERR_NO_VALID_PARTICLE) throw std::runtime_error("No particle is valid!");
    throw std::runtime_error(tsl_last_error());
}

// Offset of a named field of a PointCloud2; PointCloud2ConstIterator throws
// the same message for a missing field (point_cloud2_iterator.h).
inline uint32_t cloud_field(const sensor_msgs::PointCloud2& cloud, const std::string& name)
{
    for (const auto& f : cloud.fields)
        if (f.name == name) return f.offset;
    throw std::runtime_error("Field " + name + " does not exist");
}

inline geometry_msgs::PoseWithCovariance to_pose(const tsl_pose& p)
{
    geometry_msgs::PoseWithCovariance out;
    out.pose.position.x = p.position[0];
    out.pose.position.y = p.position[1];
    out.pose.position.z = p.position[2];
    out.pose.orientation.x = p.orientation[0];
    out.pose.orientation.y = p.orientation[1];
    out.pose.orientation.z = p.orientation[2];
    out.pose.orientation.w = p.orientation[3];
    for (int i = 0; i < 36; ++i) out.covariance[i] = p.covariance[i];
    return out;
}
}  // namespace tsl_detail

class HipEvaluator : public Evaluator
{
public:
    // CudaEvaluator(map, per_point, a_hit, a_range, a_max, max_range)
    // (cuda_evaluator.h:112): uploads the map once.
    explicit HipEvaluator(CudaSubVoxelMap<float, float>& map, bool per_point = false, float a_hit = 0.9f,
                          float a_range = 0.1f, float a_max = 0.0f, float max_range = 100.0f)
    {
        tsl_map_coef c;
        static_assert(sizeof(c) == 120, "MapCoef layout");
        std::memcpy(&c, &map.coef(), sizeof(c));
        tsl_detail::throw_on(tsl_map_create(&c, map.rawGridOcc(), map.rawData(), &map_));
        const int rc = tsl_evaluator_create(map_, per_point ? 1 : 0, a_hit, a_range, a_max, max_range, &ev_);
        if (rc != TSL_OK) {
            tsl_map_destroy(map_);
            tsl_detail::throw_on(rc);
        }
    }

    // Map given as TSDF millimetres (one int16 per voxel, INT16_MIN = unset)
    // plus the likelihood parameters of createTSDFMap (map_util.h:17-154):
    // half the device bytes of the float map, identical voxel values.
    HipEvaluator(const tsl_map_coef& coef, const int32_t* grid_occ, const int16_t* tsdf_mm, int truncation_mm,
                 float sigma, bool per_point = false, float a_hit = 0.9f, float a_range = 0.1f, float a_max = 0.0f,
                 float max_range = 100.0f)
    {
        tsl_detail::throw_on(tsl_map_create_tsdf(&coef, grid_occ, tsdf_mm, truncation_mm, sigma, &map_));
        const int rc = tsl_evaluator_create(map_, per_point ? 1 : 0, a_hit, a_range, a_max, max_range, &ev_);
        if (rc != TSL_OK) {
            tsl_map_destroy(map_);
            tsl_detail::throw_on(rc);
        }
    }

    HipEvaluator(const HipEvaluator&) = delete;
    HipEvaluator& operator=(const HipEvaluator&) = delete;

    // The reference's Evaluator has no virtual destructor (cuda_evaluator.h:94-107).
    virtual ~HipEvaluator()
    {
        tsl_evaluator_destroy(ev_);
        tsl_map_destroy(map_);
    }

    // Evaluator::evaluate(particles, points, tf) (cuda_evaluator.h:102)
    geometry_msgs::PoseWithCovariance evaluate(std::vector<Particle>& particles, const std::vector<CudaPoint>& points,
                                               float tf_matrix[16]) override
    {
        tsl_pose p;
        tsl_detail::throw_on(tsl_evaluate(ev_, reinterpret_cast<tsl_particle*>(particles.data()), particles.size(),
                                          reinterpret_cast<const tsl_point*>(points.data()), points.size(), tf_matrix,
                                          &p));
        return tsl_detail::to_pose(p);
    }

    // Evaluator::evaluate(particles, cloud, tf) (cuda_evaluator.h:97-100,
    // src/cuda/cuda_evaluator.cu:78-116): the cloud's x, y, z floats and
    // its ring field read as int, as PointCloud2ConstIterator<float>(cloud,
    // "x") and <int>(cloud, "ring") read them (one point per point_step
    // bytes of data), then the 0.064 m centre reduction on the device and the
    // evaluation.  The reference's centres come in hash-set order; here
    // they are ordered by ring, then index (tsl_filter_set_scan_cloud).
    geometry_msgs::PoseWithCovariance evaluate(std::vector<Particle>& particles,
                                               const sensor_msgs::PointCloud2& cloud, float tf_matrix[16]) override
    {
        const uint32_t ox = tsl_detail::cloud_field(cloud, "x");
        const uint32_t oring = tsl_detail::cloud_field(cloud, "ring");
        const size_t step = cloud.point_step;
        const size_t n = step ? cloud.data.size() / step : 0;
        std::vector<tsl_point> pts(n);
        std::vector<int32_t> rings(n);
        for (size_t i = 0; i < n; ++i) {
            const uint8_t* rec = cloud.data.data() + i * step;
            if (ox + 12 <= step) std::memcpy(&pts[i], rec + ox, 12);
            if (oring + 4 <= step) std::memcpy(&rings[i], rec + oring, 4);
        }
        tsl_pose p;
        tsl_detail::throw_on(tsl_evaluate_cloud(ev_, reinterpret_cast<tsl_particle*>(particles.data()),
                                                particles.size(), pts.data(), rings.data(), n, tf_matrix, &p));
        return tsl_detail::to_pose(p);
    }

    // sin/cos of the particle angles (TSL_TRIG_*, default glibc's sinf/cosf
    // as on x86-64 with FMA3, the reference CPU evaluator's binding).
    void setTrig(int mode) { tsl_detail::throw_on(tsl_evaluator_set_trig(ev_, mode)); }

    tsl_evaluator* handle() { return ev_; }

private:
    tsl_map* map_ = nullptr;
    tsl_evaluator* ev_ = nullptr;
};

class HipSystematicResampler : public Resampler
{
public:
    // SystematicResampler::resample (novel_resampling.h:38-73): U drawn from
    // uniform_real_distribution<float>(0, 1/M) on the base class's
    // generator, then the device scan and draw.
    void resample(ParticleCloud& particle_cloud) override
    {
        const size_t M = particle_cloud.size();
        const double inverse_M = 1.0 / M;
        std::uniform_real_distribution<float> uniform_distribution(0.0, inverse_M);
        const float U = uniform_distribution(*m_generator_ptr);
        resample_with(particle_cloud, U);
    }

    // The same with the uniform supplied by the caller.
    static void resample_with(ParticleCloud& particle_cloud, float U)
    {
        std::vector<Particle>& in = particle_cloud.particles();
        const size_t M = in.size();
        if (M == 0) return;
        size_t cap = 0;
        tsl_detail::throw_on(tsl_resample_capacity(M, U, &cap));
        std::vector<Particle> out(cap);
        size_t n = 0;
        tsl_detail::throw_on(tsl_resample_systematic(reinterpret_cast<const tsl_particle*>(in.data()), M, U,
                                                     reinterpret_cast<tsl_particle*>(out.data()), cap, nullptr,
                                                     &n));
        out.resize(n);
        particle_cloud.particles() = std::move(out);
    }
};

}  // namespace tsdf_localization

#endif  // TSL_HIP_EVALUATOR_H
