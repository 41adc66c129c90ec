// Shape of include/tsdf_localization/resampling/resampler.h:16-31 (no
// virtual destructor).
#pragma once
#include <memory>
#include <random>
#include <tsdf_localization/particle_cloud.h>
namespace tsdf_localization {
class Resampler {
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
}  // namespace tsdf_localization
