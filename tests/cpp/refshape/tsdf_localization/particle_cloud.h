// Shape of include/tsdf_localization/particle_cloud.h:46-259, the members a
// resampler uses.
#pragma once
#include <vector>
#include <tsdf_localization/particle.h>
namespace tsdf_localization {
class ParticleCloud {
public:
    Particle& operator[](unsigned int index) { return m_particles[index]; }
    std::size_t size() const { return m_particles.size(); }
    std::vector<Particle>& particles() { return m_particles; }
private:
    std::vector<Particle> m_particles;
};
}  // namespace tsdf_localization
