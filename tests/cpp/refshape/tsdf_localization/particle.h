// Shape of include/tsdf_localization/particle.h:26-55.
#pragma once
namespace tsdf_localization {
using FLOAT_T = float;
struct Particle { FLOAT_T first[6]; FLOAT_T second; };
}  // namespace tsdf_localization
