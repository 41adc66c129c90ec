// Shape of geometry_msgs/PoseWithCovariance (ROS): position, orientation and
// a fixed-size 36-entry covariance array.
#pragma once
#include <array>
namespace geometry_msgs {
struct Point { double x = 0, y = 0, z = 0; };
struct Quaternion { double x = 0, y = 0, z = 0, w = 0; };
struct Pose { Point position; Quaternion orientation; };
struct PoseWithCovariance { Pose pose; std::array<double, 36> covariance{}; };
}  // namespace geometry_msgs
