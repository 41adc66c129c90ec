// Shape of sensor_msgs/PointCloud2 and PointField (ROS), the members a cloud
// reader uses.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
namespace sensor_msgs {
struct PointField { std::string name; uint32_t offset = 0; uint8_t datatype = 0; uint32_t count = 0; };
struct PointCloud2 {
    uint32_t height = 0, width = 0;
    std::vector<PointField> fields;
    uint8_t is_bigendian = 0;
    uint32_t point_step = 0, row_step = 0;
    std::vector<uint8_t> data;
    uint8_t is_dense = 0;
};
}  // namespace sensor_msgs
