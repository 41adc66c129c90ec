// Motion update on the device: ParticleCloud::apply_model
// (src/particle_cloud.cpp:499-617) over a device-resident particle set, with
// the six normal draws of every particle taken from a counter-based stream
// (seed, particle index, draw index) instead of a sequential std::mt19937,
// so the particles of one filter move the same way whether the filter is
// updated alone or inside a batch.
#include "tsl_internal.h"

namespace tsl {

namespace {

__global__ void k_motion_m(const Member* __restrict__ members, const tsl_motion* __restrict__ motions)
{
    const Member& m = members[blockIdx.y];
    const tsl_motion& mo = motions[blockIdx.y];
    const uint32_t n = m.st->count;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tsl_particle p = m.particles[i];
    double d[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = motion_normal(mo.seed, i, (uint32_t)k) * mo.stddev[k] + mo.mean[k];
    float out[6];
    apply_motion(p.first, d, out, m.trig);
#pragma unroll
    for (int k = 0; k < 6; ++k) p.first[k] = out[k];
    m.particles[i] = p;
}

}  // namespace

int launch_motion_members(const Member* d_members, const tsl_motion* d_motions, uint32_t n, uint32_t grid_cap,
                          hipStream_t stream)
{
    if (n == 0) return TSL_OK;
    const unsigned gx = (grid_cap + 255) / 256 ? (grid_cap + 255) / 256 : 1;
    hipLaunchKernelGGL(k_motion_m, dim3(gx, n), dim3(256), 0, stream, d_members, d_motions);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

}  // namespace tsl

using namespace tsl;

extern "C" {

int tsl_motion_from_delta(float dx, float dy, float dz, float roll, float pitch, float yaw, float lin_base_noise,
                          float ang_base_noise, const float a[12], uint64_t seed, tsl_motion* out)
{
    if (!a || !out) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    // particle_cloud.cpp:471-492: squares and spreads in float, then the
    // double parameters of std::normal_distribution<>
    const float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    const float roll2 = roll * roll, pitch2 = pitch * pitch, yaw2 = yaw * yaw;
    out->mean[0] = dx;
    out->mean[1] = dy;
    out->mean[2] = dz;
    out->mean[3] = roll;
    out->mean[4] = pitch;
    out->mean[5] = yaw;
    out->stddev[0] = a[0] * dx2 + a[1] * yaw2 + lin_base_noise;
    out->stddev[1] = a[2] * dy2 + a[3] * yaw2 + lin_base_noise;
    out->stddev[2] = a[4] * dz2 + a[5] * yaw2 + lin_base_noise;
    out->stddev[3] = a[6] * dx2 + a[7] * roll2 + ang_base_noise;
    out->stddev[4] = a[8] * dx2 + a[9] * pitch2 + ang_base_noise;
    out->stddev[5] = a[10] * dx2 + a[11] * yaw2 + ang_base_noise;
    out->seed = seed;
    return TSL_OK;
}

int tsl_motion_from_odom(float linear_velocity, float angular_velocity, float time_diff, const float a[12],
                         uint64_t seed, tsl_motion* out)
{
    if (!a || !out) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    // particle_cloud.cpp:181-200
    const float d = linear_velocity * time_diff;
    const float d2 = d * d;
    const float theta = angular_velocity * time_diff;
    const float theta2 = theta * theta;
    for (int k = 0; k < 6; ++k) out->mean[k] = 0.0;
    out->mean[0] = d;
    out->mean[5] = theta;
    for (int k = 0; k < 6; ++k) out->stddev[k] = a[2 * k] * d2 + a[2 * k + 1] * theta2;
    out->seed = seed;
    return TSL_OK;
}

int tsl_filter_motion_async(tsl_filter* f, const tsl_motion* motion, void* stream)
{
    if (!f || !motion) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    TSL_TRY(f->motion.reserve(sizeof(tsl_motion)));
    // the parameters travel through the stream's own copy (pinned staging
    // kept until that copy is done)
    if (f->motion_pending) {
        TSL_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(f->motion_copied)));
        f->motion_pending = false;
    }
    if (!f->motion_host) {
        if (hipHostMalloc(&f->motion_host, sizeof(tsl_motion), hipHostMallocDefault) != hipSuccess)
            return set_error(TSL_ERR_OUT_OF_MEMORY, "hipHostMalloc failed");
        hipEvent_t e;
        TSL_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        f->motion_copied = e;
    }
    *static_cast<tsl_motion*>(f->motion_host) = *motion;
    TSL_HIP_CHECK(hipMemcpyAsync(f->motion.ptr, f->motion_host, sizeof(tsl_motion), hipMemcpyHostToDevice, s));
    TSL_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(f->motion_copied), s));
    f->motion_pending = true;
    Member* dm;
    TSL_TRY(filter_member(f, &dm));
    return launch_motion_members(dm, f->motion.as<tsl_motion>(), 1, (uint32_t)f->count_bound, s);
}

}  // extern "C"
