// Drives the C++ drop-in classes (HipEvaluator, HipSystematicResampler,
// the CudaSubVoxelMap stand-in) the way mcl_3d drives the reference's:
// build a map with setData, evaluate a particle set against a scan, then
// resample it.  Inputs come from a binary file written by the test; outputs
// go to another binary file that the test checks against the oracle.
//
// usage: dropin_check <input.bin> <output.bin>
#include <tsdf_localization/hip_evaluator.h>

#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

using namespace tsdf_localization;

namespace {

struct Reader {
    FILE* f;
    template <typename T>
    T get()
    {
        T v;
        if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("short input");
        return v;
    }
    template <typename T>
    void get(T* p, size_t n)
    {
        if (n && std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input");
    }
};

template <typename T>
void put(FILE* f, const T* p, size_t n)
{
    if (n && std::fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("short write");
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input.bin> <output.bin>\n", argv[0]);
        return 2;
    }
    try {
        FILE* fi = std::fopen(argv[1], "rb");
        if (!fi) throw std::runtime_error("cannot open input");
        Reader r{fi};
        float box[8];
        r.get(box, 8);  // min xyz, max xyz, resolution, init value
        const uint64_t ncells = r.get<uint64_t>();
        std::vector<float> cells(4 * ncells);
        r.get(cells.data(), cells.size());
        const uint64_t P = r.get<uint64_t>();
        std::vector<Particle> particles(P);
        r.get(reinterpret_cast<float*>(particles.data()), 7 * P);
        const uint64_t N = r.get<uint64_t>();
        std::vector<CudaPoint> points(N);
        r.get(reinterpret_cast<float*>(points.data()), 3 * N);
        float tf[16];
        r.get(tf, 16);
        const uint32_t seed = r.get<uint32_t>();
        float params[4];
        r.get(params, 4);  // a_hit, a_range, a_max, max_range
        std::fclose(fi);

        CudaSubVoxelMap<float, float> map(box[0], box[1], box[2], box[3], box[4], box[5], box[6], box[7]);
        std::vector<std::tuple<float, float, float, float>> data;
        data.reserve(ncells);
        for (uint64_t i = 0; i < ncells; ++i)
            data.emplace_back(cells[4 * i], cells[4 * i + 1], cells[4 * i + 2], cells[4 * i + 3]);
        map.setData(data);

        // TSDFEvaluator's constructor, with HipEvaluator in place of CudaEvaluator
        std::shared_ptr<Evaluator> evaluator(new HipEvaluator(map, false, params[0], params[1], params[2], params[3]));
        const geometry_msgs::PoseWithCovariance pose = evaluator->evaluate(particles, points, tf);

        // SystematicResampler with a seeded generator; a second generator with
        // the same seed reproduces the uniform it draws, for the checker
        struct SeededResampler : HipSystematicResampler {
            explicit SeededResampler(uint32_t s) { m_generator_ptr.reset(new std::mt19937(s)); }
        } resampler(seed);
        std::mt19937 twin(seed);
        std::uniform_real_distribution<float> dist(0.0, 1.0 / P);
        const float U = dist(twin);
        ParticleCloud cloud(particles);
        resampler.resample(cloud);

        // Evaluator::evaluate(particles, PointCloud2, tf): the same points as
        // a cloud of 16-byte records (x, y, z float32, ring uint16 at 12, two
        // zero bytes), ring = point index % 16, evaluated on the initial weights
        sensor_msgs::PointCloud2 pc2;
        pc2.height = 1;
        pc2.width = (uint32_t)N;
        pc2.point_step = 16;
        pc2.row_step = 16 * (uint32_t)N;
        const char* names[4] = {"x", "y", "z", "ring"};
        for (int k = 0; k < 4; ++k) {
            sensor_msgs::PointField fld;
            fld.name = names[k];
            fld.offset = 4 * k;
            fld.datatype = k < 3 ? sensor_msgs::PointField::FLOAT32 : sensor_msgs::PointField::UINT16;
            fld.count = 1;
            pc2.fields.push_back(fld);
        }
        pc2.data.assign(16 * N, 0);
        for (uint64_t i = 0; i < N; ++i) {
            std::memcpy(pc2.data.data() + 16 * i, &points[i], 12);
            const uint16_t ring = (uint16_t)(i % 16);
            std::memcpy(pc2.data.data() + 16 * i + 12, &ring, 2);
        }
        std::vector<Particle> cloud_particles(P);
        {
            FILE* fr = std::fopen(argv[1], "rb");
            Reader rr{fr};
            float skip[8];
            rr.get(skip, 8);
            const uint64_t nc = rr.get<uint64_t>();
            std::vector<float> tmp(4 * nc);
            rr.get(tmp.data(), tmp.size());
            (void)rr.get<uint64_t>();
            rr.get(reinterpret_cast<float*>(cloud_particles.data()), 7 * P);
            std::fclose(fr);
        }
        const geometry_msgs::PoseWithCovariance cloud_pose = evaluator->evaluate(cloud_particles, pc2, tf);

        // error path: every weight zero (no range term, points in unknown space)
        int error_ok = 0;
        {
            HipEvaluator zero(map, false, 1.0f, 0.0f, 0.0f, params[3]);
            std::vector<CudaPoint> far(8, CudaPoint(1e4f, 1e4f, 1e4f));
            std::vector<Particle> ps(particles.begin(), particles.begin() + (P < 4 ? P : 4));
            try {
                zero.evaluate(ps, far, tf);
            } catch (const std::runtime_error& e) {
                error_ok = std::string(e.what()) == "No particle is valid!" ? 1 : -1;
            }
        }

        FILE* fo = std::fopen(argv[2], "wb");
        if (!fo) throw std::runtime_error("cannot open output");
        const double pv[7] = {pose.pose.position.x, pose.pose.position.y, pose.pose.position.z,
                              pose.pose.orientation.x, pose.pose.orientation.y, pose.pose.orientation.z,
                              pose.pose.orientation.w};
        put(fo, pv, 7);
        put(fo, pose.covariance.data(), 36);
        std::vector<float> w(P);
        for (uint64_t i = 0; i < P; ++i) w[i] = particles[i].second;
        put(fo, w.data(), P);
        put(fo, &U, 1);
        const uint64_t n = cloud.size();
        put(fo, &n, 1);
        put(fo, reinterpret_cast<const float*>(cloud.particles().data()), 7 * n);
        const int32_t e = error_ok;
        put(fo, &e, 1);
        for (uint64_t i = 0; i < P; ++i) w[i] = cloud_particles[i].second;
        put(fo, w.data(), P);
        const double cp[3] = {cloud_pose.pose.position.x, cloud_pose.pose.position.y, cloud_pose.pose.position.z};
        put(fo, cp, 3);
        std::fclose(fo);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "dropin_check: %s\n", ex.what());
        return 1;
    }
    return 0;
}
