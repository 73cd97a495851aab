// kfusion.hpp — C++ class surface of the per-frame path, over the C ABI in
// include/kfusion.h.  It mirrors the reference's classes and free functions
// (isikmustafa/kinectfusion-cuda, project/src) so a caller of the reference's
// per-frame path can switch by changing includes:
//
//   reference                                 here
//   CudaGridMap (cuda_grid_map.h:13-29)       kfusion::GridMap (alias CudaGridMap)
//   DepthMap (depth_map.h:8-15)               kfusion::DepthMap
//   GridMapPyramid<T> (grid_map_pyramid.h)    kfusion::GridMapPyramid<T>
//   Sensor (sensor.h:7-53)                    kfusion::Sensor
//   RigidTransform3D (rigid_transform_3d.h)   kfusion::RigidTransform3D
//   Voxel / VoxelGrid (voxel_grid.h:8-61)     kfusion::Voxel / kfusion::VoxelGrid
//   kernel::* (measurement.cuh, tsdf.cuh,     kfusion::kernel::* (same names, same
//     raycast.cuh, icp.cuh)                     "returns elapsed ms" contract)
//   IcpConfig / ICP (icp.h:11-53)             kfusion::IcpConfig / kfusion::ICP
//   LinearLeastSquares (linear_least_squares.h) kfusion::LinearLeastSquares
//
// Differences a caller sees: maps live in plain device buffers (data()) instead
// of cudaArray surfaces; matrices are glm-layout std::arrays (Mat3 / Mat4, column
// major); copying a map or a volume is a compile error (the reference's implicit
// copies would double-free).  Everything throws std::runtime_error on failure,
// with the reference's messages where the reference throws.
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../kfusion.h"

namespace kfusion {

using Mat3 = std::array<float, 9>;   // glm::mat3, column-major
using Mat4 = std::array<float, 16>;  // glm::mat4, column-major
using Vec3 = std::array<float, 3>;

inline Mat4 identity4()
{
    Mat4 m{};
    m[0] = m[5] = m[10] = m[15] = 1.0f;
    return m;
}

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

inline void check(int status, const char* what)
{
    if (status == KF_OK) return;
    std::string detail = kf_last_error();
    throw Error(status, std::string(what) + ": " + kf_status_string(status) + (detail.empty() ? "" : " (" + detail + ")"));
}

// Channel formats the reference creates with cudaCreateChannelDesc
// (kinect_fusion.h:51-53): 16-bit raw depth, 32-bit depth, 4x32-bit vectors.
enum class MapFormat { Raw16 = 2, Depth32 = 4, Vector4 = 16 };

// ------------------------------------------------------------------ maps
class GridMap {
public:
    GridMap(unsigned int width, unsigned int height, MapFormat format)
        : m_width(width), m_height(height), m_format(format)
    {
        check(kf_malloc(&m_data, bytes()), "CudaGridMap");
    }
    ~GridMap() { kf_free(m_data); }
    GridMap(const GridMap&) = delete;
    GridMap& operator=(const GridMap&) = delete;
    GridMap(GridMap&& o) noexcept : m_data(o.m_data), m_width(o.m_width), m_height(o.m_height), m_format(o.m_format)
    {
        o.m_data = nullptr;
    }

    std::array<unsigned int, 2> getGridDims() const { return {m_width, m_height}; }
    MapFormat getChannelDescription() const { return m_format; }
    void* data() const { return m_data; }
    float* floats() const { return static_cast<float*>(m_data); }
    size_t bytes() const { return (size_t)m_width * m_height * (size_t)m_format; }
    void upload(const void* host) { check(kf_memcpy_h2d(m_data, host, bytes(), nullptr), "upload"); }
    void download(void* host) const { check(kf_memcpy_d2h(host, m_data, bytes(), nullptr), "download"); }

protected:
    void* m_data = nullptr;
    unsigned int m_width, m_height;
    MapFormat m_format;
};
using CudaGridMap = GridMap;

class DepthMap : public GridMap {
public:
    DepthMap(unsigned int width, unsigned int height, MapFormat format = MapFormat::Raw16)
        : GridMap(width, height, format)
    {}
    // depth_map.cpp:30-35: copies one frame of the map's format from host memory
    void update(const void* data_ptr) { upload(data_ptr); }
};

template <class T>
class GridMapPyramid {
public:
    GridMapPyramid(unsigned int width, unsigned int height, unsigned int n_layers, MapFormat format)
        : m_base_width(width), m_base_height(height), m_n_layers(n_layers), m_format(format)
    {
        // grid_map_pyramid.h:63-72
        const unsigned int e = 1u << (n_layers - 1);
        if (width % e != 0 || height % e != 0)
            throw std::runtime_error{"Can't construct pyramid. Dimensions must be divisible by 2^(m_n_layers - 1)."};
        for (unsigned int i = 0; i < n_layers; ++i)
            m_pyramid.push_back(std::make_unique<T>(width >> i, height >> i, format));
    }
    T& operator[](int i) { return *m_pyramid[i]; }
    const T& operator[](int i) const { return *m_pyramid[i]; }
    MapFormat getChannelDescription() const { return m_format; }
    unsigned int getBaseWidth() const { return m_base_width; }
    unsigned int getBaseHeight() const { return m_base_height; }
    unsigned int size() const { return m_n_layers; }

private:
    std::vector<std::unique_ptr<T>> m_pyramid;
    unsigned int m_base_width, m_base_height, m_n_layers;
    MapFormat m_format;
};

// ------------------------------------------------------------------ sensor / pose
class Sensor {
public:
    // Sensor(float fov) (sensor.h:10-20): fx = fy = fov, c = (319.5, 239.5)
    explicit Sensor(float fov) : Sensor(fov, fov, 319.5f, 239.5f) {}
    Sensor(float fx, float fy, float cx, float cy) { check(kf_sensor_init(&m_s, fx, fy, cx, cy), "Sensor"); }

    Mat4 getPose() const { return to4(m_s.pose); }
    Mat4 getInversePose() const { return to4(m_s.inv_pose); }
    void setPose(const Mat4& pose) { check(kf_sensor_set_pose(&m_s, pose.data()), "Sensor::setPose"); }
    Mat3 getIntr(int level) const { return to3(m_s.intr[level]); }
    Mat3 getInverseIntr(int level) const { return to3(m_s.inv_intr[level]); }
    Vec3 getPosition() const { return {m_s.pose[12], m_s.pose[13], m_s.pose[14]}; }
    const kf_sensor& raw() const { return m_s; }

private:
    static Mat4 to4(const float* p)
    {
        Mat4 m;
        for (int i = 0; i < 16; ++i) m[i] = p[i];
        return m;
    }
    static Mat3 to3(const float* p)
    {
        Mat3 m;
        for (int i = 0; i < 9; ++i) m[i] = p[i];
        return m;
    }
    kf_sensor m_s{};
};

struct RigidTransform3D {
    Mat3 rot_mat{};
    Vec3 transl_vec{};

    RigidTransform3D() = default;
    explicit RigidTransform3D(const Mat4& t)
        : rot_mat{t[0], t[1], t[2], t[4], t[5], t[6], t[8], t[9], t[10]}, transl_vec{t[12], t[13], t[14]}
    {}
    Mat4 getTransformation() const
    {
        Mat4 p = identity4();
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) p[c * 4 + r] = rot_mat[c * 3 + r];
        for (int r = 0; r < 3; ++r) p[12 + r] = transl_vec[r];
        return p;
    }
};

// ------------------------------------------------------------------ volume
struct Voxel {
    float f{2.0f};  // voxel_grid.h:10-16
    float w{0.0f};
};

class VoxelGrid {
public:
    // VoxelGrid(width, n, mue_scale) (voxel_grid.h:43-44): resolution = width / (n - 1), mue = scale * resolution
    VoxelGrid(float total_width_in_meters, int n, float mue_scale, int format = KF_VOXEL_F32)
    {
        check(kf_volume_create_scaled(n, total_width_in_meters, mue_scale, format, &m_v), "VoxelGrid");
    }
    // explicit truncation distance in metres
    static VoxelGrid withMue(float total_width_in_meters, int n, float mue_m, int format = KF_VOXEL_F32)
    {
        return VoxelGrid(total_width_in_meters, n, mue_m, format, 0);
    }
    ~VoxelGrid() { kf_volume_destroy(m_v); }
    VoxelGrid(const VoxelGrid&) = delete;
    VoxelGrid& operator=(const VoxelGrid&) = delete;
    VoxelGrid(VoxelGrid&& o) noexcept : m_v(o.m_v) { o.m_v = nullptr; }

    kf_volume* handle() const { return m_v; }
    int n() const { int v; kf_volume_info(m_v, &v, nullptr, nullptr, nullptr, nullptr); return v; }
    float resolution() const { float r; kf_volume_info(m_v, nullptr, nullptr, &r, nullptr, nullptr); return r; }
    float mue() const { float m; kf_volume_info(m_v, nullptr, nullptr, nullptr, &m, nullptr); return m; }
    void saveVoxelGrid(const std::string& file_name) const { check(kf_volume_save(m_v, file_name.c_str()), "saveVoxelGrid"); }
    void loadVoxelGrid(const std::string& file_name) { check(kf_volume_load(m_v, file_name.c_str()), "loadVoxelGrid"); }
    // n^3 (f, w) pairs, x fastest (voxel_grid.cpp:43-80 layout)
    std::vector<float> download() const
    {
        const int k = n();
        std::vector<float> fw((size_t)k * k * k * 2);
        check(kf_volume_download(m_v, fw.data()), "VoxelGrid download");
        return fw;
    }

private:
    VoxelGrid(float width, int n, float mue_m, int format, int) { check(kf_volume_create(n, width, mue_m, format, &m_v), "VoxelGrid"); }
    kf_volume* m_v = nullptr;
};

// ------------------------------------------------------------------ kernels
// Every wrapper blocks and returns its elapsed milliseconds (cuda_event.cpp:14-30).
namespace kernel {

constexpr float cMinDistance = 0.4f;  // tsdf.cuh:11
constexpr float cMaxDistance = 6.0f;  // tsdf.cuh:12

inline void same_dims(const GridMap& a, const GridMap& b, const char* msg)
{
    if (a.getGridDims() != b.getGridDims()) throw std::runtime_error(msg);
}

// measurement.cu:139-158
inline float convertToDepthMeters(const DepthMap& raw, const GridMap& out, float scale, bool unmirror)
{
    same_dims(raw, out, "convertToDepthMeters: input and output surface objects are not of same size!");
    float ms = 0;
    auto d = raw.getGridDims();
    check(kf_convert_to_depth_meters(static_cast<const uint16_t*>(raw.data()), out.floats(), d[0], d[1], scale,
                                     unmirror, nullptr, &ms), "convertToDepthMeters");
    return ms;
}
// measurement.cu:160-179
inline float applyBilateralFilter(const GridMap& in, const GridMap& out)
{
    same_dims(in, out, "applyBilateralFilter: input and output surface objects are not of same size!");
    float ms = 0;
    auto d = in.getGridDims();
    check(kf_apply_bilateral_filter(in.floats(), out.floats(), d[0], d[1], nullptr, &ms), "applyBilateralFilter");
    return ms;
}
// measurement.cu:181-200
inline float downSample(const GridMap& in, const GridMap& out)
{
    auto di = in.getGridDims(), dout = out.getGridDims();
    if (di[0] / dout[0] != 2 || di[1] / dout[1] != 2)
        throw std::runtime_error("downSample: output has to be half size of the input surface object!");
    float ms = 0;
    check(kf_down_sample(in.floats(), di[0], di[1], out.floats(), dout[0], dout[1], nullptr, &ms), "downSample");
    return ms;
}
// measurement.cu:202-221
inline float createVertexMap(const GridMap& depth, const GridMap& vmap, const Mat3& inv_cam_k)
{
    same_dims(depth, vmap, "createVertexMap: input and output surface objects are not of same size!");
    float ms = 0;
    auto d = depth.getGridDims();
    check(kf_create_vertex_map(depth.floats(), vmap.floats(), d[0], d[1], inv_cam_k.data(), nullptr, &ms),
          "createVertexMap");
    return ms;
}
// measurement.cu:223-242
inline float createNormalMap(const GridMap& vmap, const GridMap& nmap)
{
    same_dims(vmap, nmap, "computeNormalMap: input and output surface objects are not of same size!");
    float ms = 0;
    auto d = vmap.getGridDims();
    check(kf_create_normal_map(vmap.floats(), nmap.floats(), d[0], d[1], nullptr, &ms), "createNormalMap");
    return ms;
}
// tsdf.cu:79-90
inline float initializeGrid(VoxelGrid& grid, const Voxel& value)
{
    float ms = 0;
    check(kf_initialize_grid(grid.handle(), value.f, value.w, nullptr, &ms), "initializeGrid");
    return ms;
}
// tsdf.cu:92-103
inline float fuse(const GridMap& depth_m, VoxelGrid& grid, const Sensor& sensor)
{
    float ms = 0;
    auto d = depth_m.getGridDims();
    check(kf_fuse(depth_m.floats(), d[0], d[1], grid.handle(), &sensor.raw(), nullptr, &ms), "fuse");
    return ms;
}
// raycast.cu:149-169
inline float raycast(const VoxelGrid& grid, const Sensor& sensor, const GridMap& vmap, const GridMap& nmap, int level)
{
    same_dims(vmap, nmap, "raycast: vertex and normal surface objects are not of same size!");
    float ms = 0;
    auto d = vmap.getGridDims();
    check(kf_raycast(grid.handle(), &sensor.raw(), vmap.floats(), nmap.floats(), d[0], d[1], level, nullptr, &ms),
          "raycast");
    return ms;
}
// icp.cu:86-106
inline float constructIcpResiduals(const GridMap& vmap, const GridMap& tvmap, const GridMap& tnmap,
                                   const Mat3& prev_rot, const Vec3& prev_transl, const Mat3& cur_rot,
                                   const Vec3& cur_transl, const Mat3& intrinsics, float distance_thresh,
                                   float angle_thresh, float* mat_a, float* vec_b)
{
    float ms = 0;
    auto d = vmap.getGridDims();
    check(kf_construct_icp_residuals(vmap.floats(), tvmap.floats(), tnmap.floats(), d[0], d[1], prev_rot.data(),
                                     prev_transl.data(), cur_rot.data(), cur_transl.data(), intrinsics.data(),
                                     distance_thresh, angle_thresh, mat_a, vec_b, nullptr, &ms),
          "constructIcpResiduals");
    return ms;
}

}  // namespace kernel

// ------------------------------------------------------------------ ICP
class LinearLeastSquares {
public:
    // linear_least_squares.cpp:23-44: A (n x 6 row-major) and b on the device
    std::array<float, 6> solve(const float* mat_a_device, const float* vec_b_device, unsigned int n_equations)
    {
        std::array<float, 6> x{};
        check(kf_solve_normal_equations(mat_a_device, vec_b_device, n_equations, x.data(), nullptr, nullptr),
              "LinearLeastSquares::solve");
        return x;
    }
};

struct IcpConfig {  // icp.h:11-20
    std::vector<unsigned int> iters_per_layer;
    unsigned int width;
    unsigned int height;
    float distance_thresh;
    float angle_thresh;
    float iteration_stop_thresh_angle = 0.0f;
    float iteration_stop_thresh_distance = 0.0f;
    bool compensated = false;  // Kahan-compensated FP32 sums (precision form)
};

class ICP {
public:
    explicit ICP(const IcpConfig& config) : m_config(config)
    {
        kf_icp_config c{};
        c.levels = (int)config.iters_per_layer.size();
        if (c.levels < 1 || c.levels > 3) throw std::runtime_error("ICP: 1 to 3 pyramid layers");
        for (int l = 0; l < c.levels; ++l) c.iters_per_layer[l] = (int)config.iters_per_layer[l];
        c.width = (int)config.width;
        c.height = (int)config.height;
        c.distance_thresh = config.distance_thresh;
        c.angle_thresh = config.angle_thresh;
        c.iteration_stop_thresh_angle = config.iteration_stop_thresh_angle;
        c.iteration_stop_thresh_distance = config.iteration_stop_thresh_distance;
        c.compensated = config.compensated ? 1 : 0;
        check(kf_icp_create(&c, &m_icp), "ICP");
    }
    ~ICP() { kf_icp_destroy(m_icp); }
    ICP(const ICP&) = delete;
    ICP& operator=(const ICP&) = delete;

    // icp.cpp:25-58; throws like icp.cpp:67-73 when the update is not finite
    RigidTransform3D computePose(GridMapPyramid<GridMap>& vertex_pyramid, GridMapPyramid<GridMap>& target_vertex_pyramid,
                                 GridMapPyramid<GridMap>& target_normal_pyramid, const Sensor& sensor)
    {
        const int L = (int)m_config.iters_per_layer.size();
        const float* vp[3] = {nullptr, nullptr, nullptr};
        const float* tv[3] = {nullptr, nullptr, nullptr};
        const float* tn[3] = {nullptr, nullptr, nullptr};
        for (int l = 0; l < L; ++l) {
            vp[l] = vertex_pyramid[l].floats();
            tv[l] = target_vertex_pyramid[l].floats();
            tn[l] = target_normal_pyramid[l].floats();
        }
        Mat4 out = identity4();
        int st = kf_icp_compute_pose(m_icp, vp, tv, tn, &sensor.raw(), out.data(), nullptr, nullptr);
        if (st == KF_ERR_NONFINITE || st == KF_ERR_NOT_PD)
            throw Error(st, "Error: an element in the result vector is not finite or NAN.");
        check(st, "ICP::computePose");
        return RigidTransform3D(out);
    }
    std::array<float, 2> getExecutionTimes() const
    {
        std::array<float, 2> t{};
        check(kf_icp_execution_times(m_icp, t.data()), "ICP::getExecutionTimes");
        return t;
    }

private:
    IcpConfig m_config;
    kf_icp* m_icp = nullptr;
};

// ------------------------------------------------------------------ per frame
// KinectFusion's per-frame loop (kinect_fusion.cpp:44-183) on one volume.
struct KinectFusionConfig {
    unsigned int frame_width = 640, frame_height = 480;
    float depth_scale = 1.0f / 1000.0f;
    float fx = 525.0f, fy = 525.0f, cx = 319.5f, cy = 239.5f;
    bool unmirror = false;
    int volume_n = 256;
    float volume_size = 3.0f;
    float mue = 0.03f;
    int voxel_format = KF_VOXEL_F32;
    Mat4 initial_pose = identity4();
    // the visualisation volume (kinect_fusion.h:32-33, 0 = none)
    int visualization_volume_n = 0;
    float visualization_volume_size = 4.0f;
    float visualization_mue = 0.03f;
};

class KinectFusion {
public:
    KinectFusion(const KinectFusionConfig& kf, const IcpConfig& icp)
    {
        kf_pipeline_config c{};
        c.frame_width = (int)kf.frame_width;
        c.frame_height = (int)kf.frame_height;
        c.depth_scale = kf.depth_scale;
        c.fx = kf.fx;
        c.fy = kf.fy;
        c.cx = kf.cx;
        c.cy = kf.cy;
        c.unmirror = kf.unmirror ? 1 : 0;
        c.volume_n = kf.volume_n;
        c.volume_size = kf.volume_size;
        c.mue = kf.mue;
        c.voxel_format = kf.voxel_format;
        for (int i = 0; i < 16; ++i) c.initial_pose[i] = kf.initial_pose[i];
        c.vis_volume_n = kf.visualization_volume_n;
        c.vis_volume_size = kf.visualization_volume_size;
        c.vis_mue = kf.visualization_mue;
        c.icp.levels = (int)icp.iters_per_layer.size();
        for (int l = 0; l < c.icp.levels && l < 3; ++l) c.icp.iters_per_layer[l] = (int)icp.iters_per_layer[l];
        c.icp.width = (int)icp.width;
        c.icp.height = (int)icp.height;
        c.icp.distance_thresh = icp.distance_thresh;
        c.icp.angle_thresh = icp.angle_thresh;
        c.icp.iteration_stop_thresh_angle = icp.iteration_stop_thresh_angle;
        c.icp.iteration_stop_thresh_distance = icp.iteration_stop_thresh_distance;
        c.icp.compensated = icp.compensated ? 1 : 0;
        check(kf_pipeline_create(&c, nullptr, &m_p), "KinectFusion");
    }
    ~KinectFusion() { kf_pipeline_destroy(m_p); }
    KinectFusion(const KinectFusion&) = delete;
    KinectFusion& operator=(const KinectFusion&) = delete;

    // one frame of raw depth codes already on the device
    void processFrame(const DepthMap& raw) { check(kf_pipeline_frame(m_p, static_cast<const uint16_t*>(raw.data())), "frame"); }
    // waits; throws when a frame's ICP failed since the last call
    void synchronize()
    {
        int st = 0;
        check(kf_pipeline_sync(m_p, &st), "synchronize");
        if (st == KF_ERR_NONFINITE || st == KF_ERR_NOT_PD)
            throw Error(st, "Error: an element in the result vector is not finite or NAN.");
    }
    Mat4 getPose()
    {
        Mat4 p{};
        check(kf_pipeline_pose(m_p, p.data()), "getPose");
        return p;
    }
    // visualizeCurrentModel (kinect_fusion.cpp:185-201) into an RGBA8 device image of the frame size:
    // shading (use_shading) or the normal image; from the current pose unless one is given
    void visualize(uint32_t* rgba_device, bool shading, const Mat4* pose = nullptr)
    {
        check(kf_pipeline_render(m_p, pose ? pose->data() : nullptr, shading ? 1 : 0, rgba_device), "visualize");
    }
    kf_pipeline* handle() const { return m_p; }

private:
    kf_pipeline* m_p = nullptr;
};

}  // namespace kfusion
