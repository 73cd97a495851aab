// TUM RGB-D sequences on the host (project/src/rgbd_dataset.cpp:6-97) and the
// pose error the reference reports (general_helper.cpp:11-21).
//
// A sequence directory holds depth.txt ("timestamp depth/<name>.png" lines,
// three header lines), groundtruth.txt ("timestamp tx ty tz qx qy qz qw", three
// header lines) and the 16-bit grayscale PNG depth images.  Each depth frame is
// paired with the ground-truth line nearest in time by the reference's forward
// search.  PNG decoding is written here (zlib inflate + the five PNG row
// filters, PNG specification 1.2 section 6), since stb is not available.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <array>
#include <fstream>
#include <iterator>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "kfusion.h"

namespace kf {
// error plumbing of the library (runtime.hip): records kf_last_error() text, returns status
int fail(int status, const std::string& msg);
}  // namespace kf

struct kf_tum {
    std::string dir;
    std::vector<std::string> images;
    std::vector<std::array<float, 16>> poses;
    int width = 0, height = 0;
};

namespace {

uint32_t be32(const unsigned char* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// Paeth predictor (PNG 1.2, 6.6)
int paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

// Decodes a non-interlaced 16-bit grayscale PNG into host uint16 samples.
int decode_png16(const std::string& path, std::vector<uint16_t>& out, int& w, int& h)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return kf::fail(KF_ERR_IO, "TUM: cannot open " + path);
    std::vector<unsigned char> file((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (file.size() < 8 || memcmp(file.data(), sig, 8) != 0) return kf::fail(KF_ERR_IO, "TUM: not a PNG: " + path);
    std::vector<unsigned char> idat;
    int depth = 0, ctype = -1, interlace = 0;
    w = h = 0;
    for (size_t p = 8; p + 12 <= file.size();) {
        const uint32_t len = be32(&file[p]);
        if (p + 12 + (size_t)len > file.size()) return kf::fail(KF_ERR_IO, "TUM: truncated PNG: " + path);
        const char* type = reinterpret_cast<const char*>(&file[p + 4]);
        const unsigned char* data = &file[p + 8];
        if (!memcmp(type, "IHDR", 4) && len >= 13) {
            w = (int)be32(data);
            h = (int)be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        p += 12 + (size_t)len;
    }
    if (w <= 0 || h <= 0 || w > 65536 || h > 65536 || depth != 16 || ctype != 0 || interlace != 0)
        return kf::fail(KF_ERR_IO, "TUM: expected a non-interlaced 16-bit grayscale PNG: " + path);
    const size_t stride = (size_t)w * 2;
    std::vector<unsigned char> raw((stride + 1) * (size_t)h);
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return kf::fail(KF_ERR_IO, "TUM: inflateInit");
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_OK) || got != raw.size())
        return kf::fail(KF_ERR_IO, "TUM: corrupt image data: " + path);
    // unfilter in place (bytes per pixel = 2)
    std::vector<unsigned char> prev(stride, 0), cur(stride);
    out.resize((size_t)w * h);
    for (int y = 0; y < h; ++y) {
        const unsigned char ft = raw[(size_t)y * (stride + 1)];
        const unsigned char* src = &raw[(size_t)y * (stride + 1) + 1];
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= 2 ? cur[x - 2] : 0, b = prev[x], c = x >= 2 ? prev[x - 2] : 0;
            int v = src[x];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) / 2; break;
                case 4: v += paeth(a, b, c); break;
                default: return kf::fail(KF_ERR_IO, "TUM: bad PNG filter type: " + path);
            }
            cur[x] = (unsigned char)v;
        }
        for (int x = 0; x < w; ++x) out[(size_t)y * w + x] = (uint16_t)(cur[2 * x] << 8 | cur[2 * x + 1]);
        prev.swap(cur);
    }
    return KF_OK;
}

// glm::mat4_cast of quaternion (x, y, z, w) with pose[3] = (t, 1) (rgbd_dataset.cpp:67-69), column-major.
std::array<float, 16> quat_pose(float tx, float ty, float tz, float qx, float qy, float qz, float qw)
{
    const float xx = qx * qx, yy = qy * qy, zz = qz * qz, xz = qx * qz, xy = qx * qy, yz = qy * qz;
    const float wx = qw * qx, wy = qw * qy, wz = qw * qz;
    std::array<float, 16> m{};
    m[0] = 1.0f - 2.0f * (yy + zz);
    m[1] = 2.0f * (xy + wz);
    m[2] = 2.0f * (xz - wy);
    m[4] = 2.0f * (xy - wz);
    m[5] = 1.0f - 2.0f * (xx + zz);
    m[6] = 2.0f * (yz + wx);
    m[8] = 2.0f * (xz + wy);
    m[9] = 2.0f * (yz - wx);
    m[10] = 1.0f - 2.0f * (xx + yy);
    m[12] = tx;
    m[13] = ty;
    m[14] = tz;
    m[15] = 1.0f;
    return m;
}

// the three header lines of a TUM text file
void skip_header(std::istream& in)
{
    std::string line;
    for (int k = 0; k < 3; ++k) std::getline(in, line);
}

}  // namespace

extern "C" {

int kf_tum_open(const char* dir, kf_tum** out)
{
    if (!dir || !out) return kf::fail(KF_ERR_ARG, "kf_tum_open: null argument");
    *out = nullptr;
    const std::string d(dir);
    std::ifstream gt(d + "/groundtruth.txt"), dp(d + "/depth.txt");
    if (!gt || !dp) return kf::fail(KF_ERR_IO, "kf_tum_open: depth.txt and groundtruth.txt are required in " + d);
    skip_header(gt);
    std::vector<double> ts;
    std::vector<std::array<float, 7>> gtv;
    double t;
    while (gt >> t) {
        std::array<float, 7> v;
        for (float& x : v) gt >> x;
        if (!gt) return kf::fail(KF_ERR_IO, "kf_tum_open: malformed groundtruth.txt");
        ts.push_back(t);
        gtv.push_back(v);
    }
    if (ts.empty()) return kf::fail(KF_ERR_IO, "kf_tum_open: empty groundtruth.txt");
    skip_header(dp);
    kf_tum* s = new kf_tum();
    s->dir = d;
    // rgbd_dataset.cpp:46-64: forward search from the last match; the search stops at the
    // first increase of |dt| and only then moves its start
    int last = 0;
    const int n = (int)ts.size();
    std::string name;
    while (dp >> t >> name) {
        double best = std::numeric_limits<double>::max();
        int bi = -1;
        for (int i = last; i < n; ++i) {
            const double diff = fabs(t - ts[i]);
            if (diff < best) {
                best = diff;
                bi = i;
            } else {
                last = bi;
                break;
            }
        }
        if (bi < 0) {
            delete s;
            return kf::fail(KF_ERR_IO, "kf_tum_open: no ground truth for a depth frame");
        }
        const auto& g = gtv[bi];
        s->images.push_back(d + "/" + name);
        s->poses.push_back(quat_pose(g[0], g[1], g[2], g[3], g[4], g[5], g[6]));
    }
    if (s->images.empty()) {
        delete s;
        return kf::fail(KF_ERR_IO, "kf_tum_open: empty depth.txt");
    }
    std::vector<uint16_t> first;
    int st = decode_png16(s->images[0], first, s->width, s->height);
    if (st) {
        delete s;
        return st;
    }
    *out = s;
    return KF_OK;
}

int kf_tum_close(kf_tum* s)
{
    delete s;
    return KF_OK;
}

int kf_tum_size(const kf_tum* s, int* frames, int* width, int* height)
{
    if (!s) return kf::fail(KF_ERR_ARG, "kf_tum_size: null sequence");
    if (frames) *frames = (int)s->images.size();
    if (width) *width = s->width;
    if (height) *height = s->height;
    return KF_OK;
}

int kf_tum_frame(kf_tum* s, int i, int gt_offset, uint16_t* depth, float pose[16])
{
    if (!s || i < 0 || i >= (int)s->images.size() || gt_offset < 0)
        return kf::fail(KF_ERR_ARG, "kf_tum_frame: bad argument");
    if (depth) {
        std::vector<uint16_t> img;
        int w = 0, h = 0;
        int st = decode_png16(s->images[i], img, w, h);
        if (st) return st;
        if (w != s->width || h != s->height) return kf::fail(KF_ERR_DIMS, "kf_tum_frame: image size changed");
        memcpy(depth, img.data(), img.size() * sizeof(uint16_t));
    }
    if (pose) {
        const int k = i + gt_offset < (int)s->poses.size() ? i + gt_offset : (int)s->poses.size() - 1;
        memcpy(pose, s->poses[k].data(), sizeof(float) * 16);
    }
    return KF_OK;
}

int kf_pose_error(const float p1[16], const float p2[16], float* angle, float* distance)
{
    if (!p1 || !p2) return kf::fail(KF_ERR_ARG, "kf_pose_error: null argument");
    // v = normalize(1, 1, 1); angle = |R1 v - R2 v|; distance = |t1 - t2|
    const float inv = 1.0f / sqrtf((1.0f * 1.0f + 1.0f * 1.0f) + 1.0f * 1.0f);
    const float v[3] = {inv, inv, inv};
    float a[3], b[3];
    for (int r = 0; r < 3; ++r) {
        a[r] = (p1[r] * v[0] + p1[4 + r] * v[1]) + p1[8 + r] * v[2];
        b[r] = (p2[r] * v[0] + p2[4 + r] * v[1]) + p2[8 + r] * v[2];
    }
    const float da[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    const float dt[3] = {p1[12] - p2[12], p1[13] - p2[13], p1[14] - p2[14]};
    if (angle) *angle = sqrtf((da[0] * da[0] + da[1] * da[1]) + da[2] * da[2]);
    if (distance) *distance = sqrtf((dt[0] * dt[0] + dt[1] * dt[1]) + dt[2] * dt[2]);
    return KF_OK;
}

}  // extern "C"
