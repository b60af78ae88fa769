/*
 * kf.hpp -- C++17 host mirror of the reference's hot-path classes over the C ABI (kf_api.h).
 *
 * Same class and member names as the reference, so code written against
 * KillingFusion/include/SDF.h, DisplacementField.h and SimpleMesh.h reads the same:
 *   kf::SDF                 SDF.h:16-200         (device-resident; FP32, BF16 or FP16 storage)
 *   kf::DisplacementField   DisplacementField.h:9-66
 *   kf::SimpleMesh          SimpleMesh.h:20-145  (triangle soup; OFF read / write)
 *   kf::KillingFusion       KillingFusion::processNextFrame, KillingFusion.cpp:109-175
 *   kf::computeDisplacementField   KillingFusion::computeDisplacementField (:240-397)
 * and the reference's on-disk formats: SDF::dumpToBinFile (SDF.cpp:432-460),
 * DisplacementField::dumpToBinFile (DisplacementField.cpp:349-366), SimpleMesh::WriteMesh and
 * the OFF-reading constructor (SimpleMesh.h:25-58, 113-141).
 *
 * Errors: every failing call throws kf::Error (status + kf_last_error()).  A diverged
 * optimisation throws too, as the reference does (`throw -1`, KillingFusion.cpp:327-331).
 * Link with killingfusion_amd/libkf.so.
 */
#ifndef KF_HPP
#define KF_HPP

#include "kf_api.h"

#include <array>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace kf {

class Error : public std::runtime_error {
public:
    Error(kf_status s, const std::string &where)
        : std::runtime_error(where + ": " + kf_status_string(s) + " (" + kf_last_error() + ")"), status(s)
    {
    }
    kf_status status;
};

inline void check(kf_status s, const char *where)
{
    if (s != KF_OK)
        throw Error(s, where);
}

/* A device and its stream (kf_context). */
class Context {
public:
    explicit Context(int device = 0) { check(kf_context_create(device, nullptr, &h_), "kf_context_create"); }
    ~Context() { kf_context_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    void synchronize() { check(kf_context_synchronize(h_), "kf_context_synchronize"); }
    kf_context *handle() const { return h_; }

private:
    kf_context *h_ = nullptr;
};

/* Page-locked host memory for depth frames (kf_host_alloc): uploads from it are DMA without a
 * staging copy.  Move-only. */
template <class T>
class HostBuffer {
public:
    explicit HostBuffer(size_t count) : n_(count)
    {
        void *p = nullptr;
        check(kf_host_alloc(sizeof(T) * (count ? count : 1), &p), "kf_host_alloc");
        p_ = static_cast<T *>(p);
    }
    ~HostBuffer() { kf_host_free(p_); }
    HostBuffer(const HostBuffer &) = delete;
    HostBuffer &operator=(const HostBuffer &) = delete;
    HostBuffer(HostBuffer &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
    T *data() { return p_; }
    const T *data() const { return p_; }
    size_t size() const { return n_; }
    T &operator[](size_t i) { return p_[i]; }

private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

/* Optimiser parameters: the config.cpp:4-57 globals; `sweep` (kf_sweep) picks the adjoint
 * optimiser's dense or active-set sweep (KF_SWEEP_AUTO by default). */
struct Params : kf_params {
    Params() { kf_params_default(this); }
    /* The values quoted in config.cpp:32-41 (alpha 0.1, omega_k 0.5, gamma 0.1, omega_s 0.2)
     * with Max = clip = trunc_voxels voxels. */
    static Params paper(double voxel, double trunc_voxels = 3.0, kf_method m = KF_METHOD_ADJOINT)
    {
        Params p;
        p.voxel_size = voxel;
        p.max_surface = p.clip_distance = trunc_voxels * voxel;
        p.alpha = 0.1;
        p.omega_killing = 0.5;
        p.gamma_killing = 0.1;
        p.omega_levelset = 0.2;
        p.method = m;
        return p;
    }
};

/* Triangle soup in the reference's order: vertex 3t + k is corner k of triangle t, y already
 * negated (SimpleMesh::AddVertex); faces are (3t, 3t + 1, 3t + 2). */
class SimpleMesh {
public:
    SimpleMesh() = default;
    /* SimpleMesh(std::string filepath), SimpleMesh.h:25-58: reads an OFF file. */
    explicit SimpleMesh(const std::string &filepath)
    {
        std::ifstream in(filepath);
        if (!in.is_open())
            throw std::runtime_error("failed to open " + filepath);
        std::string tmp;
        size_t nv = 0, nt = 0;
        in >> tmp >> nv >> nt >> tmp;
        std::vector<double> v(3 * nv);
        for (size_t i = 0; i < 3 * nv; i++)
            in >> v[i];
        for (size_t t = 0; t < nt; t++) {
            unsigned int a, b, c;
            in >> tmp >> a >> b >> c;
            if (a >= nv || b >= nv || c >= nv)
                throw std::runtime_error("face index out of range in " + filepath);
            for (unsigned int k : {a, b, c})
                for (int d = 0; d < 3; d++)
                    verts_.push_back((float)v[3 * (size_t)k + d]);
        }
    }
    size_t triangles() const { return verts_.size() / 9; }
    const std::vector<float> &vertices() const { return verts_; }
    std::vector<float> &vertices() { return verts_; }
    /* SimpleMesh::WriteMesh, SimpleMesh.h:113-141: OFF, vertices with the stream's default
     * 6 significant digits, faces "3 i j k". */
    bool WriteMesh(const std::string &filename) const
    {
        std::ofstream out(filename);
        if (!out.is_open())
            return false;
        const size_t nt = triangles();
        out << "OFF" << std::endl;
        out << 3 * nt << " " << nt << " 0" << std::endl;
        for (size_t i = 0; i < 3 * nt; i++)
            out << (double)verts_[3 * i] << " " << (double)verts_[3 * i + 1] << " " << (double)verts_[3 * i + 2]
                << std::endl;
        for (size_t t = 0; t < nt; t++)
            out << "3 " << 3 * t << " " << 3 * t + 1 << " " << 3 * t + 2 << std::endl;
        return true;
    }

private:
    std::vector<float> verts_;
};

/* Welded marching-cubes mesh (kf_marching_cubes_indexed): vertices, unit normals, and
 * triangles as vertex indices in the soup's order. */
struct IndexedMesh {
    std::vector<float> vertices, normals;  // 3 per vertex
    std::vector<uint32_t> triangles;       // 3 per triangle
    size_t numVertices() const { return vertices.size() / 3; }
    size_t numTriangles() const { return triangles.size() / 3; }
    bool WriteMesh(const std::string &filename) const
    {
        std::ofstream out(filename);
        if (!out.is_open())
            return false;
        out << "OFF" << std::endl << numVertices() << " " << numTriangles() << " 0" << std::endl;
        for (size_t i = 0; i < numVertices(); i++)
            out << (double)vertices[3 * i] << " " << (double)vertices[3 * i + 1] << " " << (double)vertices[3 * i + 2]
                << std::endl;
        for (size_t t = 0; t < numTriangles(); t++)
            out << "3 " << triangles[3 * t] << " " << triangles[3 * t + 1] << " " << triangles[3 * t + 2] << std::endl;
        return true;
    }
};

class DisplacementField {
public:
    /* DisplacementField::DisplacementField, DisplacementField.cpp:8-16: a zero field. */
    DisplacementField(Context &ctx, const std::array<int32_t, 3> &dims, double voxel_size = 0.0)
        : dims_(dims), voxel_(voxel_size)
    {
        check(kf_field_create(ctx.handle(), dims_.data(), &h_), "kf_field_create");
    }
    ~DisplacementField() { kf_field_destroy(h_); }
    DisplacementField(const DisplacementField &) = delete;
    DisplacementField &operator=(const DisplacementField &) = delete;

    /* DisplacementField::initializeAllVoxels, DisplacementField.cpp:82-88. */
    void initializeAllVoxels(double x, double y, double z)
    {
        const double v[3] = {x, y, z};
        check(kf_field_fill(h_, v), "kf_field_fill");
    }
    /* DisplacementField::operator+, DisplacementField.cpp:73-80 (adds in place, as the
     * reference does, and returns *this). */
    DisplacementField &operator+(const DisplacementField &o)
    {
        check(kf_field_add(h_, o.h_), "kf_field_add");
        return *this;
    }
    void upload(const std::vector<float> &soa)
    {
        if (soa.size() != 3 * n())
            throw std::invalid_argument("field upload: need 3 N floats (SoA)");
        check(kf_field_upload(h_, soa.data()), "kf_field_upload");
    }
    std::vector<float> download() const
    {
        std::vector<float> soa(3 * n());
        check(kf_field_download(h_, soa.data()), "kf_field_download");
        return soa;
    }
    /* DisplacementField::getDisplacementAt(x, y, z), DisplacementField.cpp:30-38: the vector at
     * linear index (z ny + y) nx + x, zero when that index is outside [0, N).  Downloads the
     * field, so it is for checks, not loops. */
    std::array<double, 3> getDisplacementAt(int x, int y, int z) const
    {
        const int64_t i = ((int64_t)z * dims_[1] + y) * dims_[0] + x;
        if (i < 0 || i >= (int64_t)n())
            return {0.0, 0.0, 0.0};
        const std::vector<float> s = download();
        return {s[(size_t)i], s[n() + (size_t)i], s[2 * n() + (size_t)i]};
    }
    /* DisplacementField::dumpToBinFile, DisplacementField.cpp:349-366: int32 dims[3], double
     * voxel size, then N x 3 doubles (x, y, z per voxel). */
    void dumpToBinFile(const std::string &path) const { writeFieldBin(h_, dims_, voxel_, path); }
    /* The same dump of any field handle (SoA float32 on the device, dims voxels). */
    static void writeFieldBin(kf_field *h, const std::array<int32_t, 3> &dims, double voxel, const std::string &path)
    {
        const size_t n = (size_t)dims[0] * dims[1] * dims[2];
        std::vector<float> s(3 * n);
        check(kf_field_download(h, s.data()), "kf_field_download");
        std::ofstream out(path, std::ios::binary);
        if (!out.is_open())
            throw std::runtime_error("cannot write " + path);
        out.write((const char *)dims.data(), 3 * sizeof(int32_t));
        out.write((const char *)&voxel, sizeof(double));
        std::vector<double> aos(3 * n);
        for (size_t i = 0; i < n; i++)
            for (int a = 0; a < 3; a++)
                aos[3 * i + a] = s[(size_t)a * n + i];
        out.write((const char *)aos.data(), (std::streamsize)(aos.size() * sizeof(double)));
    }
    size_t n() const { return (size_t)dims_[0] * dims_[1] * dims_[2]; }
    const std::array<int32_t, 3> &dims() const { return dims_; }
    kf_field *handle() const { return h_; }

private:
    std::array<int32_t, 3> dims_;
    double voxel_;
    kf_field *h_ = nullptr;
};

class SDF {
public:
    /* SDF::SDF(voxelSize, min3dLoc, max3dLoc, unknownClipDistance), SDF.cpp:14-30: dims from
     * the bounds (computeVoxelGridSize, SDF.cpp:117-124), phi = Max + eps, w = 0. */
    SDF(Context &ctx, double voxelSize, const std::array<double, 3> &min3dLoc, const std::array<double, 3> &max3dLoc,
        double unknownClipDistance, double maxSurface, kf_dtype dtype = KF_F32, double epsilon = 1e-5)
    {
        std::array<int32_t, 3> dims{};
        check(kf_dims_from_bounds(min3dLoc.data(), max3dLoc.data(), voxelSize, dims.data()), "kf_dims_from_bounds");
        create(ctx, dims, min3dLoc, voxelSize, unknownClipDistance, maxSurface, dtype, epsilon);
    }
    /* The same with the dims given. */
    SDF(Context &ctx, const std::array<int32_t, 3> &dims, const std::array<double, 3> &min3dLoc, double voxelSize,
        double unknownClipDistance, double maxSurface, kf_dtype dtype = KF_F32, double epsilon = 1e-5)
    {
        create(ctx, dims, min3dLoc, voxelSize, unknownClipDistance, maxSurface, dtype, epsilon);
    }
    ~SDF() { kf_volume_destroy(h_); }
    SDF(const SDF &) = delete;
    SDF &operator=(const SDF &) = delete;

    /* SDF::integrateDepthFrame(depthFrame, C2W, K, minDepth, maxDepth), SDF.cpp:137-197;
     * depth: row-major float32 metres, 0 = invalid; C2W and K row-major. */
    void integrateDepthFrame(const std::vector<float> &depth, int width, int height, const std::array<double, 16> &C2W,
                             const std::array<double, 9> &K, double minDepth, double maxDepth)
    {
        if (depth.size() != (size_t)width * height)
            throw std::invalid_argument("depth frame size");
        check(kf_integrate_depth(h_, depth.data(), width, height, K.data(), C2W.data(), minDepth, maxDepth),
              "kf_integrate_depth");
    }
    /* The same from a page-locked frame. */
    void integrateDepthFrame(const HostBuffer<float> &depth, int width, int height, const std::array<double, 16> &C2W,
                             const std::array<double, 9> &K, double minDepth, double maxDepth)
    {
        if (depth.size() != (size_t)width * height)
            throw std::invalid_argument("depth frame size");
        check(kf_integrate_depth(h_, depth.data(), width, height, K.data(), C2W.data(), minDepth, maxDepth),
              "kf_integrate_depth");
    }
    /* SDF::fuse(otherSdf), SDF.cpp:199-241. */
    void fuse(const SDF *other) { check(kf_fuse(h_, other->h_), "kf_fuse"); }
    /* SDF::fuse(otherSdf, displacementField), SDF.cpp:244-298. */
    void fuse(const SDF *other, const DisplacementField *u)
    {
        check(kf_fuse_warped(h_, other->h_, u->handle()), "kf_fuse_warped");
    }
    /* SDF::update(displacementField), SDF.cpp:300-328. */
    void update(const DisplacementField *u) { check(kf_warp(h_, u->handle()), "kf_warp"); }
    /* SDF::getMesh(), SDF.cpp:396-414. */
    SimpleMesh getMesh(double iso = 0.0) const
    {
        uint64_t nt = 0;
        check(kf_marching_cubes(h_, iso, &nt, nullptr, 0), "kf_marching_cubes");
        SimpleMesh m;
        m.vertices().resize((size_t)nt * 9);
        if (nt)
            check(kf_marching_cubes(h_, iso, &nt, m.vertices().data(), nt), "kf_marching_cubes");
        return m;
    }
    /* The welded mesh of getMesh with per-vertex normals. */
    IndexedMesh getMeshIndexed(double iso = 0.0) const
    {
        uint64_t nv = 0, nt = 0;
        check(kf_marching_cubes_indexed(h_, iso, &nv, &nt, nullptr, nullptr, nullptr, 0, 0), "kf_marching_cubes_indexed");
        IndexedMesh m;
        m.vertices.resize(3 * (size_t)nv);
        m.normals.resize(3 * (size_t)nv);
        m.triangles.resize(3 * (size_t)nt);
        if (nt)
            check(kf_marching_cubes_indexed(h_, iso, &nv, &nt, m.vertices.data(), m.normals.data(), m.triangles.data(), nv,
                                            nt),
                  "kf_marching_cubes_indexed");
        return m;
    }
    void upload(const std::vector<float> &phi, const std::vector<int32_t> &w)
    {
        if (phi.size() != n() || (!w.empty() && w.size() != n()))
            throw std::invalid_argument("volume upload: need N values");
        check(kf_volume_upload(h_, phi.data(), w.empty() ? nullptr : w.data()), "kf_volume_upload");
    }
    std::pair<std::vector<float>, std::vector<int32_t>> download() const
    {
        std::vector<float> phi(n());
        std::vector<int32_t> w(n());
        check(kf_volume_download(h_, phi.data(), w.data()), "kf_volume_download");
        return {phi, w};
    }
    /* SDF::dumpToBinFile(path, truncationDistanceInVoxelSizeUnit, minimumWeightThreshold),
     * SDF.cpp:432-460: int32 dims[3], double min[3], double voxel size, double truncation (as
     * given), then N doubles of phi.  The weight threshold is unused there too. */
    void dumpToBinFile(const std::string &path, double truncationDistanceInVoxelSizeUnit,
                       double /*minimumWeightThreshold*/) const
    {
        const auto pw = download();
        std::ofstream out(path, std::ios::binary);
        if (!out.is_open())
            throw std::runtime_error("cannot write " + path);
        out.write((const char *)desc_.dims, 3 * sizeof(int32_t));
        out.write((const char *)desc_.min_corner, 3 * sizeof(double));
        out.write((const char *)&desc_.voxel_size, sizeof(double));
        out.write((const char *)&truncationDistanceInVoxelSizeUnit, sizeof(double));
        std::vector<double> phi(pw.first.begin(), pw.first.end());
        out.write((const char *)phi.data(), (std::streamsize)(phi.size() * sizeof(double)));
    }
    size_t n() const { return (size_t)desc_.dims[0] * desc_.dims[1] * desc_.dims[2]; }
    std::array<int32_t, 3> dims() const { return {desc_.dims[0], desc_.dims[1], desc_.dims[2]}; }
    kf_volume *handle() const { return h_; }

private:
    void create(Context &ctx, const std::array<int32_t, 3> &dims, const std::array<double, 3> &mn, double voxel,
                double clip, double maxSurface, kf_dtype dtype, double epsilon)
    {
        desc_ = kf_volume_desc{};
        for (int a = 0; a < 3; a++) {
            desc_.dims[a] = dims[a];
            desc_.min_corner[a] = mn[a];
        }
        desc_.voxel_size = voxel;
        desc_.clip_distance = clip;
        desc_.max_surface = maxSurface;
        desc_.epsilon = epsilon;
        desc_.dtype = dtype;
        check(kf_volume_create(ctx.handle(), &desc_, &h_), "kf_volume_create");
    }
    kf_volume_desc desc_{};
    kf_volume *h_ = nullptr;
};

/* KillingFusion::computeDisplacementField(src, dest, srcToDest), all-voxels mode.  Throws
 * kf::Error with status KF_E_DIVERGED on a non-finite displacement. */
inline kf_deform_stats computeDisplacementField(const SDF *src, const SDF *dest, DisplacementField *srcToDest,
                                                const Params &p)
{
    kf_deform_stats st{};
    check(kf_deform(src->handle(), dest->handle(), srcToDest->handle(), &p, &st), "kf_deform");
    return st;
}

/* Rigid pre-registration (no reference counterpart): the pose (row-major 4x4, canonical world
 * -> live world) minimising the squared SDF difference over the canonical band, from `pose`. */
inline kf_rigid_stats rigidAlign(const SDF *live, const SDF *can, std::array<double, 16> &pose, int maxIters = 30)
{
    kf_rigid_stats st{};
    check(kf_rigid_align(live->handle(), can->handle(), maxIters, pose.data(), &st), "kf_rigid_align");
    return st;
}

/* The displacement field of a rigid pose: a starting point for computeDisplacementField. */
inline void fieldFromRigid(DisplacementField *u, const SDF *can, const SDF *live, const std::array<double, 16> &pose)
{
    check(kf_field_from_rigid(u->handle(), can->handle(), live->handle(), pose.data()), "kf_field_from_rigid");
}

/* Independent volumes in one batch (kf_batch_deform): one stats entry per volume. */
inline std::vector<kf_deform_stats> batchDeform(const std::vector<const SDF *> &live, const std::vector<const SDF *> &can,
                                                const std::vector<DisplacementField *> &u, const Params &p)
{
    if (live.size() != can.size() || live.size() != u.size())
        throw std::invalid_argument("batch lists differ in length");
    std::vector<const kf_volume *> l, c;
    std::vector<kf_field *> f;
    for (size_t b = 0; b < live.size(); b++) {
        l.push_back(live[b]->handle());
        c.push_back(can[b]->handle());
        f.push_back(u[b]->handle());
    }
    std::vector<kf_deform_stats> st(live.size());
    check(kf_batch_deform(l.data(), c.data(), f.data(), (int32_t)live.size(), &p, st.data()), "kf_batch_deform");
    return st;
}

/* The per-frame driver: one processNextFrame per depth frame over a device-resident
 * canonical SDF (frame 1 becomes the canonical SDF). */
class KillingFusion {
public:
    /* fuseMode: KF_FUSE_MERGE (FUSE_BY_MERGE = true, config.cpp:7) or KF_FUSE_NEAREST. */
    KillingFusion(Context &ctx, const std::array<int32_t, 3> &dims, const std::array<double, 3> &min3d,
                  const Params &p, const std::array<double, 9> &K, const std::array<double, 16> &C2W,
                  double minDepth, double maxDepth, int width, int height, kf_dtype dtype = KF_F32,
                  int32_t fuseMode = KF_FUSE_MERGE)
        : width_(width), height_(height), dims_(dims), voxel_(p.voxel_size),
          deforms_(p.use_data || p.use_levelset || p.use_killing)
    {
        kf_pipeline_desc d{};
        for (int a = 0; a < 3; a++) {
            d.volume.dims[a] = dims[a];
            d.volume.min_corner[a] = min3d[a];
        }
        d.volume.voxel_size = p.voxel_size;
        d.volume.clip_distance = p.clip_distance;
        d.volume.max_surface = p.max_surface;
        d.volume.epsilon = p.epsilon;
        d.volume.dtype = dtype;
        d.params = p;
        for (int i = 0; i < 9; i++)
            d.K[i] = K[i];
        for (int i = 0; i < 16; i++)
            d.C2W[i] = C2W[i];
        d.min_depth = minDepth;
        d.max_depth = maxDepth;
        d.width = width;
        d.height = height;
        d.fuse_mode = fuseMode;
        check(kf_pipeline_create(ctx.handle(), &d, &h_), "kf_pipeline_create");
    }
    ~KillingFusion() { kf_pipeline_destroy(h_); }
    KillingFusion(const KillingFusion &) = delete;
    KillingFusion &operator=(const KillingFusion &) = delete;

    kf_frame_stats processNextFrame(const std::vector<float> &depth)
    {
        if (depth.size() != (size_t)width_ * height_)
            throw std::invalid_argument("depth frame size");
        kf_frame_stats st{};
        check(kf_pipeline_process_frame(h_, depth.data(), &st), "kf_pipeline_process_frame");
        // KillingFusion.cpp:146-151: after the deform of a frame (every frame but the first,
        // when some energy is on) its field goes to <prefix><frame, 3 digits>.bin
        if (!dumpPrefix_.empty() && st.frame_index > 0 && deforms_) {
            char num[16];
            std::snprintf(num, sizeof(num), "%03d", dumpStart_ + dumpStride_ * (int)st.frame_index);
            DisplacementField::writeFieldBin(kf_pipeline_field(h_), dims_, voxel_, dumpPrefix_ + num + ".bin");
        }
        return st;
    }
    /* The per-frame displacement dump of the reference (OUTPUT_DIR + dataset directory, then the
     * dataset frame number m_currFrameIndex = startFrame + stride x the frame's index, ".bin");
     * an empty prefix (the default) writes nothing. */
    void setFieldDumpPrefix(const std::string &prefix, int startFrame = 0, int stride = 1)
    {
        dumpPrefix_ = prefix;
        dumpStart_ = startFrame;
        dumpStride_ = stride;
    }
    /* Which of the frame's meshes each processNextFrame extracts (KF_MESH_* bits, iso value):
     * the live SDF before the warp (:157), after it (:160), the canonical SDF (:170). */
    void setMeshes(int32_t which, double iso = 0.0)
    {
        check(kf_pipeline_set_meshes(h_, which, iso), "kf_pipeline_set_meshes");
    }
    /* One of the last frame's meshes (a KF_MESH_* value chosen with setMeshes). */
    SimpleMesh frameMesh(int32_t which) const
    {
        uint64_t nt = 0;
        check(kf_pipeline_mesh(h_, which, &nt, nullptr, 0), "kf_pipeline_mesh");
        SimpleMesh m;
        m.vertices().resize((size_t)nt * 9);
        if (nt)
            check(kf_pipeline_mesh(h_, which, &nt, m.vertices().data(), nt), "kf_pipeline_mesh");
        return m;
    }
    /* Canonical mesh (SDF::getMesh on m_canonicalSdf, KillingFusion.cpp:170). */
    SimpleMesh canonicalMesh(double iso = 0.0) const
    {
        uint64_t nt = 0;
        kf_volume *v = kf_pipeline_canonical(h_);
        check(kf_marching_cubes(v, iso, &nt, nullptr, 0), "kf_marching_cubes");
        SimpleMesh m;
        m.vertices().resize((size_t)nt * 9);
        if (nt)
            check(kf_marching_cubes(v, iso, &nt, m.vertices().data(), nt), "kf_marching_cubes");
        return m;
    }
    kf_volume *canonical() const { return kf_pipeline_canonical(h_); }
    kf_pipeline *handle() const { return h_; }

private:
    kf_pipeline *h_ = nullptr;
    int width_, height_;
    std::array<int32_t, 3> dims_;
    double voxel_;
    bool deforms_;
    std::string dumpPrefix_;
    int dumpStart_ = 0, dumpStride_ = 1;
};

}  // namespace kf

#endif
