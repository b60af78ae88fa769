// kf_hpp_check.cpp -- drives include/kf.hpp (the C++ mirror of SDF, DisplacementField,
// SimpleMesh and the per-frame driver) on the GPU with closed-form expectations, in the style
// of the reference's own self-tests (SDF.cpp:526-840, KillingFusion::processTest,
// KillingFusion.cpp:177-207).  Writes plane.sdf.bin, field.bin and plane.off into the
// directory given as argv[1] for the Python side to read back.  Prints "OK <checks>" and
// exits 0 when every check holds.
#include "kf.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace {

int g_checks = 0, g_fails = 0;

void expect(bool ok, const std::string &what)
{
    g_checks++;
    if (!ok) {
        g_fails++;
        std::printf("FAIL %s\n", what.c_str());
    }
}

const std::array<double, 9> K = {525.0, 0.0, 319.5, 0.0, 525.0, 239.5, 0.0, 0.0, 1.0};
const std::array<double, 16> I4 = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
constexpr int W = 640, H = 480;

}  // namespace

int main(int argc, char **argv)
{
    const std::string out = argc > 1 ? argv[1] : ".";
    try {
        kf::Context ctx;
        const std::array<int32_t, 3> dims = {32, 24, 40};
        const std::array<double, 3> mn = {-0.16, -0.12, 0.5};
        const double voxel = 0.01, trunc = 0.03;
        const size_t n = (size_t)dims[0] * dims[1] * dims[2];

        // 1. A fronto-parallel plane at 0.7 m: every voxel centre projects into the image, so
        // phi = D - Z_c (D the float depth), w = 1, where D - Z_c >= -clip; untouched elsewhere.
        kf::SDF plane(ctx, dims, mn, voxel, trunc, trunc);
        std::vector<float> depth((size_t)W * H, 0.7f);
        plane.integrateDepthFrame(depth, W, H, I4, K, 0.1, 3.0);
        auto pw = plane.download();
        size_t bad = 0, seen = 0;
        for (int z = 0; z < dims[2]; z++) {
            const double Zc = voxel * (z + 0.5) + mn[2];
            const double dist = (double)0.7f - Zc;
            for (int y = 0; y < dims[1]; y++)
                for (int x = 0; x < dims[0]; x++) {
                    const size_t i = ((size_t)z * dims[1] + y) * dims[0] + x;
                    const bool hit = dist >= -trunc;
                    const float ephi = hit ? (float)dist : (float)(trunc + 1e-5);
                    seen += hit;
                    bad += pw.first[i] != ephi || pw.second[i] != (hit ? 1 : 0);
                }
        }
        expect(bad == 0 && seen > 0, "plane integration: phi = D - Z_c, w = 1");

        // 2. fuse(self) keeps phi and doubles w (processTest(1), KillingFusion.cpp:182-188).
        kf::SDF copy(ctx, dims, mn, voxel, trunc, trunc);
        copy.upload(pw.first, pw.second);
        kf::SDF fused(ctx, dims, mn, voxel, trunc, trunc);
        fused.upload(pw.first, pw.second);
        fused.fuse(&copy);
        auto fw = fused.download();
        bad = 0;
        for (size_t i = 0; i < n; i++)
            bad += fw.first[i] != pw.first[i] || fw.second[i] != 2 * pw.second[i];
        expect(bad == 0, "fuse(self): phi unchanged, w doubled");

        // 3. update(u = 0) is the identity at the nodes.
        kf::DisplacementField zero(ctx, dims, voxel);
        kf::SDF warped(ctx, dims, mn, voxel, trunc, trunc);
        warped.upload(pw.first, pw.second);
        warped.update(&zero);
        auto ww = warped.download();
        expect(ww.first == pw.first && ww.second == pw.second, "update with a zero field is the identity");

        // 4. Field arithmetic and the reference's linear-index reads (DisplacementField.cpp:30-38).
        kf::DisplacementField a(ctx, dims, voxel), b(ctx, dims, voxel);
        a.initializeAllVoxels(0.25, -0.5, 1.0);
        b.initializeAllVoxels(0.5, 0.5, 0.5);
        a + b;
        auto v = a.getDisplacementAt(1, 2, 3);
        expect(v[0] == 0.75 && v[1] == 0.0 && v[2] == 1.5, "operator+ and initializeAllVoxels");
        auto wrap = a.getDisplacementAt(dims[0], 0, 0);  // linear index nx: the start of row 1
        expect(wrap[0] == 0.75 && wrap[2] == 1.5, "x = nx reads the next row (linear index)");
        auto none = a.getDisplacementAt(-1, 0, 0);
        expect(none[0] == 0.0 && none[1] == 0.0 && none[2] == 0.0, "negative linear index reads zero");

        // 5. computeDisplacementField: live plane 5 mm behind the canonical one moves samples
        // toward +z; more iterations move them further; a NaN field throws like the reference.
        std::vector<float> depth2((size_t)W * H, 0.705f);
        kf::SDF live(ctx, dims, mn, voxel, trunc, trunc);
        live.integrateDepthFrame(depth2, W, H, I4, K, 0.1, 3.0);
        kf::Params p = kf::Params::paper(voxel, 3.0, KF_METHOD_ADJOINT);
        double uz[2] = {0.0, 0.0};
        for (int r = 0; r < 2; r++) {
            kf::DisplacementField u(ctx, dims, voxel);
            p.max_iters = r == 0 ? 10 : 40;
            kf_deform_stats st = kf::computeDisplacementField(&live, &plane, &u, p);
            expect(st.iters == p.max_iters && st.diverged_index < 0 && st.active_voxels > 0, "deform stats");
            uz[r] = u.getDisplacementAt(16, 12, 20)[2];
        }
        expect(uz[0] > 0.0 && uz[1] > uz[0], "displacement moves the live samples toward the canonical plane");
        {
            // the dense and active-set sweeps give the same field
            kf::DisplacementField ud(ctx, dims, voxel), ua(ctx, dims, voxel);
            kf::Params pd = p, pa = p;
            pd.max_iters = pa.max_iters = 25;
            pd.sweep = KF_SWEEP_DENSE;
            pa.sweep = KF_SWEEP_ACTIVE;
            kf_deform_stats sd = kf::computeDisplacementField(&live, &plane, &ud, pd);
            kf_deform_stats sa = kf::computeDisplacementField(&live, &plane, &ua, pa);
            expect(sd.sweep == KF_SWEEP_DENSE && sa.sweep == KF_SWEEP_ACTIVE && sa.listed_voxels > 0 &&
                       sa.iters == sd.iters && sa.max_update == sd.max_update && ud.download() == ua.download(),
                   "dense and active-set sweeps give the same field");
        }
        {
            // a page-locked frame integrates as a std::vector one
            kf::HostBuffer<float> pinned(depth2.size());
            for (size_t i = 0; i < depth2.size(); i++)
                pinned[i] = depth2[i];
            kf::SDF again(ctx, dims, mn, voxel, trunc, trunc);
            again.integrateDepthFrame(pinned, W, H, I4, K, 0.1, 3.0);
            const auto x1 = again.download(), x2 = live.download();
            expect(x1.first == x2.first && x1.second == x2.second, "integration from a page-locked frame");
        }
        {
            kf::DisplacementField u(ctx, dims, voxel);
            std::vector<float> s = u.download();
            s[(size_t)(20 * dims[1] + 12) * dims[0] + 16] = NAN;
            u.upload(s);
            bool threw = false;
            try {
                kf::computeDisplacementField(&live, &plane, &u, p);
            } catch (const kf::Error &e) {
                threw = e.status == KF_E_DIVERGED;
            }
            expect(threw, "a non-finite displacement throws KF_E_DIVERGED");
        }

        // 6. The per-frame driver: the same plane twice leaves phi and doubles w (data and
        // Killing gradients vanish at u = 0 on identical SDFs; the level-set term is off because
        // out-of-grid corners, Max + eps, give it a gradient on the grid's last rows); a shifted
        // plane then runs the deform step; the canonical mesh is the plane.
        kf::Params pp = kf::Params::paper(voxel, 3.0);
        pp.max_iters = 100;
        pp.use_levelset = 0;
        kf::KillingFusion kfu(ctx, dims, mn, pp, K, I4, 0.1, 3.0, W, H);
        kfu.setFieldDumpPrefix(out + "/frame_", 10, 2);
        kfu.setMeshes(KF_MESH_LIVE | KF_MESH_LIVE_WARPED | KF_MESH_CANONICAL);
        kfu.processNextFrame(depth);
        const size_t m0 = kfu.frameMesh(KF_MESH_CANONICAL).triangles();
        expect(m0 > 100 && kfu.frameMesh(KF_MESH_LIVE).triangles() == m0 &&
                   kfu.frameMesh(KF_MESH_LIVE_WARPED).triangles() == m0,
               "frame 0: all three meshes are the canonical mesh");
        kf_frame_stats f2 = kfu.processNextFrame(depth);
        kf_volume_desc d{};
        kf::check(kf_volume_describe(kfu.canonical(), &d), "kf_volume_describe");
        std::vector<float> cphi(n);
        std::vector<int32_t> cw(n);
        kf::check(kf_volume_download(kfu.canonical(), cphi.data(), cw.data()), "kf_volume_download");
        bad = 0;
        for (size_t i = 0; i < n; i++)
            bad += cphi[i] != pw.first[i] || cw[i] != 2 * pw.second[i];
        expect(bad == 0 && f2.deform.iters == 100, "pipeline: identical frames fuse to the same phi, w doubled");
        kf_frame_stats f3 = kfu.processNextFrame(depth2);
        expect(f3.frame_index == 2 && f3.deform.diverged_index < 0 && f3.ms_total > 0.0, "pipeline frame 3");
        kf::SimpleMesh mesh = kfu.canonicalMesh();
        expect(mesh.triangles() > 100, "canonical mesh");
        expect(kfu.frameMesh(KF_MESH_CANONICAL).vertices() == mesh.vertices(), "frame mesh = canonical mesh");
        // the frame's field dumps: frames 1 and 2 at dataset numbers 12 and 14, none for frame 0
        kf::DisplacementField::writeFieldBin(kf_pipeline_field(kfu.handle()), dims, voxel, out + "/last_field.bin");
        expect(!std::ifstream(out + "/frame_010.bin").good() && std::ifstream(out + "/frame_012.bin").good() &&
                   std::ifstream(out + "/frame_014.bin").good(),
               "per-frame field dumps");

        // Welded mesh of the same plane: the soup's triangles over shared vertices.
        kf::IndexedMesh wm = plane.getMeshIndexed();
        kf::SimpleMesh pm = plane.getMesh();
        bool welded = wm.numTriangles() == pm.triangles() && wm.numVertices() < 3 * wm.numTriangles();
        for (size_t k = 0; welded && k < wm.triangles.size(); k++) {
            const uint32_t vi = wm.triangles[k];
            welded = vi < wm.numVertices() &&
                     std::fabs(wm.vertices[3 * vi] - pm.vertices()[3 * k]) <= 1e-6f &&
                     std::fabs(wm.vertices[3 * vi + 1] - pm.vertices()[3 * k + 1]) <= 1e-6f &&
                     std::fabs(wm.vertices[3 * vi + 2] - pm.vertices()[3 * k + 2]) <= 1e-6f;
        }
        expect(welded && pm.triangles() > 0, "welded mesh indexes the soup's vertices");
        expect(wm.WriteMesh(out + "/plane_welded.off"), "welded WriteMesh");

        // Rigid pre-registration of the plane onto itself keeps the identity; the batch of one
        // volume equals computeDisplacementField.
        std::array<double, 16> pose = I4;
        kf_rigid_stats rs = kf::rigidAlign(&plane, &plane, pose, 5);
        bool ident = rs.voxels > 0 && rs.energy < 1e-20;
        for (int k = 0; k < 16; k++)
            ident = ident && std::fabs(pose[k] - I4[k]) < 1e-12;
        expect(ident, "rigid alignment of a volume with itself is the identity");
        {
            kf::DisplacementField u1(ctx, dims, voxel), u2(ctx, dims, voxel);
            kf::Params pb = kf::Params::paper(voxel, 3.0, KF_METHOD_ADJOINT);
            pb.max_iters = 12;
            kf::computeDisplacementField(&live, &plane, &u1, pb);
            auto bst = kf::batchDeform({&live}, {&plane}, {&u2}, pb);
            expect(bst.size() == 1 && bst[0].iters == 12 && u1.download() == u2.download(),
                   "a batch of one equals computeDisplacementField");
        }

        // 7. Formats: OFF round trip and the two .bin dumps.
        const std::string off = out + "/plane.off";
        expect(mesh.WriteMesh(off), "WriteMesh");
        kf::SimpleMesh back(off);
        bool same = back.triangles() == mesh.triangles();
        for (size_t i = 0; same && i < mesh.vertices().size(); i++)
            same = std::fabs(back.vertices()[i] - mesh.vertices()[i]) <= 1e-5f * (1.0f + std::fabs(mesh.vertices()[i]));
        expect(same, "OFF round trip (6 significant digits)");
        plane.dumpToBinFile(out + "/plane.sdf.bin", 3.0, 0.0);
        a.dumpToBinFile(out + "/field.bin");
        std::ifstream fin(out + "/plane.sdf.bin", std::ios::binary | std::ios::ate);
        expect((size_t)fin.tellg() == 3 * 4 + 5 * 8 + n * 8, "SDF .bin size");
        std::ifstream ffin(out + "/field.bin", std::ios::binary | std::ios::ate);
        expect((size_t)ffin.tellg() == 3 * 4 + 8 + n * 24, "field .bin size");
    } catch (const std::exception &e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    if (g_fails) {
        std::printf("FAILED %d of %d checks\n", g_fails, g_checks);
        return 1;
    }
    std::printf("OK %d checks\n", g_checks);
    return 0;
}
