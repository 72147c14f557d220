// Drives the reference-named C++ layer (semantic_icpslam_amd/include/sicp_ref)
// the way the reference's callers do (Core/Frame.cc:68-98,
// Tracking/Tracking.cc:433-447, Mapping/Mapping.cc:76-161), on host inputs,
// and hands the results back to the Python parity tests.
#include <cstring>
#include <vector>

#include "Reduction.h"
#include "RenderScene.h"

namespace {

template <class T> void down(const DeviceArray2D<T>& a, T* out) {
    if (out) a.download(out, sizeof(T) * a.cols);
}

Matrix3f mat(const float* R) {
    Matrix3f m;
    m.rowx = make_float3(R[0], R[1], R[2]);
    m.rowy = make_float3(R[3], R[4], R[5]);
    m.rowz = make_float3(R[6], R[7], R[8]);
    return m;
}

template <class F> int guarded(F f) {
    SafeCallModeRef() = SafeCallMode::Throw;
    try {
        f();
    } catch (const SafeCallError& e) {
        return e.status;
    }
    return 0;
}

}  // namespace

extern "C" {

// Frame::FillImages + ResizeImages (Frame.cc:68-98) through the mirror.
// Level-l outputs are (cols >> l) x (rows >> l); any output may be NULL.
int mirror_frame(const uint16_t* depth_h, const uint8_t* rgb_h, int cols, int rows, float fx,
                 float fy, float cx, float cy, float depth_scale, float cutoff, float* raw,
                 float* depth_l[3], uint8_t* gray_l[3], int16_t* dx_l[3], int16_t* dy_l[3],
                 float* vmap_l[3], float* nmap_l[3], float* vres_l[3], float* nres_l[3]) {
    return guarded([&] {
        DeviceArray2D<unsigned short> depth(cols, rows);
        depth.upload(depth_h, sizeof(uint16_t) * cols);
        DeviceArray2D<uchar3> rgb(cols, rows);
        rgb.upload(rgb_h, 3 * cols);
        DeviceArray2D<float> range, d[3];
        DeviceArray2D<unsigned char> img[3];
        DeviceArray2D<short> dx[3], dy[3];
        DeviceArray2D<float4> vmap[3], nmap[3], vres[3], nres[3];
        Intrinsics K(fx, fy, cx, cy);
        FilterDepth(depth, range, d[0], depth_scale, cutoff);
        ImageToIntensity(rgb, img[0]);
        for (int i = 1; i < 3; ++i) {
            PyrDownGauss(d[i - 1], d[i]);
            PyrDownGauss(img[i - 1], img[i]);
        }
        for (int i = 0; i < 3; ++i) {
            Intrinsics k = K(i);
            ComputeVMap(d[i], vmap[i], k.fx, k.fy, k.cx, k.cy, cutoff);
            ComputeNMap(vmap[i], nmap[i]);
            ComputeDerivativeImage(img[i], dx[i], dy[i]);
        }
        vres[0] = vmap[0];
        nres[0] = nmap[0];
        for (int i = 1; i < 3; ++i) ResizeMap(vres[i - 1], nres[i - 1], vres[i], nres[i]);
        down(range, raw);
        for (int i = 0; i < 3; ++i) {
            down(d[i], depth_l[i]);
            down(img[i], gray_l[i]);
            down(dx[i], dx_l[i]);
            down(dy[i], dy_l[i]);
            if (vmap_l[i]) vmap[i].download(vmap_l[i], 16 * vmap[i].cols);
            if (nmap_l[i]) nmap[i].download(nmap_l[i], 16 * nmap[i].cols);
            if (vres_l[i]) vres[i].download(vres_l[i], 16 * vres[i].cols);
            if (nres_l[i]) nres[i].download(nres_l[i], 16 * nres[i].cols);
        }
    });
}

// One ICPStep (Tracking.cc:433-447) through the mirror: pose_curr / pose_last
// are {R row-major[9], t[3]}; RlastInv is the transpose, as Frame keeps it.
int mirror_icp(const float* vcurr, const float* vlast, const float* ncurr, const float* nlast,
               int cols, int rows, float fx, float fy, float cx, float cy, const float* pose_curr,
               const float* pose_last, double* A, double* b, float* residual) {
    return guarded([&] {
        DeviceArray2D<float4> vc(cols, rows), vl(cols, rows), nc(cols, rows), nl(cols, rows);
        vc.upload(vcurr, 16 * cols);
        vl.upload(vlast, 16 * cols);
        nc.upload(ncurr, 16 * cols);
        nl.upload(nlast, 16 * cols);
        DeviceArray2D<float> sum;
        DeviceArray<float> out;
        Matrix3f Rc = mat(pose_curr), Rl = mat(pose_last);
        ICPStep(vc, vl, nc, nl, Rc, make_float3(pose_curr[9], pose_curr[10], pose_curr[11]), Rl,
                sicp_transpose(Rl), make_float3(pose_last[9], pose_last[10], pose_last[11]),
                Intrinsics(fx, fy, cx, cy), sum, out, residual, A, b);
    });
}

// Tracker::ComputeSE3's RGBStep call (Tracking.cc:402-423) through the
// mirror: the photometric sums of one iteration, and the correspondence
// count / sum of squared differences read back from outRes.
int mirror_rgb(const uint8_t* nimg, const uint8_t* limg, const float* nvmap, const float* lvmap,
               const int16_t* dx, const int16_t* dy, int cols, int rows, float fx, float fy,
               float cx, float cy, const float* pose_curr, const float* pose_last, double* A,
               double* b, float* residual, int* corresp) {
    return guarded([&] {
        DeviceArray2D<unsigned char> ni(cols, rows), li(cols, rows);
        DeviceArray2D<float4> nv(cols, rows), lv(cols, rows);
        DeviceArray2D<short> gx(cols, rows), gy(cols, rows);
        ni.upload(nimg, cols);
        li.upload(limg, cols);
        nv.upload(nvmap, 16 * cols);
        lv.upload(lvmap, 16 * cols);
        gx.upload(dx, 2 * cols);
        gy.upload(dy, 2 * cols);
        DeviceArray2D<float> sum;
        DeviceArray<float> out;
        DeviceArray2D<int> sumRes;
        DeviceArray<int> outRes;
        Matrix3f Rc = mat(pose_curr), Rl = mat(pose_last);
        RGBStep(ni, li, nv, lv, gx, gy, Rc, sicp_transpose(Rc), Rl, sicp_transpose(Rl),
                make_float3(pose_curr[9], pose_curr[10], pose_curr[11]),
                make_float3(pose_last[9], pose_last[10], pose_last[11]),
                Intrinsics(fx, fy, cx, cy), sum, out, sumRes, outRes, residual, A, b);
        outRes.download(corresp, 2);
    });
}

// Tracker::ComputeSO3's SO3Step call (Tracking.cc:339-350) through the
// mirror.
int mirror_so3(const uint8_t* nimg, const uint8_t* limg, const int16_t* dx, const int16_t* dy,
               int cols, int rows, float fx, float fy, float cx, float cy,
               const float* pose_curr, const float* pose_last, double* A, double* b,
               float* residual) {
    return guarded([&] {
        DeviceArray2D<unsigned char> ni(cols, rows), li(cols, rows);
        DeviceArray2D<short> gx(cols, rows), gy(cols, rows);
        ni.upload(nimg, cols);
        li.upload(limg, cols);
        gx.upload(dx, 2 * cols);
        gy.upload(dy, 2 * cols);
        DeviceArray2D<float> sum;
        DeviceArray<float> out;
        Matrix3f Rc = mat(pose_curr), Rl = mat(pose_last);
        SO3Step(ni, li, gx, gy, sicp_transpose(Rc), Rl, Intrinsics(fx, fy, cx, cy), sum, out,
                residual, A, b);
    });
}

// Mapping::FuseColor then Mapping::RayTrace (Mapping.cc:76-97,142-161)
// through the mirror, for `n` frames of (raw depth, RGB, normals, labels,
// pose {R[9], t[3]}) fused into a fresh map; then the rendering prepass and
// the raycast at the last pose. Outputs: visible count of each fusion, the
// visible count of a separate CheckBlockVisibility, zRange, vmap / nmap.
int mirror_map(const sicp_map_config* cfg, int n, const float* depth, const uint8_t* rgb,
               const float* nmap_in, const uint8_t* mask, const float* poses, int cols, int rows,
               float fx, float fy, float cx, float cy, uint32_t* visible_out,
               uint32_t* check_visible, int* rendered, float* zmin, float* zmax, float* vmap,
               float* nmap, uint32_t max_tris, float* mesh_v, uint8_t* mesh_c,
               uint32_t* n_tris) {
    return guarded([&] {
        DeviceMap map(*cfg);
        ResetMap(map);
        DeviceArray2D<float> d(cols, rows);
        DeviceArray2D<uchar3> c(cols, rows);
        DeviceArray2D<float4> nm(cols, rows);
        DeviceArray2D<unsigned char> mk(cols, rows);
        DeviceArray<int> labels;
        DeviceArray<uint> noVisible(1);
        Matrix3f R, Rinv;
        float3 t = make_float3(0, 0, 0);
        const size_t px = (size_t)cols * rows;
        for (int f = 0; f < n; ++f) {
            d.upload(depth + f * px, 4 * cols);
            c.upload(rgb + f * px * 3, 3 * cols);
            nm.upload(nmap_in + f * px * 4, 16 * cols);
            mk.upload(mask + f * px, cols);
            const float* p = poses + f * 12;
            R = mat(p);
            Rinv = sicp_transpose(R);
            t = make_float3(p[9], p[10], p[11]);
            uint host = 0;
            FuseMapColor(d, c, nm, mk, labels, 1, noVisible, R, Rinv, t, map, fx, fy, cx, cy,
                         cfg->depth_max, cfg->depth_min, &host);
            visible_out[f] = host;
        }
        uint host = 0;
        CheckBlockVisibility(map, noVisible, R, Rinv, t, cols, rows, fx, fy, cx, cy,
                             cfg->depth_max, cfg->depth_min, &host);
        *check_visible = host;
        const int zc = (cols + 7) / 8, zr = (rows + 7) / 8;
        DeviceArray2D<float> zx(zc, zr), zy(zc, zr);
        DeviceArray<RenderingBlock> list;
        DeviceArray<uint> nlist;
        *rendered = CreateRenderingBlocks(map.visibleEntries, zx, zy, cfg->depth_max,
                                          cfg->depth_min, list, nlist, Rinv, t, host, fx, fy, cx,
                                          cy)
                        ? 1
                        : 0;
        zx.download(zmin, 4 * zc);
        zy.download(zmax, 4 * zc);
        DeviceArray2D<float4> vm(cols, rows), nv(cols, rows);
        Raycast(map, vm, nv, zx, zy, R, Rinv, t, (float)(1.0 / fx), (float)(1.0 / fy), cx, cy);
        vm.download(vmap, 16 * cols);
        nv.download(nmap, 16 * cols);
        // System::WriteMeshToDisk's CreateModel (Mapping.cc:183-192)
        DeviceArray<uint> nOcc(1), nTri(1);
        DeviceArray<int> edgeTable, vertexTable;
        DeviceArray2D<int> triangleTable;
        DeviceArray<float3> mv(3 * max_tris), mn(3 * max_tris);
        DeviceArray<uchar3> mc(3 * max_tris);
        DeviceArray<int3> blockPoses;
        *n_tris = MeshScene(nOcc, nTri, map, edgeTable, vertexTable, triangleTable, mn, mv, mc,
                            blockPoses);
        mv.download(mesh_v, 3 * (size_t)*n_tris);
        mc.download(mesh_c, 3 * (size_t)*n_tris);
    });
}

}  // extern "C"
