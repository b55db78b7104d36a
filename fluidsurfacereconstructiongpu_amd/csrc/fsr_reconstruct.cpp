// fsr_reconstruct: the reference's frame-sequence driver on libfsr.
//
// Mirrors ReconstructorGPU::reconstruct (GPU/ReconstructorGPU.cpp:25-79): for
// every frame in [from, to) it reads <directory><pattern % frame>.xyz, derives
// the grids and constants as the chosen reconstructor class does, times the
// reconstruction from spatial hashing to the end of marching cubes, writes the
// mesh as <...>.obj, with --vis writes the frame's intermediates as <...>.vis
// (saveMiddleDataToVisFile), and at the end appends the per-frame times to
// "<algorithm type>.time" and, with --config, writes "<algorithm type>.config"
// (GPU/ReconstructorGPU.cpp:315-393: mean triangle count, surface particle and
// surface vertex ratios, cumulative stage times).
//
// Frames are pipelined (the reference runs them strictly one after another):
// while frame k is reconstructed and its files are written, a host thread
// parses frame k+1's .xyz into a pinned buffer, which is then uploaded on a
// copy stream; --serial runs them strictly one after another instead
// (read, upload, reconstruct, write), as the reference does.  Per-frame times
// cover the reconstruction only, as the reference's timer does; the end-to-end
// rate, from the first read to the last write, is printed at the end.
//
//   fsr_reconstruct <directory/> <pattern> <from> <to>
//                   [--method zb05|akinci12|yu13] [--no-obj] [--indexed]
//                   [--vis] [--config] [--serial] [--device N]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <future>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/fsr.h"

namespace {

#define CHECK_FSR(expr)                                                                  \
    do {                                                                                 \
        fsr_status _s = (expr);                                                          \
        if (_s != FSR_OK) {                                                              \
            std::fprintf(stderr, "%s failed: %s (%s)\n", #expr, fsr_status_string(_s),   \
                         fsr_last_error());                                              \
            std::exit(1);                                                                \
        }                                                                                \
    } while (0)

#define CHECK_HIP(expr)                                                                  \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            std::fprintf(stderr, "%s failed: %s\n", #expr, hipGetErrorString(_e));       \
            std::exit(1);                                                                \
        }                                                                                \
    } while (0)

// Device buffer that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void* get(size_t want) {
        if (want > bytes) {
            if (p) CHECK_HIP(hipFree(p));
            CHECK_HIP(hipMalloc(&p, want));
            bytes = want;
        }
        return p;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Host copy of one parsed frame, in pinned memory that only grows.
struct HostFrame {
    std::string stem, base;
    fsr_xyz_header h{};
    uint32_t n = 0;
    float* pos = nullptr;
    float* dens = nullptr;
    uint32_t cap = 0;
    std::string error;  // empty when the frame was read
    ~HostFrame() {
        if (pos) (void)hipHostFree(pos);
        if (dens) (void)hipHostFree(dens);
    }
};

// Parses <dir><pattern % frame>.xyz into f (runs on the loader thread).
void load_frame(HostFrame& f, const std::string& dir, const std::string& pattern, long frame) {
    char base[512];
    std::snprintf(base, sizeof(base), pattern.c_str(), (unsigned)frame);
    f.base = base;
    f.stem = dir + base;
    f.error.clear();
    const std::string path = f.stem + ".xyz";
    uint32_t n = 0;
    fsr_status st = fsr_xyz_info(path.c_str(), &f.h, &n);
    if (st == FSR_OK && n > f.cap) {
        if (f.pos) (void)hipHostFree(f.pos);
        if (f.dens) (void)hipHostFree(f.dens);
        f.pos = f.dens = nullptr;
        f.cap = 0;
        if (hipHostMalloc(reinterpret_cast<void**>(&f.pos), 12ull * n) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&f.dens), 4ull * n) != hipSuccess) {
            f.error = "hipHostMalloc failed";
            return;
        }
        f.cap = n;
    }
    if (st == FSR_OK) st = fsr_xyz_read(path.c_str(), &f.h, f.pos, f.dens, f.cap, &f.n);
    if (st != FSR_OK) f.error = std::string(fsr_status_string(st)) + " (" + fsr_last_error() + ")";
}

const char* algorithm_type(fsr_method m) {
    // getAlgorithmType of the three reconstructor classes
    switch (m) {
        case FSR_METHOD_ZB05: return "Our Algorithm using ZB05 kernel";
        case FSR_METHOD_AKINCI12: return "Akinci12 using ZB05 kernel";
        default: return "Our Algorithm using anisotropic kernel";
    }
}

void usage() {
    std::fprintf(stderr,
                 "usage: fsr_reconstruct <directory/> <pattern> <from> <to> "
                 "[--method zb05|akinci12|yu13] [--no-obj] [--indexed] [--vis] [--config] "
                 "[--serial] [--device N]\n");
    std::exit(2);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) usage();
    const std::string dir = argv[1];
    const std::string pattern = argv[2];
    const long from = std::strtol(argv[3], nullptr, 10), to = std::strtol(argv[4], nullptr, 10);
    fsr_method method = FSR_METHOD_ZB05;
    bool save_obj = true, indexed = false, save_cfg = false, save_vis = false, serial = false;
    int device = 0;
    for (int i = 5; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--method" && i + 1 < argc) {
            const std::string m = argv[++i];
            if (m == "zb05") method = FSR_METHOD_ZB05;
            else if (m == "akinci12") method = FSR_METHOD_AKINCI12;
            else if (m == "yu13") method = FSR_METHOD_YU13;
            else usage();
        } else if (a == "--no-obj") {
            save_obj = false;
        } else if (a == "--indexed") {
            indexed = true;
        } else if (a == "--config") {
            save_cfg = true;
        } else if (a == "--vis") {
            save_vis = true;
        } else if (a == "--serial") {
            serial = true;
        } else if (a == "--device" && i + 1 < argc) {
            device = std::atoi(argv[++i]);
        } else {
            usage();
        }
    }
    if (from >= to) {
        std::printf("Wrong settings occur for frame range.\n");
        return 1;
    }
    fsr_context* ctx = nullptr;
    CHECK_FSR(fsr_create(&ctx, device));
    CHECK_FSR(fsr_set_profiling(ctx, 1));
    hipStream_t compute = nullptr, copy = nullptr;
    CHECK_HIP(hipStreamCreateWithFlags(&compute, hipStreamNonBlocking));
    CHECK_HIP(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
    hipEvent_t uploaded[2];
    for (auto& e : uploaded) CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HostFrame host[2];
    DevBuf d_pos[2], d_dens[2], d_phi, d_tri, d_nrm, d_idx, d_stats;
    std::vector<double> times;
    double sum_tri = 0, sum_spr = 0, sum_svr = 0, sum_es = 0, sum_sf = 0, sum_mc = 0;
    uint32_t cap = 0, vcap = 0;
    // frame `from`: read and uploaded before the loop
    auto upload = [&](int slot) {
        const HostFrame& f = host[slot];
        const uint32_t n = f.n;
        float* pp = static_cast<float*>(d_pos[slot].get(12ull * (n ? n : 1)));
        float* pd = static_cast<float*>(d_dens[slot].get(4ull * (n ? n : 1)));
        if (n) {
            CHECK_HIP(hipMemcpyAsync(pp, f.pos, 12ull * n, hipMemcpyHostToDevice, copy));
            CHECK_HIP(hipMemcpyAsync(pd, f.dens, 4ull * n, hipMemcpyHostToDevice, copy));
        }
        CHECK_HIP(hipEventRecord(uploaded[slot], copy));
    };
    auto check_loaded = [&](const HostFrame& f) {
        if (!f.error.empty()) {
            std::fprintf(stderr, "reading %s.xyz failed: %s\n", f.stem.c_str(), f.error.c_str());
            std::exit(1);
        }
    };
    const auto wall0 = std::chrono::steady_clock::now();
    load_frame(host[0], dir, pattern, from);
    check_loaded(host[0]);
    upload(0);
    for (long frame = from; frame < to; ++frame) {
        const int slot = serial ? 0 : (int)((frame - from) & 1);
        HostFrame& hf = host[slot];
        if (serial && frame != from) {
            load_frame(hf, dir, pattern, frame);
            check_loaded(hf);
            upload(slot);
        }
        // pipelined: the next frame is parsed while this one runs
        std::future<void> next;
        if (!serial && frame + 1 < to)
            next = std::async(std::launch::async, load_frame, std::ref(host[slot ^ 1]), dir,
                              pattern, frame + 1);
        const std::string& stem = hf.stem;
        const char* base = hf.base.c_str();
        std::printf("\n*************************** Frame %ld ***************************\n", frame);
        const fsr_xyz_header h = hf.h;
        const uint32_t n = hf.n;
        fsr_sim_param sim;
        CHECK_FSR(fsr_sim_param_for(method, &h, &sim));
        fsr_grid_info sp, sc;
        CHECK_FSR(fsr_grid_from_aabb(h.aabb_min, h.aabb_max, sim.smoothing_radius,
                                     sim.spatial_cell_size_scale, sim.sc_sp_grid_res_ratio, &sp,
                                     &sc));
        const uint64_t V = (uint64_t)sc.resolution.x * sc.resolution.y * sc.resolution.z;
        std::printf("Number of particles: %g k.\n", n / 1000.0);
        std::printf("Scalar Field Resolution: %u * %u * %u\n", sc.resolution.x, sc.resolution.y,
                    sc.resolution.z);
        if (!cap) cap = 6u * n + 3072u;
        if (!vcap) vcap = 2u * n + 1024u;
        float* pp = static_cast<float*>(d_pos[slot].get(12ull * (n ? n : 1)));
        float* pd = static_cast<float*>(d_dens[slot].get(4ull * (n ? n : 1)));
        float* phi = static_cast<float*>(d_phi.get(4ull * V));
        fsr_frame_stats* st = static_cast<fsr_frame_stats*>(d_stats.get(sizeof(fsr_frame_stats)));
        fsr_frame_stats hs;
        double ms = 0;
        for (int attempt = 0; attempt < 2; ++attempt) {
            if (attempt) {
                // the frame sorted its particles in place: upload them again
                upload(slot);
            }
            CHECK_HIP(hipStreamWaitEvent(compute, uploaded[slot], 0));
            CHECK_HIP(hipEventSynchronize(uploaded[slot]));
            const auto t0 = std::chrono::steady_clock::now();
            if (indexed) {
                CHECK_FSR(fsr_reconstruct_frame_indexed(
                    ctx, method, &sim, nullptr, &sp, &sc, pp, pd, n, phi,
                    static_cast<float*>(d_tri.get(12ull * vcap)),
                    static_cast<float*>(d_nrm.get(12ull * vcap)), vcap,
                    static_cast<uint32_t*>(d_idx.get(4ull * cap)), cap, st, compute));
            } else {
                CHECK_FSR(fsr_reconstruct_frame(ctx, method, &sim, nullptr, &sp, &sc, pp, pd, n,
                                                phi, static_cast<float*>(d_tri.get(12ull * cap)),
                                                static_cast<float*>(d_nrm.get(12ull * cap)), cap,
                                                st, compute));
            }
            CHECK_HIP(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, compute));
            CHECK_HIP(hipStreamSynchronize(compute));
            const auto t1 = std::chrono::steady_clock::now();
            ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            if (!hs.overflow) break;
            cap = hs.num_mesh_vertices + 3072u;
            vcap = hs.num_unique_vertices + 1024u;
        }
        float stage[FSR_NUM_STAGES];
        uint32_t nf = 0;
        CHECK_FSR(fsr_stage_times(ctx, stage, 1, &nf));
        times.push_back(ms);
        std::printf("Reconstructing Frame%ld took %g milliseconds.\n", frame, ms);
        if (save_vis) {
            const std::string path = stem + ".vis";
            std::printf("Writing to %s...\n", path.c_str());
            CHECK_FSR(fsr_vis_write(ctx, path.c_str(), method, &sp, &sc, &sim, pp, n, &hs,
                                    (std::string(base) + ".obj").c_str()));
            std::printf("Finish writing %s.\n", path.c_str());
        }
        const uint32_t T3 = hs.num_mesh_vertices;
        sum_tri += T3 / 3;
        sum_spr += n ? (double)hs.num_surface_particles / n : 0.0;
        sum_svr += V ? (double)hs.num_band_vertices / (double)V : 0.0;
        if (nf) {
            sum_es += stage[FSR_STAGE_SURFACE] + stage[FSR_STAGE_BAND];
            sum_sf += stage[FSR_STAGE_PREP] + stage[FSR_STAGE_FIELD];
            sum_mc += stage[FSR_STAGE_MC];
        }
        if (save_obj && T3) {
            const std::string path = stem + ".obj";
            if (indexed) {
                const uint32_t U = hs.num_unique_vertices;
                std::vector<float> vp(3ull * U), vn(3ull * U);
                std::vector<uint32_t> idx(T3);
                CHECK_HIP(hipMemcpy(vp.data(), d_tri.p, 12ull * U, hipMemcpyDeviceToHost));
                CHECK_HIP(hipMemcpy(vn.data(), d_nrm.p, 12ull * U, hipMemcpyDeviceToHost));
                CHECK_HIP(hipMemcpy(idx.data(), d_idx.p, 4ull * T3, hipMemcpyDeviceToHost));
                CHECK_FSR(fsr_obj_write_indexed(path.c_str(), vp.data(), vn.data(), U, idx.data(),
                                                T3));
            } else {
                std::vector<float> tp(3ull * T3), tn(3ull * T3);
                CHECK_HIP(hipMemcpy(tp.data(), d_tri.p, 12ull * T3, hipMemcpyDeviceToHost));
                CHECK_HIP(hipMemcpy(tn.data(), d_nrm.p, 12ull * T3, hipMemcpyDeviceToHost));
                CHECK_FSR(fsr_obj_write_soup(path.c_str(), tp.data(), tn.data(), T3));
            }
            std::printf("Writing to %s\n", path.c_str());
        }
        // the next frame was parsed while this one ran and its files were
        // written; its upload overlaps the start of the next iteration
        if (next.valid()) {
            next.get();
            check_loaded(host[slot ^ 1]);
            upload(slot ^ 1);
        }
    }
    const double wall_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    std::printf("\n%zu frame(s) in %g ms end to end (%g frames/s; reading, uploading and writing "
                "included).\n", times.size(), wall_ms, times.size() * 1000.0 / wall_ms);
    const double nfr = (double)times.size();
    {
        const std::string path = dir + algorithm_type(method) + ".time";
        std::ofstream f(path, std::ios::out | std::ios::app);
        if (f) {
            for (double t : times) f << t << ' ';
            f << std::endl;
        } else {
            std::fprintf(stderr, "Failed to save the file: %s\n", path.c_str());
        }
    }
    if (save_cfg) {
        const std::string path = dir + algorithm_type(method) + ".config";
        std::ofstream f(path);
        const double es = sum_es / nfr, sf = sum_sf / nfr, tri = sum_mc / nfr;
        if (f) {
            f << (long long)(sum_tri / nfr) << std::endl;
            f << sum_spr / nfr << std::endl;
            f << sum_svr / nfr << std::endl;
            f << es << std::endl;
            f << es + sf << std::endl;
            f << tri << std::endl;
            f << es + sf + tri << std::endl;
        } else {
            std::fprintf(stderr, "Failed to save the file: %s\n", path.c_str());
        }
    }
    for (auto& e : uploaded) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(copy);
    (void)hipStreamDestroy(compute);
    fsr_destroy(ctx);
    return 0;
}
