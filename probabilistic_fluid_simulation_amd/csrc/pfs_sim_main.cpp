// The simulation driver of the reference (src/main.cpp:71-280) on the GPU:
// the same command line, the same input files and messages, every step one
// timestep of src/main.cpp:222-225, every frame <dir>/<step>.png.  The steps
// up to the next written frame (all of them without frames) are one
// pfs_fluid_run call, replayed from hipGraphs when there are enough.
//
// Frames leave the device without stalling the steps: after step i the
// current image is copied on a second stream into one of a ring of pinned
// host buffers, and a writer thread encodes it while the GPU runs on.  Steps
// that would overwrite an image buffer (image and itmp swap every step)
// first wait, on the device, for the copy that reads it.
//
// Besides the reference's arguments:
//   --iters K     Jacobi sweeps per solve (the reference's compile-time
//                 NUM_JACOBI_ITERS, 30)
//   --every N     write every N-th frame only (default 1)
//   --npy         frames as .npy float32 [h][w][4] instead of PNG
//   --dump DIR    after the run, write vp, tmp, image (current) and itmp as
//                 DIR/{vp,tmp,image,itmp}.npy
//   --per-op      one kernel per operator and sweep (PFS_PATH_PER_OP)
//   --buffers N   frames in flight between the GPU and the writer (4)
//   --quiet       no per-frame lines
// Inputs ending in .npy are read as float32 [h][w][4]; a .npy velocity
// field is the rescaled field itself.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <iostream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pfs.h"

namespace {

constexpr int kChannels = 4;      // src/main.cpp:25
constexpr double kVpRange = 2.0;  // src/main.cpp:26

struct HostField {
    int w = 0, h = 0;
    std::vector<float> data;  // [h][w][4]
};

// ---- .npy (numpy format 1.0, '<f4', C order, shape (h, w, 4)) -------------
bool write_npy(const std::string &path, const float *data, int w, int h) {
    std::string hdr = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(h) +
                      ", " + std::to_string(w) + ", 4), }";
    const size_t pre = 10;  // magic(6) + version(2) + header length(2)
    size_t total = pre + hdr.size() + 1;
    hdr.append((64 - total % 64) % 64, ' ');
    hdr.push_back('\n');
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    const unsigned short hl = (unsigned short)hdr.size();
    const unsigned char hlb[2] = {(unsigned char)(hl & 0xff), (unsigned char)(hl >> 8)};
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hlb, 1, 2, f) == 2 &&
              std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    const size_t n = (size_t)w * h * kChannels;
    ok = ok && std::fwrite(data, sizeof(float), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

bool read_npy(const std::string &path, HostField *out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    unsigned char pre[10];
    bool ok = std::fread(pre, 1, 10, f) == 10 && std::memcmp(pre, "\x93NUMPY", 6) == 0 && pre[6] == 1;
    std::string hdr;
    if (ok) {
        hdr.resize(pre[8] | (pre[9] << 8));
        ok = std::fread(&hdr[0], 1, hdr.size(), f) == hdr.size();
    }
    int h = 0, w = 0, c = 0;
    if (ok) {
        const size_t s = hdr.find("'shape': (");
        ok = hdr.find("'descr': '<f4'") != std::string::npos &&
             hdr.find("'fortran_order': False") != std::string::npos && s != std::string::npos &&
             std::sscanf(hdr.c_str() + s + 10, "%d, %d, %d", &h, &w, &c) == 3 && c == kChannels &&
             h > 0 && w > 0;
    }
    if (ok) {
        out->w = w;
        out->h = h;
        out->data.resize((size_t)w * h * kChannels);
        ok = std::fread(out->data.data(), sizeof(float), out->data.size(), f) == out->data.size();
    }
    std::fclose(f);
    return ok;
}

bool ends_with(const std::string &s, const char *suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// Reads an input file: PNG (byte/255.0, includes/utils.hpp:32-104) or .npy.
bool read_field(const std::string &path, HostField *out) {
    if (ends_with(path, ".npy")) return read_npy(path, out);
    if (pfs_png_size(path.c_str(), &out->w, &out->h) != PFS_OK) return false;
    out->data.resize((size_t)out->w * out->h * kChannels);
    return pfs_png_read(path.c_str(), out->data.data(), out->data.size()) == PFS_OK;
}

std::string join(const std::string &dir, const std::string &name) {
    if (dir.empty() || dir.back() == '/') return dir + name;
    return dir + "/" + name;
}

// ---- frame writer ----------------------------------------------------------
struct Writer {
    int w, h, nbuf;
    bool npy, quiet;
    std::string dir;
    std::vector<float *> host;       // pinned, one image each
    std::vector<hipEvent_t> copied;  // recorded after the copy into host[b]
    std::vector<int> busy;           // slot holds a frame not yet written
    std::deque<std::pair<int, int>> queue;  // (slot, step)
    std::mutex mu;
    std::condition_variable cv;
    bool done = false, failed = false;
    std::string error;
    std::thread thread;

    int acquire() {  // a free slot, waiting for the writer if needed
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            for (int b = 0; b < nbuf; ++b)
                if (!busy[b]) {
                    busy[b] = 1;
                    return b;
                }
            cv.wait(lk);
        }
    }
    void submit(int slot, int step) {
        {
            std::lock_guard<std::mutex> lk(mu);
            queue.emplace_back(slot, step);
        }
        cv.notify_all();
    }
    void run() {
        for (;;) {
            std::pair<int, int> job;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return done || !queue.empty(); });
                if (queue.empty()) return;
                job = queue.front();
                queue.pop_front();
            }
            const int b = job.first;
            bool ok = hipEventSynchronize(copied[b]) == hipSuccess;
            const std::string path =
                join(dir, std::to_string(job.second) + (npy ? ".npy" : ".png"));
            if (!quiet) std::cout << "[" << job.second << "] Writing to : " << path << std::endl;
            if (ok) ok = npy ? write_npy(path, host[b], w, h)
                             : pfs_png_write(path.c_str(), host[b], w, h) == PFS_OK;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!ok && !failed) {
                    failed = true;
                    error = "writing " + path + " failed";
                }
                busy[b] = 0;
            }
            cv.notify_all();
        }
    }
    void finish() {
        {
            std::lock_guard<std::mutex> lk(mu);
            done = true;
        }
        cv.notify_all();
        if (thread.joinable()) thread.join();
    }
};

void usage(const char *prog) {
    std::cerr << "Usage: " << prog
              << " <n_timesteps> <delta_t> <viscosity> <input_image> <velocity_field> [output_dir]"
              << " [--iters K] [--every N] [--npy] [--dump DIR] [--per-op] [--buffers N] [--quiet]"
              << std::endl;
}

#define SIM_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            std::cerr << #expr << ": " << hipGetErrorString(e_) << std::endl;              \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)
#define SIM_PFS(expr)                                                                      \
    do {                                                                                   \
        if ((expr) != PFS_OK) {                                                            \
            std::cerr << #expr << ": " << pfs_last_error() << std::endl;                   \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

struct Device {
    float *p[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t comp = nullptr, copy = nullptr;
    std::vector<hipEvent_t> events;
    ~Device() {
        if (comp) (void)hipStreamSynchronize(comp);
        if (copy) (void)hipStreamSynchronize(copy);
        for (float *q : p)
            if (q) (void)hipFree(q);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (comp) (void)hipStreamDestroy(comp);
        if (copy) (void)hipStreamDestroy(copy);
    }
};

int run(int argc, char **argv) {
    std::vector<std::string> pos;
    int iters = 30, every = 1, nbuf = 4;
    bool npy = false, per_op = false, quiet = false;
    std::string dump;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](int *dst) {
            if (i + 1 >= argc) return false;
            *dst = std::atoi(argv[++i]);
            return true;
        };
        if (a == "--iters") {
            if (!value(&iters)) return usage(argv[0]), 1;
        } else if (a == "--every") {
            if (!value(&every)) return usage(argv[0]), 1;
        } else if (a == "--buffers") {
            if (!value(&nbuf)) return usage(argv[0]), 1;
        } else if (a == "--dump") {
            if (i + 1 >= argc) return usage(argv[0]), 1;
            dump = argv[++i];
        } else if (a == "--npy") {
            npy = true;
        } else if (a == "--per-op") {
            per_op = true;
        } else if (a == "--quiet") {
            quiet = true;
        } else if (a.size() > 2 && a.compare(0, 2, "--") == 0) {
            return usage(argv[0]), 1;
        } else {
            pos.push_back(a);
        }
    }
    // src/main.cpp:104-123
    if (pos.size() != 5 && pos.size() != 6) return usage(argv[0]), 1;
    const bool timing_only = pos.size() == 5;
    const int n_timesteps = std::atoi(pos[0].c_str());
    const float delta_t = (float)std::atof(pos[1].c_str());
    const float viscosity = (float)std::atof(pos[2].c_str());
    if (n_timesteps <= 0) {
        std::cerr << "Timesteps must be greater than 0." << std::endl;
        return 1;
    }
    if (delta_t <= 0) {
        std::cerr << "Delta T must be greater than 0." << std::endl;
        return 1;
    }
    if (iters < 0 || every < 1 || nbuf < 1) return usage(argv[0]), 1;

    // src/main.cpp:126-140
    HostField image, vp;
    if (!read_field(pos[3], &image)) {
        std::cerr << "Something went wrong reading the input image..." << std::endl;
        return 1;
    }
    if (!read_field(pos[4], &vp)) {
        std::cerr << "Something went wrong reading the initial velocity field..." << std::endl;
        return 1;
    }
    // src/main.cpp:170-179: every channel rescaled to [-1, 1] (PNG input).
    if (!ends_with(pos[4], ".npy"))
        for (float &v : vp.data) v = (float)((double)v * kVpRange - kVpRange / 2.0);
    // src/main.cpp:184-195: tmp = (-1, -1, -1, +1); itmp is fully written
    // before it is read (zeroed here, uninitialised there).
    std::vector<float> vtmp(vp.data.size());
    for (size_t i = 0; i < vtmp.size(); ++i) vtmp[i] = (i % 4) == 3 ? 1.0f : -1.0f;

    const size_t ibytes = image.data.size() * sizeof(float);
    const size_t vbytes = vp.data.size() * sizeof(float);
    Device dev;
    SIM_HIP(hipStreamCreateWithFlags(&dev.comp, hipStreamNonBlocking));
    SIM_HIP(hipStreamCreateWithFlags(&dev.copy, hipStreamNonBlocking));
    SIM_HIP(hipMalloc((void **)&dev.p[0], vbytes));
    SIM_HIP(hipMalloc((void **)&dev.p[1], vbytes));
    SIM_HIP(hipMalloc((void **)&dev.p[2], ibytes));
    SIM_HIP(hipMalloc((void **)&dev.p[3], ibytes));
    SIM_HIP(hipMemcpy(dev.p[0], vp.data.data(), vbytes, hipMemcpyHostToDevice));
    SIM_HIP(hipMemcpy(dev.p[1], vtmp.data(), vbytes, hipMemcpyHostToDevice));
    SIM_HIP(hipMemcpy(dev.p[2], image.data.data(), ibytes, hipMemcpyHostToDevice));
    SIM_HIP(hipMemset(dev.p[3], 0, ibytes));
    pfs_field f_vp{vp.w, vp.h, kChannels, dev.p[0]}, f_tmp{vp.w, vp.h, kChannels, dev.p[1]};
    pfs_field f_img{image.w, image.h, kChannels, dev.p[2]};
    pfs_field f_itmp{image.w, image.h, kChannels, dev.p[3]};
    struct PathGuard {  // the step path is process-wide: restore it on every exit
        int prev = pfs_get_step_path();
        ~PathGuard() { (void)pfs_set_step_path(prev); }
    } path_guard;
    if (per_op) SIM_PFS(pfs_set_step_path(PFS_PATH_PER_OP));

    Writer wr{image.w, image.h, timing_only ? 0 : nbuf, npy, quiet, timing_only ? "" : pos[5]};
    wr.busy.assign(wr.nbuf, 0);
    for (int b = 0; b < wr.nbuf; ++b) {
        float *h = nullptr;
        SIM_HIP(hipHostMalloc((void **)&h, ibytes, hipHostMallocDefault));
        wr.host.push_back(h);
        hipEvent_t e;
        SIM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        wr.copied.push_back(e);
        dev.events.push_back(e);  // destroyed with the device state, after the writer
    }
    struct HostFree {
        std::vector<float *> *v;
        ~HostFree() {
            for (float *h : *v) (void)hipHostFree(h);
        }
    } host_free{&wr.host};
    hipEvent_t step_done;
    SIM_HIP(hipEventCreateWithFlags(&step_done, hipEventDisableTiming));
    dev.events.push_back(step_done);
    if (!timing_only) wr.thread = std::thread([&wr] { wr.run(); });
    struct Join {
        Writer *w;
        ~Join() { w->finish(); }
    } join_writer{&wr};

    // The slot whose copy last read each image buffer, -1: none.
    int reader_of[2] = {-1, -1};
    auto buffer_id = [&](const float *p) { return p == dev.p[2] ? 0 : 1; };

    std::cout << "Simulating [" << vp.h << " x " << vp.w << "] domain for " << n_timesteps
              << " timesteps at dt=" << delta_t << "..." << std::endl;
    const auto t0 = std::chrono::high_resolution_clock::now();
    for (int i = 0; i < n_timesteps;) {
        // Steps i..last run in one pfs_fluid_run call (hipGraph replay when
        // long): up to the next written frame, or to the end.
        int last = n_timesteps - 1;
        if (!timing_only) last = std::min(last, (i + every - 1) / every * every);
        const int count = last - i + 1;
        // The steps write itmp (one step) or both image buffers (more): wait
        // for the copies still reading them.
        for (int target = 0; target < 2; ++target)
            if (reader_of[target] >= 0 && (count > 1 || target == buffer_id(f_itmp.data))) {
                SIM_HIP(hipStreamWaitEvent(dev.comp, wr.copied[reader_of[target]], 0));
                reader_of[target] = -1;
            }
        SIM_PFS(pfs_fluid_run(&f_vp, &f_tmp, &f_img, &f_itmp, delta_t, viscosity, iters, count,
                              dev.comp));
        i = last + 1;
        if (timing_only || last % every != 0) continue;
        const int slot = wr.acquire();
        if (reader_of[0] == slot) reader_of[0] = -1;
        if (reader_of[1] == slot) reader_of[1] = -1;
        SIM_HIP(hipEventRecord(step_done, dev.comp));
        SIM_HIP(hipStreamWaitEvent(dev.copy, step_done, 0));
        SIM_HIP(hipMemcpyAsync(wr.host[slot], f_img.data, ibytes, hipMemcpyDeviceToHost, dev.copy));
        SIM_HIP(hipEventRecord(wr.copied[slot], dev.copy));
        reader_of[buffer_id(f_img.data)] = slot;
        wr.submit(slot, last);
    }
    SIM_HIP(hipStreamSynchronize(dev.comp));
    wr.finish();
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (wr.failed) {
        std::cerr << wr.error << std::endl;
        return 1;
    }
    std::cout << n_timesteps << " timesteps took "
              << std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() << " us."
              << std::endl;

    if (!dump.empty()) {
        const struct {
            const char *name;
            const pfs_field *f;
        } out[4] = {{"vp.npy", &f_vp}, {"tmp.npy", &f_tmp}, {"image.npy", &f_img},
                    {"itmp.npy", &f_itmp}};
        for (const auto &o : out) {
            std::vector<float> h((size_t)o.f->x * o.f->y * kChannels);
            SIM_HIP(hipMemcpy(h.data(), o.f->data, h.size() * sizeof(float), hipMemcpyDeviceToHost));
            if (!write_npy(join(dump, o.name), h.data(), o.f->x, o.f->y)) {
                std::cerr << "writing " << join(dump, o.name) << " failed" << std::endl;
                return 1;
            }
        }
    }
    return 0;
}

}  // namespace

extern "C" int pfs_sim_main(int argc, char **argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        std::cerr << "pfs_sim: " << e.what() << std::endl;
        return 1;
    }
}
