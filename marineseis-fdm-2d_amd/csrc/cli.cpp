// Process boundary: the bin/main_debug run (src/main.cu:58-339) on top of the C ABI.
//
// Output layout under <out_root> (cleared first, src/main.cu:82-89):
//   <shot>/snapshot/solid/{vz,vx}/{vz,vx}_%06d.bin   full [nz_s][nx] arrays
//   <shot>/snapshot/fluid/{vz,vx}/{vz,vx}_%06d.bin   full [nz_f][nx] arrays
//   <shot>/record/solid_vx.bin, solid_vz.bin          solid row 0, [nrec][nx]
//   <shot>/record/fluid_p.bin                         fluid row nz_f-1, [nrec][nx]
// A solid-only model (grid.calc_type = 0) writes the solid snapshots and
// <shot>/record/solid_vz.bin only, vz of the free-surface row 3
// (src/main.cu:120-128, 297-322, 327-339).
// Snapshots are taken after steps with it % snapshot == 0, records after steps
// with it % gap == 0 (src/main.cu:194-211).  Each shot starts from a zero
// wavefield (the reference carries the previous shot's field over).
//
// Snapshot files are written by a writer thread while the GPU steps on to the
// next snapshot (the reference synchronises and writes inline, src/main.cu:196-203);
// at most two snapshots wait in host memory.
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "model.hpp"

namespace fs = std::filesystem;
using namespace msfdm;

namespace {

struct ModelDel { void operator()(msfdm_model *m) const { msfdm_model_destroy(m); } };
struct SolverDel { void operator()(msfdm_solver *s) const { msfdm_solver_destroy(s); } };

int fail(const std::string &msg) {
    set_error(MSFDM_EIO, msg);
    return 1;
}

bool make_dirs(const fs::path &p, std::string &err) {
    std::error_code ec;
    fs::create_directories(p, ec);
    if (ec) {
        err = "cannot create " + p.string() + ": " + ec.message();
        return false;
    }
    return true;
}

// write_field, src/main.cu:225-243
bool write_field(const fs::path &dir, const char *field, int it, const std::vector<float> &data, std::string &err) {
    if (!make_dirs(dir, err)) return false;
    std::ostringstream name;
    name << field << "_" << std::setw(6) << std::setfill('0') << it << ".bin";
    const fs::path file = dir / name.str();
    std::ofstream fp(file, std::ios::binary);
    if (!fp) { err = "cannot write " + file.string(); return false; }
    fp.write(reinterpret_cast<const char *>(data.data()), (std::streamsize)(data.size() * sizeof(float)));
    if (!fp) { err = "cannot write " + file.string(); return false; }
    return true;
}

// Background writer of snapshot files: jobs are queued by the stepping thread
// (blocking while two are pending) and written in order; the first error stops
// the writer and is reported by push() / finish().
class SnapshotWriter {
public:
    struct File {
        fs::path dir;
        const char *field;
        int it;
        std::vector<float> data;
    };
    SnapshotWriter() : worker_([this] { run(); }) {}
    ~SnapshotWriter() { finish(); }
    bool push(std::vector<File> job, std::string &err) {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return q_.size() < 2 || !err_.empty(); });
        if (!err_.empty()) { err = err_; return false; }
        q_.push_back(std::move(job));
        cv_.notify_all();
        return true;
    }
    bool finish(std::string *err = nullptr) {
        {
            std::lock_guard<std::mutex> lk(m_);
            done_ = true;
        }
        cv_.notify_all();
        if (worker_.joinable()) worker_.join();
        if (err && !err_.empty()) *err = err_;
        return err_.empty();
    }

private:
    void run() {
        for (;;) {
            std::vector<File> job;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return !q_.empty() || done_; });
                if (q_.empty()) return;
                job = std::move(q_.front());
            }
            std::string e;
            for (auto &f : job)
                if (!write_field(f.dir, f.field, f.it, f.data, e)) break;
            std::lock_guard<std::mutex> lk(m_);
            q_.pop_front();
            if (!e.empty() && err_.empty()) err_ = e;
            cv_.notify_all();
            if (!err_.empty()) return;
        }
    }
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::vector<File>> q_;
    std::string err_;
    bool done_ = false;
    std::thread worker_;
};

}  // namespace

extern "C" int msfdm_run_cli(const char *json_path, const char *out_root_c) {
    clear_error();
    if (!json_path || !out_root_c) return fail("usage: main_debug <model.json> <output_dir>");
    const std::string out_root_s = out_root_c;
    if (out_root_s.empty() || out_root_s == "/" || out_root_s == "." || out_root_s == "..")
        return fail("Invalid output dir: " + out_root_s);
    const fs::path out_root = out_root_s;

    msfdm_model *mraw = nullptr;
    if (msfdm_model_load_json(json_path, &mraw) != MSFDM_OK) return 1;
    std::unique_ptr<msfdm_model, ModelDel> model(mraw);
    msfdm_desc d{};
    msfdm_model_get_desc(model.get(), &d);

    msfdm_solver *sraw = nullptr;
    if (msfdm_solver_create(model.get(), 0, &sraw) != MSFDM_OK) return 1;
    std::unique_ptr<msfdm_solver, SolverDel> solver(sraw);

    std::error_code ec;
    fs::remove_all(out_root, ec);
    if (ec) return fail("cannot clear " + out_root.string() + ": " + ec.message());
    std::string err;
    if (!make_dirs(out_root, err)) return fail(err);
    std::printf("[output] cleared %s\n", out_root_s.c_str());

    const int nx = d.nx;
    const bool fluid = d.calc_type == 1;
    const size_t nf = fluid ? (size_t)d.nz_fluid * nx : 0, ns = (size_t)d.nz_solid * nx;
    SnapshotWriter writer;
    for (int isrc = 0; isrc < d.nsrc; ++isrc) {
        const fs::path shot = out_root / std::to_string(isrc);
        if (!make_dirs(shot / "snapshot", err) || !make_dirs(shot / "record", err)) return fail(err);
        std::ofstream rvx, rp;
        std::ofstream rvz(shot / "record" / "solid_vz.bin", std::ios::binary | std::ios::trunc);
        if (fluid) {
            rvx.open(shot / "record" / "solid_vx.bin", std::ios::binary | std::ios::trunc);
            rp.open(shot / "record" / "fluid_p.bin", std::ios::binary | std::ios::trunc);
            if (!rp) return fail("cannot open fluid record files");
        }
        if ((fluid && !rvx) || !rvz) return fail("cannot open solid record files");
        if (msfdm_reset_shot(solver.get(), isrc) != MSFDM_OK) return 1;

        int it = 0;
        while (it < d.nt) {
            // next snapshot step at or after it
            const int snap = ((it + d.snapshot - 1) / d.snapshot) * d.snapshot;
            const int stop = snap < d.nt ? snap + 1 : d.nt;
            if (msfdm_run(solver.get(), stop - it) != MSFDM_OK) return 1;
            it = stop;
            if (snap < d.nt) {
                std::printf("finished %0.2f%%\r", 100.0 * snap / d.nt);
                std::fflush(stdout);
                const fs::path sd = shot / "snapshot";
                std::vector<SnapshotWriter::File> job;
                job.push_back({sd / "solid" / "vz", "vz", snap, std::vector<float>(ns)});
                job.push_back({sd / "solid" / "vx", "vx", snap, std::vector<float>(ns)});
                if (msfdm_get_field(solver.get(), MSFDM_SVZ, job[0].data.data()) ||
                    msfdm_get_field(solver.get(), MSFDM_SVX, job[1].data.data()))
                    return 1;
                if (fluid) {
                    job.push_back({sd / "fluid" / "vz", "vz", snap, std::vector<float>(nf)});
                    job.push_back({sd / "fluid" / "vx", "vx", snap, std::vector<float>(nf)});
                    if (msfdm_get_field(solver.get(), MSFDM_FVZ, job[2].data.data()) ||
                        msfdm_get_field(solver.get(), MSFDM_FVX, job[3].data.data()))
                        return 1;
                }
                if (!writer.push(std::move(job), err)) return fail(err);
            }
        }
        if (msfdm_synchronize(solver.get()) != MSFDM_OK) return 1;
        // records: time-major [nrec][nx] (output_record, src/main.cu:245-295)
        int nrec = 0;
        msfdm_num_records(solver.get(), &nrec);
        const int chunk = 256;
        std::vector<float> bx((size_t)chunk * nx), bz((size_t)chunk * nx), bp((size_t)chunk * nx);
        for (int r0 = 0; r0 < nrec; r0 += chunk) {
            const int n = std::min(chunk, nrec - r0);
            if (msfdm_get_records(solver.get(), r0, n, fluid ? bx.data() : nullptr, bz.data(), fluid ? bp.data() : nullptr) !=
                MSFDM_OK)
                return 1;
            const std::streamsize bytes = (std::streamsize)((size_t)n * nx * sizeof(float));
            rvz.write(reinterpret_cast<const char *>(bz.data()), bytes);
            if (fluid) {
                rvx.write(reinterpret_cast<const char *>(bx.data()), bytes);
                rp.write(reinterpret_cast<const char *>(bp.data()), bytes);
                if (!rp) return fail("failed to write fluid record data");
            }
            if ((fluid && !rvx) || !rvz) return fail("failed to write solid record data");
        }
    }
    if (!writer.finish(&err)) return fail(err);
    std::printf("\n");
    return 0;
}
