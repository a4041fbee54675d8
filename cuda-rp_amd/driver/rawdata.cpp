// rawdata.cpp -- see rawdata.h.
#include "rawdata.h"

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace brp_driver {

std::string step_name(const std::string &dir, const char *prefix, int step, const char *ext) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "%05d", step);
    return dir + prefix + buf + ext;
}

bool save_sph_file(const std::string &path, int nx, int ny, float dx, float dy, float time,
                   int step, const std::vector<float> &data) {
    std::ofstream os(path, std::ios::binary);
    if (!os) return false;
    auto w32 = [&](const void *p) { os.write((const char *)p, 4); };
    int len = 8, sv = 1, dt = 1;  // scalar, float
    w32(&len); w32(&sv); w32(&dt); w32(&len);
    int dims[3] = {nx, ny, 1};
    len = 12;
    w32(&len); os.write((const char *)dims, 12); w32(&len);
    float orig[3] = {0.0f, 0.0f, 0.0f};
    w32(&len); os.write((const char *)orig, 12); w32(&len);
    float pitch[3] = {dx, dy, 1.0f};
    w32(&len); os.write((const char *)pitch, 12); w32(&len);
    len = 8;
    w32(&len); w32(&step); w32(&time); w32(&len);
    len = 4 * nx * ny;
    w32(&len); os.write((const char *)data.data(), (std::streamsize)len); w32(&len);
    return (bool)os;
}

bool prepare_directory(const std::string &dir, std::string *err) {
    std::string path;
    for (size_t k = 0; k <= dir.size(); k++) {
        if (k == dir.size() || dir[k] == '/') {
            if (!path.empty() && mkdir(path.c_str(), 0755) != 0 && errno != EEXIST) {
                *err = "cannot create directory '" + path + "': " + std::strerror(errno);
                return false;
            }
        }
        if (k < dir.size()) path += dir[k];
    }
    static const char *prefixes[] = {"fg_step_", "p0_step_", "T_step_", "vx_step_", "vy_step_",
                                     "bubble_radii_", "bubble_PG_"};
    static const char *fixed[] = {"focalpoint_actual_T.txt", "focalpoint_controller_T.txt",
                                  "T_minmax.txt", "p0_minmax.txt", "fg_minmax.txt", "runtime.txt"};
    DIR *d = opendir(dir.c_str());
    if (!d) {
        *err = "cannot open directory '" + dir + "'";
        return false;
    }
    std::vector<std::string> victims;
    while (dirent *e = readdir(d)) {
        std::string n = e->d_name;
        bool ours = false;
        for (const char *p : prefixes)
            if (n.compare(0, std::strlen(p), p) == 0 &&
                (n.size() > 4 && (n.compare(n.size() - 4, 4, ".txt") == 0 ||
                                  n.compare(n.size() - 4, 4, ".sph") == 0)))
                ours = true;
        for (const char *f : fixed)
            if (n == f) ours = true;
        if (ours) victims.push_back(dir + n);
    }
    closedir(d);
    for (auto &v : victims) std::remove(v.c_str());
    return true;
}

RawWriter::RawWriter(std::string dir, Geometry geo, bool sph)
    : dir_(std::move(dir)), geo_(geo), sph_(sph) {
    th_ = std::thread([this] { run(); });
}

RawWriter::~RawWriter() { finish(); }

void RawWriter::submit(Snapshot &&s) {
    std::unique_lock<std::mutex> lk(m_);
    cv_.wait(lk, [this] { return q_.size() < 2; });
    q_.push_back(std::move(s));
    cv_.notify_all();
}

void RawWriter::finish() {
    {
        std::unique_lock<std::mutex> lk(m_);
        if (stop_ && !th_.joinable()) return;
        stop_ = true;
        cv_.notify_all();
    }
    if (th_.joinable()) th_.join();
}

std::string RawWriter::error() const {
    std::lock_guard<std::mutex> lk(m_);
    return err_;
}

void RawWriter::run() {
    for (;;) {
        Snapshot s;
        {
            std::unique_lock<std::mutex> lk(m_);
            cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;
            s = std::move(q_.front());
        }
        if (sph_)
            write_sph(s);
        else
            write_ascii(s);
        std::lock_guard<std::mutex> lk(m_);
        q_.pop_front();
        cv_.notify_all();
    }
}

// One `x<TAB>y<TAB>value` text line per point, formatted as the default
// ostream insertion of a double: the %g conversion with six significant
// digits, which std::to_chars(general, 6) produces without a locale or a
// format string.
static inline void put_g(std::string &out, double v) {
    char b[32];
    const std::to_chars_result r = std::to_chars(b, b + sizeof b, v, std::chars_format::general, 6);
    out.append(b, (size_t)(r.ptr - b));
}

// xs: the x column's text, shared by every block (built once per field: a
// per-block array above the allocator's mmap threshold is mapped and
// unmapped by every call, and unmapping stalls every thread of the process).
// The extrema are kept in locals and stored once: threads updating
// neighbouring slots per value share a cache line and stop scaling.
static void format_ascii_rows(const Field &f, const std::vector<std::string> &xs, int i_lo, int X,
                              double dy, int j_begin, int j_end, std::string &out, double *mn,
                              double *mx) {
    double lo = *mn, hi = *mx;
    for (int j = j_begin; j < j_end; j++) {
        std::string y;
        put_g(y, (double)j * dy);
        for (int i = i_lo; i <= X; i++) {
            const double v = f.at(std::abs(i), j);
            out += xs[(size_t)(i - i_lo)];
            out += '\t';
            out += y;
            out += '\t';
            put_g(out, v);
            out += '\n';
            hi = std::max(hi, v);
            lo = std::min(lo, v);
        }
        out += '\n';
    }
    *mn = lo;
    *mx = hi;
}

// Rows are formatted in blocks of kBlockRows by `threads` worker threads that
// take blocks in order from a shared counter into a ring of 4 x threads
// buffers; the calling thread writes the blocks in order as they complete,
// so the text is produced and written at the same time and a worker waits
// only when the ring is full.  `fp` null: format only (the text is
// discarded; for timing).  Returns the bytes produced, or -1 on a write
// error.
static long long ascii_pipeline(FILE *fp, const Field &f, int X, int Y, bool cylindrical, double dx,
                                double dy, int threads, double *mn, double *mx) {
    const int i_lo = cylindrical ? -X : 0;
    const int rows = Y + 1;
    constexpr int kBlockRows = 16;
    const int nblk = (rows + kBlockRows - 1) / kBlockRows;
    threads = std::max(1, std::min(threads, nblk));
    const int ring = 4 * threads;
    std::vector<std::string> buf((size_t)ring);
    std::vector<int> ready((size_t)ring, -1);  // block held by each slot, -1 while being formatted
    std::vector<std::string> xs((size_t)(X - i_lo + 1));
    for (int i = i_lo; i <= X; i++) put_g(xs[(size_t)(i - i_lo)], (double)i * dx);
    std::mutex m;
    std::condition_variable cv_ready, cv_free;
    int next = 0, written = 0;  // next block to format, blocks written (under m)
    std::vector<double> lo((size_t)threads, 1e300), hi((size_t)threads, -1e300);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&, t] {
            double l = 1e300, h = -1e300;
            for (;;) {
                int b;
                {
                    std::unique_lock<std::mutex> lk(m);
                    if (next >= nblk) break;
                    b = next++;
                    cv_free.wait(lk, [&] { return b - written < ring; });  // slot b % ring is free
                }
                std::string &out = buf[(size_t)(b % ring)];
                out.clear();
                const int j0 = b * kBlockRows, j1 = std::min(rows, j0 + kBlockRows);
                format_ascii_rows(f, xs, i_lo, X, dy, j0, j1, out, &l, &h);
                {
                    std::lock_guard<std::mutex> lk(m);
                    ready[(size_t)(b % ring)] = b;
                }
                cv_ready.notify_all();
            }
            lo[(size_t)t] = l;
            hi[(size_t)t] = h;
        });
    long long bytes = 0;
    bool ok = true;
    for (int b = 0; b < nblk; b++) {
        {
            std::unique_lock<std::mutex> lk(m);
            cv_ready.wait(lk, [&] { return ready[(size_t)(b % ring)] == b; });
        }
        const std::string &out = buf[(size_t)(b % ring)];
        bytes += (long long)out.size();
        if (fp && ok) ok = std::fwrite(out.data(), 1, out.size(), fp) == out.size();
        {
            std::lock_guard<std::mutex> lk(m);
            ready[(size_t)(b % ring)] = -1;
            written = b + 1;
        }
        cv_free.notify_all();
    }
    for (auto &th : pool) th.join();
    if (mn)
        for (int t = 0; t < threads; t++) {
            *mn = std::min(*mn, lo[(size_t)t]);
            *mx = std::max(*mx, hi[(size_t)t]);
        }
    return ok ? bytes : -1;
}

bool write_ascii_field(const std::string &path, const Field &f, int X, int Y, bool cylindrical,
                       double dx, double dy, int threads, double *mn, double *mx) {
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = ascii_pipeline(fp, f, X, Y, cylindrical, dx, dy, threads, mn, mx) >= 0;
    return (std::fclose(fp) == 0) && ok;
}

long long format_ascii_field(const Field &f, int X, int Y, bool cylindrical, double dx, double dy,
                             int threads) {
    return ascii_pipeline(nullptr, f, X, Y, cylindrical, dx, dy, threads, nullptr, nullptr);
}

static int writer_threads() {
    const char *e = std::getenv("OMP_NUM_THREADS");
    const int n = e ? std::atoi(e) : 0;
    return n > 0 ? n : std::max(1, (int)std::thread::hardware_concurrency());
}

// save_ascii (src/output_styles.cpp:182-327): one `x<TAB>y<TAB>value` line per
// point, j = 0..Y outer, i = (cylindrical ? -X : 0)..X inner with the value at
// |i|, a blank line after each row; default ostream formatting.  At 4096^2 one
// field is ~0.5 GB of text, so rows are formatted on several threads.
void RawWriter::write_ascii(const Snapshot &s) {
    const int X = geo_.X, Y = geo_.Y;
    const int nt = writer_threads();
    auto field_file = [&](const char *prefix, const Field &f, double *mn, double *mx) {
        const std::string path = step_name(dir_, prefix, s.step, ".txt");
        if (!write_ascii_field(path, f, X, Y, geo_.cylindrical, geo_.dx, geo_.dy, nt, mn, mx)) {
            std::lock_guard<std::mutex> lk(m_);
            err_ = "cannot write " + path;
        }
    };
    if (s.has_fg) field_file("fg_step_", s.fg, &fg_min, &fg_max);
    if (s.has_p0) field_file("p0_step_", s.p0, &p0_min, &p0_max);
    if (s.has_T) field_file("T_step_", s.T, &T_min, &T_max);
    if (s.has_v) {
        field_file("vx_step_", s.vx, nullptr, nullptr);
        field_file("vy_step_", s.vy, nullptr, nullptr);
    }
    if (s.has_bubbles) {
        std::ofstream r(step_name(dir_, "bubble_radii_", s.step, ".txt"));
        for (size_t k = 0; k < s.bx.size(); k++) r << s.bx[k] << "\t" << s.by[k] << "\t" << s.R_t[k] << "\n";
        std::ofstream g(step_name(dir_, "bubble_PG_", s.step, ".txt"));
        for (size_t k = 0; k < s.bx.size(); k++) g << s.bx[k] << "\t" << s.by[k] << "\t" << s.PG[k] << "\n";
    }
}

// save_sph (src/output_styles.cpp:36-137): fg, p0 and T on (X+1) x (Y+1)
// points as float32 SPH records.
void RawWriter::write_sph(const Snapshot &s) {
    const int X = geo_.X, Y = geo_.Y;
    auto one = [&](const char *prefix, const Field &f, double *mn, double *mx) {
        std::vector<float> d((size_t)(X + 1) * (Y + 1));
        for (int j = 0; j <= Y; j++)
            for (int i = 0; i <= X; i++) {
                const double v = f.at(i, j);
                d[(size_t)j * (X + 1) + i] = (float)v;
                *mx = std::max(*mx, v);
                *mn = std::min(*mn, v);
            }
        if (!save_sph_file(step_name(dir_, prefix, s.step, ".sph"), X + 1, Y + 1, (float)geo_.dx,
                           (float)geo_.dy, s.tstep, s.step, d)) {
            std::lock_guard<std::mutex> lk(m_);
            err_ = "cannot write " + step_name(dir_, prefix, s.step, ".sph");
        }
    };
    if (s.has_fg) one("fg_step_", s.fg, &fg_min, &fg_max);
    if (s.has_p0) one("p0_step_", s.p0, &p0_min, &p0_max);
    if (s.has_T) one("T_step_", s.T, &T_min, &T_max);
}

}  // namespace brp_driver
