// rawdata.cpp -- see rawdata.h.
#include "rawdata.h"

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
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

// save_ascii (src/output_styles.cpp:182-327): one `x<TAB>y<TAB>value` line per
// point, j = 0..Y outer, i = (cylindrical ? -X : 0)..X inner with the value at
// |i|, a blank line after each row; default ostream formatting.
void RawWriter::write_ascii(const Snapshot &s) {
    const int X = geo_.X, Y = geo_.Y;
    const int i_lo = geo_.cylindrical ? -X : 0;
    auto field_file = [&](const char *prefix, const Field &f, double *mn, double *mx) {
        std::ofstream out(step_name(dir_, prefix, s.step, ".txt"));
        if (!out) {
            std::lock_guard<std::mutex> lk(m_);
            err_ = "cannot write " + step_name(dir_, prefix, s.step, ".txt");
            return;
        }
        for (int j = 0; j <= Y; j++) {
            for (int i = i_lo; i <= X; i++) {
                const double v = f.at(std::abs(i), j);
                out << (double)i * geo_.dx << "\t" << (double)j * geo_.dy << "\t" << v << "\n";
                if (mn) {
                    *mx = std::max(*mx, v);
                    *mn = std::min(*mn, v);
                }
            }
            out << "\n";
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
