// rawdata.h -- the `rawdata/` output of the `bubbles` driver.
//
// File names and tuple formats of the reference's save_ascii / save_sph
// (src/output_styles.cpp:36-341, src/SphDataType.cpp:142-192).  Snapshots are
// copied out of the simulation and written by one background thread, so the
// run continues while a save is written and the min/max bookkeeping has a
// single owner.
#pragma once

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace brp_driver {

struct Field {
    int nx = 0, ny = 0, i0 = 0, j0 = 0;  // logical extents of the array
    std::vector<double> v;               // row-major, (ny, nx)
    double at(int i, int j) const { return v[(size_t)(j - j0) * nx + (i - i0)]; }
};

struct Snapshot {
    int step = 0;
    float tstep = 0.0f;
    bool has_fg = false, has_p0 = false, has_T = false, has_v = false, has_bubbles = false;
    Field fg, p0, T, vx, vy;
    std::vector<double> bx, by, R_t, PG;
};

struct Geometry {
    int X = 0, Y = 0;
    double dx = 0, dy = 0;
    bool cylindrical = false;
};

class RawWriter {
   public:
    RawWriter(std::string dir, Geometry geo, bool sph);
    ~RawWriter();
    void submit(Snapshot &&s);  // blocks while two snapshots are pending
    void finish();              // waits for every pending write
    std::string error() const;
    // running extrema of the written fields (initial 1e10 / -1e10 as the reference)
    double fg_min = 1e10, fg_max = -1e10, p0_min = 1e10, p0_max = -1e10, T_min = 1e10,
           T_max = -1e10;

   private:
    void run();
    void write_ascii(const Snapshot &s);
    void write_sph(const Snapshot &s);
    std::string dir_;
    Geometry geo_;
    bool sph_;
    std::deque<Snapshot> q_;
    mutable std::mutex m_;
    std::condition_variable cv_;
    bool stop_ = false;
    std::string err_;
    std::thread th_;
};

std::string step_name(const std::string &dir, const char *prefix, int step, const char *ext);
// Writes one field as save_ascii's text (x, y, value lines, rows j = 0..Y,
// columns (cylindrical ? -X : 0)..X with the value at |i|, a blank line per
// row), formatting row blocks on `threads` threads while the previous blocks
// are written; the bytes equal the serial ostream output.  mn / mx (may be
// null) take the running extrema.
bool write_ascii_field(const std::string &path, const Field &f, int X, int Y, bool cylindrical,
                       double dx, double dy, int threads, double *mn, double *mx);
// The same text formatted on `threads` threads and discarded (for timing the
// formatting apart from the file system); returns its size in bytes.
long long format_ascii_field(const Field &f, int X, int Y, bool cylindrical, double dx, double dy,
                             int threads);
// Writes one float32 SPH scalar record file (SphData::SaveSph layout).
bool save_sph_file(const std::string &path, int nx, int ny, float dx, float dy, float time,
                   int step, const std::vector<float> &data);
// Creates the directory (and parents) and deletes the driver's own output
// files in it; nothing else is removed.
bool prepare_directory(const std::string &dir, std::string *err);

}  // namespace brp_driver
