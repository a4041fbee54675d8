// bubbles.cpp -- the `bubbles <conf>` driver over libbubbles_rp.
//
// Same command line, configuration keys and rawdata/ output as the
// reference's executable (src/bubbles.cpp:18-310 with solve_bubbles'
// loop and save logic, src/bubble_CUDA.cu:56-476); the coupled step itself
// runs on the GPU through the C ABI of include/bubbles_rp.h.
//
// Differences that are deliberate (each one is a reference defect):
//   * the output directory is created if missing and only this program's own
//     files in it are removed (the reference runs `rm -rf <dir>` through a
//     shell, output_styles.cpp:19-31);
//   * snapshots are written by one writer thread from copies, so the min/max
//     files are not raced (output_styles.cpp:69-100, 199-247).
// Focal-point tracking and the PID controller (bubble_CUDA.cu:310-352) run
// here on the host, as in the reference: the library reports each step's raw
// focal point (the cell of the largest temperature increment), the driver
// filters it, runs the controller and moves the source focus for the next
// step with brp_set_focus.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/bubbles_rp.h"
#include "options.h"
#include "rawdata.h"

using namespace brp_driver;

namespace {

// 4th-order Butterworth low-pass filter (filter_loop,
// bubble_CUDA_kernel.cuh:118-135).  Its state is process-wide, as the
// reference's file-scope statics, so it carries over between passes.
struct Butterworth {
    double xv[5] = {0, 0, 0, 0, 0}, yv[5] = {0, 0, 0, 0, 0};
    double operator()(double input) {
        const double gain = 6.522112693e+08;
        for (int k = 0; k < 4; k++) {
            xv[k] = xv[k + 1];
            yv[k] = yv[k + 1];
        }
        xv[4] = input / gain;
        yv[4] = (xv[0] + xv[4]) + 4 * (xv[1] + xv[3]) + 6 * xv[2] + (-0.9676955438 * yv[0]) +
                (3.9025587848 * yv[1]) + (-5.9020258615 * yv[2]) + (3.9671625959 * yv[3]);
        return yv[4];
    }
};
Butterworth g_filter;

struct Vec2 {
    double x = 0.0, y = 0.0;
};

// The controller state handed from one pass to the next (solve_bubbles'
// returned tuple, bubble_CUDA.cu:470-475).
struct PidCarry {
    bool valid = false;
    Vec2 control, err_total, d_err;
};

// focal_PID (bubble_CUDA_kernel.cuh:1168-1184), componentwise.
Vec2 focal_pid(Vec2 target, Vec2 actual, Vec2 prev, Vec2 *d_err, Vec2 *err_total) {
    const double K = 0.2 * 0.60;
    const double Pc = 1000.0;
    const double D = K * 0.125 * Pc;
    const double I = K / (0.5 * Pc);
    const Vec2 err{target.x - actual.x, target.y - actual.y};
    err_total->x = err_total->x + err.x;
    err_total->y = err_total->y + err.y;
    d_err->x = prev.x - actual.x;
    d_err->y = prev.y - actual.y;
    return Vec2{target.x + (K * err.x + D * d_err->x + I * err_total->x),
                target.y + (K * err.y + D * d_err->y + I * err_total->y)};
}

bool get(brp_handle *h, int id, Field *f) {
    int nx, ny, i0, j0;
    if (brp_field_shape(h, id, &nx, &ny, &i0, &j0) != BRP_OK) return false;
    f->nx = nx;
    f->ny = ny;
    f->i0 = i0;
    f->j0 = j0;
    f->v.resize((size_t)nx * ny);
    return brp_get_field(h, id, f->v.data(), nx) == BRP_OK;
}

int fail(brp_handle *h, const char *what) {
    std::fprintf(stderr, "\nbubbles: %s: %s\n", what, brp_last_error(h));
    return 1;
}

// One pass of solve_bubbles (bubble_CUDA.cu:56-476).  focalpoint / control
// receive the pass's filtered focal points and controller outputs.
int run_pass(const Options &o, const std::string &dir, RawWriter **writer_out, int *steps_done,
             PidCarry *pid, std::vector<Vec2> *focalpoint, std::vector<Vec2> *control) {
    brp_handle *h = nullptr;
    std::printf("Computing Simulation Variables...");
    if (brp_create(&o.cfg, &h) != BRP_OK) return fail(nullptr, "create");
    std::printf("\tdone\n");
    if (o.cfg.tr_pid && !o.cfg.thermal)
        std::printf("Note: focal-point tracking needs the temperature solve "
                    "([extensions] thermal = 1); the focus stays fixed.\n");
    std::printf("Preparing folders...");
    std::string err;
    if (!prepare_directory(dir, &err)) {
        std::fprintf(stderr, "\nbubbles: %s\n", err.c_str());
        brp_destroy(h);
        return 1;
    }
    std::printf("\t\t\tdone\n");
    std::printf("Running Initialization Kernels...");
    if (brp_prepare(h) != BRP_OK) {
        int rc = fail(h, "prepare");
        brp_destroy(h);
        return rc;
    }
    std::printf("\tdone\n");
    brp_info info;
    brp_get_info(h, &info);
    Geometry geo{o.cfg.grid_x, o.cfg.grid_y, info.dx, info.dy, o.cfg.tr_cylindrical != 0};
    // the transducer's focal point fp (init_transducer) and the initial
    // control point (bubble_CUDA.cu:125-153)
    const Vec2 fp{info.fp_x, info.fp_y};
    focalpoint->assign(1, Vec2{0.0, 0.0});
    control->clear();
    Vec2 err_total, d_err;
    if (o.cfg.tr_pid && pid->valid) {
        control->push_back(pid->control);
        err_total = pid->err_total;
        d_err = pid->d_err;
    } else if (o.cfg.tr_pid && o.cfg.tr_init_control) {
        control->push_back(Vec2{fp.x, o.cfg.tr_init_control});
    } else {
        control->push_back(fp);
    }
    const bool tracking = o.cfg.thermal != 0;
    if (o.cfg.tr_pid && pid->valid)  // later passes start from the previous pass's control
        brp_set_focus(h, control->back().x, control->back().y,
                      control->back().y / 0.5 / std::sqrt(3.0));
    const int save_function = (o.ascii ? 1 : 0) | (o.sph ? 2 : 0);
    RawWriter *writer = *writer_out;
    if (!writer) writer = new RawWriter(dir, geo, (save_function & 2) != 0);
    *writer_out = writer;

    std::printf("Running the simulation\t\t");
    const brp_config &c = o.cfg;
    int nstep = 0;
    double tstep = 0.0;
    while (((c.sim_n_step_max != 0) && (nstep < c.sim_n_step_max)) ||
           ((c.sim_t_step_max != 0) && (tstep < c.sim_t_step_max))) {
        brp_step_stats st;
        if (brp_step(h, 1, &st) != BRP_OK) {
            int rc = fail(h, "step");
            writer->finish();
            brp_destroy(h);
            return rc;
        }
        nstep = (int)st.nstep;
        tstep = st.time;
        if (tracking) {
            // bubble_CUDA.cu:310-348: filtered focal y, then the controller
            focalpoint->push_back(Vec2{0.0, g_filter(st.focal_y)});
            if (o.cfg.tr_pid && nstep > o.cfg.tr_pid_start_step) {
                const size_t k = focalpoint->size();
                Vec2 u = focal_pid(fp, (*focalpoint)[k - 1], (*focalpoint)[k - 2], &d_err,
                                   &err_total);
                u.x = fp.x;  // the reference sets 0, its fp.x without the focus_x extension
                u.y = std::min(std::max(u.y, fp.y), o.cfg.grid_ly);
                control->push_back(u);
                if (brp_set_focus(h, u.x, u.y, u.y / 0.5 / std::sqrt(3.0)) != BRP_OK) {
                    int rc = fail(h, "set_focus");
                    writer->finish();
                    brp_destroy(h);
                    return rc;
                }
            }
        }
        if (c.sim_n_step_max != 0)
            std::printf("\rRunning the simulation...\t\tnstep : %5i / %i\t", nstep, c.sim_n_step_max);
        else
            std::printf("\rRunning the simulation...\t\ttstep : %4.2E / %4.2E\t", tstep, c.sim_t_step_max);
        std::fflush(stdout);
        if (c.sim_save_interval > 0 && nstep % c.sim_save_interval == 0 && save_function) {
            Snapshot s;
            s.step = nstep;
            s.tstep = (float)tstep;
            bool ok = true;
            if (o.dbg_p0) ok &= s.has_p0 = get(h, BRP_FIELD_P0, &s.p0);
            if (o.dbg_fg) ok &= s.has_fg = get(h, BRP_FIELD_FG, &s.fg);
            if (o.dbg_T) ok &= s.has_T = get(h, BRP_FIELD_T, &s.T);
            if (o.dbg_vxy && !(save_function & 2)) {
                ok &= get(h, BRP_FIELD_VX, &s.vx) && get(h, BRP_FIELD_VY, &s.vy);
                s.has_v = true;
            }
            if (o.dbg_bubbles && !(save_function & 2) && c.bub_enable && info.n_bubbles > 0) {
                const size_t n = (size_t)info.n_bubbles;
                s.bx.resize(n);
                s.by.resize(n);
                s.R_t.resize(n);
                s.PG.resize(n);
                ok &= brp_get_bubbles(h, BRP_BUB_X, s.bx.data()) == BRP_OK &&
                      brp_get_bubbles(h, BRP_BUB_Y, s.by.data()) == BRP_OK &&
                      brp_get_bubbles(h, BRP_BUB_R_T, s.R_t.data()) == BRP_OK &&
                      brp_get_bubbles(h, BRP_BUB_PG, s.PG.data()) == BRP_OK;
                s.has_bubbles = true;
            }
            if (!ok) {
                int rc = fail(h, "save");
                writer->finish();
                brp_destroy(h);
                return rc;
            }
            writer->submit(std::move(s));
        }
    }
    std::printf("\rRunning the simulation...\t\tdone\n\n");
    *steps_done = nstep;
    pid->valid = true;
    pid->control = control->back();
    pid->err_total = err_total;
    pid->d_err = d_err;
    brp_destroy(h);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    const time_t start_time = std::time(nullptr);
    Options o;
    std::string err;
    if (parse_command_line(argc, argv, &o, &err)) {
        std::cerr << "bubbles: " << err << std::endl;
        return 1;
    }
    if (o.help_all) {
        std::cout << help_text(true) << std::endl;
        return 0;
    }
    if (o.help) {
        std::cout << help_text(false) << std::endl;
        return 0;
    }
    if (o.input_file.empty()) {
        std::cout << "Please specify an input configuration file" << std::endl;
        return 1;
    }
    if (parse_config_file(o.input_file, &o, &err)) {
        std::cerr << "bubbles: " << o.input_file << ": " << err << std::endl;
        return 1;
    }
    if (o.dry_run) {
        std::cout << options_json(o) << std::endl;
        return 0;
    }
    const int save_function = (o.ascii ? 1 : 0) | (o.sph ? 2 : 0);
    std::cout << "Input file is " << o.input_file << std::endl;
    std::cout << "Saving results to " << o.target_dir << std::endl;
    std::cout << "Save format is " << (save_function == 2 ? "sph" : "ASCII plaintext");
    std::cout << std::endl << std::endl;

    double fg_min = 1e10, fg_max = -1e10, p0_min = 1e10, p0_max = -1e10, T_min = 1e10, T_max = -1e10;
    std::string target_dir = o.target_dir;
    PidCarry pid;
    for (int j = 0; j < o.num_iters; j++) {
        // multi-pass runs get directory_I/, directory_II/, ... (bubbles.cpp:224-246)
        if (o.num_iters > 1) {
            target_dir = o.directory;
            if (!target_dir.empty() && target_dir.back() == '/') target_dir.pop_back();
            target_dir += "_" + std::string(j + 1, 'I') + "/";
        }
        RawWriter *writer = nullptr;
        int steps = 0;
        std::vector<Vec2> focalpoint, control;
        int rc = run_pass(o, target_dir, &writer, &steps, &pid, &focalpoint, &control);
        if (writer) {
            writer->finish();
            fg_min = std::min(fg_min, writer->fg_min);
            fg_max = std::max(fg_max, writer->fg_max);
            p0_min = std::min(p0_min, writer->p0_min);
            p0_max = std::max(p0_max, writer->p0_max);
            T_min = std::min(T_min, writer->T_min);
            T_max = std::max(T_max, writer->T_max);
            std::string werr = writer->error();
            delete writer;
            if (!werr.empty()) {
                std::cerr << "bubbles: " << werr << std::endl;
                return 1;
            }
        }
        if (rc) return rc;
        // focal point files (bubbles.cpp:258-270)
        std::ofstream fa(target_dir + "focalpoint_actual_T.txt");
        for (size_t i = 0; i < focalpoint.size(); i++) fa << i << "\t" << focalpoint[i].y << std::endl;
        std::ofstream fc(target_dir + "focalpoint_controller_T.txt");
        for (size_t i = 0; i < control.size(); i++)
            fc << i + o.cfg.tr_pid_start_step << "\t" << control[i].y << std::endl;
        if (o.dbg_T) {
            std::ofstream f(target_dir + "T_minmax.txt");
            f << (o.has_T_min ? o.T_min : T_min) << std::endl << (o.has_T_max ? o.T_max : T_max) << std::endl;
        }
        if (o.dbg_p0) {
            std::ofstream f(target_dir + "p0_minmax.txt");
            f << p0_min << std::endl << (o.has_p_max ? o.p_max : p0_max) << std::endl;
        }
        if (o.dbg_fg) {
            std::ofstream f(target_dir + "fg_minmax.txt");
            f << (o.has_fg_min ? o.fg_min : fg_min) << std::endl << (o.has_fg_max ? o.fg_max : fg_max) << std::endl;
        }
        const long secs = (long)(std::time(nullptr) - start_time);
        if (o.num_iters == 1)
            std::cout << "The program took ";
        else
            std::cout << j << " pass took ";
        std::cout << secs << " seconds to run" << std::endl;
        std::ofstream rt(target_dir + "runtime.txt");
        rt << "Finished in " << secs << " seconds" << std::endl;
        (void)steps;
    }
    return 0;
}
