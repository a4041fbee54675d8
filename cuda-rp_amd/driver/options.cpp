// options.cpp -- see options.h.
#include "options.h"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <set>
#include <sstream>

namespace brp_driver {
namespace {

enum Kind { INT, DOUBLE, BOOL, STRING, SWITCH };

struct Opt {
    const char *name;
    Kind kind;
    std::function<void(Options &, const std::string &)> set;
    const char *help;
};

std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

bool to_int(const std::string &s, int *v) {
    if (s.empty()) return false;
    char *end = nullptr;
    errno = 0;
    long x = std::strtol(s.c_str(), &end, 10);
    if (*end || errno || x < -2147483647L - 1 || x > 2147483647L) return false;
    *v = (int)x;
    return true;
}

bool to_double(const std::string &s, double *v) {
    if (s.empty()) return false;
    char *end = nullptr;
    errno = 0;
    double x = std::strtod(s.c_str(), &end);
    if (*end || errno == ERANGE) return false;
    *v = x;
    return true;
}

bool to_bool(std::string s, bool *v) {
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    if (s.empty() || s == "on" || s == "yes" || s == "1" || s == "true") {
        *v = true;
        return true;
    }
    if (s == "off" || s == "no" || s == "0" || s == "false") {
        *v = false;
        return true;
    }
    return false;
}

#define I32(field) [](Options &o, const std::string &v) { int x = 0; to_int(v, &x); o.cfg.field = x; }
#define F64(field) [](Options &o, const std::string &v) { double x = 0; to_double(v, &x); o.cfg.field = x; }
#define B32(field) [](Options &o, const std::string &v) { bool x = false; to_bool(v, &x); o.cfg.field = x ? 1 : 0; }

// The registered options: file_opts, generic_opts and hidden_opts of
// src/bubbles.cpp:66-136 (note "debug p_min" is registered with a space, so
// no configuration line can ever set it), then the keys of the
// library's extensions E1-E4 and run controls under [extensions].
const std::vector<Opt> &file_options() {
    static const std::vector<Opt> v = {
        {"grid.x", INT, I32(grid_x), ""},
        {"grid.y", INT, I32(grid_y), ""},
        {"grid.lx", DOUBLE, F64(grid_lx), ""},
        {"grid.ly", DOUBLE, F64(grid_ly), ""},
        {"pml.x0", BOOL, B32(pml_x0), ""},
        {"pml.x1", BOOL, B32(pml_x1), ""},
        {"pml.y0", BOOL, B32(pml_y0), ""},
        {"pml.y1", BOOL, B32(pml_y1), ""},
        {"pml.grid_number", INT, I32(pml_grid_number), ""},
        {"pml.sigma", DOUBLE, F64(pml_sigma), ""},
        {"pml.order", INT, I32(pml_order), ""},
        {"simulation.cfl", DOUBLE, F64(sim_cfl), ""},
        {"simulation.dt0", DOUBLE, F64(sim_dt0), ""},
        {"simulation.n_step_max", INT, I32(sim_n_step_max), ""},
        {"simulation.t_step_max", DOUBLE, F64(sim_t_step_max), ""},
        {"simulation.save_interval", INT, I32(sim_save_interval), ""},
        {"simulation.order", INT, I32(sim_order), ""},
        {"simulation.smooth_delta_band", INT, I32(sim_smooth_delta_band), ""},
        {"transducer.amp", DOUBLE, F64(tr_amp), ""},
        {"transducer.freq", DOUBLE, F64(tr_freq), ""},
        {"transducer.focal_dist", DOUBLE, F64(tr_focal_dist), ""},
        {"transducer.pid", BOOL, B32(tr_pid), " (=0)"},
        {"transducer.pid_start_step", INT, I32(tr_pid_start_step), " (=1)"},
        {"transducer.init_control", DOUBLE, F64(tr_init_control), " (=0)"},
        {"transducer.pid_iterate", INT,
         [](Options &o, const std::string &v) { to_int(v, &o.num_iters); o.cfg.tr_pid_iterate = o.num_iters; },
         " (=1)"},
        {"transducer.box_size", DOUBLE, F64(tr_box_size), ""},
        {"transducer.cylindrical", BOOL, B32(tr_cylindrical), ""},
        {"transducer.on_wave", INT, I32(tr_on_wave), ""},
        {"transducer.off_wave", INT, I32(tr_off_wave), ""},
        {"transducer.transducer_v", BOOL, B32(tr_plane_v), ""},
        {"transducer.transducer_p", BOOL, B32(tr_plane_p), ""},
        {"transducer.focused_wave_v", BOOL, B32(tr_focused_v), ""},
        {"transducer.focused_wave_p", BOOL, B32(tr_focused_p), ""},
        {"bubbles.enable", BOOL, B32(bub_enable), " (=0)"},
        {"bubbles.fg0", DOUBLE, F64(bub_fg0), ""},
        {"bubbles.R0", DOUBLE, F64(bub_R0), ""},
        {"bubbles.P0", DOUBLE, F64(bub_P0), ""},
        {"bubbles.T0", DOUBLE, F64(bub_T0), ""},
        {"bubbles.gamma", DOUBLE, F64(bub_gamma), ""},
        {"bubbles.liquid_surface_tension", DOUBLE, F64(bub_sigma), ""},
        {"bubbles.liquid_viscosity", DOUBLE, F64(bub_mu), ""},
        {"bubbles.liquid_density", DOUBLE, F64(bub_rho), ""},
        {"bubbles.K0", DOUBLE, F64(bub_K0), ""},
        {"debug.display_dim", INT, [](Options &o, const std::string &v) { to_int(v, &o.display); }, " (=4)"},
        {"debug.fg", BOOL, [](Options &o, const std::string &v) { to_bool(v, &o.dbg_fg); }, " (=1)"},
        {"debug.fg_min", DOUBLE, [](Options &o, const std::string &v) { o.has_fg_min = to_double(v, &o.fg_min); }, ""},
        {"debug.fg_max", DOUBLE, [](Options &o, const std::string &v) { o.has_fg_max = to_double(v, &o.fg_max); }, ""},
        {"debug.T", BOOL, [](Options &o, const std::string &v) { to_bool(v, &o.dbg_T); }, " (=1)"},
        {"debug.T_min", DOUBLE, [](Options &o, const std::string &v) { o.has_T_min = to_double(v, &o.T_min); }, ""},
        {"debug.T_max", DOUBLE, [](Options &o, const std::string &v) { o.has_T_max = to_double(v, &o.T_max); }, ""},
        {"debug.p", BOOL, [](Options &o, const std::string &v) { to_bool(v, &o.dbg_p0); }, " (=0)"},
        {"debug p_min", DOUBLE, [](Options &, const std::string &) {}, ""},
        {"debug.p_max", DOUBLE, [](Options &o, const std::string &v) { o.has_p_max = to_double(v, &o.p_max); }, ""},
        {"debug.v", BOOL, [](Options &o, const std::string &v) { to_bool(v, &o.dbg_vxy); }, " (=0)"},
        {"debug.v_min", DOUBLE, [](Options &, const std::string &) {}, ""},
        {"debug.v_max", DOUBLE, [](Options &, const std::string &) {}, ""},
        {"debug.bubbles", BOOL, [](Options &o, const std::string &v) { to_bool(v, &o.dbg_bubbles); }, " (=1)"},
        // the library's extensions (ignored by the reference as unregistered)
        {"extensions.precision", STRING,
         [](Options &o, const std::string &v) { o.cfg.precision = (v == "mixed" || v == "1") ? BRP_PREC_MIXED : BRP_PREC_FP64; },
         " fp64|mixed (=fp64)"},
        {"extensions.source_in_pml", BOOL, B32(ext_source_in_pml), " E1 (=0)"},
        {"extensions.focus_x", DOUBLE,
         [](Options &o, const std::string &v) { to_double(v, &o.cfg.ext_focus_x); o.cfg.ext_focus_x_set = 1; },
         " E2"},
        {"extensions.max_picard", INT, I32(max_picard), " (=1000)"},
        {"extensions.max_substeps", INT, I32(max_substeps), " (=1000000)"},
        {"extensions.rp_exact", BOOL, B32(rp_exact), " (=0)"},
        {"extensions.thermal", BOOL, B32(thermal), " (=1)"},
    };
    return v;
}

const Opt *find_file_option(const std::string &name) {
    for (auto &o : file_options())
        if (name == o.name) return &o;
    return nullptr;
}

bool valid_value(Kind k, const std::string &v) {
    int i;
    double d;
    bool b;
    switch (k) {
        case INT: return to_int(v, &i);
        case DOUBLE: return to_double(v, &d);
        case BOOL: return to_bool(v, &b);
        default: return true;
    }
}

int store_generic(Options &o, const std::string &name, const std::string &value, bool from_file,
                  std::set<std::string> &seen, std::string *err) {
    if (name != "ascii" && name != "sph" && name != "directory" && name != "input-file")
        return -1;
    if (seen.count(name)) {
        *err = "multiple occurrences of option '" + name + "'";
        return 1;
    }
    seen.insert(name);
    if (o.stored_names.count(name)) return 0;
    if (name == "ascii") {
        o.ascii = true;
    } else if (name == "sph") {
        o.sph = true;
    } else if (name == "directory") {
        o.directory = value;
        o.directory_defaulted = false;
    } else if (name == "input-file") {
        if (!from_file || o.input_file.empty()) o.input_file = value;
    } else {
        return -1;
    }
    o.stored_names.insert(name);
    return 0;
}

}  // namespace

int parse_command_line(int argc, char **argv, Options *o, std::string *err) {
    brp_config_default(&o->cfg);
    o->cfg.thermal = 1;  // the reference computes the temperature every step
    o->stored_names.clear();
    std::set<std::string> seen;
    struct Long {
        const char *name;
        char short_name;
        bool takes_value;
    };
    const Long longs[] = {{"help", 'h', false},       {"help-all", 0, false},
                          {"ascii", 'a', false},      {"sph", 's', false},
                          {"directory", 'd', true},   {"input-file", 0, true},
                          {"dry-run", 0, false}};
    std::vector<std::string> positional;
    bool only_positional = false;
    for (int k = 1; k < argc; k++) {
        std::string t = argv[k];
        if (only_positional || t.size() < 2 || t[0] != '-') {
            positional.push_back(t);
            continue;
        }
        if (t == "--") {
            only_positional = true;
            continue;
        }
        std::string name, value;
        bool has_value = false;
        const Long *opt = nullptr;
        if (t[1] == '-') {  // --name[=value], unique-prefix guessing
            std::string body = t.substr(2);
            size_t eq = body.find('=');
            name = body.substr(0, eq);
            if (eq != std::string::npos) {
                value = body.substr(eq + 1);
                has_value = true;
            }
            int matches = 0;
            for (auto &l : longs) {
                if (name == l.name) {
                    opt = &l;
                    matches = 1;
                    break;
                }
                if (std::strncmp(l.name, name.c_str(), name.size()) == 0) {
                    opt = &l;
                    matches++;
                }
            }
            if (matches > 1) {
                *err = "option '--" + name + "' is ambiguous";
                return 1;
            }
        } else {  // -x[value]
            for (auto &l : longs)
                if (l.short_name == t[1]) opt = &l;
            if (opt && t.size() > 2) {
                value = t.substr(2);
                has_value = true;
            }
        }
        if (!opt) {
            o->ignored.push_back(t);  // allow_unregistered
            continue;
        }
        if (opt->takes_value && !has_value) {
            if (k + 1 >= argc) {
                *err = std::string("option '") + opt->name + "' requires an argument";
                return 1;
            }
            value = argv[++k];
        } else if (!opt->takes_value && has_value && t[1] == '-') {
            *err = std::string("option '") + opt->name + "' does not take an argument";
            return 1;
        }
        std::string n = opt->name;
        if (n == "help") {
            o->help = true;
        } else if (n == "help-all") {
            o->help_all = true;
        } else if (n == "dry-run") {
            o->dry_run = true;
        } else if (store_generic(*o, n, value, false, seen, err) > 0) {
            return 1;
        }
    }
    if (positional.size() > 1) {
        *err = "multiple occurrences of option 'input-file'";
        return 1;
    }
    if (!positional.empty() && store_generic(*o, "input-file", positional[0], false, seen, err) > 0)
        return 1;
    // target_dir is fixed here, before the configuration file is read
    // (bubbles.cpp:181-188): a `directory` key in the file does not move a
    // single-pass run's output.
    o->target_dir = o->directory;
    if (o->target_dir.empty() || o->target_dir.back() != '/') o->target_dir += "/";
    return 0;
}

int parse_config_text(const std::string &text, Options *o, std::string *err) {
    std::istringstream in(text);
    std::string line, prefix;
    std::set<std::string> seen;
    int lineno = 0;
    while (std::getline(in, line)) {
        lineno++;
        size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        std::string s = trim(line);
        if (s.empty()) continue;
        if (s.front() == '[' && s.back() == ']') {
            prefix = s.substr(1, s.size() - 2);
            if (prefix.empty() || prefix.back() != '.') prefix += '.';
            continue;
        }
        size_t eq = s.find('=');
        if (eq == std::string::npos) {
            *err = "line " + std::to_string(lineno) + ": unrecognized line '" + s + "'";
            return 1;
        }
        std::string name = prefix + trim(s.substr(0, eq));
        std::string value = trim(s.substr(eq + 1));
        int g = store_generic(*o, name, value, true, seen, err);
        if (g > 0) return 1;
        if (g == 0) continue;
        const Opt *opt = find_file_option(name);
        if (!opt) {
            o->ignored.push_back(name);
            continue;
        }
        if (seen.count(name)) {
            *err = "multiple occurrences of option '" + name + "'";
            return 1;
        }
        seen.insert(name);
        if (!valid_value(opt->kind, value)) {
            *err = "invalid value '" + value + "' for option '" + name + "'";
            return 1;
        }
        if (o->stored_names.count(name)) continue;
        opt->set(*o, value);
        o->stored_names.insert(name);
    }
    return 0;
}

int parse_config_file(const std::string &path, Options *o, std::string *err) {
    std::ifstream f(path);
    if (!f) {
        *err = "cannot read configuration file '" + path + "'";
        return 1;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_config_text(ss.str(), o, err);
}

std::string help_text(bool all) {
    std::string s =
        "Options:\n\n"
        "Command line options:\n"
        "  -h [ --help ]                 Output this help message.\n"
        "  --help-all                    Output a complete list of options.\n\n"
        "Generic options:\n"
        "  -a [ --ascii ]                Save plaintext files.\n"
        "  -s [ --sph ]                  Save sph files. This option has\n"
        "                                precedence over --ascii.\n"
        "  -d [ --directory ] arg (=./rawdata/)\n"
        "                                Save into a target directory.\n";
    if (all) {
        s += "\nConfiguration file options:\n";
        for (auto &o : file_options()) {
            s += "  --";
            s += o.name;
            s += " arg";
            s += o.help;
            s += "\n";
        }
    }
    return s;
}

std::string options_json(const Options &o) {
    std::ostringstream j;
    j.precision(17);
    const brp_config &c = o.cfg;
    j << "{\"grid\": [" << c.grid_x << ", " << c.grid_y << ", " << c.grid_lx << ", " << c.grid_ly
      << "], \"pml\": [" << c.pml_x0 << ", " << c.pml_x1 << ", " << c.pml_y0 << ", " << c.pml_y1
      << ", " << c.pml_grid_number << ", " << c.pml_sigma << ", " << c.pml_order
      << "], \"simulation\": [" << c.sim_cfl << ", " << c.sim_dt0 << ", " << c.sim_n_step_max
      << ", " << c.sim_t_step_max << ", " << c.sim_save_interval << ", " << c.sim_order << ", "
      << c.sim_smooth_delta_band << "], \"transducer\": [" << c.tr_amp << ", " << c.tr_freq << ", "
      << c.tr_focal_dist << ", " << c.tr_pid << ", " << c.tr_pid_start_step << ", "
      << c.tr_init_control << ", " << o.num_iters << ", " << c.tr_box_size << ", "
      << c.tr_cylindrical << ", " << c.tr_on_wave << ", " << c.tr_off_wave << ", " << c.tr_plane_v
      << ", " << c.tr_plane_p << ", " << c.tr_focused_v << ", " << c.tr_focused_p
      << "], \"bubbles\": [" << c.bub_enable << ", " << c.bub_fg0 << ", " << c.bub_R0 << ", "
      << c.bub_P0 << ", " << c.bub_T0 << ", " << c.bub_gamma << ", " << c.bub_sigma << ", "
      << c.bub_mu << ", " << c.bub_rho << ", " << c.bub_K0 << "], \"debug\": [" << o.display
      << ", " << o.dbg_fg << ", " << o.dbg_T << ", " << o.dbg_p0 << ", " << o.dbg_vxy << ", "
      << o.dbg_bubbles << "], \"extensions\": [" << c.precision << ", " << c.ext_source_in_pml
      << ", " << c.ext_focus_x_set << ", " << c.ext_focus_x << ", " << c.max_picard << ", "
      << c.max_substeps << ", " << c.rp_exact << ", " << c.thermal << "], \"ascii\": " << o.ascii
      << ", \"sph\": " << o.sph << ", \"directory\": \"" << o.directory
      << "\", \"target_dir\": \"" << o.target_dir << "\", \"input_file\": \"" << o.input_file
      << "\", \"ignored\": [";
    for (size_t k = 0; k < o.ignored.size(); k++)
        j << (k ? ", " : "") << "\"" << o.ignored[k] << "\"";
    j << "]}";
    return j.str();
}

}  // namespace brp_driver
