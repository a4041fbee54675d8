// options.h -- command line and configuration file of the `bubbles` driver.
//
// Reproduces the option surface of the reference's runSimulation
// (src/bubbles.cpp:61-205), which uses Boost.Program_options 1.40; Boost is
// not a dependency here, so its config-file and command-line semantics are
// implemented directly:
//   * `[section]` lines prefix the following keys with "section.";
//   * `#` starts a comment anywhere on a line; names and values are trimmed;
//   * `name = value` lines only; any other non-empty line is a syntax error;
//   * unregistered names are ignored (allow_unregistered);
//   * bool values: 1/0, true/false, yes/no, on/off (any case), empty = true;
//   * an option given twice in one source is an error ("multiple
//     occurrences"); across sources the first stored value wins, except that
//     a default value is replaced by a later explicit one;
//   * switches (ascii, sph) count as given whatever their value.
#pragma once

#include <set>
#include <string>
#include <vector>

#include "../../include/bubbles_rp.h"

namespace brp_driver {

struct Options {
    brp_config cfg{};
    // [debug] (debug_t, bubble_CUDA.h:48-57)
    int display = 4;
    bool dbg_fg = true, dbg_T = true, dbg_p0 = false, dbg_vxy = false, dbg_bubbles = true;
    bool has_fg_min = false, has_fg_max = false, has_T_min = false, has_T_max = false;
    bool has_p_max = false;
    double fg_min = 0, fg_max = 0, T_min = 0, T_max = 0, p_max = 0;
    // generic options
    bool ascii = false, sph = false;
    std::string directory = "./rawdata/";
    bool directory_defaulted = true;
    std::string input_file;
    int num_iters = 1;  // transducer.pid_iterate
    bool help = false, help_all = false, dry_run = false;
    std::string target_dir;  // directory with a trailing '/', fixed before the file is read
    std::vector<std::string> ignored;  // unregistered names seen
    std::set<std::string> stored_names;  // explicitly stored by an earlier source
};

// Parses argv then the configuration file named by the positional argument.
// Returns 0 on success; on error fills *err and returns nonzero.
int parse_command_line(int argc, char **argv, Options *o, std::string *err);
int parse_config_file(const std::string &path, Options *o, std::string *err);
int parse_config_text(const std::string &text, Options *o, std::string *err);

std::string help_text(bool all);
std::string options_json(const Options &o);

}  // namespace brp_driver
