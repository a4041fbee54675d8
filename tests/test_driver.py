"""The `bubbles` driver's command line and configuration-file semantics
(Boost.Program_options 1.40 as used by the reference's src/bubbles.cpp:61-205).
CPU only: --dry-run parses and prints the resolved options without a GPU."""
import json
import os
import subprocess

import pytest

import cases

BIN = os.path.join(cases.ROOT, "cuda-rp_amd", "bin", "bubbles")
SAMPLE = os.path.join(os.path.dirname(__file__), "golden", "Sample.txt")


def run(*args, cwd=None):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (make -C cuda-rp_amd)")
    return subprocess.run([BIN, *args], capture_output=True, text=True, cwd=cwd)


def dry(*args, text=None, tmp_path=None):
    if text is not None:
        path = os.path.join(str(tmp_path), "c.txt")
        open(path, "w").write(text)
        args = (*args, path)
    r = run("--dry-run", *args)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_sample_config_parses_like_the_reference():
    o = dry(SAMPLE)
    assert o["grid"] == [300, 600, 30e-3, 60e-3]
    assert o["pml"] == [0, 1, 0, 1, 10, 0.1, 2]
    assert o["simulation"] == [0.1, 0.1, 20000, 0, 100, 2, 2]
    # amp, freq, focal_dist, pid, pid_start_step, init_control, pid_iterate, box,
    # cylindrical, on_wave, off_wave, plane_v, plane_p, focused_v, focused_p
    assert o["transducer"] == [1.5e-2, 1e6, 40e-3, 1, 5000, 0, 2, 30e-3, 1, 2, 0, 0, 0, 0, 1]
    assert o["bubbles"] == [1, 1e-7, 6.5e-7, 0.1e6, 293.15, 1.4, 7.2e-2, 891e-6, 1e3, 26.2e-3]
    assert o["debug"] == [7, 1, 1, 1, 0, 0]
    # `sph = 1` stores the sph switch; plane_wave_v/p are not registered keys
    assert o["sph"] == 1 and o["ascii"] == 0
    assert o["ignored"] == ["transducer.plane_wave_v", "transducer.plane_wave_p"]
    assert o["target_dir"] == "./rawdata/"


def test_defaults_of_declared_keys(tmp_path):
    o = dry(text="[grid]\nx = 8\n", tmp_path=tmp_path)
    assert o["debug"] == [4, 1, 1, 0, 0, 1]   # display_dim 4, fg, T, !p, !v, bubbles (=10 -> true)
    t = o["transducer"]
    assert t[3] == 0 and t[4] == 1 and t[5] == 0 and t[6] == 1  # pid, pid_start_step, init_control, iterate
    assert o["bubbles"][0] == 0


@pytest.mark.parametrize("val,want", [("1", 1), ("0", 0), ("true", 1), ("FALSE", 0), ("yes", 1),
                                      ("No", 0), ("on", 1), ("off", 0), ("", 1)])
def test_bool_values(tmp_path, val, want):
    o = dry(text=f"[pml]\nx0 = {val}\n", tmp_path=tmp_path)
    assert o["pml"][0] == want


def test_comments_sections_and_whitespace(tmp_path):
    txt = ("# leading comment\n  [grid]   \n x=12 # trailing\n\ty\t=\t34\t\n"
           "[ simulation ]\ncfl = 0.25\n[simulation]\norder = 2 #x\n")
    o = dry(text=txt, tmp_path=tmp_path)
    assert o["grid"][:2] == [12, 34]
    # "[ simulation ]" gives the prefix " simulation ." -- not a registered section
    assert o["simulation"][0] == 0
    assert " simulation .cfl" in o["ignored"]


@pytest.mark.parametrize("txt,msg", [
    ("[grid]\nx = 1\nx = 2\n", "multiple occurrences"),
    ("[grid]\nx = 1.5\n", "invalid value"),
    ("[pml]\nx0 = maybe\n", "invalid value"),
    ("[grid]\njust a line\n", "unrecognized line"),
])
def test_config_errors(tmp_path, txt, msg):
    path = tmp_path / "bad.txt"
    path.write_text(txt)
    r = run("--dry-run", str(path))
    assert r.returncode != 0 and msg in r.stderr


def test_command_line_forms(tmp_path):
    path = tmp_path / "c.txt"
    path.write_text("[grid]\nx = 4\n")
    for args in (("-d", "out"), ("--directory", "out"), ("--directory=out",), ("-dout",),
                 ("--dir", "out")):
        o = json.loads(run("--dry-run", *args, str(path)).stdout)
        assert o["target_dir"] == "out/" and o["directory"] == "out"
    o = json.loads(run("--dry-run", "-a", "-s", str(path)).stdout)
    assert o["ascii"] == 1 and o["sph"] == 1
    o = json.loads(run("--dry-run", "--unknown-flag", str(path)).stdout)  # allow_unregistered
    assert o["input_file"] == str(path)
    r = run("--dry-run", str(path), str(path))
    assert r.returncode != 0 and "multiple occurrences" in r.stderr
    r = run()
    assert r.returncode == 1 and "Please specify an input configuration file" in r.stdout
    assert "--directory" in run("--help").stdout
    assert "bubbles.liquid_viscosity" in run("--help-all").stdout


def test_first_stored_value_wins_and_directory_quirk(tmp_path):
    path = tmp_path / "c.txt"
    path.write_text("directory = fromfile\nascii = 0\n[grid]\nx = 4\n")
    o = json.loads(run("--dry-run", "-d", "fromcli", str(path)).stdout)
    assert o["directory"] == "fromcli"            # the command line is stored first
    o = json.loads(run("--dry-run", str(path)).stdout)
    assert o["directory"] == "fromfile"           # a default is replaced by the file
    assert o["target_dir"] == "./rawdata/"        # ...but target_dir was fixed before the file
    assert o["ascii"] == 1                        # a switch counts as given whatever its value


def test_debug_p_min_key_cannot_be_set(tmp_path):
    """The reference registers "debug p_min" with a space (bubbles.cpp:127),
    so [debug] p_min is an unregistered key."""
    o = dry(text="[debug]\np_min = -5\np_max = 7\n", tmp_path=tmp_path)
    assert "debug.p_min" in o["ignored"]


def test_extension_keys(tmp_path):
    o = dry(text="[extensions]\nprecision = mixed\nsource_in_pml = 1\nfocus_x = 0.0125\n"
                 "rp_exact = yes\nmax_picard = 7\n", tmp_path=tmp_path)
    assert o["extensions"] == [1, 1, 1, 0.0125, 7, 0, 1, 1]
    o = dry(text="[extensions]\nthermal = 0\n", tmp_path=tmp_path)
    assert o["extensions"][-1] == 0
