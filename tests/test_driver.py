"""The test_dvecchia_batch driver: the reference's command line (flags,
checks, exit codes of include/vecchia_helper_data.h:256-739).  CPU only: these
cases stop before any device call."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "blockvecchia_amd", "bin", "test_dvecchia_batch")


def run(*args, cwd=None):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    return subprocess.run([DRIVER] + list(args), capture_output=True, text=True, cwd=cwd,
                          timeout=60)


def test_help_and_unknown_option():
    r = run("--help")
    assert r.returncode == 0 and "Usage: test_dvecchia_batch" in r.stdout
    r = run("--bogus")
    assert r.returncode == 1 and "unrecognized option --bogus" in r.stderr


@pytest.mark.parametrize("args,msg", [
    (["--num_loc", "0"], "--num_loc 0 is invalid"),
    (["--num_loc", "x"], "--num_loc x is invalid"),
    (["--vecchia_bc", "0"], "--vecchia_bc 0 is invalid"),
    (["--vecchia_cs", "-1"], "--vecchia_cs -1 is invalid"),
    (["--dist_scale", "0"], "--dist_scale 0 is invalid"),
    (["--maxiter", "0"], "--maxiter 0 is invalid"),
    (["--tol", "0"], "--tol 0 is invalid"),
    (["--lower_bound", "0"], "--lower_bound 0 is invalid"),
    (["--upper_bound", "100"], "--upper_bound 100 is invalid"),
    (["--kernel", "nope"], "Unsupported kernel type: nope"),
])
def test_invalid_values_exit_1(args, msg):
    r = run(*args)
    assert r.returncode == 1 and msg in r.stderr, (r.returncode, r.stderr)


def test_flags_need_a_following_argument():
    # as the reference's parser: a switch is only recognised before another argument
    r = run("--num_loc", "100", "--perf")
    assert r.returncode == 1 and "unrecognized option --perf" in r.stderr


def test_only_random_ordering_and_single_method():
    r = run("--permutation", "morton", "--help")
    assert r.returncode == 0  # help wins, as it is parsed after
    r = run("--permutation", "morton", "--num_loc", "100")
    assert r.returncode == 1 and "only the random block ordering" in r.stderr


def test_missing_input_file(tmp_path):
    r = run("--xy_path", str(tmp_path / "none.csv"), "--obs_path", str(tmp_path / "none2.csv"),
            "--num_loc", "100")
    assert r.returncode == 0 and "does not exist" in r.stderr
