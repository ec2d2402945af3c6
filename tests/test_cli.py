"""svm-train's command line without a device: usage, the reference's input
and option errors (svm-train/svm-train.c), and a loud failure when training
cannot run."""
import os
import subprocess

import pytest

from cli_util import SVM_TRAIN


def run(args, **kw):
    return subprocess.run([SVM_TRAIN] + args, capture_output=True, text=True, timeout=60, **kw)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(SVM_TRAIN):
        pytest.fail(f"{SVM_TRAIN} not built (make -C libsvm-cuda_amd)")


def test_usage_without_arguments():
    r = run([])
    assert r.returncode == 1 and r.stdout.startswith("Usage: svm-train [options] training_set_file")


def test_wrong_input_format_names_the_line(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("1 1:0.5 2:1\n-1 3:x\n")
    r = run([str(f)])
    assert r.returncode == 1 and "Wrong input format at line 2" in r.stderr


def test_indices_must_ascend(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("1 2:0.5 1:1\n")
    assert "Wrong input format at line 1" in run([str(f)]).stderr


def test_cross_validation_needs_two_folds(tmp_path):
    f = tmp_path / "ok.txt"
    f.write_text("1 1:0.5\n-1 1:1\n")
    r = run(["-v", "1", str(f)])
    assert r.returncode == 1 and "n-fold cross validation: n must >= 2" in r.stderr


def test_parameter_errors_from_check_parameter(tmp_path):
    f = tmp_path / "ok.txt"
    f.write_text("1 1:0.5\n-1 1:1\n")
    r = run(["-s", "7", str(f)])
    assert r.returncode == 1 and r.stderr.strip() == "ERROR: unknown svm type"
    r = run(["-s", "1", "-n", "1.5", str(f)])
    assert r.stderr.strip() == "ERROR: nu <= 0 or nu > 1"


def test_missing_file(tmp_path):
    r = run([str(tmp_path / "none.txt")])
    assert r.returncode == 1 and "can't open input file" in r.stderr


def test_no_device_fails_loudly(tmp_path):
    import libsvm_hip
    if libsvm_hip.device_count() > 0:
        pytest.skip("a device is visible: covered by the GPU tests")
    f = tmp_path / "ok.txt"
    f.write_text("1 1:0.5\n-1 1:1\n")
    r = run([str(f), str(tmp_path / "m")])
    assert r.returncode == 1 and not (tmp_path / "m").exists()
