"""MAT v5 reader/writer of the drop-in executable, checked against scipy.io on the
reference's own files (J written by scipy, trained weights written by libmatio) and
on a zlib-compressed file."""
import os
import subprocess

import numpy as np
import pytest
import scipy.io as sio

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    exe = tmp_path_factory.mktemp("mat") / "matv5_roundtrip"
    csrc = os.path.join(REPO, "nqes-trapped-ion_amd", "csrc")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", csrc, os.path.join(REPO, "tests", "cpp", "matv5_roundtrip.cpp"),
                    os.path.join(csrc, "matv5.cpp"), "-lz", "-o", str(exe)], check=True)
    return str(exe)


def roundtrip(tool, src, dst):
    r = subprocess.run([tool, src, dst], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return {line.split()[0]: line.split()[1:] for line in r.stdout.splitlines()}


def test_weights_file_from_libmatio(tool, tmp_path):
    src = os.path.join(GOLDEN, "ionh0.595train1N50M150K2.mat")
    out = str(tmp_path / "w.mat")
    info = roundtrip(tool, src, out)
    assert info["a"] == ["1", "2", "2", "50"] and info["w"] == ["1", "3", "2", "50", "150"]
    ref, got = sio.loadmat(src), sio.loadmat(out)
    for k in ("a", "b", "w"):
        assert got[k].dtype == np.complex128 and np.array_equal(got[k], ref[k])


def test_scipy_written_file_with_int_and_compression(tool, tmp_path):
    J = np.load(os.path.join(GOLDEN, "J_N8.npy"))
    src = str(tmp_path / "j.mat")
    sio.savemat(src, {"N": np.array([[8]], dtype=np.int64), "J": J, "c": (J + 1j * J.T)[:3]},
                do_compression=True)
    out = str(tmp_path / "j2.mat")
    info = roundtrip(tool, src, out)
    assert info["J"] == ["0", "2", "8", "8"]
    got = sio.loadmat(out)
    assert np.array_equal(got["J"], J) and got["N"][0, 0] == 8
    assert np.array_equal(got["c"], (J + 1j * J.T)[:3])


def test_malformed_file_is_an_error(tool, tmp_path):
    bad = tmp_path / "bad.mat"
    bad.write_bytes(b"MATLAB 5.0" + b" " * 106 + b"\0" * 8 + b"\x00\x01IM" + b"\x0e\x00\x00\x00\xff\xff\x00\x00")
    r = subprocess.run([tool, str(bad), str(tmp_path / "o.mat")], capture_output=True, text=True)
    assert r.returncode == 1 and "malformed" in r.stderr
