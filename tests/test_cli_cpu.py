"""The NIfTI-1 volume reader/writer (Python and the command line tool's C++
one) and rec_spect's command line handling, without a GPU: every case here
ends before the tool's first library call."""
import gzip
import os
import struct
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from niftyrec_amd import nifti

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REC_SPECT = os.path.join(ROOT, "niftyrec_amd", "rec_spect")


@pytest.mark.parametrize("name", ["v.nii", "v.nii.gz"])
def test_nifti_round_trip(tmp_path, name):
    rng = np.random.default_rng(0)
    v = rng.random((5, 4, 3), dtype=np.float32)
    p = str(tmp_path / name)
    nifti.write(p, v, (1.5, 2.0, 2.5))
    got, pix = nifti.read(p)
    assert got.shape == (5, 4, 3) and np.array_equal(got, v)
    assert pix == (1.5, 2.0, 2.5)
    raw = gzip.open(p).read() if name.endswith(".gz") else open(p, "rb").read()
    assert len(raw) == 352 + v.size * 4
    assert struct.unpack("<i", raw[:4])[0] == 348 and raw[344:348] == b"n+1\0"
    assert struct.unpack("<8h", raw[40:56])[:4] == (3, 3, 4, 5)  # dim: nx, ny, nz


def test_nifti_int16_scaled_and_big_endian(tmp_path):
    """Other scalar types are converted with scl_slope / scl_inter; headers and
    data written big-endian are read too."""
    vals = np.arange(24, dtype=">i2").reshape(2, 3, 4)
    h = bytearray(352)
    struct.pack_into(">i", h, 0, 348)
    struct.pack_into(">8h", h, 40, 3, 4, 3, 2, 1, 1, 1, 1)
    struct.pack_into(">2h", h, 70, 4, 16)
    struct.pack_into(">8f", h, 76, 1, 1, 1, 1, 1, 1, 1, 1)
    struct.pack_into(">3f", h, 108, 352.0, 0.5, 2.0)
    h[344:348] = b"n+1\0"
    p = str(tmp_path / "i16.nii")
    with open(p, "wb") as f:
        f.write(bytes(h) + vals.tobytes())
    got, _ = nifti.read(p)
    assert np.array_equal(got, vals.astype(np.float32) * 0.5 + 2.0)


def test_nifti_rejects_non_nifti(tmp_path):
    p = str(tmp_path / "x.nii")
    open(p, "wb").write(b"\0" * 400)
    with pytest.raises(ValueError):
        nifti.read(p)


def _run(*args):
    return subprocess.run([REC_SPECT] + list(args), capture_output=True, text=True, timeout=60)


def test_rec_spect_built_and_describes_itself():
    assert os.path.exists(REC_SPECT), "rec_spect is not built (python -m niftyrec_amd.build)"
    r = _run("--xml")
    assert r.returncode == 0
    root = ET.fromstring(r.stdout)
    names = [e.findtext("name") for e in root.find("parameters")]
    assert {"sinogram", "output", "attenuation", "iterations", "subsets", "legacy-subsets"} <= set(names)
    assert _run("--help").returncode == 0


def test_rec_spect_argument_errors(tmp_path):
    assert _run("-o", "out.nii").returncode == 1                     # sinogram missing
    assert _run("-s", "a.img", "-o", "b.nii").returncode == 1        # unknown format
    assert _run("-s", str(tmp_path / "none.nii"), "-o", "b.nii").returncode == 1  # no such file
    assert _run("-s", "a.nii", "-o", "b.nii", "-i", "ten").returncode == 1        # bad number
    assert _run("-s", "a.nii", "-o", "b.nii", "--bogus").returncode == 1


def test_rec_spect_reads_python_written_volumes(tmp_path):
    """The C++ reader parses the sizes of Python-written volumes: an
    attenuation map whose depth differs from the sinogram's width is refused
    before any reconstruction starts."""
    s = str(tmp_path / "s.nii.gz")
    m = str(tmp_path / "m.nii")
    nifti.write(s, np.ones((5, 6, 8), np.float32))    # [N=8, m=6, 5 cameras]
    nifti.write(m, np.ones((7, 6, 8), np.float32))    # [8, 6, 7]: depth 7 != N
    r = _run("-s", s, "-o", str(tmp_path / "o.nii"), "-a", m)
    assert r.returncode == 1
    assert "attenuation map must be [N x m x N]" in r.stderr
