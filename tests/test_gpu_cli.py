"""rec_spect on the GPU: NIfTI sinogram in, reconstruction through the
library's drivers, NIfTI activity out; the output equals the same driver
called in-process (bit for bit) and the oracle's updates (parity bounds).
The PSF is the reference tool's: 5 x 5 ones from et_array_make_psf."""
import os
import subprocess

import numpy as np
import pytest

from conftest import report_parity
from niftyrec_amd import api, nifti, phantoms
from oracle import orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REC_SPECT = os.path.join(ROOT, "niftyrec_amd", "rec_spect")
N, M, V = 32, 24, 16
LAST = 360.0 * (1 - 1.0 / V)


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    act, mu = phantoms.phantom((N, M, N), 3, 7, lesion_axes=(1.5, 3.0))
    mu = (mu / 0.15 * 0.08).astype(np.float32)
    cams = np.radians(np.linspace(0.0, LAST, V)).astype(np.float32)
    psf = np.ones((N, 5, 5), np.float32)
    y = orc.project(act, mu, cams, psf)
    y = np.random.default_rng(3).poisson(y * 2.0).astype(np.float32) / 2.0
    nifti.write(str(d / "sino.nii.gz"), y)
    nifti.write(str(d / "mu.nii"), mu)
    cams3 = np.zeros((3, V), np.float32)
    cams3[1] = cams
    return d, y, mu, psf, cams3


def _rec(d, out, *args):
    print("rec_spect", " ".join(args), flush=True)
    r = subprocess.run([REC_SPECT, "-s", str(d / "sino.nii.gz"), "-o", str(d / out), "-x", "1", "-c", "0",
                        "-C", repr(LAST)] + list(args), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got, _ = nifti.read(str(d / out))
    assert got.shape == (N, M, N)
    return got


def test_rec_spect_mlem(data):
    d, y, mu, psf, cams3 = data
    got = _rec(d, "mlem.nii", "-i", "3", "-u", "1", "-a", str(d / "mu.nii"))
    lib = api.mlem(y, 0.0, LAST, 3, psf=psf, attenuation=mu, epsilon=1e-7)
    assert np.array_equal(got, lib)
    x32, x64 = np.ones_like(mu), np.ones(mu.shape)
    for _ in range(3):
        x32 = orc.mlem_step(x32, y, mu, cams3, psf, eps=1e-7)
        x64 = orc.mlem_step(x64, y, mu, cams3, psf, eps=1e-7, dtype=np.float64)
    report_parity("rec_spect MLEM (32x24x32, 16 views, 3 iterations)", got, x32, x64)


def test_rec_spect_ordered_osem(data):
    d, y, mu, psf, cams3 = data
    got = _rec(d, "osem.nii.gz", "-i", "2", "-u", "4", "-a", str(d / "mu.nii"))
    lib = api.osem_ordered(y, 4, 0.0, LAST, 2, psf=psf, attenuation=mu, epsilon=1e-7)
    assert np.array_equal(got, lib)
    x, x64 = np.ones_like(mu), np.ones(mu.shape)
    for _ in range(2):
        for s in range(4):
            x = orc.mlem_step(x, y[s::4], mu, cams3[:, s::4], psf, eps=1e-7)
            x64 = orc.mlem_step(x64, y[s::4], mu, cams3[:, s::4], psf, eps=1e-7, dtype=np.float64)
    report_parity("rec_spect ordered OSEM (4 subsets, 2 passes)", got, x, x64)


def test_rec_spect_legacy_subsets_without_attenuation(data):
    """--legacy-subsets: the reference tool's et_array_osem on subsets drawn
    with glibc rand() seeded with --seed: the same volume as et_array_osem
    called in this process after the same srand (bit for bit), the default
    seed 1 the same as an unseeded process's draws, and two seeds give
    different volumes."""
    import ctypes
    d, y, mu, psf, cams3 = data
    got = _rec(d, "legacy.nii", "-i", "2", "-u", "4", "-L", "-S", "5")
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(5)
    lib = api.osem(y, 4, 0.0, LAST, 2, psf=psf, epsilon=1e-7)
    assert np.array_equal(got, lib)
    other = _rec(d, "legacy6.nii", "-i", "2", "-u", "4", "-L", "-S", "6")
    assert not np.array_equal(got, other)
    default = _rec(d, "legacy1.nii", "-i", "2", "-u", "4", "-L")
    libc.srand(1)
    assert np.array_equal(default, api.osem(y, 4, 0.0, LAST, 2, psf=psf, epsilon=1e-7))
    assert np.isfinite(got).all() and got.min() >= 0
