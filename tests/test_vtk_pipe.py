"""The reference's own fluid mask, build/scenarios/pipe.vtk (read through
build/scenarios/pipe.xml:4-11 by include/io/vtk.hpp:94-157): both readers of
this build, lbm_amd.scenario.read_vtk_mask and the C++ front-end's
lbm::io::read_vtk_mask, return its DIMENSIONS 250 54 54 as the interior
lengths, its non-unit ORIGIN and SPACING, and 408,838 fluid points of
729,000 in x-fastest order, through the file's blank header lines and padded
keyword columns. The committed block tests/golden/pipe_crop.vtk.gz
(x = 80 .. 119 around the pipe's constriction, written by
tests/golden/make_pipe_crop.py) reads the same through both, and equals the
same block of the full mask where the reference is present."""
import gzip
import subprocess
from pathlib import Path

import numpy as np
import pytest

from lbm_amd import scenario

ROOT = Path(__file__).resolve().parents[1]
PIPE = Path("/root/reference/build/scenarios/pipe.vtk")
CROP = ROOT / "tests" / "golden" / "pipe_crop.vtk.gz"


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("probe") / "vtk_mask_probe"
    r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "vtk_mask_probe.cpp"),
                        "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def read_cpp(probe, path):
    r = subprocess.run([str(probe), str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = dict(line.split(" ", 1) for line in r.stdout.splitlines())
    dims = tuple(int(v) for v in out["dims"].split())
    mask = np.frombuffer(out["mask"].encode(), dtype=np.uint8) - ord("0")
    return (dims, tuple(float(v) for v in out["origin"].split()), tuple(float(v) for v in out["spacing"].split()),
            int(out["fluid"]), mask.reshape(dims[::-1]))


@pytest.mark.skipif(not PIPE.exists(), reason="reference sources not present on this machine")
def test_reference_pipe_mask_through_both_readers(probe):
    m, origin, spacing = scenario.read_vtk_mask(PIPE)
    assert m.shape == (54, 54, 250)
    assert int(m.sum()) == 408838 and m.size == 729000
    assert origin == (-597.205589, -32.860778, -32.860778)
    assert spacing == (5.588822, 1.22844, 1.22844)
    dims, o2, s2, nf, m2 = read_cpp(probe, PIPE)
    assert dims == (250, 54, 54) and nf == 408838
    assert np.allclose(o2, origin, rtol=0, atol=1e-9) and np.allclose(s2, spacing, rtol=0, atol=1e-9)
    assert np.array_equal(m2, m)
    # the committed block is this mask's x = 80 .. 119
    cm, corigin, _ = scenario.read_vtk_mask(_unpacked(CROP, probe.parent))
    assert np.array_equal(cm, m[:, :, 80:120])
    assert abs(corigin[0] - (origin[0] + 80 * spacing[0])) < 1e-6


def _unpacked(gz: Path, where: Path) -> Path:
    dst = where / gz.name[:-3]
    dst.write_bytes(gzip.decompress(gz.read_bytes()))
    return dst


def test_pipe_block_fixture_through_both_readers(probe, tmp_path):
    path = _unpacked(CROP, tmp_path)
    m, origin, spacing = scenario.read_vtk_mask(path)
    assert m.shape == (54, 54, 40) and int(m.sum()) == 29649
    assert spacing == (5.588822, 1.22844, 1.22844)
    dims, o2, s2, nf, m2 = read_cpp(probe, path)
    assert dims == (40, 54, 54) and nf == 29649
    assert np.array_equal(m2, m)
    assert np.allclose(o2, origin, rtol=0, atol=1e-9) and np.allclose(s2, spacing, rtol=0, atol=1e-9)
    # the constriction: the fluid cross-section narrows along x inside the block
    area = m.sum(axis=(0, 1))
    assert area.max() > 4 * area.min()
