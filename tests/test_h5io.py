"""The reference's HDF5 file layout (libdlbc_h5.so), on the CPU.

* write/read round trips of float64 and int32 fields with the time step;
* files the reference wrote (its checkpoint and output golden data, read
  where the reference tree is present) read back equal to the committed
  fixtures, and files written here have the reference's structure
  (h5dump -H: dataset, types, groups and attributes).
The restart of a lattice from such files is a GPU test (test_gpu_parity.py).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from cases import golden

REF = "/root/reference/tests/runnable/tests/technical/checkpoint-3d/reference-data"


@pytest.fixture(scope="module")
def H():
    from dlbc_amd import h5io
    if not h5io.available():
        pytest.skip("libdlbc_h5.so not built (no HDF5 library)")
    return h5io


def test_roundtrip_f64_and_i32(H, tmp_path):
    rng = np.random.default_rng(3)
    a = rng.random((5, 6, 7, 19))
    p = str(tmp_path / "cp-fluids0-x-t00000042.h5")
    H.write_field(p, a, 42, ["[lb]", "timesteps = 10"], checkpoint=True)
    b, t = H.read_field(p)
    assert t == 42 and b.dtype == np.float64
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    m = (rng.random((5, 6, 7)) < 0.3).astype(np.int32)
    p2 = str(tmp_path / "mask.h5")
    H.write_field(p2, m)
    c, t2 = H.read_field(p2)
    assert t2 is None and c.dtype == np.int32 and np.array_equal(c, m)


def test_shape_mismatch_is_an_error(H, tmp_path):
    import ctypes as C
    p = str(tmp_path / "f.h5")
    H.write_field(p, np.zeros((4, 4, 4)))
    out = np.zeros((4, 4, 5))
    dims = (C.c_int64 * 3)(4, 4, 5)
    assert H.lib().dlbc_h5_read(p.encode(), 0, out.ctypes.data, 3, dims, None) == -3
    assert H.lib().dlbc_h5_read(b"/nonexistent/f.h5", 0, out.ctypes.data, 3, dims, None) == -2


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_reads_the_reference_files(H):
    """The reference's checkpoint (t = 0) and output (t = 5) of checkpoint-3d
    read back bit for bit equal to the fixtures exported with h5dump."""
    g, _ = golden("checkpoint_3d")
    for f, col in ((0, "red"), (1, "blue")):
        a, t = H.read_field(os.path.join(REF, "cp", "cp-fluids%d-checkpoint-3d-t00000000.h5" % f))
        assert t == 0
        assert np.array_equal(a.view(np.uint64), g["cp_fluids%d_t0" % f].view(np.uint64))
        b, t5 = H.read_field(os.path.join(REF, "population-%s-checkpoint-3d-t00000005.h5" % col))
        assert t5 is None
        assert np.array_equal(b.view(np.uint64), g["population_%s_t5" % col].view(np.uint64))
    m, _ = H.read_field(os.path.join(REF, "cp", "cp-mask-checkpoint-3d-t00000000.h5"))
    assert m.dtype == np.int32 and np.array_equal(m, g["cp_mask_t0"])


def _structure(path):
    """Dataset, type, dims, groups and attribute names of an HDF5 file."""
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    txt = subprocess.run([h5dump, "-H", path], capture_output=True, text=True, check=True).stdout
    txt = txt.split("\n", 1)[1]  # drop the file name line
    return [re.sub(r"\s+", " ", l.strip()) for l in txt.splitlines() if l.strip()]


@pytest.mark.skipif(not os.path.isdir(REF) or not (shutil.which("h5dump") or
                                                   os.path.exists("/opt/conda/bin/h5dump")),
                    reason="reference tree or h5dump not present")
def test_checkpoint_structure_matches_the_reference(H, tmp_path):
    g, _ = golden("checkpoint_3d")
    ref = os.path.join(REF, "cp", "cp-fluids0-checkpoint-3d-t00000000.h5")
    ref_input = _structure(ref)
    n_input = int(re.search(r"\( (\d+) \)", ref_input[8]).group(1))
    p = str(tmp_path / "cp-fluids0-checkpoint-3d-t00000000.h5")
    H.write_field(p, g["cp_fluids0_t0"], 0, ["line"] * n_input, checkpoint=True)
    assert _structure(p) == ref_input
