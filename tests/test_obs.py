"""Observable files in the reference's HDF5 layout (CPU; needs the HDF5 C library).

The reference's own golden flows are written through the writer and read back:
the hyperkagome SU(2) flow (test_reference1.ref, float oracle = bit-exact
reference arithmetic) must reproduce every measurement bit for bit, the TRI
flow (test_reference3.ref) every TRICor* group to rounding; the file structure
(meta datasets of float[3] array type, measurement groups with a float cutoff
attribute, duplicate cutoffs discarded) is checked with the HDF5 tools.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN, ref1_problem, tri_ref3_problem, xyz_ref2_problem

obs = pytest.importorskip("spinparser_amd.obs")
if not os.path.exists(obs.OBS_PATH):
    pytest.skip("observable writer not built (no HDF5 C library)", allow_module_level=True)

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def test_su2_golden_flow_roundtrip(tmp_path):
    from oracle import Oracle
    from spinparser_amd.frgcore import CorrelationRecord
    p = ref1_problem()
    zz, dd, n, _ = Oracle(p).run_flow(dtype=np.float32, nthreads=8)
    recs = [CorrelationRecord(float(p.cutoffs[k]), zz[k], dd[k]) for k in range(n)]
    path = str(tmp_path / "ref1.obs")
    assert obs.write_su2_records(path, p.lattice, recs) == n
    g = np.load(GOLDEN + "/reference1_su2.npz")
    for grp in ("SU2CorZZ", "SU2CorDD"):
        cut, data = obs.read_correlations(path, grp, (p.B, p.Rb))
        assert len(cut) == n == 34
        for k in range(n):
            assert cut[k] == g["%s/data/measurement_%d/cutoff" % (grp, k)][0]
            assert np.array_equal(data[k], g["%s/data/measurement_%d/data" % (grp, k)])
    # duplicate cutoff: discarded, as the reference does
    assert obs.write_correlation(path, "SU2CorZZ", recs[3].cutoff, p.lattice, recs[3].zz) is False
    assert len(obs.read_correlations(path, "SU2CorZZ", (p.B, p.Rb))[0]) == n


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
def test_su2_file_structure(tmp_path):
    from spinparser_amd.frgcore import CorrelationRecord
    p = ref1_problem()
    path = str(tmp_path / "s.obs")
    rec = CorrelationRecord(float(p.cutoffs[0]), np.zeros((p.B, p.Rb)), np.ones((p.B, p.Rb)))
    obs.write_su2_records(path, p.lattice, [rec])
    head = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    for grp in ("SU2CorZZ", "SU2CorDD"):
        assert 'GROUP "%s"' % grp in head
    assert head.count('DATASET "latticeVectors"') == 2 and head.count('DATASET "sites"') == 2
    assert "H5T_ARRAY { [3] H5T_IEEE_F32LE }" in head
    assert "DATASPACE  SIMPLE { ( %d, %d ) / ( %d, %d ) }" % (p.B, p.Rb, p.B, p.Rb) in head
    assert 'ATTRIBUTE "cutoff"' in head and 'GROUP "measurement_0"' in head
    # meta/sites holds the golden site positions
    g = np.load(GOLDEN + "/reference1_su2.npz")
    _, _, sites = obs.lattice_geometry(p.lattice)
    assert np.allclose(sites, g["SU2CorZZ/meta/sites"], atol=1e-6)


def test_tri_golden_flow_roundtrip(tmp_path):
    from oracle import TRI_GROUPS, TriOracle
    from spinparser_amd.frgcore import TRICorrelationRecord
    p = tri_ref3_problem()
    rec, n, _ = TriOracle(p).run_flow(dtype=np.float32, nthreads=8)
    recs = [TRICorrelationRecord(float(p.cutoffs[k]), {g: rec[k, i] for i, g in enumerate(TRI_GROUPS)})
            for k in range(n)]
    path = str(tmp_path / "ref3.obs")
    assert obs.write_tri_records(path, p.lattice, recs) == n
    g = np.load(GOLDEN + "/reference3_tri.npz")
    for grp in TRI_GROUPS:
        cut, data = obs.read_correlations(path, grp, (p.B, p.Rb))
        assert len(cut) == n
        for k in range(n):
            assert cut[k] == g["%s/data/measurement_%d/cutoff" % (grp, k)][0]
            assert np.abs(data[k] - g["%s/data/measurement_%d/data" % (grp, k)]).max() < 1e-6


def test_xyz_golden_flow_bit_exact_file(tmp_path):
    """XYZ flow of test_reference2 written as XYZCor* groups: every dataset and
    cutoff attribute equals the golden file bit for bit, and the meta group
    holds its site positions."""
    from oracle import XYZ_GROUPS, XyzOracle
    from spinparser_amd.frgcore import TRICorrelationRecord
    p = xyz_ref2_problem()
    rec, n, _ = XyzOracle(p).run_flow(dtype=np.float32, nthreads=8)
    recs = [TRICorrelationRecord(float(p.cutoffs[k]), {g: rec[k, i] for i, g in enumerate(XYZ_GROUPS)})
            for k in range(n)]
    path = str(tmp_path / "ref2.obs")
    assert obs.write_xyz_records(path, p.lattice, recs) == n
    g = np.load(GOLDEN + "/reference2_xyz.npz")
    for grp in XYZ_GROUPS:
        cut, data = obs.read_correlations(path, grp, (p.B, p.Rb))
        assert len(cut) == n
        for k in range(n):
            assert cut[k] == g["%s/data/measurement_%d/cutoff" % (grp, k)][0]
            assert np.array_equal(data[k], g["%s/data/measurement_%d/data" % (grp, k)])
    _, _, sites = obs.lattice_geometry(p.lattice)
    assert np.allclose(sites, g["XYZCorXX/meta/sites"], atol=1e-6)


def test_writer_exports_header_symbols():
    import ctypes
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "include", "spobs.h")) as fh:
        declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(sp_obs_\w+)\s*\(", fh.read(), flags=re.M))
    assert declared == {"sp_obs_write_correlation", "sp_obs_count", "sp_obs_read_correlation", "sp_obs_last_error"}
    lib = ctypes.CDLL(obs.OBS_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_checkpoint_roundtrip_and_duplicates(tmp_path):
    path = str(tmp_path / "flow.data")
    rng = np.random.default_rng(3)
    a = {"v2": rng.standard_normal(14), "v4dd": rng.standard_normal(300), "v4ss": rng.standard_normal(300)}
    b = {k: v + 1.0 for k, v in a.items()}
    assert obs.write_checkpoint(path, 10.0, a, append=False) == 0
    assert obs.write_checkpoint(path, 9.0, b, append=True) == 1
    assert obs.write_checkpoint(path, 9.0, b, append=True) == -1          # same cutoff: skipped
    sizes = {k: v.size for k, v in a.items()}
    cut, got = obs.read_checkpoint(path, sizes)                             # last checkpoint
    assert cut == np.float32(9.0)
    for k in a:
        assert np.array_equal(got[k], b[k].astype(np.float32))
    cut, got = obs.read_checkpoint(path, sizes, 0)
    assert cut == np.float32(10.0) and np.array_equal(got["v4ss"], a["v4ss"].astype(np.float32))
    # a new file without append replaces the old one
    assert obs.write_checkpoint(path, 5.0, a, append=False) == 0
    with pytest.raises(obs.ObsError):
        obs.read_checkpoint(path, sizes, 1)
