"""Extract the reference's golden vectors into small committed fixtures.

Run in the development container (where the reference checkout and the HDF5
command-line tools exist); the fixtures it writes are plain numpy/JSON/XML data
and need neither at test time:

* ``reference1_su2.npz``  -- every dataset and ``cutoff`` attribute of the SU2
  golden flow ``test/scripted/assets/test_reference1.ref`` (hyperkagome range 3,
  Heisenberg J=1, 14 manual frequencies, cutoff 10 -> 0.3 x0.9), float32;
* ``reference2_xyz.npz``  -- the same for the XYZ golden flow
  ``test_reference2.ref`` (hyperhoneycomb range 3, Heisenberg-Kitaev j=0.1,
  k=-1, same frequencies and cutoffs, task in ``test_reference2.sh``);
* ``reference3_tri.npz``  -- the same for the TRI golden flow
  ``test_reference3.ref`` (honeycomb Kitaev-Gamma range 3);
* ``python_obs_su2.json`` -- the literal honeycomb SU2 flow values held in
  ``test/scripted/assets/test_python_obs.py`` (task in ``test_pythonObs.sh``);
* ``res/*.xml``           -- the unit cells and models those two flows use, cut
  out of the reference's ``res/`` resource files unchanged (input data).

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
H5DUMP = "/opt/conda/bin/h5dump"


def _list_objects(path):
    out = subprocess.run([H5DUMP, "-n", path], capture_output=True, text=True, check=True).stdout
    return re.findall(r"^\s*(group|dataset)\s+(\S+)\s*$", out, flags=re.M)


def _read_dataset(path, name, tmp):
    fn = os.path.join(tmp, "d.bin")
    hdr = subprocess.run([H5DUMP, "-H", "-d", name, path], capture_output=True, text=True, check=True).stdout
    dims = re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)", hdr).group(1)
    shape = tuple(int(x) for x in dims.split(","))
    if "H5T_ARRAY" in hdr:
        # h5dump cannot write array types in binary; parse the 9-digit text form
        txt = subprocess.run([H5DUMP, "-m", "%.9g", "-d", name, path], capture_output=True, text=True,
                             check=True).stdout
        body = txt[txt.index("DATA {") + 6:]
        body = re.sub(r"\(\s*\d+(,\s*\d+)*\s*\):", " ", body)
        vals = [float(x) for x in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", body)]
        return np.array(vals, dtype=np.float32).reshape(shape + (3,))
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", fn, path], capture_output=True, check=True)
    data = np.fromfile(fn, dtype="<f4")
    return data.reshape(shape)


def _read_cutoff_attr(path, group):
    out = subprocess.run([H5DUMP, "-m", "%.9g", "-a", group + "/cutoff", path], capture_output=True, text=True,
                         check=True).stdout
    return np.float32(float(re.search(r"\(0\):\s*(\S+)", out).group(1)))


def dump_ref(path, out_npz):
    arrays = {}
    with tempfile.TemporaryDirectory() as tmp:
        for kind, name in _list_objects(path):
            if kind == "dataset":
                arrays[name.lstrip("/")] = _read_dataset(path, name, tmp)
            elif re.search(r"/measurement_\d+$", name):
                arrays[name.lstrip("/") + "/cutoff"] = np.array([_read_cutoff_attr(path, name)], dtype=np.float32)
    np.savez_compressed(out_npz, **arrays)
    return len(arrays)


def _lists_in(py_text, var="reference"):
    vals = []
    for m in re.finditer(r"^%s\s*=\s*(\[.*\])\s*$" % var, py_text, flags=re.M):
        vals.append(json.loads(m.group(1)))
    return vals


def dump_python_obs(ref_root, out_json):
    with open(os.path.join(ref_root, "test/scripted/assets/test_python_obs.py")) as fh:
        text = fh.read()
    lists = _lists_in(text)
    # order in the source: basis, primitives, sites (ref 0), sites (ref 1),
    # correlation all / XX at site (1,0,0), site-resolved ZZ at 0.309031,
    # two-site ZZ at [0.381520, 2.058910], structure factor
    keys = ["basis", "primitives", "sites_ref0", "sites_ref1", "corr_site100_all", "corr_site100_xx",
            "corr_zz_cutoff0309031", "corr_zz_two_cutoffs", "structure_factor"]
    fixture = dict(zip(keys, lists))
    with open(out_json, "w") as fh:
        json.dump(fixture, fh, indent=1)


def cut_resources(ref_root, out_dir, lattices, models):
    os.makedirs(out_dir, exist_ok=True)
    for fname, tag, names in (("lattices.xml", "unitcell", lattices), ("models.xml", "model", models)):
        with open(os.path.join(ref_root, "res", fname)) as fh:
            text = fh.read()
        parts = []
        for n in names:
            m = re.search(r'<%s name="%s">.*?</%s>' % (tag, re.escape(n), tag), text, flags=re.S)
            parts.append(m.group(0))
        with open(os.path.join(out_dir, fname), "w") as fh:
            fh.write("\n\n".join(parts) + "\n")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    assets = os.path.join(ref, "test/scripted/assets")
    n = dump_ref(os.path.join(assets, "test_reference1.ref"), os.path.join(HERE, "reference1_su2.npz"))
    print("reference1_su2.npz:", n, "arrays")
    n = dump_ref(os.path.join(assets, "test_reference2.ref"), os.path.join(HERE, "reference2_xyz.npz"))
    print("reference2_xyz.npz:", n, "arrays")
    n = dump_ref(os.path.join(assets, "test_reference3.ref"), os.path.join(HERE, "reference3_tri.npz"))
    print("reference3_tri.npz:", n, "arrays")
    dump_python_obs(ref, os.path.join(HERE, "python_obs_su2.json"))
    cut_resources(ref, os.path.join(HERE, "res"), ["hyperkagome", "honeycomb", "hyperhoneycomb"],
                  ["hyperkagome-heisenberg", "honeycomb-heisenberg", "honeycomb-kitaev-gamma",
                   "hyperhoneycomb-kitaev"])
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
