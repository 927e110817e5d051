"""Writes the host-side fixtures under tests/golden/host/ and their expected
outputs (host_golden.json) from the reference's own readers and writers,
compiled where they lie by `make -C oracle ref` (oracle/_ref/: ref_config
around src/configreader.cpp, ref_pdb around src/pdbio.cpp, ref_dcd around
src/dcdio.cpp, ref_xyz around src/xyzio.cpp).

The inputs are this repository's own: a configuration file exercising the
reader's quirks, a PDB pair of a 13-protofilament x 3-dimer microtubule with
two free dimers in chain X (lattice of mt_amd/lattice.py), an edge-case PDB,
a small XYZ file and DCD frames.  Run from the repository root in a checkout
where /root/reference and oracle/_ref exist:

    python tests/golden/make_host_fixtures.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(ROOT, "tests", "golden", "host")
REF = os.path.join(ROOT, "oracle", "_ref")

QUIRKS = """# configuration quirks (lines starting with '#', ' ' or tab are skipped)
 indented_is_skipped 5
\tthis_too 6
dt 200 #200
rseed\t1234567
steps    1000000
float_zero 0
float_zero_spelled 0.000000
float_zero_bad 0.00
float_sci 2.85e4
int_bad abc
int_lead 12abc
yes1 yes
yes2 On
yes3 TRUE
no1 1
no2 Nope
dir out
file <dir>/traj_<run>.dcd
twice <dir>/<dir>
first_wins 1
first_wins 2
comment_only #hidden
tabbed\t\t 7 8
mask_words a b c
"""

# (kind, name, default or None): each query runs in its own process so a
# required-missing or malformed value (the reference dies) is recorded alone.
QUERIES = [
    ("float", "dt", None), ("int", "rseed", None), ("long", "steps", None), ("long", "stride", "-1"),
    ("float", "float_zero", None), ("float", "float_zero_spelled", None), ("float", "float_zero_bad", None),
    ("float", "float_sci", None), ("int", "int_bad", None), ("int", "int_lead", None),
    ("yesno", "yes1", None), ("yesno", "yes2", None), ("yesno", "yes3", None), ("yesno", "no1", None),
    ("yesno", "no2", None), ("yesno", "absent_yes", "1"), ("yesno", "absent_no", "0"), ("yesno", "absent_req", None),
    ("masked", "file", None), ("masked", "twice", None), ("str", "file", None), ("str", "dt", None),
    ("int", "first_wins", None), ("str", "comment_only", None), ("str", "tabbed", None), ("masked", "tabbed", None),
    ("masked", "mask_words", None), ("str", "indented_is_skipped", "none"), ("float", "absent_float", "2.5"),
    ("int", "absent_int", "7"), ("masked", "absent_masked", "<dir>/x"), ("int", "override_me", "1"),
]
OVERRIDES = ["override_me=42", "dt=1.5"]


def ref_query(path, kind, name, default, overrides=()):
    q = f"{kind} {name}" + (f" {default}" if default is not None else "") + "\n"
    r = subprocess.run([os.path.join(REF, "ref_config"), path, *overrides], input=q, capture_output=True, text=True)
    for line in r.stderr.splitlines():
        if line.startswith("@@"):
            k, v = line[2:].split("\t", 1)
            if k == name:
                return {"ok": True, "value": v}
    return {"ok": False}


def pdb_line(serial, name, chain, resid, x, y, z, occ=0.0, beta=0.0, extra=""):
    return "ATOM  %5d %-4s GLY %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f%s\n" % (serial, name, chain, resid, x, y, z, occ,
                                                                        beta, extra)


def write_run_files(d):
    from mt_amd import lattice
    os.makedirs(d, exist_ok=True)
    mt = lattice.microtubule(3)
    rng = np.random.default_rng(7)
    fd = lattice.free_dimers(2, 60.0, rng, first_id=mt.n + 1, chain="X", first_resid=500)
    fd.xyz += np.array([30.0, 30.0, 0.0])
    a = lattice.concat(mt, fd)
    with open(os.path.join(d, "tube_xyz.pdb"), "w") as fx, open(os.path.join(d, "tube_ang.pdb"), "w") as fa:
        for k in range(a.n):
            nm, ch = a.names[k], chr(a.chains[k])
            fx.write(pdb_line(int(a.ids[k]), " " + nm, ch, int(a.resid[k]), *a.xyz[k], extra="         " + ch))
            fa.write(pdb_line(int(a.ids[k]), " " + nm, ch, int(a.resid[k]), *a.ang[k], extra="         " + ch))
        fx.write("END\n")
        fa.write("END\n")
    with open(os.path.join(d, "main.conf"), "w") as f:
        f.write("""#run of a 13x3-dimer tube with two free dimers
device			0
rseed 			1234567
dt 				200 #ps
LJPairsCutoff		8
LJPairsUpdateFreq	100
stride 			100
forcefield      ff.conf
conditions      cond.conf
coordinates_xyz	<dir>/tube_xyz.pdb
coordinates_ang	<dir>/tube_ang.pdb
dir .
restart_xyz 		restart_xyz_<run>.xyz
restart_ang  		restart_ang_<run>.xyz
restartkey          restart_key.txt
dcd_xyz 		traj_<run>.dcd
dcd_ang 		traj_<run>.dcd_ang
steps 		    1000
fix 			1
runnum		   2
tubule_length no
""")
    with open(os.path.join(d, "ff.conf"), "w") as f:
        f.write("""C   300.0
D_long 11.5
A_long 3.0
D_lat 6.5
A_lat 4.0
LJ_on yes
LJSigma 4.0
LJScale 0.5
B_psi 9125
B_fi  9125
B_theta 90
psi0   0
fi0    0
theta0_gdp 		0.2
theta0_gtp		0.0
barrier yes
a_barr_long 3.4
r_barr_long 0.4
w_barr_long 0.10
a_barr_lat 1.9
r_barr_lat 0.4
w_barr_lat 0.15
rep_h 				160
rep_r				80.0
rep_eps     		0.005
repulsive_walls  	no
rep_leftborder  	0.0
seam_coeff	1
""")
    with open(os.path.join(d, "cond.conf"), "w") as f:
        f.write("""Temp 		    300
viscosity 		2.85e4
is_const_conc 	no
conc			30
hydrolysis 		no
khydro			1000000
freeze_temp		1.0
alpha			5.0
""")


EDGE_PDB = (
    "REMARK edge cases of the ATOM parser\n"
    "HETATM    1  O   HOH W   1       1.000   2.000   3.000  1.00  0.00\n"
    + pdb_line(7, " CA", "A", 1, 1.25, -2.5, 3.125, 0.5, 9.75)
    + pdb_line(8, " CB", "B", 1234, -100.125, 200.5, -0.001, 1.0, 99.99, "           C  ")
    + "ATOM      9  CA  GLY C   3      -1.5     2.25    -3.0\n"
    + pdb_line(10, "CAX", "D", 2, 0.0, 0.0, 0.0)
    + "END\n"
)

XYZ = """3
comment line
C 1.5 -2.25 3.125
N\t0.000001 1e3 -7
O 10 20 30
"""


def main():
    os.makedirs(OUT, exist_ok=True)
    golden = {}
    qpath = os.path.join(OUT, "quirks.conf")
    with open(qpath, "w") as f:
        f.write(QUIRKS)
    golden["config"] = [{"kind": k, "name": n, "default": d, **ref_query(qpath, k, n, d)} for k, n, d in QUERIES]
    golden["config_overrides"] = OVERRIDES
    golden["config_with_overrides"] = [{"kind": k, "name": n, "default": d, **ref_query(qpath, k, n, d, OVERRIDES)}
                                       for k, n, d in QUERIES]
    run_dir = os.path.join(OUT, "run")
    write_run_files(run_dir)
    epath = os.path.join(OUT, "edge.pdb")
    with open(epath, "w") as f:
        f.write(EDGE_PDB)
    golden["pdb"] = {}
    for name in ("edge.pdb", "run/tube_xyz.pdb"):
        r = subprocess.run([os.path.join(REF, "ref_pdb"), os.path.join(OUT, name)], capture_output=True, text=True)
        golden["pdb"][name] = [l[2:] for l in r.stderr.splitlines() if l.startswith("@@")]
    # DCD: 5 atoms, 3 frames; the bytes except the 80-byte dated remark
    rng = np.random.default_rng(11)
    frames = rng.normal(0, 10, size=(3, 3, 5)).astype(np.float32)
    dpath = os.path.join(OUT, "_ref_out.dcd")
    text = "\n".join(" ".join("%.9g" % v for v in fr.ravel()) for fr in frames) + "\n"
    subprocess.run([os.path.join(REF, "ref_dcd"), dpath, "5", "10", "1", "200", "100"], input=text, text=True,
                   check=True, capture_output=True)
    raw = open(dpath, "rb").read()
    os.remove(dpath)
    golden["dcd"] = {"natoms": 5, "nframes_header": 10, "first": 1, "dt": 200.0, "freq": 100,
                     "frames": frames.tolist(), "hex": raw.hex(), "remark2_offset": 4 * 25 + 80}
    xpath = os.path.join(OUT, "sample.xyz")
    with open(xpath, "w") as f:
        f.write(XYZ)
    wpath = os.path.join(OUT, "_ref_out.xyz")
    r = subprocess.run([os.path.join(REF, "ref_xyz"), xpath, wpath], capture_output=True, text=True, check=True)
    golden["xyz"] = {"rows": [l[2:] for l in r.stderr.splitlines() if l.startswith("@@")],
                     "written": open(wpath).read()}
    os.remove(wpath)
    with open(os.path.join(ROOT, "tests", "golden", "host_golden.json"), "w") as f:
        json.dump(golden, f, indent=1)
    print("wrote", os.path.join("tests", "golden", "host_golden.json"))


if __name__ == "__main__":
    main()
