"""The configuration-file driver on the host (no GPU): the reader, the
object factory against its restatement (oracle/clami_build.py) bit for bit,
and the restatement pinned by what the geometry fixes.

Golden input: tests/golden/volvox_sim.cfg, the file the reference's
generator writes for the paper's swimmer (made by make_volvox_cfg.py).
"""
import math
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import clami_build as CB
from mcm_amd import MPCDError
from mcm_amd import sim as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "volvox_sim.cfg")


@pytest.fixture(scope="module")
def volvox():
    with S.Simulation(GOLDEN) as s:
        yield s.info(), s.objects(), s.model()


@pytest.fixture(scope="module")
def volvox_oracle():
    pi = CB.PI
    return CB.clami(8.0, [0.0, 0.0, 0.0], [-45.0, 0.0, 45.0], [7, 10, 7], [pi / 10, 0.0, pi / 10], 26, 0.0,
                    _golden_delays())


def _golden_delays():
    txt = open(GOLDEN).read()
    line = [l for l in txt.splitlines() if "start_delays" in l][0]
    return [float(x) for x in line.split("[")[1].split("]")[0].split(",")]


def test_library_exports_driver_symbols(mcm):
    import ctypes
    lib = ctypes.CDLL(mcm.LIB_PATH)
    syms = S.declared_symbols()
    assert len(syms) >= 12
    assert not [s for s in syms if not hasattr(lib, s)]
    assert os.access(S.EXE, os.X_OK), "the cilia executable is built beside the library"


def test_golden_file_objects(volvox):
    info, objs, m = volvox
    assert info["name"] == "volvox" and info["size"] == (100, 100, 100)
    assert info["tstep"] == 0.05 and info["desiredt"] == 225000.0 and info["globalt"] == 0.0
    # Simulation::md_objects: the sphere, 24 cilia, then the Clami (src/clami.cpp:52, 128)
    assert [o["type"] for o in objs] == ["Sphere"] + ["ConstantBeatCilia"] * 24 + ["Clami"]
    assert objs[0]["name"] == "volvox_sphere" and objs[1]["name"] == "volvox_cilia0"
    assert objs[24]["name"] == "volvox_cilia23" and objs[-1]["name"] == "volvox"
    assert not any(o["write_data"] for o in objs[:-1]) and objs[-1]["write_data"]
    assert info["nbeads"] == 643 + 24 * 3 * 26 and objs[-1]["nbeads"] == info["nbeads"]
    assert info["nattach"] == objs[-1]["nattach"] == len(m["attach"])
    assert np.all(m["vel4"][:, :3] == 0.0) and np.all(m["vel4"][:, 3] == CB.PARTMASS)
    delays = _golden_delays()
    for i, c in enumerate(m["cilia"]):
        assert c["first"] == 643 + 78 * i and c["rodlength"] == 26 and c["num_anker"] == 2
        assert c["constant_beat"] and c["midpnt"] == 2.0 and c["Tstroke"] == delays[i]
        assert (c["bending"], c["beatstrength"], c["m_speed"], c["min_bend"], c["max_bend"]) == (0.01, 0.17, 5.3, -0.7,
                                                                                                 0.7)
        assert (c["power_stroke_duration"], c["recovery_stroke_duration"], c["springconstant"]) == (75.0, 150.0,
                                                                                                   20000.0)
        assert np.all(c["newlength"] == 0.5) and np.all(c["oldlength"] == 0.5)
    sp = m["spheres"][0]
    assert sp["first"] == 0 and sp["npart"] == 643 and sp["radius"] == 8.0 and sp["springconstant"] == 20000.0
    assert m["att_k"] == 20000.0 and m["ststep"] == 0.001


def test_factory_matches_restatement_bit_for_bit(volvox, volvox_oracle):
    _, _, m = volvox
    o = volvox_oracle
    assert np.array_equal(m["pos3"], np.array(o["beads"]))
    sp = m["spheres"][0]
    assert sp["headbond"] == o["headbond"]
    for i in range(643):
        assert list(sp["nei"][sp["nei_start"][i]:sp["nei_start"][i + 1]]) == o["neighbours"][i]
    assert [(a, b) for a, b, _ in m["attach"]] == [(a, b) for a, b, _ in o["springs"]]
    assert [l for _, _, l in m["attach"]] == [l for _, _, l in o["springs"]]


# ---------------------------------------------------------------- pinning the restatement

@pytest.mark.parametrize("fill,nsurf", [(1, 42), (2, 162), (3, 642)])
def test_shell_mesh_counts_degrees_radius(fill, nsurf):
    c = (3.0, -2.0, 5.5)
    p, nl, headbond = CB.shell(fill, 4.0, c)
    assert len(p) == nsurf + 1 and p[0] == list(c) and nl[0] == []
    deg = np.array([len(n) for n in nl[1:]])
    assert (deg == 5).sum() == 12 and (deg == 6).sum() == nsurf - 12  # a geodesic sphere
    edges = deg.sum() // 2
    assert nsurf - edges + 2 * edges // 3 == 2  # Euler: V - E + F with F = 2E/3
    for i in range(1, nsurf + 1):
        assert nl[i] == sorted(nl[i]) and all(i in nl[j] for j in nl[i])
    r = np.linalg.norm(np.array(p[1:]) - np.array(c), axis=1)
    assert np.abs(r - 4.0).max() < 1e-13
    lens = [np.linalg.norm(np.subtract(p[i], p[j])) for i in range(1, nsurf + 1) for j in nl[i]]
    assert abs(headbond - np.mean(lens)) < 1e-12


def test_shell_refinement_order():
    """Each level appends the midpoints of the previous mesh's edges (j, k),
    j < k, in increasing (j, k): a bead new at the last level has exactly two
    neighbours from the coarser mesh, lies on the great circle between them,
    and the parent pairs increase.  Coarser shells are prefixes."""
    shells = {f: CB.shell(f, 1.0, (0.0, 0.0, 0.0)) for f in (1, 2, 3)}
    for f in (2, 3):
        assert shells[f][0][:len(shells[f - 1][0])] == shells[f - 1][0]
    for f, ncoarse in ((1, 12), (2, 42), (3, 162)):
        p, nl, _ = shells[f]
        u = np.array(p)
        pairs = []
        for m in range(ncoarse + 1, len(p)):
            par = [j for j in nl[m] if j <= ncoarse]
            assert len(par) == 2, (f, m, par)
            j, k = par
            assert abs(np.linalg.det(np.array([u[m], u[j], u[k]]))) < 1e-12
            ang = lambda a, b: math.acos(min(1.0, float(np.dot(u[a], u[b]))))
            assert abs(ang(j, m) + ang(m, k) - ang(j, k)) < 1e-12
            pairs.append((j, k))
        assert pairs == sorted(pairs) and len(set(pairs)) == len(pairs)


def test_shell_icosahedron_vertices():
    """Beads 1-12: the poles at +x and -x after the quarter turn about y,
    upper ring (beads 2-6) and lower ring (7-11) of a pentagonal antiprism at
    azimuths (about x, from -z towards +y) 72, 144, 216, 288, 0 and 36, 108,
    180, 252, 324 degrees."""
    p, _, _ = CB.shell(0, 1.0, (0.0, 0.0, 0.0))
    u = np.array(p[1:])
    assert np.allclose(u[0], [1, 0, 0], atol=1e-15) and np.allclose(u[11], [-1, 0, 0], atol=1e-15)
    az = np.degrees(np.arctan2(u[1:11, 1], -u[1:11, 2])) % 360.0
    assert np.allclose(az, [72, 144, 216, 288, 0, 36, 108, 180, 252, 324], atol=1e-9)
    polar = np.degrees(np.arccos(u[1:11, 0]))
    assert np.allclose(polar[:5], polar[0]) and np.allclose(polar[5:], 180.0 - polar[0])


@pytest.mark.parametrize("angles", [(0.0, 0.0, 0.0), (0.3, -0.7, 1.1), (2.5, 0.9, -0.4)])
def test_cilium_geometry(angles):
    tx, ty, tz = angles
    R, w = 7, np.array([1.0, 2.0, -3.0])
    p = np.array(CB.cilium(R, list(w), tx, ty, tz)).reshape(3, R, 3)
    bl = CB.BONDLENGTH
    h = bl * math.sqrt(3.0) / 2.0
    # the rods of an equilateral triangle of side bl centred on the axis, along z, turned by
    # Rx(thetaX) Ry(thetaY) Rz(thetaZ) (scipy's intrinsic XYZ) and moved to w
    tri = np.array([[-2 * h / 3, 0.0], [h / 3, bl / 2], [h / 3, -bl / 2]])
    body = np.array([[[tri[r, 0], tri[r, 1], bl * i] for i in range(R)] for r in range(3)])
    rot = Rotation.from_euler("XYZ", [tx, ty, tz]).as_matrix()
    assert np.abs(p - (body @ rot.T + w)).max() < 1e-14


def test_cilium_placement_and_attachment(volvox_oracle):
    """Cilium c of ring i: its axis through the shell's centre at latitude
    y_theta[i] (asin of the axis' x component) and azimuth theta_offset[i] +
    2 pi j / n[i] about x, its base radius - 3 bondlengths from the centre;
    rod beads 0 and 3 are held by springs to their closest shell bead (the
    first of equal distance, brute force) and that bead's neighbours but the
    last, at their present lengths."""
    o = volvox_oracle
    beads = np.array(o["beads"])
    shell = beads[:643]
    c = shell[0]
    k = 0
    for i, (yt, n, off) in enumerate(zip([-45.0, 0.0, 45.0], [7, 10, 7], [CB.PI / 10, 0.0, CB.PI / 10])):
        for j in range(n):
            first = 643 + 78 * k
            axis = beads[first + 1:first + 26].mean(0) - beads[first:first + 25].mean(0)
            rods = beads[first:first + 78].reshape(3, 26, 3)
            d = rods.mean(0)[1] - rods.mean(0)[0]
            d /= np.linalg.norm(d)
            assert abs(math.degrees(math.asin(d[0])) - yt) < 1e-9
            az = (math.atan2(-d[1], d[2]) - (off + 2 * math.pi * j / n)) % (2 * math.pi)
            assert min(az, 2 * math.pi - az) < 1e-9
            base = rods.mean(0)[0]
            assert abs(np.linalg.norm(base - c) - (8.0 - 1.5)) < 1e-12
            assert np.allclose(np.cross(base - c, d), 0.0, atol=1e-12)
            del axis
            k += 1
    springs = o["springs"]
    expect = []
    for k in range(24):
        first = 643 + 78 * k
        for wo in (0, 3):
            for rod in range(3):
                b = first + 26 * rod + wo
                dist = ((shell - beads[b]) ** 2).sum(1)
                idx = int(np.argmin(dist))  # first of equal distances
                for s in [idx] + o["neighbours"][idx][:-1]:
                    expect.append((s, b))
    assert [(a, b) for a, b, _ in springs] == expect
    for a, b, l0 in springs:
        assert abs(l0 - np.linalg.norm(beads[a] - beads[b])) < 1e-14


# ---------------------------------------------------------------- reader and factory errors

BASE = """name = "t"; globalt = 0.0; desiredt = 1.0; omega = 0.0; sizex = 16; sizey = 16; sizez = 16;
runtime = 100.0; tstep = 0.05;
"""
CILIUM = """{ type = "ConstantBeatCilia"; name = "c"; rodlength = 6; beatstrength = 0.17; bending = 0.01;
  min_bend = -0.7; max_bend = 0.7; m_speed = 5.3; springconstant = 20000.0; position = [8.0, 8.0, 4.0];
  thetaX = 0.0; thetaY = 0.0; thetaZ = 0.5; write_data = true; power_stroke_duration = 75.0;
  recovery_stroke_duration = 150.0; start_delay = 12.5; }"""


def _load(tmp_path, text):
    f = tmp_path / "sim.cfg"
    f.write_text(text)
    return S.Simulation(str(f))


def test_standalone_objects_and_skipped_entries(tmp_path):
    text = BASE + """md_objects = (
  { type = "Sphere"; name = "ball"; radius = 3.0; springconstant = 1000.0; position = [8.0, 8.0, 8.0];
    npart = 163; },
  { name = "no_type"; },
  """ + CILIUM + """ );  // a comment
/* block
   comment */ # and another
"""
    with _load(tmp_path, text) as s:
        objs = s.objects()
        m = s.model()
    assert [o["type"] for o in objs] == ["Sphere", "ConstantBeatCilia"]
    assert objs[0]["nbeads"] == 163 and objs[0]["write_data"] and objs[1]["write_data"]
    p, _, hb = CB.shell(2, 3.0, (8.0, 8.0, 8.0))
    assert np.array_equal(m["pos3"][:163], np.array(p)) and m["spheres"][0]["headbond"] == hb
    assert np.array_equal(m["pos3"][163:], np.array(CB.cilium(6, [8.0, 8.0, 4.0], 0.0, 0.0, 0.5)))
    assert m["cilia"][0]["Tstroke"] == 12.5 and m["attach"] == []


@pytest.mark.parametrize("body,msg", [
    ("md_objects = ( { type = \"Blob\"; name = \"x\"; } );", "Unkown type:Blob"),
    ("md_objects = ( { type = \"Sphere\"; name = \"b\"; radius = 3.0; springconstant = 1.0; "
     "position = [1.0, 1.0, 1.0]; npart = 100; } );", "Wrong nparts"),
    ("md_objects = ( { type = \"Sphere\"; name = \"b\"; radius = 3.0; springconstant = 1.0; npart = 163; } );",
     r"No md_objects\.\[0\]\.position setting"),
    ("md_objects = ( { type = \"Clami\"; name = \"v\"; radius = 4.0; position = [8.0, 8.0, 8.0]; "
     "y_thetas = [0.0]; num_equators = [3, 2]; theta_offset = [0.0]; cilia = " + CILIUM + "; } );",
     "differ in length"),
    ("md_objects = ( { type = \"Clami\"; name = \"v\"; radius = 4.0; position = [8.0, 8.0, 8.0]; "
     "y_thetas = [0.0]; num_equators = [3]; theta_offset = [0.0]; start_delays = [1.0]; cilia = " + CILIUM +
     "; } );", "one delay per cilium"),
    ("md_objects = ( { type = \"Clami\"; name = \"v\"; radius = -4.0; position = [8.0, 8.0, 8.0]; "
     "y_thetas = [0.0]; num_equators = [3]; theta_offset = [0.0]; cilia = " + CILIUM + "; } );",
     "radius must be positive"),
    ("md_objects = ( { type = \"Clami\"; name = \"v\"; radius = 4.0; position = [8.0, 8.0, 8.0]; "
     "y_thetas = [0.0]; num_equators = [3]; theta_offset = [0.0]; cilia = " + CILIUM.replace("rodlength = 6",
                                                                                              "rodlength = 3") +
     "; } );", "rodlength must exceed 3"),
    ("md_objects = ( " + CILIUM.replace("rodlength = 6", "rodlength = 6.0") + " );", "an integer was expected"),
    ("md_objects = ( { type = = \"Sphere\"; name = \"b\"; } );", "parse error at"),
    ("md_objects = ( { type = \"Sphere\"; name = \"b\"; ", "parse error at"),
])
def test_errors_name_their_setting(tmp_path, body, msg):
    with pytest.raises(MPCDError, match=msg):
        _load(tmp_path, BASE + body)


def test_missing_root_setting_and_file(tmp_path):
    with pytest.raises(MPCDError, match="No tstep setting"):
        _load(tmp_path, BASE.replace("tstep = 0.05;", "") + "md_objects = ();")
    with pytest.raises(MPCDError, match="I/O error"):
        S.Simulation(str(tmp_path / "absent.cfg"))


def test_parse_error_line_number(tmp_path):
    with pytest.raises(MPCDError, match=r"sim\.cfg:3 "):
        _load(tmp_path, BASE + "md_objects = ( { name = \"b\"; type = ; } );")


def test_two_swimmers_in_one_file(tmp_path):
    """Two Clamis in one file: one bead model, each swimmer's beads and
    springs in its own range; different attachment springconstants are
    refused (one value serves every attachment spring)."""
    clami = lambda name, x, k: (
        f'{{ type = "Clami"; name = "{name}"; radius = 4.0; springconstant = {k}; position = [{x}, 8.0, 8.0]; '
        f'y_thetas = [0.0]; num_equators = [3]; theta_offset = [0.0]; cilia = ' + CILIUM + '; }')
    with _load(tmp_path, BASE + "md_objects = ( " + clami("a", 4.0, 20000.0) + ", " + clami("b", 12.0, 20000.0)
               + " );") as s:
        objs = s.objects()
        m = s.model()
    cl = [o for o in objs if o["type"] == "Clami"]
    assert [o["name"] for o in cl] == ["a", "b"]
    assert cl[1]["first"] == cl[0]["first"] + cl[0]["nbeads"] and cl[0]["nbeads"] == 643 + 3 * 3 * 6
    assert cl[1]["first_attach"] == cl[0]["nattach"] and len(m["attach"]) == cl[0]["nattach"] + cl[1]["nattach"]
    for o in cl:
        for a, b, _ in m["attach"][o["first_attach"]:o["first_attach"] + o["nattach"]]:
            assert o["first"] <= a < o["first"] + 643 and o["first"] + 643 <= b < o["first"] + o["nbeads"]
    assert np.allclose(m["pos3"][cl[1]["first"]] - m["pos3"][cl[0]["first"]], [8.0, 0.0, 0.0])
    with pytest.raises(MPCDError, match="share its springconstant"):
        _load(tmp_path, BASE + "md_objects = ( " + clami("a", 4.0, 20000.0) + ", " + clami("b", 12.0, 1000.0) + " );")


def test_cilia_executable_usage():
    import subprocess
    for args in ([], ["--bogus", "x.cfg"], ["--steps"], ["--interval", "0", "x.cfg"]):
        r = subprocess.run([S.EXE, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([S.EXE, "/nonexistent/sim.cfg"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "I/O error" in r.stderr
