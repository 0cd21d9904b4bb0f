"""The configuration-file driver on the GPU.

* Its steps are mpcd_step on a context holding the factory's bead model:
  bit-identical to a context set up by hand with the same seed, fluid and
  swimmer (beads exactly; solvent diagnostics to 1e-12).
* Its outputs are the reference's: <name>.san (src/clami.cpp:262-276: globalt,
  centre of mass, mean velocity, shell axes to beads 1, 32 and 137,
  dissipation per time) and <name>.xyz (src/clami.cpp:208-258: every other
  cilium bead, H for rod 0, O for rods 1-2; src/libcilia/mdobject.cpp:83-107
  for objects that write data), recomputed here from the bead state.
* A run that stops and restarts from its checkpoint writes the same lines
  as one that does not stop; the `cilia` executable writes what the API
  does; configurations run together as one ensemble write what each writes
  alone.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from mcm_amd import sim as S

pytestmark = pytest.mark.gpu

CFG = """name = "{name}";
globalt = 0.0;
desiredt = 1000.0;
omega = 0.0;
sizex = {L[0]};
sizey = {L[1]};
sizez = {L[2]};
runtime = {runtime};
tstep = 0.05;
md_objects = ( {{
  name = "{name}";
  type = "Clami";
  radius = {radius};
  springconstant = 20000.0;
  position = [ {c[0]}, {c[1]}, {c[2]} ];
  y_thetas = [ -30.0, 30.0 ];
  num_equators = [ 3, 3 ];
  theta_offset = [ 0.0, 0.5 ];
  start_delays = [ 0.0, 20.0, 40.0, 60.0, 80.0, 100.0 ];
  cilia = {{
    type = "ConstantBeatCilia"; name = "c"; rodlength = {R}; beatstrength = 0.17; bending = 0.01;
    min_bend = -0.7; max_bend = 0.7; m_speed = 5.3; springconstant = 20000.0;
    position = [ 0.0, 0.0, 0.0 ]; thetaX = 0.0; thetaY = 0.0; thetaZ = 0.3; write_data = {wd};
    power_stroke_duration = 75.0; recovery_stroke_duration = 150.0; start_delay = 0.0;
  }};
}} );
"""


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _write_cfg(d, name="swim", L=(24, 24, 24), radius=4.0, R=8, wd="false", runtime=3600.0):
    os.makedirs(d, exist_ok=True)
    c = tuple(x / 2.0 for x in L)
    path = os.path.join(str(d), "sim.cfg")
    with open(path, "w") as f:
        f.write(CFG.format(name=name, L=L, radius=radius, c=c, R=R, wd=wd, runtime=runtime))
    return path


def _diag(mcm, ctx):
    d = mcm.Diag()
    assert mcm.lib().mpcd_diagnostics(C.c_void_p(ctx), C.byref(d)) == 0
    return d


def test_driver_steps_are_mpcd_steps(mcm, tmp_path):
    _print_case("driver_vs_context")
    path = _write_cfg(tmp_path)
    with S.Simulation(path) as s:
        model = s.model()
        s.start(seed=11, measureinterval=5, restart=False)
        s.next_tstep(6)
        d1 = _diag(mcm, s.context())
        with mcm.MPCD((24, 24, 24), precision=mcm.FP64, seed=11) as m:
            m.create_fluid(10 * 24 ** 3)
            m.md_attach(model, globalt=model["globalt"], ststep=model["ststep"])
            for _ in range(6):
                m.step(0.05)
            ref = m.md_fetch()
            d2 = m.diagnostics()
        got = np.zeros((model["pos3"].shape[0], 3))
        gv = np.zeros((model["pos3"].shape[0], 4))
        g = C.c_double()
        assert mcm.lib().mpcd_md_fetch(C.c_void_p(s.context()), got.ctypes.data_as(C.c_void_p),
                                       gv.ctypes.data_as(C.c_void_p), None, None, None, None, C.byref(g)) == 0
    assert np.array_equal(got, ref["pos"]) and np.array_equal(gv, ref["vel"]) and g.value == ref["globalt"]
    assert d1.flags == 0 and d2["flags"] == 0
    np.testing.assert_allclose(list(d1.momentum), d2["momentum"], rtol=0, atol=1e-9)
    assert abs(d1.twice_kinetic - d2["twice_kinetic"]) <= 1e-12 * d2["twice_kinetic"]


def _read_xyz(path):
    frames, lines = [], open(path).read().splitlines()
    i = 0
    while i < len(lines):
        n = int(lines[i])
        t = float(lines[i + 1].split()[1])
        rows = [l.split() for l in lines[i + 2:i + 2 + n]]
        frames.append((t, [r[0] for r in rows], np.array([[float(x) for x in r[1:]] for r in rows])))
        i += 2 + n
    return frames


def test_run_writes_the_reference_outputs(mcm, tmp_path):
    _print_case("driver_outputs")
    path = _write_cfg(tmp_path, wd="true")
    with S.Simulation(path) as s:
        model = s.model()
        s.start(seed=5, measureinterval=4, max_steps=8, restart=False)
        s.run()
        nb = model["pos3"].shape[0]
        pos, vel, g = np.zeros((nb, 3)), np.zeros((nb, 4)), C.c_double()
        assert mcm.lib().mpcd_md_fetch(C.c_void_p(s.context()), pos.ctypes.data_as(C.c_void_p),
                                       vel.ctypes.data_as(C.c_void_p), None, None, None, None, C.byref(g)) == 0
        info = s.info()
    assert info["steps"] == 8
    san = np.loadtxt(tmp_path / "swim.san")
    assert san.shape == (2, 17)
    R = 8
    expect = np.concatenate([[g.value], pos.mean(0), vel[:, :3].mean(0), pos[1] - pos[0], pos[32] - pos[0],
                             pos[137] - pos[0]])
    np.testing.assert_allclose(san[-1, :16], expect, rtol=6e-6, atol=1e-12)  # 6 significant digits
    assert abs(san[-1, 0] - 8 * 0.05) < 1e-9 and abs(san[0, 0] - 4 * 0.05) < 1e-9
    assert np.all(san[:, 16] > 0.0)  # the thermostat's dissipation per unit time
    frames = _read_xyz(tmp_path / "swim.xyz")
    assert len(frames) == 2
    t, names, xyz = frames[-1]
    assert len(names) == 6 * 3 * (R // 2) and names[:R // 2] == ["H"] * (R // 2) and names[R // 2] == "O"
    sel = [c["first"] + rod * R + j for c in model["cilia"] for rod in range(3) for j in range(R) if (j + 1) % 2 == 0]
    np.testing.assert_allclose(xyz, pos[sel], rtol=6e-6, atol=1e-12)
    cil = _read_xyz(tmp_path / "swim_cilia0.xyz")  # a cilium that writes data: all its beads
    assert len(cil) == 2 and cil[-1][2].shape == (3 * R, 3) and set(cil[-1][1]) == {"C"}
    np.testing.assert_allclose(cil[-1][2], pos[model["cilia"][0]["first"]:model["cilia"][0]["first"] + 3 * R],
                               rtol=6e-6, atol=1e-12)
    # the end of the run: checkpoint, parameter file, the file's globalt written back
    assert (tmp_path / "swim.ckpt").exists() and (tmp_path / "parameter").exists()
    with S.Simulation(path) as s2:
        assert s2.info()["globalt"] == g.value


def test_restart_writes_the_same_lines(mcm, tmp_path):
    _print_case("driver_restart")
    a = _write_cfg(tmp_path / "a")
    b = _write_cfg(tmp_path / "b")
    with S.Simulation(a) as s:
        s.start(seed=8, measureinterval=3, max_steps=9, restart=False)
        s.run()
    with S.Simulation(b) as s:
        s.start(seed=8, measureinterval=3, max_steps=6, restart=False)
        s.run()
    with S.Simulation(b) as s:  # picks up b/swim.ckpt
        s.start(seed=8, measureinterval=3, max_steps=3, restart=True)
        s.run()
    la = open(tmp_path / "a" / "swim.san").read().splitlines()
    lb = open(tmp_path / "b" / "swim.san").read().splitlines()
    assert len(la) == 3 and la == lb
    assert open(tmp_path / "a" / "swim.xyz").read() == open(tmp_path / "b" / "swim.xyz").read()


def test_cilia_executable_and_ensemble(mcm, tmp_path):
    _print_case("driver_executable_ensemble")
    cfgs = {"one": dict(L=(24, 24, 24), radius=4.0, R=8), "two": dict(L=(20, 28, 24), radius=3.5, R=6)}
    alone, together = {}, {}
    for k, kw in cfgs.items():  # each alone through the API, seed 21 + index as the executable gives
        alone[k] = _write_cfg(tmp_path / "alone" / k, name=k, **kw)
        together[k] = _write_cfg(tmp_path / "ens" / k, name=k, **kw)
    for i, k in enumerate(cfgs):
        with S.Simulation(alone[k]) as s:
            s.start(seed=21 + i, measureinterval=4, max_steps=8, restart=False)
            s.run()
    r = subprocess.run([S.EXE, "--steps", "8", "--interval", "4", "--seed", "21", "--fresh", "--quiet",
                        together["one"], together["two"]], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    for k in cfgs:
        for ext in (".san", ".xyz"):
            x = open(os.path.join(os.path.dirname(alone[k]), k + ext)).read()
            y = open(os.path.join(os.path.dirname(together[k]), k + ext)).read()
            assert x == y, f"{k}{ext} differs between the ensemble and the run alone"


def test_fp32_bulk_driver_run(mcm, tmp_path):
    _print_case("driver_fp32")
    path = _write_cfg(tmp_path)
    with S.Simulation(path) as s:
        s.start(precision=mcm.FP32BULK, seed=2, measureinterval=5, max_steps=10, restart=False)
        s.run()
        d = _diag(mcm, s.context())
        mcm.lib().mpcd_count.restype = C.c_longlong
        n = mcm.lib().mpcd_count(C.c_void_p(s.context()))
    assert d.flags == 0 and n == 10 * 24 ** 3
    san = np.loadtxt(tmp_path / "swim.san")
    assert san.shape == (2, 17) and np.all(np.isfinite(san))


def test_refusals_before_start(tmp_path):
    path = _write_cfg(tmp_path)
    with S.Simulation(path) as s:
        with pytest.raises(Exception, match="not started"):
            s.next_tstep(1)


def test_debug_line_and_swimmer_statistics(mcm, tmp_path):
    """mpcd.san at the start holds MPCD::debug's totals over the solvent and
    the beads; the .san series feeds the swimmer statistics."""
    from mcm_amd import analysis as A
    _print_case("driver_debug_analysis")
    path = _write_cfg(tmp_path)
    with S.Simulation(path) as s:
        nb = s.info()["nbeads"]
        s.start(seed=4, measureinterval=2, max_steps=96, restart=False)
        s.run()
    line = np.loadtxt(tmp_path / "mpcd.san")
    n = 10 * 24 ** 3
    assert line[0] == 0.0 and line[2] == n + nb and line[3] == n + 5 * nb
    assert np.abs(line[4:7]).max() < 1e-6  # the fluid's mean velocity removed, the beads at rest
    assert abs(line[7] * (n + nb) / n - 1.0) < 0.01  # kT of the solvent, diluted by the resting beads
    r, nn, v = A.analyze(A.load_san(tmp_path / "swim.san"), eq_time=0.0, k=4)
    assert len(nn) == 48 and np.allclose(np.linalg.norm(nn, axis=1), 1.0)
    assert all(np.isfinite(x) for x in r.values())
    assert r["power_mean"] > 0.0 and r["phi_max"] < 0.5 and abs(r["v_normal_m"]) < 0.05


def _md_state(mcm, ctx, nb, ncil, nlen):
    pos, vel = np.zeros((nb, 3)), np.zeros((nb, 4))
    beat, nl, ol, g = np.zeros((ncil, 2)), np.zeros(nlen), np.zeros(nlen), C.c_double()
    assert mcm.lib().mpcd_md_fetch(C.c_void_p(ctx), pos.ctypes.data_as(C.c_void_p), vel.ctypes.data_as(C.c_void_p),
                                   None, beat.ctypes.data_as(C.c_void_p), nl.ctypes.data_as(C.c_void_p),
                                   ol.ctypes.data_as(C.c_void_p), C.byref(g)) == 0
    return pos, vel, beat, ol, g.value


def _fluid_of(mcm, ctx):
    n = mcm.lib().mpcd_count(C.c_void_p(ctx))
    p, v = np.zeros((n, 3)), np.zeros((n, 4))
    assert mcm.lib().mpcd_download_fluid(C.c_void_p(ctx), p.ctypes.data_as(C.c_void_p),
                                         v.ctypes.data_as(C.c_void_p), C.c_longlong(n)) == 0
    order = np.lexsort(p.T[::-1])
    return p[order], v[order]


@pytest.mark.parametrize("prec_name", ["FP64", "FP32BULK"])
def test_reference_restart_files_round_trip(mcm, tmp_path, prec_name):
    """A run that saves the reference's restart files (fluid_n.dat, mpcd.dat,
    <name>.dat of the sphere, each cilium and the Clami's springs) and a new
    run that starts from them alone (no checkpoint): the shift, the solvent
    (as a set), the beads, rest lengths, beat state and globalt are those
    the first run saved, digit for digit (FP32-bulk positions, stored again
    as cell offsets, to an FP32 rounding)."""
    import shutil
    _print_case("reference_restart_files", prec=prec_name)
    prec = getattr(mcm, prec_name)
    a_dir, b_dir = tmp_path / "a", tmp_path / "b"
    path = _write_cfg(a_dir)
    with S.Simulation(path) as s:
        info = s.info()
        model = s.model()
        s.start(precision=prec, seed=3, measureinterval=4, max_steps=8, restart=False, ref_files=True)
        s.run()
        ctx = s.context()
        nb, ncil, nlen = info["nbeads"], info["ncilia"], sum(c["rodlength"] for c in model["cilia"])
        ref_md = _md_state(mcm, ctx, nb, ncil, nlen)
        ref_fluid = _fluid_of(mcm, ctx)
        a = (C.c_double * 3)()
        assert mcm.lib().mpcd_get_shift(C.c_void_p(ctx), C.cast(a, C.c_void_p)) == 0
        ref_shift = list(a)
    names = {"fluid_n.dat", "mpcd.dat", "swim.dat", "swim_sphere.dat"} | {f"swim_cilia{i}.dat" for i in range(6)}
    assert names <= set(os.listdir(a_dir))
    shutil.copytree(a_dir, b_dir)
    os.remove(b_dir / "swim.ckpt")
    with S.Simulation(str(b_dir / "sim.cfg")) as s:
        assert s.info()["globalt"] == ref_md[4]  # the file's globalt, written back at the end of the run
        s.start(precision=prec, seed=99, measureinterval=4, max_steps=4, restart=True)
        ctx = s.context()
        got = _md_state(mcm, ctx, nb, ncil, nlen)
        for x, y in zip(got[:4], ref_md[:4]):
            assert np.array_equal(x, y)
        assert got[4] == ref_md[4]
        gp, gv = _fluid_of(mcm, ctx)
        assert np.array_equal(gv, ref_fluid[1])
        if prec_name == "FP64":
            assert np.array_equal(gp, ref_fluid[0])
        else:  # positions re-encoded as cell + FP32 offset: a last-ulp move of an offset at a cell face at most
            assert np.abs(gp - ref_fluid[0]).max() < 1e-6
        a = (C.c_double * 3)()
        assert mcm.lib().mpcd_get_shift(C.c_void_p(ctx), C.cast(a, C.c_void_p)) == 0
        assert list(a) == ref_shift
        s.run()  # and it continues
        assert _diag(mcm, ctx).flags == 0
    if prec_name == "FP64":
        # a sphere file whose neighbour table is not the configuration's is
        # refused (Sphere::load would spring the beads by the file's table)
        c_dir = tmp_path / "c"
        shutil.copytree(a_dir, c_dir)
        os.remove(c_dir / "swim.ckpt")
        lines = open(c_dir / "swim_sphere.dat").read().split("\n")
        k = lines.index("#neilist") + 5
        row = lines[k].split()
        row[1], row[2] = row[2], row[1]
        lines[k] = " ".join(row)
        open(c_dir / "swim_sphere.dat", "w").write("\n".join(lines))
        with S.Simulation(str(c_dir / "sim.cfg")) as s:
            with pytest.raises(mcm.MPCDError, match="neilist row 4"):
                s.start(precision=prec, seed=99, measureinterval=4, max_steps=4, restart=True)


def test_runtime_limit_stops_after_a_block(mcm, tmp_path):
    """runtime = 0 wall seconds: the run stops after its first block of
    outputs, saves, and writes the file's globalt back (the reference's
    break at src/simulation.cpp:124-127 and its end-of-run save)."""
    _print_case("driver_runtime_limit")
    path = _write_cfg(tmp_path, runtime=0.0)
    with S.Simulation(path) as s:
        s.start(seed=6, measureinterval=3, restart=False)
        s.run()
        info = s.info()
    assert info["steps"] == 3
    assert len(open(tmp_path / "swim.san").read().splitlines()) == 1
    assert (tmp_path / "swim.ckpt").exists()
    with S.Simulation(path) as s2:
        assert abs(s2.info()["globalt"] - 3 * 0.05) < 1e-9
