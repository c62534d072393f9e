"""The C++ drivers (fuller_LJ_npt_md, fuller_airebo_npt_md, fuller_LJ_npt_mmmd)
end to end.

* The reference's smoke script (src/Test_fuller.sh:197-279) runs each program
  with --step=200 --mon=100, plus the OVITO and restart variants, and passes
  on exit 0 with "Done" in the output; the same invocations here.
* Output-level parity: the reference programs (built from their sources in
  oracle/_ref, where available) and these drivers started from identical
  states -- the same seeded initial state, or the same restart file in either
  direction -- print the same monitor rows to their printed precision.
* Restart round trip, stop control (abort.md / stop.md), COLD/WARM phases.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fuller_md_en_amd", "bin")
LIB = os.path.join(ROOT, "fuller_md_en_amd", "FullereneLib")
REF = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.gpu


def run(exe, args, cwd, timeout=300):
    print(f"run {os.path.basename(exe)} {' '.join(args)}", flush=True)
    r = subprocess.run([exe, *args, f"--libdir={LIB}"], cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def rows(out):
    """Monitor rows: (step, phase, numbers...)."""
    res = []
    for line in out.splitlines():
        m = re.match(r"\s+(\d+)\s+(COLD|WARM|PROD)\s+(.*)$", line)
        if m:
            res.append((int(m.group(1)), m.group(2), [float(x) for x in m.group(3).split()]))
    return res


LJ = os.path.join(BIN, "fuller_LJ_npt_md")
AB = os.path.join(BIN, "fuller_airebo_npt_md")
MMD = os.path.join(BIN, "fuller_LJ_npt_mmmd")


@pytest.mark.parametrize("exe,extra", [
    (LJ, []), (LJ, ["--ovito=100"]), (LJ, ["--restart=100"]),
    (AB, []), (AB, ["--ovito=100"]), (AB, ["--restart=100"]),
    (MMD, []), (MMD, ["--ovito=100"]), (MMD, ["--restart=100"])])
def test_smoke_like_reference_script(tmp_path, exe, extra):
    rc, out = run(exe, ["--step=200", "--mon=100", *extra], tmp_path)
    assert rc == 0 and "Done" in out, out[-2000:]
    names = os.listdir(tmp_path)
    if "--ovito=100" in extra:
        xyz = [n for n in names if n.startswith("ovito_traj")]
        assert xyz
        lines = open(tmp_path / xyz[0]).read().splitlines()
        nat = int(lines[0])
        assert lines[1].startswith("Lattice=") and len(lines) == 2 * (nat + 2)
    if "--restart=100" in extra:
        rst = sorted(n for n in names if n.endswith(".rst"))
        assert len(rst) == 2, rst  # step 100 and the final step


def _compare(ours, theirs, cols_tol):
    assert len(ours) == len(theirs) and len(ours) > 0
    for (s1, p1, a), (s2, p2, b) in zip(ours, theirs):
        assert s1 == s2 and p1 == p2
        for k, tol in cols_tol.items():
            assert abs(a[k] - b[k]) <= tol, (s1, k, a[k], b[k])


needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "fuller_LJ_npt_md_serial")),
                               reason="reference build (oracle/_ref) not present")

# printed columns: T %7.1f, P %9.3f, a %8.3f, Ecoh %10.5f  (one unit in the last digit)
LJ_TOL = {0: 0.11, 1: 0.0011, 2: 0.0011, 3: 1.1e-5}


@needs_ref
def test_lj_driver_rows_equal_reference(tmp_path):
    args = ["--cell=2", "--temp=300", "--step=60", "--mon=10", "--seed=7"]
    rc1, ours = run(LJ, args, tmp_path)
    rc2, theirs = run(os.path.join(REF, "fuller_LJ_npt_md_serial"), args, tmp_path)
    assert rc1 == 0 and rc2 == 0
    _compare(rows(ours), rows(theirs), LJ_TOL)


@needs_ref
def test_lj_cold_warm_phases_equal_reference(tmp_path):
    args = ["--cell=2", "--temp=300", "--coldstart=20", "--warmup=20", "--step=40",
            "--mon=10", "--warmup_mon=freq"]
    rc1, ours = run(LJ, args, tmp_path)
    rc2, theirs = run(os.path.join(REF, "fuller_LJ_npt_md_serial"), args, tmp_path)
    assert rc1 == 0 and rc2 == 0
    r1, r2 = rows(ours), rows(theirs)
    assert {p for _, p, _ in r1} == {"COLD", "WARM", "PROD"}
    _compare(r1, r2, LJ_TOL)


@needs_ref
def test_restart_files_cross_load(tmp_path):
    """Our restart resumed by the reference and theirs by us: same next rows."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    base = ["--cell=2", "--temp=300", "--step=40", "--mon=10", "--restart=20", "--seed=11"]
    ref_exe = os.path.join(REF, "fuller_LJ_npt_md_serial")
    assert run(LJ, base, a)[0] == 0
    assert run(ref_exe, base, b)[0] == 0
    ours = sorted(p for p in os.listdir(a) if p.endswith("_20.rst"))[0]
    theirs = sorted(p for p in os.listdir(b) if p.endswith("_20.rst"))[0]
    # the two step-20 states are the same run to within FP32 force rounding
    r_ours = [l for l in open(a / ours) if l.startswith("MOL")]
    r_theirs = [l for l in open(b / theirs) if l.startswith("MOL")]
    va = np.array([[float(x) for x in l.split()[2:]] for l in r_ours])
    vb = np.array([[float(x) for x in l.split()[2:]] for l in r_theirs])
    assert np.abs(va[:, :3] - vb[:, :3]).max() < 1e-6  # positions, A
    cont = ["--cell=2", "--temp=300", "--step=40", "--mon=10", "--seed=11"]
    _, o1 = run(ref_exe, [*cont, f"--resfile={a / ours}"], tmp_path)
    _, o2 = run(LJ, [*cont, f"--resfile={b / theirs}"], tmp_path)
    _, o3 = run(LJ, [*cont, f"--resfile={a / ours}"], tmp_path)
    assert "Restarting from global step 20" in o1 and "Restarting from global step 20" in o2
    _compare(rows(o3), rows(o1), LJ_TOL)  # same file, two programs
    _compare(rows(o2), rows(o1), LJ_TOL)


def test_restart_round_trip_continues_the_run(tmp_path):
    base = ["--cell=2", "--temp=300", "--step=40", "--mon=10", "--restart=20", "--seed=5"]
    rc, full = run(LJ, base, tmp_path)
    assert rc == 0
    rst = sorted(p for p in os.listdir(tmp_path) if p.endswith("_20.rst"))[0]
    rc, resumed = run(LJ, ["--cell=2", "--temp=300", "--step=40", "--mon=10", "--seed=5",
                           f"--resfile={tmp_path / rst}"], tmp_path)
    assert rc == 0
    r_full = [r for r in rows(full) if r[0] > 20]
    _compare(rows(resumed), r_full, LJ_TOL)


def test_stop_control(tmp_path):
    os.mkdir(tmp_path / "abort.md")
    rc, out = run(LJ, ["--cell=2", "--step=100", "--mon=10", "--restart=50"], tmp_path)
    assert rc == 0 and "abort.md detected at step 10" in out and "Done" in out
    assert any(p.endswith("_010.rst") for p in os.listdir(tmp_path))  # 3 digits of 100
    shutil.rmtree(tmp_path / "abort.md")
    os.mkdir(tmp_path / "stop.md")
    rc, out = run(LJ, ["--cell=2", "--step=100", "--mon=10", "--restart=50"], tmp_path)
    assert "stop.md detected at step 10" in out and "Stopped at restart checkpoint (step 50)" in out
    rc, out = run(AB, ["--cell=1", "--step=100", "--mon=10"], tmp_path)
    assert "stop.md detected at step 10" in out and "Done" in out


AB_TOL = {0: 0.11, 1: 0.0011, 2: 0.0011, 3: 1.1e-4, 4: 1.1e-4, 5: 1.1e-4}


@needs_ref
def test_airebo_driver_rows_equal_reference(tmp_path):
    args = ["--cell=2", "--temp=300", "--step=20", "--mon=5", "--seed=3"]
    rc1, ours = run(AB, args, tmp_path)
    rc2, theirs = run(os.path.join(REF, "fuller_airebo_npt_md_serial"), args, tmp_path)
    assert rc1 == 0 and rc2 == 0
    _compare(rows(ours), rows(theirs), AB_TOL)


# printed columns: T %7.1f, P %9.3f, a %8.3f, five energies %9.4f
MM_TOL = {0: 0.11, 1: 0.0011, 2: 0.0011, 3: 1.1e-4, 4: 1.1e-4, 5: 1.1e-4, 6: 1.1e-4,
          7: 1.1e-4}
MM_REF = os.path.join(REF, "fuller_LJ_npt_mmmd_serial")
needs_mm_ref = pytest.mark.skipif(not os.path.exists(MM_REF),
                                  reason="reference MM build (oracle/_ref) not present")


@needs_mm_ref
def test_mm_driver_rows_equal_reference(tmp_path):
    """The reference's (compat) bonded forces, 30 steps from the same seeded
    state: the same monitor rows to their printed precision."""
    args = ["--cell=2", "--temp=300", "--step=30", "--mon=5", "--seed=3"]
    rc1, ours = run(MMD, args, tmp_path)
    rc2, theirs = run(MM_REF, args, tmp_path)
    assert rc1 == 0 and rc2 == 0
    assert "Topology/mol: 90 bonds, 180 angles, 360 dihedrals, 60 impropers" in ours
    _compare(rows(ours), rows(theirs), MM_TOL)


@needs_mm_ref
def test_mm_restart_files_cross_load(tmp_path):
    """Our MM restart resumed by the reference, and the reference's by us."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    base = ["--cell=1", "--temp=300", "--step=20", "--mon=5", "--restart=10", "--seed=4",
            "--fullerene=C70", "--crystal=hcp"]
    assert run(MMD, base, a)[0] == 0
    assert run(MM_REF, base, b)[0] == 0
    ours = sorted(p for p in os.listdir(a) if p.endswith("_10.rst"))[0]
    theirs = sorted(p for p in os.listdir(b) if p.endswith("_10.rst"))[0]
    va = np.array([[float(x) for x in l.split()[2:]] for l in open(a / ours)
                   if l.startswith("ATOM")])
    vb = np.array([[float(x) for x in l.split()[2:]] for l in open(b / theirs)
                   if l.startswith("ATOM")])
    assert va.shape == vb.shape and np.abs(va[:, :3] - vb[:, :3]).max() < 1e-6
    cont = ["--cell=1", "--temp=300", "--step=20", "--mon=5", "--seed=4", "--fullerene=C70",
            "--crystal=hcp"]
    _, o1 = run(MM_REF, [*cont, f"--resfile={a / ours}"], tmp_path)
    _, o2 = run(MMD, [*cont, f"--resfile={b / theirs}"], tmp_path)
    _, o3 = run(MMD, [*cont, f"--resfile={a / ours}"], tmp_path)
    assert "Restarting from global step 10" in o1 and "Restarting from global step 10" in o2
    _compare(rows(o3), rows(o1), MM_TOL)
    _compare(rows(o2), rows(o1), MM_TOL)


def test_mm_driver_physical_stays_bounded(tmp_path):
    """--mm=physical: the exact-gradient field keeps the crystal near its
    target temperature where the reference's forces heat it without bound."""
    rc, out = run(MMD, ["--cell=2", "--temp=300", "--step=400", "--mon=100", "--mm=physical"],
                  tmp_path)
    assert rc == 0 and "Done" in out
    r = rows(out)
    assert len(r) == 4 and max(x[2][0] for x in r) < 600.0


AB_REF = os.path.join(REF, "fuller_airebo_npt_md_serial")
needs_ab_ref = pytest.mark.skipif(not os.path.exists(AB_REF),
                                  reason="reference AIREBO build (oracle/_ref) not present")


@needs_ab_ref
def test_airebo_restart_files_cross_load(tmp_path):
    """write_restart_airebo / read_restart_airebo (AB:629-677): our AIREBO
    restart resumed by the reference, and the reference's by us, give the
    same next rows; the two step-10 files hold the same atoms."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    base = ["--cell=2", "--temp=300", "--step=20", "--mon=5", "--restart=10", "--seed=4"]
    assert run(AB, base, a)[0] == 0
    assert run(AB_REF, base, b)[0] == 0
    ours = sorted(p for p in os.listdir(a) if p.endswith("_10.rst"))[0]
    theirs = sorted(p for p in os.listdir(b) if p.endswith("_10.rst"))[0]
    la = [l.split() for l in open(a / ours) if l.startswith("ATOM")]
    lb = [l.split() for l in open(b / theirs) if l.startswith("ATOM")]
    assert len(la) == len(lb) == 1920
    # "ATOM index mol_id mass x y z vx vy vz": tag, index, molecule and mass
    # text equal, then positions and velocities
    assert [x[:4] for x in la] == [x[:4] for x in lb] and {len(x) for x in la} == {10}
    va = np.array([[float(v) for v in x[4:]] for x in la])
    vb = np.array([[float(v) for v in x[4:]] for x in lb])
    assert np.abs(va[:, :3] - vb[:, :3]).max() < 1e-6  # positions, A
    assert np.abs(va[:, 3:] - vb[:, 3:]).max() < 1e-5 * np.abs(vb[:, 3:]).max()
    tags = lambda p: [l.split()[0] for l in open(p) if l.strip() and not l.startswith("#")]
    assert tags(a / ours) == tags(b / theirs)
    cont = ["--cell=2", "--temp=300", "--step=20", "--mon=5", "--seed=4"]
    _, o1 = run(AB_REF, [*cont, f"--resfile={a / ours}"], tmp_path)
    _, o2 = run(AB, [*cont, f"--resfile={b / theirs}"], tmp_path)
    _, o3 = run(AB, [*cont, f"--resfile={a / ours}"], tmp_path)
    assert "Restarting from global step 10" in o1 and "Restarting from global step 10" in o2
    _compare(rows(o3), rows(o1), AB_TOL)
    _compare(rows(o2), rows(o1), AB_TOL)


def _xyz_frames(path):
    lines = open(path).read().splitlines()
    frames, i = [], 0
    while i < len(lines):
        n = int(lines[i])
        frames.append((n, lines[i + 1], lines[i + 2:i + 2 + n]))
        i += 2 + n
    return frames


ATOM_LINE = re.compile(r"^C +-?\d+\.\d{8} +-?\d+\.\d{8} +-?\d+\.\d{8} +\d+ +"
                       r"-?\d\.\d{8}e[-+]\d\d +-?\d\.\d{8}e[-+]\d\d +-?\d\.\d{8}e[-+]\d\d$")


def _compare_ovito(ours, theirs, pos_tol, vel_rel):
    """Field by field: atom count, header text apart from the lattice
    numbers, lattice, species, positions (minimum image), molecule ids and
    velocities, each line in the reference's format."""
    fa, fb = _xyz_frames(ours), _xyz_frames(theirs)
    assert len(fa) == len(fb) > 0
    for (na, ha, ba), (nb, hb, bb) in zip(fa, fb):
        assert na == nb == len(ba) == len(bb)
        ma = re.match(r'Lattice="([^"]+)" (.*)$', ha)
        mb = re.match(r'Lattice="([^"]+)" (.*)$', hb)
        assert ma and mb and ma.group(2) == mb.group(2)  # Properties, Time, Step, pbc
        lat_a = np.array([float(x) for x in ma.group(1).split()])
        lat_b = np.array([float(x) for x in mb.group(1).split()])
        assert np.abs(lat_a - lat_b).max() < 1e-6
        assert all(ATOM_LINE.match(l) for l in ba), ba[0]
        A = np.array([[float(x) for x in l.split()[1:]] for l in ba])
        B = np.array([[float(x) for x in l.split()[1:]] for l in bb])
        assert np.array_equal(A[:, 3], B[:, 3])  # molecule ids
        box = lat_b.reshape(3, 3)  # rows: the cell vectors
        d = A[:, :3] - B[:, :3]
        s = d @ np.linalg.inv(box)
        d = (s - np.round(s)) @ box
        assert np.abs(d).max() < pos_tol
        assert np.abs(A[:, 4:] - B[:, 4:]).max() < vel_rel * np.abs(B[:, 4:]).max()


@needs_ref
def test_lj_ovito_file_equals_reference(tmp_path):
    """write_ovito (LJ:541-555): the same frames as the reference's file."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    args = ["--cell=2", "--temp=300", "--step=40", "--mon=20", "--ovito=20", "--seed=9"]
    assert run(LJ, args, a)[0] == 0
    assert run(os.path.join(REF, "fuller_LJ_npt_md_serial"), args, b)[0] == 0
    fa = [p for p in os.listdir(a) if p.startswith("ovito_traj")]
    fb = [p for p in os.listdir(b) if p.startswith("ovito_traj")]
    assert len(fa) == len(fb) == 1
    _compare_ovito(a / fa[0], b / fb[0], 1e-6, 1e-5)


@needs_ab_ref
def test_airebo_ovito_file_equals_reference(tmp_path):
    """write_ovito (AB:593-601): the same frames as the reference's file."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    args = ["--cell=2", "--temp=300", "--step=20", "--mon=10", "--ovito=10", "--seed=6"]
    assert run(AB, args, a)[0] == 0
    assert run(AB_REF, args, b)[0] == 0
    fa = [p for p in os.listdir(a) if p.startswith("ovito_traj")]
    fb = [p for p in os.listdir(b) if p.startswith("ovito_traj")]
    assert len(fa) == len(fb) == 1
    _compare_ovito(a / fa[0], b / fb[0], 1e-6, 1e-5)
