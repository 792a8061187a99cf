"""The Fortran host (advectionhpctester_amd/fortran: iso_c_binding module
mpdata_hip, the dwarf driver, the CE_advection and density-step checks) end
to end on the GPU: the 59^3 rotating sphere through the drop-in advec_dwarf,
against the reference's recorded trajectory and norms; eight fields on one
grid through the drop-in CE_advection, against one step per field; three
substepped density steps through the drop-in mpdata3d_rho_gauge_gpubc,
against the oracle on the inputs the program wrote."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from oracle import GAUGE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "advectionhpctester_amd", "fortran", "build", "mpdata_dwarf")
FDIR = os.path.join(ROOT, "advectionhpctester_amd", "fortran")
CE_EXE = os.path.join(FDIR, "build", "mpdata_ce_check")
# the same programs built with -DFLOAT_PRECISION=4 (the _f32 entry points)
BUILDS = [("build", np.float64), ("build32", np.float32)]
GOLDEN = os.path.join(ROOT, "tests", "golden", "dwarf59_reference.txt")

pytestmark = pytest.mark.gpu


def _golden():
    steps, norms = {}, {}
    for line in open(GOLDEN):
        p = line.split()
        if p and p[0] == "step":
            steps[int(p[1])] = p[3]
        elif p[:2] == ["Energy", "L2"] or p[:2] == ["Error", "L2"]:
            norms[p[0]] = float(p[-1])
    return steps, norms


def _run(nt, dump, t_adv=1):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C advectionhpctester_amd/fortran)")
    print(f"fortran dwarf nt={nt} t_adv={t_adv}", flush=True)
    r = subprocess.run([EXE, str(nt), dump, str(t_adv)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = {}
    for line in r.stdout.splitlines():
        p = line.split()
        if p[:2] in (["Energy", "L2"], ["Error", "L2"]):
            vals[p[0]] = float(p[-1])
    return vals, r.stdout


def _snapshots(path, n=59):
    raw = np.fromfile(path, dtype=np.float64)
    return raw.reshape(-1, n, n, n)


def test_fortran_dwarf_matches_reference_trajectory(tmp_path, orc64):
    """Steps 1, 50 and 100 bit-identical to the reference's recorded hashes;
    all 556 steps bit-identical to the oracle, whose norms equal the
    reference's printed ones."""
    want, norms = _golden()
    dump = str(tmp_path / "x.bin")
    vals, out = _run(556, dump)
    snaps = _snapshots(dump)
    assert snaps.shape[0] == 4, out
    for s, step in zip(snaps[:3], (1, 50, 100)):
        assert hashlib.sha256(s.tobytes()).hexdigest() == want[step], f"step {step}"
    # the reference's printed norms (8 decimals)
    assert abs(vals["Energy"] - norms["Energy"]) < 6e-9, out
    assert abs(vals["Error"] - norms["Error"]) < 6e-9, out
    assert round(vals["Error"], 4) == 0.0028
    # the whole run against the oracle (the reference's ground-plane defect
    # first changes a value at step 115, so beyond step 100 the oracle is
    # the bit-exact check)
    from advectionhpctester_amd import testdata
    c = testdata.dwarf_setup()
    x = c.x.copy()
    dims = (c.n, c.n, c.n, 1)
    for _ in range(556):
        orc64.step(GAUGE, 2, 1, x, c.u1, c.u2, c.u3, c.rhr, c.h, c.bcx, c.bcy, (0, 0, 0), dims)
    assert np.array_equal(snaps[3], x[1:-1, 1:-1, 1:-1])


def test_fortran_dwarf_standard_variant(tmp_path, orc64):
    """t_adv = 2 (standard antidiffusion) through the same host, 20 steps,
    bit-identical to the oracle."""
    dump = str(tmp_path / "x.bin")
    _run(20, dump, t_adv=2)
    snaps = _snapshots(dump)
    from advectionhpctester_amd import testdata
    c = testdata.dwarf_setup()
    x = c.x.copy()
    dims = (c.n, c.n, c.n, 1)
    for it in range(1, 21):
        orc64.step(2, 2, 1, x, c.u1, c.u2, c.u3, c.rhr, c.h, c.bcx, c.bcy, (0, 0, 0), dims)
        if it == 1:
            assert np.array_equal(snaps[0], x[1:-1, 1:-1, 1:-1])
    assert np.array_equal(snaps[-1], x[1:-1, 1:-1, 1:-1])


def _timer_rows(out):
    """{item id: (label, calls, seconds)} of the timer tables printed."""
    rows = {}
    for line in out.splitlines():
        p = line.split()
        if len(p) == 7 and p[0] == "0" and p[1].isdigit() and p[3].isdigit():
            rows.setdefault(int(p[1]), (p[2], int(p[3]), float(p[4])))
    return rows


def test_fortran_dwarf_size_sweep_and_timers(tmp_path, orc64):
    """The reference's size sweep (n = 59*iscale, dt/iscale): iscale 2 on
    118^3, 20 steps, bit-identical to the oracle on the same initial state;
    each size prints the reference's timer table with advec_dwarf (item 300)
    and its two-step driver (55) called once per step."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C advectionhpctester_amd/fortran)")
    dump = str(tmp_path / "x2.bin")
    print("fortran dwarf sweep iscale 2..2, 20 steps", flush=True)
    r = subprocess.run([EXE, "20", dump, "1", "2:2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "iscale 2, 20 steps on 118 x 118 x 118" in r.stdout, r.stdout
    rows = _timer_rows(r.stdout)
    assert rows[300][:2] == ("eulagdwarf", 20), r.stdout
    assert rows[55][:2] == ("twos_gau_gpubc", 20), r.stdout
    assert rows[300][2] >= rows[55][2] > 0.0
    snaps = np.fromfile(dump, dtype=np.float64).reshape(-1, 118, 118, 118)
    from advectionhpctester_amd import testdata
    c = testdata.dwarf_setup(iscale=2)
    x = c.x.copy()
    dims = (c.n, c.n, c.n, 1)
    for _ in range(20):
        orc64.step(GAUGE, 2, 1, x, c.u1, c.u2, c.u3, c.rhr, c.h, c.bcx, c.bcy, (0, 0, 0), dims)
    assert np.array_equal(snaps[-1], x[1:-1, 1:-1, 1:-1])
    # two sizes in one run: both tables, the second on 118^3
    r = subprocess.run([EXE, "3", "-", "1", "1:2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Dwarf statistics summary") == 2
    assert "iscale 1, 3 steps on 59 x 59 x 59" in r.stdout


@pytest.mark.parametrize("build", ["build", "build32"])
def test_fortran_ce_advection_equals_single_steps(build):
    """CE_advection with th, u, v, w, tht, pstr, qv, qc (five gauge, three
    standard fields) gives each field the bits of mpdata_step on it, in the
    FP64 and the FP32 build of the host."""
    exe = os.path.join(FDIR, build, "mpdata_ce_check")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    print(f"fortran CE_advection check ({build})", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "CE check: PASS" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("build,npdt", BUILDS)
def test_fortran_rho_gauge_matches_oracle(tmp_path, build, npdt):
    """mpdata3d_rho_gauge_gpubc through the Fortran module (FP64 and FP32
    builds), three calls with substepping: every array and the INOUT
    scalars after each call equal the oracle's on the inputs the program
    wrote."""
    from oracle import Oracle, shape
    exe = os.path.join(FDIR, build, "mpdata_rho_check")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    orc = Oracle(npdt)
    dims = (26, 21, 12, 1)
    np_, mp, lp, ih = dims
    dump = str(tmp_path / "rho")
    print(f"fortran density step check ({build})", flush=True)
    r = subprocess.run([exe, dump], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "rho check: done" in r.stdout, r.stdout + r.stderr
    states = [line.split()[3:] for line in r.stdout.splitlines() if line.startswith("call ")]
    shapes = [(3,) + shape(*dims)] * 6 + [shape(*dims, 1), shape(*dims, 2), shape(*dims, 3)] + \
        [shape(*dims)] * 8 + [(2, lp, mp), (2, lp, np_)]

    def load(path):
        raw = np.fromfile(path, dtype=npdt)
        out, o = {}, 0
        for name, sh in zip(Oracle.RHO_FIELDS, shapes):
            n = int(np.prod(sh))
            out[name] = raw[o:o + n].reshape(sh).copy()
            o += n
        assert o == raw.size
        return out

    f = load(dump + ".in")
    got = load(dump + ".out")
    f.update(dxi=0.01, dyi=0.01, dzi=0.01, dt=40.0)
    st = dict(nsubsteps=1, nsubsteps_o=1, dtn=40.0, dtnold=40.0)
    uint = np.uint64 if npdt == np.float64 else np.uint32
    for c in range(3):
        st = orc.rho_gauge(f, {"lsubstepping": 1}, st, 0.9, 80.0, 0, dims)
        assert st["status"] == 0
        assert [int(states[c][0]), int(states[c][1])] == [st["nsubsteps"], st["nsubsteps_o"]]
        assert [float(states[c][2]), float(states[c][3])] == [st["dtn"], st["dtnold"]]
    assert st["nsubsteps"] == 2
    for name in Oracle.RHO_FIELDS:
        assert np.array_equal(got[name].view(uint), f[name].view(uint)), name


def test_fortran_dwarf_fp32_build(tmp_path):
    """The rotating-sphere dwarf through the FP32 build of the host (the
    reference's -DFLOAT_PRECISION=4 switch, the _f32 entry points): 556 steps,
    norms within FP32 rounding of the reference's FP64 ones."""
    exe = os.path.join(FDIR, "build32", "mpdata_dwarf")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    _, norms = _golden()
    print("fortran dwarf FP32 nt=556", flush=True)
    r = subprocess.run([exe, "556", "-", "1"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = {}
    for line in r.stdout.splitlines():
        p = line.split()
        if p[:2] in (["Energy", "L2"], ["Error", "L2"]):
            vals[p[0]] = float(p[-1])
    print("FP32 norms", vals, "FP64 reference", norms, flush=True)
    # measured: Energy 4.065172736e-3, Error 2.810699666e-3 against the
    # reference's printed 0.00406518 and 0.0028107 (FP64)
    assert abs(vals["Energy"] - norms["Energy"]) < 1e-7, r.stdout
    assert abs(vals["Error"] - norms["Error"]) < 1e-7, r.stdout


DROPIN_SO = os.path.join(ROOT, "oracle", "_ref", "libmpdata_ref_dropin.so")


def test_reference_advec_dwarf_through_literal_dropins(orc64):
    """The reference's own advec_dwarf (advec.f90, compiled from its source
    with its procedure pointers bound to the literal drop-in modules of
    advectionhpctester_amd/fortran/mpdata_dropin.F90) on the reference's own
    59^3 initial state (set_initial15) with host arrays: the stages run in
    the HIP library, and the trajectory equals the reference's recorded
    hashes at steps 1, 50 and 100; the standard variant and mode 6 equal the
    oracle."""
    if not os.path.exists(DROPIN_SO):
        pytest.skip("oracle/_ref/libmpdata_ref_dropin.so not built (needs the reference tree)")
    from oracle import Reference, shape
    from advectionhpctester_amd import testdata
    want, _ = _golden()
    ref = Reference(path=DROPIN_SO)
    c = testdata.dwarf_setup()
    n = c.n
    ref.setup(n, n, n, 1, dt=c.dt, nt=c.nt)
    u = [np.zeros(shape(n, n, n, 1, k)) for k in (1, 2, 3)]
    x0 = np.zeros(shape(n, n, n, 1))
    ref.set_initial15(*u, x0, np.ones(shape(n, n, n, 1)))
    x = x0.copy()
    print("reference advec_dwarf through the drop-ins, 100 gauge steps", flush=True)
    for it in range(1, 101):
        ref.advec_dwarf(4, 1, x, c.bcx, c.bcy, *u, c.rhr, c.h)
        if it in want:
            got = hashlib.sha256(np.ascontiguousarray(x[1:-1, 1:-1, 1:-1]).tobytes()).hexdigest()
            assert got == want[it], f"step {it}"
    dims = (n, n, n, 1)
    for mode, t_adv, variant, steps in ((4, 2, 2, 10), (6, 1, 0, 5)):
        print(f"reference advec_dwarf through the drop-ins, mode {mode} t_adv {t_adv}", flush=True)
        xr, xo = x0.copy(), x0.copy()
        for _ in range(steps):
            ref.advec_dwarf(mode, t_adv, xr, c.bcx, c.bcy, *u, c.rhr, c.h)
            orc64.step(variant, 2, 1, xo, *u, c.rhr, c.h, c.bcx, c.bcy, (0, 0, 0), dims)
        I = (slice(1, -1),) * 3
        assert np.array_equal(xr[I], xo[I]), (mode, t_adv)
