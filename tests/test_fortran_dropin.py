"""The literal Fortran drop-ins (advectionhpctester_amd/fortran/
mpdata_dropin.F90): modules with the reference's stage-module names whose
subroutines carry the reference's abstract interfaces.  oracle/build_ref.sh
compiles the reference's own advec.f90 against them (its procedure-pointer
assignments upwind3d_ptr => upwind3d_GPUBC, antidiff3d_ptr =>
antidiff3d_{gauge,standard}_GPUBC, advec.f90:143-144, 196-203, and its calls
through them); that build exists only where the reference tree and the HIP
library are present.  A mismatched interface fails that compilation (e.g. a
missing optional argument: "Procedure pointer 'antidiff3d_ptr' associated
with incompatible procedure designator ... distinct numbers of dummy
arguments")."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN_SO = os.path.join(ROOT, "oracle", "_ref", "libmpdata_ref_dropin.so")


def _symbols(path):
    nm = shutil.which("nm") or "/usr/bin/nm"
    out = subprocess.run([nm, "-D", path], capture_output=True, text=True, check=True).stdout
    return out


def test_reference_advec_compiled_against_dropins():
    if not os.path.isdir("/root/reference/advection"):
        pytest.skip("the reference tree is not present (the build needs its sources)")
    if not os.path.exists(DROPIN_SO):
        pytest.fail(f"{DROPIN_SO} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    syms = _symbols(DROPIN_SO)
    # the reference's driver, and the drop-in modules in place of its stages
    # flang's names of module procedures: _QM<module>P<procedure>
    assert "_QMadvec_driverPadvec_dwarf" in syms
    for mod, sub in (("module_upwind3d_gpubc", "upwind3d_gpubc"),
                     ("module_antidiff3d_gauge_gpubc", "antidiff3d_gauge_gpubc"),
                     ("module_antidiff3d_standard_gpubc", "antidiff3d_standard_gpubc")):
        assert f"_QM{mod}P{sub}" in syms, mod
    assert "_QMmpdata_dropin_supportPdev_of" in syms
    # the stages run in the HIP library, which the build links
    readelf = shutil.which("readelf") or "/usr/bin/readelf"
    dyn = subprocess.run([readelf, "-d", DROPIN_SO], capture_output=True, text=True,
                         check=True).stdout
    assert "libmpdata_hip.so" in dyn


def test_dropin_interfaces_match_the_reference_declarations():
    """The dummy declarations of the drop-ins, argument for argument, as
    the reference's CPU build declares them (kind, intent, shape; read from
    the source text)."""
    src = open(os.path.join(ROOT, "advectionhpctester_amd", "fortran", "mpdata_dropin.F90")).read()
    for sub, args in (
            ("upwind3d_gpubc", "u1, u2, u3, x_in, xant, xforc_impl, xforc_expl, rhr, h, bcx, bcy, "
                               "ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, dt, "
                               "do_serialization_in, do_serialization_out"),
            ("antidiff3d_gauge_gpubc", "lupdatemulti, u1, u2, u3, xant, x_in, rhr, h, iflip, "
                                       "ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, "
                                       "do_serialization_out"),
            ("antidiff3d_standard_gpubc", "lupdatemulti, u1, u2, u3, xant, x_in, rhr, h, iflip, "
                                          "ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, "
                                          "do_serialization_out")):
        i = src.index(f"SUBROUTINE {sub}(")
        head = src[i:src.index(")", i)]
        got = [a.strip() for a in head[len(f"SUBROUTINE {sub}("):].replace("&", "").split(",")]
        assert got == [a.strip() for a in args.split(",")], sub
    assert "x_in(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)" in src
    assert "u1(1 - ih:np + 1 + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)" in src
    assert "bcx(mp, lp, 2)" in src and "bcy(np, lp, 2)" in src
