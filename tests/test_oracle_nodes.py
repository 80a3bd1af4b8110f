"""The oracle's gas optics, cloud optics, setcoef weights and SW two-stream, pinned on table nodes and
mathematical limits (tests/oracle_checks.py), and a guard that these pins have teeth: copies of the
oracle with one planted mistake each (a constant, an index, an operand) are built with gcc and each
must fail at least one pin.  CPU only.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

import oracle_checks as oc
from oracle import pyoracle

ORACLE = Path(pyoracle.__file__).resolve().parent
INCLUDE = ORACLE.parent / "include"


@pytest.mark.parametrize("check", oc.ALL, ids=lambda f: f.__name__)
def test_oracle_pin(check):
    check(pyoracle.lib())


# (file, text, replacement, what): each is a plausible slip in the restatement of the reference
MUTATIONS = [
    ("oracle_lw.c", "1.0 - 0.15 * (250.0 - pp) / 154.4", "1.0 - 0.25 * (250.0 - pp) / 154.4",
     "band-1 corradj constant (phys_rrtm_lw_taugb01.cu)"),
    ("oracle_sw.c", "const double zgam2 = 0.75 * zssa1", "const double zgam2 = 0.70 * zssa1",
     "PIFM gamma2 (phys_rrtm_sw_spcvrtc.cu:99-102)"),
    ("oracle_lw.c", "f10 * T2(tab, (i0) + 1, g, ng) + f01 * T2(tab, i1, g, ng)",
     "f01 * T2(tab, (i0) + 1, g, ng) + f10 * T2(tab, i1, g, ng)", "LW interpolation weights on swapped corners"),
    ("oracle_sw.c", "B->rayla[(js - 1) * 8 + g], ray1 = B->rayla[js * 8 + g]",
     "B->rayla[js * 8 + g], ray1 = B->rayla[(js - 1) * 8 + g]", "band-24 Rayleigh eta row"),
    ("oracle_lw.c", "const double a1 = t->absliq1[b * 58 + (index - 1)];",
     "const double a1 = t->absliq1[b * 58 + index];", "LW liquid cloud radius index"),
    ("oracle_sw.c", "(dgeice - 2.0) / 3.0", "(dgeice - 1.0) / 3.0", "Fu ice generalized-size index"),
    ("oracle_lw.c", "c->fac10[k] = tem1 * ft;", "c->fac10[k] = tem1 * ft1;", "LW setcoef fac10 temperature weight"),
    ("oracle_sw.c", "tr[g] = cm * 6.86e-10", "tr[g] = cm * 6.86e-11", "band-17 Rayleigh coefficient"),
    ("oracle_lw.c", "0.943", "0.934", "band-4 upper g-point factor"),
    ("oracle_sw.c", "id[2] = imin(nrow, id[0] + d1);", "id[2] = imin(nrow, id[0] + d1 - 1);",
     "SW eta-table temperature stride"),
    ("oracle_lw.c", "r->aerlw[0][b][kin] * (1.0 - r->aerlw[1][b][kin])", "r->aerlw[0][b][kin] * r->aerlw[1][b][kin]",
     "LW aerosol absorption tau (1 - ssa)"),
    ("oracle_sw.c", "const double zasy0 = a_y * a_s * a_t;", "const double zasy0 = a_y * a_t;",
     "SW aerosol asymmetry weight"),
    ("oracle_sw.c", "const double zssa0 = tr[k - 1] + a_t * a_s;", "const double zssa0 = tr[k - 1] + a_t;",
     "SW aerosol scattering depth"),
]


def _build_copy(tmp: Path, fname: str, old: str, new: str) -> Path:
    src, inc = tmp / "oracle", tmp / "include"
    src.mkdir()
    inc.mkdir()
    for p in ORACLE.iterdir():
        if p.suffix in (".c", ".h"):
            shutil.copy(p, src / p.name)
    shutil.copy(INCLUDE / "woofrad.h", inc / "woofrad.h")
    text = (src / fname).read_text()
    assert text.count(old) == 1, f"mutation anchor {old!r} must occur once in {fname}"
    (src / fname).write_text(text.replace(old, new))
    srcs = sorted(p.name for p in src.glob("*.c"))
    lib = src / "liboracle_mut.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-shared", "-o", str(lib),
                    *srcs, "-lm"], cwd=src, check=True, capture_output=True)
    return lib


@pytest.mark.parametrize("fname,old,new,what", MUTATIONS, ids=[m[3] for m in MUTATIONS])
def test_planted_mistake_fails_a_pin(tmp_path, fname, old, new, what):
    L = pyoracle.load(_build_copy(tmp_path, fname, old, new))
    failed = []
    for check in oc.ALL:
        try:
            check(L)
        except AssertionError as e:
            failed.append(f"{check.__name__}: {e}")
    assert failed, f"no pin catches the planted mistake: {what}"
