"""CalculiX path on the host: the deck oracle against decks the reference
wrote (tests/golden/make_golden_inp.py), the reference's node-count KATs,
and the product's .frd reader on the reference's own synthetic .frd
fixtures (tests/test_voxel_fem.py:139-213)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import inp_oracle

GOLD = Path(__file__).resolve().parent / "golden" / "inp"


def _cases():
    cases = json.loads((GOLD / "cases.json").read_text())
    with np.load(GOLD / "fields.npz", allow_pickle=False) as d:
        fields = {k: np.array(d[k]) for k in d.files}
    return cases, fields


def test_oracle_deck_equals_reference_decks():
    cases, fields = _cases()
    assert sum(c["error"] is None for c in cases.values()) >= 5
    for name, c in cases.items():
        if c["error"] is None:
            assert inp_oracle.deck(fields[name], **c["kwargs"]) == (GOLD / f"{name}.inp").read_text(), name
        else:
            with pytest.raises(ValueError, match=c["error"].split("'")[0][:20]):
                inp_oracle.deck(fields[name], **c["kwargs"])


def test_oracle_node_count_kats():
    """tests/test_voxel_fem.py:107-135: a 2x2x2 solid cube has (2+1)^3
    nodes, one voxel 8."""
    assert len(inp_oracle.mesh(np.ones((2, 2, 2)))[1]) == 27
    assert len(inp_oracle.mesh(np.ones((1, 1, 1)))[1]) == 8


def _fmt(v):
    return f"{v:12.6E}"


def _rec(node, *vals):
    return f" -1{node:10d}" + "".join(_fmt(v) for v in vals)


MINIMAL_FRD = "\n".join([
    "     2C  minimal_test", " -4  DISP        4    1",
    " -5  D1          1    1    0    0", " -5  D2          1    2    0    0",
    " -5  D3          1    3    0    0", " -5  ALL         1    0    0    0",
    _rec(1, 1.0e-2, 2.0e-2, 3.0e-2), _rec(2, 4.0e-2, 5.0e-2, 6.0e-2), " -3",
    " -4  STRESS      6    1", " -5  SXX         1    4    1    1",
    " -5  SYY         1    4    1    2", " -5  SZZ         1    4    1    3",
    " -5  SXY         1    4    1    4", " -5  SYZ         1    4    1    5",
    " -5  SZX         1    4    1    6", _rec(1, 1.0e2, 5.0e1, 2.5e1, 1.0e1, 5.0e0, 0.0), " -3",
    "  9999"])


def test_parse_frd_minimal(tmp_path):
    from genpipeline_amd.fem.voxel_hex import VoxelHexMesher
    p = tmp_path / "t.frd"
    p.write_text(MINIMAL_FRD)
    r = VoxelHexMesher._parse_frd(str(p))
    vm = math.sqrt(0.5 * ((100 - 50) ** 2 + (50 - 25) ** 2 + (25 - 100) ** 2 + 6 * (10 ** 2 + 5 ** 2)))
    assert r["stress_max"] == pytest.approx(vm, rel=1e-6) and abs(vm - 68.9) < 0.1
    assert r["displacement_max"] == pytest.approx(math.sqrt(77) * 1e-2, rel=1e-6)
    assert r["compliance"] == pytest.approx((math.sqrt(14) + math.sqrt(77)) * 1e-2, rel=1e-6)
    assert r.get("failure_reason") is None


def test_parse_frd_malformed_and_missing(tmp_path):
    from genpipeline_amd.fem.voxel_hex import VoxelHexMesher
    from genpipeline_amd.pipeline_utils import FEM_SENTINEL
    bad = "\n".join(["     2C  corrupt_test", "-4  DISP        4    1",
                     "-5  D1          1    1    0    0", "-1         1 NOT_A_NUMBER BAD BAD", "-3",
                     "-4  STRESS      6    1", "-5  SXX         1    4    1    1",
                     "-1         1 " + "@" * 82, "-3", " 9999"]) + "\n"
    p = tmp_path / "bad.frd"
    p.write_text(bad)
    r = VoxelHexMesher._parse_frd(str(p))
    assert r["stress_max"] == 0.0 and r["displacement_max"] == 0.0
    r = VoxelHexMesher._parse_frd(str(tmp_path / "none.frd"))
    assert r["stress_max"] == FEM_SENTINEL and r["failure_reason"] == "no_frd"


def test_run_ccx_without_ccx_raises(tmp_path, monkeypatch):
    from genpipeline_amd.fem import voxel_hex
    monkeypatch.setattr(voxel_hex, "find_ccx", lambda: None)
    with pytest.raises(RuntimeError, match="not found"):
        voxel_hex.VoxelHexMesher.run_ccx(str(tmp_path / "x.inp"))


def _fake_ccx(tmp_path, body):
    exe = tmp_path / "ccx_stub.py"
    exe.write_text("#!" + __import__("sys").executable + "\n" + body)
    exe.chmod(0o755)
    return str(exe)


def test_run_ccx_parses_the_frd_it_writes(tmp_path):
    """run_ccx's process handling (voxel_fem.py:228-302) with a stand-in
    solver that writes the minimal .frd beside the deck: parsed results,
    then the ccx_error, timeout and no_frd sentinels."""
    from genpipeline_amd.fem.voxel_hex import VoxelHexMesher
    from genpipeline_amd.pipeline_utils import FEM_SENTINEL
    inp = tmp_path / "deck.inp"
    inp.write_text("*HEADING\n")
    ok = _fake_ccx(tmp_path, "import sys\nopen(sys.argv[1] + '.frd', 'w').write(%r)\n" % MINIMAL_FRD)
    r = VoxelHexMesher.run_ccx(str(inp), ccx_cmd=ok)
    assert r["failure_reason"] is None and abs(r["stress_max"] - 68.9) < 0.1
    bad = _fake_ccx(tmp_path, "import sys\nsys.exit(3)\n")
    assert VoxelHexMesher.run_ccx(str(inp), ccx_cmd=bad)["failure_reason"] == "ccx_error"
    slow = _fake_ccx(tmp_path, "import time\ntime.sleep(30)\n")
    r = VoxelHexMesher.run_ccx(str(inp), ccx_cmd=slow, timeout=1)
    assert r["failure_reason"] == "timeout" and r["stress_max"] == FEM_SENTINEL
    (tmp_path / "deck.frd").unlink()
    quiet = _fake_ccx(tmp_path, "pass\n")
    assert VoxelHexMesher.run_ccx(str(inp), ccx_cmd=quiet)["failure_reason"] == "no_frd"
