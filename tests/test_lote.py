"""The batch program lote.py (GPU counterpart of the reference's rodar_lote.py).

CPU: the batch tables give the reference's case counts, and the case directories it writes hold the
same input files, byte for byte, as the reference's committed case directories
(tests/golden/driver_cases.json); the resume rule; --lista without a device.
GPU: a small batch runs as one batched invocation and every case finishes with resumo.dat.
"""
import importlib.util
import os
import subprocess
import sys
from pathlib import Path

import pytest

import _cases as K

PKG = Path(K.PKG)
spec = importlib.util.spec_from_file_location("lote", PKG / "lote.py")
lote = importlib.util.module_from_spec(spec)
spec.loader.exec_module(lote)


def test_batch_sizes_match_reference():
    sizes = {n: len(f()) for n, f in lote.BATCHES.items()}
    assert sizes == {"E1": 49, "E2": 168, "E2chi": 30, "E4": 280, "E4v": 35, "E3": 42, "E3chi": 15, "S168": 168}


def test_sweep_radii_and_boxes():
    cs = lote.batch_S168()
    radii = [c["radius"] for c in cs]
    assert min(radii) >= 10 and max(radii) <= 50
    assert all(c["nx"] == 4 * c["radius"] and c["steps"] == 20000 for c in cs)
    assert len({(c["tau"], c["sigma"], c["op"]) for c in cs}) == 168


@pytest.mark.parametrize("c", K.golden_cases(), ids=lambda c: c["label"])
def test_case_inputs_equal_reference(tmp_path, c):
    label = c["label"]
    batch = label.split("/")[0]
    match = [x for x in lote.BATCHES[batch]() if str(x["folder"]) == label]
    assert len(match) == 1, label
    d = lote.build(match[0], tmp_path, "f64")
    for name, text in c["inputs"].items():
        assert (d / name).read_text() == text, name
    assert not (d / "itolbm.txt").exists()
    head = (d / "meio.vtk").read_text().splitlines()[4]
    assert head == f"DIMENSIONS {c['nx']} {c['nx']} 1"


def test_resume_rule(tmp_path):
    c = lote.batch_E1()[0]
    d = lote.build(c, tmp_path, "f32")
    assert (d / "itolbm.txt").read_text() == "precisao f32\n"
    assert not lote.done(c, tmp_path)
    (d / "resumo.dat").write_text("# operador estado\n")
    assert not lote.done(c, tmp_path)
    (d / "resumo.dat").write_text("# operador estado\nP1 estavel\n")
    assert lote.done(c, tmp_path) and lote.state_of(d) == "estavel"


def test_list_runs_without_device(tmp_path):
    r = subprocess.run([sys.executable, str(PKG / "lote.py"), "E3chi", "--lista", "--saida", str(tmp_path)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "15 cases requested, 15 to run" in r.stdout


@pytest.mark.gpu
def test_small_batch_runs_batched(tmp_path):
    print("[gpu] lote.py E1 R=10, 1000 steps, f32", flush=True)
    r = subprocess.run([sys.executable, str(PKG / "lote.py"), "E1", "--filtro", "R=10", "--passos", "1000",
                        "--precisao", "f32", "--saida", str(tmp_path)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Imm_OPS: 7 cases" in r.stdout and "7 estavel" in r.stdout
    for c in [x for x in lote.batch_E1() if "R=10" in str(x["folder"])]:
        assert lote.state_of(tmp_path / c["folder"]) == "estavel"
    # resume: a second call finds nothing to do
    r = subprocess.run([sys.executable, str(PKG / "lote.py"), "E1", "--filtro", "R=10", "--passos", "1000",
                        "--saida", str(tmp_path)], capture_output=True, text=True, timeout=600)
    assert "nothing to do" in r.stdout


def test_single_case_directory_gets_its_log(tmp_path):
    """A program given one case directory (a lote.py chunk of one case) writes that case's console output
    to <dir>/execucao.log, as the reference's batch script keeps it (rodar_lote.py:234-235).  Without a
    device the program stops at lbm_create, after the log is open, which is all this checks."""
    d = tmp_path / "um"
    K.make_case(str(d), 16, "P3 0.5", 1.0, 0.225, 10)
    subprocess.run([K.program("Imm_OPS"), str(d)], capture_output=True, text=True, timeout=60)
    assert (d / "execucao.log").exists()
    assert "dat_out.txt" in (d / "execucao.log").read_text()
